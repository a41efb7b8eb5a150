"""CPU restatements of STAT_PARTICLE's level-1 statistics
(stat_particle.f90:169-450) and of STAT_PARTFLUID (stat_partfluid.f90:
113-381) -- TEST INFRASTRUCTURE: the checkers of the library's
gp_particle_stats and gp_partfluid_stats, never part of the product.

For each species row s (one per class J and size index IDP) it returns the
41 entries of MEAN_PART(:, J, IDP): sums over the particles of the row, two
passes (means first, then the products of the fluctuations about them),
entries 10, 39, 40, 41 formed from the sums, everything divided by a
count, entry 23 that count (the reference's count: the last class's, see
particle_stats).  The angular velocity is also rotated to the lab
frame and the inertia tensor to R diag(I) R^T as transform_basis.f90:43-77
does.  Pinned against the reference's own STAT_PARTICLE, compiled and run by
oracle/_ref/part_io_ref (tests/test_particle_stats.py)."""
from __future__ import annotations

import numpy as np

from gazpart_amd.workload import SP_APR, SP_EMAJ, SP_GRAV, SP_IXX, SP_IYY, SP_IZZ, SP_RHOP

NSTAT = 41


def rotation(q: np.ndarray) -> np.ndarray:
    """R (n, 3, 3) of transform_basis.f90:43-53 for quaternions q (4, n)."""
    q0, q1, q2, q3 = q
    R = np.empty((q.shape[1], 3, 3))
    R[:, 0, 0] = 1.0 - 2.0 * (q2 * q2 + q3 * q3)
    R[:, 0, 1] = 2.0 * (q1 * q2 - q0 * q3)
    R[:, 0, 2] = 2.0 * (q1 * q3 + q0 * q2)
    R[:, 1, 0] = 2.0 * (q1 * q2 + q0 * q3)
    R[:, 1, 1] = 1.0 - 2.0 * (q1 * q1 + q3 * q3)
    R[:, 1, 2] = 2.0 * (q2 * q3 - q0 * q1)
    R[:, 2, 0] = 2.0 * (q1 * q3 - q0 * q2)
    R[:, 2, 1] = 2.0 * (q2 * q3 + q0 * q1)
    R[:, 2, 2] = 1.0 - 2.0 * (q1 * q1 + q2 * q2)
    return R


def particle_stats(state: np.ndarray, species_id: np.ndarray, species: np.ndarray,
                   npoly: int | None = None, ref_norm: bool = True) -> np.ndarray:
    """state (13, np) [x y z u v w qa qb qc qd ox oy oz], species_id (np,),
    species rows (nspec, stride), class-major with npoly size indices per
    class -> (nspec, 41).  ref_norm: every class divided by the last class's
    count per size index, as the reference's NPM(IDP) after its class loop
    (stat_particle.f90:436-449)."""
    nspec = species.shape[0]
    npoly = npoly or nspec
    counts = np.bincount(species_id, minlength=nspec).astype(np.float64)
    u = state[3:6]
    w = state[10:13]
    R = rotation(state[6:10])
    g = np.einsum("nij,jn->in", R, w)                       # R omega (:216-225)
    out = np.zeros((nspec, NSTAT))
    for s in range(nspec):
        m = species_id == s
        n = float(m.sum())
        a = np.zeros(NSTAT)
        a[22] = n
        a[0:3] = u[:, m].sum(axis=1)
        a[10:13] = w[:, m].sum(axis=1)
        a[13:16] = g[:, m].sum(axis=1)
        um, wm, gm = a[0:3] / n, a[10:13] / n, a[13:16] / n
        uf = u[:, m] - um[:, None]                          # :270-303
        a[3], a[4], a[5] = (uf[0] * uf[0]).sum(), (uf[1] * uf[1]).sum(), (uf[2] * uf[2]).sum()
        a[6], a[7], a[8] = (uf[0] * uf[1]).sum(), (uf[0] * uf[2]).sum(), (uf[1] * uf[2]).sum()
        a[9] = 0.5 * (a[3] + a[4] + a[5])                   # :306-312
        of = w[:, m] - wm[:, None]                          # :318-352
        ixx, iyy, izz = species[s, SP_IXX], species[s, SP_IYY], species[s, SP_IZZ]
        a[16], a[17], a[18] = (of[0] * of[0]).sum(), (of[1] * of[1]).sum(), (of[2] * of[2]).sum()
        a[19], a[20], a[21] = (of[0] * of[1]).sum(), (of[0] * of[2]).sum(), (of[1] * of[2]).sum()
        a[23] = (ixx * of[0] * of[0]).sum()
        a[24] = (iyy * of[1] * of[1]).sum()
        a[25] = (izz * of[2] * of[2]).sum()
        gf = g[:, m] - gm[:, None]                          # :356-409
        Rm = R[m]
        ib = np.array([ixx, iyy, izz])

        def il(i, j):
            return (Rm[:, i, 0] * (ib[0] * Rm[:, j, 0]) + Rm[:, i, 1] * (ib[1] * Rm[:, j, 1]) +
                    Rm[:, i, 2] * (ib[2] * Rm[:, j, 2]))

        a[26], a[27], a[28] = (gf[0] * gf[0]).sum(), (gf[1] * gf[1]).sum(), (gf[2] * gf[2]).sum()
        a[29], a[30], a[31] = (gf[0] * gf[1]).sum(), (gf[0] * gf[2]).sum(), (gf[1] * gf[2]).sum()
        a[32] = (il(0, 0) * gf[0] * gf[0]).sum()
        a[33] = (il(1, 1) * gf[1] * gf[1]).sum()
        a[34] = (il(2, 2) * gf[2] * gf[2]).sum()
        a[35] = (il(0, 1) * gf[0] * gf[1]).sum()
        a[36] = (il(0, 2) * gf[0] * gf[2]).sum()
        a[37] = (il(1, 2) * gf[1] * gf[2]).sum()
        a[38] = 0.5 * (ixx * a[16] + iyy * a[17] + izz * a[18])   # :414-427
        a[39] = 0.5 * (a[23] + a[24] + a[25])
        a[40] = a[9] + a[39]
        nn = counts[nspec - npoly + s % npoly] if ref_norm else n
        a = a / nn                                          # :436-449
        a[22] = nn
        out[s] = a
    return out


ZERO = 1.0e-33     # mod_dns_dim.f90:136
# Zastawny et al. (2012) fits fixed by stat_partfluid.f90:63-100
PF_A = (1.95, 18.12, 1.023, 4.26, 0.384, 21.52, 0.990, 2.86, 0.260)
PF_B = (None, 0.083, -0.21, 1.582, 0.851, 1.842, -0.802, -0.006, 0.874, 0.009, 0.570)
PF_C = (None, 0.935, 0.146, -0.469, 0.145, 0.116, 0.748, 0.041, 0.221, 0.657, 0.044)
PF_R = (0.573, -0.154, 116.61, 1.0)
PF_R_ = (1.244, 0.239, 378.12, 0.789)
NPFSTAT = 36


def partfluid_stats(state: np.ndarray, fluid: np.ndarray, species_id: np.ndarray,
                    species: np.ndarray, visc: float, rhof: float, npoly: int | None = None,
                    ref_norm: bool = True) -> np.ndarray:
    """STAT_PARTFLUID: state (13, np), fluid (6, np) = UFAP..VORTZAP of the
    last advance, species rows (nspec, stride) class-major -> (nspec, 36) =
    MEAN_PARTFLUID(1:36, J, IDP), each row's sums started from ZERO (:114)
    and divided by NPM(IDP) (the last class's counts with ref_norm,
    :362-381)."""
    a, b, c = PF_A, PF_B, PF_C
    nspec = species.shape[0]
    npoly = npoly or nspec
    sp = species[species_id]
    rhop, emaj, apr = sp[:, SP_RHOP], sp[:, SP_EMAJ], sp[:, SP_APR]
    pm = rhop * (4.0 / 3.0) * np.pi * emaj ** 3.0 / apr ** 2.0              # :124
    dq = 2.0 * emaj / apr ** (2.0 / 3.0)                                    # :135
    R = rotation(state[6:10])
    up, om = state[3:6], state[10:13]
    ur = fluid[0:3] - up
    rep = np.sqrt((ur * ur).sum(axis=0)) * dq / visc                        # :138
    u = np.einsum("nji,jn->in", R, ur)                                      # by conj_q
    s23 = np.sqrt(u[1] * u[1] + u[2] * u[2])
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = np.abs(np.arctan(s23 / u[0]))                                 # :148
    um = np.sqrt((u * u).sum(axis=0))
    sphi, cphi = np.sin(phi), np.cos(phi)
    cd0 = a[1] / rep ** a[2] + a[3] / rep ** a[4]
    cd90 = a[5] / rep ** a[6] + a[7] / rep ** a[8]
    cd = cd0 + (cd90 - cd0) * sphi ** a[0]                                  # :160
    cl = ((b[1] / rep ** b[2] + b[3] / rep ** b[4]) * sphi ** (b[5] + b[6] * rep ** b[7]) *
          cphi ** (b[8] + b[9] * rep ** b[10]))
    k = 0.5 * (rhof / pm) * 0.25 * (np.pi * dq ** 2)
    fd = k * cd * um * u                                                    # :169
    fl = np.stack([k * cl * um ** 2 * sphi * np.copysign(1.0, -u[0]),
                   k * cl * um ** 2 * cphi * (u[1] / s23),
                   k * cl * um ** 2 * cphi * (u[2] / s23)])
    fdl = np.einsum("nij,jn->in", R, fd)
    fll = np.einsum("nij,jn->in", R, fl)
    cu = np.einsum("nji,jn->in", R, fluid[3:6])                             # :220-225
    orl = 0.5 * cu - om
    orm = np.sqrt((orl * orl).sum(axis=0))
    oyz = np.sqrt(orl[2] ** 2 + orl[1] ** 2)
    rer = orm * dq ** 2 / visc
    ctp = ((c[1] / rep ** c[2] + c[3] / rep ** c[4]) * sphi ** (c[5] + c[6] * rep ** c[7]) *
           cphi ** (c[8] + c[9] * rep ** c[10]))
    ctr = PF_R[0] * rer ** PF_R[1] + PF_R[2] / rer ** PF_R[3]
    ctr2 = PF_R_[0] * rer ** PF_R_[1] + PF_R_[2] / rer ** PF_R_[3]
    kp = 0.25 * (rhof / pm) * 0.25 * np.pi * dq ** 3 * ctp * um * um        # :253-258
    tp = np.stack([np.zeros_like(kp), kp * (np.abs(u[2]) / s23) * np.copysign(1.0, u[0] * u[2]),
                   kp * (np.abs(u[1]) / s23) * np.copysign(1.0, u[0] * u[1])])
    kr = (rhof / pm) / 2.0 * (dq / 2.0) ** 5                                # :262-266
    tr = np.stack([ctr * kr * np.abs(orl[0]) * orl[0], ctr2 * kr * oyz * orl[1],
                   ctr2 * kr * oyz * orl[2]])
    minax = emaj / apr                                                      # :299-314
    with np.errstate(invalid="ignore"):
        fbeta = np.where(apr == 1.0, 1.0,
                         apr * np.log(apr + np.sqrt(apr ** 2 - 1)) / np.sqrt(apr ** 2 - 1))
    taupf = (2.0 / 9.0) * (rhop * minax ** 2) / (rhof * visc) * fbeta
    e = np.zeros((NPFSTAT, state.shape[1]))
    e[0], e[1], e[2] = rep, cd, cl
    e[3:6], e[6:9] = fdl, fll
    e[9], e[10] = np.sqrt((fdl * fdl).sum(axis=0)), np.sqrt((fll * fll).sum(axis=0))
    e[11], e[12], e[13] = rer, ctp, ctr
    e[14:17], e[17:20] = tp, tr
    e[20], e[21] = np.sqrt((tp * tp).sum(axis=0)), np.sqrt((tr * tr).sum(axis=0))
    e[22], e[23], e[24] = taupf, sp[:, SP_GRAV], phi
    e[25:28] = R[:, :, 0].T                                                 # R (1, 0, 0)
    e[28] = (0.5 * rhof * cd * (np.pi / 4.0) * dq ** 2 * um) / pm
    e[29:32], e[32:35] = up, om
    e[35] = 1.0
    acc = np.full((nspec, NPFSTAT), ZERO)
    for s in range(nspec):
        acc[s] += e[:, species_id == s].sum(axis=1)
    out = np.empty_like(acc)
    for s in range(nspec):
        n = acc[nspec - npoly + s % npoly, 35] if ref_norm else acc[s, 35]
        out[s] = acc[s] / n
        out[s, 35] = n
    return out


# ---------------------------------------------------------------------------
# STAT_PARTICLE's PDFs and level 2
# ---------------------------------------------------------------------------
NPDF = 100


def _bins(v, lo, hi):
    """int((v - lo) / ((hi - lo) / NPDF)) + 1 clamped to [1, NPDF]
    (stat_particle.f90:562-645), 0-based"""
    t = (v - lo) / ((hi - lo) / float(NPDF))
    t = np.clip(np.trunc(t), -1e9, 1e9).astype(np.int64) + 1
    return np.clip(t, 1, NPDF) - 1


def particle_pdf(state: np.ndarray, species_id: np.ndarray, nspec: int) -> np.ndarray:
    """MEAN_PART_PDF(1:12, J, IDP, 1:NPDF) of one call, (nspec, 12, NPDF),
    counts added to the routine's ZERO start:
    u, v, w in [-2, 2]; body spin in [-2000, 2000]; R(q) omega in [-2000,
    2000]; R(q) e_x in [-1, 1] (stat_particle.f90:536-645).  R omega is
    summed left to right as the routine's matmul does."""
    R = rotation(state[6:10])
    u, w = state[3:6], state[10:13]
    g = [(R[:, k, 0] * w[0] + R[:, k, 1] * w[1]) + R[:, k, 2] * w[2] for k in range(3)]
    vals = [(u[k], -2.0, 2.0) for k in range(3)] + [(w[k], -2000.0, 2000.0) for k in range(3)] + \
        [(g[k], -2000.0, 2000.0) for k in range(3)] + [(R[:, k, 0], -1.0, 1.0) for k in range(3)]
    out = np.full((nspec, 12, NPDF), 1.0e-33)     # MEAN_PART_PDF = ZERO (:164)
    for k, (v, lo, hi) in enumerate(vals):
        b = _bins(v, lo, hi)
        np.add.at(out, (species_id, k, b), 1.0)
    return out


def lagrangian_origin(state, species_id, mean):
    """UPT0..WPT0 (stat_particle.f90:760-770): u - MEAN_PART(1:3, J, IDP),
    (3, np)"""
    return state[3:6] - mean[species_id].T


def lagrangian_corr(state, species_id, mean, npm, u0):
    """The sums over each species row of (u - MEAN_PART(1:3)) * u_t0 /
    NPM(IDP) (stat_particle.f90:790-805), (nspec, 3)"""
    nspec = mean.shape[0]
    t = ((state[3:6] - mean[species_id].T) * u0) / npm[species_id]
    out = np.zeros((nspec, 3))
    for c in range(3):
        out[:, c] = np.bincount(species_id, weights=t[c], minlength=nspec)
    return out


def npm_of(species_id, nspec, npoly):
    """NPM(IDP) as STAT_PARTICLE holds it after its class loop: the count of
    the last class's row of size IDP, per species row"""
    counts = np.bincount(species_id, minlength=nspec).astype(np.float64)
    last = counts[nspec - npoly:]
    return np.tile(last, nspec // npoly)
