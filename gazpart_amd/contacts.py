"""Contact decisions and their distance to the threshold (host side, numpy).

A candidate pair (HOME, NEAR) collides in the reference when three tests
hold at the contact solve's final iterate and the pair approaches
(ellipsoid_contact_detection.f90:120-125, collision_interaction.f90:348):

  depth = (PT1 - PT2) . cnormal > 0
  fm1   = margin of PT2 in ellipsoid 1 (mod_particle_parallel.f90:145-201) < 1
  fm2   = margin of PT1 in ellipsoid 2 < 1
  nvrel = (v2 - v1) . cnormal < ZERO   (velocities of the contact points,
          collision_interaction.f90:269-324)

Two runs (the GPU and the oracle on the same state, the mixed and the FP64
form, a network-seeded and a centre-line solve) can decide a pair
differently only when one of these sits near its threshold in at least one
of them, or when the contact solve ended in another local minimum.  This
module computes the four quantities from a state and the solve's outputs
(PT1, PT2, cnormal, as ``Context.ellipsoid_contact_detection`` returns
them), lists the pairs two runs decide differently, and gives each listed
pair its distance to the threshold

  closeness = min(|depth| / s, |fm1 - 1|, |fm2 - 1|, |nvrel| / |v2 - v1|)

(s the sum of the two largest semi-axes), the smallest relative distance of
any of the four tests to its switching point in either run.
"""
from __future__ import annotations

import numpy as np

ZERO = 1.0e-33


def rotation(q):
    """R(q) of transform_basis.f90:44-54 for quaternions q (4, n) -> (n, 3, 3)."""
    a, b, c, d = q
    R = np.empty((q.shape[1], 3, 3))
    R[:, 0, 0] = 1.0 - 2.0 * (c * c + d * d)
    R[:, 0, 1] = 2.0 * (b * c - a * d)
    R[:, 0, 2] = 2.0 * (b * d + a * c)
    R[:, 1, 0] = 2.0 * (b * c + a * d)
    R[:, 1, 1] = 1.0 - 2.0 * (b * b + d * d)
    R[:, 1, 2] = 2.0 * (c * d - a * b)
    R[:, 2, 0] = 2.0 * (b * d - a * c)
    R[:, 2, 1] = 2.0 * (c * d + a * b)
    R[:, 2, 2] = 1.0 - 2.0 * (b * b + c * c)
    return R


def margin(axes, q, centre, point):
    """vec^T R diag(1/a^2, 1/b^2, 1/c^2) R^T vec with vec = point - centre:
    below 1 inside the ellipsoid (mod_particle_parallel.f90:145-201, with
    the semi-axes of the species row instead of b = c = a/APR)."""
    vb = np.einsum("nji,nj->ni", rotation(q), point - centre)
    return ((vb / axes) ** 2).sum(axis=1)


def pair_geometry(state, L, home, near):
    """Centres b1, b2 (b2 the minimum image, collision_interaction.f90:
    221-237) of 1-based pairs; state is (13, np)."""
    i = np.asarray(home, np.int64) - 1
    j = np.asarray(near, np.int64) - 1
    b1 = state[0:3, i].T
    d = state[0:3, j].T - b1
    d = np.where(np.abs(d) > L / 2.0, d - np.copysign(L, d), d)
    return i, j, b1, b1 + d


def decision_quantities(state, species, species_id, L, home, near, ept1, ept2, cnormal):
    """depth, fm1, fm2, nvrel, |v2 - v1| and the scale s for every pair."""
    i, j, b1, b2 = pair_geometry(state, L, home, near)
    si, sj = species[species_id[i]], species[species_id[j]]
    ai, aj = si[:, 0:3], sj[:, 0:3]
    qi, qj = state[6:10, i], state[6:10, j]
    depth = ((ept1 - ept2) * cnormal).sum(axis=1)
    fm1 = margin(ai, qi, b1, ept2)
    fm2 = margin(aj, qj, b2, ept1)
    # lab-frame spins and contact-point velocities
    w1 = np.einsum("nij,nj->ni", rotation(qi), state[10:13, i].T)
    w2 = np.einsum("nij,nj->ni", rotation(qj), state[10:13, j].T)
    v1 = state[3:6, i].T + np.cross(w1, ept1 - b1)
    v2 = state[3:6, j].T + np.cross(w2, ept2 - b2)
    dv = v2 - v1
    nvrel = (dv * cnormal).sum(axis=1)
    s = ai.max(axis=1) + aj.max(axis=1)
    return dict(depth=depth, fm1=fm1, fm2=fm2, nvrel=nvrel, vrel=np.linalg.norm(dv, axis=1),
                scale=s)


def flag_of(qd):
    return (qd["depth"] > 0.0) & (qd["fm1"] < 1.0) & (qd["fm2"] < 1.0)


def contact_of(flag, qd):
    """the approach test on top of the flag: a collision when reached"""
    return np.asarray(flag, bool) & (qd["nvrel"] < ZERO)


def closeness(qd):
    """smallest relative distance of the four tests to their thresholds"""
    return np.minimum.reduce([
        np.abs(qd["depth"]) / qd["scale"], np.abs(qd["fm1"] - 1.0), np.abs(qd["fm2"] - 1.0),
        np.abs(qd["nvrel"]) / np.maximum(qd["vrel"], 1e-300)])


def list_differences(home, near, flag_a, qd_a, flag_b, qd_b, step=0):
    """Pairs decided differently by runs a and b (flag, or flag and the
    approach test).  One record per pair with both runs' quantities and the
    pair's closeness (the smaller of the two runs')."""
    ca, cb = contact_of(flag_a, qd_a), contact_of(flag_b, qd_b)
    fa, fb = np.asarray(flag_a, bool), np.asarray(flag_b, bool)
    idx = np.nonzero((fa != fb) | (ca != cb))[0]
    close = np.minimum(closeness(qd_a), closeness(qd_b))
    out = []
    for k in idx:
        out.append(dict(
            step=int(step), home=int(home[k]), near=int(near[k]),
            flag=(int(fa[k]), int(fb[k])), contact=(int(ca[k]), int(cb[k])),
            depth=(float(qd_a["depth"][k]), float(qd_b["depth"][k])),
            fm1=(float(qd_a["fm1"][k]), float(qd_b["fm1"][k])),
            fm2=(float(qd_a["fm2"][k]), float(qd_b["fm2"][k])),
            nvrel=(float(qd_a["nvrel"][k]), float(qd_b["nvrel"][k])),
            closeness=float(close[k])))
    return out


def format_differences(recs, limit=20):
    lines = []
    for r in recs[:limit]:
        lines.append(
            f"step {r['step']} pair ({r['home']},{r['near']}) flag {r['flag']} contact "
            f"{r['contact']} depth {r['depth'][0]:.3e}/{r['depth'][1]:.3e} fm1 "
            f"{r['fm1'][0]:.9f}/{r['fm1'][1]:.9f} fm2 {r['fm2'][0]:.9f}/{r['fm2'][1]:.9f} "
            f"nvrel {r['nvrel'][0]:.3e}/{r['nvrel'][1]:.3e} closeness {r['closeness']:.2e}")
    if len(recs) > limit:
        lines.append(f"... {len(recs) - limit} more")
    return "\n".join(lines)
