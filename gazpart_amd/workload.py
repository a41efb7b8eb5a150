"""Synthetic inputs for the Lagrangian ellipsoid step (host side, numpy).

The reference has no random-Fourier-mode field and no seeded particle
initialisation (its INITIATION_FROZEN_FLOW modes are zero / linear shear /
file profile, initiation_frozen_flow.f90:51-158, and particle positions use
the unseeded ``random_number``).  The workloads therefore come from this
module, with one integer seed per workload, and are fed identically to the
HIP library and to the CPU oracle.

Conventions kept from the reference:
  * grid spacing ``DX = L/(N-1)`` and node ``i`` at ``(i-1)*DX``
    (meshing.f90:43-82), field index-periodic with period N
    (interp_lag3.f90:132-154, MPI_structures.f90:27-132);
  * field array ``UAIN(1:N, -1:N+2, -1:N+2)`` with x fastest
    (allocate_arrays.f90:108-118); here as numpy ``[c, k+1, j+1, i-1]``;
  * initial velocities and spins ``ZERO = 1e-33``
    (initiation_particle_velocity.f90:127-132);
  * inertia ``IPXX = (b^2+c^2)/5`` (initiation_particle_position.f90:620-622).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field as dfield

import numpy as np

TWOPI = 2.0 * math.pi
ZERO = 1.0e-33

# species row layout, identical to the C-ABI's GP_SP_* indices
SP_A, SP_B, SP_C, SP_PMASS, SP_DEQ, SP_IXX, SP_IYY, SP_IZZ, SP_GRAV, SP_ECP, \
    SP_MUP, SP_RHOP, SP_EMAJ, SP_APR, SP_STRIDE = range(15)

# state component order (x y z u v w qa qb qc qd ox oy oz)
NSTATE = 13


def rng(seed: int) -> np.random.Generator:
    """Counter-based generator (Philox) for every seeded draw."""
    return np.random.Generator(np.random.Philox(int(seed)))


# ---------------------------------------------------------------------------
# field
# ---------------------------------------------------------------------------
def random_fourier_field(n: int, k0: float, seed: int, urms: float = 1.0,
                         dtype=np.float64) -> tuple[np.ndarray, float]:
    """Divergence-free random Fourier-mode velocity and its vorticity.

    Spectrum ``E(k) ~ k^4 exp(-2 (k/k0)^2)`` on integer wavenumbers of the
    2*pi-periodic box (modes with any |k_i| >= n/2 dropped), random complex
    Gaussian amplitudes projected onto the plane normal to k, velocity
    rescaled so that sqrt(<|u|^2>/3) = urms, vorticity = curl u spectrally.

    The transform grid x_j = j*2pi/n is mapped onto the reference nodes
    X_i = (i-1)*L/(n-1); the field seen on the reference grid is
    U(X) = u(X (n-1)/n), whose curl is (n-1)/n times the spectral vorticity,
    so the vorticity is scaled by that factor to stay consistent with U.

    Returns (field, max |u|) with field shaped (6, n+4, n+4, n) in the
    reference layout, ghosts filled with index period n.
    """
    g = rng(seed)
    kz = np.fft.fftfreq(n, 1.0 / n)[:, None, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :, None]
    kx = np.fft.rfftfreq(n, 1.0 / n)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    kk = np.sqrt(k2)
    amp = kk * np.exp(-(kk / k0) ** 2) / math.sqrt(4.0 * math.pi)
    amp[(np.abs(kx) >= n // 2) | (np.abs(ky) >= n // 2) | (np.abs(kz) >= n // 2)] = 0.0
    amp[0, 0, 0] = 0.0
    shape = k2.shape
    w = [g.standard_normal(shape) + 1j * g.standard_normal(shape) for _ in range(3)]
    k2s = np.where(k2 == 0.0, 1.0, k2)
    kdw = (kx * w[0] + ky * w[1] + kz * w[2]) / k2s
    uh = [amp * (w[0] - kx * kdw), amp * (w[1] - ky * kdw), amp * (w[2] - kz * kdw)]
    del w, kdw
    # vorticity = i k x u
    oh = [1j * (ky * uh[2] - kz * uh[1]),
          1j * (kz * uh[0] - kx * uh[2]),
          1j * (kx * uh[1] - ky * uh[0])]
    comps = [np.fft.irfftn(c, s=(n, n, n), axes=(0, 1, 2)) for c in uh + oh]
    del uh, oh
    rms = math.sqrt(float(sum(np.mean(c * c) for c in comps[:3])) / 3.0)
    scale = urms / rms
    fld = np.empty((6, n + 4, n + 4, n), dtype=dtype)
    umax2 = np.zeros((n, n, n))
    for c in range(6):
        v = comps[c] * scale
        if c >= 3:
            v = v * ((n - 1) / n)
        else:
            umax2 += v * v
        fld[c, 2:n + 2, 2:n + 2, :] = v.astype(dtype)
        comps[c] = None
    fill_ghosts(fld, n)
    return fld, float(math.sqrt(umax2.max()))


def fill_ghosts(fld: np.ndarray, n: int) -> None:
    """Ghost planes j,k in {-1, 0, n+1, n+2} copy planes {n-1, n, 1, 2}."""
    src = {0: n, 1: n + 1, n + 2: 2, n + 3: 3}   # array rows (index + 1)
    for dst, s in src.items():
        fld[:, :, dst, :] = fld[:, :, s, :]
    for dst, s in src.items():
        fld[:, dst, :, :] = fld[:, s, :, :]


# ---------------------------------------------------------------------------
# species (one row per particle class/size)
# ---------------------------------------------------------------------------
def species_spheroid(emaj, apr, rhop, grav=0.0, ecp=1.0, mup=0.0) -> np.ndarray:
    """Species rows with the reference's own formulas: PMASS and dequiv of
    fluid_force_ellipsoid.f90:59,70, semi-axes of
    calculate_ellipsoid_matrix.f90:33-35 (b = c = a/APR), inertia per unit
    mass of initiation_particle_position.f90:620-622.  Scalars give one row
    (shape (SP_STRIDE,)), arrays give one row per element."""
    scalar = np.ndim(emaj) == 0 and np.ndim(apr) == 0 and np.ndim(rhop) == 0
    emaj, apr, rhop = np.broadcast_arrays(np.asarray(emaj, float), np.asarray(apr, float),
                                          np.asarray(rhop, float))
    a = emaj
    b = a / apr
    c = b
    row = np.zeros(emaj.shape + (SP_STRIDE,))
    row[..., SP_A], row[..., SP_B], row[..., SP_C] = a, b, c
    # EMAJ**3.0 and APR**(2.0) evaluate as products in the reference build
    row[..., SP_PMASS] = rhop * (4.0 / 3.0) * math.pi * ((emaj * emaj) * emaj) / (apr * apr)
    row[..., SP_DEQ] = (2.0 * emaj) / np.power(apr, 2.0 / 3.0)
    row[..., SP_IXX] = (b * b + c * c) / 5.0
    row[..., SP_IYY] = (a * a + c * c) / 5.0
    row[..., SP_IZZ] = (a * a + b * b) / 5.0
    row[..., SP_GRAV], row[..., SP_ECP], row[..., SP_MUP] = grav, ecp, mup
    row[..., SP_RHOP], row[..., SP_EMAJ], row[..., SP_APR] = rhop, emaj, apr
    return row[()] if not scalar else row.reshape(SP_STRIDE)


def species_triaxial(a, b, c, rhop, grav=0.0, ecp=1.0, mup=0.0) -> np.ndarray:
    """Tri-axial extension (no reference counterpart): mass and
    volume-equivalent diameter of a general ellipsoid; for b = c it equals
    the spheroid row up to rounding.  Scalars or arrays, as above."""
    scalar = all(np.ndim(v) == 0 for v in (a, b, c, rhop))
    a, b, c, rhop = np.broadcast_arrays(*(np.asarray(v, float) for v in (a, b, c, rhop)))
    row = np.zeros(a.shape + (SP_STRIDE,))
    row[..., SP_A], row[..., SP_B], row[..., SP_C] = a, b, c
    row[..., SP_PMASS] = rhop * (4.0 / 3.0) * math.pi * (a * b * c)
    row[..., SP_DEQ] = 2.0 * np.cbrt(a * b * c)
    row[..., SP_IXX] = (b * b + c * c) / 5.0
    row[..., SP_IYY] = (a * a + c * c) / 5.0
    row[..., SP_IZZ] = (a * a + b * b) / 5.0
    row[..., SP_GRAV], row[..., SP_ECP], row[..., SP_MUP] = grav, ecp, mup
    row[..., SP_RHOP], row[..., SP_EMAJ], row[..., SP_APR] = rhop, a, a / b
    return row.reshape(SP_STRIDE) if scalar else row


def rhop_for_stokes(st: float, deq: float, tau_flow: float, rhof: float,
                    visc: float) -> float:
    """Particle density giving Stokes number st = tau_p / tau_flow with the
    Stokes-limit response time tau_p = rhop deq^2 / (18 rhof visc)."""
    return st * tau_flow * 18.0 * rhof * visc / (deq * deq)


# ---------------------------------------------------------------------------
# particles
# ---------------------------------------------------------------------------
def random_particles(np_: int, L: float, margin: float, seed: int) -> np.ndarray:
    """State (13, np): positions uniform in [margin, L-margin]^3 (as
    initiation_particle_position.f90:310-316), unit quaternions from a
    normalised Gaussian 4-vector, u = omega = ZERO."""
    g = rng(seed)
    st = np.empty((NSTATE, np_))
    st[0:3] = margin + (L - 2.0 * margin) * g.random((3, np_))
    q = g.standard_normal((4, np_))
    q /= np.sqrt((q * q).sum(axis=0))
    st[6:10] = q
    st[3:6] = ZERO
    st[10:13] = ZERO
    return st


@dataclass
class Workload:
    """One configured run: field, species table, particles and step params."""
    name: str
    n: int
    L: float
    k0: float
    seed: int
    np_: int
    scheme: int
    steps: int
    visc: float
    rhof: float
    dtime: float
    field: np.ndarray
    umax: float
    species: np.ndarray           # (nspec, SP_STRIDE)
    species_id: np.ndarray        # (np,) int32
    state: np.ndarray             # (13, np)
    ecw: float = 1.0
    muw: float = 0.0
    dpmax: float = 0.0            # 2 * max(EMAJ), collision_initiation.f90:50
    meta: dict = dfield(default_factory=dict)

    @property
    def dx(self) -> float:
        return self.L / (self.n - 1)


def build_workload(name: str, n: int, np_: int, *, k0: float = 4.0, seed: int = 1234,
                   scheme: int = 3, steps: int = 100, visc: float = 0.01,
                   rhof: float = 1.0, cfl: float = 0.1, apr=2.0, emaj_dx: float | None = 0.2,
                   phi: float | None = None, st=1.0, field_dtype=np.float64,
                   triaxial_ratio: tuple | None = None) -> Workload:
    """Generic workload builder.

    apr / st may be scalars (one species, reference parity) or (lo, hi)
    tuples drawn log-uniformly per particle (one species row per particle).
    The size is set either by emaj_dx (semi-major axis in grid spacings) or
    by the volume fraction phi (volume-equivalent radius from phi).
    triaxial_ratio = (lo, hi) draws b/a... as uniform ratios a/b, a/c in
    [lo, hi] (tri-axial extension).
    """
    L = TWOPI
    dx = L / (n - 1)
    fld, umax = random_fourier_field(n, k0, seed, dtype=field_dtype)
    dtime = cfl * dx / umax
    tau_flow = 1.0 / (k0 * 1.0)
    g = rng(seed + 1)
    per_particle = isinstance(apr, tuple) or isinstance(st, tuple) or triaxial_ratio is not None
    if phi is not None:
        req = (phi * L ** 3 * 3.0 / (4.0 * math.pi * np_)) ** (1.0 / 3.0)
    else:
        req = None
    if not per_particle:
        if req is not None:
            emaj = req * apr ** (2.0 / 3.0)
        else:
            emaj = emaj_dx * dx
        deq = 2.0 * emaj / apr ** (2.0 / 3.0)
        rhop = rhop_for_stokes(st, deq, tau_flow, rhof, visc)
        spec = species_spheroid(emaj, apr, rhop)[None, :]
        sid = np.zeros(np_, dtype=np.int32)
        dpmax = 2.0 * emaj
        amax = max(emaj, emaj / apr)
    else:
        def draw(v):
            if isinstance(v, tuple):
                lo, hi = v
                return np.exp(g.uniform(math.log(lo), math.log(hi), np_))
            return np.full(np_, float(v))
        sts = draw(st)
        if triaxial_ratio is not None:
            lo, hi = triaxial_ratio
            rb = g.uniform(lo, hi, np_)
            rc = g.uniform(lo, hi, np_)
            # a = r (rb rc)^(1/3) keeps a b c = r^3
            a = req * np.cbrt(rb * rc)
            deq = np.full(np_, 2.0 * req)
            spec = species_triaxial(a, a / rb, a / rc,
                                    rhop_for_stokes(sts, deq, tau_flow, rhof, visc))
        else:
            aprs = draw(apr)
            emaj = req * aprs ** (2.0 / 3.0) if req is not None else np.full(np_, emaj_dx * dx)
            deq = 2.0 * emaj / aprs ** (2.0 / 3.0)
            spec = species_spheroid(emaj, aprs, rhop_for_stokes(sts, deq, tau_flow, rhof, visc))
        sid = np.arange(np_, dtype=np.int32)
        dpmax = 2.0 * float(spec[:, SP_EMAJ].max())
        amax = float(spec[:, [SP_A, SP_B, SP_C]].max())
    state = random_particles(np_, L, amax, seed + 2)
    return Workload(name=name, n=n, L=L, k0=k0, seed=seed, np_=np_, scheme=scheme,
                    steps=steps, visc=visc, rhof=rhof, dtime=dtime, field=fld,
                    umax=umax, species=spec, species_id=sid, state=state, dpmax=dpmax,
                    meta=dict(tau_flow=tau_flow, cfl=cfl, phi=phi))


def build_ensembles(name: str, n: int, counts, *, k0: float = 4.0, seed: int = 77,
                    scheme: int = 3, steps: int = 100, visc: float = 0.01, rhof: float = 1.0,
                    cfl: float = 0.1, req: float = 3.9e-3, apr=(0.2, 5.0), st=(0.1, 10.0),
                    field_dtype=np.float64, field=None) -> Workload:
    """Independent ensembles on one shared field (no reference counterpart:
    the reference runs one ensemble per process).  Ensemble e has counts[e]
    spheroids with its own aspect ratio and Stokes number (log-uniform in
    the given ranges, one draw per ensemble) and volume-equivalent radius
    req; Workload.meta["group"] holds each particle's ensemble and
    meta["ranges"] the particle index range of each ensemble."""
    L = TWOPI
    dx = L / (n - 1)
    if field is None:
        fld, umax = random_fourier_field(n, k0, seed, dtype=field_dtype)
    else:
        fld, umax = field
    dtime = cfl * dx / umax
    tau_flow = 1.0 / k0
    g = rng(seed + 5)
    ne = len(counts)
    aprs = np.exp(g.uniform(math.log(apr[0]), math.log(apr[1]), ne))
    sts = np.exp(g.uniform(math.log(st[0]), math.log(st[1]), ne))
    emaj = req * aprs ** (2.0 / 3.0)
    deq = 2.0 * emaj / aprs ** (2.0 / 3.0)
    spec = species_spheroid(emaj, aprs, rhop_for_stokes(sts, deq, tau_flow, rhof, visc))
    np_ = int(sum(counts))
    state = np.empty((NSTATE, np_))
    sid = np.empty(np_, dtype=np.int32)
    grp = np.empty(np_, dtype=np.int32)
    ranges = []
    o = 0
    for e, c in enumerate(counts):
        amax = max(emaj[e], emaj[e] / aprs[e])
        state[:, o:o + c] = random_particles(int(c), L, amax, seed + 100 + e)
        sid[o:o + c] = e
        grp[o:o + c] = e
        ranges.append((o, o + int(c)))
        o += int(c)
    return Workload(name=name, n=n, L=L, k0=k0, seed=seed, np_=np_, scheme=scheme, steps=steps,
                    visc=visc, rhof=rhof, dtime=dtime, field=fld, umax=umax,
                    species=np.atleast_2d(spec), species_id=sid, state=state,
                    dpmax=2.0 * float(emaj.max()),
                    meta=dict(group=grp, ranges=ranges, apr=aprs, st=sts, tau_flow=tau_flow))


def c3(n: int = 512, np_: int = 8_000_000, steps: int = 500, seed: int = 3072,
       scheme: int = 3) -> Workload:
    """Precision: 512^3 field (k0 = 8) for FP32 storage, 8,000,000 tri-axial
    ellipsoids with semi-axis ratios a/b, a/c uniform in [1, 4]; the
    volume fraction is not given for this configuration, 1e-3 is assumed
    (as C2); St = 1."""
    return build_workload("C3", n, np_, k0=8.0, seed=seed, scheme=scheme, steps=steps,
                          phi=1e-3, st=1.0, triaxial_ratio=(1.0, 4.0), emaj_dx=None)


def c5(n: int = 256, np_: int = 2_000_000, steps: int = 100, seed: int = 5120,
       scheme: int = 3) -> Workload:
    """Learned contact seed: 2,000,000 tri-axial ellipsoids (semi-axis ratios
    uniform in [1, 4]) at volume fraction 5 % on 256^3 (collision-heavy),
    St = 1."""
    return build_workload("C5", n, np_, k0=4.0, seed=seed, scheme=scheme, steps=steps,
                          phi=0.05, st=1.0, triaxial_ratio=(1.0, 4.0), emaj_dx=None)


def c4(n: int = 256, nens: int = 64, steps: int = 100, seed: int = 4096,
       field_dtype=np.float64) -> Workload:
    """Batching: nens independent ensembles on one shared 256^3 field,
    particle counts log-uniform in [1e3, 5e5] (seeded), per-ensemble St in
    [0.1, 10] and aspect ratio in [0.2, 5]."""
    g = rng(seed)
    counts = np.exp(g.uniform(math.log(1e3), math.log(5e5), nens)).astype(int)
    return build_ensembles("C4", n, counts, seed=seed, steps=steps, field_dtype=field_dtype)


def c1(n: int = 64, np_: int = 10_000, steps: int = 100, seed: int = 1234,
       scheme: int = 3) -> Workload:
    """Plain FP64, CPU-runnable: periodic 2*pi box at 64^3, k0 = 4, 10,000
    prolate spheroids (APR 2, a = 0.2 DX), St = 1, lag3, dt at CFL 0.1,
    collisions on with the reference cell rule, 100 steps."""
    return build_workload("C1", n, np_, k0=4.0, seed=seed, scheme=scheme, steps=steps,
                          apr=2.0, emaj_dx=0.2, st=1.0)


def c2(n: int = 256, np_: int = 1_000_000, steps: int = 200, seed: int = 2024,
       scheme: int = 3) -> Workload:
    """Plain FP64 at scale: 256^3, 1,000,000 spheroids, APR log-uniform in
    [0.2, 5], volume fraction 1e-3, St log-uniform in [0.1, 10]."""
    return build_workload("C2", n, np_, k0=4.0, seed=seed, scheme=scheme, steps=steps,
                          apr=(0.2, 5.0), phi=1e-3, st=(0.1, 10.0))
