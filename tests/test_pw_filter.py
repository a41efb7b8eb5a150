"""The Perram-Wertheim disjointness test used to skip contact solves
(gazpart_amd/csrc/gp_contact.h, pw_disjoint) against the oracle's contact
detection (a restatement of ellipsoid_contact_detection.f90).

The GPU kernel's claim is: F(l) = l (1-l) r^T [(1-l) E1 + l E2]^-1 r above
1 + 1e-6 for some l proves the pair disjoint, and the reference's flag is
then false.  Pinned here by
  * the closed form for spheres, max F = d^2 / (r1 + r2)^2;
  * containment: F < 1 at every l when one centre lies inside the other
    ellipsoid;
  * the oracle's LM flag on a few thousand random close pairs of spheroids
    and tri-axial ellipsoids (every pair proven disjoint has flag 0, and the
    test proves a sizeable share of the unflagged pairs disjoint).
The bit-identity of GPU trajectories with the filter on and off is a GPU
test (test_gpu_parity.py::test_prefilters_change_no_result)."""
import numpy as np

import oracle as O
from gazpart_amd import workload as W

MARGIN = 1.0 + 1.0e-6


def rotm(q):
    q0, q1, q2, q3 = q
    return np.array([[1 - 2 * (q2 * q2 + q3 * q3), 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)],
                     [2 * (q1 * q2 + q0 * q3), 1 - 2 * (q1 * q1 + q3 * q3), 2 * (q2 * q3 - q0 * q1)],
                     [2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 1 - 2 * (q1 * q1 + q2 * q2)]])


def shape(q, abc):
    R = rotm(q)
    return R @ np.diag(np.asarray(abc) ** 2) @ R.T


def F(E1, E2, r, lam):
    M = (1 - lam) * E1 + lam * E2
    return lam * (1 - lam) * r @ np.linalg.solve(M, r)


def Fmax(E1, E2, r, n=199):
    return max(F(E1, E2, r, l) for l in np.linspace(0.005, 0.995, n))


def test_spheres_closed_form():
    g = np.random.default_rng(1)
    for _ in range(50):
        r1, r2 = g.uniform(0.1, 2.0, 2)
        d = g.uniform(0.1, 5.0)
        r = d * g.normal(size=3)
        r *= d / np.linalg.norm(r)
        E1, E2 = np.eye(3) * r1 * r1, np.eye(3) * r2 * r2
        lam = r1 / (r1 + r2)
        assert np.isclose(F(E1, E2, r, lam), d * d / (r1 + r2) ** 2, rtol=1e-12)
        assert Fmax(E1, E2, r, 999) <= d * d / (r1 + r2) ** 2 * (1 + 1e-12)


def test_centre_inside_other_is_never_disjoint():
    g = np.random.default_rng(2)
    for _ in range(50):
        q1, q2 = (v / np.linalg.norm(v) for v in g.normal(size=(2, 4)))
        a1, a2 = g.uniform(0.2, 2.0, (2, 3))
        E1, E2 = shape(q1, a1), shape(q2, a2)
        u = g.normal(size=3)
        u /= np.linalg.norm(u)
        # a point strictly inside ellipsoid 1: scale u to 0.9 of its boundary
        r = 0.9 * u / np.sqrt(u @ np.linalg.solve(E1, u))
        assert Fmax(E1, E2, r) < 1.0


def _pairs(nrand, seed):
    g = np.random.default_rng(seed)
    emaj = 0.05
    sp = [W.species_spheroid(emaj, a, 2.0) for a in (0.2, 0.5, 1.0, 2.0, 5.0)]
    sp.append(W.species_triaxial(emaj, 0.6 * emaj, 0.3 * emaj, 2.0))
    for _ in range(nrand):
        s1, s2 = sp[g.integers(len(sp))], sp[g.integers(len(sp))]
        q1, q2 = (v / np.linalg.norm(v) for v in g.normal(size=(2, 4)))
        u = g.normal(size=3)
        u /= np.linalg.norm(u)
        rmax = s1[0:3].max() + s2[0:3].max()
        b1 = g.uniform(1.0, 2.0, 3)
        b2 = b1 + g.uniform(0.2, 1.0) * rmax * u
        yield s1, s2, q1, q2, b1, b2


def test_disjoint_pairs_have_no_contact_flag():
    proven = flagged = unflagged = 0
    for s1, s2, q1, q2, b1, b2 in _pairs(3000, 3):
        E1, E2 = shape(q1, s1[0:3]), shape(q2, s2[0:3])
        fm = Fmax(E1, E2, b2 - b1, 49)
        flag = O.contact(s1, s2, b1, q1, b2, q2)[0]
        if fm > MARGIN:
            proven += 1
            assert flag == 0, (s1[0:3], s2[0:3], q1, q2, b1, b2, fm)
        if flag:
            flagged += 1
        else:
            unflagged += 1
    assert flagged > 300
    assert proven > 0.5 * unflagged


def test_prefilter_float_slack():
    """The bounding-sphere prefilter runs on float copies of the centres
    (gp_lib.hip, k_prefilter) with radii rounded up to float, a slack of
    1e-5 of the largest coordinate magnitude (prefilter_tol) and a 1e-4
    relative factor.  With float32 arithmetic on pairs spread over the whole
    box (the periodic wrap included) at distances from 0.999 to 1.01 times
    the sum of radii, it must keep every pair whose exact centre distance is below the
    sum: a dropped pair has to be disjoint for the reference's flag to be
    0 without a solve."""
    g = np.random.default_rng(11)
    L = 2.0 * np.pi
    ext = max(1.0, L + 0.05)               # largest coordinate magnitude of the grid
    tol = np.float32(1.0e-5 * ext)
    n = 200_000
    ri = g.uniform(0.001, 0.05, n)
    rj = g.uniform(0.001, 0.05, n)
    rs = ri + rj
    d = rs * (1.0 + np.where(g.random(n) < 0.5, g.uniform(-1e-3, 1e-3, n), g.uniform(3e-3, 1e-2, n)))
    u = g.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    xi = g.uniform(-0.03, L + 0.03, (n, 3))
    xj = xi + d[:, None] * u
    xj = np.where(xj >= L, xj - L, np.where(xj < 0.0, xj + L, xj))   # stored wrapped
    # exact (float64) minimum-image distance
    dv = xj - xi
    dv -= L * np.round(dv / L)
    dex = np.sqrt((dv * dv).sum(axis=1))
    # the kernel's float test
    f32 = np.float32
    rmi = np.nextafter(ri.astype(f32), f32(np.inf))   # float_up
    rmj = np.nextafter(rj.astype(f32), f32(np.inf))
    Lf = f32(L)
    Lh = f32(0.5) * Lf
    df = xj.astype(f32) - xi.astype(f32)
    df = np.where(np.abs(df) > Lh, df - np.copysign(Lf, df), df).astype(f32)
    rsf = (rmi + tol) + rmj
    keep = (df[:, 0] * df[:, 0] + df[:, 1] * df[:, 1] + df[:, 2] * df[:, 2]) <= rsf * rsf * f32(1.0001)
    must = dex < rs
    assert must.sum() > n // 5
    assert np.all(keep[must]), int((~keep & must).sum())
    # the slack is small: far pairs are still dropped
    far = dex > rs * (1.0 + 2e-3) + 2e-4
    assert far.sum() > n // 5
    assert not np.any(keep[far])
