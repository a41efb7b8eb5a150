"""The decision quantities of gazpart_amd.contacts (host side) against the
oracle's own contact decisions and closed forms (CPU only).

* The flag rebuilt from depth and the two margins equals the flag of the
  oracle's solve (ellipsoid_contact_detection.f90:120-125) on every pair
  but those within rounding of a threshold.
* margin() of a sphere is |p - c|^2 / r^2, and of an ellipsoid rotated by
  q the margin of c + R(q) (a e_k) is 1 on each axis.
* NVREL of two particles in pure translation is (u2 - u1) . n.
"""
import numpy as np

import oracle as O
from gazpart_amd import contacts as CT
from gazpart_amd import workload as W


def test_margin_closed_forms():
    g = np.random.default_rng(3)
    n = 50
    q = g.normal(size=(4, n))
    q /= np.linalg.norm(q, axis=0)
    c = g.normal(size=(n, 3))
    p = g.normal(size=(n, 3))
    r = np.full((n, 3), 0.7)
    np.testing.assert_allclose(CT.margin(r, q, c, p), ((p - c) ** 2).sum(1) / 0.49, rtol=1e-12)
    axes = np.abs(g.normal(size=(n, 3))) + 0.1
    R = CT.rotation(q)
    for k in range(3):
        pt = c + axes[:, k:k + 1] * R[:, :, k]
        np.testing.assert_allclose(CT.margin(axes, q, c, pt), 1.0, rtol=1e-12)
    # R(q) is a rotation
    np.testing.assert_allclose(np.einsum("nij,nkj->nik", R, R), np.broadcast_to(np.eye(3), (n, 3, 3)),
                               atol=1e-12)


def test_nvrel_translation():
    st = np.zeros((13, 2))
    st[0:3, 1] = [0.1, 0.0, 0.0]
    st[6, :] = 1.0
    st[3:6, 0] = [1.0, 0.5, 0.0]
    st[3:6, 1] = [-1.0, 0.0, 2.0]
    sp = W.species_spheroid(0.1, 2.0, 1.0)[None, :]
    cn = np.array([[1.0, 0.0, 0.0]])
    qd = CT.decision_quantities(st, sp, np.zeros(2, np.int32), 2 * np.pi, [1], [2],
                                np.array([[0.05, 0, 0]]), np.array([[0.06, 0, 0]]), cn)
    assert qd["nvrel"][0] == -2.0
    np.testing.assert_allclose(qd["depth"][0], -0.01)


def test_flags_rebuilt_from_quantities_match_oracle():
    wl = W.c5(n=32, np_=600, seed=11)
    st = wl.state
    p = O.params(wl)
    g = O.collision_grid(p, wl.dpmax)
    con, ncon = O.connectivity(g)
    h, nr, nc, err = O.detect(g, con, ncon, st[0], st[1], st[2], 10 ** 6)
    assert err == 0
    i, j, b1, b2 = CT.pair_geometry(st, wl.L, h, nr)
    near = np.linalg.norm(b1 - b2, axis=1) < 1.2 * wl.species[:, 0:3].max() * 2
    h, nr, b1, b2 = h[near], nr[near], b1[near], b2[near]
    n = len(h)
    assert n > 300
    flag = np.zeros(n, np.int32)
    e1, e2, cn = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    for k in range(n):
        a, b = h[k] - 1, nr[k] - 1
        flag[k], e1[k], e2[k], cn[k], _, _ = O.contact(wl.species[wl.species_id[a]],
                                                       wl.species[wl.species_id[b]], b1[k],
                                                       st[6:10, a], b2[k], st[6:10, b])
    qd = CT.decision_quantities(st, wl.species, wl.species_id, wl.L, h, nr, e1, e2, cn)
    rebuilt = CT.flag_of(qd)
    assert flag.sum() > 20
    mism = np.nonzero(rebuilt != flag.astype(bool))[0]
    assert (CT.closeness(qd)[mism] < 1e-12).all(), mism
    # a listing of identical decisions is empty; flipping one run lists it
    assert CT.list_differences(h, nr, flag, qd, flag, qd) == []
    f2 = flag.copy()
    f2[np.argmax(flag)] = 0
    recs = CT.list_differences(h, nr, flag, qd, f2, qd, step=4)
    assert len(recs) == 1 and recs[0]["step"] == 4 and recs[0]["flag"] == (1, 0)
