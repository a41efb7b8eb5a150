"""The CPU oracle against the reference's own outputs (tests/golden/*.npz,
written by tests/golden/make_golden.py from the compiled reference Fortran).

The oracle restates the reference arithmetic operation for operation, so the
comparison is bit-exact (same glibc libm on the same inputs)."""
import numpy as np
import pytest

import oracle as O
from gazpart_amd import workload as W

GOLD = "tests/golden/"


@pytest.fixture(scope="module")
def leaf():
    return np.load(GOLD + "golden_leaf.npz")


def _params(d, n=None):
    p = O.Params()
    p.n = int(d["n"]) if n is None else n
    p.interp_scheme, p.renorm_quat = 3, 0
    p.L = float(d["L"])
    p.dx = p.L / (p.n - 1)
    p.visc, p.rhof, p.dtime = float(d["visc"]), float(d["rhof"]), float(d["dtime"])
    p.ecw, p.muw = 1.0, 0.0
    return p


def _sp(d, cls, idp):
    return W.species_spheroid(d["emaj"][cls - 1, idp - 1], d["apr"][cls - 1],
                              d["rhop"][cls - 1, idp - 1])


def test_interp_lag3_lag4_bitexact(leaf):
    n, L = int(leaf["n"]), float(leaf["L"])
    pts = leaf["interp_pts"]
    for scheme, key in ((3, "interp3"), (4, "interp4")):
        for c in range(6):
            out, err = O.interph(scheme, n, L / (n - 1), leaf["field"][c], *pts)
            assert err == 0
            np.testing.assert_array_equal(out, leaf[key][c])


def test_force_torque_quaternion_bitexact(leaf):
    p = _params(leaf)
    for t in range(len(leaf["ft_cls"])):
        sp = _sp(leaf, leaf["ft_cls"][t], leaf["ft_idp"][t])
        up, fl6, q, om = leaf["ft_up"][t], leaf["ft_fl6"][t], leaf["ft_q"][t], leaf["ft_om"][t]
        fd, fl, e = O.force(p, sp, *up, *fl6[:3], q)
        assert e == 0
        np.testing.assert_array_equal(fd, leaf["f_drag"][t])
        np.testing.assert_array_equal(fl, leaf["f_lift"][t])
        tp, tr, e = O.torque(p, sp, fl6, up, *om, q)
        np.testing.assert_array_equal(tp, leaf["t_pitch"][t])
        np.testing.assert_array_equal(tr, leaf["t_rot"][t])
        np.testing.assert_array_equal(O.quat(q, *om, leaf["q_dt"][t]), leaf["q_new"][t])


def test_wall_test_and_rebound_bitexact(leaf):
    p = _params(leaf)
    assert leaf["w_hit"].sum() > 50
    for t in range(len(leaf["w_cls"])):
        sp = _sp(leaf, leaf["w_cls"][t], leaf["w_idp"][t])
        nrm, arm, dxw, hit = O.wall_test(p, sp, leaf["w_x"][t], *leaf["w_yz"][t], leaf["w_q"][t])
        assert hit == leaf["w_hit"][t]
        np.testing.assert_array_equal(nrm, leaf["w_normal"][t])
        if hit:
            np.testing.assert_array_equal(arm, leaf["w_arm"][t])
            assert dxw == leaf["w_dx"][t]
            u, o, e = O.wall_rebound(p, sp, leaf["w_u"][t], leaf["w_o"][t], leaf["w_q"][t], arm,
                                     nrm)
            assert e == 0
            np.testing.assert_array_equal(u, leaf["w_u_after"][t])
            np.testing.assert_array_equal(o, leaf["w_o_after"][t])


def test_contact_detection_bitexact(leaf):
    assert 50 < leaf["c_flag"].sum() < len(leaf["c_flag"]) - 50
    for t in range(len(leaf["c_cls"])):
        c = leaf["c_cls"][t]
        s1, s2 = _sp(leaf, c, leaf["c_idp1"][t]), _sp(leaf, c, leaf["c_idp2"][t])
        f, e1, e2, cn, nfev, xs = O.contact(s1, s2, leaf["c_b1"][t], leaf["c_q1"][t],
                                            leaf["c_b2"][t], leaf["c_q2"][t])
        assert f == leaf["c_flag"][t], t
        np.testing.assert_array_equal(e1, leaf["c_e1"][t])
        np.testing.assert_array_equal(e2, leaf["c_e2"][t])
        np.testing.assert_array_equal(cn, leaf["c_cn"][t])


def test_collision_response_bitexact(leaf):
    for t in range(len(leaf["r_cls"])):
        sp = _sp(leaf, leaf["r_cls"][t], leaf["r_idp"][t])
        a, b, e = O.response(sp, leaf["r_s1"][t], leaf["r_s2"][t], leaf["r_r1"][t],
                             leaf["r_r2"][t], leaf["r_n"][t])
        assert e == 0
        np.testing.assert_array_equal(a, leaf["r_s1_after"][t])
        np.testing.assert_array_equal(b, leaf["r_s2_after"][t])


def test_cell_list_bitexact():
    d = np.load(GOLD + "golden_cells.npz")
    p = O.Params()
    p.n, p.L = int(d["n"]), float(d["L"])
    p.dx = p.L / (p.n - 1)
    g = O.collision_grid(p, 2.0 * float(d["emaj"]))
    assert [g.nxbox, g.nybox, g.nzbox, g.nbox] == list(d["grid_i"])
    assert [g.xminbox, g.yminbox, g.zminbox, g.dxbox, g.dybox, g.dzbox] == list(d["grid_r"])
    con, ncon = O.connectivity(g)
    np.testing.assert_array_equal(con, d["con"])
    np.testing.assert_array_equal(ncon, d["ncon"])
    st = d["state"]
    home, near, nc, err = O.detect(g, con, ncon, st[0], st[1], st[2], 10 ** 7)
    assert err == 0 and nc == len(d["home"])
    np.testing.assert_array_equal(home, d["home"])
    np.testing.assert_array_equal(near, d["near"])


@pytest.mark.parametrize("name", ["golden_trajectory.npz", "golden_trajectory_lag4.npz"])
def test_full_steps_bitexact(name):
    d = np.load(GOLD + name)
    wl = W.Workload(name="golden", n=int(d["n"]), L=float(d["L"]), k0=2.0, seed=0,
                    np_=d["state0"].shape[1], scheme=int(d["scheme"]), steps=0,
                    visc=float(d["visc"]), rhof=float(d["rhof"]), dtime=float(d["dtime"]),
                    field=d["field"], umax=0.0,
                    species=W.species_spheroid(float(d["emaj"]), float(d["apr"]),
                                               float(d["rhop"]))[None, :],
                    species_id=np.zeros(d["state0"].shape[1], dtype=np.int32),
                    state=d["state0"], dpmax=2.0 * float(d["emaj"]))
    stp = O.Stepper(wl)
    st = wl.state.copy()
    keep = list(d["keep_steps"])
    for s in range(1, int(max(keep)) + 1):
        st, err, nc, ncoll = stp.step(st)
        assert err == 0
        assert nc == d["nclose"][s - 1]
        assert 2 * ncoll == d["collided"][s - 1].sum()
        if s in keep:
            np.testing.assert_array_equal(st, d["states"][keep.index(s)])
