"""HIP path (through the C-ABI) against the reference's own outputs
(tests/golden/) and the CPU oracle, on the same seeded inputs.

Tolerances
  FP64 form   : interpolation 1e-13 of the field scale; one step 1e-12 of
                each quantity's scale; trajectories 1e-9 of scale after
                excluding particles whose collision history differs from the
                reference (they are listed in the assertion message).
  mixed form  : FP32 field, gather and drag; relative error per quantity
                (max |gpu - oracle| / max |oracle|) <= 1e-5.
  integer     : cell grid, candidate pairs (HOME/NEAR) bit-exact.
"""
import numpy as np
import pytest

import gazpart_amd as G
import oracle as O
from gazpart_amd import workload as W

pytestmark = pytest.mark.gpu
GOLD = "tests/golden/"
QUANT = {"x": slice(0, 3), "u": slice(3, 6), "q": slice(6, 10), "omega": slice(10, 13)}


def relerr(a, b, sl):
    return float(np.abs(a[sl] - b[sl]).max() / max(np.abs(b[sl]).max(), 1e-300))


@pytest.fixture(scope="module")
def c1():
    return W.c1(np_=4000)


@pytest.fixture(scope="module")
def leaf():
    return np.load(GOLD + "golden_leaf.npz")


def golden_workload(d):
    np_ = d["state0"].shape[1]
    return W.Workload(name="golden", n=int(d["n"]), L=float(d["L"]), k0=2.0, seed=0, np_=np_,
                      scheme=int(d["scheme"]), steps=0, visc=float(d["visc"]),
                      rhof=float(d["rhof"]), dtime=float(d["dtime"]), field=d["field"], umax=0.0,
                      species=W.species_spheroid(float(d["emaj"]), float(d["apr"]),
                                                 float(d["rhop"]))[None, :],
                      species_id=np.zeros(np_, dtype=np.int32), state=d["state0"],
                      dpmax=2.0 * float(d["emaj"]))


@pytest.mark.parametrize("prec,tol", [(G.PREC_FP64, 1e-13), (G.PREC_MIXED, 2e-6)])
def test_interph_against_reference_fixture(leaf, prec, tol):
    n, L = int(leaf["n"]), float(leaf["L"])
    ctx = G.Context(n, L, 0.01, 1.0, 1e-3, precision=prec)
    ctx.bind_field(leaf["field"])
    pts = leaf["interp_pts"]
    for scheme, key in ((3, "interp3"), (4, "interp4")):
        for c in range(6):
            out = ctx.interph(scheme, c, *pts)
            ref = leaf[key][c]
            assert np.abs(out - ref).max() <= tol * np.abs(ref).max(), (scheme, c)


def test_interph_field_argument_against_reference_fixture(leaf):
    """INTERPH with UAIN passed in (gp_interph_field), one component at a
    time, no bound field."""
    n, L = int(leaf["n"]), float(leaf["L"])
    ctx = G.Context(n, L, 0.01, 1.0, 1e-3)
    pts = leaf["interp_pts"]
    for scheme, key in ((3, "interp3"), (4, "interp4")):
        for c in range(6):
            out = ctx.interph_field(scheme, leaf["field"][c], *pts)
            ref = leaf[key][c]
            assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max(), (scheme, c)


@pytest.mark.parametrize("prec,tol", [(G.PREC_FP64, 1e-13), (G.PREC_MIXED, 2e-6)])
def test_interph_against_oracle(c1, prec, tol):
    ctx = G.Context.from_workload(c1, precision=prec)
    g = W.rng(3)
    x, y, z = g.uniform(0, c1.L, (3, 20000))
    for scheme in (3, 4):
        for c in range(6):
            out = ctx.interph(scheme, c, x, y, z)
            ref, err = O.interph(scheme, c1.n, c1.dx, c1.field[c], x, y, z)
            assert np.abs(out - ref).max() <= tol * np.abs(ref).max(), (scheme, c)


def test_cell_list_bitexact_against_reference_fixture():
    d = np.load(GOLD + "golden_cells.npz")
    st = d["state"]
    sp = W.species_spheroid(float(d["emaj"]), float(d["apr"]), 10.0)
    ctx = G.Context(int(d["n"]), float(d["L"]), 0.01, 1.0, 1e-3, dpmax=2.0 * float(d["emaj"]))
    ctx.bind_field(np.zeros((6, int(d["n"]) + 4, int(d["n"]) + 4, int(d["n"]))))
    ctx.set_species(sp)
    ctx.set_particles(st, np.zeros(st.shape[1], dtype=np.int32))
    nb, bm = ctx.collision_grid()
    assert list(nb) == list(d["grid_i"])
    assert list(bm) == list(d["grid_r"])
    home, near = ctx.collision_detection_connectivity()
    np.testing.assert_array_equal(home, d["home"])
    np.testing.assert_array_equal(near, d["near"])


def test_cell_list_bitexact_edge_positions(c1):
    """Particles on box faces, at 0, just below L and in the L+DY strip of
    y/z (x = L exactly lands outside the grid in the reference, which then
    aliases into the next row; both sides report it as an error instead)."""
    p = O.params(c1)
    g = O.collision_grid(p, c1.dpmax)
    st = c1.state.copy()
    m = 400
    gen = W.rng(17)
    for k in range(m):
        i = k
        ax = k % 3
        face = gen.integers(0, g.nxbox)
        d = (g.dxbox, g.dybox, g.dzbox)[ax]
        st[ax, i] = [0.0, 1e-300, face * d, face * d * (1 + 2e-16), c1.L - 1e-15][k % 5]
    ctx = G.Context.from_workload(c1)
    ctx.set_particles(st, c1.species_id)
    home, near = ctx.collision_detection_connectivity()
    con, ncon = O.connectivity(g)
    ho, no, nc, err = O.detect(g, con, ncon, st[0], st[1], st[2], 10 ** 7)
    assert err == 0
    np.testing.assert_array_equal(home, ho)
    np.testing.assert_array_equal(near, no)


def test_contact_detection_against_reference_fixture(leaf):
    """LM contact solve on the fixture's 600 pairs (three aspect ratios,
    oblate included): flags identical, contact points and normals within the
    solver's own tolerance of the reference."""
    k = len(leaf["c_cls"])
    apr, emaj, rhop = leaf["apr"], leaf["emaj"], leaf["rhop"]
    rows = np.array([W.species_spheroid(emaj[c, i], apr[c], rhop[c, i]) for c in range(3)
                     for i in range(2)])
    st = np.zeros((13, 2 * k))
    st[0:3, 0::2] = leaf["c_b1"].T
    st[0:3, 1::2] = leaf["c_b2"].T
    st[6:10, 0::2] = leaf["c_q1"].T
    st[6:10, 1::2] = leaf["c_q2"].T
    sid = np.zeros(2 * k, dtype=np.int32)
    sid[0::2] = (leaf["c_cls"] - 1) * 2 + leaf["c_idp1"] - 1
    sid[1::2] = (leaf["c_cls"] - 1) * 2 + leaf["c_idp2"] - 1
    n = int(leaf["n"])
    ctx = G.Context(n, float(leaf["L"]), 0.01, 1.0, 1e-3)
    ctx.bind_field(leaf["field"])
    ctx.set_species(rows)
    ctx.set_particles(st, sid)
    home = np.arange(1, 2 * k, 2, dtype=np.int32)
    flag, e1, e2, cn, nfev = ctx.ellipsoid_contact_detection(home, home + 1)
    np.testing.assert_array_equal(flag, leaf["c_flag"])
    f = flag.astype(bool)
    scale = emaj.max()
    assert np.abs(e1[f] - leaf["c_e1"][f]).max() < 1e-5 * scale
    assert np.abs(e2[f] - leaf["c_e2"][f]).max() < 1e-5 * scale
    assert np.abs(cn[f] - leaf["c_cn"][f]).max() < 1e-4
    assert (nfev > 0).all() and (nfev <= 300).all()


@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
def test_one_advance_against_oracle(c1, prec):
    ctx = G.Context.from_workload(c1, precision=prec, keep_fluid=True)
    ctx.adv_particle_ellp(1)
    st, fl = ctx.get_particles(), ctx.get_fluid_at_particles()
    so, flo, err = O.adv(O.params(c1), c1.field, c1.species, c1.state, c1.species_id,
                         want_fluid=True)
    assert err == 0
    tol = 1e-12 if prec == G.PREC_FP64 else 1e-5
    assert np.abs(fl - flo).max() <= (1e-13 if prec == G.PREC_FP64 else 2e-6) * np.abs(flo).max()
    for name, sl in QUANT.items():
        assert relerr(st, so, sl) <= tol, name


def _history_divergence(coll_gpu, coll_ref):
    """particles whose set of collision steps differs"""
    return np.nonzero((coll_gpu != coll_ref).any(axis=0))[0]


@pytest.mark.parametrize("name", ["golden_trajectory.npz", "golden_trajectory_lag4.npz"])
def test_trajectory_against_reference_fixture(name):
    """Full ADV_PARTICLE_ELLP + COLLISION steps of a dense ensemble (5 %
    volume fraction, wall bounces and ~20 collisions per step) against the
    reference Fortran's own trajectory."""
    d = np.load(GOLD + name)
    wl = golden_workload(d)
    ctx = G.Context.from_workload(wl)
    np_ = wl.np_
    keep = list(d["keep_steps"])
    coll = np.zeros_like(d["collided"])
    prev = wl.state.copy()
    for s in range(1, int(max(keep)) + 1):
        ctx.adv_particle_ellp(s)
        mid = ctx.get_particles()
        ctx.collision(s)
        cur = ctx.get_particles()
        # a particle took part in a collision iff its velocity changed there
        coll[s - 1] = np.any(cur[3:6] != mid[3:6], axis=0) | np.any(cur[10:13] != mid[10:13], axis=0)
        prev = cur
        if s in keep:
            ref = d["states"][keep.index(s)]
            bad = _history_divergence(coll[:s], d["collided"][:s])
            ok = np.setdiff1d(np.arange(np_), bad)
            assert len(bad) <= 0.05 * np_, f"step {s}: collision history differs for {bad}"
            for q, sl in QUANT.items():
                e = np.abs(cur[sl][:, ok] - ref[sl][:, ok]).max() / np.abs(ref[sl]).max()
                assert e <= 1e-9, (s, q, e, "excluded:", bad)
    assert coll.sum() > 100
    # the per-particle collision counters count the same collisions
    np.testing.assert_array_equal(ctx.collision_counts(), coll.sum(axis=0))


@pytest.mark.parametrize("prec,tol", [(G.PREC_FP64, 1e-9), (G.PREC_MIXED, 1e-5)])
def test_trajectory_against_oracle_c1(c1, prec, tol):
    """20 steps of the C1 configuration (64^3, k0 = 4, APR 2, St 1) with
    collisions on: per-quantity relative error vs the FP64 oracle."""
    ctx = G.Context.from_workload(c1, precision=prec)
    stp = O.Stepper(c1)
    so = c1.state.copy()
    ncoll_o = ncoll_g = 0
    for s in range(20):
        ctx.step(1)
        so, err, nc, ncoll = stp.step(so)
        assert err == 0
        ncoll_o += ncoll
        ncoll_g += ctx.stats()["ncollision"]
    sg = ctx.get_particles()
    assert ncoll_o == ncoll_g
    for name, sl in QUANT.items():
        assert relerr(sg, so, sl) <= tol, name


@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
def test_prefilters_change_no_result(prec):
    """The bounding-sphere and Perram-Wertheim prefilters only skip pairs
    whose flag is false: 10 steps of the dense reference ensemble give the
    same state bit for bit with prefilter 0 (LM on every candidate), 1 and
    2, while the LM sees fewer pairs at each level."""
    d = np.load(GOLD + "golden_trajectory.npz")
    wl = golden_workload(d)
    out, st = {}, {}
    for lvl in (0, 1, 2):
        ctx = G.Context.from_workload(wl, precision=prec, prefilter=lvl)
        nlm = ncoll = 0
        for s in range(10):
            ctx.step(1)
            t = ctx.stats()
            nlm += t["nlm"]
            ncoll += t["ncollision"]
            if lvl == 0:
                assert t["nlm"] == t["nclose"]
            assert t["nlm"] <= t["nprefilter"] <= t["nclose"]
        out[lvl] = ctx.get_particles()
        st[lvl] = (nlm, ncoll)
        ctx.close()
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])
    assert st[0][1] == st[1][1] == st[2][1] > 0
    assert st[0][0] > st[1][0] > st[2][0]


@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
def test_visiting_order_changes_no_result(c1, prec):
    """The advance's visiting order (reference cell order, or bricks of 2,
    8 or 16 collision cells a side) decides only where particles are stored:
    the state after 6 steps is the same bit for bit, dense ensemble and C1.
    Field tiles of 3, 7 or 8 cells (gathered from LDS by a kernel of their
    own, then advanced by the tile mode's advance kernel) give the same bits
    as each other; against the direct path they differ by the rounding of
    that kernel's own fused multiply-adds only: within 1e-11 of the
    largest magnitude (FP64) and 1e-5 (mixed: FP32 velocities)."""
    d = np.load(GOLD + "golden_trajectory.npz")
    tol = 1e-11 if prec == G.PREC_FP64 else 1e-5

    def run(wl, **kw):
        ctx = G.Context.from_workload(wl, precision=prec, **kw)
        ctx.step(3)
        ctx.adv_particle_ellp(4)
        ctx.collision(4)
        ctx.step(2)
        o = ctx.get_particles()
        ctx.close()
        return o

    for wl in (golden_workload(d), c1):
        out = [run(wl, brick=b) for b in (-1, 2, 0, 16)]
        for o in out[1:]:
            np.testing.assert_array_equal(out[0], o)
        tiles = [run(wl, tile=t) for t in (3, 7, 8)]
        for o in tiles[1:]:
            np.testing.assert_array_equal(tiles[0], o)
        for q in (slice(0, 3), slice(3, 6), slice(6, 10), slice(10, 13)):
            scale = np.abs(out[0][q]).max()
            assert np.abs(tiles[0][q] - out[0][q]).max() <= tol * scale


def test_step_equals_adv_plus_collision_and_is_deterministic(c1):
    a = G.Context.from_workload(c1)
    b = G.Context.from_workload(c1)
    c = G.Context.from_workload(c1)
    for s in range(5):
        a.adv_particle_ellp(s + 1)
        a.collision(s + 1)
    b.step(5)
    c.step(5)
    sa, sb, sc = a.get_particles(), b.get_particles(), c.get_particles()
    np.testing.assert_array_equal(sa, sb)
    np.testing.assert_array_equal(sb, sc)


@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
def test_host_handover_between_steps_changes_no_result(c1, prec):
    """The Fortran drop-in copies the particles out to the reference's host
    arrays and back in between steps (GPU_GET_PART / GPU_PUT_PART): the
    library carries no per-particle state across steps beyond the 13
    components of the record, so handing them over changes nothing -- the
    dense ensemble and C1, in eager and in graph mode."""
    d = np.load(GOLD + "golden_trajectory.npz")
    for wl in (golden_workload(d), c1):
        for graph in (False, True):
            a = G.Context.from_workload(wl, precision=prec, graph=graph)
            b = G.Context.from_workload(wl, precision=prec, graph=graph)
            a.step(6)
            b.step(3)
            b.set_particles(b.get_particles(), wl.species_id)
            b.step(1)
            b.set_particles(b.get_particles(), wl.species_id)
            b.step(2)
            np.testing.assert_array_equal(a.get_particles(), b.get_particles())
            a.close()
            b.close()


@pytest.mark.parametrize("mode,overlap,tile", [(1, 0, 0), (1, 1, 0), (2, 1, 0), (2, 2, 0),
                                               (1, 1, 4), (2, 1, 4)])
@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
def test_graph_replays_equal_eager_steps(c1, prec, mode, overlap, tile):
    """The captured hipGraph (2 steps, or 8 pipelined steps with the next
    advance beside the contact solves) replayed several times, with eager
    work and synchronisations in between, gives the eager result bit for
    bit after every call; on C1 and on the dense reference ensemble (many
    contact solves per step).  Mode 2 launches the same pipelined steps
    directly, without capture.  With field tiles the fluid at every particle
    is gathered beside the solves and the solves' particles read it after
    the collision pass.  overlap = 1 also solves each step's contact pairs
    ahead, from the pairs of the step before (2: pipelined without it)."""
    d = np.load(GOLD + "golden_trajectory.npz")
    # the dense ensemble stops at step 53 in the reference (NaN drag, see
    # test_error_at_the_reference_step): keep it below that
    for wl, counts in ((c1, (9, 16, 3, 17, 2)), (golden_workload(d), (9, 16, 3, 9, 2))):
        eager = G.Context.from_workload(wl, precision=prec, tile=tile)
        graph = G.Context.from_workload(wl, precision=prec, graph=mode, overlap=overlap,
                                        tile=tile)
        ahead = 0
        for n in counts:
            eager.step(n)
            graph.step(n)
            a, b = eager.get_particles(), graph.get_particles()
            np.testing.assert_array_equal(a, b)
            se, sg = eager.stats(), graph.stats()
            assert se["ncollision"] == sg["ncollision"]
            # every pair reached the solve, in the step or ahead of it
            assert se["nlm"] == sg["nlm"] and 0 <= sg["nahead"] <= sg["nlm"]
            ahead += sg["nahead"]
            # eager calls on the graph context between replays
            graph.collision_detection_connectivity()
            graph.sync()
        # the pairs of one pass are nearly all the pairs of the pass before
        assert (ahead > 0) == (overlap == 1)
        graph.adv_particle_ellp(1)
        graph.collision(1)
        eager.adv_particle_ellp(1)
        eager.collision(1)
        graph.step(8)
        eager.step(8)
        np.testing.assert_array_equal(eager.get_particles(), graph.get_particles())
        eager.close()
        graph.close()


def ensemble_alone(wl, a, b):
    """The particles [a, b) of an ensemble workload as a workload of their
    own (same field, species and collision grid)."""
    return W.Workload(name="e", n=wl.n, L=wl.L, k0=wl.k0, seed=0, np_=b - a, scheme=wl.scheme,
                      steps=0, visc=wl.visc, rhof=wl.rhof, dtime=wl.dtime, field=wl.field,
                      umax=wl.umax, species=wl.species, species_id=wl.species_id[a:b],
                      state=wl.state[:, a:b], dpmax=wl.dpmax)


def test_ensemble_pair_lists_are_each_ensembles_own():
    """HOME/NEAR of three ensembles sharing one collision grid (a cell's
    members of other ensembles skipped): no pair crosses ensembles, and each
    ensemble's pairs, in order, are the list that a context of that
    ensemble alone gives (collision_detection_connectivity.f90:145-207 per
    ensemble), after steps have mixed the ensembles within the cells."""
    wl = W.build_ensembles("E3", 32, [700, 1500, 400], k0=3.0, seed=5, req=0.05)
    grp = np.asarray(wl.meta["group"])
    both = G.Context.from_workload(wl, precision=G.PREC_FP64)
    both.set_groups(grp)
    both.step(3)
    state = both.get_particles()
    home, near = both.collision_detection_connectivity()
    assert len(home) > 100
    assert np.all(grp[home - 1] == grp[near - 1])
    for a, b in wl.meta["ranges"]:
        sub = ensemble_alone(wl, a, b)
        one = G.Context.from_workload(sub, precision=G.PREC_FP64)
        one.set_particles(state[:, a:b], wl.species_id[a:b])
        h1, n1 = one.collision_detection_connectivity()
        sel = (home - 1 >= a) & (home - 1 < b)
        np.testing.assert_array_equal(home[sel], h1 + a)
        np.testing.assert_array_equal(near[sel], n1 + a)
        one.close()
    both.close()


@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
def test_ensembles_share_a_field_and_never_collide(prec):
    """Three ensembles (different aspect ratios, sizes and counts) in one
    context give, particle for particle, the bits of three separate
    contexts; each matches its own oracle run."""
    wl = W.build_ensembles("E3", 32, [700, 1500, 400], k0=3.0, seed=5, req=0.05)
    grp = wl.meta["group"]
    both = G.Context.from_workload(wl, precision=prec, graph=True)
    both.set_groups(grp)
    both.step(8)
    got = both.get_particles()
    ncoll = 0
    for e, (a, b) in enumerate(wl.meta["ranges"]):
        sub = ensemble_alone(wl, a, b)
        one = G.Context.from_workload(sub, precision=prec)
        one.step(8)
        ncoll += one.stats()["ncollision"]
        np.testing.assert_array_equal(got[:, a:b], one.get_particles())
        if prec == G.PREC_FP64:
            stp = O.Stepper(sub)
            so = sub.state.copy()
            for _ in range(8):
                so, err, _, _ = stp.step(so)
                assert err == 0
            for name, sl in QUANT.items():
                assert relerr(got[:, a:b], so, sl) <= 1e-9, (e, name)
    assert ncoll > 0


def test_error_at_the_reference_step():
    """The dense reference ensemble runs into a NaN drag at step 53 in the
    reference's own algorithm (the FP64 oracle, bit-exact with the compiled
    reference over the fixture's 30 steps, stops there with error 1): the
    FP64 GPU path stops at the same step with the same error."""
    d = np.load(GOLD + "golden_trajectory.npz")
    wl = golden_workload(d)
    stp = O.Stepper(wl)
    s = wl.state.copy()
    ref_step = None
    for k in range(1, 60):
        s, err, _, _ = stp.step(s)
        if err:
            ref_step = (k, err)
            break
    assert ref_step == (53, 1)
    ctx = G.Context.from_workload(wl, precision=G.PREC_FP64)
    for k in range(1, 53):
        ctx.step(1)
        ctx.sync()
    ctx.step(1)
    with pytest.raises(G.GpError) as e:
        ctx.sync()
    assert e.value.code == 1


def test_error_nan_drag(c1):
    """u_p = u_f exactly gives Re_p = 0 and C_D = NaN: the reference stops
    (fluid_force_ellipsoid.f90:150-158); the library returns the code."""
    ctx = G.Context.from_workload(c1)
    st = c1.state.copy()
    for c in range(3):
        st[3 + c] = ctx.interph(c1.scheme, c, st[0], st[1], st[2])
    st[10:13] = 0.0
    ctx.set_particles(st, c1.species_id)
    with pytest.raises(G.GpError) as e:
        ctx.adv_particle_ellp(1)
    assert e.value.code == 1


def test_error_ke_collision():
    """Restitution below 1 breaks the kinetic-energy balance the reference
    enforces (collision_response.f90:408-436)."""
    d = np.load(GOLD + "golden_trajectory.npz")
    wl = golden_workload(d)
    wl.species = W.species_spheroid(float(d["emaj"]), float(d["apr"]), float(d["rhop"]),
                                    ecp=0.5)[None, :]
    ctx = G.Context.from_workload(wl)
    ctx.adv_particle_ellp(1)
    with pytest.raises(G.GpError) as e:
        ctx.collision(1)
    assert e.value.code == 3


def test_error_out_of_grid_and_interp_range(c1):
    ctx = G.Context.from_workload(c1)
    st = c1.state.copy()
    st[0, 7] = c1.L + 0.5
    ctx.set_particles(st, c1.species_id)
    with pytest.raises(G.GpError) as e:
        ctx.collision_detection_connectivity()
    assert e.value.code == 5
    with pytest.raises(G.GpError) as e:
        ctx.interph(3, 0, np.array([-1.0]), np.array([1.0]), np.array([1.0]))
    assert e.value.code == 7
    with pytest.raises(G.GpError) as e:
        ctx.interph(5, 0, np.array([1.0]), np.array([1.0]), np.array([1.0]))
    assert e.value.code == 10


@pytest.mark.parametrize("dpmax", [0.9, 0.06, 0.018])
def test_cell_list_bitexact_short_and_long_scans(dpmax):
    """The cell-count scan runs in one workgroup for grids of at most 32768
    cells, as two launches above, and with the tile totals scanned in a
    launch of their own from 2048 tiles of 2048 cells: each gives the
    reference's HOME/NEAR lists (C oracle) bit for bit (dpmax 0.9: 27
    cells; 0.06: 140608; 0.018: 5.6M, cells narrower than C1's particles,
    so the cell rule goes unchecked on both sides)."""
    wl = W.c1(np_=2000 if dpmax > 0.05 else 20000)
    wl.dpmax = dpmax
    p = O.params(wl)
    g = O.collision_grid(p, dpmax)
    assert (g.nbox <= 32768) == (dpmax > 0.5)
    assert (g.nbox >= 2048 * 2048) == (dpmax < 0.05)
    ctx = G.Context.from_workload(wl, cell_rule=2 if dpmax < 0.05 else 0)
    home, near = ctx.collision_detection_connectivity()
    con, ncon = O.connectivity(g)
    st = wl.state
    ho, no, nc, err = O.detect(g, con, ncon, st[0], st[1], st[2], 10 ** 7)
    assert err == 0 and nc > 0
    np.testing.assert_array_equal(home, ho)
    np.testing.assert_array_equal(near, no)


@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
def test_collision_advance_equals_collision_then_advance(c1, prec):
    """gp_collision_advance (COLLISION, then the next ADV_PARTICLE_ELLP beside
    the contact solves, in one call) gives the bits of gp_collision followed
    by gp_adv_particle_ellp: the dense reference ensemble and C1, with a
    host handover and eager calls in between."""
    d = np.load(GOLD + "golden_trajectory.npz")
    for wl, tile in ((golden_workload(d), 0), (c1, 0), (c1, 8)):
        a = G.Context.from_workload(wl, precision=prec, tile=tile)
        b = G.Context.from_workload(wl, precision=prec, tile=tile)
        a.adv_particle_ellp(1)
        b.adv_particle_ellp(1)
        for s in range(1, 12):
            a.collision(s)
            a.adv_particle_ellp(s + 1)
            b.collision_advance(s, 0.0, s + 1)
            if s == 5:
                np.testing.assert_array_equal(a.get_particles(), b.get_particles())
            if s == 7:
                b.set_particles(b.get_particles(), wl.species_id)
                a.set_particles(a.get_particles(), wl.species_id)
        a.collision(12)
        b.collision(12)
        np.testing.assert_array_equal(a.get_particles(), b.get_particles())
        assert a.stats()["ncollision"] == b.stats()["ncollision"]
        np.testing.assert_array_equal(a.collision_counts(), b.collision_counts())
        a.close()
        b.close()


def _ctx_rule(wl, rule, **kw):
    """a context whose collision cells come from its cell rule (no DPMAX
    given)"""
    ctx = G.Context(wl.n, wl.L, wl.visc, wl.rhof, wl.dtime, scheme=wl.scheme, dpmax=0.0,
                    cell_rule=rule, **kw)
    ctx.bind_field(wl.field)
    ctx.set_species(wl.species)
    ctx.set_particles(wl.state, wl.species_id)
    return ctx


def test_cell_rule_guard_for_oblate_species():
    """All-oblate ensemble (APR 0.3: b = c = 3.3 a): the reference rule
    DPMAX = 2 max(EMAJ) (collision_initiation.f90:50-75) makes cells of
    4 max(EMAJ), narrower than two largest semi-axes, and loses pairs.
    cell_rule 0 refuses it (GP_ERR_CELL_RULE); 2 keeps the reference's lists
    bit for bit (oracle on the same grid); 1 sizes the cells from the
    largest semi-axis, matches the oracle on that grid and contains every
    pair whose centres are closer than the sum of the largest semi-axes
    (brute force), which the reference rule misses some of."""
    wl = W.build_workload("oblate", 48, 3000, k0=3.0, seed=21, apr=0.3, emaj_dx=0.6, st=1.0)
    with pytest.raises(G.GpError) as e:
        _ctx_rule(wl, 0)
    assert e.value.code == 8
    st = wl.state
    p = O.params(wl)
    sp = wl.species[0]
    rmax = sp[[W.SP_A, W.SP_B, W.SP_C]].max()
    # brute force: pairs that could touch
    d = st[0:3, :, None] - st[0:3, None, :]
    d = np.where(np.abs(d) > wl.L / 2, d - np.copysign(wl.L, d), d)
    close = np.linalg.norm(d, axis=0) < 2.0 * rmax
    iu = np.triu_indices(wl.np_, 1)
    need = {(int(a) + 1, int(b) + 1) for a, b in zip(*iu) if close[a, b]}
    assert len(need) > 20
    found = {}
    for rule, dpmax in ((2, 2.0 * sp[W.SP_EMAJ]), (1, 2.0 * rmax)):
        ctx = _ctx_rule(wl, rule)
        h, n = ctx.collision_detection_connectivity()
        g = O.collision_grid(p, dpmax)
        con, ncon = O.connectivity(g)
        ho, no, nc, err = O.detect(g, con, ncon, st[0], st[1], st[2], 10 ** 7)
        assert err == 0
        np.testing.assert_array_equal(h, ho)
        np.testing.assert_array_equal(n, no)
        found[rule] = {(min(a, b), max(a, b)) for a, b in zip(h.tolist(), n.tolist())}
        ctx.close()
    assert need <= found[1]
    assert not need <= found[2]


@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
def test_cell_rule_changes_nothing_when_safe(c1, prec):
    """Prolate species (the largest semi-axis is EMAJ): cell_rule 0, 1 and 2
    build the same grid and give the same steps bit for bit."""
    out = []
    for rule in (0, 1, 2):
        ctx = _ctx_rule(c1, rule, precision=prec)
        ctx.step(4)
        out.append(ctx.get_particles())
        ctx.close()
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])
