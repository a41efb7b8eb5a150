"""STAT_PARTICLE level 1 (stat_particle.f90:169-450), its PDFs
(LEVEL1_STPDF, :536-645) and level 2 (the Lagrangian autocorrelation,
:737-826): the library's gp_particle_stats, gp_particle_pdf and
gp_lagrangian_origin / gp_lagrangian_corr against the CPU restatement
oracle/stats_ref.py, which is pinned against the reference's own
STAT_PARTICLE (compiled unmodified, run on one MPI rank by
oracle/_ref/part_io_ref; MEAN_PART is read back at full precision through
the STAT_TIME accumulator, the PDFs through MEAN_TIME_PART_PDF, level 2 as
RPX/RPY/RPZ_LOC).

Tolerances: PDF counts exactly; sums run in another order than the
reference's particle loop (numpy pairwise sums, GPU atomics), so they agree
to 1e-12 of each entry's scale rather than bit for bit."""
import os
import subprocess

import numpy as np
import pytest

import gazpart_amd as G
from gazpart_amd import workload as W
from oracle import stats_ref as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "part_io_ref")


def _close(a, b, tol=1e-12):
    scale = np.maximum(np.abs(b), 1e-300)
    bad = np.abs(a - b) > tol * np.maximum(scale, np.abs(b).max(axis=0, keepdims=True) * 1e-3)
    assert not bad.any(), (np.argwhere(bad)[:5], a[bad][:5], b[bad][:5])


def _classes(seed, nig=2, npoly=3, npart=600):
    g = np.random.default_rng(seed)
    L = 2.0 * np.pi
    emaj = g.uniform(1e-3, 1e-2, (nig, npoly))
    apr = g.uniform(0.2, 5.0, nig)
    rhop = g.uniform(500.0, 3000.0, (nig, npoly))
    states, idps = [], []
    for j in range(nig):
        st = np.empty((13, npart))
        st[0:3] = g.uniform(0.05, L - 0.05, (3, npart))
        st[3:6] = g.normal(size=(3, npart)) + g.normal(size=(3, 1))
        q = g.normal(size=(4, npart))
        st[6:10] = q / np.linalg.norm(q, axis=0)
        st[10:13] = g.normal(size=(3, npart)) * 50.0 + 10.0
        states.append(st)
        idps.append(g.integers(1, npoly + 1, npart).astype(np.int32))
    species = np.concatenate([W.species_spheroid(emaj[j], apr[j], rhop[j]) for j in range(nig)])
    return dict(L=L, emaj=emaj, apr=apr, rhop=rhop, states=states, idps=idps, species=species,
                nig=nig, npoly=npoly, npart=npart)


def _flat(c):
    st = np.concatenate(c["states"], axis=1)
    sid = np.concatenate([(j * c["npoly"] + c["idps"][j] - 1) for j in range(c["nig"])])
    return st, sid.astype(np.int32)


@pytest.mark.skipif(not os.path.exists(EXE), reason="reference build missing")
def test_oracle_matches_reference_stat_particle(tmp_path):
    c = _classes(4)
    raw = tmp_path / "in.raw"
    with open(raw, "wb") as f:
        np.array([c["nig"], c["npoly"], c["npart"], 16], dtype=np.int32).tofile(f)
        np.array([c["L"]]).tofile(f)
        ncl = np.array([np.bincount(c["idps"][j], minlength=c["npoly"] + 1)[1:]
                        for j in range(c["nig"])], dtype=np.int32)
        ncl.tofile(f)
        c["rhop"].tofile(f)
        c["emaj"].tofile(f)
        c["apr"].tofile(f)
        for j in range(c["nig"]):
            np.ascontiguousarray(c["states"][j]).tofile(f)
            c["idps"][j].tofile(f)
    r = subprocess.run([EXE, "pstat", str(raw), "out.raw"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = np.fromfile(tmp_path / "out.raw", np.float64).reshape(c["nig"] * c["npoly"], 41)
    st, sid = _flat(c)
    _close(S.particle_stats(st, sid, c["species"], npoly=c["npoly"]), ref)
    # the reference's normalisation is the last class's count per size index
    # (so the rows of the last class are the plain per-row means)
    own = S.particle_stats(st, sid, c["species"], npoly=c["npoly"], ref_norm=False)
    last = slice((c["nig"] - 1) * c["npoly"], None)
    _close(own[last], ref[last])
    assert not np.allclose(own[:c["npoly"], 0], ref[:c["npoly"], 0])


def _write_classes(f, c, states=None):
    np.array([c["nig"], c["npoly"], c["npart"], 16], dtype=np.int32).tofile(f)
    np.array([c["L"]]).tofile(f)
    ncl = np.array([np.bincount(c["idps"][j], minlength=c["npoly"] + 1)[1:]
                    for j in range(c["nig"])], dtype=np.int32)
    ncl.tofile(f)
    c["rhop"].tofile(f)
    c["emaj"].tofile(f)
    c["apr"].tofile(f)
    for sts in (states or [c["states"]]):
        for j in range(c["nig"]):
            np.ascontiguousarray(sts[j]).tofile(f)
            c["idps"][j].tofile(f)


def _spread(c, seed):
    """velocities and spins over the PDF ranges (and beyond, for the end
    bins)"""
    g = np.random.default_rng(seed)
    for st in c["states"]:
        n = st.shape[1]
        st[3:6] = g.uniform(-2.5, 2.5, (3, n))
        st[10:13] = g.uniform(-2500.0, 2500.0, (3, n))


@pytest.mark.skipif(not os.path.exists(EXE), reason="reference build missing")
def test_oracle_pdf_matches_reference(tmp_path):
    c = _classes(5)
    _spread(c, 6)
    with open(tmp_path / "in.raw", "wb") as f:
        _write_classes(f, c)
    r = subprocess.run([EXE, "ppdf", "in.raw", "out.raw"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = np.fromfile(tmp_path / "out.raw", np.float64).reshape(c["nig"] * c["npoly"], S.NPDF, 12)
    ref = ref.transpose(0, 2, 1)
    st, sid = _flat(c)
    got = S.particle_pdf(st, sid, c["nig"] * c["npoly"])
    np.testing.assert_array_equal(got, ref)
    # every particle in one bin of each histogram, end bins used
    cnt = np.round(got)
    np.testing.assert_array_equal(cnt.sum(axis=2), np.repeat(np.bincount(sid)[:, None], 12, 1))
    assert (cnt[:, 0:6, 0] > 0).all() and (cnt[:, 0:6, -1] > 0).all()


@pytest.mark.skipif(not os.path.exists(EXE), reason="reference build missing")
def test_oracle_lagrangian_matches_reference(tmp_path):
    c = _classes(7)
    g = np.random.default_rng(8)
    states = []
    for k in range(3):
        cyc = [st.copy() for st in c["states"]]
        for st in cyc:
            st[3:6] = 0.7 * c["states"][0][3:6].mean() + (1.0 - 0.2 * k) * st[3:6] + \
                0.3 * g.normal(size=st[3:6].shape)
        states.append(cyc)
    with open(tmp_path / "in.raw", "wb") as f:
        _write_classes(f, c, states)
    r = subprocess.run([EXE, "plag", "in.raw", "out.raw"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nspec = c["nig"] * c["npoly"]
    ref = np.fromfile(tmp_path / "out.raw", np.float64).reshape(nspec, 3, 3)   # [s, xyz, ntlgr]
    sid = _flat(c)[1]
    npm = S.npm_of(sid, nspec, c["npoly"])
    acc = np.zeros((nspec, 3, 3))
    u0 = None
    for k in range(3):
        st = np.concatenate(states[k], axis=1)
        mean = S.particle_stats(st, sid, c["species"], npoly=c["npoly"])[:, 0:3]
        if k == 0:
            u0 = S.lagrangian_origin(st, sid, mean)
        acc[:, :, k] = S.lagrangian_corr(st, sid, mean, npm, u0)
    _close(acc.reshape(nspec, 9), ref.reshape(nspec, 9))
    # R(0) is the (reference-normalised) variance: positive
    assert (ref[:, :, 0] > 0).all()


def test_oracle_closed_forms():
    """equal velocities: no fluctuations; two particles: the variance is the
    half-difference squared; a lab-frame quantity at the identity rotation
    equals the body-frame one"""
    c = _classes(9, nig=1, npoly=1, npart=2)
    st, sid = _flat(c)
    st[6:10] = [[1.0, 1.0], [0.0, 0.0], [0.0, 0.0], [0.0, 0.0]]
    st[3:6, 1] = st[3:6, 0] + np.array([2.0, 0.0, -4.0])
    out = S.particle_stats(st, sid, c["species"])[0]
    assert out[22] == 2.0
    np.testing.assert_allclose(out[3:6], [1.0, 0.0, 4.0], rtol=1e-14)
    np.testing.assert_allclose(out[7], -2.0, rtol=1e-14)             # <u'w'>
    np.testing.assert_allclose(out[26:32], out[16:22], rtol=1e-14)   # R = I
    np.testing.assert_allclose(out[9], 0.5 * (out[3] + out[4] + out[5]), rtol=1e-15)
    st[3:6, 1] = st[3:6, 0]
    out = S.particle_stats(st, sid, c["species"])[0]
    np.testing.assert_array_equal(out[3:10], 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,nig,npoly", [(G.PREC_FP64, 2, 3), (G.PREC_MIXED, 2, 3),
                                            (G.PREC_FP64, 4, 5)])
def test_gpu_particle_stats_match_oracle(prec, nig, npoly):
    """(2, 3): six species, accumulated in LDS; (4, 5): twenty, global
    atomics.  The oracle reads the state back from the context (in the
    mixed form velocities are stored in float)."""
    wl = W.c1(np_=2000)
    c = _classes(17 + nig, nig=nig, npoly=npoly, npart=wl.np_ // nig)
    st, sid = _flat(c)
    ctx = G.Context(wl.n, wl.L, wl.visc, wl.rhof, wl.dtime, scheme=wl.scheme, dpmax=0.05,
                    precision=prec)
    try:
        ctx.bind_field(wl.field)
        ctx.set_species(c["species"])
        ctx.set_particles(st, sid)
        for steps in (0, 3):
            ctx.step(steps)
            cur = ctx.get_particles()
            for ref_norm in (True, False):
                got = ctx.particle_stats(npoly=npoly, ref_norm=ref_norm)
                _close(got, S.particle_stats(cur, sid, c["species"], npoly=npoly,
                                             ref_norm=ref_norm))
    finally:
        ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,nig,npoly", [(G.PREC_FP64, 2, 3), (G.PREC_MIXED, 2, 3),
                                            (G.PREC_FP64, 4, 5)])
def test_gpu_pdf_and_lagrangian_match_oracle(prec, nig, npoly):
    """(2, 3): six species rows, LDS counters; (4, 5): twenty, global ones.
    PDF counts exactly; the correlation sums over three statistics cycles
    (the origin at the first) to 1e-12."""
    wl = W.c1(np_=3000)
    c = _classes(31 + nig, nig=nig, npoly=npoly, npart=wl.np_ // nig)
    _spread(c, 3)
    st, sid = _flat(c)
    nspec = nig * npoly
    ctx = G.Context(wl.n, wl.L, wl.visc, wl.rhof, wl.dtime, scheme=wl.scheme, dpmax=0.05,
                    precision=prec)
    try:
        ctx.bind_field(wl.field)
        ctx.set_species(c["species"])
        ctx.set_particles(st, sid)
        npm = S.npm_of(sid, nspec, npoly)
        u0 = None
        for k in range(3):
            cur = ctx.get_particles()
            np.testing.assert_array_equal(ctx.particle_pdf(), S.particle_pdf(cur, sid, nspec))
            mean = ctx.particle_stats(npoly=npoly)[:, 0:3]
            if k == 0:
                ctx.lagrangian_origin(0, mean)
                u0 = S.lagrangian_origin(cur, sid, mean)
            got = ctx.lagrangian_corr(0, mean, npm)
            _close(got, S.lagrangian_corr(cur, sid, mean, npm, u0))
            ctx.step(2)
        with pytest.raises(G.GpError):
            ctx.lagrangian_corr(5, mean, npm)
    finally:
        ctx.close()
