"""The Fortran boundary end to end: the reference's own ADV_PARTICLE_ELLP +
COLLISION and the drop-ins GPU_ADV_PARTICLE_ELLP + GPU_COLLISION of
gazpart_amd/fortran/gp_dropin.f90 (iso_c_binding over the C-ABI) run on the
same module state inside one Fortran program (oracle/ref/dropin_test.f90,
built by `make -C oracle/ref dropin` where the reference sources exist).

Tolerance: 1e-9 of each quantity's scale after 10 steps of a dense
ensemble, excluding particles whose trajectories left that band because a
contact or wall decision went the other way (at most 5 %, listed)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "gp_dropin_test")


def _write_input(path, d, nsteps, mirror, fout2=0, interph=0, evolve=0):
    n, np_ = int(d["n"]), d["state0"].shape[1]
    emaj = np.atleast_1d(np.asarray(d["emaj"], dtype=np.float64))
    rhop = np.atleast_1d(np.asarray(d["rhop"], dtype=np.float64)) * np.ones(len(emaj))
    idp = d["idp"] if "idp" in d else np.ones(np_, dtype=np.int32)
    with open(path, "wb") as f:
        np.array([n, np_, nsteps, int(d["scheme"]), mirror, len(emaj), fout2, interph, evolve],
                 dtype=np.int32).tofile(f)
        np.array([float(d["L"]), float(d["visc"]), float(d["rhof"]), float(d["dtime"]),
                  float(d["apr"])]).tofile(f)
        emaj.tofile(f)
        rhop.tofile(f)
        np.ascontiguousarray(d["field"], dtype=np.float64).tofile(f)
        np.ascontiguousarray(d["state0"], dtype=np.float64).tofile(f)
        np.ascontiguousarray(idp, dtype=np.int32).tofile(f)
    return np_


def _far(ref, gpu):
    bad = set()
    for sl in (slice(0, 3), slice(3, 6), slice(6, 10), slice(10, 13)):
        scale = np.abs(ref[sl]).max()
        far = np.nonzero((np.abs(gpu[sl] - ref[sl]) > 1e-9 * scale).any(axis=0))[0]
        bad.update(far.tolist())
    return bad


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran drop-in test driver not built")
@pytest.mark.parametrize("mirror", [1, 0])
def test_fortran_dropin_matches_reference(tmp_path, mirror):
    d = np.load(os.path.join(ROOT, "tests", "golden", "golden_trajectory.npz"))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np_ = _write_input(fin, d, 10, mirror)
    subprocess.run([EXE, str(fin), str(fout)], check=True, timeout=120)
    out = np.fromfile(fout, dtype=np.float64).reshape(2, 13, np_)
    ref, gpu = out[0], out[1]
    assert not np.array_equal(ref, d["state0"])
    bad = _far(ref, gpu)
    assert len(bad) <= 0.05 * np_, sorted(bad)


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran drop-in test driver not built")
def test_fortran_dropin_collision_statistics(tmp_path):
    """With LEVEL1_STPAR, STAT_TIME and FOUT2 = 2, GPU_COLLISION writes the
    same per-pass statistics lines as COLLISION_INTERACTION (three sizes,
    nine units fort.561-569, e17.7) and leaves the same MEAN_TIME_PART_COL:
    counts exactly, sums within 1e-9 (summation order); the printed lines
    (7 significant digits) within one unit of the last digit."""
    d = np.load(os.path.join(ROOT, "tests", "golden", "golden_collision_stats.npz"))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np_ = _write_input(fin, d, int(d["nsteps"]), 1, fout2=int(d["fout2"]))
    subprocess.run([EXE, str(fin), str(fout)], check=True, timeout=120, cwd=tmp_path)
    out = np.fromfile(fout, dtype=np.float64)
    ref, gpu = out[:13 * np_].reshape(13, np_), out[13 * np_:26 * np_].reshape(13, np_)
    assert len(_far(ref, gpu)) == 0
    mref, mgpu = out[26 * np_:].reshape(2, 3, 3, 8)   # (IDP2, IDP1, k)
    np.testing.assert_array_equal(mgpu[..., :3], mref[..., :3])
    np.testing.assert_allclose(mgpu, mref, rtol=1e-9, atol=1e-30)
    # the sums of the reference's own passes in the fixture (from 0 here)
    np.testing.assert_allclose(mref.transpose(1, 0, 2), d["col"].sum(axis=0), rtol=1e-15)
    npass = len(d["stat_steps"])
    for unit in range(561, 570):
        rows = np.loadtxt(tmp_path / f"fort.{unit}")
        assert rows.shape == (2 * npass, 9)
        np.testing.assert_allclose(rows[npass:], rows[:npass], rtol=2e-6, atol=1e-30)


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran drop-in test driver not built")
def test_fortran_interph_before_init(tmp_path):
    """GPU_INTERPH where INITIATION_PARTICLE_VELOCITY calls INTERPH
    (main.f90:286), before GPU_INIT: the six components at the particles
    equal the reference's INTERPH to 1e-13 of their scale."""
    d = np.load(os.path.join(ROOT, "tests", "golden", "golden_trajectory.npz"))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np_ = _write_input(fin, d, 1, 1, interph=1)
    subprocess.run([EXE, str(fin), str(fout)], check=True, timeout=120)
    out = np.fromfile(fout, dtype=np.float64)
    iref = out[26 * np_:32 * np_].reshape(6, np_)
    igpu = out[32 * np_:38 * np_].reshape(6, np_)
    for c in range(6):
        scale = np.abs(iref[c]).max()
        assert scale > 0
        assert np.abs(igpu[c] - iref[c]).max() <= 1e-13 * scale, c


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran drop-in test driver not built")
@pytest.mark.parametrize("mirror", [1, 0])
def test_fortran_dropin_evolving_field(tmp_path, mirror):
    """SOLVE_FLUID = 1 with the field and DTIME changed half-way through (as
    ADV_FLUID and CHECK_CFL change them every cycle): the drop-ins take the
    new field and time step, in mirror and in resident mode, and match the
    reference as in the frozen case."""
    d = np.load(os.path.join(ROOT, "tests", "golden", "golden_trajectory.npz"))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np_ = _write_input(fin, d, 10, mirror, evolve=1)
    subprocess.run([EXE, str(fin), str(fout)], check=True, timeout=120)
    out = np.fromfile(fout, dtype=np.float64).reshape(2, 13, np_)
    ref, gpu = out[0], out[1]
    bad = _far(ref, gpu)
    assert len(bad) <= 0.05 * np_, sorted(bad)
    # the same run with a frozen field ends elsewhere: the change was seen
    fin2, fout2 = tmp_path / "in2.bin", tmp_path / "out2.bin"
    _write_input(fin2, d, 10, mirror)
    subprocess.run([EXE, str(fin2), str(fout2)], check=True, timeout=120)
    frozen = np.fromfile(fout2, dtype=np.float64).reshape(2, 13, np_)[0]
    assert np.abs(frozen[0:3] - ref[0:3]).max() > 1e-6
