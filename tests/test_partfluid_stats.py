"""STAT_PARTFLUID (stat_partfluid.f90:113-381): particle-fluid statistics per
class and size from the particles and the fluid at them of the last advance.

The fixture (tests/golden/golden_partfluid_stats.npz, written by
tests/golden/make_golden.py from the reference's own Fortran) holds two
classes (APR 2 and 5) of two sizes: the state before and after one
ADV_PARTICLE_ELLP, the fluid at the particles that advance left
(UFAP .. VORTZAP) and the routine's MEAN_PARTFLUID(1:36, J, IDP).

Tolerances
  oracle (numpy, CPU) : 1e-13 relative per entry (pow, atan and the
                        summation order differ from the Fortran's).
  FP64 GPU path       : the advance 1e-12 of each quantity's scale; the
                        statistics 1e-10 relative per entry.
  mixed GPU path      : FP32 field and gather, so the fluid at the particles
                        carries ~1e-7 relative error: statistics within
                        1e-4 of each entry's largest magnitude over the rows.
"""
import numpy as np
import pytest

import gazpart_amd as G
from gazpart_amd import workload as W
from oracle import stats_ref as S

GOLD = "tests/golden/golden_partfluid_stats.npz"


@pytest.fixture(scope="module")
def fx():
    return np.load(GOLD)


def _inputs(d):
    rows = W.species_spheroid(d["emaj"].ravel(), np.repeat(d["apr"], 2), d["rhop"].ravel())
    sid = np.concatenate([j * 2 + d["idp"][j] - 1 for j in range(2)]).astype(np.int32)
    cat = lambda a: np.concatenate(list(a), axis=1)  # noqa: E731
    return rows, sid, cat(d["state0"]), cat(d["state1"]), cat(d["fluid1"])


def test_oracle_against_reference_fixture(fx):
    rows, sid, _, st1, fl1 = _inputs(fx)
    got = S.partfluid_stats(st1, fl1, sid, rows, float(fx["visc"]), float(fx["rhof"]), npoly=2)
    ref = fx["mean"].reshape(4, 36)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)


def test_fixture_normalisation_quirk(fx):
    """Every class is divided by the LAST class's count per size (the
    reference's NPM(IDP) after its class loop), entry 36 is that count,
    and a sum that only ever adds zeros keeps the ZERO it started from."""
    ref = fx["mean"].reshape(4, 36)
    last = np.bincount(fx["idp"][1], minlength=3)[1:]
    np.testing.assert_array_equal(ref[:, 35], np.tile(last, 2))
    np.testing.assert_array_equal(ref[:, 14], 1.0e-33 / np.tile(last, 2))   # pitching x = 0
    # the oracle without the quirk divides by each row's own count
    rows, sid, _, st1, fl1 = _inputs(fx)
    own = S.partfluid_stats(st1, fl1, sid, rows, float(fx["visc"]), float(fx["rhof"]), npoly=2,
                            ref_norm=False)
    n_own = np.bincount(sid, minlength=4)
    np.testing.assert_array_equal(own[:, 35], n_own)
    np.testing.assert_allclose(own[:, :35] * n_own[:, None], ref[:, :35] * ref[:, 35:36],
                               rtol=1e-13)
    # unit major axis: the mean of unit vectors has norm <= 1
    assert (np.linalg.norm(own[:, 25:28], axis=1) <= 1.0 + 1e-12).all()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
def test_gpu_against_reference_fixture(fx, prec):
    rows, sid, st0, st1, fl1 = _inputs(fx)
    n, L = int(fx["n"]), float(fx["L"])
    ctx = G.Context(n, L, float(fx["visc"]), float(fx["rhof"]), float(fx["dtime"]),
                    precision=prec, keep_fluid=True)
    ctx.bind_field(fx["field"])
    ctx.set_species(rows)
    ctx.set_particles(st0, sid)
    with pytest.raises(G.GpError):
        ctx.partfluid_stats(npoly=2)          # no advance yet
    ctx.adv_particle_ellp(1)
    got = ctx.partfluid_stats(npoly=2)
    ref = fx["mean"].reshape(4, 36)
    if prec == G.PREC_FP64:
        st = ctx.get_particles()
        for sl in (slice(0, 3), slice(3, 6), slice(6, 10), slice(10, 13)):
            assert np.abs(st[sl] - st1[sl]).max() <= 1e-12 * np.abs(st1[sl]).max()
        fl = ctx.get_fluid_at_particles()
        assert np.abs(fl - fl1).max() <= 1e-12 * np.abs(fl1).max()
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=0)
    else:
        scale = np.abs(ref).max(axis=0)
        assert (np.abs(got - ref) <= 1e-4 * scale).all(), np.abs(got - ref) / scale
        np.testing.assert_array_equal(got[:, 35], ref[:, 35])
