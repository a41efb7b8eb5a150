"""Launch shapes that must not change results, and the autotuner that picks
among them (gazpart_amd/autotune.py)."""
import numpy as np
import pytest
import torch

import gazpart_amd as G
from gazpart_amd import workload as W
from gazpart_amd.autotune import tune

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dense():
    # 4000 prolate spheroids on 64^3 with a = 0.5 DX: ~tens of contacts per step
    return W.build_workload("dense", 64, 4000, k0=4.0, seed=9, apr=2.0, emaj_dx=0.5, st=1.0)


@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
def test_contact_launch_width_changes_no_result(dense, prec):
    out, ncoll = [], []
    for waves in (0, 1, 64, 5000):
        ctx = G.Context.from_workload(dense, precision=prec, lm_waves=waves)
        n = 0
        for _ in range(5):
            ctx.step(1)
            n += ctx.stats()["ncollision"]
        # the width used: as asked, or (0) one wavefront per SIMD while the
        # solves (tens here) leave lanes idle
        nsimd = 4 * torch.cuda.get_device_properties(0).multi_processor_count
        assert ctx.stats()["lm_waves"] == (waves if waves > 0 else nsimd)
        out.append(ctx.get_particles())
        ncoll.append(n)
        ctx.close()
    assert ncoll[0] > 0 and len(set(ncoll)) == 1
    for o in out[1:]:
        np.testing.assert_array_equal(out[0], o)


@pytest.mark.parametrize("graph", [0, 2])
def test_xcd_block_order_changes_no_result(dense, graph):
    """Workgroups placed round-robin over the XCDs (the default) or in one
    contiguous range per XCD (xcd_remap = 1): the same bits, eager and in
    pipelined steps with the pairs solved ahead."""
    out = []
    for x in (0, 1):
        ctx = G.Context.from_workload(dense, precision=G.PREC_MIXED, graph=graph, overlap=1,
                                      xcd_remap=x)
        ctx.step(7)
        out.append(ctx.get_particles())
        ctx.close()
    np.testing.assert_array_equal(out[0], out[1])


def test_autotune_picks_the_fastest_of_its_table(dense):
    r = tune(dense, G.PREC_MIXED, steps=2, warmup=1, bricks=(-1, 4), tiles=(8,),
             lm_waves=(0, 256))
    assert len(r["table"]) == 12
    assert {t["tile"] for t in r["table"]} == {0, 8}
    assert {t["graph"] for t in r["table"]} == {1, 2}
    assert r["best"] == min(r["table"], key=lambda t: t["ms_per_step"])
    assert all(t["ms_per_step"] > 0 for t in r["table"])
