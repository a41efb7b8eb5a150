"""gp_interph_noalloc (INTERPH through the library's fixed device staging,
no allocation) against gp_interph on the same points: bit-identical, on two
calls in a row (the second with fewer points, so the first call's staging
holds other values), both schemes and both precisions; more points than
the staging holds are refused."""
import numpy as np
import pytest

import gazpart_amd as G
from gazpart_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl():
    return W.c1(np_=1000)


@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
def test_noalloc_equals_interph(wl, prec):
    ctx = G.Context.from_workload(wl, precision=prec)
    g = W.rng(41)
    for npts in (G.INTERPH_NOALLOC_MAX, 1234):
        x, y, z = g.uniform(0.0, wl.L, (3, npts))
        for scheme in (3, 4):
            for c in (0, 4):
                a = ctx.interph(scheme, c, x, y, z)
                b = ctx.interph_noalloc(scheme, c, x, y, z)
                np.testing.assert_array_equal(a, b)
                assert np.abs(a).max() > 0
    ctx.close()


def test_noalloc_refuses_more_points(wl):
    ctx = G.Context.from_workload(wl)
    n = G.INTERPH_NOALLOC_MAX + 1
    x = np.full(n, 1.0)
    with pytest.raises(G.GpError) as e:
        ctx.interph_noalloc(3, 0, x, x, x)
    assert e.value.code == 10
    # a second context on the same GPU uses the same staging
    other = G.Context.from_workload(wl)
    pts = np.linspace(0.1, 6.0, 50)
    np.testing.assert_array_equal(other.interph_noalloc(3, 1, pts, pts, pts),
                                  ctx.interph(3, 1, pts, pts, pts))
    other.close()
    ctx.close()
