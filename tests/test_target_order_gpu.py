"""Direct EXP calls of at least 32,768 targets visit the targets in Morton order
(a wave of neighbouring targets leaves the exponential branch together).  A
target's sums depend only on itself, the sources and the tile, so the ordered call
gives the same bits as the same targets evaluated in index order: here, the ordered
call against two calls of half the targets each (below the size, so in index
order), all with the same explicit tile, in every precision and with the self pair
excluded.  The ordered call is also checked against the FP64 kernel.
"""
import numpy as np
import pytest

from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def infl():
    from omega3d_amd import _lib, influence
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return influence


@pytest.mark.parametrize("prec", ["fp32", "comp"])
@pytest.mark.parametrize("grads", [True, False])
def test_ordered_exp_call_is_bitwise_index_order(infl, prec, grads):
    n = 40000
    p = scenes.ring(n, sigma=0.02)
    kw = dict(grads=grads, precision=prec, tile=256, tpt=1, method="direct")
    v, g = infl.p2p("exp", p.pos, p.str, p.rad, p.pos, p.rad, **kw)
    h = n // 2
    v0, g0 = infl.p2p("exp", p.pos, p.str, p.rad, p.pos[:, :h], p.rad[:h], **kw)
    v1, g1 = infl.p2p("exp", p.pos, p.str, p.rad, p.pos[:, h:], p.rad[h:], **kw)
    np.testing.assert_array_equal(v, np.concatenate([v0, v1], 1))
    if grads:
        np.testing.assert_array_equal(g, np.concatenate([g0, g1], 1))


def test_ordered_exp_call_self_excluded_and_against_fp64(infl):
    n = 36000
    p = scenes.ring(n, sigma=0.02, seed=5)
    kw = dict(grads=True, precision="comp", self_term=1, tile=256, tpt=1, method="direct")
    v, g = infl.p2p("exp", p.pos, p.str, p.rad, p.pos, p.rad, **kw)
    # self exclusion needs one collection, so no split into halves here: the
    # ordered call against the FP64 kernel, and against itself
    v64, g64 = infl.p2p("exp", p.pos, p.str, p.rad, p.pos, p.rad, grads=True, precision="fp64", self_term=1)
    assert rel_errors(v, v64)[0] <= 1e-6 and rel_errors(g, g64)[0] <= 1e-6
    # repeated calls: the same bits
    v2, g2 = infl.p2p("exp", p.pos, p.str, p.rad, p.pos, p.rad, **kw)
    np.testing.assert_array_equal(v, v2)
    np.testing.assert_array_equal(g, g2)
