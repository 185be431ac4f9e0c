"""Direct EXP calls that read their sources as rows in the targets' Morton order.

A direct EXP call on one collection of >= 1,048,576 particles (too many targets for
the source split) with tile and lane count left to the library sorts its sources
into the targets' Morton order and sums a tile whose bounding sphere lies beyond
every target's far-branch distance with no per-pair test.  Checked here on a cloud
whose core radii vary by a factor of three (so tiles carry different largest core
terms and far-branch distances), both precisions, self pair included and excluded:
  * against the FP64 kernel on 1,024 spread targets: compensated within 1e-6, plain
    FP32 within twice the unsorted plain call's own error plus 1e-6;
  * against the unsorted one-pass kernel (explicit tile 512, one target per lane):
    compensated within 1e-6, and not bitwise equal (the sums run in another order,
    which shows that the sorted path ran);
  * repeated calls give the same bits, and a host-array call that passes the same
    position arrays as sources and targets (copied to the device once, so it too
    acts on one collection) gives the device call's bits.
"""
import numpy as np
import pytest

from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors, self_term_gradient

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cloud():
    from omega3d_amd import _lib
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    c = scenes.colliding_rings(n_each=540000, seed=scenes.SEED_BASE + 71)
    rng = np.random.default_rng(71)
    rad = (0.01 + 0.02 * rng.uniform(size=c.n)).astype(np.float32)
    sel = np.linspace(0, c.n - 1, 1024).astype(np.int64)
    return c, rad, sel


def _dev_call(pos, st, rad, prec, self_term, tile=0, tpt=0):
    """one device-array call with sources and targets the same arrays (one
    collection, as the sorted path needs)"""
    import ctypes as C
    from omega3d_amd import _lib
    n = pos.shape[1]
    D = _lib.DeviceArray.from_host
    dpos, dstr, drad = D(pos), D(st), D(rad)
    dv, dg = _lib.DeviceArray((3, n)), _lib.DeviceArray((9, n))
    dv.zero()
    dg.zero()
    s = _lib.Points()
    s.n = n
    for d in range(3):
        s.x[d], s.s[d], s.u[d] = dpos.row(d), dstr.row(d), dv.row(d)
    s.r = drad.ptr
    for k in range(9):
        s.ug[k] = dg.row(k)
    o = _lib.make_opts("exp", True, prec, self_term=self_term, tile=tile, tpt=tpt, method="direct")
    _lib.check(_lib.lib.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(o), None))
    _lib.synchronize()
    return dv.to_host(), dg.to_host()


@pytest.mark.parametrize("self_term", [0, 1])
def test_sorted_direct_exp_against_fp64_and_unsorted(cloud, self_term):
    from omega3d_amd import influence as infl
    c, rad, sel = cloud
    print(f"sorted sources: N={c.n} self_term={self_term}", flush=True)
    v64, g64 = infl.p2p("exp", c.pos, c.str, rad, c.pos[:, sel], rad[sel], True, "fp64")
    if self_term:
        g64 = g64 - self_term_gradient("exp", c.str[:, sel], rad[sel])
    res = {}
    for prec in ("comp", "fp32"):
        vs, gs = _dev_call(c.pos, c.str, rad, prec, self_term)
        vu, gu = _dev_call(c.pos, c.str, rad, prec, self_term, tile=512, tpt=1)
        es = rel_errors(vs[:, sel], v64), rel_errors(gs[:, sel], g64)
        eu = rel_errors(vu[:, sel], v64), rel_errors(gu[:, sel], g64)
        print(f"  {prec}: sorted vel {es[0]} grad {es[1]}; unsorted vel {eu[0]} grad {eu[1]}", flush=True)
        if prec == "comp":
            assert es[0][0] <= 1e-6 and es[1][0] <= 1e-6
            assert rel_errors(vs, vu)[0] <= 1e-6 and rel_errors(gs, gu)[0] <= 1e-6
        else:
            assert es[0][0] <= 2 * eu[0][0] + 1e-6 and es[1][0] <= 2 * eu[1][0] + 1e-6
        assert not (np.array_equal(vs, vu) and np.array_equal(gs, gu))
        res[prec] = (vs, gs)
    v2, g2 = _dev_call(c.pos, c.str, rad, "comp", self_term)
    np.testing.assert_array_equal(v2, res["comp"][0])
    np.testing.assert_array_equal(g2, res["comp"][1])
    vh, gh = infl.p2p("exp", c.pos, c.str, rad, c.pos, rad, True, "comp", self_term=self_term,
                      method="direct")
    np.testing.assert_array_equal(vh, res["comp"][0])
    np.testing.assert_array_equal(gh, res["comp"][1])
