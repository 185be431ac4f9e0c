"""Size-independent properties at the full two-ring size (N = 1,048,576, EXP core,
one collection): the velocity and gradient are linear in the strengths, and every
operation on a strength is a multiplication or an addition of such products, so
scaling all strengths by 2 or by -1 scales every output by exactly the same factor
(rounding commutes with both: no value here comes near overflow or underflow).
Checked bit for bit through the product path of a direct call (targets in Morton
order, sources read as sorted rows, far tiles with no per-pair test), in both FP32
modes.  The factored method takes the factor 2 exactly (its per-tile scales are
powers of two, so the matrix cores see the same inputs), but not the factor -1: a
v_mfma_f32_32x32x16_f16 product of -A and B is not the exact negation of A B in a
few outputs (4 of 1,024 for random FP16 inputs), so its sums round differently by
sign, within its error bound.  The oracle cannot sum 1.1e12 pairs; these identities
hold whatever the order of the sums, so they check the whole launch at its real
size."""
import numpy as np
import pytest

from omega3d_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rings():
    from omega3d_amd import _lib
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    c = scenes.colliding_rings()
    assert c.n == 1048576
    return c


def _dev_call(pos, st, rad, prec, method):
    """one device-array call with sources and targets the same arrays"""
    import ctypes as C
    from omega3d_amd import _lib
    n = pos.shape[1]
    D = _lib.DeviceArray.from_host
    dpos, dstr, drad = D(pos), D(np.ascontiguousarray(st, np.float32)), D(rad)
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
    o = _lib.make_opts("exp", True, prec, method=method)
    _lib.check(_lib.lib.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(o), None))
    _lib.synchronize()
    return dv.to_host(), dg.to_host()


@pytest.mark.parametrize("method,prec", [("direct", "comp"), ("direct", "fp32"), ("factored", "comp")])
def test_two_rings_scale_exactly_with_the_strengths(rings, method, prec):
    c = rings
    print(f"full-size properties: N={c.n} EXP {method} {prec}", flush=True)
    v, g = _dev_call(c.pos, c.str, c.rad, prec, method)
    assert np.isfinite(v).all() and np.isfinite(g).all()
    assert np.abs(v).max() > 0 and np.abs(g).max() > 0
    v2, g2 = _dev_call(c.pos, 2.0 * c.str, c.rad, prec, method)
    assert np.array_equal(v2, 2.0 * v) and np.array_equal(g2, 2.0 * g), "scaling by 2"
    if method == "direct":
        vn, gn = _dev_call(c.pos, -c.str, c.rad, prec, method)
        assert np.array_equal(vn, -v) and np.array_equal(gn, -g), "negation"


@pytest.mark.parametrize("rule,core", [("recursive", "wl"), ("gauss7", "rm")])
def test_sphere_panels_scale_exactly_with_the_strengths(rings, rule, core):
    """The same identities for the 20,480-panel sphere acting on 524,288 points in
    1.02 < r < 1.5, by the reference's recursive centroid rule and by the 7-point
    rule with its closed-form near field: panel sums are FP32 (the near field FP64)
    products and sums of the sheet strengths, so both factors are exact."""
    from omega3d_amd import influence
    pan = scenes.icosphere_panels()
    tp = scenes.shell_points()
    assert pan.np == 20480 and tp.shape[1] == 524288
    ss = pan.ss if rule == "recursive" else None
    print(f"full-size properties: {pan.np} panels -> {tp.shape[1]} points, {rule} rule", flush=True)

    def call(f):
        return influence.pan(core, pan.nodes, pan.idx, f * pan.ts, pan.area, None if ss is None else f * ss, tp,
                             True, "comp", rule=rule)
    v, g = call(1.0)
    assert np.isfinite(v).all() and np.isfinite(g).all() and np.abs(v).max() > 0
    v2, g2 = call(2.0)
    assert np.array_equal(v2, 2.0 * v) and np.array_equal(g2, 2.0 * g), "scaling by 2"
    vn, gn = call(-1.0)
    assert np.array_equal(vn, -v) and np.array_equal(gn, -g), "negation"


def test_particles_on_sphere_panels_scale_exactly_with_the_strengths(rings):
    """points_affect_panels (the BEM right-hand side's sum) from 524,288 vortex
    particles in 1.02 < r < 1.5 onto the 20,480-panel sphere: both factors exact."""
    from omega3d_amd import influence
    pan = scenes.icosphere_panels()
    pos = scenes.shell_points(524288)
    rng = np.random.default_rng(5)
    st = (rng.standard_normal((3, pos.shape[1])) / pos.shape[1]).astype(np.float32)
    print(f"full-size properties: {pos.shape[1]} particles -> {pan.np} panels", flush=True)

    def call(f):
        return influence.ptpan(pos, f * st, pan.nodes, pan.idx, pan.area, "comp")
    v = call(1.0)
    assert np.isfinite(v).all() and np.abs(v).max() > 0
    assert np.array_equal(call(2.0), 2.0 * v), "scaling by 2"
    assert np.array_equal(call(-1.0), -v), "negation"
