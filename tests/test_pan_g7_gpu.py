"""GPU parity of the 7-point panel rule with the closed-form near field
(o3d_panels_affect_points with method O3D_PANEL_GAUSS7) against the FP64 oracle
(oracle/o3d_panel_quad.cpp, itself pinned by mathematics in test_panel_quad.py).

Bars (normwise relative, omega3d_amd.metrics): compensated FP32 1e-6; plain FP32
2e-5.  FP64 mode sums in FP64 and rounds once into the caller's float arrays, so it
is held to one float ulp of the oracle's FP64 result, elementwise (plus 1e-12 of
the largest entry).  The near/far decision is the same float computation on
both sides, so the GPU's count of closed-form pairs equals the oracle's exactly.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu

BARS = {"comp": 1e-6, "fp32": 2e-5}


def _within_ulp(a, ref):
    """float results within one float ulp of the FP64 reference, elementwise (plus
    1e-12 of the largest entry, for entries that are small by cancellation)"""
    r32 = ref.astype(np.float32)
    tol = np.spacing(np.abs(r32)).astype(np.float64) + 1e-12 * np.abs(ref).max()
    return bool(np.all(np.abs(a.astype(np.float64) - ref) <= tol))


@pytest.fixture(scope="module")
def infl():
    from omega3d_amd import _lib, influence
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return influence


def _case(levels=3, nt=3000, seed=5, r0=1.02, r1=1.5):
    pan = scenes.icosphere_panels(levels=levels)
    rng = np.random.default_rng(seed)
    ss = (0.3 * rng.normal(size=pan.area.shape[0])).astype(np.float32)
    tp = scenes.shell_points(nt, r0, r1, seed=seed)
    return pan, ss, tp


@pytest.mark.parametrize("prec", ["comp", "fp32", "fp64"])
def test_g7_matches_oracle(infl, prec):
    pan, ss, tp = _case()
    print(f"case: gauss7 {prec} panels={pan.area.shape[0]} targets={tp.shape[1]}", flush=True)
    v, g = infl.pan("rm", pan.nodes, pan.idx, pan.ts, pan.area, ss, tp, True, prec, rule="gauss7")
    ov, og, nnear = oracle.pan_g7(pan.nodes, pan.idx, pan.ts, pan.area, ss, tp)
    assert nnear > 0
    ev, eg = rel_errors(v, ov), rel_errors(g, og)
    print(prec, "vel", ev, "grad", eg, "near pairs", nnear)
    if prec == "fp64":
        assert _within_ulp(v, ov) and _within_ulp(g, og)
    else:
        assert ev[0] <= BARS[prec] and eg[0] <= BARS[prec]


def test_g7_velocity_only_tiles_and_accumulation(infl):
    pan, ss, tp = _case(levels=2, nt=1001, seed=6)
    ov, _, _ = oracle.pan_g7(pan.nodes, pan.idx, pan.ts, pan.area, ss, tp, grads=False)
    for tile in (64, 128, 256):
        v, _ = infl.pan("rm", pan.nodes, pan.idx, pan.ts, pan.area, ss, tp, False, "comp", rule="gauss7", tile=tile)
        assert rel_errors(v, ov)[0] <= 1e-6, tile
    v0 = np.random.default_rng(3).normal(size=(3, tp.shape[1])).astype(np.float32)
    v, _ = infl.pan("rm", pan.nodes, pan.idx, pan.ts, pan.area, ss, tp, False, "comp", rule="gauss7", vel=v0.copy())
    np.testing.assert_allclose(v, v0 + ov, rtol=1e-6, atol=1e-6 * np.abs(ov).max())


@pytest.mark.parametrize("np_", [1, 7, 333])
def test_g7_ragged_and_tiny(infl, np_):
    pan, ss, tp = _case(levels=3, nt=257, seed=8, r0=1.01, r1=1.2)
    P = pan.area.shape[0]
    keep = np.random.default_rng(np_).choice(P, np_, replace=False)
    idx = pan.idx.reshape(P, 3)[keep].reshape(-1)
    ts, area, s2 = pan.ts[:, keep].copy(), pan.area[keep].copy(), ss[keep].copy()
    for nt in (1, 3, 257):
        t = tp[:, :nt].copy()
        v, g = infl.pan("rm", pan.nodes, idx, ts, area, s2, t, True, "fp64", rule="gauss7")
        ov, og, _ = oracle.pan_g7(pan.nodes, idx, ts, area, s2, t)
        assert _within_ulp(v, ov) and _within_ulp(g, og), (np_, nt)


def _dev_panels(_lib, pan, ss):
    D = _lib.DeviceArray.from_host
    keep = [D(pan.nodes), D(pan.ts), D(pan.area), D(ss), D(pan.idx.view(np.float32))]
    p = _lib.Panels()
    p.np = pan.area.shape[0]
    p.nn = pan.nodes.shape[1]
    for d in range(3):
        p.x[d] = keep[0].row(d)
        p.ts[d] = keep[1].row(d)
    p.area, p.ss, p.idx = keep[2].ptr, keep[3].ptr, keep[4].ptr
    return p, keep


def test_g7_profile_counts_near_pairs_like_the_oracle(infl):
    from omega3d_amd import _lib
    pan, ss, tp = _case(levels=3, nt=4096, seed=9)
    p, keep = _dev_panels(_lib, pan, ss)
    D = _lib.DeviceArray.from_host
    dt = D(tp)
    dv, dg = _lib.DeviceArray((3, tp.shape[1])), _lib.DeviceArray((9, tp.shape[1]))
    dv.zero(); dg.zero()
    t = _lib.Points()
    t.n = tp.shape[1]
    for d in range(3):
        t.x[d] = dt.row(d)
        t.u[d] = dv.row(d)
    for k in range(9):
        t.ug[k] = dg.row(k)
    o = _lib.make_opts("rm", True, "comp", method="gauss7")
    prof = _lib.PanelProfile()
    _lib.check(_lib.lib.o3d_panels_g7_profile_dev(C.byref(p), C.byref(t), C.byref(o), C.byref(prof)))
    _lib.synchronize()
    ov, og, nnear = oracle.pan_g7(pan.nodes, pan.idx, pan.ts, pan.area, ss, tp)
    assert prof.near_pairs == nnear and prof.far_pairs + prof.near_pairs == prof.pairs
    assert prof.total_ms > 0 and prof.far_ms > 0
    # the profile's full evaluation landed in the caller's arrays, once
    assert rel_errors(dv.to_host(), ov)[0] <= 1e-6 and rel_errors(dg.to_host(), og)[0] <= 1e-6


def test_g7_full_sphere_subset(infl):
    """The 20,480-panel sphere with 524,288 targets in 1.02 < r < 1.5 (compensated):
    1,024 targets against the oracle."""
    pan = scenes.icosphere_panels()
    tp = scenes.shell_points()
    v, g = infl.pan("rm", pan.nodes, pan.idx, pan.ts, pan.area, None, tp, True, "comp", rule="gauss7")
    sel = np.arange(0, tp.shape[1], tp.shape[1] // 1024)[:1024]
    ov, og, nnear = oracle.pan_g7(pan.nodes, pan.idx, pan.ts, pan.area, None, tp[:, sel].copy())
    ev, eg = rel_errors(v[:, sel], ov), rel_errors(g[:, sel], og)
    print("sphere gauss7 comp vs fp64 oracle", ev, eg, "near pairs in the subset", nnear)
    assert ev[0] <= 1e-6 and eg[0] <= 1e-6


def test_g7_rejections(infl):
    from omega3d_amd import _lib
    pan, ss, tp = _case(levels=1, nt=10)
    with pytest.raises(_lib.O3DError) as ei:     # 160 B of LDS per staged panel: at most 256
        infl.pan("rm", pan.nodes, pan.idx, pan.ts, pan.area, ss, tp, True, "comp", rule="gauss7", tile=512)
    assert ei.value.code == 1
    p = scenes.ring(100)
    with pytest.raises(_lib.O3DError) as ei:     # a panel rule, not a point->point method
        infl.p2p("rm", p.pos, p.str, p.rad, p.pos, p.rad, True, "comp", method="gauss7")
    assert ei.value.code == 1
