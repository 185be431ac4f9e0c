"""Serving: Ralston RK2 steps of device-resident sessions by continuous batching.

Checks, through the C ABI:
  * one step against the oracle's restatement of advect_2nd_ralston (oracle/rk2.py):
    positions and strengths within 4 float ulps (plus 2e-5 of the increment) of the
    reference's float semantics, increments within 1e-4 normwise of the FP64
    restatement (float storage of positions ~1 puts a 6e-8 absolute floor under
    increments of ~1e-3);
  * the same session stepped alone and in a mix of concurrent requests ends in
    bit-identical state (fixed tile, per-target summation order);
  * admission by arena footprint, cancellation, deadlines, injected failures that
    leave the other requests' results unchanged, and the latency statistics.
"""
import numpy as np
import pytest

from oracle import rk2
from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def serve():
    from omega3d_amd import _lib, serve
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return serve


def _ring(n, seed, core_sigma=0.05):
    return scenes.ring(n, sigma=core_sigma, seed=seed)


@pytest.mark.parametrize("core,prec", [("rm", "comp"), ("exp", "comp"), ("wl", "fp32")])
def test_step_matches_oracle(serve, core, prec):
    p = _ring(3000, 11)
    dt, fs = 0.02, (0.05, 0.0, -0.02)
    srv = serve.Server(hbm_budget=256 << 20, tick_interactions=2e6, adaptive_tick=False)  # several ticks/stage
    try:
        sid = srv.open_session(p, core)
        info = srv.wait(srv.submit(sid, dt, prec, fs))
        assert info["state"] == "done" and info["rc"] == 0, info
        x, s, e = srv.read_session(sid)
    finally:
        srv.close()
    x32, s32, e32 = rk2.rk2_step(core, p.pos, p.str, p.rad, np.ones(p.n, np.float32), dt, fs, "f32d")
    x64, s64, e64 = rk2.rk2_step(core, p.pos, p.str, p.rad, np.ones(p.n), dt, fs, "f64")
    dx, dx64 = x - p.pos, x64 - p.pos.astype(np.float64)
    ds, ds64 = s - p.str, s64 - p.str.astype(np.float64)
    print(core, prec, "dx vs f64", rel_errors(dx, dx64), "ds vs f64", rel_errors(ds, ds64))
    _close_float(x, x32, x32 - p.pos)
    _close_float(s, s32, s32 - p.str)
    assert rel_errors(dx, dx64)[0] <= 1e-4
    assert rel_errors(ds, ds64)[0] <= 1e-4
    assert rel_errors((e - 1)[None, :], (e64 - 1)[None, :])[0] <= 1e-4


def _close_float(a, b, inc):
    """a within 4 ulps of b plus 2e-5 of the increment, elementwise"""
    tol = 4 * np.spacing(np.abs(b)) + 2e-5 * np.abs(inc).max()
    bad = np.abs(a.astype(np.float64) - b) > tol
    assert not bad.any(), (np.abs(a - b).max(), tol.max())


def test_two_steps_in_order(serve):
    p = _ring(1500, 12)
    srv = serve.Server(hbm_budget=128 << 20)
    try:
        sid = srv.open_session(p, "rm")
        r1 = srv.submit(sid, 0.01, "comp")
        r2 = srv.submit(sid, 0.01, "comp")
        assert srv.wait(r2)["state"] == "done"
        assert srv.wait(r1)["state"] == "done"
        x, s, e = srv.read_session(sid)
    finally:
        srv.close()
    x1, s1, e1 = rk2.rk2_step("rm", p.pos, p.str, p.rad, np.ones(p.n), 0.01, (0, 0, 0), "f64")
    x2, s2, e2 = rk2.rk2_step("rm", x1, s1, p.rad, e1, 0.01, (0, 0, 0), "f64")
    assert rel_errors(x - p.pos, x2 - p.pos)[0] <= 1e-4
    assert rel_errors(s - p.str, s2 - p.str)[0] <= 1e-4


def test_alone_vs_mix_bitwise(serve):
    """A session's step gives the same bits alone and among concurrent requests of
    other sessions (different sizes, cores and precisions)."""
    target = _ring(2500, 21)
    srv = serve.Server(hbm_budget=512 << 20, tile=256, tpt=1)
    try:
        sid = srv.open_session(target, "rm")
        assert srv.wait(srv.submit(sid, 0.015, "comp"))["state"] == "done"
        alone = srv.read_session(sid)
    finally:
        srv.close()
    srv = serve.Server(hbm_budget=512 << 20, tile=256, tpt=1, tick_interactions=5e6, adaptive_tick=False)
    try:
        others = []
        for k in range(6):
            q = _ring(400 + 700 * k, 30 + k)
            others.append(srv.open_session(q, ["rm", "exp", "wl"][k % 3]))
        sid = srv.open_session(target, "rm")
        rids = [srv.submit(o, 0.01, ["comp", "fp32"][k % 2]) for k, o in enumerate(others)]
        mine = srv.submit(sid, 0.015, "comp")
        rids += [srv.submit(o, 0.01, "comp") for o in others[:3]]
        for r in rids + [mine]:
            assert srv.wait(r)["state"] == "done"
        mixed = srv.read_session(sid)
        st = srv.stats()
        assert st["completed"] == len(rids) + 1 and st["ticks"] > 1
    finally:
        srv.close()
    for a, b in zip(alone, mixed):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("core,prec", [("rm", "comp"), ("exp", "fp32")])
def test_factored_step_matches_oracle(serve, core, prec):
    """method FACTORED: both stages by the factored far field on the second stream;
    the same bars as the ticks' steps against the oracle's RK2 restatement."""
    p = _ring(3000, 11)
    dt, fs = 0.02, (0.05, 0.0, -0.02)
    srv = serve.Server(hbm_budget=256 << 20, method="factored")
    try:
        sid = srv.open_session(p, core)
        info = srv.wait(srv.submit(sid, dt, prec, fs))
        assert info["state"] == "done" and info["rc"] == 0, info
        x, s, e = srv.read_session(sid)
        st = srv.stats()
        assert st["factored_steps"] == 1 and st["factored_scratch"] > 0 and st["ticks"] == 0
    finally:
        srv.close()
    x32, s32, e32 = rk2.rk2_step(core, p.pos, p.str, p.rad, np.ones(p.n, np.float32), dt, fs, "f32d")
    x64, s64, e64 = rk2.rk2_step(core, p.pos, p.str, p.rad, np.ones(p.n), dt, fs, "f64")
    dx, dx64 = x - p.pos, x64 - p.pos.astype(np.float64)
    ds, ds64 = s - p.str, s64 - p.str.astype(np.float64)
    print(core, prec, "factored: dx vs f64", rel_errors(dx, dx64), "ds vs f64", rel_errors(ds, ds64))
    _close_float(x, x32, x32 - p.pos)
    _close_float(s, s32, s32 - p.str)
    assert rel_errors(dx, dx64)[0] <= 1e-4
    assert rel_errors(ds, ds64)[0] <= 1e-4
    assert rel_errors((e - 1)[None, :], (e64 - 1)[None, :])[0] <= 1e-4


def test_large_session_factored_alone_vs_mix(serve):
    """A 70,000-particle RM session takes the factored way by default (AUTO): its step
    agrees with the ticks' step (method DIRECT) to float rounding of the state, and
    gives the same bits alone and among concurrent small requests."""
    big = _ring(70000, 41, core_sigma=0.03)
    dt = 0.01

    def step_alone(method):
        srv = serve.Server(hbm_budget=1 << 30, method=method)
        try:
            sid = srv.open_session(big, "rm")
            assert srv.wait(srv.submit(sid, dt, "comp"))["state"] == "done"
            st = srv.stats()
            return srv.read_session(sid), st
        finally:
            srv.close()

    alone, st = step_alone("auto")
    assert st["factored_steps"] == 1 and st["ticks"] == 0
    direct, st_d = step_alone("direct")
    assert st_d["factored_steps"] == 0 and st_d["ticks"] > 0
    _close_float(alone[0], direct[0], direct[0] - big.pos)
    _close_float(alone[1], direct[1], direct[1] - big.str)
    srv = serve.Server(hbm_budget=1 << 30, tile=256, tpt=1, tick_interactions=5e6, adaptive_tick=False)
    try:
        others = [srv.open_session(_ring(500 + 900 * k, 50 + k), ["rm", "exp", "wl"][k % 3]) for k in range(5)]
        sid = srv.open_session(big, "rm")
        rids = [srv.submit(o, 0.01, ["comp", "fp32"][k % 2]) for k, o in enumerate(others)]
        mine = srv.submit(sid, dt, "comp")
        rids += [srv.submit(o, 0.01, "comp") for o in others]
        for r in rids + [mine]:
            assert srv.wait(r)["state"] == "done"
        mixed = srv.read_session(sid)
        st = srv.stats()
        assert st["factored_steps"] == 1 and st["completed"] == len(rids) + 1
    finally:
        srv.close()
    for a, b in zip(alone, mixed):
        np.testing.assert_array_equal(a, b)


def test_admission_cancel_deadline(serve):
    from omega3d_amd import _lib
    big = _ring(20000, 41)
    small = _ring(500, 42)
    # arena: 20000 * 32 B state fits, the 20000 * 120 B step does not
    srv = serve.Server(hbm_budget=2 << 20)
    try:
        sid = srv.open_session(big, "rm")
        with pytest.raises(_lib.O3DError) as ei:
            srv.submit(sid, 0.01)
        assert ei.value.code == 5   # O3D_EADMIT
        with pytest.raises(_lib.O3DError) as ei:
            srv.open_session(_ring(70000, 43), "rm")      # 2.24 MB of state
        assert ei.value.code == 5
        srv.close_session(sid)
        s2 = srv.open_session(small, "rm")
        # a queued request can be cancelled; one whose deadline passes in the queue ends cancelled
        r1 = srv.submit(s2, 0.01)
        r2 = srv.submit(s2, 0.01)          # waits behind r1 (same session)
        assert srv.cancel(r2) in (0, 6)
        w1, w2 = srv.wait(r1), srv.wait(r2)
        assert w1["state"] == "done"
        assert w2["state"] in ("cancelled", "done")
        if w2["state"] == "cancelled":
            assert w2["status"] == 7
    finally:
        srv.close()


def test_injected_failure_leaves_others_intact(serve):
    pa, pb, pc = _ring(1200, 51), _ring(900, 52), _ring(700, 53)
    ref = {}
    srv = serve.Server(hbm_budget=256 << 20)
    try:
        for name, p in (("b", pb), ("c", pc)):
            sid = srv.open_session(p, "exp")
            assert srv.wait(srv.submit(sid, 0.01, "comp"))["state"] == "done"
            ref[name] = srv.read_session(sid)
    finally:
        srv.close()
    srv = serve.Server(hbm_budget=256 << 20)
    try:
        sa, sb, sc = (srv.open_session(p, "exp") for p in (pa, pb, pc))
        serve.inject_fault(1)
        ra = srv.submit(sa, 0.01, "comp")
        rb = srv.submit(sb, 0.01, "comp")
        rc = srv.submit(sc, 0.01, "comp")
        wa = srv.wait(ra)
        assert wa["state"] == "failed" and wa["status"] == 4   # O3D_EINJECTED
        assert srv.wait(rb)["state"] == "done" and srv.wait(rc)["state"] == "done"
        for name, sid in (("b", sb), ("c", sc)):
            for a, b in zip(ref[name], srv.read_session(sid)):
                np.testing.assert_array_equal(a, b)
        # the failed step left session a untouched
        xa, sa_, ea = srv.read_session(sa)
        np.testing.assert_array_equal(xa, pa.pos)
        np.testing.assert_array_equal(sa_, pa.str)
        st = srv.stats()
        assert st["failed"] == 1 and st["completed"] == 2 and st["p99_s"] >= st["p50_s"] > 0
    finally:
        serve.inject_fault(0)
        srv.close()


def test_rejects_non_finite_input(serve):
    from omega3d_amd import _lib
    p = _ring(100, 61)
    p.pos[1, 7] = np.nan
    srv = serve.Server(hbm_budget=16 << 20)
    try:
        with pytest.raises(_lib.O3DError) as ei:
            srv.open_session(p, "rm")
        assert ei.value.code == 1
    finally:
        srv.close()
