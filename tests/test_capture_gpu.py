"""Device entry points under HIP graph capture.

A small direct call splits its sources over workgroups and takes its partial sums
from the stream's device workspace, which cannot grow while the stream is being
captured.  Captured after a call of the same size on that stream, the call replays
bitwise equal to its eager result, on every replay, with eager work in between.
Captured on a stream whose workspace is too small, it takes the one-pass kernel
instead (the same sum in another order).  The same holds for the target order and
sorted source rows of large EXP calls, and for the factored method.
"""
import ctypes as C

import numpy as np
import pytest

from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    from omega3d_amd import _lib
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    hip = C.CDLL("libamdhip64.so")
    for f in ("hipStreamBeginCapture", "hipStreamEndCapture", "hipGraphInstantiate", "hipGraphLaunch",
              "hipGraphExecDestroy", "hipGraphDestroy"):
        getattr(hip, f).restype = C.c_int
    return _lib, hip


def _setup(_lib, n, stream):
    p = scenes.ring(n, sigma=0.05)
    D = _lib.DeviceArray.from_host
    keep = [D(p.pos), D(p.str), D(p.rad)]
    dv, dg = _lib.DeviceArray((3, n)), _lib.DeviceArray((9, n))
    s = _lib.Points()
    s.n = n
    for d in range(3):
        s.x[d], s.s[d], s.u[d] = keep[0].row(d), keep[1].row(d), dv.row(d)
    s.r = keep[2].ptr
    for k in range(9):
        s.ug[k] = dg.row(k)
    o = _lib.make_opts("rm", True, "comp", self_term=1, stream=stream)
    return keep, dv, dg, s, o


def _capture(_lib, hip, stream, fn):
    assert hip.hipStreamBeginCapture(C.c_void_p(stream), 0) == 0   # hipStreamCaptureModeGlobal
    rc = fn()
    graph = C.c_void_p()
    end = hip.hipStreamEndCapture(C.c_void_p(stream), C.byref(graph))
    assert rc == 0, _lib.last_error()
    assert end == 0
    exe = C.c_void_p()
    assert hip.hipGraphInstantiate(C.byref(exe), graph, None, None, C.c_size_t(0)) == 0
    return graph, exe


def test_captured_split_call_replays_bitwise(env):
    _lib, hip = env
    from omega3d_amd.batch import stream_create, stream_destroy
    st = stream_create()
    try:
        keep, dv, dg, s, o = _setup(_lib, 8192, st)
        call = lambda: _lib.lib.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(o), None)
        # eager, which also sizes the stream's workspace
        dv.zero(); dg.zero(); _lib.synchronize()
        _lib.check(call())
        _lib.check(_lib.lib.o3d_stream_synchronize(st))
        v0, g0 = dv.to_host(), dg.to_host()
        graph, exe = _capture(_lib, hip, st, call)
        try:
            other = _setup(_lib, 3000, None)
            for rep in range(3):
                dv.zero(); dg.zero(); _lib.synchronize()
                assert hip.hipGraphLaunch(exe, C.c_void_p(st)) == 0
                _lib.check(_lib.lib.o3d_stream_synchronize(st))
                np.testing.assert_array_equal(dv.to_host(), v0)
                np.testing.assert_array_equal(dg.to_host(), g0)
                # eager work on the library's own stream between replays
                _lib.check(_lib.lib.o3d_points_affect_points_dev(C.byref(other[3]), C.byref(other[3]),
                                                                 C.byref(other[4]), None))
                _lib.synchronize()
        finally:
            hip.hipGraphExecDestroy(exe)
            hip.hipGraphDestroy(graph)
    finally:
        stream_destroy(st)


def test_captured_ordered_exp_split_replays_bitwise(env):
    """An EXP call on one collection of 40,000 particles splits its sources, visits
    its targets in Morton order and reads its sources as sorted rows, all in the
    stream's workspace after the partial sums.  Captured after the eager call sized
    that workspace, every replay gives the eager bits."""
    _lib, hip = env
    from omega3d_amd.batch import stream_create, stream_destroy
    st = stream_create()
    try:
        keep, dv, dg, s, _ = _setup(_lib, 40000, st)
        o = _lib.make_opts("exp", True, "comp", self_term=1, stream=st, method="direct")
        call = lambda: _lib.lib.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(o), None)
        dv.zero(); dg.zero(); _lib.synchronize()
        _lib.check(call())
        _lib.check(_lib.lib.o3d_stream_synchronize(st))
        v0, g0 = dv.to_host(), dg.to_host()
        graph, exe = _capture(_lib, hip, st, call)
        try:
            for rep in range(3):
                dv.zero(); dg.zero(); _lib.synchronize()
                assert hip.hipGraphLaunch(exe, C.c_void_p(st)) == 0
                _lib.check(_lib.lib.o3d_stream_synchronize(st))
                np.testing.assert_array_equal(dv.to_host(), v0)
                np.testing.assert_array_equal(dg.to_host(), g0)
        finally:
            hip.hipGraphExecDestroy(exe)
            hip.hipGraphDestroy(graph)
    finally:
        stream_destroy(st)


def test_capture_on_a_fresh_stream_falls_back_to_one_pass(env):
    _lib, hip = env
    from omega3d_amd.batch import stream_create, stream_destroy
    st = stream_create()
    try:
        # 16,384 targets: a partial buffer four times that of the 8,192 call above,
        # so the capture cannot use a workspace left on a reused stream handle
        keep, dv, dg, s, o = _setup(_lib, 16384, st)
        call = lambda: _lib.lib.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(o), None)
        dv.zero(); dg.zero(); _lib.synchronize()
        _lib.check(_lib.lib.o3d_stream_synchronize(st))
        graph, exe = _capture(_lib, hip, st, call)
        try:
            assert hip.hipGraphLaunch(exe, C.c_void_p(st)) == 0
            _lib.check(_lib.lib.o3d_stream_synchronize(st))
            v1, g1 = dv.to_host(), dg.to_host()
        finally:
            hip.hipGraphExecDestroy(exe)
            hip.hipGraphDestroy(graph)
        dv.zero(); dg.zero(); _lib.synchronize()
        _lib.check(call())
        _lib.check(_lib.lib.o3d_stream_synchronize(st))
        v0, g0 = dv.to_host(), dg.to_host()
        assert rel_errors(v1, v0)[0] <= 2e-6 and rel_errors(g1, g0)[0] <= 2e-6
    finally:
        stream_destroy(st)


def test_bem_matrix_refuses_capture(env):
    """The BEM matrix sizes its pair list from a count read back between its two
    passes, so under capture it returns a status before it enqueues anything and
    the capture ends cleanly."""
    _lib, hip = env
    from omega3d_amd.batch import stream_create, stream_destroy
    st = stream_create()
    try:
        D = _lib.DeviceArray.from_host
        keep = [D(np.eye(3, dtype=np.float32)), D(np.zeros(3, np.float32)), D(np.ones(1, np.float32))]
        p = _lib.Panels()
        p.np, p.nn = 1, 3
        for d in range(3):
            p.x[d] = keep[0].row(d)
        p.idx, p.area = keep[1].ptr, keep[2].ptr
        out = _lib.DeviceArray((4,))
        assert hip.hipStreamBeginCapture(C.c_void_p(st), 0) == 0
        rc = _lib.lib.o3d_bem_matrix_dev(C.byref(p), C.byref(p), 2, 1, 3, C.c_void_p(out.ptr), C.c_void_p(st))
        graph = C.c_void_p()
        end = hip.hipStreamEndCapture(C.c_void_p(st), C.byref(graph))
        assert rc == 1 and "cannot be captured" in _lib.last_error()
        assert end == 0
        hip.hipGraphDestroy(graph)
    finally:
        stream_destroy(st)


def test_captured_exp_call_skips_target_order_with_same_bits(env):
    """A direct EXP call of 40,000 targets visits them in Morton order (a target's
    sums do not depend on the order).  Captured after the eager call has sized the
    stream's workspace, the order's sort -- kernels only -- is recorded too, and
    the replays give the eager result's bits."""
    _lib, hip = env
    from omega3d_amd.batch import stream_create, stream_destroy
    st = stream_create()
    try:
        n = 40000
        p = scenes.ring(n, sigma=0.02)
        D = _lib.DeviceArray.from_host
        keep = [D(p.pos), D(p.str), D(p.rad)]
        dv, dg = _lib.DeviceArray((3, n)), _lib.DeviceArray((9, n))
        s = _lib.Points()
        s.n = n
        for d in range(3):
            s.x[d], s.s[d], s.u[d] = keep[0].row(d), keep[1].row(d), dv.row(d)
        s.r = keep[2].ptr
        for k in range(9):
            s.ug[k] = dg.row(k)
        o = _lib.make_opts("exp", True, "comp", tile=256, tpt=1, stream=st, method="direct")
        call = lambda: _lib.lib.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(o), None)
        dv.zero(); dg.zero(); _lib.synchronize()
        _lib.check(call())
        _lib.check(_lib.lib.o3d_stream_synchronize(st))
        v0, g0 = dv.to_host(), dg.to_host()
        graph, exe = _capture(_lib, hip, st, call)
        try:
            for rep in range(2):
                dv.zero(); dg.zero(); _lib.synchronize()
                assert hip.hipGraphLaunch(exe, C.c_void_p(st)) == 0
                _lib.check(_lib.lib.o3d_stream_synchronize(st))
                np.testing.assert_array_equal(dv.to_host(), v0)
                np.testing.assert_array_equal(dg.to_host(), g0)
        finally:
            hip.hipGraphExecDestroy(exe)
            hip.hipGraphDestroy(graph)
    finally:
        stream_destroy(st)


def test_captured_sorted_exp_call_replays_bitwise(env):
    """A direct EXP call on one collection of 1,048,576 particles, configuration left
    to the library, orders its targets and reads its sources as sorted rows (sort,
    gather and tile bounds: kernels only).  Captured after the eager call has sized
    the stream's workspace, every replay gives the eager result's bits, with eager
    work in between; and those bits differ from the index-order sum of an explicit
    tile, which shows that the sorted path was the one captured."""
    _lib, hip = env
    from omega3d_amd.batch import stream_create, stream_destroy
    st = stream_create()
    try:
        p = scenes.colliding_rings(n_each=524288)
        n = p.n
        D = _lib.DeviceArray.from_host
        keep = [D(p.pos), D(p.str), D(p.rad)]
        dv, dg = _lib.DeviceArray((3, n)), _lib.DeviceArray((9, n))
        s = _lib.Points()
        s.n = n
        for d in range(3):
            s.x[d], s.s[d], s.u[d] = keep[0].row(d), keep[1].row(d), dv.row(d)
        s.r = keep[2].ptr
        for k in range(9):
            s.ug[k] = dg.row(k)
        o = _lib.make_opts("exp", True, "comp", stream=st, method="direct")
        call = lambda: _lib.lib.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(o), None)
        dv.zero(); dg.zero(); _lib.synchronize()
        _lib.check(call())
        _lib.check(_lib.lib.o3d_stream_synchronize(st))
        v0, g0 = dv.to_host(), dg.to_host()
        print(f"sorted capture: N={n} eager done", flush=True)
        graph, exe = _capture(_lib, hip, st, call)
        try:
            for rep in range(3):
                dv.zero(); dg.zero(); _lib.synchronize()
                assert hip.hipGraphLaunch(exe, C.c_void_p(st)) == 0
                _lib.check(_lib.lib.o3d_stream_synchronize(st))
                np.testing.assert_array_equal(dv.to_host(), v0)
                np.testing.assert_array_equal(dg.to_host(), g0)
                print(f"sorted capture: replay {rep} equal", flush=True)
        finally:
            hip.hipGraphExecDestroy(exe)
            hip.hipGraphDestroy(graph)
        ou = _lib.make_opts("exp", True, "comp", tile=512, tpt=1, stream=st, method="direct")
        dv.zero(); dg.zero(); _lib.synchronize()
        _lib.check(_lib.lib.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(ou), None))
        _lib.check(_lib.lib.o3d_stream_synchronize(st))
        vu, gu = dv.to_host(), dg.to_host()
        assert not (np.array_equal(vu, v0) and np.array_equal(gu, g0))
        assert rel_errors(v0, vu)[0] <= 1e-6 and rel_errors(g0, gu)[0] <= 1e-6
    finally:
        stream_destroy(st)


def test_captured_factored_call_replays_bitwise(env):
    """An AUTO call of 70,000 particles (RM, compensated) takes the factored method:
    Morton keys, the library's radix sorts and scans, far and near kernels -- all
    kernel launches.  Captured after an eager call of the same size on the stream,
    it records the whole pipeline and replays the eager bits, with eager work in
    between; captured on a fresh stream, whose workspace cannot grow, it falls back
    to the direct kernel and replays that kernel's bits."""
    _lib, hip = env
    from omega3d_amd.batch import stream_create, stream_destroy
    n = 70000
    for fresh in (False, True):
        st = stream_create()
        try:
            keep, dv, dg, s, o = _setup(_lib, n, st)
            o.self_term = 0
            call = lambda: _lib.lib.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(o), None)
            if fresh:
                od = _lib.make_opts("rm", True, "comp", tile=512, tpt=1, stream=st, method="direct")
                ref_call = lambda: _lib.lib.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(od), None)
                st2 = stream_create()
                try:
                    od.stream = st2
                    dv.zero(); dg.zero(); _lib.synchronize()
                    _lib.check(ref_call())
                    _lib.check(_lib.lib.o3d_stream_synchronize(st2))
                finally:
                    stream_destroy(st2)
            else:
                dv.zero(); dg.zero(); _lib.synchronize()
                _lib.check(call())
                _lib.check(_lib.lib.o3d_stream_synchronize(st))
            v0, g0 = dv.to_host(), dg.to_host()
            print(f"capture: factored AUTO n={n} fresh={fresh}", flush=True)
            graph, exe = _capture(_lib, hip, st, call)
            try:
                for rep in range(3):
                    dv.zero(); dg.zero(); _lib.synchronize()
                    assert hip.hipGraphLaunch(exe, C.c_void_p(st)) == 0
                    _lib.check(_lib.lib.o3d_stream_synchronize(st))
                    np.testing.assert_array_equal(dv.to_host(), v0)
                    np.testing.assert_array_equal(dg.to_host(), g0)
                    p = scenes.ring(3000, sigma=0.05)
                    from omega3d_amd import influence
                    influence.p2p("rm", p.pos, p.str, p.rad, p.pos, p.rad, True, "comp")   # eager work
            finally:
                hip.hipGraphExecDestroy(exe)
                hip.hipGraphDestroy(graph)
        finally:
            stream_destroy(st)
