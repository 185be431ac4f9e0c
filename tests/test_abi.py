"""CPU tests of the C ABI: the library loads, exports every declared entry point,
reports the reference's flop counts and rejects bad arguments with status codes
(no compute: nothing here needs a GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "o3d_biot.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|float|void|double|char)\s*\*?\s*(\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from omega3d_amd import _lib
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(_lib.lib, n)]
    assert not missing, missing


def test_reference_plugin_symbols_present():
    from omega3d_amd import _lib
    for n in ("external_vel_solver_f_", "external_vel_solver_d_"):
        assert hasattr(_lib.lib, n)


def test_flops_per_interaction_matches_reference():
    from omega3d_amd import _lib
    f = np.load(os.path.join(ROOT, "tests", "golden", "flops.npz"))
    for core in ("rm", "wl", "exp", "v2"):
        got = [_lib.flops_per_interaction(core, True, True), _lib.flops_per_interaction(core, True, False),
               _lib.flops_per_interaction(core, False, True), _lib.flops_per_interaction(core, False, False),
               _lib.flops_per_interaction(core, True, False, True),
               _lib.flops_per_interaction(core, False, False, True)]
        assert got == list(f[core])
    assert _lib.flops_per_interaction(9) == -1


def test_default_opts():
    from omega3d_amd import _lib
    o = _lib.Opts()
    _lib.lib.o3d_default_opts(C.byref(o))
    assert (o.core, o.results, o.precision, o.self_term, o.tile, o.tpt, o.maxlev, o.method) == (2, 2, 1, 0, 0, 0, -1, 0)
    assert _lib.lib.o3d_abi_version() == 1


def _pts(n, grads=True, thick=True):
    from omega3d_amd import _lib
    from omega3d_amd.influence import _pts_struct
    pos = np.zeros((3, n), np.float32)
    st = np.zeros((3, n), np.float32)
    rad = np.ones(n, np.float32)
    vel = np.zeros((3, n), np.float32)
    grad = np.zeros((9, n), np.float32) if grads else None
    keep = (pos, st, rad, vel, grad)
    return _pts_struct(pos, st, rad if thick else None, vel, grad), keep


def test_argument_errors_are_status_codes():
    from omega3d_amd import _lib
    L = _lib.lib
    s, k1 = _pts(10)
    t, k2 = _pts(10)
    o = _lib.make_opts("rm")
    assert L.o3d_points_affect_points(C.byref(s), C.byref(t), None, None) == 1
    bad = _lib.make_opts("rm"); bad.core = 7
    assert L.o3d_points_affect_points(C.byref(s), C.byref(t), C.byref(bad), None) == 1
    assert "core" in _lib.last_error()
    bad = _lib.make_opts("rm"); bad.tile = 100
    assert L.o3d_points_affect_points(C.byref(s), C.byref(t), C.byref(bad), None) == 1
    bad = _lib.make_opts("rm"); bad.tpt = 3
    assert L.o3d_points_affect_points(C.byref(s), C.byref(t), C.byref(bad), None) == 1
    bad = _lib.make_opts("rm"); bad.results = 3      # psionly is not served
    assert L.o3d_points_affect_points(C.byref(s), C.byref(t), C.byref(bad), None) == 1
    t2, k3 = _pts(10, grads=False)
    assert L.o3d_points_affect_points(C.byref(s), C.byref(t2), C.byref(o), None) == 1
    assert "gradient" in _lib.last_error()
    t3, k4 = _pts(7)
    ex = _lib.make_opts("rm", self_term=1)
    assert L.o3d_points_affect_points(C.byref(s), C.byref(t3), C.byref(ex), None) == 1
    s.r = None
    assert L.o3d_points_affect_points(C.byref(s), C.byref(t), C.byref(o), None) == 1
    assert L.o3d_set_external_core(0) == 1


def _gpu_visible():
    from omega3d_amd import _lib
    return _lib.lib.o3d_init(-1) == 0


def test_no_device_is_reported_not_crashed():
    if _gpu_visible():
        pytest.skip("a GPU is visible")
    from omega3d_amd import _lib
    s, k1 = _pts(10)
    t, k2 = _pts(10)
    o = _lib.make_opts("rm")
    assert _lib.lib.o3d_points_affect_points(C.byref(s), C.byref(t), C.byref(o), None) == 8
    # empty collections are a no-op that still validates, and need no device
    s0, k3 = _pts(0)
    assert _lib.lib.o3d_points_affect_points(C.byref(s0), C.byref(t), C.byref(o), None) in (0, 8)


def test_out_of_range_node_index_is_rejected_before_any_launch():
    """Host-memory panels are checked on the host: a node index at or beyond the
    node count is a status code, not a device access out of bounds."""
    from omega3d_amd import _lib
    L = _lib.lib
    nodes = np.ascontiguousarray(np.eye(3, 4, dtype=np.float32))
    idx = np.array([[0, 1, 2], [1, 2, 4]], np.uint32)            # node 4 of 4
    area = np.ones(2, np.float32)
    ts = np.zeros((3, 2), np.float32)
    p = _lib.Panels()
    p.np, p.nn = 2, 4
    for d in range(3):
        p.x[d] = nodes[d].ctypes.data
        p.ts[d] = ts[d].ctypes.data
    p.idx, p.area = idx.ctypes.data, area.ctypes.data
    t, keep = _pts(5)
    o = _lib.make_opts("rm")
    assert L.o3d_panels_affect_points(C.byref(p), C.byref(t), C.byref(o), None) == 1
    assert "panel 1 names node 4 of 4" in _lib.last_error()
    vel = np.zeros((3, 2), np.float32)
    pv = (C.c_void_p * 3)(*[vel[d].ctypes.data for d in range(3)])
    ov = _lib.make_opts("rm", False)
    assert L.o3d_points_affect_panels(C.byref(t), C.byref(p), pv, C.byref(ov), None) == 1
    assert "names node 4" in _lib.last_error()
    m = np.zeros(2 * 2 * 2 * 2, np.float32)
    assert L.o3d_bem_matrix(C.byref(p), C.byref(p), 2, 1, 3, m.ctypes.data_as(C.POINTER(C.c_float))) == 1
    assert "names node 4" in _lib.last_error()


def test_no_allocation_call_refuses_before_any_launch():
    """The no-allocation entry point refuses what its fixed scratch cannot serve --
    the factored method, FP64, more than O3D_NOALLOC_MAX_TARGETS targets -- on the
    host, before it touches a device."""
    from omega3d_amd import _lib
    L = _lib.lib
    s, k1 = _pts(10)
    t, k2 = _pts(10)
    na = L.o3d_points_affect_points_noalloc_dev
    assert na(C.byref(s), C.byref(t), C.byref(_lib.make_opts("rm", method="factored")), None) == 1
    assert "factored" in _lib.last_error()
    assert na(C.byref(s), C.byref(t), C.byref(_lib.make_opts("rm", precision="fp64")), None) == 1
    big, k3 = _pts(10)
    big.n = 4194305
    assert na(C.byref(s), C.byref(big), C.byref(_lib.make_opts("rm", method="direct")), None) == 1
    assert "at most 4194304" in _lib.last_error()
    assert "O3D_NOALLOC_MAX_TARGETS 4194304LL" in open(HEADER).read()
