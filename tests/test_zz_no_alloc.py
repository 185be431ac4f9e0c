"""The no-allocation entry point, o3d_points_affect_points_noalloc_dev, against
o3d_points_affect_points_dev (method DIRECT) through the C ABI, on device arrays.

Three calls in a row, each smaller than the one before, so that the static scratch
keeps a larger call's contents when the next one starts:
  * two rings of 524,288 particles (EXP, compensated): the Morton target order runs
    in the static scratch; the _dev call also reads the sources in that order (from
    sorted rows in its own workspace), this call in index order, so the sums run in
    another order: within 1e-6 (normwise relative) of the _dev call;
  * a ring of 20,000 particles (RM, compensated): the _dev call splits the sources
    into more slices than the static scratch holds partial sums for, so this call
    takes fewer slices: within 1e-6 (normwise relative) of the _dev call, and
    within 1e-6 of the FP64 kernel on sampled targets;
  * a ring of 4,096 particles (WL, plain FP32, self pair excluded): both calls
    split the same way, so the bits must agree.
"""
import ctypes as C

import numpy as np
import pytest

from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from omega3d_amd import _lib
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return _lib


def _run(L, fn, p, core, prec, self_term=0):
    """one call of fn on device copies of p (sources = targets), outputs zeroed first"""
    D = L.DeviceArray.from_host
    dpos, dstr, drad = D(p.pos), D(p.str), D(p.rad)
    dv, dg = L.DeviceArray((3, p.n)), L.DeviceArray((9, p.n))
    dv.zero()
    dg.zero()
    s = L.Points()
    s.n = p.n
    for d in range(3):
        s.x[d], s.s[d], s.u[d] = dpos.row(d), dstr.row(d), dv.row(d)
    s.r = drad.ptr
    for k in range(9):
        s.ug[k] = dg.row(k)
    o = L.make_opts(core, True, prec, self_term=self_term, method="direct")
    L.check(fn(C.byref(s), C.byref(s), C.byref(o), None))
    L.synchronize()
    return dv.to_host(), dg.to_host()


def _fp64(L, p, core, sel):
    from omega3d_amd.workloads import _fp64_self
    return _fp64_self(core, p, sel)


def test_noalloc_matches_dev_on_three_calls(L):
    na, dev = L.lib.o3d_points_affect_points_noalloc_dev, L.lib.o3d_points_affect_points_dev
    cases = [(scenes.colliding_rings(n_each=524288), "exp", "comp", 0, False),
             (scenes.ring(20000, sigma=0.03, seed=scenes.SEED_BASE + 7), "rm", "comp", 0, False),
             (scenes.ring(4096, sigma=0.05, seed=scenes.SEED_BASE + 8), "wl", "fp32", 1, True)]
    for p, core, prec, self_term, bitwise in cases:
        print(f"noalloc: n={p.n} core={core} precision={prec} self_term={self_term}", flush=True)
        v0, g0 = _run(L, dev, p, core, prec, self_term)
        v1, g1 = _run(L, na, p, core, prec, self_term)
        assert np.isfinite(v1).all() and np.isfinite(g1).all()
        if bitwise:
            assert np.array_equal(v0, v1) and np.array_equal(g0, g1), (p.n, core)
        else:
            ev, eg = rel_errors(v1, v0), rel_errors(g1, g0)
            print(f"noalloc: n={p.n} against the _dev call: vel {ev} grad {eg}", flush=True)
            assert ev[0] <= 1e-6 and eg[0] <= 1e-6
            sel = np.linspace(0, p.n - 1, 512).astype(np.int64)
            vr, gr = _fp64(L, p, core, sel)
            ev, eg = rel_errors(v1[:, sel], vr), rel_errors(g1[:, sel], gr)
            print(f"noalloc: n={p.n} against the FP64 kernel: vel {ev} grad {eg}", flush=True)
            assert ev[0] <= 1e-6 and eg[0] <= 1e-6


def test_noalloc_refusals(L):
    na = L.lib.o3d_points_affect_points_noalloc_dev
    p = scenes.ring(1024, sigma=0.05)
    D = L.DeviceArray.from_host
    dpos, dstr, drad = D(p.pos), D(p.str), D(p.rad)
    dv, dg = L.DeviceArray((3, p.n)), L.DeviceArray((9, p.n))
    s = L.Points()
    s.n = p.n
    for d in range(3):
        s.x[d], s.s[d], s.u[d] = dpos.row(d), dstr.row(d), dv.row(d)
    s.r = drad.ptr
    for k in range(9):
        s.ug[k] = dg.row(k)
    for method, prec in (("factored", "comp"), ("direct", "fp64")):
        o = L.make_opts("rm", True, prec, method=method)
        assert na(C.byref(s), C.byref(s), C.byref(o), None) == 1, method
    t = L.Points()
    C.memmove(C.byref(t), C.byref(s), C.sizeof(s))
    t.n = 4194305                           # one past the maximum: refused before any launch
    o = L.make_opts("rm", True, "comp", method="direct")
    assert na(C.byref(s), C.byref(t), C.byref(o), None) == 1
    assert "at most 4194304" in L.last_error()
