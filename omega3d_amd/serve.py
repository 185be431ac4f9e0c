"""Serving device-resident simulations: Python side of o3d_server_* (C ABI).

    srv = Server(hbm_budget=8 << 30)
    sid = srv.open_session(particles, core="rm")
    rid = srv.submit(sid, dt=0.01, precision="comp")
    info = srv.wait(rid)          # {'state': 'done', 'latency_s': ..., ...}
    pos, st, elong = srv.read_session(sid)

Each request is one Ralston RK2 convection step with vortex stretching of the
session (Convection.h:349-425) evaluated on the GPU; the session's particles stay
in HBM between steps.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import CORES, PREC, check, lib


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", C.c_int32), ("tile", C.c_int32), ("tpt", C.c_int32),
                ("use_graph", C.c_int32), ("adaptive_tick", C.c_int32), ("method", C.c_int32),
                ("tick_interactions", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("submit_s", C.c_double), ("admit_s", C.c_double),
                ("done_s", C.c_double), ("latency_s", C.c_double), ("interactions", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("cancelled", C.c_int64), ("queued", C.c_int64), ("running", C.c_int64),
                ("p50_s", C.c_double), ("p99_s", C.c_double), ("mean_s", C.c_double),
                ("interactions", C.c_double), ("wall_s", C.c_double), ("ginter_per_s", C.c_double),
                ("hbm_used", C.c_size_t), ("hbm_budget", C.c_size_t), ("hbm_peak", C.c_size_t),
                ("ticks", C.c_int64), ("factored_steps", C.c_int64), ("factored_scratch", C.c_size_t)]


STATES = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "cancelled"}

lib.o3d_server_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(ServerOpts)]
lib.o3d_server_destroy.argtypes = [C.c_void_p]
lib.o3d_server_open_session.argtypes = [C.c_void_p, C.POINTER(_lib.Points), C.c_int32, C.POINTER(C.c_int64)]
lib.o3d_server_close_session.argtypes = [C.c_void_p, C.c_int64]
lib.o3d_server_read_session.argtypes = [C.c_void_p, C.c_int64, C.c_void_p * 3, C.c_void_p * 3, C.c_void_p]
lib.o3d_server_submit.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.c_int32, C.c_void_p, C.c_double,
                                  C.POINTER(C.c_int64)]
lib.o3d_server_wait.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.POINTER(RequestInfo)]
lib.o3d_server_cancel.argtypes = [C.c_void_p, C.c_int64]
lib.o3d_server_get_stats.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
lib.o3d_server_reset_stats.argtypes = [C.c_void_p]
lib.o3d_inject_fault.argtypes = [C.c_int32, C.c_int32]
for _n in ("o3d_server_create", "o3d_server_destroy", "o3d_server_open_session", "o3d_server_close_session",
           "o3d_server_read_session", "o3d_server_submit", "o3d_server_wait", "o3d_server_cancel",
           "o3d_server_get_stats", "o3d_server_reset_stats", "o3d_inject_fault"):
    getattr(lib, _n).restype = C.c_int


def inject_fault(count, kind=1):
    check(lib.o3d_inject_fault(kind if count else 0, count))


class Server:
    def __init__(self, hbm_budget=8 << 30, max_active=64, tile=0, tpt=0, use_graph=True, tick_interactions=0.0,
                 adaptive_tick=True, method="auto"):
        m = _lib.METHODS[method] if isinstance(method, str) else int(method)
        o = ServerOpts(hbm_budget, max_active, tile, tpt, int(use_graph), int(adaptive_tick), m, tick_interactions)
        h = C.c_void_p()
        check(lib.o3d_server_create(C.byref(h), C.byref(o)))
        self.h = h.value
        self.sizes = {}

    def open_session(self, p, core="rm"):
        pos = np.ascontiguousarray(p.pos, np.float32)
        st = np.ascontiguousarray(p.str, np.float32)
        rad = np.ascontiguousarray(p.rad, np.float32)
        pts = _lib.Points()
        pts.n = pos.shape[1]
        for d in range(3):
            pts.x[d] = pos[d].ctypes.data
            pts.s[d] = st[d].ctypes.data
        pts.r = rad.ctypes.data
        sid = C.c_int64()
        check(lib.o3d_server_open_session(self.h, C.byref(pts), CORES[core] if isinstance(core, str) else core,
                                          C.byref(sid)))
        self.sizes[sid.value] = pts.n
        return sid.value

    def close_session(self, sid):
        check(lib.o3d_server_close_session(self.h, sid))
        self.sizes.pop(sid, None)

    def read_session(self, sid):
        n = self.sizes[sid]
        pos = np.zeros((3, n), np.float32)
        st = np.zeros((3, n), np.float32)
        el = np.zeros(n, np.float32)
        xp = (C.c_void_p * 3)(*[pos[d].ctypes.data for d in range(3)])
        sp = (C.c_void_p * 3)(*[st[d].ctypes.data for d in range(3)])
        check(lib.o3d_server_read_session(self.h, sid, xp, sp, el.ctypes.data))
        return pos, st, el

    def submit(self, sid, dt, precision="comp", fs=(0.0, 0.0, 0.0), timeout_s=0.0):
        f = np.ascontiguousarray(fs, np.float64)
        rid = C.c_int64()
        check(lib.o3d_server_submit(self.h, sid, dt, PREC[precision] if isinstance(precision, str) else precision,
                                    f.ctypes.data, timeout_s, C.byref(rid)))
        return rid.value

    def wait(self, rid, timeout_s=-1.0):
        info = RequestInfo()
        rc = lib.o3d_server_wait(self.h, rid, timeout_s, C.byref(info))
        return {"rc": rc, "state": STATES.get(info.state, info.state), "status": info.status,
                "submit_s": info.submit_s, "admit_s": info.admit_s, "done_s": info.done_s,
                "latency_s": info.latency_s, "interactions": info.interactions}

    def cancel(self, rid):
        return lib.o3d_server_cancel(self.h, rid)

    def stats(self):
        s = ServerStats()
        check(lib.o3d_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        check(lib.o3d_server_reset_stats(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib.o3d_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
