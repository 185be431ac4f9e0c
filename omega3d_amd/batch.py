"""Many independent particle collections evaluated together on the GPU.

Wraps o3d_batch_* of the C ABI: the instances' particles are concatenated into
one set of device SoA arrays; one persistent launch (or its captured HIP graph)
evaluates every instance on itself (velocity + gradient).  run_separate() does
the same work as one ordinary launch per instance, for comparison.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import CORES, PREC, check, lib

lib.o3d_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p]
lib.o3d_batch_destroy.argtypes = [C.c_void_p]
for _n in ("o3d_batch_run", "o3d_batch_run_separate", "o3d_batch_capture"):
    getattr(lib, _n).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    getattr(lib, _n).restype = C.c_int
lib.o3d_batch_replay.argtypes = [C.c_void_p, C.c_void_p]
lib.o3d_batch_tune.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                               C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_float)]
lib.o3d_batch_config.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int64)]
lib.o3d_batch_factored.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
for _n in ("o3d_batch_create", "o3d_batch_destroy", "o3d_batch_replay", "o3d_batch_tune", "o3d_batch_config",
           "o3d_batch_factored",
           "o3d_tune_p2p", "o3d_tune_save", "o3d_tune_load"):
    getattr(lib, _n).restype = C.c_int
lib.o3d_tune_save.argtypes = [C.c_char_p]
lib.o3d_tune_load.argtypes = [C.c_char_p]


class Batch:
    def __init__(self, instances, core="rm", precision="comp", self_term=0, tile=0, tpt=0,
                 cores=None, precisions=None, self_terms=None, method="auto"):
        self.counts = np.array([p.n for p in instances], np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)[:-1]]).astype(np.int64)
        self.total = int(self.counts.sum())
        pos = np.concatenate([p.pos for p in instances], axis=1) if instances else np.zeros((3, 0), np.float32)
        st = np.concatenate([p.str for p in instances], axis=1) if instances else np.zeros((3, 0), np.float32)
        rad = np.concatenate([p.rad for p in instances]) if instances else np.zeros(0, np.float32)
        D = _lib.DeviceArray.from_host
        self.dpos, self.dstr, self.drad = D(pos), D(st), D(rad)
        self.dvel = _lib.DeviceArray((3, max(self.total, 1)))
        self.dgrad = _lib.DeviceArray((9, max(self.total, 1)))
        self.all = _lib.Points()
        self.all.n = self.total
        for d in range(3):
            self.all.x[d] = self.dpos.row(d) if self.total else self.dpos.ptr
            self.all.s[d] = self.dstr.row(d) if self.total else self.dstr.ptr
            self.all.u[d] = self.dvel.row(d)
        self.all.r = self.drad.ptr
        for k in range(9):
            self.all.ug[k] = self.dgrad.row(k)
        o = _lib.make_opts(core, True, precision, self_term, tile, tpt, method=method)
        k = len(instances)
        arr = lambda v, f: None if v is None else np.ascontiguousarray([f(x) for x in v], np.int32)
        self._cores = arr(cores, lambda c: CORES[c] if isinstance(c, str) else int(c))
        self._precs = arr(precisions, lambda p: PREC[p] if isinstance(p, str) else int(p))
        self._selfs = arr(self_terms, int)
        P = lambda a: None if a is None else a.ctypes.data
        h = C.c_void_p()
        check(lib.o3d_batch_create(C.byref(h), k, self.offsets.ctypes.data, self.counts.ctypes.data,
                                   P(self._cores), P(self._precs), P(self._selfs), C.byref(o)))
        self.h = h.value

    def run(self, stream=None):
        check(lib.o3d_batch_run(self.h, C.byref(self.all), stream))

    def run_separate(self, stream=None):
        check(lib.o3d_batch_run_separate(self.h, C.byref(self.all), stream))

    def capture(self, stream):
        check(lib.o3d_batch_capture(self.h, C.byref(self.all), stream))

    def replay(self, stream):
        check(lib.o3d_batch_replay(self.h, stream))

    def tune(self, stream=None, tiles=(128, 256, 512), tpts=(1, 2, 4), reps=1):
        t = np.ascontiguousarray(tiles, np.int32)
        p = np.ascontiguousarray(tpts, np.int32)
        bt, bp, ms = C.c_int32(), C.c_int32(), C.c_float()
        check(lib.o3d_batch_tune(self.h, C.byref(self.all), stream, t.ctypes.data, len(t), p.ctypes.data, len(p),
                                 reps, C.byref(bt), C.byref(bp), C.byref(ms)))
        return bt.value, bp.value, ms.value

    def config(self):
        t, p, g, n = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int64()
        check(lib.o3d_batch_config(self.h, C.byref(t), C.byref(p), C.byref(g), C.byref(n)))
        nf, wb = C.c_int32(), C.c_int64()
        check(lib.o3d_batch_factored(self.h, C.byref(nf), C.byref(wb)))
        return {"tile": t.value, "tpt": p.value, "grid": g.value, "items": n.value,
                "factored_instances": nf.value, "factored_scratch_bytes": wb.value}

    def results(self):
        v = self.dvel.to_host()[:, : self.total]
        g = self.dgrad.to_host()[:, : self.total]
        return [(v[:, o:o + n], g[:, o:o + n]) for o, n in zip(self.offsets, self.counts)]

    def close(self):
        if getattr(self, "h", None):
            lib.o3d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_create():
    s = C.c_void_p()
    check(lib.o3d_stream_create(C.byref(s)))
    return s.value


def stream_destroy(s):
    lib.o3d_stream_destroy(s)
