"""ctypes binding of libo3dbiot.so (the C ABI in include/o3d_biot.h).

The library is loaded from this package directory.  There is no CPU fallback:
if the HIP library is missing, importing this module raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libo3dbiot.so")

CORES = {"rm": 1, "wl": 2, "exp": 3, "v2": 4}
VELONLY, VELANDGRAD = 1, 2
PREC = {"fp32": 0, "comp": 1, "compensated": 1, "fp64": 2}
STATUS = {0: "O3D_OK", 1: "O3D_EINVAL", 2: "O3D_ENOMEM", 3: "O3D_EDEVICE", 4: "O3D_EINJECTED",
          5: "O3D_EADMIT", 6: "O3D_EBUSY", 7: "O3D_ECANCELLED", 8: "O3D_ENODEV", 9: "O3D_ENOTFOUND"}


class O3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Points(C.Structure):
    _fields_ = [("n", C.c_int64), ("x", C.c_void_p * 3), ("s", C.c_void_p * 3), ("r", C.c_void_p),
                ("u", C.c_void_p * 3), ("ug", C.c_void_p * 9)]


class Points64(C.Structure):
    _fields_ = Points._fields_


class Panels(C.Structure):
    _fields_ = [("np", C.c_int64), ("nn", C.c_int64), ("x", C.c_void_p * 3), ("idx", C.c_void_p),
                ("ts", C.c_void_p * 3), ("area", C.c_void_p), ("ss", C.c_void_p)]


class Opts(C.Structure):
    _fields_ = [("core", C.c_int32), ("results", C.c_int32), ("precision", C.c_int32),
                ("self_term", C.c_int32), ("tile", C.c_int32), ("tpt", C.c_int32),
                ("maxlev", C.c_int32), ("method", C.c_int32), ("stream", C.c_void_p)]


class BemRhsOpts(C.Structure):
    _fields_ = [("fs", C.c_double * 3), ("vel", C.c_double * 3), ("rotvel", C.c_double * 3), ("tc", C.c_float * 3),
                ("body", C.c_int32), ("nunk", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("flops", C.c_double), ("interactions", C.c_double)]


if not os.path.exists(LIBPATH):
    raise ImportError(f"{LIBPATH} is not built: run `python -m omega3d_amd.build` "
                      "(there is no CPU fallback for the HIP path)")
lib = C.CDLL(LIBPATH)

_int_fns = ["o3d_init", "o3d_abi_version", "o3d_points_affect_points", "o3d_points_affect_points_dev",
            "o3d_points_affect_points64", "o3d_points_affect_points64_dev",
            "o3d_panels_affect_points", "o3d_panels_affect_points_dev",
            "o3d_flops_per_interaction", "o3d_set_external_core",
            "o3d_dev_malloc", "o3d_dev_free", "o3d_memcpy_htod", "o3d_memcpy_dtoh", "o3d_memset_dev",
            "o3d_device_synchronize", "o3d_stream_create", "o3d_stream_destroy", "o3d_stream_synchronize",
            "o3d_event_create", "o3d_event_destroy", "o3d_event_record", "o3d_event_elapsed_ms",
            "o3d_device_info", "o3d_copy_probe", "o3d_valu_probe", "o3d_factored_profile_dev",
            "o3d_panels_g7_profile_dev",
            "o3d_points_affect_panels", "o3d_points_affect_panels_dev", "o3d_bem_rhs_dev",
            "o3d_bem_rhs_from_vels_dev", "o3d_bem_matrix", "o3d_bem_matrix_dev", "o3d_bem_matrix_profile_dev",
            "o3d_morton_order_dev", "o3d_panels_affect_panels", "o3d_panels_affect_panels_dev",
            "o3d_points_affect_points_noalloc_dev"]
for _n in _int_fns:
    if hasattr(lib, _n):
        getattr(lib, _n).restype = C.c_int
lib.o3d_last_error.restype = C.c_char_p
lib.o3d_default_opts.restype = None
lib.external_vel_solver_f_.restype = C.c_float
lib.external_vel_solver_d_.restype = C.c_float
lib.o3d_dev_malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
lib.o3d_dev_free.argtypes = [C.c_void_p]
lib.o3d_memcpy_htod.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
lib.o3d_memcpy_dtoh.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
lib.o3d_memset_dev.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
lib.o3d_stream_create.argtypes = [C.POINTER(C.c_void_p)]
lib.o3d_stream_destroy.argtypes = [C.c_void_p]
lib.o3d_stream_synchronize.argtypes = [C.c_void_p]
lib.o3d_event_create.argtypes = [C.POINTER(C.c_void_p)]
lib.o3d_event_destroy.argtypes = [C.c_void_p]
lib.o3d_event_record.argtypes = [C.c_void_p, C.c_void_p]
lib.o3d_event_elapsed_ms.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
lib.o3d_flops_per_interaction.argtypes = [C.c_int] * 4
lib.o3d_copy_probe.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]


def check(rc):
    if rc != 0:
        raise O3DError(rc, lib.o3d_last_error().decode(errors="replace"))
    return rc


def last_error():
    return lib.o3d_last_error().decode(errors="replace")


METHODS = {"auto": 0, "direct": 1, "factored": 2, "gauss7": 3}


class FactoredProfile(C.Structure):
    _fields_ = [("prep_ms", C.c_double), ("far_ms", C.c_double), ("near_ms", C.c_double), ("total_ms", C.c_double),
                ("far_pairs", C.c_double), ("pairs", C.c_double), ("far_tile_fraction", C.c_double)]


class PanelProfile(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("far_ms", C.c_double), ("pairs", C.c_double),
                ("near_pairs", C.c_double), ("far_pairs", C.c_double)]


class BemMatrixProfile(C.Structure):
    _fields_ = [("count_ms", C.c_double), ("far_ms", C.c_double), ("near_ms", C.c_double), ("total_ms", C.c_double),
                ("near_pairs", C.c_double)]


def make_opts(core="wl", grads=True, precision="comp", self_term=0, tile=0, tpt=0, maxlev=-1, stream=None,
              method="auto"):
    o = Opts()
    lib.o3d_default_opts(C.byref(o))
    o.core = CORES[core] if isinstance(core, str) else int(core)
    o.results = VELANDGRAD if grads else VELONLY
    o.precision = PREC[precision] if isinstance(precision, str) else int(precision)
    o.self_term = int(self_term)
    o.tile = int(tile)
    o.tpt = int(tpt)
    o.maxlev = int(maxlev)
    o.method = METHODS[method] if isinstance(method, str) else int(method)
    o.stream = stream
    return o


def flops_per_interaction(core, grads=True, thick=True, vs=False):
    c = CORES[core] if isinstance(core, str) else int(core)
    return lib.o3d_flops_per_interaction(c, VELANDGRAD if grads else VELONLY, int(thick), int(vs))


def device_info():
    name = C.create_string_buffer(256)
    cus, clk, hbm = C.c_int(), C.c_int(), C.c_size_t()
    check(lib.o3d_device_info(name, 256, C.byref(cus), C.byref(clk), C.byref(hbm)))
    return {"name": name.value.decode(), "cus": cus.value, "clock_khz": clk.value, "hbm_bytes": hbm.value}


# ---- device memory (plumbing) ---------------------------------------------------

class DeviceArray:
    """A device buffer holding a numpy array's bytes (shape and dtype kept host-side)."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib.o3d_dev_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(lib.o3d_memcpy_htod(d.ptr, a.ctypes.data, d.nbytes))
        return d

    def to_host(self):
        a = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib.o3d_memcpy_dtoh(a.ctypes.data, self.ptr, self.nbytes))
        return a

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib.o3d_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes))

    def zero(self):
        check(lib.o3d_memset_dev(self.ptr, 0, self.nbytes))

    def row(self, k):
        """device address of row k of a 2-D array"""
        return self.ptr + k * self.shape[1] * self.dtype.itemsize

    def free(self):
        if self.ptr:
            lib.o3d_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib.o3d_event_create(C.byref(p)))
        self.ptr = p.value

    def record(self, stream=None):
        check(lib.o3d_event_record(self.ptr, stream))

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(lib.o3d_event_elapsed_ms(C.byref(ms), self.ptr, end.ptr))
        return ms.value

    def __del__(self):
        try:
            lib.o3d_event_destroy(self.ptr)
        except Exception:
            pass


def synchronize():
    check(lib.o3d_device_synchronize())
