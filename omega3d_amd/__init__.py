"""omega3d_amd -- Omega3D's direct Biot-Savart velocity / velocity-gradient path on
AMD Instinct MI355X (gfx950).

The compute lives in libo3dbiot.so (hand-written HIP kernels behind the C ABI of
include/o3d_biot.h).  This package is the Python side: ctypes bindings (_lib),
a mirror of the reference's influence interface (influence), seeded synthetic
workloads (scenes) and the build script (build).
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libo3dbiot.so")

__all__ = ["LIBPATH", "load"]


def load():
    """Import the ctypes binding (raises ImportError if the HIP library is not built)."""
    from . import _lib
    return _lib
