"""Writes tests/golden/vtk_*.vtu with the reference's own VTK XML writer
(oracle/_ref/libo3dref_vtk.so, built by oracle/build_ref.sh from VtkXmlWriter.h and
the reference's vendored tinyxml2/cppcodec, driven through the call sequences of
Points::write_vtk and Surfaces::write_vtk) and the inputs beside them in
vtk_inputs.npz.  Run from the repository
root in the container that holds the reference checkout:

    python tests/golden/make_vtk_golden.py
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libo3dref_vtk.so"))
L.ref_write_points_vtk.argtypes = [C.c_char_p, C.c_int64] + [C.c_void_p] * 10 + [C.c_double, C.c_int]
L.ref_write_points_vtk.restype = C.c_int
L.ref_write_panels_vtk.argtypes = [C.c_char_p, C.c_int64] + [C.c_void_p] * 3 + [C.c_int64] + [C.c_void_p] * 5 + [
    C.c_double]
L.ref_write_panels_vtk.restype = C.c_int


def write(fn, pos, st, rad, vel, time, inert):
    arrs = [np.ascontiguousarray(a, np.float32) for a in (pos[0], pos[1], pos[2], st[0], st[1], st[2], rad,
                                                         vel[0], vel[1], vel[2])]
    rc = L.ref_write_points_vtk(os.path.join(HERE, fn).encode(), pos.shape[1],
                                *[a.ctypes.data_as(C.c_void_p) for a in arrs], time, int(inert))
    assert rc == 0


rng = np.random.default_rng(20261202)
cases = {}
for name, n, inert, time in (("vtk_part", 7, 0, 0.375), ("vtk_fldpt", 5, 1, 12.0), ("vtk_one", 1, 0, 0.0)):
    pos = rng.normal(size=(3, n)).astype(np.float32)
    st = (1e-3 * rng.normal(size=(3, n))).astype(np.float32)
    rad = rng.uniform(0.01, 0.05, n).astype(np.float32)
    vel = rng.normal(size=(3, n)).astype(np.float32)
    write(name + ".vtu", pos, st, rad, vel, time, inert)
    cases.update({f"{name}_pos": pos, f"{name}_str": st, f"{name}_rad": rad, f"{name}_vel": vel,
                  f"{name}_time": np.float64(time), f"{name}_inert": np.int32(inert)})
# panels: the 12 triangles of a unit cube; an active surface with source strengths,
# one without, and an inert one
nodes = np.array([((i >> 2) & 1, (i >> 1) & 1, i & 1) for i in range(8)], np.float32).T.copy()
idx = np.array([0, 4, 5, 0, 5, 1, 4, 6, 7, 4, 7, 5, 1, 5, 7, 1, 7, 3, 6, 2, 3, 6, 3, 7, 2, 0, 1, 2, 1, 3,
                2, 6, 4, 2, 4, 0], np.uint32)
for name, vort, src, time in (("vtk_pan_src", True, True, 1.25), ("vtk_pan", True, False, 0.0),
                              ("vtk_pan_inert", False, False, 3.5)):
    vs = rng.normal(size=(3, 12)).astype(np.float32)
    ss = rng.normal(size=12).astype(np.float32)
    keep = [np.ascontiguousarray(a) for a in (nodes[0], nodes[1], nodes[2], idx, vs[0], vs[1], vs[2], ss)]
    ptr = [a.ctypes.data_as(C.c_void_p) for a in keep]
    rc = L.ref_write_panels_vtk(os.path.join(HERE, name + ".vtu").encode(), 8, ptr[0], ptr[1], ptr[2], 12, ptr[3],
                                ptr[4] if vort else None, ptr[5] if vort else None, ptr[6] if vort else None,
                                ptr[7] if src else None, time)
    assert rc == 0
    cases.update({f"{name}_nodes": nodes, f"{name}_idx": idx, f"{name}_vs": vs, f"{name}_ss": ss,
                  f"{name}_time": np.float64(time), f"{name}_vort": np.int32(vort), f"{name}_src": np.int32(src)})
np.savez(os.path.join(HERE, "vtk_inputs.npz"), **cases)
print("wrote", sorted(f for f in os.listdir(HERE) if f.startswith("vtk_")))
