"""Writes tests/golden/bem_matrix.npz: BEM influence-matrix blocks computed by the
reference's own panel-panel recursion (rkernel_2vs_2p, Kernels.h:1213-1311, built
in place by oracle/build_ref.sh into oracle/_ref/libo3dref_<core>.so) through the
loop of panels_on_panels_coeff (Coefficients.h:332-433) in oracle/ref_harness.cpp,
with the surfaces beside them.  The bases and areas are Surfaces::compute_bases
(Surfaces.h:766-816) as the oracle restates it: Surfaces.h needs Eigen3 and is not
built.  Run from the repository root in the container that holds the reference
checkout:

    python tests/golden/make_bem_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402
from omega3d_amd import scenes  # noqa: E402


def surface(levels, radius, centre):
    v, f = scenes.icosphere(levels, radius)
    v = v + np.asarray(centre, float)[None, :]
    return oracle.surface(np.ascontiguousarray(v.T, np.float32), f.astype(np.uint32))


SURF = {"A": surface(1, 1.0, (0, 0, 0)), "B": surface(0, 0.3, (1.6, 0.2, 0.1))}
# (core, src, targ, nunk, same, maxlev)
CASES = [("wl", "A", "A", 2, 1, 3), ("wl", "B", "A", 3, 0, 3), ("wl", "A", "B", 3, 0, 3), ("wl", "B", "B", 2, 1, 3),
         ("rm", "B", "A", 2, 0, 1), ("exp", "A", "B", 2, 0, 2)]

out = {}
for name, (nodes, idx, area, b) in SURF.items():
    out.update({f"{name}_nodes": nodes, f"{name}_idx": idx, f"{name}_area": area, f"{name}_bases": b})
for n, (core, s, t, nunk, same, maxlev) in enumerate(CASES):
    m = oracle.ref_bem_matrix(core, SURF[s], SURF[t], nunk, bool(same), maxlev, mode="f32")
    out[f"case{n}"] = np.array([("rm", "wl", "exp", "v2").index(core) + 1, ord(s), ord(t), nunk, same, maxlev],
                               np.int32)
    out[f"case{n}_matrix"] = m
np.savez_compressed(os.path.join(HERE, "bem_matrix.npz"), **out)
print("wrote bem_matrix.npz with", len(CASES), "blocks")
