"""The C++ mirror of the reference interface (include/omega3d/Influence.hpp), driven
like Convection::find_vels and like solve_bem's right-hand-side loop
(tests/cpp/drive_influence.cpp), checked against the oracle: particles and field
points, vortex and panel sources, finalize_vels; particles -> panels and panels ->
panels through InfluenceVisitor."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "omega3d_amd", "_build", "drive_influence")


def test_cpp_influence_shim(tmp_path):
    assert os.path.exists(DRIVER), "build() compiles the C++ driver"
    p = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0
    rd = lambda name, shape: np.fromfile(tmp_path / f"{name}.f32", np.float32).reshape(shape)
    n, nf = 3000, 500
    x, s, r, xf = rd("x", (3, n)), rd("s", (3, n)), rd("r", (n,)), rd("xf", (3, nf))
    nodes = np.array([[-0.3, 0.3, 0.3, -0.3], [-0.2, -0.2, 0.2, 0.2], [0.6] * 4], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    area = np.array([0.12, 0.12], np.float32)
    ts = np.array([[0.012, -0.006], [0.003, 0.009], [0.0, 0.0]], np.float32)
    fs = np.array([[0.1], [0.0], [-0.05]])
    f = 0.25 / np.pi
    for tpos, trad, uname in ((x, r, "vort"), (xf, None, "fld")):
        v1, g1 = oracle.p2p("exp", x, s, r, tpos, trad, True, "f64")
        v2, g2 = oracle.pan("exp", nodes, idx, ts, area, None, tpos, True, "f64")
        want_v = fs + (v1 + v2) * f
        want_g = (g1 + g2) * f
        got_v = rd(uname + "_u", (3, tpos.shape[1]))
        got_g = rd(uname + "_ug", (9, tpos.shape[1]))
        assert rel_errors(got_v, want_v)[0] <= 1e-6, uname
        assert rel_errors(got_g, want_g)[0] <= 1e-6, uname
    # solve_bem's loop: zero, particles -> panels (subtracted), finalize
    pu = rd("bdry_pu", (3, 2))
    want = fs + oracle.ptpan("exp", x, s, nodes, idx, area, "f64") * f
    assert rel_errors(pu, want)[0] <= 1e-6
    # panels -> panels: the strip's velocity at the colocation points of the strip
    # 0.15 above it, overwriting (no finalize)
    nodes2 = nodes.copy()
    nodes2[2] += np.float32(0.15)
    pu2 = rd("bdry2_pu", (3, 2))
    r64 = oracle.pap("exp", nodes, idx, ts, area, None, nodes2, idx, "f64")
    r32 = oracle.pap("exp", nodes, idx, ts, area, None, nodes2, idx, "f32d")
    assert rel_errors(pu2, r32)[0] <= 1e-5 and rel_errors(pu2, r64)[0] <= max(2e-6, 3 * rel_errors(r32, r64)[0])
