"""GPU parity of the panel -> point kernels (recursive centroid rule) against the
oracle and the reference's golden outputs, through the C ABI.

The tree decisions are computed in the reference's exact IEEE arithmetic, so the
GPU subdivides exactly where the reference does; a different decision would move
a target's result by the quadrature error (~1e-3), far above these bars:
  * against the reference's float outputs (golden / oracle 'f32d'): max rel <= 1e-5
  * FP32 / compensated against the FP64 oracle: within the reference float path's
    own distance from FP64 (which takes its tree decisions in double and has
    singular near-field leaves): <= max(2e-6, 3 x |f32d - f64|)
  * the FP64 panel kernel against the CPU FP64 oracle: max rel <= 1.2e-7 (the host
    call accumulates the FP64 result into the caller's float arrays: half-ulp floor)
"""
import os

import numpy as np
import pytest

import oracle
from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CORES = ["rm", "wl", "exp", "v2"]


def _band(ref32, ref64):
    return max(2e-6, 3.0 * rel_errors(ref32, ref64)[0])


@pytest.fixture(scope="module")
def infl():
    from omega3d_amd import _lib, influence
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return influence


@pytest.mark.parametrize("core", CORES)
def test_panels_vs_reference_golden(infl, core):
    g = np.load(os.path.join(GOLD, f"pan_{core}.npz"))
    for with_ss in (True, False):
        for grads in (True, False):
            key = f"{'vs' if with_ss else 'v'}_{'g' if grads else 'v'}"
            ss = g["ss"] if with_ss else None
            for prec in ("fp32", "comp"):
                v, gr = infl.pan(core, g["nodes"], g["idx"], g["ts"], g["area"], ss, g["tpos"], grads, prec)
                assert rel_errors(v, g["vel_f32d_" + key])[0] <= 1e-5, (key, prec)
                assert rel_errors(v, g["vel_f64_" + key])[0] <= _band(g["vel_f32d_" + key], g["vel_f64_" + key])
                if grads:
                    assert rel_errors(gr, g["grad_f32d_" + key])[0] <= 1e-5, (key, prec)
                    assert rel_errors(gr, g["grad_f64_" + key])[0] <= _band(g["grad_f32d_" + key],
                                                                            g["grad_f64_" + key])
            v, gr = infl.pan(core, g["nodes"], g["idx"], g["ts"], g["area"], ss, g["tpos"], grads, "fp64")
            assert rel_errors(v, g["vel_f64_" + key])[0] <= 1.2e-7
            if grads:
                assert rel_errors(gr, g["grad_f64_" + key])[0] <= 1.2e-7


@pytest.mark.parametrize("maxlev", [0, 1, 2, 3, 4])
def test_panels_maxlev(infl, maxlev):
    pan = scenes.icosphere_panels(levels=2)
    tp = scenes.shell_points(700, 1.0, 1.3, seed=5)
    v, g = infl.pan("rm", pan.nodes, pan.idx, pan.ts, pan.area, pan.ss, tp, True, "comp", maxlev=maxlev)
    ov, og = oracle.pan("rm", pan.nodes, pan.idx, pan.ts, pan.area, pan.ss, tp, True, "f64", maxlev=maxlev)
    rv, rg = oracle.pan("rm", pan.nodes, pan.idx, pan.ts, pan.area, pan.ss, tp, True, "f32d", maxlev=maxlev)
    assert rel_errors(v, rv)[0] <= 1e-5 and rel_errors(g, rg)[0] <= 1e-5
    assert rel_errors(v, ov)[0] <= _band(rv, ov) and rel_errors(g, og)[0] <= _band(rg, og)


def test_icosphere_full_size_subset(infl):
    """Unit icosphere of 20,480 panels with Gaussian tangential sheet strength, targets
    uniform in 1.02 < r < 1.5: a 600-target subset (many near the surface, so every
    recursion depth occurs) against the oracle in FP64 and in the reference's float
    semantics."""
    pan = scenes.icosphere_panels()
    assert pan.np == 20480
    tp = scenes.shell_points(600, 1.02, 1.06, seed=7)
    v, g = infl.pan("wl", pan.nodes, pan.idx, pan.ts, pan.area, pan.ss, tp, True, "comp")
    ov, og = oracle.pan("wl", pan.nodes, pan.idx, pan.ts, pan.area, pan.ss, tp, True, "f64")
    rv, rg = oracle.pan("wl", pan.nodes, pan.idx, pan.ts, pan.area, pan.ss, tp, True, "f32d")
    print("icosphere comp vs fp64", rel_errors(v, ov), rel_errors(g, og), "reference float vs fp64",
          rel_errors(rv, ov), rel_errors(rg, og))
    assert rel_errors(v, ov)[0] <= _band(rv, ov) and rel_errors(g, og)[0] <= _band(rg, og)
    assert rel_errors(v, rv)[0] <= 1e-5 and rel_errors(g, rg)[0] <= 1e-5


def test_panels_edge_cases(infl):
    pan = scenes.icosphere_panels(levels=1)
    tp = scenes.shell_points(33, 1.0, 1.2, seed=3)
    v, g = infl.pan("rm", pan.nodes, pan.idx, pan.ts, pan.area, pan.ss, tp[:, :0], True, "comp")
    assert v.shape == (3, 0)
    one = pan.idx[:3]
    v, g = infl.pan("rm", pan.nodes, one, pan.ts[:, :1], pan.area[:1], None, tp, True, "comp")
    ov, og = oracle.pan("rm", pan.nodes, one, pan.ts[:, :1], pan.area[:1], None, tp, True, "f64")
    rv, rg = oracle.pan("rm", pan.nodes, one, pan.ts[:, :1], pan.area[:1], None, tp, True, "f32d")
    assert rel_errors(v, rv)[0] <= 1e-5 and rel_errors(g, rg)[0] <= 1e-5
    assert rel_errors(v, ov)[0] <= _band(rv, ov) and rel_errors(g, og)[0] <= _band(rg, og)
    # velocity only, with a source strength
    ss = np.linspace(-1, 1, pan.np).astype(np.float32)
    v, _ = infl.pan("exp", pan.nodes, pan.idx, pan.ts, pan.area, ss, tp, False, "fp32")
    ov, _ = oracle.pan("exp", pan.nodes, pan.idx, pan.ts, pan.area, ss, tp, False, "f64")
    rv, _ = oracle.pan("exp", pan.nodes, pan.idx, pan.ts, pan.area, ss, tp, False, "f32d")
    assert rel_errors(v, rv)[0] <= 1e-5 and rel_errors(v, ov)[0] <= _band(rv, ov)
