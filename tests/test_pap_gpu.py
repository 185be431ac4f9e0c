"""panels_affect_panels (Influence.h:1225-1245) on the GPU: the source panels'
velocity at the target surface's colocation points (Surfaces::represent_as_particles
(0.0001, -1.0): centroids pushed 1e-4 along the normals), by the recursive panel
rule, velocity only, overwriting the target's velocities.

Checked against the oracle (its colocation restatement, oracle.colocation, then its
pinned panels_affect_points restatement) with the bars of the panel kernels:
<= 1e-5 against the reference's float semantics ('f32d') and within 3x the float
path's own distance from FP64; the overwrite semantics; a surface on itself (the
colocation points sit 1e-4 off their own panels, the deepest recursion); the
error paths (gradients requested, FP64)."""
import numpy as np
import pytest

import oracle
from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from omega3d_amd import _lib, influence
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return _lib, influence


def _band(r32, r64):
    return max(2e-6, 3.0 * rel_errors(r32, r64)[0])


@pytest.mark.parametrize("core", ["rm", "wl", "exp", "v2"])
def test_sphere_on_shifted_sphere(mods, core):
    _lib, infl = mods
    src = scenes.icosphere_panels(levels=3)
    tg = scenes.icosphere_panels(levels=2)
    tnodes = (tg.nodes * np.float32(1.25) + np.float32(0.05)).astype(np.float32)
    ss = np.linspace(-0.5, 0.5, src.np).astype(np.float32)
    for prec in ("fp32", "comp"):
        v = infl.pap(core, src.nodes, src.idx, src.ts, src.area, ss, tnodes, tg.idx, prec)
        r32 = oracle.pap(core, src.nodes, src.idx, src.ts, src.area, ss, tnodes, tg.idx, "f32d")
        r64 = oracle.pap(core, src.nodes, src.idx, src.ts, src.area, ss, tnodes, tg.idx, "f64")
        print(core, prec, rel_errors(v, r32), rel_errors(v, r64), flush=True)
        assert rel_errors(v, r32)[0] <= 1e-5 and rel_errors(v, r64)[0] <= _band(r32, r64)


def test_surface_on_itself_and_overwrite(mods):
    _lib, infl = mods
    s = scenes.icosphere_panels(levels=2)
    prior = np.full((3, s.np), 7.0, np.float32)
    v = infl.pap("rm", s.nodes, s.idx, s.ts, s.area, s.ss, s.nodes, s.idx, "comp", vel=prior)
    assert v is prior
    r32 = oracle.pap("rm", s.nodes, s.idx, s.ts, s.area, s.ss, s.nodes, s.idx, "f32d")
    r64 = oracle.pap("rm", s.nodes, s.idx, s.ts, s.area, s.ss, s.nodes, s.idx, "f64")
    print("self", rel_errors(v, r32), rel_errors(v, r64), flush=True)
    assert rel_errors(v, r32)[0] <= 1e-5 and rel_errors(v, r64)[0] <= _band(r32, r64)


def test_pap_errors_and_empty(mods):
    _lib, infl = mods
    s = scenes.icosphere_panels(levels=1)
    with pytest.raises(_lib.O3DError):
        infl.pap("rm", s.nodes, s.idx, s.ts, s.area, None, s.nodes, s.idx, "fp64")
    # no source panels: the target velocities are overwritten with zeros
    v = infl.pap("rm", s.nodes, s.idx[:0], s.ts[:, :0], s.area[:0], None, s.nodes, s.idx, "comp",
                 vel=np.ones((3, s.np), np.float32))
    assert not v.any()
