"""GPU parity of the BEM right-hand side (o3d_bem_rhs_dev, o3d_bem_rhs_from_vels_dev)
against the oracle (oracle/o3d_bem.cpp after oracle.ptpan):

  * from given panel velocities the per-panel stage (finalize, body motion, bases,
    projection) is the same float/double sequence as the oracle's: bitwise equal,
    for every unknown count, with and without a body;
  * the whole chain from two particle collections meets points_affect_panels' own
    bar against the oracle (compensated: the float-accumulation band of
    test_ptpan_gpu.py, here 1e-6 normwise on the velocities and the rhs).
"""
import ctypes as C

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


BODY = {"vel": (0.1, -0.3, 0.05), "rotvel": (0.4, 0.2, -0.6), "tc": (0.01, -0.02, 0.03)}


@pytest.mark.parametrize("nunk", [1, 2, 3])
@pytest.mark.parametrize("body", [None, BODY])
def test_rhs_from_velocities_bitwise(mods, nunk, body):
    _lib, infl = mods
    pan = scenes.icosphere_panels(levels=3)
    P = pan.area.shape[0]
    raw = np.random.default_rng(nunk).normal(size=(3, P)).astype(np.float32) * 5
    fs = (0.7, -0.1, 0.25)
    D = _lib.DeviceArray.from_host
    dn, di = D(pan.nodes), D(pan.idx.view(np.float32))
    p = _lib.Panels()
    p.np, p.nn = P, pan.nodes.shape[1]
    for d in range(3):
        p.x[d] = dn.row(d)
    p.idx = di.ptr
    dpu, drhs = D(raw), _lib.DeviceArray((nunk * P,))
    pu = (C.c_void_p * 3)(*[dpu.row(d) for d in range(3)])
    b = infl.bem_opts(fs, body, nunk)
    _lib.check(_lib.lib.o3d_bem_rhs_from_vels_dev(C.byref(p), C.byref(b), pu, C.c_void_p(drhs.ptr), None))
    _lib.synchronize()
    kw = {} if body is None else dict(body=True, **body)
    opu, orhs = oracle.bem_rhs(pan.nodes, pan.idx, raw, fs, nunk=nunk, **kw)
    np.testing.assert_array_equal(dpu.to_host(), opu)
    np.testing.assert_array_equal(drhs.to_host(), orhs)


def test_rhs_from_two_particle_collections(mods):
    _lib, infl = mods
    pan = scenes.icosphere_panels(levels=3, radius=0.5)
    a = scenes.ring(4000, R=0.9, sigma=0.03, seed=3)
    b = scenes.ring(2500, R=1.2, sigma=0.03, seed=4)
    b.pos[2] += 0.3
    fs = (0.0, 0.0, 1.0)
    pu, rhs = infl.bem_rhs([(a.pos, a.str), (b.pos, b.str)], pan.nodes, pan.idx, pan.area, fs, BODY, nunk=3)
    raw = oracle.ptpan("rm", a.pos, a.str, pan.nodes, pan.idx, pan.area, mode="f64")
    raw = raw + oracle.ptpan("rm", b.pos, b.str, pan.nodes, pan.idx, pan.area, mode="f64")
    opu, orhs = oracle.bem_rhs(pan.nodes, pan.idx, raw.astype(np.float32), fs, nunk=3, body=True, **BODY)
    ev = rel_errors(pu, opu)
    er = rel_errors(rhs.reshape(-1, 3).T, orhs.reshape(-1, 3).T)
    print("bem rhs vs oracle: panel velocities", ev, "rhs", er)
    assert ev[0] <= 1e-6 and er[0] <= 1e-6


def test_rhs_rejections(mods):
    _lib, infl = mods
    pan = scenes.icosphere_panels(levels=1)
    a = scenes.ring(100)
    with pytest.raises(_lib.O3DError):
        infl.bem_rhs([(a.pos, a.str)], pan.nodes, pan.idx, pan.area, (0, 0, 1), nunk=4)
    with pytest.raises(_lib.O3DError):
        infl.bem_rhs([(a.pos, a.str)], pan.nodes, pan.idx, pan.area, (0, 0, 1), precision="fp64")
