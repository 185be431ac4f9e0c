"""An example scene of the reference (two colliding singular rings) through the
library: velocities of its particles against the oracle (WL core, the reference's
compiled default), then device-resident RK2 steps by the server, written as the
reference's VTU files and read back."""
import os

import numpy as np
import pytest

import oracle
from omega3d_amd import scene
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def mods():
    from omega3d_amd import _lib, influence, serve
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return influence, serve


def test_example_scene_velocities_and_steps(mods, tmp_path):
    infl, serve = mods
    s = scene.Scene(os.path.join(GOLD, "scenes", "colliding_vortex_rings_nv.json"))
    v, g = infl.p2p("wl", s.pos, s.str, s.rad, s.pos, s.rad, True, "comp")
    ov, og = oracle.p2p("wl", s.pos, s.str, s.rad, s.pos, s.rad, True, "f64")
    # 316 particles 0.02 apart with cores of 0.03: the compensated per-pair FP32 floor
    # of small rings (2e-6, as in test_batch_gpu.py); measured 8.3e-7 and 1.4e-6
    assert rel_errors(v, ov)[0] <= 2e-6 and rel_errors(g, og)[0] <= 2e-6
    srv = serve.Server(hbm_budget=64 << 20)
    try:
        sid = srv.open_session(s, "wl")
        for k in range(3):
            assert srv.wait(srv.submit(sid, s.dt, "comp", s.fs))["state"] == "done"
            x, st, el = srv.read_session(sid)
            fn = scene.write_points_vtu(str(tmp_path / scene.vtu_name(0, k + 1)), x, (k + 1) * s.dt, st, s.rad)
            d = scene.read_points_vtu(fn)
            np.testing.assert_array_equal(d["position"], x)
            np.testing.assert_array_equal(d["circulation"], st)
    finally:
        srv.close()
    # the rings moved: a finite displacement, of the order of dt x the ring speed
    disp = np.linalg.norm(x - s.pos, axis=0)
    assert 0 < disp.max() < 1.0


@pytest.mark.parametrize("name", ["flow_over_sphere.json", "flow_over_cube.json", "driven_cavity.json"])
def test_example_scene_bodies_through_the_panel_kernel(mods, name):
    """The meshed bodies of the reference's example scenes (icosphere, refined cube,
    cavity quads), given a random tangential sheet strength, acting on 400 points
    near their surfaces on the fluid side (every recursion depth of the panel rule
    occurs): the GPU against the oracle in FP64 and in the reference's float
    semantics, with the bars of the full-sphere panel test."""
    infl, _ = mods
    s = scene.Scene(os.path.join(GOLD, "scenes", name))
    pan, bc = s.surfaces()
    nodes = pan.nodes.T.astype(np.float64)
    faces = pan.idx.reshape(-1, 3).astype(np.int64)
    area, n = scene.panel_records(pan.nodes.T, faces)
    rng = np.random.default_rng(11)
    g = rng.standard_normal((pan.np, 3))
    gam = g - (g * n).sum(1)[:, None] * n
    pan.ts = np.ascontiguousarray((gam * area[:, None]).T, dtype=np.float32)
    # points off the centroids of random panels along their normals (into the fluid)
    pick = rng.integers(0, pan.np, 400)
    cen = nodes[faces[pick]].mean(axis=1)
    h = np.sqrt(area[pick]) * rng.uniform(0.05, 3.0, 400)
    tp = np.ascontiguousarray((cen + h[:, None] * n[pick]).T, dtype=np.float32)
    v, gr = infl.pan("wl", pan.nodes, pan.idx, pan.ts, pan.area, pan.ss, tp, True, "comp")
    ov, og = oracle.pan("wl", pan.nodes, pan.idx, pan.ts, pan.area, pan.ss, tp, True, "f64")
    rv, rg = oracle.pan("wl", pan.nodes, pan.idx, pan.ts, pan.area, pan.ss, tp, True, "f32d")
    ev, eg = rel_errors(v, ov), rel_errors(gr, og)
    print(f"{name}: {pan.np} panels, comp vs fp64 {ev} {eg}; reference float vs fp64 {rel_errors(rv, ov)}")
    band = lambda r32, r64: max(2e-6, 3.0 * rel_errors(r32, r64)[0])
    assert ev[0] <= band(rv, ov) and eg[0] <= band(rg, og)
    assert rel_errors(v, rv)[0] <= 1e-5 and rel_errors(gr, rg)[0] <= 1e-5


def test_example_sphere_scene_bem(mods):
    """The BEM pieces on the meshed body of flow_over_sphere.json (320 panels): the
    influence matrix (3 unknowns per panel) against the oracle in FP64 tree
    decisions and in the reference's float semantics, with the bars of the matrix
    tests, and the right-hand side of the scene's freestream (no particles yet)
    bitwise against the oracle."""
    import ctypes as C
    from omega3d_amd import _lib
    infl, _ = mods
    s = scene.Scene(os.path.join(GOLD, "scenes", "flow_over_sphere.json"))
    pan, bc = s.surfaces()
    assert pan.np == 320
    S = oracle.surface(pan.nodes, pan.idx)
    g = infl.bem_matrix(S[:3], None, 3, -1)
    m64 = oracle.bem_matrix("wl", S, S, 3, True, 3, mode="f64t")
    m32 = oracle.bem_matrix("wl", S, S, 3, True, 3, mode="f32")
    nerr = lambda a, b: float(np.abs(a.astype(np.float64) - b).max() / np.abs(b).max())
    print(f"sphere scene matrix: gpu vs f64t {nerr(g, m64):.3g}, reference float vs f64t {nerr(m32, m64):.3g}")
    assert nerr(g, m64) <= 1e-6 and nerr(g, m32) <= 3 * nerr(m32, m64) + 1e-7
    D = _lib.DeviceArray.from_host
    dn, di = D(pan.nodes), D(pan.idx.view(np.float32))
    p = _lib.Panels()
    p.np, p.nn = pan.np, pan.nodes.shape[1]
    for d in range(3):
        p.x[d] = dn.row(d)
    p.idx = di.ptr
    raw = bc.copy()                       # zero panel velocities: no particles in the scene
    dpu, drhs = D(raw), _lib.DeviceArray((3 * pan.np,))
    pu = (C.c_void_p * 3)(*[dpu.row(d) for d in range(3)])
    b = infl.bem_opts(tuple(s.fs), None, 3)
    _lib.check(_lib.lib.o3d_bem_rhs_from_vels_dev(C.byref(p), C.byref(b), pu, C.c_void_p(drhs.ptr), None))
    _lib.synchronize()
    opu, orhs = oracle.bem_rhs(pan.nodes, pan.idx, raw, tuple(s.fs), nunk=3)
    np.testing.assert_array_equal(dpu.to_host(), opu)
    np.testing.assert_array_equal(drhs.to_host(), orhs)
    assert np.abs(orhs).max() > 0.1                  # the freestream (1, 0, 0) projected on the panels


def test_example_scene_measurement_points(mods):
    """An example scene with a measurement quad and line added: the velocity and its
    gradient at the inert field points (point targets, kernel_0v_0pg) from the
    scene's particles, against the oracle."""
    import json
    infl, _ = mods
    doc = json.load(open(os.path.join(GOLD, "scenes", "colliding_vortex_rings_nv.json")))
    doc["measurements"] = [{"type": "measurement quad", "start": [-1.0, -1.0, 0.0], "axis1": [2.0, 0.0, 0.0],
                            "axis2": [0.0, 0.0, 2.0], "dx": [0.05, 0.05]},
                           {"type": "measurement line", "center": [0.0, 0.0, -1.0], "end": [0.0, 0.0, 1.0]}]
    s = scene.Scene(doc)
    assert s.fieldpts.shape[1] > 1600
    v, g = infl.p2p("wl", s.pos, s.str, s.rad, s.fieldpts, None, True, "comp")
    ov, og = oracle.p2p("wl", s.pos, s.str, s.rad, s.fieldpts, None, True, "f64")
    ev, eg = rel_errors(v, ov), rel_errors(g, og)
    print(f"{s.fieldpts.shape[1]} field points: comp vs fp64 {ev} {eg}")
    assert ev[0] <= 2e-6 and eg[0] <= 2e-6
