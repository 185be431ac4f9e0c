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
