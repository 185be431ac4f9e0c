"""The panel root kernels' separated-pair loops at their edges, against the oracle.

k_pan (panels -> points) and k_ptpan (particles -> panels) take the pairs of a
staged tile four at a time while every lane's pairs are surely separated at the
root, and fall back to the per-pair decision otherwise; the tail of a tile (its
count modulo 4) always takes the per-pair path.  Here the counts are chosen so
that every case occurs: 317 panels (317 = 4 * 79 + 1, so the last tile ends in a
tail) with the default tile and a tile of 64, 1,000 targets (not a multiple of 64, so the last wave has idle
lanes) and 1,001 particles, half of them far from the sphere (waves whose pairs
are all separated) and half within reach of the panels (waves that take the
per-pair path).  Bars as for the other panel tests: within 1e-5 of the
reference's float path restated by the oracle, and within max(2e-6, 3 x that
path's own distance from FP64) of the FP64 oracle.
"""
import numpy as np
import pytest

import oracle
from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu


def _band(ref32, ref64):
    return max(2e-6, 3.0 * rel_errors(ref32, ref64)[0])


@pytest.fixture(scope="module")
def setup():
    from omega3d_amd import _lib, influence
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    pan = scenes.icosphere_panels(levels=2)
    keep = 317
    idx = np.ascontiguousarray(pan.idx.reshape(-1, 3)[:keep]).reshape(-1)
    area, ts = pan.area[:keep].copy(), np.ascontiguousarray(pan.ts[:, :keep])
    ss = None if pan.ss is None else pan.ss[:keep].copy()
    rng = np.random.default_rng(97)

    def cloud(n):
        d = rng.normal(size=(3, n))
        r = np.where(np.arange(n) % 128 < 64, rng.uniform(2.0, 3.0, n), rng.uniform(1.02, 1.15, n))
        return (d / np.linalg.norm(d, axis=0) * r).astype(np.float32)

    return influence, pan.nodes, idx, ts, area, ss, cloud(1000), cloud(1001), rng


@pytest.mark.parametrize("tile", [0, 64])
def test_panels_to_points_tails(setup, tile):
    infl, nodes, idx, ts, area, ss, tp, _, _ = setup
    r64 = oracle.pan("rm", nodes, idx, ts, area, ss, tp, True, "f64")
    r32 = oracle.pan("rm", nodes, idx, ts, area, ss, tp, True, "f32d")
    for prec in ("fp32", "comp"):
        v, g = infl.pan("rm", nodes, idx, ts, area, ss, tp, True, prec, tile=tile)
        print(f"tile {tile} {prec}: vel {rel_errors(v, r32[0])} grad {rel_errors(g, r32[1])}", flush=True)
        assert rel_errors(v, r32[0])[0] <= 1e-5 and rel_errors(g, r32[1])[0] <= 1e-5
        assert rel_errors(v, r64[0])[0] <= _band(r32[0], r64[0])
        assert rel_errors(g, r64[1])[0] <= _band(r32[1], r64[1])


def test_particles_to_panels_tails(setup):
    infl, nodes, idx, _, area, _, _, sp, rng = setup
    st = (rng.normal(size=(3, sp.shape[1])) / sp.shape[1]).astype(np.float32)
    r64 = oracle.ptpan("rm", sp, st, nodes, idx, area, "f64")
    r32 = oracle.ptpan("rm", sp, st, nodes, idx, area, "f32d")
    for prec in ("fp32", "comp"):
        v = infl.ptpan(sp, st, nodes, idx, area, prec)
        print(f"ptpan {prec}: {rel_errors(v, r32)}", flush=True)
        assert rel_errors(v, r64)[0] <= _band(r32, r64)
