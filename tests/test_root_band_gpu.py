"""The root decision of the recursive panel rule at its threshold.

The reference takes a pair as one leaf when its float distance from the panel
centroid exceeds 4 sqrt(area) (Kernels.h:1140-1160).  The GPU decides most pairs
from a fused squared distance against bands 1e-6 either side of that threshold
and takes the reference's own float expression inside the bands.  Targets (and,
for particles -> panels, particles) placed within a few ulps of the threshold in
eight directions exercise both sides and the exact path; a decision different
from the reference's moves that target's result by the subdivision's correction,
which the test first shows to be far above the 1e-5 bar.
"""
import numpy as np
import pytest

import oracle
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu

F = np.float32


def _panel():
    nodes = np.array([[0.1, 1.3, 0.4], [0.2, 0.1, 1.1], [0.05, -0.2, 0.3]], F)   # [coord, node]
    idx = np.array([[0, 1, 2]], np.uint32)
    e1 = nodes[:, 1].astype(np.float64) - nodes[:, 0]
    e2 = nodes[:, 2].astype(np.float64) - nodes[:, 0]
    area = np.array([0.5 * np.linalg.norm(np.cross(e1, e2))], F)
    # the reference's float centroid and threshold
    c = ((nodes[:, 0] + nodes[:, 1]) + nodes[:, 2]) / F(3)
    lim = np.sqrt(area[0]) * F(4)
    return nodes, idx, area, c, lim


def _hugging_points(c, lim, ndir=8, kmax=48):
    rng = np.random.default_rng(11)
    u = rng.normal(size=(3, ndir))
    u /= np.linalg.norm(u, axis=0)
    k = np.arange(-kmax, kmax + 1)
    r = lim.astype(np.float64) * (1.0 + k * 2.0 ** -24)
    p = c.astype(np.float64)[:, None, None] + u[:, :, None] * r[None, None, :]
    p = p.reshape(3, -1).astype(F)
    # the reference's float distance and decision for each point
    e = p - c[:, None]
    d = np.sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2])
    return np.ascontiguousarray(p), d > lim, np.abs(d.astype(np.float64) / lim - 1.0) < 1e-6


@pytest.fixture(scope="module")
def infl():
    from omega3d_amd import _lib, influence
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return influence


def test_panels_on_points_at_the_threshold(infl):
    nodes, idx, area, c, lim = _panel()
    tp, leaf, inband = _hugging_points(c, lim)
    assert leaf.any() and (~leaf).any() and inband.sum() >= 64
    ts = np.array([[0.3], [-0.7], [0.5]], F)
    ss = np.array([0.4], F)
    rv, rg = oracle.pan("rm", nodes, idx, ts, area, ss, tp, True, "f32d")
    # sensitivity: one leaf everywhere differs from the recursion by far more than the bar
    lv, lg = oracle.pan("rm", nodes, idx, ts, area, ss, tp, True, "f32d", maxlev=0)
    d = np.linalg.norm(lv - rv, axis=0) / np.linalg.norm(rv, axis=0).max()
    assert d[~leaf].min() > 1e-4
    for prec in ("fp32", "comp"):
        v, g = infl.pan("rm", nodes, idx, ts, area, ss, tp, True, prec)
        assert rel_errors(v, rv)[0] <= 1e-5, prec
        assert rel_errors(g, rg)[0] <= 1e-5, prec


def test_particles_on_panels_at_the_threshold(infl):
    nodes, idx, area, c, lim = _panel()
    sp, leaf, inband = _hugging_points(c, lim, ndir=4, kmax=20)
    assert leaf.any() and (~leaf).any() and inband.sum() >= 16
    a = np.array([[0.3], [-0.7], [0.5]], F)
    for j in range(sp.shape[1]):
        p = np.ascontiguousarray(sp[:, j:j + 1])
        want = oracle.ptpan("rm", p, a, nodes, idx, area, "f32d")
        if not leaf[j]:
            one = oracle.ptpan("rm", p, a, nodes, idx, area, "f32d", maxlev=0)
            assert rel_errors(one, want)[0] > 1e-4
        got = infl.ptpan(p, a, nodes, idx, area, "comp")
        assert rel_errors(got, want)[0] <= 1e-5, (j, bool(leaf[j]), bool(inband[j]))
