"""The oracle's panel method with a 7-point rule per panel and a closed-form near
field (oracle/o3d_panel_quad.cpp), pinned by mathematics (the reference has no such
method, so there is nothing of its to compare with):

  * the rule integrates every monomial of degree <= 5 on a triangle exactly, and
    not all of degree 6;
  * the closed-form I = int_T (x - y)/|x - y|^3 dA agrees with subdivided quadrature
    (Richardson-extrapolated), and dI with central differences of I; dI is
    symmetric and trace-free (I is the gradient of a harmonic potential);
  * I jumps by 4 pi n across the sheet, its tangential part is continuous;
  * far away a panel acts as a point of strength area at its centroid;
  * the whole method (near/far split, strength conventions, source term,
    gradient assembly) agrees with a brute-force integral of the sheet over a
    320-panel sphere, whose gradient is taken by central differences.
"""
import math

import numpy as np
import pytest

import oracle
from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors


def test_rule_exact_to_degree_5():
    bary, w = oracle.g7_rule()
    assert abs(w.sum() - 1.0) < 1e-15 and np.allclose(bary.sum(1), 1.0, atol=1e-15)
    x, y = bary[:, 1], bary[:, 2]       # reference triangle (0,0), (1,0), (0,1), area 1/2
    worst6 = 0.0
    for deg in range(7):
        for a in range(deg + 1):
            b = deg - a
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            got = 0.5 * (w * x ** a * y ** b).sum()
            if deg <= 5:
                assert abs(got - exact) <= 1e-15 * max(1.0, 1 / exact) * exact + 1e-17, (a, b)
            else:
                worst6 = max(worst6, abs(got - exact) / exact)
    assert worst6 > 1e-4      # degree 5, not 6


def _subdivide(tris, levels):
    for _ in range(levels):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
        tris = np.concatenate([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1),
                               np.stack([ca, bc, c], 1), np.stack([ab, bc, ca], 1)])
    return tris


def _quad_I(v, x, levels):
    """int_T (x - y)/|x - y|^3 dA by the centroid rule on 4^levels sub-triangles,
    Richardson-extrapolated over the last two levels (error O(h^2) -> O(h^4))"""
    out = []
    for L in (levels - 1, levels):
        t = _subdivide(v[None], L)
        c = t.mean(1)
        ar = 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)
        d = x - c
        out.append((ar[:, None] * d / np.linalg.norm(d, axis=1)[:, None] ** 3).sum(0))
    return out[1] + (out[1] - out[0]) / 3


TRIS = [np.array([[0, 0, 0], [1, 0.1, 0], [0.3, 0.9, 0.05]]),
        np.array([[0.2, -0.3, 0.4], [-0.5, 0.6, 0.1], [0.7, 0.4, -0.2]]),
        np.array([[1, 1, 1], [1.05, 1.0, 1.02], [1.01, 1.04, 0.99]])]


def _probe_points(v):
    c = v.mean(0)
    n = np.cross(v[1] - v[0], v[2] - v[0])
    n /= np.linalg.norm(n)
    size = max(np.linalg.norm(v[1] - v[0]), np.linalg.norm(v[2] - v[1]), np.linalg.norm(v[0] - v[2]))
    e = (v[1] - v[0]) / np.linalg.norm(v[1] - v[0])
    return size, [c + 0.5 * size * n, c - 0.4 * size * n,            # above / below the interior
                  v[0] + 0.3 * size * n - 0.2 * size * e,          # near a vertex, off the plane
                  0.5 * (v[1] + v[2]) + 0.35 * size * n,           # above an edge
                  c + 1.5 * size * e + 0.1 * size * n,             # beside, almost in plane
                  c + 3.0 * size * (n + e)]                        # farther away


@pytest.mark.parametrize("k", range(len(TRIS)))
def test_tri_integrals_against_quadrature(k):
    v = TRIS[k]
    size, pts = _probe_points(v)
    for x in pts:
        I, _ = oracle.tri_integrals(v, x)
        q = _quad_I(v, x, 7)
        assert np.abs(I - q).max() <= 2e-7 * np.abs(q).max(), (x, I, q)


@pytest.mark.parametrize("k", range(len(TRIS)))
def test_tri_gradient_symmetric_tracefree_and_fd(k):
    v = TRIS[k]
    size, pts = _probe_points(v)
    h = 1e-5 * size
    for x in pts:
        I, dI = oracle.tri_integrals(v, x)
        fd = np.array([(oracle.tri_integrals(v, x + h * e)[0] - oracle.tri_integrals(v, x - h * e)[0]) / (2 * h)
                       for e in np.eye(3)])
        scale = np.abs(fd).max()
        assert np.abs(dI - fd).max() <= 1e-6 * scale, x
        assert np.abs(dI - dI.T).max() <= 1e-12 * scale
        assert abs(np.trace(dI)) <= 1e-12 * scale


def test_jump_across_the_sheet():
    v = TRIS[1]
    n = np.cross(v[1] - v[0], v[2] - v[0])
    n /= np.linalg.norm(n)
    p = 0.2 * v[0] + 0.5 * v[1] + 0.3 * v[2]        # inside the triangle
    for eps in (1e-6, 1e-8):
        up, _ = oracle.tri_integrals(v, p + eps * n, grads=False)
        dn, _ = oracle.tri_integrals(v, p - eps * n, grads=False)
        jump = up - dn
        assert abs(jump @ n - 4 * math.pi) < 1e-4
        assert np.linalg.norm(jump - (jump @ n) * n) < 1e-5      # tangential part continuous


def test_far_away_panel_is_a_point():
    v = TRIS[0]
    c = v.mean(0)
    area = 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
    x = c + np.array([300.0, -200.0, 500.0])
    d = x - c
    I, dI = oracle.tri_integrals(v, x)
    pt = area * d / np.linalg.norm(d) ** 3
    assert np.abs(I - pt).max() <= 1e-6 * np.abs(pt).max()


def test_near_decision_at_twice_the_longest_edge():
    v = np.array([[0, 0, 0], [1, 0, 0], [0.5, math.sqrt(3) / 2, 0]], np.float32)   # longest edge 1
    c = v.astype(np.float64).mean(0)
    assert oracle.pan_g7_near(v, c + [0, 0, 1.999])
    assert not oracle.pan_g7_near(v, c + [0, 0, 2.001])
    assert not oracle.pan_g7_near(v, c + [2.5, 0.0, 0.0])


def _brute_sheet_vel(verts, ts, area, ss, x, levels=5):
    """velocity of the panel sheets at points x (n, 3): densities ts/area and ss on
    each flat triangle, centroid rule on 4^levels sub-triangles, Richardson over the
    last two levels; FP64"""
    out = []
    for L in (levels - 1, levels):
        t = _subdivide(verts, L)                                    # (P * 4^L, 3, 3), panel-major blocks
        P = verts.shape[0]
        m = t.shape[0] // P
        c = t.mean(1)
        ar = 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)
        owner = np.tile(np.arange(P), m)                            # _subdivide stacks children by block
        gam = (ts / area[:, None])[owner] * ar[:, None]
        sig = ss[owner] * ar
        u = np.zeros((x.shape[0], 3))
        for i, xi in enumerate(x):
            d = xi - c
            r3 = np.linalg.norm(d, axis=1) ** 3
            u[i] = ((np.cross(gam, d) + sig[:, None] * d) / r3[:, None]).sum(0)
        out.append(u)
    return out[1] + (out[1] - out[0]) / 3


def test_whole_method_against_brute_force_sheet():
    pan = scenes.icosphere_panels(levels=2)
    P = pan.area.shape[0]
    rng = np.random.default_rng(7)
    ss = rng.normal(size=P).astype(np.float32) * 0.3
    verts = pan.nodes.T.astype(np.float64)[pan.idx.reshape(P, 3)]   # (P, 3, 3)
    dirs = rng.normal(size=(40, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    r = np.concatenate([np.full(20, 1.12), rng.uniform(1.2, 1.5, 20)])
    x = (dirs * r[:, None]).astype(np.float32)
    vel, grad, nnear = oracle.pan_g7(pan.nodes, pan.idx, pan.ts, pan.area, ss, x.T.copy())
    assert nnear > 0                                                # both branches ran
    ref = _brute_sheet_vel(verts, pan.ts.T.astype(np.float64), pan.area.astype(np.float64), ss.astype(np.float64),
                           x.astype(np.float64))
    assert rel_errors(vel, ref.T)[0] <= 1e-5          # measured 7.4e-7: the rule's own error beyond 2 D
    h = 1e-4
    fd = np.zeros((9, x.shape[0]))
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        up = _brute_sheet_vel(verts, pan.ts.T.astype(np.float64), pan.area.astype(np.float64),
                              ss.astype(np.float64), x.astype(np.float64) + e)
        dn = _brute_sheet_vel(verts, pan.ts.T.astype(np.float64), pan.area.astype(np.float64),
                              ss.astype(np.float64), x.astype(np.float64) - e)
        fd[3 * j:3 * j + 3] = ((up - dn) / (2 * h)).T
    assert rel_errors(grad, fd)[0] <= 2e-5            # measured 2.6e-6
    # off the sheet the field is divergence-free only without sources; with them
    # the trace is still zero (each point source's field is harmonic there)
    tr = grad[0] + grad[4] + grad[8]
    assert np.abs(tr).max() <= 1e-9 * np.abs(grad).max()


def test_whole_method_detects_a_wrong_convention():
    """the comparison above is sharp: the sheet density taken as ts instead of
    ts/area changes the field by O(1)"""
    pan = scenes.icosphere_panels(levels=1)
    P = pan.area.shape[0]
    x = np.array([[0.0, 0.0, 1.3], [0.8, 0.6, 0.3]], np.float32)
    verts = pan.nodes.T.astype(np.float64)[pan.idx.reshape(P, 3)]
    vel, _, _ = oracle.pan_g7(pan.nodes, pan.idx, pan.ts, pan.area, None, x.T.copy(), grads=False)
    wrong = _brute_sheet_vel(verts, pan.ts.T.astype(np.float64) * pan.area[:, None], pan.area.astype(np.float64),
                             np.zeros(P), x.astype(np.float64))
    assert rel_errors(vel, wrong.T)[0] > 0.1
