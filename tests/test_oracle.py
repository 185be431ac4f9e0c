"""CPU tests that pin the oracle (oracle/o3d_oracle.cpp) before it is trusted.

Pins, strongest first:
  1. golden fixtures written from the reference's OWN kernels (Kernels.h /
     CoreFunc.h compiled per core, tests/golden/make_golden.py): bit-for-bit in
     every arithmetic mode, including the reference templates instantiated in
     double, which pins the FP64 oracle;
  2. the live reference build in oracle/_ref when present, on fresh inputs;
  3. mathematics: finite-difference Jacobian, zero divergence, closed-form self
     term and on-axis ring velocity, far-field limit, rotation/translation
     equivariance and linearity; for panels, the far-field point vortex and the
     convergence of the subdivision to the panel integral.
"""
import os

import numpy as np
import pytest

import oracle
from omega3d_amd import scenes

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CORES = ["rm", "wl", "exp", "v2"]
MODES = ["f32f", "f32d", "f64"]


def gold(name):
    return np.load(os.path.join(GOLD, name))


# ---- 1. golden fixtures -----------------------------------------------------------

@pytest.mark.parametrize("core", CORES)
def test_p2p_matches_reference_golden_bitwise(core):
    g = gold(f"p2p_{core}.npz")
    for mode in MODES:
        for thick in (True, False):
            for grads in (True, False):
                key = f"{mode}_{'tv' if thick else 'tp'}_{'g' if grads else 'v'}"
                v, gr = oracle.p2p(core, g["spos"], g["sstr"], g["srad"], g["tpos"],
                                   g["trad"] if thick else None, grads, mode)
                np.testing.assert_array_equal(v, g["vel_" + key], err_msg=key)
                if grads:
                    np.testing.assert_array_equal(gr, g["grad_" + key], err_msg=key)


@pytest.mark.parametrize("core", CORES)
def test_panels_match_reference_golden_bitwise(core):
    g = gold(f"pan_{core}.npz")
    for mode in MODES:
        for with_ss in (True, False):
            for grads in (True, False):
                key = f"{mode}_{'vs' if with_ss else 'v'}_{'g' if grads else 'v'}"
                v, gr = oracle.pan(core, g["nodes"], g["idx"], g["ts"], g["area"],
                                   g["ss"] if with_ss else None, g["tpos"], grads, mode)
                np.testing.assert_array_equal(v, g["vel_" + key], err_msg=key)
                if grads:
                    np.testing.assert_array_equal(gr, g["grad_" + key], err_msg=key)


def test_flop_counts_match_reference():
    f = gold("flops.npz")
    for core in CORES:
        # order: 0v_0bg, 0v_0pg, 0v_0b, 0v_0p, 0vs_0pg, 0vs_0p
        want = [oracle.flops_pair(core, True, True), oracle.flops_pair(core, True, False),
                oracle.flops_pair(core, False, True), oracle.flops_pair(core, False, False),
                oracle.flops_pair(core, True, False, True), oracle.flops_pair(core, False, False, True)]
        assert list(f[core]) == want
    # kernel_0v_0bg's count per interaction (Kernels.h:155 plus each core's flops, CoreFunc.h): RM 63, WL 70, EXP 68, V2 67
    assert [oracle.flops_pair(c) for c in CORES] == [63, 70, 68, 67]


# ---- 2. the live reference build ----------------------------------------------------

@pytest.mark.skipif(any(oracle.ref_lib(c) is None for c in CORES), reason="oracle/_ref not built")
@pytest.mark.parametrize("core", CORES)
def test_oracle_matches_live_reference(core):
    rng = np.random.default_rng(77)
    p = scenes.ring(600, sigma=0.03, seed=5)
    tp = np.concatenate([p.pos[:, :50], rng.uniform(-1.2, 1.2, (3, 50)).astype(np.float32)], 1)
    tr = rng.uniform(0.01, 0.05, 100).astype(np.float32)
    for mode in MODES:
        for thick in (True, False):
            a = oracle.p2p(core, p.pos, p.str, p.rad, tp, tr if thick else None, True, mode)
            b = oracle.ref_p2p(core, p.pos, p.str, p.rad, tp, tr if thick else None, True, mode)
            np.testing.assert_array_equal(a[0], b[0])
            np.testing.assert_array_equal(a[1], b[1])


# ---- 3. mathematics -----------------------------------------------------------------

def _kself(core, s):
    """K(0; s, s) for each core (CoreFunc.h formulas at rho = 0)"""
    return {"rm": (2 * s * s) ** -1.5, "wl": 2.5 * (2 * s * s) ** -1.5,
            "exp": 1.0 / (2 * s ** 3), "v2": (2 * s ** 4) ** -0.75}[core]


def _fd_check(core, sp, ss, sr, tp, tr, rtol, atol_frac):
    _, g = oracle.p2p(core, sp, ss, sr, tp, tr, True, "f64")
    h = 1e-5
    for j in range(3):
        e = np.zeros((3, 1)); e[j] = h
        up, _ = oracle.p2p(core, sp, ss, sr, tp + e, tr, False, "f64")
        um, _ = oracle.p2p(core, sp, ss, sr, tp - e, tr, False, "f64")
        fd = (up - um) / (2 * h)
        for i in range(3):
            np.testing.assert_allclose(g[3 * j + i], fd[i], rtol=rtol, atol=atol_frac * np.abs(g).max())


@pytest.mark.parametrize("core", ["rm", "wl", "exp"])
@pytest.mark.parametrize("thick", [True, False])
def test_gradient_is_jacobian_of_velocity(core, thick):
    """du_i/dx_j from the gradient output equals central differences of the velocity
    output (pins B against K and the index order ug[3j+i] = du_i/dx_j)."""
    rng = np.random.default_rng(3)
    ns = 40
    sp = rng.uniform(-1, 1, (3, ns))
    ss = rng.normal(size=(3, ns))
    sr = rng.uniform(0.2, 0.6, ns)
    tp = rng.uniform(-1, 1, (3, 5))
    tr = rng.uniform(0.2, 0.6, 5) if thick else None
    _fd_check(core, sp, ss, sr, tp, tr, 2e-6, 1e-6)


@pytest.mark.parametrize("thick", [True, False])
def test_v2_gradient_quirk(thick):
    """Vatistas n=2: the reference's gradient factor is B = -3 K D^-1/2
    (CoreFunc.h:327), not the exact derivative -3 K rho^2/D of K = D^-3/4
    (D = rho^4 + ss^4 + st^4).  The two agree when rho^4 >> ss^4 + st^4: the
    gradient is the velocity's Jacobian in the far field and not near the core."""
    rng = np.random.default_rng(3)
    ns = 30
    sp = rng.uniform(-1, 1, (3, ns))
    ss = rng.normal(size=(3, ns))
    sr = np.full(ns, 0.002)
    tp = rng.uniform(-1, 1, (3, 5))
    tr = np.full(5, 0.002) if thick else None
    _fd_check("v2", sp, ss, sr, tp, tr, 1e-5, 1e-6)
    # near the core the gradient is not the Jacobian
    sr = np.full(ns, 0.8)
    tr = np.full(5, 0.8) if thick else None
    with pytest.raises(AssertionError):
        _fd_check("v2", sp, ss, sr, tp, tr, 1e-3, 1e-4)


@pytest.mark.parametrize("core", CORES)
def test_divergence_free_and_self_term(core):
    rng = np.random.default_rng(4)
    p = scenes.ring(300, sigma=0.08, seed=9)
    v, g = oracle.p2p(core, p.pos, p.str, p.rad, p.pos, p.rad, True, "f64")
    tr = g[0] + g[4] + g[8]
    assert np.abs(tr).max() < 1e-12 * np.abs(g).max()
    # a lone particle on itself: zero velocity, gradient K(0) skew(alpha)
    s = 0.07
    a = rng.normal(size=3)
    v1, g1 = oracle.p2p(core, np.zeros((3, 1)), a[:, None], np.array([s]), np.zeros((3, 1)), np.array([s]), True, "f64")
    k = _kself(core, s)
    want = np.zeros(9)
    want[1], want[2], want[3], want[5], want[6], want[7] = a[2] * k, -a[1] * k, -a[2] * k, a[0] * k, a[1] * k, -a[0] * k
    assert np.abs(v1).max() == 0.0
    np.testing.assert_allclose(g1[:, 0], want, rtol=1e-13)


@pytest.mark.parametrize("core", CORES)
def test_singular_ring_on_axis_closed_form(core):
    """n equal singular vortex elements alpha = Gamma 2 pi R/n e_theta on a circle:
    on the axis the sum is exactly the continuous ring's u_z = Gamma R^2 / (2 (R^2+z^2)^1.5)
    (times 4 pi: outputs are unscaled); every core reduces to 1/rho^3 at sigma = 0."""
    n, R, Gam = 64, 1.3, 0.7
    th = 2 * np.pi * np.arange(n) / n
    sp = np.stack([R * np.cos(th), R * np.sin(th), np.zeros(n)])
    ss = Gam * 2 * np.pi * R / n * np.stack([-np.sin(th), np.cos(th), np.zeros(n)])
    z = np.array([0.0, 0.4, -1.7, 3.0])
    tp = np.stack([np.zeros(4), np.zeros(4), z])
    v, _ = oracle.p2p(core, sp, ss, np.zeros(n), tp, None, True, "f64")
    want = 4 * np.pi * Gam * R * R / (2 * (R * R + z * z) ** 1.5)
    np.testing.assert_allclose(v[2], want, rtol=1e-13)
    np.testing.assert_allclose(v[:2], 0.0, atol=1e-13 * want.max())


def test_rm_ring_on_axis_with_core():
    """RM with core sigma: the same sum is Gamma R^2/(2 (R^2+z^2+2 sigma^2)^1.5) x 4 pi
    for thick targets of the same sigma (r2 = rho^2 + ss^2 + st^2, CoreFunc.h:68)."""
    n, R, Gam, s = 48, 0.9, 1.1, 0.13
    th = 2 * np.pi * np.arange(n) / n
    sp = np.stack([R * np.cos(th), R * np.sin(th), np.zeros(n)])
    ss = Gam * 2 * np.pi * R / n * np.stack([-np.sin(th), np.cos(th), np.zeros(n)])
    z = np.array([0.0, 0.5, 2.0])
    tp = np.stack([np.zeros(3), np.zeros(3), z])
    v, _ = oracle.p2p("rm", sp, ss, np.full(n, s), tp, np.full(3, s), True, "f64")
    want = 4 * np.pi * Gam * R * R / (2 * (R * R + z * z + 2 * s * s) ** 1.5)
    np.testing.assert_allclose(v[2], want, rtol=1e-13)


@pytest.mark.parametrize("core,tol", [("rm", 2e-4), ("wl", 1e-6), ("exp", 1e-14), ("v2", 1e-6)])
def test_far_field_is_singular(core, tol):
    """Far from the core (rho = 100 sigma) each core approaches 1/rho^3."""
    rng = np.random.default_rng(5)
    a = rng.normal(size=(3, 1))
    s = 0.01
    tp = rng.normal(size=(3, 6))
    tp = tp / np.linalg.norm(tp, axis=0) * 100 * s
    v, _ = oracle.p2p(core, np.zeros((3, 1)), a, np.array([s]), tp, None, False, "f64")
    rho = np.linalg.norm(tp, axis=0)
    want = np.cross(a[:, 0], tp.T).T / rho ** 3
    np.testing.assert_allclose(v, want, rtol=tol, atol=0)


@pytest.mark.parametrize("core", CORES)
def test_rotation_translation_linearity(core):
    rng = np.random.default_rng(6)
    ns, nt = 50, 20
    sp = rng.uniform(-1, 1, (3, ns)); ss = rng.normal(size=(3, ns)); sr = rng.uniform(0.1, 0.3, ns)
    tp = rng.uniform(-1, 1, (3, nt)); tr = rng.uniform(0.1, 0.3, nt)
    v, g = oracle.p2p(core, sp, ss, sr, tp, tr, True, "f64")
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    shift = rng.normal(size=(3, 1))
    v2, g2 = oracle.p2p(core, q @ sp + shift, q @ ss, sr, q @ tp + shift, tr, True, "f64")
    np.testing.assert_allclose(v2, q @ v, rtol=1e-10, atol=1e-12 * np.abs(v).max())
    G = g.T.reshape(nt, 3, 3).transpose(0, 2, 1)          # G[t, i, j] = du_i/dx_j
    G2 = g2.T.reshape(nt, 3, 3).transpose(0, 2, 1)
    np.testing.assert_allclose(G2, q @ G @ q.T, rtol=1e-9, atol=1e-11 * np.abs(G).max())
    v3, g3 = oracle.p2p(core, sp, 2.5 * ss, sr, tp, tr, True, "f64")
    np.testing.assert_allclose(v3, 2.5 * v, rtol=1e-12, atol=1e-14 * np.abs(v3).max())
    np.testing.assert_allclose(g3, 2.5 * g, rtol=1e-12, atol=1e-14 * np.abs(g3).max())


def _single_panel(rng, far=True):
    nodes = np.array([[0.0, 0.0, 0.0], [0.3, 0.0, 0.0], [0.1, 0.25, 0.0]]).T
    area = 0.5 * 0.3 * 0.25
    gam = np.array([rng.normal(), rng.normal(), 0.0])
    ts = (gam * area)[:, None]
    return nodes, np.array([0, 1, 2], np.uint32), ts, np.array([area]), gam


@pytest.mark.parametrize("core", CORES)
def test_panel_far_field_is_point_vortex_at_centroid(core):
    """A well-separated panel (dist > 4 sqrt(area)) is one singular vortex + source
    point at its centroid carrying the absolute strengths (Kernels.h:1162-1172)."""
    rng = np.random.default_rng(8)
    nodes, idx, ts, area, _ = _single_panel(rng)
    ss = np.array([0.37])
    c = nodes.mean(axis=1)[:, None]
    d = rng.normal(size=(3, 7))
    tp = c + d / np.linalg.norm(d, axis=0) * (4.5 * np.sqrt(area[0]) + rng.uniform(0, 1, 7))
    v, _ = oracle.pan(core, nodes, idx, ts, area, ss, tp, True, "f64")
    dd = tp - c
    rho = np.linalg.norm(dd, axis=0)
    want = (np.cross(ts[:, 0], dd.T).T + ss[0] * area[0] * dd) / rho ** 3
    np.testing.assert_allclose(v, want, rtol=1e-12)


def _split(tris, levels):
    for _ in range(levels):
        nxt = []
        for t in tris:
            m01, m02, m12 = (t[0] + t[1]) / 2, (t[0] + t[2]) / 2, (t[1] + t[2]) / 2
            nxt += [np.array([t[0], m01, m02]), np.array([m01, t[1], m12]),
                    np.array([m01, m12, m02]), np.array([m02, m12, t[2]])]
        tris = nxt
    return tris


def _centroid_rule(gam, area, tris, x):
    cent = np.array([t.mean(axis=0) for t in tris])
    dd = x[None, :] - cent
    rho = np.linalg.norm(dd, axis=1)
    return (np.cross(gam, dd) / rho[:, None] ** 3).sum(0) * area / len(tris)


def test_panel_subdivision_is_centroid_rule_of_sheet_integral():
    """The recursion is a centroid rule of u(x) = int gamma x (x - y)/|x - y|^3 dA over
    midpoint sub-triangles: a target close enough to force depth 3 everywhere sees
    exactly the 64-leaf rule (built here independently); farther targets see a
    coarser adaptive rule that still beats the single-point rule against a 4^7-leaf
    reference quadrature."""
    rng = np.random.default_rng(9)
    nodes, idx, ts, area, gam = _single_panel(rng)
    P = nodes.T
    c = P.mean(axis=0)
    near = c + np.array([0.01, -0.02, 0.05])            # dist < 4 sqrt(area/16): full depth
    v3, _ = oracle.pan("rm", nodes, idx, ts, area, None, near[:, None], False, "f64", maxlev=3)
    np.testing.assert_allclose(v3[:, 0], _centroid_rule(gam, area[0], _split([P], 3), near), rtol=1e-12)
    fine = _split([P], 7)
    tp = np.array([[0.12, 0.08, 0.35], [0.2, 0.1, 0.5], [0.05, 0.2, 0.3]]).T
    v3, _ = oracle.pan("rm", nodes, idx, ts, area, None, tp, False, "f64", maxlev=3)
    v0, _ = oracle.pan("rm", nodes, idx, ts, area, None, tp, False, "f64", maxlev=0)
    for k in range(tp.shape[1]):
        ref = _centroid_rule(gam, area[0], fine, tp[:, k])
        e3 = np.linalg.norm(v3[:, k] - ref) / np.linalg.norm(ref)
        e0 = np.linalg.norm(v0[:, k] - ref) / np.linalg.norm(ref)
        assert e3 < 2e-2 and e3 < 0.5 * e0, (k, e3, e0)


def test_vcstyle_baseline_close_to_exact():
    """The Vc-style loop (AVX2 rsqrt/rcp estimates, 8 float lanes) agrees with the exact
    float oracle to the estimates' precision; Vc's own bit patterns are unpinned."""
    if not oracle.lib().orc_have_avx2():
        pytest.skip("no AVX2 on this host")
    p = scenes.ring(1000, sigma=0.05, seed=11)
    for core in CORES:
        a, ga = oracle.p2p_vcstyle(core, p.pos, p.str, p.rad, p.pos, p.rad)
        b, gb = oracle.p2p(core, p.pos, p.str, p.rad, p.pos, p.rad, True, "f32d")
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 3e-3, core
        assert np.linalg.norm(ga - gb) / np.linalg.norm(gb) < 3e-3, core
        # subset of targets
        c, gc = oracle.p2p_vcstyle(core, p.pos, p.str, p.rad, p.pos, p.rad, t0=100, t1=200)
        np.testing.assert_array_equal(c[:, 100:200], a[:, 100:200])
        assert np.abs(c[:, :100]).max() == 0


# ---- the RK2 step restatement (oracle/rk2.py) ------------------------------------------

@pytest.mark.parametrize("core", CORES)
def test_rk2_corotating_pair(core):
    """Two equal particles with strength a z_hat: by symmetry u_1 = -u_0 and
    u_0 = K(|r|) a z_hat x r / (4 pi) with r = x_0 - x_1, so one step is Ralston's
    method (stage at 2/3 h, weights 1/4 and 3/4) on this two-body ODE:
        r1 = r0 + (4/3) h u_0(r0),   x_0 <- x_0 + h (u_0(r0)/4 + 3 u_0(r1)/4).
    K comes from the pinned pair oracle.  The strength is not stretched
    ((a . grad) u = 0 for a || z_hat; the self pair adds only K(0) skew(a)) and
    the elongation stays 1.  A wrong stage point or weight fails this."""
    from oracle import rk2
    a, rho, sig, h = 0.8, 0.3, 0.1, 0.05
    half = rho / 2
    pos = np.array([[half, -half], [0.0, 0.0], [0.0, 0.0]])
    st = np.array([[0.0, 0.0], [0.0, 0.0], [a, a]])
    rad = np.array([sig, sig])

    def u0(r):
        """velocity of particle 0 for relative vector r (finalized: / 4 pi)"""
        p2 = np.stack([r / 2, -r / 2], axis=1)
        v, _ = oracle.p2p(core, p2, st, rad, p2, rad, False, "f64")
        return v[:, 0] * 0.25 / np.pi

    r0 = pos[:, 0] - pos[:, 1]
    k1 = u0(r0)
    r1 = r0 + (4.0 / 3.0) * h * k1
    k2 = u0(r1)
    want0 = pos[:, 0] + h * (0.25 * k1 + 0.75 * k2)
    x, s, e = rk2.rk2_step(core, pos, st, rad, np.ones(2), h, (0.0, 0.0, 0.0), "f64")
    np.testing.assert_allclose(x[:, 0], want0, rtol=0, atol=1e-14)
    np.testing.assert_allclose(x[:, 1], -want0, rtol=0, atol=1e-14)
    np.testing.assert_allclose(s, st, rtol=0, atol=1e-14)
    np.testing.assert_allclose(e, 1.0, rtol=0, atol=1e-14)
    x32, s32, e32 = rk2.rk2_step(core, pos.astype(np.float32), st.astype(np.float32), rad.astype(np.float32),
                                 np.ones(2, np.float32), h, (0.0, 0.0, 0.0), "f32d")
    np.testing.assert_allclose(x32[:, 0], want0, rtol=0, atol=2e-7)


@pytest.mark.parametrize("core", CORES)
def test_rk2_stretching_sign(core):
    """Stretching: a weak particle with alpha along z on the axis of a ring changes
    its strength and elongation with the sign of dw/dz there, as (alpha . grad) u
    prescribes (Points.h:376-423)."""
    from oracle import rk2
    ring = scenes.ring(400, R=1.0, a=0.01, sigma=0.05, seed=4)
    probe_pos = np.array([[0.0], [0.0], [0.5]], np.float32)
    pos = np.concatenate([ring.pos, probe_pos], axis=1)
    st = np.concatenate([ring.str, np.array([[0.0], [0.0], [1e-6]], np.float32)], axis=1)
    rad = np.concatenate([ring.rad, np.array([0.05], np.float32)])
    x, s, e = rk2.rk2_step(core, pos, st, rad, np.ones(pos.shape[1]), 0.05, (0, 0, 0), "f64")
    # on the axis of a ring the axial velocity decays away from the ring plane, so
    # an axial filament there is stretched or compressed by the sign of dw/dz
    vv, gg = oracle.p2p(core, pos, st, rad, probe_pos, rad[-1:], True, "f64")
    dwdz = gg[8, 0]
    grow = s[2, -1] - 1e-6
    assert np.sign(grow) == np.sign(dwdz) and abs(grow) > 0
    assert np.sign(e[-1] - 1.0) == np.sign(dwdz)


def _rm_field(src, alpha, sig_s, sig_t, x):
    """Velocity (finalized: / 4 pi) and its gradient g[3j+i] = du_i/dx_j at x of one
    Rosenhead-Moore particle, written from the formulas: u = K(d) alpha x d / 4 pi,
    K = (|d|^2 + ss^2 + st^2)^-3/2, d = x - src, so
    du_i/dx_j = (dK/dx_j (alpha x d)_i + K eps_ikj alpha_k) / 4 pi."""
    d = x - src
    r2 = d @ d + sig_s ** 2 + sig_t ** 2
    K = r2 ** -1.5
    c = np.cross(alpha, d)
    u = K * c / (4 * np.pi)
    dK = -3.0 * d * r2 ** -2.5
    eps = np.zeros((3, 3, 3))
    eps[0, 1, 2] = eps[1, 2, 0] = eps[2, 0, 1] = 1.0
    eps[0, 2, 1] = eps[2, 1, 0] = eps[1, 0, 2] = -1.0
    J = np.zeros((3, 3))      # J[i, j] = du_i/dx_j
    for i in range(3):
        for j in range(3):
            J[i, j] = (dK[j] * c[i] + K * sum(eps[i, k, j] * alpha[k] for k in range(3))) / (4 * np.pi)
    return u, J


def _probe_step_expected(J_use, weights, h, src, alpha_s, sig, p0, a0, e0):
    """One step of the reference's scheme for a weak probe near one stationary source,
    from the closed-form field: Ralston's position update (x1 = x0 + 2h/3 u(x0),
    x = x0 + h (w1 u(x0) + w2 u(x1))) and the stretch with the ORIGINAL strength at
    both stages (Points.h:380-416): wdu_k = (a0 . grad) u at x0 and x1, combined with
    the same weights; elongation e0 (1 + h a0 . wdu / |a0|^2)."""
    u1, J1 = _rm_field(src, alpha_s, sig, sig, p0)
    x1 = p0 + (2.0 / 3.0) * h * u1
    u2, J2 = _rm_field(src, alpha_s, sig, sig, x1)
    w1, w2 = weights
    x = p0 + h * (w1 * u1 + w2 * u2)
    wdu = w1 * J_use(J1) @ a0 + w2 * J_use(J2) @ a0
    a = a0 + h * wdu
    e = e0 * (1.0 + h * (a0 @ wdu) / (a0 @ a0))
    return x, a, e


def test_rk2_stretching_off_axis_probe():
    """The stretch term and the elongation of one Ralston step, pinned by a case where
    the velocity gradient is not symmetric: a weak probe particle off the axis of a
    single Rosenhead-Moore source, with its strength along no eigenvector of the
    gradient.  The source does not move (its self pair has no velocity) and keeps its
    strength (the self pair's gradient K(0) skew(alpha) gives (alpha . grad) u = 0);
    the probe's strength is 1e-12 of the source's, so its influence back is far below
    the bar.  The expected step comes from the closed-form field of one particle.  A
    transposed gradient ((grad u)^T a instead of (a . grad) u, the alternative the
    reference's comment names, Points.h:380-383) or swapped weights 3/4, 1/4 fail it
    by orders of magnitude more than the bar."""
    from oracle import rk2
    src = np.array([0.0, 0.0, 0.0])
    alpha_s = np.array([0.3, -0.2, 1.0])
    sig = 0.1
    p0 = np.array([0.25, 0.12, 0.07])
    a0 = np.array([0.4, 0.9, -0.3]) * 1e-12
    h = 0.02
    pos = np.stack([src, p0], axis=1)
    st = np.stack([alpha_s, a0], axis=1)
    rad = np.array([sig, sig])
    x, s, e = rk2.rk2_step("rm", pos, st, rad, np.ones(2), h, (0.0, 0.0, 0.0), "f64")
    ident = lambda J: J
    want_x, want_a, want_e = _probe_step_expected(ident, (0.25, 0.75), h, src, alpha_s, sig, p0, a0, 1.0)
    da = s[:, 1] - a0
    dwant = want_a - a0
    np.testing.assert_allclose(x[:, 1], want_x, rtol=0, atol=1e-12)
    assert np.linalg.norm(da - dwant) <= 1e-9 * np.linalg.norm(dwant)
    assert abs((e[1] - 1.0) - (want_e - 1.0)) <= 1e-9 * abs(want_e - 1.0)
    # the source: neither moved nor stretched beyond the probe's O(1e-12) influence
    assert np.linalg.norm(x[:, 0] - src) <= 1e-12 and np.linalg.norm(s[:, 0] - alpha_s) <= 1e-12
    # the case discriminates: the plausible mistakes are far outside the bar
    for J_use, wts in ((lambda J: J.T, (0.25, 0.75)), (ident, (0.75, 0.25)), (ident, (0.5, 0.5))):
        _, bad_a, bad_e = _probe_step_expected(J_use, wts, h, src, alpha_s, sig, p0, a0, 1.0)
        assert np.linalg.norm((bad_a - a0) - dwant) > 1e-3 * np.linalg.norm(dwant)
    # (a . (grad u) a is the same for the transpose; the weights move the elongation)
    _, bad_a, bad_e = _probe_step_expected(ident, (0.75, 0.25), h, src, alpha_s, sig, p0, a0, 1.0)
    assert abs((bad_e - 1.0) - (want_e - 1.0)) > 1e-3 * abs(want_e - 1.0)
    # float storage with double accumulation (the served step's mode) within FP32 rounding
    x32, s32, e32 = rk2.rk2_step("rm", pos.astype(np.float32), st.astype(np.float32), rad.astype(np.float32),
                                 np.ones(2, np.float32), h, (0.0, 0.0, 0.0), "f32d")
    assert np.linalg.norm(s32[:, 1].astype(np.float64) - want_a) <= 1e-5 * np.linalg.norm(dwant) + 1e-7 * np.linalg.norm(a0)


def test_rk2_stretching_planted_transpose_fails(monkeypatch):
    """Planting the transposed stretch term in the oracle makes the off-axis check fail."""
    from oracle import rk2
    T = rk2._wdu
    monkeypatch.setattr(rk2, "_wdu", lambda s, ug, dt: T(s, ug[[0, 3, 6, 1, 4, 7, 2, 5, 8]], dt))
    with pytest.raises(AssertionError):
        test_rk2_stretching_off_axis_probe()


def test_rk2_freestream_only():
    """With zero strengths the step moves every particle by dt * U_inf exactly."""
    from oracle import rk2
    p = scenes.ring(50, seed=3)
    st = np.zeros_like(p.str)
    fs = (0.3, -0.1, 0.05)
    x, s, e = rk2.rk2_step("rm", p.pos, st, p.rad, np.ones(p.n), 0.2, fs, "f64")
    np.testing.assert_allclose(x, p.pos.astype(np.float64) + 0.2 * np.array(fs)[:, None], rtol=0, atol=1e-15)
    assert np.all(s == 0) and np.all(e == 1)


# ---- points_affect_panels (Influence.h:1190-1210) ---------------------------------

@pytest.mark.parametrize("core", CORES)
def test_particles_on_panels_match_reference_golden_bitwise(core):
    g = gold(f"ptpan_{core}.npz")
    for mode in MODES:
        v = oracle.ptpan(core, g["spos"], g["sstr"], g["nodes"], g["idx"], g["area"], mode)
        np.testing.assert_array_equal(v, g["vel_" + mode], err_msg=mode)


@pytest.mark.skipif(any(oracle.ref_lib(c) is None for c in CORES), reason="oracle/_ref not built")
def test_particles_on_panels_match_live_reference():
    pan = scenes.icosphere_panels(levels=2)
    rng = np.random.default_rng(78)
    d = rng.normal(size=(3, 90))
    sp = (d / np.linalg.norm(d, axis=0) * rng.uniform(0.95, 2.0, 90)).astype(np.float32)
    ss = (rng.normal(size=(3, 90)) * 0.02).astype(np.float32)
    for core in CORES:
        for mode in MODES:
            a = oracle.ptpan(core, sp, ss, pan.nodes, pan.idx, pan.area, mode)
            b = oracle.ref_ptpan(core, sp, ss, pan.nodes, pan.idx, pan.area, mode)
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("core", CORES)
def test_particles_on_panels_far_field_is_point_vortex_velocity(core):
    """A particle more than 4 sqrt(area) from a panel's centroid moves the panel at the
    point-vortex velocity it induces at the centroid: u = alpha x (c - p) / |c - p|^3
    (the singular kernel; the reference's 1/(4 pi) comes later)."""
    rng = np.random.default_rng(12)
    nodes, idx, _, area, _ = _single_panel(rng)
    c = nodes.mean(axis=1)
    d = rng.normal(size=(3, 5))
    p = c[:, None] + d / np.linalg.norm(d, axis=0) * (4.5 * np.sqrt(area[0]) + rng.uniform(0, 1, 5))
    a = rng.normal(size=(3, 5))
    for j in range(5):
        v = oracle.ptpan(core, p[:, j:j + 1], a[:, j:j + 1], nodes, idx, area, "f64")
        dd = c - p[:, j]
        want = np.cross(a[:, j], dd) / np.linalg.norm(dd) ** 3
        np.testing.assert_allclose(v[:, 0], want, rtol=1e-12)


def test_particles_on_panels_is_panel_kernel_backwards():
    """Particle j on panel i equals minus the panel (carrying alpha_j) on the point p_j,
    operation for operation, at every recursion depth."""
    rng = np.random.default_rng(13)
    nodes, idx, _, area, _ = _single_panel(rng)
    c = nodes.mean(axis=1)
    pts = c[:, None] + rng.normal(size=(3, 40)) * np.geomspace(0.005, 2.0, 40)
    for mode in MODES:
        for j in range(40):
            a = rng.normal(size=(3, 1))
            v = oracle.ptpan("rm", pts[:, j:j + 1], a, nodes, idx, area, mode)
            w, _ = oracle.pan("rm", nodes, idx, a, area, None, pts[:, j:j + 1], False, mode)
            np.testing.assert_array_equal(v, -w)
