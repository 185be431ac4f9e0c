"""Source split of small direct calls (o3d_api.hip split_plan, k_p2p_split and
k_p2p_combine): a call whose targets make too few workgroups to fill the GPU, with
the tile and lane count left to the library, spreads its sources over workgroups
and adds the slices in a fixed order.  Checked against the oracle (FP64) and the
one-pass kernel on every core, thick and thin targets, with and without the
gradient and the self pair, at ragged sizes (slices of unequal length, sources
shorter than one tile, targets beyond one workgroup), for accumulation into
non-zero outputs and for repeatability; and the EXP split of one collection, which
visits its targets in Morton order and reads its sources as sorted rows."""
import numpy as np
import pytest

import oracle
from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def infl():
    from omega3d_amd import _lib, influence
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return influence


def cloud(n, seed):
    return scenes.ring(n, R=1.1, sigma=0.04, seed=seed, jitter=0.05)


# (targets, sources): one tile; several slices with a short last one; one slice of
# each tile; targets of several workgroups with two slices; sources under one tile
SIZES = [(1000, 1000), (300, 5000), (8192, 8192), (70000, 300), (4000, 100)]


@pytest.mark.parametrize("nt,ns", SIZES)
@pytest.mark.parametrize("core", ["rm", "wl", "exp", "v2"])
def test_split_matches_oracle_and_one_pass(infl, core, nt, ns):
    s = cloud(ns, 21)
    same = nt == ns
    t = s if same else cloud(nt, 22)
    for prec in ("comp", "fp32"):
        for grads in (True, False):
            selfx = 1 if same and grads else 0
            v, g = infl.p2p(core, s.pos, s.str, s.rad, t.pos, t.rad, grads, prec, self_term=selfx)
            v1, g1 = infl.p2p(core, s.pos, s.str, s.rad, t.pos, t.rad, grads, prec, self_term=selfx, tile=256, tpt=1)
            sel = np.arange(0, nt, max(1, nt // 700))
            ov, og = oracle.p2p(core, s.pos, s.str, s.rad, t.pos[:, sel], t.rad[sel], grads, "f64")
            if selfx:
                # the oracle includes the self pair; take it out through a one-target call
                for k, i in enumerate(sel):
                    _, gi = oracle.p2p(core, s.pos[:, i:i + 1], s.str[:, i:i + 1], s.rad[i:i + 1],
                                       t.pos[:, i:i + 1], t.rad[i:i + 1], True, "f64")
                    og[:, k] -= gi[:, 0]
            tol = 2e-6 if prec == "comp" else 2e-5
            assert rel_errors(v[:, sel], ov)[0] <= tol, (prec, grads)
            # the one-pass kernel is the same sum in another order
            assert rel_errors(v, v1)[0] <= tol
            if grads:
                assert rel_errors(g[:, sel], og)[0] <= tol, (prec, grads)
                assert rel_errors(g, g1)[0] <= tol


def test_split_thick_accumulates_and_repeats(infl):
    s = cloud(6000, 23)
    t = cloud(2000, 24)
    rng = np.random.default_rng(5)
    for prec in ("comp", "fp32"):
        v0 = rng.normal(size=(3, t.n)).astype(np.float32)
        g0 = rng.normal(size=(9, t.n)).astype(np.float32)
        vz, gz = infl.p2p("exp", s.pos, s.str, s.rad, t.pos, t.rad, True, prec)
        v, g = infl.p2p("exp", s.pos, s.str, s.rad, t.pos, t.rad, True, prec, vel=v0.copy(), grad=g0.copy())
        np.testing.assert_allclose(v, v0 + vz, rtol=1e-6, atol=1e-6 * np.abs(vz).max())
        np.testing.assert_allclose(g, g0 + gz, rtol=1e-6, atol=1e-6 * np.abs(gz).max())
        v2, g2 = infl.p2p("exp", s.pos, s.str, s.rad, t.pos, t.rad, True, prec)
        np.testing.assert_array_equal(vz, v2)
        np.testing.assert_array_equal(gz, g2)
        ov, og = oracle.p2p("exp", s.pos, s.str, s.rad, t.pos, t.rad, True, "f64")
        tol = 2e-6 if prec == "comp" else 2e-5
        assert rel_errors(vz, ov)[0] <= tol and rel_errors(gz, og)[0] <= tol
    # thin targets (no radius): the same path without the target core term
    v, g = infl.p2p("rm", s.pos, s.str, s.rad, t.pos, None, True, "comp")
    ov, og = oracle.p2p("rm", s.pos, s.str, s.rad, t.pos, None, True, "f64")
    assert rel_errors(v, ov)[0] <= 2e-6 and rel_errors(g, og)[0] <= 2e-6


def test_split_single_ring_differs_only_in_order(infl):
    """The single 8,192-particle ring: the split (32 slices) and the one-pass kernel
    agree to the FP32 bar; plain FP32 sums in another order must differ somewhere,
    which shows that the split path ran."""
    p = scenes.ring(8192, sigma=0.05)
    v, g = infl.p2p("rm", p.pos, p.str, p.rad, p.pos, p.rad, True, "fp32", self_term=1)
    v1, g1 = infl.p2p("rm", p.pos, p.str, p.rad, p.pos, p.rad, True, "fp32", self_term=1, tile=256, tpt=1)
    assert not (np.array_equal(v, v1) and np.array_equal(g, g1))
    assert rel_errors(v, v1)[0] <= 2e-5 and rel_errors(g, g1)[0] <= 2e-5


@pytest.mark.parametrize("n", [40000, 262144])
def test_split_exp_one_collection_in_morton_order(infl, n):
    """An EXP call on one collection of 32,768 or more particles that splits its
    sources also visits its targets in Morton order and reads its sources as sorted
    rows with tile spheres: the partial sums are in that order and the combine adds
    each target's slices into its original index.  Against FP64 on spread targets,
    self pair included and excluded, both precisions; into non-zero outputs; and
    the same bits on a repeated call."""
    from omega3d_amd.metrics import self_term_gradient
    p = scenes.colliding_rings(n_each=n // 2, seed=scenes.SEED_BASE + 73)
    sel = np.linspace(0, n - 1, 512).astype(np.int64)
    rng = np.random.default_rng(73)
    v0 = rng.normal(size=(3, n)).astype(np.float32)
    g0 = rng.normal(size=(9, n)).astype(np.float32)
    ov, og = oracle.p2p("exp", p.pos, p.str, p.rad, p.pos[:, sel], p.rad[sel], True, "f64")
    for selfx in (0, 1):
        ogx = og - self_term_gradient("exp", p.str[:, sel], p.rad[sel]) if selfx else og
        for prec in ("comp", "fp32"):
            v, g = infl.p2p("exp", p.pos, p.str, p.rad, p.pos, p.rad, True, prec, self_term=selfx, method="direct")
            ev, eg = rel_errors(v[:, sel], ov), rel_errors(g[:, sel], ogx)
            print(f"split EXP N={n} self_term={selfx} {prec}: vel {ev} grad {eg}", flush=True)
            tol = 1e-6 if prec == "comp" else 1e-2   # plain FP32: the float sum's own error
            assert ev[0] <= tol and eg[0] <= tol
            v2, g2 = infl.p2p("exp", p.pos, p.str, p.rad, p.pos, p.rad, True, prec, self_term=selfx,
                              method="direct")
            np.testing.assert_array_equal(v, v2)
            np.testing.assert_array_equal(g, g2)
            va, ga = infl.p2p("exp", p.pos, p.str, p.rad, p.pos, p.rad, True, prec, self_term=selfx,
                              method="direct", vel=v0.copy(), grad=g0.copy())
            np.testing.assert_allclose(va, v0 + v, rtol=1e-6, atol=1e-6 * np.abs(v).max())
            np.testing.assert_allclose(ga, g0 + g, rtol=1e-6, atol=1e-6 * np.abs(g).max())


def test_split_exp_targets_in_morton_order_same_bits():
    """EXP sources acting on another collection of 32,768 or more targets: the split
    visits the targets in Morton order, which does not change any target's sums, so
    the result has the bits of the index-order split that the no-allocation entry
    point takes (a caller's scratch keeps index order; 4,096 sources are at most 16
    slices, whose partial sums fit its scratch, so both calls split alike)."""
    import ctypes as C
    from omega3d_amd import _lib
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    src = scenes.ring(4096, R=1.0, sigma=0.02, seed=scenes.SEED_BASE + 74)
    tgt = scenes.shell_points(40000, 0.9, 1.1, seed=scenes.SEED_BASE + 75)
    D = _lib.DeviceArray.from_host
    keep = [D(src.pos), D(src.str), D(src.rad), D(tgt)]
    s = _lib.Points()
    s.n = src.n
    for d in range(3):
        s.x[d], s.s[d] = keep[0].row(d), keep[1].row(d)
    s.r = keep[2].ptr
    res = {}
    for name in ("o3d_points_affect_points_dev", "o3d_points_affect_points_noalloc_dev"):
        for prec in ("comp", "fp32"):
            dv, dg = _lib.DeviceArray((3, tgt.shape[1])), _lib.DeviceArray((9, tgt.shape[1]))
            dv.zero()
            dg.zero()
            t = _lib.Points()
            t.n = tgt.shape[1]
            for d in range(3):
                t.x[d], t.u[d] = keep[3].row(d), dv.row(d)
            for k in range(9):
                t.ug[k] = dg.row(k)
            o = _lib.make_opts("exp", True, prec, method="direct")
            _lib.check(getattr(_lib.lib, name)(C.byref(s), C.byref(t), C.byref(o), None))
            _lib.synchronize()
            res[name, prec] = dv.to_host(), dg.to_host()
    for prec in ("comp", "fp32"):
        a, b = res["o3d_points_affect_points_dev", prec], res["o3d_points_affect_points_noalloc_dev", prec]
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
