"""The BEM influence matrix oracle (oracle/o3d_oracle.cpp orc_bem_matrix, the
restatement of panels_on_panels_coeff, Coefficients.h:169-444, over
rkernel_2vs_2p, Kernels.h:1213-1311).

Pinned by:
  * the reference's own recursion: the float mode is bitwise the blocks in
    tests/golden/bem_matrix.npz (made by make_bem_golden.py through
    oracle/_ref) and, where oracle/_ref is built, the live reference on more
    geometry and every depth; the FP64 mode equals the reference instantiated in
    double wherever the two trees coincide;
  * a well-separated pair is one singular leaf: fac sa ((b x d) . t) / |d|^3 in
    closed form;
  * the diagonal blocks of a surface on itself are the reference's constants;
  * a unit source sheet on a closed sphere has unit outward flux: the normal rows
    of the source columns sum to ~1 (jump 1/2 plus principal value 1/2), and the
    sum converges toward 1 as the recursion deepens.
The GPU's division by 3 for the centroids of its tree decisions (a multiply and
two FMAs) is checked against IEEE division on every positive float whose
quotient is normal (the negative half mirrors it: every step is sign-symmetric).
"""
import os

import numpy as np
import pytest

import oracle
from omega3d_amd import scenes

GOLD = os.path.join(os.path.dirname(__file__), "golden", "bem_matrix.npz")
CORE_NAMES = ("rm", "wl", "exp", "v2")


def surface(levels, radius, centre, jitter=0.0, seed=0):
    v, f = scenes.icosphere(levels, radius)
    if jitter:
        v = v + jitter * np.random.default_rng(seed).normal(size=v.shape) * radius / 2 ** levels
    v = v + np.asarray(centre, float)[None, :]
    return oracle.surface(np.ascontiguousarray(v.T, np.float32), f.astype(np.uint32))


def test_float_mode_is_bitwise_the_reference_golden():
    g = np.load(GOLD)
    surf = {k: (g[f"{k}_nodes"], g[f"{k}_idx"], g[f"{k}_area"], g[f"{k}_bases"]) for k in "AB"}
    # the golden bases and areas are the oracle's compute_bases
    for k in "AB":
        b, a = oracle.bem_bases(surf[k][0], surf[k][1])
        np.testing.assert_array_equal(b, surf[k][3])
        np.testing.assert_array_equal(a, surf[k][2])
    n = 0
    while f"case{n}" in g:
        core, s, t, nunk, same, maxlev = (int(x) for x in g[f"case{n}"])
        m = oracle.bem_matrix(CORE_NAMES[core - 1], surf[chr(s)], surf[chr(t)], nunk, bool(same), maxlev, mode="f32")
        np.testing.assert_array_equal(m, g[f"case{n}_matrix"])
        n += 1
    assert n == 6


@pytest.mark.skipif(oracle.ref_lib("wl") is None, reason="oracle/_ref not built")
@pytest.mark.parametrize("core", ["wl", "rm", "exp", "v2"])
def test_float_mode_is_bitwise_the_live_reference(core):
    A = surface(1, 1.0, (0, 0, 0), jitter=0.2, seed=1)
    B = surface(0, 0.4, (1.3, -0.3, 0.2), jitter=0.1, seed=2)
    for (s, t, nunk, same) in ((A, A, 3, True), (B, A, 2, False), (A, B, 3, False)):
        for maxlev in (0, 1, 2, 3, 4) if core == "wl" else (3,):
            m = oracle.bem_matrix(core, s, t, nunk, same, maxlev, mode="f32")
            r = oracle.ref_bem_matrix(core, s, t, nunk, same, maxlev, mode="f32")
            np.testing.assert_array_equal(m, r, err_msg=f"{core} nunk {nunk} maxlev {maxlev}")
            m64 = oracle.bem_matrix(core, s, t, nunk, same, maxlev, mode="f64t")
            r64 = oracle.ref_bem_matrix(core, s, t, nunk, same, maxlev, mode="f64")
            # equal where the double tree is the float tree: compare the columns
            # whose float and double blocks agree to the float rounding
            close = np.abs(m - r64).max(0) <= 1e-5 * np.abs(r64).max()
            assert close.mean() > 0.9
            np.testing.assert_allclose(m64[:, close], r64[:, close], rtol=0, atol=1e-13 * np.abs(r64).max())


def test_well_separated_pair_is_one_leaf():
    """Two small panels far apart: each coefficient is fac * sa * ((b_k x d) . t_m)
    / |d|^3 (vortex unknowns) or fac * sa * (d . t_m) / |d|^3 (source), d from the
    source centroid to the target centroid."""
    rng = np.random.default_rng(7)
    for trial in range(5):
        ns = rng.normal(size=(3, 3)) * 0.05
        nt = rng.normal(size=(3, 3)) * 0.05 + rng.normal(size=(3, 1)) * 3.0
        S = oracle.surface(np.ascontiguousarray(ns, np.float32), np.array([[0, 1, 2]], np.uint32))
        T = oracle.surface(np.ascontiguousarray(nt, np.float32), np.array([[0, 1, 2]], np.uint32))
        m = oracle.bem_matrix("wl", S, T, 3, False, 3, mode="f64t")
        sc = S[0].astype(np.float64).mean(1)
        tc = T[0].astype(np.float64).mean(1)
        d = tc - sc
        k = 1.0 / np.linalg.norm(d) ** 3
        sa = float(S[2][0])
        sb, tb = S[3].astype(np.float64), T[3].astype(np.float64)
        want = np.zeros((3, 3))
        for kk in range(3):
            u = np.cross(sb[3 * kk:3 * kk + 3, 0], d) if kk < 2 else d
            for mm in range(3):
                want[mm, kk] = sa * k * (u @ tb[3 * mm:3 * mm + 3, 0]) / (4 * np.pi)
        np.testing.assert_allclose(m, want, rtol=2e-6, atol=1e-9 * np.abs(want).max())


def test_self_blocks_are_the_reference_constants():
    S = surface(1, 1.0, (0, 0, 0))
    tp = np.float32(2.0 * np.pi) * np.float32(1.0 / (4.0 * np.pi))
    for nunk in (2, 3):
        m = oracle.bem_matrix("wl", S, S, nunk, True, 3, mode="f32")
        want = np.zeros((nunk, nunk), np.float32)
        want[1, 0], want[0, 1] = tp, -tp          # [row m, column k]
        if nunk == 3:
            want[2, 2] = tp
        for j in range(0, 80, 7):
            np.testing.assert_array_equal(m[j * nunk:(j + 1) * nunk, j * nunk:(j + 1) * nunk], want)


def test_source_sheet_flux_on_a_sphere():
    v, f = scenes.icosphere(2, 1.0)
    S = oracle.surface(np.ascontiguousarray(v.T, np.float32), f.astype(np.uint32))
    dev = []
    for maxlev in (0, 3):
        m = oracle.bem_matrix("wl", S, S, 3, True, maxlev, mode="f64t")
        rs = m[2::3, 2::3].sum(1)
        dev.append(np.abs(rs - 1.0).max())
    assert dev[1] < 1.5e-2 and dev[1] < 0.5 * dev[0], dev


def test_gpu_division_by_three_is_ieee():
    import ctypes as C
    L = oracle.lib()
    L.orc_check_div3.restype = C.c_int64
    L.orc_check_div3.argtypes = [C.c_int64, C.c_int64]
    assert L.orc_check_div3(0, 1 << 31) == 0
