"""GPU parity of the BEM influence matrix (o3d_bem_matrix[_dev], bemmat.hip)
against the oracle's restatement of panels_on_panels_coeff (Coefficients.h:169-444)
through the C ABI.

Bars (normwise over the block: max |a - b| / max |b|):
  * against the same tree with FP64 leaves and sums (oracle mode f64t): <= 1e-6;
  * against the reference's float arithmetic (oracle mode f32, bitwise the
    reference's own recursion): within 3x that path's own distance from f64t plus
    1e-7, since the GPU takes the same tree and sums a recursed pair in FP64;
  * the self blocks bitwise the reference's constants; the number of recursed
    pairs equal to the root decisions taken in numpy float32; repeated calls
    bitwise equal.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from omega3d_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from omega3d_amd import _lib, influence
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return _lib, influence


def surface(levels, radius, centre, jitter=0.0, seed=0):
    v, f = scenes.icosphere(levels, radius)
    if jitter:
        v = v + jitter * np.random.default_rng(seed).normal(size=v.shape) * radius / 2 ** levels
    v = v + np.asarray(centre, float)[None, :]
    return oracle.surface(np.ascontiguousarray(v.T, np.float32), f.astype(np.uint32))


def nerr(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / np.abs(b).max())


A = surface(2, 1.0, (0, 0, 0), jitter=0.1, seed=3)          # 320 panels
B = surface(0, 0.35, (1.45, 0.2, -0.1), jitter=0.1, seed=4)  # 20 panels, close to A
CASES = [("A", None, 3), ("A", None, 2), ("B", "A", 3), ("A", "B", 2), ("B", None, 3)]


@pytest.mark.parametrize("maxlev", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_matrix_matches_oracle(mods, case, maxlev):
    _lib, infl = mods
    sname, tname, nunk = CASES[case]
    S = {"A": A, "B": B}[sname]
    T = S if tname is None else {"A": A, "B": B}[tname]
    same = tname is None
    print(f"case: src {sname} targ {tname or sname} nunk {nunk} maxlev {maxlev}", flush=True)
    g = infl.bem_matrix(S[:3], None if same else T[:3], nunk, maxlev)
    m64 = oracle.bem_matrix("wl", S, T, nunk, same, maxlev, mode="f64t")
    m32 = oracle.bem_matrix("wl", S, T, nunk, same, maxlev, mode="f32")
    e64, eref = nerr(g, m64), nerr(m32, m64)
    print(f"gpu vs f64t {e64:.3g}; reference float vs f64t {eref:.3g}; gpu vs reference {nerr(g, m32):.3g}")
    assert e64 <= 1e-6
    assert nerr(g, m32) <= 3 * eref + 1e-7
    if same:
        tp = np.float32(2.0 * np.pi) * np.float32(1.0 / (4.0 * np.pi))
        want = np.zeros((nunk, nunk), np.float32)
        want[1, 0], want[0, 1] = tp, -tp
        if nunk == 3:
            want[2, 2] = tp
        for j in range(S[1].shape[0]):
            np.testing.assert_array_equal(g[j * nunk:(j + 1) * nunk, j * nunk:(j + 1) * nunk], want)


def _root_near_count(S, T, same):
    """pairs the root decision sends into the recursion, in numpy float32 (each
    operation correctly rounded, as the reference's float path)"""
    def cen(s):
        v = s[0][:, s[1]]                      # (3, P, 3)
        return (v[:, :, 0] + v[:, :, 1] + v[:, :, 2]) / np.float32(3.0)
    cs, ct = cen(S), cen(T)
    rs, rt = np.sqrt(S[2]), np.sqrt(T[2])
    ex = ct[0][:, None] - cs[0][None, :]
    ey = ct[1][:, None] - cs[1][None, :]
    ez = ct[2][:, None] - cs[2][None, :]
    dist = np.sqrt(ex * ex + ey * ey + ez * ez)
    near = ~(dist > (rt[:, None] + rs[None, :]) * np.float32(4.0))
    if same:
        np.fill_diagonal(near, False)
    return int(near.sum())


def test_device_entry_profile_and_repeatability(mods):
    _lib, infl = mods
    S, T = A, B
    D = _lib.DeviceArray.from_host
    dev = {}
    for key, s in (("s", S), ("t", T)):
        dn, di, da = D(s[0]), D(s[1].reshape(-1).view(np.float32)), D(s[2])
        p = _lib.Panels()
        p.np, p.nn = s[2].shape[0], s[0].shape[1]
        for d in range(3):
            p.x[d] = dn.row(d)
        p.idx, p.area = di.ptr, da.ptr
        dev[key] = (p, dn, di, da)
    nunk = 3
    rows, cols = nunk * T[2].shape[0], nunk * S[2].shape[0]
    out = _lib.DeviceArray((rows * cols,))
    prof = _lib.BemMatrixProfile()
    _lib.check(_lib.lib.o3d_bem_matrix_profile_dev(C.byref(dev["s"][0]), C.byref(dev["t"][0]), nunk, 0, -1,
                                                   C.c_void_p(out.ptr), None, C.byref(prof)))
    _lib.synchronize()
    assert prof.near_pairs == _root_near_count(S, T, False) > 0
    assert prof.total_ms > 0
    g1 = out.to_host().reshape(cols, rows).T
    _lib.check(_lib.lib.o3d_bem_matrix_dev(C.byref(dev["s"][0]), C.byref(dev["t"][0]), nunk, 0, -1,
                                           C.c_void_p(out.ptr), None))
    _lib.synchronize()
    g2 = out.to_host().reshape(cols, rows).T
    np.testing.assert_array_equal(g1, g2)
    # the host entry point gives the same bits
    np.testing.assert_array_equal(g1, infl.bem_matrix(S[:3], T[:3], nunk, -1))
    # the same-surface count excludes the diagonal
    prof2 = _lib.BemMatrixProfile()
    out2 = _lib.DeviceArray((nunk * nunk * S[2].shape[0] ** 2,))
    _lib.check(_lib.lib.o3d_bem_matrix_profile_dev(C.byref(dev["s"][0]), C.byref(dev["s"][0]), nunk, 1, -1,
                                                   C.c_void_p(out2.ptr), None, C.byref(prof2)))
    _lib.synchronize()
    assert prof2.near_pairs == _root_near_count(S, S, True)


def test_larger_surface_columns(mods):
    """1,280 panels on themselves (nunk 3, 3,840 x 3,840): every column of a
    subset against the oracle's FP64 tree."""
    _lib, infl = mods
    v, f = scenes.icosphere(3, 1.0)
    S = oracle.surface(np.ascontiguousarray(v.T, np.float32), f.astype(np.uint32))
    g = infl.bem_matrix(S[:3], None, 3, -1)
    m64 = oracle.bem_matrix("wl", S, S, 3, True, 3, mode="f64t")
    assert nerr(g, m64) <= 1e-6


def test_rejects_bad_arguments(mods):
    _lib, infl = mods
    with pytest.raises(_lib.O3DError):
        infl.bem_matrix(A[:3], B[:3], 1, -1)
    with pytest.raises(_lib.O3DError):
        infl.bem_matrix(A[:3], B[:3], 2, 5)
