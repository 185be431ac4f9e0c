"""The factored far-field identity used by the matrix-core kernel, checked on the CPU
in float64 against the direct pair sums (Kernels.h:155-193 with the RM core and the
far branch of the compact-exponential core, CoreFunc.h:65-71, 121, 169).

The kernel (omega3d_amd/csrc/ff_kernels.hpp) sums, per target, 29 moment columns of
each source weighted by Bp = r2^-5/2 and recovers the velocity and the gradient
from them.  This test builds the same columns in the same order, applies the same
recovery, and requires agreement with the direct sums to 1e-12: a wrong column, a
dropped term, a sign or a transposed index fails it.
"""
import numpy as np
import pytest


def _cross(a, b):
    return np.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def _direct(t, tt, s, ss, a, core):
    """u = sum K (a x d), D_ji = sum Bp d_j (a x d)_i, W = sum K a with d = t - s."""
    d = t[:, :, None] - s[:, None, :]
    d2 = (d * d).sum(0)
    r2 = d2 + (tt[:, None] + ss[None, :] if core == "rm" else 0.0)
    K = r2 ** -1.5
    Bp = K / r2
    c = _cross(a[:, None, :], d)
    u = (K[None] * c).sum(2)
    D = np.einsum("jts,its->jit", d * Bp[None], c)
    W = (K[None] * a[:, None, :]).sum(2)
    return u, D, W


def _columns(sp, ss, a, core):
    """the 29 moment columns of each source, in the kernel's order"""
    m = _cross(a, sp)
    b = (sp * sp).sum(0) + (ss if core == "rm" else 0.0)
    cols = [a[0], a[1], a[2], m[0], m[1], m[2]]
    cols += [sp[l] * a[k] for l in range(3) for k in range(3)]
    cols += [sp[e // 3] * m[e % 3] for e in range(8)]
    cols += [b * a[k] for k in range(3)] + [b * m[k] for k in range(3)]
    return np.stack(cols)


def _recover(c, t, al):
    """velocity, dyad D and W of each target from its 29 moment sums c (29, nt)"""
    P, Q = c[0:3], c[3:6]
    R = c[6:15].reshape(3, 3, -1)
    S = np.zeros((3, 3, c.shape[1]))
    for e in range(8):
        S[e // 3, e % 3] = c[15 + e]
    S[2, 2] = -S[0, 0] - S[1, 1]
    U, V = c[23:26], c[26:29]
    W = al * P - 2 * np.einsum("lt,lkt->kt", t, R) + U
    M = al * Q - 2 * np.einsum("lt,lkt->kt", t, S) + V
    u = _cross(W, t) - M
    Px = _cross(P, t)
    D = np.empty((3, 3, t.shape[1]))
    for j in range(3):
        D[j] = t[j] * Px - t[j] * Q - _cross(R[j], t) + S[j]
    return u, D, W


@pytest.mark.parametrize("core", ["rm", "exp"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_factored_identity(core, seed):
    rng = np.random.default_rng(seed)
    nt, ns = 40, 70
    c = rng.uniform(-1, 1, 3)
    t = c[:, None] + rng.uniform(-0.1, 0.1, (3, nt))            # target block, radius ~0.17
    s = c[:, None] + np.array([[0.9], [0.3], [-0.4]]) + rng.uniform(-0.15, 0.15, (3, ns))
    tt = rng.uniform(0.01, 0.05, nt) ** 2
    ss = rng.uniform(0.01, 0.05, ns) ** 2
    a = rng.normal(size=(3, ns))
    u0, D0, W0 = _direct(t, tt, s, ss, a, core)
    tp, sp = t - c[:, None], s - c[:, None]
    d = tp[:, :, None] - sp[:, None, :]
    r2 = (d * d).sum(0) + (tt[:, None] + ss[None, :] if core == "rm" else 0.0)
    Bp = r2 ** -2.5
    sums = _columns(sp, ss, a, core) @ Bp.T                      # (29, nt)
    al = (tp * tp).sum(0) + (tt if core == "rm" else 0.0)
    u1, D1, W1 = _recover(sums, tp, al)
    for x0, x1 in ((u0, u1), (D0, D1), (W0, W1)):
        assert np.abs(x1 - x0).max() <= 1e-12 * np.abs(x0).max()


def test_identity_detects_a_dropped_term():
    """the check above is sharp: dropping one moment column (U) changes W visibly"""
    rng = np.random.default_rng(5)
    t = rng.uniform(-0.1, 0.1, (3, 8))
    s = np.array([[1.0], [0.0], [0.0]]) + rng.uniform(-0.1, 0.1, (3, 9))
    ss = np.full(9, 1e-3)
    a = rng.normal(size=(3, 9))
    _, _, W0 = _direct(t, np.full(8, 1e-3), s, ss, a, "rm")
    d = t[:, :, None] - s[:, None, :]
    Bp = ((d * d).sum(0) + 2e-3) ** -2.5
    cols = _columns(s, ss, a, "rm")
    cols[23:26] = 0.0
    _, _, W1 = _recover(cols @ Bp.T, t, (t * t).sum(0) + 1e-3)
    assert np.abs(W1 - W0).max() > 1e-3 * np.abs(W0).max()
