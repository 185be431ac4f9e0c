"""THE ORACLE (test infrastructure only): one inviscid Ralston RK2 convection step
with vortex stretching, restated from Omega3D for checking the served GPU step.

Reference (file:line):
  Convection::advect_2nd_ralston        Convection.h:349-425
  Convection::find_vels                 Convection.h:131-182  (zero, accumulate, finalize)
  ElementBase::finalize_vels            ElementBase.h:187-208 (u = fs + u/(4 pi), double)
  Points::finalize_vels                 Points.h:265-277      (ug *= (float)(1/(4 pi)))
  ElementBase::move, one weight         ElementBase.h:252-266 (x += (S)dt * wt1 * u)
  Points::move, one weight              Points.h:288-351      (stretch and elongation)
  ElementBase::move, two weights        ElementBase.h:276-304 (u = wt1 u1 + wt2 u2; x += (S)dt * u)
  Points::move, two weights             Points.h:356-439      (stretch with the ORIGINAL strength)

The pair sums come from the oracle's points_affect_points restatement (sources =
targets = the collection, self pair included as in the reference); the O(N)
updates follow the reference's float/double types operation by operation.  A test
may pass its own pair sums (``pair``) to check a large step whose N^2 sums the
CPU cannot afford: the O(N) updates are still this restatement's.
"""
import numpy as np

from . import p2p

F32, F64 = np.float32, np.float64


def _find_vels(core, pos, st, rad, fs, mode, pair=None):
    if pair is None:
        v, g = p2p(core, pos, st, rad, pos, rad, True, mode)
    else:
        v, g = pair(core, pos, st, rad)
    if mode == "f64":
        factor = 0.25 / np.pi
        return np.asarray(fs, F64)[:, None] + v * factor, g * factor
    factor = 0.25 / np.pi
    u = (np.asarray(fs, F64)[:, None] + v.astype(F64) * factor).astype(F32)
    ug = (g * F32(factor)).astype(F32)
    return u, ug


def _wdu(s, ug, T):
    # wdu_i = s0 ug[i] + s1 ug[3+i] + s2 ug[6+i], in the storage type
    return np.stack([(s[0] * ug[i] + s[1] * ug[3 + i]) + s[2] * ug[6 + i] for i in range(3)]).astype(T)


def rk2_step(core, pos, st, rad, elong, dt, fs=(0.0, 0.0, 0.0), mode="f32d", pair=None):
    """Returns (pos, str, elong) after one Ralston step.  mode 'f32d' restates the
    reference's float storage with double accumulation; 'f64' does every operation
    in double on the same inputs.  pair(core, pos, str, rad) -> (vel, grad), the
    unscaled sums of the collection on itself, replaces the oracle's own."""
    T = F64 if mode == "f64" else F32
    x0 = np.asarray(pos, T)
    s0 = np.asarray(st, T)
    r = np.asarray(rad, T)
    e0 = np.asarray(elong, T)
    dt23 = 2.0 / 3.0 * dt
    # stage 1
    u1, ug1 = _find_vels(core, x0, s0, r, fs, mode, pair)
    if mode == "f64":
        x1 = x0 + dt23 * 1.0 * u1
        s1 = s0 + dt23 * 1.0 * _wdu(s0, ug1, F64)
    else:
        x1 = (x0.astype(F64) + (F64(F32(dt23)) * 1.0) * u1.astype(F64)).astype(F32)
        w1 = _wdu(s0, ug1, F32)
        s1 = (s0.astype(F64) + dt23 * 1.0 * w1.astype(F64)).astype(F32)
    # stage 2 at the interim state
    u2, ug2 = _find_vels(core, x1, s1, r, fs, mode, pair)
    wdu1 = _wdu(s0, ug1, T)
    wdu2 = _wdu(s0, ug2, T)
    if mode == "f64":
        u = 0.25 * u1 + 0.75 * u2
        x = x0 + dt * u
        wdu = 0.25 * wdu1 + 0.75 * wdu2
        c2 = (s0 * s0).sum(0)
        dots = (s0 * wdu).sum(0)
        ef = np.where(c2 > 0, dt * dots / np.where(c2 > 0, c2, 1.0), 0.0)
        e = e0 * (1.0 + ef)
        s = s0 + dt * wdu
        return x, s, e
    u = (0.25 * u1.astype(F64) + 0.75 * u2.astype(F64)).astype(F32)
    x = (x0 + F32(dt) * u).astype(F32)
    wdu = (0.25 * wdu1.astype(F64) + 0.75 * wdu2.astype(F64)).astype(F32)
    c2 = ((s0[0] * s0[0] + s0[1] * s0[1]) + s0[2] * s0[2]).astype(F32)
    dots = ((s0[0] * wdu[0] + s0[1] * wdu[1]) + s0[2] * wdu[2]).astype(F32)
    safe = np.where(c2 > 0, c2, F32(1))
    ef = np.where(c2 > 0, (F32(dt) * dots) / safe, F32(0)).astype(F32)
    e = np.where(c2 > 0, (e0.astype(F64) * (1.0 + ef.astype(F64))).astype(F32), e0)
    s = (s0.astype(F64) + dt * wdu.astype(F64)).astype(F32)
    return x, s, e
