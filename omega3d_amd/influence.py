"""Python mirror of Omega3D's influence interface for the direct Biot-Savart path.

Reference shapes (Influence.h, ExecEnv.h, ResultsType.h, ElementBase.h, Points.h):
  points_affect_points(src, targ, restype, env)   Influence.h:67-551
  panels_affect_points(src, targ, restype, env)   Influence.h:557-1099
  ExecEnv / accel_t                               ExecEnv.h:26-123
  ResultsType / results_t                         ResultsType.h:27-95
  zero_vels / finalize_vels (1/4pi, freestream)   ElementBase.h:170-208, Points.h:252-277

Every call goes to the gfx950 kernels through libo3dbiot.so; there is no CPU
path.  Results accumulate into the target's vel/grad arrays, unscaled, exactly
as in the reference; finalize_vels applies 1/(4 pi) and the freestream.

Array conventions: pos (3, n), str (3, n), rad (n,), vel (3, n), grad (9, n)
with grad[3*j + i] = du_i/dx_j; panels as in scenes.Panels.
"""
import ctypes as C
import enum
import math

import numpy as np

from . import _lib
from ._lib import check, lib, make_opts


class accel_t(enum.IntEnum):
    cpu_x86 = 1
    cpu_vc = 2
    gpu_opengl = 3
    gpu_cuda = 4
    gpu_hip = 5      # this library


class results_t(enum.IntEnum):
    velonly = 1
    velandgrad = 2
    psionly = 3
    velandvort = 4
    velandshear = 5


class ResultsType:
    def __init__(self, rtype=results_t.velonly):
        self.rtype = results_t(rtype)

    def compute_grad(self):
        return self.rtype == results_t.velandgrad

    def get_type(self):
        return self.rtype


class ExecEnv:
    """Execution environment; the default accelerator here is gpu_hip.  `core`,
    `precision`, `self_term`, `tile`, `tpt` carry what the reference fixes at
    compile time or does not have."""

    def __init__(self, internal=True, useomp=True, summation=1, accel=accel_t.gpu_hip,
                 core="wl", precision="comp", self_term=0, tile=0, tpt=0, maxlev=-1):
        self.internal, self.useomp, self.summation, self.accel = internal, useomp, summation, accel_t(accel)
        self.core, self.precision, self.self_term = core, precision, self_term
        self.tile, self.tpt, self.maxlev = tile, tpt, maxlev

    def to_string(self):
        return " HIP-accelerated direct sums" if self.accel == accel_t.gpu_hip else " unsupported accelerator"


class Points:
    """SoA particles or field points.  rad None = inert field points (point targets);
    grad is allocated when with_grad (the reference gives inert+lagrangian points no
    gradient storage, Points.h:131-139)."""

    def __init__(self, pos, str_=None, rad=None, with_grad=True, dtype=np.float32):
        self.pos = np.ascontiguousarray(pos, dtype=dtype)
        n = self.pos.shape[1]
        self.str = None if str_ is None else np.ascontiguousarray(str_, dtype=dtype)
        self.rad = None if rad is None else np.ascontiguousarray(rad, dtype=dtype)
        self.vel = np.zeros((3, n), dtype)
        self.grad = np.zeros((9, n), dtype) if with_grad else None

    @property
    def n(self):
        return self.pos.shape[1]

    def is_inert(self):
        return self.rad is None

    def zero_vels(self):
        self.vel[:] = 0
        if self.grad is not None:
            self.grad[:] = 0

    def finalize_vels(self, fs=(0.0, 0.0, 0.0)):
        # ElementBase.h:187-208 (velocity, factor in double) and Points.h:265-277
        # (gradient, factor as the storage type)
        factor = 0.25 / math.pi
        for d in range(3):
            self.vel[d] = (fs[d] + self.vel[d].astype(np.float64) * factor).astype(self.vel.dtype)
        if self.grad is not None:
            self.grad *= self.grad.dtype.type(factor)


def _pts_struct(pos, str_, rad, vel, grad, cls=_lib.Points):
    p = cls()
    p.n = pos.shape[1]
    for d in range(3):
        p.x[d] = pos[d].ctypes.data
        if str_ is not None:
            p.s[d] = str_[d].ctypes.data
        if vel is not None:
            p.u[d] = vel[d].ctypes.data
    p.r = rad.ctypes.data if rad is not None else None
    if grad is not None:
        for k in range(9):
            p.ug[k] = grad[k].ctypes.data
    return p


def _c32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _c64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def p2p(core, spos, sstr, srad, tpos, trad=None, grads=True, precision="comp", self_term=0,
        tile=0, tpt=0, vel=None, grad=None, stats=False, method="auto"):
    """Host-array point->point sum on the GPU.  Accumulates into vel/grad (new zero
    arrays when not given) and returns them (and the o3d_stats when stats=True)."""
    f = _c64 if precision == "fp64" else _c32
    spos, sstr, srad, tpos, trad = f(spos), f(sstr), f(srad), f(tpos), f(trad)
    dt = spos.dtype
    nt = tpos.shape[1]
    vel = np.zeros((3, nt), dt) if vel is None else vel
    grad = (np.zeros((9, nt), dt) if grad is None else grad) if grads else None
    o = make_opts(core, grads, precision, self_term, tile, tpt, method=method)
    st = _lib.Stats()
    if dt == np.float64:
        s = _pts_struct(spos, sstr, srad, None, None, _lib.Points64)
        t = _pts_struct(tpos, None, trad, vel, grad, _lib.Points64)
        check(lib.o3d_points_affect_points64(C.byref(s), C.byref(t), C.byref(o), C.byref(st)))
    else:
        s = _pts_struct(spos, sstr, srad, None, None)
        t = _pts_struct(tpos, None, trad, vel, grad)
        check(lib.o3d_points_affect_points(C.byref(s), C.byref(t), C.byref(o), C.byref(st)))
    return (vel, grad, st) if stats else (vel, grad)


def ptpan(spos, sstr, nodes, idx, area, precision="comp", maxlev=-1, vel=None, stats=False):
    """Host-array particles -> panels (points_affect_panels, Influence.h:1107-1221):
    panel velocities, decremented by the particles' influence (new zero arrays when
    vel is not given)."""
    spos, sstr, nodes, area = _c32(spos), _c32(sstr), _c32(nodes), _c32(area)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    np_ = area.shape[0]
    vel = np.zeros((3, np_), np.float32) if vel is None else vel
    s = _lib.Points()
    s.n = spos.shape[1]
    for d in range(3):
        s.x[d] = spos[d].ctypes.data
        s.s[d] = sstr[d].ctypes.data
    p = _lib.Panels()
    p.np, p.nn = np_, nodes.shape[1]
    for d in range(3):
        p.x[d] = nodes[d].ctypes.data
    p.idx, p.area = idx.ctypes.data, area.ctypes.data
    vp = (C.c_void_p * 3)(*[vel[d].ctypes.data for d in range(3)])
    o = make_opts("rm", False, precision, 0, 0, 0, maxlev)
    st = _lib.Stats()
    check(lib.o3d_points_affect_panels(C.byref(s), C.byref(p), vp, C.byref(o), C.byref(st)))
    return (vel, st) if stats else vel


def bem_opts(fs, body=None, nunk=2):
    """o3d_bem_rhs_opts: body = None (no body, or the ground) or a dict with vel,
    rotvel (double 3-vectors) and tc (the surface's transformed centre)"""
    b = _lib.BemRhsOpts()
    for d in range(3):
        b.fs[d] = float(fs[d])
    if body is not None:
        b.body = 1
        for d in range(3):
            b.vel[d] = float(body.get("vel", (0, 0, 0))[d])
            b.rotvel[d] = float(body.get("rotvel", (0, 0, 0))[d])
            b.tc[d] = float(body.get("tc", (0, 0, 0))[d])
    b.nunk = int(nunk)
    return b


def bem_rhs(vorts, nodes, idx, area, fs, body=None, nunk=2, precision="comp", maxlev=-1):
    """BEM right-hand side of one boundary surface (BEMHelper.h:83-117) on the GPU:
    vorts is a list of (pos, str) particle collections; returns the panel-centre
    velocities (finalized, body motion removed) and the rhs (nunk * P,)."""
    D = _lib.DeviceArray.from_host
    nodes, area = _c32(nodes), _c32(area)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    P = area.shape[0]
    dn, di, da = D(nodes), D(idx.view(np.float32)), D(area)
    pan = _lib.Panels()
    pan.np, pan.nn = P, nodes.shape[1]
    for d in range(3):
        pan.x[d] = dn.row(d)
    pan.idx, pan.area = di.ptr, da.ptr
    keep, pts = [], (_lib.Points * max(1, len(vorts)))()
    for k, (pos, st) in enumerate(vorts):
        dp, ds = D(_c32(pos)), D(_c32(st))
        keep += [dp, ds]
        pts[k].n = pos.shape[1]
        for d in range(3):
            pts[k].x[d] = dp.row(d)
            pts[k].s[d] = ds.row(d)
    dpu = _lib.DeviceArray((3, max(P, 1)))
    drhs = _lib.DeviceArray((max(nunk * P, 1),))
    pu = (C.c_void_p * 3)(*[dpu.row(d) for d in range(3)])
    o = make_opts("rm", False, precision, 0, 0, 0, maxlev)
    b = bem_opts(fs, body, nunk)
    check(lib.o3d_bem_rhs_dev(pts, len(vorts), C.byref(pan), C.byref(b), C.byref(o), pu, C.c_void_p(drhs.ptr)))
    _lib.synchronize()
    return dpu.to_host()[:, :P], drhs.to_host()[:nunk * P]


def _surface_struct(nodes, idx, area):
    p = _lib.Panels()
    p.np, p.nn = area.shape[0], nodes.shape[1]
    for d in range(3):
        p.x[d] = nodes[d].ctypes.data
    p.idx, p.area = idx.ctypes.data, area.ctypes.data
    return p


def bem_matrix(src, targ=None, nunk=2, maxlev=-1):
    """BEM influence matrix block (panels_on_panels_coeff, Coefficients.h:169-444)
    on the GPU.  src, targ: (nodes (3, N), idx (P, 3), area (P,)) surfaces; targ
    None means the surface on itself (the reference's &src == &targ).  Returns a
    (nunk * targ panels, nunk * src panels) float32 array [row, column]."""
    same = targ is None
    sn, si, sa = _c32(src[0]), np.ascontiguousarray(src[1], dtype=np.uint32), _c32(src[2])
    tn, ti, ta = (sn, si, sa) if same else (_c32(targ[0]), np.ascontiguousarray(targ[1], dtype=np.uint32),
                                            _c32(targ[2]))
    ps, pt = _surface_struct(sn, si, sa), _surface_struct(tn, ti, ta)
    ns, nt = sa.shape[0], ta.shape[0]
    out = np.zeros(nunk * nunk * ns * nt, np.float32)
    check(lib.o3d_bem_matrix(C.byref(ps), C.byref(pt), int(nunk), int(same), int(maxlev),
                             out.ctypes.data_as(C.c_void_p)))
    return out.reshape(nunk * ns, nunk * nt).T


def pan(core, nodes, idx, ts, area, ss, tpos, grads=True, precision="comp", maxlev=-1,
        vel=None, grad=None, stats=False, rule="recursive", tile=0):
    """Host-array panel->point sum on the GPU: rule "recursive" is the reference's
    centroid subdivision, "gauss7" the 7-point rule with the closed-form near field."""
    nodes, ts, area, ss, tpos = _c32(nodes), _c32(ts), _c32(area), _c32(ss), _c32(tpos)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    nt = tpos.shape[1]
    vel = np.zeros((3, nt), np.float32) if vel is None else vel
    grad = (np.zeros((9, nt), np.float32) if grad is None else grad) if grads else None
    p = _lib.Panels()
    p.np = area.shape[0]
    p.nn = nodes.shape[1]
    for d in range(3):
        p.x[d] = nodes[d].ctypes.data
        p.ts[d] = ts[d].ctypes.data
    p.idx = idx.ctypes.data
    p.area = area.ctypes.data
    p.ss = ss.ctypes.data if ss is not None else None
    t = _pts_struct(tpos, None, None, vel, grad)
    o = make_opts(core, grads, precision, 0, tile, 0, maxlev, method={"recursive": "auto", "gauss7": "gauss7"}[rule])
    st = _lib.Stats()
    check(lib.o3d_panels_affect_points(C.byref(p), C.byref(t), C.byref(o), C.byref(st)))
    return (vel, grad, st) if stats else (vel, grad)


def pap(core, nodes, idx, ts, area, ss, tnodes, tidx, precision="comp", maxlev=-1, vel=None, stats=False):
    """Host-array panels -> panels on the GPU (panels_affect_panels, Influence.h:1225-1245):
    the source panels' velocity at the target surface's colocation points (centroids
    pushed 1e-4 along the normals), velocity only; vel (3, Pt) is overwritten (new
    when not given) and returned."""
    nodes, ts, area, ss, tnodes = _c32(nodes), _c32(ts), _c32(area), _c32(ss), _c32(tnodes)
    idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    tidx = np.ascontiguousarray(tidx, dtype=np.uint32).reshape(-1)
    npt = tidx.shape[0] // 3
    vel = np.zeros((3, npt), np.float32) if vel is None else vel
    p = _lib.Panels()
    p.np, p.nn = area.shape[0], nodes.shape[1]
    for d in range(3):
        p.x[d] = nodes[d].ctypes.data
        p.ts[d] = ts[d].ctypes.data
    p.idx, p.area = idx.ctypes.data, area.ctypes.data
    p.ss = ss.ctypes.data if ss is not None else None
    t = _lib.Panels()
    t.np, t.nn = npt, tnodes.shape[1]
    for d in range(3):
        t.x[d] = tnodes[d].ctypes.data
    t.idx = tidx.ctypes.data
    vp = (C.c_void_p * 3)(*[vel[d].ctypes.data for d in range(3)])
    o = make_opts(core, False, precision, 0, 0, 0, maxlev)
    st = _lib.Stats()
    check(lib.o3d_panels_affect_panels(C.byref(p), C.byref(t), vp, C.byref(o), C.byref(st)))
    return (vel, st) if stats else vel


def points_affect_points(src: Points, targ: Points, restype: ResultsType, env: ExecEnv):
    """Influence.h:67-551 on the GPU: accumulate src's influence into targ."""
    grads = restype.compute_grad()
    if grads and targ.grad is None:
        raise _lib.O3DError(1, "targets have no velocity-gradient storage (Influence.h:369)")
    if restype.get_type() not in (results_t.velonly, results_t.velandgrad):
        raise _lib.O3DError(1, f"results type {restype.get_type().name} is not served by this path")
    if env.accel != accel_t.gpu_hip:
        raise _lib.O3DError(1, f"accelerator {env.accel.name} is not provided by this library")
    p2p(env.core, src.pos, src.str, src.rad, targ.pos, targ.rad, grads, env.precision,
        env.self_term, env.tile, env.tpt, vel=targ.vel, grad=targ.grad if grads else None)


def panels_affect_points(src, targ: Points, restype: ResultsType, env: ExecEnv):
    """Influence.h:557-1099 on the GPU (point kernels for every target)."""
    grads = restype.compute_grad() and targ.grad is not None
    pan(env.core, src.nodes, src.idx, src.ts, src.area, src.ss, targ.pos, grads, env.precision,
        env.maxlev, vel=targ.vel, grad=targ.grad if grads else None)


def external_vel_solver_f(spos, sstr, srad, tpos, vel, grad):
    """The reference's external-solver ABI (Influence.h:33-48) through ctypes."""
    ns = C.c_int(spos.shape[1])
    nt = C.c_int(tpos.shape[1])
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    return lib.external_vel_solver_f_(C.byref(ns), P(spos[0]), P(spos[1]), P(spos[2]), P(sstr[0]), P(sstr[1]),
                                      P(sstr[2]), P(srad), C.byref(nt), P(tpos[0]), P(tpos[1]), P(tpos[2]),
                                      P(vel[0]), P(vel[1]), P(vel[2]), *[P(grad[k]) for k in range(9)])
