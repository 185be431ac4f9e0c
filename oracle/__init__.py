"""THE ORACLE -- test infrastructure only.

ctypes front-end over ``liboracle.so`` (a CPU restatement of Omega3D's direct
Biot-Savart path, see ``o3d_oracle.cpp``) and, when it has been built, over the
reference's own scalar kernels in ``_ref/libo3dref_<core>.so``.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker / the timed CPU baseline.
The product (``omega3d_amd``) never imports it.

Array conventions (shared with the product's Python mirror):
  source points : pos (3, N) float32, str (3, N) float32, rad (N,) float32
  targets       : pos (3, M) float32, rad (M,) float32 or None for point targets
  results       : vel (3, M), grad (9, M) with grad[3*j + i] = du_i/dx_j,
                  accumulated and unscaled (no 1/4pi), as Influence.h:296-307.
  panels        : nodes (3, Nn) float32, idx (3*P,) uint32, ts (3, P) absolute
                  vortex strength, area (P,), ss (P,) sheet source strength or None.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CORES = {"rm": 1, "wl": 2, "exp": 3, "v2": 4}
RECURSIVE_LEVELS = 3  # Influence.h:23

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i64 = C.c_int64


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (run make -C oracle)")
    return C.CDLL(path)


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load(os.path.join(HERE, "liboracle.so"))
        for name in ("orc_p2p_f", "orc_p2p_d", "orc_pan_f", "orc_pan_d", "orc_p2p_vcstyle", "orc_tri_integrals",
                     "orc_pan_g7_near", "orc_pan_g7_d", "orc_bem_rhs", "orc_bem_bases", "orc_bem_matrix",
                     "orc_colocation"):
            getattr(_lib, name).restype = C.c_int
        _lib.orc_flops_pair.restype = C.c_int
        _lib.orc_flops_pair.argtypes = [C.c_int] * 4
        _lib.orc_max_threads.restype = C.c_int
        _lib.orc_have_avx2.restype = C.c_int
    return _lib


def ref_path(core):
    return os.path.join(HERE, "_ref", f"libo3dref_{core}.so")


_refs = {}


def ref_lib(core):
    """The reference's own kernels for ``core`` (None when oracle/_ref is not built)."""
    if core not in _refs:
        p = ref_path(core)
        _refs[core] = C.CDLL(p) if os.path.exists(p) else None
        if _refs[core] is not None:
            for name in ("ref_p2p_f", "ref_p2p_d", "ref_pan_f", "ref_pan_d", "ref_flops", "ref_bem_matrix_f",
                         "ref_bem_matrix_d"):
                getattr(_refs[core], name).restype = C.c_int
    return _refs[core]


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def _ptrs9(g):
    if g is None:
        return None
    arr = (_vp * 9)(*[g[k].ctypes.data for k in range(9)])
    return arr


def _core_id(core):
    return CORES[core] if isinstance(core, str) else int(core)


def _as(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def _prep(dt, spos, sstr, srad, tpos, trad, grads):
    spos, sstr, srad = _as(spos, dt), _as(sstr, dt), _as(srad, dt)
    tpos, trad = _as(tpos, dt), _as(trad, dt)
    nt = tpos.shape[1]
    vel = np.zeros((3, nt), dt)
    grad = np.zeros((9, nt), dt) if grads else None
    return spos, sstr, srad, tpos, trad, vel, grad


def flops_pair(core, grads=True, thick=True, vs=False):
    return lib().orc_flops_pair(_core_id(core), int(grads), int(thick), int(vs))


def p2p(core, spos, sstr, srad, tpos, trad=None, grads=True, mode="f32d", nthreads=0):
    """points_affect_points restated.  mode: 'f32f' (S=A=float), 'f32d' (S=float,
    A=double), 'f64' (FP64 oracle: every operation in double)."""
    cid = _core_id(core)
    if mode == "f64":
        spos, sstr, srad, tpos, trad, vel, grad = _prep(np.float64, spos, sstr, srad, tpos, trad, grads)
        rc = lib().orc_p2p_d(cid, int(grads), _i64(spos.shape[1]), _ptr(spos[0]), _ptr(spos[1]), _ptr(spos[2]),
                             _ptr(srad), _ptr(sstr[0]), _ptr(sstr[1]), _ptr(sstr[2]),
                             _i64(tpos.shape[1]), _ptr(tpos[0]), _ptr(tpos[1]), _ptr(tpos[2]), _ptr(trad),
                             _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2]), _ptrs9(grad), int(nthreads))
    else:
        m = {"f32f": 0, "f32d": 1}[mode]
        spos, sstr, srad, tpos, trad, vel, grad = _prep(np.float32, spos, sstr, srad, tpos, trad, grads)
        rc = lib().orc_p2p_f(cid, int(grads), m, _i64(spos.shape[1]), _ptr(spos[0]), _ptr(spos[1]), _ptr(spos[2]),
                             _ptr(srad), _ptr(sstr[0]), _ptr(sstr[1]), _ptr(sstr[2]),
                             _i64(tpos.shape[1]), _ptr(tpos[0]), _ptr(tpos[1]), _ptr(tpos[2]), _ptr(trad),
                             _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2]), _ptrs9(grad), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle p2p failed rc={rc}")
    return vel, grad


def p2p_vcstyle(core, spos, sstr, srad, tpos, trad, t0=0, t1=None, nthreads=0):
    """The Vc-style CPU baseline (thick targets, velocity + gradient) over targets [t0, t1)."""
    spos, sstr, srad, tpos, trad, vel, grad = _prep(np.float32, spos, sstr, srad, tpos, trad, True)
    nt = tpos.shape[1]
    t1 = nt if t1 is None else t1
    rc = lib().orc_p2p_vcstyle(_core_id(core), _i64(spos.shape[1]), _ptr(spos[0]), _ptr(spos[1]), _ptr(spos[2]),
                               _ptr(srad), _ptr(sstr[0]), _ptr(sstr[1]), _ptr(sstr[2]),
                               _i64(nt), _ptr(tpos[0]), _ptr(tpos[1]), _ptr(tpos[2]), _ptr(trad),
                               _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2]), _ptrs9(grad),
                               _i64(t0), _i64(t1), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle vcstyle failed rc={rc}")
    return vel, grad


def pan(core, nodes, idx, ts, area, ss, tpos, grads=True, mode="f32d", maxlev=RECURSIVE_LEVELS, nthreads=0):
    """panels_affect_points restated (recursive centroid rule, Kernels.h:1120-1211)."""
    cid = _core_id(core)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    dt = np.float64 if mode == "f64" else np.float32
    nodes, ts, area, ss, tpos = _as(nodes, dt), _as(ts, dt), _as(area, dt), _as(ss, dt), _as(tpos, dt)
    nt = tpos.shape[1]
    vel = np.zeros((3, nt), dt)
    grad = np.zeros((9, nt), dt) if grads else None
    args = [_i64(area.shape[0]), _ptr(nodes[0]), _ptr(nodes[1]), _ptr(nodes[2]), _ptr(idx),
            _ptr(ts[0]), _ptr(ts[1]), _ptr(ts[2]), _ptr(area), _ptr(ss),
            _i64(nt), _ptr(tpos[0]), _ptr(tpos[1]), _ptr(tpos[2]),
            _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2]), _ptrs9(grad), int(nthreads)]
    if mode == "f64":
        rc = lib().orc_pan_d(cid, int(grads), int(maxlev), *args)
    else:
        rc = lib().orc_pan_f(cid, int(grads), {"f32f": 0, "f32d": 1}[mode], int(maxlev), *args)
    if rc != 0:
        raise RuntimeError(f"oracle pan failed rc={rc}")
    return vel, grad


def ptpan(core, spos, sstr, nodes, idx, area, mode="f32d", maxlev=RECURSIVE_LEVELS, nthreads=0):
    """points_affect_panels restated (Influence.h:1190-1210): velocity of each panel,
    decremented by the panel kernel run backwards from every particle."""
    cid = _core_id(core)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    dt = np.float64 if mode == "f64" else np.float32
    spos, sstr, nodes, area = _as(spos, dt), _as(sstr, dt), _as(nodes, dt), _as(area, dt)
    np_ = area.shape[0]
    vel = np.zeros((3, np_), dt)
    args = [_i64(spos.shape[1]), _ptr(spos[0]), _ptr(spos[1]), _ptr(spos[2]),
            _ptr(sstr[0]), _ptr(sstr[1]), _ptr(sstr[2]), _i64(np_), _ptr(nodes[0]), _ptr(nodes[1]),
            _ptr(nodes[2]), _ptr(idx), _ptr(area), _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2]), int(nthreads)]
    if mode == "f64":
        rc = lib().orc_ptpan_d(cid, int(maxlev), *args)
    else:
        rc = lib().orc_ptpan_f(cid, {"f32f": 0, "f32d": 1}[mode], int(maxlev), *args)
    if rc != 0:
        raise RuntimeError(f"oracle ptpan failed rc={rc}")
    return vel


# ---- panels by the 7-point rule with the closed-form near field (o3d_panel_quad.cpp) ----

def g7_rule():
    """(bary (7, 3), weights (7,)) of the 7-point degree-5 triangle rule, weights summing to 1."""
    bary, w = np.zeros((7, 3)), np.zeros(7)
    lib().orc_g7_rule(bary.ctypes.data_as(_vp), w.ctypes.data_as(_vp))
    return bary, w


def tri_integrals(v, x, grads=True):
    """I = int_T (x - y)/|x - y|^3 dA and dI[j, l] = dI_l/dx_j over the triangle v (3, 3) (rows = vertices), FP64."""
    v = np.ascontiguousarray(v, np.float64).reshape(9)
    x = np.ascontiguousarray(x, np.float64).reshape(3)
    I, dI = np.zeros(3), np.zeros(9)
    lib().orc_tri_integrals(v.ctypes.data_as(_vp), x.ctypes.data_as(_vp), I.ctypes.data_as(_vp),
                            dI.ctypes.data_as(_vp) if grads else None)
    return I, dI.reshape(3, 3)


def pan_g7_near(v, t):
    """the float near/far decision for one (target, panel): True = closed form"""
    v = np.ascontiguousarray(v, np.float32).reshape(9)
    t = np.asarray(t, np.float32)
    return bool(lib().orc_pan_g7_near(v.ctypes.data_as(_vp), C.c_float(t[0]), C.c_float(t[1]), C.c_float(t[2])))


def pan_g7(nodes, idx, ts, area, ss, tpos, grads=True, nthreads=0):
    """panels -> points by the 7-point rule per panel with the closed-form near field
    (within twice the longest edge of the centroid), FP64; returns (vel, grad, near_pairs)."""
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    nodes, ts, area, ss, tpos = (_as(a, np.float32) for a in (nodes, ts, area, ss, tpos))
    nt = tpos.shape[1]
    vel = np.zeros((3, nt))
    grad = np.zeros((9, nt)) if grads else None
    nnear = C.c_int64(0)
    rc = lib().orc_pan_g7_d(int(grads), _i64(area.shape[0]), _ptr(nodes[0]), _ptr(nodes[1]), _ptr(nodes[2]),
                            _ptr(idx), _ptr(ts[0]), _ptr(ts[1]), _ptr(ts[2]), _ptr(area), _ptr(ss),
                            _i64(nt), _ptr(tpos[0]), _ptr(tpos[1]), _ptr(tpos[2]),
                            _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2]), _ptrs9(grad), C.byref(nnear), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle pan_g7 failed rc={rc}")
    return vel, grad, nnear.value


# ---- BEM right-hand side from panel velocities (o3d_bem.cpp) ------------------

def bem_rhs(nodes, idx, pu, fs, body=False, vel=(0, 0, 0), rotvel=(0, 0, 0), tc=(0, 0, 0), nunk=2):
    """finalize_vels + add_body_motion(-1) + vels_to_rhs_panels (Surfaces.h, RHS.h) on
    the panel velocities pu (3, P) as points_affect_panels left them; returns
    (finalized pu, rhs (nunk * P,)), float32"""
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    nodes = _as(nodes, np.float32)
    pu = np.array(pu, dtype=np.float32, copy=True)
    P = pu.shape[1]
    rhs = np.zeros(nunk * P, np.float32)
    d3 = lambda v: np.ascontiguousarray(v, np.float64)
    f3 = lambda v: np.ascontiguousarray(v, np.float32)
    fs_, vel_, rv_, tc_ = d3(fs), d3(vel), d3(rotvel), f3(tc)
    rc = lib().orc_bem_rhs(_i64(P), _ptr(nodes[0]), _ptr(nodes[1]), _ptr(nodes[2]), _ptr(idx),
                           _ptr(pu[0]), _ptr(pu[1]), _ptr(pu[2]), _ptr(fs_), int(bool(body)), _ptr(vel_), _ptr(rv_),
                           _ptr(tc_), int(nunk), _ptr(rhs))
    if rc != 0:
        raise RuntimeError(f"oracle bem_rhs failed rc={rc}")
    return pu, rhs


def bem_bases(nodes, idx):
    """Surfaces::compute_bases (Surfaces.h:766-816): (bases (9, P) -- x1, x2, normal
    rows --, area (P,)), float32"""
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    nodes = _as(nodes, np.float32)
    P = idx.shape[0]
    b = np.zeros((9, P), np.float32)
    area = np.zeros(P, np.float32)
    rc = lib().orc_bem_bases(_i64(P), _ptr(nodes[0]), _ptr(nodes[1]), _ptr(nodes[2]), _ptr(idx), _ptrs9(b),
                             _ptr(area))
    if rc != 0:
        raise RuntimeError(f"oracle bem_bases failed rc={rc}")
    return b, area


def colocation(nodes, idx, offset=1e-4):
    """Surfaces::represent_as_particles(offset, ...) positions (Surfaces.h:959-1010):
    each panel's centroid pushed offset along its normal, (3, P) float32"""
    idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    nodes = _as(nodes, np.float32)
    P = idx.shape[0] // 3
    out = np.zeros((3, P), np.float32)
    rc = lib().orc_colocation(_i64(P), _ptr(nodes[0]), _ptr(nodes[1]), _ptr(nodes[2]), _ptr(idx), C.c_float(offset),
                              _ptr(out[0]), _ptr(out[1]), _ptr(out[2]))
    if rc != 0:
        raise RuntimeError(f"oracle colocation failed rc={rc}")
    return out


def pap(core, nodes, idx, ts, area, ss, tnodes, tidx, mode="f32d", maxlev=RECURSIVE_LEVELS):
    """panels_affect_panels (Influence.h:1225-1245): the velocity of the source panels
    at the target surface's colocation points (offset 1e-4), velocity only, (3, Pt)"""
    v, _ = pan(core, nodes, idx, ts, area, ss, colocation(tnodes, tidx, 1e-4), False, mode, maxlev)
    return v


def surface(nodes, idx):
    """A boundary surface for bem_matrix: (nodes, idx, area, bases) with the area
    and bases of compute_bases."""
    idx = np.ascontiguousarray(np.asarray(idx).reshape(-1, 3), dtype=np.uint32)   # (panels, 3), flat accepted
    b, area = bem_bases(nodes, idx)
    return (_as(nodes, np.float32), idx, area, b)


def _bem_args(surf, dt):
    nodes, idx, area, b = surf
    nodes, area, b = _as(nodes, dt), _as(area, dt), _as(b, dt)
    return [nodes, idx, area, b], [_i64(idx.shape[0]), _ptr(nodes[0]), _ptr(nodes[1]), _ptr(nodes[2]), _ptr(idx),
                                   _ptr(area), _ptrs9(b)]


def bem_matrix(core, src, targ, nunk=2, same=False, maxlev=RECURSIVE_LEVELS, mode="f32", nthreads=0):
    """panels_on_panels_coeff (Coefficients.h:169-444) over rkernel_2vs_2p: the
    BEM influence matrix of surface src on surface targ (each from surface()), as a
    (nunk * targ panels, nunk * src panels) array [row, column].  mode 'f32': the
    reference's float arithmetic (float32); 'f64t': the same tree (decided in
    float) with FP64 leaves and sums (float64)."""
    keep_s, a_s = _bem_args(src, np.float32)
    keep_t, a_t = _bem_args(targ, np.float32)
    ns, nt = src[1].shape[0], targ[1].shape[0]
    m = {"f32": 0, "f64t": 1}[mode]
    out = np.zeros(nunk * nunk * ns * nt, np.float32 if m == 0 else np.float64)
    rc = lib().orc_bem_matrix(_core_id(core), m, int(maxlev), int(nunk), int(bool(same)), *a_s, *a_t,
                              _ptr(out) if m == 0 else None, _ptr(out) if m == 1 else None, int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle bem_matrix failed rc={rc}")
    return out.reshape(nunk * ns, nunk * nt).T


# ---- the reference's own kernels (oracle/_ref), same conventions -------------

def ref_p2p(core, spos, sstr, srad, tpos, trad=None, grads=True, mode="f32d"):
    L = ref_lib(core)
    if L is None:
        raise FileNotFoundError(ref_path(core))
    if mode == "f64":
        spos, sstr, srad, tpos, trad, vel, grad = _prep(np.float64, spos, sstr, srad, tpos, trad, grads)
        rc = L.ref_p2p_d(int(grads), _i64(spos.shape[1]), _ptr(spos[0]), _ptr(spos[1]), _ptr(spos[2]),
                         _ptr(srad), _ptr(sstr[0]), _ptr(sstr[1]), _ptr(sstr[2]),
                         _i64(tpos.shape[1]), _ptr(tpos[0]), _ptr(tpos[1]), _ptr(tpos[2]), _ptr(trad),
                         _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2]), _ptrs9(grad))
    else:
        spos, sstr, srad, tpos, trad, vel, grad = _prep(np.float32, spos, sstr, srad, tpos, trad, grads)
        rc = L.ref_p2p_f(int(grads), {"f32f": 0, "f32d": 1}[mode], _i64(spos.shape[1]),
                         _ptr(spos[0]), _ptr(spos[1]), _ptr(spos[2]),
                         _ptr(srad), _ptr(sstr[0]), _ptr(sstr[1]), _ptr(sstr[2]),
                         _i64(tpos.shape[1]), _ptr(tpos[0]), _ptr(tpos[1]), _ptr(tpos[2]), _ptr(trad),
                         _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2]), _ptrs9(grad))
    if rc != 0:
        raise RuntimeError(f"ref p2p failed rc={rc}")
    return vel, grad


def ref_pan(core, nodes, idx, ts, area, ss, tpos, grads=True, mode="f32d", maxlev=RECURSIVE_LEVELS):
    L = ref_lib(core)
    if L is None:
        raise FileNotFoundError(ref_path(core))
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    dt = np.float64 if mode == "f64" else np.float32
    nodes, ts, area, ss, tpos = _as(nodes, dt), _as(ts, dt), _as(area, dt), _as(ss, dt), _as(tpos, dt)
    nt = tpos.shape[1]
    vel = np.zeros((3, nt), dt)
    grad = np.zeros((9, nt), dt) if grads else None
    args = [_i64(area.shape[0]), _ptr(nodes[0]), _ptr(nodes[1]), _ptr(nodes[2]), _ptr(idx),
            _ptr(ts[0]), _ptr(ts[1]), _ptr(ts[2]), _ptr(area), _ptr(ss),
            _i64(nt), _ptr(tpos[0]), _ptr(tpos[1]), _ptr(tpos[2]),
            _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2]), _ptrs9(grad)]
    if mode == "f64":
        rc = L.ref_pan_d(int(grads), int(maxlev), *args)
    else:
        rc = L.ref_pan_f(int(grads), {"f32f": 0, "f32d": 1}[mode], int(maxlev), *args)
    if rc != 0:
        raise RuntimeError(f"ref pan failed rc={rc}")
    return vel, grad


def ref_ptpan(core, spos, sstr, nodes, idx, area, mode="f32d", maxlev=RECURSIVE_LEVELS):
    L = ref_lib(core)
    if L is None:
        raise FileNotFoundError(ref_path(core))
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    dt = np.float64 if mode == "f64" else np.float32
    spos, sstr, nodes, area = _as(spos, dt), _as(sstr, dt), _as(nodes, dt), _as(area, dt)
    np_ = area.shape[0]
    vel = np.zeros((3, np_), dt)
    args = [_i64(spos.shape[1]), _ptr(spos[0]), _ptr(spos[1]), _ptr(spos[2]),
            _ptr(sstr[0]), _ptr(sstr[1]), _ptr(sstr[2]), _i64(np_), _ptr(nodes[0]), _ptr(nodes[1]),
            _ptr(nodes[2]), _ptr(idx), _ptr(area), _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2])]
    if mode == "f64":
        rc = L.ref_ptpan_d(int(maxlev), *args)
    else:
        rc = L.ref_ptpan_f({"f32f": 0, "f32d": 1}[mode], int(maxlev), *args)
    if rc != 0:
        raise RuntimeError(f"ref ptpan failed rc={rc}")
    return vel


def ref_flops(core, which):
    return ref_lib(core).ref_flops(int(which))


def ref_bem_matrix(core, src, targ, nunk=2, same=False, maxlev=RECURSIVE_LEVELS, mode="f32"):
    """The same matrix through the reference's own rkernel_2vs_2p (S = float for
    'f32', S = double for 'f64', with its own double-precision tree)."""
    L = ref_lib(core)
    if L is None:
        raise FileNotFoundError(ref_path(core))
    dt = np.float32 if mode == "f32" else np.float64
    keep_s, a_s = _bem_args(src, dt)
    keep_t, a_t = _bem_args(targ, dt)
    ns, nt = src[1].shape[0], targ[1].shape[0]
    out = np.zeros(nunk * nunk * ns * nt, dt)
    f = L.ref_bem_matrix_f if mode == "f32" else L.ref_bem_matrix_d
    rc = f(int(maxlev), int(nunk), int(bool(same)), *a_s, *a_t, _ptr(out))
    if rc != 0:
        raise RuntimeError(f"reference bem_matrix failed rc={rc}")
    return out.reshape(nunk * ns, nunk * nt).T
