// o3d_oracle.cpp -- THE ORACLE.  Test infrastructure only.
//
// A clean-room CPU restatement of Omega3D's direct Biot-Savart path, used as
// the checker for the HIP kernels in omega3d_amd/ and as the timed CPU
// baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load this library; the product path never does.
//
// What it restates (reference file:line):
//   * the four regularised cores          CoreFunc.h:43-341
//   * the point->point pair kernels       Kernels.h:50-345
//   * the panel recursion                 Kernels.h:979-1211 (RECURSIVE_LEVELS=3, Influence.h:23)
//   * the scalar target loops             Influence.h:202-545 (points), 729-1091 (panels)
//   * the Vc/OpenMP loop shape            Influence.h:388-442, MathHelper.h:36-181
//
// Three arithmetic modes:
//   S=float,  A=float   "Vc-build accumulation, scalar math"   (Simulation.h:41-43)
//   S=float,  A=double  "non-Vc build"                          (Simulation.h:44-46)
//   S=double, A=double  "FP64 oracle": the same formulas evaluated in double on
//                       the float inputs (promoted exactly).
// The float modes follow the reference's operation order one operation at a
// time so that, compiled with -ffp-contract=off, they agree bit for bit with
// the reference's own scalar kernels (checked against oracle/_ref and the
// golden fixtures in tests/golden/).
//
// The Vc-style loop (orc_p2p_vcstyle) is the CPU performance baseline: 8 float
// lanes per source block with AVX2 rsqrt/rcp *estimates*, lane-wise float
// partial sums and one horizontal sum per target.  Vc 1.4.5 itself is not
// available, so the exact Vc results (estimate bit patterns, the .sum() order,
// Vc::exp) are "parity unpinned"; this loop is only timed and checked against
// the exact oracle within the estimates' tolerance.
//
// Parity status of each function is stated at its definition.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <immintrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum { C_RM = 1, C_WL = 2, C_EXP = 3, C_V2 = 4 };

// ---------------------------------------------------------------------------
// cores.  K is the velocity kernel ("r3" in the reference), B its gradient
// factor ("bbb"): u += K (s x d), du_i/dx_j += d_j B (s x d)_i + skew(K s).
// Each function lists the reference lines whose operation order it follows.
// ---------------------------------------------------------------------------

template <class S> inline S one() { return S(1); }

// Rosenhead-Moore, CoreFunc.h:47-82
template <class S> inline S rm_tv(S d2, S sr, S tr) {
  const S r2 = d2 + sr*sr + tr*tr;
  return S(1) / (r2*std::sqrt(r2));
}
template <class S> inline S rm_tp(S d2, S sr) {
  const S r2 = d2 + sr*sr;
  return S(1) / (r2*std::sqrt(r2));
}
template <class S> inline void rm_tvg(S d2, S sr, S tr, S& k, S& b) {
  const S r2 = d2 + sr*sr + tr*tr;
  k = S(1) / (r2*std::sqrt(r2));
  b = S(-3) * k * (S(1) / r2);
}
template <class S> inline void rm_tpg(S d2, S sr, S& k, S& b) {
  const S r2 = d2 + sr*sr;
  k = S(1) / (r2*std::sqrt(r2));
  b = S(-3) * k * (S(1) / r2);
}

// compact exponential, CoreFunc.h:119-237 (the non-Vc generic templates,
// which the Vc build's float specialisations at :98-107,144-154 equal)
template <class S> inline S exp_k(S ood3, S cf, S q) {
  if (q > S(16)) return ood3;
  if (q < S(0.001f)) return cf;
  return ood3 * (S(1) - std::exp(-q));
}
template <class S> inline S exp_b(S k, S cf, S q, S dist, S d2) {
  if (q > S(16)) return S(-3) * k / d2;
  if (q < S(0.001f)) return S(-1.5) * dist * k * k;
  const S e = std::exp(-q);
  return S(3) * (cf*e - k) / d2;
}
template <class S> inline S exp_tp(S d2, S sr) {
  const S dist = std::sqrt(d2);
  const S ood3 = S(1) / (d2*dist);
  const S cf = S(1) / (sr*sr*sr);
  const S q = cf / ood3;
  return exp_k(ood3, cf, q);
}
template <class S> inline S exp_tv(S d2, S sr, S tr) {
  const S dist = std::sqrt(d2);
  const S ood3 = S(1) / (d2*dist);
  const S cf = S(1) / (sr*sr*sr + tr*tr*tr);
  const S q = cf / ood3;
  return exp_k(ood3, cf, q);
}
template <class S> inline void exp_tpg(S d2, S sr, S& k, S& b) {
  const S dist = std::sqrt(d2);
  const S cf = S(1) / (sr*sr*sr);
  const S d3 = d2 * dist;
  const S q = d3 * cf;
  const S ood3 = S(1) / d3;
  k = exp_k(ood3, cf, q);
  b = exp_b(k, cf, q, dist, d2);
}
template <class S> inline void exp_tvg(S d2, S sr, S tr, S& k, S& b) {
  const S dist = std::sqrt(d2);
  const S cf = S(1) / (sr*sr*sr + tr*tr*tr);
  const S d3 = d2 * dist;
  const S q = d3 * cf;
  const S ood3 = S(1) / d3;
  k = exp_k(ood3, cf, q);
  b = exp_b(k, cf, q, dist, d2);
}

// Winckelmans-Leonard, CoreFunc.h:245-288; oor2p5 = 1/(x*x*sqrt(x)), MathHelper.h:132-134
template <class S> inline S oor2p5(S x) { return S(1) / (x*x*std::sqrt(x)); }
template <class S> inline S wl_tv(S d2, S sr, S tr) {
  const S r2 = sr*sr + tr*tr;
  const S dd = d2 + r2;
  return (d2 + S(2.5)*r2) * oor2p5(dd);
}
template <class S> inline S wl_tp(S d2, S sr) {
  const S r2 = sr*sr;
  const S dd = d2 + r2;
  return (d2 + S(2.5)*r2) * oor2p5(dd);
}
template <class S> inline void wl_g(S d2, S r2, S& k, S& b) {
  const S dd = d2 + r2;
  const S top = d2 + S(2.5)*r2;
  const S dn5 = oor2p5(dd);
  k = top * dn5;
  b = S(2)*dn5 - S(5)*top*dn5/dd;
}
template <class S> inline void wl_tvg(S d2, S sr, S tr, S& k, S& b) { wl_g(d2, sr*sr + tr*tr, k, b); }
template <class S> inline void wl_tpg(S d2, S sr, S& k, S& b) { wl_g(d2, sr*sr, k, b); }

// Vatistas n=2, CoreFunc.h:296-340; oor0p75 = 1/(sqrt(x)*sqrt(sqrt(x))), MathHelper.h:175-178
template <class S> inline S oor0p75(S x) { const S q = std::sqrt(x); return S(1) / (q*std::sqrt(q)); }
template <class S> inline S v2_tv(S d2, S sr, S tr) {
  const S s2 = sr*sr, t2 = tr*tr;
  return oor0p75(d2*d2 + s2*s2 + t2*t2);
}
template <class S> inline S v2_tp(S d2, S sr) {
  const S s2 = sr*sr;
  return oor0p75(d2*d2 + s2*s2);
}
template <class S> inline void v2_tvg(S d2, S sr, S tr, S& k, S& b) {
  const S s2 = sr*sr, t2 = tr*tr;
  const S den = d2*d2 + s2*s2 + t2*t2;
  k = oor0p75(den);
  b = S(-3) * k * (S(1) / std::sqrt(den));
}
template <class S> inline void v2_tpg(S d2, S sr, S& k, S& b) {
  const S s2 = sr*sr;
  const S den = d2*d2 + s2*s2;
  k = oor0p75(den);
  b = S(-3) * k * (S(1) / std::sqrt(den));
}

template <int CORE, class S> inline S core_tv(S d2, S sr, S tr) {
  if (CORE == C_RM) return rm_tv(d2, sr, tr);
  if (CORE == C_WL) return wl_tv(d2, sr, tr);
  if (CORE == C_EXP) return exp_tv(d2, sr, tr);
  return v2_tv(d2, sr, tr);
}
template <int CORE, class S> inline S core_tp(S d2, S sr) {
  if (CORE == C_RM) return rm_tp(d2, sr);
  if (CORE == C_WL) return wl_tp(d2, sr);
  if (CORE == C_EXP) return exp_tp(d2, sr);
  return v2_tp(d2, sr);
}
template <int CORE, class S> inline void core_tvg(S d2, S sr, S tr, S& k, S& b) {
  if (CORE == C_RM) rm_tvg(d2, sr, tr, k, b);
  else if (CORE == C_WL) wl_tvg(d2, sr, tr, k, b);
  else if (CORE == C_EXP) exp_tvg(d2, sr, tr, k, b);
  else v2_tvg(d2, sr, tr, k, b);
}
template <int CORE, class S> inline void core_tpg(S d2, S sr, S& k, S& b) {
  if (CORE == C_RM) rm_tpg(d2, sr, k, b);
  else if (CORE == C_WL) wl_tpg(d2, sr, k, b);
  else if (CORE == C_EXP) exp_tpg(d2, sr, k, b);
  else v2_tpg(d2, sr, k, b);
}

// ---------------------------------------------------------------------------
// pair kernels.  acc[0..2] = velocity, acc[3..11] = gradient in the order
// ux vx wx uy vy wy uz vz wz, i.e. acc[3+3j+i] = du_i/dx_j  (Kernels.h:155-193)
// THICK: target has a core radius (tv cores) or is a point (tp cores).
// VS:    vortex + scalar source strength (Kernels.h:115-133, 294-345).
// ---------------------------------------------------------------------------
template <int CORE, bool THICK, bool GRAD, bool VS, class S, class A>
inline void pair(S sx, S sy, S sz, S sr, S ssx, S ssy, S ssz, S ss,
                 S tx, S ty, S tz, S tr, A* acc) {
  const S dx = tx - sx;
  const S dy = ty - sy;
  const S dz = tz - sz;
  const S d2 = dx*dx + dy*dy + dz*dz;
  if (!GRAD) {
    const S k = THICK ? core_tv<CORE>(d2, sr, tr) : core_tp<CORE>(d2, sr);
    if (VS) {   // Kernels.h:115-133
      const S cx = dz*ssy - dy*ssz + dx*ss;
      const S cy = dx*ssz - dz*ssx + dy*ss;
      const S cz = dy*ssx - dx*ssy + dz*ss;
      acc[0] += k * cx; acc[1] += k * cy; acc[2] += k * cz;
    } else {    // Kernels.h:50-112
      const S cx = dz*ssy - dy*ssz;
      const S cy = dx*ssz - dz*ssx;
      const S cz = dy*ssx - dx*ssy;
      acc[0] += k * cx; acc[1] += k * cy; acc[2] += k * cz;
    }
    return;
  }
  S k, b;
  if (THICK) core_tvg<CORE>(d2, sr, tr, k, b); else core_tpg<CORE>(d2, sr, k, b);
  S cx = dz*ssy - dy*ssz;
  S cy = dx*ssz - dz*ssx;
  S cz = dy*ssx - dx*ssy;
  if (VS) {     // Kernels.h:320-322
    acc[0] += k * (cx + dx*ss);
    acc[1] += k * (cy + dy*ss);
    acc[2] += k * (cz + dz*ss);
  } else {      // Kernels.h:173-175
    acc[0] += k * cx;
    acc[1] += k * cy;
    acc[2] += k * cz;
  }
  cx *= b; cy *= b; cz *= b;          // Kernels.h:179-181
  acc[3]  += dx*cx;                   // Kernels.h:183-192
  acc[4]  += dx*cy + ssz*k;
  acc[5]  += dx*cz - ssy*k;
  acc[6]  += dy*cx - ssz*k;
  acc[7]  += dy*cy;
  acc[8]  += dy*cz + ssx*k;
  acc[9]  += dz*cx + ssy*k;
  acc[10] += dz*cy - ssx*k;
  acc[11] += dz*cz;
  if (VS) {     // Kernels.h:334-344
    const S px = dx*b*ss;
    const S py = dy*b*ss;
    const S pz = dz*b*ss;
    const S pk = ss*k;
    acc[3]  += dx*px + pk;
    acc[4]  += dx*py;
    acc[5]  += dx*pz;
    acc[6]  += dy*px;
    acc[7]  += dy*py + pk;
    acc[8]  += dy*pz;
    acc[9]  += dz*px;
    acc[10] += dz*py;
    acc[11] += dz*pz + pk;
  }
}

// ---------------------------------------------------------------------------
// point -> point loop (Influence.h:213-545, scalar branches).  For each target
// the accumulators start at zero, all sources are visited in index order, then
// the sums are added into the caller's (pre-zeroed, unscaled) outputs.
// ---------------------------------------------------------------------------
template <class T> struct Src { int64_t n; const T *x, *y, *z, *r, *sx, *sy, *sz; };
template <class T> struct Tgt { int64_t n; const T *x, *y, *z, *r; T *u, *v, *w; T* const* g; };

template <int CORE, bool THICK, bool GRAD, class S, class A, class T>
void p2p_loop(const Src<T>& s, const Tgt<T>& t, int nthreads) {
  const int64_t nt = t.n, ns = s.n;
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t i = 0; i < nt; ++i) {
    A acc[12] = {0};
    const S tx = S(t.x[i]), ty = S(t.y[i]), tz = S(t.z[i]);
    const S tr = THICK ? S(t.r[i]) : S(0);
    for (int64_t j = 0; j < ns; ++j)
      pair<CORE, THICK, GRAD, false, S, A>(S(s.x[j]), S(s.y[j]), S(s.z[j]), S(s.r[j]),
                                           S(s.sx[j]), S(s.sy[j]), S(s.sz[j]), S(0),
                                           tx, ty, tz, tr, acc);
    t.u[i] += acc[0]; t.v[i] += acc[1]; t.w[i] += acc[2];
    if (GRAD) for (int k = 0; k < 9; ++k) t.g[k][i] += acc[3 + k];
  }
}

template <class S, class A, class T>
int p2p_dispatch(int core, int grads, int thick, const Src<T>& s, const Tgt<T>& t, int nth) {
#define O3D_P2P_CASE(CID) \
  case CID: \
    if (grads) { if (thick) p2p_loop<CID, true, true, S, A>(s, t, nth); else p2p_loop<CID, false, true, S, A>(s, t, nth); } \
    else       { if (thick) p2p_loop<CID, true, false, S, A>(s, t, nth); else p2p_loop<CID, false, false, S, A>(s, t, nth); } \
    return 0;
  switch (core) {
    O3D_P2P_CASE(C_RM) O3D_P2P_CASE(C_WL) O3D_P2P_CASE(C_EXP) O3D_P2P_CASE(C_V2)
  }
#undef O3D_P2P_CASE
  return 1;
}

// ---------------------------------------------------------------------------
// panel -> point recursion (Kernels.h:1028-1211).  Strength arguments are
// sheet strengths at level 0 (converted to absolute with the area once), then
// absolute; each split quarters strength and area.  Leaves are singular
// (sigma = 0) vortex+source points at the sub-triangle centroid.
// ---------------------------------------------------------------------------
template <int CORE, bool GRAD, class S, class A>
void rpanel(S x0, S y0, S z0, S x1, S y1, S z1, S x2, S y2, S z2,
            S gx, S gy, S gz, S gs, S tx, S ty, S tz, S area, int lev, int maxlev, A* acc) {
  if (lev == 0) { gx *= area; gy *= area; gz *= area; gs *= area; }
  const S cx = (x0 + x1 + x2) / S(3);
  const S cy = (y0 + y1 + y2) / S(3);
  const S cz = (z0 + z1 + z2) / S(3);
  const S size = std::sqrt(area);
  const S ex = tx - cx, ey = ty - cy, ez = tz - cz;
  const S dist = std::sqrt(ex*ex + ey*ey + ez*ez);
  if (dist > size*S(4) || lev == maxlev) {
    pair<CORE, false, GRAD, true, S, A>(cx, cy, cz, S(0), gx, gy, gz, gs, tx, ty, tz, S(0), acc);
    return;
  }
  gx *= S(0.25); gy *= S(0.25); gz *= S(0.25); gs *= S(0.25);
  const S qa = S(0.25) * area;
  // six nodes: corners 0,2,5 and edge midpoints 1 (0-1), 3 (0-2), 4 (1-2)
  const S nx[6] = {x0, S(0.5)*(x0+x1), x1, S(0.5)*(x0+x2), S(0.5)*(x1+x2), x2};
  const S ny[6] = {y0, S(0.5)*(y0+y1), y1, S(0.5)*(y0+y2), S(0.5)*(y1+y2), y2};
  const S nz[6] = {z0, S(0.5)*(z0+z1), z1, S(0.5)*(z0+z2), S(0.5)*(z1+z2), z2};
  static const int child[4][3] = {{0,1,3}, {1,2,4}, {1,4,3}, {3,4,5}};
  for (int c = 0; c < 4; ++c) {
    const int a = child[c][0], b = child[c][1], d = child[c][2];
    rpanel<CORE, GRAD, S, A>(nx[a], ny[a], nz[a], nx[b], ny[b], nz[b], nx[d], ny[d], nz[d],
                             gx, gy, gz, gs, tx, ty, tz, qa, lev + 1, maxlev, acc);
  }
}

template <class T> struct Pan {
  int64_t np; const T *x, *y, *z; const uint32_t* idx;
  const T *tsx, *tsy, *tsz, *area, *ss;   // ts = absolute vortex strength, ss = sheet source strength (may be null)
};

// Influence.h:729-760 (scalar branch): sheet strength passed as ts/area, the
// recursion multiplies it back by the area at level 0.
template <int CORE, bool GRAD, class S, class A, class T>
void pan_loop(const Pan<T>& p, const Tgt<T>& t, int maxlev, int nthreads) {
  const int64_t nt = t.n, np = p.np;
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t i = 0; i < nt; ++i) {
    A acc[12] = {0};
    const S tx = S(t.x[i]), ty = S(t.y[i]), tz = S(t.z[i]);
    for (int64_t j = 0; j < np; ++j) {
      const uint32_t a = p.idx[3*j], b = p.idx[3*j+1], c = p.idx[3*j+2];
      const S ar = S(p.area[j]);
      rpanel<CORE, GRAD, S, A>(S(p.x[a]), S(p.y[a]), S(p.z[a]), S(p.x[b]), S(p.y[b]), S(p.z[b]),
                               S(p.x[c]), S(p.y[c]), S(p.z[c]),
                               S(p.tsx[j])/ar, S(p.tsy[j])/ar, S(p.tsz[j])/ar,
                               p.ss ? S(p.ss[j]) : S(0), tx, ty, tz, ar, 0, maxlev, acc);
    }
    t.u[i] += acc[0]; t.v[i] += acc[1]; t.w[i] += acc[2];
    if (GRAD) for (int k = 0; k < 9; ++k) t.g[k][i] += acc[3 + k];
  }
}

// points_affect_panels (Influence.h:1190-1210, scalar branch): the same recursion
// used backwards -- panel i carries particle j's strength divided by the panel's
// area, the "target" is the particle, and the sum is subtracted from the panel's
// velocity.  Velocity only (the reference asserts on gradients, :1111).
template <int CORE, class S, class A, class T>
void ptpan_loop(const Src<T>& s, const Pan<T>& p, T* u, T* v, T* w, int maxlev, int nthreads) {
  const int64_t np = p.np, ns = s.n;
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t i = 0; i < np; ++i) {
    A acc[12] = {0};
    const uint32_t a = p.idx[3*i], b = p.idx[3*i+1], c = p.idx[3*i+2];
    const S ar = S(p.area[i]);
    for (int64_t j = 0; j < ns; ++j)
      rpanel<CORE, false, S, A>(S(p.x[a]), S(p.y[a]), S(p.z[a]), S(p.x[b]), S(p.y[b]), S(p.z[b]),
                                S(p.x[c]), S(p.y[c]), S(p.z[c]),
                                S(s.sx[j])/ar, S(s.sy[j])/ar, S(s.sz[j])/ar, S(0),
                                S(s.x[j]), S(s.y[j]), S(s.z[j]), ar, 0, maxlev, acc);
    u[i] -= acc[0]; v[i] -= acc[1]; w[i] -= acc[2];
  }
}

template <class S, class A, class T>
int ptpan_dispatch(int core, const Src<T>& s, const Pan<T>& p, T* u, T* v, T* w, int maxlev, int nth) {
  switch (core) {
    case C_RM:  ptpan_loop<C_RM, S, A>(s, p, u, v, w, maxlev, nth); return 0;
    case C_WL:  ptpan_loop<C_WL, S, A>(s, p, u, v, w, maxlev, nth); return 0;
    case C_EXP: ptpan_loop<C_EXP, S, A>(s, p, u, v, w, maxlev, nth); return 0;
    case C_V2:  ptpan_loop<C_V2, S, A>(s, p, u, v, w, maxlev, nth); return 0;
  }
  return 1;
}

template <class S, class A, class T>
int pan_dispatch(int core, int grads, const Pan<T>& p, const Tgt<T>& t, int maxlev, int nth) {
#define O3D_PAN_CASE(CID) \
  case CID: if (grads) pan_loop<CID, true, S, A>(p, t, maxlev, nth); else pan_loop<CID, false, S, A>(p, t, maxlev, nth); return 0;
  switch (core) {
    O3D_PAN_CASE(C_RM) O3D_PAN_CASE(C_WL) O3D_PAN_CASE(C_EXP) O3D_PAN_CASE(C_V2)
  }
#undef O3D_PAN_CASE
  return 1;
}

// ---------------------------------------------------------------------------
// BEM influence matrix: panels_on_panels_coeff (Coefficients.h:169-444, scalar
// branch :332-397, self block :399-421, scaling :428-433) over the panel-panel
// recursion rkernel_2vs_2p (Kernels.h:1213-1311).  Per (source panel, target
// panel, source unknown) the recursion splits both triangles into 4 (16 child
// pairs, strength / 16, areas / 4) until the centroids are more than
// 4 (sqrt(sa) + sqrt(ta)) apart or maxlev is reached; a leaf is kernel_0vs_0p, a
// singular vortex + source point at the source centroid acting on the target
// centroid.  The tree is decided in float (the reference's S); the leaves run in
// L (float: the reference's arithmetic; double: the same tree in FP64).
// ---------------------------------------------------------------------------
template <int CORE, class L, class A>
void rpanpan(const float* sf, const float* tf, const L* sl, const L* tl, L gx, L gy, L gz, L gs, float saf,
             float taf, L sal, L tal, int lev, int maxlev, A* acc) {
  if (lev == 0) { gx *= sal; gy *= sal; gz *= sal; gs *= sal; }
  const float scx = (sf[0] + sf[3] + sf[6]) / 3.0f;
  const float scy = (sf[1] + sf[4] + sf[7]) / 3.0f;
  const float scz = (sf[2] + sf[5] + sf[8]) / 3.0f;
  const float tcx = (tf[0] + tf[3] + tf[6]) / 3.0f;
  const float tcy = (tf[1] + tf[4] + tf[7]) / 3.0f;
  const float tcz = (tf[2] + tf[5] + tf[8]) / 3.0f;
  const float trisize = std::sqrt(saf) + std::sqrt(taf);
  const float ex = tcx - scx, ey = tcy - scy, ez = tcz - scz;
  const float dist = std::sqrt(ex*ex + ey*ey + ez*ez);
  if (dist > trisize*4.0f || lev == maxlev) {
    const L sx = (sl[0] + sl[3] + sl[6]) / L(3), sy = (sl[1] + sl[4] + sl[7]) / L(3), sz = (sl[2] + sl[5] + sl[8]) / L(3);
    const L tx = (tl[0] + tl[3] + tl[6]) / L(3), ty = (tl[1] + tl[4] + tl[7]) / L(3), tz = (tl[2] + tl[5] + tl[8]) / L(3);
    pair<CORE, false, false, true, L, A>(sx, sy, sz, L(0), gx, gy, gz, gs, tx, ty, tz, L(0), acc);
    return;
  }
  gx *= L(0.0625); gy *= L(0.0625); gz *= L(0.0625); gs *= L(0.0625);
  const float scaf = 0.25f * saf, tcaf = 0.25f * taf;
  const L scal = L(0.25) * sal, tcal = L(0.25) * tal;
  // six nodes per triangle: corners 0, 2, 5 and edge midpoints 1 (0-1), 3 (0-2), 4 (1-2)
  float snf[18], tnf[18];
  L snl[18], tnl[18];
  for (int d = 0; d < 3; ++d) {
    snf[d] = sf[d]; snf[3+d] = 0.5f*(sf[d]+sf[3+d]); snf[6+d] = sf[3+d];
    snf[9+d] = 0.5f*(sf[d]+sf[6+d]); snf[12+d] = 0.5f*(sf[3+d]+sf[6+d]); snf[15+d] = sf[6+d];
    tnf[d] = tf[d]; tnf[3+d] = 0.5f*(tf[d]+tf[3+d]); tnf[6+d] = tf[3+d];
    tnf[9+d] = 0.5f*(tf[d]+tf[6+d]); tnf[12+d] = 0.5f*(tf[3+d]+tf[6+d]); tnf[15+d] = tf[6+d];
    snl[d] = sl[d]; snl[3+d] = L(0.5)*(sl[d]+sl[3+d]); snl[6+d] = sl[3+d];
    snl[9+d] = L(0.5)*(sl[d]+sl[6+d]); snl[12+d] = L(0.5)*(sl[3+d]+sl[6+d]); snl[15+d] = sl[6+d];
    tnl[d] = tl[d]; tnl[3+d] = L(0.5)*(tl[d]+tl[3+d]); tnl[6+d] = tl[3+d];
    tnl[9+d] = L(0.5)*(tl[d]+tl[6+d]); tnl[12+d] = L(0.5)*(tl[3+d]+tl[6+d]); tnl[15+d] = tl[6+d];
  }
  static const int child[4][3] = {{0,1,3}, {1,2,4}, {1,4,3}, {3,4,5}};
  for (int i = 0; i < 4; ++i) {
    float csf[9]; L csl[9];
    for (int v = 0; v < 3; ++v)
      for (int d = 0; d < 3; ++d) { csf[3*v+d] = snf[3*child[i][v]+d]; csl[3*v+d] = snl[3*child[i][v]+d]; }
    for (int j = 0; j < 4; ++j) {
      float ctf[9]; L ctl[9];
      for (int v = 0; v < 3; ++v)
        for (int d = 0; d < 3; ++d) { ctf[3*v+d] = tnf[3*child[j][v]+d]; ctl[3*v+d] = tnl[3*child[j][v]+d]; }
      rpanpan<CORE, L, A>(csf, ctf, csl, ctl, gx, gy, gz, gs, scaf, tcaf, scal, tcal, lev + 1, maxlev, acc);
    }
  }
}

// one boundary surface as the matrix builder sees it: nodes, panels, areas and
// the per-panel basis (x1, x2, normal) that Surfaces::compute_bases keeps
struct BemSurf {
  int64_t np; const float *x, *y, *z; const uint32_t* idx; const float* area;
  const float *b1[3], *b2[3], *nm[3];
};

// column-major block: column j*nunk + k (source panel j, unknown k: vortex along
// x1, along x2, source), row i*nunk + m (target panel i, component on x1, x2,
// normal).  O = float (L = A = float: the reference) or double.
template <int CORE, class L, class A, class O>
void bem_matrix_loop(const BemSurf& s, const BemSurf& t, int nunk, int same, int maxlev, O* coeffs, int nthreads) {
  const int64_t ns = s.np, nt = t.np, R = nunk * nt;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t j = 0; j < ns; ++j) {
    float sf[9]; L sl[9];
    for (int v = 0; v < 3; ++v) {
      const uint32_t n = s.idx[3*j+v];
      sf[3*v] = s.x[n]; sf[3*v+1] = s.y[n]; sf[3*v+2] = s.z[n];
      for (int d = 0; d < 3; ++d) sl[3*v+d] = L(sf[3*v+d]);
    }
    const float saf = s.area[j];
    for (int64_t i = 0; i < nt; ++i) {
      float tf[9]; L tl[9];
      for (int v = 0; v < 3; ++v) {
        const uint32_t n = t.idx[3*i+v];
        tf[3*v] = t.x[n]; tf[3*v+1] = t.y[n]; tf[3*v+2] = t.z[n];
        for (int d = 0; d < 3; ++d) tl[3*v+d] = L(tf[3*v+d]);
      }
      const float taf = t.area[i];
      for (int k = 0; k < nunk; ++k) {
        const L gx = k == 0 ? L(s.b1[0][j]) : k == 1 ? L(s.b2[0][j]) : L(0);
        const L gy = k == 0 ? L(s.b1[1][j]) : k == 1 ? L(s.b2[1][j]) : L(0);
        const L gz = k == 0 ? L(s.b1[2][j]) : k == 1 ? L(s.b2[2][j]) : L(0);
        const L gs = k == 2 ? L(1) : L(0);
        A r[3] = {0, 0, 0};
        rpanpan<CORE, L, A>(sf, tf, sl, tl, gx, gy, gz, gs, saf, taf, L(saf), L(taf), 0, maxlev, r);
        O* c = coeffs + (j*nunk + k) * R + i*nunk;
        c[0] = r[0]*A(t.b1[0][i]) + r[1]*A(t.b1[1][i]) + r[2]*A(t.b1[2][i]);
        c[1] = r[0]*A(t.b2[0][i]) + r[1]*A(t.b2[1][i]) + r[2]*A(t.b2[2][i]);
        if (nunk == 3) c[2] = r[0]*A(t.nm[0][i]) + r[1]*A(t.nm[1][i]) + r[2]*A(t.nm[2][i]);
      }
    }
    if (same) {
      // Coefficients.h:399-421: the panel on itself, set before the scaling
      const O tp = O(2.0 * M_PI);
      O* c = coeffs + (j*nunk) * R + j*nunk;
      c[0] = O(0); c[1] = tp; if (nunk == 3) c[2] = O(0);
      c += R;
      c[0] = -tp; c[1] = O(0); if (nunk == 3) c[2] = O(0);
      if (nunk == 3) { c += R; c[0] = O(0); c[1] = O(0); c[2] = tp; }
    }
  }
  const O fac = O(1.0 / (4.0 * M_PI));
  const int64_t total = R * (int64_t)(nunk * ns);
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t e = 0; e < total; ++e) coeffs[e] = coeffs[e] * fac;
}

template <class L, class A, class O>
int bem_matrix_dispatch(int core, const BemSurf& s, const BemSurf& t, int nunk, int same, int maxlev, O* c, int nth) {
  switch (core) {
    case C_RM:  bem_matrix_loop<C_RM, L, A>(s, t, nunk, same, maxlev, c, nth); return 0;
    case C_WL:  bem_matrix_loop<C_WL, L, A>(s, t, nunk, same, maxlev, c, nth); return 0;
    case C_EXP: bem_matrix_loop<C_EXP, L, A>(s, t, nunk, same, maxlev, c, nth); return 0;
    case C_V2:  bem_matrix_loop<C_V2, L, A>(s, t, nunk, same, maxlev, c, nth); return 0;
  }
  return 1;
}

int clamp_threads(int n) {
#ifdef _OPENMP
  if (n <= 0) n = omp_get_max_threads();
#else
  n = 1;
#endif
  return n < 1 ? 1 : n;
}

// ---------------------------------------------------------------------------
// Vc-style loop: the CPU performance baseline (Influence.h:388-442 shape).
// AVX2, 8 float lanes; rsqrt/rcp are the hardware estimates, as Vc's
// rsqrt()/reciprocal() are on x86 (MathHelper.h:56-60,116-141).  exp is this
// file's own 8-lane implementation: Cody-Waite reduction by ln2 and a degree-5
// minimax polynomial for 2^f on [-0.5,0.5] (the Cephes expf scheme, S. Moshier).
// ---------------------------------------------------------------------------
#define O3D_AVX __attribute__((target("avx2,fma")))

O3D_AVX inline __m256 v8_exp(__m256 x) {
  const __m256 hi = _mm256_set1_ps(88.3762626647949f), lo = _mm256_set1_ps(-87.3365447504f);
  x = _mm256_min_ps(_mm256_max_ps(x, lo), hi);
  __m256 fx = _mm256_round_ps(_mm256_mul_ps(x, _mm256_set1_ps(1.44269504088896341f)),
                              _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC);
  // x - fx*ln2 in two pieces
  x = _mm256_sub_ps(x, _mm256_mul_ps(fx, _mm256_set1_ps(0.693359375f)));
  x = _mm256_sub_ps(x, _mm256_mul_ps(fx, _mm256_set1_ps(-2.12194440e-4f)));
  const __m256 z = _mm256_mul_ps(x, x);
  __m256 y = _mm256_set1_ps(1.9875691500e-4f);
  y = _mm256_add_ps(_mm256_mul_ps(y, x), _mm256_set1_ps(1.3981999507e-3f));
  y = _mm256_add_ps(_mm256_mul_ps(y, x), _mm256_set1_ps(8.3334519073e-3f));
  y = _mm256_add_ps(_mm256_mul_ps(y, x), _mm256_set1_ps(4.1665795894e-2f));
  y = _mm256_add_ps(_mm256_mul_ps(y, x), _mm256_set1_ps(1.6666665459e-1f));
  y = _mm256_add_ps(_mm256_mul_ps(y, x), _mm256_set1_ps(5.0000001201e-1f));
  y = _mm256_add_ps(_mm256_mul_ps(y, z), _mm256_add_ps(x, _mm256_set1_ps(1.0f)));
  __m256i e = _mm256_add_epi32(_mm256_cvtps_epi32(fx), _mm256_set1_epi32(127));
  e = _mm256_slli_epi32(e, 23);
  return _mm256_mul_ps(y, _mm256_castsi256_ps(e));
}

O3D_AVX inline float v8_hsum(__m256 v) {
  const __m128 a = _mm_add_ps(_mm256_castps256_ps128(v), _mm256_extractf128_ps(v, 1));
  const __m128 b = _mm_add_ps(a, _mm_movehl_ps(a, a));
  const __m128 c = _mm_add_ss(b, _mm_shuffle_ps(b, b, 1));
  return _mm_cvtss_f32(c);
}

template <int CORE>
O3D_AVX inline void v8_core(__m256 d2, __m256 sr, __m256 tr, __m256& k, __m256& b) {
  const __m256 m3 = _mm256_set1_ps(-3.0f);
  if (CORE == C_RM) {
    const __m256 r2 = _mm256_add_ps(_mm256_add_ps(d2, _mm256_mul_ps(sr, sr)), _mm256_mul_ps(tr, tr));
    const __m256 rc = _mm256_rcp_ps(r2);
    k = _mm256_mul_ps(_mm256_rsqrt_ps(r2), rc);
    b = _mm256_mul_ps(_mm256_mul_ps(m3, k), rc);
  } else if (CORE == C_WL) {
    const __m256 r2 = _mm256_add_ps(_mm256_mul_ps(sr, sr), _mm256_mul_ps(tr, tr));
    const __m256 dd = _mm256_add_ps(d2, r2);
    const __m256 top = _mm256_add_ps(d2, _mm256_mul_ps(_mm256_set1_ps(2.5f), r2));
    const __m256 dn5 = _mm256_mul_ps(_mm256_rsqrt_ps(dd), _mm256_rcp_ps(_mm256_mul_ps(dd, dd)));
    k = _mm256_mul_ps(top, dn5);
    b = _mm256_sub_ps(_mm256_mul_ps(_mm256_set1_ps(2.0f), dn5),
                      _mm256_div_ps(_mm256_mul_ps(_mm256_mul_ps(_mm256_set1_ps(5.0f), top), dn5), dd));
  } else if (CORE == C_EXP) {
    const __m256 dist = _mm256_sqrt_ps(d2);
    const __m256 cs = _mm256_add_ps(_mm256_mul_ps(_mm256_mul_ps(sr, sr), sr), _mm256_mul_ps(_mm256_mul_ps(tr, tr), tr));
    const __m256 cf = _mm256_rcp_ps(cs);
    const __m256 d3 = _mm256_mul_ps(d2, dist);
    const __m256 q = _mm256_mul_ps(d3, cf);
    const __m256 ood3 = _mm256_rcp_ps(d3);
    const __m256 e = v8_exp(_mm256_sub_ps(_mm256_setzero_ps(), q));
    const __m256 mid = _mm256_cmp_ps(q, _mm256_set1_ps(16.0f), _CMP_LT_OQ);
    const __m256 near = _mm256_cmp_ps(q, _mm256_set1_ps(0.001f), _CMP_LT_OQ);
    const __m256 far = _mm256_cmp_ps(q, _mm256_set1_ps(16.0f), _CMP_GT_OQ);
    // exp_cond: ood3, then (q<16) ood3*(1-e), then (q<0.001) cf   (CoreFunc.h:92-96)
    k = ood3;
    k = _mm256_blendv_ps(k, _mm256_mul_ps(ood3, _mm256_sub_ps(_mm256_set1_ps(1.0f), e)), mid);
    k = _mm256_blendv_ps(k, cf, near);
    // exp_bbb (CoreFunc.h:136-142)
    b = _mm256_setzero_ps();
    b = _mm256_blendv_ps(b, _mm256_div_ps(_mm256_mul_ps(m3, k), d2), far);
    b = _mm256_blendv_ps(b, _mm256_div_ps(_mm256_mul_ps(_mm256_set1_ps(3.0f),
                                          _mm256_sub_ps(_mm256_mul_ps(cf, e), k)), d2), mid);
    b = _mm256_blendv_ps(b, _mm256_mul_ps(_mm256_mul_ps(_mm256_mul_ps(_mm256_set1_ps(-1.5f), dist), k), k), near);
  } else {
    const __m256 s2 = _mm256_mul_ps(sr, sr), t2 = _mm256_mul_ps(tr, tr);
    const __m256 den = _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(d2, d2), _mm256_mul_ps(s2, s2)), _mm256_mul_ps(t2, t2));
    const __m256 rs = _mm256_rsqrt_ps(den);
    k = _mm256_mul_ps(_mm256_mul_ps(rs, rs), _mm256_rsqrt_ps(rs));
    b = _mm256_mul_ps(_mm256_mul_ps(m3, k), _mm256_rsqrt_ps(den));
  }
}

template <int CORE>
O3D_AVX void vc_loop(const Src<float>& s, const Tgt<float>& t, int64_t t0, int64_t t1, int nthreads) {
  // pad sources to a whole number of 8-lane vectors: position 0, radius 1,
  // strength 0, so pads contribute exactly zero (VectorHelper.h:42-51, Influence.h:232-238)
  const int64_t nv = (s.n + 7) / 8;
  std::vector<float> buf(7 * nv * 8);
  float* px = buf.data();      float* py = px + nv*8;  float* pz = py + nv*8;  float* pr = pz + nv*8;
  float* qx = pr + nv*8;       float* qy = qx + nv*8;  float* qz = qy + nv*8;
  for (int64_t j = 0; j < nv*8; ++j) {
    const bool in = j < s.n;
    px[j] = in ? s.x[j] : 0.f;  py[j] = in ? s.y[j] : 0.f;  pz[j] = in ? s.z[j] : 0.f;
    pr[j] = in ? s.r[j] : 1.f;
    qx[j] = in ? s.sx[j] : 0.f; qy[j] = in ? s.sy[j] : 0.f; qz[j] = in ? s.sz[j] : 0.f;
  }
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t i = t0; i < t1; ++i) {
    const __m256 tx = _mm256_set1_ps(t.x[i]), ty = _mm256_set1_ps(t.y[i]), tz = _mm256_set1_ps(t.z[i]);
    const __m256 tr = _mm256_set1_ps(t.r[i]);
    __m256 a[12];
    for (int k = 0; k < 12; ++k) a[k] = _mm256_setzero_ps();
    for (int64_t v = 0; v < nv; ++v) {
      const __m256 dx = _mm256_sub_ps(tx, _mm256_loadu_ps(px + 8*v));
      const __m256 dy = _mm256_sub_ps(ty, _mm256_loadu_ps(py + 8*v));
      const __m256 dz = _mm256_sub_ps(tz, _mm256_loadu_ps(pz + 8*v));
      const __m256 sr = _mm256_loadu_ps(pr + 8*v);
      const __m256 sx = _mm256_loadu_ps(qx + 8*v), sy = _mm256_loadu_ps(qy + 8*v), sz = _mm256_loadu_ps(qz + 8*v);
      const __m256 d2 = _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(dx, dx), _mm256_mul_ps(dy, dy)), _mm256_mul_ps(dz, dz));
      __m256 k, b;
      v8_core<CORE>(d2, sr, tr, k, b);
      __m256 cx = _mm256_sub_ps(_mm256_mul_ps(dz, sy), _mm256_mul_ps(dy, sz));
      __m256 cy = _mm256_sub_ps(_mm256_mul_ps(dx, sz), _mm256_mul_ps(dz, sx));
      __m256 cz = _mm256_sub_ps(_mm256_mul_ps(dy, sx), _mm256_mul_ps(dx, sy));
      a[0] = _mm256_add_ps(a[0], _mm256_mul_ps(k, cx));
      a[1] = _mm256_add_ps(a[1], _mm256_mul_ps(k, cy));
      a[2] = _mm256_add_ps(a[2], _mm256_mul_ps(k, cz));
      cx = _mm256_mul_ps(cx, b); cy = _mm256_mul_ps(cy, b); cz = _mm256_mul_ps(cz, b);
      const __m256 kx = _mm256_mul_ps(sx, k), ky = _mm256_mul_ps(sy, k), kz = _mm256_mul_ps(sz, k);
      a[3]  = _mm256_add_ps(a[3],  _mm256_mul_ps(dx, cx));
      a[4]  = _mm256_add_ps(a[4],  _mm256_add_ps(_mm256_mul_ps(dx, cy), kz));
      a[5]  = _mm256_add_ps(a[5],  _mm256_sub_ps(_mm256_mul_ps(dx, cz), ky));
      a[6]  = _mm256_add_ps(a[6],  _mm256_sub_ps(_mm256_mul_ps(dy, cx), kz));
      a[7]  = _mm256_add_ps(a[7],  _mm256_mul_ps(dy, cy));
      a[8]  = _mm256_add_ps(a[8],  _mm256_add_ps(_mm256_mul_ps(dy, cz), kx));
      a[9]  = _mm256_add_ps(a[9],  _mm256_add_ps(_mm256_mul_ps(dz, cx), ky));
      a[10] = _mm256_add_ps(a[10], _mm256_sub_ps(_mm256_mul_ps(dz, cy), kx));
      a[11] = _mm256_add_ps(a[11], _mm256_mul_ps(dz, cz));
    }
    t.u[i] += v8_hsum(a[0]); t.v[i] += v8_hsum(a[1]); t.w[i] += v8_hsum(a[2]);
    for (int k = 0; k < 9; ++k) t.g[k][i] += v8_hsum(a[3 + k]);
  }
}

}  // namespace

// ===========================================================================
// C ABI of the oracle (ctypes, tests only)
// ===========================================================================
extern "C" {

// Flop counts of the reference (CoreFunc.h flops_t{v,p}_{no,}grads and
// Kernels.h flops_0v_0b / 0v_0p / 0v_0bg / 0v_0pg / 0vs_0p / 0vs_0pg).
int orc_flops_pair(int core, int grads, int thick, int vs) {
  static const int tvg[5] = {0, 9, 16, 14, 13}, tpg[5] = {0, 7, 14, 11, 10};
  static const int tvn[5] = {0, 7, 10, 12, 11}, tpn[5] = {0, 5, 8, 9, 8};
  if (core < 1 || core > 4) return -1;
  if (grads) return (vs ? 79 : 54) + (thick ? tvg[core] : tpg[core]);
  return (vs ? 29 : 23) + (thick ? tvn[core] : tpn[core]);
}

// mode 0: S=float A=float;  mode 1: S=float A=double.  Float arrays in and out.
// tr may be NULL (point targets); tug is 9 pointers or NULL (velocity only).
int orc_p2p_f(int core, int grads, int mode, int64_t ns, const float* sx, const float* sy, const float* sz,
              const float* sr, const float* ssx, const float* ssy, const float* ssz,
              int64_t nt, const float* tx, const float* ty, const float* tz, const float* tr,
              float* tu, float* tv, float* tw, float* const* tug, int nthreads) {
  if (grads && !tug) return 2;
  Src<float> s{ns, sx, sy, sz, sr, ssx, ssy, ssz};
  Tgt<float> t{nt, tx, ty, tz, tr, tu, tv, tw, tug};
  const int nth = clamp_threads(nthreads);
  if (mode == 0) return p2p_dispatch<float, float>(core, grads, tr != nullptr, s, t, nth);
  if (mode == 1) return p2p_dispatch<float, double>(core, grads, tr != nullptr, s, t, nth);
  return 3;
}

// FP64 oracle: every operation in double.  Double arrays in and out.
int orc_p2p_d(int core, int grads, int64_t ns, const double* sx, const double* sy, const double* sz,
              const double* sr, const double* ssx, const double* ssy, const double* ssz,
              int64_t nt, const double* tx, const double* ty, const double* tz, const double* tr,
              double* tu, double* tv, double* tw, double* const* tug, int nthreads) {
  if (grads && !tug) return 2;
  Src<double> s{ns, sx, sy, sz, sr, ssx, ssy, ssz};
  Tgt<double> t{nt, tx, ty, tz, tr, tu, tv, tw, tug};
  return p2p_dispatch<double, double>(core, grads, tr != nullptr, s, t, clamp_threads(nthreads));
}

int orc_pan_f(int core, int grads, int mode, int maxlev, int64_t np, const float* x, const float* y, const float* z,
              const uint32_t* idx, const float* tsx, const float* tsy, const float* tsz, const float* area,
              const float* ss, int64_t nt, const float* tx, const float* ty, const float* tz,
              float* tu, float* tv, float* tw, float* const* tug, int nthreads) {
  if (grads && !tug) return 2;
  Pan<float> p{np, x, y, z, idx, tsx, tsy, tsz, area, ss};
  Tgt<float> t{nt, tx, ty, tz, nullptr, tu, tv, tw, tug};
  const int nth = clamp_threads(nthreads);
  if (mode == 0) return pan_dispatch<float, float>(core, grads, p, t, maxlev, nth);
  if (mode == 1) return pan_dispatch<float, double>(core, grads, p, t, maxlev, nth);
  return 3;
}

int orc_pan_d(int core, int grads, int maxlev, int64_t np, const double* x, const double* y, const double* z,
              const uint32_t* idx, const double* tsx, const double* tsy, const double* tsz, const double* area,
              const double* ss, int64_t nt, const double* tx, const double* ty, const double* tz,
              double* tu, double* tv, double* tw, double* const* tug, int nthreads) {
  if (grads && !tug) return 2;
  Pan<double> p{np, x, y, z, idx, tsx, tsy, tsz, area, ss};
  Tgt<double> t{nt, tx, ty, tz, nullptr, tu, tv, tw, tug};
  return pan_dispatch<double, double>(core, grads, p, t, maxlev, clamp_threads(nthreads));
}

// points_affect_panels: particles (position, strength; radii unused) onto panels;
// u/v/w per panel, decremented as the reference does.  mode 0: S=A=float,
// 1: S=float, A=double.
int orc_ptpan_f(int core, int mode, int maxlev, int64_t ns, const float* sx, const float* sy, const float* sz,
                const float* ssx, const float* ssy, const float* ssz, int64_t np, const float* x, const float* y,
                const float* z, const uint32_t* idx, const float* area, float* tu, float* tv, float* tw,
                int nthreads) {
  Src<float> s{ns, sx, sy, sz, nullptr, ssx, ssy, ssz};
  Pan<float> p{np, x, y, z, idx, nullptr, nullptr, nullptr, area, nullptr};
  const int nth = clamp_threads(nthreads);
  if (mode == 0) return ptpan_dispatch<float, float>(core, s, p, tu, tv, tw, maxlev, nth);
  if (mode == 1) return ptpan_dispatch<float, double>(core, s, p, tu, tv, tw, maxlev, nth);
  return 3;
}

int orc_ptpan_d(int core, int maxlev, int64_t ns, const double* sx, const double* sy, const double* sz,
                const double* ssx, const double* ssy, const double* ssz, int64_t np, const double* x,
                const double* y, const double* z, const uint32_t* idx, const double* area, double* tu, double* tv,
                double* tw, int nthreads) {
  Src<double> s{ns, sx, sy, sz, nullptr, ssx, ssy, ssz};
  Pan<double> p{np, x, y, z, idx, nullptr, nullptr, nullptr, area, nullptr};
  return ptpan_dispatch<double, double>(core, s, p, tu, tv, tw, maxlev, clamp_threads(nthreads));
}

// BEM influence matrix (Coefficients.h:169-444).  Surfaces: nodes, panels,
// areas and the bases b1, b2, nm (3 rows of np floats each, as compute_bases
// leaves them).  nunk 2 or 3 for both; same: the two are one surface (the
// diagonal blocks of Coefficients.h:399-421).  mode 0: the reference's float
// arithmetic, float output c32; mode 1: the same tree with FP64 leaves and
// sums, double output c64.  Sizes (nunk*ns) columns x (nunk*nt) rows, column-major.
int orc_bem_matrix(int core, int mode, int maxlev, int nunk, int same, int64_t ns, const float* sx, const float* sy,
                   const float* sz, const uint32_t* sidx, const float* sarea, const float* const* sbases,
                   int64_t nt, const float* tx, const float* ty, const float* tz, const uint32_t* tidx,
                   const float* tarea, const float* const* tbases, float* c32, double* c64, int nthreads) {
  if (nunk != 2 && nunk != 3) return 2;
  if (same && ns != nt) return 4;
  BemSurf s{ns, sx, sy, sz, sidx, sarea, {sbases[0], sbases[1], sbases[2]}, {sbases[3], sbases[4], sbases[5]},
            {sbases[6], sbases[7], sbases[8]}};
  BemSurf t{nt, tx, ty, tz, tidx, tarea, {tbases[0], tbases[1], tbases[2]}, {tbases[3], tbases[4], tbases[5]},
            {tbases[6], tbases[7], tbases[8]}};
  const int nth = clamp_threads(nthreads);
  if (mode == 0) return bem_matrix_dispatch<float, float>(core, s, t, nunk, same, maxlev, c32, nth);
  if (mode == 1) return bem_matrix_dispatch<double, double>(core, s, t, nunk, same, maxlev, c64, nth);
  return 3;
}

// 1 if the host has AVX2+FMA (the Vc-style loop needs them)
int orc_have_avx2(void) { return __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma"); }

// Vc-style velocity+gradient loop on thick targets, over targets [t0, t1).
int orc_p2p_vcstyle(int core, int64_t ns, const float* sx, const float* sy, const float* sz,
                    const float* sr, const float* ssx, const float* ssy, const float* ssz,
                    int64_t nt, const float* tx, const float* ty, const float* tz, const float* tr,
                    float* tu, float* tv, float* tw, float* const* tug,
                    int64_t t0, int64_t t1, int nthreads) {
  if (!orc_have_avx2()) return 4;
  if (!tr || !tug) return 2;
  if (t0 < 0 || t1 > nt || t0 > t1) return 5;
  Src<float> s{ns, sx, sy, sz, sr, ssx, ssy, ssz};
  Tgt<float> t{nt, tx, ty, tz, tr, tu, tv, tw, tug};
  const int nth = clamp_threads(nthreads);
  switch (core) {
    case C_RM:  vc_loop<C_RM>(s, t, t0, t1, nth); return 0;
    case C_WL:  vc_loop<C_WL>(s, t, t0, t1, nth); return 0;
    case C_EXP: vc_loop<C_EXP>(s, t, t0, t1, nth); return 0;
    case C_V2:  vc_loop<C_V2>(s, t, t0, t1, nth); return 0;
  }
  return 1;
}

int orc_max_threads(void) { return clamp_threads(0); }

}  // extern "C"
