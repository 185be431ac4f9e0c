// pan_g7.hpp -- panels -> points by a 7-point rule per panel with a closed-form
// near field (method O3D_PANEL_GAUSS7).  No counterpart in the reference, whose
// panels_affect_points subdivides recursively (Kernels.h:1028-1211, pan_kernels.hpp);
// the panel conventions are the reference's (Influence.h:557-1099): ts the
// absolute vortex strength, ss the sheet source strength, point targets, the
// singular kernel K = 1/r^3, results accumulated and unscaled.
//
// Per (target t, panel p), c the centroid and D the longest edge:
//   far  (|t - c|^2 >= 4 D^2, decided in float, left to right, no contraction):
//        seven singular vortex + source points y_q, strengths w_q ts and w_q ss area
//        (Strang & Fix 1973; Dunavant 1985, the degree-5 rule);
//   near: the exact integral of the densities gamma = ts/area, sigma = ss over the
//        flat triangle, in FP64: u += gamma x I + sigma I with
//        I = int_T (t - y)/|t - y|^3 dA = sum_e m_e f_e + Omega n (edge logarithms
//        f_e, outward in-plane edge normals m_e, signed solid angle Omega).
//
// GPU form: one lane per target, panels staged per tile in LDS with their seven
// points precomputed.  The far part of a pair sums, over the seven points, only
// the geometric moments P = sum w K d, Q = sum w K d d^T / r^2 (symmetric) and
// S = sum w K, and applies the panel's strengths once:
//   u      += g x P + gs P
//   du_i/dx_j += -3 ((g x Q_j)_i + gs Q_ji) + S (eps_ikj g_k + gs delta_ij),
// ~22 VALU operations per point and ~55 per panel instead of ~39 per point.  Near
// pairs are rare (a target within two panel sizes of the surface, a few dozen
// panels): the far kernel lists the targets that have one, and a second kernel
// gives each listed target a wave whose lanes share out the panels, take the
// decision again and evaluate the near pairs in FP64, adding after the far kernel.
#pragma once
#include "pan_kernels.hpp"

namespace o3d {

constexpr int G7_F4 = 10;   // float4 per staged panel

// the rule: point 0 the centroid, 1-3 the permutations of (1 - 2 a1, a1, a1),
// 4-6 those of (1 - 2 a2, a2, a2); weights sum to 1
constexpr double G7_A1 = 0.101286507323456338800987361915;   // (6 - sqrt 15)/21
constexpr double G7_A2 = 0.470142064105115089770441209513;   // (6 + sqrt 15)/21
constexpr double G7_W0 = 0.225;
constexpr double G7_W1 = 0.1259391805448271525956839455;     // (155 - sqrt 15)/1200
constexpr double G7_W2 = 0.132394152788506180737649387833;   // (155 + sqrt 15)/1200

__host__ __device__ constexpr double g7_bary(int q, int k) {
  return q == 0 ? 1.0 / 3.0
                : (q < 4 ? ((q - 1) == k ? 1.0 - 2.0 * G7_A1 : G7_A1) : ((q - 4) == k ? 1.0 - 2.0 * G7_A2 : G7_A2));
}
__host__ __device__ constexpr double g7_weight(int q) { return q == 0 ? G7_W0 : (q < 4 ? G7_W1 : G7_W2); }

// the decision's panel geometry {centroid, 4 D^2}: float, left to right, no contraction
__device__ __forceinline__ float4 g7_geometry(const float* v) {
#pragma clang fp contract(off)
  const float cx = (v[0] + v[3] + v[6]) / 3.0f;
  const float cy = (v[1] + v[4] + v[7]) / 3.0f;
  const float cz = (v[2] + v[5] + v[8]) / 3.0f;
  float e[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int b = 3 * ((k + 1) % 3), c = 3 * k;
    const float dx = v[b] - v[c], dy = v[b + 1] - v[c + 1], dz = v[b + 2] - v[c + 2];
    e[k] = dx * dx + dy * dy + dz * dz;
  }
  return make_float4(cx, cy, cz, 4.0f * fmaxf(fmaxf(e[0], e[1]), e[2]));
}

__device__ __forceinline__ void panel_vertices(const PanArgs& a, int64_t p, float* v) {
  const uint32_t i0 = a.idx[3 * p], i1 = a.idx[3 * p + 1], i2 = a.idx[3 * p + 2];
  v[0] = a.nx[i0]; v[1] = a.ny[i0]; v[2] = a.nz[i0];
  v[3] = a.nx[i1]; v[4] = a.ny[i1]; v[5] = a.nz[i1];
  v[6] = a.nx[i2]; v[7] = a.ny[i2]; v[8] = a.nz[i2];
}

// stage panels [p0, p0 + cnt):
//   f0..f5.x  the seven points (21 floats, computed in FP64 and rounded)
//   f5.yzw    centroid (float, the decision's arithmetic)
//   f6        {tsx, tsy, tsz, ss * area}   absolute vortex and source strengths
//   f7        {4 D^2, v0x, v0y, v0z}       f8 {v1x, v1y, v1z, v2x}
//   f9        {v2y, v2z, area, ss}
__device__ __forceinline__ void stage_g7(const PanArgs& a, int64_t p0, int cnt, float4* sh) {
  for (int j = threadIdx.x; j < cnt; j += BLOCK) {
    const int64_t p = p0 + j;
    float v[9];
    panel_vertices(a, p, v);
    float pts[21];
#pragma unroll
    for (int q = 0; q < 7; ++q)
#pragma unroll
      for (int d = 0; d < 3; ++d)
        pts[3 * q + d] = (float)(g7_bary(q, 0) * (double)v[d] + g7_bary(q, 1) * (double)v[3 + d] +
                                 g7_bary(q, 2) * (double)v[6 + d]);
    const float4 geo = g7_geometry(v);
    const float cx = geo.x, cy = geo.y, cz = geo.z, lim = geo.w;
    const float ar = a.area[p];
    const float sg = a.ss ? a.ss[p] : 0.0f;
    float4* s = sh + G7_F4 * j;
    s[0] = make_float4(pts[0], pts[1], pts[2], pts[3]);
    s[1] = make_float4(pts[4], pts[5], pts[6], pts[7]);
    s[2] = make_float4(pts[8], pts[9], pts[10], pts[11]);
    s[3] = make_float4(pts[12], pts[13], pts[14], pts[15]);
    s[4] = make_float4(pts[16], pts[17], pts[18], pts[19]);
    s[5] = make_float4(pts[20], cx, cy, cz);
    s[6] = make_float4(a.tsx[p], a.tsy[p], a.tsz[p], sg * ar);
    s[7] = make_float4(lim, v[0], v[1], v[2]);
    s[8] = make_float4(v[3], v[4], v[5], v[6]);
    s[9] = make_float4(v[7], v[8], ar, sg);
  }
}

__device__ __forceinline__ bool g7_is_near(float tx, float ty, float tz, const float4* s) {
  const float4 f5 = s[5];
  float d2;
  {
#pragma clang fp contract(off)
    const float ex = tx - f5.y, ey = ty - f5.z, ez = tz - f5.w;
    d2 = ex * ex + ey * ey + ez * ez;
  }
  return d2 < s[7].x;
}

__device__ __forceinline__ void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// I = int_T (x - y)/|x - y|^3 dA and dI[3j+l] = dI_l/dx_j over the triangle v, FP64
template <bool GRAD>
__device__ __forceinline__ void tri_integrals(const double* v, const double* x, double* I, double* dI) {
  double e1[3], e2[3], n[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    e1[d] = v[3 + d] - v[d];
    e2[d] = v[6 + d] - v[d];
  }
  cross3(e1, e2, n);
  const double ia = 1.0 / sqrt(dot3(n, n));
#pragma unroll
  for (int d = 0; d < 3; ++d) n[d] *= ia;
  double R[3][3], r[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int d = 0; d < 3; ++d) R[k][d] = v[3 * k + d] - x[d];
    r[k] = sqrt(dot3(R[k], R[k]));
  }
  double c12[3];
  cross3(R[1], R[2], c12);
  const double num = dot3(R[0], c12);
  const double den = r[0] * r[1] * r[2] + dot3(R[0], R[1]) * r[2] + dot3(R[0], R[2]) * r[1] + dot3(R[1], R[2]) * r[0];
  const double om = -2.0 * atan2(num, den);   // signed solid angle (Van Oosterom & Strackee)
#pragma unroll
  for (int d = 0; d < 3; ++d) I[d] = om * n[d];
  double gom[3] = {0.0, 0.0, 0.0};
  if (GRAD) {
#pragma unroll
    for (int k = 0; k < 9; ++k) dI[k] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double* a = v + 3 * k;
    const double* b = v + 3 * ((k + 1) % 3);
    double t[3], m[3], q[3], da[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) t[d] = b[d] - a[d];
    const double len = sqrt(dot3(t, t));
    const double il = 1.0 / len;
#pragma unroll
    for (int d = 0; d < 3; ++d) t[d] *= il;
    cross3(t, n, m);
#pragma unroll
    for (int d = 0; d < 3; ++d) da[d] = x[d] - a[d];
    const double s = dot3(da, t);
    const double um = -s, up = len - s;
#pragma unroll
    for (int d = 0; d < 3; ++d) q[d] = da[d] - s * t[d];
    const double r02 = dot3(q, q);
    const double rm = r[k], rp = r[(k + 1) % 3];
    // int_e ds/|x - y| in cancellation-free forms
    double f;
    if (um >= 0.0) f = log((rp + up) / (rm + um));
    else if (up <= 0.0) f = log((rm - um) / (rp - up));
    else f = log((rp + up) * (rm - um) / r02);
#pragma unroll
    for (int d = 0; d < 3; ++d) I[d] += m[d] * f;
    if (GRAD) {
      const double c = um * up > 0.0 ? (up * up - um * um) / (rp * rm * (up * rm + um * rp)) : (up / rp - um / rm) / r02;
      const double e = 1.0 / rp - 1.0 / rm;
      double tq[3];
      cross3(t, q, tq);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double df = -(q[j] * c + t[j] * e);
#pragma unroll
        for (int l = 0; l < 3; ++l) dI[3 * j + l] += df * m[l];
        gom[j] -= tq[j] * c;
      }
    }
  }
  if (GRAD) {
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int l = 0; l < 3; ++l) dI[3 * j + l] += gom[j] * n[l];
  }
}

// the closed form of one near pair, FP64: adds to u[3] and g[9]; vf the vertices,
// ts the absolute vortex strength, area and ss the sheet source strength
template <bool GRAD>
__device__ __forceinline__ void g7_near_pair(double tx, double ty, double tz, const float* vf, const float* ts,
                                             float area, float ss, double* u, double* g) {
  double v[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) v[k] = vf[k];
  const double x[3] = {tx, ty, tz};
  double I[3], dI[9];
  tri_integrals<GRAD>(v, x, I, dI);
  const double ar = area, sg = ss;
  const double gam[3] = {(double)ts[0] / ar, (double)ts[1] / ar, (double)ts[2] / ar};
  double gi[3];
  cross3(gam, I, gi);
#pragma unroll
  for (int d = 0; d < 3; ++d) u[d] += gi[d] + sg * I[d];
  if (GRAD) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double gj[3];
      cross3(gam, dI + 3 * j, gj);
#pragma unroll
      for (int d = 0; d < 3; ++d) g[3 * j + d] += gj[d] + sg * dI[3 * j + d];
    }
  }
}

// far part of one pair: the seven points' moments, then the panel's strengths;
// m = 0 drops the pair (near, or an inactive lane).  acc: u[0..2], grad[3..11]
template <class R, bool GRAD>
__device__ __forceinline__ void g7_far_pair(R tx, R ty, R tz, const float4* s, R m, R* acc) {
  R P0 = 0, P1 = 0, P2 = 0, S = 0;
  R Q00 = 0, Q01 = 0, Q02 = 0, Q11 = 0, Q12 = 0, Q22 = 0;
  if constexpr (sizeof(R) == 4) {
    const float4 f0 = s[0], f1 = s[1], f2 = s[2], f3 = s[3], f4 = s[4], f5 = s[5];
    const float pt[21] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w, f2.x, f2.y, f2.z,
                          f2.w, f3.x, f3.y, f3.z, f3.w, f4.x, f4.y, f4.z, f4.w, f5.x};
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const float dx = tx - pt[3 * q], dy = ty - pt[3 * q + 1], dz = tz - pt[3 * q + 2];
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      const float ir = frsq(r2);
      const float ir2 = ir * ir;
      const float k = ((float)g7_weight(q) * ir) * ir2;
      const float kb = k * ir2;
      P0 = fmaf(k, dx, P0); P1 = fmaf(k, dy, P1); P2 = fmaf(k, dz, P2);
      const float ex = kb * dx, ey = kb * dy, ez = kb * dz;
      if (GRAD) {
        Q00 = fmaf(ex, dx, Q00); Q01 = fmaf(ex, dy, Q01); Q02 = fmaf(ex, dz, Q02);
        Q11 = fmaf(ey, dy, Q11); Q12 = fmaf(ey, dz, Q12); Q22 = fmaf(ez, dz, Q22);
        S += k;
      }
    }
  } else {
    const float4 f7 = s[7], f8 = s[8], f9 = s[9];
    const double v[9] = {f7.y, f7.z, f7.w, f8.x, f8.y, f8.z, f8.w, f9.x, f9.y};
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      double d[3];
#pragma unroll
      for (int k = 0; k < 3; ++k)
        d[k] = (k == 0 ? tx : (k == 1 ? ty : tz)) -
               (g7_bary(q, 0) * v[k] + g7_bary(q, 1) * v[3 + k] + g7_bary(q, 2) * v[6 + k]);
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      const double k = g7_weight(q) / (r2 * sqrt(r2));
      const double kb = k / r2;
      P0 += k * d[0]; P1 += k * d[1]; P2 += k * d[2];
      if (GRAD) {
        Q00 += kb * d[0] * d[0]; Q01 += kb * d[0] * d[1]; Q02 += kb * d[0] * d[2];
        Q11 += kb * d[1] * d[1]; Q12 += kb * d[1] * d[2]; Q22 += kb * d[2] * d[2];
        S += k;
      }
    }
  }
  const float4 f6 = s[6];
  // FP64 forms the absolute source strength ss * area in double; FP32 takes the staged float product
  const R gs0 = sizeof(R) == 8 ? R(s[9].w) * R(s[9].z) : R(f6.w);
  const R gx = R(f6.x) * m, gy = R(f6.y) * m, gz = R(f6.z) * m, gs = gs0 * m;
  acc[0] += gy * P2 - gz * P1 + gs * P0;
  acc[1] += gz * P0 - gx * P2 + gs * P1;
  acc[2] += gx * P1 - gy * P0 + gs * P2;
  if (GRAD) {
    // Q holds sum w K d d^T / r^2; the gradient weight is -3 times that
    const R Q[3][3] = {{Q00, Q01, Q02}, {Q01, Q11, Q12}, {Q02, Q12, Q22}};
    const R hx = gx * R(-3), hy = gy * R(-3), hz = gz * R(-3), hs = gs * R(-3);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      acc[3 + 3 * j] += hy * Q[j][2] - hz * Q[j][1] + hs * Q[j][0];
      acc[4 + 3 * j] += hz * Q[j][0] - hx * Q[j][2] + hs * Q[j][1];
      acc[5 + 3 * j] += hx * Q[j][1] - hy * Q[j][0] + hs * Q[j][2];
    }
    const R sx = S * gx, sy = S * gy, sz = S * gz, sd = S * gs;
    acc[3] += sd; acc[7] += sd; acc[11] += sd;
    acc[4] += sz; acc[5] -= sy;
    acc[6] -= sz; acc[8] += sx;
    acc[9] += sy; acc[10] -= sx;
  }
}

// far pairs of every target (near pairs contribute nothing here)
template <bool GRAD, int PREC>
__global__ __launch_bounds__(BLOCK) void k_pan_g7(PanArgs a) {
  using R = typename std::conditional<PREC == PREC_FP64, double, float>::type;
  extern __shared__ float4 shp[];
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const bool active = i < a.nt;
  const int64_t ic = active ? i : a.nt - 1;
  const float txf = a.tx[ic], tyf = a.ty[ic], tzf = a.tz[ic];
  const R tx = txf, ty = tyf, tz = tzf;
  R acc[12];
  float hi[3] = {0.f, 0.f, 0.f}, lo[3] = {0.f, 0.f, 0.f};
  double gd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool any_near = false;
#pragma unroll
  for (int c = 0; c < 12; ++c) acc[c] = R(0);
  for (int64_t p0 = 0; p0 < a.np; p0 += a.tile) {
    const int cnt = (int)((a.np - p0) < a.tile ? (a.np - p0) : a.tile);
    __syncthreads();
    stage_g7(a, p0, cnt, shp);
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      const float4* s = shp + G7_F4 * j;
      const bool near = g7_is_near(txf, tyf, tzf, s);
      any_near |= near;
      g7_far_pair<R, GRAD>(tx, ty, tz, s, (active && !near) ? R(1) : R(0), acc);
    }
    if (PREC == PREC_COMP) {
      two_sum(hi[0], lo[0], (float)acc[0]);
      two_sum(hi[1], lo[1], (float)acc[1]);
      two_sum(hi[2], lo[2], (float)acc[2]);
      if (GRAD) {
#pragma unroll
        for (int k = 0; k < 9; ++k) gd[k] += (double)acc[3 + k];
      }
#pragma unroll
      for (int c = 0; c < 12; ++c) acc[c] = R(0);
    }
  }
  if (!active) return;
  if (any_near && a.near_list) a.near_list[atomicAdd(a.near_count, 1u)] = (int32_t)i;
  if (PREC == PREC_FP64) {
    ((double*)a.u)[i] += acc[0]; ((double*)a.v)[i] += acc[1]; ((double*)a.w)[i] += acc[2];
    if (GRAD) {
#pragma unroll
      for (int k = 0; k < 9; ++k) ((double*)a.g[k])[i] += acc[3 + k];
    }
  } else if (PREC == PREC_FP32) {
    ((float*)a.u)[i] += (float)acc[0]; ((float*)a.v)[i] += (float)acc[1]; ((float*)a.w)[i] += (float)acc[2];
    if (GRAD) {
#pragma unroll
      for (int k = 0; k < 9; ++k) ((float*)a.g[k])[i] += (float)acc[3 + k];
    }
  } else {
    float* u = (float*)a.u; float* v = (float*)a.v; float* w = (float*)a.w;
    u[i] = add_comp(u[i], hi[0], lo[0]);
    v[i] = add_comp(v[i], hi[1], lo[1]);
    w[i] = add_comp(w[i], hi[2], lo[2]);
    if (GRAD) {
#pragma unroll
      for (int k = 0; k < 9; ++k) ((float*)a.g[k])[i] = (float)((double)((float*)a.g[k])[i] + gd[k]);
    }
  }
}

// the decision's geometry of every panel, for k_pan_g7_near
__global__ __launch_bounds__(BLOCK) void k_pan_g7_geom(PanArgs a, float4* geom) {
  const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (p >= a.np) return;
  float v[9];
  panel_vertices(a, p, v);
  geom[p] = g7_geometry(v);
}


// near pairs of the targets k_pan_g7 listed, in FP64, added to the outputs after it.
// One wave per listed target: lane l takes the decision for panels l, l + 64, ...
// (the same float arithmetic as the far kernel) and evaluates the closed form of
// its near ones; a fixed butterfly sums the lanes, so results do not depend on
// scheduling.  Grid-stride over the list, whose length the far kernel left in
// near_count; nnear (optional) counts the near pairs.
template <bool GRAD, int PREC>
__global__ __launch_bounds__(BLOCK) void k_pan_g7_near(PanArgs a, const float4* geom, unsigned long long* nnear) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * BLOCK) >> 6;
  const int64_t listed = *a.near_count;
  unsigned long long mine = 0;
  for (int64_t e = wave; e < listed; e += nwaves) {
    const int64_t i = a.near_list[e];
    const float tx = a.tx[i], ty = a.ty[i], tz = a.tz[i];
    double u[3] = {0, 0, 0}, g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t p = lane; p < a.np; p += 64) {
      const float4 ge = geom[p];
      float d2;
      {
#pragma clang fp contract(off)
        const float ex = tx - ge.x, ey = ty - ge.y, ez = tz - ge.z;
        d2 = ex * ex + ey * ey + ez * ez;
      }
      if (d2 < ge.w) {
        float v[9];
        panel_vertices(a, p, v);
        const float ts[3] = {a.tsx[p], a.tsy[p], a.tsz[p]};
        g7_near_pair<GRAD>(tx, ty, tz, v, ts, a.area[p], a.ss ? a.ss[p] : 0.0f, u, g);
        ++mine;
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = wave_sum(u[d]);
    if (GRAD) {
#pragma unroll
      for (int k = 0; k < 9; ++k) g[k] = wave_sum(g[k]);
    }
    if (lane != 0) continue;
    if (PREC == PREC_FP64) {
      ((double*)a.u)[i] += u[0]; ((double*)a.v)[i] += u[1]; ((double*)a.w)[i] += u[2];
      if (GRAD) {
#pragma unroll
        for (int k = 0; k < 9; ++k) ((double*)a.g[k])[i] += g[k];
      }
    } else {
      float* uu = (float*)a.u; float* vv = (float*)a.v; float* ww = (float*)a.w;
      uu[i] = (float)((double)uu[i] + u[0]);
      vv[i] = (float)((double)vv[i] + u[1]);
      ww[i] = (float)((double)ww[i] + u[2]);
      if (GRAD) {
#pragma unroll
        for (int k = 0; k < 9; ++k) ((float*)a.g[k])[i] = (float)((double)((float*)a.g[k])[i] + g[k]);
      }
    }
  }
  if (nnear && mine) atomicAdd(nnear, mine);
}

__global__ void k_pan_g7_zero(unsigned int* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0u;
}

}  // namespace o3d
