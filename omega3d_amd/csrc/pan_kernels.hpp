// pan_kernels.hpp -- panel -> point direct sums (velocity and gradient) for gfx950:
// the recursive centroid rule of Omega3D's rkernel_2vs_0p / rkernel_2vs_0pg
// (Kernels.h:1028-1211) over flat triangles, as driven by panels_affect_points
// (Influence.h:557-1099).
//
// Per (target, panel) pair the reference walks a 4-ary subdivision tree: a node
// (sub-triangle) is a leaf when its centroid is more than 4 sqrt(area) from the
// target or at depth maxlev; a leaf is a singular (sigma = 0) vortex + source
// point at the centroid carrying the node's share of the panel strength.
//
// GPU form: k_pan takes one lane per target with the panels staged per tile in
// LDS, decides every root pair and evaluates the pairs that are leaves at the
// root, and lists the targets that have a pair needing the recursion (those near
// the surface).  k_pan_near then gives each listed target one wave whose lanes
// take the panels in turn and walk the near ones: a lane walks the tree in
// depth-first order with an explicit (level, path) cursor -- no recursion, no
// stack -- recomputing a node's vertices from the panel's corners along its path.
// Walking inside k_pan instead held a whole wave of targets for the deepest walk
// of any of its lanes.  Every quantity that decides the tree
// shape (midpoints, centroid, sqrt(area), distance) is computed in correctly
// rounded IEEE arithmetic with FMA contraction off, in the reference's order,
// so the GPU subdivides exactly where the reference does.  At the root the
// decision is taken from a fast distance estimate whenever it is unambiguous
// and from the exact expression otherwise.
//
// At sigma = 0 every core (RM, WL, EXP, V2) is the singular kernel K = 1/rho^3,
// B = -3/rho^5 (CoreFunc.h formulas with sr = 0), so the panel path has one
// leaf kernel whatever core the caller selected.
#pragma once
#include "o3d_common.hpp"

namespace o3d {

// Surfaces::compute_bases (Surfaces.h:766-816) for the panel with nodes id0, id1,
// id2 of xs: x1 along edge 0-1, x2 toward node 2 and orthogonal to x1, nm = x1 x x2;
// the scalings by 1/base and 1/height in double, stored float; contraction off.
__device__ __forceinline__ void panel_basis_f(const float* const xs[3], uint32_t id0, uint32_t id1, uint32_t id2,
                                              float* x1, float* x2, float* nm) {
#pragma clang fp contract(off)
#pragma unroll
  for (int j = 0; j < 3; ++j) x1[j] = xs[j][id1] - xs[j][id0];
  const float base = sqrtf(x1[0] * x1[0] + x1[1] * x1[1] + x1[2] * x1[2]);
#pragma unroll
  for (int j = 0; j < 3; ++j) x1[j] = (float)((double)x1[j] * (1.0 / (double)base));
#pragma unroll
  for (int j = 0; j < 3; ++j) x2[j] = xs[j][id2] - xs[j][id0];
  const float dp = x2[0] * x1[0] + x2[1] * x1[1] + x2[2] * x1[2];
#pragma unroll
  for (int j = 0; j < 3; ++j) x2[j] = x2[j] - dp * x1[j];
  const float height = sqrtf(x2[0] * x2[0] + x2[1] * x2[1] + x2[2] * x2[2]);
#pragma unroll
  for (int j = 0; j < 3; ++j) x2[j] = (float)((double)x2[j] * (1.0 / (double)height));
  nm[0] = x1[1] * x2[2] - x1[2] * x2[1];
  nm[1] = x1[2] * x2[0] - x1[0] * x2[2];
  nm[2] = x1[0] * x2[1] - x1[1] * x2[0];
}

struct PanArgs {
  const float* nx; const float* ny; const float* nz;    // nodes
  const uint32_t* idx;                                  // 3 per panel
  const float* tsx; const float* tsy; const float* tsz; // absolute vortex strength
  const float* area;
  const float* ss;                                      // sheet source strength or nullptr
  int64_t np;
  const float* tx; const float* ty; const float* tz;    // targets
  void* u; void* v; void* w; void* g[9];                // float (FP32/COMP) or double (FP64) outputs
  int64_t nt;
  int32_t tile;
  int32_t maxlev;
  int32_t* near_list;          // 7-point rule: targets with a near pair (or nullptr)
  unsigned int* near_count;
};

constexpr int PAN_F4 = 6;   // float4 per staged panel

// exact reference arithmetic for the tree decisions (Kernels.h:1140-1160)
template <class R>
__device__ __forceinline__ void centroid_exact(R x0, R y0, R z0, R x1, R y1, R z1, R x2, R y2, R z2,
                                               R& cx, R& cy, R& cz) {
#pragma clang fp contract(off)
  cx = (x0 + x1 + x2) / R(3);
  cy = (y0 + y1 + y2) / R(3);
  cz = (z0 + z1 + z2) / R(3);
}
template <class R>
__device__ __forceinline__ R dist_exact(R tx, R ty, R tz, R cx, R cy, R cz) {
#pragma clang fp contract(off)
  const R ex = tx - cx, ey = ty - cy, ez = tz - cz;
  return sqrt(ex * ex + ey * ey + ez * ez);
}
template <class R>
__device__ __forceinline__ R mid(R a, R b) {
#pragma clang fp contract(off)
  return R(0.5) * (a + b);
}

// child c of triangle (v0, v1, v2) over the six nodes [v0, m01, v1, m02, m12, v2]:
// {0,1,3}, {1,2,4}, {1,4,3}, {3,4,5}   (Kernels.h:1185-1192)
template <class R>
__device__ __forceinline__ void child_of(R* a, int c) {
  R n1[3], n3[3], n4[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    n1[d] = mid(a[d], a[3 + d]);
    n3[d] = mid(a[d], a[6 + d]);
    n4[d] = mid(a[3 + d], a[6 + d]);
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const R v0 = a[d], v1 = a[3 + d], v2 = a[6 + d];
    const R p = c == 0 ? v0 : (c == 3 ? n3[d] : n1[d]);
    const R q = c == 0 ? n1[d] : (c == 1 ? v1 : n4[d]);
    const R r = c == 0 ? n3[d] : (c == 1 ? n4[d] : (c == 2 ? n3[d] : v2));
    a[d] = p; a[3 + d] = q; a[6 + d] = r;
  }
}

// leaf: singular vortex (g) + source (gs) point at c acting on target t.
// acc: [0..2] u, [3..11] sum B d_j (g x d + gs d)_i, [12..14] sum K g, [15] sum K gs
template <class R, bool GRAD>
__device__ __forceinline__ void leaf_acc(R tx, R ty, R tz, R cx, R cy, R cz, R gx, R gy, R gz, R gs,
                                         R* acc) {
  const R dx = tx - cx, dy = ty - cy, dz = tz - cz;
  R K, B;
  if constexpr (sizeof(R) == 4) {
    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float ir = frsq(d2);
    const float ir2 = ir * ir;
    K = ir2 * ir;
    B = K * (ir2 * -3.0f);
  } else {
    const double d2 = dx * dx + dy * dy + dz * dz;
    K = 1.0 / (d2 * sqrt(d2));
    B = -3.0 * K / d2;
  }
  const R cvx = dz * gy - dy * gz + dx * gs;
  const R cvy = dx * gz - dz * gx + dy * gs;
  const R cvz = dy * gx - dx * gy + dz * gs;
  acc[0] += K * cvx; acc[1] += K * cvy; acc[2] += K * cvz;
  if (GRAD) {
    const R bx = B * cvx, by = B * cvy, bz = B * cvz;
    acc[3] += dx * bx; acc[4] += dx * by; acc[5] += dx * bz;
    acc[6] += dy * bx; acc[7] += dy * by; acc[8] += dy * bz;
    acc[9] += dz * bx; acc[10] += dz * by; acc[11] += dz * bz;
    acc[12] += K * gx; acc[13] += K * gy; acc[14] += K * gz;
    acc[15] += K * gs;
  }
}

// gradient from the 16 accumulators: dyad + skew(W) + diag(S)
template <class R>
__device__ __forceinline__ void grad_of(const R* a, R* g) {
  g[0] = a[3] + a[15];
  g[1] = a[4] + a[14];
  g[2] = a[5] - a[13];
  g[3] = a[6] - a[14];
  g[4] = a[7] + a[15];
  g[5] = a[8] + a[12];
  g[6] = a[9] + a[13];
  g[7] = a[10] - a[12];
  g[8] = a[11] + a[15];
}

// a panel's record (written once per call by k_pan_geom, staged into LDS by k_pan):
//   f0 = {x0,y0,z0,x1}  f1 = {y1,z1,x2,y2}
//   f2 = {level-0 absolute vortex strength (ts/area)*area (3 floats),
//         level-0 absolute source strength ss*area}
//   f3 = {cx, cy, cz (root centroid), far2}   f4 = {z2, area, sqrt(area), near2}
//   f5 = {tsx, tsy, tsz, ss} (raw inputs)
// so that a surely separated pair reads f2 and f3 only.
// The reference passes the sheet strength ts/area and multiplies it back by the
// area at level 0 (Influence.h:737, Kernels.h:1136-1146); that round trip is kept.
// squared-distance bands of the root decision (the reference's float test
// dist > 4 sqrt(area)): d2 > far2 is surely a leaf, d2 < near2 surely not.
// The bands are 1e-6 of the distance either side of 4 sqrt(area), about five
// times the rounding of the fast d2 and of the reference's distance together.
__device__ __forceinline__ void root_bands(float sqa, float& far2, float& near2) {
  const float lim = sqa * 4.0f;
  const float lh = lim * 1.000001f, ll = lim * 0.999999f;
  far2 = lh * lh;
  near2 = ll * ll;
}

__device__ __forceinline__ void panel_record(const PanArgs& a, int64_t p, float4* s) {
  {
    const uint32_t i0 = a.idx[3 * p], i1 = a.idx[3 * p + 1], i2 = a.idx[3 * p + 2];
    const float x0 = a.nx[i0], y0 = a.ny[i0], z0 = a.nz[i0];
    const float x1 = a.nx[i1], y1 = a.ny[i1], z1 = a.nz[i1];
    const float x2 = a.nx[i2], y2 = a.ny[i2], z2 = a.nz[i2];
    const float ar = a.area[p];
    const float tx = a.tsx[p], ty = a.tsy[p], tz = a.tsz[p];
    const float sg = a.ss ? a.ss[p] : 0.0f;
    float cx, cy, cz, gx, gy, gz, gs;
    centroid_exact(x0, y0, z0, x1, y1, z1, x2, y2, z2, cx, cy, cz);
    {
#pragma clang fp contract(off)
      gx = (tx / ar) * ar; gy = (ty / ar) * ar; gz = (tz / ar) * ar; gs = sg * ar;
    }
    s[0] = make_float4(x0, y0, z0, x1);
    s[1] = make_float4(y1, z1, x2, y2);
    s[2] = make_float4(gx, gy, gz, gs);
    float far2, near2;
    root_bands(sqrtf(ar), far2, near2);
    s[3] = make_float4(cx, cy, cz, far2);
    s[4] = make_float4(z2, ar, sqrtf(ar), near2);
    s[5] = make_float4(tx, ty, tz, sg);
  }
}

// copy the records of panels [p0, p0 + cnt) into LDS
__device__ __forceinline__ void stage_panels(const float4* rec, int64_t p0, int cnt, float4* sh) {
  const float4* r = rec + PAN_F4 * p0;
  for (int f = threadIdx.x; f < PAN_F4 * cnt; f += BLOCK) sh[f] = r[f];
}

// the records of all panels in global memory, for the root and near kernels, and
// (cen, when given) each panel's root centroid with its far band, the near
// kernel's prefilter
static __global__ __launch_bounds__(BLOCK) void k_pan_geom(PanArgs a, float4* rec, float4* cen) {
  const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (p >= a.np) return;
  float4* s = rec + PAN_F4 * p;
  panel_record(a, p, s);
  if (cen) cen[p] = s[3];
}

// a panel's root quantities: level-0 absolute strengths (Kernels.h:1136-1146),
// root centroid and 4 sqrt(area) with its squared bands (FP32 only)
struct RootLim {
  float lim, far2, near2;
};
template <class R>
__device__ __forceinline__ void panel_root(const float4* s, R& gx, R& gy, R& gz, R& gs, R& cx0, R& cy0, R& cz0,
                                           R& lim, RootLim& rl) {
  const float4 f0 = s[0], f1 = s[1], f2 = s[2], f3 = s[3], f4 = s[4];
  const R ar = f4.y;
  if constexpr (sizeof(R) == 4) {
    gx = f2.x; gy = f2.y; gz = f2.z; gs = f2.w;
    cx0 = f3.x; cy0 = f3.y; cz0 = f3.z;
    lim = f4.z * 4.0f;
    rl = RootLim{lim, f3.w, f4.w};
  } else {
    rl = RootLim{0.0f, 0.0f, 0.0f};
    const float4 f5 = s[5];
    {
#pragma clang fp contract(off)
      gx = (R(f5.x) / ar) * ar; gy = (R(f5.y) / ar) * ar; gz = (R(f5.z) / ar) * ar; gs = R(f5.w) * ar;
    }
    centroid_exact<R>(f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w, f4.x, cx0, cy0, cz0);
    lim = sqrt(ar) * R(4);
  }
}

// the root decision of one (target, panel) pair: the reference's float test
// (dist > 4 sqrt(area)), taken from the fused d2 against the bands of root_bands
// when unambiguous and from the exact expression otherwise (both start from the
// same float differences); every caller takes it through here, so the far and
// near kernels agree on it.  The empty volatile asm keeps the exact expression
// behind its branch: without it the compiler evaluates the IEEE square root for
// every pair and selects.
template <class R>
__device__ __forceinline__ bool root_leaf(R tx, R ty, R tz, R cx0, R cy0, R cz0, R lim, const RootLim& rl,
                                          int maxlev, bool active) {
  bool leaf0;
  if constexpr (sizeof(R) == 4) {
    const float ex = tx - cx0, ey = ty - cy0, ez = tz - cz0;
    const float d2 = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
    const bool sure_far = d2 > rl.far2;
    const bool sure_near = d2 < rl.near2;
    leaf0 = sure_far;
    const bool unsure = !sure_far && !sure_near;
    if (wave_any(unsure && active)) {
      asm volatile("" ::: "memory");
      if (unsure) leaf0 = dist_exact<float>(tx, ty, tz, cx0, cy0, cz0) > rl.lim;
    }
  } else {
    leaf0 = dist_exact<double>(tx, ty, tz, cx0, cy0, cz0) > lim;
  }
  return leaf0 || maxlev == 0;
}

// one (target, panel) tree walk, accumulating into acc
template <class R, bool GRAD>
__device__ __forceinline__ void walk_panel(R tx, R ty, R tz, const float4* s, int maxlev, bool active, R* acc) {
  const float4 f0 = s[0], f1 = s[1], f4 = s[4];
  const R ar = f4.y;
  R gx, gy, gz, gs, cx0, cy0, cz0, lim;
  RootLim rl;
  panel_root<R>(s, gx, gy, gz, gs, cx0, cy0, cz0, lim, rl);
  const bool leaf0 = root_leaf<R>(tx, ty, tz, cx0, cy0, cz0, lim, rl, maxlev, active);
  if (!wave_any(!leaf0 && active)) {
    if (active) leaf_acc<R, GRAD>(tx, ty, tz, cx0, cy0, cz0, gx, gy, gz, gs, acc);
    return;
  }
  // depth-first walk; lev = 0 is the root (already decided above)
  int lev = 0, path = 0;
  bool done = !active;
  bool first = true;
  while (wave_any(!done)) {
    if (!done) {
      bool leaf;
      R cx, cy, cz;
      R scale = R(1);
      if (first) {
        leaf = leaf0;
        cx = cx0; cy = cy0; cz = cz0;
      } else {
        R t[9] = {R(f0.x), R(f0.y), R(f0.z), R(f0.w), R(f1.x), R(f1.y), R(f1.z), R(f1.w), R(f4.x)};
        R sa = ar;
        for (int l = 1; l <= lev; ++l) {
          child_of<R>(t, (path >> (2 * (l - 1))) & 3);
          sa = R(0.25) * sa;
          scale = R(0.25) * scale;
        }
        centroid_exact<R>(t[0], t[1], t[2], t[3], t[4], t[5], t[6], t[7], t[8], cx, cy, cz);
        leaf = lev == maxlev || dist_exact<R>(tx, ty, tz, cx, cy, cz) > sqrt(sa) * R(4);
      }
      first = false;
      if (leaf) {
        if (lev == 0) leaf_acc<R, GRAD>(tx, ty, tz, cx, cy, cz, gx, gy, gz, gs, acc);
        else leaf_acc<R, GRAD>(tx, ty, tz, cx, cy, cz, gx * scale, gy * scale, gz * scale, gs * scale, acc);
        // advance to the next node in preorder after this subtree
        while (lev > 0) {
          const int sh = 2 * (lev - 1);
          const int c = (path >> sh) & 3;
          if (c < 3) { path += 1 << sh; break; }
          path &= ~(3 << sh);
          --lev;
        }
        if (lev == 0) done = true;
      } else {
        ++lev;   // first child: digit 0
      }
    }
  }
}

// the root decision of one pair: a leaf is evaluated, a recursed pair noted
template <class R, bool GRAD>
__device__ __forceinline__ void pan_root_pair(const float4* s, R tx, R ty, R tz, int maxlev, bool active, R* acc,
                                              bool& has_near) {
  R gx, gy, gz, gs, cx0, cy0, cz0, lim;
  RootLim rl;
  panel_root<R>(s, gx, gy, gz, gs, cx0, cy0, cz0, lim, rl);
  if (root_leaf<R>(tx, ty, tz, cx0, cy0, cz0, lim, rl, maxlev, active)) {
    if (active) leaf_acc<R, GRAD>(tx, ty, tz, cx0, cy0, cz0, gx, gy, gz, gs, acc);
  } else {
    has_near = true;
  }
}

constexpr int PAN_UNROLL = 4;   // panels per step of k_pan's separated-pair loop

// rec: the records of all panels (k_pan_geom).  At least 5 waves per SIMD (96
// VGPRs; the FP64 gradient sums spill once per tile, outside the panel loop): on
// the icosphere 1.7% faster than the 103 VGPRs and 4 waves the compiler chooses,
// and 6 waves (80 VGPRs) 2% slower.
template <bool GRAD, int PREC>
__global__ __launch_bounds__(BLOCK, 5) void k_pan(PanArgs a, const float4* rec) {
  using R = typename std::conditional<PREC == PREC_FP64, double, float>::type;
  extern __shared__ float4 shp[];
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const bool active = i < a.nt;
  const int64_t ic = active ? i : a.nt - 1;
  const R tx = a.tx[ic], ty = a.ty[ic], tz = a.tz[ic];
  R acc[16];
  float hi[3] = {0.f, 0.f, 0.f}, lo[3] = {0.f, 0.f, 0.f};
  double gd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool has_near = false;
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c] = R(0);
  for (int64_t p0 = 0; p0 < a.np; p0 += a.tile) {
    const int cnt = (int)((a.np - p0) < a.tile ? (a.np - p0) : a.tile);
    __syncthreads();
    stage_panels(rec, p0, cnt, shp);
    __syncthreads();
    int j = 0;
    if constexpr (sizeof(R) == 4) {
      // PAN_UNROLL panels at a time while every lane's pairs are surely leaves at the
      // root (d2 beyond the far band): one wave-uniform branch, the leaves with no
      // per-pair mask, the LDS address in a vector register (made so by an opaque
      // zero) with immediate offsets.  The same leaves, in the same order and with
      // the same bits, as the per-pair path below; an idle lane's sums are dropped.
      int vz;
      asm volatile("v_mov_b32 %0, 0" : "=v"(vz));
      const float4* q = shp + vz;
      for (; j + PAN_UNROLL <= cnt; j += PAN_UNROLL, q += PAN_F4 * PAN_UNROLL) {
        float4 g[PAN_UNROLL], c[PAN_UNROLL];
        uint64_t near = 0;   // lanes with a pair not surely a leaf, as a scalar mask
#pragma unroll
        for (int u = 0; u < PAN_UNROLL; ++u) {
          g[u] = q[PAN_F4 * u + 2];
          c[u] = q[PAN_F4 * u + 3];
          const float ex = tx - c[u].x, ey = ty - c[u].y, ez = tz - c[u].z;
          near |= __builtin_amdgcn_ballot_w64(!(fmaf(ez, ez, fmaf(ey, ey, ex * ex)) > c[u].w));
        }
        if (near == 0) {
#pragma unroll
          for (int u = 0; u < PAN_UNROLL; ++u)
            leaf_acc<R, GRAD>(tx, ty, tz, c[u].x, c[u].y, c[u].z, g[u].x, g[u].y, g[u].z, g[u].w, acc);
        } else {
#pragma unroll
          for (int u = 0; u < PAN_UNROLL; ++u) pan_root_pair<R, GRAD>(q + PAN_F4 * u, tx, ty, tz, a.maxlev, active,
                                                                      acc, has_near);
        }
      }
    }
    for (; j < cnt; ++j) pan_root_pair<R, GRAD>(shp + PAN_F4 * j, tx, ty, tz, a.maxlev, active, acc, has_near);
    if (PREC == PREC_COMP) {
      two_sum(hi[0], lo[0], (float)acc[0]);
      two_sum(hi[1], lo[1], (float)acc[1]);
      two_sum(hi[2], lo[2], (float)acc[2]);
      if (GRAD) {
        R g[9];
        grad_of<R>(acc, g);
#pragma unroll
        for (int k = 0; k < 9; ++k) gd[k] += (double)g[k];
      }
#pragma unroll
      for (int c = 0; c < 16; ++c) acc[c] = R(0);
    }
  }
  if (!active) return;
  if (has_near) a.near_list[atomicAdd(a.near_count, 1u)] = (int32_t)i;
  if (PREC == PREC_FP64) {
    double* u = (double*)a.u; double* v = (double*)a.v; double* w = (double*)a.w;
    u[i] += acc[0]; v[i] += acc[1]; w[i] += acc[2];
    if (GRAD) {
      R g[9];
      grad_of<R>(acc, g);
#pragma unroll
      for (int k = 0; k < 9; ++k) ((double*)a.g[k])[i] += g[k];
    }
  } else if (PREC == PREC_FP32) {
    float* u = (float*)a.u; float* v = (float*)a.v; float* w = (float*)a.w;
    u[i] += (float)acc[0]; v[i] += (float)acc[1]; w[i] += (float)acc[2];
    if (GRAD) {
      R g[9];
      grad_of<R>(acc, g);
#pragma unroll
      for (int k = 0; k < 9; ++k) ((float*)a.g[k])[i] += (float)g[k];
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

// The pairs of the listed targets that the root decision sends into the
// recursion, added to the outputs after k_pan.  One wave per 4 listed targets, one
// per 16-lane quarter: lane l of a quarter takes panels l, l + 16, ..., so the
// quarters load the same panels in the same instruction.  A panel whose centroid
// lies beyond its far band (cen: the fused d2 of root_leaf against the same band)
// is a leaf at the root, which k_pan evaluated; the others are decided as k_pan
// did and the near ones walked (Kernels.h:1028-1211).  A fixed butterfly adds each
// quarter's sums in FP64, so results do not depend on scheduling.  Grid-stride
// over the list, whose length k_pan left in near_count.
template <bool GRAD, int PREC>
__global__ __launch_bounds__(BLOCK) void k_pan_near(PanArgs a, const float4* rec, const float4* cen) {
  using R = typename std::conditional<PREC == PREC_FP64, double, float>::type;
  const int lane = threadIdx.x & 63;
  const int ql = lane & 15;
  const int64_t wave = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * BLOCK) >> 6;
  const int64_t listed = *a.near_count;
  const int64_t groups = (listed + 3) / 4;
  for (int64_t e = wave; e < groups; e += nwaves) {
    const int64_t li = 4 * e + (lane >> 4);
    const bool has = li < listed;
    const int64_t i = a.near_list[has ? li : 4 * e];
    const R tx = a.tx[i], ty = a.ty[i], tz = a.tz[i];
    R acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = R(0);
    if (has) {
      // four panels per step: their prefilter loads are issued together, then the
      // candidates are decided and walked in ascending order
      for (int64_t p = ql; p < a.np; p += 64) {
        unsigned m = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int64_t pk = p + 16 * k;
          if (pk >= a.np) continue;
          if constexpr (sizeof(R) == 4) {
            const float4 c = cen[pk];
            const float ex = tx - c.x, ey = ty - c.y, ez = tz - c.z;
            if (fmaf(ez, ez, fmaf(ey, ey, ex * ex)) > c.w) continue;
          }
          m |= 1u << k;
        }
        while (m) {
          const int k = __builtin_ctz(m);
          m &= m - 1;
          const float4* s = rec + PAN_F4 * (p + 16 * k);
          R gx, gy, gz, gs, cx0, cy0, cz0, lim;
          RootLim rl;
          panel_root<R>(s, gx, gy, gz, gs, cx0, cy0, cz0, lim, rl);
          if (!root_leaf<R>(tx, ty, tz, cx0, cy0, cz0, lim, rl, a.maxlev, true))
            walk_panel<R, GRAD>(tx, ty, tz, s, a.maxlev, true, acc);
        }
      }
    }
    double u[3], g[9];
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = quarter_sum((double)acc[d]);
    if (GRAD) {
      R gl[9];
      grad_of<R>(acc, gl);
#pragma unroll
      for (int k = 0; k < 9; ++k) g[k] = quarter_sum((double)gl[k]);
    }
    if (ql != 0 || !has) continue;
    if (PREC == PREC_FP64) {
      double* uu = (double*)a.u; double* vv = (double*)a.v; double* ww = (double*)a.w;
      uu[i] += u[0]; vv[i] += u[1]; ww[i] += u[2];
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
}

}  // namespace o3d
