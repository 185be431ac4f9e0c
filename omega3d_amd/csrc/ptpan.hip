// ptpan.hip -- particles -> panels: Omega3D's points_affect_panels
// (Influence.h:1107-1221), the velocity of each panel induced by vortex particles,
// as used for the boundary-element right-hand side.
//
// The reference runs the panel kernel backwards: for target panel i and particle
// j it evaluates rkernel_2vs_0p with the panel carrying alpha_j / area_i (divided
// in S, multiplied back by the area at level 0), the particle's position as the
// "target", and SUBTRACTS the sum from the panel's velocity (Influence.h:1196-1210).
// Velocity only: the reference asserts on gradient requests (:1111).
//
// GPU form: k_ptpan takes one lane per target panel (its corners, area,
// sqrt(area) and exact centroid in registers) and a slice of the particles per
// workgroup row, staged per tile in LDS and broadcast (a boundary has too few
// panels to fill the GPU otherwise); it takes the root decision of every pair and
// evaluates the one-leaf pairs.  k_ptpan_combine adds the slices in order and
// lists the panels that have particles near enough for the recursion; k_ptpan
// also keeps the first few such particles of each (slice, panel) in ascending
// order.  k_ptpan_near gives each listed panel one wave whose lanes walk those
// pairs with the depth-first tree walk of panels_affect_points (pan_kernels.hpp),
// so the subdivision decisions are the reference's.
#include "o3d_api_internal.hpp"
#include "pan_kernels.hpp"

#include <algorithm>
#include <chrono>
#include <vector>

namespace o3d {

struct PtPanArgs {
  const float* sx; const float* sy; const float* sz;    // particles
  const float* ax; const float* ay; const float* az;
  int64_t ns;
  const float* nx; const float* ny; const float* nz;    // panel nodes
  const uint32_t* idx;
  const float* area;
  int64_t np;
  void* u; void* v; void* w;   // float (decremented in place) or double (FP64: zeroed scratch)
  int32_t tile;
  int32_t maxlev;
  int32_t* near_list;          // panels with a pair that needs the recursion
  unsigned int* near_count;
  double* part;                // per slice: 3 sums per panel
  int32_t* cnt;                // [panel][slice]: particles that need the recursion
  int32_t* ent;                // [panel][slice][PTPAN_CAP]: the first of them, ascending
  int64_t slen;                // particles per slice
  int32_t nslice;
};

// particles kept per (slice, panel) for the near kernel; a slice with more is
// scanned again there
constexpr int PTPAN_CAP = 8;

// this lane's panel in the staged-panel layout of pan_kernels.hpp (the strength
// fields are filled per particle)
__device__ __forceinline__ void ptpan_record(const PtPanArgs& a, int64_t i, float4* s) {
  const uint32_t i0 = a.idx[3 * i], i1 = a.idx[3 * i + 1], i2 = a.idx[3 * i + 2];
  const float x0 = a.nx[i0], y0 = a.ny[i0], z0 = a.nz[i0];
  const float x1 = a.nx[i1], y1 = a.ny[i1], z1 = a.nz[i1];
  const float x2 = a.nx[i2], y2 = a.ny[i2], z2 = a.nz[i2];
  const float ar = a.area[i];
  float cx, cy, cz;
  centroid_exact(x0, y0, z0, x1, y1, z1, x2, y2, z2, cx, cy, cz);
  s[0] = make_float4(x0, y0, z0, x1);
  s[1] = make_float4(y1, z1, x2, y2);
  s[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float far2, near2;
  root_bands(sqrtf(ar), far2, near2);
  s[3] = make_float4(cx, cy, cz, far2);
  s[4] = make_float4(z2, ar, sqrtf(ar), near2);
  s[5] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// particle j's strength on panel record s at level 0.  FP64: the walk forms
// (alpha / area) * area in double from s[5].  FP32 and compensated: alpha itself;
// the reference's float round trip (alpha / area) * area (Influence.h:1198,
// Kernels.h:1136-1146) moves it by at most one float ulp and costs three IEEE
// divisions per pair.
template <int PREC>
__device__ __forceinline__ void ptpan_strength(float4* s, float4 A) {
  if (PREC == PREC_FP64) s[5] = make_float4(A.x, A.y, A.z, 0.0f);
  else { s[2].x = A.x; s[2].y = A.y; s[2].z = A.z; }
}

// One lane per target panel, particles [y*slen, (y+1)*slen) of slice y =
// blockIdx.y staged per tile in LDS: the root decision of every (panel, particle)
// pair -- the reference's, through root_leaf -- and the pairs that are leaves at
// the root.  Each slice leaves its sum (FP64) in part[y][0..2][panel], the count
// of pairs that need the recursion in cnt and the first PTPAN_CAP of their
// particles in ent; k_ptpan_combine adds the slices in order.
constexpr int PTPAN_UNROLL = 4;   // particles per step of k_ptpan's separated-pair loop

template <int PREC>
__global__ __launch_bounds__(BLOCK) void k_ptpan(PtPanArgs a) {
  using R = typename std::conditional<PREC == PREC_FP64, double, float>::type;
  extern __shared__ float4 shs[];
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const bool active = i < a.np;
  const int64_t ic = active ? i : a.np - 1;
  float4 s[PAN_F4];
  ptpan_record(a, ic, s);
  const float ar = s[4].y;
  R acc[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c] = R(0);
  float hi[3] = {0.f, 0.f, 0.f}, lo[3] = {0.f, 0.f, 0.f};
  int nnear = 0;
  int32_t* ent = a.ent + ((int64_t)ic * a.nslice + blockIdx.y) * PTPAN_CAP;
  R gx, gy, gz, gs, cx0, cy0, cz0, lim;
  RootLim rl;
  panel_root<R>(s, gx, gy, gz, gs, cx0, cy0, cz0, lim, rl);
  const int64_t q0 = (int64_t)blockIdx.y * a.slen;
  const int64_t q1 = q0 + a.slen < a.ns ? q0 + a.slen : a.ns;

  for (int64_t p0 = q0; p0 < q1; p0 += a.tile) {
    const int cnt = (int)((q1 - p0) < a.tile ? (q1 - p0) : a.tile);
    __syncthreads();
    for (int j = threadIdx.x; j < cnt; j += BLOCK) {
      const int64_t g = p0 + j;
      shs[2 * j] = make_float4(a.sx[g], a.sy[g], a.sz[g], 0.0f);
      shs[2 * j + 1] = make_float4(a.ax[g], a.ay[g], a.az[g], 0.0f);
    }
    __syncthreads();
    // the root decision of particle j of the tile: a leaf is evaluated, a recursed
    // pair listed
    auto pair = [&](const float4 P, const float4 A, int j) {
      if (root_leaf<R>(P.x, P.y, P.z, cx0, cy0, cz0, lim, rl, a.maxlev, active)) {
        if (!active) return;
        R ax, ay, az;
        if (PREC == PREC_FP64) {
#pragma clang fp contract(off)
          ax = (R(A.x) / R(ar)) * R(ar); ay = (R(A.y) / R(ar)) * R(ar); az = (R(A.z) / R(ar)) * R(ar);
        } else {
          ax = A.x; ay = A.y; az = A.z;
        }
        leaf_acc<R, false>(P.x, P.y, P.z, cx0, cy0, cz0, ax, ay, az, R(0), acc);
      } else if (active) {
        if (nnear < PTPAN_CAP) ent[nnear] = (int32_t)(p0 + j);
        ++nnear;
      }
    };
    int j = 0;
    if constexpr (sizeof(R) == 4) {
      // PTPAN_UNROLL particles at a time while they are surely beyond every lane's
      // far band: one scalar branch on the OR of the ballots, the leaves with no
      // per-pair mask, the LDS address in a vector register (made so by an opaque
      // zero); the same leaves in the same order as the per-pair path, and an idle
      // lane's sums are dropped
      int vz;
      asm volatile("v_mov_b32 %0, 0" : "=v"(vz));
      const float4* q = shs + vz;
      for (; j + PTPAN_UNROLL <= cnt; j += PTPAN_UNROLL, q += 2 * PTPAN_UNROLL) {
        float4 P[PTPAN_UNROLL], A[PTPAN_UNROLL];
        uint64_t near = 0;
#pragma unroll
        for (int u = 0; u < PTPAN_UNROLL; ++u) {
          P[u] = q[2 * u];
          A[u] = q[2 * u + 1];
          const float ex = P[u].x - cx0, ey = P[u].y - cy0, ez = P[u].z - cz0;
          near |= __builtin_amdgcn_ballot_w64(!(fmaf(ez, ez, fmaf(ey, ey, ex * ex)) > rl.far2));
        }
        if (near == 0) {
#pragma unroll
          for (int u = 0; u < PTPAN_UNROLL; ++u)
            leaf_acc<R, false>(P[u].x, P[u].y, P[u].z, cx0, cy0, cz0, A[u].x, A[u].y, A[u].z, R(0), acc);
        } else {
#pragma unroll
          for (int u = 0; u < PTPAN_UNROLL; ++u) pair(P[u], A[u], j + u);
        }
      }
    }
    for (; j < cnt; ++j) pair(shs[2 * j], shs[2 * j + 1], j);
    if (PREC == PREC_COMP) {
      two_sum(hi[0], lo[0], (float)acc[0]);
      two_sum(hi[1], lo[1], (float)acc[1]);
      two_sum(hi[2], lo[2], (float)acc[2]);
      acc[0] = acc[1] = acc[2] = R(0);
    }
  }
  if (!active) return;
  double* pt = a.part + (int64_t)blockIdx.y * 3 * a.np + i;
  if (PREC == PREC_COMP) {
#pragma unroll
    for (int d = 0; d < 3; ++d) pt[d * a.np] = (double)hi[d] + (double)lo[d];
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) pt[d * a.np] = (double)acc[d];
  }
  a.cnt[i * a.nslice + blockIdx.y] = nnear;
}

// per panel: the slices' sums added in slice order and subtracted from the
// panel's velocity (Influence.h:1206-1210); a panel with a pair that needs the
// recursion in any slice is listed for k_ptpan_near
template <int PREC>
__global__ __launch_bounds__(BLOCK) void k_ptpan_combine(PtPanArgs a) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= a.np) return;
  double u[3] = {0.0, 0.0, 0.0};
  bool near = false;
  for (int y = 0; y < a.nslice; ++y) {
    const double* pt = a.part + (int64_t)y * 3 * a.np + i;
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] += pt[d * a.np];
    near = near || a.cnt[i * a.nslice + y] != 0;
  }
  if (near) a.near_list[atomicAdd(a.near_count, 1u)] = (int32_t)i;
  if (PREC == PREC_FP64) {
    ((double*)a.u)[i] = -u[0];
    ((double*)a.v)[i] = -u[1];
    ((double*)a.w)[i] = -u[2];
  } else {
    float* uu = (float*)a.u; float* vv = (float*)a.v; float* ww = (float*)a.w;
    uu[i] = (float)((double)uu[i] - u[0]);
    vv[i] = (float)((double)vv[i] - u[1]);
    ww[i] = (float)((double)ww[i] - u[2]);
  }
}

// The near pairs of the listed panels, after k_ptpan_combine.  One wave per listed
// panel: lane l takes slices l, l + 64, ... and walks the particles k_ptpan kept
// for its slice, in ascending order; a slice with more than PTPAN_CAP such
// particles is scanned again by the whole wave (lane l takes its particles l,
// l + 64, ..., decides each root pair as k_ptpan did).  The walk is the depth-first
// tree walk of panels_affect_points (Kernels.h:1028-1211).  A fixed butterfly
// adds the lanes' sums in FP64, subtracted from the panel's velocity.
template <int PREC>
__global__ __launch_bounds__(BLOCK) void k_ptpan_near(PtPanArgs a) {
  using R = typename std::conditional<PREC == PREC_FP64, double, float>::type;
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * BLOCK) >> 6;
  const int64_t listed = *a.near_count;
  for (int64_t e = wave; e < listed; e += nwaves) {
    const int64_t i = a.near_list[e];
    float4 s[PAN_F4];
    ptpan_record(a, i, s);
    R acc[3] = {R(0), R(0), R(0)};
    // one walk of particle p on this panel, added to acc
    auto walk = [&](int64_t p) {
      const float px = a.sx[p], py = a.sy[p], pz = a.sz[p];
      ptpan_strength<PREC>(s, make_float4(a.ax[p], a.ay[p], a.az[p], 0.0f));
      R t16[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) t16[c] = R(0);
      walk_panel<R, false>(px, py, pz, s, a.maxlev, true, t16);
      acc[0] += t16[0]; acc[1] += t16[1]; acc[2] += t16[2];
    };
    const int32_t* cnt = a.cnt + i * a.nslice;
    for (int64_t y0 = 0; y0 < a.nslice; y0 += 64) {
      const int64_t y = y0 + lane;
      const int n = y < a.nslice ? cnt[y] : 0;
      if (n <= PTPAN_CAP) {
        const int32_t* ent = a.ent + (i * a.nslice + y) * PTPAN_CAP;
        for (int c = 0; c < n; ++c) walk(ent[c]);
      }
      // slices whose particles did not all fit: the whole wave, one slice at a time
      uint64_t over = __builtin_amdgcn_ballot_w64(n > PTPAN_CAP);
      while (over) {
        const int l = __builtin_ctzll(over);
        over &= over - 1;
        const int64_t q0 = (y0 + l) * a.slen;
        const int64_t q1 = q0 + a.slen < a.ns ? q0 + a.slen : a.ns;
        for (int64_t p = q0 + lane; p < q1; p += 64) {
          R gx, gy, gz, gs, cx0, cy0, cz0, lim;
          RootLim rl;
          panel_root<R>(s, gx, gy, gz, gs, cx0, cy0, cz0, lim, rl);
          if (!root_leaf<R>(a.sx[p], a.sy[p], a.sz[p], cx0, cy0, cz0, lim, rl, a.maxlev, true)) walk(p);
        }
      }
    }
    double u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = wave_sum((double)acc[d]);
    if (lane != 0) continue;
    if (PREC == PREC_FP64) {
      ((double*)a.u)[i] -= u[0];
      ((double*)a.v)[i] -= u[1];
      ((double*)a.w)[i] -= u[2];
    } else {
      float* uu = (float*)a.u; float* vv = (float*)a.v; float* ww = (float*)a.w;
      uu[i] = (float)((double)uu[i] - u[0]);
      vv[i] = (float)((double)vv[i] - u[1]);
      ww[i] = (float)((double)ww[i] - u[2]);
    }
  }
}

__global__ void k_ptpan_zero(unsigned int* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0u;
}

// root pairs over particle slices, their combination, then the near pairs of the
// listed panels; the partial sums, the list and its counter in the stream's
// workspace.  Enough slices of whole tiles for about 16 workgroups per CU.
static int launch_ptpan(PtPanArgs a, int prec, hipStream_t st) {
  if (a.np <= 0 || a.ns <= 0) return O3D_OK;
  if (a.np > INT32_MAX) return fail(O3D_EINVAL, "points_affect_panels takes at most %d panels", INT32_MAX);
  if (a.ns > INT32_MAX) return fail(O3D_EINVAL, "points_affect_panels takes at most %d particles", INT32_MAX);
  const int64_t nbx = (a.np + BLOCK - 1) / BLOCK;
  const int64_t ntiles = (a.ns + a.tile - 1) / a.tile;
  int64_t nsl = (16LL * num_cus() + nbx - 1) / nbx;
  nsl = std::max<int64_t>(1, std::min<int64_t>({nsl, ntiles, 65535}));
  a.slen = ((ntiles + nsl - 1) / nsl) * a.tile;
  a.nslice = (int32_t)((a.ns + a.slen - 1) / a.slen);
  std::lock_guard<std::mutex> lk(stream_ws_mutex());
  void* ws = nullptr;
  const size_t list = (sizeof(int32_t) * (size_t)a.np + 255) & ~(size_t)255;
  const size_t part = sizeof(double) * 3 * (size_t)a.nslice * (size_t)a.np;
  const size_t cnt = (sizeof(int32_t) * (size_t)a.nslice * (size_t)a.np + 255) & ~(size_t)255;
  const size_t ent = sizeof(int32_t) * PTPAN_CAP * (size_t)a.nslice * (size_t)a.np;
  int rc = stream_workspace(st, 256 + list + part + cnt + ent, &ws);
  if (rc != O3D_OK) return rc;
  a.near_count = (unsigned int*)ws;
  a.near_list = (int32_t*)((char*)ws + 256);
  a.part = (double*)((char*)ws + 256 + list);
  a.cnt = (int32_t*)((char*)ws + 256 + list + part);
  a.ent = (int32_t*)((char*)ws + 256 + list + part + cnt);
  const size_t lds = (size_t)a.tile * 2 * sizeof(float4);
  const dim3 grid((unsigned)nbx, (unsigned)a.nslice);
  const dim3 grid1((unsigned)nbx);
  // one wave per listed panel, grid-stride: at most 16 waves per CU
  const dim3 grid_near((unsigned)std::min<int64_t>((a.np + 3) / 4, 4 * (int64_t)num_cus()));
  hipLaunchKernelGGL(k_ptpan_zero, dim3(1), dim3(64), 0, st, a.near_count);
#define O3D_PP(P)                                                                  \
  if (prec == P) {                                                                 \
    hipLaunchKernelGGL((k_ptpan<P>), grid, dim3(BLOCK), lds, st, a);               \
    hipLaunchKernelGGL((k_ptpan_combine<P>), grid1, dim3(BLOCK), 0, st, a);        \
    hipLaunchKernelGGL((k_ptpan_near<P>), grid_near, dim3(BLOCK), 0, st, a);       \
  }
  O3D_PP(PREC_FP32) O3D_PP(PREC_COMP) O3D_PP(PREC_FP64)
#undef O3D_PP
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(O3D_EDEVICE, "points_affect_panels launch failed: %s", hipGetErrorString(e));
  return O3D_OK;
}

static int check_ptpan(const o3d_points* src, const o3d_panels* targ, float* const* vel, const o3d_opts* o) {
  if (o->results != O3D_VELONLY)
    return fail(O3D_EINVAL, "points_affect_panels computes velocity only (the reference asserts on gradients, "
                            "Influence.h:1111)");
  if (!src) return fail(O3D_EINVAL, "src is NULL");
  if (src->n < 0) return fail(O3D_EINVAL, "src: negative count");
  if (src->n > 0)
    for (int d = 0; d < 3; ++d)
      if (!src->x[d] || !src->s[d]) return fail(O3D_EINVAL, "src: position or strength component %d is NULL", d);
  if (!targ) return fail(O3D_EINVAL, "targ panels are NULL");
  if (targ->np < 0 || targ->nn < 0) return fail(O3D_EINVAL, "negative panel or node count");
  if (targ->np > 0) {
    if (targ->nn <= 0) return fail(O3D_EINVAL, "panels without nodes");
    for (int d = 0; d < 3; ++d)
      if (!targ->x[d]) return fail(O3D_EINVAL, "panel node component %d is NULL", d);
    if (!targ->idx || !targ->area) return fail(O3D_EINVAL, "panel indices or areas are NULL");
    if (!vel) return fail(O3D_EINVAL, "panel velocity arrays are NULL");
    for (int d = 0; d < 3; ++d)
      if (!vel[d]) return fail(O3D_EINVAL, "panel velocity component %d is NULL", d);
  }
  if (o->maxlev > 8) return fail(O3D_EINVAL, "maxlev %d: at most 8 levels", o->maxlev);
  return O3D_OK;
}

static PtPanArgs ptpan_args(const o3d_points* src, const o3d_panels* targ, const o3d_opts* o) {
  PtPanArgs a{};
  a.sx = src->x[0]; a.sy = src->x[1]; a.sz = src->x[2];
  a.ax = src->s[0]; a.ay = src->s[1]; a.az = src->s[2];
  a.ns = src->n;
  a.nx = targ->x[0]; a.ny = targ->x[1]; a.nz = targ->x[2];
  a.idx = targ->idx; a.area = targ->area; a.np = targ->np;
  a.tile = o->tile > 0 ? o->tile : 256;
  a.maxlev = o->maxlev < 0 ? 3 : o->maxlev;   // RECURSIVE_LEVELS, Influence.h:23
  return a;
}

// ---- BEM right-hand side of one boundary surface (BEMHelper.h:83-117) ----------
// Per panel, in the reference's own float/double mixing, contraction off:
//   finalize_vels      Surfaces.h:877-887   pu = fs + pu * (0.25/pi)              (double -> float)
//   add_body_motion    Surfaces.h:556-597   pu += -1 * (float)vel                 (float)
//                                           pu += -1 * (rotvel x (c - tc))        (double -> float),
//                                           c - tc = float((float sum)/3.0 - tc)
//   compute_bases      Surfaces.h:766-816   x1 = (n1 - n0)*(1.0/base), x2 = ((n2 - n0) - dp x1)
//                                           *(1.0/height) (scalings in double), norm = x1 x x2
//   vels_to_rhs_panels RHS.h:47-98          rhs = -(pu.x1), -(pu.x2) [, -(pu.norm)]; 1 unknown: -(pu.norm)
struct BemArgs {
  const float* nx; const float* ny; const float* nz;
  const uint32_t* idx;
  float* pu[3];
  int64_t np;
  double fs[3], vel[3], rotvel[3];
  float tc[3];
  int32_t body;
  int32_t nunk;
  float* rhs;
};

__global__ __launch_bounds__(BLOCK) void k_bem_rhs(BemArgs a) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= a.np) return;
  const double factor = 0.25 / M_PI;
  float u[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) u[d] = (float)(a.fs[d] + (double)a.pu[d][i] * factor);
  const uint32_t id0 = a.idx[3 * i], id1 = a.idx[3 * i + 1], id2 = a.idx[3 * i + 2];
  const float* xs[3] = {a.nx, a.ny, a.nz};
  if (a.body) {
    const float f = -1.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = u[d] + f * (float)a.vel[d];
    float c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] = (float)((double)(xs[d][id0] + xs[d][id1] + xs[d][id2]) / 3.0 - (double)a.tc[d]);
    u[0] = (float)((double)u[0] + (double)f * (a.rotvel[1] * c[2] - a.rotvel[2] * c[1]));
    u[1] = (float)((double)u[1] + (double)f * (a.rotvel[2] * c[0] - a.rotvel[0] * c[2]));
    u[2] = (float)((double)u[2] + (double)f * (a.rotvel[0] * c[1] - a.rotvel[1] * c[0]));
  }
  float x1[3], x2[3], nm[3];
  panel_basis_f(xs, id0, id1, id2, x1, x2, nm);
  const float t1 = -(u[0] * x1[0] + u[1] * x1[1] + u[2] * x1[2]);
  const float t2 = -(u[0] * x2[0] + u[1] * x2[1] + u[2] * x2[2]);
  const float tn = -(u[0] * nm[0] + u[1] * nm[1] + u[2] * nm[2]);
#pragma unroll
  for (int d = 0; d < 3; ++d) a.pu[d][i] = u[d];
  if (a.nunk == 1) {
    a.rhs[i] = tn;
  } else if (a.nunk == 2) {
    a.rhs[2 * i] = t1;
    a.rhs[2 * i + 1] = t2;
  } else {
    a.rhs[3 * i] = t1;
    a.rhs[3 * i + 1] = t2;
    a.rhs[3 * i + 2] = tn;
  }
}

__global__ void k_bem_zero(float* p0, float* p1, float* p2, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    p0[i] = 0.f; p1[i] = 0.f; p2[i] = 0.f;
  }
}

static int bem_rhs_args(const o3d_panels* bdry, const o3d_bem_rhs_opts* b, float* const* pu, float* rhs,
                        BemArgs* a) {
  if (!bdry || !b) return fail(O3D_EINVAL, "boundary panels or BEM options are NULL");
  if (b->nunk < 1 || b->nunk > 3) return fail(O3D_EINVAL, "unknowns per panel %d: must be 1, 2 or 3", b->nunk);
  if (bdry->np < 0 || bdry->nn < 0) return fail(O3D_EINVAL, "negative panel or node count");
  if (bdry->np > 0) {
    if (bdry->nn <= 0 || !bdry->idx) return fail(O3D_EINVAL, "panels without nodes or indices");
    for (int d = 0; d < 3; ++d)
      if (!bdry->x[d]) return fail(O3D_EINVAL, "panel node component %d is NULL", d);
    if (!pu || !pu[0] || !pu[1] || !pu[2] || !rhs) return fail(O3D_EINVAL, "panel velocity or rhs arrays are NULL");
  }
  *a = BemArgs{};
  a->nx = bdry->x[0]; a->ny = bdry->x[1]; a->nz = bdry->x[2];
  a->idx = bdry->idx;
  for (int d = 0; d < 3; ++d) {
    a->pu[d] = bdry->np > 0 ? pu[d] : nullptr;
    a->fs[d] = b->fs[d]; a->vel[d] = b->vel[d]; a->rotvel[d] = b->rotvel[d]; a->tc[d] = b->tc[d];
  }
  a->np = bdry->np;
  a->body = b->body ? 1 : 0;
  a->nunk = b->nunk;
  a->rhs = rhs;
  return O3D_OK;
}

static double ptpan_flops(int64_t np, int64_t ns) {
  // one leaf kernel_0vs_0p per (panel, particle) pair plus 3 per panel (Influence.h:1212)
  const int f = o3d_flops_per_interaction(CORE_RM, O3D_VELONLY, 0, 1);
  return (double)np * (3.0 + (double)f * (double)ns);
}

}  // namespace o3d

using namespace o3d;

extern "C" {

int o3d_points_affect_panels(const o3d_points* src, const o3d_panels* targ, float* const vel[3],
                             const o3d_opts* o, o3d_stats* stats) {
  clear_error();
  int rc;
  if ((rc = check_opts(o)) != O3D_OK) return rc;
  if ((rc = check_ptpan(src, targ, vel, o)) != O3D_OK) return rc;
  if (targ->np > 0 && (rc = check_node_indices(targ, "targ")) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t ns = src->n, np = targ->np, nn = targ->nn;
  if (ns > 0 && np > 0) {
    std::lock_guard<std::mutex> lk(pool_mutex());
    hipStream_t st = lib_stream();
    const bool f64 = o->precision == PREC_FP64;
    void *dp, *ds;
    // panels: nodes 3*nn, area np (floats), idx 3*np (uint32), velocities 3*np (float or double)
    const size_t vbytes = (f64 ? sizeof(double) : sizeof(float)) * 3 * np;
    if ((rc = pool_buffer(2, sizeof(float) * (3 * nn + 4 * np) + vbytes + 256, &dp)) != O3D_OK) return rc;
    if ((rc = pool_buffer(3, sizeof(float) * 6 * ns, &ds)) != O3D_OK) return rc;
    float* f = (float*)dp;
    o3d_panels d{};
    d.np = np; d.nn = nn;
    for (int k = 0; k < 3; ++k) d.x[k] = f + k * nn;
    d.area = f + 3 * nn;
    d.idx = (const uint32_t*)(f + 3 * nn + np);
    char* vb = (char*)(f + 3 * nn + 4 * np);
    vb += (256 - ((uintptr_t)vb & 255)) & 255;
    for (int k = 0; k < 3; ++k)
      HIPCHK(hipMemcpyAsync((void*)d.x[k], targ->x[k], sizeof(float) * nn, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync((void*)d.area, targ->area, sizeof(float) * np, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync((void*)d.idx, targ->idx, sizeof(uint32_t) * 3 * np, hipMemcpyHostToDevice, st));
    o3d_points dsrc{};
    dsrc.n = ns;
    float* sf = (float*)ds;
    for (int k = 0; k < 3; ++k) {
      dsrc.x[k] = sf + k * ns;
      dsrc.s[k] = sf + (3 + k) * ns;
      HIPCHK(hipMemcpyAsync((void*)dsrc.x[k], src->x[k], sizeof(float) * ns, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync((void*)dsrc.s[k], src->s[k], sizeof(float) * ns, hipMemcpyHostToDevice, st));
    }
    PtPanArgs a = ptpan_args(&dsrc, &d, o);
    const size_t vsz = f64 ? sizeof(double) : sizeof(float);
    a.u = vb; a.v = vb + vsz * np; a.w = vb + 2 * vsz * np;
    if (!f64)
      for (int k = 0; k < 3; ++k)
        HIPCHK(hipMemcpyAsync(vb + vsz * np * k, vel[k], sizeof(float) * np, hipMemcpyHostToDevice, st));
    if ((rc = launch_ptpan(a, o->precision, st)) != O3D_OK) return rc;
    if (f64) {
      std::vector<double> h((size_t)3 * np);
      HIPCHK(hipMemcpyAsync(h.data(), vb, vbytes, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      for (int k = 0; k < 3; ++k)
        for (int64_t i = 0; i < np; ++i) vel[k][i] = (float)((double)vel[k][i] + h[k * np + i]);
    } else {
      for (int k = 0; k < 3; ++k)
        HIPCHK(hipMemcpyAsync(vel[k], vb + vsz * np * k, sizeof(float) * np, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
    }
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(stats, secs, ptpan_flops(np, ns), (double)np * (double)ns);
  return O3D_OK;
}

int o3d_points_affect_panels_dev(const o3d_points* src, const o3d_panels* targ, float* const vel[3],
                                 const o3d_opts* o, o3d_stats* stats) {
  clear_error();
  int rc;
  if ((rc = check_opts(o)) != O3D_OK) return rc;
  if ((rc = check_ptpan(src, targ, vel, o)) != O3D_OK) return rc;
  if (o->precision == PREC_FP64)
    return fail(O3D_EINVAL, "FP64 particles-on-panels on device float arrays are not offered; use the host call");
  if ((rc = ensure_init()) != O3D_OK) return rc;
  if (src->n > 0 && targ->np > 0) {
    PtPanArgs a = ptpan_args(src, targ, o);
    a.u = vel[0]; a.v = vel[1]; a.w = vel[2];
    if ((rc = launch_ptpan(a, o->precision, (hipStream_t)o->stream)) != O3D_OK) return rc;
  }
  fill_stats(stats, 0.0, ptpan_flops(targ->np, src->n), (double)targ->np * (double)src->n);
  return O3D_OK;
}

}  // extern "C"

extern "C" {

int o3d_bem_rhs_from_vels_dev(const o3d_panels* bdry, const o3d_bem_rhs_opts* b, float* const pu[3], float* rhs,
                              void* stream) {
  clear_error();
  int rc;
  BemArgs a;
  if ((rc = bem_rhs_args(bdry, b, pu, rhs, &a)) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  if (a.np == 0) return O3D_OK;
  hipLaunchKernelGGL(k_bem_rhs, dim3((unsigned)((a.np + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, (hipStream_t)stream, a);
  HIPCHK(hipGetLastError());
  return O3D_OK;
}

int o3d_bem_rhs_dev(const o3d_points* vort, int32_t nvort, const o3d_panels* bdry, const o3d_bem_rhs_opts* b,
                    const o3d_opts* opts, float* const pu[3], float* rhs) {
  clear_error();
  int rc;
  if (nvort < 0 || (nvort > 0 && !vort)) return fail(O3D_EINVAL, "bad particle collection list");
  if ((rc = check_opts(opts)) != O3D_OK) return rc;
  if (opts->results != O3D_VELONLY)
    return fail(O3D_EINVAL, "the BEM right-hand side uses velocities only (BEMHelper.h:73: ResultsType(velonly))");
  if (opts->precision == PREC_FP64) return fail(O3D_EINVAL, "device float arrays: FP32 or compensated precision");
  BemArgs a;
  if ((rc = bem_rhs_args(bdry, b, pu, rhs, &a)) != O3D_OK) return rc;
  for (int32_t k = 0; k < nvort; ++k)
    if ((rc = check_ptpan(&vort[k], bdry, pu, opts)) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  if (a.np == 0) return O3D_OK;
  hipStream_t st = (hipStream_t)opts->stream;
  const unsigned zb = (unsigned)std::min<int64_t>((a.np + BLOCK - 1) / BLOCK, 1024);
  hipLaunchKernelGGL(k_bem_zero, dim3(zb), dim3(BLOCK), 0, st, pu[0], pu[1], pu[2], a.np);   // zero_vels
  HIPCHK(hipGetLastError());
  for (int32_t k = 0; k < nvort; ++k) {   // accumulate from every particle collection
    if (vort[k].n == 0) continue;
    PtPanArgs pa = ptpan_args(&vort[k], bdry, opts);
    pa.u = pu[0]; pa.v = pu[1]; pa.w = pu[2];
    if ((rc = launch_ptpan(pa, opts->precision, st)) != O3D_OK) return rc;
  }
  hipLaunchKernelGGL(k_bem_rhs, dim3((unsigned)((a.np + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, a);
  HIPCHK(hipGetLastError());
  return O3D_OK;
}

}  // extern "C"
