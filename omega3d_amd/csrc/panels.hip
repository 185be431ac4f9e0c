// panels.hip -- panel -> point kernels and the panels_affect_points entry points.
// o3d_opts.method selects the panel rule: AUTO/DIRECT the reference's recursive
// centroid rule (pan_kernels.hpp), O3D_PANEL_GAUSS7 the 7-point rule with the
// closed-form near field (pan_g7.hpp).
#include "o3d_api_internal.hpp"
#include "pan_g7.hpp"
#include "pan_kernels.hpp"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <vector>

namespace o3d {

__global__ void k_ff_zero_u64_pan(unsigned long long* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

// root pairs (k_pan), then the near pairs of the targets it listed (k_pan_near);
// a.near_list / a.near_count must point at scratch for nt entries and a counter
// (zeroed here), rec at np panel records followed by np centroids
static hipError_t launch_pan(const PanArgs& a, bool grad, int prec, hipStream_t st, float4* rec) {
  if (a.nt <= 0 || a.np <= 0) return hipSuccess;
  if (!a.near_list || !a.near_count || !rec) return hipErrorInvalidValue;
  const size_t lds = (size_t)a.tile * PAN_F4 * sizeof(float4);
  const dim3 grid((unsigned)((a.nt + BLOCK - 1) / BLOCK));
  // one wave per 4 listed targets, grid-stride: at most 16 waves per CU
  const dim3 grid_near((unsigned)std::min<int64_t>((a.nt + 15) / 16, 4 * (int64_t)num_cus()));
  float4* cen = rec + PAN_F4 * a.np;
  hipLaunchKernelGGL(k_pan_g7_zero, dim3(1), dim3(64), 0, st, a.near_count);
  hipLaunchKernelGGL(k_pan_geom, dim3((unsigned)((a.np + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, a, rec, cen);
  hipError_t e = hipErrorInvalidValue;
#define O3D_PL(G, P)                                                                      \
  if (grad == G && prec == P) {                                                           \
    hipLaunchKernelGGL((k_pan<G, P>), grid, dim3(BLOCK), lds, st, a, (const float4*)rec);  \
    e = hipGetLastError();                                                                \
    if (e == hipSuccess) {                                                                \
      hipLaunchKernelGGL((k_pan_near<G, P>), grid_near, dim3(BLOCK), 0, st, a, (const float4*)rec, \
                         (const float4*)cen);                                             \
      e = hipGetLastError();                                                              \
    }                                                                                     \
  }
  O3D_PL(true, PREC_FP32) O3D_PL(true, PREC_COMP) O3D_PL(true, PREC_FP64)
  O3D_PL(false, PREC_FP32) O3D_PL(false, PREC_COMP) O3D_PL(false, PREC_FP64)
#undef O3D_PL
  return e;
}

// far kernel, then (near) the near kernel over the targets the far kernel listed;
// with near, a.near_list / a.near_count must point at scratch for nt entries and a
// counter (zeroed here), and nnear, if not NULL, counts the near pairs
static hipError_t launch_pan_g7(PanArgs a, bool grad, int prec, bool near, hipStream_t st, float4* geom = nullptr,
                                unsigned long long* nnear = nullptr) {
  if (a.nt <= 0 || a.np <= 0) return hipSuccess;
  if (!near) a.near_list = nullptr;
  const size_t lds = (size_t)a.tile * G7_F4 * sizeof(float4);
  const dim3 grid((unsigned)((a.nt + BLOCK - 1) / BLOCK));
  // one wave per listed target, grid-stride: at most 16 waves per CU
  const dim3 grid_near((unsigned)std::min<int64_t>((a.nt + 3) / 4, 4 * (int64_t)num_cus()));
  if (near) {
    if (!a.near_list || !a.near_count || !geom) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pan_g7_zero, dim3(1), dim3(64), 0, st, a.near_count);
    hipLaunchKernelGGL(k_pan_g7_geom, dim3((unsigned)((a.np + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, a, geom);
  }
  hipError_t e = hipErrorInvalidValue;
#define O3D_PL(G, P)                                                                           \
  if (grad == G && prec == P) {                                                                \
    hipLaunchKernelGGL((k_pan_g7<G, P>), grid, dim3(BLOCK), lds, st, a);                       \
    e = hipGetLastError();                                                                     \
    if (e == hipSuccess && near) {                                                             \
      hipLaunchKernelGGL((k_pan_g7_near<G, P>), grid_near, dim3(BLOCK), 0, st, a, geom, nnear); \
      e = hipGetLastError();                                                                   \
    }                                                                                          \
  }
  O3D_PL(true, PREC_FP32) O3D_PL(true, PREC_COMP) O3D_PL(true, PREC_FP64)
  O3D_PL(false, PREC_FP32) O3D_PL(false, PREC_COMP) O3D_PL(false, PREC_FP64)
#undef O3D_PL
  return e;
}

// the near-target list, its counter and f4_per_panel float4 of geometry per panel
// in the stream's workspace (caller holds stream_ws_mutex() until the launches are
// enqueued)
static int g7_scratch(PanArgs* a, hipStream_t st, float4** geom, int f4_per_panel = 1, size_t extra = 0,
                      void** tail = nullptr) {
  if (a->nt > INT32_MAX) return fail(O3D_EINVAL, "the panel kernels take at most %d targets", INT32_MAX);
  void* ws = nullptr;
  const size_t list = (sizeof(int32_t) * (size_t)a->nt + 255) & ~(size_t)255;
  const size_t gbytes = (sizeof(float4) * f4_per_panel * (size_t)a->np + 255) & ~(size_t)255;
  const int rc = stream_workspace(st, 256 + list + gbytes + extra, &ws);
  if (rc != O3D_OK) return rc;
  a->near_count = (unsigned int*)ws;
  a->near_list = (int32_t*)((char*)ws + 256);
  *geom = (float4*)((char*)ws + 256 + list);
  if (tail) *tail = (char*)ws + 256 + list + gbytes;
  return O3D_OK;
}

// the panel rule and the options check_opts validates (method is the panel rule here)
static int panel_opts(const o3d_opts* o, o3d_opts* oc, bool* g7) {
  if (!o) return fail(O3D_EINVAL, "opts is NULL");
  *oc = *o;
  *g7 = o->method == O3D_PANEL_GAUSS7;
  if (o->method == O3D_METHOD_FACTORED) return fail(O3D_EINVAL, "the factored method is for point sources");
  if (*g7) {
    oc->method = O3D_METHOD_AUTO;
    if (o->tile > 256) return fail(O3D_EINVAL, "tile %d: the 7-point rule stages at most 256 panels per tile", o->tile);
  }
  return check_opts(oc);
}

static int check_panels(const o3d_panels* p) {
  if (!p) return fail(O3D_EINVAL, "panels are NULL");
  if (p->np < 0 || p->nn < 0) return fail(O3D_EINVAL, "negative panel or node count");
  if (p->np == 0) return O3D_OK;
  if (p->nn <= 0) return fail(O3D_EINVAL, "panels without nodes");
  for (int d = 0; d < 3; ++d)
    if (!p->x[d] || !p->ts[d]) return fail(O3D_EINVAL, "panel node or strength component %d is NULL", d);
  if (!p->idx || !p->area) return fail(O3D_EINVAL, "panel indices or areas are NULL");
  return O3D_OK;
}

static int check_targets(const o3d_points* t, bool grad) {
  if (!t) return fail(O3D_EINVAL, "targ is NULL");
  if (t->n < 0) return fail(O3D_EINVAL, "targ: negative count");
  if (t->n == 0) return O3D_OK;
  for (int d = 0; d < 3; ++d)
    if (!t->x[d] || !t->u[d]) return fail(O3D_EINVAL, "targ: position or velocity component %d is NULL", d);
  if (grad)
    for (int k = 0; k < 9; ++k)
      if (!t->ug[k]) return fail(O3D_EINVAL, "targ: velocity-gradient component %d is NULL", k);
  return O3D_OK;
}

static int maxlev_of(const o3d_opts* o, int* ml) {
  *ml = o->maxlev < 0 ? 3 : o->maxlev;   // RECURSIVE_LEVELS, Influence.h:23
  if (*ml > 8) return fail(O3D_EINVAL, "maxlev %d: at most 8 levels", o->maxlev);
  return O3D_OK;
}

static double pan_flops(bool grad, int64_t nt, int64_t np, bool g7) {
  // recursive rule: leaf-only count, one flops_0vs_0p(g) per panel-target pair plus
  // the per-target term (subdivision beyond the first leaf is not counted);
  // 7-point rule: seven of them per pair (near pairs included at that count)
  const int f = o3d_flops_per_interaction(CORE_RM, grad ? O3D_VELANDGRAD : O3D_VELONLY, 0, 1);
  return (double)nt * ((grad ? 12.0 : 3.0) + (g7 ? 7.0 : 1.0) * (double)f * (double)np);
}

static PanArgs pan_args(const o3d_panels* p, const o3d_points* t, bool grad, void* const* out64) {
  PanArgs a{};
  a.nx = p->x[0]; a.ny = p->x[1]; a.nz = p->x[2];
  a.idx = p->idx;
  a.tsx = p->ts[0]; a.tsy = p->ts[1]; a.tsz = p->ts[2];
  a.area = p->area; a.ss = p->ss; a.np = p->np;
  a.tx = t->x[0]; a.ty = t->x[1]; a.tz = t->x[2];
  if (out64) {
    a.u = out64[0]; a.v = out64[1]; a.w = out64[2];
    for (int k = 0; k < 9; ++k) a.g[k] = grad ? out64[3 + k] : nullptr;
  } else {
    a.u = t->u[0]; a.v = t->u[1]; a.w = t->u[2];
    for (int k = 0; k < 9; ++k) a.g[k] = grad ? t->ug[k] : nullptr;
  }
  a.nt = t->n;
  return a;
}

// float device arrays (FP32 or compensated), or FP64 scratch rows out64
static int pan_dev(const o3d_panels* p, o3d_points* t, const o3d_opts* o, int prec, void* const* out64,
                   hipStream_t st, bool g7) {
  const bool grad = o->results == O3D_VELANDGRAD;
  int ml, rc;
  if ((rc = maxlev_of(o, &ml)) != O3D_OK) return rc;
  PanArgs a = pan_args(p, t, grad, out64);
  a.tile = o->tile > 0 ? o->tile : 128;
  a.maxlev = ml;
  if (g7) {
    std::lock_guard<std::mutex> lk(stream_ws_mutex());
    float4* geom = nullptr;
    if ((rc = g7_scratch(&a, st, &geom)) != O3D_OK) return rc;
    const hipError_t e = launch_pan_g7(a, grad, prec, true, st, geom);
    if (e != hipSuccess) return fail(O3D_EDEVICE, "panel launch failed: %s", hipGetErrorString(e));
    return O3D_OK;
  }
  std::lock_guard<std::mutex> lk(stream_ws_mutex());
  float4* rec = nullptr;
  if ((rc = g7_scratch(&a, st, &rec, PAN_F4 + 1)) != O3D_OK) return rc;
  const hipError_t e = launch_pan(a, grad, prec, st, rec);
  if (e != hipSuccess) return fail(O3D_EDEVICE, "panel launch failed: %s", hipGetErrorString(e));
  return O3D_OK;
}

// Colocation points of a surface as Surfaces::represent_as_particles(offset, ips)
// places them (Surfaces.h:959-1010): the centroid, the float sum of the three nodes
// times (1./3.) in double, stored float; then pushed out by offset along the
// panel's normal from compute_bases (Surfaces.h:766-816), in float.
__global__ __launch_bounds__(BLOCK) void k_colocation(const float* nx, const float* ny, const float* nz,
                                                      const uint32_t* idx, int64_t np, float offset, float* px,
                                                      float* py, float* pz) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= np) return;
  const uint32_t id0 = idx[3 * i], id1 = idx[3 * i + 1], id2 = idx[3 * i + 2];
  const float* xs[3] = {nx, ny, nz};
  float c[3], x1[3], x2[3], nm[3];
  panel_basis_f(xs, id0, id1, id2, x1, x2, nm);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    c[j] = (float)((1. / 3.) * (double)(xs[j][id0] + xs[j][id1] + xs[j][id2]));
    c[j] = c[j] + offset * nm[j];
  }
  px[i] = c[0];
  py[i] = c[1];
  pz[i] = c[2];
}

__global__ __launch_bounds__(BLOCK) void k_zero3(float* a, float* b, float* c, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
    a[i] = 0.0f; b[i] = 0.0f; c[i] = 0.0f;
  }
}

// the offset Influence.h:1233 asks of represent_as_particles
constexpr float COLOCATION_OFFSET = 0.0001f;

// panels_affect_panels (Influence.h:1225-1245) on device arrays: the target
// surface's colocation points, zeroed velocities (the reference copies the
// temporary points' velocities over the target's, :1240-1244), then the recursive
// panel rule of panels_affect_points, velocity only.
static int pap_dev(const o3d_panels* src, const o3d_panels* targ, float* const vel[3], const o3d_opts* o,
                   hipStream_t st) {
  const int64_t nt = targ->np;
  if (nt <= 0) return O3D_OK;
  const unsigned g = (unsigned)((nt + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_zero3, dim3(std::min(g, 4096u)), dim3(BLOCK), 0, st, vel[0], vel[1], vel[2], nt);
  HIPCHK(hipGetLastError());
  if (src->np <= 0) return O3D_OK;
  int ml, rc;
  if ((rc = maxlev_of(o, &ml)) != O3D_OK) return rc;
  o3d_points t{};
  t.n = nt;
  for (int d = 0; d < 3; ++d) t.u[d] = vel[d];
  PanArgs a = pan_args(src, &t, false, nullptr);
  a.tile = o->tile > 0 ? o->tile : 128;
  a.maxlev = ml;
  std::lock_guard<std::mutex> lk(stream_ws_mutex());
  float4* rec = nullptr;
  void* tail = nullptr;
  if ((rc = g7_scratch(&a, st, &rec, PAN_F4 + 1, sizeof(float) * 3 * (size_t)nt, &tail)) != O3D_OK) return rc;
  float* px = (float*)tail;
  a.tx = px; a.ty = px + nt; a.tz = px + 2 * nt;
  hipLaunchKernelGGL(k_colocation, dim3(g), dim3(BLOCK), 0, st, targ->x[0], targ->x[1], targ->x[2], targ->idx, nt,
                     COLOCATION_OFFSET, px, px + nt, px + 2 * nt);
  HIPCHK(hipGetLastError());
  const hipError_t e = launch_pan(a, false, o->precision, st, rec);
  if (e != hipSuccess) return fail(O3D_EDEVICE, "panel launch failed: %s", hipGetErrorString(e));
  return O3D_OK;
}

static int check_pap(const o3d_panels* src, const o3d_panels* targ, float* const vel[3], const o3d_opts* o) {
  int rc;
  if ((rc = check_panels(src)) != O3D_OK) return rc;
  if (!targ) return fail(O3D_EINVAL, "target panels are NULL");
  if (targ->np < 0 || targ->nn < 0) return fail(O3D_EINVAL, "negative target panel or node count");
  if (targ->np > 0) {
    if (targ->nn <= 0 || !targ->idx) return fail(O3D_EINVAL, "target panels without nodes or indices");
    for (int d = 0; d < 3; ++d)
      if (!targ->x[d]) return fail(O3D_EINVAL, "target node component %d is NULL", d);
    if (!vel) return fail(O3D_EINVAL, "vel is NULL");
    for (int d = 0; d < 3; ++d)
      if (!vel[d]) return fail(O3D_EINVAL, "velocity component %d is NULL", d);
  }
  if (o->results != O3D_VELONLY)
    return fail(O3D_EINVAL, "panels_affect_panels computes velocities only (the reference asserts, Influence.h:1227)");
  if (o->precision == PREC_FP64) return fail(O3D_EINVAL, "panels_affect_panels runs in FP32 or compensated precision");
  if (o->method == O3D_PANEL_GAUSS7) return fail(O3D_EINVAL, "panels_affect_panels uses the recursive panel rule");
  return O3D_OK;
}

}  // namespace o3d

using namespace o3d;

extern "C" {

int o3d_panels_affect_points(const o3d_panels* src, o3d_points* targ, const o3d_opts* opts, o3d_stats* stats) {
  clear_error();
  int rc;
  o3d_opts oc;
  bool g7;
  if ((rc = panel_opts(opts, &oc, &g7)) != O3D_OK) return rc;
  const o3d_opts* o = &oc;
  const bool grad = o->results == O3D_VELANDGRAD;
  if ((rc = check_panels(src)) != O3D_OK) return rc;
  if ((rc = check_targets(targ, grad)) != O3D_OK) return rc;
  if (src->np > 0 && (rc = check_node_indices(src, "src")) != O3D_OK) return rc;
  int ml;
  if ((rc = maxlev_of(o, &ml)) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t np = src->np, nn = src->nn, nt = targ->n;
  if (np > 0 && nt > 0) {
    std::lock_guard<std::mutex> lk(pool_mutex());
    hipStream_t st = lib_stream();
    const int nout = grad ? 12 : 3;
    const size_t osz = o->precision == PREC_FP64 ? sizeof(double) : sizeof(float);
    void *dp, *dt;
    // panel block: nodes 3*nn, ts 3*np, area np, ss np (floats) + idx 3*np (uint32)
    if ((rc = pool_buffer(2, sizeof(float) * (3 * nn + 8 * np), &dp)) != O3D_OK) return rc;
    if ((rc = pool_buffer(3, sizeof(float) * 3 * nt + osz * nout * nt, &dt)) != O3D_OK) return rc;
    float* f = (float*)dp;
    o3d_panels d{};
    d.np = np; d.nn = nn;
    for (int k = 0; k < 3; ++k) { d.x[k] = f + k * nn; d.ts[k] = f + 3 * nn + k * np; }
    d.area = f + 3 * nn + 3 * np;
    d.ss = src->ss ? f + 3 * nn + 4 * np : nullptr;
    d.idx = (const uint32_t*)(f + 3 * nn + 5 * np);
    for (int k = 0; k < 3; ++k) {
      HIPCHK(hipMemcpyAsync((void*)d.x[k], src->x[k], sizeof(float) * nn, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync((void*)d.ts[k], src->ts[k], sizeof(float) * np, hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync((void*)d.area, src->area, sizeof(float) * np, hipMemcpyHostToDevice, st));
    if (src->ss) HIPCHK(hipMemcpyAsync((void*)d.ss, src->ss, sizeof(float) * np, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync((void*)d.idx, src->idx, sizeof(uint32_t) * 3 * np, hipMemcpyHostToDevice, st));
    o3d_points dtg{};
    dtg.n = nt;
    float* tf = (float*)dt;
    for (int k = 0; k < 3; ++k) {
      dtg.x[k] = tf + k * nt;
      HIPCHK(hipMemcpyAsync((void*)dtg.x[k], targ->x[k], sizeof(float) * nt, hipMemcpyHostToDevice, st));
    }
    char* ob = (char*)(tf + 3 * nt);
    void* outs[12];
    for (int k = 0; k < nout; ++k) outs[k] = ob + osz * nt * k;
    if (o->precision == PREC_FP64) {
      HIPCHK(hipMemsetAsync(ob, 0, osz * nout * nt, st));
      rc = pan_dev(&d, &dtg, o, PREC_FP64, outs, st, g7);
    } else {
      for (int k = 0; k < 3; ++k) {
        dtg.u[k] = (float*)outs[k];
        HIPCHK(hipMemcpyAsync(dtg.u[k], targ->u[k], sizeof(float) * nt, hipMemcpyHostToDevice, st));
      }
      if (grad)
        for (int k = 0; k < 9; ++k) {
          dtg.ug[k] = (float*)outs[3 + k];
          HIPCHK(hipMemcpyAsync(dtg.ug[k], targ->ug[k], sizeof(float) * nt, hipMemcpyHostToDevice, st));
        }
      rc = pan_dev(&d, &dtg, o, o->precision, nullptr, st, g7);
    }
    if (rc != O3D_OK) return rc;
    if (o->precision == PREC_FP64) {
      std::vector<double> h((size_t)nout * nt);
      HIPCHK(hipMemcpyAsync(h.data(), ob, osz * nout * nt, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      for (int k = 0; k < 3; ++k)
        for (int64_t i = 0; i < nt; ++i) targ->u[k][i] = (float)((double)targ->u[k][i] + h[k * nt + i]);
      if (grad)
        for (int k = 0; k < 9; ++k)
          for (int64_t i = 0; i < nt; ++i)
            targ->ug[k][i] = (float)((double)targ->ug[k][i] + h[(3 + k) * nt + i]);
    } else {
      for (int k = 0; k < 3; ++k)
        HIPCHK(hipMemcpyAsync(targ->u[k], outs[k], sizeof(float) * nt, hipMemcpyDeviceToHost, st));
      if (grad)
        for (int k = 0; k < 9; ++k)
          HIPCHK(hipMemcpyAsync(targ->ug[k], outs[3 + k], sizeof(float) * nt, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
    }
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(stats, secs, pan_flops(grad, nt, np, g7), (double)nt * (double)np);
  return O3D_OK;
}

int o3d_panels_affect_points_dev(const o3d_panels* src, o3d_points* targ, const o3d_opts* opts, o3d_stats* stats) {
  clear_error();
  int rc;
  o3d_opts oc;
  bool g7;
  if ((rc = panel_opts(opts, &oc, &g7)) != O3D_OK) return rc;
  const o3d_opts* o = &oc;
  const bool grad = o->results == O3D_VELANDGRAD;
  if ((rc = check_panels(src)) != O3D_OK) return rc;
  if ((rc = check_targets(targ, grad)) != O3D_OK) return rc;
  if (o->precision == PREC_FP64)
    return fail(O3D_EINVAL, "FP64 panels on device float arrays are not offered; use the host call");
  if ((rc = ensure_init()) != O3D_OK) return rc;
  if (src->np > 0 && targ->n > 0) {
    rc = pan_dev(src, targ, o, o->precision, nullptr, (hipStream_t)o->stream, g7);
    if (rc != O3D_OK) return rc;
  }
  fill_stats(stats, 0.0, pan_flops(grad, targ->n, src->np, g7), (double)targ->n * (double)src->np);
  return O3D_OK;
}

// The 7-point rule on device arrays, timed: one full evaluation into targ (as
// o3d_panels_affect_points_dev), one far-field-only pass into library scratch, and
// a count of the near pairs.  Velocity and gradient, FP32 or compensated.
int o3d_panels_g7_profile_dev(const o3d_panels* src, o3d_points* targ, const o3d_opts* opts,
                              o3d_panel_profile* out) {
  clear_error();
  int rc;
  if (!out) return fail(O3D_EINVAL, "out is NULL");
  *out = o3d_panel_profile{};
  o3d_opts oc;
  bool g7;
  if ((rc = panel_opts(opts, &oc, &g7)) != O3D_OK) return rc;
  if (!g7) return fail(O3D_EINVAL, "the profile is for the 7-point rule (method O3D_PANEL_GAUSS7)");
  if (oc.results != O3D_VELANDGRAD || oc.precision == PREC_FP64)
    return fail(O3D_EINVAL, "the profile runs velocity and gradient in FP32 or compensated precision");
  if ((rc = check_panels(src)) != O3D_OK) return rc;
  if ((rc = check_targets(targ, true)) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  const int64_t nt = targ->n, np = src->np;
  if (nt == 0 || np == 0) return O3D_OK;
  hipStream_t st = (hipStream_t)oc.stream;
  std::lock_guard<std::mutex> lk(pool_mutex());
  void* scratch;
  if ((rc = pool_buffer(4, sizeof(float) * 12 * (size_t)nt + 256, &scratch)) != O3D_OK) return rc;
  float* f = (float*)scratch;
  o3d_points t2 = *targ;
  for (int d = 0; d < 3; ++d) t2.u[d] = f + d * nt;
  for (int k = 0; k < 9; ++k) t2.ug[k] = f + (3 + k) * nt;
  unsigned long long* cnt = (unsigned long long*)(f + 12 * nt);
  PanArgs full = pan_args(src, targ, true, nullptr);
  PanArgs far = pan_args(src, &t2, true, nullptr);
  full.tile = far.tile = oc.tile > 0 ? oc.tile : 128;
  std::lock_guard<std::mutex> lk2(stream_ws_mutex());
  float4* geom = nullptr;
  if ((rc = g7_scratch(&full, st, &geom)) != O3D_OK) return rc;
  hipEvent_t ev[4];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  hipLaunchKernelGGL(k_ff_zero_u64_pan, dim3(1), dim3(64), 0, st, cnt);
  HIPCHK(hipEventRecord(ev[0], st));
  HIPCHK(launch_pan_g7(full, true, oc.precision, true, st, geom, cnt));
  HIPCHK(hipEventRecord(ev[1], st));
  HIPCHK(launch_pan_g7(far, true, oc.precision, false, st));
  HIPCHK(hipEventRecord(ev[2], st));
  unsigned long long nnear = 0;
  HIPCHK(hipMemcpyAsync(&nnear, cnt, sizeof nnear, hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(ev[3], st));
  HIPCHK(hipEventSynchronize(ev[3]));
  float ms0 = 0.f, ms1 = 0.f;
  HIPCHK(hipEventElapsedTime(&ms0, ev[0], ev[1]));
  HIPCHK(hipEventElapsedTime(&ms1, ev[1], ev[2]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  out->total_ms = ms0;
  out->far_ms = ms1;
  out->pairs = (double)nt * (double)np;
  out->near_pairs = (double)nnear;
  out->far_pairs = out->pairs - out->near_pairs;
  return O3D_OK;
}

int o3d_panels_affect_panels_dev(const o3d_panels* src, const o3d_panels* targ, float* const vel[3],
                                 const o3d_opts* opts, o3d_stats* stats) {
  clear_error();
  int rc;
  if ((rc = check_opts(opts)) != O3D_OK) return rc;
  if ((rc = check_pap(src, targ, vel, opts)) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  if ((rc = pap_dev(src, targ, vel, opts, (hipStream_t)opts->stream)) != O3D_OK) return rc;
  fill_stats(stats, 0.0, pan_flops(false, targ->np, src->np, false), (double)targ->np * (double)src->np);
  return O3D_OK;
}

int o3d_panels_affect_panels(const o3d_panels* src, const o3d_panels* targ, float* const vel[3],
                             const o3d_opts* opts, o3d_stats* stats) {
  clear_error();
  int rc;
  if ((rc = check_opts(opts)) != O3D_OK) return rc;
  if ((rc = check_pap(src, targ, vel, opts)) != O3D_OK) return rc;
  if (src->np > 0 && (rc = check_node_indices(src, "src")) != O3D_OK) return rc;
  if (targ->np > 0 && (rc = check_node_indices(targ, "targ")) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t np = src->np, nn = src->nn, nt = targ->np, tn = targ->nn;
  if (nt > 0) {
    std::lock_guard<std::mutex> lk(pool_mutex());
    hipStream_t st = lib_stream();
    void *dp, *dt;
    // source panels: nodes 3*nn, ts 3*np, area np, ss np (floats) + idx 3*np (uint32);
    // target panels: nodes 3*tn (floats) + idx 3*nt (uint32), velocities 3*nt (floats)
    if ((rc = pool_buffer(2, sizeof(float) * (3 * nn + 8 * np) + 256, &dp)) != O3D_OK) return rc;
    if ((rc = pool_buffer(3, sizeof(float) * (3 * tn + 6 * nt) + 256, &dt)) != O3D_OK) return rc;
    float* f = (float*)dp;
    o3d_panels d{};
    d.np = np; d.nn = nn;
    for (int k = 0; k < 3; ++k) { d.x[k] = f + k * nn; d.ts[k] = f + 3 * nn + k * np; }
    d.area = f + 3 * nn + 3 * np;
    d.ss = src->ss ? f + 3 * nn + 4 * np : nullptr;
    d.idx = (const uint32_t*)(f + 3 * nn + 5 * np);
    if (np > 0) {
      for (int k = 0; k < 3; ++k) {
        HIPCHK(hipMemcpyAsync((void*)d.x[k], src->x[k], sizeof(float) * nn, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync((void*)d.ts[k], src->ts[k], sizeof(float) * np, hipMemcpyHostToDevice, st));
      }
      HIPCHK(hipMemcpyAsync((void*)d.area, src->area, sizeof(float) * np, hipMemcpyHostToDevice, st));
      if (src->ss) HIPCHK(hipMemcpyAsync((void*)d.ss, src->ss, sizeof(float) * np, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync((void*)d.idx, src->idx, sizeof(uint32_t) * 3 * np, hipMemcpyHostToDevice, st));
    }
    float* g = (float*)dt;
    o3d_panels dtg{};
    dtg.np = nt; dtg.nn = tn;
    for (int k = 0; k < 3; ++k) dtg.x[k] = g + k * tn;
    dtg.idx = (const uint32_t*)(g + 3 * tn);
    float* dv[3] = {g + 3 * tn + 3 * nt, g + 3 * tn + 4 * nt, g + 3 * tn + 5 * nt};
    for (int k = 0; k < 3; ++k)
      HIPCHK(hipMemcpyAsync((void*)dtg.x[k], targ->x[k], sizeof(float) * tn, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync((void*)dtg.idx, targ->idx, sizeof(uint32_t) * 3 * nt, hipMemcpyHostToDevice, st));
    o3d_opts oc = *opts;
    oc.stream = st;
    if ((rc = pap_dev(&d, &dtg, dv, &oc, st)) != O3D_OK) return rc;
    for (int k = 0; k < 3; ++k) HIPCHK(hipMemcpyAsync(vel[k], dv[k], sizeof(float) * nt, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(stats, secs, pan_flops(false, nt, np, false), (double)nt * (double)np);
  return O3D_OK;
}

}  // extern "C"
