// panels.hip -- panel -> point kernels and the panels_affect_points entry points.
#include "o3d_api_internal.hpp"
#include "pan_kernels.hpp"

#include <chrono>
#include <cstring>
#include <vector>

namespace o3d {

static hipError_t launch_pan(const PanArgs& a, bool grad, int prec, hipStream_t st) {
  if (a.nt <= 0 || a.np <= 0) return hipSuccess;
  const size_t lds = (size_t)a.tile * PAN_F4 * sizeof(float4);
  const dim3 grid((unsigned)((a.nt + BLOCK - 1) / BLOCK));
#define O3D_PL(G, P) \
  if (grad == G && prec == P) { hipLaunchKernelGGL((k_pan<G, P>), grid, dim3(BLOCK), lds, st, a); return hipGetLastError(); }
  O3D_PL(true, PREC_FP32) O3D_PL(true, PREC_COMP) O3D_PL(true, PREC_FP64)
  O3D_PL(false, PREC_FP32) O3D_PL(false, PREC_COMP) O3D_PL(false, PREC_FP64)
#undef O3D_PL
  return hipErrorInvalidValue;
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

static double pan_flops(bool grad, int64_t nt, int64_t np) {
  // leaf-only count: one flops_0vs_0p(g) per panel-target pair plus the per-target
  // term; the subdivision work beyond the first leaf is not counted
  const int f = o3d_flops_per_interaction(CORE_RM, grad ? O3D_VELANDGRAD : O3D_VELONLY, 0, 1);
  return (double)nt * ((grad ? 12.0 : 3.0) + (double)f * (double)np);
}

// float device arrays; FP32 or compensated
static int pan_dev(const o3d_panels* p, o3d_points* t, const o3d_opts* o, int prec, void* const* out64,
                   hipStream_t st) {
  const bool grad = o->results == O3D_VELANDGRAD;
  int ml, rc;
  if ((rc = maxlev_of(o, &ml)) != O3D_OK) return rc;
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
  a.tile = o->tile > 0 ? o->tile : 128;
  a.maxlev = ml;
  hipError_t e = launch_pan(a, grad, prec, st);
  if (e != hipSuccess) return fail(O3D_EDEVICE, "panel launch failed: %s", hipGetErrorString(e));
  return O3D_OK;
}

}  // namespace o3d

using namespace o3d;

extern "C" {

int o3d_panels_affect_points(const o3d_panels* src, o3d_points* targ, const o3d_opts* o, o3d_stats* stats) {
  clear_error();
  int rc;
  if ((rc = check_opts(o)) != O3D_OK) return rc;
  const bool grad = o->results == O3D_VELANDGRAD;
  if ((rc = check_panels(src)) != O3D_OK) return rc;
  if ((rc = check_targets(targ, grad)) != O3D_OK) return rc;
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
      rc = pan_dev(&d, &dtg, o, PREC_FP64, outs, st);
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
      rc = pan_dev(&d, &dtg, o, o->precision, nullptr, st);
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
  fill_stats(stats, secs, pan_flops(grad, nt, np), (double)nt * (double)np);
  return O3D_OK;
}

int o3d_panels_affect_points_dev(const o3d_panels* src, o3d_points* targ, const o3d_opts* o, o3d_stats* stats) {
  clear_error();
  int rc;
  if ((rc = check_opts(o)) != O3D_OK) return rc;
  const bool grad = o->results == O3D_VELANDGRAD;
  if ((rc = check_panels(src)) != O3D_OK) return rc;
  if ((rc = check_targets(targ, grad)) != O3D_OK) return rc;
  if (o->precision == PREC_FP64)
    return fail(O3D_EINVAL, "FP64 panels on device float arrays are not offered; use the host call");
  if ((rc = ensure_init()) != O3D_OK) return rc;
  if (src->np > 0 && targ->n > 0) {
    rc = pan_dev(src, targ, o, o->precision, nullptr, (hipStream_t)o->stream);
    if (rc != O3D_OK) return rc;
  }
  fill_stats(stats, 0.0, pan_flops(grad, targ->n, src->np), (double)targ->n * (double)src->np);
  return O3D_OK;
}

}  // extern "C"
