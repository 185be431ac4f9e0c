// ff.hip -- host side of the factored far-field evaluation (method "factored").
//
// One call enqueues, on the caller's stream and without synchronising:
//   bounding box -> Morton keys -> radix sort (hipcub) of sources (and of targets
//   when they are a different collection) -> gather into sorted float4 rows ->
//   bounding spheres of source tiles and target blocks -> far/near bit per
//   (block, tile) -> far-field kernel (matrix cores) -> near-field kernel, which
//   also adds the far part and accumulates into the caller's arrays.
// Scratch lives in one device workspace per stream (calls on one stream are
// ordered, so they can share it); it grows when a larger call comes, after the
// stream has drained.
#include <hipcub/hipcub.hpp>

#include "ff_kernels.hpp"
#include "o3d_api_internal.hpp"

#include <algorithm>
#include <climits>
#include <map>
#include <mutex>

namespace o3d {
namespace {

struct Work {
  void* p = nullptr;
  size_t bytes = 0;
};
std::mutex g_ws_mu;
std::map<hipStream_t, Work> g_ws;

// caller holds g_ws_mu for the whole enqueue, so that no other thread frees or
// grows this stream's workspace between its sizing and the last launch using it
int workspace(hipStream_t st, size_t bytes, void** out) {
  Work& w = g_ws[st];
  if (w.bytes < bytes) {
    if (w.p) {
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipFree(w.p));
      w.p = nullptr;
      w.bytes = 0;
    }
    const size_t want = bytes + bytes / 4;
    if (hipMalloc(&w.p, want) != hipSuccess) {
      w.p = nullptr;
      return fail(O3D_ENOMEM, "factored path: device allocation of %zu bytes failed", want);
    }
    w.bytes = want;
  }
  *out = w.p;
  return O3D_OK;
}

// sequential sub-allocation of the workspace, 256-byte aligned
struct Carve {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(size_t count) {
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += (count * sizeof(T) + 255) & ~(size_t)255;
    return p;
  }
};

unsigned grid_of(int64_t n, int per) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

// workgroups of the persistent near-field kernel: every slot of every CU
template <int CORE, int PREC>
int64_t near_grid(size_t lds) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ff_near<CORE, PREC>, FF_TB, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return (int64_t)per_cu * num_cus();
}

// RAII pair of events for the optional stage timing
struct Ev {
  hipEvent_t e = nullptr;
  ~Ev() { if (e) (void)hipEventDestroy(e); }
};

template <int CORE, bool THICK, int PREC>
int run(const o3d_points* src, o3d_points* targ, bool self_excl, int T, hipStream_t st, o3d_factored_profile* prof) {
  const int64_t ns = src->n, nt = targ->n;
  const bool same = ns == nt && src->x[0] == targ->x[0] && src->x[1] == targ->x[1] && src->x[2] == targ->x[2];
  const int64_t nb = (nt + FF_TB - 1) / FF_TB;
  const int ntiles = (int)((ns + T - 1) / T);
  const int mwords = (ntiles + 31) / 32;

  size_t tb_s = 0, tb_t = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb_s, (const unsigned*)nullptr, (unsigned*)nullptr,
                                            (const int*)nullptr, (int*)nullptr, (int)ns, 0, 30, st));
  if (!same)
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb_t, (const unsigned*)nullptr, (unsigned*)nullptr,
                                              (const int*)nullptr, (int*)nullptr, (int)nt, 0, 30, st));
  size_t tb_scan = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb_scan, (const int*)nullptr, (int*)nullptr, (int)nb, st));
  size_t tmp_bytes = std::max(std::max(tb_s, tb_t), tb_scan);

  // layout (first pass sizes it, second pass places it)
  void* base = nullptr;
  for (int pass = 0; pass < 2; ++pass) {
    Carve cv{static_cast<char*>(base)};
    unsigned* bb = cv.take<unsigned>(8);
    unsigned* ks_in = cv.take<unsigned>(ns);
    unsigned* ks_out = cv.take<unsigned>(ns);
    int* is_in = cv.take<int>(ns);
    int* is_out = cv.take<int>(ns);
    unsigned* kt_in = same ? nullptr : cv.take<unsigned>(nt);
    unsigned* kt_out = same ? nullptr : cv.take<unsigned>(nt);
    int* it_in = same ? nullptr : cv.take<int>(nt);
    int* it_out = same ? nullptr : cv.take<int>(nt);
    void* tmp = cv.take<char>(tmp_bytes);
    float4* sp = cv.take<float4>(ns);
    float4* sa = cv.take<float4>(ns);
    float4* tq = cv.take<float4>(nt);
    float4* tsph = cv.take<float4>(ntiles);
    float* tcs = cv.take<float>(ntiles);
    float* tam = cv.take<float>(ntiles);
    float4* bsph = cv.take<float4>(nb);
    float* bcs = cv.take<float>(nb);
    unsigned* mask = cv.take<unsigned>((size_t)nb * mwords);
    double* farv = cv.take<double>((size_t)12 * nt);
    unsigned long long* npair = cv.take<unsigned long long>(1);
    int* nn = cv.take<int>(nb);
    int* noff = cv.take<int>(nb);
    int* items = cv.take<int>(nb);
    int* ioff = cv.take<int>(nb);
    int* mfold = cv.take<int>(1);
    unsigned* ctr = cv.take<unsigned>(1);
    double* part = cv.take<double>((size_t)3 * nb * 12 * FF_TB);   // at most 3 items per block
    if (pass == 0) {
      int rc = workspace(st, cv.off, &base);
      if (rc != O3D_OK) return rc;
      continue;
    }

    SrcView sv{src->x[0], src->x[1], src->x[2], src->r, src->s[0], src->s[1], src->s[2]};
    TgtView tv{};
    tv.x = targ->x[0]; tv.y = targ->x[1]; tv.z = targ->x[2]; tv.r = targ->r;
    tv.u = targ->u[0]; tv.v = targ->u[1]; tv.w = targ->u[2];
    for (int k = 0; k < 9; ++k) tv.g[k] = targ->ug[k];

    Ev ev[4];
    if (prof) {
      for (auto& e : ev) HIPCHK(hipEventCreate(&e.e));
      HIPCHK(hipEventRecord(ev[0].e, st));
    }
    hipLaunchKernelGGL(k_ff_bbox_init, dim3(1), dim3(64), 0, st, bb);
    hipLaunchKernelGGL(k_ff_bbox, dim3(std::min(grid_of(ns, 256), 2048u)), dim3(256), 0, st, sv.x, sv.y, sv.z, ns, bb);
    if (!same)
      hipLaunchKernelGGL(k_ff_bbox, dim3(std::min(grid_of(nt, 256), 2048u)), dim3(256), 0, st, tv.x, tv.y, tv.z, nt,
                         bb);
    hipLaunchKernelGGL(k_ff_keys, dim3(grid_of(ns, 256)), dim3(256), 0, st, sv.x, sv.y, sv.z, ns, bb, ks_in, is_in);
    HIPCHK(hipGetLastError());
    size_t tb = tmp_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, ks_in, ks_out, is_in, is_out, (int)ns, 0, 30, st));
    const int* tperm = is_out;
    if (!same) {
      hipLaunchKernelGGL(k_ff_keys, dim3(grid_of(nt, 256)), dim3(256), 0, st, tv.x, tv.y, tv.z, nt, bb, kt_in, it_in);
      HIPCHK(hipGetLastError());
      tb = tmp_bytes;
      HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, kt_in, kt_out, it_in, it_out, (int)nt, 0, 30, st));
      tperm = it_out;
    }
    hipLaunchKernelGGL((k_ff_gather_src<CORE>), dim3(grid_of(ns, 256)), dim3(256), 0, st, sv, is_out, ns, sp, sa);
    hipLaunchKernelGGL((k_ff_gather_tgt<CORE, THICK>), dim3(grid_of(nt, 256)), dim3(256), 0, st, tv, tperm, nt, tq);
    hipLaunchKernelGGL(k_ff_bounds, dim3((unsigned)ntiles), dim3(256), 0, st, sp, sa, ns, T, tsph, tcs, tam);
    hipLaunchKernelGGL(k_ff_bounds, dim3((unsigned)nb), dim3(256), 0, st, tq, (const float4*)nullptr, nt, FF_TB, bsph,
                       bcs, (float*)nullptr);
    hipLaunchKernelGGL((k_ff_classify<CORE>), dim3(grid_of(nb * mwords, 256)), dim3(256), 0, st, bsph, bcs, nb, tsph,
                       tcs, ntiles, mwords, mask);
    hipLaunchKernelGGL(k_ff_plan_count, dim3(grid_of(nb, 256)), dim3(256), 0, st, mask, nb, mwords, ntiles, nn);
    HIPCHK(hipGetLastError());
    tb = tmp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, nn, noff, (int)nb, st));
    hipLaunchKernelGGL(k_ff_plan_items, dim3(grid_of(nb, 256)), dim3(256), 0, st, nn, noff, nb, items, mfold);
    HIPCHK(hipGetLastError());
    tb = tmp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, items, ioff, (int)nb, st));
    hipLaunchKernelGGL(k_ff_zero_u32, dim3(1), dim3(64), 0, st, ctr);
    HIPCHK(hipGetLastError());

    FFArgs a{};
    a.sp = sp; a.sa = sa; a.tq = tq; a.tperm = tperm;
    a.tsph = tsph; a.tcs = tcs; a.tam = tam; a.bsph = bsph; a.bcs = bcs;
    a.mask = mask; a.farv = farv;
    a.nn = nn; a.items = items; a.ioff = ioff; a.mfold = mfold; a.ctr = ctr; a.part = part;
    a.s = sv; a.t = tv;
    a.ns = ns; a.nt = nt;
    a.tile = T; a.mwords = mwords; a.ntiles = ntiles;
    a.self_excl = self_excl ? 1 : 0;
    const size_t lds_far = (size_t)256 * FF_SL * sizeof(__bf16) + (size_t)2 * 32 * FF_TP * sizeof(_Float16);
    if (prof) HIPCHK(hipEventRecord(ev[1].e, st));
    hipLaunchKernelGGL((k_ff_far<CORE, THICK, PREC>), dim3((unsigned)nb), dim3(FF_FAR_THREADS), lds_far, st, a);
    HIPCHK(hipGetLastError());
    if (prof) HIPCHK(hipEventRecord(ev[2].e, st));
    const size_t lds_near = (size_t)T * 2 * sizeof(float4);
    const unsigned gnear = (unsigned)std::min<int64_t>(3 * nb, near_grid<CORE, PREC>(lds_near));
    hipLaunchKernelGGL((k_ff_near<CORE, PREC>), dim3(gnear), dim3(FF_TB), lds_near, st, a);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL((k_ff_combine<CORE, PREC>), dim3((unsigned)nb), dim3(FF_TB), 0, st, a);
    HIPCHK(hipGetLastError());
    if (prof) {
      HIPCHK(hipEventRecord(ev[3].e, st));
      hipLaunchKernelGGL(k_ff_zero_u64, dim3(1), dim3(64), 0, st, npair);
      hipLaunchKernelGGL(k_ff_count, dim3(grid_of(nb, 256)), dim3(256), 0, st, mask, nb, mwords, ntiles, T, ns, nt,
                         npair);
      HIPCHK(hipGetLastError());
      unsigned long long far_pairs = 0;
      HIPCHK(hipMemcpyAsync(&far_pairs, npair, sizeof far_pairs, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      float ms[3];
      for (int k = 0; k < 3; ++k) HIPCHK(hipEventElapsedTime(&ms[k], ev[k].e, ev[k + 1].e));
      prof->prep_ms = ms[0];
      prof->far_ms = ms[1];
      prof->near_ms = ms[2];
      prof->total_ms = (double)ms[0] + ms[1] + ms[2];
      prof->far_pairs = (double)far_pairs;
      prof->pairs = (double)ns * (double)nt;
      prof->far_tile_fraction = 0.0;
      if (nb > 0 && ntiles > 0) prof->far_tile_fraction = prof->far_pairs / prof->pairs;
    }
  }
  return O3D_OK;
}

template <int CORE>
int run_core(const o3d_points* src, o3d_points* targ, bool thick, int prec, bool self_excl, int T, hipStream_t st,
             o3d_factored_profile* prof) {
  if (thick)
    return prec == PREC_COMP ? run<CORE, true, PREC_COMP>(src, targ, self_excl, T, st, prof)
                             : run<CORE, true, PREC_FP32>(src, targ, self_excl, T, st, prof);
  return prec == PREC_COMP ? run<CORE, false, PREC_COMP>(src, targ, self_excl, T, st, prof)
                           : run<CORE, false, PREC_FP32>(src, targ, self_excl, T, st, prof);
}

}  // namespace

bool ff_supported(int core, bool grad, int prec) {
  return (core == CORE_RM || core == CORE_EXP) && grad && (prec == PREC_FP32 || prec == PREC_COMP);
}

// device-pointer point->point by the factored far field; arguments already checked
int p2p_factored(const o3d_points* src, o3d_points* targ, const o3d_opts* o, hipStream_t st,
                 o3d_factored_profile* prof) {
  const bool grad = o->results == O3D_VELANDGRAD;
  if (!ff_supported(o->core, grad, o->precision))
    return fail(O3D_EINVAL, "factored method: needs the RM or EXP core, velocity and gradient, and FP32 or "
                            "compensated precision");
  if (src->n > INT_MAX || targ->n > INT_MAX)
    return fail(O3D_EINVAL, "factored method: at most %d sources and %d targets", INT_MAX, INT_MAX);
  const int T = o->tile > 0 ? o->tile : 256;
  if (T != 64 && T != 128 && T != 256) return fail(O3D_EINVAL, "factored method: tile must be 0, 64, 128 or 256");
  const bool thick = targ->r != nullptr;
  const bool self_excl = o->self_term == O3D_SELF_EXCLUDE;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  if (o->core == CORE_RM) return run_core<CORE_RM>(src, targ, thick, o->precision, self_excl, T, st, prof);
  return run_core<CORE_EXP>(src, targ, thick, o->precision, self_excl, T, st, prof);
}

}  // namespace o3d
