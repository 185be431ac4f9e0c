// ff.hip -- host side of the factored far-field evaluation (method "factored").
//
// One call enqueues, on the caller's stream and without synchronising:
//   bounding box -> Morton keys -> radix sort (sort.hip) of sources (and of targets
//   when they are a different collection) -> gather into sorted float4 rows ->
//   bounding spheres of source tiles and target blocks -> far/near bit per
//   (block, tile) -> far-field kernel (matrix cores) -> near-field kernel, which
//   also adds the far part and accumulates into the caller's arrays.
// Scratch lives in one device workspace per stream (calls on one stream are
// ordered, so they can share it); it grows when a larger call comes, after the
// stream has drained.  A caller that owns its scratch (a batch) passes it in
// instead, sized by ff_workspace_bytes.
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

bool same_points(const o3d_points* src, const o3d_points* targ) {
  return src->n == targ->n && src->x[0] == targ->x[0] && src->x[1] == targ->x[1] && src->x[2] == targ->x[2];
}

// scratch of the radix sorts of one call (the scans need none)
size_t lib_tmp_bytes(int64_t ns, int64_t nt, bool same) {
  return radix_tmp_bytes(same ? ns : std::max(ns, nt));
}

// the workspace's sub-arrays in order; with base == nullptr only .bytes is meaningful
struct Layout {
  unsigned *bb, *ks_in, *ks_out, *kt_in, *kt_out, *mask, *ctr;
  int *is_in, *is_out, *it_in, *it_out, *nn, *noff, *items, *ioff, *mfold;
  void* tmp;
  float4 *sp, *sa, *tq, *tsph, *bsph;
  float *tcs, *tam, *bcs;
  double *farv, *part;
  unsigned long long* npair;
  size_t bytes;
};

Layout layout(char* base, int64_t ns, int64_t nt, bool same, int T, size_t tmp_bytes) {
  const int64_t nb = (nt + FF_TB - 1) / FF_TB;
  const int64_t ntiles = (ns + T - 1) / T;
  const int64_t mwords = (ntiles + 31) / 32;
  Carve cv{base};
  Layout L{};
  L.bb = cv.take<unsigned>(8);
  L.ks_in = cv.take<unsigned>(ns);
  L.ks_out = cv.take<unsigned>(ns);
  L.is_in = cv.take<int>(ns);
  L.is_out = cv.take<int>(ns);
  L.kt_in = same ? nullptr : cv.take<unsigned>(nt);
  L.kt_out = same ? nullptr : cv.take<unsigned>(nt);
  L.it_in = same ? nullptr : cv.take<int>(nt);
  L.it_out = same ? nullptr : cv.take<int>(nt);
  L.tmp = cv.take<char>(tmp_bytes);
  L.sp = cv.take<float4>(ns);
  L.sa = cv.take<float4>(ns);
  L.tq = cv.take<float4>(nt);
  L.tsph = cv.take<float4>(ntiles);
  L.tcs = cv.take<float>(ntiles);
  L.tam = cv.take<float>(ntiles);
  L.bsph = cv.take<float4>(nb);
  L.bcs = cv.take<float>(nb);
  L.mask = cv.take<unsigned>((size_t)(nb * mwords));
  L.farv = cv.take<double>((size_t)12 * nt);
  L.npair = cv.take<unsigned long long>(1);
  L.nn = cv.take<int>(nb);
  L.noff = cv.take<int>(nb);
  L.items = cv.take<int>(nb);
  L.ioff = cv.take<int>(nb);
  L.mfold = cv.take<int>(1);
  L.ctr = cv.take<unsigned>(1);
  L.part = cv.take<double>((size_t)3 * nb * 12 * FF_TB);   // at most 3 items per block
  L.bytes = cv.off;
  return L;
}

// ws == nullptr: the stream's own workspace (caller holds g_ws_mu); otherwise the
// caller's, of ws_bytes
template <int CORE, bool THICK, int PREC>
int run(const o3d_points* src, o3d_points* targ, bool self_excl, int T, hipStream_t st, o3d_factored_profile* prof,
        void* ws, size_t ws_bytes) {
  const int64_t ns = src->n, nt = targ->n;
  const bool same = same_points(src, targ);
  const int64_t nb = (nt + FF_TB - 1) / FF_TB;
  const int ntiles = (int)((ns + T - 1) / T);
  const int mwords = (ntiles + 31) / 32;

  const size_t tmp_bytes = lib_tmp_bytes(ns, nt, same);

  Layout L = layout(nullptr, ns, nt, same, T, tmp_bytes);
  char* base = static_cast<char*>(ws);
  if (ws) {
    if (ws_bytes < L.bytes) return fail(O3D_EINVAL, "factored path: workspace of %zu bytes, %zu needed", ws_bytes, L.bytes);
  } else {
    void* p = nullptr;
    int rc = stream_workspace(st, L.bytes, &p);
    if (rc != O3D_OK) return rc;
    base = static_cast<char*>(p);
  }
  L = layout(base, ns, nt, same, T, tmp_bytes);

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
  hipLaunchKernelGGL(k_ff_bbox_init, dim3(1), dim3(64), 0, st, L.bb);
  hipLaunchKernelGGL(k_ff_bbox, dim3(std::min(grid_of(ns, 256), FF_BBOX_GRID)), dim3(256), 0, st, sv.x, sv.y, sv.z, ns, L.bb);
  if (!same)
    hipLaunchKernelGGL(k_ff_bbox, dim3(std::min(grid_of(nt, 256), FF_BBOX_GRID)), dim3(256), 0, st, tv.x, tv.y, tv.z, nt,
                       L.bb);
  hipLaunchKernelGGL(k_ff_keys, dim3(grid_of(ns, 256)), dim3(256), 0, st, sv.x, sv.y, sv.z, ns, L.bb, L.ks_in, L.is_in);
  HIPCHK(hipGetLastError());
  unsigned* skeys = nullptr;
  int* sperm = nullptr;
  int rc = radix_sort_pairs(L.ks_in, L.is_in, L.ks_out, L.is_out, ns, 30, L.tmp, st, &skeys, &sperm);
  if (rc != O3D_OK) return rc;
  const int* tperm = sperm;
  if (!same) {
    hipLaunchKernelGGL(k_ff_keys, dim3(grid_of(nt, 256)), dim3(256), 0, st, tv.x, tv.y, tv.z, nt, L.bb, L.kt_in, L.it_in);
    HIPCHK(hipGetLastError());
    unsigned* tkeys = nullptr;
    int* tp = nullptr;
    if ((rc = radix_sort_pairs(L.kt_in, L.it_in, L.kt_out, L.it_out, nt, 30, L.tmp, st, &tkeys, &tp)) != O3D_OK)
      return rc;
    tperm = tp;
  }
  hipLaunchKernelGGL((k_ff_gather_src<CORE>), dim3(grid_of(ns, 256)), dim3(256), 0, st, sv, sperm, ns, L.sp, L.sa);
  hipLaunchKernelGGL((k_ff_gather_tgt<CORE, THICK>), dim3(grid_of(nt, 256)), dim3(256), 0, st, tv, tperm, nt, L.tq);
  hipLaunchKernelGGL(k_ff_bounds, dim3((unsigned)ntiles), dim3(256), 0, st, L.sp, L.sa, ns, T, L.tsph, L.tcs, L.tam);
  hipLaunchKernelGGL(k_ff_bounds, dim3((unsigned)nb), dim3(256), 0, st, L.tq, (const float4*)nullptr, nt, FF_TB, L.bsph,
                     L.bcs, (float*)nullptr);
  hipLaunchKernelGGL((k_ff_classify<CORE>), dim3(grid_of(nb * mwords, 256)), dim3(256), 0, st, L.bsph, L.bcs, nb, L.tsph,
                     L.tcs, ntiles, mwords, L.mask);
  hipLaunchKernelGGL(k_ff_plan_count, dim3(grid_of(nb, 256)), dim3(256), 0, st, L.mask, nb, mwords, ntiles, L.nn);
  HIPCHK(hipGetLastError());
  if ((rc = exclusive_scan_i32(L.nn, L.noff, nb, st)) != O3D_OK) return rc;
  hipLaunchKernelGGL(k_ff_plan_items, dim3(grid_of(nb, 256)), dim3(256), 0, st, L.nn, L.noff, nb, L.items, L.mfold);
  HIPCHK(hipGetLastError());
  if ((rc = exclusive_scan_i32(L.items, L.ioff, nb, st)) != O3D_OK) return rc;
  hipLaunchKernelGGL(k_ff_zero_u32, dim3(1), dim3(64), 0, st, L.ctr);
  HIPCHK(hipGetLastError());

  FFArgs a{};
  a.sp = L.sp; a.sa = L.sa; a.tq = L.tq; a.tperm = tperm;
  a.tsph = L.tsph; a.tcs = L.tcs; a.tam = L.tam; a.bsph = L.bsph; a.bcs = L.bcs;
  a.mask = L.mask; a.farv = L.farv;
  a.nn = L.nn; a.items = L.items; a.ioff = L.ioff; a.mfold = L.mfold; a.ctr = L.ctr; a.part = L.part;
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
    hipLaunchKernelGGL(k_ff_zero_u64, dim3(1), dim3(64), 0, st, L.npair);
    hipLaunchKernelGGL(k_ff_count, dim3(grid_of(nb, 256)), dim3(256), 0, st, L.mask, nb, mwords, ntiles, T, ns, nt,
                       L.npair);
    HIPCHK(hipGetLastError());
    unsigned long long far_pairs = 0;
    HIPCHK(hipMemcpyAsync(&far_pairs, L.npair, sizeof far_pairs, hipMemcpyDeviceToHost, st));
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
  return O3D_OK;
}

template <int CORE>
int run_core(const o3d_points* src, o3d_points* targ, bool thick, int prec, bool self_excl, int T, hipStream_t st,
             o3d_factored_profile* prof, void* ws, size_t wsb) {
  if (thick)
    return prec == PREC_COMP ? run<CORE, true, PREC_COMP>(src, targ, self_excl, T, st, prof, ws, wsb)
                             : run<CORE, true, PREC_FP32>(src, targ, self_excl, T, st, prof, ws, wsb);
  return prec == PREC_COMP ? run<CORE, false, PREC_COMP>(src, targ, self_excl, T, st, prof, ws, wsb)
                           : run<CORE, false, PREC_FP32>(src, targ, self_excl, T, st, prof, ws, wsb);
}

}  // namespace

// Morton order of n points (the direct kernel's target order): workspace bytes,
// and the order itself, written into ws (bounding box, keys, indices and the
// sort's scratch; *perm points into ws)
size_t point_order_bytes(int64_t n) {
  const size_t a = (sizeof(int32_t) * (size_t)n + 255) & ~(size_t)255;
  return 256 + 4 * a + ((radix_tmp_bytes(n) + 255) & ~(size_t)255);
}

int point_order(const float* x, const float* y, const float* z, int64_t n, void* ws, hipStream_t st,
                const int32_t** perm, const unsigned** keys) {
  const size_t a = (sizeof(int32_t) * (size_t)n + 255) & ~(size_t)255;
  char* p = (char*)ws;
  unsigned* bb = (unsigned*)p;
  unsigned* kin = (unsigned*)(p + 256);
  unsigned* kout = (unsigned*)(p + 256 + a);
  int* iin = (int*)(p + 256 + 2 * a);
  int* iout = (int*)(p + 256 + 3 * a);
  void* tmp = p + 256 + 4 * a;
  hipLaunchKernelGGL(k_ff_bbox_init, dim3(1), dim3(64), 0, st, bb);
  hipLaunchKernelGGL(k_ff_bbox, dim3(std::min(grid_of(n, 256), FF_BBOX_GRID)), dim3(256), 0, st, x, y, z, n, bb);
  hipLaunchKernelGGL(k_ff_keys, dim3(grid_of(n, 256)), dim3(256), 0, st, x, y, z, n, bb, kin, iin);
  HIPCHK(hipGetLastError());
  unsigned* sk = nullptr;
  int* sp = nullptr;
  const int rc = radix_sort_pairs(kin, iin, kout, iout, n, 30, tmp, st, &sk, &sp);
  if (rc != O3D_OK) return rc;
  *perm = sp;
  if (keys) *keys = sk;
  return O3D_OK;
}

bool ff_same_points(const o3d_points* src, const o3d_points* targ) { return same_points(src, targ); }

// Sources of a direct EXP call as rows sorted by perm (the targets' Morton order of
// the same collection), and each tile's bounding sphere and largest core term, in
// ws (sorted_sources_bytes) after the point order's own point_order_bytes(n)
size_t sorted_sources_bytes(int64_t n, int tile) {
  const size_t ntiles = (size_t)((n + tile - 1) / tile);
  const auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
  return 2 * r(sizeof(float4) * (size_t)n) + r(sizeof(float4) * ntiles) + r(sizeof(float) * ntiles);
}

int sorted_sources(const o3d_points* src, const int32_t* perm, int tile, void* ws, hipStream_t st, P2PArgs* a) {
  const int64_t n = src->n;
  const int ntiles = (int)((n + tile - 1) / tile);
  const auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
  char* p = (char*)ws;
  float4* sp = (float4*)p;
  float4* sa = (float4*)(p + r(sizeof(float4) * (size_t)n));
  float4* sph = (float4*)(p + 2 * r(sizeof(float4) * (size_t)n));
  float* cs = (float*)(p + 2 * r(sizeof(float4) * (size_t)n) + r(sizeof(float4) * (size_t)ntiles));
  const SrcView sv{src->x[0], src->x[1], src->x[2], src->r, src->s[0], src->s[1], src->s[2]};
  hipLaunchKernelGGL((k_ff_gather_src<CORE_EXP>), dim3(grid_of(n, 256)), dim3(256), 0, st, sv, perm, n, sp, sa);
  hipLaunchKernelGGL(k_ff_bounds, dim3((unsigned)ntiles), dim3(256), 0, st, (const float4*)sp, (const float4*)nullptr,
                     n, tile, sph, cs, (float*)nullptr);
  HIPCHK(hipGetLastError());
  a->sp = sp;
  a->sa = sa;
  a->tsph = sph;
  a->tcs = cs;
  return O3D_OK;
}

std::mutex& stream_ws_mutex() { return g_ws_mu; }

// caller holds stream_ws_mutex() for the whole enqueue, so that no other thread
// frees or grows this stream's workspace between its sizing and the last launch
// using it; a larger request waits for the stream to drain, then reallocates.  A
// stream being captured into a graph cannot do that: the workspace must already
// be large enough (a call of the same size outside the capture sizes it).
bool stream_capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

bool stream_workspace_fits(hipStream_t st, size_t bytes) {
  auto it = g_ws.find(st);
  return it != g_ws.end() && it->second.bytes >= bytes;
}

void release_stream_workspace(hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto it = g_ws.find(st);
  if (it == g_ws.end()) return;
  if (it->second.p) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(it->second.p);
  }
  g_ws.erase(it);
}

int stream_workspace(hipStream_t st, size_t bytes, void** out) {
  Work& w = g_ws[st];
  if (w.bytes < bytes) {
    if (stream_capturing(st))
      return fail(O3D_EINVAL, "the stream is being captured and its device workspace would have to grow to %zu "
                              "bytes; make the same call once outside the capture first", bytes);
    if (w.p) {
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipFree(w.p));
      w.p = nullptr;
      w.bytes = 0;
    }
    const size_t want = bytes + bytes / 4;
    if (hipMalloc(&w.p, want) != hipSuccess) {
      w.p = nullptr;
      return fail(O3D_ENOMEM, "device workspace of %zu bytes: allocation failed", want);
    }
    w.bytes = want;
  }
  *out = w.p;
  return O3D_OK;
}

bool ff_supported(int core, bool grad, int prec) {
  return (core == CORE_RM || core == CORE_EXP) && grad && (prec == PREC_FP32 || prec == PREC_COMP);
}

namespace {
int check_factored(const o3d_points* src, const o3d_points* targ, const o3d_opts* o, int* T) {
  const bool grad = o->results == O3D_VELANDGRAD;
  if (!ff_supported(o->core, grad, o->precision))
    return fail(O3D_EINVAL, "factored method: needs the RM or EXP core, velocity and gradient, and FP32 or "
                            "compensated precision");
  if (src->n > INT_MAX || targ->n > INT_MAX)
    return fail(O3D_EINVAL, "factored method: at most %d sources and %d targets", INT_MAX, INT_MAX);
  *T = o->tile > 0 ? o->tile : 256;
  if (*T != 64 && *T != 128 && *T != 256) return fail(O3D_EINVAL, "factored method: tile must be 0, 64, 128 or 256");
  return O3D_OK;
}

int dispatch(const o3d_points* src, o3d_points* targ, const o3d_opts* o, int T, hipStream_t st,
             o3d_factored_profile* prof, void* ws, size_t wsb) {
  const bool thick = targ->r != nullptr;
  const bool self_excl = o->self_term == O3D_SELF_EXCLUDE;
  if (o->core == CORE_RM) return run_core<CORE_RM>(src, targ, thick, o->precision, self_excl, T, st, prof, ws, wsb);
  return run_core<CORE_EXP>(src, targ, thick, o->precision, self_excl, T, st, prof, ws, wsb);
}
}  // namespace

// device-pointer point->point by the factored far field; arguments already checked
int p2p_factored(const o3d_points* src, o3d_points* targ, const o3d_opts* o, hipStream_t st,
                 o3d_factored_profile* prof) {
  int T = 0, rc = check_factored(src, targ, o, &T);
  if (rc != O3D_OK) return rc;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  return dispatch(src, targ, o, T, st, prof, nullptr, 0);
}

// the same on scratch the caller owns (ws_bytes >= ff_workspace_bytes of this shape);
// nothing is allocated and no lock is taken
int p2p_factored_ws(const o3d_points* src, o3d_points* targ, const o3d_opts* o, hipStream_t st, void* ws,
                    size_t ws_bytes) {
  int T = 0, rc = check_factored(src, targ, o, &T);
  if (rc != O3D_OK) return rc;
  if (!ws) return fail(O3D_EINVAL, "factored method: NULL workspace");
  return dispatch(src, targ, o, T, st, nullptr, ws, ws_bytes);
}

size_t ff_workspace_bytes(int64_t ns, int64_t nt, bool same, int tile) {
  const int T = tile > 0 ? tile : 256;
  return layout(nullptr, ns, nt, same, T, lib_tmp_bytes(ns, nt, same)).bytes;
}

}  // namespace o3d

extern "C" int o3d_morton_order_dev(const float* x, const float* y, const float* z, int64_t n, uint32_t* keys,
                                    int32_t* perm, void* stream) {
  using namespace o3d;
  clear_error();
  int rc = ensure_init();
  if (rc != O3D_OK) return rc;
  if (n < 0 || n > INT32_MAX) return fail(O3D_EINVAL, "morton order: n %lld out of range", (long long)n);
  if (n == 0) return O3D_OK;
  if (!x || !y || !z || !keys || !perm) return fail(O3D_EINVAL, "morton order: NULL array");
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  void* ws = nullptr;
  if ((rc = stream_workspace(st, point_order_bytes(n), &ws)) != O3D_OK) return rc;
  const int32_t* p = nullptr;
  if ((rc = point_order(x, y, z, n, ws, st, &p)) != O3D_OK) return rc;
  HIPCHK(hipMemcpyAsync(perm, p, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, st));
  // the keys in index order, from the same box (point_order left it at the workspace's start)
  hipLaunchKernelGGL(k_ff_keys, dim3(grid_of(n, 256)), dim3(256), 0, st, x, y, z, n, (const unsigned*)ws,
                     (unsigned*)keys, (int*)nullptr);
  HIPCHK(hipGetLastError());
  return O3D_OK;
}
