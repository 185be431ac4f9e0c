// o3d_api.hip -- the C ABI of libo3dbiot.so (declared in include/o3d_biot.h).
//
// Host layer between callers shaped like Omega3D's Influence.h and the gfx950
// kernels: argument checks (status codes instead of the reference's asserts,
// Influence.h:369,537), staging of host arrays through a grow-only device pool,
// tile / targets-per-thread selection, and the flop accounting of
// Influence.h:75,310,475 (N_t * (12 + F * N_s) with F from Kernels.h).
#include "o3d_api_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace o3d {

// ---- error reporting ---------------------------------------------------------
static thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
void clear_error() { g_err.clear(); }
const char* last_error() { return g_err.c_str(); }

// ---- device context ------------------------------------------------------------
struct DevPool {
  std::vector<std::pair<void*, size_t>> bufs;  // slot -> (ptr, bytes)
};

struct Context {
  std::mutex mu;
  bool ready = false;
  int device = -1;
  int status = O3D_OK;
  hipDeviceProp_t prop{};
  hipStream_t stream = nullptr;
  DevPool pool;
};

static Context g_ctx;
static int g_ext_core = CORE_WL;

static int init_locked(int device) {
  if (g_ctx.ready && (device < 0 || device == g_ctx.device)) return O3D_OK;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return fail(O3D_ENODEV, "no HIP device visible (%s)", hipGetErrorString(e));
  if (device < 0) {
    if (hipGetDevice(&device) != hipSuccess) device = 0;
  }
  if (device >= ndev) return fail(O3D_ENODEV, "device %d out of range (%d devices)", device, ndev);
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipGetDeviceProperties(&g_ctx.prop, device));
  if (std::strncmp(g_ctx.prop.gcnArchName, "gfx950", 6) != 0)
    return fail(O3D_ENODEV, "device %d is %s; this library is built for gfx950 only", device, g_ctx.prop.gcnArchName);
  if (g_ctx.stream == nullptr) HIPCHK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  g_ctx.device = device;
  g_ctx.ready = true;
  noalloc_setup(device);   // a failure is kept for the no-allocation entry point alone
  return O3D_OK;
}

int ensure_init() {
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  if (g_ctx.ready) {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != g_ctx.device) HIPCHK(hipSetDevice(g_ctx.device));
    return O3D_OK;
  }
  return init_locked(-1);
}

int num_cus() { return g_ctx.ready ? g_ctx.prop.multiProcessorCount : 256; }

// grow-only pool of device scratch buffers, one per slot; callers hold g_pool_mu
static std::mutex g_pool_mu;
static int pool_get(int slot, size_t bytes, void** out) {
  auto& b = g_ctx.pool.bufs;
  if ((int)b.size() <= slot) b.resize(slot + 1, {nullptr, 0});
  if (b[slot].second < bytes) {
    if (b[slot].first) {
      HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipFree(b[slot].first));
      b[slot] = {nullptr, 0};
    }
    size_t want = bytes + bytes / 4 + 256;
    if (hipMalloc(&b[slot].first, want) != hipSuccess) {
      b[slot] = {nullptr, 0};
      return fail(O3D_ENOMEM, "device allocation of %zu bytes failed", want);
    }
    b[slot].second = want;
  }
  *out = b[slot].first;
  return O3D_OK;
}

// ---- launch configuration -------------------------------------------------------
// Targets per thread: enough workgroups to cover every CU several times, more
// register reuse of each staged source when the target count allows it.
void choose_config(const Tuning& req, int64_t nt, int64_t ns, bool grad, int prec, int* tile, int* tpt) {
  // measured at N = 1M (velocity + gradient, both FP32 modes): tile 512 with one
  // target per lane is fastest; more targets per lane cost occupancy
  int tl = req.tile > 0 ? req.tile : (ns >= 65536 ? 512 : 256);
  int tp = req.tpt;
  if (tp <= 0) {
    const int64_t cus = num_cus();
    tp = 1;
    if (!grad && prec != PREC_FP64 && nt >= (int64_t)BLOCK * 4 * cus * 4) tp = 4;
  }
  *tile = tl;
  *tpt = tp;
}

hipError_t dispatch_p2p(int core, const P2PArgs& a, bool thick, bool grad, int prec, int tpt, hipStream_t st) {
  switch (core) {
    case CORE_RM: return launch_p2p_rm(a, thick, grad, prec, tpt, st);
    case CORE_WL: return launch_p2p_wl(a, thick, grad, prec, tpt, st);
    case CORE_EXP: return launch_p2p_exp(a, thick, grad, prec, tpt, st);
    case CORE_V2: return launch_p2p_v2(a, thick, grad, prec, tpt, st);
  }
  return hipErrorInvalidValue;
}
hipError_t dispatch_p2p_split(int core, const P2PArgs& a, bool thick, bool grad, int prec, int nslice, int64_t slen,
                              void* part, hipStream_t st) {
  switch (core) {
    case CORE_RM: return launch_p2p_split_rm(a, thick, grad, prec, nslice, slen, part, st);
    case CORE_WL: return launch_p2p_split_wl(a, thick, grad, prec, nslice, slen, part, st);
    case CORE_EXP: return launch_p2p_split_exp(a, thick, grad, prec, nslice, slen, part, st);
    case CORE_V2: return launch_p2p_split_v2(a, thick, grad, prec, nslice, slen, part, st);
  }
  return hipErrorInvalidValue;
}

// Source split of a call whose targets make fewer workgroups than the GPU holds at
// once: slices of whole tiles, enough of them for about SPLIT_WGS_PER_CU
// workgroups per compute unit, and a partial buffer of at most SPLIT_MAX_BYTES.
// cap > 0 (a caller's fixed scratch of cap bytes): fewer slices when the partials
// of the wanted slice count do not fit.  Returns the slice count (1: no split) and
// the slice length.
constexpr int SPLIT_WGS_PER_CU = 16;
constexpr size_t SPLIT_MAX_BYTES = (size_t)256 << 20;
int split_plan(int64_t nt, int64_t ns, int tile, bool grad, int prec, int64_t* slen, size_t* bytes, size_t cap) {
  const int64_t nb = (nt + BLOCK - 1) / BLOCK;
  const int64_t want = (int64_t)SPLIT_WGS_PER_CU * num_cus();
  if (nb >= want) return 1;
  int64_t ntiles = (ns + tile - 1) / tile;
  int64_t nsl = (want + nb - 1) / nb;
  if (nsl > ntiles) nsl = ntiles;
  if (nsl > 65535) nsl = 65535;
  const size_t cols = prec == PREC_COMP ? (grad ? 12 : 3) * sizeof(double) : (grad ? 14 : 3) * sizeof(float);
  if (cap > 0) {
    const int64_t fit = (int64_t)(cap / ((size_t)nt * cols));
    if (nsl > fit) nsl = fit;
  }
  if (nsl < 2) return 1;
  const int64_t tiles_per = (ntiles + nsl - 1) / nsl;
  *slen = tiles_per * tile;
  nsl = (ns + *slen - 1) / *slen;
  *bytes = (size_t)nsl * (size_t)nt * cols;
  if (nsl < 2 || *bytes > (cap > 0 ? cap : SPLIT_MAX_BYTES)) return 1;
  return (int)nsl;
}

hipError_t dispatch_p2p64(int core, const P2PArgs64& a, bool thick, bool grad, hipStream_t st) {
  switch (core) {
    case CORE_RM: return launch_p2p64_rm(a, thick, grad, st);
    case CORE_WL: return launch_p2p64_wl(a, thick, grad, st);
    case CORE_EXP: return launch_p2p64_exp(a, thick, grad, st);
    case CORE_V2: return launch_p2p64_v2(a, thick, grad, st);
  }
  return hipErrorInvalidValue;
}

int check_opts(const o3d_opts* o) {
  if (!o) return fail(O3D_EINVAL, "opts is NULL");
  if (o->core < 1 || o->core > 4) return fail(O3D_EINVAL, "unknown core %d", o->core);
  if (o->results != O3D_VELONLY && o->results != O3D_VELANDGRAD)
    return fail(O3D_EINVAL, "results type %d is not served (velonly=1, velandgrad=2)", o->results);
  if (o->precision < 0 || o->precision > 2) return fail(O3D_EINVAL, "unknown precision %d", o->precision);
  if (o->self_term != O3D_SELF_INCLUDE && o->self_term != O3D_SELF_EXCLUDE)
    return fail(O3D_EINVAL, "self_term must be 0 or 1");
  if (o->tile != 0 && (o->tile < 64 || o->tile > 1024 || o->tile % 64 != 0))
    return fail(O3D_EINVAL, "tile %d: must be 0 or a multiple of 64 in [64, 1024]", o->tile);
  if (o->tpt != 0 && o->tpt != 1 && o->tpt != 2 && o->tpt != 4) return fail(O3D_EINVAL, "tpt %d: must be 0, 1, 2 or 4", o->tpt);
  if (o->method < O3D_METHOD_AUTO || o->method > O3D_METHOD_FACTORED)
    return fail(O3D_EINVAL, "unknown method %d", o->method);
  if (o->method == O3D_METHOD_FACTORED && o->precision == PREC_FP64)
    return fail(O3D_EINVAL, "the factored method computes in FP32 (plain or compensated), not FP64");
  return O3D_OK;
}

int check_node_indices(const o3d_panels* p, const char* what) {
  const uint32_t nn = (uint32_t)p->nn;
  for (int64_t i = 0; i < 3 * p->np; ++i)
    if (p->idx[i] >= nn)
      return fail(O3D_EINVAL, "%s: panel %lld names node %u of %lld", what, (long long)(i / 3), p->idx[i],
                  (long long)p->nn);
  return O3D_OK;
}

template <class P>
int check_points(const P* p, bool as_src, bool as_tgt, bool grad, const char* what) {
  if (!p) return fail(O3D_EINVAL, "%s is NULL", what);
  if (p->n < 0) return fail(O3D_EINVAL, "%s: negative count", what);
  if (p->n == 0) return O3D_OK;
  for (int d = 0; d < 3; ++d)
    if (!p->x[d]) return fail(O3D_EINVAL, "%s: position component %d is NULL", what, d);
  if (as_src) {
    for (int d = 0; d < 3; ++d)
      if (!p->s[d]) return fail(O3D_EINVAL, "%s: strength component %d is NULL", what, d);
    if (!p->r) return fail(O3D_EINVAL, "%s: source core radii are NULL", what);
  }
  if (as_tgt) {
    for (int d = 0; d < 3; ++d)
      if (!p->u[d]) return fail(O3D_EINVAL, "%s: velocity component %d is NULL", what, d);
    if (grad)
      for (int k = 0; k < 9; ++k)
        if (!p->ug[k])
          return fail(O3D_EINVAL, "%s: velocity-gradient component %d is NULL but results = velandgrad "
                                  "(the reference asserts here, Influence.h:369)", what, k);
  }
  return O3D_OK;
}

template <class P>
int check_self(const P* src, const P* targ, const o3d_opts* o) {
  if (o->self_term == O3D_SELF_EXCLUDE && src->n != targ->n)
    return fail(O3D_EINVAL, "self_term exclude needs sources and targets to be one collection (n %lld vs %lld)",
                (long long)src->n, (long long)targ->n);
  return O3D_OK;
}

// the argument checks of a float point -> point call
int check_p2p_args(const o3d_points* src, const o3d_points* targ, const o3d_opts* o) {
  int rc;
  if ((rc = check_opts(o)) != O3D_OK) return rc;
  const bool grad = o->results == O3D_VELANDGRAD;
  if ((rc = check_points(src, true, false, grad, "src")) != O3D_OK) return rc;
  if ((rc = check_points(targ, false, true, grad, "targ")) != O3D_OK) return rc;
  return check_self(src, targ, o);
}

double flops_of(int core, bool grad, bool thick, int64_t nt, int64_t ns) {
  const int f = o3d_flops_per_interaction(core, grad ? O3D_VELANDGRAD : O3D_VELONLY, thick, 0);
  return (double)nt * ((grad ? 12.0 : 3.0) + (double)f * (double)ns);
}

// device-pointer point->point; all pointers device, float
int p2p_dev(const o3d_points* src, o3d_points* targ, const o3d_opts* o, hipStream_t st, const Tuning* tun) {
  const bool grad = o->results == O3D_VELANDGRAD;
  if (src->n == 0 || targ->n == 0) return O3D_OK;
  if (o->precision == PREC_FP64) return fail(O3D_EINVAL, "internal: FP64 routed to p2p_dev");
  if (o->method == O3D_METHOD_FACTORED) return p2p_factored(src, targ, o, st);
  if (o->method == O3D_METHOD_AUTO && !tun && o->tile == 0 && o->tpt == 0 && ff_supported(o->core, grad, o->precision) &&
      src->n >= FF_AUTO_MIN_N && targ->n >= FF_AUTO_MIN_N && src->n <= INT32_MAX && targ->n <= INT32_MAX) {
    // under graph capture the stream's workspace cannot grow: when it is too small
    // for the factored method, the direct kernel (which needs none) takes the call
    bool ff = true;
    if (stream_capturing(st)) {
      std::lock_guard<std::mutex> lk(stream_ws_mutex());
      ff = stream_workspace_fits(st, ff_workspace_bytes(src->n, targ->n, ff_same_points(src, targ), 0));
    }
    if (ff) return p2p_factored(src, targ, o, st);
  }
  return p2p_direct(src, targ, o, st, tun, nullptr, 0);
}

int p2p_direct(const o3d_points* src, o3d_points* targ, const o3d_opts* o, hipStream_t st, const Tuning* tun,
               void* ws, size_t ws_bytes) {
  const bool grad = o->results == O3D_VELANDGRAD;
  const bool thick = targ->r != nullptr;
  if (src->n == 0 || targ->n == 0) return O3D_OK;
  P2PArgs a{};
  a.s = SrcView{src->x[0], src->x[1], src->x[2], src->r, src->s[0], src->s[1], src->s[2]};
  a.t.x = targ->x[0]; a.t.y = targ->x[1]; a.t.z = targ->x[2]; a.t.r = targ->r;
  a.t.u = targ->u[0]; a.t.v = targ->u[1]; a.t.w = targ->u[2];
  for (int k = 0; k < 9; ++k) a.t.g[k] = grad ? targ->ug[k] : nullptr;
  a.ns = src->n;
  a.nt = targ->n;
  a.self_excl = o->self_term == O3D_SELF_EXCLUDE;
  Tuning req{o->tile, o->tpt};
  const bool auto_cfg = !tun && o->tile == 0 && o->tpt == 0;
  if (tun) req = *tun;
  else if (auto_cfg) {
    Tuning t;
    if (tuned_lookup(o->core, o->results, o->precision, thick, a.nt, &t)) req = t;
  }
  int tile, tpt;
  choose_config(req, a.nt, a.ns, grad, o->precision, &tile, &tpt);
  a.tile = tile;
  // too few targets to fill the GPU: split the sources (a configuration left to the
  // library only; an explicit tile or lane count keeps the one-pass order that the
  // batched and served kernels share)
  int64_t slen = 0;
  size_t pbytes = 0;
  const int nsl = auto_cfg && tpt == 1 ? split_plan(a.nt, a.ns, tile, grad, o->precision, &slen, &pbytes,
                                                    ws ? ws_bytes : 0)
                                       : 1;
  if (nsl > 1) {
    void* part = ws;
    // an EXP call also takes its targets in Morton order (the same bits), and on one
    // collection its sources as sorted rows, as the one-pass kernel below does (in
    // the stream's workspace after the partial sums); the partials are then in that
    // order
    const bool ord = !ws && o->core == CORE_EXP && a.nt >= TARGET_ORDER_MIN_N && a.nt <= INT32_MAX;
    const bool srt = ord && ff_same_points(src, targ);
    const size_t pal = (pbytes + 255) & ~(size_t)255;
    const size_t obytes = ord ? point_order_bytes(a.nt) + (srt ? sorted_sources_bytes(a.ns, tile) : 0) : 0;
    bool ordered = false;
    std::unique_lock<std::mutex> lk(stream_ws_mutex(), std::defer_lock);
    if (!ws) {
      lk.lock();
      // under graph capture the workspace cannot grow: without room for the order,
      // index order; without room for the partials, the one-pass kernel
      const bool cap = stream_capturing(st);
      if (ord && (!cap || stream_workspace_fits(st, pal + obytes))) {
        const int rc = stream_workspace(st, pal + obytes, &part);
        if (rc != O3D_OK) return rc;
        ordered = true;
      } else if (cap && !stream_workspace_fits(st, pbytes)) {
        part = nullptr;
      } else {
        const int rc = stream_workspace(st, pbytes, &part);
        if (rc != O3D_OK) return rc;
      }
    }
    if (part) {
      if (ordered) {
        char* ow = (char*)part + pal;
        int rc = point_order(a.t.x, a.t.y, a.t.z, a.nt, ow, st, &a.tperm);
        if (rc != O3D_OK) return rc;
        if (srt && (rc = sorted_sources(src, a.tperm, tile, ow + point_order_bytes(a.nt), st, &a)) != O3D_OK)
          return rc;
      }
      hipError_t e = dispatch_p2p_split(o->core, a, thick, grad, o->precision, nsl, slen, part, st);
      if (e != hipSuccess) return fail(O3D_EDEVICE, "p2p launch failed: %s", hipGetErrorString(e));
      return O3D_OK;
    }
  }
  // EXP core: the kernel takes the exponential branch for a whole wave when any of
  // its targets needs it for the current source, so large calls visit the targets
  // in Morton order, where a wave's targets are neighbours.  A target's sums do
  // not depend on the order: the results are the same bits.  Under graph capture
  // the order needs the stream's workspace to be large enough already; otherwise
  // the targets are visited in index order.
  //
  // A call on one collection with the configuration left to the library also reads
  // its sources as rows sorted in that order: a tile of sorted sources is compact,
  // and a tile beyond the far-branch distance of every target of a wave is summed
  // with no per-pair test.  The sums then run in another order (the pair values
  // are the same bits).  Not for a caller's fixed scratch (ws).
  if (o->core == CORE_EXP && a.nt >= TARGET_ORDER_MIN_N && a.nt <= INT32_MAX) {
    const size_t bytes = point_order_bytes(a.nt);
    const bool want_sorted = !ws && auto_cfg && tpt == 1 && ff_same_points(src, targ);
    const size_t sbytes = want_sorted ? bytes + sorted_sources_bytes(a.ns, tile) : bytes;
    bool sorted = false;
    void* ow = ws;
    std::unique_lock<std::mutex> lk(stream_ws_mutex(), std::defer_lock);
    if (ws) {
      if (bytes > ws_bytes)
        return fail(O3D_EINVAL, "the target order of %lld targets needs %zu bytes of scratch, %zu given",
                    (long long)a.nt, bytes, ws_bytes);
    } else {
      lk.lock();
      const bool cap = stream_capturing(st);
      if (!cap || stream_workspace_fits(st, sbytes)) {
        const int rc = stream_workspace(st, sbytes, &ow);
        if (rc != O3D_OK) return rc;
        sorted = want_sorted;
      } else if (stream_workspace_fits(st, bytes)) {
        const int rc = stream_workspace(st, bytes, &ow);
        if (rc != O3D_OK) return rc;
      } else {
        ow = nullptr;
      }
    }
    if (ow) {
      int rc = point_order(a.t.x, a.t.y, a.t.z, a.nt, ow, st, &a.tperm);
      if (rc != O3D_OK) return rc;
      if (sorted && (rc = sorted_sources(src, a.tperm, tile, (char*)ow + bytes, st, &a)) != O3D_OK) return rc;
      hipError_t e = dispatch_p2p(o->core, a, thick, grad, o->precision, tpt, st);
      if (e != hipSuccess) return fail(O3D_EDEVICE, "p2p launch failed: %s", hipGetErrorString(e));
      return O3D_OK;
    }
  }
  hipError_t e = dispatch_p2p(o->core, a, thick, grad, o->precision, tpt, st);
  if (e != hipSuccess) return fail(O3D_EDEVICE, "p2p launch failed: %s", hipGetErrorString(e));
  return O3D_OK;
}

int p2p64_dev(const o3d_points64* src, o3d_points64* targ, const o3d_opts* o, hipStream_t st) {
  const bool grad = o->results == O3D_VELANDGRAD;
  const bool thick = targ->r != nullptr;
  if (src->n == 0 || targ->n == 0) return O3D_OK;
  P2PArgs64 a{};
  a.s = SrcView64{src->x[0], src->x[1], src->x[2], src->r, src->s[0], src->s[1], src->s[2]};
  a.t.x = targ->x[0]; a.t.y = targ->x[1]; a.t.z = targ->x[2]; a.t.r = targ->r;
  a.t.u = targ->u[0]; a.t.v = targ->u[1]; a.t.w = targ->u[2];
  for (int k = 0; k < 9; ++k) a.t.g[k] = grad ? targ->ug[k] : nullptr;
  a.ns = src->n;
  a.nt = targ->n;
  a.tile = 128;
  a.self_excl = o->self_term == O3D_SELF_EXCLUDE;
  hipError_t e = dispatch_p2p64(o->core, a, thick, grad, st);
  if (e != hipSuccess) return fail(O3D_EDEVICE, "p2p64 launch failed: %s", hipGetErrorString(e));
  return O3D_OK;
}

// host arrays -> pool -> kernel -> host arrays (accumulate)
template <class T, class P>
int p2p_host_impl(const P* src, P* targ, const o3d_opts* o, o3d_stats* stats) {
  const bool grad = o->results == O3D_VELANDGRAD;
  const bool thick = targ->r != nullptr;
  const int64_t ns = src->n, nt = targ->n;
  if (ns == 0 || nt == 0) return O3D_OK;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  hipStream_t st = g_ctx.stream;
  const size_t sb = sizeof(T) * (size_t)ns, tb = sizeof(T) * (size_t)nt;
  // a collection acting on itself (the same host position arrays) has its
  // positions copied once, so the device call sees one collection; the radii
  // are shared too when they are the same host array
  const bool alias = ns == nt && src->x[0] == targ->x[0] && src->x[1] == targ->x[1] && src->x[2] == targ->x[2];
  const bool alias_r = alias && targ->r == src->r;
  void* dsrc;
  void* dtgt;
  int rc;
  if ((rc = pool_get(0, 7 * sb, &dsrc)) != O3D_OK) return rc;
  if ((rc = pool_get(1, (grad ? 16 : 7) * tb, &dtgt)) != O3D_OK) return rc;
  T* s = (T*)dsrc;
  T* t = (T*)dtgt;
  P ds{}, dt{};
  ds.n = ns;
  for (int d = 0; d < 3; ++d) { ds.x[d] = s + d * ns; ds.s[d] = s + (3 + d) * ns; }
  ds.r = s + 6 * ns;
  for (int d = 0; d < 3; ++d) HIPCHK(hipMemcpyAsync((void*)ds.x[d], src->x[d], sb, hipMemcpyHostToDevice, st));
  for (int d = 0; d < 3; ++d) HIPCHK(hipMemcpyAsync((void*)ds.s[d], src->s[d], sb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void*)ds.r, src->r, sb, hipMemcpyHostToDevice, st));
  dt.n = nt;
  for (int d = 0; d < 3; ++d) { dt.x[d] = alias ? ds.x[d] : t + d * nt; dt.u[d] = t + (3 + d) * nt; }
  dt.r = thick ? (alias_r ? ds.r : t + 6 * nt) : nullptr;
  if (grad) for (int k = 0; k < 9; ++k) dt.ug[k] = t + (7 + k) * nt;
  if (!alias)
    for (int d = 0; d < 3; ++d) HIPCHK(hipMemcpyAsync((void*)dt.x[d], targ->x[d], tb, hipMemcpyHostToDevice, st));
  if (thick && !alias_r) HIPCHK(hipMemcpyAsync((void*)dt.r, targ->r, tb, hipMemcpyHostToDevice, st));
  for (int d = 0; d < 3; ++d) HIPCHK(hipMemcpyAsync(dt.u[d], targ->u[d], tb, hipMemcpyHostToDevice, st));
  if (grad) for (int k = 0; k < 9; ++k) HIPCHK(hipMemcpyAsync(dt.ug[k], targ->ug[k], tb, hipMemcpyHostToDevice, st));
  if constexpr (sizeof(T) == 4) rc = p2p_dev((const o3d_points*)&ds, (o3d_points*)&dt, o, st, nullptr);
  else rc = p2p64_dev((const o3d_points64*)&ds, (o3d_points64*)&dt, o, st);
  if (rc != O3D_OK) return rc;
  for (int d = 0; d < 3; ++d) HIPCHK(hipMemcpyAsync(targ->u[d], dt.u[d], tb, hipMemcpyDeviceToHost, st));
  if (grad) for (int k = 0; k < 9; ++k) HIPCHK(hipMemcpyAsync(targ->ug[k], dt.ug[k], tb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  (void)stats;
  return O3D_OK;
}

// float host arrays through the FP64 kernel: promote, run, accumulate back
int p2p_host_fp64_from_f32(const o3d_points* src, o3d_points* targ, const o3d_opts* o) {
  const bool grad = o->results == O3D_VELANDGRAD;
  const int64_t ns = src->n, nt = targ->n;
  std::vector<double> sb(7 * (size_t)ns), tb(16 * (size_t)nt, 0.0);
  o3d_points64 s{}, t{};
  s.n = ns;
  for (int d = 0; d < 3; ++d) {
    s.x[d] = sb.data() + d * ns;
    s.s[d] = sb.data() + (3 + d) * ns;
    for (int64_t j = 0; j < ns; ++j) {
      sb[d * ns + j] = src->x[d][j];
      sb[(3 + d) * ns + j] = src->s[d][j];
    }
  }
  s.r = sb.data() + 6 * ns;
  for (int64_t j = 0; j < ns; ++j) sb[6 * ns + j] = src->r[j];
  t.n = nt;
  for (int d = 0; d < 3; ++d) {
    t.x[d] = tb.data() + d * nt;
    t.u[d] = tb.data() + (3 + d) * nt;
    for (int64_t j = 0; j < nt; ++j) tb[d * nt + j] = targ->x[d][j];
  }
  t.r = nullptr;
  if (targ->r) {
    t.r = tb.data() + 6 * nt;
    for (int64_t j = 0; j < nt; ++j) tb[6 * nt + j] = targ->r[j];
  }
  if (grad) for (int k = 0; k < 9; ++k) t.ug[k] = tb.data() + (7 + k) * nt;
  int rc = p2p_host_impl<double>(&s, &t, o, nullptr);
  if (rc != O3D_OK) return rc;
  for (int d = 0; d < 3; ++d)
    for (int64_t j = 0; j < nt; ++j) targ->u[d][j] = (float)((double)targ->u[d][j] + t.u[d][j]);
  if (grad)
    for (int k = 0; k < 9; ++k)
      for (int64_t j = 0; j < nt; ++j) targ->ug[k][j] = (float)((double)targ->ug[k][j] + t.ug[k][j]);
  return O3D_OK;
}

void fill_stats(o3d_stats* stats, double secs, double flops, double inter) {
  if (!stats) return;
  stats->seconds = secs;
  stats->flops = flops;
  stats->interactions = inter;
}

int ext_core() { return g_ext_core; }
void set_ext_core(int c) { g_ext_core = c; }
Context& ctx() { return g_ctx; }
hipStream_t lib_stream() { return g_ctx.stream; }
int lib_device() { return g_ctx.device; }
std::mutex& pool_mutex() { return g_pool_mu; }
int pool_buffer(int slot, size_t bytes, void** out) { return pool_get(slot, bytes, out); }

__global__ void k_copy_probe(float4* __restrict__ d, const float4* __restrict__ s, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) d[i] = s[i];
}

// VALU issue-rate probe: 16 independent chains of one FP32 instruction per lane
// (op 0 v_fmac_f32 with a scalar-register operand, 1 v_add_f32, 2 v_mul_f32,
// 3 v_rsq_f32, 4 v_fmac_f32 with vector-register operands only), four steps of
// each chain per loop iteration (64 instructions between two loop branches, so the
// branch and counter cost ~1/20 of the loop), iters times
template <int OP>
__global__ __launch_bounds__(BLOCK) void k_valu_probe(float* out, int iters, float b) {
  float x[16], y[16], z[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    x[j] = threadIdx.x * 1e-3f + j;
    y[j] = 1.0f + (j + 1) * 1e-4f;   // never 1: every chain keeps its instruction
    z[j] = 1e-7f * (float)(threadIdx.x + j + 1);   // per lane: not folded into a literal
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int jj = 0; jj < 64; ++jj) {
      const int j = jj & 15;
      if (OP == 0) x[j] = fmaf(y[j], b, x[j]);
      else if (OP == 1) x[j] = x[j] + y[j];
      else if (OP == 2) x[j] = x[j] * y[j];
      else if (OP == 3) x[j] = __builtin_amdgcn_rsqf(x[j]);
      else x[j] = fmaf(z[j], z[(j + 1) & 15], x[j]);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int j = 0; j < 16; ++j) s += x[j];
  out[(int64_t)blockIdx.x * BLOCK + threadIdx.x] = s;
}

// matrix-core and VALU issue together: per iteration four independent
// v_mfma_f32_32x32x16_f16 (accumulators in registers) and/or 64 v_fmac_f32 with
// vector-register operands (16 independent chains).  MODE 0: FMAs only; 1: MFMAs
// only; 2: both in every wave; 3: waves 0-3 of the block MFMAs, waves 4-7 FMAs.
typedef float o3d_f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 o3d_f16x8 __attribute__((ext_vector_type(8)));
template <int MODE>
__global__ __launch_bounds__(512, 1) void k_mix_probe(float* out, int iters) {
  const int wave = threadIdx.x >> 6;
  const bool mf = MODE == 1 || MODE == 2 || (MODE == 3 && wave < 4);
  const bool va = MODE == 0 || MODE == 2 || (MODE == 3 && wave >= 4);
  o3d_f32x16 acc[4];
  o3d_f16x8 A, B;
  float x[16], z[16];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[m][r] = 0.0f;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    A[e] = (_Float16)(1e-3f * (float)(threadIdx.x + e));
    B[e] = (_Float16)(1e-3f * (float)e);
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    x[j] = threadIdx.x * 1e-3f + j;
    z[j] = 1e-7f * (float)(threadIdx.x + j + 1);
  }
  for (int it = 0; it < iters; ++it) {
    if (mf) {
#pragma unroll
      for (int m = 0; m < 4; ++m) acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_f16(A, B, acc[m], 0, 0, 0);
    }
    if (va) {
#pragma unroll
      for (int rep = 0; rep < 4; ++rep)
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = fmaf(z[j], z[(j + 1) & 15], x[j]);
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int r = 0; r < 16; ++r) s += acc[m][r];
#pragma unroll
  for (int j = 0; j < 16; ++j) s += x[j];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace o3d

using namespace o3d;

extern "C" {

int o3d_abi_version(void) { return O3D_ABI_VERSION; }
const char* o3d_last_error(void) { return last_error(); }

void o3d_default_opts(o3d_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->core = O3D_CORE_WL;
  o->results = O3D_VELANDGRAD;
  o->precision = O3D_PREC_COMPENSATED;
  o->self_term = O3D_SELF_INCLUDE;
  o->maxlev = -1;
}

int o3d_init(int device) {
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  clear_error();
  return init_locked(device);
}

int o3d_flops_per_interaction(int core, int results, int thick, int vs) {
  // CoreFunc.h flops_tv_grads / tp_grads / tv_nograds / tp_nograds, and the
  // Kernels.h pair counts 54 (0v_*g), 79 (0vs_*g), 23 (0v_*), 29 (0vs_*)
  static const int tvg[5] = {0, 9, 16, 14, 13}, tpg[5] = {0, 7, 14, 11, 10};
  static const int tvn[5] = {0, 7, 10, 12, 11}, tpn[5] = {0, 5, 8, 9, 8};
  if (core < 1 || core > 4) return -1;
  if (results == O3D_VELANDGRAD) return (vs ? 79 : 54) + (thick ? tvg[core] : tpg[core]);
  if (results == O3D_VELONLY) return (vs ? 29 : 23) + (thick ? tvn[core] : tpn[core]);
  return -1;
}

int o3d_points_affect_points(const o3d_points* src, o3d_points* targ, const o3d_opts* o, o3d_stats* stats) {
  clear_error();
  int rc;
  if ((rc = check_opts(o)) != O3D_OK) return rc;
  const bool grad = o->results == O3D_VELANDGRAD;
  if ((rc = check_points(src, true, false, grad, "src")) != O3D_OK) return rc;
  if ((rc = check_points(targ, false, true, grad, "targ")) != O3D_OK) return rc;
  if ((rc = check_self(src, targ, o)) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  if (o->precision == PREC_FP64) rc = src->n && targ->n ? p2p_host_fp64_from_f32(src, targ, o) : O3D_OK;
  else rc = p2p_host_impl<float>(src, targ, o, stats);
  if (rc != O3D_OK) return rc;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(stats, secs, flops_of(o->core, grad, targ->r != nullptr, targ->n, src->n),
             (double)targ->n * (double)src->n);
  return O3D_OK;
}

int o3d_points_affect_points_dev(const o3d_points* src, o3d_points* targ, const o3d_opts* o, o3d_stats* stats) {
  clear_error();
  int rc;
  if ((rc = check_p2p_args(src, targ, o)) != O3D_OK) return rc;
  const bool grad = o->results == O3D_VELANDGRAD;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  if (o->precision == PREC_FP64)
    return fail(O3D_EINVAL, "FP64 on device float arrays: use o3d_points_affect_points64_dev with double arrays");
  rc = p2p_dev(src, targ, o, (hipStream_t)o->stream, nullptr);
  if (rc != O3D_OK) return rc;
  fill_stats(stats, 0.0, flops_of(o->core, grad, targ->r != nullptr, targ->n, src->n), (double)targ->n * (double)src->n);
  return O3D_OK;
}

int o3d_factored_profile_dev(const o3d_points* src, o3d_points* targ, const o3d_opts* o, o3d_factored_profile* out) {
  clear_error();
  int rc;
  if (!out) return fail(O3D_EINVAL, "out is NULL");
  *out = o3d_factored_profile{};
  if ((rc = check_opts(o)) != O3D_OK) return rc;
  const bool grad = o->results == O3D_VELANDGRAD;
  if ((rc = check_points(src, true, false, grad, "src")) != O3D_OK) return rc;
  if ((rc = check_points(targ, false, true, grad, "targ")) != O3D_OK) return rc;
  if ((rc = check_self(src, targ, o)) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  if (src->n == 0 || targ->n == 0) return O3D_OK;
  return p2p_factored(src, targ, o, (hipStream_t)o->stream, out);
}

int o3d_points_affect_points64(const o3d_points64* src, o3d_points64* targ, const o3d_opts* o, o3d_stats* stats) {
  clear_error();
  int rc;
  if ((rc = check_opts(o)) != O3D_OK) return rc;
  const bool grad = o->results == O3D_VELANDGRAD;
  if ((rc = check_points(src, true, false, grad, "src")) != O3D_OK) return rc;
  if ((rc = check_points(targ, false, true, grad, "targ")) != O3D_OK) return rc;
  if ((rc = check_self(src, targ, o)) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  rc = p2p_host_impl<double>(src, targ, o, stats);
  if (rc != O3D_OK) return rc;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(stats, secs, flops_of(o->core, grad, targ->r != nullptr, targ->n, src->n), (double)targ->n * (double)src->n);
  return O3D_OK;
}

int o3d_points_affect_points64_dev(const o3d_points64* src, o3d_points64* targ, const o3d_opts* o, o3d_stats* stats) {
  clear_error();
  int rc;
  if ((rc = check_opts(o)) != O3D_OK) return rc;
  const bool grad = o->results == O3D_VELANDGRAD;
  if ((rc = check_points(src, true, false, grad, "src")) != O3D_OK) return rc;
  if ((rc = check_points(targ, false, true, grad, "targ")) != O3D_OK) return rc;
  if ((rc = check_self(src, targ, o)) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  rc = p2p64_dev(src, targ, o, (hipStream_t)o->stream);
  if (rc != O3D_OK) return rc;
  fill_stats(stats, 0.0, flops_of(o->core, grad, targ->r != nullptr, targ->n, src->n), (double)targ->n * (double)src->n);
  return O3D_OK;
}

int o3d_set_external_core(int core) {
  if (core < 1 || core > 4) return fail(O3D_EINVAL, "unknown core %d", core);
  set_ext_core(core);
  return O3D_OK;
}

float external_vel_solver_f_(int* ns, const float* sx, const float* sy, const float* sz,
                             const float* ssx, const float* ssy, const float* ssz, const float* sr,
                             int* nt, const float* tx, const float* ty, const float* tz,
                             float* tu, float* tv, float* tw, float* tux, float* tvx, float* twx,
                             float* tuy, float* tvy, float* twy, float* tuz, float* tvz, float* twz) {
  if (!ns || !nt) { fail(O3D_EINVAL, "counts are NULL"); return -1.0f; }
  o3d_points s{}, t{};
  s.n = *ns; s.x[0] = sx; s.x[1] = sy; s.x[2] = sz; s.s[0] = ssx; s.s[1] = ssy; s.s[2] = ssz; s.r = sr;
  t.n = *nt; t.x[0] = tx; t.x[1] = ty; t.x[2] = tz; t.r = nullptr;
  t.u[0] = tu; t.u[1] = tv; t.u[2] = tw;
  float* g[9] = {tux, tvx, twx, tuy, tvy, twy, tuz, tvz, twz};
  for (int k = 0; k < 9; ++k) t.ug[k] = g[k];
  o3d_opts o;
  o3d_default_opts(&o);
  o.core = ext_core();
  o3d_stats st{};
  if (o3d_points_affect_points(&s, &t, &o, &st) != O3D_OK) return -1.0f;
  return (float)st.flops;
}

float external_vel_solver_d_(int* ns, const double* sx, const double* sy, const double* sz,
                             const double* ssx, const double* ssy, const double* ssz, const double* sr,
                             int* nt, const double* tx, const double* ty, const double* tz,
                             double* tu, double* tv, double* tw, double* tux, double* tvx, double* twx,
                             double* tuy, double* tvy, double* twy, double* tuz, double* tvz, double* twz) {
  if (!ns || !nt) { fail(O3D_EINVAL, "counts are NULL"); return -1.0f; }
  o3d_points64 s{}, t{};
  s.n = *ns; s.x[0] = sx; s.x[1] = sy; s.x[2] = sz; s.s[0] = ssx; s.s[1] = ssy; s.s[2] = ssz; s.r = sr;
  t.n = *nt; t.x[0] = tx; t.x[1] = ty; t.x[2] = tz; t.r = nullptr;
  t.u[0] = tu; t.u[1] = tv; t.u[2] = tw;
  double* g[9] = {tux, tvx, twx, tuy, tvy, twy, tuz, tvz, twz};
  for (int k = 0; k < 9; ++k) t.ug[k] = g[k];
  o3d_opts o;
  o3d_default_opts(&o);
  o.core = ext_core();
  o.precision = O3D_PREC_FP64;
  o3d_stats st{};
  if (o3d_points_affect_points64(&s, &t, &o, &st) != O3D_OK) return -1.0f;
  return (float)st.flops;
}

// ---- device plumbing --------------------------------------------------------------
int o3d_dev_malloc(void** p, size_t bytes) {
  clear_error();
  int rc = ensure_init();
  if (rc) return rc;
  if (!p) return fail(O3D_EINVAL, "NULL out pointer");
  if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) return fail(O3D_ENOMEM, "hipMalloc(%zu) failed", bytes);
  return O3D_OK;
}
int o3d_dev_free(void* p) { HIPCHK(hipFree(p)); return O3D_OK; }
int o3d_memcpy_htod(void* d, const void* s, size_t b) { HIPCHK(hipMemcpy(d, s, b, hipMemcpyHostToDevice)); return O3D_OK; }
int o3d_memcpy_dtoh(void* d, const void* s, size_t b) { HIPCHK(hipMemcpy(d, s, b, hipMemcpyDeviceToHost)); return O3D_OK; }
int o3d_memset_dev(void* d, int v, size_t b) { HIPCHK(hipMemset(d, v, b)); return O3D_OK; }
int o3d_device_synchronize(void) { HIPCHK(hipDeviceSynchronize()); return O3D_OK; }
int o3d_stream_create(void** s) {
  int rc = ensure_init();
  if (rc) return rc;
  HIPCHK(hipStreamCreateWithFlags((hipStream_t*)s, hipStreamNonBlocking));
  return O3D_OK;
}
int o3d_stream_destroy(void* s) {
  // the stream's device workspace goes with it (a later stream may reuse the handle)
  release_stream_workspace((hipStream_t)s);
  HIPCHK(hipStreamDestroy((hipStream_t)s));
  return O3D_OK;
}
int o3d_stream_synchronize(void* s) { HIPCHK(hipStreamSynchronize((hipStream_t)s)); return O3D_OK; }
int o3d_event_create(void** e) {
  int rc = ensure_init();
  if (rc) return rc;
  HIPCHK(hipEventCreate((hipEvent_t*)e));
  return O3D_OK;
}
int o3d_event_destroy(void* e) { HIPCHK(hipEventDestroy((hipEvent_t)e)); return O3D_OK; }
int o3d_event_record(void* e, void* s) { HIPCHK(hipEventRecord((hipEvent_t)e, (hipStream_t)s)); return O3D_OK; }
int o3d_event_elapsed_ms(float* ms, void* a, void* b) {
  HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  return O3D_OK;
}
int o3d_copy_probe(void* dst, const void* src, size_t bytes, void* stream) {
  clear_error();
  if (!dst || !src || bytes % 16) return fail(O3D_EINVAL, "copy probe needs device buffers and bytes %% 16 == 0");
  int rc = ensure_init();
  if (rc) return rc;
  const int64_t n = (int64_t)(bytes / 16);
  const int blocks = (int)std::min<int64_t>((n + BLOCK - 1) / BLOCK, 8192);
  if (n > 0)
    hipLaunchKernelGGL(k_copy_probe, dim3(std::max(blocks, 1)), dim3(BLOCK), 0, (hipStream_t)stream, (float4*)dst,
                       (const float4*)src, n);
  HIPCHK(hipGetLastError());
  return O3D_OK;
}

int o3d_valu_probe(int32_t op, float* ms, double* wave_instructions) {
  clear_error();
  if (op < 0 || op > 8 || !ms || !wave_instructions) return fail(O3D_EINVAL, "valu probe: op 0..8 and outputs");
  int rc = ensure_init();
  if (rc) return rc;
  // ops 0-4: 8 workgroups of 256 lanes per CU; ops 5-7: one workgroup of 256 lanes
  // per CU (one wave per SIMD); op 8: one of 512 lanes per CU (two waves per SIMD)
  const int blocks = op < 5 ? 8 * num_cus() : num_cus(), iters = op == 3 ? 256 : op < 5 ? 1024 : 8192;
  const int threads = op == 8 ? 512 : BLOCK;
  std::lock_guard<std::mutex> lk(pool_mutex());
  void* buf = nullptr;
  if ((rc = pool_buffer(4, sizeof(float) * (size_t)blocks * BLOCK, &buf)) != O3D_OK) return rc;
  hipStream_t st = lib_stream();
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (int rep = 0; rep < 2; ++rep) {   // the first launch warms up
    HIPCHK(hipEventRecord(e0, st));
    if (op == 0) hipLaunchKernelGGL(k_valu_probe<0>, dim3(blocks), dim3(BLOCK), 0, st, (float*)buf, iters, 1e-7f);
    if (op == 1) hipLaunchKernelGGL(k_valu_probe<1>, dim3(blocks), dim3(BLOCK), 0, st, (float*)buf, iters, 0.f);
    if (op == 2) hipLaunchKernelGGL(k_valu_probe<2>, dim3(blocks), dim3(BLOCK), 0, st, (float*)buf, iters, 0.f);
    if (op == 3) hipLaunchKernelGGL(k_valu_probe<3>, dim3(blocks), dim3(BLOCK), 0, st, (float*)buf, iters, 0.f);
    if (op == 4) hipLaunchKernelGGL(k_valu_probe<4>, dim3(blocks), dim3(BLOCK), 0, st, (float*)buf, iters, 0.f);
    if (op == 5) hipLaunchKernelGGL(k_mix_probe<0>, dim3(blocks), dim3(threads), 0, st, (float*)buf, iters);
    if (op == 6) hipLaunchKernelGGL(k_mix_probe<1>, dim3(blocks), dim3(threads), 0, st, (float*)buf, iters);
    if (op == 7) hipLaunchKernelGGL(k_mix_probe<2>, dim3(blocks), dim3(threads), 0, st, (float*)buf, iters);
    if (op == 8) hipLaunchKernelGGL(k_mix_probe<3>, dim3(blocks), dim3(threads), 0, st, (float*)buf, iters);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
  }
  HIPCHK(hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // ops 0-4: the probed instruction; 5: the FMAs; 6-8: the MFMAs (op 7 also issues
  // 16 FMAs per MFMA in the same waves, op 8 as many in the other waves)
  const double waves = (double)blocks * (op < 5 ? BLOCK / 64 : 4);
  *wave_instructions = waves * (double)iters * (op < 5 ? 64.0 : op == 5 ? 64.0 : 4.0);
  return O3D_OK;
}

int o3d_device_info(char* name, int len, int* cus, int* clk, size_t* hbm) {
  clear_error();
  int rc = ensure_init();
  if (rc) return rc;
  if (name && len > 0) {
    std::snprintf(name, (size_t)len, "%s %s", g_ctx.prop.name, g_ctx.prop.gcnArchName);
  }
  if (cus) *cus = g_ctx.prop.multiProcessorCount;
  if (clk) *clk = g_ctx.prop.clockRate;
  if (hbm) *hbm = g_ctx.prop.totalGlobalMem;
  return O3D_OK;
}

}  // extern "C"
