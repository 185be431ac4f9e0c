// batch.hip -- many independent collections in one persistent launch, its HIP
// graph, the per-instance-launch baseline and the (tile, tpt) autotuners.
//
// Work items are (instance, chunk of BLOCK*TPT targets).  They are ordered by
// instance size, largest first (longest-processing-time first), and drained by a
// persistent grid: each workgroup takes the next item from a device counter.  An
// item runs the same p2p_chunk as the single-collection kernel, so a target's
// sums do not depend on the launch that computed it.
//
// Instances large enough for the factored far field (method AUTO: 32,768 particles
// and more, where a single call takes it from 65,536; or all of them with method
// FACTORED) leave the queue and
// run that pipeline on FF_LANES (8) streams of the batch (longest first, each to the
// least-loaded lane), each lane with scratch of its own.  The lanes fork from the
// caller's stream after the zero fill and join it at the end, so they overlap
// one another and the persistent kernel: an instance of 65,536 particles has
// only 256 far-field workgroups, half the GPU's slots.  Every step of a run is a
// kernel launch or an event (no memory fills, no allocation), so the whole run,
// lanes included, is captured into one HIP graph.
#include "o3d_api_internal.hpp"
#include "p2p_kernels.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <vector>

namespace o3d {

struct BatchInst {
  int64_t off;
  int64_t n;
  int32_t core;
  int32_t prec;
  int32_t self_excl;
  int32_t pad;
};
struct BatchItem {
  int32_t inst;
  int32_t chunk;
};
struct BatchArgs {
  SrcView s;            // shared SoA, offset by the instance
  TgtView t;
  const BatchInst* inst;
  const BatchItem* items;
  int32_t nitems;
  int32_t tile;
  unsigned int* counter;
};

template <int CORE, int PREC, int TPT>
__device__ __forceinline__ void run_item(const BatchArgs& a, const BatchInst& I, int chunk, float4* sh) {
  SrcView s{a.s.x + I.off, a.s.y + I.off, a.s.z + I.off, a.s.r + I.off,
            a.s.ax + I.off, a.s.ay + I.off, a.s.az + I.off};
  TgtView t;
  t.x = a.t.x + I.off; t.y = a.t.y + I.off; t.z = a.t.z + I.off; t.r = a.t.r + I.off;
  t.u = a.t.u + I.off; t.v = a.t.v + I.off; t.w = a.t.w + I.off;
#pragma unroll
  for (int k = 0; k < 9; ++k) t.g[k] = a.t.g[k] + I.off;
  p2p_chunk<CORE, true, true, PREC, TPT>(s, t, I.n, I.n, (int64_t)chunk * (BLOCK * TPT), a.tile, I.self_excl != 0, sh);
}

template <int TPT>
__global__ __launch_bounds__(BLOCK) void k_batch(BatchArgs a) {
  extern __shared__ float4 sh[];
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.counter, 1u);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= a.nitems) break;   // every workgroup reaches this once the queue is empty
    const BatchItem w = a.items[it];
    const BatchInst I = a.inst[w.inst];
    const int key = I.core * 2 + (I.prec == PREC_COMP ? 1 : 0);
    switch (key) {
      case CORE_RM * 2: run_item<CORE_RM, PREC_FP32, TPT>(a, I, w.chunk, sh); break;
      case CORE_RM * 2 + 1: run_item<CORE_RM, PREC_COMP, TPT>(a, I, w.chunk, sh); break;
      case CORE_WL * 2: run_item<CORE_WL, PREC_FP32, TPT>(a, I, w.chunk, sh); break;
      case CORE_WL * 2 + 1: run_item<CORE_WL, PREC_COMP, TPT>(a, I, w.chunk, sh); break;
      case CORE_EXP * 2: run_item<CORE_EXP, PREC_FP32, TPT>(a, I, w.chunk, sh); break;
      case CORE_EXP * 2 + 1: run_item<CORE_EXP, PREC_COMP, TPT>(a, I, w.chunk, sh); break;
      case CORE_V2 * 2: run_item<CORE_V2, PREC_FP32, TPT>(a, I, w.chunk, sh); break;
      default: run_item<CORE_V2, PREC_COMP, TPT>(a, I, w.chunk, sh); break;
    }
  }
}

__global__ void k_reset_counter(unsigned int* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0u;
}

// zero the 12 output rows over [0, n) (a fill kernel, so it can live in a graph)
__global__ void k_zero_outputs(TgtView t, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    t.u[i] = 0.f; t.v[i] = 0.f; t.w[i] = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) t.g[k][i] = 0.f;
  }
}

}  // namespace o3d

using namespace o3d;

struct o3d_batch {
  std::vector<BatchInst> inst;
  int64_t total_n = 0;
  int tile = 256;
  int tpt = 1;
  int grid = 0;            // persistent workgroups for the current tpt
  BatchInst* d_inst = nullptr;
  BatchItem* d_items[3] = {nullptr, nullptr, nullptr};  // per tpt 1, 2, 4
  int32_t nitems[3] = {0, 0, 0};
  unsigned int* d_counter = nullptr;
  std::vector<int> ff;     // instances evaluated by the factored method, largest first
  struct Lane {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    void* ws = nullptr;      // reused instance after instance (stream order)
    size_t bytes = 0;
    std::vector<int> inst;
  };
  std::vector<Lane> lanes; // lane 0 holds the largest instance, so its scratch fits any
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  o3d_points cap_all{};    // the arrays the captured graph is bound to
};

namespace {

constexpr size_t FF_LANES = 8;

int tpt_slot(int tpt) { return tpt == 1 ? 0 : (tpt == 2 ? 1 : 2); }

hipError_t batch_launch_kernel(o3d_batch* b, const BatchArgs& a, hipStream_t st) {
  const size_t lds = (size_t)b->tile * 2 * sizeof(float4);
  const dim3 grid((unsigned)b->grid);
  if (b->tpt == 1) hipLaunchKernelGGL(k_batch<1>, grid, dim3(BLOCK), lds, st, a);
  else if (b->tpt == 2) hipLaunchKernelGGL(k_batch<2>, grid, dim3(BLOCK), lds, st, a);
  else hipLaunchKernelGGL(k_batch<4>, grid, dim3(BLOCK), lds, st, a);
  return hipGetLastError();
}

int persistent_grid(int tpt, int tile) {
  int per_cu = 0;
  const size_t lds = (size_t)tile * 2 * sizeof(float4);
  hipError_t e;
  if (tpt == 1) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch<1>, BLOCK, lds);
  else if (tpt == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch<2>, BLOCK, lds);
  else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch<4>, BLOCK, lds);
  if (e != hipSuccess || per_cu <= 0) per_cu = 2;
  return per_cu * num_cus();
}

int check_all(const o3d_points* all, int64_t total) {
  if (!all) return fail(O3D_EINVAL, "all is NULL");
  if (all->n < total) return fail(O3D_EINVAL, "shared arrays hold %lld particles, instances need %lld",
                                  (long long)all->n, (long long)total);
  for (int d = 0; d < 3; ++d)
    if (!all->x[d] || !all->s[d] || !all->u[d]) return fail(O3D_EINVAL, "shared array component %d is NULL", d);
  if (!all->r) return fail(O3D_EINVAL, "shared radii are NULL");
  for (int k = 0; k < 9; ++k)
    if (!all->ug[k]) return fail(O3D_EINVAL, "shared gradient component %d is NULL", k);
  return O3D_OK;
}

BatchArgs args_of(o3d_batch* b, const o3d_points* all) {
  BatchArgs a{};
  a.s = SrcView{all->x[0], all->x[1], all->x[2], all->r, all->s[0], all->s[1], all->s[2]};
  a.t.x = all->x[0]; a.t.y = all->x[1]; a.t.z = all->x[2]; a.t.r = all->r;
  a.t.u = all->u[0]; a.t.v = all->u[1]; a.t.w = all->u[2];
  for (int k = 0; k < 9; ++k) a.t.g[k] = all->ug[k];
  a.inst = b->d_inst;
  const int sl = tpt_slot(b->tpt);
  a.items = b->d_items[sl];
  a.nitems = b->nitems[sl];
  a.tile = b->tile;
  a.counter = b->d_counter;
  return a;
}

// instance I's slice of the shared arrays, as one collection
o3d_points slice(const o3d_points* all, const BatchInst& I) {
  o3d_points p = *all;
  p.n = I.n;
  for (int d = 0; d < 3; ++d) {
    p.x[d] = all->x[d] + I.off;
    p.s[d] = all->s[d] + I.off;
    p.u[d] = all->u[d] + I.off;
  }
  p.r = all->r + I.off;
  for (int k = 0; k < 9; ++k) p.ug[k] = all->ug[k] + I.off;
  return p;
}

o3d_opts factored_opts(const BatchInst& I) {
  o3d_opts o{};
  o.core = I.core;
  o.results = O3D_VELANDGRAD;
  o.precision = I.prec;
  o.self_term = I.self_excl;
  o.method = O3D_METHOD_FACTORED;
  return o;
}

void zero_outputs(o3d_batch* b, const TgtView& t, hipStream_t st) {
  const int zb = (int)std::min<int64_t>((b->total_n + BLOCK - 1) / BLOCK, 4096);
  if (b->total_n > 0) hipLaunchKernelGGL(k_zero_outputs, dim3(std::max(zb, 1)), dim3(BLOCK), 0, st, t, b->total_n);
}

// zero fill, queue reset and the persistent kernel: kernels only, so it can be captured
int enqueue_queue(o3d_batch* b, const o3d_points* all, hipStream_t st) {
  BatchArgs a = args_of(b, all);
  zero_outputs(b, a.t, st);
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, b->d_counter);
  HIPCHK(hipGetLastError());
  if (a.nitems > 0) HIPCHK(batch_launch_kernel(b, a, st));
  return O3D_OK;
}

int factored_one(o3d_batch* b, const o3d_points* all, int i, hipStream_t st, void* ws, size_t bytes) {
  const BatchInst& I = b->inst[i];
  o3d_points p = slice(all, I);
  const o3d_opts o = factored_opts(I);
  return p2p_factored_ws(&p, &p, &o, st, ws, bytes);
}

// the factored instances on the lanes, accumulating into outputs zeroed on st
// before this call; st waits for them at join_factored
int fork_factored(o3d_batch* b, const o3d_points* all, hipStream_t st) {
  if (b->ff.empty()) return O3D_OK;
  HIPCHK(hipEventRecord(b->fork, st));
  for (auto& L : b->lanes) {
    if (L.inst.empty()) continue;
    HIPCHK(hipStreamWaitEvent(L.st, b->fork, 0));
    for (int i : L.inst) {
      const int rc = factored_one(b, all, i, L.st, L.ws, L.bytes);
      if (rc != O3D_OK) return rc;
    }
    HIPCHK(hipEventRecord(L.done, L.st));
  }
  return O3D_OK;
}

int join_factored(o3d_batch* b, hipStream_t st) {
  for (auto& L : b->lanes)
    if (!L.inst.empty()) HIPCHK(hipStreamWaitEvent(st, L.done, 0));
  return O3D_OK;
}

// the factored instances one after another on st (the per-instance baseline)
int factored_serial(o3d_batch* b, const o3d_points* all, hipStream_t st) {
  for (int i : b->ff) {
    const int rc = factored_one(b, all, i, st, b->lanes[0].ws, b->lanes[0].bytes);
    if (rc != O3D_OK) return rc;
  }
  return O3D_OK;
}

int enqueue_run(o3d_batch* b, const o3d_points* all, hipStream_t st) {
  BatchArgs a = args_of(b, all);
  zero_outputs(b, a.t, st);
  int rc = fork_factored(b, all, st);
  if (rc != O3D_OK) return rc;
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, b->d_counter);
  HIPCHK(hipGetLastError());
  if (a.nitems > 0) HIPCHK(batch_launch_kernel(b, a, st));
  return join_factored(b, st);
}

void drop_graph(o3d_batch* b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int set_config(o3d_batch* b, int tile, int tpt) {
  if (tile < 64 || tile > 1024 || tile % 64) return fail(O3D_EINVAL, "tile %d", tile);
  if (tpt != 1 && tpt != 2 && tpt != 4) return fail(O3D_EINVAL, "tpt %d", tpt);
  if (tile != b->tile || tpt != b->tpt) drop_graph(b);
  b->tile = tile;
  b->tpt = tpt;
  b->grid = persistent_grid(tpt, tile);
  return O3D_OK;
}

// ---- single-collection tuning table ------------------------------------------------
struct TuneKey {
  int core, results, prec, thick, bucket;
  bool operator<(const TuneKey& o) const {
    return std::tie(core, results, prec, thick, bucket) < std::tie(o.core, o.results, o.prec, o.thick, o.bucket);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, Tuning> g_tune;

int bucket_of(int64_t nt) { return nt <= 1 ? 0 : (int)std::floor(std::log2((double)nt)); }

}  // namespace

namespace o3d {
// consulted by choose_config when the caller leaves tile and tpt at 0
bool tuned_lookup(int core, int results, int prec, bool thick, int64_t nt, Tuning* out) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(TuneKey{core, results, prec, thick ? 1 : 0, bucket_of(nt)});
  if (it == g_tune.end()) return false;
  *out = it->second;
  return true;
}
}  // namespace o3d

extern "C" {

int o3d_batch_create(o3d_batch** out, int32_t ninst, const int64_t* off, const int64_t* n, const int32_t* cores,
                     const int32_t* precs, const int32_t* selfs, const o3d_opts* opts) {
  clear_error();
  int rc;
  if (!out) return fail(O3D_EINVAL, "NULL handle pointer");
  *out = nullptr;
  if ((rc = check_opts(opts)) != O3D_OK) return rc;
  if (ninst < 0 || (ninst > 0 && (!off || !n))) return fail(O3D_EINVAL, "bad instance table");
  if (opts->results != O3D_VELANDGRAD) return fail(O3D_EINVAL, "batches evaluate velocity and gradient");
  if ((rc = ensure_init()) != O3D_OK) return rc;
  auto* b = new o3d_batch();
  for (int i = 0; i < ninst; ++i) {
    BatchInst I{};
    I.off = off[i];
    I.n = n[i];
    I.core = cores ? cores[i] : opts->core;
    I.prec = precs ? precs[i] : opts->precision;
    I.self_excl = selfs ? selfs[i] : opts->self_term;
    if (I.off < 0 || I.n < 0 || I.core < 1 || I.core > 4 || (I.prec != PREC_FP32 && I.prec != PREC_COMP) ||
        (I.self_excl != 0 && I.self_excl != 1)) {
      delete b;
      return fail(O3D_EINVAL, "instance %d: bad offset/count/core/precision/self_term", i);
    }
    if ((I.n + BLOCK - 1) / BLOCK > (int64_t)INT32_MAX) {
      delete b;
      return fail(O3D_EINVAL, "instance %d too large", i);
    }
    const bool ffok = ff_supported(I.core, true, I.prec) && I.n <= INT32_MAX;
    if (opts->method == O3D_METHOD_FACTORED && !ffok) {
      delete b;
      return fail(O3D_EINVAL, "instance %d: the factored method needs the RM or EXP core and FP32 or compensated "
                              "precision", i);
    }
    b->total_n = std::max(b->total_n, I.off + I.n);
    b->inst.push_back(I);
  }
  // LPT order: largest instances first
  std::vector<int> order(ninst);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->inst[x].n > b->inst[y].n; });
  std::vector<char> factored(ninst, 0);
  for (int i : order) {
    const BatchInst& I = b->inst[i];
    const bool ffok = ff_supported(I.core, true, I.prec) && I.n <= INT32_MAX;
    if (I.n > 0 && ffok &&
        (opts->method == O3D_METHOD_FACTORED || (opts->method == O3D_METHOD_AUTO && I.n >= LANES_FF_AUTO_MIN_N))) {
      factored[i] = 1;
      b->ff.push_back(i);
    }
  }
  const int tpts[3] = {1, 2, 4};
  std::vector<BatchItem> items[3];
  for (int s = 0; s < 3; ++s) {
    const int64_t per = (int64_t)BLOCK * tpts[s];
    for (int i : order)
      if (!factored[i])
        for (int64_t c = 0; c * per < b->inst[i].n; ++c) items[s].push_back(BatchItem{i, (int32_t)c});
    b->nitems[s] = (int32_t)items[s].size();
  }
  auto cleanup = [&](int code) { o3d_batch_destroy(b); return code; };
  if (!b->ff.empty()) {
    // longest first, each to the least-loaded lane (load = pairs)
    b->lanes.resize(std::min<size_t>(FF_LANES, b->ff.size()));
    std::vector<double> load(b->lanes.size(), 0.0);
    for (int i : b->ff) {
      const size_t k = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
      load[k] += (double)b->inst[i].n * (double)b->inst[i].n;
      b->lanes[k].inst.push_back(i);
      const size_t need = ff_workspace_bytes(b->inst[i].n, b->inst[i].n, true, 0);
      if (need == 0) return cleanup(fail(O3D_EDEVICE, "instance %d: factored scratch query failed", i));
      b->lanes[k].bytes = std::max(b->lanes[k].bytes, need);
    }
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
      return cleanup(fail(O3D_EDEVICE, "event"));
    for (auto& L : b->lanes) {
      if (hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(O3D_EDEVICE, "factored lane stream"));
      if (hipMalloc(&L.ws, L.bytes) != hipSuccess)
        return cleanup(fail(O3D_ENOMEM, "factored scratch of %zu bytes", L.bytes));
    }
  }
  if (hipMalloc(&b->d_counter, sizeof(unsigned int)) != hipSuccess) return cleanup(fail(O3D_ENOMEM, "counter"));
  if (ninst > 0) {
    if (hipMalloc(&b->d_inst, sizeof(BatchInst) * ninst) != hipSuccess) return cleanup(fail(O3D_ENOMEM, "instances"));
    if (hipMemcpy(b->d_inst, b->inst.data(), sizeof(BatchInst) * ninst, hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(O3D_EDEVICE, "instance upload"));
  }
  for (int s = 0; s < 3; ++s) {
    if (items[s].empty()) continue;
    if (hipMalloc(&b->d_items[s], sizeof(BatchItem) * items[s].size()) != hipSuccess)
      return cleanup(fail(O3D_ENOMEM, "work items"));
    if (hipMemcpy(b->d_items[s], items[s].data(), sizeof(BatchItem) * items[s].size(), hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(O3D_EDEVICE, "work item upload"));
  }
  const int tile = opts->tile > 0 ? opts->tile : 256;
  const int tpt = opts->tpt > 0 ? opts->tpt : 1;
  b->tile = 0;
  if ((rc = set_config(b, tile, tpt)) != O3D_OK) return cleanup(rc);
  *out = b;
  return O3D_OK;
}

int o3d_batch_destroy(o3d_batch* b) {
  if (!b) return O3D_OK;
  drop_graph(b);
  if (b->d_inst) hipFree(b->d_inst);
  for (auto* p : b->d_items) if (p) hipFree(p);
  if (b->d_counter) hipFree(b->d_counter);
  for (auto& L : b->lanes) {
    if (L.st) {
      hipStreamSynchronize(L.st);
      hipStreamDestroy(L.st);
    }
    if (L.done) hipEventDestroy(L.done);
    if (L.ws) hipFree(L.ws);
  }
  if (b->fork) hipEventDestroy(b->fork);
  delete b;
  return O3D_OK;
}

int o3d_batch_config(const o3d_batch* b, int32_t* tile, int32_t* tpt, int32_t* grid, int64_t* items) {
  if (!b) return fail(O3D_EINVAL, "NULL batch");
  if (tile) *tile = b->tile;
  if (tpt) *tpt = b->tpt;
  if (grid) *grid = b->grid;
  if (items) *items = b->nitems[tpt_slot(b->tpt)];
  return O3D_OK;
}

int o3d_batch_factored(const o3d_batch* b, int32_t* ninst, int64_t* scratch_bytes) {
  if (!b) return fail(O3D_EINVAL, "NULL batch");
  if (ninst) *ninst = (int32_t)b->ff.size();
  size_t bytes = 0;
  for (const auto& L : b->lanes) bytes += L.bytes;
  if (scratch_bytes) *scratch_bytes = (int64_t)bytes;
  return O3D_OK;
}

int o3d_batch_run(o3d_batch* b, const o3d_points* all, void* stream) {
  clear_error();
  if (!b) return fail(O3D_EINVAL, "NULL batch");
  int rc = check_all(all, b->total_n);
  if (rc) return rc;
  return enqueue_run(b, all, (hipStream_t)stream);
}

int o3d_batch_run_separate(o3d_batch* b, const o3d_points* all, void* stream) {
  clear_error();
  if (!b) return fail(O3D_EINVAL, "NULL batch");
  int rc = check_all(all, b->total_n);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  BatchArgs a = args_of(b, all);
  zero_outputs(b, a.t, st);
  // the factored instances one after another, each an ordinary single-collection call
  if ((rc = factored_serial(b, all, st)) != O3D_OK) return rc;
  std::vector<char> factored(b->inst.size(), 0);
  for (int i : b->ff) factored[i] = 1;
  for (size_t i = 0; i < b->inst.size(); ++i) {
    const BatchInst& I = b->inst[i];
    if (I.n == 0 || factored[i]) continue;
    P2PArgs p{};
    p.s = SrcView{a.s.x + I.off, a.s.y + I.off, a.s.z + I.off, a.s.r + I.off, a.s.ax + I.off, a.s.ay + I.off, a.s.az + I.off};
    p.t.x = a.t.x + I.off; p.t.y = a.t.y + I.off; p.t.z = a.t.z + I.off; p.t.r = a.t.r + I.off;
    p.t.u = a.t.u + I.off; p.t.v = a.t.v + I.off; p.t.w = a.t.w + I.off;
    for (int k = 0; k < 9; ++k) p.t.g[k] = a.t.g[k] + I.off;
    p.ns = p.nt = I.n;
    p.tile = b->tile;
    p.self_excl = I.self_excl;
    hipError_t e = dispatch_p2p(I.core, p, true, true, I.prec, b->tpt, st);
    if (e != hipSuccess) return fail(O3D_EDEVICE, "per-instance launch failed: %s", hipGetErrorString(e));
  }
  return O3D_OK;
}

int o3d_batch_capture(o3d_batch* b, const o3d_points* all, void* stream) {
  clear_error();
  if (!b) return fail(O3D_EINVAL, "NULL batch");
  if (!stream) return fail(O3D_EINVAL, "graph capture needs a non-null stream");
  int rc = check_all(all, b->total_n);
  if (rc) return rc;
  drop_graph(b);
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  rc = enqueue_run(b, all, st);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (rc != O3D_OK) { if (g) hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return fail(O3D_EDEVICE, "end capture: %s", hipGetErrorString(e));
  b->graph = g;
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) { drop_graph(b); return fail(O3D_EDEVICE, "graph instantiate: %s", hipGetErrorString(e)); }
  b->cap_all = *all;
  return O3D_OK;
}

int o3d_batch_replay(o3d_batch* b, void* stream) {
  clear_error();
  if (!b || !b->exec) return fail(O3D_EINVAL, "no captured graph");
  HIPCHK(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return O3D_OK;
}

int o3d_batch_tune(o3d_batch* b, const o3d_points* all, void* stream, const int32_t* tiles, int32_t ntiles,
                   const int32_t* tpts, int32_t ntpts, int32_t reps, int32_t* best_tile, int32_t* best_tpt,
                   float* best_ms) {
  clear_error();
  if (!b) return fail(O3D_EINVAL, "NULL batch");
  int rc = check_all(all, b->total_n);
  if (rc) return rc;
  static const int32_t dt[3] = {128, 256, 512}, dp[3] = {1, 2, 4};
  if (!tiles || ntiles <= 0) { tiles = dt; ntiles = 3; }
  if (!tpts || ntpts <= 0) { tpts = dp; ntpts = 3; }
  if (reps <= 0) reps = 1;
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  float best = 1e30f;
  int bt = b->tile, bp = b->tpt;
  for (int i = 0; i < ntiles; ++i)
    for (int j = 0; j < ntpts; ++j) {
      if ((rc = set_config(b, tiles[i], tpts[j])) != O3D_OK) break;
      if ((rc = enqueue_queue(b, all, st)) != O3D_OK) break;   // warm-up
      HIPCHK(hipEventRecord(e0, st));
      for (int r = 0; r < reps && rc == O3D_OK; ++r) rc = enqueue_queue(b, all, st);
      HIPCHK(hipEventRecord(e1, st));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      ms /= reps;
      if (ms < best) { best = ms; bt = tiles[i]; bp = tpts[j]; }
    }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != O3D_OK) return rc;
  if ((rc = set_config(b, bt, bp)) != O3D_OK) return rc;
  // the candidates only change the queue part; complete the outputs once
  if ((rc = fork_factored(b, all, st)) != O3D_OK || (rc = join_factored(b, st)) != O3D_OK) return rc;
  if (best_tile) *best_tile = bt;
  if (best_tpt) *best_tpt = bp;
  if (best_ms) *best_ms = best;
  return O3D_OK;
}

int o3d_tune_p2p(const o3d_points* src, o3d_points* targ, const o3d_opts* opts, int32_t reps, int32_t* best_tile,
                 int32_t* best_tpt, float* best_ms) {
  clear_error();
  int rc;
  if ((rc = check_opts(opts)) != O3D_OK) return rc;
  if (opts->precision == PREC_FP64) return fail(O3D_EINVAL, "the FP64 kernel has one configuration");
  if (!src || !targ || src->n <= 0 || targ->n <= 0) return fail(O3D_EINVAL, "empty collections");
  if ((rc = ensure_init()) != O3D_OK) return rc;
  const bool grad = opts->results == O3D_VELANDGRAD;
  // the targets' outputs are left untouched: the candidates write into pool scratch
  std::lock_guard<std::mutex> lk(pool_mutex());
  void* scratch;
  if ((rc = pool_buffer(4, sizeof(float) * 12 * (size_t)targ->n, &scratch)) != O3D_OK) return rc;
  o3d_points t = *targ;
  float* f = (float*)scratch;
  for (int d = 0; d < 3; ++d) t.u[d] = f + d * targ->n;
  for (int k = 0; k < 9; ++k) t.ug[k] = grad ? f + (3 + k) * targ->n : nullptr;
  hipStream_t st = (hipStream_t)opts->stream;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  static const int tiles[3] = {128, 256, 512}, tpts[3] = {1, 2, 4};
  float best = 1e30f;
  Tuning bestc{256, 1};
  if (reps <= 0) reps = 1;
  for (int tile : tiles)
    for (int tpt : tpts) {
      Tuning c{tile, tpt};
      if ((rc = p2p_dev(src, &t, opts, st, &c)) != O3D_OK) break;
      HIPCHK(hipEventRecord(e0, st));
      for (int r = 0; r < reps && rc == O3D_OK; ++r) rc = p2p_dev(src, &t, opts, st, &c);
      HIPCHK(hipEventRecord(e1, st));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      ms /= reps;
      if (ms < best) { best = ms; bestc = c; }
    }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != O3D_OK) return rc;
  {
    std::lock_guard<std::mutex> lk2(g_tune_mu);
    g_tune[TuneKey{opts->core, opts->results, opts->precision, targ->r ? 1 : 0, bucket_of(targ->n)}] = bestc;
  }
  if (best_tile) *best_tile = bestc.tile;
  if (best_tpt) *best_tpt = bestc.tpt;
  if (best_ms) *best_ms = best;
  return O3D_OK;
}

int o3d_tune_save(const char* path) {
  if (!path) return fail(O3D_EINVAL, "NULL path");
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(O3D_EINVAL, "cannot write %s", path);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  std::fprintf(f, "# core results precision thick log2_targets tile tpt\n");
  for (auto& kv : g_tune)
    std::fprintf(f, "%d %d %d %d %d %d %d\n", kv.first.core, kv.first.results, kv.first.prec, kv.first.thick,
                 kv.first.bucket, kv.second.tile, kv.second.tpt);
  std::fclose(f);
  return O3D_OK;
}

int o3d_tune_load(const char* path) {
  if (!path) return fail(O3D_EINVAL, "NULL path");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(O3D_EINVAL, "cannot read %s", path);
  char line[256];
  int n = 0;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    TuneKey k{};
    Tuning t{};
    if (std::sscanf(line, "%d %d %d %d %d %d %d", &k.core, &k.results, &k.prec, &k.thick, &k.bucket, &t.tile, &t.tpt) == 7 &&
        t.tile >= 64 && t.tile <= 1024 && t.tile % 64 == 0 && (t.tpt == 1 || t.tpt == 2 || t.tpt == 4)) {
      g_tune[k] = t;
      ++n;
    }
  }
  std::fclose(f);
  return n > 0 ? O3D_OK : fail(O3D_EINVAL, "no valid entries in %s", path);
}

}  // extern "C"
