// server.hip -- device-resident simulations stepped by continuous batching.
//
// A request is one inviscid Ralston RK2 convection step with vortex stretching of
// a session (Convection.h:349-425):
//   stage 1: u1, ug1 = finalize(sum over the session at x0, s0)
//            x1 = x0 + (2/3) dt u1,  s1 = s0 + (2/3) dt (s0 . grad) u1
//   stage 2: u2, ug2 = finalize(sum over the interim state x1, s1)
//            x  = x0 + dt (u1/4 + 3 u2/4)
//            s  = s0 + dt (s0 . grad)(u1/4 + 3 u2/4),  elongation updated
// (ElementBase.h:187-304, Points.h:265-439, with the reference's float/double types).
//
// The host loop thread builds ticks.  A tick = one persistent evaluation launch over
// a budget of (request, 256*tpt-target chunk) items drawn from every admitted
// request, followed by one launch that applies the stage updates of the requests
// whose stage finished in that tick.  Requests are admitted whenever a slot and
// arena space are free, so new work joins ticks while earlier work is running.
// The evaluation of a chunk is p2p_chunk, the same code as the single-collection
// kernel, so a step's result does not depend on its companions.
//
// A step of a large session with the RM or EXP core (by default >= 32,768
// particles) skips the ticks: both of its stages run the factored far-field
// method, each followed by its stage update, on one of a few further streams
// ("lanes") that overlap the ticks and one another, each lane with scratch the
// server sizes for its largest such session.  Which way a
// session's steps go depends only on the session and the server's options, so a
// step's result still does not depend on its companions.
#include "o3d_api_internal.hpp"
#include "p2p_kernels.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace o3d {

struct ServeInst {
  SrcView s;
  TgtView t;
  int64_t n;
  int32_t core;
  int32_t prec;
};
struct ServeItem {
  int32_t inst;
  int32_t chunk;
};
struct MoveInst {
  int64_t n;
  int32_t stage;
  int32_t pad;
  double dt;
  double fs[3];
  float* x[3];
  float* s[3];
  float* elong;
  float* x1[3];
  float* s1[3];
  float* u1[3];
  float* g1[9];
  float* u2[3];
  float* g2[9];
};
struct ServeTables {      // what one tick reads, in device memory
  int32_t n_eval;
  int32_t n_move;
  unsigned int counter;
  int32_t pad;
};

template <int CORE, int PREC, int TPT>
__device__ __forceinline__ void serve_item(const ServeInst& I, int chunk, int tile, float4* sh) {
  p2p_chunk<CORE, true, true, PREC, TPT, false>(I.s, I.t, I.n, I.n, (int64_t)chunk * (BLOCK * TPT), tile, false, sh);
}

template <int TPT>
__global__ __launch_bounds__(BLOCK) void k_serve_eval(const ServeInst* inst, const ServeItem* items,
                                                      ServeTables* tab, int tile) {
  extern __shared__ float4 sh[];
  __shared__ int s_item;
  const int nitems = tab->n_eval;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(&tab->counter, 1u);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= nitems) break;
    const ServeItem w = items[it];
    const ServeInst I = inst[w.inst];
    const int key = I.core * 2 + (I.prec == PREC_COMP ? 1 : 0);
    switch (key) {
      case CORE_RM * 2: serve_item<CORE_RM, PREC_FP32, TPT>(I, w.chunk, tile, sh); break;
      case CORE_RM * 2 + 1: serve_item<CORE_RM, PREC_COMP, TPT>(I, w.chunk, tile, sh); break;
      case CORE_WL * 2: serve_item<CORE_WL, PREC_FP32, TPT>(I, w.chunk, tile, sh); break;
      case CORE_WL * 2 + 1: serve_item<CORE_WL, PREC_COMP, TPT>(I, w.chunk, tile, sh); break;
      case CORE_EXP * 2: serve_item<CORE_EXP, PREC_FP32, TPT>(I, w.chunk, tile, sh); break;
      case CORE_EXP * 2 + 1: serve_item<CORE_EXP, PREC_COMP, TPT>(I, w.chunk, tile, sh); break;
      case CORE_V2 * 2: serve_item<CORE_V2, PREC_FP32, TPT>(I, w.chunk, tile, sh); break;
      default: serve_item<CORE_V2, PREC_COMP, TPT>(I, w.chunk, tile, sh); break;
    }
  }
}

__global__ void k_serve_reset(ServeTables* tab) {
  if (threadIdx.x == 0 && blockIdx.x == 0) tab->counter = 0u;
}

// wdu_i = s0 ug[i] + s1 ug[3+i] + s2 ug[6+i] in float, as Points.h:310-312, 384-386
__device__ __forceinline__ float wdu_of(float s0, float s1, float s2, const float* g, int i) {
#pragma clang fp contract(off)
  return (s0 * g[i] + s1 * g[3 + i]) + s2 * g[6 + i];
}

// one particle of one request's stage update
__device__ void move_particle(const MoveInst& m, int64_t i) {
  const double factor = 0.25 / M_PI;
  const float ffac = (float)factor;
  if (m.stage == 1) {
    float g1[9];
    float u1[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      u1[d] = (float)(m.fs[d] + (double)m.u1[d][i] * factor);          // ElementBase.h:190
      m.u1[d][i] = u1[d];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      g1[k] = m.g1[k][i] * ffac;                                        // Points.h:270
      m.g1[k][i] = g1[k];
    }
    const double dt23 = 2.0 / 3.0 * m.dt;
    const float s0x = m.s[0][i], s0y = m.s[1][i], s0z = m.s[2][i];
#pragma unroll
    for (int d = 0; d < 3; ++d)                                         // ElementBase.h:260-262
      m.x1[d][i] = (float)((double)m.x[d][i] + ((double)(float)dt23 * 1.0) * (double)u1[d]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {                                       // Points.h:326-328
      const float w = wdu_of(s0x, s0y, s0z, g1, d);
      m.s1[d][i] = (float)((double)m.s[d][i] + dt23 * 1.0 * (double)w);
    }
  } else {
    float u2[3], g1[9], g2[9];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      u2[d] = (float)(m.fs[d] + (double)m.u2[d][i] * factor);
      m.u2[d][i] = u2[d];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      g1[k] = m.g1[k][i];                                               // finalised in stage 1
      g2[k] = m.g2[k][i] * ffac;
      m.g2[k][i] = g2[k];
    }
    const float s0x = m.s[0][i], s0y = m.s[1][i], s0z = m.s[2][i];
    // ElementBase.h:286-296: u = 0.25 u1 + 0.75 u2 (double, stored float); x += (S)dt * u
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float u = (float)(0.25 * (double)m.u1[d][i] + 0.75 * (double)u2[d]);
      float x = m.x[d][i];
      {
#pragma clang fp contract(off)
        x = x + (float)m.dt * u;
      }
      m.x[d][i] = x;
    }
    // Points.h:376-423
    float w[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float w1 = wdu_of(s0x, s0y, s0z, g1, d);
      const float w2 = wdu_of(s0x, s0y, s0z, g2, d);
      w[d] = (float)(0.25 * (double)w1 + 0.75 * (double)w2);
    }
    float c2, dot;
    {
#pragma clang fp contract(off)
      c2 = (s0x * s0x + s0y * s0y) + s0z * s0z;
      dot = (s0x * w[0] + s0y * w[1]) + s0z * w[2];
    }
    if (c2 > 0.0f) {
      float ef;
      {
#pragma clang fp contract(off)
        ef = ((float)m.dt * dot) / c2;
      }
      m.elong[i] = (float)((double)m.elong[i] * (1.0 + (double)ef));
    }
    m.s[0][i] = (float)((double)s0x + m.dt * (double)w[0]);
    m.s[1][i] = (float)((double)s0y + m.dt * (double)w[1]);
    m.s[2][i] = (float)((double)s0z + m.dt * (double)w[2]);
  }
}

// one request's stage update over all its particles (the factored steps)
__global__ __launch_bounds__(BLOCK) void k_serve_move_one(MoveInst m) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < m.n; i += stride) move_particle(m, i);
}

__global__ void k_serve_zero(float* p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.f;
}

// persistent: move items are (request, chunk of BLOCK particles)
__global__ __launch_bounds__(BLOCK) void k_serve_move(const MoveInst* mi, const ServeItem* items, const ServeTables* tab) {
  const int nitems = tab->n_move;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const ServeItem w = items[it];
    const MoveInst& m = mi[w.inst];
    const int64_t i = (int64_t)w.chunk * BLOCK + threadIdx.x;
    if (i < m.n) move_particle(m, i);
  }
}

}  // namespace o3d

using namespace o3d;

namespace {

using Clock = std::chrono::steady_clock;
constexpr size_t ALIGN = 256;
constexpr size_t STATE_FLOATS = 8;    // x3 s3 r elong
constexpr size_t WORK_FLOATS = 30;    // x1 3, s1 3, u1 3, g1 9, u2 3, g2 9

std::atomic<int> g_fault_count{0};

// first-fit arena bookkeeping over one device allocation
struct Arena {
  std::map<size_t, size_t> free_;   // offset -> size
  size_t size = 0, used = 0, peak = 0;
  void init(size_t n) { size = n; free_.clear(); free_[0] = n; used = peak = 0; }
  bool alloc(size_t bytes, size_t* off) {
    bytes = (bytes + ALIGN - 1) / ALIGN * ALIGN;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t noff = it->first + bytes;
        free_.erase(it);
        if (rest) free_[noff] = rest;
        used += bytes;
        peak = std::max(peak, used);
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + ALIGN - 1) / ALIGN * ALIGN;
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Session {
  int64_t id;
  int64_t n;
  int32_t core;
  size_t off, bytes;
  float* base;     // x[3] s[3] r elong, each n floats
  int pending = 0; // queued or running requests
  bool closing = false;
  float* x(int d) const { return base + d * n; }
  float* s(int d) const { return base + (3 + d) * n; }
  float* r() const { return base + 6 * n; }
  float* elong() const { return base + 7 * n; }
};

struct Request {
  int64_t id;
  int64_t sid;
  double dt;
  int32_t prec;
  double fs[3];
  double deadline;      // server clock; <= 0 none
  int32_t state = O3D_RQ_QUEUED;
  int32_t status = O3D_OK;
  double submit_s = 0, admit_s = 0, done_s = 0;
  double interactions = 0;
  // running
  int stage = 0;
  int64_t next_chunk = 0;   // next unscheduled chunk of the current stage
  int64_t nchunks = 0;
  size_t woff = 0, wbytes = 0;
  float* work = nullptr;
  bool cancel = false;
  bool finished_planning = false;
  bool factored = false;    // both stages by the factored method on the second stream
  bool ff_launched = false;
  float* w(int k, int64_t n) const { return work + (size_t)k * n; }   // row k of 30
};

struct Tick {
  hipEvent_t done = nullptr;
  std::vector<int64_t> completed;   // requests whose stage 2 ran in this tick
  int slot = 0;
};

struct FFJob {
  hipEvent_t done = nullptr;
  int64_t rid = 0;
  double work = 0;     // the step's interactions
};

// Factored steps run on SRV_FF_LANES streams of their own, each with scratch for
// the largest factored session; a step goes to the lane with the least work
// queued.  A session has at most one step admitted at a time, so its steps stay
// in submission order whichever lanes they take.  A 65,536-particle session gives
// the far-field kernel half the GPU's workgroup slots, so steps of several
// sessions must overlap to fill it.
constexpr int SRV_FF_LANES = 4;
struct FFLane {
  hipStream_t st = nullptr;
  void* ws = nullptr;
  size_t bytes = 0;
  std::deque<FFJob> inflight;
  double pending = 0;
};

}  // namespace

struct o3d_server {
  o3d_server_opts opts{};
  int tile = 256, tpt = 1, grid = 0, move_grid = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  char* arena_base = nullptr;
  Arena arena;
  std::mutex mu;
  std::condition_variable cv;      // wakes the loop
  std::condition_variable done_cv; // wakes waiters
  std::thread loop;
  bool stop = false;
  int64_t next_sid = 1, next_rid = 1;
  std::map<int64_t, Session> sessions;
  std::map<int64_t, Request> requests;
  std::deque<int64_t> queue;       // queued request ids in arrival order
  std::vector<int64_t> active;     // admitted, in admission order
  std::deque<Tick> inflight;
  // factored steps: their lanes (stream, scratch, steps enqueued on it)
  FFLane lanes[SRV_FF_LANES];
  int64_t n_factored = 0;
  bool ff_idle() const {
    for (const FFLane& l : lanes)
      if (!l.inflight.empty()) return false;
    return true;
  }
  Clock::time_point t0 = Clock::now();
  // per-tick device tables, double buffered
  size_t inst_cap = 0, item_cap = 0, move_cap = 0;
  ServeInst* d_inst[2] = {nullptr, nullptr};
  ServeItem* d_items[2] = {nullptr, nullptr};
  MoveInst* d_minst[2] = {nullptr, nullptr};
  ServeItem* d_mitems[2] = {nullptr, nullptr};
  ServeTables* d_tab[2] = {nullptr, nullptr};
  ServeInst* h_inst[2] = {nullptr, nullptr};       // pinned staging
  ServeItem* h_items[2] = {nullptr, nullptr};
  MoveInst* h_minst[2] = {nullptr, nullptr};
  ServeItem* h_mitems[2] = {nullptr, nullptr};
  ServeTables* h_tab[2] = {nullptr, nullptr};
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  hipGraph_t graph_src[2] = {nullptr, nullptr};
  int64_t tick_no = 0;
  // statistics
  std::vector<double> lat;
  double inter_done = 0, first_submit = -1, last_done = 0;
  int64_t n_sub = 0, n_done = 0, n_fail = 0, n_cancel = 0;
  int loop_status = O3D_OK;
  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

template <class T>
int grow_pair(size_t need, size_t& cap, T** d, T** h) {
  if (need <= cap) return O3D_OK;
  const size_t nc = std::max<size_t>(need + need / 2, 64);
  for (int b = 0; b < 2; ++b) {
    if (d[b]) hipFree(d[b]);
    if (h[b]) hipHostFree(h[b]);
    d[b] = nullptr;
    h[b] = nullptr;
    if (hipMalloc(&d[b], sizeof(T) * nc) != hipSuccess) return fail(O3D_ENOMEM, "server tables");
    if (hipHostMalloc(&h[b], sizeof(T) * nc, hipHostMallocDefault) != hipSuccess)
      return fail(O3D_ENOMEM, "server pinned tables");
  }
  cap = nc;
  return O3D_OK;
}

void drop_graphs(o3d_server* S) {
  for (int b = 0; b < 2; ++b) {
    if (S->graph[b]) { hipGraphExecDestroy(S->graph[b]); S->graph[b] = nullptr; }
    if (S->graph_src[b]) { hipGraphDestroy(S->graph_src[b]); S->graph_src[b] = nullptr; }
  }
}

void retire(o3d_server* S, bool block_one);

// Grow the per-tick tables.  Both buffers are replaced, so the stream is drained
// first: no in-flight tick may still read the old ones.  The captured graphs hold
// the old addresses and are dropped.
int grow_tables(o3d_server* S, size_t ninst, size_t nitems, size_t nmove) {
  if (ninst <= S->inst_cap && nitems <= S->item_cap && nmove <= S->move_cap) return O3D_OK;
  HIPCHK(hipStreamSynchronize(S->stream));
  retire(S, false);
  drop_graphs(S);
  size_t c0 = S->inst_cap, c1 = S->inst_cap;
  int rc;
  if ((rc = grow_pair(ninst, c0, S->d_inst, S->h_inst))) return rc;
  if ((rc = grow_pair(ninst, c1, S->d_minst, S->h_minst))) return rc;
  S->inst_cap = std::min(c0, c1);
  if ((rc = grow_pair(nitems, S->item_cap, S->d_items, S->h_items))) return rc;
  if ((rc = grow_pair(nmove, S->move_cap, S->d_mitems, S->h_mitems))) return rc;
  return O3D_OK;
}

hipError_t launch_tick_kernels(o3d_server* S, int b, hipStream_t st) {
  hipLaunchKernelGGL(k_serve_reset, dim3(1), dim3(64), 0, st, S->d_tab[b]);
  const size_t lds = (size_t)S->tile * 2 * sizeof(float4);
  if (S->tpt == 1)
    hipLaunchKernelGGL(k_serve_eval<1>, dim3(S->grid), dim3(BLOCK), lds, st, S->d_inst[b], S->d_items[b], S->d_tab[b], S->tile);
  else if (S->tpt == 2)
    hipLaunchKernelGGL(k_serve_eval<2>, dim3(S->grid), dim3(BLOCK), lds, st, S->d_inst[b], S->d_items[b], S->d_tab[b], S->tile);
  else
    hipLaunchKernelGGL(k_serve_eval<4>, dim3(S->grid), dim3(BLOCK), lds, st, S->d_inst[b], S->d_items[b], S->d_tab[b], S->tile);
  hipLaunchKernelGGL(k_serve_move, dim3(S->move_grid), dim3(BLOCK), 0, st, S->d_minst[b], S->d_mitems[b], S->d_tab[b]);
  return hipGetLastError();
}

int enqueue_tick(o3d_server* S, int b, size_t ninst, size_t nitems, size_t nmove) {
  hipStream_t st = S->stream;
  HIPCHK(hipMemcpyAsync(S->d_inst[b], S->h_inst[b], sizeof(ServeInst) * std::max<size_t>(ninst, 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(S->d_items[b], S->h_items[b], sizeof(ServeItem) * std::max<size_t>(nitems, 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(S->d_minst[b], S->h_minst[b], sizeof(MoveInst) * std::max<size_t>(ninst, 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(S->d_mitems[b], S->h_mitems[b], sizeof(ServeItem) * std::max<size_t>(nmove, 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(S->d_tab[b], S->h_tab[b], sizeof(ServeTables), hipMemcpyHostToDevice, st));
  if (S->opts.use_graph) {
    if (!S->graph[b]) {
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      hipError_t e = launch_tick_kernels(S, b, st);
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(st, &g);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return fail(O3D_EDEVICE, "tick capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
      }
      S->graph_src[b] = g;
      HIPCHK(hipGraphInstantiate(&S->graph[b], g, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(S->graph[b], st));
  } else {
    HIPCHK(launch_tick_kernels(S, b, st));
  }
  return O3D_OK;
}

// the pair-evaluation cost of one chunk of request r
double chunk_cost(const o3d_server* S, const Session& ses, int64_t chunk) {
  const int64_t per = (int64_t)BLOCK * S->tpt;
  const int64_t lo = chunk * per, hi = std::min(ses.n, lo + per);
  return (double)(hi - lo) * (double)ses.n;
}

void finish_request(o3d_server* S, Request& r, int state, int status, double when) {
  if (r.state >= O3D_RQ_DONE) return;   // already ended (a failure path ended it first)
  r.state = state;
  r.status = status;
  r.done_s = when;
  auto sit = S->sessions.find(r.sid);
  if (sit != S->sessions.end()) sit->second.pending--;
  if (r.work) {
    S->arena.release(r.woff, r.wbytes);
    r.work = nullptr;
  }
  if (state == O3D_RQ_DONE) {
    S->n_done++;
    S->lat.push_back(r.done_s - r.submit_s);
    S->inter_done += r.interactions;
    S->last_done = std::max(S->last_done, when);
  } else if (state == O3D_RQ_FAILED) {
    S->n_fail++;
  } else {
    S->n_cancel++;
  }
}

// steps of this session go to the factored method
bool session_factored(const o3d_server* S, const Session& ses) {
  if (ses.core != CORE_RM && ses.core != CORE_EXP) return false;
  if (S->opts.method == O3D_METHOD_FACTORED) return true;
  return S->opts.method == O3D_METHOD_AUTO && ses.n >= LANES_FF_AUTO_MIN_N;
}

// admit queued requests (arrival order; a request of a session with an admitted
// step waits for it; others may pass a request that does not fit yet)
void admit(o3d_server* S) {
  const double t = S->now();
  const int max_active = S->opts.max_active;
  for (auto it = S->queue.begin(); it != S->queue.end();) {
    Request& r = S->requests.at(*it);
    if (r.cancel || (r.deadline > 0 && t > r.deadline)) {
      finish_request(S, r, O3D_RQ_CANCELLED, O3D_ECANCELLED, t);
      it = S->queue.erase(it);
      continue;
    }
    if ((int)S->active.size() >= max_active) break;
    bool blocked = false;
    for (int64_t a : S->active)
      if (S->requests.at(a).sid == r.sid) { blocked = true; break; }
    for (auto jt = S->queue.begin(); jt != it && !blocked; ++jt)
      if (S->requests.at(*jt).sid == r.sid) blocked = true;
    if (blocked) { ++it; continue; }
    const Session& ses = S->sessions.at(r.sid);
    if (g_fault_count.load() > 0) {
      if (g_fault_count.fetch_sub(1) > 0) {
        r.admit_s = t;
        finish_request(S, r, O3D_RQ_FAILED, O3D_EINJECTED, t);
        it = S->queue.erase(it);
        continue;
      }
    }
    size_t off;
    const size_t bytes = sizeof(float) * WORK_FLOATS * (size_t)ses.n;
    if (!S->arena.alloc(bytes, &off)) { ++it; continue; }
    r.woff = off;
    r.wbytes = bytes;
    r.work = (float*)(S->arena_base + off);
    r.state = O3D_RQ_RUNNING;
    r.admit_s = t;
    r.stage = 1;
    r.next_chunk = 0;
    r.nchunks = (ses.n + (int64_t)BLOCK * S->tpt - 1) / ((int64_t)BLOCK * S->tpt);
    r.factored = session_factored(S, ses);
    r.finished_planning = r.factored;   // the ticks never see it
    r.ff_launched = false;
    S->active.push_back(r.id);
    it = S->queue.erase(it);
  }
}

ServeInst eval_inst(const Session& ses, const Request& r) {
  ServeInst I{};
  const int64_t n = ses.n;
  const bool st1 = r.stage == 1;
  const float* x[3] = {st1 ? ses.x(0) : r.w(0, n), st1 ? ses.x(1) : r.w(1, n), st1 ? ses.x(2) : r.w(2, n)};
  const float* s[3] = {st1 ? ses.s(0) : r.w(3, n), st1 ? ses.s(1) : r.w(4, n), st1 ? ses.s(2) : r.w(5, n)};
  I.s = SrcView{x[0], x[1], x[2], ses.r(), s[0], s[1], s[2]};
  I.t.x = x[0]; I.t.y = x[1]; I.t.z = x[2]; I.t.r = ses.r();
  const int ub = st1 ? 6 : 18;   // u1 rows 6..8, g1 9..17; u2 18..20, g2 21..29
  I.t.u = r.w(ub, n); I.t.v = r.w(ub + 1, n); I.t.w = r.w(ub + 2, n);
  for (int k = 0; k < 9; ++k) I.t.g[k] = r.w(ub + 3 + k, n);
  I.n = n;
  I.core = ses.core;
  I.prec = r.prec;
  return I;
}

MoveInst move_inst(const Session& ses, const Request& r) {
  MoveInst m{};
  const int64_t n = ses.n;
  m.n = n;
  m.stage = r.stage;
  m.dt = r.dt;
  for (int d = 0; d < 3; ++d) {
    m.fs[d] = r.fs[d];
    m.x[d] = ses.x(d);
    m.s[d] = ses.s(d);
    m.x1[d] = r.w(d, n);
    m.s1[d] = r.w(3 + d, n);
    m.u1[d] = r.w(6 + d, n);
    m.u2[d] = r.w(18 + d, n);
  }
  for (int k = 0; k < 9; ++k) {
    m.g1[k] = r.w(9 + k, n);
    m.g2[k] = r.w(21 + k, n);
  }
  m.elong = ses.elong();
  return m;
}

// Both stages of a factored step on the least loaded lane: zero the stage's 12
// result rows (the factored method accumulates), evaluate the session on itself
// (self pair included, as in the ticks), then the stage update.
int enqueue_factored_step(o3d_server* S, Request& r, const Session& ses) {
  int li = 0;
  for (int k = 1; k < SRV_FF_LANES; ++k)
    if (S->lanes[k].pending < S->lanes[li].pending) li = k;
  FFLane& lane = S->lanes[li];
  hipStream_t st = lane.st;
  const int64_t n = ses.n;
  const unsigned zgrid = (unsigned)std::min<int64_t>((12 * n + BLOCK - 1) / BLOCK, 4096);
  const unsigned mgrid = (unsigned)std::min<int64_t>((n + BLOCK - 1) / BLOCK, (int64_t)S->move_grid);
  for (int stage = 1; stage <= 2; ++stage) {
    r.stage = stage;
    const ServeInst I = eval_inst(ses, r);
    hipLaunchKernelGGL(k_serve_zero, dim3(zgrid), dim3(BLOCK), 0, st, I.t.u, 12 * n);   // rows u, v, w, g[0..8]
    HIPCHK(hipGetLastError());
    o3d_points p{};
    p.n = n;
    p.x[0] = I.s.x; p.x[1] = I.s.y; p.x[2] = I.s.z;
    p.s[0] = I.s.ax; p.s[1] = I.s.ay; p.s[2] = I.s.az;
    p.r = I.s.r;
    p.u[0] = I.t.u; p.u[1] = I.t.v; p.u[2] = I.t.w;
    for (int k = 0; k < 9; ++k) p.ug[k] = I.t.g[k];
    o3d_opts o{};
    o.core = ses.core;
    o.results = O3D_VELANDGRAD;
    o.precision = r.prec;
    o.self_term = O3D_SELF_INCLUDE;
    o.method = O3D_METHOD_FACTORED;
    const int rc = p2p_factored_ws(&p, &p, &o, st, lane.ws, lane.bytes);
    if (rc != O3D_OK) return rc;
    hipLaunchKernelGGL(k_serve_move_one, dim3(mgrid), dim3(BLOCK), 0, st, move_inst(ses, r));
    HIPCHK(hipGetLastError());
  }
  FFJob j;
  j.rid = r.id;
  j.work = r.interactions;
  HIPCHK(hipEventCreateWithFlags(&j.done, hipEventDisableTiming));
  HIPCHK(hipEventRecord(j.done, st));
  lane.inflight.push_back(j);
  lane.pending += j.work;
  r.ff_launched = true;
  S->n_factored++;
  return O3D_OK;
}

// Size the factored scratch for a new session (caller holds S->mu, under which alone
// the loop enqueues factored steps).  The new block is allocated before the old one
// is freed, so a failure leaves the sessions already open working.
int ensure_ff_scratch(o3d_server* S, int64_t n, int32_t core) {
  Session probe{};
  probe.n = n;
  probe.core = core;
  if (!session_factored(S, probe)) return O3D_OK;
  const size_t need = ff_workspace_bytes(n, n, true, 0);
  if (need == 0) return fail(O3D_EDEVICE, "factored scratch query failed");
  for (FFLane& lane : S->lanes) {
    if (need <= lane.bytes) continue;
    void* p = nullptr;
    if (hipMalloc(&p, need) != hipSuccess) return fail(O3D_ENOMEM, "factored scratch of %zu bytes", need);
    if (lane.ws) {
      const hipError_t e = hipStreamSynchronize(lane.st);
      if (e != hipSuccess) {
        hipFree(p);
        return fail(O3D_EDEVICE, "factored stream: %s", hipGetErrorString(e));
      }
      hipFree(lane.ws);
    }
    lane.ws = p;
    lane.bytes = need;
  }
  return O3D_OK;
}

int launch_factored(o3d_server* S) {
  for (int64_t id : S->active) {
    Request& r = S->requests.at(id);
    if (!r.factored || r.ff_launched) continue;
    const int rc = enqueue_factored_step(S, r, S->sessions.at(r.sid));
    if (rc != O3D_OK) return rc;
  }
  return O3D_OK;
}

// finished factored steps (in each lane's stream order)
void retire_factored(o3d_server* S) {
  for (FFLane& lane : S->lanes)
  while (!lane.inflight.empty()) {
    FFJob& j = lane.inflight.front();
    const hipError_t q = hipEventQuery(j.done);
    if (q == hipErrorNotReady) break;
    const double when = S->now();
    if (q != hipSuccess) S->loop_status = fail(O3D_EDEVICE, "factored step failed: %s", hipGetErrorString(q));
    Request& r = S->requests.at(j.rid);
    S->active.erase(std::remove(S->active.begin(), S->active.end(), j.rid), S->active.end());
    finish_request(S, r, q == hipSuccess ? O3D_RQ_DONE : O3D_RQ_FAILED, q == hipSuccess ? O3D_OK : O3D_EDEVICE, when);
    hipEventDestroy(j.done);
    lane.pending -= j.work;
    lane.inflight.pop_front();
    S->done_cv.notify_all();
  }
}

// plan and enqueue one tick; returns false when there is nothing to run
int plan_tick(o3d_server* S, bool* ran) {
  *ran = false;
  std::vector<Request*> act;
  for (int64_t id : S->active) {
    Request& r = S->requests.at(id);
    if (!r.finished_planning) act.push_back(&r);
  }
  if (act.empty()) return O3D_OK;
  // A chunk of a large session keeps one workgroup busy for a long time, so the
  // tick must hold enough such chunks to occupy every workgroup slot: the budget
  // is at least slots x the largest active chunk.
  double budget = S->opts.tick_interactions;
  if (S->opts.adaptive_tick) {
    double biggest = 0;
    for (Request* r : act) biggest = std::max(biggest, chunk_cost(S, S->sessions.at(r->sid), 0));
    budget = std::max(budget, biggest * (double)S->grid);
  }
  struct Take { Request* r; int64_t from, to; };
  std::vector<Take> take;
  auto remaining = [&](Request* r) {
    const Session& ses = S->sessions.at(r->sid);
    return (double)(ses.n - std::min(ses.n, r->next_chunk * (int64_t)BLOCK * S->tpt)) * (double)ses.n;
  };
  auto grab = [&](Request* r, double limit) {
    const Session& ses = S->sessions.at(r->sid);
    int64_t from = r->next_chunk, to = from;
    double used = 0;
    while (to < r->nchunks) {
      const double c = chunk_cost(S, ses, to);
      if (used + c > limit && to > from) break;
      if (used + c > limit && !take.empty()) break;
      used += c;
      ++to;
    }
    if (to > from) {
      take.push_back({r, from, to});
      r->next_chunk = to;
    }
    return used;
  };
  // the oldest admitted request always advances (up to half the budget), then the
  // others smallest-remaining-first, then whatever budget is left in age order
  budget -= grab(act[0], budget * 0.5);
  std::vector<Request*> rest(act.begin() + 1, act.end());
  std::stable_sort(rest.begin(), rest.end(), [&](Request* a, Request* b) { return remaining(a) < remaining(b); });
  for (Request* r : rest) {
    if (budget <= 0) break;
    if (remaining(r) <= budget) budget -= grab(r, remaining(r));
  }
  for (Request* r : act) {
    if (budget <= 0) break;
    if (r->next_chunk < r->nchunks) budget -= grab(r, budget);
  }
  // tables: instance slot per taken request (eval) and per stage finish (move)
  std::map<Request*, int> slot;
  size_t nitems = 0, nmove = 0;
  for (auto& tk : take) {
    if (!slot.count(tk.r)) { int s = (int)slot.size(); slot[tk.r] = s; }
    nitems += (size_t)(tk.to - tk.from);
  }
  std::vector<Request*> finishing;
  for (auto& kv : slot)
    if (kv.first->next_chunk >= kv.first->nchunks) finishing.push_back(kv.first);
  for (Request* r : finishing) nmove += (size_t)((S->sessions.at(r->sid).n + BLOCK - 1) / BLOCK);
  // buffer b was last read by tick tick_no - 2, which has retired: plan_tick runs
  // only with at most one tick in flight
  const int b = (int)(S->tick_no & 1);
  int rc = grow_tables(S, std::max<size_t>(slot.size(), 1), std::max<size_t>(nitems, 1), std::max<size_t>(nmove, 1));
  if (rc) return rc;
  for (auto& kv : slot) {
    const Session& ses = S->sessions.at(kv.first->sid);
    S->h_inst[b][kv.second] = eval_inst(ses, *kv.first);
    S->h_minst[b][kv.second] = move_inst(ses, *kv.first);
  }
  // longest items first: with the queue drained greedily this is LPT scheduling
  std::stable_sort(take.begin(), take.end(), [&](const Take& x, const Take& y) {
    return S->sessions.at(x.r->sid).n > S->sessions.at(y.r->sid).n;
  });
  size_t k = 0;
  for (auto& tk : take)
    for (int64_t c = tk.from; c < tk.to; ++c) S->h_items[b][k++] = ServeItem{slot[tk.r], (int32_t)c};
  size_t m = 0;
  for (Request* r : finishing) {
    const int64_t nb = (S->sessions.at(r->sid).n + BLOCK - 1) / BLOCK;
    for (int64_t c = 0; c < nb; ++c) S->h_mitems[b][m++] = ServeItem{slot[r], (int32_t)c};
  }
  S->h_tab[b]->n_eval = (int32_t)nitems;
  S->h_tab[b]->n_move = (int32_t)nmove;
  S->h_tab[b]->counter = 0;
  if ((rc = enqueue_tick(S, b, slot.size(), nitems, nmove)) != O3D_OK) return rc;
  Tick t;
  t.slot = b;
  HIPCHK(hipEventCreateWithFlags(&t.done, hipEventDisableTiming));
  HIPCHK(hipEventRecord(t.done, S->stream));
  // advance stages now: later ticks are ordered after this one on the stream
  for (Request* r : finishing) {
    if (r->stage == 1) {
      r->stage = 2;
      r->next_chunk = 0;
    } else {
      r->finished_planning = true;
      t.completed.push_back(r->id);
    }
  }
  S->inflight.push_back(std::move(t));
  S->tick_no++;
  *ran = true;
  return O3D_OK;
}

void retire(o3d_server* S, bool block_one) {
  while (!S->inflight.empty()) {
    Tick& t = S->inflight.front();
    hipError_t q = block_one ? hipEventSynchronize(t.done) : hipEventQuery(t.done);
    if (q == hipErrorNotReady) return;
    block_one = false;
    const double when = S->now();
    const int status = q == hipSuccess ? O3D_OK : O3D_EDEVICE;
    if (q != hipSuccess) S->loop_status = fail(O3D_EDEVICE, "tick failed: %s", hipGetErrorString(q));
    for (int64_t id : t.completed) {
      Request& r = S->requests.at(id);
      S->active.erase(std::remove(S->active.begin(), S->active.end(), id), S->active.end());
      finish_request(S, r, status == O3D_OK ? O3D_RQ_DONE : O3D_RQ_FAILED, status, when);
    }
    hipEventDestroy(t.done);
    S->inflight.pop_front();
    S->done_cv.notify_all();
  }
}

void loop_main(o3d_server* S) {
  hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  while (true) {
    retire(S, false);
    retire_factored(S);
    if (S->stop && S->queue.empty() && S->active.empty() && S->inflight.empty() && S->ff_idle()) break;
    admit(S);
    bool ran = false;
    int rc = launch_factored(S);
    if (rc == O3D_OK && S->inflight.size() < 2) rc = plan_tick(S, &ran);
    if (rc != O3D_OK) {
      S->loop_status = rc;
      // fail everything that was admitted; the stream state is unknown (ticks and
      // factored steps still in flight retire later and find their requests ended)
      for (int64_t id : S->active) finish_request(S, S->requests.at(id), O3D_RQ_FAILED, rc, S->now());
      S->active.clear();
      S->done_cv.notify_all();
    }
    if (ran) continue;
    if (!S->inflight.empty() || !S->ff_idle()) {
      // let submissions in while the GPU works
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(200));
      lk.lock();
      continue;
    }
    S->done_cv.notify_all();
    S->cv.wait_for(lk, std::chrono::milliseconds(5));
  }
  S->done_cv.notify_all();
}

}  // namespace

extern "C" {

int o3d_inject_fault(int32_t kind, int32_t count) {
  if (kind == 0) { g_fault_count = 0; return O3D_OK; }
  if (kind != 1 || count < 0) return fail(O3D_EINVAL, "fault kind %d", kind);
  g_fault_count = count;
  return O3D_OK;
}

int o3d_server_create(o3d_server** out, const o3d_server_opts* opts) {
  clear_error();
  if (!out) return fail(O3D_EINVAL, "NULL handle pointer");
  *out = nullptr;
  int rc = ensure_init();
  if (rc) return rc;
  auto S = std::make_unique<o3d_server>();
  if (opts) S->opts = *opts;
  if (S->opts.hbm_budget == 0) S->opts.hbm_budget = (size_t)8 << 30;
  if (S->opts.max_active <= 0) S->opts.max_active = 64;
  if (S->opts.tick_interactions <= 0) S->opts.tick_interactions = 4e10;
  S->tile = S->opts.tile > 0 ? S->opts.tile : 256;
  S->tpt = S->opts.tpt > 0 ? S->opts.tpt : 1;
  if (S->tile % 64 || S->tile > 1024 || (S->tpt != 1 && S->tpt != 2 && S->tpt != 4))
    return fail(O3D_EINVAL, "bad tile/tpt");
  if (hipMalloc(&S->arena_base, S->opts.hbm_budget) != hipSuccess)
    return fail(O3D_ENOMEM, "arena of %zu bytes", S->opts.hbm_budget);
  S->arena.init(S->opts.hbm_budget);
  HIPCHK(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
  for (FFLane& lane : S->lanes) HIPCHK(hipStreamCreateWithFlags(&lane.st, hipStreamNonBlocking));
  for (int b = 0; b < 2; ++b) {
    HIPCHK(hipMalloc(&S->d_tab[b], sizeof(ServeTables)));
    HIPCHK(hipHostMalloc(&S->h_tab[b], sizeof(ServeTables), hipHostMallocDefault));
  }
  int per_cu = 0;
  const size_t lds = (size_t)S->tile * 2 * sizeof(float4);
  if (S->tpt == 1) hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_serve_eval<1>, BLOCK, lds);
  else if (S->tpt == 2) hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_serve_eval<2>, BLOCK, lds);
  else hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_serve_eval<4>, BLOCK, lds);
  S->grid = std::max(1, per_cu) * num_cus();
  S->move_grid = 4 * num_cus();
  HIPCHK(hipGetDevice(&S->device));
  o3d_server* raw = S.release();
  raw->loop = std::thread(loop_main, raw);
  *out = raw;
  return O3D_OK;
}

int o3d_server_destroy(o3d_server* S) {
  if (!S) return O3D_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (int64_t id : S->queue) finish_request(S, S->requests.at(id), O3D_RQ_CANCELLED, O3D_ECANCELLED, S->now());
    S->queue.clear();
    S->stop = true;
  }
  S->cv.notify_all();
  if (S->loop.joinable()) S->loop.join();
  hipStreamSynchronize(S->stream);
  for (FFLane& lane : S->lanes) hipStreamSynchronize(lane.st);
  for (int b = 0; b < 2; ++b) {
    if (S->graph[b]) hipGraphExecDestroy(S->graph[b]);
    if (S->graph_src[b]) hipGraphDestroy(S->graph_src[b]);
    hipFree(S->d_inst[b]); hipFree(S->d_items[b]); hipFree(S->d_minst[b]); hipFree(S->d_mitems[b]); hipFree(S->d_tab[b]);
    hipHostFree(S->h_inst[b]); hipHostFree(S->h_items[b]); hipHostFree(S->h_minst[b]); hipHostFree(S->h_mitems[b]);
    hipHostFree(S->h_tab[b]);
  }
  hipStreamDestroy(S->stream);
  for (FFLane& lane : S->lanes) {
    hipStreamDestroy(lane.st);
    if (lane.ws) hipFree(lane.ws);
  }
  hipFree(S->arena_base);
  delete S;
  return O3D_OK;
}

int o3d_server_open_session(o3d_server* S, const o3d_points* p, int32_t core, int64_t* sid) {
  clear_error();
  if (!S || !p || !sid) return fail(O3D_EINVAL, "NULL argument");
  if (core < 1 || core > 4) return fail(O3D_EINVAL, "unknown core %d", core);
  if (p->n <= 0) return fail(O3D_EINVAL, "empty session");
  if (p->n > ((int64_t)1 << 31)) return fail(O3D_EINVAL, "session too large");
  for (int d = 0; d < 3; ++d)
    if (!p->x[d] || !p->s[d]) return fail(O3D_EINVAL, "position/strength component %d is NULL", d);
  if (!p->r) return fail(O3D_EINVAL, "radii are NULL");
  const int64_t n = p->n;
  for (int64_t i = 0; i < n; ++i) {
    bool ok = std::isfinite(p->r[i]) && p->r[i] > 0.0f;
    for (int d = 0; d < 3; ++d) ok = ok && std::isfinite(p->x[d][i]) && std::isfinite(p->s[d][i]);
    if (!ok) return fail(O3D_EINVAL, "particle %lld has a non-finite position/strength or a radius <= 0", (long long)i);
  }
  std::vector<float> host(STATE_FLOATS * (size_t)n);
  for (int d = 0; d < 3; ++d) {
    std::copy(p->x[d], p->x[d] + n, host.begin() + d * n);
    std::copy(p->s[d], p->s[d] + n, host.begin() + (3 + d) * n);
  }
  std::copy(p->r, p->r + n, host.begin() + 6 * n);
  std::fill(host.begin() + 7 * n, host.end(), 1.0f);   // elongation starts at 1
  Session ses{};
  {
    std::lock_guard<std::mutex> lk(S->mu);
    size_t off;
    if (!S->arena.alloc(sizeof(float) * host.size(), &off))
      return fail(O3D_EADMIT, "session of %lld particles (%zu bytes) does not fit the arena (%zu of %zu used)",
                  (long long)n, sizeof(float) * host.size(), S->arena.used, S->arena.size);
    const int frc = ensure_ff_scratch(S, n, core);
    if (frc != O3D_OK) {
      S->arena.release(off, sizeof(float) * host.size());
      return frc;
    }
    ses.id = S->next_sid++;
    ses.n = n;
    ses.core = core;
    ses.off = off;
    ses.bytes = sizeof(float) * host.size();
    ses.base = (float*)(S->arena_base + off);
    S->sessions[ses.id] = ses;
  }
  HIPCHK(hipMemcpy(ses.base, host.data(), ses.bytes, hipMemcpyHostToDevice));
  *sid = ses.id;
  return O3D_OK;
}

int o3d_server_close_session(o3d_server* S, int64_t sid) {
  clear_error();
  if (!S) return fail(O3D_EINVAL, "NULL server");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->sessions.find(sid);
  if (it == S->sessions.end()) return fail(O3D_ENOTFOUND, "no session %lld", (long long)sid);
  it->second.closing = true;
  S->done_cv.wait(lk, [&] { return S->sessions.at(sid).pending == 0; });
  S->arena.release(it->second.off, it->second.bytes);
  S->sessions.erase(sid);
  return O3D_OK;
}

int o3d_server_read_session(o3d_server* S, int64_t sid, float* const x[3], float* const str[3], float* elong) {
  clear_error();
  if (!S || !x || !str) return fail(O3D_EINVAL, "NULL argument");
  Session ses;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->sessions.find(sid);
    if (it == S->sessions.end()) return fail(O3D_ENOTFOUND, "no session %lld", (long long)sid);
    S->done_cv.wait(lk, [&] { return S->sessions.at(sid).pending == 0; });
    ses = S->sessions.at(sid);
  }
  for (int d = 0; d < 3; ++d) {
    HIPCHK(hipMemcpy(x[d], ses.x(d), sizeof(float) * ses.n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(str[d], ses.s(d), sizeof(float) * ses.n, hipMemcpyDeviceToHost));
  }
  if (elong) HIPCHK(hipMemcpy(elong, ses.elong(), sizeof(float) * ses.n, hipMemcpyDeviceToHost));
  return O3D_OK;
}

int o3d_server_submit(o3d_server* S, int64_t sid, double dt, int32_t prec, const double* fs, double timeout_s,
                      int64_t* rid) {
  clear_error();
  if (!S || !rid) return fail(O3D_EINVAL, "NULL argument");
  if (prec != PREC_FP32 && prec != PREC_COMP) return fail(O3D_EINVAL, "steps run in FP32 or compensated FP32");
  if (!std::isfinite(dt)) return fail(O3D_EINVAL, "non-finite dt");
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->loop_status != O3D_OK) return fail(S->loop_status, "server stopped after a device error");
  auto it = S->sessions.find(sid);
  if (it == S->sessions.end() || it->second.closing) return fail(O3D_ENOTFOUND, "no session %lld", (long long)sid);
  const size_t need = sizeof(float) * WORK_FLOATS * (size_t)it->second.n;
  if (need > S->arena.size - it->second.bytes)
    return fail(O3D_EADMIT, "a step of %lld particles needs %zu bytes; the arena holds %zu",
                (long long)it->second.n, need, S->arena.size);
  Request r{};
  r.id = S->next_rid++;
  r.sid = sid;
  r.dt = dt;
  r.prec = prec;
  for (int d = 0; d < 3; ++d) r.fs[d] = fs ? fs[d] : 0.0;
  r.submit_s = S->now();
  r.deadline = timeout_s > 0 ? r.submit_s + timeout_s : 0.0;
  r.interactions = 2.0 * (double)it->second.n * (double)it->second.n;
  it->second.pending++;
  if (S->first_submit < 0) S->first_submit = r.submit_s;
  S->n_sub++;
  S->requests[r.id] = r;
  S->queue.push_back(r.id);
  *rid = r.id;
  S->cv.notify_all();
  return O3D_OK;
}

int o3d_server_wait(o3d_server* S, int64_t rid, double timeout_s, o3d_request_info* info) {
  clear_error();
  if (!S) return fail(O3D_EINVAL, "NULL server");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->requests.find(rid);
  if (it == S->requests.end()) return fail(O3D_ENOTFOUND, "no request %lld", (long long)rid);
  auto ended = [&] { const int st = S->requests.at(rid).state; return st >= O3D_RQ_DONE; };
  if (timeout_s < 0) S->done_cv.wait(lk, ended);
  else S->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), ended);
  const Request& r = S->requests.at(rid);
  if (info) {
    info->state = r.state;
    info->status = r.status;
    info->submit_s = r.submit_s;
    info->admit_s = r.admit_s;
    info->done_s = r.done_s;
    info->latency_s = r.state >= O3D_RQ_DONE ? r.done_s - r.submit_s : 0.0;
    info->interactions = r.interactions;
  }
  if (r.state < O3D_RQ_DONE) return fail(O3D_EBUSY, "request %lld still running", (long long)rid);
  return r.status;
}

int o3d_server_cancel(o3d_server* S, int64_t rid) {
  clear_error();
  if (!S) return fail(O3D_EINVAL, "NULL server");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->requests.find(rid);
  if (it == S->requests.end()) return fail(O3D_ENOTFOUND, "no request %lld", (long long)rid);
  if (it->second.state != O3D_RQ_QUEUED) return fail(O3D_EBUSY, "request %lld is already running or done", (long long)rid);
  it->second.cancel = true;
  S->cv.notify_all();
  return O3D_OK;
}

int o3d_server_get_stats(o3d_server* S, o3d_server_stats* st) {
  clear_error();
  if (!S || !st) return fail(O3D_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *st = o3d_server_stats{};
  st->submitted = S->n_sub;
  st->completed = S->n_done;
  st->failed = S->n_fail;
  st->cancelled = S->n_cancel;
  st->queued = (int64_t)S->queue.size();
  st->running = (int64_t)S->active.size();
  std::vector<double> l = S->lat;
  std::sort(l.begin(), l.end());
  auto pct = [&](double p) {
    if (l.empty()) return 0.0;
    const double k = p * (double)(l.size() - 1);
    const size_t a = (size_t)std::floor(k), b = std::min(l.size() - 1, a + 1);
    return l[a] + (l[b] - l[a]) * (k - (double)a);
  };
  st->p50_s = pct(0.5);
  st->p99_s = pct(0.99);
  double sum = 0;
  for (double v : l) sum += v;
  st->mean_s = l.empty() ? 0.0 : sum / (double)l.size();
  st->interactions = S->inter_done;
  st->wall_s = S->first_submit >= 0 ? S->last_done - S->first_submit : 0.0;
  st->ginter_per_s = st->wall_s > 0 ? S->inter_done / st->wall_s / 1e9 : 0.0;
  st->hbm_used = S->arena.used;
  st->hbm_budget = S->arena.size;
  st->hbm_peak = S->arena.peak;
  st->ticks = S->tick_no;
  st->factored_steps = S->n_factored;
  st->factored_scratch = 0;
  for (const FFLane& lane : S->lanes) st->factored_scratch += lane.bytes;
  return O3D_OK;
}

int o3d_server_reset_stats(o3d_server* S) {
  if (!S) return fail(O3D_EINVAL, "NULL server");
  std::lock_guard<std::mutex> lk(S->mu);
  S->lat.clear();
  S->inter_done = 0;
  S->first_submit = -1;
  S->last_done = 0;
  S->n_sub = S->n_done = S->n_fail = S->n_cancel = 0;
  S->n_factored = 0;
  S->arena.peak = S->arena.used;
  return O3D_OK;
}

}  // extern "C"
