// o3d_api_internal.hpp -- declarations shared by the host-side translation units.
#pragma once

#include "o3d_biot.h"
#include "p2p_kernels_decl.hpp"

#include <mutex>

namespace o3d {

int fail(int code, const char* fmt, ...);
void clear_error();
const char* last_error();
int ensure_init();
int num_cus();
hipStream_t lib_stream();
int lib_device();   // the context's device ordinal
std::mutex& pool_mutex();
int pool_buffer(int slot, size_t bytes, void** out);

struct Tuning {
  int tile;
  int tpt;
};
bool tuned_lookup(int core, int results, int prec, bool thick, int64_t nt, Tuning* out);
void choose_config(const Tuning& req, int64_t nt, int64_t ns, bool grad, int prec, int* tile, int* tpt);
hipError_t dispatch_p2p(int core, const P2PArgs& a, bool thick, bool grad, int prec, int tpt, hipStream_t st);
double flops_of(int core, bool grad, bool thick, int64_t nt, int64_t ns);
int check_opts(const o3d_opts* o);
// host-memory panels: every node index below nn (device panels are the caller's to keep in range)
int check_node_indices(const o3d_panels* p, const char* what);
void fill_stats(o3d_stats* stats, double secs, double flops, double inter);
bool ff_supported(int core, bool grad, int prec);
int p2p_factored(const o3d_points* src, o3d_points* targ, const o3d_opts* o, hipStream_t st,
                 o3d_factored_profile* prof = nullptr);
int p2p_factored_ws(const o3d_points* src, o3d_points* targ, const o3d_opts* o, hipStream_t st, void* ws,
                    size_t ws_bytes);
size_t ff_workspace_bytes(int64_t ns, int64_t nt, bool same, int tile);
// sources and targets are one collection (the factored method sorts it once)
bool ff_same_points(const o3d_points* src, const o3d_points* targ);
// one growable device workspace per stream (calls on a stream are ordered, so they
// share it); hold stream_ws_mutex() from stream_workspace until the last enqueue
std::mutex& stream_ws_mutex();
int stream_workspace(hipStream_t st, size_t bytes, void** out);
// whether st is being captured into a graph; whether its workspace holds bytes
// already (callers hold stream_ws_mutex())
bool stream_capturing(hipStream_t st);
// frees st's workspace (before the stream is destroyed); takes stream_ws_mutex()
void release_stream_workspace(hipStream_t st);
bool stream_workspace_fits(hipStream_t st, size_t bytes);
// Morton order of n points into a workspace of point_order_bytes(n): *perm gets the
// permutation (slot -> point index), *keys (optional) the sorted 30-bit keys
size_t point_order_bytes(int64_t n);
// a direct EXP call's sources as rows in the targets' Morton order (one collection)
// with tile spheres, for k_p2p (P2PArgs::sp, sa, tsph, tcs): scratch bytes, and the
// enqueue into ws
size_t sorted_sources_bytes(int64_t n, int tile);
int sorted_sources(const o3d_points* src, const int32_t* perm, int tile, void* ws, hipStream_t st, P2PArgs* a);
int point_order(const float* x, const float* y, const float* z, int64_t n, void* ws, hipStream_t st,
                const int32_t** perm, const unsigned** keys = nullptr);
// stable radix sort of (key, value) pairs by the low `bits` bits of the keys
// (sort.hip): (ka, va) and (kb, vb) are ping-pong buffers of n entries, both
// overwritten; *keys / *vals name the one that holds the result.  tmp holds
// radix_tmp_bytes(n).  Kernels only: capturable into a graph.
size_t radix_tmp_bytes(int64_t n);
int radix_sort_pairs(unsigned* ka, int* va, unsigned* kb, int* vb, int64_t n, int bits, void* tmp, hipStream_t st,
                     unsigned** keys, int** vals);
// exclusive prefix sum of n int32 values (one workgroup; meant for small n)
int exclusive_scan_i32(const int* in, int* out, int64_t n, hipStream_t st);
// a collection large enough for the factored method to pay when the caller leaves the choice open
constexpr int64_t FF_AUTO_MIN_N = 65536;
// the same for an instance of a batch or a session of the server: their factored work runs on
// streams of its own while the persistent kernel sums the rest, so the split point that balances
// the two is lower.  256 RM rings of 2,048-262,144 particles, compensated, graph replay: 682-685 ms
// at 65,536, 667-668 at 49,152, 665-666 at 40,960, 662-664 at 32,768, 741 at 98,304.  16 served
// sessions of 1,986-296,912 particles: at 32,768 against 65,536 saturated 2,881-2,905 against
// 2,853-2,856 G interactions/s, Poisson p50 7.8 against 10.6-11.0 ms, p99 78.5 against 80.3 ms
constexpr int64_t LANES_FF_AUTO_MIN_N = 32768;
// direct EXP calls with at least this many targets visit them in Morton order
constexpr int64_t TARGET_ORDER_MIN_N = 32768;
int p2p_dev(const o3d_points* src, o3d_points* targ, const o3d_opts* o, hipStream_t st, const Tuning* tun);
// the direct method of p2p_dev (one-pass kernel, source split, EXP target order).
// ws == nullptr: scratch from the stream's workspace; otherwise ws_bytes of the
// caller's scratch, never grown: a split takes fewer slices to fit, and a target
// order that does not fit is an error
int p2p_direct(const o3d_points* src, o3d_points* targ, const o3d_opts* o, hipStream_t st, const Tuning* tun,
               void* ws, size_t ws_bytes);
int check_p2p_args(const o3d_points* src, const o3d_points* targ, const o3d_opts* o);
// the no-allocation entry point's scratch on `device` (the current device): its
// address is looked up once; called by the library's init
int noalloc_setup(int device);
int bem_matrix_dev(const o3d_panels* src, const o3d_panels* targ, int32_t nunk, int32_t same, int32_t maxlev,
                   float* coeffs, hipStream_t st, o3d_bem_matrix_profile* prof);

}  // namespace o3d

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return ::o3d::fail(O3D_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
