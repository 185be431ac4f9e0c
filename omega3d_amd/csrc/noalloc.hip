// noalloc.hip -- o3d_points_affect_points_noalloc_dev: the direct point -> point
// evaluation on device arrays for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The call allocates nothing: the scratch of its kernels (the source
// split's per-slice partial sums, the Morton target order of EXP calls) lives in
// one static device array, `scratch`, that every GPU the library runs on holds
// until the process ends.
//
// Host code never passes `scratch` itself to a kernel or a copy: in host code the
// name is a host-side stand-in, not a device address.  The library's init looks
// the array's device address up once per device (hipGetSymbolAddress, with that
// device current) and every call uses that address.
//
// Every call on a GPU shares that GPU's array and it keeps the last call's
// contents: calls that overlap on one GPU (two streams, two threads) race on it.
// Nothing in it needs to start at zero: the sort's counts and the bounding box are
// written by the kernels that use them, and the split writes every partial sum it
// later reads.
#include "o3d_api_internal.hpp"

#include <mutex>

// 66 MiB: the Morton order of O3D_NOALLOC_MAX_TARGETS targets (16 B per target for
// keys and indices, twice, plus the radix sort's counts) fits
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)66 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace o3d {
namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_na_mu;
void* g_na_ptr[MAX_DEVICES] = {};
hipError_t g_na_err[MAX_DEVICES] = {};

}  // namespace

int noalloc_setup(int device) {
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (device < 0 || device >= MAX_DEVICES) return O3D_ENODEV;
  if (g_na_ptr[device]) return O3D_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    g_na_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
    return O3D_EDEVICE;
  }
  g_na_ptr[device] = p;
  return O3D_OK;
}

}  // namespace o3d

using namespace o3d;

extern "C" int o3d_points_affect_points_noalloc_dev(const o3d_points* src, o3d_points* targ, const o3d_opts* opts,
                                                    o3d_stats* stats) {
  clear_error();
  int rc;
  if ((rc = check_p2p_args(src, targ, opts)) != O3D_OK) return rc;
  if (opts->precision == PREC_FP64)
    return fail(O3D_EINVAL, "no-allocation call: FP64 is not served (use o3d_points_affect_points64_dev)");
  if (opts->method == O3D_METHOD_FACTORED)
    return fail(O3D_EINVAL, "no-allocation call: the factored method needs a workspace of its own; "
                            "method AUTO and DIRECT both run the direct kernels here");
  if (targ->n > O3D_NOALLOC_MAX_TARGETS)
    return fail(O3D_EINVAL, "no-allocation call: %lld targets, at most %lld (its scratch is a fixed %zu bytes)",
                (long long)targ->n, (long long)O3D_NOALLOC_MAX_TARGETS, NOALLOC_SCRATCH_BYTES);
  if ((rc = ensure_init()) != O3D_OK) return rc;
  const int dev = lib_device();
  void* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (dev >= 0 && dev < MAX_DEVICES) ptr = g_na_ptr[dev];
    if (!ptr)
      return fail(O3D_EDEVICE, "no-allocation call: the scratch array of device %d has no device address (%s)", dev,
                  dev >= 0 && dev < MAX_DEVICES ? hipGetErrorString(g_na_err[dev]) : "device out of range");
  }
  const bool grad = opts->results == O3D_VELANDGRAD;
  rc = p2p_direct(src, targ, opts, (hipStream_t)opts->stream, nullptr, ptr, NOALLOC_SCRATCH_BYTES);
  if (rc != O3D_OK) return rc;
  fill_stats(stats, 0.0, flops_of(opts->core, grad, targ->r != nullptr, targ->n, src->n),
             (double)targ->n * (double)src->n);
  return O3D_OK;
}
