// sort.hip -- stable LSD radix sort of (key, index) pairs and an exclusive scan
// of int32 counts, for gfx950.
//
// Used for the Morton order of particles (the factored method's sources and
// targets, the direct EXP kernel's target order) and for the factored method's
// work plan.  Everything here is plain kernel launches on the caller's stream:
// no memory fills, no allocation, no host synchronisation, so a call can be
// recorded into a HIP graph.
//
// One pass per RS_BITS-bit digit, three kernels per pass:
//   k_rs_hist     per workgroup tile of RS_TILE keys: digit histogram ->
//                 cnt[digit][tile] (digit-major);
//   k_rs_rowscan  one workgroup per digit: exclusive scan of its row over the
//                 tiles, and the digit's total;
//   k_rs_scatter  per tile: the digit bases (exclusive scan of the totals) plus
//                 the row offset, then the tile's keys in RS_I rounds of RS_T
//                 consecutive keys.  Within a round, a key's rank among equal
//                 digits of its wave comes from RS_BITS ballots (the lanes that
//                 agree on every bit), and the waves of the workgroup are
//                 ordered through LDS counters; so equal digits keep their
//                 input order and the sort is stable.  A stable sort of
//                 (key, index) pairs whose indices start in ascending order
//                 gives one permutation, whatever the algorithm: ties stay in
//                 index order.
#include "o3d_api_internal.hpp"

#include <algorithm>

namespace o3d {
namespace {

constexpr int RS_T = 256;             // threads per workgroup (4 waves)
constexpr int RS_I = 16;              // keys per thread
constexpr int RS_TILE = RS_T * RS_I;  // keys per workgroup tile
constexpr int RS_BITS = 8;
constexpr int RS_B = 1 << RS_BITS;    // digits per pass (== RS_T: one digit per thread)
constexpr int RS_W = RS_T / 64;       // waves per workgroup
static_assert(RS_B == RS_T, "one digit per thread");

// exclusive prefix of v over the RS_T threads of the workgroup; *total gets the sum.
// sh: RS_W words of LDS.  Every thread must call it.
__device__ __forceinline__ unsigned wg_excl_scan(unsigned v, unsigned* sh, unsigned* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  __syncthreads();              // sh is free (an earlier call's reads are done)
  if (lane == 63) sh[w] = x;
  __syncthreads();
  unsigned before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < RS_W; ++k) {
    const unsigned s = sh[k];
    before += k < w ? s : 0u;
    all += s;
  }
  *total = all;
  return before + x - v;
}

__global__ __launch_bounds__(RS_T) void k_rs_hist(const unsigned* __restrict__ key, int64_t n, int shift,
                                                  unsigned* __restrict__ cnt, int nblk) {
  __shared__ unsigned h[RS_B];
  h[threadIdx.x] = 0u;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int r = 0; r < RS_I; ++r) {
    const int64_t i = base + (int64_t)r * RS_T + threadIdx.x;
    if (i < n) atomicAdd(&h[(key[i] >> shift) & (RS_B - 1)], 1u);
  }
  __syncthreads();
  cnt[(int64_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// grid = RS_B workgroups, one per digit row of cnt
__global__ __launch_bounds__(RS_T) void k_rs_rowscan(unsigned* __restrict__ cnt, int nblk,
                                                     unsigned* __restrict__ tot) {
  __shared__ unsigned sh[RS_W];
  unsigned* row = cnt + (int64_t)blockIdx.x * nblk;
  unsigned carry = 0u;
  for (int b0 = 0; b0 < nblk; b0 += RS_T) {
    const int b = b0 + (int)threadIdx.x;
    const unsigned v = b < nblk ? row[b] : 0u;
    unsigned t;
    const unsigned ex = wg_excl_scan(v, sh, &t);
    if (b < nblk) row[b] = carry + ex;
    carry += t;
  }
  if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

__global__ __launch_bounds__(RS_T) void k_rs_scatter(const unsigned* __restrict__ kin, const int* __restrict__ vin,
                                                     unsigned* __restrict__ kout, int* __restrict__ vout, int64_t n,
                                                     int shift, const unsigned* __restrict__ cnt,
                                                     const unsigned* __restrict__ tot, int nblk) {
  __shared__ unsigned off[RS_B];        // next free slot of each digit for this tile
  __shared__ unsigned wc[RS_W][RS_B];   // this round: keys per (wave, digit)
  __shared__ unsigned wo[RS_W][RS_B];   // this round: first slot per (wave, digit)
  __shared__ unsigned sh[RS_W];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;

  unsigned k[RS_I];
  int v[RS_I];
#pragma unroll
  for (int r = 0; r < RS_I; ++r) {
    const int64_t i = base + (int64_t)r * RS_T + tid;
    k[r] = i < n ? kin[i] : 0u;
    v[r] = i < n ? vin[i] : 0;
  }
  unsigned t;
  const unsigned dbase = wg_excl_scan(tot[tid], sh, &t);
  off[tid] = dbase + cnt[(int64_t)tid * nblk + blockIdx.x];
#pragma unroll
  for (int q = 0; q < RS_W; ++q) wc[q][tid] = 0u;
  __syncthreads();

  const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int r = 0; r < RS_I; ++r) {
    const int64_t i = base + (int64_t)r * RS_T + tid;
    const bool valid = i < n;
    const unsigned d = (k[r] >> shift) & (RS_B - 1);
    unsigned long long m = __builtin_amdgcn_ballot_w64(valid);
#pragma unroll
    for (int b = 0; b < RS_BITS; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long bm = __builtin_amdgcn_ballot_w64(bit);
      m &= bit ? bm : ~bm;
    }
    const unsigned rank = (unsigned)__popcll(m & lt);
    if (valid && rank == 0u) wc[w][d] = (unsigned)__popcll(m);
    __syncthreads();
    {   // digit tid: slots of the waves in wave order; clear the counts for the next round
      unsigned s = off[tid];
#pragma unroll
      for (int q = 0; q < RS_W; ++q) {
        const unsigned c = wc[q][tid];
        wo[q][tid] = s;
        wc[q][tid] = 0u;
        s += c;
      }
      off[tid] = s;
    }
    __syncthreads();
    if (valid) {
      const unsigned dst = wo[w][d] + rank;
      kout[dst] = k[r];
      vout[dst] = v[r];
    }
    // the next round writes wc (cleared above) before its first barrier and wo only
    // after it, when every lane has read this round's wo
  }
}

// exclusive scan of n int32 values by one workgroup, RS_T * 4 values per step
__global__ __launch_bounds__(RS_T) void k_scan_i32(const int* __restrict__ in, int* __restrict__ out, int64_t n) {
  __shared__ unsigned sh[RS_W];
  unsigned carry = 0u;
  for (int64_t b0 = 0; b0 < n; b0 += (int64_t)RS_T * 4) {
    const int64_t i0 = b0 + (int64_t)threadIdx.x * 4;
    unsigned v[4], s = 0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j] = i0 + j < n ? (unsigned)in[i0 + j] : 0u;
      s += v[j];
    }
    unsigned t;
    unsigned ex = carry + wg_excl_scan(s, sh, &t);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (i0 + j < n) out[i0 + j] = (int)ex;
      ex += v[j];
    }
    carry += t;
  }
}

int blocks_of(int64_t n) { return (int)std::max<int64_t>(1, (n + RS_TILE - 1) / RS_TILE); }

}  // namespace

size_t radix_tmp_bytes(int64_t n) {
  return sizeof(unsigned) * ((size_t)RS_B * (size_t)blocks_of(n) + RS_B);
}

int radix_sort_pairs(unsigned* ka, int* va, unsigned* kb, int* vb, int64_t n, int bits, void* tmp, hipStream_t st,
                     unsigned** keys, int** vals) {
  if (n < 0 || n > INT32_MAX || bits < 1 || bits > 32 || !tmp)
    return fail(O3D_EINVAL, "radix sort: n %lld, bits %d", (long long)n, bits);
  unsigned* kin = ka;
  int* vin = va;
  unsigned* kout = kb;
  int* vout = vb;
  if (n > 0) {
    const int nblk = blocks_of(n);
    unsigned* cnt = static_cast<unsigned*>(tmp);
    unsigned* tot = cnt + (size_t)RS_B * nblk;
    for (int shift = 0; shift < bits; shift += RS_BITS) {
      hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(RS_T), 0, st, kin, n, shift, cnt, nblk);
      hipLaunchKernelGGL(k_rs_rowscan, dim3(RS_B), dim3(RS_T), 0, st, cnt, nblk, tot);
      hipLaunchKernelGGL(k_rs_scatter, dim3(nblk), dim3(RS_T), 0, st, kin, vin, kout, vout, n, shift, cnt, tot, nblk);
      HIPCHK(hipGetLastError());
      std::swap(kin, kout);
      std::swap(vin, vout);
    }
  }
  *keys = kin;
  *vals = vin;
  return O3D_OK;
}

int exclusive_scan_i32(const int* in, int* out, int64_t n, hipStream_t st) {
  if (n <= 0) return O3D_OK;
  hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(RS_T), 0, st, in, out, n);
  HIPCHK(hipGetLastError());
  return O3D_OK;
}

}  // namespace o3d
