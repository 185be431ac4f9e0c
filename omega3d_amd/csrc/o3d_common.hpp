// o3d_common.hpp -- shared types and device math for the gfx950 Biot-Savart kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace o3d {

enum { CORE_RM = 1, CORE_WL = 2, CORE_EXP = 3, CORE_V2 = 4 };
enum { PREC_FP32 = 0, PREC_COMP = 1, PREC_FP64 = 2 };

constexpr int BLOCK = 256;          // 4 wave64s per workgroup
constexpr int WAVE = 64;

// device SoA views (all pointers are device pointers)
struct SrcView {
  const float* x; const float* y; const float* z; const float* r;
  const float* ax; const float* ay; const float* az;
};
struct TgtView {
  const float* x; const float* y; const float* z; const float* r;   // r == nullptr: point targets
  float* u; float* v; float* w;
  float* g[9];                                                       // nullptr when velocity only
};

struct P2PArgs {
  SrcView s;
  TgtView t;
  int64_t ns;
  int64_t nt;
  int32_t tile;        // sources per LDS tile (multiple of 64, <= 1024)
  int32_t self_excl;   // subtract the j == i pair (src and targ are one collection)
  const int32_t* tperm;   // k_p2p only: target order (nullptr: index order)
  // k_p2p only, EXP core: sources staged from rows sorted in the targets' Morton
  // order ({x, y, z, core term}, {alpha, 0}), with each tile's bounding sphere and
  // largest core term (nullptr: the SoA arrays in index order)
  const float4* sp;
  const float4* sa;
  const float4* tsph;
  const float* tcs;
};

// double-precision views for the FP64 oracle kernel
struct SrcView64 {
  const double* x; const double* y; const double* z; const double* r;
  const double* ax; const double* ay; const double* az;
};
struct TgtView64 {
  const double* x; const double* y; const double* z; const double* r;
  double* u; double* v; double* w; double* g[9];
};
struct P2PArgs64 {
  SrcView64 s; TgtView64 t; int64_t ns; int64_t nt; int32_t tile; int32_t self_excl;
};

// ---------------------------------------------------------------------------
// fast FP32 transcendentals: single VALU ops (about 1 ulp), no IEEE fix-up code
// ---------------------------------------------------------------------------
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }
__device__ __forceinline__ bool wave_all(bool p) { return __builtin_amdgcn_ballot_w64(!p) == 0; }

// sum over the 64 lanes of a wave by a fixed butterfly (the same order on every run)
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// sum over each group of 16 lanes (lanes 16q .. 16q+15) by a fixed butterfly
__device__ __forceinline__ double quarter_sum(double x) {
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// error-free transformation a + b = s + e (Knuth's TwoSum)
__device__ __forceinline__ void two_sum(float& hi, float& lo, float b) {
  const float s = hi + b;
  const float bb = s - hi;
  const float e = (hi - (s - bb)) + (b - bb);
  hi = s;
  lo += e;
}

// u + (hi + lo) in FP32: TwoSum of u and hi, the two low parts added, one rounding
__device__ __forceinline__ float add_comp(float u, float hi, float lo) {
  float s = u, e = lo;
  two_sum(s, e, hi);
  return s + e;
}

// per-particle core quantity kept in LDS / registers so that each pair needs
// one add to form the pair's core parameter:
//   RM, WL: sigma^2  (r2 = rho^2 + ss^2 + st^2;  R^2 = ss^2 + st^2)
//   EXP:    sigma^3  (c = 1/(ss^3 + st^3))
//   V2:     sigma^4  (den = rho^4 + ss^4 + st^4)
template <int CORE> __device__ __forceinline__ float core_term(float r) {
  if (CORE == CORE_EXP) return r * r * r;
  if (CORE == CORE_V2) { const float r2 = r * r; return r2 * r2; }
  return r * r;
}
template <int CORE> __device__ __forceinline__ double core_term64(double r) {
  if (CORE == CORE_EXP) return r * r * r;
  if (CORE == CORE_V2) { const double r2 = r * r; return r2 * r2; }
  return r * r;
}

}  // namespace o3d
