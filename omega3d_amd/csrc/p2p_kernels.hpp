// p2p_kernels.hpp -- point -> point direct sums (velocity and velocity gradient)
// for gfx950.  Included once per core by p2p_<core>.hip.
//
// Math follows Omega3D's pair kernels and cores (Kernels.h:50-112, 155-290;
// CoreFunc.h:43-341); the restructuring for the GPU is:
//   * the skew part of the gradient, sum_j K_j skew(alpha_j), is accumulated as
//     one 3-vector W = sum_j K_j alpha_j and spread into the six off-diagonal
//     entries once per target (3 FMAs per pair instead of 6);
//   * RM / WL / V2 use one v_rsq_f32 per pair; the compact-exponential core
//     evaluates its far branch (q = rho^3/(ss^3+st^3) > 16, CoreFunc.h:121,169)
//     first and enters the exp/near-field code only when some lane of the wave
//     needs it; every lane's result is independent of the other lanes;
//   * for q < 1 the exponential core uses Taylor series of (1-e^-q)/q and of
//     (e^-q - (1-e^-q)/q)/q instead of 1 - exp(-q), which cancels in FP32.
//
// Layout: sources are staged per tile in LDS as two float4 per source,
// {x, y, z, core_term(sigma)} and {ax, ay, az, 0}; every lane of the workgroup
// reads the same source (LDS broadcast, no bank conflicts).  Targets live in
// VGPRs: TPT targets per lane, strided by the workgroup size so that loads and
// stores stay coalesced.
#pragma once
#include "o3d_common.hpp"

namespace o3d {

// ---- compact exponential: the rare path (exp, series and near branch) -------
__device__ __forceinline__ float horner_f(float q) {
  // (1 - e^-q)/q = sum_n (-1)^n q^n/(n+1)!,  n = 0..10   (|trunc| < 3e-9 for q < 1)
  float p = (float)(1.0 / 39916800.0);
  p = fmaf(p, q, (float)(-1.0 / 3628800.0));
  p = fmaf(p, q, (float)(1.0 / 362880.0));
  p = fmaf(p, q, (float)(-1.0 / 40320.0));
  p = fmaf(p, q, (float)(1.0 / 5040.0));
  p = fmaf(p, q, (float)(-1.0 / 720.0));
  p = fmaf(p, q, (float)(1.0 / 120.0));
  p = fmaf(p, q, (float)(-1.0 / 24.0));
  p = fmaf(p, q, (float)(1.0 / 6.0));
  p = fmaf(p, q, -0.5f);
  return fmaf(p, q, 1.0f);
}
__device__ __forceinline__ float horner_h(float q) {
  // (f(q) - e^-q)/q = sum_m (-1)^m q^m (m+1)/(m+2)!,  m = 0..10
  float p = (float)(1.0 / 43545600.0);
  p = fmaf(p, q, (float)(-1.0 / 3991680.0));
  p = fmaf(p, q, (float)(1.0 / 403200.0));
  p = fmaf(p, q, (float)(-1.0 / 45360.0));
  p = fmaf(p, q, (float)(1.0 / 5760.0));
  p = fmaf(p, q, (float)(-1.0 / 840.0));
  p = fmaf(p, q, (float)(1.0 / 144.0));
  p = fmaf(p, q, (float)(-1.0 / 30.0));
  p = fmaf(p, q, 0.125f);
  p = fmaf(p, q, (float)(-1.0 / 3.0));
  return fmaf(p, q, 0.5f);
}

// Every core returns K and Bp = -B/3, where u += K (a x d) and the gradient's
// dyad part is B d (x) (a x d); the factor -3 is applied once per target.
//
// Compact exponential, far branch: q = rho^3/cs > 16  <=>  rho^6 > 256 cs^2, where
// K = 1/rho^3 and B = -3K/rho^2 (CoreFunc.h:121,169).  Elsewhere (CoreFunc.h:124-130,
// 172-178), with c = 1/cs and e = exp(-q):
//   q < 0.001:  K = c,                 B = -1.5 rho c^2
//   q < 1:      K = c f(q),            B = -3 c^2 rho h(q)      (series, no cancellation)
//   otherwise:  K = (1 - e)/rho^3,     B = 3 (c e - K)/rho^2
__device__ __forceinline__ void exp_far(float d2, float& K, float& Bp) {
  const float ir = frsq(d2);
  const float ir2 = ir * ir;
  K = ir2 * ir;
  Bp = K * ir2;
}
__device__ __forceinline__ void exp_any(float d2, float cs, float& K, float& Bp) {
  const bool far = d2 * d2 * d2 > 256.0f * (cs * cs);
  // d2m differs from d2 only below FLT_MIN, where no pair is far: the far values
  // below are bit-identical to exp_far(d2)
  const float d2m = fmaxf(d2, 1.17549435e-38f);
  const float ir = frsq(d2m);
  const float ir2 = ir * ir;
  const float ood3 = ir2 * ir;
  const float kf = ood3, bf = ood3 * ir2;
  const float c = frcp(cs);
  const float dist = d2m * ir;
  const float q = (d2m * dist) * c;
  const float e = fexp2(q * -1.44269504088896341f);
  const float kg = fmaf(-ood3, e, ood3);
  const float bg = fmaf(-c, e, kg) * ir2;
  const bool small = q < 1.0f;
  float kn = kg, bn = bg;
  if (wave_any(small)) {          // the series only where some lane needs them
    const float ks = c * horner_f(q);
    const float bs = c * (c * dist) * horner_h(q);
    kn = small ? ks : kg;
    bn = small ? bs : bg;
  }
  const bool near = q < 0.001f;
  kn = near ? c : kn;
  bn = near ? (0.5f * dist) * c * c : bn;
  K = far ? kf : kn;
  Bp = far ? bf : bn;
}

// K and Bp of one pair from rho^2 (d2) and the pair's core parameter st (sum of the
// two particles' core_term).  thr (EXP only): rho^2 above which every pair of this
// target with the current source tile is in the far branch.
template <int CORE>
__device__ __forceinline__ void kb(float d2, float st, float thr, float& K, float& Bp) {
  if (CORE == CORE_RM) {                       // CoreFunc.h:65-71
    const float r2 = d2 + st;
    const float ir = frsq(r2);
    const float ir2 = ir * ir;
    K = ir2 * ir;
    Bp = K * ir2;
  } else if (CORE == CORE_WL) {                // CoreFunc.h:265-274
    const float dd = d2 + st;
    const float top = fmaf(2.5f, st, d2);
    const float ir = frsq(dd);
    const float ir2 = ir * ir;
    const float dn5 = ir2 * ir2 * ir;
    K = top * dn5;
    Bp = dn5 * fmaf((float)(5.0 / 3.0), top * ir2, (float)(-2.0 / 3.0));
  } else if (CORE == CORE_V2) {                // CoreFunc.h:317-328
    const float den = fmaf(d2, d2, st);
    const float rs = frsq(den);
    K = rs * fsqrt(rs);
    Bp = K * rs;
  } else {                                     // CoreFunc.h:225-236
    if (__builtin_expect(wave_any(!(d2 > thr)), 0)) exp_any(d2, st, K, Bp);
    else exp_far(d2, K, Bp);
  }
}

// acc: [0..2] velocity; [3..10] D(j,i) = sum_j Bp d_j (a x d)_i for
// (j,i) = 00 01 02 10 11 12 20 21 (D22 = -D00 - D11: d . (a x d) = 0);
// [11..13] W = sum_j K_j a_j
constexpr int NACC = 14;

// accumulate one pair given its separation d = t - s and K, Bp.
// VALU operations per pair with the gradient (far branch of EXP, or RM): d 3,
// |d|^2 3, v_rsq 1, K and Bp 3, branch test 1 (EXP), c = alpha x d 6, velocity 3,
// b = Bp c 3, dyad 8, W 3: 34; the LDS reads of the source are shared by the lane's
// targets.
template <bool GRAD>
__device__ __forceinline__ void acc_pair(float dx, float dy, float dz, float K, float Bp, const float4 A, float* acc) {
  const float cx = fmaf(dz, A.y, -dy * A.z);
  const float cy = fmaf(dx, A.z, -dz * A.x);
  const float cz = fmaf(dy, A.x, -dx * A.y);
  acc[0] = fmaf(K, cx, acc[0]);
  acc[1] = fmaf(K, cy, acc[1]);
  acc[2] = fmaf(K, cz, acc[2]);
  if (GRAD) {
    const float bx = Bp * cx, by = Bp * cy, bz = Bp * cz;
    acc[3] = fmaf(dx, bx, acc[3]);
    acc[4] = fmaf(dx, by, acc[4]);
    acc[5] = fmaf(dx, bz, acc[5]);
    acc[6] = fmaf(dy, bx, acc[6]);
    acc[7] = fmaf(dy, by, acc[7]);
    acc[8] = fmaf(dy, bz, acc[8]);
    acc[9] = fmaf(dz, bx, acc[9]);
    acc[10] = fmaf(dz, by, acc[10]);
    acc[11] = fmaf(K, A.x, acc[11]);
    acc[12] = fmaf(K, A.y, acc[12]);
    acc[13] = fmaf(K, A.z, acc[13]);
  }
}

// K and Bp of one pair from its separation d and core parameter st.  RM and WL need
// rho^2 only as rho^2 + st: that sum is the FMA chain of |d|^2 started at st (one
// add fewer per pair than |d|^2 then + st; at d = 0 it is st exactly, the self
// pair's value of kb).  WL's numerator rho^2 + 2.5 st is then (rho^2 + st) + 1.5 st.
template <int CORE>
__device__ __forceinline__ void kb_d(float dx, float dy, float dz, float st, float thr, float& K, float& Bp) {
  if (CORE == CORE_RM || CORE == CORE_WL) {
    const float dd = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, st)));
    const float ir = frsq(dd);
    const float ir2 = ir * ir;
    if (CORE == CORE_RM) {                     // CoreFunc.h:65-71
      K = ir2 * ir;
      Bp = K * ir2;
    } else {                                   // CoreFunc.h:265-274
      const float top = fmaf(1.5f, st, dd);
      const float dn5 = ir2 * ir2 * ir;
      K = top * dn5;
      Bp = dn5 * fmaf((float)(5.0 / 3.0), top * ir2, (float)(-2.0 / 3.0));
    }
  } else {
    kb<CORE>(fmaf(dz, dz, fmaf(dy, dy, dx * dx)), st, thr, K, Bp);
  }
}

// one source against the TPT targets of this lane.  For the exponential core the
// far-branch test of all TPT pairs is combined into a single wave-uniform branch,
// so that the common case is one basic block holding all TPT pair evaluations.
template <int CORE, bool GRAD, int TPT>
__device__ __forceinline__ void source_acc(const float* tx, const float* ty, const float* tz, const float* tt,
                                           const float* thr, const float4 P, const float4 A, float (*acc)[NACC]) {
  float dx[TPT], dy[TPT], dz[TPT], d2[TPT];
#pragma unroll
  for (int k = 0; k < TPT; ++k) {
    dx[k] = tx[k] - P.x;
    dy[k] = ty[k] - P.y;
    dz[k] = tz[k] - P.z;
    d2[k] = fmaf(dz[k], dz[k], fmaf(dy[k], dy[k], dx[k] * dx[k]));
  }
  float K[TPT], Bp[TPT];
  if (CORE == CORE_EXP) {
    bool allfar = true;
#pragma unroll
    for (int k = 0; k < TPT; ++k) allfar = allfar && (d2[k] > thr[k]);
    if (__builtin_expect(wave_any(!allfar), 0)) {
#pragma unroll
      for (int k = 0; k < TPT; ++k) exp_any(d2[k], P.w + tt[k], K[k], Bp[k]);
    } else {
#pragma unroll
      for (int k = 0; k < TPT; ++k) exp_far(d2[k], K[k], Bp[k]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < TPT; ++k) kb_d<CORE>(dx[k], dy[k], dz[k], P.w + tt[k], thr[k], K[k], Bp[k]);
  }
#pragma unroll
  for (int k = 0; k < TPT; ++k) acc_pair<GRAD>(dx[k], dy[k], dz[k], K[k], Bp[k], A, acc[k]);
}

// one source against the lane's targets when every pair of the tile is known to be
// in the exponential core's far branch: exp_far for all, no test (the same values,
// bit for bit, as source_acc gives those pairs)
template <bool GRAD, int TPT>
__device__ __forceinline__ void source_acc_far(const float* tx, const float* ty, const float* tz, const float4 P,
                                               const float4 A, float (*acc)[NACC]) {
#pragma unroll
  for (int k = 0; k < TPT; ++k) {
    const float dx = tx[k] - P.x, dy = ty[k] - P.y, dz = tz[k] - P.z;
    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    float K, Bp;
    exp_far(d2, K, Bp);
    acc_pair<GRAD>(dx, dy, dz, K, Bp, A, acc[k]);
  }
}

// source_acc_far over the staged sources [j0, j1), U at a time.  The LDS address is
// kept in a vector register (made so by an opaque zero from the assembler), so the
// reads of the U sources take it with immediate offsets and one add advances it;
// an address known to be uniform lives in a scalar register and is moved to a
// vector register before every read.
template <bool GRAD, int TPT, int U>
__device__ __forceinline__ void sources_acc_far(const float* tx, const float* ty, const float* tz, const float4* sh,
                                                int j0, int j1, float (*acc)[NACC]) {
  int vz;
  asm volatile("v_mov_b32 %0, 0" : "=v"(vz));
  const float4* q = sh + 2 * j0 + vz;
  int j = j0;
  for (; j + U <= j1; j += U, q += 2 * U) {
#pragma unroll
    for (int u = 0; u < U; ++u) source_acc_far<GRAD, TPT>(tx, ty, tz, q[2 * u], q[2 * u + 1], acc);
  }
  for (; j < j1; ++j, q += 2) source_acc_far<GRAD, TPT>(tx, ty, tz, q[0], q[1], acc);
}

// source_acc over the staged sources [j0, j1), U at a time, with the LDS address in
// a vector register as in sources_acc_far: for the cores whose pair code has no
// branch (RM, WL, V2)
template <int CORE, bool GRAD, int TPT, int U>
__device__ __forceinline__ void sources_acc_u(const float* tx, const float* ty, const float* tz, const float* tt,
                                              const float* thr, const float4* sh, int j0, int j1,
                                              float (*acc)[NACC]) {
  int vz;
  asm volatile("v_mov_b32 %0, 0" : "=v"(vz));
  const float4* q = sh + 2 * j0 + vz;
  int j = j0;
  for (; j + U <= j1; j += U, q += 2 * U) {
#pragma unroll
    for (int u = 0; u < U; ++u) source_acc<CORE, GRAD, TPT>(tx, ty, tz, tt, thr, q[2 * u], q[2 * u + 1], acc);
  }
  for (; j < j1; ++j, q += 2) source_acc<CORE, GRAD, TPT>(tx, ty, tz, tt, thr, q[0], q[1], acc);
}

template <int CORE, bool GRAD>
__device__ __forceinline__ void pair_acc(float tx, float ty, float tz, float tt, float thr, const float4 P,
                                         const float4 A, float* acc) {
  const float dx = tx - P.x;
  const float dy = ty - P.y;
  const float dz = tz - P.z;
  float K, Bp;
  kb_d<CORE>(dx, dy, dz, P.w + tt, thr, K, Bp);
  const float cx = fmaf(dz, A.y, -dy * A.z);
  const float cy = fmaf(dx, A.z, -dz * A.x);
  const float cz = fmaf(dy, A.x, -dx * A.y);
  acc[0] = fmaf(K, cx, acc[0]);
  acc[1] = fmaf(K, cy, acc[1]);
  acc[2] = fmaf(K, cz, acc[2]);
  if (GRAD) {
    const float bx = Bp * cx, by = Bp * cy, bz = Bp * cz;
    acc[3] = fmaf(dx, bx, acc[3]);
    acc[4] = fmaf(dx, by, acc[4]);
    acc[5] = fmaf(dx, bz, acc[5]);
    acc[6] = fmaf(dy, bx, acc[6]);
    acc[7] = fmaf(dy, by, acc[7]);
    acc[8] = fmaf(dy, bz, acc[8]);
    acc[9] = fmaf(dz, bx, acc[9]);
    acc[10] = fmaf(dz, by, acc[10]);
    acc[11] = fmaf(K, A.x, acc[11]);
    acc[12] = fmaf(K, A.y, acc[12]);
    acc[13] = fmaf(K, A.z, acc[13]);
  }
}

// gradient in ug order from the accumulators (Kernels.h:183-192):
// g = -3 D + skew(W), skew: g1 += Wz, g2 -= Wy, g3 -= Wz, g5 += Wx, g6 += Wy, g7 -= Wx
template <class T>
__device__ __forceinline__ void grad_from(const T* a, T* g) {
  g[0] = T(-3) * a[3];
  g[1] = T(-3) * a[4] + a[13];
  g[2] = T(-3) * a[5] - a[12];
  g[3] = T(-3) * a[6] - a[13];
  g[4] = T(-3) * a[7];
  g[5] = T(-3) * a[8] + a[11];
  g[6] = T(-3) * a[9] + a[12];
  g[7] = T(-3) * a[10] - a[11];
  g[8] = T(3) * (a[3] + a[7]);
}

// Compensated accumulation in two parts.  The velocity's FP32 partial of a tile
// goes into a TwoSum (hi, lo) pair.  The gradient's 11 raw FP32 sums (dyad D and
// W, acc[3..13]) go into FP64 sums gr[3..13] every P2P_GFOLD sources: a close
// source's large dyad term sets the rounding of every later FP32 addition until
// the next fold, and the terms of a target's nearest sources cancel, so the FP32
// partials are kept short (at a 512-source tile the gradient of a thin ring was
// 4e-6 from FP64, at 64 sources 2e-7).  The gradient is formed from the FP64 sums
// once per target (grad_from).
constexpr int P2P_GFOLD = 64;
// sources per gradient fold in a tile known to be far from every target of the
// wave (sorted sources only): no term of it is a nearest source's
constexpr int P2P_GFOLD_FAR = 512;
constexpr int P2P_FAR_UNROLL = 4;   // sources per step of the far-tile loop (two rings, compensated: 704.5 ms at 1, 681.9 at 2, 680.5 at 4)
__device__ __forceinline__ void fold_vel_comp(float* acc, float* hi, float* lo) {
  two_sum(hi[0], lo[0], acc[0]);
  two_sum(hi[1], lo[1], acc[1]);
  two_sum(hi[2], lo[2], acc[2]);
  acc[0] = acc[1] = acc[2] = 0.0f;
}
__device__ __forceinline__ void fold_grad_raw(float* acc, double* gr) {
#pragma unroll
  for (int c = 3; c < NACC; ++c) {
    gr[c] += (double)acc[c];
    acc[c] = 0.0f;
  }
}

// far-branch threshold of the exponential core for one target and one source
// tile: rho^2 > (256 cs_max^2)^(1/3) (1 + 2^-16) implies rho^6 > 256 cs^2 for every
// source of the tile (cs <= cs_max); the margin covers the log2/exp2 error.
__device__ __forceinline__ float exp_threshold(float csmax) {
  const float l = __builtin_amdgcn_logf(256.0f * (csmax * csmax));
  return fexp2(l * (float)(1.0 / 3.0)) * 1.0000153f;
}

// stage sources [s0, s0+tile) into LDS as {x,y,z,term} {ax,ay,az,0}; *tmax
// (LDS) receives the largest core term of the tile (float bits compare as ints)
template <int CORE>
__device__ __forceinline__ void stage_tile(const SrcView& s, int64_t s0, int cnt, int tile, float4* sh,
                                           unsigned int* tmax) {
  unsigned int mx = 0u;
  for (int j = threadIdx.x; j < tile; j += BLOCK) {
    float4 P = make_float4(0.f, 0.f, 0.f, 1.f), A = make_float4(0.f, 0.f, 0.f, 0.f);
    if (j < cnt) {
      const int64_t g = s0 + j;
      P = make_float4(s.x[g], s.y[g], s.z[g], core_term<CORE>(s.r[g]));
      A = make_float4(s.ax[g], s.ay[g], s.az[g], 0.f);
      mx = max(mx, __float_as_uint(fmaxf(P.w, 0.0f)));
    }
    sh[2 * j] = P;
    sh[2 * j + 1] = A;
  }
  if (CORE == CORE_EXP) atomicMax(tmax, mx);
}

// stage sorted rows [s0, s0 + cnt) of sp/sa into LDS (pads: zero strength, unit core)
__device__ __forceinline__ void stage_sorted(const float4* sp, const float4* sa, int64_t s0, int cnt, int tile,
                                             float4* sh) {
  for (int j = threadIdx.x; j < tile; j += BLOCK) {
    float4 P = make_float4(0.f, 0.f, 0.f, 1.f), A = make_float4(0.f, 0.f, 0.f, 0.f);
    if (j < cnt) {
      P = sp[s0 + j];
      A = sa[s0 + j];
    }
    sh[2 * j] = P;
    sh[2 * j + 1] = A;
  }
}

// what p2p_span does with a chunk's sums: overwrite the outputs, add into them (the
// reference's semantics), or store them as one slice of a source-split evaluation
// (no self term; k_p2p_combine adds the slices in order and finishes).
enum { OUT_SET = 0, OUT_ADD = 1, OUT_PART = 2 };

// partial sums per target and slice: FP32 keeps the raw accumulators (NACC floats,
// 3 without the gradient), compensated keeps each velocity component's (hi, lo)
// FP32 pair in one 8-byte column and the FP64 gradient (12 or 3 columns of 8
// bytes).  Layout [slice][column][target].
template <bool GRAD, int PREC>
struct PartCols {
  static constexpr int n = PREC == PREC_COMP ? (GRAD ? 12 : 3) : (GRAD ? NACC : 3);
};

// One chunk of BLOCK*TPT targets [t_base, t_base + BLOCK*TPT) against sources
// [s_lo, s_hi) of (s, t), in tiles that start at s_lo.  Shared by the
// single-collection, batched and serving kernels so that a target's result does not
// depend on which launch computed it.  part/slice: OUT_PART's destination.
//
// ssp, ssa, stsph, stcs (EXP core only; s_lo a multiple of tile): the sources come from rows sorted in the
// targets' Morton order with tile spheres, and a tile whose sphere is beyond the
// far-branch distance of every target of the wave is summed with no per-pair test.
template <int CORE, bool THICK, bool GRAD, int PREC, int TPT, int OUT>
__device__ __forceinline__ void p2p_span(const SrcView& s, const TgtView& t, int64_t s_lo, int64_t s_hi, int64_t nt,
                                         int64_t t_base, int tile, bool self_excl, float4* sh, void* part,
                                         int64_t slice, const int32_t* tperm = nullptr,
                                         const float4* ssp = nullptr, const float4* ssa = nullptr,
                                         const float4* stsph = nullptr, const float* stcs = nullptr) {
  __shared__ unsigned int s_tmax;
  const int tid = threadIdx.x;
  float tx[TPT], ty[TPT], tz[TPT], tt[TPT];
  int64_t ti[TPT], tix[TPT];
#pragma unroll
  for (int k = 0; k < TPT; ++k) {
    const int64_t i = t_base + tid + (int64_t)k * BLOCK;
    ti[k] = i;
    const int64_t ic = i < nt ? i : nt - 1;    // idle lanes mirror a real target
    const int64_t ix = tperm ? (int64_t)tperm[ic] : ic;   // the target's own index
    tix[k] = ix;
    tx[k] = t.x[ix]; ty[k] = t.y[ix]; tz[k] = t.z[ix];
    tt[k] = THICK ? core_term<CORE>(t.r[ix]) : 0.0f;
  }
  float acc[TPT][NACC];
  float hi[TPT][3], lo[TPT][3];
  double gr[TPT][NACC];    // compensated: FP64 sums of the raw gradient accumulators [3..13]
#pragma unroll
  for (int k = 0; k < TPT; ++k) {
#pragma unroll
    for (int c = 0; c < NACC; ++c) { acc[k][c] = 0.0f; gr[k][c] = 0.0; }
#pragma unroll
    for (int c = 0; c < 3; ++c) { hi[k][c] = 0.0f; lo[k][c] = 0.0f; }
  }

  const bool srt = CORE == CORE_EXP && ssp != nullptr;
  for (int64_t s0 = s_lo; s0 < s_hi; s0 += tile) {
    const int cnt = (int)((s_hi - s0) < tile ? (s_hi - s0) : tile);
    __syncthreads();
    if (CORE == CORE_EXP && tid == 0) s_tmax = 0u;
    __syncthreads();
    if (srt) stage_sorted(ssp, ssa, s0, cnt, tile, sh);
    else stage_tile<CORE>(s, s0, cnt, tile, sh, &s_tmax);
    __syncthreads();
    const int64_t ti = srt ? s0 / tile : 0;   // s_lo is a whole number of tiles
    const float tmax = srt ? stcs[ti] : __uint_as_float(s_tmax);
    float thr[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) thr[k] = CORE == CORE_EXP ? exp_threshold(tmax + tt[k]) : 0.0f;
    bool far_tile = false;
    if (srt) {
      // every pair of the tile is farther than the tile sphere's gap; with a margin
      // of 1e-5 on the squared gap the exact far test (rho^6 > 256 cs^2) holds for
      // each of them, so they take the far formulas in any case
      const float4 c = stsph[ti];
      bool all = true;
#pragma unroll
      for (int k = 0; k < TPT; ++k) {
        const float ex = tx[k] - c.x, ey = ty[k] - c.y, ez = tz[k] - c.z;
        const float gap = __fsqrt_rn(fmaf(ez, ez, fmaf(ey, ey, ex * ex))) - c.w;
        all = all && gap > 0.0f && gap * gap > thr[k] * 1.00001f;
      }
      far_tile = wave_all(all);
    }
    if (PREC == PREC_COMP) {
      const int fold = far_tile ? P2P_GFOLD_FAR : P2P_GFOLD;
      for (int j0 = 0; j0 < cnt; j0 += fold) {
        const int j1 = j0 + fold < cnt ? j0 + fold : cnt;
        if (far_tile) {
          sources_acc_far<GRAD, TPT, P2P_FAR_UNROLL>(tx, ty, tz, sh, j0, j1, acc);
        } else {
          if (CORE != CORE_EXP && TPT == 1) {
            sources_acc_u<CORE, GRAD, TPT, P2P_FAR_UNROLL>(tx, ty, tz, tt, thr, sh, j0, j1, acc);
          } else {
            for (int j = j0; j < j1; ++j) {
              const float4 P = sh[2 * j];
              const float4 A = sh[2 * j + 1];
              source_acc<CORE, GRAD, TPT>(tx, ty, tz, tt, thr, P, A, acc);
            }
          }
          // a tile that may hold a target's nearest sources (in sorted order they
          // gather in few tiles): its velocity partials are folded as often
#pragma unroll
          for (int k = 0; k < TPT; ++k) fold_vel_comp(acc[k], hi[k], lo[k]);
        }
        if (GRAD) {
#pragma unroll
          for (int k = 0; k < TPT; ++k) fold_grad_raw(acc[k], gr[k]);
        }
      }
    } else if (far_tile) {
      sources_acc_far<GRAD, TPT, P2P_FAR_UNROLL>(tx, ty, tz, sh, 0, cnt, acc);
    } else if (CORE != CORE_EXP && TPT == 1) {
      sources_acc_u<CORE, GRAD, TPT, P2P_FAR_UNROLL>(tx, ty, tz, tt, thr, sh, 0, cnt, acc);
    } else {
      for (int j = 0; j < cnt; ++j) {
        const float4 P = sh[2 * j];
        const float4 A = sh[2 * j + 1];
        source_acc<CORE, GRAD, TPT>(tx, ty, tz, tt, thr, P, A, acc);
      }
    }
    if (PREC == PREC_COMP) {
#pragma unroll
      for (int k = 0; k < TPT; ++k) fold_vel_comp(acc[k], hi[k], lo[k]);
    }
  }
  // compensated: the gradient in ug order from the FP64 sums
  double gd[TPT][9];
#pragma unroll
  for (int k = 0; k < TPT; ++k) {
    if (PREC == PREC_COMP && GRAD) grad_from<double>(gr[k], gd[k]);
    else {
#pragma unroll
      for (int c = 0; c < 9; ++c) gd[k][c] = 0.0;
    }
  }

  if (OUT == OUT_PART) {
    constexpr int NC = PartCols<GRAD, PREC>::n;
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
      const int64_t i = ti[k];
      if (i >= nt) continue;
      if (PREC == PREC_FP32) {
        float* pf = (float*)part + slice * NC * nt + i;
#pragma unroll
        for (int c = 0; c < NC; ++c) pf[c * nt] = acc[k][c];
      } else {
        double* pd = (double*)part + slice * NC * nt + i;
#pragma unroll
        for (int c = 0; c < 3; ++c) reinterpret_cast<float2*>(pd)[c * nt] = make_float2(hi[k][c], lo[k][c]);
        if (GRAD) {
#pragma unroll
          for (int c = 0; c < 9; ++c) pd[(3 + c) * nt] = gd[k][c];
        }
      }
    }
    return;
  }

  // self pair j == i: velocity and dyad parts vanish at d = 0; only K(0) alpha_i
  // entered W.  Evaluated with the same code as in the loop.
  float ks[TPT], sax[TPT], say[TPT], saz[TPT];
#pragma unroll
  for (int k = 0; k < TPT; ++k) {
    ks[k] = 0.0f; sax[k] = say[k] = saz[k] = 0.0f;
    if (GRAD && self_excl) {
      const int64_t ic = tix[k];
      float Bp;
      kb<CORE>(0.0f, core_term<CORE>(s.r[ic]) + tt[k], 0.0f, ks[k], Bp);
      sax[k] = s.ax[ic]; say[k] = s.ay[ic]; saz[k] = s.az[ic];
    }
  }

#pragma unroll
  for (int k = 0; k < TPT; ++k) {
    if (ti[k] >= nt) continue;
    const int64_t i = tix[k];
    if (PREC == PREC_FP32) {
      float r[12];
      r[0] = acc[k][0];
      r[1] = acc[k][1];
      r[2] = acc[k][2];
      if (GRAD) {
        acc[k][11] = fmaf(-ks[k], sax[k], acc[k][11]);
        acc[k][12] = fmaf(-ks[k], say[k], acc[k][12]);
        acc[k][13] = fmaf(-ks[k], saz[k], acc[k][13]);
        grad_from<float>(acc[k], r + 3);
      }
      if (OUT == OUT_ADD) {
        t.u[i] += r[0]; t.v[i] += r[1]; t.w[i] += r[2];
        if (GRAD) {
#pragma unroll
          for (int c = 0; c < 9; ++c) t.g[c][i] += r[3 + c];
        }
      } else {
        t.u[i] = r[0]; t.v[i] = r[1]; t.w[i] = r[2];
        if (GRAD) {
#pragma unroll
          for (int c = 0; c < 9; ++c) t.g[c][i] = r[3 + c];
        }
      }
    } else {
      double r[12];
      if (GRAD) {
        const double kx = (double)ks[k] * sax[k], ky = (double)ks[k] * say[k], kz = (double)ks[k] * saz[k];
        r[3] = gd[k][0];
        r[4] = gd[k][1] - kz;
        r[5] = gd[k][2] + ky;
        r[6] = gd[k][3] + kz;
        r[7] = gd[k][4];
        r[8] = gd[k][5] - kx;
        r[9] = gd[k][6] - ky;
        r[10] = gd[k][7] + kx;
        r[11] = gd[k][8];
      }
      // velocity: the TwoSum pair added in FP32; gradient: its FP64 reduction rounded once
      if (OUT == OUT_ADD) {
        t.u[i] = add_comp(t.u[i], hi[k][0], lo[k][0]);
        t.v[i] = add_comp(t.v[i], hi[k][1], lo[k][1]);
        t.w[i] = add_comp(t.w[i], hi[k][2], lo[k][2]);
        if (GRAD) {
#pragma unroll
          for (int c = 0; c < 9; ++c) t.g[c][i] = (float)((double)t.g[c][i] + r[3 + c]);
        }
      } else {
        t.u[i] = hi[k][0] + lo[k][0]; t.v[i] = hi[k][1] + lo[k][1]; t.w[i] = hi[k][2] + lo[k][2];
        if (GRAD) {
#pragma unroll
          for (int c = 0; c < 9; ++c) t.g[c][i] = (float)r[3 + c];
        }
      }
    }
  }
}

// targets [t_base, t_base + BLOCK*TPT) against all ns sources.  ACC: add into the
// outputs (the reference's semantics); false: overwrite them.
template <int CORE, bool THICK, bool GRAD, int PREC, int TPT, bool ACC = true>
__device__ __forceinline__ void p2p_chunk(const SrcView& s, const TgtView& t, int64_t ns, int64_t nt,
                                          int64_t t_base, int tile, bool self_excl, float4* sh) {
  p2p_span<CORE, THICK, GRAD, PREC, TPT, ACC ? OUT_ADD : OUT_SET>(s, t, 0, ns, nt, t_base, tile, self_excl, sh,
                                                                  nullptr, 0);
}

// a.tperm (optional): slot i of the launch evaluates target a.tperm[i]; a target's
// sums do not depend on the slot, so any order gives the same bits
// one target per lane: at least 6 waves per SIMD (at most 80 VGPRs).  Measured on
// the two-ring workload (compensated, sorted sources): 6 waves 704.5 ms, 7 waves
// (72 VGPRs, a few values of the rare exponential branch in scratch) 710, 8 waves
// 709.  More targets per lane need their registers.
template <int CORE, bool THICK, bool GRAD, int PREC, int TPT>
__global__ __launch_bounds__(BLOCK, TPT == 1 ? 6 : 1) void k_p2p(P2PArgs a) {
  extern __shared__ float4 sh[];
  const int64_t t_base = (int64_t)blockIdx.x * (BLOCK * TPT);
  p2p_span<CORE, THICK, GRAD, PREC, TPT, OUT_ADD>(a.s, a.t, 0, a.ns, a.nt, t_base, a.tile, a.self_excl != 0, sh,
                                                  nullptr, 0, a.tperm, a.sp, a.sa, a.tsph, a.tcs);
}

// ---------------------------------------------------------------------------
// Source split for calls with too few targets to fill the GPU: blockIdx.y takes
// sources [y*slen, (y+1)*slen), and its sums go to slice y of part.
// ---------------------------------------------------------------------------
template <int CORE, bool THICK, bool GRAD, int PREC>
__global__ __launch_bounds__(BLOCK) void k_p2p_split(P2PArgs a, void* part, int64_t slen) {
  extern __shared__ float4 sh[];
  const int64_t t_base = (int64_t)blockIdx.x * BLOCK;
  const int64_t lo = (int64_t)blockIdx.y * slen;
  const int64_t hi = lo + slen < a.ns ? lo + slen : a.ns;
  p2p_span<CORE, THICK, GRAD, PREC, 1, OUT_PART>(a.s, a.t, lo, hi, a.nt, t_base, a.tile, false, sh, part,
                                                 blockIdx.y, a.tperm, a.sp, a.sa, a.tsph, a.tcs);
}

// adds the nslice slices of each target in slice order -- velocity as FP32 TwoSum
// pairs, the gradient's sums in FP64 (its cross-slice reduction) -- takes out the
// self pair as p2p_span does, forms the gradient and adds into the outputs; with
// a.tperm the slices are in that order and the outputs go to the original index.
// A block takes COMBINE_TG targets, one per lane of each wave: wave w sums columns
// w, w + COMBINE_NW, ..., each over every slice in order (so a column's sum is the
// same chain of operations whatever wave runs it), loading COMBINE_CH slices of its
// columns before adding them; the first wave finishes the targets from the column
// sums in LDS.
constexpr int COMBINE_TG = 64, COMBINE_NW = BLOCK / COMBINE_TG, COMBINE_CH = 8;
// the columns W, W + COMBINE_NW, ... below NC of one target: slices [y0, y0 + ny)
// loaded, then added in order
template <int NC, int PREC, int W, int CPT, bool FULL, typename Elem>
__device__ __forceinline__ void combine_chunk(const Elem* pe, int64_t nt, int y0, int ny, float* h, float* l,
                                              double* sm) {
  Elem v[CPT][COMBINE_CH];
#pragma unroll
  for (int m = 0; m < CPT; ++m)
#pragma unroll
    for (int k = 0; k < COMBINE_CH; ++k)
      if (FULL || k < ny) v[m][k] = pe[((int64_t)(y0 + k) * NC + W + m * COMBINE_NW) * nt];
#pragma unroll
  for (int m = 0; m < CPT; ++m) {
    constexpr int c0 = W;
    const int c = c0 + m * COMBINE_NW;
#pragma unroll
    for (int k = 0; k < COMBINE_CH; ++k) {
      if (!FULL && k >= ny) break;
      if (c < 3) {
        if (PREC == PREC_FP32) {
          two_sum(h[m], l[m], (float)v[m][k]);
        } else {
          const float2 hl = __builtin_bit_cast(float2, (double)v[m][k]);
          two_sum(h[m], l[m], hl.x);
          l[m] += hl.y;
        }
      } else {
        sm[m] += (double)v[m][k];
      }
    }
  }
}
// one wave's share of k_p2p_combine: the sums of its columns into LDS
template <int NC, int PREC, int W>
__device__ __forceinline__ void combine_wave(const void* part, int64_t nt, int64_t i, int nslice, int lane,
                                             double (*cs)[COMBINE_TG], float2 (*vs)[COMBINE_TG]) {
  constexpr int CPT = W < NC ? (NC - W + COMBINE_NW - 1) / COMBINE_NW : 0;
  if constexpr (CPT > 0) {
    using Elem = typename std::conditional<PREC == PREC_FP32, float, double>::type;
    const Elem* pe = (const Elem*)part + i;
    float h[CPT], l[CPT];
    double sm[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) { h[m] = 0.f; l[m] = 0.f; sm[m] = 0.0; }
    int y0 = 0;
    for (; y0 + COMBINE_CH <= nslice; y0 += COMBINE_CH)
      combine_chunk<NC, PREC, W, CPT, true>(pe, nt, y0, COMBINE_CH, h, l, sm);
    if (y0 < nslice) combine_chunk<NC, PREC, W, CPT, false>(pe, nt, y0, nslice - y0, h, l, sm);
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int c = W + m * COMBINE_NW;
      if (c < 3) vs[c][lane] = make_float2(h[m], l[m]);
      else cs[c][lane] = sm[m];
    }
  }
}
template <int CORE, bool THICK, bool GRAD, int PREC>
__global__ __launch_bounds__(BLOCK) void k_p2p_combine(P2PArgs a, const void* part, int nslice) {
  constexpr int NC = PartCols<GRAD, PREC>::n;
  static_assert(COMBINE_NW == 4, "one column group per wave of the block");
  __shared__ double cs[NC][COMBINE_TG];
  __shared__ float2 vs[3][COMBINE_TG];
  const int64_t nt = a.nt;
  const int lane = threadIdx.x % COMBINE_TG;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / COMBINE_TG));
  const int64_t i = (int64_t)blockIdx.x * COMBINE_TG + lane;
  if (i < nt) {
    if (w == 0) combine_wave<NC, PREC, 0>(part, nt, i, nslice, lane, cs, vs);
    else if (w == 1) combine_wave<NC, PREC, 1>(part, nt, i, nslice, lane, cs, vs);
    else if (w == 2) combine_wave<NC, PREC, 2>(part, nt, i, nslice, lane, cs, vs);
    else combine_wave<NC, PREC, 3>(part, nt, i, nslice, lane, cs, vs);
  }
  __syncthreads();
  if (w != 0 || i >= nt) return;
  const int64_t io = a.tperm ? (int64_t)a.tperm[i] : i;
  float vh[3], vl[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) { vh[c] = vs[c][lane].x; vl[c] = vs[c][lane].y; }
  double sum[NC];
#pragma unroll
  for (int c = 3; c < NC; ++c) sum[c] = cs[c][lane];
  double r[12];
  if constexpr (GRAD) {
    double kx = 0.0, ky = 0.0, kz = 0.0;
    if (a.self_excl) {
      float ks, Bp;
      kb<CORE>(0.0f, core_term<CORE>(a.s.r[io]) + (THICK ? core_term<CORE>(a.t.r[io]) : 0.0f), 0.0f, ks, Bp);
      kx = (double)ks * a.s.ax[io]; ky = (double)ks * a.s.ay[io]; kz = (double)ks * a.s.az[io];
    }
    if constexpr (PREC == PREC_FP32) {
      sum[11] -= kx; sum[12] -= ky; sum[13] -= kz;
      grad_from<double>(sum, r + 3);
    } else {
      r[3] = sum[3];
      r[4] = sum[4] - kz;
      r[5] = sum[5] + ky;
      r[6] = sum[6] + kz;
      r[7] = sum[7];
      r[8] = sum[8] - kx;
      r[9] = sum[9] - ky;
      r[10] = sum[10] + kx;
      r[11] = sum[11];
    }
  }
  a.t.u[io] = add_comp(a.t.u[io], vh[0], vl[0]);
  a.t.v[io] = add_comp(a.t.v[io], vh[1], vl[1]);
  a.t.w[io] = add_comp(a.t.w[io], vh[2], vl[2]);
  if (GRAD) {
#pragma unroll
    for (int c = 0; c < 9; ++c) a.t.g[c][io] = (float)((double)a.t.g[c][io] + r[3 + c]);
  }
}

// ===========================================================================
// FP64 oracle kernel: the reference's formulas (CoreFunc.h, Kernels.h) with
// every operation in double, including its branches; double arrays in and out.
// ===========================================================================

// r2/d2 etc. exactly as CoreFunc.h orders them; sr, tr are the two core radii
template <int CORE, bool THICK, bool GRAD>
__device__ __forceinline__ void kb64(double d2, double sr, double tr, double& K, double& B) {
  if (CORE == CORE_RM) {
    const double r2 = THICK ? d2 + sr * sr + tr * tr : d2 + sr * sr;
    K = 1.0 / (r2 * sqrt(r2));
    B = -3.0 * K * (1.0 / r2);
  } else if (CORE == CORE_WL) {
    const double r2 = THICK ? sr * sr + tr * tr : sr * sr;
    const double dd = d2 + r2;
    const double top = d2 + 2.5 * r2;
    const double dn5 = 1.0 / (dd * dd * sqrt(dd));
    K = top * dn5;
    B = 2.0 * dn5 - 5.0 * top * dn5 / dd;
  } else if (CORE == CORE_V2) {
    const double s2 = sr * sr, t2 = tr * tr;
    const double den = THICK ? d2 * d2 + s2 * s2 + t2 * t2 : d2 * d2 + s2 * s2;
    const double q = sqrt(den);
    K = 1.0 / (q * sqrt(q));
    B = -3.0 * K * (1.0 / sqrt(den));
  } else {
    const double dist = sqrt(d2);
    const double cf = 1.0 / (THICK ? sr * sr * sr + tr * tr * tr : sr * sr * sr);
    if (GRAD) {
      const double d3 = d2 * dist;
      const double q = d3 * cf;
      const double ood3 = 1.0 / d3;
      if (q > 16.0) { K = ood3; B = -3.0 * K / d2; }
      else if (q < (double)0.001f) { K = cf; B = -1.5 * dist * K * K; }
      else { const double e = exp(-q); K = ood3 * (1.0 - e); B = 3.0 * (cf * e - K) / d2; }
    } else {
      const double ood3 = 1.0 / (d2 * dist);
      const double q = cf / ood3;
      B = 0.0;
      if (q > 16.0) K = ood3;
      else if (q < (double)0.001f) K = cf;
      else K = ood3 * (1.0 - exp(-q));
    }
  }
}

template <int CORE, bool THICK, bool GRAD>
__global__ __launch_bounds__(BLOCK) void k_p2p64(P2PArgs64 a) {
  extern __shared__ double4 sh64[];
  const int tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * BLOCK + tid;
  const int64_t ic = i < a.nt ? i : a.nt - 1;
  const double tx = a.t.x[ic], ty = a.t.y[ic], tz = a.t.z[ic];
  const double tr = THICK ? a.t.r[ic] : 0.0;
  double acc[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) acc[c] = 0.0;
  for (int64_t s0 = 0; s0 < a.ns; s0 += a.tile) {
    const int cnt = (int)((a.ns - s0) < a.tile ? (a.ns - s0) : a.tile);
    __syncthreads();
    for (int j = tid; j < cnt; j += BLOCK) {
      const int64_t g = s0 + j;
      sh64[2 * j] = make_double4(a.s.x[g], a.s.y[g], a.s.z[g], a.s.r[g]);
      sh64[2 * j + 1] = make_double4(a.s.ax[g], a.s.ay[g], a.s.az[g], 0.0);
    }
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      if (a.self_excl && s0 + j == i) continue;
      const double4 P = sh64[2 * j];
      const double4 A = sh64[2 * j + 1];
      const double dx = tx - P.x, dy = ty - P.y, dz = tz - P.z;
      const double d2 = dx * dx + dy * dy + dz * dz;
      double K, B;
      kb64<CORE, THICK, GRAD>(d2, P.w, tr, K, B);
      double cx = dz * A.y - dy * A.z;
      double cy = dx * A.z - dz * A.x;
      double cz = dy * A.x - dx * A.y;
      acc[0] += K * cx; acc[1] += K * cy; acc[2] += K * cz;
      if (GRAD) {
        cx *= B; cy *= B; cz *= B;
        acc[3] += dx * cx;
        acc[4] += dx * cy + A.z * K;
        acc[5] += dx * cz - A.y * K;
        acc[6] += dy * cx - A.z * K;
        acc[7] += dy * cy;
        acc[8] += dy * cz + A.x * K;
        acc[9] += dz * cx + A.y * K;
        acc[10] += dz * cy - A.x * K;
        acc[11] += dz * cz;
      }
    }
  }
  if (i < a.nt) {
    a.t.u[i] += acc[0]; a.t.v[i] += acc[1]; a.t.w[i] += acc[2];
    if (GRAD) {
#pragma unroll
      for (int c = 0; c < 9; ++c) a.t.g[c][i] += acc[3 + c];
    }
  }
}

// ---------------------------------------------------------------------------
// launchers (one translation unit per core instantiates them)
// ---------------------------------------------------------------------------
template <int CORE, bool THICK, bool GRAD, int PREC>
hipError_t launch_p2p_prec(const P2PArgs& a, int tpt, hipStream_t st) {
  const size_t lds = (size_t)a.tile * 2 * sizeof(float4);
  const int64_t per = (int64_t)BLOCK * tpt;
  const dim3 grid((unsigned)((a.nt + per - 1) / per));
  if (tpt == 1) hipLaunchKernelGGL((k_p2p<CORE, THICK, GRAD, PREC, 1>), grid, dim3(BLOCK), lds, st, a);
  else if (tpt == 2) hipLaunchKernelGGL((k_p2p<CORE, THICK, GRAD, PREC, 2>), grid, dim3(BLOCK), lds, st, a);
  else if (tpt == 4) hipLaunchKernelGGL((k_p2p<CORE, THICK, GRAD, PREC, 4>), grid, dim3(BLOCK), lds, st, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int CORE>
hipError_t launch_p2p(const P2PArgs& a, bool thick, bool grad, int prec, int tpt, hipStream_t st) {
  if (a.nt <= 0 || a.ns <= 0) return hipSuccess;
#define O3D_L(TH, GR)                                                                        \
  if (thick == TH && grad == GR) {                                                           \
    if (prec == PREC_FP32) return launch_p2p_prec<CORE, TH, GR, PREC_FP32>(a, tpt, st);     \
    if (prec == PREC_COMP) return launch_p2p_prec<CORE, TH, GR, PREC_COMP>(a, tpt, st);     \
    return hipErrorInvalidValue;                                                             \
  }
  O3D_L(true, true) O3D_L(true, false) O3D_L(false, true) O3D_L(false, false)
#undef O3D_L
  return hipErrorInvalidValue;
}

// the source split: nslice slices of slen sources each, part holding
// nslice * nt * PartCols columns
template <int CORE, bool THICK, bool GRAD, int PREC>
hipError_t launch_p2p_split_prec(const P2PArgs& a, int nslice, int64_t slen, void* part, hipStream_t st) {
  const size_t lds = (size_t)a.tile * 2 * sizeof(float4);
  const unsigned nb = (unsigned)((a.nt + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL((k_p2p_split<CORE, THICK, GRAD, PREC>), dim3(nb, (unsigned)nslice), dim3(BLOCK), lds, st, a,
                     part, slen);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const unsigned nbc = (unsigned)((a.nt + COMBINE_TG - 1) / COMBINE_TG);
  hipLaunchKernelGGL((k_p2p_combine<CORE, THICK, GRAD, PREC>), dim3(nbc), dim3(BLOCK), 0, st, a, (const void*)part,
                     nslice);
  return hipGetLastError();
}

template <int CORE>
hipError_t launch_p2p_split(const P2PArgs& a, bool thick, bool grad, int prec, int nslice, int64_t slen, void* part,
                            hipStream_t st) {
  if (a.nt <= 0 || a.ns <= 0) return hipSuccess;
  if (nslice < 1 || nslice > 65535 || slen < 1 || (int64_t)nslice * slen < a.ns || (int64_t)(nslice - 1) * slen >= a.ns)
    return hipErrorInvalidValue;
#define O3D_L(TH, GR)                                                                                 \
  if (thick == TH && grad == GR) {                                                                    \
    if (prec == PREC_FP32) return launch_p2p_split_prec<CORE, TH, GR, PREC_FP32>(a, nslice, slen, part, st); \
    if (prec == PREC_COMP) return launch_p2p_split_prec<CORE, TH, GR, PREC_COMP>(a, nslice, slen, part, st); \
    return hipErrorInvalidValue;                                                                      \
  }
  O3D_L(true, true) O3D_L(true, false) O3D_L(false, true) O3D_L(false, false)
#undef O3D_L
  return hipErrorInvalidValue;
}

template <int CORE>
hipError_t launch_p2p64(const P2PArgs64& a, bool thick, bool grad, hipStream_t st) {
  if (a.nt <= 0 || a.ns <= 0) return hipSuccess;
  const size_t lds = (size_t)a.tile * 2 * sizeof(double4);
  const dim3 grid((unsigned)((a.nt + BLOCK - 1) / BLOCK));
  if (thick && grad) hipLaunchKernelGGL((k_p2p64<CORE, true, true>), grid, dim3(BLOCK), lds, st, a);
  else if (thick) hipLaunchKernelGGL((k_p2p64<CORE, true, false>), grid, dim3(BLOCK), lds, st, a);
  else if (grad) hipLaunchKernelGGL((k_p2p64<CORE, false, true>), grid, dim3(BLOCK), lds, st, a);
  else hipLaunchKernelGGL((k_p2p64<CORE, false, false>), grid, dim3(BLOCK), lds, st, a);
  return hipGetLastError();
}

}  // namespace o3d

#define O3D_INSTANTIATE_P2P(CORE, NAME)                                                        \
  namespace o3d {                                                                              \
  hipError_t launch_p2p_##NAME(const P2PArgs& a, bool thick, bool grad, int prec, int tpt,     \
                               hipStream_t st) {                                               \
    return launch_p2p<CORE>(a, thick, grad, prec, tpt, st);                                    \
  }                                                                                            \
  hipError_t launch_p2p64_##NAME(const P2PArgs64& a, bool thick, bool grad, hipStream_t st) {  \
    return launch_p2p64<CORE>(a, thick, grad, st);                                             \
  }                                                                                            \
  hipError_t launch_p2p_split_##NAME(const P2PArgs& a, bool thick, bool grad, int prec, int nslice, \
                                     int64_t slen, void* part, hipStream_t st) {               \
    return launch_p2p_split<CORE>(a, thick, grad, prec, nslice, slen, part, st);               \
  }                                                                                            \
  }
