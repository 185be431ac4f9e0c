// ff_kernels.hpp -- factored far field for the point->point sums with velocity
// gradient, Rosenhead-Moore and compact-exponential cores (CoreFunc.h:65-71,
// 225-236), on the gfx950 matrix cores.
//
// The sums are the same direct sums over every pair (Kernels.h:155-193); only
// their arithmetic is reorganised for pairs of a target block and a source tile
// that are well separated.  Relative to the block centre c, with t' = t - c,
// s' = s - c, d = t' - s', m = alpha x s', and
//   RM:          r2 = |d|^2 + ss^2 + st^2,   K = r2^-3/2,  Bp = r2^-5/2
//   EXP (q>16):  r2 = |d|^2,                 K = r2^-3/2,  Bp = r2^-5/2
// every pair has K = Bp r2 = Bp (a - 2 t'.s' + b) with a = |t'|^2 + st^2 and
// b = |s'|^2 + ss^2 (no core terms for EXP).  Hence, summing over the tile,
//   W   = sum K alpha     = a P - 2 t'_l R_l + U
//   M   = sum K m         = a Q - 2 t'_l S_l + V
//   u   = sum K alpha x d = W x t' - M
//   D_ji = sum Bp d_j (alpha x d)_i = t'_j (P x t')_i - t'_j Q_i - (R_j x t')_i + S_ji
// with 29 moment sums, all of the form sum_k Bp_k f(source k):
//   P = sum Bp alpha (3), Q = sum Bp m (3), R_lk = sum Bp s'_l alpha_k (9),
//   S_lk = sum Bp s'_l m_k (8: s'.m = 0 so S_22 = -S_00 - S_11),
//   U = sum Bp b alpha (3), V = sum Bp b m (3).
// That is a product of the pair matrix Bp (sources x targets) with a per-source
// moment table (29 columns padded to 32): one v_mfma_f32_32x32x2f32 per two
// sources and 32 targets, leaving about ten VALU operations per pair (d, r2,
// one v_rsq_f32 and three products) instead of about thirty-three.
//
// Separation: a tile is "far" for a block when the gap between their bounding
// spheres is at least one block radius (then every term above is at most
// (1+1/1)^2 = 4 times the pair's own contribution; measured on the two-ring
// workload, gaps of 1 and 2 radii give the same error, 1.5e-7, against FP64,
// and 0.5 radii 2.1e-7, with hardly fewer near pairs) and,
// for EXP, when that gap guarantees q = rho^3/(ss^3+st^3) > 16 for every pair,
// i.e. the reference's far branch (CoreFunc.h:121,169).  Other tiles go through
// the direct per-pair code in original coordinates, so near pairs keep exact
// differences t - s.
//
// Particles are put in Morton order first (keys on a 10-bit grid per axis) so
// that blocks and tiles are compact; the decision per (block, tile) is computed
// once into a bit mask read by both kernels.
#pragma once
#include "p2p_kernels.hpp"

namespace o3d {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int FF_TB = 256;          // targets per workgroup: 4 waves x 64
constexpr float FF_KAPPA = 1.0f;    // minimum gap, in block radii
constexpr int FF_FOLD = 32;         // near field, compensated: sources per FP32 gradient partial
constexpr int FF_M0 = 16;           // near field: fewest tiles per work item
constexpr int FF_FAR_THREADS = 256; // far field: 4 waves of 64 targets per block
constexpr int FF_TP = 264;          // far field: LDS row of the FP16 moment tables (halves, tile <= 256)
constexpr int FF_SL = 40;           // far field: LDS row of a source's BF16 r2 slots (32 used, 80 bytes)
constexpr int FF_MERGE = 8;         // far field: most consecutive tiles of equal scales per fold

struct FFArgs {
  const float4* sp;    // sorted sources {x, y, z, core_term}
  const float4* sa;    // sorted sources {ax, ay, az, 0}
  const float4* tq;    // sorted targets {x, y, z, core_term or 0 for point targets}
  const int* tperm;    // sorted target -> original target index
  const float4* tsph;  // per source tile: bounding sphere {cx, cy, cz, r}
  const float* tcs;    // per source tile: largest core term
  const float* tam;    // per source tile: largest |alpha|
  const float4* bsph;  // per target block: bounding sphere
  const float* bcs;    // per target block: largest core term
  const unsigned* mask;  // far bit per (block, tile), mwords words per block
  double* farv;        // 12 x nt, sorted target order: far-field velocity and gradient
  const int* nn;       // near tiles per block
  const int* items;    // near-field work items per block
  const int* ioff;     // exclusive scan of items
  const int* mfold;    // near tiles per item
  unsigned* ctr;       // work-queue counter
  double* part;        // per item: 12 x 256 partial sums
  SrcView s;           // original sources (self-pair exclusion)
  TgtView t;           // original targets (outputs, accumulated)
  int64_t ns, nt;
  int32_t tile;        // sources per tile (multiple of 32, <= 512)
  int32_t mwords;
  int32_t ntiles;
  int32_t self_excl;
};

// ---- order-preserving float <-> uint (for atomic min / max of floats) ---------
__device__ __forceinline__ unsigned fenc(float f) {
  const unsigned b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fdec(unsigned e) {
  return __uint_as_float((e & 0x80000000u) ? (e & 0x7fffffffu) : ~e);
}

__global__ void k_ff_bbox_init(unsigned* bb) {
  if (threadIdx.x < 3) bb[threadIdx.x] = 0xffffffffu;
  else if (threadIdx.x < 6) bb[threadIdx.x] = 0u;
}

// bounding box of one or two point sets (x, y, z SoA) into bb[0..5] (encoded).
// Each workgroup reduces its strided share through shuffles and LDS and then
// issues 6 atomics, so a launch of FF_BBOX_GRID workgroups makes few atomics on
// the 6 shared words (one per wave took ~0.5 ms at N = 1M).
constexpr unsigned FF_BBOX_GRID = 512;
__global__ __launch_bounds__(256) void k_ff_bbox(const float* x, const float* y, const float* z, int64_t n,
                                                 unsigned* bb) {
  __shared__ float red[4][6];
  float lo[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
  float hi[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float v[3] = {x[i], y[i], z[i]};
#pragma unroll
    for (int d = 0; d < 3; ++d) { lo[d] = fminf(lo[d], v[d]); hi[d] = fmaxf(hi[d], v[d]); }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = fminf(lo[d], __shfl_xor(lo[d], off));
      hi[d] = fmaxf(hi[d], __shfl_xor(hi[d], off));
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) { red[w][d] = lo[d]; red[w][3 + d] = hi[d]; }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int d = threadIdx.x;
    const int nw = (int)(blockDim.x >> 6);
    float v = red[0][d];
    for (int k = 1; k < nw; ++k) v = d < 3 ? fminf(v, red[k][d]) : fmaxf(v, red[k][d]);
    if (d < 3) atomicMin(&bb[d], fenc(v));
    else atomicMax(&bb[d], fenc(v));
  }
}

__device__ __forceinline__ unsigned spread10(unsigned v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// 30-bit Morton keys on the common bounding box, and the identity permutation
__global__ __launch_bounds__(256) void k_ff_keys(const float* x, const float* y, const float* z, int64_t n,
                                                 const unsigned* bb, unsigned* keys, int* idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float lx = fdec(bb[0]), ly = fdec(bb[1]), lz = fdec(bb[2]);
  const float ext = fmaxf(fmaxf(fdec(bb[3]) - lx, fdec(bb[4]) - ly), fdec(bb[5]) - lz);
  const float sc = ext > 0.0f ? 1023.0f / ext : 0.0f;
  const float q[3] = {(x[i] - lx) * sc, (y[i] - ly) * sc, (z[i] - lz) * sc};
  unsigned k = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float c = fminf(fmaxf(q[d], 0.0f), 1023.0f);   // NaN -> 0
    k |= spread10((unsigned)c) << d;
  }
  keys[i] = k;
  if (idx) idx[i] = (int)i;
}

template <int CORE>
__global__ __launch_bounds__(256) void k_ff_gather_src(SrcView s, const int* perm, int64_t n, float4* sp,
                                                       float4* sa) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  sp[i] = make_float4(s.x[j], s.y[j], s.z[j], core_term<CORE>(s.r[j]));
  sa[i] = make_float4(s.ax[j], s.ay[j], s.az[j], 0.0f);
}

template <int CORE, bool THICK>
__global__ __launch_bounds__(256) void k_ff_gather_tgt(TgtView t, const int* perm, int64_t n, float4* tq) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  tq[i] = make_float4(t.x[j], t.y[j], t.z[j], THICK ? core_term<CORE>(t.r[j]) : 0.0f);
}

// workgroup-wide reductions over 256 threads (4 waves), result in every thread
__device__ __forceinline__ float wg_max(float v, float* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}
__device__ __forceinline__ float wg_min(float v, float* red) { return -wg_max(-v, red); }

// bounding sphere (bbox centre, largest distance), largest core term and (when pa
// is given) largest |alpha| of each group of `group` consecutive sorted points;
// one workgroup per group
__global__ __launch_bounds__(256) void k_ff_bounds(const float4* p, const float4* pa, int64_t n, int group,
                                                   float4* sph, float* cs, float* amax) {
  __shared__ float red[4];
  const int64_t b0 = (int64_t)blockIdx.x * group;
  const int64_t b1 = (b0 + group) < n ? (b0 + group) : n;
  float lo[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
  float hi[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
  float cmax = 0.0f;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    const float4 q = p[i];
    lo[0] = fminf(lo[0], q.x); hi[0] = fmaxf(hi[0], q.x);
    lo[1] = fminf(lo[1], q.y); hi[1] = fmaxf(hi[1], q.y);
    lo[2] = fminf(lo[2], q.z); hi[2] = fmaxf(hi[2], q.z);
    cmax = fmaxf(cmax, q.w);
  }
  float c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float l = wg_min(lo[d], red);
    const float h = wg_max(hi[d], red);
    c[d] = 0.5f * l + 0.5f * h;
  }
  cmax = wg_max(cmax, red);
  float r2 = 0.0f;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    const float4 q = p[i];
    const float dx = q.x - c[0], dy = q.y - c[1], dz = q.z - c[2];
    r2 = fmaxf(r2, fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
  }
  r2 = wg_max(r2, red);
  float a2 = 0.0f;
  if (pa)
    for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
      const float4 q = pa[i];
      a2 = fmaxf(a2, fmaf(q.z, q.z, fmaf(q.y, q.y, q.x * q.x)));
    }
  a2 = wg_max(a2, red);
  if (threadIdx.x == 0) {
    // margin for the rounding of the distances above
    sph[blockIdx.x] = make_float4(c[0], c[1], c[2], __fsqrt_rn(r2) * 1.00001f + 1e-30f);
    cs[blockIdx.x] = cmax;
    if (amax) amax[blockIdx.x] = __fsqrt_rn(a2) * 1.00001f;
  }
}

// far bit of every (target block, source tile): one thread per 32 tiles of a block
template <int CORE>
__global__ __launch_bounds__(256) void k_ff_classify(const float4* bsph, const float* bcs, int64_t nb,
                                                     const float4* tsph, const float* tcs, int ntiles, int mwords,
                                                     unsigned* mask) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nb * mwords) return;
  const int64_t b = g / mwords;
  const int w = (int)(g - b * mwords);
  const float4 cb = bsph[b];
  const float csb = bcs[b];
  unsigned bits = 0u;
  for (int k = 0; k < 32; ++k) {
    const int ti = w * 32 + k;
    if (ti >= ntiles) break;
    const float4 ct = tsph[ti];
    const float dx = cb.x - ct.x, dy = cb.y - ct.y, dz = cb.z - ct.z;
    const float gap = __fsqrt_rn(fmaf(dz, dz, fmaf(dy, dy, dx * dx))) - cb.w - ct.w;
    bool far = gap > 0.0f && gap > FF_KAPPA * cb.w;
    if (CORE == CORE_EXP) far = far && gap * gap > exp_threshold(csb + tcs[ti]);
    if (far) bits |= 1u << k;
  }
  mask[g] = bits;
}

// pair weight Bp = r2^-5/2 from a (scaled) squared distance
__device__ __forceinline__ float bp_of(float r2) {
  const float ir = frsq(r2);
  const float ir2 = ir * ir;
  return ir2 * ir2 * ir;
}

// velocity and gradient (ug order, g = -3 D + skew(W)) of one target from its 32
// moment sums c[] and its offset t from the block centre; al = |t|^2 + st^2 (RM)
__device__ __forceinline__ void ff_epilogue(const double* c, double tx, double ty, double tz, double al,
                                            double* out) {
  const double P[3] = {c[0], c[1], c[2]};
  const double Q[3] = {c[3], c[4], c[5]};
  double R[3][3], S[3][3];
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int k = 0; k < 3; ++k) R[l][k] = c[6 + 3 * l + k];
#pragma unroll
  for (int e = 0; e < 8; ++e) S[e / 3][e % 3] = c[15 + e];
  S[2][2] = -S[0][0] - S[1][1];
  const double U[3] = {c[23], c[24], c[25]};
  const double V[3] = {c[26], c[27], c[28]};
  const double t[3] = {tx, ty, tz};
  double W[3], M[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    W[k] = al * P[k] - 2.0 * (t[0] * R[0][k] + t[1] * R[1][k] + t[2] * R[2][k]) + U[k];
    M[k] = al * Q[k] - 2.0 * (t[0] * S[0][k] + t[1] * S[1][k] + t[2] * S[2][k]) + V[k];
  }
  out[0] = W[1] * t[2] - W[2] * t[1] - M[0];
  out[1] = W[2] * t[0] - W[0] * t[2] - M[1];
  out[2] = W[0] * t[1] - W[1] * t[0] - M[2];
  const double Px[3] = {P[1] * t[2] - P[2] * t[1], P[2] * t[0] - P[0] * t[2], P[0] * t[1] - P[1] * t[0]};
  double a[NACC];   // a[3 + 3j + i] = D_ji for 3j + i < 8 (D_22 = -D_00 - D_11)
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double Rx[3] = {R[j][1] * t[2] - R[j][2] * t[1], R[j][2] * t[0] - R[j][0] * t[2],
                          R[j][0] * t[1] - R[j][1] * t[0]};
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (3 * j + i < 8) a[3 + 3 * j + i] = t[j] * Px[i] - t[j] * Q[i] - Rx[i] + S[j][i];
  }
  a[11] = W[0]; a[12] = W[1]; a[13] = W[2];
  grad_from<double>(a, out + 3);
}

// moment column -> scale group: alpha (0), m and s' alpha (1), s' m (2), b alpha (3),
// b m (4); padding columns 29-31 hold zeros
__device__ __forceinline__ int ff_group(int col) {
  return col < 3 ? 0 : col < 15 ? 1 : col < 23 ? 2 : col < 26 ? 3 : 4;
}

// E with x < 2^E (x > 0), in double so that bounds like gap^-5 cannot overflow
__device__ __forceinline__ int ff_exp2_above(double x) {
  int e;
  (void)frexp(x, &e);
  return e;
}

// FP16 hi/lo split of two scaled FP32 values (|x| < 2^15) into elements e, e + 1
// of packed FP16 vectors: hi by one round-toward-zero pair conversion (the top 11
// significant bits, exact in FP16 for |x| >= 2^-14); lo = x - hi, exact in FP32,
// rounded to FP16 by one mixed-precision FMA (a relative error of 2^-11 of lo,
// about 2^-22 of x).  The moment columns of the staging take the same split.
__device__ __forceinline__ void split16x2(float x0, float x1, f16x8& hv, f16x8& lv, int e) {
  const auto h = __builtin_amdgcn_cvt_pkrtz(x0, x1);
  hv[e] = (_Float16)h[0];
  hv[e + 1] = (_Float16)h[1];
  lv[e] = (_Float16)(x0 - (float)h[0]);
  lv[e + 1] = (_Float16)(x1 - (float)h[1]);
}

// BF16 three-way split x = hi + mid + lo (about 24 significant bits): hi and mid
// take the leading bits, each remainder is exact in FP32
__device__ __forceinline__ void split_bf3(float x, __bf16& hi, __bf16& mid, __bf16& lo) {
  hi = (__bf16)x;
  const float r = x - (float)hi;
  mid = (__bf16)r;
  lo = (__bf16)(r - (float)mid);
}

// squared distances on the matrix cores: r2 = sum_c S_c T_c over the five terms
// S = (-2 s'x, -2 s'y, -2 s'z, b, 1) (per source) and T = (t'x, t'y, t'z, 1, a) (per
// target), both pre-scaled by the tile's r2 scale, each split into BF16 hi/mid/lo
// and multiplied as hi hi + hi mid + mid hi + hi lo + lo hi + mid mid (about 24 bits:
// the accuracy of the FP32 expression it replaces).  Slot q < 30 of the K = 32
// product holds term q / 5 of component q % 5; the parts taken per term:
__device__ __forceinline__ int ff_src_part(int term) { return term == 2 || term == 5 ? 1 : (term == 4 ? 2 : 0); }
__device__ __forceinline__ int ff_tgt_part(int term) { return term == 1 || term == 5 ? 1 : (term == 3 ? 2 : 0); }

// source position of tile slot p in the FP16 moment tables: inside each group of 16
// sources, slot 16g + 8h + e holds source 16g + 4h + (e % 4) + 8 (e / 4), so that the
// squared distances a lane gets from the r2 product (rows 8(r/4) + 4h + r%4) are, in
// register order, the sources of its contiguous 8 K-slots in the moment product.
__device__ __forceinline__ int ff_perm(int j) {
  const int g = j & ~15, s = j & 15;
  return g + 8 * ((s >> 2) & 1) + (s & 3) + 4 * (s >> 3);
}

// Far field: one workgroup per 256 sorted targets (4 waves of 64); lane l of wave w
// owns target w*64 + l at the end, and during the sums works for targets
// w*64 + (l & 31) (set 0) and w*64 + 32 + (l & 31) (set 1).  Per chunk of 32 sources
// and set: two BF16 products give the 32 x 32 squared distances (lane l: target
// l & 31, 16 sources), the lane turns its 16 into pair weights Bp (one v_rsq_f32,
// three products) split into FP16 hi + lo, and six FP16 products (two 16-source
// K blocks, three split terms each) add the weights times the 29 moment columns.
//
// The moment products keep about 22 significant bits per product (hi hi + lo hi +
// hi lo), the accuracy of the FP32 products they replace.  Per tile, pair weights
// are scaled by 2^(5 kw) (from Bp <= gap^-5), applied as r2 -> r2 2^(-2 kw) through
// the staged source terms and the targets' a, and each moment group by 2^eg (from
// the tile's largest |alpha| and source distance), so that every FP16 factor stays
// below 2^14; the scales are removed when the tile's FP32 sums are folded into FP64.
//
// 32x32x16 FP16/BF16 matrix operand layout (A: 32 x 16, B: 16 x 32, D: 32 x 32):
//   A[i][k] in lane i + 32 (k / 8), element k % 8;  B[k][j] in lane j + 32 (k / 8),
//   element k % 8;  D[i][j] in lane j + 32 ((i / 4) % 2), register (i % 4) + 4 (i / 8).
// r2 product: A = source slots (row = source), B = target slots, D = r2.
// Moment product: A = moment table transposed (row = column), B = Bp (source, target).
// LDS: r2 source slots [source][FF_SL] BF16 (rows of 80 bytes), and the scaled
// moments as FP16 hi and lo tables [column][FF_TP], rows of 264 halves, so that the
// 16-byte reads of 16 lanes cover all 64 banks.  Each tile's matrix sums are folded
// at once through the FP64 epilogue into the target's 12 results, which keeps the
// kernel within 256 registers (two waves per SIMD) with every matrix result in
// ordinary VGPRs.
template <int CORE, bool THICK, int PREC>
__global__ __launch_bounds__(FF_FAR_THREADS, 2) void k_ff_far(FFArgs a) {
  extern __shared__ float4 sh[];
  const int T = a.tile;
  constexpr int TP = FF_TP;
  __bf16* sl = reinterpret_cast<__bf16*>(sh);
  _Float16* mh = reinterpret_cast<_Float16*>(sl + 256 * FF_SL);
  _Float16* ml = mh + 32 * TP;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int half = lane >> 5, col = lane & 31;
  const int64_t blk = blockIdx.x;
  const int64_t base = blk * FF_TB + w * 64;
  const float4 cb = a.bsph[blk];

  float tx[2], ty[2], tz[2], tt[2], al[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    int64_t i = base + 32 * s + col;
    i = i < a.nt ? i : a.nt - 1;
    const float4 q = a.tq[i];
    tx[s] = q.x - cb.x; ty[s] = q.y - cb.y; tz[s] = q.z - cb.z;
    tt[s] = CORE == CORE_RM ? q.w : 0.0f;
    al[s] = fmaf(tz[s], tz[s], fmaf(ty[s], ty[s], fmaf(tx[s], tx[s], tt[s])));
  }
  f32x16 acc0, acc1;
#pragma unroll
  for (int r = 0; r < 16; ++r) { acc0[r] = 0.0f; acc1[r] = 0.0f; }
  // lane l < 32 owns target w*64 + l (set 0), lane l >= 32 target w*64 + l (set 1)
  const double ox = half ? tx[1] : tx[0], oy = half ? ty[1] : ty[0], oz = half ? tz[1] : tz[0];
  const double oa = ox * ox + oy * oy + oz * oz + (double)(half ? tt[1] : tt[0]);
  double sum[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) sum[k] = 0.0;

  const unsigned* mrow = a.mask + blk * a.mwords;
  // consecutive far tiles with the same scales share one fold: acc keeps summing
  // (at most FF_MERGE tiles, which bounds its FP32 sum) and the epilogue runs when
  // the scales change, the run is full, or the tiles end
  int run = 0, pew = 0, peg[5] = {0, 0, 0, 0, 0};
  bf16x8 tb[2][2];
  // fold: one v_permlane32_swap per register exchanges the upper 32 lanes of acc0[r]
  // with the lower 32 of acc1[r], after which every lane holds its own target's rows
  // of half 0 in acc0[r] and of half 1 in acc1[r] (row of register r in half h:
  // 8 (r / 4) + 4 h + r % 4); the scales are removed and the epilogue adds the run's
  // velocity and gradient in FP64
  auto fold = [&]() {
    double c[32];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc0[r]), __float_as_uint(acc1[r]),
                                                       false, false);
      const int i0 = 8 * (r / 4) + (r % 4);
      c[i0] = (double)__uint_as_float(sw[0]) * ldexp(1.0, -(pew + peg[ff_group(i0)]));
      c[i0 + 4] = (double)__uint_as_float(sw[1]) * ldexp(1.0, -(pew + peg[ff_group(i0 + 4)]));
      acc0[r] = 0.0f;
      acc1[r] = 0.0f;
    }
    double out[12];
    ff_epilogue(c, ox, oy, oz, oa, out);
#pragma unroll
    for (int k = 0; k < 12; ++k) sum[k] += out[k];
  };
  for (int ti = 0; ti < a.ntiles; ++ti) {
    if (!((mrow[ti >> 5] >> (ti & 31)) & 1u)) continue;
    const int64_t s0 = (int64_t)ti * T;
    const int cnt = (int)((a.ns - s0) < T ? (a.ns - s0) : T);
    // scales of this (block, tile), the same in every lane
    const float4 ct = a.tsph[ti];
    const double ddx = (double)cb.x - ct.x, ddy = (double)cb.y - ct.y, ddz = (double)cb.z - ct.z;
    const double dc = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    const double gap = dc - cb.w - ct.w;                   // > 0 for a far tile
    const double g2 = gap * gap;
    // weight scale 2^(5 kw) <= 2^14 / max Bp, applied as r2 -> r2 2^(-2 kw)
    const int kw = (int)floor((14 - ff_exp2_above(1.0 / (g2 * g2 * gap))) / 5.0);
    const int ew = 5 * kw;
    const float rsc = ldexpf(1.0f, -2 * kw);
    const double rs = dc + ct.w;                           // |s'| bound
    const double am = (double)a.tam[ti] + 1e-300;
    const double bm = rs * rs + (CORE == CORE_RM ? (double)a.tcs[ti] : 0.0);
    int eg[5];
    eg[0] = 14 - ff_exp2_above(am);
    eg[1] = 14 - ff_exp2_above(am * rs);
    eg[2] = 14 - ff_exp2_above(am * rs * rs);
    eg[3] = 14 - ff_exp2_above(am * bm);
    eg[4] = 14 - ff_exp2_above(am * bm * rs);
    float gs[5];
#pragma unroll
    for (int g = 0; g < 5; ++g) gs[g] = ldexpf(1.0f, eg[g]);
    bool same = run > 0 && run < FF_MERGE && ew == pew;
#pragma unroll
    for (int g = 0; g < 5; ++g) same = same && eg[g] == peg[g];
    if (run > 0 && !same) {
      fold();
      run = 0;
    }
    pew = ew;
#pragma unroll
    for (int g = 0; g < 5; ++g) peg[g] = eg[g];
    ++run;

    // this lane's target slots for both sets and both r2 products (K slots 8 half + e
    // and 16 + 8 half + e); they depend on the tile only through rsc
    if (run == 1) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const float tv[5] = {tx[s], ty[s], tz[s], 1.0f, al[s] * rsc};
      __bf16 part[3][5];
#pragma unroll
      for (int c = 0; c < 5; ++c) split_bf3(tv[c], part[0][c], part[1][c], part[2][c]);
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int q0 = 16 * m + e, q1 = 16 * m + 8 + e;     // slot for half 0 / half 1
          const __bf16 v0 = q0 < 30 ? part[ff_tgt_part(q0 / 5)][q0 % 5] : (__bf16)0.0f;
          const __bf16 v1 = q1 < 30 ? part[ff_tgt_part(q1 / 5)][q1 % 5] : (__bf16)0.0f;
          tb[s][m][e] = half ? v1 : v0;
        }
    }
    }

    __syncthreads();
    for (int j = tid; j < T; j += FF_FAR_THREADS) {
      // the 32 moment columns, scaled, then split into FP16 hi/lo two at a time
      const int pj = ff_perm(j);
      __bf16* srow = sl + j * FF_SL;
      float sv5[5];
      float col[32];
#pragma unroll
      for (int c = 0; c < 32; ++c) col[c] = 0.0f;
      if (j < cnt) {
        const float4 P = a.sp[s0 + j];
        const float4 A = a.sa[s0 + j];
        const float sx = P.x - cb.x, sy = P.y - cb.y, sz = P.z - cb.z;
        const float cw = CORE == CORE_RM ? P.w : 0.0f;
        const float m0 = fmaf(A.y, sz, -A.z * sy);
        const float m1 = fmaf(A.z, sx, -A.x * sz);
        const float m2 = fmaf(A.x, sy, -A.y * sx);
        const float b = fmaf(sz, sz, fmaf(sy, sy, fmaf(sx, sx, cw)));
        sv5[0] = -2.0f * sx * rsc; sv5[1] = -2.0f * sy * rsc; sv5[2] = -2.0f * sz * rsc;   // exact scalings
        sv5[3] = b * rsc; sv5[4] = 1.0f;
        const float sv[3] = {sx, sy, sz}, av[3] = {A.x, A.y, A.z}, mv[3] = {m0, m1, m2};
#pragma unroll
        for (int k = 0; k < 3; ++k) { col[k] = av[k]; col[3 + k] = mv[k]; }
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
          for (int k = 0; k < 3; ++k) col[6 + 3 * l + k] = sv[l] * av[k];
#pragma unroll
        for (int e = 0; e < 8; ++e) col[15 + e] = sv[e / 3] * mv[e % 3];
#pragma unroll
        for (int k = 0; k < 3; ++k) { col[23 + k] = b * av[k]; col[26 + k] = b * mv[k]; }
      } else {
        // padding: zero moments and a large finite distance, so that Bp * moment = 0
        sv5[0] = 0.0f; sv5[1] = 0.0f; sv5[2] = 0.0f; sv5[3] = 1e20f; sv5[4] = 1.0f;
      }
#pragma unroll
      for (int c = 0; c < 32; c += 2) {
        const float x0 = col[c] * gs[ff_group(c)], x1 = col[c + 1] * gs[ff_group(c + 1)];
        const auto h = __builtin_amdgcn_cvt_pkrtz(x0, x1);
        mh[c * TP + pj] = (_Float16)h[0];
        mh[(c + 1) * TP + pj] = (_Float16)h[1];
        ml[c * TP + pj] = (_Float16)(x0 - (float)h[0]);
        ml[(c + 1) * TP + pj] = (_Float16)(x1 - (float)h[1]);
      }
      __bf16 part[3][5];
#pragma unroll
      for (int c = 0; c < 5; ++c) split_bf3(sv5[c], part[0][c], part[1][c], part[2][c]);
#pragma unroll
      for (int q = 0; q < 32; ++q) srow[q] = q < 30 ? part[ff_src_part(q / 5)][q % 5] : (__bf16)0.0f;
    }
    __syncthreads();
    for (int k0 = 0; k0 < T; k0 += 32) {
      // r2 of sources k0 + 0..31 with this lane's two targets
      const __bf16* srow = sl + (k0 + col) * FF_SL + 8 * half;
      const bf16x8 sa0 = *reinterpret_cast<const bf16x8*>(srow);
      const bf16x8 sa1 = *reinterpret_cast<const bf16x8*>(srow + 16);
      f32x16 r20 = {}, r21 = {};
      r20 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sa0, tb[0][0], r20, 0, 0, 0);
      r21 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sa0, tb[1][0], r21, 0, 0, 0);
      r20 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sa1, tb[0][1], r20, 0, 0, 0);
      r21 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sa1, tb[1][1], r21, 0, 0, 0);
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        const int pbase = k0 + 16 * kb + 8 * half;
        const f16x8 Ah = *reinterpret_cast<const f16x8*>(mh + col * TP + pbase);
        const f16x8 Al = *reinterpret_cast<const f16x8*>(ml + col * TP + pbase);
        f16x8 Bh0, Bl0, Bh1, Bl1;
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
          split16x2(bp_of(r20[8 * kb + e]), bp_of(r20[8 * kb + e + 1]), Bh0, Bl0, e);
          split16x2(bp_of(r21[8 * kb + e]), bp_of(r21[8 * kb + e + 1]), Bh1, Bl1, e);
        }
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah, Bh0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah, Bh1, acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah, Bl0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah, Bl1, acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(Al, Bh0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(Al, Bh1, acc1, 0, 0, 0);
      }
    }
  }
  if (run > 0) fold();
  const int64_t i = base + lane;
  if (i < a.nt) {
#pragma unroll
    for (int k = 0; k < 12; ++k) a.farv[(int64_t)k * a.nt + i] = sum[k];
  }
}

// ---- near field -------------------------------------------------------------
// The near tiles of a block are split into items of at most M tiles (M chosen on
// the device so that there are at most 3 items per block on average), so that a block
// whose targets straddle a jump of the Morton order, and which therefore has
// nearly every tile near, is spread over many workgroups.  Items are taken from a
// queue by a persistent kernel and write FP64 partial sums; the combine kernel
// adds a block's partials in item order, so results do not depend on scheduling.

__device__ __forceinline__ int near_tiles_of(const unsigned* mrow, int mwords, int ntiles) {
  int far = 0;
  for (int w = 0; w < mwords; ++w) far += __popc(mrow[w]);
  return ntiles - far;
}

// nn[b] = number of near tiles of block b
__global__ __launch_bounds__(256) void k_ff_plan_count(const unsigned* mask, int64_t nb, int mwords, int ntiles,
                                                       int* nn) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) nn[b] = near_tiles_of(mask + b * mwords, mwords, ntiles);
}

// after the exclusive scan noff of nn: M = max(FF_M0, ceil(total / (2 nb))) and the
// item count of every block (at most 3 nb items in all)
__global__ __launch_bounds__(256) void k_ff_plan_items(const int* nn, const int* noff, int64_t nb, int* items,
                                                       int* mfold) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)noff[nb - 1] + nn[nb - 1];
  int64_t m = (total + 2 * nb - 1) / (2 * nb);
  m = m < FF_M0 ? FF_M0 : m;
  if (b == 0) *mfold = (int)m;
  if (b < nb) items[b] = (int)((nn[b] + m - 1) / m);
}

__global__ void k_ff_zero_u32(unsigned* p) {
  if (threadIdx.x == 0) *p = 0u;
}

// per-target partial of one item: near tiles of rank [r0, r1) among block b's
template <int CORE, int PREC>
__device__ __forceinline__ void near_item(const FFArgs& a, int64_t b, int r0, int r1, float4* sh, double* r) {
  const int T = a.tile;
  const int tid = threadIdx.x;
  const int64_t i = b * FF_TB + tid;
  const int64_t ic = i < a.nt ? i : a.nt - 1;
  const float4 q = a.tq[ic];
  float tx[1] = {q.x}, ty[1] = {q.y}, tz[1] = {q.z}, tt[1] = {q.w};
  float acc[1][NACC];
  float hi[3] = {0.f, 0.f, 0.f}, lo[3] = {0.f, 0.f, 0.f};
  double gr[NACC];
#pragma unroll
  for (int c = 0; c < NACC; ++c) { acc[0][c] = 0.0f; gr[c] = 0.0; }

  const unsigned* mrow = a.mask + b * a.mwords;
  int rank = 0;
  for (int w = 0; w < a.mwords && rank < r1; ++w) {
    const int base = w * 32;
    const unsigned valid = (a.ntiles - base) >= 32 ? 0xffffffffu : ((1u << (a.ntiles - base)) - 1u);
    unsigned near = ~mrow[w] & valid;
    const int c = __popc(near);
    if (rank + c <= r0) { rank += c; continue; }
    while (near && rank < r1) {
      const int ti = base + __builtin_ctz(near);
      near &= near - 1u;
      if (rank++ < r0) continue;
      const int64_t s0 = (int64_t)ti * T;
      const int cnt = (int)((a.ns - s0) < T ? (a.ns - s0) : T);
      __syncthreads();
      for (int j = tid; j < cnt; j += FF_TB) {
        sh[2 * j] = a.sp[s0 + j];
        sh[2 * j + 1] = a.sa[s0 + j];
      }
      __syncthreads();
      float thr[1] = {CORE == CORE_EXP ? exp_threshold(a.tcs[ti] + tt[0]) : 0.0f};
      if (PREC == PREC_COMP) {
        // In Morton order a near tile holds the target's closest sources, whose large
        // terms cancel; the gradient's FP32 partials are therefore folded into FP64
        // every FF_FOLD sources, the velocity's into its TwoSum pair once per tile.
        for (int j0 = 0; j0 < cnt; j0 += FF_FOLD) {
          const int j1 = j0 + FF_FOLD < cnt ? j0 + FF_FOLD : cnt;
          for (int j = j0; j < j1; ++j)
            source_acc<CORE, true, 1>(tx, ty, tz, tt, thr, sh[2 * j], sh[2 * j + 1], acc);
          fold_grad_raw(acc[0], gr);
        }
        fold_vel_comp(acc[0], hi, lo);
      } else {
        for (int j = 0; j < cnt; ++j)
          source_acc<CORE, true, 1>(tx, ty, tz, tt, thr, sh[2 * j], sh[2 * j + 1], acc);
      }
    }
  }
  if (PREC == PREC_COMP) {
    r[0] = (double)hi[0] + (double)lo[0];
    r[1] = (double)hi[1] + (double)lo[1];
    r[2] = (double)hi[2] + (double)lo[2];
    grad_from<double>(gr, r + 3);
  } else {
    float g[9];
    grad_from<float>(acc[0], g);
    r[0] = acc[0][0]; r[1] = acc[0][1]; r[2] = acc[0][2];
#pragma unroll
    for (int c = 0; c < 9; ++c) r[3 + c] = g[c];
  }
}

// persistent: every workgroup takes items from the queue until none is left
template <int CORE, int PREC>
__global__ __launch_bounds__(256) void k_ff_near(FFArgs a) {
  extern __shared__ float4 sh[];
  __shared__ int s_item;
  const int64_t nb = (a.nt + FF_TB - 1) / FF_TB;
  const int total = a.ioff[nb - 1] + a.items[nb - 1];
  const int m = *a.mfold;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.ctr, 1u);
    __syncthreads();
    const int item = s_item;
    if (item >= total) break;
    // block b: the last with ioff[b] <= item (blocks without items share the offset
    // of the next block, so the last of a run of equal offsets owns the items)
    int64_t lo = 0, hi = nb;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) / 2;
      if (a.ioff[mid] <= item) lo = mid; else hi = mid;
    }
    const int64_t b = lo;
    const int k = item - a.ioff[b];
    const int r0 = k * m;
    const int nnb = a.nn[b];
    const int r1 = r0 + m < nnb ? r0 + m : nnb;
    double r[12];
    near_item<CORE, PREC>(a, b, r0, r1, sh, r);
    double* slot = a.part + (int64_t)item * 12 * FF_TB;
#pragma unroll
    for (int c = 0; c < 12; ++c) slot[c * FF_TB + threadIdx.x] = r[c];
  }
}

// one workgroup per block: near partials in item order, far part, self-pair
// exclusion, then accumulate into the caller's arrays at the original index
template <int CORE, int PREC>
__global__ __launch_bounds__(256) void k_ff_combine(FFArgs a) {
  const int64_t b = blockIdx.x;
  const int64_t i = b * FF_TB + threadIdx.x;
  if (i >= a.nt) return;
  double r[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) r[c] = a.farv[(int64_t)c * a.nt + i];
  const int n_it = a.items[b], off = a.ioff[b];
  for (int k = 0; k < n_it; ++k) {
    const double* slot = a.part + (int64_t)(off + k) * 12 * FF_TB;
#pragma unroll
    for (int c = 0; c < 12; ++c) r[c] += slot[c * FF_TB + threadIdx.x];
  }
  const int io = a.tperm[i];
  if (a.self_excl) {
    // the self pair entered only W (velocity and dyad vanish at d = 0)
    float ks, Bp;
    kb<CORE>(0.0f, core_term<CORE>(a.s.r[io]) + a.tq[i].w, 0.0f, ks, Bp);
    const double kx = (double)ks * a.s.ax[io], ky = (double)ks * a.s.ay[io], kz = (double)ks * a.s.az[io];
    r[4] -= kz; r[5] += ky; r[6] += kz; r[8] -= kx; r[9] -= ky; r[10] += kx;
  }
  a.t.u[io] = (float)((double)a.t.u[io] + r[0]);
  a.t.v[io] = (float)((double)a.t.v[io] + r[1]);
  a.t.w[io] = (float)((double)a.t.w[io] + r[2]);
#pragma unroll
  for (int c = 0; c < 9; ++c) a.t.g[c][io] = (float)((double)a.t.g[c][io] + r[3 + c]);
}

// number of (target, source) pairs the far-field kernel takes: one thread per block
__global__ __launch_bounds__(256) void k_ff_count(const unsigned* mask, int64_t nb, int mwords, int ntiles, int T,
                                                  int64_t ns, int64_t nt, unsigned long long* out) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int64_t bt = (nt - b * FF_TB) < FF_TB ? (nt - b * FF_TB) : FF_TB;
  unsigned long long sum = 0;
  for (int w = 0; w < mwords; ++w) {
    unsigned bits = mask[b * mwords + w];
    while (bits) {
      const int ti = w * 32 + __builtin_ctz(bits);
      bits &= bits - 1u;
      const int64_t cnt = (ns - (int64_t)ti * T) < T ? (ns - (int64_t)ti * T) : T;
      if (ti < ntiles) sum += (unsigned long long)(cnt * bt);
    }
  }
  atomicAdd(out, sum);
}

__global__ void k_ff_zero_u64(unsigned long long* p) {
  if (threadIdx.x == 0) *p = 0ull;
}

}  // namespace o3d
