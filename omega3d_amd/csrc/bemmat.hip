// bemmat.hip -- the BEM influence matrix on the GPU: panels_on_panels_coeff
// (Coefficients.h:169-444) over the panel-panel recursion rkernel_2vs_2p
// (Kernels.h:1213-1311).
//
// Entry (source panel j, unknown k) -> (target panel i, component m) of a
// column-major block with nunk*nt rows: the velocity that panel j induces at the
// centroid of panel i carrying unit sheet strength along its x1 (k = 0), its x2
// (k = 1) or a unit source sheet (k = 2), averaged over the recursion's target
// sub-triangles, projected on the target's x1, x2, normal (m = 0, 1, 2), times
// 1/(4 pi).  The recursion splits both triangles into 4 (16 child pairs) until the
// centroids are more than 4 (sqrt(sa) + sqrt(ta)) apart or maxlev is reached; each
// leaf is a singular point at the source sub-centroid acting on the target
// sub-centroid, with strength (sheet strength) x area / 16^level.
//
// By linearity every column of a pair comes from one vector
//   G = sum over leaves of 16^-level K(d) d,   K = 1/|d|^3,  d = t_leaf - s_leaf,
// as u_k = sa (b_k x G) for the vortex unknowns and u_2 = sa G for the source, so
// the tree is walked once per pair, not once per unknown as in the reference.
//
// GPU form.  A prologue stores per panel its centroid, sqrt(area), area and basis
// (Surfaces::compute_bases, Surfaces.h:766-816, with its float/double mixing) and
// its corners.  k_bm_far: one lane per target, sources staged in LDS; it takes
// the root decision for every pair in the reference's float arithmetic (FMA
// contraction off, IEEE sqrt where the cheap test is ambiguous), writes the
// coefficients of well-separated pairs (one leaf) -- for a fixed source the lanes
// of a wave write consecutive rows, so the columns stream out coalesced, which is
// what bounds this kernel (HBM writes) -- and lists the others.  A first pass of
// the same kernel only counts them, to size the list.  (Running the recursed
// pairs on a second stream beside the writes was measured slower: the write
// kernel's workgroups hold the CUs.)  k_bm_near: one wave per
// listed pair: 16 lanes take the level-1 decisions and leave the open pairs'
// triangles in LDS, the children of the open level-1 pairs are spread over the
// wave, each lane walks what lies below its own, and a fixed
// butterfly adds the lanes' G in FP64, so results do not depend on scheduling.
// Every tree decision below the root is the reference's float arithmetic too, so
// the GPU subdivides exactly where the reference does; leaves below the last
// decision level take FP64 centroids of the float sub-triangle nodes.
//
// The panel on itself (same surface, i == j) gets the reference's constants
// (Coefficients.h:399-421) instead of a recursion.
#include "o3d_api_internal.hpp"
#include "o3d_common.hpp"

#include <cmath>
#include <mutex>

namespace o3d {
namespace {

// per-panel record, 8 float4: [0] centroid, sqrt(area); [1] area, x1; [2] x2, -;
// [3] normal, -; [4] corner 0, corner 1 x; [5] corner 1 y z, corner 2 x y;
// [6] corner 2 z, -, -, -; [7] -
constexpr int BM_F4 = 8;
constexpr int BM_MAXLEV = 4;

struct BmGeomArgs {
  const float* x; const float* y; const float* z;
  const uint32_t* idx;
  const float* area;
  int64_t np;
  float4* rec;
};

__global__ __launch_bounds__(BLOCK) void k_bm_geom(BmGeomArgs a) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= a.np) return;
  const uint32_t id0 = a.idx[3 * i], id1 = a.idx[3 * i + 1], id2 = a.idx[3 * i + 2];
  const float* xs[3] = {a.x, a.y, a.z};
  float v[9];
#pragma unroll
  for (int d = 0; d < 3; ++d) { v[d] = xs[d][id0]; v[3 + d] = xs[d][id1]; v[6 + d] = xs[d][id2]; }
  // Kernels.h:1243-1248: (x0 + x1 + x2) / 3 in float
  const float cx = (v[0] + v[3] + v[6]) / 3.0f;
  const float cy = (v[1] + v[4] + v[7]) / 3.0f;
  const float cz = (v[2] + v[5] + v[8]) / 3.0f;
  const float ar = a.area[i];
  // Surfaces.h:783-806
  float x1[3], x2[3], nm[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) x1[d] = v[3 + d] - v[d];
  const float base = sqrtf(x1[0] * x1[0] + x1[1] * x1[1] + x1[2] * x1[2]);
#pragma unroll
  for (int d = 0; d < 3; ++d) x1[d] = (float)((double)x1[d] * (1.0 / (double)base));
#pragma unroll
  for (int d = 0; d < 3; ++d) x2[d] = v[6 + d] - v[d];
  const float dp = x2[0] * x1[0] + x2[1] * x1[1] + x2[2] * x1[2];
#pragma unroll
  for (int d = 0; d < 3; ++d) x2[d] = x2[d] - dp * x1[d];
  const float height = sqrtf(x2[0] * x2[0] + x2[1] * x2[1] + x2[2] * x2[2]);
#pragma unroll
  for (int d = 0; d < 3; ++d) x2[d] = (float)((double)x2[d] * (1.0 / (double)height));
  nm[0] = x1[1] * x2[2] - x1[2] * x2[1];
  nm[1] = x1[2] * x2[0] - x1[0] * x2[2];
  nm[2] = x1[0] * x2[1] - x1[1] * x2[0];
  float4* r = a.rec + i * BM_F4;
  r[0] = make_float4(cx, cy, cz, sqrtf(ar));
  r[1] = make_float4(ar, x1[0], x1[1], x1[2]);
  r[2] = make_float4(x2[0], x2[1], x2[2], 0.f);
  r[3] = make_float4(nm[0], nm[1], nm[2], 0.f);
  r[4] = make_float4(v[0], v[1], v[2], v[3]);
  r[5] = make_float4(v[4], v[5], v[6], v[7]);
  r[6] = make_float4(v[8], 0.f, 0.f, 0.f);
  r[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// my_well_sep (Kernels.h:986-999) on the reference's float values: dist > 4 size
// with dist = sqrt(d2) correctly rounded.  The square comparison decides unless
// d2 lies within 1e-4 of (4 size)^2, where the IEEE square root decides.
__device__ __forceinline__ bool well_sep(float d2, float t4) {
#pragma clang fp contract(off)
  const float q = t4 * t4;
  if (d2 > q * 1.0001f) return true;
  if (d2 < q * 0.9999f) return false;
  return sqrtf(d2) > t4;
}

// d2 of two centroids in the reference's order (Kernels.h:979-981)
__device__ __forceinline__ float dist2(float tx, float ty, float tz, float sx, float sy, float sz, float& ex,
                                       float& ey, float& ez) {
#pragma clang fp contract(off)
  ex = tx - sx; ey = ty - sy; ez = tz - sz;
  return ex * ex + ey * ey + ez * ez;
}

struct BmFarArgs {
  const float4* srec; const float4* trec;
  int64_t ns, nt;
  int64_t schunk;            // sources per blockIdx.y
  int32_t nunk, same, maxlev, pad;
  float* coeffs;             // (nunk*ns) x (nunk*nt), column-major
  uint2* list;               // near pairs (target, source)
  unsigned long long* count; // near pairs found
  unsigned long long cap;    // list capacity
  int32_t vec4;              // coeffs and every column 16-byte aligned
};

constexpr int BM_TILE = BLOCK;   // sources staged per LDS pass

// BM_COUNT: count the pairs that need the recursion; BM_WRITE: write the one-leaf
// and self coefficients and list the others.  The same decisions in both, so the
// list has exactly the counted length.
enum { BM_COUNT = 0, BM_WRITE = 1 };

template <int MODE>
__global__ __launch_bounds__(BLOCK) void k_bm_far(BmFarArgs a) {
  __shared__ float4 sh[BM_TILE * 3];
  // per wave: its 3 column segments of 64 rows x 3 floats (float4 stores)
  __shared__ __align__(16) float wbs[(BLOCK / 64) * 3 * 64 * 3];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  float* wb = wbs + (tid >> 6) * 3 * 64 * 3;
  const int64_t i = (int64_t)blockIdx.x * BLOCK + tid;
  const int64_t i0 = i - lane;   // the wave's first row
  // float4 stores: every row of the wave valid and the columns 16-byte aligned
  const bool vec = MODE == BM_WRITE && a.vec4 && i0 + 64 <= a.nt;
  const bool valid = i < a.nt;
  const int64_t ic = valid ? i : a.nt - 1;
  const float4 t0 = a.trec[ic * BM_F4];
  const float4 t1 = a.trec[ic * BM_F4 + 1];
  const float4 t2 = a.trec[ic * BM_F4 + 2];
  const float4 t3 = a.trec[ic * BM_F4 + 3];
  const float tb[3][3] = {{t1.y, t1.z, t1.w}, {t2.x, t2.y, t2.z}, {t3.x, t3.y, t3.z}};
  const int nunk = a.nunk;
  const int64_t R = (int64_t)nunk * a.nt;
  const float fac = (float)(1.0 / (4.0 * M_PI));
  const float tp = (float)(2.0 * M_PI) * fac;
  const int64_t j0 = (int64_t)blockIdx.y * a.schunk;
  const int64_t j1 = j0 + a.schunk < a.ns ? j0 + a.schunk : a.ns;
  unsigned long long cnt = 0;
  for (int64_t jt = j0; jt < j1; jt += BM_TILE) {
    const int n = (int)(j1 - jt < BM_TILE ? j1 - jt : BM_TILE);
    __syncthreads();
    if (tid < n) {
      const float4* r = a.srec + (jt + tid) * BM_F4;
      sh[3 * tid] = r[0];
      sh[3 * tid + 1] = r[1];
      sh[3 * tid + 2] = r[2];
    }
    __syncthreads();
    for (int jj = 0; jj < n; ++jj) {
      const int64_t j = jt + jj;
      const float4 s0 = sh[3 * jj], s1 = sh[3 * jj + 1], s2 = sh[3 * jj + 2];
      float ex, ey, ez;
      const float d2 = dist2(t0.x, t0.y, t0.z, s0.x, s0.y, s0.z, ex, ey, ez);
      float t4;
      {
#pragma clang fp contract(off)
        t4 = (s0.w + t0.w) * 4.0f;
      }
      const bool diag = a.same && j == i;
      const bool leaf = a.maxlev == 0 || well_sep(d2, t4);
      const bool nearp = valid && !diag && !leaf;
      if (MODE == BM_COUNT) {
        cnt += nearp ? 1ull : 0ull;
        continue;
      }
      const unsigned long long m = __builtin_amdgcn_ballot_w64(nearp);
      if (m) {
        const int first = __builtin_ctzll(m);
        unsigned long long base = 0;
        if (lane == first) base = atomicAdd(a.count, (unsigned long long)__builtin_popcountll(m));
        base = __shfl(base, first, 64);
        const unsigned long long slot = base + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                      __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (nearp && slot < a.cap) a.list[slot] = make_uint2((unsigned)i, (unsigned)j);
      }
      if (!vec && (!valid || nearp)) continue;
      // this pair's coefficients [unknown k][component m]
      float cv[3][3];
      if (diag) {
        // Coefficients.h:399-421, then the scaling of :428-433
        cv[0][0] = 0.f; cv[0][1] = tp; cv[0][2] = 0.f;
        cv[1][0] = -tp; cv[1][1] = 0.f; cv[1][2] = 0.f;
        cv[2][0] = 0.f; cv[2][1] = 0.f; cv[2][2] = tp;
      } else {
        // one leaf: kernel_0vs_0p at the centroids (Kernels.h:115-133), sigma = 0
        const float rr = frsq(d2);
        const float w = rr * rr * rr * s1.x * fac;     // K sa / (4 pi)
        const float gx = w * ex, gy = w * ey, gz = w * ez;
        const float sb[2][3] = {{s1.y, s1.z, s1.w}, {s2.x, s2.y, s2.z}};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          float ux, uy, uz;
          if (k < 2) {   // (b x G): the reference's s x d order, Kernels.h:128-130
            ux = gz * sb[k][1] - gy * sb[k][2];
            uy = gx * sb[k][2] - gz * sb[k][0];
            uz = gy * sb[k][0] - gx * sb[k][1];
          } else {
            ux = gx; uy = gy; uz = gz;
          }
          cv[k][0] = ux * tb[0][0] + uy * tb[0][1] + uz * tb[0][2];
          cv[k][1] = ux * tb[1][0] + uy * tb[1][1] + uz * tb[1][2];
          cv[k][2] = ux * tb[2][0] + uy * tb[2][1] + uz * tb[2][2];
        }
      }
      if (vec && nunk == 3) {
        // three floats per lane and column: the wave's 64 consecutive rows are 768
        // contiguous bytes, one 12-byte store per lane (7% faster on the icosphere than
        // the transposition through LDS below, which 1 or 2 unknowns still take); the
        // entries of a listed pair are written too and replaced by k_bm_near
        typedef float f3v __attribute__((ext_vector_type(3)));
        float* colx = a.coeffs + (j * nunk) * R + i * nunk;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          f3v v3 = {cv[k][0], cv[k][1], cv[k][2]};
          __builtin_memcpy(colx + k * R, &v3, 12);
        }
        continue;
      }
      if (vec) {
        // the wave's column segments (64 rows x nunk floats each) through LDS, out as
        // float4 stores; the entries of a listed pair are written too and replaced
        // by k_bm_near, which runs after this kernel
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if (k == nunk) break;
#pragma unroll
          for (int m = 0; m < 3; ++m)
            if (m < nunk) wb[k * 64 * 3 + lane * nunk + m] = cv[k][m];
        }
        __builtin_amdgcn_wave_barrier();
        float* col0 = a.coeffs + (j * nunk) * R + i0 * nunk;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if (k == nunk) break;
          if (lane < 16 * nunk)
            reinterpret_cast<float4*>(col0 + k * R)[lane] = reinterpret_cast<const float4*>(wb + k * 64 * 3)[lane];
        }
        __builtin_amdgcn_wave_barrier();
        continue;
      }
      float* col = a.coeffs + (j * nunk) * R + i * nunk;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k == nunk) break;
        col[0] = cv[k][0];
        col[1] = cv[k][1];
        if (nunk == 3) col[2] = cv[k][2];
        col += R;
      }
    }
  }
  if (MODE == BM_COUNT) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if ((tid & 63) == 0 && cnt) atomicAdd(a.count, cnt);
  }
}

// child c of triangle v (9 floats) over the six nodes [v0, m01, v1, m02, m12, v2]:
// {0,1,3}, {1,2,4}, {1,4,3}, {3,4,5} (Kernels.h:1284-1292), midpoints 0.5 (a + b).
// Written as three selects per coordinate: an indexed node table with a runtime
// child index compiles to hundreds of compare/select instructions.
__device__ __forceinline__ void bm_child(const float* v, int c, float* o) {
#pragma clang fp contract(off)
  const bool c0 = c == 0, c1 = c == 1, c3 = c == 3;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float m01 = 0.5f * (v[d] + v[3 + d]);
    const float m02 = 0.5f * (v[d] + v[6 + d]);
    const float m12 = 0.5f * (v[3 + d] + v[6 + d]);
    o[d] = c0 ? v[d] : (c3 ? m02 : m01);            // {0, 1, 1, 3}
    o[3 + d] = c0 ? m01 : (c1 ? v[3 + d] : m12);    // {1, 2, 4, 4}
    o[6 + d] = c1 ? m12 : (c3 ? v[6 + d] : m02);    // {3, 4, 3, 5}
  }
}

// s / 3 correctly rounded: q = s RN(1/3), r = fma(-3, q, s) (exact), q + r RN(1/3)
// rounded once (Markstein's correction step).  Checked against IEEE division for
// every float whose quotient is normal (oracle orc_check_div3).
__device__ __forceinline__ float div3(float s) {
#pragma clang fp contract(off)
  const float y = 0x1.555556p-2f;
  const float q = s * y;
  return __builtin_fmaf(__builtin_fmaf(-3.0f, q, s), y, q);
}

// (x0 + x1 + x2) / 3 in float, as Kernels.h:1243-1248
__device__ __forceinline__ void bm_centroid(const float* v, float& x, float& y, float& z) {
#pragma clang fp contract(off)
  x = div3(v[0] + v[3] + v[6]);
  y = div3(v[1] + v[4] + v[7]);
  z = div3(v[2] + v[5] + v[8]);
}

// the decision of one node pair, with the centroid separation it was taken on
// (the reference's leaf uses those centroids).  t4 = 4 (sqrt(sa_l) + sqrt(ta_l))
// at the pair's level l: the root's value times 2^-l exactly, since the areas are
// quartered exactly and sqrt(x / 4) = sqrt(x) / 2 in IEEE arithmetic.
__device__ __forceinline__ bool bm_sep(const float* sv, const float* tv, float t4, float& ex, float& ey, float& ez,
                                       float& d2) {
  float sx, sy, sz, tx, ty, tz;
  bm_centroid(sv, sx, sy, sz);
  bm_centroid(tv, tx, ty, tz);
  d2 = dist2(tx, ty, tz, sx, sy, sz, ex, ey, ez);
  return well_sep(d2, t4);
}

// a leaf on a separation already formed: G += w K d
__device__ __forceinline__ void bm_leaf_d(float ex, float ey, float ez, float d2, double w, double* G) {
  const float rr = frsq(d2);
  const double k = w * (double)(rr * rr * rr);
  G[0] += k * ex; G[1] += k * ey; G[2] += k * ez;
}

// the six nodes [v0, m01, v1, m02, m12, v2] of triangle v as the reference forms
// them (Kernels.h:1276-1281)
__device__ __forceinline__ void bm_nodes(const float* v, float* n) {
#pragma clang fp contract(off)
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    n[d] = v[d];
    n[3 + d] = 0.5f * (v[d] + v[3 + d]);
    n[6 + d] = v[3 + d];
    n[9 + d] = 0.5f * (v[d] + v[6 + d]);
    n[12 + d] = 0.5f * (v[3 + d] + v[6 + d]);
    n[15 + d] = v[6 + d];
  }
}

// centroids of the 4 children of a triangle with nodes n, in FP64
__device__ __forceinline__ void bm_child_centroids(const float* n, double (*c)[3]) {
  constexpr int ID[4][3] = {{0, 1, 3}, {1, 2, 4}, {1, 4, 3}, {3, 4, 5}};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
      c[k][d] = ((double)n[3 * ID[k][0] + d] + (double)n[3 * ID[k][1] + d] + (double)n[3 * ID[k][2] + d]) *
                (1.0 / 3.0);
  }
}

// the 16 child pairs of node pair (sv, tv) as leaves (their level is maxlev)
__device__ __forceinline__ void bm_leaves16(const float* sv, const float* tv, double w, double* G) {
  constexpr int ID[4][3] = {{0, 1, 3}, {1, 2, 4}, {1, 4, 3}, {3, 4, 5}};
  double ct[4][3];
  {
    float tn[18];
    bm_nodes(tv, tn);
    bm_child_centroids(tn, ct);
  }
  float sn[18];
  bm_nodes(sv, sn);
#pragma unroll 1
  for (int p = 0; p < 4; ++p) {
    double cs[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      cs[d] = ((double)sn[3 * ID[p][0] + d] + (double)sn[3 * ID[p][1] + d] + (double)sn[3 * ID[p][2] + d]) *
              (1.0 / 3.0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double ex = ct[q][0] - cs[0], ey = ct[q][1] - cs[1], ez = ct[q][2] - cs[2];
      const float rr = frsq((float)(ex * ex + ey * ey + ez * ez));
      const double k = w * (double)(rr * rr * rr);
      G[0] += k * ex; G[1] += k * ey; G[2] += k * ez;
    }
  }
}

struct BmNearArgs {
  const float4* srec; const float4* trec;
  const uint2* list;
  const unsigned long long* count;
  unsigned long long cap;    // list capacity (entries beyond it were not written)
  int64_t nt;
  int32_t nunk, maxlev;
  float* coeffs;
};

__device__ __forceinline__ void bm_corners(const float4* r, float* v) {
  const float4 a = r[4], b = r[5], c = r[6];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w; v[8] = c.x;
}

// One wave per listed pair.  Level 1: lane c < 16 decides child pair c (source
// child c/4, target child c%4) and evaluates it if separated.  Level 2: the 16
// children of each open level-1 pair are spread over the wave (item k: the
// (k/16)-th open pair, child k%16); a separated one is a leaf, an open one below
// maxlev is walked by its lane (level 3 all leaves when maxlev is 3).  The
// level-1 triangles of the open pairs pass through LDS.  The tree
// decisions are the reference's float arithmetic; leaves below a decision level
// take FP64 centroids of the float nodes.
constexpr int BM_L1 = 20;   // floats per level-1 pair in LDS (18 used; rows 16-byte aligned)

template <int MAXLEV>
__global__ __launch_bounds__(BLOCK) void k_bm_near(BmNearArgs a) {
  // each wave's 16 level-1 triangle pairs, shared with its items through LDS; a
  // wave reads only its own rows, and LDS operations of one wave complete in order
  __shared__ float4 l1s[(BLOCK / 64) * 16 * BM_L1 / 4];
  const int lane = threadIdx.x & 63;
  float* l1 = (float*)l1s + (threadIdx.x >> 6) * 16 * BM_L1;
  const int64_t wave = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * BLOCK) >> 6;
  const int64_t listed = (int64_t)(*a.count < a.cap ? *a.count : a.cap);
  const int nunk = a.nunk;
  constexpr int maxlev = MAXLEV;
  const int64_t R = (int64_t)nunk * a.nt;
  for (int64_t e = wave; e < listed; e += nwaves) {
    const uint2 p = a.list[e];
    const int64_t i = p.x, j = p.y;
    const float4* sr = a.srec + j * BM_F4;
    const float4* tr = a.trec + i * BM_F4;
    float s0[9], t0[9];
    bm_corners(sr, s0);
    bm_corners(tr, t0);
    const float sa = sr[1].x;
    float t4;
    {
#pragma clang fp contract(off)
      t4 = (sr[0].w + tr[0].w) * 4.0f;
    }
    const float t41 = 0.5f * t4, t42 = 0.25f * t4, t43 = 0.125f * t4;
    double G[3] = {0.0, 0.0, 0.0};
    if constexpr (maxlev == 1) {
      if (lane == 0) bm_leaves16(s0, t0, 0.0625, G);
    } else {
      // level 1 on lanes 0..15; the open pairs' triangles go to LDS for the items
      bool open1 = false;
      float sv1[9], tv1[9];
      if (lane < 16) {
        bm_child(s0, lane >> 2, sv1);
        bm_child(t0, lane & 3, tv1);
        float ex, ey, ez, d2;
        if (bm_sep(sv1, tv1, t41, ex, ey, ez, d2)) bm_leaf_d(ex, ey, ez, d2, 0.0625, G);
        else open1 = true;
      }
      const unsigned long long m1 = __builtin_amdgcn_ballot_w64(open1);
      if (open1) {
        // the open pairs' rows in lane order, so that item k reads row k / 16
        const unsigned rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m1 >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((unsigned)m1, 0u));
        float4* row = (float4*)(l1 + rank * BM_L1);
        row[0] = make_float4(sv1[0], sv1[1], sv1[2], sv1[3]);
        row[1] = make_float4(sv1[4], sv1[5], sv1[6], sv1[7]);
        row[2] = make_float4(sv1[8], tv1[0], tv1[1], tv1[2]);
        row[3] = make_float4(tv1[3], tv1[4], tv1[5], tv1[6]);
        row[4] = make_float4(tv1[7], tv1[8], 0.f, 0.f);
      }
      __builtin_amdgcn_wave_barrier();
      const int items = 16 * __builtin_popcountll(m1);
      for (int k = lane; k < items; k += 64) {
        const int c2 = k & 15;
        float s1v[9], t1v[9];
        {
          const float4* row = (const float4*)(l1 + (k >> 4) * BM_L1);
          const float4 r0 = row[0], r1 = row[1], r2 = row[2], r3 = row[3], r4 = row[4];
          s1v[0] = r0.x; s1v[1] = r0.y; s1v[2] = r0.z; s1v[3] = r0.w;
          s1v[4] = r1.x; s1v[5] = r1.y; s1v[6] = r1.z; s1v[7] = r1.w; s1v[8] = r2.x;
          t1v[0] = r2.y; t1v[1] = r2.z; t1v[2] = r2.w; t1v[3] = r3.x; t1v[4] = r3.y;
          t1v[5] = r3.z; t1v[6] = r3.w; t1v[7] = r4.x; t1v[8] = r4.y;
        }
        float s2v[9], t2v[9];
        bm_child(s1v, c2 >> 2, s2v);
        bm_child(t1v, c2 & 3, t2v);
        float ex, ey, ez, d2;
        if constexpr (maxlev == 2) {
          float sx, sy, sz, tx, ty, tz;
          bm_centroid(s2v, sx, sy, sz);
          bm_centroid(t2v, tx, ty, tz);
          d2 = dist2(tx, ty, tz, sx, sy, sz, ex, ey, ez);
          bm_leaf_d(ex, ey, ez, d2, 0.00390625, G);
          continue;
        }
        if (bm_sep(s2v, t2v, t42, ex, ey, ez, d2)) {
          bm_leaf_d(ex, ey, ez, d2, 0.00390625, G);
          continue;
        }
        if constexpr (maxlev == 3) {
          bm_leaves16(s2v, t2v, 0.000244140625, G);
          continue;
        } else {
        for (int c3 = 0; c3 < 16; ++c3) {
          float s3v[9], t3v[9];
          bm_child(s2v, c3 >> 2, s3v);
          bm_child(t2v, c3 & 3, t3v);
          if (bm_sep(s3v, t3v, t43, ex, ey, ez, d2)) bm_leaf_d(ex, ey, ez, d2, 0.000244140625, G);
          else bm_leaves16(s3v, t3v, 0.0000152587890625, G);
        }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) G[d] += __shfl_xor(G[d], o, 64);
    }
    // lane k*nunk + m writes coefficient (unknown k, component m)
    if (lane < nunk * nunk) {
      const int k = lane / nunk, m = lane % nunk;
      const float4 s1 = sr[1], s2 = sr[2];
      const float4 t1 = tr[1], t2 = tr[2], t3 = tr[3];
      const double w = (double)sa * (1.0 / (4.0 * M_PI));
      double ux, uy, uz;
      if (k < 2) {
        const double bx = k == 0 ? s1.y : s2.x, by = k == 0 ? s1.z : s2.y, bz = k == 0 ? s1.w : s2.z;
        ux = G[2] * by - G[1] * bz;
        uy = G[0] * bz - G[2] * bx;
        uz = G[1] * bx - G[0] * by;
      } else {
        ux = G[0]; uy = G[1]; uz = G[2];
      }
      double tx, ty, tz;
      if (m == 0) { tx = t1.y; ty = t1.z; tz = t1.w; }
      else if (m == 1) { tx = t2.x; ty = t2.y; tz = t2.z; }
      else { tx = t3.x; ty = t3.y; tz = t3.z; }
      a.coeffs[(j * nunk + k) * R + i * nunk + m] = (float)(w * (ux * tx + uy * ty + uz * tz));
    }
  }
}

__global__ void k_bm_zero(unsigned long long* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

int check_surface(const o3d_panels* p, const char* what) {
  if (!p) return fail(O3D_EINVAL, "%s panels are NULL", what);
  if (p->np < 0 || p->nn < 0) return fail(O3D_EINVAL, "%s: negative panel or node count", what);
  if (p->np == 0) return O3D_OK;
  if (p->nn <= 0 || !p->idx || !p->area) return fail(O3D_EINVAL, "%s: panels without nodes, indices or areas", what);
  for (int d = 0; d < 3; ++d)
    if (!p->x[d]) return fail(O3D_EINVAL, "%s: node component %d is NULL", what, d);
  if (p->np > INT32_MAX) return fail(O3D_EINVAL, "%s: at most %d panels", what, INT32_MAX);
  return O3D_OK;
}

void launch_geom(const o3d_panels* p, float4* rec, hipStream_t st) {
  BmGeomArgs g{p->x[0], p->x[1], p->x[2], p->idx, p->area, p->np, rec};
  hipLaunchKernelGGL(k_bm_geom, dim3((unsigned)((p->np + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, g);
}

}  // namespace

// the whole block on stream st; prof (optional) gets the stage times and the
// listed pair count (it synchronises per stage)
int bem_matrix_dev(const o3d_panels* src, const o3d_panels* targ, int32_t nunk, int32_t same, int32_t maxlev,
                   float* coeffs, hipStream_t st, o3d_bem_matrix_profile* prof) {
  int rc;
  if ((rc = check_surface(src, "src")) != O3D_OK) return rc;
  if ((rc = check_surface(targ, "targ")) != O3D_OK) return rc;
  if (nunk != 2 && nunk != 3)
    return fail(O3D_EINVAL, "unknowns per panel %d: must be 2 or 3 (Surfaces.h:261)", nunk);
  if (maxlev < 0) maxlev = 3;
  if (maxlev > BM_MAXLEV) return fail(O3D_EINVAL, "maxlev %d: at most %d", maxlev, BM_MAXLEV);
  if (same && src->np != targ->np)
    return fail(O3D_EINVAL, "same surface with %lld source and %lld target panels", (long long)src->np,
                (long long)targ->np);
  const int64_t ns = src->np, nt = targ->np;
  if (ns == 0 || nt == 0) return O3D_OK;
  if (!coeffs) return fail(O3D_EINVAL, "coeffs is NULL");
  if ((rc = ensure_init()) != O3D_OK) return rc;
  // the pair list is sized from a count read back to the host between the passes
  if (stream_capturing(st))
    return fail(O3D_EINVAL, "the BEM matrix reads its pair count back between passes and cannot be captured");
  const size_t recb = ((sizeof(float4) * BM_F4 * (size_t)(ns + nt)) + 255) & ~(size_t)255;
  std::lock_guard<std::mutex> lk(stream_ws_mutex());
  void* ws = nullptr;
  if ((rc = stream_workspace(st, recb + 256, &ws)) != O3D_OK) return rc;
  // ev: start, counted, one-leaf pass done, recursed pairs done
  struct Events {
    hipEvent_t e[4] = {};
    ~Events() {
      for (auto x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } evs;
  hipEvent_t* ev = evs.e;
  if (prof) {
    for (int k = 0; k < 4; ++k) HIPCHK(hipEventCreate(&ev[k]));
    HIPCHK(hipEventRecord(ev[0], st));
  }
  auto layout = [&](void* base, float4** srec, float4** trec, unsigned long long** cnt, uint2** list) {
    *srec = (float4*)base;
    *trec = *srec + BM_F4 * ns;
    *cnt = (unsigned long long*)((char*)base + recb);
    *list = (uint2*)((char*)base + recb + 256);
  };
  float4 *srec, *trec;
  unsigned long long* cnt;
  uint2* list;
  layout(ws, &srec, &trec, &cnt, &list);
  launch_geom(src, srec, st);
  launch_geom(targ, trec, st);
  hipLaunchKernelGGL(k_bm_zero, dim3(1), dim3(64), 0, st, cnt);
  BmFarArgs fa{};
  fa.srec = srec; fa.trec = trec; fa.ns = ns; fa.nt = nt;
  fa.nunk = nunk; fa.same = same ? 1 : 0; fa.maxlev = maxlev;
  fa.coeffs = coeffs; fa.count = cnt;
  fa.vec4 = ((uintptr_t)coeffs % 16 == 0) && ((int64_t)nunk * nt) % 4 == 0;
  // about 8 workgroups per CU: target blocks x source chunks of whole LDS tiles
  const int64_t nbx = (nt + BLOCK - 1) / BLOCK;
  const int64_t ntiles = (ns + BM_TILE - 1) / BM_TILE;
  int64_t ny = (8LL * num_cus() + nbx - 1) / nbx;
  if (ny > ntiles) ny = ntiles;
  if (ny > 65535) ny = 65535;
  if (ny < 1) ny = 1;
  fa.schunk = ((ntiles + ny - 1) / ny) * BM_TILE;
  ny = (ns + fa.schunk - 1) / fa.schunk;
  const dim3 grid((unsigned)nbx, (unsigned)ny);
  hipLaunchKernelGGL(k_bm_far<BM_COUNT>, grid, dim3(BLOCK), 0, st, fa);
  HIPCHK(hipGetLastError());
  unsigned long long nnear = 0;
  HIPCHK(hipMemcpyAsync(&nnear, cnt, sizeof(nnear), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (prof) HIPCHK(hipEventRecord(ev[1], st));
  // room for the list; the workspace may move, so the records are made again
  if ((rc = stream_workspace(st, recb + 256 + sizeof(uint2) * (size_t)nnear, &ws)) != O3D_OK) return rc;
  layout(ws, &srec, &trec, &cnt, &list);
  launch_geom(src, srec, st);
  launch_geom(targ, trec, st);
  hipLaunchKernelGGL(k_bm_zero, dim3(1), dim3(64), 0, st, cnt);
  fa.srec = srec; fa.trec = trec; fa.count = cnt; fa.list = list; fa.cap = nnear;
  hipLaunchKernelGGL(k_bm_far<BM_WRITE>, grid, dim3(BLOCK), 0, st, fa);
  HIPCHK(hipGetLastError());
  if (prof) HIPCHK(hipEventRecord(ev[2], st));
  if (nnear > 0 && maxlev > 0) {
    BmNearArgs na{srec, trec, list, cnt, nnear, nt, nunk, maxlev, coeffs};
    const int64_t waves_per_block = BLOCK / 64;
    int64_t nb = ((int64_t)nnear + waves_per_block - 1) / waves_per_block;
    const int64_t maxb = 64LL * num_cus();
    if (nb > maxb) nb = maxb;
    if (maxlev == 1) hipLaunchKernelGGL(k_bm_near<1>, dim3((unsigned)nb), dim3(BLOCK), 0, st, na);
    else if (maxlev == 2) hipLaunchKernelGGL(k_bm_near<2>, dim3((unsigned)nb), dim3(BLOCK), 0, st, na);
    else if (maxlev == 3) hipLaunchKernelGGL(k_bm_near<3>, dim3((unsigned)nb), dim3(BLOCK), 0, st, na);
    else hipLaunchKernelGGL(k_bm_near<4>, dim3((unsigned)nb), dim3(BLOCK), 0, st, na);
    HIPCHK(hipGetLastError());
  }
  if (prof) {
    HIPCHK(hipEventRecord(ev[3], st));
    HIPCHK(hipEventSynchronize(ev[3]));
    float a = 0, b = 0, c = 0;
    HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
    HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
    HIPCHK(hipEventElapsedTime(&c, ev[2], ev[3]));
    prof->count_ms = a;
    prof->far_ms = b;
    prof->near_ms = c;
    prof->total_ms = (double)a + b + c;
    prof->near_pairs = (double)nnear;
  }
  return O3D_OK;
}

}  // namespace o3d

using namespace o3d;

extern "C" {

int o3d_bem_matrix_dev(const o3d_panels* src, const o3d_panels* targ, int32_t nunk, int32_t same, int32_t maxlev,
                       float* coeffs, void* stream) {
  clear_error();
  return bem_matrix_dev(src, targ, nunk, same, maxlev, coeffs, (hipStream_t)stream, nullptr);
}

int o3d_bem_matrix_profile_dev(const o3d_panels* src, const o3d_panels* targ, int32_t nunk, int32_t same,
                               int32_t maxlev, float* coeffs, void* stream, o3d_bem_matrix_profile* prof) {
  clear_error();
  if (!prof) return fail(O3D_EINVAL, "profile is NULL");
  *prof = o3d_bem_matrix_profile{};
  return bem_matrix_dev(src, targ, nunk, same, maxlev, coeffs, (hipStream_t)stream, prof);
}

int o3d_bem_matrix(const o3d_panels* src, const o3d_panels* targ, int32_t nunk, int32_t same, int32_t maxlev,
                   float* coeffs) {
  clear_error();
  int rc;
  if ((rc = check_surface(src, "src")) != O3D_OK) return rc;
  if ((rc = check_surface(targ, "targ")) != O3D_OK) return rc;
  if (nunk != 2 && nunk != 3)
    return fail(O3D_EINVAL, "unknowns per panel %d: must be 2 or 3 (Surfaces.h:261)", nunk);
  if (src->np == 0 || targ->np == 0) return O3D_OK;
  if (!coeffs) return fail(O3D_EINVAL, "coeffs is NULL");
  if ((rc = check_node_indices(src, "src")) != O3D_OK) return rc;
  if ((rc = check_node_indices(targ, "targ")) != O3D_OK) return rc;
  if ((rc = ensure_init()) != O3D_OK) return rc;
  const size_t nel = (size_t)nunk * nunk * (size_t)src->np * (size_t)targ->np;
  // device copies of both surfaces (nodes 3*nn floats, areas np floats, indices
  // 3*np), the block, then the block back to the host
  auto bytes = [](const o3d_panels* p) { return sizeof(float) * (3 * p->nn + 4 * p->np); };
  void* d = nullptr;
  const size_t total = bytes(src) + bytes(targ) + sizeof(float) * nel;
  if (hipMalloc(&d, total) != hipSuccess) return fail(O3D_ENOMEM, "device buffers of %zu bytes", total);
  hipStream_t st = lib_stream();
  auto up = [&](const o3d_panels* p, float* f, o3d_panels* o) -> int {
    *o = o3d_panels{};
    o->np = p->np; o->nn = p->nn;
    for (int k = 0; k < 3; ++k) {
      o->x[k] = f + k * p->nn;
      HIPCHK(hipMemcpyAsync((void*)o->x[k], p->x[k], sizeof(float) * p->nn, hipMemcpyHostToDevice, st));
    }
    o->area = f + 3 * p->nn;
    HIPCHK(hipMemcpyAsync((void*)o->area, p->area, sizeof(float) * p->np, hipMemcpyHostToDevice, st));
    o->idx = (const uint32_t*)(f + 3 * p->nn + p->np);
    HIPCHK(hipMemcpyAsync((void*)o->idx, p->idx, sizeof(uint32_t) * 3 * p->np, hipMemcpyHostToDevice, st));
    return O3D_OK;
  };
  o3d_panels ds, dt;
  float* f = (float*)d;
  rc = up(src, f, &ds);
  if (rc == O3D_OK) rc = up(targ, f + 3 * src->nn + 4 * src->np, &dt);
  float* dc = f + 3 * src->nn + 4 * src->np + 3 * targ->nn + 4 * targ->np;
  if (rc == O3D_OK) rc = bem_matrix_dev(&ds, same ? &ds : &dt, nunk, same, maxlev, dc, st, nullptr);
  if (rc == O3D_OK && hipMemcpyAsync(coeffs, dc, sizeof(float) * nel, hipMemcpyDeviceToHost, st) != hipSuccess)
    rc = fail(O3D_EDEVICE, "copying the block back failed");
  if (hipStreamSynchronize(st) != hipSuccess && rc == O3D_OK) rc = fail(O3D_EDEVICE, "stream synchronise failed");
  (void)hipFree(d);
  return rc;
}

}  // extern "C"
