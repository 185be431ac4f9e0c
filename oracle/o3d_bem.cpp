// o3d_bem.cpp -- THE ORACLE, third part.  Test infrastructure only.
//
// The BEM right-hand side of one boundary surface from its panel-centre
// velocities, restated from the reference with its own float/double mixing:
//   finalize_vels       Surfaces.h:877-887     pu = fs + pu * (0.25/pi)      (double, stored float)
//   add_body_motion     Surfaces.h:556-597     pu += f * (float)vel          (float)
//                                              pu += f * (rotvel x (c - tc)) (double, stored float);
//                                              c = (x0 + x1 + x2)/3.0 in double from a float sum,
//                                              minus the float tc, rounded to float
//   compute_bases       Surfaces.h:766-816     x1 = (n1 - n0) * (1.0/base), x2 = ((n2 - n0) - dp x1)
//                                              * (1.0/height) (the scalings in double), norm = x1 x x2
//                                              (also exported alone: orc_bem_bases, with the areas)
//   represent_as_particles  Surfaces.h:959-1010  colocation points: centroid + offset * normal
//                                              (orc_colocation, for panels_affect_panels)
//   vels_to_rhs_panels  RHS.h:47-98            rhs = -(pu . x1), -(pu . x2) [, -(pu . norm)], or
//                                              -(pu . norm) alone for one unknown
// as BEMHelper.h:83-117 chains them (finalize, then the body motion with factor -1).
// Float operations are left to right with no contraction (the Makefile builds with
// -ffp-contract=off).
//
// Parity status: Surfaces.h cannot be compiled here (it needs Eigen3 through
// Body.h), so the rounding sequence is pinned by reading only; the values are
// pinned by mathematics in tests/test_bem_rhs.py (orthonormal bases along the
// first edge, the freestream alone, a rigid rotation, a body translating with
// the stream, and the unknown counts 1, 2, 3).
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

// Surfaces::compute_bases (Surfaces.h:766-816) for panel i: x1 along edge 0-1,
// x2 toward node 2 and orthogonal to x1, nm = x1 x x2; the scalings by 1/base and
// 1/height in double, stored float; area = 0.5 * base * height in double
void panel_basis(const float* const xs[3], const uint32_t* idx, int64_t i, float* x1, float* x2, float* nm,
                 float* area) {
  const uint32_t id0 = idx[3 * i], id1 = idx[3 * i + 1], id2 = idx[3 * i + 2];
  for (int j = 0; j < 3; ++j) x1[j] = xs[j][id1] - xs[j][id0];
  const float base = std::sqrt(x1[0] * x1[0] + x1[1] * x1[1] + x1[2] * x1[2]);
  for (int j = 0; j < 3; ++j) x1[j] = (float)((double)x1[j] * (1.0 / (double)base));
  for (int j = 0; j < 3; ++j) x2[j] = xs[j][id2] - xs[j][id0];
  const float dp = x2[0] * x1[0] + x2[1] * x1[1] + x2[2] * x1[2];
  for (int j = 0; j < 3; ++j) x2[j] -= dp * x1[j];
  const float height = std::sqrt(x2[0] * x2[0] + x2[1] * x2[1] + x2[2] * x2[2]);
  for (int j = 0; j < 3; ++j) x2[j] = (float)((double)x2[j] * (1.0 / (double)height));
  if (area) *area = (float)(0.5 * (double)base * (double)height);
  nm[0] = x1[1] * x2[2] - x1[2] * x2[1];
  nm[1] = x1[2] * x2[0] - x1[0] * x2[2];
  nm[2] = x1[0] * x2[1] - x1[1] * x2[0];
}

}  // namespace

extern "C" {

// the bases and areas of np panels: b[0..2] x1, b[3..5] x2, b[6..8] normal (np
// floats each); area may be NULL
int orc_bem_bases(int64_t np, const float* x, const float* y, const float* z, const uint32_t* idx, float* const* b,
                  float* area) {
  const float* xs[3] = {x, y, z};
  for (int64_t i = 0; i < np; ++i) {
    float x1[3], x2[3], nm[3];
    panel_basis(xs, idx, i, x1, x2, nm, area ? area + i : nullptr);
    for (int d = 0; d < 3; ++d) { b[d][i] = x1[d]; b[3 + d][i] = x2[d]; b[6 + d][i] = nm[d]; }
  }
  return 0;
}

// Surfaces::represent_as_particles (Surfaces.h:959-1010) positions, the colocation
// points panels_affect_panels evaluates at (Influence.h:1233-1237): centroid as
// (1./3.) * (float sum of the nodes) in double, stored float, plus offset times
// the compute_bases normal in float
int orc_colocation(int64_t np, const float* x, const float* y, const float* z, const uint32_t* idx, float offset,
                   float* px, float* py, float* pz) {
  const float* xs[3] = {x, y, z};
  float* out[3] = {px, py, pz};
  for (int64_t i = 0; i < np; ++i) {
    float x1[3], x2[3], nm[3];
    panel_basis(xs, idx, i, x1, x2, nm, nullptr);
    const uint32_t id0 = idx[3 * i], id1 = idx[3 * i + 1], id2 = idx[3 * i + 2];
    for (int j = 0; j < 3; ++j) {
      float c = (float)((1. / 3.) * (double)(xs[j][id0] + xs[j][id1] + xs[j][id2]));
      c += offset * nm[j];
      out[j][i] = c;
    }
  }
  return 0;
}

// pu: panel velocities as points_affect_panels left them (3 rows of np floats),
// overwritten with the finalized velocities minus the body motion.  body: 0 =
// no body or the ground (no motion term, Surfaces.h:561-562).  rhs: nunk * np.
int orc_bem_rhs(int64_t np, const float* x, const float* y, const float* z, const uint32_t* idx, float* pu0,
                float* pu1, float* pu2, const double* fs, int body, const double* vel, const double* rotvel,
                const float* tc, int nunk, float* rhs) {
  if (nunk < 1 || nunk > 3) return 2;
  float* pu[3] = {pu0, pu1, pu2};
  const double factor = 0.25 / M_PI;
  for (int d = 0; d < 3; ++d)
    for (int64_t i = 0; i < np; ++i) pu[d][i] = (float)(fs[d] + (double)pu[d][i] * factor);
  const float* xs[3] = {x, y, z};
  for (int64_t i = 0; i < np; ++i) {
    const uint32_t id0 = idx[3 * i], id1 = idx[3 * i + 1], id2 = idx[3 * i + 2];
    if (body) {
      const float f = -1.0f;
      for (int d = 0; d < 3; ++d) pu[d][i] += f * (float)vel[d];
      const float xc = (float)((double)(xs[0][id0] + xs[0][id1] + xs[0][id2]) / 3.0 - (double)tc[0]);
      const float yc = (float)((double)(xs[1][id0] + xs[1][id1] + xs[1][id2]) / 3.0 - (double)tc[1]);
      const float zc = (float)((double)(xs[2][id0] + xs[2][id1] + xs[2][id2]) / 3.0 - (double)tc[2]);
      pu[0][i] = (float)((double)pu[0][i] + (double)f * (rotvel[1] * zc - rotvel[2] * yc));
      pu[1][i] = (float)((double)pu[1][i] + (double)f * (rotvel[2] * xc - rotvel[0] * zc));
      pu[2][i] = (float)((double)pu[2][i] + (double)f * (rotvel[0] * yc - rotvel[1] * xc));
    }
    float x1[3], x2[3], nm[3];
    panel_basis(xs, idx, i, x1, x2, nm, nullptr);
    const float u0 = pu[0][i], u1 = pu[1][i], u2 = pu[2][i];
    const float t1 = -(u0 * x1[0] + u1 * x1[1] + u2 * x1[2]);
    const float t2 = -(u0 * x2[0] + u1 * x2[1] + u2 * x2[2]);
    const float tn = -(u0 * nm[0] + u1 * nm[1] + u2 * nm[2]);
    if (nunk == 1) {
      rhs[i] = tn;
    } else if (nunk == 2) {
      rhs[2 * i] = t1;
      rhs[2 * i + 1] = t2;
    } else {
      rhs[3 * i] = t1;
      rhs[3 * i + 1] = t2;
      rhs[3 * i + 2] = tn;
    }
  }
  return 0;
}

}  // extern "C"

extern "C" {

// Checks the division by 3 that the GPU uses for centroids in its tree decisions
// against IEEE division: q = s * RN(1/3), r = fma(-3, q, s), q' = fma(r, RN(1/3), q)
// (Markstein's correction step), over the float bit patterns [lo, hi) whose value
// and quotient are normal.  Returns the number of mismatches.
int64_t orc_check_div3(int64_t lo, int64_t hi) {
  const float y = 1.0f / 3.0f;
  int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (int64_t b = lo; b < hi; ++b) {
    const uint32_t u = (uint32_t)b;
    float s;
    std::memcpy(&s, &u, 4);
    const float as = std::fabs(s);
    if (!(as >= 3.0f * 1.17549435e-38f) || !(as <= 3.4028234e38f)) continue;
    const float q = s * y;
    const float r = std::fma(-3.0f, q, s);
    const float q2 = std::fma(r, y, q);
    const float ref = s / 3.0f;
    if (std::memcmp(&q2, &ref, 4) != 0) ++bad;
  }
  return bad;
}

}  // extern "C"
