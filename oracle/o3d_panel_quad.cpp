// o3d_panel_quad.cpp -- THE ORACLE, second part.  Test infrastructure only.
//
// Panels -> points by a 7-point Gauss rule per panel with a closed-form near
// field: the panel method of the 20,480-panel sphere workload.  It has no
// counterpart in the reference, whose panels_affect_points uses the recursive
// centroid rule (Kernels.h:1028-1211, restated in o3d_oracle.cpp); it keeps the
// reference's panel conventions (Influence.h:557-1099):
//   * ts is the ABSOLUTE vortex strength (sheet strength x area), ss the sheet
//     source strength (NULL = none); targets are points;
//   * results accumulate (+=), unscaled (no 1/(4 pi)); grad[3j+i] = du_i/dx_j;
//   * the kernel is the singular Biot-Savart one, K = 1/r^3 (every core at
//     sigma = 0, CoreFunc.h with sr = 0).
//
// Per (target t, panel p), with c the centroid and D the longest edge:
//   far  when |t - c|^2 >= 4 D^2 (decided in float, left to right, no contraction,
//        so that any implementation decides alike):
//        the 7-point degree-5 rule (Strang & Fix 1973; Dunavant 1985, rule 5):
//        a singular vortex + source point at y_q with strengths w_q ts and
//        w_q ss area (the point kernel of Kernels.h:294-345 at sigma = 0);
//   near otherwise: the exact integral over the flat triangle of the sheet
//        densities gamma = ts / area and sigma = ss:
//          u     += gamma x I + sigma I,
//          du_i/dx_j += (gamma x dI_j)_i + sigma dI_j,i,
//        with I(t) = int_T (t - y)/|t - y|^3 dA(y) = sum_e m_e f_e + Omega n and
//        dI_j = d I / d t_j, where for each edge e (tangent tau_e, outward
//        in-plane normal m_e = tau_e x n):
//          f_e   = int_e ds/|t - y|           (a logarithm of the end distances),
//          Omega = signed solid angle of T    (Van Oosterom & Strackee 1983),
//          grad f_e    = -(q_e c_e + tau_e (1/R+ - 1/R-)),
//          grad Omega  = -sum_e (tau_e x q_e) c_e,
//        q_e the perpendicular from the edge line to t, u-/u+ the edge ends'
//        positions along tau_e relative to t, R-/R+ their distances and
//        c_e = (u+/R+ - u-/R-)/|q_e|^2 (both in cancellation-free forms).
//        All of it in FP64.  Singular on the sheet itself, as the reference's
//        singular kernels are at a leaf centroid.
//
// Parity status: no reference to compare with; pinned by mathematics in
// tests/test_panel_quad.py -- the rule integrates monomials of degree <= 5 on a
// triangle exactly (and not degree 6); I agrees with subdivided quadrature and
// dI with finite differences of I; dI is symmetric and trace-free; I jumps by
// 4 pi n across the sheet; far away a panel is a point vortex; and the whole
// method agrees with a brute-force integral of the sheet over a sphere mesh.
#include <cmath>
#include <cstdint>

#include <omp.h>

namespace {

constexpr int G7 = 7;

// barycentric coordinates (3 per point) and weights (sum 1) of the 7-point rule
void g7_rule(double* bary, double* w) {
  const double s15 = std::sqrt(15.0);
  const double a1 = (6.0 - s15) / 21.0, a2 = (6.0 + s15) / 21.0;
  const double w1 = (155.0 - s15) / 1200.0, w2 = (155.0 + s15) / 1200.0;
  const double pts[G7][3] = {{1.0 / 3, 1.0 / 3, 1.0 / 3},
                             {1 - 2 * a1, a1, a1}, {a1, 1 - 2 * a1, a1}, {a1, a1, 1 - 2 * a1},
                             {1 - 2 * a2, a2, a2}, {a2, 1 - 2 * a2, a2}, {a2, a2, 1 - 2 * a2}};
  const double wts[G7] = {9.0 / 40.0, w1, w1, w1, w2, w2, w2};
  for (int q = 0; q < G7; ++q) {
    for (int k = 0; k < 3; ++k) bary[3 * q + k] = pts[q][k];
    w[q] = wts[q];
  }
}

inline void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
inline double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// I = int_T (x - y)/|x - y|^3 dA and, if dI, dI[3j+l] = dI_l/dx_j; v[3k+d] = vertex k
void tri_integrals(const double* v, const double* x, double* I, double* dI) {
  double e1[3], e2[3], n[3];
  for (int d = 0; d < 3; ++d) {
    e1[d] = v[3 + d] - v[d];
    e2[d] = v[6 + d] - v[d];
  }
  cross(e1, e2, n);
  const double a2 = std::sqrt(dot(n, n));
  for (int d = 0; d < 3; ++d) n[d] /= a2;
  double R[3][3], r[3];
  for (int k = 0; k < 3; ++k) {
    for (int d = 0; d < 3; ++d) R[k][d] = v[3 * k + d] - x[d];
    r[k] = std::sqrt(dot(R[k], R[k]));
  }
  double c12[3];
  cross(R[1], R[2], c12);
  const double num = dot(R[0], c12);
  const double den = r[0] * r[1] * r[2] + dot(R[0], R[1]) * r[2] + dot(R[0], R[2]) * r[1] + dot(R[1], R[2]) * r[0];
  const double om = -2.0 * std::atan2(num, den);
  for (int d = 0; d < 3; ++d) I[d] = om * n[d];
  double gom[3] = {0, 0, 0};
  if (dI)
    for (int k = 0; k < 9; ++k) dI[k] = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double* a = v + 3 * k;
    const double* b = v + 3 * ((k + 1) % 3);
    double t[3], m[3], q[3];
    for (int d = 0; d < 3; ++d) t[d] = b[d] - a[d];
    const double len = std::sqrt(dot(t, t));
    for (int d = 0; d < 3; ++d) t[d] /= len;
    cross(t, n, m);
    double da[3];
    for (int d = 0; d < 3; ++d) da[d] = x[d] - a[d];
    const double s = dot(da, t);
    const double um = -s, up = len - s;
    for (int d = 0; d < 3; ++d) q[d] = da[d] - s * t[d];
    const double r02 = dot(q, q);
    const double rm = r[k], rp = r[(k + 1) % 3];
    double f;
    if (um >= 0) f = std::log((rp + up) / (rm + um));
    else if (up <= 0) f = std::log((rm - um) / (rp - up));
    else f = std::log((rp + up) * (rm - um) / r02);
    for (int d = 0; d < 3; ++d) I[d] += m[d] * f;
    if (!dI) continue;
    const double c = um * up > 0 ? (up * up - um * um) / (rp * rm * (up * rm + um * rp)) : (up / rp - um / rm) / r02;
    const double e = 1.0 / rp - 1.0 / rm;
    double tq[3];
    cross(t, q, tq);
    for (int j = 0; j < 3; ++j) {
      const double df = -(q[j] * c + t[j] * e);
      for (int l = 0; l < 3; ++l) dI[3 * j + l] += df * m[l];
      gom[j] -= tq[j] * c;
    }
  }
  if (dI)
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < 3; ++l) dI[3 * j + l] += gom[j] * n[l];
}

// the near/far decision, in float with the operations in this order and no contraction
bool pair_is_near(const float* v, float tx, float ty, float tz) {
  const float cx = (v[0] + v[3] + v[6]) / 3.0f;
  const float cy = (v[1] + v[4] + v[7]) / 3.0f;
  const float cz = (v[2] + v[5] + v[8]) / 3.0f;
  float e[3];
  for (int k = 0; k < 3; ++k) {
    const float* a = v + 3 * k;
    const float* b = v + 3 * ((k + 1) % 3);
    const float dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    e[k] = dx * dx + dy * dy + dz * dz;
  }
  const float d2max = std::fmax(std::fmax(e[0], e[1]), e[2]);
  const float ex = tx - cx, ey = ty - cy, ez = tz - cz;
  const float dist2 = ex * ex + ey * ey + ez * ez;
  return dist2 < 4.0f * d2max;
}

}  // namespace

extern "C" {

void orc_g7_rule(double* bary, double* w) { g7_rule(bary, w); }

// one triangle (v[9], FP64) at one point x[3]: I[3] and dI[9] (dI may be NULL)
int orc_tri_integrals(const double* v, const double* x, double* I, double* dI) {
  tri_integrals(v, x, I, dI);
  return 0;
}

// 1 where the float decision takes the closed form for (target i, panel p)
int orc_pan_g7_near(const float* v, float tx, float ty, float tz) { return pair_is_near(v, tx, ty, tz) ? 1 : 0; }

// panels -> points by the 7-point rule with the closed-form near field, FP64.
// Inputs are the product's float arrays; outputs double, accumulated.
int orc_pan_g7_d(int grads, int64_t np, const float* x, const float* y, const float* z, const uint32_t* idx,
                 const float* tsx, const float* tsy, const float* tsz, const float* area, const float* ss, int64_t nt,
                 const float* tx, const float* ty, const float* tz, double* tu, double* tv, double* tw,
                 double* const* tug, int64_t* near_pairs, int nthreads) {
  if (grads && !tug) return 2;
  double bary[3 * G7], w[G7];
  g7_rule(bary, w);
  int64_t nnear = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nnear)
  for (int64_t i = 0; i < nt; ++i) {
    const double t[3] = {tx[i], ty[i], tz[i]};
    double u[3] = {0, 0, 0}, g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t p = 0; p < np; ++p) {
      float vf[9];
      double v[9];
      for (int k = 0; k < 3; ++k) {
        const uint32_t n = idx[3 * p + k];
        vf[3 * k] = x[n]; vf[3 * k + 1] = y[n]; vf[3 * k + 2] = z[n];
      }
      for (int k = 0; k < 9; ++k) v[k] = vf[k];
      const double ar = area[p];
      const double ts[3] = {tsx[p], tsy[p], tsz[p]};
      const double sg = ss ? (double)ss[p] : 0.0;
      if (pair_is_near(vf, tx[i], ty[i], tz[i])) {
        ++nnear;
        double I[3], dI[9];
        tri_integrals(v, t, I, grads ? dI : nullptr);
        const double gam[3] = {ts[0] / ar, ts[1] / ar, ts[2] / ar};
        double gi[3];
        cross(gam, I, gi);
        for (int d = 0; d < 3; ++d) u[d] += gi[d] + sg * I[d];
        if (grads)
          for (int j = 0; j < 3; ++j) {
            double gj[3];
            cross(gam, dI + 3 * j, gj);
            for (int d = 0; d < 3; ++d) g[3 * j + d] += gj[d] + sg * dI[3 * j + d];
          }
        continue;
      }
      for (int q = 0; q < G7; ++q) {
        double d[3];
        for (int k = 0; k < 3; ++k)
          d[k] = t[k] - (bary[3 * q] * v[k] + bary[3 * q + 1] * v[3 + k] + bary[3 * q + 2] * v[6 + k]);
        const double r2 = dot(d, d);
        const double K = 1.0 / (r2 * std::sqrt(r2));
        const double B = -3.0 * K / r2;
        const double gv[3] = {w[q] * ts[0], w[q] * ts[1], w[q] * ts[2]};
        const double gs = w[q] * sg * ar;
        double c[3];
        cross(gv, d, c);
        for (int k = 0; k < 3; ++k) c[k] += gs * d[k];
        for (int k = 0; k < 3; ++k) u[k] += K * c[k];
        if (!grads) continue;
        // du_i/dx_j = B d_j c_i + K (eps_ikj g_k + gs delta_ij)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) g[3 * j + k] += B * d[j] * c[k];
        g[1] += K * gv[2]; g[2] -= K * gv[1];
        g[3] -= K * gv[2]; g[5] += K * gv[0];
        g[6] += K * gv[1]; g[7] -= K * gv[0];
        g[0] += K * gs; g[4] += K * gs; g[8] += K * gs;
      }
    }
    tu[i] += u[0]; tv[i] += u[1]; tw[i] += u[2];
    if (grads)
      for (int k = 0; k < 9; ++k) tug[k][i] += g[k];
  }
  if (near_pairs) *near_pairs = nnear;
  return 0;
}

}  // extern "C"
