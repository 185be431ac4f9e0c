// ref_harness.cpp -- drives the reference's OWN scalar kernels (test infrastructure).
//
// Compiled by oracle/build_ref.sh against Omega3D's src/Kernels.h, src/CoreFunc.h
// and src/MathHelper.h where they lie under the reference checkout (one build per
// core; the core is a compile-time #define in CoreFunc.h:35-39).  Nothing of the
// reference is copied into this repository: the build reads the headers in place
// (CoreFunc.h through a temporary, core-selected copy outside the repository) and
// writes only oracle/_ref/libo3dref_<core>.so.
//
// The loops below restate the scalar branches of points_affect_points
// (Influence.h:405-440 thick targets, :280-307 point targets) and
// panels_affect_points (Influence.h:729-775, 968-1006): zero accumulators of type A
// per target, visit every source in index order, add into the caller's outputs;
// and of points_affect_panels (Influence.h:1190-1210), which runs the panel kernel
// backwards and subtracts from the panels' velocities; and of panels_on_panels_coeff
// (Coefficients.h:169-444), the BEM influence matrix over rkernel_2vs_2p.

#include "Kernels.h"
#include <cmath>
#include <cstdint>

namespace {

template <class S, class A>
void ref_p2p_loop(int grads, int thick, int64_t ns, const S* sx, const S* sy, const S* sz, const S* sr,
                  const S* ssx, const S* ssy, const S* ssz, int64_t nt, const S* tx, const S* ty,
                  const S* tz, const S* tr, S* tu, S* tv, S* tw, S* const* g) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nt; ++i) {
    A u = 0, v = 0, w = 0, ux = 0, vx = 0, wx = 0, uy = 0, vy = 0, wy = 0, uz = 0, vz = 0, wz = 0;
    for (int64_t j = 0; j < ns; ++j) {
      if (grads && thick)
        kernel_0v_0bg<S, A>(sx[j], sy[j], sz[j], sr[j], ssx[j], ssy[j], ssz[j], tx[i], ty[i], tz[i], tr[i],
                            &u, &v, &w, &ux, &vx, &wx, &uy, &vy, &wy, &uz, &vz, &wz);
      else if (grads)
        kernel_0v_0pg<S, A>(sx[j], sy[j], sz[j], sr[j], ssx[j], ssy[j], ssz[j], tx[i], ty[i], tz[i],
                            &u, &v, &w, &ux, &vx, &wx, &uy, &vy, &wy, &uz, &vz, &wz);
      else if (thick)
        kernel_0v_0b<S, A>(sx[j], sy[j], sz[j], sr[j], ssx[j], ssy[j], ssz[j], tx[i], ty[i], tz[i], tr[i],
                           &u, &v, &w);
      else
        kernel_0v_0p<S, A>(sx[j], sy[j], sz[j], sr[j], ssx[j], ssy[j], ssz[j], tx[i], ty[i], tz[i],
                           &u, &v, &w);
    }
    tu[i] += u; tv[i] += v; tw[i] += w;
    if (grads) {
      g[0][i] += ux; g[1][i] += vx; g[2][i] += wx;
      g[3][i] += uy; g[4][i] += vy; g[5][i] += wy;
      g[6][i] += uz; g[7][i] += vz; g[8][i] += wz;
    }
  }
}

template <class S, class A>
void ref_pan_loop(int grads, int maxlev, int64_t np, const S* x, const S* y, const S* z, const uint32_t* idx,
                  const S* tsx, const S* tsy, const S* tsz, const S* area, const S* ss, int64_t nt,
                  const S* tx, const S* ty, const S* tz, S* tu, S* tv, S* tw, S* const* g) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nt; ++i) {
    A u = 0, v = 0, w = 0, ux = 0, vx = 0, wx = 0, uy = 0, vy = 0, wy = 0, uz = 0, vz = 0, wz = 0;
    for (int64_t j = 0; j < np; ++j) {
      const size_t a = idx[3*j], b = idx[3*j+1], c = idx[3*j+2];
      const S s0 = ss ? ss[j] : S(0.0);
      if (grads)
        (void) rkernel_2vs_0pg<S, A>(x[a], y[a], z[a], x[b], y[b], z[b], x[c], y[c], z[c],
                                     tsx[j]/area[j], tsy[j]/area[j], tsz[j]/area[j], s0,
                                     tx[i], ty[i], tz[i], area[j], 0, maxlev,
                                     &u, &v, &w, &ux, &vx, &wx, &uy, &vy, &wy, &uz, &vz, &wz);
      else
        (void) rkernel_2vs_0p<S, A>(x[a], y[a], z[a], x[b], y[b], z[b], x[c], y[c], z[c],
                                    tsx[j]/area[j], tsy[j]/area[j], tsz[j]/area[j], s0,
                                    tx[i], ty[i], tz[i], area[j], 0, maxlev, &u, &v, &w);
    }
    tu[i] += u; tv[i] += v; tw[i] += w;
    if (grads) {
      g[0][i] += ux; g[1][i] += vx; g[2][i] += wx;
      g[3][i] += uy; g[4][i] += vy; g[5][i] += wy;
      g[6][i] += uz; g[7][i] += vz; g[8][i] += wz;
    }
  }
}

template <class S, class A>
void ref_ptpan_loop(int maxlev, int64_t ns, const S* sx, const S* sy, const S* sz, const S* ssx, const S* ssy,
                    const S* ssz, int64_t np, const S* x, const S* y, const S* z, const uint32_t* idx,
                    const S* area, S* tu, S* tv, S* tw) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < np; ++i) {
    A u = 0, v = 0, w = 0;
    const size_t a = idx[3*i], b = idx[3*i+1], c = idx[3*i+2];
    for (int64_t j = 0; j < ns; ++j)
      (void) rkernel_2vs_0p<S, A>(x[a], y[a], z[a], x[b], y[b], z[b], x[c], y[c], z[c],
                                  ssx[j]/area[i], ssy[j]/area[i], ssz[j]/area[i], S(0.0),
                                  sx[j], sy[j], sz[j], area[i], 0, maxlev, &u, &v, &w);
    tu[i] -= u; tv[i] -= v; tw[i] -= w;
  }
}

// panels_on_panels_coeff's scalar branch (Coefficients.h:332-397), its self
// block (:399-421) and its scaling (:428-433), driving the reference's own
// rkernel_2vs_2p; bases as Surfaces::compute_bases leaves them (b[0..2] x1,
// b[3..5] x2, b[6..8] normal), column-major output (nunk*ns) x (nunk*nt)
template <class S>
void ref_bem_loop(int maxlev, int nunk, int same, int64_t ns, const S* sx, const S* sy, const S* sz,
                  const uint32_t* si, const S* sa, const S* const* sb, int64_t nt, const S* tx, const S* ty,
                  const S* tz, const uint32_t* ti, const S* ta, const S* const* tb, S* coeffs) {
  const int64_t R = nunk * nt;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t j = 0; j < ns; ++j) {
    const size_t a = si[3*j], b = si[3*j+1], c = si[3*j+2];
    for (int64_t i = 0; i < nt; ++i) {
      const size_t p = ti[3*i], q = ti[3*i+1], r = ti[3*i+2];
      for (int k = 0; k < nunk; ++k) {
        const S gx = k == 0 ? sb[0][j] : k == 1 ? sb[3][j] : S(0.0);
        const S gy = k == 0 ? sb[1][j] : k == 1 ? sb[4][j] : S(0.0);
        const S gz = k == 0 ? sb[2][j] : k == 1 ? sb[5][j] : S(0.0);
        const S gs = k == 2 ? S(1.0) : S(0.0);
        S u = 0.0, v = 0.0, w = 0.0;
        (void) rkernel_2vs_2p<S, S>(sx[a], sy[a], sz[a], sx[b], sy[b], sz[b], sx[c], sy[c], sz[c],
                                    gx, gy, gz, gs,
                                    tx[p], ty[p], tz[p], tx[q], ty[q], tz[q], tx[r], ty[r], tz[r],
                                    sa[j], ta[i], 0, maxlev, &u, &v, &w);
        S* o = coeffs + (j*nunk + k) * R + i*nunk;
        o[0] = u*tb[0][i] + v*tb[1][i] + w*tb[2][i];
        o[1] = u*tb[3][i] + v*tb[4][i] + w*tb[5][i];
        if (nunk == 3) o[2] = u*tb[6][i] + v*tb[7][i] + w*tb[8][i];
      }
    }
    if (same) {
      S* o = coeffs + (j*nunk) * R + j*nunk;
      o[0] = 0.0; o[1] = 2.0*M_PI; if (nunk == 3) o[2] = 0.0;
      o += R;
      o[0] = -2.0*M_PI; o[1] = 0.0; if (nunk == 3) o[2] = 0.0;
      if (nunk == 3) { o += R; o[0] = 0.0; o[1] = 0.0; o[2] = 2.0*M_PI; }
    }
  }
  const S fac = 1.0 / (4.0 * M_PI);
  const int64_t total = R * (int64_t)(nunk * ns);
  for (int64_t e = 0; e < total; ++e) coeffs[e] = coeffs[e] * fac;
}

}  // namespace

extern "C" {

// BEM influence matrix block through the reference's rkernel_2vs_2p, S = float
int ref_bem_matrix_f(int maxlev, int nunk, int same, int64_t ns, const float* sx, const float* sy, const float* sz,
                     const uint32_t* si, const float* sa, const float* const* sb, int64_t nt, const float* tx,
                     const float* ty, const float* tz, const uint32_t* ti, const float* ta, const float* const* tb,
                     float* coeffs) {
  if (nunk != 2 && nunk != 3) return 2;
  ref_bem_loop<float>(maxlev, nunk, same, ns, sx, sy, sz, si, sa, sb, nt, tx, ty, tz, ti, ta, tb, coeffs);
  return 0;
}

// the same with the reference's templates instantiated in double (its own
// double-precision tree decisions)
int ref_bem_matrix_d(int maxlev, int nunk, int same, int64_t ns, const double* sx, const double* sy,
                     const double* sz, const uint32_t* si, const double* sa, const double* const* sb, int64_t nt,
                     const double* tx, const double* ty, const double* tz, const uint32_t* ti, const double* ta,
                     const double* const* tb, double* coeffs) {
  if (nunk != 2 && nunk != 3) return 2;
  ref_bem_loop<double>(maxlev, nunk, same, ns, sx, sy, sz, si, sa, sb, nt, tx, ty, tz, ti, ta, tb, coeffs);
  return 0;
}

// points_affect_panels; accum: 0 = A float, 1 = A double (S = float)
int ref_ptpan_f(int accum, int maxlev, int64_t ns, const float* sx, const float* sy, const float* sz,
                const float* ssx, const float* ssy, const float* ssz, int64_t np, const float* x, const float* y,
                const float* z, const uint32_t* idx, const float* area, float* tu, float* tv, float* tw) {
  if (accum) ref_ptpan_loop<float, double>(maxlev, ns, sx, sy, sz, ssx, ssy, ssz, np, x, y, z, idx, area, tu, tv, tw);
  else       ref_ptpan_loop<float, float>(maxlev, ns, sx, sy, sz, ssx, ssy, ssz, np, x, y, z, idx, area, tu, tv, tw);
  return 0;
}

int ref_ptpan_d(int maxlev, int64_t ns, const double* sx, const double* sy, const double* sz, const double* ssx,
                const double* ssy, const double* ssz, int64_t np, const double* x, const double* y, const double* z,
                const uint32_t* idx, const double* area, double* tu, double* tv, double* tw) {
  ref_ptpan_loop<double, double>(maxlev, ns, sx, sy, sz, ssx, ssy, ssz, np, x, y, z, idx, area, tu, tv, tw);
  return 0;
}


// accum: 0 = A float, 1 = A double (S = float)
int ref_p2p_f(int grads, int accum, int64_t ns, const float* sx, const float* sy, const float* sz,
              const float* sr, const float* ssx, const float* ssy, const float* ssz, int64_t nt,
              const float* tx, const float* ty, const float* tz, const float* tr,
              float* tu, float* tv, float* tw, float* const* g) {
  const int thick = tr != nullptr;
  if (accum) ref_p2p_loop<float, double>(grads, thick, ns, sx, sy, sz, sr, ssx, ssy, ssz, nt, tx, ty, tz, tr, tu, tv, tw, g);
  else       ref_p2p_loop<float, float>(grads, thick, ns, sx, sy, sz, sr, ssx, ssy, ssz, nt, tx, ty, tz, tr, tu, tv, tw, g);
  return 0;
}

// S = A = double: the reference's own templates instantiated in double
int ref_p2p_d(int grads, int64_t ns, const double* sx, const double* sy, const double* sz,
              const double* sr, const double* ssx, const double* ssy, const double* ssz, int64_t nt,
              const double* tx, const double* ty, const double* tz, const double* tr,
              double* tu, double* tv, double* tw, double* const* g) {
  ref_p2p_loop<double, double>(grads, tr != nullptr, ns, sx, sy, sz, sr, ssx, ssy, ssz, nt, tx, ty, tz, tr, tu, tv, tw, g);
  return 0;
}

int ref_pan_f(int grads, int accum, int maxlev, int64_t np, const float* x, const float* y, const float* z,
              const uint32_t* idx, const float* tsx, const float* tsy, const float* tsz, const float* area,
              const float* ss, int64_t nt, const float* tx, const float* ty, const float* tz,
              float* tu, float* tv, float* tw, float* const* g) {
  if (accum) ref_pan_loop<float, double>(grads, maxlev, np, x, y, z, idx, tsx, tsy, tsz, area, ss, nt, tx, ty, tz, tu, tv, tw, g);
  else       ref_pan_loop<float, float>(grads, maxlev, np, x, y, z, idx, tsx, tsy, tsz, area, ss, nt, tx, ty, tz, tu, tv, tw, g);
  return 0;
}

int ref_pan_d(int grads, int maxlev, int64_t np, const double* x, const double* y, const double* z,
              const uint32_t* idx, const double* tsx, const double* tsy, const double* tsz, const double* area,
              const double* ss, int64_t nt, const double* tx, const double* ty, const double* tz,
              double* tu, double* tv, double* tw, double* const* g) {
  ref_pan_loop<double, double>(grads, maxlev, np, x, y, z, idx, tsx, tsy, tsz, area, ss, nt, tx, ty, tz, tu, tv, tw, g);
  return 0;
}

// the reference's per-interaction flop counts for the core this library was compiled with
int ref_flops(int which) {
  switch (which) {
    case 0: return (int) flops_0v_0bg<float>();
    case 1: return (int) flops_0v_0pg<float>();
    case 2: return (int) flops_0v_0b<float>();
    case 3: return (int) flops_0v_0p<float>();
    case 4: return (int) flops_0vs_0pg<float>();
    case 5: return (int) flops_0vs_0p<float>();
  }
  return -1;
}

}  // extern "C"
