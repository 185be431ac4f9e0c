// Influence.hpp -- header-only C++ mirror of Omega3D's influence interface for the
// direct Biot-Savart path, backed by libo3dbiot.so (include/o3d_biot.h).
//
// Same names, argument meaning and accumulate semantics as the reference:
//   points_affect_points<S,A>(src, targ, restype, env)   src/Influence.h:67-551
//   panels_affect_points<S,A>(src, targ, restype, env)   src/Influence.h:557-1099
//   points_affect_panels<S,A>(src, targ, restype, env)   src/Influence.h:1107-1221
//   panels_affect_panels<S,A>(src, targ, restype, env)   src/Influence.h:1225-1245
//   InfluenceVisitor<S,A>                                 src/Influence.h:1251-1261
//   ExecEnv, accel_t (plus gpu_hip)                       src/ExecEnv.h:26-123
//   ResultsType, results_t                                src/ResultsType.h:27-95
//   Points / Surfaces SoA accessors and zero/finalize     src/ElementBase.h, Points.h, Surfaces.h
// Where the reference asserts (Influence.h:369,537), this header throws
// o3d::InfluenceError carrying the library's status and message.
#pragma once

#include "o3d_biot.h"

#include <array>
#include <cmath>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

namespace o3d {

enum summation_t { direct = 1, barneshut = 2, vic = 3, fmm = 4 };
enum accel_t { cpu_x86 = 1, cpu_vc = 2, gpu_opengl = 3, gpu_cuda = 4, gpu_hip = 5 };
enum results_t { velonly = 1, velandgrad = 2, psionly = 3, velandvort = 4, velandshear = 5 };

struct InfluenceError : std::runtime_error {
  int status;
  InfluenceError(int st, const std::string& m) : std::runtime_error(m), status(st) {}
};

inline void check(int st, const char* where) {
  if (st != O3D_OK) throw InfluenceError(st, std::string(where) + ": " + o3d_last_error());
}

class ExecEnv {
 public:
  ExecEnv(bool internal, bool useomp, summation_t summ, accel_t accel)
      : m_internal(internal), m_useomp(useomp), m_summ(summ), m_accel(accel) {}
  ExecEnv() : ExecEnv(true, true, direct, gpu_hip) {}
  void use_internal() { m_internal = true; }
  void use_external() { m_internal = false; }
  void set_internal(bool v) { m_internal = v; }
  bool is_internal() const { return m_internal; }
  void set_openmp(bool v) { m_useomp = v; }
  bool using_openmp() const { return m_useomp; }
  void set_summation(summation_t s) { m_summ = s; }
  void set_instrs(accel_t a) { m_accel = a; }
  accel_t get_instrs() const { return m_accel; }
  // what the reference fixes at compile time (CoreFunc.h:35-39) or does not have
  void set_core(int core) { m_core = core; }
  int get_core() const { return m_core; }
  void set_precision(int p) { m_prec = p; }
  int get_precision() const { return m_prec; }
  void set_self_term(int s) { m_self = s; }
  int get_self_term() const { return m_self; }
  void set_stream(void* s) { m_stream = s; }
  void* get_stream() const { return m_stream; }
  std::string to_string() const {
    return m_accel == gpu_hip ? " HIP-accelerated direct sums" : " unsupported accelerator";
  }

 private:
  bool m_internal;
  bool m_useomp;
  summation_t m_summ;
  accel_t m_accel;
  int m_core = O3D_CORE_WL;
  int m_prec = O3D_PREC_COMPENSATED;
  int m_self = O3D_SELF_INCLUDE;
  void* m_stream = nullptr;
};

class ResultsType {
 public:
  ResultsType(results_t t = velonly) : m_rtype(t) {}
  bool compute_psi() const { return m_rtype == psionly; }
  bool compute_vel() const { return m_rtype != psionly; }
  bool compute_grad() const { return m_rtype == velandgrad; }
  bool compute_vort() const { return m_rtype == velandvort; }
  results_t get_type() const { return m_rtype; }

 private:
  results_t m_rtype;
};

// SoA particle / field-point collection with the reference's accessors.
template <class S>
class Points {
 public:
  // active (thick-cored) particles
  Points(std::array<std::vector<S>, 3> x, std::array<std::vector<S>, 3> s, std::vector<S> r)
      : m_x(std::move(x)), m_s(std::move(s)), m_r(std::move(r)), m_inert(false) {
    init_outputs(true);
  }
  // inert field points (no radius); with_grad mirrors Points.h:131-139
  Points(std::array<std::vector<S>, 3> x, bool with_grad) : m_x(std::move(x)), m_inert(true) {
    init_outputs(with_grad);
  }
  size_t get_n() const { return m_x[0].size(); }
  bool is_inert() const { return m_inert; }
  const std::array<std::vector<S>, 3>& get_pos() const { return m_x; }
  const std::array<std::vector<S>, 3>& get_str() const { return m_s; }
  const std::vector<S>& get_rad() const { return m_r; }
  std::array<std::vector<S>, 3>& get_vel() { return m_u; }
  std::optional<std::array<std::vector<S>, 9>>& get_velgrad() { return m_ug; }
  const std::array<std::vector<S>, 3>& get_vel() const { return m_u; }
  const std::optional<std::array<std::vector<S>, 9>>& get_velgrad() const { return m_ug; }

  void zero_vels() {
    for (auto& v : m_u) std::fill(v.begin(), v.end(), S(0));
    if (m_ug) for (auto& v : *m_ug) std::fill(v.begin(), v.end(), S(0));
  }
  // ElementBase.h:187-208 and Points.h:265-277: u = fs + u/(4 pi), ug *= 1/(4 pi)
  void finalize_vels(const std::array<double, 3>& fs) {
    const double factor = 0.25 / M_PI;
    for (int d = 0; d < 3; ++d)
      for (auto& v : m_u[d]) v = (S)(fs[d] + v * factor);
    if (m_ug) {
      const S f = (S)factor;
      for (auto& g : *m_ug) for (auto& v : g) v = v * f;
    }
  }

 private:
  void init_outputs(bool with_grad) {
    const size_t n = get_n();
    for (auto& v : m_u) v.assign(n, S(0));
    if (with_grad) {
      m_ug.emplace();
      for (auto& v : *m_ug) v.assign(n, S(0));
    }
  }
  std::array<std::vector<S>, 3> m_x, m_s, m_u;
  std::vector<S> m_r;
  std::optional<std::array<std::vector<S>, 9>> m_ug;
  bool m_inert;
};

// SoA flat triangular panels: node positions, 3 node indices per panel, absolute
// vortex strength per panel, area, optional sheet source strength, and the
// panel-centre velocities pu that get_vel() returns (Surfaces.h:239-240).
template <class S>
class Surfaces {
 public:
  Surfaces(std::array<std::vector<S>, 3> x, std::vector<uint32_t> idx, std::array<std::vector<S>, 3> ts,
           std::vector<S> area, std::vector<S> ss = {})
      : m_x(std::move(x)), m_idx(std::move(idx)), m_ts(std::move(ts)), m_area(std::move(area)), m_ss(std::move(ss)) {
    for (auto& v : m_pu) v.assign(m_area.size(), S(0));
  }
  size_t get_npanels() const { return m_area.size(); }
  const std::array<std::vector<S>, 3>& get_pos() const { return m_x; }
  const std::vector<uint32_t>& get_idx() const { return m_idx; }
  const std::array<std::vector<S>, 3>& get_str() const { return m_ts; }
  const std::vector<S>& get_area() const { return m_area; }
  bool have_src_str() const { return !m_ss.empty(); }
  const std::vector<S>& get_src_str() const { return m_ss; }
  std::array<std::vector<S>, 3>& get_vel() { return m_pu; }
  const std::array<std::vector<S>, 3>& get_vel() const { return m_pu; }

  // Surfaces.h:868-887: panel-centre velocities zeroed, then pu = fs + pu/(4 pi)
  void zero_vels() {
    for (auto& v : m_pu) std::fill(v.begin(), v.end(), S(0));
  }
  void finalize_vels(const std::array<double, 3>& fs) {
    const double factor = 0.25 / M_PI;
    for (int d = 0; d < 3; ++d)
      for (auto& v : m_pu[d]) v = (S)(fs[d] + v * factor);
  }

 private:
  std::array<std::vector<S>, 3> m_x;
  std::vector<uint32_t> m_idx;
  std::array<std::vector<S>, 3> m_ts;
  std::vector<S> m_area, m_ss;
  std::array<std::vector<S>, 3> m_pu;
};

namespace detail {
inline o3d_opts opts_of(const ResultsType& rt, const ExecEnv& env) {
  o3d_opts o;
  o3d_default_opts(&o);
  o.core = env.get_core();
  o.results = rt.get_type();
  o.precision = env.get_precision();
  o.self_term = env.get_self_term();
  o.stream = env.get_stream();
  return o;
}
template <class S>
o3d_panels panels_of(const Surfaces<S>& p) {
  o3d_panels d{};
  d.np = (int64_t)p.get_npanels();
  d.nn = (int64_t)p.get_pos()[0].size();
  for (int k = 0; k < 3; ++k) {
    d.x[k] = p.get_pos()[k].data();
    d.ts[k] = p.get_str()[k].data();
  }
  d.idx = p.get_idx().data();
  d.area = p.get_area().data();
  d.ss = p.have_src_str() ? p.get_src_str().data() : nullptr;
  return d;
}
inline void check_env(const ResultsType& rt, const ExecEnv& env) {
  if (env.get_instrs() != gpu_hip) throw InfluenceError(O3D_EINVAL, "this library provides accel_t::gpu_hip only");
  if (rt.compute_psi() || rt.compute_vort())
    throw InfluenceError(O3D_EINVAL, "only velonly and velandgrad are served (Influence.h:71-72)");
}
}  // namespace detail

// Influence.h:67-551.  S is the storage type (float); A (the CPU accumulator type)
// has no meaning on the GPU, where env's precision chooses the accumulation.
template <class S, class A>
void points_affect_points(const Points<S>& src, Points<S>& targ, const ResultsType& restype, const ExecEnv& env) {
  static_assert(std::is_same<S, float>::value, "libo3dbiot.so stores float (use o3d_points64 for double)");
  detail::check_env(restype, env);
  const bool grads = restype.get_type() == velandgrad;
  auto& tug = targ.get_velgrad();
  if (grads && !tug) throw InfluenceError(O3D_EINVAL, "Optional target vel grads are not present!");
  o3d_points s{}, t{};
  s.n = (int64_t)src.get_n();
  for (int d = 0; d < 3; ++d) {
    s.x[d] = src.get_pos()[d].data();
    s.s[d] = src.get_str()[d].data();
  }
  s.r = src.get_rad().data();
  t.n = (int64_t)targ.get_n();
  for (int d = 0; d < 3; ++d) {
    t.x[d] = targ.get_pos()[d].data();
    t.u[d] = targ.get_vel()[d].data();
  }
  t.r = targ.is_inert() ? nullptr : targ.get_rad().data();
  if (grads) for (int k = 0; k < 9; ++k) t.ug[k] = (*tug)[k].data();
  const o3d_opts o = detail::opts_of(restype, env);
  check(o3d_points_affect_points(&s, &t, &o, nullptr), "points_affect_points");
}

// Influence.h:557-1099 (target radii ignored; gradients when the target has storage)
template <class S, class A>
void panels_affect_points(const Surfaces<S>& src, Points<S>& targ, const ResultsType& restype, const ExecEnv& env) {
  static_assert(std::is_same<S, float>::value, "libo3dbiot.so stores float");
  detail::check_env(restype, env);
  auto& tug = targ.get_velgrad();
  const o3d_panels p = detail::panels_of(src);
  o3d_points t{};
  t.n = (int64_t)targ.get_n();
  for (int d = 0; d < 3; ++d) {
    t.x[d] = targ.get_pos()[d].data();
    t.u[d] = targ.get_vel()[d].data();
  }
  ResultsType rt = restype;
  if (tug) for (int k = 0; k < 9; ++k) t.ug[k] = (*tug)[k].data();
  else rt = ResultsType(velonly);
  const o3d_opts o = detail::opts_of(rt, env);
  check(o3d_panels_affect_points(&p, &t, &o, nullptr), "panels_affect_points");
}

namespace detail {
inline void check_velonly(const ResultsType& rt, const char* what) {
  if (rt.compute_grad())
    throw InfluenceError(O3D_EINVAL, std::string(what) + ": velocity gradients are not computed on panels "
                                                         "(the reference asserts, Influence.h:1111,1227)");
}
}  // namespace detail

// Influence.h:1107-1221: the particles' velocity at each target panel's centre,
// SUBTRACTED from the panel velocities (the panel kernel run backwards)
template <class S, class A>
void points_affect_panels(const Points<S>& src, Surfaces<S>& targ, const ResultsType& restype, const ExecEnv& env) {
  static_assert(std::is_same<S, float>::value, "libo3dbiot.so stores float");
  detail::check_env(restype, env);
  detail::check_velonly(restype, "points_affect_panels");
  o3d_points s{};
  s.n = (int64_t)src.get_n();
  for (int d = 0; d < 3; ++d) {
    s.x[d] = src.get_pos()[d].data();
    s.s[d] = src.get_str()[d].data();
  }
  s.r = src.is_inert() ? nullptr : src.get_rad().data();
  const o3d_panels t = detail::panels_of(targ);
  float* vel[3] = {targ.get_vel()[0].data(), targ.get_vel()[1].data(), targ.get_vel()[2].data()};
  const o3d_opts o = detail::opts_of(restype, env);
  check(o3d_points_affect_panels(&s, &t, vel, &o, nullptr), "points_affect_panels");
}

// Influence.h:1225-1245: src's velocity at targ's colocation points (centroids
// pushed 1e-4 along the normals); like the reference, the result REPLACES the
// target's panel velocities
template <class S, class A>
void panels_affect_panels(const Surfaces<S>& src, Surfaces<S>& targ, const ResultsType& restype, const ExecEnv& env) {
  static_assert(std::is_same<S, float>::value, "libo3dbiot.so stores float");
  detail::check_env(restype, env);
  detail::check_velonly(restype, "panels_affect_panels");
  const o3d_panels s = detail::panels_of(src);
  const o3d_panels t = detail::panels_of(targ);
  float* vel[3] = {targ.get_vel()[0].data(), targ.get_vel()[1].data(), targ.get_vel()[2].data()};
  const o3d_opts o = detail::opts_of(restype, env);
  check(o3d_panels_affect_panels(&s, &t, vel, &o, nullptr), "panels_affect_panels");
}

template <class S>
using Collection = std::variant<Points<S>, Surfaces<S>>;

// Influence.h:1251-1261: double dispatch on (source, target) collection types
template <class S, class A>
struct InfluenceVisitor {
  ResultsType restype;
  ExecEnv env;
  void operator()(const Points<S>& src, Points<S>& targ) { points_affect_points<S, A>(src, targ, restype, env); }
  void operator()(const Surfaces<S>& src, Points<S>& targ) { panels_affect_points<S, A>(src, targ, restype, env); }
  void operator()(const Points<S>& src, Surfaces<S>& targ) { points_affect_panels<S, A>(src, targ, restype, env); }
  void operator()(const Surfaces<S>& src, Surfaces<S>& targ) { panels_affect_panels<S, A>(src, targ, restype, env); }
};

}  // namespace o3d
