// drive_influence.cpp -- drives include/omega3d/Influence.hpp the way Omega3D's
// Convection::find_vels does (Convection.h:131-182): zero the targets, visit every
// source collection with InfluenceVisitor, finalize; then the way solve_bem builds
// its right-hand side (BEMHelper.h:73-95): for every boundary surface, zero its
// panel velocities, visit every vortex collection with a velocity-only visitor
// (particles -> panels) and finalize; and one surface on another (panels ->
// panels).  Writes inputs and results as raw float32 to the directory given on the
// command line so that the test can check them against the oracle.  Exit status 0
// on success.
#include "omega3d/Influence.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>

using namespace o3d;

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static float urand() {  // xorshift64*, floats in [0, 1)
  g_state ^= g_state >> 12; g_state ^= g_state << 25; g_state ^= g_state >> 27;
  return (float)((g_state * 2685821657736338717ull) >> 40) * (1.0f / 16777216.0f);
}

static void dump(const std::string& dir, const char* name, const std::vector<float>& v) {
  FILE* f = std::fopen((dir + "/" + name + ".f32").c_str(), "wb");
  if (!f) { std::perror("fopen"); std::exit(3); }
  std::fwrite(v.data(), sizeof(float), v.size(), f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: %s outdir\n", argv[0]); return 2; }
  const std::string out = argv[1];
  const size_t n = 3000, nf = 500;
  std::array<std::vector<float>, 3> x, s, xf;
  std::vector<float> r(n);
  for (int d = 0; d < 3; ++d) {
    x[d].resize(n); s[d].resize(n); xf[d].resize(nf);
    for (size_t i = 0; i < n; ++i) { x[d][i] = urand() - 0.5f; s[d][i] = (urand() - 0.5f) / n; }
    for (size_t i = 0; i < nf; ++i) xf[d][i] = 1.5f * (urand() - 0.5f);
  }
  for (auto& v : r) v = 0.02f + 0.05f * urand();
  // one flat panel collection: a 2x1 strip of triangles
  std::array<std::vector<float>, 3> nx = {std::vector<float>{-0.3f, 0.3f, 0.3f, -0.3f},
                                          std::vector<float>{-0.2f, -0.2f, 0.2f, 0.2f},
                                          std::vector<float>{0.6f, 0.6f, 0.6f, 0.6f}};
  std::vector<uint32_t> idx = {0, 1, 2, 0, 2, 3};
  std::vector<float> area = {0.12f, 0.12f};
  std::array<std::vector<float>, 3> ts = {std::vector<float>{0.012f, -0.006f}, std::vector<float>{0.003f, 0.009f},
                                          std::vector<float>{0.0f, 0.0f}};
  std::vector<Collection<float>> vort, bdry;
  vort.emplace_back(Points<float>(x, s, r));
  bdry.emplace_back(Surfaces<float>(nx, idx, ts, area));
  std::vector<Collection<float>> fld;
  fld.emplace_back(Points<float>(xf, true));

  ExecEnv env;
  env.set_core(O3D_CORE_EXP);
  InfluenceVisitor<float, double> vis{ResultsType(velandgrad), env};
  const std::array<double, 3> fs = {0.1, 0.0, -0.05};
  std::array<std::vector<float>, 3> nx2 = nx;
  for (auto& z : nx2[2]) z += 0.15f;
  std::vector<Collection<float>> bdry2;
  bdry2.emplace_back(Surfaces<float>(nx2, idx, ts, area));
  try {
    for (auto* targs : {&vort, &fld}) {
      for (auto& t : *targs) {
        std::visit([](auto& e) { if constexpr (std::is_same<std::decay_t<decltype(e)>, Points<float>>::value) e.zero_vels(); }, t);
        for (auto& src : vort) std::visit(vis, src, t);
        for (auto& src : bdry) std::visit(vis, src, t);
        std::visit([&](auto& e) { if constexpr (std::is_same<std::decay_t<decltype(e)>, Points<float>>::value) e.finalize_vels(fs); }, t);
      }
    }
    // solve_bem's right-hand-side loop (BEMHelper.h:73-95), through std::visit
    InfluenceVisitor<float, double> ivis{ResultsType(velonly), env};
    for (auto& t : bdry) {
      std::visit([](auto& e) { e.zero_vels(); }, t);
      for (auto& src : vort) std::visit(ivis, src, t);
      std::visit([&](auto& e) { e.finalize_vels(fs); }, t);
    }
    // one surface on another: a second strip above the first
    for (auto& src : bdry) std::visit(ivis, src, bdry2[0]);

    Points<float> nograd(xf, false);
    bool threw = false;
    try { points_affect_points<float, double>(std::get<Points<float>>(vort[0]), nograd, ResultsType(velandgrad), env); }
    catch (const InfluenceError& e) { threw = e.status == O3D_EINVAL; }
    if (!threw) { std::fprintf(stderr, "missing-gradient request did not throw\n"); return 4; }
  } catch (const InfluenceError& e) {
    std::fprintf(stderr, "InfluenceError %d: %s\n", e.status, e.what());
    return 1;
  }
  auto& pv = std::get<Points<float>>(vort[0]);
  auto& pf = std::get<Points<float>>(fld[0]);
  auto flat = [](const auto& arr) { std::vector<float> v; for (auto& a : arr) v.insert(v.end(), a.begin(), a.end()); return v; };
  dump(out, "x", flat(x)); dump(out, "s", flat(s)); dump(out, "r", r); dump(out, "xf", flat(xf));
  dump(out, "vort_u", flat(pv.get_vel())); dump(out, "vort_ug", flat(*pv.get_velgrad()));
  dump(out, "fld_u", flat(pf.get_vel())); dump(out, "fld_ug", flat(*pf.get_velgrad()));
  dump(out, "bdry_pu", flat(std::get<Surfaces<float>>(bdry[0]).get_vel()));
  dump(out, "bdry2_pu", flat(std::get<Surfaces<float>>(bdry2[0]).get_vel()));
  std::printf("drive_influence: ok (%zu particles, %zu field points, 2 panels)\n", n, nf);
  return 0;
}
