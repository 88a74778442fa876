// Drop-in for the reference's interpolant study t_interp (cpu/t_interp.cpp,
// CPU only in the reference): one 0.01 rad spot (a point-source element) at
// logT 6 on an oblate 1.4 M_sun, 12 km, 600 Hz star at 0.15 kpc, observed
// through each NSX interpolant (interp_type 0..3 = 4c, 4c with use_sg, 4c_le,
// 2l2c_le; cpu/t_interp.cpp:210-219), 256 phase bins x 2 Gauss-Legendre
// nodes, 1024 fine phases, 1024 energies linspace(0.1, 5) keV, no fold.
// Writes output/<name>_flux.txt (256 rows x 1024 energies, 10 significant
// digits) as the reference does (cpu/t_interp.cpp:228-239).  All eight stars
// are evaluated in one batched call.
#include <vector>

#include "drop_in.hpp"

namespace {

struct Test {
  const char* name;
  double obs_theta, theta;
  int interp_type;
};

constexpr double kPi = dropin::kPi;
// cpu/t_interp.cpp:248-261
constexpr Test kTests[] = {
    {"test10", 0.01, 1.983, 0},        {"test11", 0.01, 1.983, 1},        {"test12", 0.01, 1.983, 2},
    {"test13", 0.01, 1.983, 3},        {"test20", 0.964, kPi - 0.01, 0}, {"test21", 0.964, kPi - 0.01, 1},
    {"test22", 0.964, kPi - 0.01, 2}, {"test23", 0.964, kPi - 0.01, 3},
};
constexpr int kInterp[] = {PPM_NSX_4C, PPM_NSX_4C_SG, PPM_NSX_4C_LE, PPM_NSX_2L2C_LE};
constexpr int kNPhase = 256, kNGl = 2, kNFine = 256 * 4, kNE = 1024;

ppm_star make_star(const Test& t) {
  ppm_star st{};
  st.mass_msun = 1.4;
  st.radius_eq_km = 12.0;
  st.spin_hz = 600.0;
  st.inclination_rad = t.obs_theta;
  st.distance_km = 0.15 * dropin::kKpcKm;
  st.point_source_below = 0.1;  // the spot is one element (cpu/grid.hpp:17-27)
  st.atmosphere = PPM_ATM_NSX;
  st.variant = PPM_VARIANT_DEVARSHI;
  st.nsx_interp = kInterp[t.interp_type];
  st.n_spots = 1;
  ppm_spot& sp = st.spot[0];
  sp.theta1 = t.theta;
  sp.angrad1 = 0.01;
  sp.kT_keV = 1e6 * dropin::kBoltzmannKeVPerK;  // logT = 6.0
  sp.membership = PPM_MEMBER_CAP;
  sp.beaming = PPM_BEAM_ISOTROPIC;
  return st;
}

}  // namespace

int main() {
  ppm_ctx* ctx = dropin::open_or_die(nullptr, "");
  dropin::ensure_output_dir();
  std::vector<double> energies(kNE);
  const double step = (5.0 - 0.1) / (kNE - 1);  // linspace, cpu/ut.hpp:7-14
  for (int i = 0; i < kNE; ++i) energies[i] = 0.1 + i * step;
  ppm_grid g{};
  g.n_side = 16;  // unused by point-source elements
  g.n_phase = kNPhase;
  g.sampling = PPM_SAMPLE_GL;
  g.n_gl = kNGl;
  g.n_fine = kNFine;
  g.n_energy = kNE;
  g.energies_keV = energies.data();
  g.response = -1;
  g.method = PPM_METHOD_BATCHED;
  const int n = (int)(sizeof kTests / sizeof kTests[0]);
  std::vector<ppm_star> stars;
  for (const Test& t : kTests) stars.push_back(make_star(t));
  std::vector<double> flux((size_t)n * kNPhase * kNE);
  ppm_stats stats{};
  dropin::check(ctx, ppm_eval(ctx, stars.data(), n, &g, flux.data(), nullptr, &stats), "ppm_eval");
  std::cout << "Time taken: " << stats.total_ms << " ms\n";
  for (int s = 0; s < n; ++s) {
    const std::string name = std::string("output/") + kTests[s].name + "_flux.txt";
    std::ofstream f = dropin::open_out(name);
    f.precision(10);
    const double* F = flux.data() + (size_t)s * kNPhase * kNE;
    for (int p = 0; p < kNPhase; ++p) {
      for (int e = 0; e < kNE; ++e) f << F[(size_t)p * kNE + e] << ' ';
      f << '\n';
    }
    std::cout << ">>> Total flux written to " << name << '\n';
  }
  ppm_close(ctx);
  return 0;
}
