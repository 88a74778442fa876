// Drop-in for the reference's devarshi24st_gpu (gpu/devarshi24st_gpu.cu:
// 431-533): ring and crescent spots (disk differences) on an oblate star
// with the NSX hydrogen atmosphere, 32 phase bins x 2 Gauss-Legendre nodes,
// N_fine 192, folded through tables/ism_inst/std.  Writes
// output/<name>_<setting>_counts_gpu.txt (N_CH rows x 32 columns) and
// output/<name>_<setting>_time_gpu.txt (ms), as the reference does.
// Star parameters follow gpu/devarshi24st_gpu.cu:22-44, 137-142 and
// cpu/devarshi24st_cpu.cpp:17-22, 48-59, 267-285.
#include <vector>

#include "drop_in.hpp"

namespace {

struct Spot {
  const char* name;
  double theta1, phi1, angrad1, theta2, phi2, angrad2;
};

constexpr double kPi = dropin::kPi;
constexpr Spot kWarmup = {"warmup", 0.5 * kPi, -0.2, 0.5, 0.5 * kPi, 0.0, 0.25};
constexpr Spot kRingEq = {"Ring_Eq", 0.5 * kPi, -0.2, 0.5, 0.5 * kPi, 0.0, 0.25};
constexpr Spot kRingPolar = {"Ring_Polar", kPi - 0.001, -0.02, 0.05, kPi - 0.001, 0.0, 0.025};
constexpr Spot kCrescentEq = {"Crescent_Eq", 1.0, -1.0, 1.4, 0.5, 0.0, 0.5 * kPi - 0.001};
constexpr Spot kCrescentPolar = {"Crescent_Polar", 1.15, -0.1, 1.3, 1.75, 0.0, 0.5 * kPi - 0.001};

constexpr int kNPhase = 32, kNGl = 2, kNFine = 32 * 6;

ppm_star make_star(const Spot& s) {
  ppm_star st{};
  st.mass_msun = 1.4;
  st.radius_eq_km = 12.0;
  st.spin_hz = 300.0;
  st.inclination_rad = 2.391;
  st.distance_km = 0.15 * dropin::kKpcKm;
  st.exposure_s = 1e6;
  st.atmosphere = PPM_ATM_NSX;
  st.variant = PPM_VARIANT_DEVARSHI;
  st.n_spots = 1;
  ppm_spot& sp = st.spot[0];
  sp.theta1 = s.theta1;
  sp.phi1 = s.phi1;
  sp.angrad1 = s.angrad1;
  sp.theta2 = s.theta2;
  sp.phi2 = s.phi2;
  sp.angrad2 = s.angrad2;
  sp.kT_keV = 1e6 * dropin::kBoltzmannKeVPerK;  // logT = 6.0
  sp.membership = PPM_MEMBER_DISKDIFF;
  sp.beaming = PPM_BEAM_ISOTROPIC;
  return st;
}

void calc_one_star(ppm_ctx* ctx, const Spot& s, int n_side, const char* setting) {
  int n_ch = 0, n_e = 0;
  dropin::check(ctx, ppm_response_info(ctx, 0, &n_ch, &n_e, nullptr), "ppm_response_info");
  ppm_grid g{};
  g.n_side = n_side;
  g.n_phase = kNPhase;
  g.sampling = PPM_SAMPLE_GL;
  g.n_gl = kNGl;
  g.n_fine = kNFine;
  g.n_energy = 0;  // the response's E_obs
  g.energies_keV = nullptr;
  g.response = 0;
  g.method = PPM_METHOD_PLAIN;
  const ppm_star st = make_star(s);
  std::vector<double> flux((size_t)kNPhase * n_e), counts((size_t)n_ch * kNPhase);
  ppm_stats stats{};
  dropin::check(ctx, ppm_eval(ctx, &st, 1, &g, flux.data(), counts.data(), &stats), "ppm_eval");
  std::cout << "Time taken: " << stats.total_ms << " ms\n";
  const std::string stem = std::string("output/") + s.name + "_" + setting;
  {
    std::ofstream t = dropin::open_out(stem + "_time_gpu.txt");
    t << stats.total_ms << '\n';
  }
  const std::string name = stem + "_counts_gpu.txt";
  std::ofstream f = dropin::open_out(name);
  for (int c = 0; c < n_ch; ++c) {
    for (int p = 0; p < kNPhase; ++p) f << counts[(size_t)c * kNPhase + p] << ' ';
    f << '\n';
  }
  f.close();
  std::cout << "Output written to file ./" << name << '\n';
}

}  // namespace

int main() {
  ppm_ctx* ctx = dropin::open_or_die(nullptr, "std");
  dropin::ensure_output_dir();
  calc_one_star(ctx, kWarmup, 173, "std");
  calc_one_star(ctx, kRingEq, 173, "std");
  calc_one_star(ctx, kRingPolar, 1708, "std");
  calc_one_star(ctx, kCrescentEq, 128, "std");
  calc_one_star(ctx, kCrescentPolar, 189, "std");
  calc_one_star(ctx, kRingEq, 345, "high");
  calc_one_star(ctx, kRingPolar, 3413, "high");
  calc_one_star(ctx, kCrescentEq, 247, "high");
  calc_one_star(ctx, kCrescentPolar, 382, "high");
  ppm_close(ctx);
  return 0;
}
