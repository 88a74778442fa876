// Drop-in for the reference's os1_gpu (gpu/os1_gpu.cu:303-420): the ten
// oblate-Schwarzschild blackbody test stars os1a..os1j at N_side 256, 128
// phase points, one observed energy of 1 keV; writes output/os1<x>_gpu.txt
// (one line, 128 values) and prints the time of each profile.  Star
// parameters follow gpu/os1_gpu.cu:18-65 and cpu/os1_cpu.cpp:17-36.
#include <vector>

#include "drop_in.hpp"

namespace {

struct Os1 {
  char key;
  double nu, theta_deg, inc_deg, angrad;
  char beaming;
};

constexpr Os1 kStars[] = {
    {'a', 600, 90, 90, 0.01, 'i'}, {'b', 600, 90, 90, 1.0, 'i'}, {'c', 200, 90, 90, 0.01, 'i'},
    {'d', 1, 90, 90, 1.0, 'i'},    {'e', 600, 60, 30, 1.0, 'i'}, {'f', 600, 20, 80, 1.0, 'i'},
    {'g', 600, 60, 30, 1.0, 'c'},  {'h', 600, 60, 30, 1.0, 's'}, {'i', 600, 20, 80, 1.0, 'c'},
    {'j', 600, 20, 80, 1.0, 's'},
};
constexpr int kNSide = 256, kNPhase = 128, kNFine = 3 * kNPhase;
constexpr double kEObs = 1.0;  // keV

ppm_star make_star(const Os1& s) {
  ppm_star st{};
  st.mass_msun = 1.4;
  st.radius_eq_km = 12.0;
  st.spin_hz = s.nu;
  st.inclination_rad = s.inc_deg * dropin::kDegree;
  st.distance_km = 0.2 * dropin::kKpcKm;
  st.exposure_s = 1.0;
  st.point_source_below = 0.1;  // gpu/os1_gpu.cu:315
  st.atmosphere = PPM_ATM_BLACKBODY;
  st.variant = PPM_VARIANT_OS1;
  st.n_spots = 1;
  ppm_spot& sp = st.spot[0];
  sp.theta1 = s.theta_deg * dropin::kDegree;
  sp.angrad1 = s.angrad;
  sp.kT_keV = 0.35;
  sp.membership = PPM_MEMBER_CAP;
  sp.beaming = s.beaming;
  return st;
}

}  // namespace

int main() {
  ppm_ctx* ctx = dropin::open_or_die("", "");
  dropin::ensure_output_dir();
  ppm_grid g{};
  g.n_side = kNSide;
  g.n_phase = kNPhase;
  g.sampling = PPM_SAMPLE_POINTS;
  g.n_gl = 1;
  g.n_fine = kNFine;
  g.n_energy = 1;
  g.energies_keV = &kEObs;
  g.response = -1;
  g.method = PPM_METHOD_PLAIN;
  std::vector<double> flux(kNPhase);
  for (const Os1& s : kStars) {
    const ppm_star st = make_star(s);
    ppm_stats stats{};
    dropin::check(ctx, ppm_eval(ctx, &st, 1, &g, flux.data(), nullptr, &stats), "ppm_eval");
    std::cout << "Time taken: " << stats.total_ms << " ms\n";
    const std::string name = std::string("output/os1") + s.key + "_gpu.txt";
    std::ofstream f = dropin::open_out(name);
    for (int p = 0; p < kNPhase; ++p) f << flux[p] << ' ';
    f.close();
    std::cout << "Output written to file ./" << name << "\n";
  }
  ppm_close(ctx);
  return 0;
}
