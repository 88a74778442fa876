// Shared pieces of the drop-in programs: context set-up, output files and
// the reference's physical constants (cpu/unit.hpp:5-25).
#pragma once

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <sys/stat.h>

#include "../../include/ppm.h"

namespace dropin {

constexpr double kPi = 3.14159265358979323846;
constexpr double kDegree = kPi / 180.0;
constexpr double kKpcKm = 3.08567758e16;              // kpc in km
constexpr double kBoltzmannKeVPerK = 8.617333262e-8;  // k_B in keV / K

// Open a context on the tables below the working directory (the reference
// programs read relative paths and run from the repository root).
inline ppm_ctx* open_or_die(const char* nsx_path, const char* responses) {
  ppm_open_opts o{};
  o.root = ".";
  o.lensing_setting = "std";
  o.nsx_path = nsx_path;
  o.response_settings = responses;
  ppm_ctx* ctx = nullptr;
  const ppm_status st = ppm_open(&o, &ctx);
  if (st != PPM_OK) {
    std::cerr << "ppm_open failed: " << ppm_strerror(st) << (ctx ? "" : "") << "\n";
    std::exit(EXIT_FAILURE);
  }
  return ctx;
}

inline void check(ppm_ctx* ctx, ppm_status st, const char* what) {
  if (st != PPM_OK) {
    std::cerr << what << " failed: " << ppm_strerror(st) << ": " << ppm_last_error(ctx) << "\n";
    ppm_close(ctx);
    std::exit(EXIT_FAILURE);
  }
}

inline void ensure_output_dir() {
  if (mkdir("output", 0755) != 0 && errno != EEXIST) {
    std::cerr << "Error creating directory output\n";
    std::exit(EXIT_FAILURE);
  }
}

inline std::ofstream open_out(const std::string& name) {
  std::ofstream f(name);
  if (!f.is_open()) {
    std::cerr << "Error opening output file " << name << "\n";
    std::exit(EXIT_FAILURE);
  }
  f << std::setprecision(16);
  return f;
}

}  // namespace dropin
