// Readers of the reference's table formats (host side, C++20).
//
//   lensing  tables/lensing/<setting>/{u,cos_psi,cos_alpha}.txt ("min max N")
//            and three row-major matrices (gpu/lensing.cuh:27-54,
//            cpu/lensing.hpp:33-41)
//   NSX      5-column text "logEkT mu logInuT3 logT logg", logT > logg >
//            logEkT > mu order, checked against the uniform axes
//            (cpu/nsx.hpp:83-148); a binary cache is written next to it
//   response tables/ism_inst/<setting>/{num,E_obs,rsp}.txt
//            (gpu/ism_inst.cuh:18-32, cpu/ism_inst.hpp:59-70)
#pragma once

#include <string>
#include <vector>

#include "../../include/ppm.h"

namespace ppm {

struct LensingHost {
  int n_u = 0, n_cp = 0, n_ca = 0;
  double u_min = 0, u_max = 0, cp_min = 0, cp_max = 0, ca_min = 0, ca_max = 0;
  std::vector<double> ca, lf, dt;
};

struct NsxHost {
  ppm_nsx_config cfg{};
  std::vector<double> mu;    // [n_mu]
  std::vector<double> data;  // 10^logInuT3, [logT][logg][logEkT][mu]
};

struct ResponseHost {
  std::string setting;  // directory name under tables/ism_inst
  int n_ch = 0, n_e = 0;
  std::vector<double> e_obs;  // [n_e]
  std::vector<double> rsp;    // [n_ch][n_e]
};

// Each returns PPM_OK or an error status with a message in `err`.
ppm_status read_lensing(const std::string& dir, LensingHost& out, std::string& err);
ppm_status read_nsx(const std::string& path, const ppm_nsx_config& cfg, NsxHost& out, std::string& err);
ppm_status read_response(const std::string& dir, ResponseHost& out, std::string& err);

ppm_nsx_config default_nsx_config();

}  // namespace ppm
