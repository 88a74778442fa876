// Batched evaluation: many stars in one persistent launch (ppm_batch.cpp).
#pragma once

#include <vector>

#include "ppm_ctx.hpp"

namespace ppm {

// Stars of one batched launch share a launch class: the atmosphere and, for
// NSX, the energy-axis evaluation of the interpolant (a kernel instance); the
// mu and (logT, log g) stencils are per spot.
inline int batch_emode(const ppm_star& s) {
  return s.atmosphere == PPM_ATM_NSX ? nsx_energy_mode(s.nsx_interp) : kECubic;
}
inline int batch_class(const ppm_star& s) { return s.atmosphere == PPM_ATM_NSX ? 1 + batch_emode(s) : 0; }

ppm_status eval_batched(ppm_ctx* c, Lane& L, const ppm_star* stars, int n, const ppm_grid& g,
                        const std::vector<double>& energies, const std::vector<PhasePoint>& pts,
                        double* d_flux, double* d_counts, ppm_stats* stats);

// Stars on one grid within a launch that may hold several grids.
struct BatchGroup {
  const ppm_star* stars = nullptr;
  int n = 0;
  ppm_grid g{};                   // validated; its energies are `energies`
  std::vector<double> energies;   // ascending
  std::vector<PhasePoint> pts;
  double* d_flux = nullptr;       // [n][n_phase][n_e]
  double* d_counts = nullptr;     // [n][n_ch][n_phase], or null (no fold)
  int fit_star0 = -1;             // fit statistics: index of stars[0] in ppm_ctx::fit (-1: its launch index)
};

// One persistent launch over the stars of several grids (one launch class,
// energies strictly increasing in every group).
ppm_status eval_batched_groups(ppm_ctx* c, Lane& L, std::vector<BatchGroup>& groups, ppm_stats* stats);

// Whether one launch over these groups fits the LDS of a workgroup.
bool batch_groups_fit(const ppm_ctx* c, const std::vector<BatchGroup>& groups);
// Energies per lane the batched kernel takes for this grid alone (2 or 6).
int batch_grid_epl(const ppm_ctx* c, const BatchGroup& G);

}  // namespace ppm
