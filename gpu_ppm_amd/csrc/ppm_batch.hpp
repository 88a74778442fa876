// Batched evaluation: many stars in one persistent launch (ppm_batch.cpp).
#pragma once

#include <vector>

#include "ppm_ctx.hpp"

namespace ppm {

ppm_status eval_batched(ppm_ctx* c, const ppm_star* stars, int n, const ppm_grid& g,
                        const std::vector<double>& energies, const std::vector<PhasePoint>& pts,
                        double* d_flux, double* d_counts, ppm_stats* stats);

}  // namespace ppm
