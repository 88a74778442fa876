// Batched evaluation (persistent work-queue kernel); see ppm_batch.hpp.
#include "ppm_batch.hpp"

namespace ppm {

ppm_status eval_batched(ppm_ctx* c, const ppm_star*, int, const ppm_grid&, const std::vector<double>&,
                        const std::vector<PhasePoint>&, double*, double*, ppm_stats*) {
  return fail(c, PPM_ERR_UNSUPPORTED, "batched path not built yet");
}

}  // namespace ppm
