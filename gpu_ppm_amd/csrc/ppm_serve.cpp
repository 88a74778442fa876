// Serving loop (continuous batching of single-profile requests).
#include "ppm_ctx.hpp"

namespace ppm {
void serve_init(ppm_ctx*) {}
void serve_shutdown(ppm_ctx*) {}
}  // namespace ppm

extern "C" {
ppm_status ppm_submit(ppm_ctx*, const ppm_star*, const ppm_grid*, uint64_t*) { return PPM_ERR_UNSUPPORTED; }
ppm_status ppm_poll(ppm_ctx*, uint64_t, double*, double*, double*) { return PPM_ERR_UNSUPPORTED; }
ppm_status ppm_cancel(ppm_ctx*, uint64_t) { return PPM_ERR_UNSUPPORTED; }
ppm_status ppm_drain(ppm_ctx*) { return PPM_ERR_UNSUPPORTED; }
}
