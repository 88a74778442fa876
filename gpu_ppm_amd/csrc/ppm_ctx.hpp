// The context behind the opaque ppm_ctx handle (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ppm_internal.hpp"
#include "ppm_tables.hpp"

namespace ppm {

// Device buffer that only grows (no per-call hipMalloc once warm).
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes);
  void release();
};

// Pinned host buffer (hipHostMalloc) for asynchronous uploads.
struct HostBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes);
  void release();
};

struct Workspace {
  DevBuf spots, tasks, desc, ph, fine, pts, energies, partial, red, scale, out_flux, out_counts;
  DevBuf fit_data, fit_part, fit_out;  // ppm_eval_fit
  void release_all() {
    for (DevBuf* b : {&spots, &tasks, &desc, &ph, &fine, &pts, &energies, &partial, &red, &scale, &out_flux,
                      &out_counts, &fit_data, &fit_part, &fit_out})
      b->release();
  }
};

// Buffers of the batched path (ppm_batch.cpp).
struct BatchWorkspace {
  DevBuf spots, tasks, desc, ph, fine, pts, energies, items, scale, counter, evals, summary;
  void release_all() {
    for (DevBuf* b : {&spots, &tasks, &desc, &ph, &fine, &pts, &energies, &items, &scale, &counter, &evals, &summary})
      b->release();
  }
};

struct Response {
  std::string setting;
  int n_ch = 0, n_e = 0;
  std::vector<double> e_obs, rsp;
  DevBuf buf;
  const uint16_t* d_hi = nullptr;  // BF16 high part of FP32(rsp)
  const uint16_t* d_lo = nullptr;  // BF16 of the remainder
};

struct ServeState;  // ppm_serve.cpp
struct BatchState;  // ppm_batch.cpp: caches shared by the lanes of a context
struct LaneBatch;   // ppm_batch.cpp: buffers of one lane's batched launches

// Accumulator bytes of one batched launch at most (further capped by the
// context's device-memory budget).
constexpr size_t kAccCapBytes = (size_t)12 << 30;

// A lane: one stream of device work with its own buffers, so that lanes run
// concurrently on the device.  The context's main lane serves the blocking
// entry points (one host thread at a time); each serving worker owns one.
struct Lane {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev_i0 = nullptr, ev_i1 = nullptr, ev_staging = nullptr;
  HostBuf staging;  // pinned host memory of uploads (free again after ev_staging)
  Workspace ws;
  BatchWorkspace batch;
  LaneBatch* lb = nullptr;
  // fold-epilogue fit statistics for the duration of one ppm_eval_fit call
  // (main lane only); fit_star is the first star of the next fold
  const FoldFit* fit = nullptr;
  int fit_star = 0;
  bool no_graph = false;                     // launch the batched sequence directly
  const ppm_launch_shape* forced = nullptr;  // shape under test by ppm_tune
  size_t acc_cap = kAccCapBytes;             // accumulator bytes of one batched launch
  int occ_wg_max = 0;  // occupancy limit of the last batched launch
  // batched launches with one workgroup per work item instead of persistent
  // workgroups (a concurrent launch on another lane gets slots as they retire)
  bool per_item = false;
  int init();
  void release();
};

inline uint16_t bf16_round(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
inline float bf16_value(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

}  // namespace ppm

struct ppm_ctx {
  int device = 0;
  int max_lds = 65536;
  int n_cu = 256;
  size_t budget = 0;
  ppm::Lane main;      // the blocking entry points' stream and buffers
  std::string err;
  std::mutex err_mu;
  std::mutex exec_mu;  // one evaluation on the device at a time

  ppm::LensingHost lens;
  ppm::DevLensing L{};
  ppm::DevBuf d_lens;
  ppm::DevNsx N{};
  ppm::DevBuf d_nsx, d_mu;
  std::vector<double> nsx_mu;
  std::vector<ppm::Response> resp;

  ppm::ServeState* serve = nullptr;
  ppm::BatchState* batch_state = nullptr;
  bool no_graph = false;  // PPM_NO_GRAPH=1: launch the batched sequence directly
  // device address of this device's ppm_eval_device_noalloc scratch
  // (kNoAllocScratchBytes), looked up at ppm_open; shared by every context
  // on the device
  void* noalloc = nullptr;
};

namespace ppm {

ppm_status fail(ppm_ctx* c, ppm_status st, const std::string& msg);
ppm_status hip_fail(ppm_ctx* c, int e, const char* what);
ppm_status validate(ppm_ctx* c, const ppm_star& st, const ppm_grid& g, int n_e_resolved);
void spot_params(const ppm_ctx* c, const ppm_star& st, const ppm_spot& sp, int n_side, SpotParams& P);
void candidate_rings(const SpotParams& P, const ppm_spot& sp, int n_side, std::vector<int>& rings);
void phase_points(const ppm_grid& g, std::vector<PhasePoint>& pts);
ppm_status resolve_energies(ppm_ctx* c, const ppm_grid& g, std::vector<double>& e);
ppm_status eval_any(ppm_ctx* c, Lane& L, const ppm_star* stars, int n, const ppm_grid* grid, double* d_flux,
                    double* d_counts, ppm_stats* stats);

void serve_init(ppm_ctx* c);
void serve_shutdown(ppm_ctx* c);
void release_batch_state(ppm_ctx* c);
void release_lane_batch(Lane& L);
// Device bytes held by a lane's buffers; *fixed receives its segment-total scratch.
size_t lane_device_bytes(const Lane& L, size_t* fixed);

// Host arrays copied to the device through a lane's pinned staging buffer on
// its stream (ppm_api.cpp).
struct Upload {
  std::vector<std::pair<const void*, size_t>> src;
  std::vector<void*> dst;
  void add(void* d, const void* s, size_t n) {
    src.emplace_back(s, n);
    dst.push_back(d);
  }
  int run(Lane& L);
};

}  // namespace ppm
