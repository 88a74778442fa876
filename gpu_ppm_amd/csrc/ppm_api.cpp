// C ABI of libppm (include/ppm.h): context, table upload, evaluation driver.
//
// The reference drives the path from calc_one_star
// (gpu/os1_gpu.cu:303-349, gpu/devarshi24st_gpu.cu:431-511): per call it
// allocates and clears the outputs, launches one kernel, synchronises, runs
// the cuBLAS fold, times with a cudaEvent pair and copies back.  Here device
// memory is owned by the context and reused across calls; each evaluation
// is a fixed sequence of launches on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ppm_batch.hpp"
#include "ppm_ctx.hpp"
#include "ppm_internal.hpp"
#include "ppm_tables.hpp"

using namespace ppm;

namespace ppm {

ppm_status fail(ppm_ctx* c, ppm_status st, const std::string& msg) {
  if (c) {
    std::lock_guard<std::mutex> lk(c->err_mu);  // the serving thread reports too
    c->err = msg;
  }
  return st;
}

ppm_status hip_fail(ppm_ctx* c, int e, const char* what) {
  return fail(c, PPM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString((hipError_t)e));
}

int DevBuf::ensure(size_t bytes) {
  if (bytes <= cap) return hipSuccess;
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  size_t want = std::max(bytes, (size_t)256);
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) return e;
  cap = want;
  return hipSuccess;
}
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}
int HostBuf::ensure(size_t bytes) {
  if (bytes <= cap) return hipSuccess;
  if (p) hipHostFree(p);
  p = nullptr;
  cap = 0;
  size_t want = std::max(bytes, (size_t)4096);
  hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
  if (e != hipSuccess) return e;
  cap = want;
  return hipSuccess;
}
void HostBuf::release() {
  if (p) hipHostFree(p);
  p = nullptr;
  cap = 0;
}

int Lane::init() {
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) || (e = hipEventCreate(&ev0)) ||
      (e = hipEventCreate(&ev1)) || (e = hipEventCreate(&ev2)) || (e = hipEventCreate(&ev_i0)) ||
      (e = hipEventCreate(&ev_i1)) || (e = hipEventCreateWithFlags(&ev_staging, hipEventDisableTiming)))
    return e;
  return hipEventRecord(ev_staging, stream);
}

void Lane::release() {
  if (stream) hipStreamSynchronize(stream);
  ws.release_all();
  batch.release_all();
  release_lane_batch(*this);
  staging.release();
  for (hipEvent_t* ev : {&ev0, &ev1, &ev2, &ev_i0, &ev_i1, &ev_staging})
    if (*ev) {
      hipEventDestroy(*ev);
      *ev = nullptr;
    }
  if (stream) hipStreamDestroy(stream);
  stream = nullptr;
}

// ------------------------------------------------------------ validation

ppm_status validate(ppm_ctx* c, const ppm_star& st, const ppm_grid& g, int n_e_resolved) {
  if (g.n_side < 1 || g.n_side > 8192) return fail(c, PPM_ERR_ARG, "n_side must be in 1..8192");
  if (g.n_phase < 1 || g.n_phase > 4096) return fail(c, PPM_ERR_ARG, "n_phase must be in 1..4096");
  if (g.n_fine < 4 || g.n_fine > 8192) return fail(c, PPM_ERR_ARG, "n_fine must be in 4..8192");
  if (g.sampling == PPM_SAMPLE_GL) {
    if (g.n_gl < 1 || g.n_gl > 6) return fail(c, PPM_ERR_ARG, "n_gl must be in 1..6");
  } else if (g.sampling != PPM_SAMPLE_POINTS) {
    return fail(c, PPM_ERR_ARG, "unknown sampling");
  }
  if (n_e_resolved < 1 || n_e_resolved > 65536) return fail(c, PPM_ERR_ARG, "bad energy count");
  if (g.response >= (int)c->resp.size()) return fail(c, PPM_ERR_ARG, "response index out of range");
  if (g.response >= 0 && n_e_resolved != c->resp[g.response].n_e)
    return fail(c, PPM_ERR_ARG, "the fold needs the flux on the response's E_obs grid");
  if (st.n_spots < 1 || st.n_spots > PPM_MAX_SPOTS) return fail(c, PPM_ERR_ARG, "n_spots must be 1..4");
  if (!(st.mass_msun > 0) || !(st.radius_eq_km > 0) || !(st.distance_km > 0) || !(st.spin_hz >= 0))
    return fail(c, PPM_ERR_ARG, "mass, radius, distance must be positive, spin non-negative");
  if (st.atmosphere != PPM_ATM_BLACKBODY && st.atmosphere != PPM_ATM_NSX)
    return fail(c, PPM_ERR_ARG, "unknown atmosphere");
  if (st.atmosphere == PPM_ATM_NSX && !c->N.data)
    return fail(c, PPM_ERR_UNSUPPORTED, "NSX atmosphere requested but no NSX grid is loaded");
  if (st.atmosphere == PPM_ATM_NSX && (st.nsx_interp < PPM_NSX_2L2C_LE || st.nsx_interp > PPM_NSX_4L))
    return fail(c, PPM_ERR_ARG, "unknown NSX interpolant");
  if (st.atmosphere == PPM_ATM_NSX && nsx_cubic_tg(st.nsx_interp) && (c->N.nT < 4 || c->N.ng < 4))
    return fail(c, PPM_ERR_ARG, "cubic logT / log g stencils need 4 nodes on each axis");
  if (st.variant != PPM_VARIANT_OS1 && st.variant != PPM_VARIANT_DEVARSHI)
    return fail(c, PPM_ERR_ARG, "unknown variant");
  // the surface must lie outside the Schwarzschild radius and rotate below c
  const double Rs = st.mass_msun * kRsSunKm, Re = st.radius_eq_km;
  const double Ob = 2 * kPi * st.spin_hz * std::sqrt(Re * Re * Re / st.mass_msun) * kSqrtKm3S2GMsun;
  const double o2 = Ob * Ob * (-0.788 + 1.030 * 0.5 * Rs / Re);
  const double r_min = Re * std::min(1.0, 1.0 + o2), r_max = Re * std::max(1.0, 1.0 + o2);
  if (!(r_min > Rs)) return fail(c, PPM_ERR_ARG, "stellar surface inside the Schwarzschild radius");
  if (!(2 * kPi * st.spin_hz * r_max / std::sqrt(1 - Rs / r_max) < kCKmS))
    return fail(c, PPM_ERR_ARG, "surface velocity at or above c");
  for (int s = 0; s < st.n_spots; ++s) {
    const ppm_spot& sp = st.spot[s];
    if (!(sp.angrad1 > 0) || !(sp.kT_keV > 0))
      return fail(c, PPM_ERR_ARG, "spot radius and temperature must be positive");
    if (sp.membership != PPM_MEMBER_CAP && sp.membership != PPM_MEMBER_DISKDIFF)
      return fail(c, PPM_ERR_ARG, "unknown membership rule");
    if (st.atmosphere == PPM_ATM_BLACKBODY && sp.beaming != PPM_BEAM_ISOTROPIC &&
        sp.beaming != PPM_BEAM_COS2 && sp.beaming != PPM_BEAM_SIN2)
      return fail(c, PPM_ERR_ARG, "unknown beaming");
    if (!std::isfinite(sp.theta1) || !std::isfinite(sp.phi1) || !std::isfinite(sp.theta2) ||
        !std::isfinite(sp.phi2) || !std::isfinite(sp.angrad2))
      return fail(c, PPM_ERR_ARG, "spot geometry must be finite");
  }
  return PPM_OK;
}

// -------------------------------------------------------- preparation

void spot_params(const ppm_ctx* c, const ppm_star& st, const ppm_spot& sp, int n_side, SpotParams& P) {
  std::memset(&P, 0, sizeof P);
  P.M = st.mass_msun;
  P.Re = st.radius_eq_km;
  P.nu = st.spin_hz;
  P.Rs = P.M * kRsSunKm;
  P.x = 0.5 * P.Rs / P.Re;
  P.Ob = 2 * kPi * P.nu * std::sqrt(P.Re * P.Re * P.Re / P.M) * kSqrtKm3S2GMsun;
  P.o2 = P.Ob * P.Ob * (-0.788 + 1.030 * P.x);
  P.inv_D2 = 1.0 / (st.distance_km * st.distance_km);
  P.ci = std::cos(st.inclination_rad);
  P.si = std::sin(st.inclination_rad);
  P.logg_override = st.logg_override;
  P.logg0_total = std::log10(P.M / (P.Re * P.Re * std::sqrt(1. - P.Rs / P.Re))) + std::log10(kGMsunCgs);
  P.atmosphere = st.atmosphere;
  P.variant = st.variant;
  P.beaming = sp.beaming;
  P.membership = sp.membership;
  P.st1 = std::sin(sp.theta1);
  P.ct1 = std::cos(sp.theta1);
  P.phi1 = sp.phi1;
  P.cr1 = std::cos(sp.angrad1);
  P.has2 = sp.angrad2 > 0 || sp.membership == PPM_MEMBER_DISKDIFF;
  P.st2 = std::sin(sp.theta2);
  P.ct2 = std::cos(sp.theta2);
  P.phi2 = sp.phi2;
  P.cr2 = std::cos(sp.angrad2);
  P.point_source = st.point_source_below > 0 && sp.angrad1 < st.point_source_below;
  {
    const double sh = std::sin(0.5 * sp.angrad1);  // cpu/grid.hpp:17-27
    P.ps_z = std::cos(sp.theta1);
    P.ps_domega = 2 * kPi * 2. * sh * sh;
    P.ps_shift = sp.phi1 / (2 * kPi);
  }
  P.kT = sp.kT_keV;
  P.inv_kT = 1.0 / sp.kT_keV;
  P.logT = std::log10(sp.kT_keV / kBoltzKeVK);
  P.kT_nsx = std::pow(10.0, P.logT) * kBoltzKeVK;  // cpu/devarshi24st_cpu.cpp:58-59
  if (st.atmosphere == PPM_ATM_NSX) {
    P.scale = std::pow(10.0, 3 * P.logT) / kPlanckErgS * P.inv_D2 / c->N.scale;
  } else {
    P.scale = 2. / (kHKeVs * kHKeVs * kHKeVs * kCCmS * kCCmS) * P.inv_D2;  // cpu/unit.hpp:19
    P.kT_nsx = P.kT;
  }
  P.n_side = n_side;
  P.nsx_interp = st.atmosphere == PPM_ATM_NSX ? st.nsx_interp : PPM_NSX_2L2C_LE;
}

// Candidate rings of a spot: rings whose colatitude is within angrad1 of the
// spot centre (plus a tiny margin); the device decides every pixel in FP64.
void candidate_rings(const SpotParams& P, const ppm_spot& sp, int n_side, std::vector<int>& rings) {
  rings.clear();
  if (P.point_source) {
    rings.push_back(0);
    return;
  }
  const double reach = sp.angrad1 + 1e-9;
  const int n_ring = 4 * n_side - 1;
  auto offset = [&](int i) {  // ring colatitude minus the spot centre's
    const bool north = i <= 2 * n_side;
    const int ii = north ? i : 4 * n_side - i;
    double z = ii < n_side ? 1. - 1. * ii * ii / (3. * n_side * n_side) : 4. / 3. - 2. * ii / (3. * n_side);
    if (!north) z = -z;
    return std::acos(std::max(-1.0, std::min(1.0, z))) - sp.theta1;
  };
  // the colatitude grows with i, so the rings within reach are a contiguous
  // range: bisect for its first ring, then test rings until one is beyond
  int lo = 1, hi = n_ring + 1;
  while (lo < hi) {
    const int m = (lo + hi) / 2;
    if (offset(m) >= -reach)
      hi = m;
    else
      lo = m + 1;
  }
  for (int i = std::max(1, lo - 2); i <= n_ring; ++i) {
    const double d = offset(i);
    if (d > reach) break;
    if (std::fabs(d) <= reach) rings.push_back(i);
  }
}

void phase_points(const ppm_grid& g, std::vector<PhasePoint>& pts) {
  static const double X[] = {0.0,
                             -0.5773502691896257645091488, 0.5773502691896257645091488,
                             -0.7745966692414833770358531, 0.0, 0.7745966692414833770358531,
                             -0.8611363115940525752239465, -0.3399810435848562648026658,
                             0.3399810435848562648026658, 0.8611363115940525752239465,
                             -0.9061798459386639927976269, -0.5384693101056830910363144, 0.0,
                             0.5384693101056830910363144, 0.9061798459386639927976269,
                             -0.9324695142031520278123016, -0.6612093864662645136613996,
                             -0.2386191860831969086305017, 0.2386191860831969086305017,
                             0.6612093864662645136613996, 0.9324695142031520278123016};
  static const double W[] = {2.0,
                             1.0, 1.0,
                             0.5555555555555555555555556, 0.8888888888888888888888889, 0.5555555555555555555555556,
                             0.3478548451374538573730639, 0.6521451548625461426269361,
                             0.6521451548625461426269361, 0.3478548451374538573730639,
                             0.2369268850561890875142640, 0.4786286704993664680412915,
                             0.5688888888888888888888889, 0.4786286704993664680412915,
                             0.2369268850561890875142640,
                             0.1713244923791703450402961, 0.3607615730481386075698335,
                             0.4679139345726910473898703, 0.4679139345726910473898703,
                             0.3607615730481386075698335, 0.1713244923791703450402961};
  const bool gl = g.sampling == PPM_SAMPLE_GL;
  const int ng = gl ? g.n_gl : 1;
  const double bw = 1. / g.n_phase;
  pts.resize((size_t)g.n_phase * ng);
  for (int b = 0; b < g.n_phase; ++b)
    for (int k = 0; k < ng; ++k) {
      PhasePoint& p = pts[(size_t)b * ng + k];
      p.bin = b;
      p.pad = 0;
      if (gl) {  // GLNodes / GLWeights (cpu/gl.hpp:52-60)
        const double lb = b * bw, ub = (b + 1) * bw;
        const int o = ng * (ng - 1) / 2 + k;
        p.p = 0.5 * (lb + ub + X[o] * (ub - lb));
        p.w = 0.5 * 1.0 * W[o];
      } else {  // linspace(0, 1 - 1/n, n) (cpu/os1_cpu.cpp:36)
        p.p = (double)b / g.n_phase;
        p.w = 1.0;
      }
    }
}

ppm_status resolve_energies(ppm_ctx* c, const ppm_grid& g, std::vector<double>& e) {
  if (g.energies_keV && g.n_energy > 0) {
    e.assign(g.energies_keV, g.energies_keV + g.n_energy);
  } else if (g.response >= 0 && g.response < (int)c->resp.size()) {
    e = c->resp[g.response].e_obs;
  } else {
    return fail(c, PPM_ERR_ARG, "grid has no energies and no response");
  }
  for (double v : e)
    if (!(v > 0) || !std::isfinite(v)) return fail(c, PPM_ERR_ARG, "energies must be positive");
  return PPM_OK;
}

int Upload::run(Lane& L) {
  size_t total = 0;
  for (auto& s : src) total += (s.second + 255) & ~(size_t)255;
  int e = L.staging.ensure(total);
  if (e) return e;
  // the previous upload from this buffer must have been consumed
  e = hipEventSynchronize(L.ev_staging);
  if (e) return e;
  char* h = (char*)L.staging.p;
  size_t off = 0;
  for (size_t k = 0; k < src.size(); ++k) {
    std::memcpy(h + off, src[k].first, src[k].second);
    e = hipMemcpyAsync(dst[k], h + off, src[k].second, hipMemcpyHostToDevice, L.stream);
    if (e) return e;
    off += (src[k].second + 255) & ~(size_t)255;
  }
  return hipEventRecord(L.ev_staging, L.stream);
}

// One star, plain path: the reference iteration space with one launch
// sequence per profile.  Outputs are device pointers.
// NSX slice rows one workgroup stages: the widest energy chunk plus room for
// the redshift spread of a ring (a ring that needs more is flagged on the
// device and the profile is recomputed with the whole slice).
int nsx_rows_for(const ppm_ctx* c, const std::vector<double>& energies) {
  const int per = energies_per_chunk();
  const double h = (c->N.logE_max - c->N.logE_min) / (c->N.nE - 1);
  double span = 0;
  for (size_t e0 = 0; e0 < energies.size(); e0 += per) {
    const size_t e1 = std::min(energies.size(), e0 + per) - 1;
    span = std::max(span, std::fabs(std::log10(energies[e1] / energies[e0])) / h);
  }
  const int rows = (int)std::ceil(span) + 32 + 8;
  return std::min(rows, c->N.nE);
}

ppm_status eval_plain_rows(ppm_ctx* c, Lane& L, const ppm_star& st, const ppm_grid& g, const std::vector<double>& energies,
                           const std::vector<PhasePoint>& pts, double* d_flux, double* d_counts,
                           double* d_scale_slot, ppm_stats* stats, int rows_override, bool no_alloc);

ppm_status eval_plain(ppm_ctx* c, Lane& L, const ppm_star& st, const ppm_grid& g, const std::vector<double>& energies,
                      const std::vector<PhasePoint>& pts, double* d_flux, double* d_counts,
                      double* d_scale_slot, ppm_stats* stats) {
  return eval_plain_rows(c, L, st, g, energies, pts, d_flux, d_counts, d_scale_slot, stats, 0, false);
}

// Device buffers of one plain evaluation, carved in order from one region
// (the no-allocation scratch) with 256-byte alignment.
struct Carve {
  unsigned char* base;
  size_t cap, used = 0;
  void* take(size_t bytes) {
    const size_t o = (used + 255) & ~(size_t)255;
    used = o + std::max<size_t>(bytes, 1);
    return used <= cap ? base + o : nullptr;
  }
};

ppm_status eval_plain_rows(ppm_ctx* c, Lane& L, const ppm_star& st, const ppm_grid& g, const std::vector<double>& energies,
                           const std::vector<PhasePoint>& pts, double* d_flux, double* d_counts,
                           double* d_scale_slot, ppm_stats* stats, int rows_override, bool no_alloc) {
  const int n_e = (int)energies.size();
  const int nf1 = g.n_fine + 1;
  std::vector<SpotParams> sp(st.n_spots);
  std::vector<RingTask> tasks;
  std::vector<int> rings;
  for (int s = 0; s < st.n_spots; ++s) {
    spot_params(c, st, st.spot[s], g.n_side, sp[s]);
    candidate_rings(sp[s], st.spot[s], g.n_side, rings);
    for (int r : rings) tasks.push_back(RingTask{s, r});
  }
  const int n_tasks = (int)tasks.size();
  int nsx_rows = st.atmosphere == PPM_ATM_NSX ? (rows_override ? rows_override : nsx_rows_for(c, energies)) : 0;
  if (nsx_rows) nsx_rows = std::max(nsx_rows, 64);  // a row set's 64 lanes read inside the allocation
  if (integral_lds_bytes(g.n_fine, st.atmosphere, nsx_rows) > (size_t)c->max_lds)
    return fail(c, PPM_ERR_ARG, "n_fine too large for the LDS of one workgroup");
  int e = 0;
  const size_t nt1 = (size_t)std::max(n_tasks, 1);
  const size_t b_spots = sizeof(SpotParams) * sp.size(), b_tasks = sizeof(RingTask) * nt1,
               b_desc = sizeof(RingDesc) * nt1, b_ph = sizeof(double) * nf1 * nt1,
               b_fine = sizeof(float) * 3 * nf1 * nt1, b_pts = sizeof(PhasePoint) * pts.size(),
               b_en = sizeof(double) * n_e, b_part = sizeof(float) * nt1 * g.n_phase * n_e,
               b_red = sizeof(double) * reduce_groups(n_tasks) * g.n_phase * n_e;
  void *p_spots, *p_tasks, *p_desc, *p_ph, *p_fine, *p_pts, *p_en, *p_part, *p_red;
  if (no_alloc) {
    Carve cv{(unsigned char*)c->noalloc, kNoAllocScratchBytes};
    p_spots = cv.take(b_spots);
    p_tasks = cv.take(b_tasks);
    p_desc = cv.take(b_desc);
    p_ph = cv.take(b_ph);
    p_fine = cv.take(b_fine);
    p_pts = cv.take(b_pts);
    p_en = cv.take(b_en);
    p_part = cv.take(b_part);
    p_red = cv.take(b_red);
    d_scale_slot = (double*)cv.take(sizeof(double));
    if (!c->noalloc || cv.used > cv.cap)
      return fail(c, PPM_ERR_NOMEM,
                  "this profile needs " + std::to_string(cv.used) + " bytes of device scratch; the no-allocation "
                  "entry point has " + std::to_string(kNoAllocScratchBytes) + " (use ppm_eval_device)");
    // host arrays copied directly (no pinned staging buffer, which would
    // have to be allocated); they stay alive until the stream is
    // synchronised below
    if (((e = hipMemcpyAsync(p_spots, sp.data(), b_spots, hipMemcpyHostToDevice, L.stream))) ||
        (n_tasks && (e = hipMemcpyAsync(p_tasks, tasks.data(), sizeof(RingTask) * n_tasks, hipMemcpyHostToDevice,
                                        L.stream))) ||
        (e = hipMemcpyAsync(p_pts, pts.data(), b_pts, hipMemcpyHostToDevice, L.stream)) ||
        (e = hipMemcpyAsync(p_en, energies.data(), b_en, hipMemcpyHostToDevice, L.stream)))
      return hip_fail(c, e, "upload");
  } else {
    Workspace& W = L.ws;
    if ((e = W.spots.ensure(b_spots)) || (e = W.tasks.ensure(b_tasks)) || (e = W.desc.ensure(b_desc)) ||
        (e = W.ph.ensure(b_ph)) || (e = W.fine.ensure(b_fine)) || (e = W.pts.ensure(b_pts)) ||
        (e = W.energies.ensure(b_en)) || (e = W.partial.ensure(b_part)) || (e = W.red.ensure(b_red)))
      return hip_fail(c, e, "device allocation");
    p_spots = W.spots.p;
    p_tasks = W.tasks.p;
    p_desc = W.desc.p;
    p_ph = W.ph.p;
    p_fine = W.fine.p;
    p_pts = W.pts.p;
    p_en = W.energies.p;
    p_part = W.partial.p;
    p_red = W.red.p;
    Upload up;
    up.add(p_spots, sp.data(), b_spots);
    if (n_tasks) up.add(p_tasks, tasks.data(), sizeof(RingTask) * n_tasks);
    up.add(p_pts, pts.data(), b_pts);
    up.add(p_en, energies.data(), b_en);
    if ((e = up.run(L))) return hip_fail(c, e, "upload");
  }

  PlainLaunch P{};
  P.d_spots = (const SpotParams*)p_spots;
  P.d_tasks = (const RingTask*)p_tasks;
  P.d_desc = (RingDesc*)p_desc;
  P.d_ph = (double*)p_ph;
  P.d_fine = (float*)p_fine;
  P.d_pts = (const PhasePoint*)p_pts;
  P.d_energies = (const double*)p_en;
  P.d_partial = (float*)p_part;
  P.d_flux = d_flux;
  P.n_tasks = n_tasks;
  P.n_fine = g.n_fine;
  P.n_pts = (int)pts.size();
  P.n_phase = g.n_phase;
  P.n_e = n_e;
  P.n_echunks = (n_e + energies_per_chunk() - 1) / energies_per_chunk();
  P.atmosphere = st.atmosphere;
  P.emode = st.atmosphere == PPM_ATM_NSX ? nsx_energy_mode(st.nsx_interp) : kECubic;
  P.nsx_rows = nsx_rows;

  if ((e = hipEventRecord(L.ev0, L.stream))) return hip_fail(c, e, "event");
  if ((e = launch_ring_setup(c->L, c->N, P, L.stream))) return hip_fail(c, e, "ring set-up launch");
  if ((e = hipEventRecord(L.ev_i0, L.stream))) return hip_fail(c, e, "event");
  if ((e = launch_integral_plain(c->L, c->N, P, L.stream))) return hip_fail(c, e, "integral launch");
  if ((e = hipEventRecord(L.ev_i1, L.stream))) return hip_fail(c, e, "event");
  if (n_tasks == 0) {
    if ((e = hipMemsetAsync(d_flux, 0, sizeof(double) * g.n_phase * n_e, L.stream)))
      return hip_fail(c, e, "clear");
  } else if ((e = launch_reduce(P, (double*)p_red, L.stream))) {
    return hip_fail(c, e, "reduce launch");
  }
  if ((e = hipEventRecord(L.ev1, L.stream))) return hip_fail(c, e, "event");
  if (d_counts && g.response >= 0) {
    const Response& R = c->resp[g.response];
    const double scale = st.exposure_s * (1.0 / g.n_phase);  // cpu/devarshi24st_cpu.cpp:255
    if ((e = hipMemcpyAsync(d_scale_slot, &scale, sizeof(double), hipMemcpyHostToDevice, L.stream)))
      return hip_fail(c, e, "scale upload");
    if ((e = launch_fold(R.d_hi, R.d_lo, R.n_ch, R.n_e, d_flux, 1, g.n_phase, d_scale_slot, d_counts, L.stream,
                         L.fit, L.fit_star)))
      return hip_fail(c, e, "fold launch");
  }
  if ((e = hipEventRecord(L.ev2, L.stream))) return hip_fail(c, e, "event");
  if ((e = hipEventSynchronize(L.ev2))) return hip_fail(c, e, "execution");
  // read back the descriptors for statistics and run-overflow detection
  std::vector<RingDesc> desc(n_tasks);
  if (n_tasks && (e = hipMemcpy(desc.data(), p_desc, sizeof(RingDesc) * n_tasks, hipMemcpyDeviceToHost)))
    return hip_fail(c, e, "descriptor read-back");
  int64_t n_act = 0, n_eval = 0, n_rings = 0;
  bool rows_overflow = false;
  for (const RingDesc& d : desc) rows_overflow |= (d.flags & 2) != 0;
  if (rows_overflow && nsx_rows < c->N.nE) {
    // a ring's redshift range needs more NSX rows than were staged: redo
    // with the whole slice (results of the first pass are discarded)
    return eval_plain_rows(c, L, st, g, energies, pts, d_flux, d_counts, d_scale_slot, stats, c->N.nE, no_alloc);
  }
  for (const RingDesc& d : desc) {
    if (d.flags & 1) return fail(c, PPM_ERR_UNSUPPORTED, "a ring has more active runs than supported");
    n_act += d.n_act;
    if (d.n_act) ++n_rings;
    if (d.kept) n_eval += (int64_t)d.n_act * (int64_t)pts.size() * n_e;
  }
  if (stats) {
    float t01 = 0, t12 = 0, ti = 0;
    hipEventElapsedTime(&t01, L.ev0, L.ev1);
    hipEventElapsedTime(&t12, L.ev1, L.ev2);
    hipEventElapsedTime(&ti, L.ev_i0, L.ev_i1);
    stats->kernel_ms += t01;
    stats->integral_ms += ti;
    stats->fold_ms += t12;
    stats->total_ms += t01 + t12;
    stats->n_active += n_act;
    stats->n_eval_ref += n_eval;
    stats->n_eval_done += n_eval;
    stats->n_rings += n_rings;
  }
  return PPM_OK;
}

// PPM_METHOD_AUTO for one star: the batched path (phase-shift dedupe, one
// star per launch) where it is the faster one on MI355X and can take the
// grid, else the plain path.  Measured per profile (64 bins x 2 GL, 360
// energies): N_side 64 plain 0.9 ms / batched 1.0 ms, 128 1.6 / 1.5, 1024
// 40 / 7.9, 2048 153 / 22; with 32 bins and 128 energies the batched path is
// ahead from N_side 32 (1.2 / 0.3 ms).
static int auto_single_method(const ppm_ctx* c, const ppm_star& star, const ppm_grid& g,
                              const std::vector<double>& energies) {
  const int n_side_min = energies.size() <= 128 ? 32 : 128;
  if (g.n_side < n_side_min) return PPM_METHOD_PLAIN;
  for (size_t q = 1; q < energies.size(); ++q)
    if (!(energies[q] > energies[q - 1])) return PPM_METHOD_PLAIN;
  BatchGroup G;
  G.stars = &star;
  G.n = 1;
  G.g = g;
  G.energies = energies;
  return batch_groups_fit(c, {G}) ? PPM_METHOD_BATCHED : PPM_METHOD_PLAIN;
}

ppm_status eval_any(ppm_ctx* c, Lane& L, const ppm_star* stars, int n, const ppm_grid* grid, double* d_flux,
                    double* d_counts, ppm_stats* stats) {
  if (!c || !stars || !grid || n < 1 || !d_flux) return fail(c, PPM_ERR_ARG, "null argument");
  std::vector<double> energies;
  ppm_status s = resolve_energies(c, *grid, energies);
  if (s) return s;
  const int n_e = (int)energies.size();
  for (int i = 0; i < n; ++i)
    if ((s = validate(c, stars[i], *grid, n_e))) return s;
  std::vector<PhasePoint> pts;
  phase_points(*grid, pts);
  if (stats) std::memset(stats, 0, sizeof *stats);
  int method = grid->method;
  if (method == PPM_METHOD_AUTO) method = n == 1 ? auto_single_method(c, stars[0], *grid, energies) : PPM_METHOD_BATCHED;
  if (method == PPM_METHOD_BATCHED)
    return eval_batched(c, L, stars, n, *grid, energies, pts, d_flux, d_counts, stats);
  if (method != PPM_METHOD_PLAIN) return fail(c, PPM_ERR_ARG, "unknown method");
  int e = L.ws.scale.ensure(sizeof(double) * n);
  if (e) return hip_fail(c, e, "device allocation");
  const size_t fstride = (size_t)grid->n_phase * n_e;
  const size_t cstride = grid->response >= 0 ? (size_t)c->resp[grid->response].n_ch * grid->n_phase : 0;
  for (int i = 0; i < n; ++i) {
    L.fit_star = i;
    s = eval_plain(c, L, stars[i], *grid, energies, pts, d_flux + i * fstride,
                   d_counts ? d_counts + i * cstride : nullptr, (double*)L.ws.scale.p + i, stats);
    if (s) return s;
  }
  return PPM_OK;
}

}  // namespace ppm

// ================================================================= C ABI

extern "C" {

int ppm_abi_version(void) { return PPM_ABI_VERSION; }

const char* ppm_strerror(ppm_status s) {
  switch (s) {
    case PPM_OK: return "ok";
    case PPM_ERR_ARG: return "invalid argument";
    case PPM_ERR_IO: return "table file missing or unreadable";
    case PPM_ERR_FORMAT: return "table file malformed or inconsistent";
    case PPM_ERR_DEVICE: return "HIP runtime error";
    case PPM_ERR_NOMEM: return "device memory budget exceeded";
    case PPM_ERR_BUSY: return "serving queue full";
    case PPM_ERR_NOT_FOUND: return "unknown ticket";
    case PPM_ERR_CANCELLED: return "request cancelled";
    case PPM_PENDING: return "request pending";
    case PPM_ERR_UNSUPPORTED: return "not supported";
  }
  return "unknown status";
}

const char* ppm_last_error(const ppm_ctx* c) { return c ? c->err.c_str() : ""; }

ppm_status ppm_open(const ppm_open_opts* opts, ppm_ctx** out) {
  if (!out) return PPM_ERR_ARG;
  *out = nullptr;
  ppm_open_opts o{};
  if (opts) o = *opts;
  const std::string root = o.root ? o.root : ".";
  const std::string lset = o.lensing_setting ? o.lensing_setting : "std";
  std::string err;
  // ---- parse every table on the host first (errors need no device)
  LensingHost lens;
  ppm_status st = read_lensing(root + "/tables/lensing/" + lset, lens, err);
  if (st) {
    std::fprintf(stderr, "ppm_open: %s\n", err.c_str());
    return st;
  }
  const std::string nsx_path = o.nsx_path ? std::string(o.nsx_path) : root + "/ext/model_data/nsx_H_v200804.out";
  NsxHost nh;
  const ppm_nsx_config cfg = o.nsx_config ? *o.nsx_config : default_nsx_config();
  if (!nsx_path.empty()) {
    st = read_nsx(nsx_path, cfg, nh, err);
    if (st) {
      std::fprintf(stderr, "ppm_open: %s\n", err.c_str());
      return st;
    }
  }
  std::vector<ResponseHost> rhs;
  if (o.response_settings && *o.response_settings) {
    const std::string list = o.response_settings;
    size_t pos = 0;
    while (pos <= list.size()) {
      size_t q = list.find(',', pos);
      if (q == std::string::npos) q = list.size();
      const std::string name = list.substr(pos, q - pos);
      pos = q + 1;
      if (name.empty()) continue;
      ResponseHost rh;
      rh.setting = name;
      st = read_response(root + "/tables/ism_inst/" + name, rh, err);
      if (st) {
        std::fprintf(stderr, "ppm_open: %s\n", err.c_str());
        return st;
      }
      rhs.push_back(std::move(rh));
    }
  }
  // ---- device
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= o.device || o.device < 0) return PPM_ERR_DEVICE;
  auto* c = new ppm_ctx();
  c->device = o.device;
  c->budget = o.hbm_budget_bytes ? o.hbm_budget_bytes : ((size_t)48 << 30);
  int e;
  if ((e = hipSetDevice(o.device)) || (e = c->main.init()) ||
      (e = hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, o.device))) {
    ppm_close(c);
    return PPM_ERR_DEVICE;
  }
  c->max_lds = max_dynamic_lds(o.device);
  if ((e = noalloc_scratch_address(o.device, &c->noalloc))) {
    ppm_close(c);
    return PPM_ERR_DEVICE;
  }
  {
    const char* ng = std::getenv("PPM_NO_GRAPH");
    c->no_graph = ng && *ng && *ng != '0';
    c->main.no_graph = c->no_graph;
    c->main.acc_cap = std::min(kAccCapBytes, c->budget);
  }
  c->lens = std::move(lens);
  {  // lensing tables, FP64 on the device
    const LensingHost& h = c->lens;
    const size_t n1 = h.ca.size(), n2 = h.lf.size(), n3 = h.dt.size();
    double* d = nullptr;
    if (!(e = c->d_lens.ensure(sizeof(double) * (n1 + n2 + n3)))) {
      d = (double*)c->d_lens.p;
      if (!(e = hipMemcpy(d, h.ca.data(), sizeof(double) * n1, hipMemcpyHostToDevice)) &&
          !(e = hipMemcpy(d + n1, h.lf.data(), sizeof(double) * n2, hipMemcpyHostToDevice)))
        e = hipMemcpy(d + n1 + n2, h.dt.data(), sizeof(double) * n3, hipMemcpyHostToDevice);
    }
    if (e) {
      ppm_close(c);
      return PPM_ERR_DEVICE;
    }
    c->L = DevLensing{d, d + n1, d + n1 + n2, h.n_u, h.n_cp, h.n_ca, 0,
                      h.u_min, h.u_max, h.cp_min, h.cp_max, h.ca_min, h.ca_max};
  }
  if (!nsx_path.empty()) {  // NSX grid, FP32 on the device scaled to a unit maximum
    double vmax = 0;
    for (double v : nh.data) vmax = std::max(vmax, std::fabs(v));
    const double scale = vmax > 0 ? std::ldexp(1.0, -std::ilogb(vmax)) : 1.0;
    std::vector<float> f(nh.data.size());
    for (size_t k = 0; k < f.size(); ++k) f[k] = (float)(nh.data[k] * scale);
    // mu nodes, then the inverse denominators of the cubic Lagrange weights
    // of the stencil starting at each node (cpu/nsx.hpp:262-275)
    const size_t nm = nh.mu.size();
    std::vector<double> mu(nm * 5, 0.0);
    std::copy(nh.mu.begin(), nh.mu.end(), mu.begin());
    for (size_t i = 0; i + 3 < nm; ++i) {
      const double* x = nh.mu.data() + i;
      for (int k = 0; k < 4; ++k) {
        double d = 1;
        for (int m = 0; m < 4; ++m)
          if (m != k) d *= x[k] - x[m];
        mu[nm + 4 * i + k] = 1.0 / d;
      }
    }
    if ((e = c->d_nsx.ensure(sizeof(float) * f.size())) || (e = c->d_mu.ensure(sizeof(double) * mu.size())) ||
        (e = hipMemcpy(c->d_nsx.p, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(c->d_mu.p, mu.data(), sizeof(double) * mu.size(), hipMemcpyHostToDevice))) {
      ppm_close(c);
      return PPM_ERR_DEVICE;
    }
    DevNsx& N = c->N;
    N.data = (const float*)c->d_nsx.p;
    N.mu = (const double*)c->d_mu.p;
    N.mu_den = N.mu + nm;
    N.nT = cfg.n_logT;
    N.ng = cfg.n_logg;
    N.nE = cfg.n_logEkT;
    N.nm = cfg.n_mu;
    N.logT_min = cfg.logT_min;
    N.logT_max = cfg.logT_max;
    N.logg_min = cfg.logg_min;
    N.logg_max = cfg.logg_max;
    N.logE_min = cfg.logEkT_min;
    N.logE_max = cfg.logEkT_max;
    N.scale = scale;
    N.mu_min = nh.mu.back();
    c->nsx_mu = nh.mu;
  }
  for (ResponseHost& rh : rhs) {  // rsp split into BF16 hi/lo for the MFMA fold
    Response R;
    R.setting = rh.setting;
    R.n_ch = rh.n_ch;
    R.n_e = rh.n_e;
    std::vector<uint16_t> hl(2 * rh.rsp.size());
    for (size_t k = 0; k < rh.rsp.size(); ++k) {
      const float f = (float)rh.rsp[k];
      hl[k] = bf16_round(f);
      hl[rh.rsp.size() + k] = bf16_round(f - bf16_value(hl[k]));
    }
    R.e_obs = std::move(rh.e_obs);
    R.rsp = std::move(rh.rsp);
    c->resp.push_back(std::move(R));
    Response& RR = c->resp.back();
    if ((e = RR.buf.ensure(sizeof(uint16_t) * hl.size())) ||
        (e = hipMemcpy(RR.buf.p, hl.data(), sizeof(uint16_t) * hl.size(), hipMemcpyHostToDevice))) {
      ppm_close(c);
      return PPM_ERR_DEVICE;
    }
    RR.d_hi = (const uint16_t*)RR.buf.p;
    RR.d_lo = RR.d_hi + RR.rsp.size();
  }
  serve_init(c);
  *out = c;
  return PPM_OK;
}

ppm_status ppm_close(ppm_ctx* c) {
  if (!c) return PPM_ERR_ARG;
  serve_shutdown(c);
  hipSetDevice(c->device);
  c->main.release();
  release_batch_state(c);
  c->d_lens.release();
  c->d_nsx.release();
  c->d_mu.release();
  for (auto& r : c->resp) r.buf.release();
  delete c;
  return PPM_OK;
}

ppm_status ppm_response_info(const ppm_ctx* c, int response, int* n_ch, int* n_energy, double* energies) {
  if (!c || response < 0 || response >= (int)c->resp.size()) return PPM_ERR_ARG;
  const Response& R = c->resp[response];
  if (n_ch) *n_ch = R.n_ch;
  if (n_energy) *n_energy = R.n_e;
  if (energies) std::memcpy(energies, R.e_obs.data(), sizeof(double) * R.n_e);
  return PPM_OK;
}

ppm_status ppm_lensing_domain(const ppm_ctx* c, double* u_min, double* u_max, double* cos_psi_min) {
  if (!c) return PPM_ERR_ARG;
  if (u_min) *u_min = c->lens.u_min;
  if (u_max) *u_max = c->lens.u_max;
  if (cos_psi_min) *cos_psi_min = c->lens.cp_min;
  return PPM_OK;
}

ppm_status ppm_eval_device(ppm_ctx* c, const ppm_star* stars, int n, const ppm_grid* grid, double* d_flux,
                           double* d_counts, ppm_stats* stats) {
  if (!c) return PPM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->exec_mu);
  hipSetDevice(c->device);
  return eval_any(c, c->main, stars, n, grid, d_flux, d_counts, stats);
}

ppm_status ppm_eval_device_noalloc(ppm_ctx* c, const ppm_star* star, const ppm_grid* grid, double* d_flux,
                                   double* d_counts, ppm_stats* stats) {
  if (!c || !star || !grid || !d_flux) return fail(c, PPM_ERR_ARG, "null argument");
  if (grid->method == PPM_METHOD_BATCHED)
    return fail(c, PPM_ERR_ARG, "the no-allocation entry point evaluates on the plain path only");
  std::lock_guard<std::mutex> lk(c->exec_mu);
  hipSetDevice(c->device);
  std::vector<double> energies;
  ppm_status s = resolve_energies(c, *grid, energies);
  if (s) return s;
  if ((s = validate(c, *star, *grid, (int)energies.size()))) return s;
  std::vector<PhasePoint> pts;
  phase_points(*grid, pts);
  if (stats) std::memset(stats, 0, sizeof *stats);
  return eval_plain_rows(c, c->main, *star, *grid, energies, pts, d_flux, grid->response >= 0 ? d_counts : nullptr,
                         nullptr,
                         stats, 0, true);
}

ppm_status ppm_eval(ppm_ctx* c, const ppm_star* stars, int n, const ppm_grid* grid, double* flux,
                    double* counts, ppm_stats* stats) {
  if (!c || !stars || !grid || n < 1 || !flux) return fail(c, PPM_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->exec_mu);
  hipSetDevice(c->device);
  Lane& L = c->main;
  std::vector<double> energies;
  ppm_status s = resolve_energies(c, *grid, energies);
  if (s) return s;
  const size_t nf = (size_t)n * grid->n_phase * energies.size();
  const bool fold = counts && grid->response >= 0 && grid->response < (int)c->resp.size();
  const size_t nc = fold ? (size_t)n * c->resp[grid->response].n_ch * grid->n_phase : 0;
  int e;
  if ((e = L.ws.out_flux.ensure(sizeof(double) * nf)) || (fold && (e = L.ws.out_counts.ensure(sizeof(double) * nc))))
    return hip_fail(c, e, "device allocation");
  s = eval_any(c, L, stars, n, grid, (double*)L.ws.out_flux.p, fold ? (double*)L.ws.out_counts.p : nullptr, stats);
  if (s) return s;
  if ((e = hipMemcpy(flux, L.ws.out_flux.p, sizeof(double) * nf, hipMemcpyDeviceToHost)))
    return hip_fail(c, e, "read-back");
  if (fold && (e = hipMemcpy(counts, L.ws.out_counts.p, sizeof(double) * nc, hipMemcpyDeviceToHost)))
    return hip_fail(c, e, "read-back");
  return PPM_OK;
}

ppm_status ppm_count_active(ppm_ctx* c, const ppm_star* star, const ppm_grid* grid, int64_t* n_active) {
  if (!c || !star || !grid || !n_active) return fail(c, PPM_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->exec_mu);
  hipSetDevice(c->device);
  Lane& L = c->main;
  ppm_grid g = *grid;
  g.response = -1;
  g.n_fine = 4;
  double e1 = 1.0;
  g.energies_keV = &e1;
  g.n_energy = 1;
  g.sampling = PPM_SAMPLE_POINTS;
  g.n_phase = 1;
  ppm_star st = *star;
  ppm_status s = validate(c, st, g, 1);
  if (s) return s;
  std::vector<RingTask> tasks;
  std::vector<SpotParams> sp(st.n_spots);
  std::vector<int> rings;
  for (int k = 0; k < st.n_spots; ++k) {
    spot_params(c, st, st.spot[k], g.n_side, sp[k]);
    candidate_rings(sp[k], st.spot[k], g.n_side, rings);
    for (int r : rings) tasks.push_back(RingTask{k, r});
  }
  *n_active = 0;
  if (tasks.empty()) return PPM_OK;
  Workspace& W = L.ws;
  int e;
  const int nt = (int)tasks.size();
  if ((e = W.spots.ensure(sizeof(SpotParams) * sp.size())) || (e = W.tasks.ensure(sizeof(RingTask) * nt)) ||
      (e = W.desc.ensure(sizeof(RingDesc) * nt)) || (e = W.ph.ensure(sizeof(double) * 5 * nt)) ||
      (e = W.fine.ensure(sizeof(float) * 15 * nt)))
    return hip_fail(c, e, "device allocation");
  Upload up;
  up.add(W.spots.p, sp.data(), sizeof(SpotParams) * sp.size());
  up.add(W.tasks.p, tasks.data(), sizeof(RingTask) * nt);
  if ((e = up.run(L))) return hip_fail(c, e, "upload");
  PlainLaunch P{};
  P.d_spots = (const SpotParams*)W.spots.p;
  P.d_tasks = (const RingTask*)W.tasks.p;
  P.d_desc = (RingDesc*)W.desc.p;
  P.d_ph = (double*)W.ph.p;
  P.d_fine = (float*)W.fine.p;
  P.n_tasks = nt;
  P.n_fine = 4;
  if ((e = launch_ring_setup(c->L, c->N, P, L.stream))) return hip_fail(c, e, "ring set-up launch");
  std::vector<RingDesc> desc(nt);
  if ((e = hipMemcpyAsync(desc.data(), W.desc.p, sizeof(RingDesc) * nt, hipMemcpyDeviceToHost, L.stream)) ||
      (e = hipStreamSynchronize(L.stream)))
    return hip_fail(c, e, "descriptor read-back");
  for (const RingDesc& d : desc) *n_active += d.n_act;
  return PPM_OK;
}

ppm_status ppm_fold_device(ppm_ctx* c, int response, int n_batch, int n_phase, const double* d_flux,
                           const double* scale, double* d_counts) {
  if (!c || !d_flux || !scale || !d_counts || n_batch < 1 || n_phase < 1) return fail(c, PPM_ERR_ARG, "bad argument");
  if (response < 0 || response >= (int)c->resp.size()) return fail(c, PPM_ERR_ARG, "response index out of range");
  std::lock_guard<std::mutex> lk(c->exec_mu);
  hipSetDevice(c->device);
  Lane& L = c->main;
  int e;
  if ((e = L.ws.scale.ensure(sizeof(double) * n_batch))) return hip_fail(c, e, "device allocation");
  if ((e = hipMemcpyAsync(L.ws.scale.p, scale, sizeof(double) * n_batch, hipMemcpyHostToDevice, L.stream)))
    return hip_fail(c, e, "upload");
  const Response& R = c->resp[response];
  if ((e = launch_fold(R.d_hi, R.d_lo, R.n_ch, R.n_e, d_flux, n_batch, n_phase, (const double*)L.ws.scale.p,
                       d_counts, L.stream)))
    return hip_fail(c, e, "fold launch");
  if ((e = hipStreamSynchronize(L.stream))) return hip_fail(c, e, "execution");
  return PPM_OK;
}

ppm_status ppm_eval_fit(ppm_ctx* c, const ppm_star* stars, int n, const ppm_grid* grid, const double* data,
                        int data_per_star, ppm_fit* fit, double* counts, ppm_stats* stats) {
  if (!c || !stars || !grid || n < 1 || !data || !fit) return fail(c, PPM_ERR_ARG, "null argument");
  if (grid->response < 0 || grid->response >= (int)c->resp.size())
    return fail(c, PPM_ERR_ARG, "a fit needs the grid's response fold");
  if (data_per_star != 0 && data_per_star != 1) return fail(c, PPM_ERR_ARG, "data_per_star must be 0 or 1");
  std::lock_guard<std::mutex> lk(c->exec_mu);
  hipSetDevice(c->device);
  Lane& L = c->main;
  std::vector<double> energies;
  ppm_status s = resolve_energies(c, *grid, energies);
  if (s) return s;
  const int n_ch = c->resp[grid->response].n_ch, blocks = fold_blocks(n_ch, grid->n_phase);
  const size_t per = (size_t)n_ch * grid->n_phase;
  const size_t nf = (size_t)n * grid->n_phase * energies.size(), nc = (size_t)n * per;
  const size_t nd = (data_per_star ? (size_t)n : 1) * per;
  Workspace& W = L.ws;
  int e;
  if ((e = W.out_flux.ensure(sizeof(double) * nf)) || (e = W.out_counts.ensure(sizeof(double) * nc)) ||
      (e = W.fit_data.ensure(sizeof(double) * nd)) || (e = W.fit_part.ensure(sizeof(double) * 3 * n * blocks)) ||
      (e = W.fit_out.ensure(sizeof(double) * 3 * n)))
    return hip_fail(c, e, "device allocation");
  if ((e = hipMemcpy(W.fit_data.p, data, sizeof(double) * nd, hipMemcpyHostToDevice)))
    return hip_fail(c, e, "data upload");
  const FoldFit F{(const double*)W.fit_data.p, data_per_star ? per : 0, (double*)W.fit_part.p, blocks, 0};
  struct Reset {
    Lane& L;
    ~Reset() {
      L.fit = nullptr;
      L.fit_star = 0;
    }
  } reset{L};
  L.fit = &F;
  L.fit_star = 0;
  if ((s = eval_any(c, L, stars, n, grid, (double*)W.out_flux.p, (double*)W.out_counts.p, stats))) return s;
  std::vector<double> out(3 * (size_t)n);
  if ((e = launch_fit_reduce((const double*)W.fit_part.p, blocks, n, (double*)W.fit_out.p, L.stream)) ||
      (e = hipMemcpyAsync(out.data(), W.fit_out.p, sizeof(double) * out.size(), hipMemcpyDeviceToHost, L.stream)) ||
      (counts && (e = hipMemcpyAsync(counts, W.out_counts.p, sizeof(double) * nc, hipMemcpyDeviceToHost, L.stream))) ||
      (e = hipStreamSynchronize(L.stream)))
    return hip_fail(c, e, "fit statistics");
  for (int b = 0; b < n; ++b) fit[b] = ppm_fit{out[3 * b], out[3 * b + 1], out[3 * b + 2]};
  return PPM_OK;
}

}  // extern "C"
