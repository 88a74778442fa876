// Internal types shared by the host driver (ppm_api.cpp) and the HIP kernels
// (ppm_kernels.hip).  Not part of the ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ppm.h"

namespace ppm {

// Physical constants (gpu/unit.cuh:3-22, cpu/unit.hpp:5-25).
constexpr double kRsSunKm = 2.9532501;
constexpr double kSqrtKm3S2GMsun = 2.745011592867327e-6;
constexpr double kPi = 3.14159265358979323846;
constexpr double kCKmS = 299792458. / 1e3;
constexpr double kCCmS = 299792458. * 1e2;
constexpr double kHKeVs = 4.135668e-18;
constexpr double kPlanckErgS = 6.62607015E-27;
constexpr double kBoltzKeVK = 8.617333262E-8;
constexpr double kGMsunCgs = 1.32712440041279419E16;

// The NSX interpolants (PPM_NSX_*) in three parts: the energy-axis
// evaluation (a kernel template parameter), the mu stencil and the
// (logT, log g) stencil (both per spot).
enum NsxEnergyMode { kECubic = 0, kECubicClip = 1, kELinear = 2 };
enum NsxMuMode { kMuCubicLinearEdge = 0, kMuCubic = 1, kMuLinear = 2 };
constexpr int nsx_energy_mode(int interp) {
  return interp == PPM_NSX_4L ? kELinear : interp == PPM_NSX_4C_SG ? kECubicClip : kECubic;
}
constexpr int nsx_mu_mode(int interp) {
  return interp == PPM_NSX_4L ? kMuLinear
         : (interp == PPM_NSX_4C || interp == PPM_NSX_4C_SG) ? kMuCubic
                                                             : kMuCubicLinearEdge;
}
constexpr bool nsx_cubic_tg(int interp) {
  return interp == PPM_NSX_4C || interp == PPM_NSX_4C_SG || interp == PPM_NSX_4C_LE;
}

constexpr int kMaxRuns = 16;      // active-pixel runs per ring kept in a descriptor
constexpr int kSliceStride = 69;  // LDS row stride of the NSX slice (67 mu + 2 zero pad)

// Per-(star, spot) parameters, all FP64, prepared on the host.
struct SpotParams {
  // star
  double M, Re, nu, Rs, x, Ob, o2, inv_D2, ci, si, logg_override;
  double logg0_total;  // log10(M / (Re^2 sqrt(1 - Rs/Re))) + log10(GM_sun cgs)
  int atmosphere, variant, beaming, membership;
  // spot
  double st1, ct1, phi1, cr1, st2, ct2, phi2, cr2;
  int has2, point_source;
  double ps_z, ps_domega, ps_shift;  // point-source element
  double kT, inv_kT, logT, kT_nsx;   // temperatures
  double scale;                      // FP64 constant applied at the reduction
  int n_side;
  int nsx_interp;                    // PPM_NSX_*
};

// One candidate ring of one spot (input of the setup kernel).
struct RingTask {
  int spot;   // index into SpotParams
  int ring;   // HEALPix ring 1..4N-1, or 0 for a point-source element
};

// Output of the setup kernel, one per RingTask.
// RingDesc::hidden: every fine node visible; some invisible (flux/I = 0
// there, the patch skip of cpu/devarshi24st_cpu.cpp:226 applies); or some
// invisible and the zeros of the interpolated flux/I are the phase interval
// [gap_lo, gap_hi] (one gap of invisible nodes, ascending phase table).
enum { kHiddenNone = 0, kHiddenLookup = 1, kHiddenInterval = 2 };

struct RingDesc {
  double z, d_omega, logg;
  double zf_min, zf_max;      // range of the redshift factor over fine nodes
  double ps_shift;            // point source: phase shift of the element
  int n_pix, s_shift2;        // pixels in the ring, 2*s of phi_j = 2pi (j - s/2)/n_pix
  int n_act, n_runs;          // active patches and runs
  int kept;                   // 1 if the ring contributes
  int flags;                  // bit0: run overflow, bit1: NSX rows beyond the staged slice
  int iT, ig;                 // first (logT, log g) node of the NSX stencil
  int n_tg;                   // stencil nodes per axis: 2 (linear) or 4 (cubic)
  int hidden;                 // kHidden*: whether and how some fine node is invisible
  double gap_lo, gap_hi;      // kHiddenInterval: flux/I = 0 exactly for phases in [gap_lo, gap_hi]
  float wtg[16];              // weight of node (iT + a, ig + b) at [4 a + b]
  int run_start[kMaxRuns];    // first active pixel j (1-based) of each run
  int run_len[kMaxRuns];
};

// Phase sampling point of the output grid.
struct PhasePoint {
  double p;   // phase in [0, 1)
  double w;   // weight (GL weight or 1)
  int bin;    // output row
  int pad;
};

// Device-resident tables.
struct DevLensing {
  const double* ca;  // [n_u][n_cp]
  const double* lf;  // [n_u][n_cp]
  const double* dt;  // [n_u][n_ca]
  int n_u, n_cp, n_ca, pad;
  double u_min, u_max, cp_min, cp_max, ca_min, ca_max;
};

struct DevNsx {
  const float* data;   // [nT][ng][nE][nm], scaled by `scale`
  const double* mu;    // [nm] decreasing
  const double* mu_den;  // [nm][4] inverse Lagrange denominators of the stencil starting at each node
  int nT, ng, nE, nm;
  double logT_min, logT_max, logg_min, logg_max, logE_min, logE_max;
  double scale;        // data = 10^logInuT3 * scale
  double mu_min;
};

struct Workspace;  // device buffers (ppm_api.cpp)

// Arguments of one integral launch sequence (plain path).
struct PlainLaunch {
  const SpotParams* d_spots;
  const RingTask* d_tasks;
  RingDesc* d_desc;
  double* d_ph;          // [n_tasks][nf1] fine phase_o
  float* d_fine;         // [n_tasks][3][nf1] fI, zf, cos_sigma'
  const PhasePoint* d_pts;
  const double* d_energies;
  float* d_partial;      // [n_tasks][n_phase][n_e]
  double* d_flux;        // [n_phase][n_e] output (one star)
  int n_tasks, n_fine, n_pts, n_phase, n_e, n_echunks;
  int atmosphere;
  int emode;             // NsxEnergyMode of the star's interpolant
  int nsx_rows;          // NSX slice rows staged per workgroup
  // ring set-up counters, or null: [0] active patches, [1] active patches of
  // kept rings, [2] rings with active patches, [3] flags of all rings (OR)
  unsigned long long* summary;
};

// Kernel launchers (ppm_kernels.hip).  Return a hipError_t as int.
int launch_ring_setup(const DevLensing& L, const DevNsx& N, const PlainLaunch& P, void* stream);
int launch_integral_plain(const DevLensing& L, const DevNsx& N, const PlainLaunch& P, void* stream);
int launch_reduce(const PlainLaunch& P, double* d_red, void* stream);
int reduce_groups(int n_tasks);
// Fit statistics against data computed in the fold's epilogue: per star
// partial sums over the fold's blocks, reduced by launch_fit_reduce.
struct FoldFit {
  const double* data;  // [stars][n_ch][n_phase] (stride between stars; 0: one shared array)
  size_t stride;
  double* part;        // [stars][blocks][3]: sum m, sum (d-m)^2/m, sum d ln m - m - lnGamma(d+1)
  int blocks;          // fold_blocks(n_ch, n_phase)
  int pad;
};
int fold_blocks(int n_ch, int n_phase);
int launch_fold(const uint16_t* rsp_hi, const uint16_t* rsp_lo, int n_ch, int n_e,
                const double* d_flux, int n_batch, int n_phase, const double* d_scale,
                double* d_counts, void* stream, const FoldFit* fit = nullptr, int star0 = 0);
int launch_fit_reduce(const double* d_part, int blocks, int n, double* d_out, void* stream);
int max_dynamic_lds(int device);
// Device scratch of ppm_eval_device_noalloc (a file-scope array in
// ppm_kernels.hip, kNoAllocScratchBytes per GPU): its address on `device`,
// looked up once per device.  Returns a hipError_t as int.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
constexpr int kMaxDevices = 64;
int noalloc_scratch_address(int device, void** ptr);
size_t integral_lds_bytes(int n_fine, int atmosphere, int nsx_rows);
int energies_per_chunk();

}  // namespace ppm
