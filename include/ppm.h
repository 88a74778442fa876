/*
 * ppm.h - C ABI of libppm, an MI355X (gfx950) implementation of the
 * pulse-profile hot path of fizban007/gpu_ppm.
 *
 * The reference exposes no library: its boundary is a process that reads
 * table files and writes output files.  The entry points below are the ones
 * a foreign-function binding of that path would need; each names the
 * reference routine it replaces.  Conventions:
 *   - every call returns a ppm_status; nothing exits or throws across the ABI;
 *   - host buffers are owned by the caller; an opaque ppm_ctx owns all device
 *     memory (tables stay resident in HBM until ppm_close);
 *   - one ctx is used from one host thread at a time, except ppm_submit /
 *     ppm_poll / ppm_cancel, which are internally synchronised;
 *   - plain C types only (no HIP or torch types).
 */
#ifndef PPM_H
#define PPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PPM_ABI_VERSION 2
#define PPM_MAX_SPOTS 4

typedef enum ppm_status {
  PPM_OK = 0,
  PPM_ERR_ARG = 1,        /* invalid argument or parameter out of domain   */
  PPM_ERR_IO = 2,         /* table file missing or unreadable              */
  PPM_ERR_FORMAT = 3,     /* table contents malformed or inconsistent      */
  PPM_ERR_DEVICE = 4,     /* HIP runtime error or no usable device         */
  PPM_ERR_NOMEM = 5,      /* request exceeds the device-memory budget      */
  PPM_ERR_BUSY = 6,       /* serving queue full                            */
  PPM_ERR_NOT_FOUND = 7,  /* unknown ticket                                */
  PPM_ERR_CANCELLED = 8,  /* request was cancelled                         */
  PPM_PENDING = 9,        /* request admitted but not finished             */
  PPM_ERR_UNSUPPORTED = 10 /* feature not available in this build / ctx    */
} ppm_status;

/* Atmosphere model of a spot. */
enum {
  PPM_ATM_BLACKBODY = 0, /* Ibb(E, kT), cpu/unit.hpp:28-31, os1 path            */
  PPM_ATM_NSX = 1        /* NSX grid, interpolant ppm_star.nsx_interp, cpu/nsx.hpp */
};

/* Interpolant of the NSX grid (ppm_star.nsx_interp).  Every axis look-up
 * clamps to the axis ends (cpu/hunt.hpp:155-178); a cubic stencil starts one
 * node below the bracket, clamped to [0, n-4].  t_interp's interp_type
 * 0, 1, 2, 3 (cpu/t_interp.cpp:210-219) are 4C, 4C_SG, 4C_LE, 2L2C_LE. */
enum {
  PPM_NSX_2L2C_LE = 0, /* linear logT, log g; cubic logEkT; cubic mu, linear
                          in the last mu brackets (Interp_IT3_2l2c_le,
                          cpu/nsx.hpp:248-296; os1/devarshi programs)        */
  PPM_NSX_4C = 1,      /* cubic in all four (Interp_IT3_4c, use_sg false,
                          cpu/nsx.hpp:170-203)                                */
  PPM_NSX_4C_SG = 2,   /* 4C with negative results set to 0 (use_sg true)   */
  PPM_NSX_4C_LE = 3,   /* cubic logT, log g, logEkT; mu as 2L2C_LE
                          (Interp_IT3_4c_le, cpu/nsx.hpp:205-246)             */
  PPM_NSX_4L = 4       /* linear in all four (Interp_IT3_4l, cpu/nsx.hpp:298-321) */
};

/* Beaming of the blackbody atmosphere (cpu/os1_cpu.cpp:128-136). */
enum { PPM_BEAM_ISOTROPIC = 'i', PPM_BEAM_COS2 = 'c', PPM_BEAM_SIN2 = 's' };

/* Patch-membership rule of a spot.  Both are FP64 tests on HEALPix pixel
 * centres; they differ in arithmetic only:
 *   CAP:      cos rho = sin th1 sqrt(1-z^2) cos(phi-phi1) + cos th1 z  (cpu/grid.hpp:94-156)
 *   DISKDIFF: theta = acos z, disk1 \ disk2 via sin/cos of theta       (cpu/grid.hpp:158-250) */
enum { PPM_MEMBER_CAP = 0, PPM_MEMBER_DISKDIFF = 1 };

/* Form of the cos(sigma) expression of the per-element kernel.  Analytically
 * equal; kept to reproduce each reference program to the last bits.
 *   OS1:      cpu/os1_cpu.cpp:95-96
 *   DEVARSHI: cpu/devarshi24st_cpu.cpp:145-147 */
enum { PPM_VARIANT_OS1 = 0, PPM_VARIANT_DEVARSHI = 1 };

/* Phase sampling: os1 samples points k/n_phase (cpu/os1_cpu.cpp:36,169-181);
 * devarshi integrates each bin with n_gl Gauss-Legendre nodes
 * (cpu/devarshi24st_cpu.cpp:214-239). */
enum { PPM_SAMPLE_POINTS = 0, PPM_SAMPLE_GL = 1 };

/* Evaluation method. */
enum {
  PPM_METHOD_AUTO = 0,    /* BATCHED for several stars; for one star BATCHED
                           * from N_side 128 (32 with <= 128 energies) when
                           * its energies ascend and it fits the LDS, else PLAIN */
  PPM_METHOD_PLAIN = 1,   /* reference iteration space, one profile per launch */
  PPM_METHOD_BATCHED = 2  /* persistent work queue + phase-shift dedupe        */
};

typedef struct ppm_spot {
  double theta1, phi1, angrad1; /* disk 1: centre colatitude, azimuth, radius (rad) */
  double theta2, phi2, angrad2; /* disk 2 (subtracted) when angrad2 > 0              */
  double kT_keV;                /* temperature; NSX uses logT = log10(kT / k_B)      */
  int membership;               /* PPM_MEMBER_*                                      */
  int beaming;                  /* PPM_BEAM_* (blackbody only)                       */
} ppm_spot;

typedef struct ppm_star {
  double mass_msun;        /* M                                         */
  double radius_eq_km;     /* equatorial radius R_eq                    */
  double spin_hz;          /* rotation frequency nu                     */
  double inclination_rad;  /* observer inclination i                    */
  double distance_km;      /* D                                         */
  double exposure_s;       /* fold scale (cpu/devarshi24st_cpu.cpp:255) */
  double logg_override;    /* > 0: use this log g for every ring        */
  double point_source_below; /* os1 shortcut: a spot with angrad1 below
                                this is one element (cpu/os1_cpu.cpp:40-42) */
  int atmosphere;          /* PPM_ATM_*                                 */
  int variant;             /* PPM_VARIANT_*                             */
  int n_spots;             /* 1..PPM_MAX_SPOTS; contributions add       */
  int nsx_interp;          /* PPM_NSX_* (NSX atmosphere; 0 = 2L2C_LE)   */
  ppm_spot spot[PPM_MAX_SPOTS];
} ppm_star;

typedef struct ppm_grid {
  int n_side;        /* HEALPix N_side                                      */
  int n_phase;       /* output phase points or bins                         */
  int sampling;      /* PPM_SAMPLE_*                                        */
  int n_gl;          /* GL nodes per bin (1..6), PPM_SAMPLE_GL only         */
  int n_fine;        /* fine rotational-phase nodes per ring                */
  int n_energy;      /* number of observed energies                         */
  const double* energies_keV; /* [n_energy]; NULL -> the response's E_obs   */
  int response;      /* index into the ctx's response list, -1: no fold      */
  int method;        /* PPM_METHOD_*                                         */
} ppm_grid;

/* Dimensions and axes of an NSX grid (cpu/nsx.hpp:11-53). */
typedef struct ppm_nsx_config {
  int n_logT, n_logg, n_logEkT, n_mu;
  double logT_min, logT_max;
  double logg_min, logg_max;
  double logEkT_min, logEkT_max;
} ppm_nsx_config;

typedef struct ppm_open_opts {
  const char* root;              /* directory holding tables/ and ext/ ("." if NULL) */
  const char* lensing_setting;   /* tables/lensing/<setting>; NULL -> "std"          */
  const char* nsx_path;          /* NULL -> ext/model_data/nsx_H_v200804.out;
                                    "" -> no NSX grid                                */
  const char* response_settings; /* comma list, tables/ism_inst/<s>; NULL/"" -> none */
  const ppm_nsx_config* nsx_config; /* NULL -> nsx_H_v200804 (35x14x166x67);
                                       nsx_H_v171019 is 35x11x166x67, log g
                                       13.7-14.7 (cpu/nsx.hpp:39-51)         */
  int device;                    /* HIP device ordinal                               */
  int reserved;
  size_t hbm_budget_bytes;       /* device memory of batched launches and of the
                                    serving queue (admission); 0 -> 48 GiB          */
} ppm_open_opts;

typedef struct ppm_stats {
  double kernel_ms;        /* set-up + integral + reduction (hipEvent)       */
  double integral_ms;      /* the integral kernel alone (hipEvent)           */
  double fold_ms;          /* response fold (hipEvent)                       */
  double total_ms;         /* event pair around integral + fold              */
  int64_t n_active;        /* active patches summed over stars and spots     */
  int64_t n_eval_ref;      /* evaluations in the reference iteration space  */
  int64_t n_eval_done;     /* (phase point, energy) evaluations performed    */
  int64_t n_rings;         /* ring tasks with active patches                 */
} ppm_stats;

typedef struct ppm_ctx ppm_ctx;

/* Version / errors. */
int ppm_abi_version(void);
const char* ppm_strerror(ppm_status status);
/* Message of the last failed call on ctx (thread-unsafe view; "" if none). */
const char* ppm_last_error(const ppm_ctx* ctx);

/* Load the reference-format tables and keep them resident in HBM.
 * Replaces Lens::Lens (gpu/lensing.cuh:27-54), NSX::NSX (gpu/nsx.cuh:83-128)
 * and IsmInst::IsmInst (gpu/ism_inst.cuh:18-32). */
ppm_status ppm_open(const ppm_open_opts* opts, ppm_ctx** out_ctx);

/* Release all device memory and the serving thread.  Replaces
 * free_gpu_memory() (gpu/lensing.cuh:56-61, gpu/nsx.cuh:130-134,
 * gpu/ism_inst.cuh:34-38). */
ppm_status ppm_close(ppm_ctx* ctx);

/* Table facts needed by callers (response sizes, lensing domain). */
ppm_status ppm_response_info(const ppm_ctx* ctx, int response, int* n_ch, int* n_energy,
                             double* energies_keV /* [n_energy] or NULL */);
ppm_status ppm_lensing_domain(const ppm_ctx* ctx, double* u_min, double* u_max,
                              double* cos_psi_min);

/* Evaluate n profiles on one grid.  flux is [n][n_phase][n_energy] (FP64);
 * counts, if non-NULL and grid->response >= 0, is [n][n_ch][n_phase].
 * Host buffers; blocking.  Replaces calc_one_star
 * (gpu/os1_gpu.cu:303-349, gpu/devarshi24st_gpu.cu:431-511, and
 * cpu/t_interp.cpp:19-224 with ppm_star.nsx_interp as its interp_type).
 * In one call, stars may use different atmospheres and interpolants. */
ppm_status ppm_eval(ppm_ctx* ctx, const ppm_star* stars, int n, const ppm_grid* grid,
                    double* flux, double* counts, ppm_stats* stats);

/* Same as ppm_eval with device-resident outputs (HBM pointers) and no host
 * copies; returns after the work is complete.  This is the form the
 * benchmark times (inputs and outputs stay in HBM). */
ppm_status ppm_eval_device(ppm_ctx* ctx, const ppm_star* stars, int n, const ppm_grid* grid,
                           double* d_flux, double* d_counts, ppm_stats* stats);

/* Same evaluation as ppm_eval_device for ONE star on the plain path, with
 * no device-memory allocation: no counterpart in the reference; for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Every device buffer beyond the caller's
 * outputs lives in one file-scope scratch array of 128 MiB per GPU, held
 * from library load until the process ends; a profile that needs more
 * (about n_ring_tasks x n_phase x n_energy x 4 bytes of per-ring partials:
 * N_side 1024, 64 bins, 360 energies fits; N_side 2048 x 128 bins does not)
 * returns PPM_ERR_NOMEM before anything is launched.  Calls on one GPU share
 * that array, from any context: calls that overlap on one GPU race (a ctx is
 * used by one thread at a time anyway).  Outputs as ppm_eval_device;
 * grid->method BATCHED is rejected. */
ppm_status ppm_eval_device_noalloc(ppm_ctx* ctx, const ppm_star* star, const ppm_grid* grid, double* d_flux,
                                   double* d_counts, ppm_stats* stats);

/* Number of active patches of one star on a grid (FP64 membership on the
 * device).  Replaces the N_active print of GenerateHealpixGrid
 * (cpu/grid.hpp:247). */
ppm_status ppm_count_active(ppm_ctx* ctx, const ppm_star* star, const ppm_grid* grid,
                            int64_t* n_active);

/* Response fold on the device: counts[b][c][p] = scale[b] * sum_e
 * rsp[c][e] * flux[b][p][e] using split-BF16 MFMA with FP32 accumulation.
 * d_flux / d_counts are device pointers.  Replaces MatMul::operator()
 * (gpu/matmul.cuh:11-35). */
ppm_status ppm_fold_device(ppm_ctx* ctx, int response, int n_batch, int n_phase,
                           const double* d_flux, const double* scale, double* d_counts);

/* Serving: continuous batching of single-profile requests.  Requests are
 * admitted against the HBM budget, packed into persistent batched launches
 * while others run, and completed out of order.  The grid's energy array is
 * copied at submit.  ppm_poll returns PPM_PENDING until the result is ready,
 * then copies flux [n_phase][n_energy] and counts [n_ch][n_phase] (if asked
 * for) and releases the ticket.  latency_ms (optional) receives submit->done. */
ppm_status ppm_submit(ppm_ctx* ctx, const ppm_star* star, const ppm_grid* grid, uint64_t* ticket);
ppm_status ppm_poll(ppm_ctx* ctx, uint64_t ticket, double* flux, double* counts, double* latency_ms);
ppm_status ppm_cancel(ppm_ctx* ctx, uint64_t ticket);
/* Block until every admitted request is finished or cancelled. */
ppm_status ppm_drain(ppm_ctx* ctx);

/* State of the serving loop.  No counterpart in the reference (it has no
 * serving layer).  device_bytes counts the device buffers the serving lanes
 * hold (grow-only, reused by later launches); the admission budget bounds
 * them (fixed_bytes, memory held outside the budget, is 0: the batched
 * kernel keeps no scratch beyond its budgeted accumulators). */
typedef struct ppm_serve_stats {
  int64_t queued, running;   /* requests waiting, and in a launch              */
  int64_t launches[2];       /* launch sequences of the express and bulk lanes */
  int64_t overlapped;        /* express launches started while a bulk launch ran */
  size_t charged_bytes;      /* admission charges of queued and running requests */
  size_t budget_bytes;       /* the context's hbm_budget_bytes                  */
  size_t device_bytes;       /* device memory held by the serving lanes          */
  size_t fixed_bytes;        /* of which outside the budget (0)                   */
} ppm_serve_stats;
ppm_status ppm_serve_stats_get(ppm_ctx* ctx, ppm_serve_stats* out);

/* Instrument response built on the device from an instrument's files,
 * replacing c_ism_inst.py:30-101 (numpy/scipy in the reference): trim to
 * channels [ch_lo, ch_hi) and energies [e_lo, e_hi), response = area x RMF,
 * E_obs log-spaced between the first lower and the last upper bin edge (ends
 * exact), C[j][i] = integral over fine bin j of the not-a-knot cubic spline
 * through the unit vector e_i on E_obs, ISM transmission (not-a-knot spline
 * at the bin centres) to the 5th power, rsp = (response x ISM^5) @ C with
 * |x| < 1e-300 set to 0.  Host arrays in, host arrays out; the result is
 * what tables/ism_inst/<setting>/{num,E_obs,rsp}.txt hold. */
typedef struct ppm_response_spec {
  const double* rmf;  /* [n_e_all][n_ch_all]: the RMF file's last column, energy-major */
  const double* arf;  /* [n_e_all][3]: E_lo, E_hi, area                              */
  const double* ism;  /* [n_ism][2]: energy, transmission (ISM file columns 0 and 2)  */
  int n_e_all, n_ch_all, n_ism;
  int e_lo, e_hi, ch_lo, ch_hi;  /* trim ranges (c_ism_inst.py:9-12: 0,1800 and 30,300) */
  int n_eobs;                    /* observed energies (std 128, high 256, ultra 512)    */
  int device;                    /* HIP device ordinal                                 */
} ppm_response_spec;
ppm_status ppm_response_build(const ppm_response_spec* spec, double* e_obs /* [n_eobs] */,
                              double* rsp /* [ch_hi - ch_lo][n_eobs] */);

/* Lensing tables built on the device, replacing c_lensing.py:47-155 (scipy
 * quad, numpy.gradient and PchipInterpolator in the reference): for u on
 * [u_min, u_max] (n_u nodes) the alpha march of step pi / (8 n_cos) until psi
 * passes acos(-0.9), psi(alpha) and c dt / R by adaptive Gauss-Kronrod 7/15
 * quadrature (relative tolerance 1e-12), the lensing factor
 * d cos(alpha) / d cos(psi) by the non-uniform second-order difference, and
 * cos(alpha), the lensing factor (on cos psi in [-0.9, 1]) and c dt / R (on
 * cos alpha in [-0.2, 1]) on n_cos uniform nodes: the three matrices of
 * tables/lensing/<setting>/ (std: 256 u in 0.3-0.5, 512 nodes).  Outputs are
 * host arrays [n_u][n_cos]; u_max < 0.568 (single images). */
typedef struct ppm_lensing_spec {
  double u_min, u_max;
  int n_u, n_cos;
  int device; /* HIP device ordinal */
  int reserved;
} ppm_lensing_spec;
ppm_status ppm_lensing_build(const ppm_lensing_spec* spec, double* cos_alpha, double* lensing_factor,
                             double* cdt_over_R);

/* Fit of one star's folded counts m[c][p] to data d[c][p]. */
typedef struct ppm_fit {
  double chi2;        /* sum (d - m)^2 / m with model and data scaled so that the model sums
                         to 1e6: cal_chi2 of p_devarshi24st.py:43-49                      */
  double loglike;     /* Poisson log-likelihood: sum d ln m - m - ln Gamma(d + 1)         */
  double model_total; /* sum of m                                                         */
} ppm_fit;

/* Evaluate n profiles and fit their counts to data, the statistics summed in
 * the response fold's epilogue on the device (FP64, fixed order).  data is
 * [n_ch][n_phase] for every star (data_per_star = 0) or [n][n_ch][n_phase]
 * (1); fit receives n entries; counts [n][n_ch][n_phase] if non-NULL.  The
 * grid needs a response.  The statistics follow the formulas literally (as
 * cal_chi2 does), except that a bin with m = 0 and d = 0 adds 0 to both (the
 * limit m -> 0; cal_chi2 would give 0/0): a bin with m < 0, which the cubic
 * energy interpolant can produce in a steep Wien tail, adds a negative chi2
 * term and makes the log-likelihood NaN; m = 0 < d gives infinities.  No
 * counterpart in the reference,
 * which computes cal_chi2 in its plotting script (p_devarshi24st.py:43-49)
 * from the written counts files. */
ppm_status ppm_eval_fit(ppm_ctx* ctx, const ppm_star* stars, int n, const ppm_grid* grid, const double* data,
                        int data_per_star, ppm_fit* fit, double* counts, ppm_stats* stats);

/* Launch shape of the batched path.  A 0 field means the default: 10
 * work-queue items per resident workgroup (each spot's rings are cut into
 * chunks of similar cost so that the queue holds about that many items per
 * workgroup), the occupancy limit of workgroups per CU, and the fewest of 2,
 * 4 or 6 energies per lane (128, 256 or 384 energies per work item) that
 * hold the grid in one energy chunk with each row set of the NSX evaluation
 * within 57 slice rows (else the most that stay within them). */
typedef struct ppm_launch_shape {
  int items_per_slot;    /* work-queue items per resident workgroup (0..4096)     */
  int wg_per_cu;         /* resident workgroups per CU (0..64, capped by occupancy) */
  int energies_per_lane; /* 0 (default), 2, 4 or 6                                  */
  int reserved;          /* 0                                                       */
} ppm_launch_shape;

typedef struct ppm_tune_result {
  ppm_launch_shape best; /* the shape kept for this grid shape            */
  double best_ms;        /* its integral time (min over repeats, hipEvent) */
  double default_ms;     /* the default shape's integral time             */
  int n_tried;           /* shapes timed                                  */
  int reserved;
} ppm_tune_result;

/* Autotuner of the batched launch.  No counterpart in the reference, whose
 * launch geometry is fixed (one block per ring, gpu/devarshi24st_gpu.cu:463,
 * gpu/os1_gpu.cu:316).  Times the batched integral of `stars` on `grid` over
 * a short coordinate search of the launch shape (items per workgroup halved
 * and doubled, one more step the better way, one workgroup per CU fewer,
 * and, beyond 128 energies, the other energies-per-lane choices), `repeats`
 * runs each (0 -> 2), and keeps the fastest for this grid
 * shape (grid fields, energy count, atmosphere of stars[0], bit length of
 * n): later batched evaluations and served launches of that shape use it.
 * Blocking; the fold is not run. */
ppm_status ppm_tune(ppm_ctx* ctx, const ppm_star* stars, int n, const ppm_grid* grid, int repeats,
                    ppm_tune_result* out);

/* Store a shape for a grid shape (for example one saved from an earlier
 * ppm_tune), or clear it (shape NULL or all zero); read the stored shape
 * (all zero: none stored, the default applies). */
ppm_status ppm_set_launch_shape(ppm_ctx* ctx, const ppm_grid* grid, int n_stars, int atmosphere,
                                const ppm_launch_shape* shape);
ppm_status ppm_get_launch_shape(ppm_ctx* ctx, const ppm_grid* grid, int n_stars, int atmosphere,
                                ppm_launch_shape* shape);

/* Phase timers of the batched integral.  No counterpart in the reference
 * (it has no instrumentation); for profiling.  In a library built with them
 * (python -m gpu_ppm_amd.build --phase-timers writes libppm_phase.so, the
 * same kernels with shader-clock reads at phase boundaries), copies the
 * wavefront-cycle totals accumulated since the last reset into
 * cycles[0..n-1]: taking work items, ring set-up, events, sort, visibility
 * pass, segment split, sweep (which includes the next two), FP64 patch
 * set-up, window ends, segment offsets, whole kernel, and the number of
 * wavefronts; reset != 0 clears them.  Other builds return
 * PPM_ERR_UNSUPPORTED. */
#define PPM_PHASE_COUNT 12
ppm_status ppm_phase_timers(ppm_ctx* ctx, uint64_t* cycles, int n, int reset);

#ifdef __cplusplus
}
#endif
#endif /* PPM_H */
