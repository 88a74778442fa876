/*
 * FP64 CPU oracle of the pulse-profile path - TEST INFRASTRUCTURE ONLY.
 *
 * A parameterised C restatement of the reference's FP64 CPU programs
 * (cpu/os1_cpu.cpp, cpu/devarshi24st_cpu.cpp and the headers they use).  It
 * is the checker the GPU library is compared against; only tests/, the
 * smoke check and the benchmark's cpu_baseline leg may load it.  Nothing in
 * gpu_ppm_amd/ links or calls it.
 */
#ifndef PPM_ORACLE_H
#define PPM_ORACLE_H

#include <stdint.h>

#include "../include/ppm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_lensing {
  int n_u, n_cp, n_ca;
  double u_min, u_max, cp_min, cp_max, ca_min, ca_max;
  double* ca_tab; /* [n_u][n_cp] cos(alpha)(u, cos psi)   */
  double* lf_tab; /* [n_u][n_cp] lensing factor           */
  double* dt_tab; /* [n_u][n_ca] c dt / R (u, cos alpha)   */
} or_lensing;

typedef struct or_nsx {
  ppm_nsx_config cfg;
  double* mu;   /* [n_mu], decreasing                       */
  double* data; /* [logT][logg][logEkT][mu] = 10^logInuT3   */
} or_nsx;

typedef struct or_response {
  int n_ch, n_e;
  double* e_obs; /* [n_e]        */
  double* rsp;   /* [n_ch][n_e]  */
} or_response;

typedef struct or_info {
  int64_t n_active;    /* active patches over spots                         */
  int64_t n_eval_ref;  /* (patch, phase point, energy) triples on kept rings */
  int64_t n_rings;     /* rings with active patches (all spots)             */
  int64_t n_rings_kept;/* of those, not dropped by the ring-skip rule       */
  int64_t n_clamped;   /* table look-ups outside the table axes (clamped)   */
} or_info;

/* Readers of the reference formats.  NULL on failure with a message in err. */
or_lensing* or_lensing_load(const char* dir, char* err, int errlen);
or_nsx* or_nsx_load(const char* path, const ppm_nsx_config* cfg, char* err, int errlen);
or_response* or_response_load(const char* dir, char* err, int errlen);
void or_lensing_free(or_lensing* t);
void or_nsx_free(or_nsx* t);
void or_response_free(or_response* t);

/* nsx_H_v200804 dimensions (cpu/nsx.hpp:25-37). */
ppm_nsx_config or_nsx_default_config(void);

/* Flux [n_phase][n_energy] of one star.  energies may be NULL when the grid
 * names a response (its E_obs is used).  n_threads <= 1: the reference's
 * serial order.  Returns 0, or a nonzero code with err set. */
int or_profile(const or_lensing* lt, const or_nsx* nsx, const or_response* rsp,
               const ppm_star* star, const ppm_grid* grid, const double* energies,
               double* flux, or_info* info, int n_threads, char* err, int errlen);

/* Active patches of one star (FP64 membership). */
int64_t or_count_active(const ppm_star* star, const ppm_grid* grid);

/* counts[ch][b] = scale * sum_e flux[b][e] rsp[ch][e]
 * (cpu/devarshi24st_cpu.cpp:249-258, scale = exposure * phase bin width). */
void or_fold(const or_response* rsp, const double* flux, int n_phase, double scale,
             double* counts);

/* Building blocks exposed for unit tests. */
void or_healpix_ring(int n_side, int i, double* z, int* n_pix, double* s_shift);
double or_healpix_phi(int n_side, int i, int j);
/* NSX interpolant mode = PPM_NSX_* at one point. */
double or_nsx_interp(const or_nsx* nsx, int mode, double logT, double logg, double logEkT, double mu);
void or_lensing_lookup(const or_lensing* lt, double u, double cos_psi, double* cos_alpha,
                       double* lf, double* cdt_over_R_at_ca);
double or_gl_node(int n, int i, double a, double b);
double or_gl_weight(int n, int i, double dx);
/* Bracket search with the reference's VecHunt result semantics. */
int or_vec_hunt(const double* xx, int n, double x, double* a);

#ifdef __cplusplus
}
#endif
#endif
