/*
 * FP64 CPU oracle of the pulse-profile path - TEST INFRASTRUCTURE ONLY.
 *
 * Restates, in C, the reference's serial FP64 path with the reference CPU's
 * semantics wherever its CPU and GPU programs differ:
 *   - table look-ups clamp to the axis ends (cpu/hunt.hpp:98-116);
 *   - bracket searches follow VecHunt (cpu/hunt.hpp:8-76), extrapolating;
 *   - a ring is dropped when its first or last fine node is invisible
 *     (cpu/os1_cpu.cpp:151, cpu/devarshi24st_cpu.cpp:195);
 *   - (patch, phase) pairs whose interpolated flux/I is exactly 0 are skipped
 *     (cpu/os1_cpu.cpp:176, cpu/devarshi24st_cpu.cpp:226);
 *   - the emitted intensity is divided by E_obs (cpu/os1_cpu.cpp:180);
 *   - patch membership and visibility are FP64.
 * Each function cites the reference lines it follows.
 *
 * Pinning (no function is left unpinned): or_profile and or_fold against the
 * outputs of the reference's own os1_cpu, devarshi24st_cpu and t_interp
 * built from their sources (tests/test_oracle_golden.py, <= 5e-10, with a
 * planted error that the check catches), which also pins the lensing
 * look-ups and the active counts they go through; or_vec_hunt,
 * or_gl_node/weight, or_healpix_ring/phi and or_nsx_interp (all five
 * interpolants, clamps included) in addition by brute force, closed forms
 * and exactness properties (tests/test_oracle_math.py); the NSX path against
 * the blackbody closed form (tests/test_blackbody_pin.py).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------ constants (cpu/unit.hpp) */
static const double RS_SUN_KM = 2.9532501;                 /* cpu/unit.hpp:6  */
static const double SQRT_KM3_S2_GMSUN = 2.745011592867327e-6; /* cpu/unit.hpp:9  */
static const double PI_ = 3.14159265358979323846;          /* cpu/unit.hpp:11 */
static const double C_KM_S = 299792458. / 1e3;             /* cpu/unit.hpp:15-16 */
static const double H_KEV_S = 4.135668e-18;                /* cpu/unit.hpp:14 */
static const double C_CM_S = 299792458. * 1e2;             /* cpu/unit.hpp:17 */
static const double PLANCK_ERG_S = 6.62607015E-27;         /* cpu/unit.hpp:25 */
static const double KB_KEV_K = 8.617333262E-8;             /* cpu/devarshi24st_cpu.cpp:59 */
static const double LOG_GMSUN_CGS_ARG = 1.32712440041279419E16; /* cpu/devarshi24st_cpu.cpp:68 */

static double ibb_constant(void) { /* cpu/unit.hpp:19 */
  return 2. / (H_KEV_S * H_KEV_S * H_KEV_S * C_CM_S * C_CM_S);
}

/* ------------------------------------------------------------ search */

typedef struct { int i; double a; } bracket;

/* Uniform axis with end clamping (cpu/hunt.hpp:155-178); x is clamped in
 * place like the reference's by-reference argument. */
static bracket uniform_hunt(double x_min, double x_max, int n, double* x, int64_t* n_clamped) {
  bracket r;
  if (*x <= x_min) {
    if (*x < x_min && n_clamped) ++*n_clamped;
    r.i = 0;
    r.a = 0;
    *x = x_min;
    return r;
  }
  if (*x >= x_max) {
    if (*x > x_max && n_clamped) ++*n_clamped;
    r.i = n - 2;
    r.a = 1;
    *x = x_max;
    return r;
  }
  double t = (*x - x_min) / (x_max - x_min) * (n - 1);
  r.i = (int)t;
  r.a = t - r.i;
  return r;
}

static double uniform_node(double x_min, double x_max, int n, int i) { /* cpu/hunt.hpp:180-187 */
  return x_min + (x_max - x_min) * i / (n - 1);
}

/* Result of VecHunt (cpu/hunt.hpp:71-123) on a strictly monotone array:
 * the lower index of the bracket holding x (its hunt from any start ends
 * there), clamped to [0, n-2], with the linear fraction, which extrapolates
 * outside the array. */
int or_vec_hunt(const double* xx, int n, double x, double* a) {
  int ascnd = xx[n - 1] >= xx[0];
  int jl = 0, ju = n - 1;
  while (ju - jl > 1) {
    int jm = (ju + jl) >> 1;
    if ((x >= xx[jm]) == ascnd)
      jl = jm;
    else
      ju = jm;
  }
  int i = jl < 0 ? 0 : (jl > n - 2 ? n - 2 : jl);
  *a = (x - xx[i]) / (xx[i + 1] - xx[i]);
  return i;
}

/* ------------------------------------------------------- Gauss-Legendre */

static const double GL_X[] = { /* cpu/gl.hpp:5-27 */
    0.0,
    -0.5773502691896257645091488, 0.5773502691896257645091488,
    -0.7745966692414833770358531, 0.0, 0.7745966692414833770358531,
    -0.8611363115940525752239465, -0.3399810435848562648026658,
    0.3399810435848562648026658, 0.8611363115940525752239465,
    -0.9061798459386639927976269, -0.5384693101056830910363144, 0.0,
    0.5384693101056830910363144, 0.9061798459386639927976269,
    -0.9324695142031520278123016, -0.6612093864662645136613996,
    -0.2386191860831969086305017, 0.2386191860831969086305017,
    0.6612093864662645136613996, 0.9324695142031520278123016};
static const double GL_W[] = { /* cpu/gl.hpp:28-50 */
    2.0,
    1.0, 1.0,
    0.5555555555555555555555556, 0.8888888888888888888888889, 0.5555555555555555555555556,
    0.3478548451374538573730639, 0.6521451548625461426269361,
    0.6521451548625461426269361, 0.3478548451374538573730639,
    0.2369268850561890875142640, 0.4786286704993664680412915, 0.5688888888888888888888889,
    0.4786286704993664680412915, 0.2369268850561890875142640,
    0.1713244923791703450402961, 0.3607615730481386075698335, 0.4679139345726910473898703,
    0.4679139345726910473898703, 0.3607615730481386075698335, 0.1713244923791703450402961};

double or_gl_node(int n, int i, double a, double b) { /* cpu/gl.hpp:52-55 */
  return 0.5 * (a + b + GL_X[n * (n - 1) / 2 + i] * (b - a));
}
double or_gl_weight(int n, int i, double dx) { /* cpu/gl.hpp:57-60 */
  return 0.5 * dx * GL_W[n * (n - 1) / 2 + i];
}

/* -------------------------------------------------------------- HEALPix */

/* Ring i (1..4N-1): z = cos(theta), pixel count and the half-pixel shift s
 * of phi_j = (pi / (2 i')) (j - s/2) (cpu/grid.hpp:106-131, :140). */
void or_healpix_ring(int n_side, int i, double* z, int* n_pix, double* s_shift) {
  int north = i <= 2 * n_side;
  int ii = north ? i : 4 * n_side - i;
  double zz, s;
  if (ii < n_side) {
    zz = 1. - 1. * ii * ii / (3. * n_side * n_side);
    s = 1.;
  } else {
    zz = 4. / 3. - 2. * ii / (3. * n_side);
    s = (ii - n_side + 1) % 2;
  }
  *z = north ? zz : -zz;
  int m = i < 4 * n_side - i ? i : 4 * n_side - i;
  if (m > n_side) m = n_side;
  *n_pix = 4 * m;
  *s_shift = s;
}

double or_healpix_phi(int n_side, int i, int j) { /* cpu/grid.hpp:107-116 */
  int ii = i <= 2 * n_side ? i : 4 * n_side - i;
  if (ii < n_side) return PI_ / (2. * ii) * (j - 1. / 2.);
  double s = (ii - n_side + 1) % 2;
  return PI_ / (2. * n_side) * (j - s / 2.);
}

/* Patch membership in FP64. */
static int in_spot(const ppm_spot* sp, double z, double phi) {
  if (sp->membership == PPM_MEMBER_CAP) { /* cpu/grid.hpp:144-148 */
    double s2 = sqrt(1. - z * z);
    int in1 = sin(sp->theta1) * s2 * cos(phi - sp->phi1) + cos(sp->theta1) * z > cos(sp->angrad1);
    if (!in1) return 0;
    if (sp->angrad2 <= 0) return 1;
    return !(sin(sp->theta2) * s2 * cos(phi - sp->phi2) + cos(sp->theta2) * z > cos(sp->angrad2));
  }
  /* DiskDifference on theta = acos z (cpu/grid.hpp:158-190, :237) */
  double th = acos(z);
  double c2 = cos(th), s2 = sin(th);
  int in1 = sin(sp->theta1) * s2 * cos(phi - sp->phi1) + cos(sp->theta1) * c2 > cos(sp->angrad1);
  if (!in1) return 0;
  int in2 = sin(sp->theta2) * s2 * cos(phi - sp->phi2) + cos(sp->theta2) * c2 > cos(sp->angrad2);
  return !in2;
}

/* ---------------------------------------------------------------- tables */

static char* slurp(const char* path, size_t* len) {
  FILE* f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  char* buf = (char*)malloc((size_t)n + 1);
  if (!buf) {
    fclose(f);
    return NULL;
  }
  size_t got = fread(buf, 1, (size_t)n, f);
  fclose(f);
  buf[got] = '\0';
  if (len) *len = got;
  return buf;
}

static int read_numbers(const char* path, double* out, long count, char* err, int errlen) {
  char* buf = slurp(path, NULL);
  if (!buf) {
    snprintf(err, errlen, "could not open %s", path);
    return 2;
  }
  char* p = buf;
  for (long k = 0; k < count; ++k) {
    char* e;
    out[k] = strtod(p, &e);
    if (e == p) {
      snprintf(err, errlen, "short or malformed file %s (value %ld)", path, k);
      free(buf);
      return 3;
    }
    p = e;
  }
  free(buf);
  return 0;
}

static int read_axis(const char* dir, const char* name, double* lo, double* hi, int* n,
                     char* err, int errlen) { /* cpu/hunt.hpp:141-153 */
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  double v[3];
  int rc = read_numbers(path, v, 3, err, errlen);
  if (rc) return rc;
  *lo = v[0];
  *hi = v[1];
  *n = (int)v[2];
  if (*n < 2) {
    snprintf(err, errlen, "%s: n must be at least 2", path);
    return 3;
  }
  return 0;
}

or_lensing* or_lensing_load(const char* dir, char* err, int errlen) { /* cpu/lensing.hpp:33-41 */
  or_lensing* t = (or_lensing*)calloc(1, sizeof *t);
  char path[4096];
  if (read_axis(dir, "cos_psi.txt", &t->cp_min, &t->cp_max, &t->n_cp, err, errlen) ||
      read_axis(dir, "u.txt", &t->u_min, &t->u_max, &t->n_u, err, errlen) ||
      read_axis(dir, "cos_alpha.txt", &t->ca_min, &t->ca_max, &t->n_ca, err, errlen))
    goto fail;
  t->ca_tab = (double*)malloc(sizeof(double) * t->n_u * t->n_cp);
  t->lf_tab = (double*)malloc(sizeof(double) * t->n_u * t->n_cp);
  t->dt_tab = (double*)malloc(sizeof(double) * t->n_u * t->n_ca);
  snprintf(path, sizeof path, "%s/cos_alpha_of_u_cos_psi.txt", dir);
  if (read_numbers(path, t->ca_tab, (long)t->n_u * t->n_cp, err, errlen)) goto fail;
  snprintf(path, sizeof path, "%s/lf_of_u_cos_psi.txt", dir);
  if (read_numbers(path, t->lf_tab, (long)t->n_u * t->n_cp, err, errlen)) goto fail;
  snprintf(path, sizeof path, "%s/cdt_over_R_of_u_cos_alpha.txt", dir);
  if (read_numbers(path, t->dt_tab, (long)t->n_u * t->n_ca, err, errlen)) goto fail;
  return t;
fail:
  or_lensing_free(t);
  return NULL;
}

void or_lensing_free(or_lensing* t) {
  if (!t) return;
  free(t->ca_tab);
  free(t->lf_tab);
  free(t->dt_tab);
  free(t);
}

ppm_nsx_config or_nsx_default_config(void) {
  ppm_nsx_config c = {35, 14, 166, 67, 5.10, 6.80, 13.7, 15.0, -1.30, 2.00};
  return c;
}

static int is_close(double a, double b) { /* cpu/nsx.hpp:150-156 */
  const double atol = 10 * 2.220446049250313e-16, rtol = 1e-8;
  double big = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
  return fabs(a - b) <= atol + rtol * big;
}

or_nsx* or_nsx_load(const char* path, const ppm_nsx_config* cfg, char* err, int errlen) {
  /* cpu/nsx.hpp:83-148 */
  or_nsx* t = (or_nsx*)calloc(1, sizeof *t);
  t->cfg = cfg ? *cfg : or_nsx_default_config();
  const ppm_nsx_config* c = &t->cfg;
  long n = (long)c->n_logT * c->n_logg * c->n_logEkT * c->n_mu;
  t->mu = (double*)calloc(c->n_mu, sizeof(double));
  t->data = (double*)malloc(sizeof(double) * n);
  char* buf = slurp(path, NULL);
  if (!buf) {
    snprintf(err, errlen, "could not open NSX file %s", path);
    goto fail;
  }
  char* p = buf;
  long idx = 0;
  for (int iT = 0; iT < c->n_logT; ++iT)
    for (int ig = 0; ig < c->n_logg; ++ig)
      for (int iE = 0; iE < c->n_logEkT; ++iE)
        for (int im = 0; im < c->n_mu; ++im, ++idx) {
          double v[5];
          for (int k = 0; k < 5; ++k) {
            char* e;
            v[k] = strtod(p, &e);
            if (e == p) {
              snprintf(err, errlen, "error reading NSX data at %d %d %d %d", iT, ig, iE, im);
              free(buf);
              goto fail;
            }
            p = e;
          }
          double logEkT = v[0], mu = v[1], logInuT3 = v[2], logT = v[3], logg = v[4];
          if (t->mu[im] == 0) {
            t->mu[im] = mu;
          } else if (!is_close(t->mu[im], mu)) {
            snprintf(err, errlen, "inconsistent mu at index %d: %g vs %g", im, t->mu[im], mu);
            free(buf);
            goto fail;
          }
          if (!is_close(uniform_node(c->logEkT_min, c->logEkT_max, c->n_logEkT, iE), logEkT) ||
              !is_close(uniform_node(c->logT_min, c->logT_max, c->n_logT, iT), logT) ||
              !is_close(uniform_node(c->logg_min, c->logg_max, c->n_logg, ig), logg)) {
            snprintf(err, errlen, "NSX row %d %d %d %d is off the uniform axes", iT, ig, iE, im);
            free(buf);
            goto fail;
          }
          t->data[idx] = pow(10, logInuT3);
        }
  free(buf);
  for (int i = 0; i < c->n_mu - 1; ++i)
    if (t->mu[i] < t->mu[i + 1]) {
      snprintf(err, errlen, "mu is not decreasing at index %d", i);
      goto fail;
    }
  return t;
fail:
  or_nsx_free(t);
  return NULL;
}

void or_nsx_free(or_nsx* t) {
  if (!t) return;
  free(t->mu);
  free(t->data);
  free(t);
}

or_response* or_response_load(const char* dir, char* err, int errlen) { /* cpu/ism_inst.hpp:59-70 */
  or_response* r = (or_response*)calloc(1, sizeof *r);
  char path[4096];
  double nums[2];
  snprintf(path, sizeof path, "%s/num.txt", dir);
  if (read_numbers(path, nums, 2, err, errlen)) goto fail;
  r->n_ch = (int)nums[0];
  r->n_e = (int)nums[1];
  if (r->n_ch < 1 || r->n_e < 1) {
    snprintf(err, errlen, "%s: bad sizes", path);
    goto fail;
  }
  r->e_obs = (double*)malloc(sizeof(double) * r->n_e);
  r->rsp = (double*)malloc(sizeof(double) * r->n_ch * r->n_e);
  snprintf(path, sizeof path, "%s/E_obs.txt", dir);
  if (read_numbers(path, r->e_obs, r->n_e, err, errlen)) goto fail;
  snprintf(path, sizeof path, "%s/rsp.txt", dir);
  if (read_numbers(path, r->rsp, (long)r->n_ch * r->n_e, err, errlen)) goto fail;
  return r;
fail:
  or_response_free(r);
  return NULL;
}

void or_response_free(or_response* r) {
  if (!r) return;
  free(r->e_obs);
  free(r->rsp);
  free(r);
}

/* --------------------------------------------------------- look-ups */

typedef struct {
  const or_lensing* lt;
  int64_t* n_clamped;
} lens_ctx;

static void lens_ca_lf(const lens_ctx* L, double u, double cos_psi, double* ca, double* lf) {
  /* cpu/lensing.hpp:43-60 */
  const or_lensing* t = L->lt;
  bracket bu = uniform_hunt(t->u_min, t->u_max, t->n_u, &u, L->n_clamped);
  bracket bc = uniform_hunt(t->cp_min, t->cp_max, t->n_cp, &cos_psi, L->n_clamped);
  double b_u = 1. - bu.a, b_c = 1. - bc.a;
  const double* A = t->ca_tab;
  const double* F = t->lf_tab;
  int n = t->n_cp, r0 = bu.i * n, r1 = (bu.i + 1) * n;
  *ca = b_u * b_c * A[r0 + bc.i] + bu.a * b_c * A[r1 + bc.i] + b_u * bc.a * A[r0 + bc.i + 1] +
        bu.a * bc.a * A[r1 + bc.i + 1];
  *lf = b_u * b_c * F[r0 + bc.i] + bu.a * b_c * F[r1 + bc.i] + b_u * bc.a * F[r0 + bc.i + 1] +
        bu.a * bc.a * F[r1 + bc.i + 1];
}

static double lens_cdt(const lens_ctx* L, double u, double ca) { /* cpu/lensing.hpp:62-72 */
  const or_lensing* t = L->lt;
  bracket bu = uniform_hunt(t->u_min, t->u_max, t->n_u, &u, L->n_clamped);
  bracket bc = uniform_hunt(t->ca_min, t->ca_max, t->n_ca, &ca, L->n_clamped);
  double b_u = 1. - bu.a, b_c = 1. - bc.a;
  const double* D = t->dt_tab;
  int n = t->n_ca, r0 = bu.i * n, r1 = (bu.i + 1) * n;
  return b_u * b_c * D[r0 + bc.i] + bu.a * b_c * D[r1 + bc.i] + b_u * bc.a * D[r0 + bc.i + 1] +
         bu.a * bc.a * D[r1 + bc.i + 1];
}

void or_lensing_lookup(const or_lensing* lt, double u, double cos_psi, double* cos_alpha,
                       double* lf, double* cdt_over_R_at_ca) {
  lens_ctx L = {lt, NULL};
  lens_ca_lf(&L, u, cos_psi, cos_alpha, lf);
  if (cdt_over_R_at_ca) *cdt_over_R_at_ca = lens_cdt(&L, u, *cos_alpha);
}

static void lagrange4(double x, double x0, double x1, double x2, double x3, double* w) {
  /* cpu/nsx.hpp:159-167 */
  w[0] = (x - x1) * (x - x2) * (x - x3) / ((x0 - x1) * (x0 - x2) * (x0 - x3));
  w[1] = (x - x0) * (x - x2) * (x - x3) / ((x1 - x0) * (x1 - x2) * (x1 - x3));
  w[2] = (x - x0) * (x - x1) * (x - x3) / ((x2 - x0) * (x2 - x1) * (x2 - x3));
  w[3] = (x - x0) * (x - x1) * (x - x2) / ((x3 - x0) * (x3 - x1) * (x3 - x2));
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Weights of one table axis: the 4-point Lagrange stencil starting one node
 * below the bracket (clamped to [0, n-4]) at the clamped coordinate, or the
 * bracket's two linear weights.  Returns the first node. */
static int axis_weights(double x_min, double x_max, int n, double x, int cubic, double* w, int* nw,
                        int64_t* n_clamped) {
  bracket b = uniform_hunt(x_min, x_max, n, &x, n_clamped);
  if (!cubic) { /* (j a + (1 - j)(1 - a)), cpu/nsx.hpp:284-289 */
    *nw = 2;
    w[0] = 1 - b.a;
    w[1] = b.a;
    return b.i;
  }
  int i = clampi(b.i - 1, 0, n - 4);
  lagrange4(x, uniform_node(x_min, x_max, n, i), uniform_node(x_min, x_max, n, i + 1),
            uniform_node(x_min, x_max, n, i + 2), uniform_node(x_min, x_max, n, i + 3), w);
  *nw = 4;
  return i;
}

/* The NSX interpolants (mode = PPM_NSX_*), each a weighted sum over a
 * tensor stencil of the grid, accumulated in the reference's loop order
 * (logT, log g, logEkT, mu) with weight wT wg wE wmu:
 *   2L2C_LE  NSX::Interp_IT3_2l2c_le (cpu/nsx.hpp:248-296): linear logT and
 *            log g, cubic logEkT, mu cubic or linear in the last brackets
 *            (bracket index >= n_mu - 5);
 *   4C, 4C_SG  NSX::Interp_IT3_4c (cpu/nsx.hpp:170-203): cubic in all four,
 *            use_sg sets a negative result to 0;
 *   4C_LE    NSX::Interp_IT3_4c_le (cpu/nsx.hpp:205-246): cubic logT, log g,
 *            logEkT, mu as 2L2C_LE;
 *   4L       NSX::Interp_IT3_4l (cpu/nsx.hpp:298-321): linear in all four.
 * Axis coordinates are clamped to the axis ends by the hunt
 * (cpu/hunt.hpp:155-178); the mu bracket extrapolates (VecHunt). */
static double nsx_interp(const or_nsx* t, int mode, double logT, double logg, double logEkT, double mu,
                         int64_t* n_clamped) {
  const ppm_nsx_config* c = &t->cfg;
  const int cubic_tg = mode == PPM_NSX_4C || mode == PPM_NSX_4C_SG || mode == PPM_NSX_4C_LE;
  const int cubic_e = mode != PPM_NSX_4L;
  double wT[4], wg[4], wE[4], wm[4];
  int nT, ng, nE, nm;
  int iT = axis_weights(c->logT_min, c->logT_max, c->n_logT, logT, cubic_tg, wT, &nT, n_clamped);
  int ig = axis_weights(c->logg_min, c->logg_max, c->n_logg, logg, cubic_tg, wg, &ng, n_clamped);
  int iE = axis_weights(c->logEkT_min, c->logEkT_max, c->n_logEkT, logEkT, cubic_e, wE, &nE, n_clamped);
  double a_mu;
  int i_mu = or_vec_hunt(t->mu, c->n_mu, mu, &a_mu);
  int linear_mu = mode == PPM_NSX_4L ||
                  ((mode == PPM_NSX_2L2C_LE || mode == PPM_NSX_4C_LE) && i_mu >= c->n_mu - 5);
  if (linear_mu) {
    nm = 2;
    wm[0] = 1 - a_mu;
    wm[1] = a_mu;
  } else {
    nm = 4;
    i_mu = clampi(i_mu - 1, 0, c->n_mu - 4);
    lagrange4(mu, t->mu[i_mu], t->mu[i_mu + 1], t->mu[i_mu + 2], t->mu[i_mu + 3], wm);
  }
  long sT = (long)c->n_logg * c->n_logEkT * c->n_mu, sg = (long)c->n_logEkT * c->n_mu, sE = c->n_mu;
  double ret = 0;
  for (int jT = 0; jT < nT; ++jT)
    for (int jg = 0; jg < ng; ++jg)
      for (int jE = 0; jE < nE; ++jE)
        for (int jm = 0; jm < nm; ++jm) {
          double v = t->data[(iT + jT) * sT + (ig + jg) * sg + (iE + jE) * sE + i_mu + jm];
          ret += v * (wT[jT] * wg[jg] * wE[jE] * wm[jm]);
        }
  if (mode == PPM_NSX_4C_SG && ret < 0) ret = 0;
  return ret;
}

double or_nsx_interp(const or_nsx* nsx, int mode, double logT, double logg, double logEkT, double mu) {
  return nsx_interp(nsx, mode, logT, logg, logEkT, mu, NULL);
}

/* ----------------------------------------------------------- profile */

typedef struct {
  double z;          /* cos(theta) of the ring            */
  double d_omega;    /* solid angle per patch             */
  int n_act;
  double* phi;       /* active patch azimuths             */
} ring_t;

typedef struct {
  const or_lensing* lt;
  const or_nsx* nsx;
  const ppm_star* star;
  const ppm_spot* spot;
  const ppm_grid* grid;
  const double* energies;
  int n_e;
  int n_pts;
  const double* pts;   /* phase point values            */
  const double* wts;   /* weight of each point          */
  const int* bin;      /* output row of each point      */
  ring_t* rings;
  int n_rings;
  double* flux;        /* [n_phase][n_e] accumulation   */
  int64_t n_eval, n_kept, n_clamped;
} job_t;

/* Ring loop body of calc_one_star (cpu/os1_cpu.cpp:62-182,
 * cpu/devarshi24st_cpu.cpp:89-241) for one ring. */
static void do_ring(job_t* J, const ring_t* ring, double* fl, double* zf, double* ph, double* cs) {
  const ppm_star* st = J->star;
  const ppm_spot* sp = J->spot;
  const int nf = J->grid->n_fine;
  const double M = st->mass_msun, Re = st->radius_eq_km, nu = st->spin_hz;
  const double Rs = M * RS_SUN_KM;
  const double Ob = 2 * PI_ * nu * sqrt(Re * Re * Re / M) * SQRT_KM3_S2_GMSUN;
  const double x = 0.5 * Rs / Re;
  const double o2 = Ob * Ob * (-0.788 + 1.030 * x);
  const double D2 = st->distance_km * st->distance_km;
  const double ci = cos(st->inclination_rad), si = sin(st->inclination_rad);
  lens_ctx L = {J->lt, &J->n_clamped};

  const double cz = ring->z, sz = sqrt(1. - cz * cz);
  const double R = Re * (1. + o2 * cz * cz);
  const double dRdth = -2. * Re * o2 * cz * sz;
  const double u = Rs / R, uu = sqrt(1. - u);
  const double f = dRdth / (R * uu), ff = sqrt(1. + f * f);
  const double dS = R * R * ring->d_omega * ff;
  const double cos_tau = 1. / ff, sin_tau = f / ff;

  /* surface gravity (cpu/devarshi24st_cpu.cpp:63-68, 103-124) */
  double logg = st->logg_override;
  if (!(logg > 0)) {
    const double ue = Rs / Re, uue = sqrt(1. - ue);
    const double logg0 = log10(M / (Re * Re * uue));
    const double ce = -0.791 + 0.776 * x, cp = 1.138 + -1.431 * x;
    const double de = -1.315 * x + 2.431 * x * x, fe = -1.172 * x;
    const double dp = 0.653 * x + -2.864 * x * x, fp = 0.975 * x;
    const double d60 = 13.47 * x + -27.13 * x * x;
    const double Ob2 = Ob * Ob, Ob4 = Ob2 * Ob2, Ob6 = Ob4 * Ob2;
    const double g_g0 = 1. + sz * sz * (ce * Ob2 + de * Ob4 + fe * Ob6) +
                        cz * cz * (cp * Ob2 + dp * Ob4 + fp * Ob6 - d60 * Ob4) + fabs(cz) * d60 * Ob4;
    logg = log10(g_g0) + logg0 + log10(LOG_GMSUN_CGS_ARG);
  }

  int inv_min = nf + 1, inv_max = -1;
  for (int k = 0; k < nf; ++k) {
    int vis = 1;
    do {
      double phase = 1. * k / nf;
      double sphi_ = 2 * PI_ * phase;
      double cphi = cos(sphi_), sphi = sin(sphi_);
      double cos_psi = ci * cz + si * sz * cphi;
      if (cos_psi < J->lt->cp_min) {
        vis = 0;
        break;
      }
      double sin_psi = sqrt(1. - cos_psi * cos_psi);
      double ca, lf;
      lens_ca_lf(&L, u, cos_psi, &ca, &lf);
      double sa = sqrt(1. - ca * ca);
      double sa_sp = cos_psi == 1. ? sqrt(lf) : sa / sin_psi;
      double cos_sigma;
      if (st->variant == PPM_VARIANT_OS1) /* cpu/os1_cpu.cpp:95-96 */
        cos_sigma = ca * cos_tau + sa_sp * sin_tau * (ci - cz * cos_psi) / sz;
      else /* cpu/devarshi24st_cpu.cpp:145-147 */
        cos_sigma = ca * cos_tau + sa_sp * sin_tau * (ci * sz - si * cz * cphi);
      if (cos_sigma <= 0) {
        vis = 0;
        break;
      }
      double beta = 2 * PI_ * nu * R * sz / uu / C_KM_S;
      double gamma = 1. / sqrt(1. - beta * beta);
      double cos_xi = -sa_sp * si * sphi;
      double delta = 1. / (gamma * (1. - beta * cos_xi));
      double delta3 = delta * delta * delta;
      double dt1 = lens_cdt(&L, u, ca) * R / C_KM_S;
      /* closed-form radial delay and turning point (cpu/os1_cpu.cpp:111-125) */
      double dt2;
      {
        double RR = R;
        double lg = log1p((Re - RR) / (RR - Rs));
        double dR = (Re - RR + Rs * lg) / C_KM_S;
        if (ca >= 0) {
          dt2 = dR;
        } else {
          double tmp = (2. * sa) / sqrt(3. * (1. - u));
          double p = -tmp * cos((acos(3. * u / tmp) + 2. * PI_) / 3.) * R;
          double lgp = log1p((Re - p) / (p - Rs));
          double dP = (Re - p + Rs * lgp) / C_KM_S;
          dt2 = 2. * dP - dR;
        }
      }
      double dphase = (dt1 + dt2) * nu;
      double csp = cos_sigma * delta;
      double beam = 1.0;
      if (st->atmosphere == PPM_ATM_BLACKBODY) { /* cpu/os1_cpu.cpp:128-136 */
        if (sp->beaming == PPM_BEAM_SIN2)
          beam = 1. - csp * csp;
        else if (sp->beaming == PPM_BEAM_COS2)
          beam = csp * csp;
      }
      fl[k] = uu * delta3 * csp * lf * (dS * gamma) / D2 * beam;
      zf[k] = 1. / (delta * uu);
      ph[k] = phase + dphase;
      cs[k] = csp;
    } while (0);
    if (!vis) {
      fl[k] = 0;
      zf[k] = -1;
      ph[k] = -1;
      cs[k] = 0;
      if (k < inv_min) inv_min = k;
      if (k > inv_max) inv_max = k;
    }
  }
  /* ring skip and gap fill (cpu/os1_cpu.cpp:151-167) */
  if (inv_min == 0 || inv_max == nf - 1) return;
  if (inv_min <= inv_max) {
    int beg = inv_min - 1, end = inv_max + 1;
    double p0 = ph[beg], p1 = ph[end], r0 = zf[beg], r1 = zf[end];
    for (int k = inv_min; k <= inv_max; ++k) {
      ph[k] = p0 + (p1 - p0) * ((double)(k - beg) / (end - beg));
      zf[k] = r0 + (r1 - r0) * ((double)(k - beg) / (end - beg));
    }
  }
  fl[nf] = fl[0];
  zf[nf] = zf[0];
  ph[nf] = 1.0 + ph[0];
  cs[nf] = cs[0];
  J->n_kept++;

  const int ne = J->n_e;
  const double kT = sp->kT_keV;
  const double logT = log10(kT / KB_KEV_K);
  const double kT_nsx = pow(10, logT) * KB_KEV_K;
  const double T3 = pow(10, 3 * logT);
  const double ibb_c = ibb_constant();
  const double mu_min = J->nsx ? J->nsx->mu[J->nsx->cfg.n_mu - 1] : 0;

  J->n_eval += (int64_t)ring->n_act * J->n_pts * ne;
  for (int q = 0; q < J->n_pts; ++q) {
    double* out = J->flux + (size_t)J->bin[q] * ne;
    for (int ia = 0; ia < ring->n_act; ++ia) {
      double shift = ring->phi[ia] / (2 * PI_);
      double pp = fmod(J->pts[q] + shift + 1., 1.);
      if (pp < ph[0]) pp += 1.;
      double a;
      int i = or_vec_hunt(ph, nf + 1, pp, &a);
      double fI = (1 - a) * fl[i] + a * fl[i + 1];
      if (fI == 0) continue;
      double zz = (1 - a) * zf[i] + a * zf[i + 1];
      double w = J->wts[q];
      if (st->atmosphere == PPM_ATM_BLACKBODY) { /* cpu/os1_cpu.cpp:178-180 */
        for (int e = 0; e < ne; ++e) {
          double E = J->energies[e], Ee = E * zz;
          double I = ibb_c * (Ee * Ee * Ee) / expm1(Ee / kT);
          out[e] += fI * I / E * w;
        }
      } else { /* cpu/devarshi24st_cpu.cpp:229-239 */
        double mu = (1 - a) * cs[i] + a * cs[i + 1];
        if (mu < mu_min) mu = mu_min;
        for (int e = 0; e < ne; ++e) {
          double E = J->energies[e], Ee = E * zz;
          double logEkT = log10(Ee / kT_nsx);
          double IT3 = nsx_interp(J->nsx, st->nsx_interp, logT, logg, logEkT, mu, &J->n_clamped);
          double I = IT3 * T3 / PLANCK_ERG_S;
          out[e] += fI * I / E * w;
        }
      }
    }
  }
}

typedef struct {
  job_t* J;
  int r0, r1;
  double* flux;
  int64_t n_eval, n_kept, n_clamped;
} worker_t;

static void* worker(void* arg) {
  worker_t* W = (worker_t*)arg;
  job_t J = *W->J;
  J.flux = W->flux;
  J.n_eval = J.n_kept = J.n_clamped = 0;
  int nf = J.grid->n_fine;
  double* buf = (double*)malloc(sizeof(double) * 4 * (nf + 1));
  for (int r = W->r0; r < W->r1; ++r)
    do_ring(&J, &J.rings[r], buf, buf + (nf + 1), buf + 2 * (nf + 1), buf + 3 * (nf + 1));
  free(buf);
  W->n_eval = J.n_eval;
  W->n_kept = J.n_kept;
  W->n_clamped = J.n_clamped;
  return NULL;
}

/* Active patches of a spot: HEALPix scan (cpu/grid.hpp:133-155, 224-245) or
 * the single-element point source (cpu/grid.hpp:17-27). */
static int build_rings(const ppm_star* st, const ppm_spot* sp, int n_side, ring_t** out) {
  if (st->point_source_below > 0 && sp->angrad1 < st->point_source_below) {
    ring_t* r = (ring_t*)calloc(1, sizeof(ring_t));
    double sh = sin(0.5 * sp->angrad1);
    r->z = cos(sp->theta1);
    r->d_omega = 2 * PI_ * 2. * sh * sh;
    r->n_act = 1;
    r->phi = (double*)malloc(sizeof(double));
    r->phi[0] = sp->phi1;
    *out = r;
    return 1;
  }
  int n_r = 4 * n_side - 1, count = 0;
  ring_t* rings = (ring_t*)calloc(n_r, sizeof(ring_t));
  double dom = PI_ / (3. * n_side * n_side);
  double* tmp = (double*)malloc(sizeof(double) * 4 * n_side);
  for (int i = 1; i <= n_r; ++i) {
    double z, s;
    int npix;
    or_healpix_ring(n_side, i, &z, &npix, &s);
    int na = 0;
    for (int j = 1; j <= npix; ++j) {
      double phi = or_healpix_phi(n_side, i, j);
      if (in_spot(sp, z, phi)) tmp[na++] = phi;
    }
    if (na) {
      ring_t* r = &rings[count++];
      r->z = z;
      r->d_omega = dom;
      r->n_act = na;
      r->phi = (double*)malloc(sizeof(double) * na);
      memcpy(r->phi, tmp, sizeof(double) * na);
    }
  }
  free(tmp);
  *out = rings;
  return count;
}

static void free_rings(ring_t* r, int n) {
  for (int i = 0; i < n; ++i) free(r[i].phi);
  free(r);
}

int64_t or_count_active(const ppm_star* star, const ppm_grid* grid) {
  int64_t total = 0;
  for (int s = 0; s < star->n_spots; ++s) {
    ring_t* rings;
    int n = build_rings(star, &star->spot[s], grid->n_side, &rings);
    for (int i = 0; i < n; ++i) total += rings[i].n_act;
    free_rings(rings, n);
  }
  return total;
}

int or_profile(const or_lensing* lt, const or_nsx* nsx, const or_response* rsp,
               const ppm_star* star, const ppm_grid* grid, const double* energies,
               double* flux, or_info* info, int n_threads, char* err, int errlen) {
  if (!lt || !star || !grid || !flux) {
    snprintf(err, errlen, "null argument");
    return 1;
  }
  if (star->n_spots < 1 || star->n_spots > PPM_MAX_SPOTS || grid->n_side < 1 || grid->n_phase < 1 ||
      grid->n_fine < 2) {
    snprintf(err, errlen, "bad star or grid");
    return 1;
  }
  if (star->atmosphere == PPM_ATM_NSX && !nsx) {
    snprintf(err, errlen, "NSX atmosphere requested without a grid");
    return 1;
  }
  if (star->atmosphere == PPM_ATM_NSX) {
    const int m = star->nsx_interp;
    if (m < PPM_NSX_2L2C_LE || m > PPM_NSX_4L) {
      snprintf(err, errlen, "unknown NSX interpolant");
      return 1;
    }
    if ((m == PPM_NSX_4C || m == PPM_NSX_4C_SG || m == PPM_NSX_4C_LE) &&
        (nsx->cfg.n_logT < 4 || nsx->cfg.n_logg < 4)) {
      snprintf(err, errlen, "cubic logT / log g stencils need 4 nodes per axis");
      return 1;
    }
  }
  int ne = grid->n_energy;
  if (!energies) {
    if (!rsp) {
      snprintf(err, errlen, "no energies and no response");
      return 1;
    }
    energies = rsp->e_obs;
    ne = rsp->n_e;
  }
  int gl = grid->sampling == PPM_SAMPLE_GL;
  int ng = gl ? grid->n_gl : 1;
  if (gl && (ng < 1 || ng > 6)) {
    snprintf(err, errlen, "n_gl must be 1..6");
    return 1;
  }
  int np = grid->n_phase * ng;
  double* pts = (double*)malloc(sizeof(double) * np);
  double* wts = (double*)malloc(sizeof(double) * np);
  int* bin = (int*)malloc(sizeof(int) * np);
  const double bw = 1. / grid->n_phase;
  /* point order of cpu/devarshi24st_cpu.cpp:214-221 (bin-major) */
  for (int q = 0; q < np; ++q) {
    int b = q / ng, g = q % ng;
    bin[q] = b;
    if (gl) {
      pts[q] = or_gl_node(ng, g, b * bw, (b + 1) * bw);
      wts[q] = or_gl_weight(ng, g, 1);
    } else {
      pts[q] = (double)b / grid->n_phase; /* linspace(0, 1 - 1/n, n), cpu/os1_cpu.cpp:36 */
      wts[q] = 1.0;
    }
  }
  memset(flux, 0, sizeof(double) * grid->n_phase * ne);
  or_info inf = {0};
  if (n_threads < 1) n_threads = 1;
  for (int s = 0; s < star->n_spots; ++s) {
    job_t J;
    memset(&J, 0, sizeof J);
    J.lt = lt;
    J.nsx = nsx;
    J.star = star;
    J.spot = &star->spot[s];
    J.grid = grid;
    J.energies = energies;
    J.n_e = ne;
    J.n_pts = np;
    J.pts = pts;
    J.wts = wts;
    J.bin = bin;
    J.n_rings = build_rings(star, J.spot, grid->n_side, &J.rings);
    for (int r = 0; r < J.n_rings; ++r) inf.n_active += J.rings[r].n_act;
    inf.n_rings += J.n_rings;
    int nt = n_threads < J.n_rings ? n_threads : (J.n_rings > 0 ? J.n_rings : 1);
    worker_t* W = (worker_t*)calloc(nt, sizeof(worker_t));
    pthread_t* th = (pthread_t*)calloc(nt, sizeof(pthread_t));
    /* contiguous ring blocks balanced by active patches */
    int64_t tot = 0;
    for (int r = 0; r < J.n_rings; ++r) tot += J.rings[r].n_act;
    int r = 0;
    int64_t acc = 0;
    for (int t = 0; t < nt; ++t) {
      W[t].J = &J;
      W[t].r0 = r;
      int64_t target = tot * (t + 1) / nt;
      while (r < J.n_rings && (acc < target || t == nt - 1)) acc += J.rings[r++].n_act;
      W[t].r1 = r;
      W[t].flux = nt == 1 ? flux : (double*)calloc((size_t)grid->n_phase * ne, sizeof(double));
    }
    if (nt == 1) {
      worker(&W[0]);
    } else {
      for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, worker, &W[t]);
      for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
      for (int t = 0; t < nt; ++t) {
        for (long k = 0; k < (long)grid->n_phase * ne; ++k) flux[k] += W[t].flux[k];
        free(W[t].flux);
      }
    }
    for (int t = 0; t < nt; ++t) {
      inf.n_eval_ref += W[t].n_eval;
      inf.n_rings_kept += W[t].n_kept;
      inf.n_clamped += W[t].n_clamped;
    }
    free(W);
    free(th);
    free_rings(J.rings, J.n_rings);
  }
  free(pts);
  free(wts);
  free(bin);
  if (info) *info = inf;
  return 0;
}

void or_fold(const or_response* rsp, const double* flux, int n_phase, double scale,
             double* counts) {
  for (int ch = 0; ch < rsp->n_ch; ++ch)
    for (int b = 0; b < n_phase; ++b) {
      double sum = 0;
      for (int e = 0; e < rsp->n_e; ++e) sum += flux[(size_t)b * rsp->n_e + e] * rsp->rsp[(size_t)ch * rsp->n_e + e];
      counts[(size_t)ch * n_phase + b] = sum * scale;
    }
}
