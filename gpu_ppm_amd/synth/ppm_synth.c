/*
 * ppm_synth - writes the synthetic inputs of the pulse-profile path in the
 * reference's own file formats:
 *
 *   lensing  OUTDIR [u_min u_max n_u n_cos]
 *       tables/lensing/<setting>/{u,cos_psi,cos_alpha}.txt ("min max N") and
 *       the three row-major matrices cos_alpha_of_u_cos_psi, lf_of_u_cos_psi,
 *       cdt_over_R_of_u_cos_alpha.  Algorithm of c_lensing.py:8-155: light
 *       bending psi(u, alpha) and the light-travel integral by quadrature, an
 *       alpha march of step pi/(8 n_cos) until psi passes acos(cos_psi_min),
 *       lensing factor d cos(alpha)/d cos(psi) by the non-uniform second-order
 *       difference, and monotone (Fritsch-Carlson / PCHIP) resampling onto the
 *       uniform cos(psi) axis.  Quadrature here is adaptive Gauss-Kronrod
 *       7/15 (QUADPACK's rule) instead of scipy's QAGS.
 *
 *   nsx      PATH [seed nT ng nE nmu [logg_max]]
 *       NSX-format text grid ("logEkT mu logInuT3 logT logg", loop order
 *       logT > logg > logEkT > mu, mu decreasing; cpu/nsx.hpp:83-148), by
 *       default of the 35x14x166x67 nsx_H_v200804 shape (log g 13.7-15.0;
 *       nsx_H_v171019 is 35 11 166 67 with logg_max 14.7), filled with a limb-darkened Planck
 *       intensity B_nu(T)/T^3 (1 + a mu) (1 + 0.02 n), n a seeded smooth
 *       noise with |n| <= 1.
 *
 *   nsx_bb   PATH [nT ng nE nmu]
 *       The same grid filled with the isotropic Planck intensity alone
 *       (a = 0, no noise): logInuT3 = log10(2 k^3 x^3 / (h^2 c^2 (e^x - 1))),
 *       x = E/kT, independent of logT, log g and mu, so that the NSX path
 *       reproduces the blackbody path up to the energy interpolation.
 *
 *   response OUTDIR n_ch n_E [seed]
 *       tables/ism_inst/<setting>/{num,E_obs,rsp}.txt (c_ism_inst.py:61-101
 *       formats): n_ch channels linear in 0.2-12 keV, n_E log-spaced E_obs,
 *       Gaussian redistribution FWHM = w0 sqrt(E), 10% exponential
 *       low-energy tail, log-normal effective area peaking at 1.5 keV, ISM
 *       factor exp(-N_H sigma0 E^(-8/3)), trapezoid energy weights folded in.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static const double PI = 3.14159265358979323846;

/* ---------------------------------------------------------------- utils */

static int mkdir_p(const char* path) {
  char buf[4096];
  size_t n = strlen(path);
  if (n >= sizeof buf) return -1;
  memcpy(buf, path, n + 1);
  for (size_t i = 1; i <= n; ++i) {
    if (buf[i] == '/' || buf[i] == '\0') {
      char c = buf[i];
      buf[i] = '\0';
      if (mkdir(buf, 0755) != 0 && errno != EEXIST) return -1;
      buf[i] = c;
    }
  }
  return 0;
}

static FILE* open_out(const char* dir, const char* name) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE* f = fopen(path, "w");
  if (!f) fprintf(stderr, "ppm_synth: cannot write %s\n", path);
  return f;
}

/* splitmix64: small seeded generator for the synthetic fills. */
static uint64_t sm_state;
static uint64_t sm_next(void) {
  uint64_t z = (sm_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double sm_uniform(double lo, double hi) {
  return lo + (hi - lo) * ((double)(sm_next() >> 11) * (1.0 / 9007199254740992.0));
}

/* ------------------------------------------- adaptive Gauss-Kronrod 7/15 */

/* Kronrod abscissae (descending, last = 0) and weights of the 15-point rule;
 * Gauss 7-point weights for abscissae xgk[1], xgk[3], xgk[5], xgk[7]. */
static const double XGK[8] = {
    0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
    0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
    0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
    0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double WGK[8] = {
    0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
    0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
    0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
    0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double WG[4] = {
    0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
    0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

typedef double (*integrand_fn)(double x, const double* p);

static double gk15(integrand_fn f, const double* p, double a, double b, double* err) {
  double c = 0.5 * (a + b), h = 0.5 * (b - a);
  double fc = f(c, p);
  double rk = fc * WGK[7], rg = fc * WG[3];
  for (int j = 0; j < 7; ++j) {
    double dx = h * XGK[j];
    double s = f(c - dx, p) + f(c + dx, p);
    rk += WGK[j] * s;
    if (j & 1) rg += WG[j / 2] * s;
  }
  *err = fabs((rk - rg) * h);
  return rk * h;
}

static double adapt(integrand_fn f, const double* p, double a, double b, double whole,
                    double err, double tol, int depth) {
  if (err <= tol || depth >= 40) return whole;
  double m = 0.5 * (a + b), e1, e2;
  double l = gk15(f, p, a, m, &e1), r = gk15(f, p, m, b, &e2);
  return adapt(f, p, a, m, l, e1, 0.5 * tol, depth + 1) +
         adapt(f, p, m, b, r, e2, 0.5 * tol, depth + 1);
}

static double integrate01(integrand_fn f, const double* p) {
  double err, whole = gk15(f, p, 0.0, 1.0, &err);
  double tol = fmax(1e-14, 1e-12 * fabs(whole));
  return adapt(f, p, 0.0, 1.0, whole, err, tol, 0);
}

/* --------------------------------------------------- light bending tables */

/* Substitution x in [0,1] of the bending integral (c_lensing.py:47-54);
 * p = {u, sin^2 alpha}. */
static double bend_integrand(double x, const double* p) {
  double u = p[0], s2 = p[1], c2 = 1.0 - s2, x2 = x * x;
  double q = (2.0 - x2 - u * (1.0 - x2) * (1.0 - x2) / (1.0 - u)) * s2;
  return x / sqrt(c2 + x2 * q);
}

/* Light-travel-time integrand (c_lensing.py:115-122). */
static double delay_integrand(double x, const double* p) {
  double u = p[0], s2 = p[1], c2 = 1.0 - s2, x2 = x * x;
  double q = (2.0 - x2 - u * (1.0 - x2) * (1.0 - x2) / (1.0 - u)) * s2;
  double t = sqrt(c2 + x2 * q);
  return x / (t * (1.0 + t));
}

static double psi_out(double u, double alpha) { /* c_lensing.py:57-61 */
  double s = sin(alpha), p[2] = {u, s * s};
  return 2.0 * s / sqrt(1.0 - u) * integrate01(bend_integrand, p);
}

static double turning_point_over_R(double u, double alpha) { /* c_lensing.py:65-67 */
  double t = 2.0 * sin(alpha) / sqrt(3.0 * (1.0 - u));
  return -t * cos((acos(3.0 * u / t) + 2.0 * PI) / 3.0);
}

static double psi_of(double u, double alpha) { /* c_lensing.py:64-77 */
  if (alpha <= 0.5 * PI) return psi_out(u, alpha);
  double pR = turning_point_over_R(u, alpha);
  return 2.0 * psi_out(u / pR, 0.5 * PI) - psi_out(u, PI - alpha);
}

static double cdt_out(double u, double alpha) { /* c_lensing.py:125-129 */
  double s = sin(alpha), p[2] = {u, s * s};
  return 2.0 * s * s / (1.0 - u) * integrate01(delay_integrand, p);
}

static double cdt_of(double u, double alpha) { /* c_lensing.py:132-145 */
  if (alpha <= 0.5 * PI) return cdt_out(u, alpha);
  double pR = turning_point_over_R(u, alpha);
  return 2.0 * cdt_out(u / pR, 0.5 * PI) - cdt_out(u, PI - alpha);
}

/* Second-order derivative on a non-uniform axis, first order at the ends
 * (numpy.gradient with coordinates, edge_order 1). */
static void gradient_nonuniform(const double* f, const double* x, int n, double* out) {
  out[0] = (f[1] - f[0]) / (x[1] - x[0]);
  out[n - 1] = (f[n - 1] - f[n - 2]) / (x[n - 1] - x[n - 2]);
  for (int i = 1; i < n - 1; ++i) {
    double d1 = x[i] - x[i - 1], d2 = x[i + 1] - x[i];
    out[i] = -d2 / (d1 * (d1 + d2)) * f[i - 1] + (d2 - d1) / (d1 * d2) * f[i] +
             d1 / (d2 * (d1 + d2)) * f[i + 1];
  }
}

static double sgn(double v) { return (v > 0) - (v < 0); }

static double pchip_end_slope(double h0, double h1, double m0, double m1) {
  double d = ((2.0 * h0 + h1) * m0 - h0 * m1) / (h0 + h1);
  if (sgn(d) != sgn(m0)) return 0.0;
  if (sgn(m0) != sgn(m1) && fabs(d) > 3.0 * fabs(m0)) return 3.0 * m0;
  return d;
}

/* Monotone piecewise-cubic Hermite (Fritsch-Carlson slopes with the
 * weighted harmonic mean) through ascending x; evaluates q[0..nq-1] with the
 * end polynomials extending past the data. */
static void pchip(const double* x, const double* y, int n, const double* q, int nq, double* out) {
  double* d = (double*)malloc(sizeof(double) * n);
  double* h = (double*)malloc(sizeof(double) * (n - 1));
  double* m = (double*)malloc(sizeof(double) * (n - 1));
  for (int k = 0; k < n - 1; ++k) {
    h[k] = x[k + 1] - x[k];
    m[k] = (y[k + 1] - y[k]) / h[k];
  }
  for (int k = 1; k < n - 1; ++k) {
    if (sgn(m[k]) != sgn(m[k - 1]) || m[k] == 0.0 || m[k - 1] == 0.0) {
      d[k] = 0.0;
    } else {
      double w1 = 2.0 * h[k] + h[k - 1], w2 = h[k] + 2.0 * h[k - 1];
      d[k] = (w1 + w2) / (w1 / m[k - 1] + w2 / m[k]);
    }
  }
  d[0] = pchip_end_slope(h[0], h[1], m[0], m[1]);
  d[n - 1] = pchip_end_slope(h[n - 2], h[n - 3], m[n - 2], m[n - 3]);
  int k = 0;
  for (int i = 0; i < nq; ++i) {
    double xi = q[i];
    while (k < n - 2 && xi > x[k + 1]) ++k;
    while (k > 0 && xi < x[k]) --k;
    double t = (xi - x[k]) / h[k], t2 = t * t, t3 = t2 * t;
    double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t, h01 = -2 * t3 + 3 * t2,
           h11 = t3 - t2;
    out[i] = h00 * y[k] + h10 * h[k] * d[k] + h01 * y[k + 1] + h11 * h[k] * d[k + 1];
  }
  free(d);
  free(h);
  free(m);
}

static int write_axis(const char* dir, const char* name, double lo, double hi, int n) {
  FILE* f = open_out(dir, name);
  if (!f) return -1;
  fprintf(f, "%.17g %.17g %d\n", lo, hi, n);
  fclose(f);
  return 0;
}

static int write_matrix(const char* dir, const char* name, const double* a, int rows, int cols) {
  FILE* f = open_out(dir, name);
  if (!f) return -1;
  for (int i = 0; i < rows; ++i) {
    for (int j = 0; j < cols; ++j) fprintf(f, j ? " %.18e" : "%.18e", a[(size_t)i * cols + j]);
    fputc('\n', f);
  }
  fclose(f);
  return 0;
}

static int cmd_lensing(const char* dir, double u_min, double u_max, int n_u, int n_cos) {
  const double cp_min = -0.9, cp_max = 1.0, ca_min = -0.2, ca_max = 1.0; /* c_lensing.py:12-16 */
  if (n_u < 2 || n_cos < 4 || !(u_min < u_max) || u_max >= 0.568 || u_min <= 0) {
    fprintf(stderr, "ppm_synth lensing: bad axes\n");
    return 2;
  }
  if (mkdir_p(dir)) return 1;
  if (write_axis(dir, "u.txt", u_min, u_max, n_u) || write_axis(dir, "cos_psi.txt", cp_min, cp_max, n_cos) ||
      write_axis(dir, "cos_alpha.txt", ca_min, ca_max, n_cos))
    return 1;
  double* ca_tab = (double*)calloc((size_t)n_u * n_cos, sizeof(double));
  double* lf_tab = (double*)calloc((size_t)n_u * n_cos, sizeof(double));
  double* dt_tab = (double*)calloc((size_t)n_u * n_cos, sizeof(double));
  double* cp_out = (double*)malloc(sizeof(double) * n_cos);
  double* ca_out = (double*)malloc(sizeof(double) * n_cos);
  for (int j = 0; j < n_cos; ++j) {
    cp_out[j] = cp_min + (cp_max - cp_min) * j / (n_cos - 1);
    ca_out[j] = ca_min + (ca_max - ca_min) * j / (n_cos - 1);
  }
  const double psi_stop = acos(cp_out[0]);
  const double d_alpha = PI / (n_cos * 8.0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int iu = 0; iu < n_u; ++iu) {
    double u = u_min + (u_max - u_min) * iu / (n_u - 1);
    int cap = 4096, n = 0;
    double* al = (double*)malloc(sizeof(double) * cap);
    double* ps = (double*)malloc(sizeof(double) * cap);
    for (double alpha = 0.0;; alpha += d_alpha) {
      if (n == cap) {
        cap *= 2;
        al = (double*)realloc(al, sizeof(double) * cap);
        ps = (double*)realloc(ps, sizeof(double) * cap);
      }
      al[n] = alpha;
      ps[n] = psi_of(u, alpha);
      if (ps[n++] > psi_stop) break;
    }
    /* ascending cos(psi): reverse the march */
    double* xs = (double*)malloc(sizeof(double) * n);
    double* ys = (double*)malloc(sizeof(double) * n);
    double* gs = (double*)malloc(sizeof(double) * n);
    double* cpg = (double*)malloc(sizeof(double) * n);
    double* cag = (double*)malloc(sizeof(double) * n);
    for (int k = 0; k < n; ++k) {
      cpg[k] = cos(ps[k]);
      cag[k] = cos(al[k]);
    }
    gradient_nonuniform(cag, cpg, n, gs); /* lensing factor, c_lensing.py:100 */
    for (int k = 0; k < n; ++k) {
      xs[k] = cpg[n - 1 - k];
      ys[k] = cag[n - 1 - k];
    }
    pchip(xs, ys, n, cp_out, n_cos, ca_tab + (size_t)iu * n_cos);
    for (int k = 0; k < n; ++k) ys[k] = gs[n - 1 - k];
    pchip(xs, ys, n, cp_out, n_cos, lf_tab + (size_t)iu * n_cos);
    for (int j = 0; j < n_cos; ++j) dt_tab[(size_t)iu * n_cos + j] = cdt_of(u, acos(ca_out[j]));
    free(al);
    free(ps);
    free(xs);
    free(ys);
    free(gs);
    free(cpg);
    free(cag);
  }
  int rc = write_matrix(dir, "cos_alpha_of_u_cos_psi.txt", ca_tab, n_u, n_cos) ||
           write_matrix(dir, "lf_of_u_cos_psi.txt", lf_tab, n_u, n_cos) ||
           write_matrix(dir, "cdt_over_R_of_u_cos_alpha.txt", dt_tab, n_u, n_cos);
  free(ca_tab);
  free(lf_tab);
  free(dt_tab);
  free(cp_out);
  free(ca_out);
  return rc;
}

/* ------------------------------------------------------------ NSX grid */

static int cmd_nsx(const char* path, uint64_t seed, int nT, int ng, int nE, int nm, double logg_max, double limb_a,
                   double noise_amp) {
  const double logT_min = 5.10, logT_max = 6.80, logg_min = 13.7;
  const double logE_min = -1.30, logE_max = 2.00; /* cpu/nsx.hpp:25-37 */
  const double kB = 1.380649e-16, h = 6.62607015e-27, c = 2.99792458e10;
  const double pref = 2.0 * kB * kB * kB / (h * h * c * c);
  sm_state = seed;
  double amp[4], kv[4][4], ph[4], asum = 0;
  for (int m = 0; m < 4; ++m) {
    amp[m] = sm_uniform(0.5, 1.0);
    asum += amp[m];
    kv[m][0] = sm_uniform(1.0, 4.0);
    kv[m][1] = sm_uniform(1.0, 4.0);
    kv[m][2] = sm_uniform(1.0, 5.0);
    kv[m][3] = sm_uniform(1.0, 6.0);
    ph[m] = sm_uniform(0.0, 2.0 * PI);
  }
  for (int m = 0; m < 4; ++m) amp[m] /= asum;
  /* mu nodes: strictly decreasing from 1 to sin(1 deg) */
  double* mu = (double*)malloc(sizeof(double) * nm);
  for (int k = 0; k < nm; ++k) {
    double t = (double)k / (nm - 1);
    mu[k] = cos(t * (0.5 * PI - 0.0175) * (0.35 + 0.65 * t));
  }
  char dir[4096];
  snprintf(dir, sizeof dir, "%s", path);
  char* slash = strrchr(dir, '/');
  if (slash) {
    *slash = '\0';
    if (mkdir_p(dir)) return 1;
  }
  FILE* f = fopen(path, "w");
  if (!f) {
    fprintf(stderr, "ppm_synth: cannot write %s\n", path);
    return 1;
  }
  static char buf[1 << 22];
  setvbuf(f, buf, _IOFBF, sizeof buf);
  for (int iT = 0; iT < nT; ++iT) {
    double logT = logT_min + (logT_max - logT_min) * iT / (nT - 1);
    for (int ig = 0; ig < ng; ++ig) {
      double logg = logg_min + (logg_max - logg_min) * ig / (ng - 1);
      for (int iE = 0; iE < nE; ++iE) {
        double logE = logE_min + (logE_max - logE_min) * iE / (nE - 1);
        double x = pow(10.0, logE);
        double planck = pref * x * x * x / expm1(x);
        for (int im = 0; im < nm; ++im) {
          double v[4] = {logT, logg, logE, mu[im]}, noise = 0;
          for (int m = 0; m < 4; ++m)
            noise += amp[m] * sin(kv[m][0] * v[0] + kv[m][1] * v[1] + kv[m][2] * v[2] +
                                  kv[m][3] * v[3] + ph[m]);
          double it3 = planck * (1.0 + limb_a * mu[im]) * (1.0 + noise_amp * noise);
          fprintf(f, "%.10g %.10g %.12g %.10g %.10g\n", logE, mu[im], log10(it3), logT, logg);
        }
      }
    }
  }
  free(mu);
  return fclose(f) ? 1 : 0;
}

/* ------------------------------------------------------------- response */

static int cmd_response(const char* dir, int n_ch, int n_e, uint64_t seed) {
  const double e_lo = 0.2, e_hi = 12.0, w0 = 0.1 /* keV FWHM at 1 keV */;
  const double nh = 1e21, sigma0 = 2.4e-22, tail_frac = 0.10, tail_len = 0.4;
  if (n_ch < 2 || n_e < 2) return 2;
  if (mkdir_p(dir)) return 1;
  sm_state = seed ^ 0x5EED5EEDull;
  double ripple_k = sm_uniform(2.0, 5.0), ripple_ph = sm_uniform(0, 2 * PI);
  double* e = (double*)malloc(sizeof(double) * n_e);
  double* w = (double*)malloc(sizeof(double) * n_e);
  double* jit = (double*)malloc(sizeof(double) * n_ch);
  for (int k = 0; k < n_e; ++k) e[k] = exp(log(e_lo) + k * (log(e_hi) - log(e_lo)) / (n_e - 1));
  e[0] = e_lo;
  e[n_e - 1] = e_hi;
  for (int k = 0; k < n_e; ++k) {
    double left = k > 0 ? e[k] - e[k - 1] : 0.0, right = k < n_e - 1 ? e[k + 1] - e[k] : 0.0;
    w[k] = 0.5 * (left + right);
  }
  for (int c = 0; c < n_ch; ++c) jit[c] = 1.0 + sm_uniform(-0.05, 0.05);
  double dch = (e_hi - e_lo) / n_ch;
  double* rsp = (double*)calloc((size_t)n_ch * n_e, sizeof(double));
  for (int k = 0; k < n_e; ++k) {
    double E = e[k];
    double sig = w0 * sqrt(E) / 2.3548200450309493;
    double lnr = log(E / 1.5);
    double aeff = 350.0 * exp(-lnr * lnr / (2 * 0.55 * 0.55)) *
                  (1.0 + 0.03 * sin(ripple_k * log(E) + ripple_ph));
    double ism = exp(-nh * sigma0 * pow(E, -8.0 / 3.0));
    /* tail normalisation over the channel range below E */
    double tail_norm = 1.0 - exp(-(E - e_lo) / tail_len);
    for (int c = 0; c < n_ch; ++c) {
      double a = e_lo + c * dch, b = a + dch;
      double g = 0.5 * (erf((b - E) / (sqrt(2.0) * sig * jit[c])) -
                        erf((a - E) / (sqrt(2.0) * sig * jit[c])));
      double t = 0.0;
      if (a < E && tail_norm > 0) {
        double bb = b < E ? b : E;
        t = (exp(-(E - bb) / tail_len) - exp(-(E - a) / tail_len)) / tail_norm;
      }
      double r = (1.0 - tail_frac) * g + tail_frac * t;
      rsp[(size_t)c * n_e + k] = r * aeff * ism * w[k];
    }
  }
  FILE* f = open_out(dir, "num.txt");
  if (!f) return 1;
  fprintf(f, "%d %d\n", n_ch, n_e);
  fclose(f);
  f = open_out(dir, "E_obs.txt");
  if (!f) return 1;
  for (int k = 0; k < n_e; ++k) fprintf(f, "%.18e\n", e[k]);
  fclose(f);
  int rc = write_matrix(dir, "rsp.txt", rsp, n_ch, n_e);
  free(e);
  free(w);
  free(jit);
  free(rsp);
  return rc;
}

int main(int argc, char** argv) {
  if (argc >= 3 && !strcmp(argv[1], "lensing")) {
    double u_min = argc > 3 ? atof(argv[3]) : 0.3, u_max = argc > 4 ? atof(argv[4]) : 0.5;
    int n_u = argc > 5 ? atoi(argv[5]) : 256, n_cos = argc > 6 ? atoi(argv[6]) : 512;
    return cmd_lensing(argv[2], u_min, u_max, n_u, n_cos);
  }
  if (argc >= 3 && !strcmp(argv[1], "nsx")) {
    uint64_t seed = argc > 3 ? strtoull(argv[3], 0, 10) : 1;
    int nT = argc > 4 ? atoi(argv[4]) : 35, ng = argc > 5 ? atoi(argv[5]) : 14;
    int nE = argc > 6 ? atoi(argv[6]) : 166, nm = argc > 7 ? atoi(argv[7]) : 67;
    double logg_max = argc > 8 ? atof(argv[8]) : 15.0;
    return cmd_nsx(argv[2], seed, nT, ng, nE, nm, logg_max, 0.5, 0.02);
  }
  if (argc >= 3 && !strcmp(argv[1], "nsx_bb")) {
    int nT = argc > 3 ? atoi(argv[3]) : 35, ng = argc > 4 ? atoi(argv[4]) : 14;
    int nE = argc > 5 ? atoi(argv[5]) : 166, nm = argc > 6 ? atoi(argv[6]) : 67;
    return cmd_nsx(argv[2], 1, nT, ng, nE, nm, 15.0, 0.0, 0.0);
  }
  if (argc >= 5 && !strcmp(argv[1], "response")) {
    uint64_t seed = argc > 5 ? strtoull(argv[5], 0, 10) : 0;
    return cmd_response(argv[2], atoi(argv[3]), atoi(argv[4]), seed);
  }
  fprintf(stderr,
          "usage: ppm_synth lensing OUTDIR [u_min u_max n_u n_cos]\n"
          "       ppm_synth nsx PATH [seed nT ng nE nmu [logg_max]]\n"
          "       ppm_synth nsx_bb PATH [nT ng nE nmu]\n"
          "       ppm_synth response OUTDIR n_ch n_E [seed]\n");
  return 2;
}
