// Instrument response on the device: rsp = (area x RMF x ISM^5) @ C, the
// steps of c_ism_inst.py:30-101 (the reference runs them in numpy/scipy):
//   - trim to channels [ch_lo, ch_hi) and energies [e_lo, e_hi), response =
//     area x RMF (c_ism_inst.py:9-12, 43-46, 58);
//   - E_obs log-spaced between the first lower and the last upper bin edge,
//     ends exact (np.geomspace, c_ism_inst.py:61);
//   - C[j][i] = integral over fine bin j of the not-a-knot cubic spline
//     through the unit vector e_i on E_obs (scipy CubicSpline,
//     c_ism_inst.py:75-82): one thread per i solves the spline's tridiagonal
//     system, one thread per (j, i) integrates;
//   - ISM: not-a-knot cubic spline of the transmission at the fine bin
//     centres, to the 5th power (c_ism_inst.py:85-89);
//   - rsp = (response x frac) @ C in FP64, one thread per (channel, E_obs),
//     the fine energies summed in ascending order; |x| < 1e-300 set to 0
//     (c_ism_inst.py:96-98).
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/ppm.h"

namespace ppm {

namespace {

// Not-a-knot spline through (x[k], y[k]), k < n (n >= 4), as piece
// coefficients c[k][0..3] of y = c0 + c1 t + c2 t^2 + c3 t^3, t = v - x_k:
// the first derivatives s solve the tridiagonal system of second-derivative
// continuity at interior knots and third-derivative continuity at x_1 and
// x_{n-2} (Thomas algorithm; w is scratch of 4 n doubles).
__device__ void not_a_knot(const double* x, const double* y, int n, double* c, double* w) {
  double *a = w, *b = w + n, *cc = w + 2 * n, *r = w + 3 * n;
  for (int k = 1; k < n - 1; ++k) {
    const double d0 = x[k] - x[k - 1], d1 = x[k + 1] - x[k];
    const double s0 = (y[k] - y[k - 1]) / d0, s1 = (y[k + 1] - y[k]) / d1;
    a[k] = d1;
    b[k] = 2 * (d0 + d1);
    cc[k] = d0;
    r[k] = 3 * (d1 * s0 + d0 * s1);
  }
  {
    const double dx0 = x[1] - x[0], dx1 = x[2] - x[1], d = x[2] - x[0];
    const double sl0 = (y[1] - y[0]) / dx0, sl1 = (y[2] - y[1]) / dx1;
    a[0] = 0;
    b[0] = dx1;
    cc[0] = d;
    r[0] = ((dx0 + 2 * d) * dx1 * sl0 + dx0 * dx0 * sl1) / d;
    const double dm2 = x[n - 1] - x[n - 2], dm3 = x[n - 2] - x[n - 3], e = x[n - 1] - x[n - 3];
    const double slm2 = (y[n - 1] - y[n - 2]) / dm2, slm3 = (y[n - 2] - y[n - 3]) / dm3;
    a[n - 1] = e;
    b[n - 1] = dm3;
    cc[n - 1] = 0;
    r[n - 1] = (dm2 * dm2 * slm3 + (2 * e + dm2) * dm3 * slm2) / e;
  }
  for (int k = 1; k < n; ++k) {
    const double m = a[k] / b[k - 1];
    b[k] -= m * cc[k - 1];
    r[k] -= m * r[k - 1];
  }
  double* s = a;  // slopes overwrite a (no longer needed)
  s[n - 1] = r[n - 1] / b[n - 1];
  for (int k = n - 2; k >= 0; --k) s[k] = (r[k] - cc[k] * s[k + 1]) / b[k];
  for (int k = 0; k < n - 1; ++k) {
    const double dx = x[k + 1] - x[k], sl = (y[k + 1] - y[k]) / dx;
    const double t = (s[k] + s[k + 1] - 2 * sl) / dx;
    c[4 * k + 0] = y[k];
    c[4 * k + 1] = s[k];
    c[4 * k + 2] = (sl - s[k]) / dx - t;
    c[4 * k + 3] = t / dx;
  }
}

// Piece of the spline holding v (the end pieces extrapolate).
__device__ inline int piece(const double* x, int n, double v) {
  int lo = 0, hi = n - 2;
  while (lo < hi) {
    const int m = (lo + hi + 1) / 2;
    if (v >= x[m])
      lo = m;
    else
      hi = m - 1;
  }
  return lo;
}

__device__ inline double antideriv(const double* c, double t) {
  return t * (c[0] + t * (c[1] / 2 + t * (c[2] / 3 + t * c[3] / 4)));
}

__device__ double integrate(const double* x, int n, const double* c, double a, double b) {
  const int ka = piece(x, n, a), kb = piece(x, n, b);
  if (ka == kb) return antideriv(c + 4 * ka, b - x[ka]) - antideriv(c + 4 * ka, a - x[ka]);
  double s = antideriv(c + 4 * ka, x[ka + 1] - x[ka]) - antideriv(c + 4 * ka, a - x[ka]);
  for (int k = ka + 1; k < kb; ++k) s += antideriv(c + 4 * k, x[k + 1] - x[k]);
  return s + antideriv(c + 4 * kb, b - x[kb]);
}

struct RespDev {
  const double *rmf, *arf, *ism;  // inputs as in ppm_response_spec
  double* e_obs;                  // [n_eobs]
  double* unit_c;                 // [n_eobs][n_eobs - 1][4] spline pieces of the unit vectors
  double* unit_w;                 // [n_eobs][4 n_eobs] solver scratch, then [n_eobs][n_eobs] right-hand sides
  double* ism_c;                  // [n_ism - 1][4]
  double* ism_w;                  // [4 n_ism]
  double* cmat;                   // [n_e][n_eobs]
  double* frac;                   // [n_e]
  double* ism_x;                  // [n_ism] knots
  double* ism_y;                  // [n_ism]
  double* rsp;                    // [n_ch][n_eobs]
  int n_e_all, n_ch_all, n_ism, e_lo, n_e, ch_lo, n_ch, n_eobs;
};

__global__ void k_eobs(RespDev R) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= R.n_eobs) return;
  const double lo = R.arf[3 * R.e_lo], hi = R.arf[3 * (R.e_lo + R.n_e - 1) + 1];
  const double l0 = log10(lo), l1 = log10(hi), step = (l1 - l0) / (R.n_eobs - 1);
  double v = pow(10.0, l0 + k * step);
  if (k == 0) v = lo;
  if (k == R.n_eobs - 1) v = hi;
  R.e_obs[k] = v;
}

// One thread per unit vector e_i (and thread n_eobs: the ISM spline).  A
// unit-vector thread keeps its solver scratch at unit_w + 4 n i and its y at
// unit_w + 4 n^2 + n i (n = n_eobs).
__global__ void k_splines(RespDev R) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = R.n_eobs;
  if (i < n) {
    double* w = R.unit_w + (size_t)i * 4 * n;
    double* y = R.unit_w + (size_t)4 * n * n + (size_t)i * n;
    for (int k = 0; k < n; ++k) y[k] = k == i ? 1.0 : 0.0;
    not_a_knot(R.e_obs, y, n, R.unit_c + (size_t)i * (n - 1) * 4, w);
  } else if (i == n) {
    for (int k = 0; k < R.n_ism; ++k) {
      R.ism_x[k] = R.ism[2 * k];
      R.ism_y[k] = R.ism[2 * k + 1];
    }
    not_a_knot(R.ism_x, R.ism_y, R.n_ism, R.ism_c, R.ism_w);
  }
}

__global__ void k_cmatrix(RespDev R) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)R.n_e * R.n_eobs) return;
  const int j = (int)(idx / R.n_eobs), i = (int)(idx - (size_t)j * R.n_eobs);
  const double* a = R.arf + 3 * (R.e_lo + j);
  R.cmat[idx] = integrate(R.e_obs, R.n_eobs, R.unit_c + (size_t)i * (R.n_eobs - 1) * 4, a[0], a[1]);
}

__global__ void k_frac(RespDev R) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R.n_e) return;
  const double* a = R.arf + 3 * (R.e_lo + j);
  const double mid = (a[0] + a[1]) * 0.5;
  const int k = piece(R.ism_x, R.n_ism, mid);
  const double* c = R.ism_c + 4 * k;
  const double t = mid - R.ism_x[k];
  R.frac[j] = pow(c[0] + t * (c[1] + t * (c[2] + t * c[3])), 5.0);
}

// rsp[c][i] = sum_j area_j rmf[e_lo + j][ch_lo + c] frac_j C[j][i], j ascending.
__global__ void k_fold_response(RespDev R) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)R.n_ch * R.n_eobs) return;
  const int c = (int)(idx / R.n_eobs), i = (int)(idx - (size_t)c * R.n_eobs);
  double s = 0.0;
  for (int j = 0; j < R.n_e; ++j) {
    const double v = R.arf[3 * (R.e_lo + j) + 2] * R.rmf[(size_t)(R.e_lo + j) * R.n_ch_all + R.ch_lo + c] * R.frac[j];
    if (v == 0.0) continue;
    s += v * R.cmat[(size_t)j * R.n_eobs + i];
  }
  R.rsp[idx] = fabs(s) < 1e-300 ? 0.0 : s;
}

}  // namespace

}  // namespace ppm

using namespace ppm;

extern "C" ppm_status ppm_response_build(const ppm_response_spec* sp, double* e_obs, double* rsp) {
  if (!sp || !e_obs || !rsp || !sp->rmf || !sp->arf || !sp->ism) return PPM_ERR_ARG;
  const ppm_response_spec S = *sp;
  if (S.n_e_all < 1 || S.n_ch_all < 1 || S.n_ism < 4 || S.n_eobs < 4 || S.e_lo < 0 || S.e_hi > S.n_e_all ||
      S.e_lo >= S.e_hi || S.ch_lo < 0 || S.ch_hi > S.n_ch_all || S.ch_lo >= S.ch_hi)
    return PPM_ERR_ARG;
  for (int k = 0; k + 1 < S.n_ism; ++k)
    if (!(S.ism[2 * (k + 1)] > S.ism[2 * k])) return PPM_ERR_ARG;  // spline knots increase strictly
  for (int j = S.e_lo; j < S.e_hi; ++j)
    if (!(S.arf[3 * j] > 0) || !(S.arf[3 * j + 1] > S.arf[3 * j])) return PPM_ERR_ARG;
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || S.device < 0 || S.device >= n_dev) return PPM_ERR_DEVICE;
  if (hipSetDevice(S.device) != hipSuccess) return PPM_ERR_DEVICE;
  const int n_e = S.e_hi - S.e_lo, n_ch = S.ch_hi - S.ch_lo, ne = S.n_eobs, ni = S.n_ism;
  const size_t n_rmf = (size_t)S.n_e_all * S.n_ch_all;
  // one device block: inputs, then every intermediate
  const size_t sizes[] = {n_rmf, 3 * (size_t)S.n_e_all, 2 * (size_t)ni, (size_t)ne, (size_t)ne * (ne - 1) * 4,
                          (size_t)ne * 5 * ne, (size_t)(ni - 1) * 4, 4 * (size_t)ni,
                          (size_t)n_e * ne, (size_t)n_e, (size_t)ni, (size_t)ni, (size_t)n_ch * ne};
  size_t total = 0;
  for (size_t v : sizes) total += v;
  double* base = nullptr;
  if (hipMalloc(&base, sizeof(double) * total) != hipSuccess) return PPM_ERR_DEVICE;
  double* p[13];
  {
    double* q = base;
    for (int k = 0; k < 13; ++k) {
      p[k] = q;
      q += sizes[k];
    }
  }
  RespDev R{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10], p[11], p[12],
            S.n_e_all, S.n_ch_all, ni, S.e_lo, n_e, S.ch_lo, n_ch, ne};
  hipError_t e = hipMemcpy(p[0], S.rmf, sizeof(double) * n_rmf, hipMemcpyHostToDevice);
  if (!e) e = hipMemcpy(p[1], S.arf, sizeof(double) * 3 * S.n_e_all, hipMemcpyHostToDevice);
  if (!e) e = hipMemcpy(p[2], S.ism, sizeof(double) * 2 * ni, hipMemcpyHostToDevice);
  if (!e) {
    hipLaunchKernelGGL(k_eobs, dim3((ne + 63) / 64), dim3(64), 0, 0, R);
    hipLaunchKernelGGL(k_splines, dim3((ne + 1 + 63) / 64), dim3(64), 0, 0, R);
    const size_t nc = (size_t)n_e * ne, nr = (size_t)n_ch * ne;
    hipLaunchKernelGGL(k_cmatrix, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, 0, R);
    hipLaunchKernelGGL(k_frac, dim3((n_e + 255) / 256), dim3(256), 0, 0, R);
    hipLaunchKernelGGL(k_fold_response, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, 0, R);
    e = hipGetLastError();
  }
  if (!e) e = hipMemcpy(e_obs, R.e_obs, sizeof(double) * ne, hipMemcpyDeviceToHost);
  if (!e) e = hipMemcpy(rsp, R.rsp, sizeof(double) * n_ch * ne, hipMemcpyDeviceToHost);
  (void)hipFree(base);
  return e ? PPM_ERR_DEVICE : PPM_OK;
}
