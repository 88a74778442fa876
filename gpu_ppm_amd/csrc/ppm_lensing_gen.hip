// Lensing tables on the device: the steps of c_lensing.py:47-155 (the
// reference runs them in Python with scipy.integrate.quad,
// scipy.interpolate.PchipInterpolator and numpy.gradient), one workgroup per
// u row:
//   - the alpha march of step pi/(8 n_cos) (c_lensing.py:84-99): chunks of
//     256 angles evaluated in parallel, until psi passes acos(cos_psi_min);
//     psi(alpha) from the bending integral (c_lensing.py:47-77) by adaptive
//     Gauss-Kronrod 7/15 bisection with the tolerance of the host generator;
//   - lensing factor d cos(alpha) / d cos(psi) by the non-uniform
//     second-order difference, first order at the ends (np.gradient,
//     c_lensing.py:100);
//   - cos(alpha) and the lensing factor resampled onto the uniform cos(psi)
//     axis by the monotone piecewise-cubic Hermite interpolant
//     (Fritsch-Carlson slopes, PchipInterpolator's end conditions,
//     c_lensing.py:102-110);
//   - c dt / R on the uniform cos(alpha) axis from the delay integral
//     (c_lensing.py:115-155).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/ppm.h"

namespace ppm {

namespace {

constexpr double kPiL = 3.14159265358979323846;
constexpr int kGenThreads = 256;
constexpr int kMaxDepth = 40;

// Kronrod abscissae (descending, last = 0) and weights of the 15-point rule;
// Gauss 7-point weights for abscissae xgk[1], xgk[3], xgk[5], xgk[7]
// (QUADPACK's qk15 tables).
__constant__ double kXGK[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
__constant__ double kWGK[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                               0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                               0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                               0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
__constant__ double kWG[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                              0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

// The two integrands on x in [0, 1] (c_lensing.py:47-54, 115-122), p = {u, sin^2 alpha}.
template <bool DELAY>
__device__ inline double integrand(double x, double u, double s2) {
  const double c2 = 1.0 - s2, x2 = x * x;
  const double q = (2.0 - x2 - u * (1.0 - x2) * (1.0 - x2) / (1.0 - u)) * s2;
  const double t = sqrt(c2 + x2 * q);
  return DELAY ? x / (t * (1.0 + t)) : x / t;
}

template <bool DELAY>
__device__ double gk15(double u, double s2, double a, double b, double& err) {
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  const double fc = integrand<DELAY>(c, u, s2);
  double rk = fc * kWGK[7], rg = fc * kWG[3];
  for (int j = 0; j < 7; ++j) {
    const double dx = h * kXGK[j];
    const double s = integrand<DELAY>(c - dx, u, s2) + integrand<DELAY>(c + dx, u, s2);
    rk += kWGK[j] * s;
    if (j & 1) rg += kWG[j / 2] * s;
  }
  err = fabs((rk - rg) * h);
  return rk * h;
}

// Adaptive bisection of [0, 1] until a piece's error estimate is within its
// share of the tolerance (halved per level, at most 40 levels), the pieces'
// results summed as the recursion left + right would: a post-order walk
// with an explicit stack.  A NaN estimate ends the piece.
template <bool DELAY>
__device__ double integrate01(double u, double s2) {
  struct Frame {
    double a, b, whole, err, tol, left, r_whole, r_err;
    int depth, stage;
  };
  Frame st[kMaxDepth + 1];
  double err0;
  const double whole0 = gk15<DELAY>(u, s2, 0.0, 1.0, err0);
  int top = 0;
  st[0] = Frame{0.0, 1.0, whole0, err0, fmax(1e-14, 1e-12 * fabs(whole0)), 0.0, 0.0, 0.0, 0, 0};
  double ret = 0.0;
  bool have_ret = false;
  for (;;) {
    Frame& F = st[top];
    if (have_ret) {  // a child finished
      have_ret = false;
      if (F.stage == 1) {
        F.left = ret;
        F.stage = 2;
        const double m = 0.5 * (F.a + F.b);
        st[top + 1] = Frame{m, F.b, F.r_whole, F.r_err, 0.5 * F.tol, 0.0, 0.0, 0.0, F.depth + 1, 0};
        ++top;
        continue;
      }
      ret = F.left + ret;  // stage 2: both halves done
    } else if (!(F.err > F.tol) || F.depth >= kMaxDepth) {  // accepted (or NaN): the piece's estimate
      ret = F.whole;
    } else {
      const double m = 0.5 * (F.a + F.b);
      double e1, e2;
      const double l = gk15<DELAY>(u, s2, F.a, m, e1);
      F.r_whole = gk15<DELAY>(u, s2, m, F.b, e2);
      F.r_err = e2;
      F.stage = 1;
      st[top + 1] = Frame{F.a, m, l, e1, 0.5 * F.tol, 0.0, 0.0, 0.0, F.depth + 1, 0};
      ++top;
      continue;
    }
    if (top == 0) return ret;
    --top;
    have_ret = true;
  }
}

__device__ double psi_out(double u, double alpha) {  // c_lensing.py:57-61
  const double s = sin(alpha);
  return 2.0 * s / sqrt(1.0 - u) * integrate01<false>(u, s * s);
}

// Turning point p / R of a photon emitted inward (c_lensing.py:65-67); NaN
// where none exists (beyond the march's end).
__device__ double turning_point_over_R(double u, double alpha) {
  const double t = 2.0 * sin(alpha) / sqrt(3.0 * (1.0 - u));
  return -t * cos((acos(3.0 * u / t) + 2.0 * kPiL) / 3.0);
}

__device__ double psi_of(double u, double alpha) {  // c_lensing.py:64-77
  if (alpha <= 0.5 * kPiL) return psi_out(u, alpha);
  const double pR = turning_point_over_R(u, alpha);
  if (!(pR > 0)) return INFINITY;
  return 2.0 * psi_out(u / pR, 0.5 * kPiL) - psi_out(u, kPiL - alpha);
}

__device__ double cdt_out(double u, double alpha) {  // c_lensing.py:125-129
  const double s = sin(alpha);
  return 2.0 * s * s / (1.0 - u) * integrate01<true>(u, s * s);
}

__device__ double cdt_of(double u, double alpha) {  // c_lensing.py:132-145
  if (alpha <= 0.5 * kPiL) return cdt_out(u, alpha);
  const double pR = turning_point_over_R(u, alpha);
  return 2.0 * cdt_out(u / pR, 0.5 * kPiL) - cdt_out(u, kPiL - alpha);
}

__device__ inline double sgn(double v) { return (v > 0) - (v < 0); }

__device__ inline double pchip_end_slope(double h0, double h1, double m0, double m1) {
  const double d = ((2.0 * h0 + h1) * m0 - h0 * m1) / (h0 + h1);
  if (sgn(d) != sgn(m0)) return 0.0;
  if (sgn(m0) != sgn(m1) && fabs(d) > 3.0 * fabs(m0)) return 3.0 * m0;
  return d;
}

struct GenArgs {
  double u_min, u_max, cp_min, cp_max, ca_min, ca_max;
  int n_u, n_cos, k_max;  // k_max: alpha steps available (the march ends before alpha = pi)
  const double* alpha;    // [k_max] the march's angles (accumulated steps)
  double* scratch;        // [n_u][5][k_max]: psi, cos psi, cos alpha, gradient, slopes
  double *ca, *lf, *dt;   // [n_u][n_cos] outputs
  int* status;            // [n_u]: 0 ok, 1 march did not pass psi_min
};

// The march's angles: alpha_0 = 0, alpha_{k+1} = alpha_k + pi / (8 n_cos),
// accumulated in order as the march does.
__global__ void k_alpha(double* alpha, int k_max, int n_cos) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double d = kPiL / (n_cos * 8.0);
    double a = 0.0;
    for (int k = 0; k < k_max; ++k) {
      alpha[k] = a;
      a += d;
    }
  }
}

__global__ __launch_bounds__(kGenThreads) void k_lensing_row(GenArgs A) {
  const int iu = blockIdx.x, tid = threadIdx.x;
  const int n_cos = A.n_cos, K = A.k_max;
  const double u = A.u_min + (A.u_max - A.u_min) * iu / (A.n_u - 1);
  const double psi_stop = acos(A.cp_min);
  double* ps = A.scratch + (size_t)iu * 5 * K;
  double* cpg = ps + K;
  double* cag = cpg + K;
  double* gs = cag + K;
  double* d = gs + K;
  __shared__ int s_stop;
  if (tid == 0) s_stop = K;
  __syncthreads();
  // ---- the march: the first angle whose psi passes psi_stop ends it
  int n = 0;
  for (int k0 = 0; k0 < K; k0 += kGenThreads) {
    const int k = k0 + tid;
    if (k < K) {
      ps[k] = psi_of(u, A.alpha[k]);
      if (ps[k] > psi_stop) atomicMin(&s_stop, k);
    }
    __syncthreads();
    const bool done = s_stop < K;
    __syncthreads();  // every thread has read s_stop before the next chunk may lower it
    if (done) break;
  }
  n = s_stop + 1;
  if (n > K) {
    if (tid == 0) A.status[iu] = 1;
    return;
  }
  if (tid == 0) A.status[iu] = 0;
  // ---- cos psi, cos alpha in march order; gradient (np.gradient, edge order 1)
  for (int k = tid; k < n; k += kGenThreads) {
    cpg[k] = cos(ps[k]);
    cag[k] = cos(A.alpha[k]);
  }
  __syncthreads();
  for (int k = tid; k < n; k += kGenThreads) {
    double g;
    if (k == 0) {
      g = (cag[1] - cag[0]) / (cpg[1] - cpg[0]);
    } else if (k == n - 1) {
      g = (cag[n - 1] - cag[n - 2]) / (cpg[n - 1] - cpg[n - 2]);
    } else {
      const double d1 = cpg[k] - cpg[k - 1], d2 = cpg[k + 1] - cpg[k];
      g = -d2 / (d1 * (d1 + d2)) * cag[k - 1] + (d2 - d1) / (d1 * d2) * cag[k] + d1 / (d2 * (d1 + d2)) * cag[k + 1];
    }
    gs[k] = g;
  }
  __syncthreads();
  // ---- two monotone cubic resamplings onto the ascending cos psi axis: the
  // data reversed (x_i = cpg[n-1-i]), y = cos alpha, then y = the gradient
  for (int pass = 0; pass < 2; ++pass) {
    const double* yv = pass == 0 ? cag : gs;
    double* out = (pass == 0 ? A.ca : A.lf) + (size_t)iu * n_cos;
    auto X = [&](int i) { return cpg[n - 1 - i]; };
    auto Y = [&](int i) { return yv[n - 1 - i]; };
    for (int i = tid; i < n; i += kGenThreads) {  // slopes at the nodes
      double di;
      if (i == 0 || i == n - 1) {
        const bool lo = i == 0;
        const int j0 = lo ? 0 : n - 2, j1 = lo ? 1 : n - 3;  // first and second interval from the end
        const double h0 = X(j0 + 1) - X(j0), h1 = X(j1 + 1) - X(j1);
        const double m0 = (Y(j0 + 1) - Y(j0)) / h0, m1 = (Y(j1 + 1) - Y(j1)) / h1;
        di = pchip_end_slope(h0, h1, m0, m1);
      } else {
        const double hp = X(i) - X(i - 1), hk = X(i + 1) - X(i);
        const double mp = (Y(i) - Y(i - 1)) / hp, mk = (Y(i + 1) - Y(i)) / hk;
        if (sgn(mk) != sgn(mp) || mk == 0.0 || mp == 0.0) {
          di = 0.0;
        } else {
          const double w1 = 2.0 * hk + hp, w2 = hk + 2.0 * hp;
          di = (w1 + w2) / (w1 / mp + w2 / mk);
        }
      }
      d[i] = di;
    }
    __syncthreads();
    for (int j = tid; j < n_cos; j += kGenThreads) {
      const double xi = A.cp_min + (A.cp_max - A.cp_min) * j / (n_cos - 1);
      int lo = 1, hi = n - 1;  // the first node >= xi among 1..n-1
      while (lo < hi) {
        const int m = (lo + hi) / 2;
        if (X(m) >= xi)
          hi = m;
        else
          lo = m + 1;
      }
      int k = lo - 1;
      if (k > n - 2) k = n - 2;
      const double h = X(k + 1) - X(k);
      const double t = (xi - X(k)) / h, t2 = t * t, t3 = t2 * t;
      const double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t, h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
      out[j] = h00 * Y(k) + h10 * h * d[k] + h01 * Y(k + 1) + h11 * h * d[k + 1];
    }
    __syncthreads();
  }
  // ---- c dt / R on the uniform cos alpha axis
  for (int j = tid; j < n_cos; j += kGenThreads) {
    const double ca = A.ca_min + (A.ca_max - A.ca_min) * j / (n_cos - 1);
    A.dt[(size_t)iu * n_cos + j] = cdt_of(u, acos(ca));
  }
}

}  // namespace

}  // namespace ppm

using namespace ppm;

extern "C" ppm_status ppm_lensing_build(const ppm_lensing_spec* sp, double* cos_alpha, double* lensing_factor,
                                        double* cdt_over_R) {
  if (!sp || !cos_alpha || !lensing_factor || !cdt_over_R) return PPM_ERR_ARG;
  const ppm_lensing_spec S = *sp;
  // single images need u < 0.568 (c_lensing.py:22-27 use 0.3-0.5)
  if (S.n_u < 2 || S.n_cos < 4 || S.n_cos > 16384 || !(S.u_min > 0) || !(S.u_min < S.u_max) || !(S.u_max < 0.568))
    return PPM_ERR_ARG;
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || S.device < 0 || S.device >= n_dev) return PPM_ERR_DEVICE;
  if (hipSetDevice(S.device) != hipSuccess) return PPM_ERR_DEVICE;
  const int K = 8 * S.n_cos;  // angles below pi
  const size_t n_tab = (size_t)S.n_u * S.n_cos;
  const size_t total = (size_t)K + (size_t)S.n_u * 5 * K + 3 * n_tab;
  double* base = nullptr;
  int* status = nullptr;
  if (hipMalloc(&base, sizeof(double) * total) != hipSuccess) return PPM_ERR_DEVICE;
  if (hipMalloc(&status, sizeof(int) * S.n_u) != hipSuccess) {
    (void)hipFree(base);
    return PPM_ERR_DEVICE;
  }
  GenArgs A{};
  A.u_min = S.u_min;
  A.u_max = S.u_max;
  A.cp_min = -0.9;  // c_lensing.py:12-16
  A.cp_max = 1.0;
  A.ca_min = -0.2;
  A.ca_max = 1.0;
  A.n_u = S.n_u;
  A.n_cos = S.n_cos;
  A.k_max = K;
  double* alpha = base;
  A.alpha = alpha;
  A.scratch = base + K;
  A.ca = A.scratch + (size_t)S.n_u * 5 * K;
  A.lf = A.ca + n_tab;
  A.dt = A.lf + n_tab;
  A.status = status;
  hipLaunchKernelGGL(k_alpha, dim3(1), dim3(64), 0, 0, alpha, K, S.n_cos);
  hipLaunchKernelGGL(k_lensing_row, dim3(S.n_u), dim3(kGenThreads), 0, 0, A);
  hipError_t e = hipGetLastError();
  std::vector<int> st(S.n_u, 0);
  if (!e) e = hipMemcpy(st.data(), status, sizeof(int) * S.n_u, hipMemcpyDeviceToHost);
  if (!e) e = hipMemcpy(cos_alpha, A.ca, sizeof(double) * n_tab, hipMemcpyDeviceToHost);
  if (!e) e = hipMemcpy(lensing_factor, A.lf, sizeof(double) * n_tab, hipMemcpyDeviceToHost);
  if (!e) e = hipMemcpy(cdt_over_R, A.dt, sizeof(double) * n_tab, hipMemcpyDeviceToHost);
  (void)hipFree(base);
  (void)hipFree(status);
  if (e) return PPM_ERR_DEVICE;
  for (int v : st)
    if (v) return PPM_ERR_ARG;  // a march did not pass cos psi_min before alpha = pi
  return PPM_OK;
}
