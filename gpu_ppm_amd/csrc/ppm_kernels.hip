// HIP kernels of the pulse-profile path for gfx950 (MI355X).
//
//   k_ring_setup      one workgroup per (spot, candidate ring): FP64 patch
//                     membership compacted into runs by wavefront ballot and
//                     prefix scan; the ring's two rows of the lensing tables
//                     staged in LDS for the bilinear look-ups; FP64
//                     fine-phase table of flux/I, redshift,
//                     cos(sigma') and time-delayed phase; ring-skip rule, gap
//                     fill, closure; NSX (logT, log g) bracket.
//   k_integral_plain  one workgroup per (ring, 128-energy chunk): NSX slice
//                     blended into LDS, per-patch set-up in FP64 by one lane
//                     each, FP32 cubic-Lagrange evaluation with two energies
//                     per lane (rows contracted over the mu stencil once per
//                     patch), per-bin Kahan accumulation in registers.
//   k_reduce          fixed-order FP64 sum of the per-ring partials.
//   k_fold            response fold on MFMA (split BF16, FP32 accumulate).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "ppm_eval.hpp"
#include "ppm_internal.hpp"

namespace ppm {

// Scratch of ppm_eval_device_noalloc: every device buffer of one plain
// evaluation beyond the caller's inputs and outputs (spot and ring tables,
// fine tables, per-ring partials, reduction groups, the fold's scale).  A
// file-scope array, so that entry point calls no allocator; each GPU the
// library runs on holds its own copy until the process ends.  In host code
// `scratch` is only a symbol: its device address comes from
// noalloc_scratch_address.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

int noalloc_scratch_address(int device, void** ptr) {
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// ------------------------------------------------------------------ helpers

// HEALPix ring i of N_side n: z, pixel count, 2*s (cpu/grid.hpp:106-131).
__device__ inline void healpix_ring(int n, int i, double& z, int& n_pix, int& s2) {
  const bool north = i <= 2 * n;
  const int ii = north ? i : 4 * n - i;
  double zz;
  if (ii < n) {
    zz = 1. - 1. * ii * ii / (3. * n * n);
    s2 = 2;
  } else {
    zz = 4. / 3. - 2. * ii / (3. * n);
    s2 = 2 * ((ii - n + 1) % 2);
  }
  z = north ? zz : -zz;
  int m = i < 4 * n - i ? i : 4 * n - i;
  if (m > n) m = n;
  n_pix = 4 * m;
}

__device__ inline double healpix_phi(int n, int i, int j) {  // cpu/grid.hpp:107-116
  const int ii = i <= 2 * n ? i : 4 * n - i;
  if (ii < n) return kPi / (2. * ii) * (j - 1. / 2.);
  const double s = (ii - n + 1) % 2;
  return kPi / (2. * n) * (j - s / 2.);
}

// Uniform axis with end clamp (cpu/hunt.hpp:155-178).
__device__ inline int uniform_hunt(double x_min, double x_max, int n, double x, double& a) {
  if (x <= x_min) {
    a = 0;
    return 0;
  }
  if (x >= x_max) {
    a = 1;
    return n - 2;
  }
  const double t = (x - x_min) / (x_max - x_min) * (n - 1);
  const int i = (int)t;
  a = t - i;
  return i;
}

// Stencil of one NSX table axis at x (FP64): the bracket's two linear
// weights (cpu/nsx.hpp:277-280), or the 4-point Lagrange weights of the nodes
// from one below the bracket, clamped to [0, n-4], at x clamped to the axis
// (cpu/nsx.hpp:159-167, 172-186).  Returns the first node.
__device__ inline int axis_stencil(double x_min, double x_max, int n, double x, int n_w, double* w) {
  double a;
  const int i = uniform_hunt(x_min, x_max, n, x, a);
  if (n_w == 2) {
    w[0] = 1 - a;
    w[1] = a;
    return i;
  }
  x = fmin(fmax(x, x_min), x_max);
  const int s = clampi(i - 1, 0, n - 4);
  double xs[4];
  for (int k = 0; k < 4; ++k) xs[k] = x_min + (x_max - x_min) * (s + k) / (n - 1);  // cpu/hunt.hpp:180-187
  for (int k = 0; k < 4; ++k) {
    double num = 1, den = 1;
    for (int l = 0; l < 4; ++l)
      if (l != k) {
        num *= x - xs[l];
        den *= xs[k] - xs[l];
      }
    w[k] = num / den;
  }
  return s;
}

// The two rows (iu, iu + 1) of the three lensing tables that one ring's
// bilinear look-ups read (every element of a ring has the same u), staged
// in LDS by the ring's workgroup: [2][n_cp] cos(alpha), [2][n_cp] lensing
// factor, [2][n_ca] c dt / R.
struct LensRows {
  const double *ca, *lf, *dt;
  double au;  // fraction of u between rows iu and iu + 1
};

__device__ inline LensRows stage_lens_rows(const DevLensing& L, double u, double* s, int tid, int nthreads) {
  LensRows r;
  const int iu = uniform_hunt(L.u_min, L.u_max, L.n_u, u, r.au);  // cpu/lensing.hpp:43-72
  double* ca = s;
  double* lf = s + 2 * L.n_cp;
  double* dt = s + 4 * L.n_cp;
  const double* gca = L.ca + (size_t)iu * L.n_cp;
  const double* glf = L.lf + (size_t)iu * L.n_cp;
  const double* gdt = L.dt + (size_t)iu * L.n_ca;
  for (int k = tid; k < 2 * L.n_cp; k += nthreads) {
    ca[k] = gca[k];
    lf[k] = glf[k];
  }
  for (int k = tid; k < 2 * L.n_ca; k += nthreads) dt[k] = gdt[k];
  r.ca = ca;
  r.lf = lf;
  r.dt = dt;
  return r;
}

__device__ inline void lens_ca_lf(const DevLensing& L, const LensRows& R, double cp, double& ca, double& lf) {
  double ac;  // cpu/lensing.hpp:43-60
  const int ic = uniform_hunt(L.cp_min, L.cp_max, L.n_cp, cp, ac);
  const double au = R.au, bu = 1. - au, bc = 1. - ac;
  const double* A = R.ca;
  const double* F = R.lf;
  const int n = L.n_cp;
  ca = bu * bc * A[ic] + au * bc * A[n + ic] + bu * ac * A[ic + 1] + au * ac * A[n + ic + 1];
  lf = bu * bc * F[ic] + au * bc * F[n + ic] + bu * ac * F[ic + 1] + au * ac * F[n + ic + 1];
}

__device__ inline double lens_cdt(const DevLensing& L, const LensRows& R, double ca) {
  double ac;  // cpu/lensing.hpp:62-72
  const int ic = uniform_hunt(L.ca_min, L.ca_max, L.n_ca, ca, ac);
  const double au = R.au, bu = 1. - au, bc = 1. - ac;
  const double* D = R.dt;
  const int n = L.n_ca;
  return bu * bc * D[ic] + au * bc * D[n + ic] + bu * ac * D[ic + 1] + au * ac * D[n + ic + 1];
}

// FP64 patch membership (cpu/grid.hpp:144-148 CAP, :158-190 DISKDIFF).
// Contraction is off so the decisions round like the oracle's.
__device__ inline bool in_spot(const SpotParams& S, double c2, double s2, double phi) {
#pragma clang fp contract(off)
  const double r1 = S.st1 * s2 * cos(phi - S.phi1) + S.ct1 * c2;
  if (!(r1 > S.cr1)) return false;
  if (!S.has2) return true;
  const double r2 = S.st2 * s2 * cos(phi - S.phi2) + S.ct2 * c2;
  return !(r2 > S.cr2);
}

__device__ inline int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_min_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_max_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ inline uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// ------------------------------------------------------------ ring set-up

constexpr int kMaskWords = 8192 / 64 * 4;  // 4 * 8192 pixels max per ring

__global__ __launch_bounds__(kBlock) void k_ring_setup(DevLensing L, DevNsx N, PlainLaunch P) {
  const int task = blockIdx.x;
  const RingTask T = P.d_tasks[task];
  const SpotParams S = P.d_spots[T.spot];
  RingDesc* D = P.d_desc + task;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nf = P.n_fine, nf1 = nf + 1;

  __shared__ uint64_t s_mask[kMaskWords];
  __shared__ int s_cnt[kWaves * 2];
  __shared__ int s_int[kWaves * 2];
  __shared__ double s_dbl[kWaves * 2];
  __shared__ int s_base[2];

  double z, d_omega;
  int n_pix, s2;
  if (T.ring == 0) {  // point-source element (cpu/grid.hpp:17-27)
    z = S.ps_z;
    d_omega = S.ps_domega;
    n_pix = 1;
    s2 = 0;
  } else {
    healpix_ring(S.n_side, T.ring, z, n_pix, s2);
    d_omega = kPi / (3. * S.n_side * S.n_side);
  }
  if (tid == 0) {
    D->z = z;
    D->d_omega = d_omega;
    D->n_pix = n_pix;
    D->s_shift2 = s2;
    D->ps_shift = S.ps_shift;
    D->flags = 0;
    s_base[0] = s_base[1] = 0;
  }
  __syncthreads();

  // ---- membership, ballot-compacted into runs of consecutive pixels
  int n_act = 0, n_runs = 0;
  if (T.ring == 0) {
    n_act = 1;
    n_runs = 0;
  } else {
    double c2, sz;
    if (S.membership == PPM_MEMBER_CAP) {
      c2 = z;
      sz = sqrt(1. - z * z);
    } else {
      const double th = acos(z);
      c2 = cos(th);
      sz = sin(th);
    }
    const int n_iter = (n_pix + kBlock - 1) / kBlock;
    int cnt = 0;
    for (int it = 0; it < n_iter; ++it) {
      const int j = 1 + it * kBlock + tid;
      const bool act = j <= n_pix && in_spot(S, c2, sz, healpix_phi(S.n_side, T.ring, j));
      const uint64_t m = __ballot(act);
      if (lane == 0) s_mask[it * kWaves + wave] = m;
      cnt += __popcll(m);
    }
    if (lane == 0) s_cnt[wave] = cnt;
    __syncthreads();
    n_act = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    // run starts/ends from neighbouring bits, placed in pixel order
    for (int it = 0; it < n_iter; ++it) {
      const int q = it * kBlock + tid;  // 0-based pixel
      auto bit = [&](int k) -> bool {
        return k >= 0 && k < n_pix && ((s_mask[k >> 6] >> (k & 63)) & 1ull);
      };
      const bool a = bit(q);
      const bool st = a && !bit(q - 1), en = a && !bit(q + 1);
      const uint64_t ms = __ballot(st), me = __ballot(en);
      if (lane == 0) {
        s_int[wave] = __popcll(ms);
        s_int[kWaves + wave] = __popcll(me);
      }
      __syncthreads();
      int ps = s_base[0], pe = s_base[1];
      for (int w = 0; w < wave; ++w) {
        ps += s_int[w];
        pe += s_int[kWaves + w];
      }
      if (st) {
        const int k = ps + __popcll(ms & lanemask_lt());
        if (k < kMaxRuns) D->run_start[k] = q + 1;
      }
      if (en) {
        const int k = pe + __popcll(me & lanemask_lt());
        if (k < kMaxRuns) D->run_len[k] = q + 1;  // end pixel for now
      }
      __syncthreads();
      if (tid == 0) {
        for (int w = 0; w < kWaves; ++w) {
          s_base[0] += s_int[w];
          s_base[1] += s_int[kWaves + w];
        }
      }
      __syncthreads();
    }
    n_runs = s_base[0];
    __syncthreads();
    if (tid < kMaxRuns && tid < n_runs) D->run_len[tid] = D->run_len[tid] - D->run_start[tid] + 1;
    if (tid == 0 && n_runs > kMaxRuns) D->flags |= 1;
  }
  if (tid == 0) {
    D->n_act = n_act;
    D->n_runs = n_runs < kMaxRuns ? n_runs : kMaxRuns;
    if (P.summary) {
      atomicAdd(&P.summary[0], (unsigned long long)n_act);
      if (n_act) atomicAdd(&P.summary[2], 1ull);
      if (n_runs > kMaxRuns) atomicOr(&P.summary[3], 1ull);
    }
  }
  if (n_act == 0) {
    if (tid == 0) D->kept = 0;
    return;
  }

  // ---- fine-phase table in FP64 (cpu/os1_cpu.cpp:62-167, cpu/devarshi24st_cpu.cpp:89-212)
  const double cz = z, sz = sqrt(1. - cz * cz);
  const double R = S.Re * (1. + S.o2 * cz * cz);
  const double dRdth = -2. * S.Re * S.o2 * cz * sz;
  const double u = S.Rs / R, uu = sqrt(1. - u);
  extern __shared__ __align__(16) unsigned char setup_smem[];
  const LensRows LR = stage_lens_rows(L, u, reinterpret_cast<double*>(setup_smem), tid, kBlock);
  __syncthreads();
  const double f = dRdth / (R * uu), ff = sqrt(1. + f * f);
  const double dS = R * R * d_omega * ff;
  const double cos_tau = 1. / ff, sin_tau = f / ff;
  const double beta = 2 * kPi * S.nu * R * sz / uu / kCKmS;
  const double gamma = 1. / sqrt(1. - beta * beta);
  const double lgR = log1p((S.Re - R) / (R - S.Rs));
  const double dtR = (S.Re - R + S.Rs * lgR) / kCKmS;

  double* ph = P.d_ph + (size_t)task * nf1;
  float* fI = P.d_fine + (size_t)task * 3 * nf1;
  float* zf = fI + nf1;
  float* cs = zf + nf1;

  int inv_min = nf + 1, inv_max = -1;
  for (int k = tid; k < nf; k += kBlock) {
    bool vis = true;
    double v_fl = 0, v_zf = -1, v_ph = -1, v_cs = 0;
    do {
      const double phase = 1. * k / nf;
      const double sp = 2 * kPi * phase;
      const double cphi = cos(sp), sphi = sin(sp);
      double cos_psi, cos_sigma;
      {
#pragma clang fp contract(off)
        cos_psi = S.ci * cz + S.si * sz * cphi;
      }
      if (cos_psi < L.cp_min) {
        vis = false;
        break;
      }
      const double sin_psi = sqrt(1. - cos_psi * cos_psi);
      double ca, lf;
      lens_ca_lf(L, LR, cos_psi, ca, lf);
      const double sa = sqrt(1. - ca * ca);
      const double sa_sp = cos_psi == 1. ? sqrt(lf) : sa / sin_psi;
      {
#pragma clang fp contract(off)
        if (S.variant == PPM_VARIANT_OS1)
          cos_sigma = ca * cos_tau + sa_sp * sin_tau * (S.ci - cz * cos_psi) / sz;
        else
          cos_sigma = ca * cos_tau + sa_sp * sin_tau * (S.ci * sz - S.si * cz * cphi);
      }
      if (cos_sigma <= 0) {
        vis = false;
        break;
      }
      const double cos_xi = -sa_sp * S.si * sphi;
      const double delta = 1. / (gamma * (1. - beta * cos_xi));
      const double delta3 = delta * delta * delta;
      const double dt1 = lens_cdt(L, LR, ca) * R / kCKmS;
      double dt2;
      if (ca >= 0) {
        dt2 = dtR;
      } else {
        const double tmp = (2. * sa) / sqrt(3. * (1. - u));
        const double p = -tmp * cos((acos(3. * u / tmp) + 2. * kPi) / 3.) * R;
        const double lgp = log1p((S.Re - p) / (p - S.Rs));
        dt2 = 2. * ((S.Re - p + S.Rs * lgp) / kCKmS) - dtR;
      }
      const double csp = cos_sigma * delta;
      double beam = 1.0;
      if (S.atmosphere == PPM_ATM_BLACKBODY) {
        if (S.beaming == PPM_BEAM_SIN2)
          beam = 1. - csp * csp;
        else if (S.beaming == PPM_BEAM_COS2)
          beam = csp * csp;
      }
      // 1/D^2 and the intensity constants are applied in FP64 at the reduction
      v_fl = uu * delta3 * csp * lf * (dS * gamma) * beam;
      v_zf = 1. / (delta * uu);
      v_ph = phase + (dt1 + dt2) * S.nu;
      v_cs = csp;
    } while (0);
    if (!vis) {
      inv_min = min(inv_min, k);
      inv_max = max(inv_max, k);
    }
    ph[k] = v_ph;
    fI[k] = (float)v_fl;
    zf[k] = (float)v_zf;
    cs[k] = (float)v_cs;
  }
  inv_min = wave_min(inv_min);
  inv_max = wave_max(inv_max);
  if (lane == 0) {
    s_int[wave] = inv_min;
    s_int[kWaves + wave] = inv_max;
  }
  __syncthreads();
  inv_min = min(min(s_int[0], s_int[1]), min(s_int[2], s_int[3]));
  inv_max = max(max(s_int[4], s_int[5]), max(s_int[6], s_int[7]));
  // ring skip (cpu/os1_cpu.cpp:151)
  if (inv_min == 0 || inv_max == nf - 1) {
    if (tid == 0) D->kept = 0;
    return;
  }
  if (tid == 0) D->hidden = inv_min <= inv_max ? kHiddenLookup : kHiddenNone;
  if (inv_min <= inv_max) {  // gap fill (cpu/os1_cpu.cpp:152-163)
    const int beg = inv_min - 1, end = inv_max + 1;
    const double p0 = ph[beg], p1 = ph[end];
    const double r0 = zf[beg], r1 = zf[end];
    __syncthreads();
    for (int k = inv_min + tid; k <= inv_max; k += kBlock) {
      ph[k] = p0 + (p1 - p0) * ((double)(k - beg) / (end - beg));
      zf[k] = (float)(r0 + (r1 - r0) * ((double)(k - beg) / (end - beg)));
    }
  }
  __syncthreads();
  if (tid == 0) {  // closure (cpu/os1_cpu.cpp:165-167)
    fI[nf] = fI[0];
    zf[nf] = zf[0];
    ph[nf] = 1.0 + ph[0];
    cs[nf] = cs[0];
  }
  __syncthreads();
  if (inv_min <= inv_max) {
    // With every node of the gap invisible and the phase table strictly
    // ascending, flux/I interpolated at phase pp is zero exactly when
    // ph[inv_min] <= pp <= ph[inv_max] (inside, both bracket nodes are zero
    // or the fraction is; outside, the bracket has a visible node with a
    // nonzero weight, up to a fraction that rounds to 1 next to the gap)
    bool ok = true;
    for (int k = tid; k < nf1; k += kBlock) {
      if (k >= inv_min && k <= inv_max) ok = ok && fI[k] == 0.f;
      if (k < nf) ok = ok && ph[k + 1] > ph[k];
    }
    ok = __all(ok);
    if (lane == 0) s_int[wave] = ok;  // (s_int was last read before the gap fill's barrier)
    __syncthreads();
    for (int w = 0; w < kWaves; ++w) ok = ok && s_int[w] != 0;
    if (tid == 0 && ok) {
      D->hidden = kHiddenInterval;
      D->gap_lo = ph[inv_min];
      D->gap_hi = ph[inv_max];
    }
  }
  double zmin = 1e300, zmax = -1e300;
  for (int k = tid; k < nf1; k += kBlock) {
    zmin = fmin(zmin, (double)zf[k]);
    zmax = fmax(zmax, (double)zf[k]);
  }
  zmin = wave_min_d(zmin);
  zmax = wave_max_d(zmax);
  if (lane == 0) {
    s_dbl[wave] = zmin;
    s_dbl[kWaves + wave] = zmax;
  }
  __syncthreads();
  if (tid == 0) {
    D->zf_min = fmin(fmin(s_dbl[0], s_dbl[1]), fmin(s_dbl[2], s_dbl[3]));
    D->zf_max = fmax(fmax(s_dbl[4], s_dbl[5]), fmax(s_dbl[6], s_dbl[7]));
    // surface gravity (cpu/devarshi24st_cpu.cpp:63-68, 103-124)
    double logg = S.logg_override;
    if (!(logg > 0)) {
      const double x = S.x;
      const double ce = -0.791 + 0.776 * x, cp = 1.138 + -1.431 * x;
      const double de = -1.315 * x + 2.431 * x * x, fe = -1.172 * x;
      const double dp = 0.653 * x + -2.864 * x * x, fp = 0.975 * x;
      const double d60 = 13.47 * x + -27.13 * x * x;
      const double Ob2 = S.Ob * S.Ob, Ob4 = Ob2 * Ob2, Ob6 = Ob4 * Ob2;
      const double g_g0 = 1. + sz * sz * (ce * Ob2 + de * Ob4 + fe * Ob6) +
                          cz * cz * (cp * Ob2 + dp * Ob4 + fp * Ob6 - d60 * Ob4) +
                          fabs(cz) * d60 * Ob4;
      logg = log10(g_g0) + S.logg0_total;
    }
    D->logg = logg;
    double wT[4] = {0, 0, 0, 0}, wg[4] = {0, 0, 0, 0};
    D->iT = D->ig = 0;
    D->n_tg = 2;
    if (N.data) {
      const int n_tg = nsx_cubic_tg(S.nsx_interp) ? 4 : 2;
      D->iT = axis_stencil(N.logT_min, N.logT_max, N.nT, S.logT, n_tg, wT);
      D->ig = axis_stencil(N.logg_min, N.logg_max, N.ng, logg, n_tg, wg);
      D->n_tg = n_tg;
    }
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) D->wtg[4 * a + b] = (float)(wT[a] * wg[b]);
    D->kept = 1;
    if (P.summary) atomicAdd(&P.summary[1], (unsigned long long)n_act);
  }
}

// ---------------------------------------------------------- plain integral

// Byte offsets of the plain integral kernel's dynamic LDS.
struct LdsLayout {
  size_t ph, mu, fI, zf, cs, S, C, total;
};
__host__ __device__ inline LdsLayout integral_layout(int nf1, int nsx_rows) {
  LdsLayout l;
  size_t o = 0;
  l.ph = o;
  o += (size_t)nf1 * sizeof(double);
  l.mu = o;
  o += 80 * sizeof(double);
  l.fI = o;
  o += (size_t)nf1 * sizeof(float);
  l.zf = o;
  o += (size_t)nf1 * sizeof(float);
  l.cs = o;
  o += (size_t)nf1 * sizeof(float);
  o = (o + 15) & ~(size_t)15;
  l.S = o;
  o += (size_t)nsx_rows * kSliceStride * sizeof(float);
  l.C = o;  // per-wavefront evaluation scratch
  o += (size_t)kWaves * kWaveScratchBytes;
  l.total = (o + 15) & ~(size_t)15;
  return l;
}

constexpr int kEPL = 2;                     // energies per lane
constexpr int kEPerWave = 64 * kEPL;        // energies per wavefront / workgroup
constexpr int kNP = 4;                      // patches per round of the per-energy path

// One workgroup per (ring, 128-energy chunk); its 4 wavefronts take the
// phase bins round-robin.  Within a wavefront each lane sets up one patch in
// FP64 and keeps it in registers; the evaluation loop broadcasts a patch's
// values with readlane, contracts its mu stencil over one slice row per lane
// and evaluates every lane's two energies from the wavefront's table of cubic
// coefficients (nsx_eval_table), as the batched kernel does.
template <int ATM, int EM>
__global__ __launch_bounds__(kBlock) void k_integral_plain(DevLensing L, DevNsx N, PlainLaunch P) {
  const int task = blockIdx.x, chunk = blockIdx.y;
  RingDesc* Dp = P.d_desc + task;
  if (!Dp->kept) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nf1 = P.n_fine + 1;
  const RingTask T = P.d_tasks[task];
  const SpotParams& S = P.d_spots[T.spot];

  extern __shared__ __align__(16) unsigned char smem[];
  const int rows_alloc = ATM == PPM_ATM_NSX ? P.nsx_rows : 0;
  const LdsLayout lay = integral_layout(nf1, rows_alloc);
  double* s_ph = reinterpret_cast<double*>(smem + lay.ph);
  double* s_mu = reinterpret_cast<double*>(smem + lay.mu);
  float* s_fI = reinterpret_cast<float*>(smem + lay.fI);
  float* s_zf = reinterpret_cast<float*>(smem + lay.zf);
  float* s_cs = reinterpret_cast<float*>(smem + lay.cs);
  float* s_S = reinterpret_cast<float*>(smem + lay.S);
  float4* s_C = reinterpret_cast<float4*>(smem + lay.C) + wave * (kWaveScratchBytes / 16);

  {
    const double* gph = P.d_ph + (size_t)task * nf1;
    const float* gf = P.d_fine + (size_t)task * 3 * nf1;
    for (int k = tid; k < nf1; k += kBlock) {
      s_ph[k] = gph[k];
      s_fI[k] = gf[k];
      s_zf[k] = gf[nf1 + k];
      s_cs[k] = gf[2 * nf1 + k];
    }
  }
  // energies of this chunk and the local energy axis
  const int e_first = chunk * kEPerWave;
  const int e_last = min(e_first + kEPerWave, P.n_e) - 1;
  const double h = (N.logE_max - N.logE_min) / (N.nE - 1);
  const double inv_h = 1.0 / h;
  const double l_ref = log10(P.d_energies[e_first]);
  int r_lo = 0, n_rows = 0;
  EAxis ax{0.f, 0.f, 0, 0, 1};
  if (ATM == PPM_ATM_NSX) {
    const double lkT = log10(S.kT_nsx);
    const double t_min = (l_ref + log10(Dp->zf_min) - lkT - N.logE_min) * inv_h;
    const double t_max = (log10(P.d_energies[e_last]) + log10(Dp->zf_max) - lkT - N.logE_min) * inv_h;
    r_lo = clampi((int)floor(t_min - 0.01) - 1, 0, N.nE - 4);
    const int r_hi = clampi((int)floor(t_max + 0.01) - 1, 0, N.nE - 4) + 3;
    n_rows = r_hi - r_lo + 1;
    if (n_rows > rows_alloc) {  // redshift range beyond the staged rows: flagged, host re-runs
      if (tid == 0) atomicOr(&Dp->flags, 2);
      n_rows = rows_alloc;
    }
    ax = energy_axis<EM>(N.nE, r_lo, n_rows, rows_alloc);
    for (int k = tid; k < N.nm; k += kBlock) s_mu[k] = N.mu[k];
    stage_nsx_rows(N, *Dp, r_lo, n_rows, s_S, tid, kBlock);
  }
  __syncthreads();
  const TxMap txm{(float)((l_ref - log10(S.kT_nsx) - N.logE_min) * inv_h - r_lo),
                  (float)(inv_h * 0.30102999566398120)};

  float le[kEPL], Ef[kEPL];
  bool ok[kEPL];
#pragma unroll
  for (int u = 0; u < kEPL; ++u) {
    const int e = e_first + u * 64 + lane;
    ok[u] = e <= e_last;
    const double E = P.d_energies[ok[u] ? e : e_last];
    Ef[u] = (float)E;
    le[u] = (float)((log10(E) - l_ref) * inv_h);
  }
  const bool asc = energies_ascend<kEPL>(le, lane);
  const WaveEnergies<1> we = wave_energies<kEPL, 1>(le, lane);
  const float inv_kT = (float)S.inv_kT;
  const int n_act = Dp->n_act, n_runs = Dp->n_runs, n_pix = Dp->n_pix, s2 = Dp->s_shift2;
  const bool point = T.ring == 0;
  const double ps_shift = Dp->ps_shift;
  const RingTab R{s_ph, s_fI, s_zf, s_cs, nf1};
  const int ng = P.n_pts / P.n_phase;
  const int mu_mode = nsx_mu_mode(S.nsx_interp);

  for (int b = wave; b < P.n_phase; b += kWaves) {
    float acc[kEPL], comp[kEPL];
#pragma unroll
    for (int u = 0; u < kEPL; ++u) acc[u] = comp[u] = 0.f;
    for (int g = 0; g < ng; ++g) {
      const PhasePoint pt = P.d_pts[b * ng + g];
      for (int base = 0; base < n_act; base += 64) {
        // ---- patch set-up, one patch per lane (FP64)
        PatchVals pv;
        pv.im = -1;
        const int k = base + lane;
        if (k < n_act) {
          double shift;
          if (point) {
            shift = ps_shift;
          } else {
            int kk = k, j = 0;
            for (int q = 0; q < n_runs; ++q) {
              const int len = Dp->run_len[q];
              if (kk < len) {
                j = Dp->run_start[q] + kk;
                break;
              }
              kk -= len;
            }
            shift = (j - 0.25 * s2) / n_pix;
          }
          // time-delayed phase (cpu/devarshi24st_cpu.cpp:218-224)
          const double v = pt.p + shift + 1.;
          double pp = v - floor(v);
          if (pp < s_ph[0]) pp += 1.;
          patch_setup<ATM>(pp, pt.w, R, s_mu, N.mu_den, N.nm, N.mu_min, mu_mode, txm, pv);
        }
        // ---- evaluation: every lane evaluates each visible patch at its
        // energies
        float bsum[kEPL];
#pragma unroll
        for (int u = 0; u < kEPL; ++u) bsum[u] = 0.f;
        uint64_t todo = __ballot(pv.im >= 0);
        PatchRows<1> pr{};
        if (ATM == PPM_ATM_NSX && pv.im >= 0) pr = patch_rows(pv.tx, pv.im, we, ax);
        const uint64_t interior = __ballot(pv.im >= 0 && pr.clamp == 0);
        if (ATM == PPM_ATM_NSX && we.fast) {
          // one patch at a time through the wavefront's coefficient table;
          // the slice rows of the next patch are read while the current one
          // is evaluated (two row buffers, used in turn)
          auto step = [&](const float (&p_cur)[1][4], float (&p_next)[1][4]) {
            const int q = __builtin_ctzll(todo);
            todo &= todo - 1;
            if (todo) nsx_load_rows(pr, __builtin_ctzll(todo), s_S, lane, p_next);
            float val[kEPL];
            if (__builtin_expect((interior >> q) & 1, 1))
              nsx_eval_table<kEPL, 0, 1, EM>(le, pv, pr, q, ax, p_cur, lane, s_C, val);
            else
              nsx_eval_table<kEPL, 1, 1, EM>(le, pv, pr, q, ax, p_cur, lane, s_C, val);
#pragma unroll
            for (int u = 0; u < kEPL; ++u) bsum[u] += val[u];
          };
          float p_a[1][4], p_b[1][4];
          if (todo) nsx_load_rows(pr, __builtin_ctzll(todo), s_S, lane, p_a);
          while (todo) {
            step(p_a, p_b);
            if (!todo) break;
            step(p_b, p_a);
          }
        }
        // energies whose rows exceed the table (NSX), or the blackbody:
        // kNP patches per round
        while (todo) {
          int qs[kNP];  // -1: no patch
#pragma unroll
          for (int j = 0; j < kNP; ++j) {
            qs[j] = todo ? __builtin_ctzll(todo) : -1;
            todo &= todo - 1;
          }
#pragma unroll
          for (int j = 0; j < kNP; ++j) {
            if (qs[j] < 0) break;
            if (ATM == PPM_ATM_NSX) {
              float w[4], val[kEPL];
#pragma unroll
              for (int m = 0; m < 4; ++m) w[m] = readlane_f(pv.w[m], qs[j]);
              nsx_eval_edge<kEPL, EM>(le, asc, readlane_f(pv.tx, qs[j]), readlane_i(pv.im, qs[j]), w, ax, s_S, s_C,
                                      lane, val);
#pragma unroll
              for (int u = 0; u < kEPL; ++u) bsum[u] += val[u];
            } else {
              const float c = readlane_f(pv.w[0], qs[j]), z = readlane_f(pv.zf, qs[j]);
#pragma unroll
              for (int u = 0; u < kEPL; ++u) {
                const float Ee = Ef[u] * z;
                bsum[u] = fmaf(c, Ee * Ee * Ee / expm1f(Ee * inv_kT), bsum[u]);
              }
            }
          }
        }
        // Kahan-compensated per-bin accumulation
#pragma unroll
        for (int u = 0; u < kEPL; ++u) {
          const float y = bsum[u] - comp[u];
          const float tt = acc[u] + y;
          comp[u] = (tt - acc[u]) - y;
          acc[u] = tt;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kEPL; ++u)
      if (ok[u]) P.d_partial[((size_t)task * P.n_phase + b) * P.n_e + e_first + u * 64 + lane] = acc[u] - comp[u];
  }
}

// ------------------------------------------------------------------ reduce

// Stage 1: group g sums its contiguous range of ring tasks, each term scaled
// by its spot's FP64 constant.  Stage 2 sums the groups in order and divides
// by E (cpu/devarshi24st_cpu.cpp:238).  Fixed order: deterministic.
__global__ void k_reduce_groups(PlainLaunch P, double* red, int n_groups) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int n_out = P.n_phase * P.n_e;
  if (idx >= n_out) return;
  const int g = blockIdx.y;
  const int t0 = (int)((long long)P.n_tasks * g / n_groups), t1 = (int)((long long)P.n_tasks * (g + 1) / n_groups);
  double total = 0.0;
  for (int t = t0; t < t1; ++t)
    if (P.d_desc[t].kept) total += P.d_spots[P.d_tasks[t].spot].scale * (double)P.d_partial[(size_t)t * n_out + idx];
  red[(size_t)g * n_out + idx] = total;
}

__global__ void k_reduce_final(PlainLaunch P, const double* red, int n_groups) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int n_out = P.n_phase * P.n_e;
  if (idx >= n_out) return;
  double total = 0.0;
  for (int g = 0; g < n_groups; ++g) total += red[(size_t)g * n_out + idx];
  P.d_flux[idx] = total / P.d_energies[idx % P.n_e];
}

// ------------------------------------------------------------------- fold

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ inline uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// counts[b][c][p] = scale[b] * sum_e rsp[c][e] * flux[b][p][e]; one wave per
// (16-phase tile, star, group of kFoldTiles 16-channel tiles), K in steps of
// 16 energies: the flux tile is read from HBM once per channel group and
// split into BF16 hi/lo once per K step, the response (L2-resident) is
// streamed per channel tile.
// v_mfma_f32_16x16x16bf16_1k lane layout: A[m = l%16][k = 4(l/16)+i],
// B[k = 4(l/16)+i][n = l%16], D[m = 4(l/16)+i][n = l%16].
constexpr int kFoldTiles = 10;

__global__ __launch_bounds__(64) void k_fold(const uint16_t* __restrict__ rsp_hi,
                                             const uint16_t* __restrict__ rsp_lo, int n_ch,
                                             int n_e, const double* __restrict__ flux, int n_phase,
                                             const double* __restrict__ scale,
                                             double* __restrict__ counts, FoldFit fit, int star0) {
  const int p0 = blockIdx.x * 16, b = blockIdx.y, t0 = blockIdx.z * kFoldTiles;
  const int l = threadIdx.x, row = l & 15, kq = (l >> 4) * 4;
  const int ph = p0 + row;
  const int n_tiles = min(kFoldTiles, (n_ch + 15) / 16 - t0);
  const double* fb = flux + ((size_t)b * n_phase + ph) * n_e;
  float4_t acc[kFoldTiles];
#pragma unroll
  for (int t = 0; t < kFoldTiles; ++t) acc[t] = float4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < n_e; k0 += 16) {
    short4_t bh, bl;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = k0 + kq + i;
      const float fv = (ph < n_phase && e < n_e) ? (float)fb[e] : 0.f;
      const uint16_t fh = bf16_rne(fv);
      bh[i] = (short)fh;
      bl[i] = (short)bf16_rne(fv - bf16_to_f(fh));
    }
#pragma unroll
    for (int t = 0; t < kFoldTiles; ++t) {
      const int ch = (t0 + t) * 16 + row;  // tiles past n_ch read zeros
      short4_t ah, al;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int e = k0 + kq + i;
        const bool ok = t < n_tiles && ch < n_ch && e < n_e;
        ah[i] = ok ? (short)rsp_hi[(size_t)ch * n_e + e] : (short)0;
        al[i] = ok ? (short)rsp_lo[(size_t)ch * n_e + e] : (short)0;
      }
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc[t], 0, 0, 0);
    }
  }
  const double sc = scale[b];
  // epilogue: the model counts, and with data the partial sums of the fit
  // statistics over this wave's tile (FP64, wave reduction in fixed order)
  const double* dat = fit.data ? fit.data + (size_t)(star0 + b) * fit.stride : nullptr;
  double s_m = 0.0, s_chi = 0.0, s_ll = 0.0;
#pragma unroll
  for (int t = 0; t < kFoldTiles; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c = (t0 + t) * 16 + kq + i, p = p0 + row;
      if (t < n_tiles && c < n_ch && p < n_phase) {
        const double m = sc * (double)acc[t][i];
        if (counts) counts[((size_t)b * n_ch + c) * n_phase + p] = m;
        if (dat) {
          const double d = dat[(size_t)c * n_phase + p], r = d - m;
          s_m += m;
          if (m != 0.0 || d != 0.0) {  // an empty bin of an empty model adds 0 (the limit m -> 0)
            s_chi += r * r / m;
            s_ll += (d == 0.0 ? 0.0 : d * log(m)) - m - lgamma(d + 1.0);
          }
        }
      }
    }
  if (dat) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      s_m += __shfl_xor(s_m, o);
      s_chi += __shfl_xor(s_chi, o);
      s_ll += __shfl_xor(s_ll, o);
    }
    if (l == 0) {
      double* out = fit.part + ((size_t)(star0 + b) * fit.blocks + blockIdx.z * gridDim.x + blockIdx.x) * 3;
      out[0] = s_m;
      out[1] = s_chi;
      out[2] = s_ll;
    }
  }
}

// Fit statistics per star from the fold's partial sums (fixed order):
// chi2 of cal_chi2 (p_devarshi24st.py:43-49: model and data scaled so the
// model sums to 1e6, i.e. 1e6 / sum(m) * sum (d - m)^2 / m), the Poisson
// log-likelihood and the model total.
__global__ void k_fit_reduce(const double* __restrict__ part, int blocks, int n, double* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  double s_m = 0.0, s_chi = 0.0, s_ll = 0.0;
  for (int k = 0; k < blocks; ++k) {
    const double* q = part + ((size_t)b * blocks + k) * 3;
    s_m += q[0];
    s_chi += q[1];
    s_ll += q[2];
  }
  out[3 * b + 0] = 1e6 / s_m * s_chi;
  out[3 * b + 1] = s_ll;
  out[3 * b + 2] = s_m;
}

}  // namespace

// ---------------------------------------------------------------- launchers

size_t integral_lds_bytes(int n_fine, int atmosphere, int nsx_rows) {
  return integral_layout(n_fine + 1, atmosphere == PPM_ATM_NSX ? nsx_rows : 0).total;
}

int energies_per_chunk() { return kEPerWave; }

int max_dynamic_lds(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
    return 65536;
  return v;
}

int launch_ring_setup(const DevLensing& L, const DevNsx& N, const PlainLaunch& P, void* stream) {
  if (P.n_tasks == 0) return hipSuccess;
  // the ring's two rows of the lensing tables, staged in LDS
  const size_t lds = sizeof(double) * (4 * (size_t)L.n_cp + 2 * (size_t)L.n_ca);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  hipFuncAttributes fa{};
  if (e == hipSuccess) e = hipFuncGetAttributes(&fa, (const void*)k_ring_setup);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_ring_setup, hipFuncAttributeMaxDynamicSharedMemorySize,
                            max_dynamic_lds(dev) - (int)fa.sharedSizeBytes);
  if (e != hipSuccess) return e;
  if (lds + fa.sharedSizeBytes > (size_t)max_dynamic_lds(dev)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ring_setup, dim3(P.n_tasks), dim3(kBlock), lds, (hipStream_t)stream, L, N, P);
  return hipGetLastError();
}

template <int ATM, int EM>
static hipError_t launch_plain_inst(const DevLensing& L, const DevNsx& N, const PlainLaunch& P, dim3 grid, size_t lds,
                                    hipStream_t st) {
  // the limit is set to the device's maximum, the same value from every lane
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  hipFuncAttributes fa{};
  if ((e = hipFuncGetAttributes(&fa, (const void*)k_integral_plain<ATM, EM>)) != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)k_integral_plain<ATM, EM>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          max_dynamic_lds(dev) - (int)fa.sharedSizeBytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_integral_plain<ATM, EM>), grid, dim3(kBlock), lds, st, L, N, P);
  return hipGetLastError();
}

int launch_integral_plain(const DevLensing& L, const DevNsx& N, const PlainLaunch& P, void* stream) {
  if (P.n_tasks == 0) return hipSuccess;
  const size_t lds = integral_lds_bytes(P.n_fine, P.atmosphere, P.nsx_rows);
  const dim3 grid(P.n_tasks, P.n_echunks);
  const hipStream_t st = (hipStream_t)stream;
  if (P.atmosphere != PPM_ATM_NSX) return launch_plain_inst<PPM_ATM_BLACKBODY, kECubic>(L, N, P, grid, lds, st);
  switch (P.emode) {
    case kECubic: return launch_plain_inst<PPM_ATM_NSX, kECubic>(L, N, P, grid, lds, st);
    case kECubicClip: return launch_plain_inst<PPM_ATM_NSX, kECubicClip>(L, N, P, grid, lds, st);
    case kELinear: return launch_plain_inst<PPM_ATM_NSX, kELinear>(L, N, P, grid, lds, st);
  }
  return hipErrorInvalidValue;
}

int reduce_groups(int n_tasks) { return n_tasks < 32 ? 1 : (n_tasks / 16 < 64 ? n_tasks / 16 : 64); }

int launch_reduce(const PlainLaunch& P, double* d_red, void* stream) {
  const int n = P.n_phase * P.n_e;
  const int G = reduce_groups(P.n_tasks);
  hipLaunchKernelGGL(k_reduce_groups, dim3((n + 63) / 64, G), dim3(64), 0, (hipStream_t)stream, P, d_red, G);
  hipLaunchKernelGGL(k_reduce_final, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, P,
                     (const double*)d_red, G);
  return hipGetLastError();
}

int fold_blocks(int n_ch, int n_phase) {
  const int n_tiles = (n_ch + 15) / 16;
  return ((n_phase + 15) / 16) * ((n_tiles + kFoldTiles - 1) / kFoldTiles);
}

int launch_fold(const uint16_t* rsp_hi, const uint16_t* rsp_lo, int n_ch, int n_e,
                const double* d_flux, int n_batch, int n_phase, const double* d_scale,
                double* d_counts, void* stream, const FoldFit* fit, int star0) {
  const int n_tiles = (n_ch + 15) / 16;
  dim3 grid((n_phase + 15) / 16, n_batch, (n_tiles + kFoldTiles - 1) / kFoldTiles);
  const FoldFit f = fit ? *fit : FoldFit{};
  hipLaunchKernelGGL(k_fold, grid, dim3(64), 0, (hipStream_t)stream, rsp_hi, rsp_lo, n_ch, n_e,
                     d_flux, n_phase, d_scale, d_counts, f, star0);
  return hipGetLastError();
}

int launch_fit_reduce(const double* d_part, int blocks, int n, double* d_out, void* stream) {
  hipLaunchKernelGGL(k_fit_reduce, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, d_part, blocks, n, d_out);
  return hipGetLastError();
}

}  // namespace ppm
