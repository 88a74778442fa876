// Device building blocks shared by the plain and the batched integral
// kernels: the per-patch set-up (FP64, one lane per patch) and the per
// (patch, energy) atmosphere evaluation (FP32, one energy per lane).
#pragma once

#include <hip/hip_runtime.h>

#include "ppm_internal.hpp"

namespace ppm {

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ inline int med3i(int v, int lo, int hi) { return min(max(v, lo), hi); }

// Bracket search with VecHunt's result (cpu/hunt.hpp:8-76): lower index of
// the bracket, clamped to [0, n-2]; the linear fraction extrapolates.
template <typename T>
__device__ inline int vec_hunt(const T* xx, int n, double x, double& a) {
  const bool ascnd = xx[n - 1] >= xx[0];
  int jl = 0, ju = n - 1;
  while (ju - jl > 1) {
    const int jm = (ju + jl) >> 1;
    if ((x >= (double)xx[jm]) == ascnd)
      jl = jm;
    else
      ju = jm;
  }
  const int i = clampi(jl, 0, n - 2);
  a = (x - (double)xx[i]) / ((double)xx[i + 1] - (double)xx[i]);
  return i;
}

// A ring's fine-phase table staged in LDS.
struct RingTab {
  const double* ph;  // time-delayed phase, [nf1]
  const float* fI;   // flux / I
  const float* zf;   // redshift factor 1/(delta sqrt(1-u))
  const float* cs;   // cos(sigma')
  int nf1;
};

// Set-up of one patch at rotational phase pp (already wrapped into
// [ph[0], ph[0] + 1)): interpolation of the fine table
// (cpu/devarshi24st_cpu.cpp:224-230), the mu stencil with its weights
// (cpu/nsx.hpp:248-282) pre-multiplied by flux/I x weight, and the patch's
// offset on the local energy axis.  im < 0 marks a skipped patch
// (flux/I == 0, cpu/devarshi24st_cpu.cpp:226).
struct PatchVals {
  float tx;       // local energy-axis offset of this patch
  float w[4];     // mu weights x flux/I x phase-point weight (NSX) | w[0] = that product (blackbody)
  float zf;       // redshift factor (blackbody)
  int im;         // first mu node, -1: skip
};

template <int ATM>
__device__ inline void patch_setup(double pp, double weight, const RingTab& R, const double* s_mu, int nm,
                                   double mu_min, double tx_base, double inv_h, PatchVals& o) {
  o.tx = 0.f;
  o.w[0] = o.w[1] = o.w[2] = o.w[3] = 0.f;
  o.zf = 0.f;
  o.im = -1;
  double a;
  const int i = vec_hunt(R.ph, R.nf1, pp, a);
  const double fI = (1 - a) * (double)R.fI[i] + a * (double)R.fI[i + 1];
  if (fI == 0) return;
  const double zz = (1 - a) * (double)R.zf[i] + a * (double)R.zf[i + 1];
  const double c = fI * weight;
  o.zf = (float)zz;
  if (ATM == PPM_ATM_NSX) {
    double mu = (1 - a) * (double)R.cs[i] + a * (double)R.cs[i + 1];
    if (mu < mu_min) mu = mu_min;
    double am;
    int im = vec_hunt(s_mu, nm, mu, am);
    if (im >= nm - 5) {  // linear at the edge
      o.w[0] = (float)(c * (1 - am));
      o.w[1] = (float)(c * am);
    } else {  // cubic in the middle
      im = clampi(im - 1, 0, nm - 4);
      const double x0 = s_mu[im], x1 = s_mu[im + 1], x2 = s_mu[im + 2], x3 = s_mu[im + 3];
      o.w[0] = (float)(c * ((mu - x1) * (mu - x2) * (mu - x3) / ((x0 - x1) * (x0 - x2) * (x0 - x3))));
      o.w[1] = (float)(c * ((mu - x0) * (mu - x2) * (mu - x3) / ((x1 - x0) * (x1 - x2) * (x1 - x3))));
      o.w[2] = (float)(c * ((mu - x0) * (mu - x1) * (mu - x3) / ((x2 - x0) * (x2 - x1) * (x2 - x3))));
      o.w[3] = (float)(c * ((mu - x0) * (mu - x1) * (mu - x2) / ((x3 - x0) * (x3 - x1) * (x3 - x2))));
    }
    o.im = im;
    o.tx = (float)(tx_base + log10(zz) * inv_h);
  } else {
    o.w[0] = (float)c;
    o.im = 0;
  }
}

// Energy-axis geometry of one workgroup: the staged rows of the (logT, log g)
// blended NSX slice and the clamps of the reference's hunt (the stencil start
// clamps to [0, nE-4] and logEkT to the axis ends, cpu/nsx.hpp:252-256).
struct EAxis {
  float t_lo, t_hi;  // local t clamped to the axis ends
  int st_lo, st_hi;  // legal local stencil starts (also the staged range)
};

// Intensity at local axis position t (cubic Lagrange over 4 rows x 4 mu
// columns; weights already carry flux/I).  `rows` points at the staged slice
// row 0 plus the patch's first mu column.
__device__ inline float nsx_eval(float t, const EAxis& ax, const float* rows, const float w[4]) {
  t = __builtin_amdgcn_fmed3f(t, ax.t_lo, ax.t_hi);
  const int st = med3i((int)floorf(t) - 1, ax.st_lo, ax.st_hi);
  const float v = t - (float)st;
  const float d1 = v - 1.f, d2 = v - 2.f, d3 = v - 3.f;
  const float p23 = d2 * d3, p01 = v * d1;
  const float e0 = -(1.f / 6.f) * d1 * p23;
  const float e1 = 0.5f * v * p23;
  const float e2 = -0.5f * p01 * d3;
  const float e3 = (1.f / 6.f) * p01 * d2;
  const float* r = rows + __mul24(st, kSliceStride);
  float acc = 0.f;
  const float ek[4] = {e0, e1, e2, e3};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float* rk = r + k * kSliceStride;
    const float m = w[0] * rk[0] + w[1] * rk[1] + w[2] * rk[2] + w[3] * rk[3];
    acc = fmaf(ek[k], m, acc);
  }
  return acc;
}

__device__ inline float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ inline int readlane_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

__device__ inline void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rows of the contracted slice one wavefront keeps per patch.
constexpr int kRowsCap = 128;

// Per-wavefront LDS scratch of the contracted evaluation.
struct WaveRows {
  float* U;    // [kRowsCap] slice rows contracted over the patch's mu stencil
  float4* C;   // [kRowsCap] cubic coefficients (Newton form) starting at each row
};

// All energies of a wavefront at one patch (uniform: tx, im, w).  The mu
// stencil is contracted once per row the wavefront touches, each row start
// gets the Newton coefficients of the cubic through 4 consecutive rows, and
// each energy is then one coefficient read and three FMAs:
//   I(v) = r0 + v (a + (v-1) (d2/2 + (v-2) d3/6)),  v = t - st in [0, 3],
// the same cubic Lagrange polynomial as cpu/nsx.hpp:159-167 on those rows.
// Returns false (nothing written) when the rows exceed the scratch.
template <int EPL>
__device__ inline bool nsx_eval_rows(const float (&le)[EPL], float tx, int im, const float (&w)[4], const EAxis& ax,
                                     const float* s_S, const WaveRows& wr, int lane, float (&out)[EPL]) {
  int st[EPL];
  float v[EPL];
#pragma unroll
  for (int u = 0; u < EPL; ++u) {
    const float t = __builtin_amdgcn_fmed3f(le[u] + tx, ax.t_lo, ax.t_hi);
    st[u] = med3i((int)floorf(t) - 1, ax.st_lo, ax.st_hi);
    v[u] = t - (float)st[u];
  }
  // energies ascend with (u, lane): the first and last lanes bound the rows
  const int r_lo = readlane_i(st[0], 0);
  const int n_row = readlane_i(st[EPL - 1], 63) + 4 - r_lo;
  if (n_row > kRowsCap) return false;
  wave_fence();  // earlier reads of the scratch are complete
  const float* base = s_S + __mul24(r_lo, kSliceStride) + im;
  for (int rr = lane; rr < n_row; rr += 64) {
    const float* p = base + __mul24(rr, kSliceStride);
    wr.U[rr] = w[0] * p[0] + w[1] * p[1] + w[2] * p[2] + w[3] * p[3];
  }
  wave_fence();
  for (int rr = lane; rr + 3 < n_row; rr += 64) {
    const float r0 = wr.U[rr], r1 = wr.U[rr + 1], r2 = wr.U[rr + 2], r3 = wr.U[rr + 3];
    const float a = r1 - r0, b = r2 - r1, c = r3 - r2;
    const float d2 = b - a, d3 = (c - b) - d2;
    wr.C[rr] = make_float4(r0, a, 0.5f * d2, (1.f / 6.f) * d3);
  }
  wave_fence();
#pragma unroll
  for (int u = 0; u < EPL; ++u) {
    const float4 k = wr.C[st[u] - r_lo];
    const float x = v[u];
    out[u] = fmaf(x, fmaf(x - 1.f, fmaf(x - 2.f, k.w, k.z), k.y), k.x);
  }
  return true;
}


// Stage the rows [r_lo, r_lo + n_rows) of the (logT, log g) bilinear blend of
// the NSX grid into LDS (row stride kSliceStride, mu padding zeroed).
__device__ inline void stage_nsx_rows(const DevNsx& N, int iT, int ig, float aT, float ag, int r_lo, int n_rows,
                                      float* s_S, int tid, int nthreads) {
  const float w00 = (1.f - aT) * (1.f - ag), w01 = (1.f - aT) * ag, w10 = aT * (1.f - ag), w11 = aT * ag;
  const size_t sT = (size_t)N.ng * N.nE * N.nm, sg = (size_t)N.nE * N.nm;
  const float* c00 = N.data + iT * sT + ig * sg + (size_t)r_lo * N.nm;
  const float* c01 = c00 + sg;
  const float* c10 = c00 + sT;
  const float* c11 = c10 + sg;
  const int total = n_rows * kSliceStride;
  for (int k = tid; k < total; k += nthreads) {
    const int r = k / kSliceStride, m = k - r * kSliceStride;
    float v = 0.f;
    if (m < N.nm) {
      const size_t o = (size_t)r * N.nm + m;
      v = w00 * c00[o] + w01 * c01[o] + w10 * c10[o] + w11 * c11[o];
    }
    s_S[k] = v;
  }
}

}  // namespace ppm
