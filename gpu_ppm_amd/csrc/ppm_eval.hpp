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

// Bucket index of an ascending table for VecHunt look-ups: the search of x
// starts from the brackets of the two bucket bounds around x instead of the
// whole table.  idx[b] is VecHunt's bracket of z_b = x0 + b / scale,
// b = 0..kHuntBuckets, with x0 = xx[0] and scale = kHuntBuckets /
// (xx[n-1] - xx[0]).  For xx[lo] <= x < xx[hi] (or hi = n - 1) the bracket
// is the largest i <= n - 2 with xx[i] <= x, which a binary search between
// lo and hi finds as the full search does.  The bounds are taken one
// millionth of a bucket outside x, far beyond the rounding of z_b and of the
// bucket of x; below the first bucket the search starts at node 0 and beyond
// the last it ends at node n - 1, so that queries outside the table (and
// tables with repeated end nodes) give the full search's bracket too.  The
// bracket and the fraction equal the full search's bit for bit.
// idx == nullptr: full search (a descending or degenerate table).
constexpr int kHuntBuckets = 64;
struct HuntIndex {
  const uint16_t* idx;
  double x0, scale;
};

template <typename T>
__device__ inline int hunt(const T* xx, int n, double x, const HuntIndex& H, double& a) {
  if (H.idx == nullptr) return vec_hunt(xx, n, x, a);
  const double t = fmin(fmax((x - H.x0) * H.scale, -1.0), (double)kHuntBuckets + 1.0);
  const double tl = t - 1e-6;  // below the first bucket: from node 0, as the full search
  int jl = tl < 0 ? 0 : H.idx[clampi((int)floor(tl), 0, kHuntBuckets)];
  const int bh = (int)floor(t + 1e-6) + 1;  // beyond the last bucket: to node n - 1
  int ju = bh > kHuntBuckets ? n - 1 : min((int)H.idx[max(bh, 0)] + 1, n - 1);
  while (ju - jl > 1) {
    const int jm = (ju + jl) >> 1;
    if (x >= (double)xx[jm])
      jl = jm;
    else
      ju = jm;
  }
  const int i = clampi(jl, 0, n - 2);
  a = (x - (double)xx[i]) / ((double)xx[i + 1] - (double)xx[i]);
  return i;
}

// Builds the bucket index of xx[0..n-1] into idx (kHuntBuckets + 1 entries,
// threads tid, tid + nthreads, ...) and returns its HuntIndex; the caller
// synchronises before the first look-up.  Tables that do not ascend, or
// whose nodes do not fit 16 bits, get the full search.
template <typename T>
__device__ inline HuntIndex build_hunt_index(const T* xx, int n, uint16_t* idx, int tid, int nthreads) {
  if (n < 2 || n > 65535) return HuntIndex{nullptr, 0.0, 0.0};
  const double x0 = (double)xx[0], span = (double)xx[n - 1] - x0;
  if (!(span > 0)) return HuntIndex{nullptr, 0.0, 0.0};
  const double scale = kHuntBuckets / span;
  for (int b = tid; b <= kHuntBuckets; b += nthreads) {
    double a;
    idx[b] = (uint16_t)vec_hunt(xx, n, x0 + b / scale, a);
  }
  return HuntIndex{idx, x0, scale};
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
// pre-multiplied by flux/I x weight, and the patch's offset on the local
// energy axis.  The mu stencil (NsxMuMode) is cubic Lagrange from one node
// below the bracket, clamped to [0, nm-4] (cpu/nsx.hpp:181, 238-242), linear
// in the bracket (cpu/nsx.hpp:305-313), or cubic with the linear bracket in
// the last five (cpu/nsx.hpp:226-243, 260-271).  im < 0 marks a skipped patch
// (flux/I == 0, cpu/devarshi24st_cpu.cpp:226).
struct PatchVals {
  float tx;       // local energy-axis offset of this patch
  float w[4];     // mu weights x flux/I x phase-point weight (NSX) | w[0] = that product (blackbody)
  float zf;       // redshift factor (blackbody)
  int im;         // first mu node, -1: skip
};

// Energy-axis offset of a patch: tx = tx_base + log10(zf) / h in slice rows,
// taken through the FP32 logarithm (the evaluation itself runs in FP32).
struct TxMap {
  float base;   // tx_base
  float scale;  // log10(2) / h
};

// flux/I of the patch at rotational phase pp, with the fine-table bracket
// (i, a) it was interpolated in.  Zero exactly where the patch is skipped.
// PH: bucket index of the ring's phase table (or the full search).
__device__ inline double patch_flux_factor(double pp, const RingTab& R, int& i, double& a,
                                           const HuntIndex& PH = HuntIndex{nullptr, 0.0, 0.0}) {
  i = hunt(R.ph, R.nf1, pp, PH, a);
  return (1 - a) * (double)R.fI[i] + a * (double)R.fI[i + 1];
}

// PH, MU: bucket indexes of the ring's phase table and of the mu nodes
// (their brackets equal the full search's).
template <int ATM>
__device__ inline void patch_setup(double pp, double weight, const RingTab& R, const double* s_mu,
                                   const double* mu_den, int nm, double mu_min, int mu_mode, TxMap txm,
                                   PatchVals& o, const HuntIndex& PH = HuntIndex{nullptr, 0.0, 0.0},
                                   const HuntIndex& MU = HuntIndex{nullptr, 0.0, 0.0}) {
  o.tx = 0.f;
  o.w[0] = o.w[1] = o.w[2] = o.w[3] = 0.f;
  o.zf = 0.f;
  o.im = -1;
  double a;
  int i;
  const double fI = patch_flux_factor(pp, R, i, a, PH);
  if (fI == 0) return;
  const double zz = (1 - a) * (double)R.zf[i] + a * (double)R.zf[i + 1];
  const double c = fI * weight;
  o.zf = (float)zz;
  if (ATM == PPM_ATM_NSX) {
    double mu = (1 - a) * (double)R.cs[i] + a * (double)R.cs[i + 1];
    if (mu < mu_min) mu = mu_min;
    double am;
    int im = hunt(s_mu, nm, mu, MU, am);
    if (mu_mode == kMuLinear || (mu_mode == kMuCubicLinearEdge && im >= nm - 5)) {  // linear
      o.w[0] = (float)(c * (1 - am));
      o.w[1] = (float)(c * am);
    } else {  // cubic
      im = clampi(im - 1, 0, nm - 4);
      const double d0 = mu - s_mu[im], d1 = mu - s_mu[im + 1], d2 = mu - s_mu[im + 2], d3 = mu - s_mu[im + 3];
      const double* den = mu_den + 4 * im;
      const double p01 = d0 * d1, p23 = d2 * d3;
      o.w[0] = (float)(c * (d1 * p23 * den[0]));
      o.w[1] = (float)(c * (d0 * p23 * den[1]));
      o.w[2] = (float)(c * (p01 * d3 * den[2]));
      o.w[3] = (float)(c * (p01 * d2 * den[3]));
    }
    o.im = im;
    o.tx = fmaf(log2f((float)zz), txm.scale, txm.base);
  } else {
    o.w[0] = (float)c;
    o.im = 0;
  }
}

// Energy-axis geometry of one workgroup: the staged rows of the (logT, log g)
// blended NSX slice and the clamps of the reference's hunt (logEkT clamps to
// the axis ends, cpu/hunt.hpp:160-172).  An energy at local axis position t
// (clamped to [t_lo, t_hi]) uses row f = floor(t) clamped to [f_lo, f_hi] and
// x = t - f: the cubic stencil rows f-1..f+2 (its start clamped to [0, nE-4],
// cpu/nsx.hpp:180, 252-256) or the linear bracket f, f+1 (cpu/nsx.hpp:304).
struct EAxis {
  float t_lo, t_hi;  // local t clamped to the axis ends
  int f_lo, f_hi;    // legal f (within the axis and the staged rows)
  int i0_max;        // largest row-set start whose 64 rows lie in the staged allocation
};

// Axis of rows [r_lo, r_lo + n_rows) of an nE-row grid, staged in rows_alloc.
template <int EM>
__device__ inline EAxis energy_axis(int nE, int r_lo, int n_rows, int rows_alloc) {
  EAxis ax;
  ax.t_lo = (float)(-r_lo);
  ax.t_hi = (float)(nE - 1 - r_lo);
  if (EM == kELinear) {
    ax.f_lo = 0;
    ax.f_hi = min(nE - 2 - r_lo, n_rows - 2);
  } else {
    ax.f_lo = 1;
    ax.f_hi = min(nE - 4 - r_lo, n_rows - 4) + 1;
  }
  ax.i0_max = rows_alloc - 63;  // rows_alloc >= 64
  return ax;
}

// Intensity at local axis position t (cubic Lagrange over 4 rows or linear
// over 2, each row contracted over 4 mu columns; weights already carry
// flux/I).  `rows` points at the staged slice row 0 plus the patch's first mu
// column.
template <int EM>
__device__ inline float nsx_eval(float t, const EAxis& ax, const float* rows, const float w[4]) {
  t = __builtin_amdgcn_fmed3f(t, ax.t_lo, ax.t_hi);
  const int f = med3i((int)floorf(t), ax.f_lo, ax.f_hi);
  if (EM == kELinear) {
    const float* r0 = rows + __mul24(f, kSliceStride);
    const float* r1 = r0 + kSliceStride;
    const float m0 = w[0] * r0[0] + w[1] * r0[1] + w[2] * r0[2] + w[3] * r0[3];
    const float m1 = w[0] * r1[0] + w[1] * r1[1] + w[2] * r1[2] + w[3] * r1[3];
    return fmaf(t - (float)f, m1 - m0, m0);
  }
  const int st = f - 1;
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

// Per-wavefront LDS scratch of nsx_eval_wave: one float4 per slice row.
constexpr int kWaveScratchBytes = 64 * 16;

// True when the lane's energies ascend with (u, lane) across the wavefront,
// the order nsx_eval_wave relies on to bound the rows it touches.
template <int EPL>
__device__ inline bool energies_ascend(const float (&le)[EPL], int lane) {
  bool ok = true;
#pragma unroll
  for (int u = 0; u < EPL; ++u) {
    const float next = __shfl_down(le[u], 1);
    if (lane < 63) ok = ok && le[u] <= next;
    if (u + 1 < EPL) ok = ok && __shfl(le[u], 63) <= __shfl(le[u + 1 < EPL ? u + 1 : u], 0);
  }
  return __all(ok);
}

// All energies of a wavefront at one patch (uniform tx, im, w).  The patch's
// mu stencil is contracted once per slice row the wavefront touches (one row
// per lane), and each row f gets the coefficients of the polynomial from f:
// cubic (rows f-1..f+2, Newton form in v = t - (f - 1) in [0, 3])
//   I(v) = r0 + v (D1 + (v-1) (D2/2 + (v-2) D3/6)),
// the cubic Lagrange polynomial of cpu/nsx.hpp:159-167 over those rows, or
// linear (rows f, f+1), I = r0 + x (r1 - r0), x = t - f; each energy reads one
// coefficient set.  Requires energies that ascend with (u, lane)
// (energies_ascend); returns false, with nothing written, when the rows
// exceed one per lane (the caller then uses nsx_eval).
template <int EPL, int EM>
__device__ inline bool nsx_eval_wave(const float (&le)[EPL], float tx, int im, const float (&w)[4], const EAxis& ax,
                                     const float* s_S, float4* s_C, int lane, float (&out)[EPL]) {
  constexpr int kSpan = EM == kELinear ? 1 : 3;  // rows beyond a coefficient set's first
  int st[EPL];
  float v[EPL];
#pragma unroll
  for (int u = 0; u < EPL; ++u) {
    const float t = __builtin_amdgcn_fmed3f(le[u] + tx, ax.t_lo, ax.t_hi);
    st[u] = med3i((int)floorf(t), ax.f_lo, ax.f_hi) - (EM == kELinear ? 0 : 1);
    v[u] = t - (float)st[u];
  }
  const int r_lo = readlane_i(st[0], 0);
  const int n_c = readlane_i(st[EPL - 1], 63) - r_lo + 1;  // coefficient sets needed
  if (n_c > 64 - kSpan) return false;
  float U = 0.f;
  if (lane < n_c + kSpan) {
    const float* p = s_S + __mul24(r_lo + lane, kSliceStride) + im;
    U = w[0] * p[0] + w[1] * p[1] + w[2] * p[2] + w[3] * p[3];
  }
  const float U1 = __shfl_down(U, 1), U2 = __shfl_down(U, 2), U3 = __shfl_down(U, 3);
  if (lane < n_c) {
    if (EM == kELinear) {
      s_C[lane] = make_float4(U, U1 - U, 0.f, 0.f);
    } else {
      const float d1 = U1 - U, d1b = U2 - U1, d1c = U3 - U2;
      const float d2 = d1b - d1, d3 = (d1c - d1b) - d2;
      s_C[lane] = make_float4(U, d1, 0.5f * d2, (1.f / 6.f) * d3);
    }
  }
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int u = 0; u < EPL; ++u) {
    const float4 k = s_C[med3i(st[u] - r_lo, 0, 63)];
    const float x = v[u];
    out[u] = EM == kELinear ? fmaf(x, k.y, k.x) : fmaf(x, fmaf(x - 1.f, fmaf(x - 2.f, k.w, k.z), k.y), k.x);
  }
  __builtin_amdgcn_wave_barrier();
  return true;
}

// Wave-uniform description of a wavefront's energies on the local axis.  A
// lane's EPL energies (u = 0..EPL-1, energy u*64 + lane) are split into SETS
// row sets of EPL/SETS consecutive u; each set spans its own window of slice
// rows, contracted separately by nsx_eval_table.
template <int SETS>
struct WaveEnergies {
  float le_first[SETS], le_last[SETS];  // lane 0's first and lane 63's last energy of each set
  bool fast;  // ascending and every set spanning at most 58 rows: nsx_eval_table applies
};

template <int EPL, int SETS>
__device__ inline WaveEnergies<SETS> wave_energies(const float (&le)[EPL], int lane) {
  static_assert(EPL % SETS == 0, "row sets split the lane's energies evenly");
  constexpr int G = EPL / SETS;
  WaveEnergies<SETS> we;
  we.fast = energies_ascend<EPL>(le, lane);
#pragma unroll
  for (int s = 0; s < SETS; ++s) {
    we.le_first[s] = readlane_f(le[s * G], 0);
    we.le_last[s] = readlane_f(le[s * G + G - 1], 63);
    we.fast = we.fast && we.le_last[s] - we.le_first[s] <= 58.f;
  }
  return we;
}

// Classification of one patch, made once in its set-up lane.  When the
// wavefront's energies allow nsx_eval_table (WaveEnergies::fast) the patch is
// evaluated there, each row set without clamps when none of its energies
// reaches a clamp of the axis or of the row f (checked on the set's first and
// last energy, the evaluation being monotone in the energy), with them
// otherwise.
// Set s holds slice rows i0[s] - 1 + lane: i0[s] is at most the row f of the
// set's first energy, and at most EAxis::i0_max so that every lane's row is
// in the staged allocation.  txs = tx - i0 puts a set's
// energies on its lane axis (floor(le + txs) is the lane holding the row),
// rb = (i0 - 1) * kSliceStride + im is the LDS offset of lane 0's row at the
// patch's first mu column.
template <int SETS>
struct PatchRows {
  int i0[SETS];
  float txs[SETS];
  int rb[SETS];
  int clamp;  // bit s: set s takes the clamped evaluation (every bit without WaveEnergies::fast)
};

template <int SETS>
__device__ inline PatchRows<SETS> patch_rows(float tx, int im, const WaveEnergies<SETS>& we, const EAxis& ax) {
  PatchRows<SETS> r;
  r.clamp = 0;
#pragma unroll
  for (int s = 0; s < SETS; ++s) {
    const float tc = __builtin_amdgcn_fmed3f(we.le_first[s] + tx, ax.t_lo, ax.t_hi);
    r.i0[s] = min(med3i((int)floorf(tc), ax.f_lo, ax.f_hi), ax.i0_max);
    r.txs[s] = tx - (float)r.i0[s];
    r.rb[s] = (r.i0[s] - 1) * kSliceStride + im;
    // the unclamped path evaluates at le + txs (rounded once): its end
    // points, taken the same way, bound every energy's lane and stencil
    // (the rounding of le + tx may differ near an integer)
    const float uf = we.le_first[s] + r.txs[s], ul = we.le_last[s] + r.txs[s];
    const float i0f = (float)r.i0[s];
    if (!(we.fast && uf >= 0.f && ul < (float)(ax.f_hi + 1) - i0f && ul <= ax.t_hi - i0f)) r.clamp |= 1 << s;
  }
  return r;
}

// Lane i receives lane i+1's value (DPP wave_shl:1); lane 63 receives 0.
__device__ inline float lane_next(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}

// Lane i receives lane i-1's value (DPP wave_shr:1); lane 0 receives 0.
__device__ inline float lane_prev(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}

// floor(t) as an integer (one instruction).
__device__ inline int floor_to_int(float t) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(t));
  return r;
}

// nsx_load_rows issues the slice-row reads of patch q (lane l: row i0 - 1 + l
// of each set at the patch's four mu columns); a caller may issue them for
// the next patch before evaluating the current one.
template <int SETS>
__device__ inline void nsx_load_rows(const PatchRows<SETS>& pr, int q, const float* s_S, int lane,
                                     float (&p)[SETS][4]) {
#pragma unroll
  for (int s = 0; s < SETS; ++s) {
    const float* r = s_S + readlane_i(pr.rb[s], q) + lane * kSliceStride;
#pragma unroll
    for (int m = 0; m < 4; ++m) p[s][m] = r[m];
  }
}

// Contracted evaluation of one patch for every energy of a wavefront with
// WaveEnergies::fast, the patch set up in lane q (pv, pr) and its slice rows
// read by nsx_load_rows into p.  For energy t = le + tx (clamped to the axis
// ends) the row is f = floor(t) (clamped to EAxis::f_lo..f_hi) and x = t - f,
// in [0, 1) away from the clamps.  For each row set, lane l contracts slice
// row i0 - 1 + l over the patch's mu stencil and forms the power-basis
// coefficients of the polynomial in x of that row f: the cubic through rows
// f-1..f+2 at the nodes x = -1, 0, 1, 2,
//   I(x) = a0 + x (a1 + x (a2 + x a3))
// (the polynomial and the clamps of cpu/nsx.hpp:159-167, 252-256), or the
// line through rows f, f+1, I(x) = a0 + x a1 (cpu/nsx.hpp:304-316).  a0 is
// the lane's own value, the first difference looks back one lane and the
// higher ones forward: the differences of lane l + 1 are the next
// differences of lane l, so each order takes one shifted subtraction.
// Each lane writes its row's four coefficients to a per-wavefront LDS table
// (one ds_write_b128) and each energy reads its row's (one ds_read_b128):
// entry l belongs to row i0 - 1 + l, an energy of row f reads entry
// f - i0 + 1, and lane 0 (no lane behind it) is never read since f >= i0.
// Away from the clamps an energy's entry and x come from t - i0 = le + txs
// directly.  Both sets' coefficients and every energy's row and x are
// formed first, then the sets go through the 64-entry table in turn with
// all EPL reads in flight before the first result is used: a wavefront's
// LDS instructions complete in issue order, so the reads of one set precede
// the next set's writes.  Bit s of CM selects the clamped evaluation for
// set s (PatchRows::clamp).  EM = kECubicClip sets negative intensities to
// 0 (use_sg, cpu/nsx.hpp:201; the weights carry the positive flux/I x
// weight, so the sign is the interpolant's).
template <int EPL, int CM, int SETS, int EM>
__device__ inline void nsx_eval_table(const float (&le)[EPL], const PatchVals& pv, const PatchRows<SETS>& pr, int q,
                                      const EAxis& ax, const float (&p)[SETS][4], int lane, float4* s_T,
                                      float (&out)[EPL]) {
  constexpr int G = EPL / SETS;
  constexpr bool kLin = EM == kELinear;
  float w[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) w[m] = readlane_f(pv.w[m], q);
  const float tx = CM ? readlane_f(pv.tx, q) : 0.f;
  float4 a[SETS];
#pragma unroll
  for (int s = 0; s < SETS; ++s) {
    const float U = w[0] * p[s][0] + w[1] * p[s][1] + w[2] * p[s][2] + w[3] * p[s][3];
    if (kLin) {
      a[s] = make_float4(U, lane_next(U) - U, 0.f, 0.f);
    } else {
      const float d1 = U - lane_prev(U);
      const float D2 = lane_next(d1) - d1;
      const float D3 = lane_next(D2) - D2;
      const float a2 = 0.5f * D2, a3 = (1.f / 6.f) * D3;
      a[s] = make_float4(U, d1 + a2 - a3, a2, a3);
    }
  }
  int src[EPL];
  float x[EPL];
  // the clamps of the row f in the float domain (floor and med3 of a float,
  // f exact), the set's first row folded into the table pointer (s_T + 1 -
  // i0 stays inside the dynamic LDS block: the table lies behind the ring
  // tables and the slice, far more than 16 i0 <= 16 EAxis::i0_max bytes)
  const float f_lo = (float)ax.f_lo, f_hi = (float)ax.f_hi;
  const float4* T[SETS];
#pragma unroll
  for (int s = 0; s < SETS; ++s) {
    const bool clamp = (CM >> s) & 1;
    const int i0 = clamp ? readlane_i(pr.i0[s], q) : 0;
    const float txs = clamp ? 0.f : readlane_f(pr.txs[s], q);
    T[s] = s_T + 1 - i0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int u = s * G + g;
      if (clamp) {
        const float t = __builtin_amdgcn_fmed3f(le[u] + tx, ax.t_lo, ax.t_hi);
        const float f = __builtin_amdgcn_fmed3f(floorf(t), f_lo, f_hi);
        x[u] = t - f;
        src[u] = (int)f;
      } else {
        const float t = le[u] + txs;
        src[u] = floor_to_int(t);
        x[u] = __builtin_amdgcn_fractf(t);
      }
    }
  }
  float4 c[EPL];
#pragma unroll
  for (int s = 0; s < SETS; ++s) {
    __builtin_amdgcn_wave_barrier();
    s_T[lane] = a[s];
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int g = 0; g < G; ++g) c[s * G + g] = T[s][src[s * G + g]];
  }
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int u = 0; u < EPL; ++u) {
    const float r = kLin ? fmaf(x[u], c[u].y, c[u].x) : fmaf(x[u], fmaf(x[u], fmaf(x[u], c[u].w, c[u].z), c[u].y), c[u].x);
    out[u] = EM == kECubicClip ? fmaxf(r, 0.f) : r;
  }
}

// One patch whose wavefront energies span too many rows for nsx_eval_table:
// contracted evaluation with clamps (nsx_eval_wave) when they fit the
// scratch, else per energy (nsx_eval).
template <int EPL, int EM>
__device__ inline void nsx_eval_edge(const float (&le)[EPL], bool asc, float tx, int im, const float (&w)[4],
                                     const EAxis& ax, const float* s_S, float4* s_C, int lane, float (&out)[EPL]) {
  if (!(asc && nsx_eval_wave<EPL, EM>(le, tx, im, w, ax, s_S, s_C, lane, out))) {
#pragma unroll
    for (int u = 0; u < EPL; ++u) {
      out[u] = nsx_eval<EM>(le[u] + tx, ax, s_S + im, w);
      __builtin_amdgcn_sched_barrier(0);  // rare path: keep its register use low
    }
  }
  if (EM == kECubicClip)
#pragma unroll
    for (int u = 0; u < EPL; ++u) out[u] = fmaxf(out[u], 0.f);
}

// Stage the rows [r_lo, r_lo + n_rows) of the (logT, log g) blend of the NSX
// grid into LDS (row stride kSliceStride, mu padding zeroed): the ring's
// 2 x 2 (bilinear) or 4 x 4 (bicubic Lagrange) stencil with the weights
// k_ring_setup computed in FP64.
__device__ inline void stage_nsx_rows(const DevNsx& N, const RingDesc& D, int r_lo, int n_rows, float* s_S, int tid,
                                      int nthreads) {
  const size_t sT = (size_t)N.ng * N.nE * N.nm, sg = (size_t)N.nE * N.nm;
  const float* c00 = N.data + D.iT * sT + D.ig * sg + (size_t)r_lo * N.nm;
  const int total = n_rows * kSliceStride;
  if (D.n_tg == 2) {
    const float w00 = D.wtg[0], w01 = D.wtg[1], w10 = D.wtg[4], w11 = D.wtg[5];
    const float* c01 = c00 + sg;
    const float* c10 = c00 + sT;
    const float* c11 = c10 + sg;
    for (int k = tid; k < total; k += nthreads) {
      const int r = k / kSliceStride, m = k - r * kSliceStride;
      float v = 0.f;
      if (m < N.nm) {
        const size_t o = (size_t)r * N.nm + m;
        v = w00 * c00[o] + w01 * c01[o] + w10 * c10[o] + w11 * c11[o];
      }
      s_S[k] = v;
    }
    return;
  }
  float w[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = D.wtg[i];
  for (int k = tid; k < total; k += nthreads) {
    const int r = k / kSliceStride, m = k - r * kSliceStride;
    float v = 0.f;
    if (m < N.nm) {
      const float* c = c00 + (size_t)r * N.nm + m;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) v = fmaf(w[4 * a + b], c[a * sT + b * sg], v);
    }
    s_S[k] = v;
  }
}

}  // namespace ppm
