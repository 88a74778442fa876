// Batched integral for gfx950: many stars in one persistent launch.
//
// Work queue: items = (star, spot, contiguous range of ring tasks), taken
// largest-first by persistent workgroups through an atomic counter.
//
// Phase-shift dedupe.  On a ring of n pixels, patch j seen at phase point p
// has rotational phase x = frac(p + (j - s/2)/n).  Writing p n = m + delta,
// points sharing delta (a "class") see the same n phases
//   x_k = frac((k + c)/n),  k = (m + j - 1) mod n,  c = delta + 1 - s/2,
// so each x_k is evaluated once per energy and every (point, run of active
// pixels) sums a window of consecutive k.  The sum is taken from an FP64
// running prefix over k at the window's two ends (sorted events), which
// only reassociates the reference's sum (cpu/devarshi24st_cpu.cpp:214-239).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ppm_batch_dev.hpp"
#include "ppm_eval.hpp"

namespace ppm {

namespace {

constexpr int kBW = 3;                  // wavefronts per workgroup
constexpr int kBThreads = 64 * kBW;
constexpr int kBEPL = 2;                // energies per lane
constexpr int kBEnergies = 64 * kBW * kBEPL;  // energies per workgroup (384)

// A window end on the position axis: at pos, acc[bin] += sw * prefix.
struct EvRec {
  int pos, bin;
  double sw;
};

struct BLds {
  size_t ph, mu, fI, zf, cs, S, cov, keys, sorted, total;
};

__host__ __device__ inline BLds batch_layout(int nf1, int nsx_rows, int cov_words, int ev_cap) {
  BLds l;
  size_t o = 0;
  l.ph = o;
  o += (size_t)nf1 * 8;
  l.mu = o;
  o += 80 * 8;
  l.fI = o;
  o += (size_t)nf1 * 4;
  l.zf = o;
  o += (size_t)nf1 * 4;
  l.cs = o;
  o += (size_t)nf1 * 4;
  o = (o + 15) & ~(size_t)15;
  l.S = o;
  o += (size_t)nsx_rows * kSliceStride * 4;
  o = (o + 15) & ~(size_t)15;
  l.cov = o;
  o += (size_t)cov_words * 4;
  o = (o + 15) & ~(size_t)15;
  l.keys = o;
  o += (size_t)ev_cap * 8;
  l.sorted = o;
  o += (size_t)ev_cap * sizeof(EvRec);
  l.total = (o + 15) & ~(size_t)15;
  return l;
}

__device__ inline uint64_t rfl64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

__global__ void k_queue_reset(int* q, unsigned long long* n_evals) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *q = 0;
    *n_evals = 0ull;
  }
}

template <int ATM>
__global__ __launch_bounds__(kBThreads) void k_integral_batched(DevNsx N, BatchLaunch B) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nf1 = B.n_fine + 1;
  extern __shared__ __align__(16) unsigned char smem[];
  const int rows_alloc = ATM == PPM_ATM_NSX ? B.nsx_rows : 0;
  const BLds lay = batch_layout(nf1, rows_alloc, B.cov_words, B.ev_cap);
  double* s_ph = reinterpret_cast<double*>(smem + lay.ph);
  double* s_mu = reinterpret_cast<double*>(smem + lay.mu);
  float* s_fI = reinterpret_cast<float*>(smem + lay.fI);
  float* s_zf = reinterpret_cast<float*>(smem + lay.zf);
  float* s_cs = reinterpret_cast<float*>(smem + lay.cs);
  float* s_S = reinterpret_cast<float*>(smem + lay.S);
  uint32_t* s_cov = reinterpret_cast<uint32_t*>(smem + lay.cov);
  uint64_t* s_keys = reinterpret_cast<uint64_t*>(smem + lay.keys);
  EvRec* s_ev = reinterpret_cast<EvRec*>(smem + lay.sorted);
  __shared__ int s_item, s_nev;

  if (ATM == PPM_ATM_NSX)
    for (int k = tid; k < N.nm; k += kBThreads) s_mu[k] = N.mu[k];

  const int n_e = B.n_e;
  const double h = (N.logE_max - N.logE_min) / (N.nE - 1);
  const double inv_h = 1.0 / h;
  unsigned int n_done = 0;  // evaluations by this lane

  for (;;) {
    __syncthreads();
    if (tid == 0) s_item = atomicAdd(B.queue, 1);
    __syncthreads();
    const int q_idx = s_item;
    if (q_idx >= B.n_items) break;
    const BatchItem It = B.items[B.order[q_idx]];
    double* acc = B.acc + (size_t)B.order[q_idx] * B.n_phase * n_e;
    // energies of this thread: e = e0 + wave*128 + u*64 + lane (ascending grid)
    const int e_hi = min(It.e0 + kBEnergies, n_e) - 1;
    const double l_ref = log10(B.energies[It.e0]);
    float le[kBEPL], Ef[kBEPL];
    int eidx[kBEPL];
#pragma unroll
    for (int u = 0; u < kBEPL; ++u) {
      const int e = It.e0 + wave * 64 * kBEPL + u * 64 + lane;
      eidx[u] = e <= e_hi ? e : -1;
      const double E = B.energies[e <= e_hi ? e : e_hi];
      Ef[u] = (float)E;
      le[u] = (float)((log10(E) - l_ref) * inv_h);
    }
    for (int b = 0; b < B.n_phase; ++b)
#pragma unroll
      for (int u = 0; u < kBEPL; ++u)
        if (eidx[u] >= 0) acc[(size_t)b * n_e + eidx[u]] = 0.0;

    for (int task = It.task0; task < It.task1; ++task) {
      const RingDesc* Dp = B.desc + task;
      if (!Dp->kept) continue;
      const RingTask T = B.tasks[task];
      const SpotParams& S = B.spots[T.spot];
      __syncthreads();  // previous ring's LDS is no longer read
      {
        const double* gph = B.ph + (size_t)task * nf1;
        const float* gf = B.fine + (size_t)task * 3 * nf1;
        for (int k = tid; k < nf1; k += kBThreads) {
          s_ph[k] = gph[k];
          s_fI[k] = gf[k];
          s_zf[k] = gf[nf1 + k];
          s_cs[k] = gf[2 * nf1 + k];
        }
      }
      int r_lo = 0;
      EAxis ax{0.f, 0.f, 0, 0};
      if (ATM == PPM_ATM_NSX) {
        const double lkT = log10(S.kT_nsx);
        const double t_min = (l_ref + log10(Dp->zf_min) - lkT - N.logE_min) * inv_h;
        const double t_max = (log10(B.energies[e_hi]) + log10(Dp->zf_max) - lkT - N.logE_min) * inv_h;
        r_lo = clampi((int)floor(t_min - 0.01) - 1, 0, N.nE - 4);
        const int r_hi = clampi((int)floor(t_max + 0.01) - 1, 0, N.nE - 4) + 3;
        int n_rows = r_hi - r_lo + 1;
        if (n_rows > rows_alloc) {
          if (tid == 0) atomicOr(&B.desc[task].flags, 2);
          n_rows = rows_alloc;
        }
        ax.t_lo = (float)(-r_lo);
        ax.t_hi = (float)(N.nE - 1 - r_lo);
        ax.st_lo = 0;
        ax.st_hi = min(N.nE - 4 - r_lo, n_rows - 4);
        stage_nsx_rows(N, Dp->iT, Dp->ig, Dp->aT, Dp->ag, r_lo, n_rows, s_S, tid, kBThreads);
      }
      __syncthreads();
      const double tx_base = (l_ref - log10(S.kT_nsx) - N.logE_min) * inv_h - r_lo;
      const float inv_kT = (float)S.inv_kT;
      const RingTab R{s_ph, s_fI, s_zf, s_cs, nf1};
      const bool point = T.ring == 0;
      const int n = point ? 1 : Dp->n_pix;
      const int n_runs = point ? 1 : Dp->n_runs;
      const ClassSet cset = B.csets[point ? 0 : n / 4];
      const double c_shift = point ? Dp->ps_shift : 1.0 - 0.25 * Dp->s_shift2;
      const int n_words = (n + 63) / 64 * 2;  // coverage words, 64-position batches

      for (int ci = 0; ci < cset.n_classes; ++ci) {
        const ClassHdr C = B.chdr[cset.first_class + ci];
        const double c_o = C.delta + c_shift;
        const int per_group = max(1, B.ev_cap / (4 * n_runs));
        for (int g0 = 0; g0 < C.n_pts; g0 += per_group) {
          const int g1 = min(C.n_pts, g0 + per_group);
          const int n_win = (g1 - g0) * n_runs;
          // ---- windows -> events and coverage
          __syncthreads();
          for (int w = tid; w < n_words; w += kBThreads) s_cov[w] = 0u;
          if (tid == 0) s_nev = 0;
          __syncthreads();
          for (int wi = tid; wi < n_win; wi += kBThreads) {
            const int ql = wi / n_runs, r = wi - ql * n_runs;
            const ClassPt cp = B.cpts[C.first_pt + g0 + ql];
            const int j0 = point ? 1 : Dp->run_start[r];
            const int L = point ? 1 : Dp->run_len[r];
            int a = cp.m + j0 - 1;
            a %= n;
            const uint64_t qr = ((uint64_t)ql << 12) | ((uint64_t)r << 2);
            int pieces[2][2];
            int np = 0;
            if (a + L <= n) {
              pieces[np][0] = a;
              pieces[np++][1] = a + L;
            } else {
              pieces[np][0] = a;
              pieces[np++][1] = n;
              pieces[np][0] = 0;
              pieces[np++][1] = a + L - n;
            }
            for (int pc = 0; pc < np; ++pc) {
              const int lo = pieces[pc][0], hi = pieces[pc][1];
              const int slot = atomicAdd(&s_nev, 2);
              s_keys[slot] = ((uint64_t)lo << 32) | qr | ((uint64_t)pc << 1) | 0u;
              s_keys[slot + 1] = ((uint64_t)hi << 32) | qr | ((uint64_t)pc << 1) | 1u;
              for (int k = lo; k < hi;) {  // coverage bits
                const int wd = k >> 5, b0 = k & 31;
                const int nb = min(32 - b0, hi - k);
                const uint32_t m = (nb == 32 ? 0xffffffffu : ((1u << nb) - 1u)) << b0;
                atomicOr(&s_cov[wd], m);
                k += nb;
              }
            }
          }
          __syncthreads();
          const int n_ev = __builtin_amdgcn_readfirstlane(s_nev);
          // ---- rank sort of the events (keys are unique) into records
          for (int i = tid; i < n_ev; i += kBThreads) {
            const uint64_t ki = s_keys[i];
            int rank = 0;
            for (int j = 0; j < n_ev; ++j) rank += s_keys[j] < ki;
            const ClassPt cp = B.cpts[C.first_pt + g0 + (int)((ki >> 12) & 0xfffff)];
            s_ev[rank] = EvRec{(int)(ki >> 32), cp.bin, (ki & 1u) ? cp.w : -cp.w};
          }
          __syncthreads();

          // ---- sweep over covered positions; each wavefront owns its energies
          double P[kBEPL];
#pragma unroll
          for (int u = 0; u < kBEPL; ++u) P[u] = 0.0;
          // the next event is held in scalar registers; contributions are
          // posted as no-return FP64 atomics to the item's private
          // accumulator (one writer per address: deterministic)
          int ev = 0;
          int ev_pos = 0x7fffffff, ev_bin = 0;
          double ev_sw = 0.0;
          auto load_ev = [&]() {
            if (ev < n_ev) {
              ev_pos = __builtin_amdgcn_readfirstlane(s_ev[ev].pos);
              ev_bin = __builtin_amdgcn_readfirstlane(s_ev[ev].bin);
              ev_sw = __builtin_bit_cast(double, rfl64(__builtin_bit_cast(uint64_t, s_ev[ev].sw)));
            } else {
              ev_pos = 0x7fffffff;
            }
          };
          auto fire = [&]() {
#pragma unroll
            for (int u = 0; u < kBEPL; ++u)
              if (eidx[u] >= 0) unsafeAtomicAdd(&acc[(size_t)ev_bin * n_e + eidx[u]], ev_sw * P[u]);
            ++ev;
            load_ev();
          };
          load_ev();
          for (int base = 0; base < n; base += 64) {
            const uint32_t c_lo = __builtin_amdgcn_readfirstlane(s_cov[base >> 5]);
            const uint32_t c_hi = __builtin_amdgcn_readfirstlane(s_cov[(base >> 5) + 1]);
            const uint64_t cov = (uint64_t)c_lo | ((uint64_t)c_hi << 32);
            if (cov == 0) continue;
            const int k = base + lane;
            PatchVals pv;
            pv.im = -1;
            if ((cov >> lane) & 1ull) {
              const double v = (k + c_o) / n + 1.0;
              double pp = v - floor(v);
              if (pp < s_ph[0]) pp += 1.;
              patch_setup<ATM>(pp, 1.0, R, s_mu, N.nm, N.mu_min, tx_base, inv_h, pv);
            }
            uint64_t todo = cov;
            while (todo) {
              const int q = __builtin_ctzll(todo);
              todo &= todo - 1;
              const int pos = base + q;
              // window ends at or before pos use the prefix before F(pos)
              while (ev_pos <= pos) fire();
              const int im = readlane_i(pv.im, q);
              if (im < 0) continue;
              if (ATM == PPM_ATM_NSX) {
                const float tx = readlane_f(pv.tx, q);
                const float w[4] = {readlane_f(pv.w[0], q), readlane_f(pv.w[1], q), readlane_f(pv.w[2], q),
                                    readlane_f(pv.w[3], q)};
                const float* rows = s_S + im;
#pragma unroll
                for (int u = 0; u < kBEPL; ++u) P[u] += (double)nsx_eval(le[u] + tx, ax, rows, w);
                n_done += (eidx[0] >= 0) + (eidx[1] >= 0);
              } else {
                const float c = readlane_f(pv.w[0], q), z = readlane_f(pv.zf, q);
#pragma unroll
                for (int u = 0; u < kBEPL; ++u) {
                  const float Ee = Ef[u] * z;
                  P[u] += (double)(c * (Ee * Ee * Ee / expm1f(Ee * inv_kT)));
                }
                n_done += (eidx[0] >= 0) + (eidx[1] >= 0);
              }
            }
          }
          while (ev < n_ev) fire();  // window ends past the last covered position
        }
      }
    }
  }
  // evaluations performed (statistics)
  unsigned long long tot = n_done;
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, 64);
  if (lane == 0) atomicAdd(B.n_evals, tot);
}

// flux[s][b][e] = sum over the star's items (in item order) of
// scale(spot) * acc / E.
__global__ void k_reduce_batched(BatchLaunch B) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)B.n_phase * B.n_e;
  if (idx >= per * B.n_stars) return;
  const int s = (int)(idx / per);
  const size_t o = idx - (size_t)s * per;
  double total = 0.0;
  for (int it = B.star_item0[s]; it < B.star_item0[s + 1]; ++it)
    total += B.spots[B.items[it].spot].scale * B.acc[(size_t)it * per + o];
  B.flux[idx] = total / B.energies[o % B.n_e];
}

}  // namespace

size_t batch_lds_bytes(int n_fine, int atmosphere, int nsx_rows, int cov_words, int ev_cap) {
  return batch_layout(n_fine + 1, atmosphere == PPM_ATM_NSX ? nsx_rows : 0, cov_words, ev_cap).total;
}

int batch_threads() { return kBThreads; }
int batch_energies() { return kBEnergies; }

int batched_prepare(int atmosphere, size_t lds) {
  const void* f = atmosphere == PPM_ATM_NSX ? (const void*)k_integral_batched<PPM_ATM_NSX>
                                            : (const void*)k_integral_batched<PPM_ATM_BLACKBODY>;
  return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

int launch_batched(const DevNsx& N, const BatchLaunch& B, int n_wg, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, B.queue, B.n_evals);
  const size_t lds = batch_lds_bytes(B.n_fine, B.atmosphere, B.nsx_rows, B.cov_words, B.ev_cap);
  if (B.atmosphere == PPM_ATM_NSX)
    hipLaunchKernelGGL(k_integral_batched<PPM_ATM_NSX>, dim3(n_wg), dim3(kBThreads), lds, st, N, B);
  else
    hipLaunchKernelGGL(k_integral_batched<PPM_ATM_BLACKBODY>, dim3(n_wg), dim3(kBThreads), lds, st, N, B);
  return hipGetLastError();
}

int launch_reduce_batched(const BatchLaunch& B, void* stream) {
  const size_t n = (size_t)B.n_phase * B.n_e * B.n_stars;
  hipLaunchKernelGGL(k_reduce_batched, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, B);
  return hipGetLastError();
}

int batched_occupancy(int atmosphere, size_t lds, int* wg_per_cu) {
  int n = 0;
  hipError_t e = atmosphere == PPM_ATM_NSX
                     ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_integral_batched<PPM_ATM_NSX>, kBThreads, lds)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_integral_batched<PPM_ATM_BLACKBODY>,
                                                                    kBThreads, lds);
  *wg_per_cu = n;
  return e;
}

}  // namespace ppm
