// Batched integral for gfx950: many stars in one persistent launch.
//
// Work queue: items = (star, spot, contiguous range of ring tasks), taken
// largest-first, either by persistent workgroups through an atomic counter
// or one item per workgroup in dispatch order (workgroups then retire every
// item, and the device can interleave another stream's launch among them).
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

#include <algorithm>
#include <cstdint>
#include <cstring>

#include "ppm_batch_dev.hpp"
#include "ppm_eval.hpp"

namespace ppm {

namespace {

// Phase timers (a build with PPM_PHASE_TIMERS=1 only): each wavefront adds
// the shader-clock cycles it spends in each phase of the integral and posts
// its totals once at the end; the default build compiles them out.
#ifndef PPM_PHASE_TIMERS
#define PPM_PHASE_TIMERS 0
#endif
enum BatchPhase {
  kPhFetch,     // taking the next work item (two barriers and the queue)
  kPhStage,     // ring set-up: fine tables and NSX slice rows into LDS
  kPhEvents,    // windows -> events and the coverage bitmap
  kPhSort,      // rank sort of the events
  kPhVisible,   // dropping covered positions whose patch is not seen
  kPhSplit,     // the split into segments and their events
  kPhSweep,     // the sweep (includes the next two)
  kPhSetup,     //   FP64 patch set-up of each batch of 64 positions
  kPhEnds,      //   window ends posted during the sweep
  kPhOffsets,   // segment offsets posted on later segments' closing events, and the closing barrier
  kPhTotal,     // the whole kernel
  kNumPhases
};
static_assert(kNumPhases + 1 == PPM_PHASE_COUNT, "ppm_phase_timers reports every phase and the wavefront count");
#if PPM_PHASE_TIMERS
__device__ unsigned long long g_phase[kNumPhases + 1];  // + the number of wavefronts
#define PT_DECL unsigned long long pt_[kNumPhases] = {}, pt_t = clock64(), pt_0 = pt_t;
#define PT_MARK() (pt_t = clock64())
#define PT_ADD(ph)                 \
  do {                             \
    const unsigned long long t_ = clock64(); \
    pt_[ph] += t_ - pt_t;          \
    pt_t = t_;                     \
  } while (0)
#define PT_SCOPE_BEGIN(v) const unsigned long long v = clock64();
#define PT_SCOPE_END(v, ph) (pt_[ph] += clock64() - (v))
#define PT_POST()                                                              \
  do {                                                                         \
    pt_[kPhTotal] = clock64() - pt_0;                                          \
    if (lane == 0) {                                                           \
      for (int i_ = 0; i_ < kNumPhases; ++i_) atomicAdd(&g_phase[i_], pt_[i_]); \
      atomicAdd(&g_phase[kNumPhases], 1ull);                                   \
    }                                                                          \
  } while (0)
#else
#define PT_DECL
#define PT_MARK() ((void)0)
#define PT_ADD(ph) ((void)0)
#define PT_SCOPE_BEGIN(v)
#define PT_SCOPE_END(v, ph) ((void)0)
#define PT_POST() ((void)0)
#endif

constexpr int kBW = 4;                  // wavefronts per workgroup, each a segment of the positions
constexpr int kBWP = kBW;
constexpr int kBThreads = 64 * kBW;
// Energies per lane: 2, 4 or 6 (128, 256 or 384 energies per work item, the
// host's choice by energy count and row span); a lane's energies form two
// row sets in every case.
constexpr int kMaxEPL = 6;
constexpr int kBSets = 2;

// A window end on the position axis: at pos, acc[bin] += sw * prefix.
// pb = pos << 12 | bin (pos <= 4 N_side <= 32768, bin < 4096), and, on a
// window's closing event (sw > 0), the sweep segment of its opening event in
// bits 28-29 (set once the segments are known).
struct EvRec {
  uint32_t pb;
  float sw;
};
constexpr int kEvSegShift = 28;
__device__ inline int ev_position(uint32_t pb) { return (int)((pb >> 12) & 0xffffu); }
__device__ inline int ev_open_segment(uint32_t pb) { return (int)((pb >> kEvSegShift) & 3u); }

// Per-wavefront LDS of the sweep: the evaluation scratch and a queue of
// covered positions (at most 63 pending + 64 new).
constexpr int kQueueCap = 128;
constexpr int kSweepBytes = kWaveScratchBytes + kQueueCap * 4;

// Position of covered position number c (0-based, ascending) in the coverage
// bitmap, or n when fewer than c + 1 positions are covered.  Wave-uniform.
__device__ inline int covered_position(const uint32_t* s_cov, int n_words, int c, int n, int lane) {
  int before = 0;
  for (int w0 = 0; w0 < n_words; w0 += 64) {
    const int w = w0 + lane;
    const uint32_t word = w < n_words ? s_cov[w] : 0u;
    const int pc = __builtin_popcount(word);
    int incl = pc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(incl, d);
      if (lane >= d) incl += v;
    }
    const int chunk = __shfl(incl, 63);
    if (before + chunk > c) {
      const int L = __builtin_ctzll(__ballot(before + incl > c));
      uint32_t wd = (uint32_t)__shfl((int)word, L);
      for (int k = c - before - __shfl(incl - pc, L); k > 0; --k) wd &= wd - 1;
      return (w0 + L) * 32 + __builtin_ctz(wd);
    }
    before += chunk;
  }
  return n;
}

// Number of covered positions.  Wave-uniform.
__device__ inline int covered_count(const uint32_t* s_cov, int n_words, int lane) {
  int c = 0;
  for (int w = lane; w < n_words; w += 64) c += __builtin_popcount(s_cov[w]);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d);
  return c;
}

// Number of (sorted) events before position p.  Wave-uniform.
__device__ inline int events_before(const EvRec* s_ev, int n_ev, int p, int lane) {
  int c = 0;
  for (int i = lane; i < n_ev; i += 64) c += ev_position(s_ev[i].pb) < p;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d);
  return c;
}

struct BLds {
  size_t ph, mu, hidx, grp, fI, zf, cs, S, cov, keys, sorted, total;
};

// LDS of the nm mu nodes and their [nm][4] inverse Lagrange denominators
// (the layout of DevNsx::mu), and the bucket indexes of the mu nodes and of
// the ring's phase table.

__host__ __device__ inline BLds batch_layout(int nf1, int nm, int nsx_rows, int cov_words, int ev_cap) {
  BLds l;
  size_t o = 0;
  l.ph = o;
  o += (size_t)nf1 * 8;
  l.mu = o;
  o += (size_t)nm * 5 * 8;
  l.hidx = o;
  o += 2 * (kHuntBuckets + 1) * sizeof(uint16_t);
  o = (o + 7) & ~(size_t)7;
  l.grp = o;  // phase offset of each class slot, then the slot of each point of a group
  o += kMaxSlots * 8 + (size_t)ev_cap / 4;
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
  l.keys = o;  // event keys; during the sweep, each wavefront's scratch and queue
  o += max((size_t)ev_cap * 8, (size_t)kBW * kSweepBytes);
  l.sorted = o;
  o += (size_t)ev_cap * sizeof(EvRec);
  l.total = (o + 15) & ~(size_t)15;
  return l;
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void k_queue_reset(int* q, unsigned long long* n_evals) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *q = 0;
    *n_evals = 0ull;
  }
}

// FL: the ring's fine-phase tables are staged in LDS (FL) or read from
// global memory where staging them would cost a resident workgroup per CU
// (large n_fine); the host picks the instance with the higher occupancy.
template <int ATM, int EPL, int EM, bool FL>
__global__ __launch_bounds__(kBThreads, 3) void k_integral_batched(DevNsx N, BatchLaunch B) {
  constexpr int kBEPL = EPL;
  constexpr int kBEnergies = 64 * EPL;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wp = wave;  // position part
  extern __shared__ __align__(16) unsigned char smem[];
  const int rows_alloc = ATM == PPM_ATM_NSX ? B.nsx_rows : 0;
  const BLds lay = batch_layout(FL ? B.n_fine_max + 1 : 0, ATM == PPM_ATM_NSX ? N.nm : 0, rows_alloc, B.cov_words,
                                B.ev_cap);
  double* s_ph = reinterpret_cast<double*>(smem + lay.ph);
  double* s_mu = reinterpret_cast<double*>(smem + lay.mu);
  uint16_t* s_mu_idx = reinterpret_cast<uint16_t*>(smem + lay.hidx);
  uint16_t* s_ph_idx = s_mu_idx + kHuntBuckets + 1;
  double* s_co = reinterpret_cast<double*>(smem + lay.grp);
  uint8_t* s_pslot = reinterpret_cast<uint8_t*>(smem + lay.grp + kMaxSlots * 8);
  float* s_fI = reinterpret_cast<float*>(smem + lay.fI);
  float* s_zf = reinterpret_cast<float*>(smem + lay.zf);
  float* s_cs = reinterpret_cast<float*>(smem + lay.cs);
  float* s_S = reinterpret_cast<float*>(smem + lay.S);
  uint32_t* s_cov = reinterpret_cast<uint32_t*>(smem + lay.cov);
  uint64_t* s_keys = reinterpret_cast<uint64_t*>(smem + lay.keys);
  EvRec* s_ev = reinterpret_cast<EvRec*>(smem + lay.sorted);
  float4* s_C = reinterpret_cast<float4*>(smem + lay.keys + wave * kSweepBytes);
  int* s_q = reinterpret_cast<int*>(smem + lay.keys + wave * kSweepBytes + kWaveScratchBytes);
  __shared__ int s_item, s_nev;
  __shared__ int s_cut[kBWP];  // first covered position of each sweep segment

  HuntIndex MU{nullptr, 0.0, 0.0};
  if (ATM == PPM_ATM_NSX) {
    for (int k = tid; k < 5 * N.nm; k += kBThreads) s_mu[k] = N.mu[k];  // nodes, then denominators
    __syncthreads();
    MU = build_hunt_index(s_mu, N.nm, s_mu_idx, tid, kBThreads);  // read after the item loop's barrier
  }

  const double h = (N.logE_max - N.logE_min) / (N.nE - 1);
  const double inv_h = 1.0 / h;
  unsigned int n_done = 0;  // evaluations by this lane
  PT_DECL

  for (int round = 0;; ++round) {
    PT_MARK();
    __syncthreads();
    if (tid == 0) s_item = B.persistent ? atomicAdd(B.queue, 1) : (round == 0 ? (int)blockIdx.x : B.n_items);
    __syncthreads();
    const int q_idx = s_item;
    if (q_idx >= B.n_items) break;
    PT_ADD(kPhFetch);
    const BatchItem It = B.items[B.order[q_idx]];
    const BatchGrid G = B.grids[It.grid];
    const int n_e = G.n_e, n_phase = G.n_phase, nf1 = G.n_fine + 1;
    const double* energies = G.energies;
    // one accumulator copy per position part: each address has one writer;
    // a copy holds the item's own energies, rows of It.en
    const int en = It.en;
    double* acc = B.acc + It.acc + (size_t)wp * n_phase * en;
    // energies of this thread: e = e0 + u*64 + lane (ascending grid), valid
    // while u*64 + lane <= e_lim
    const int e_hi = min(It.e0 + kBEnergies, n_e) - 1, e_lim = e_hi - It.e0;
    const double l_ref = log10(energies[It.e0]);
    float le[kBEPL], Ef[kBEPL];
#pragma unroll
    for (int u = 0; u < kBEPL; ++u) {
      const double E = energies[It.e0 + min(u * 64 + lane, e_lim)];
      Ef[u] = (float)E;
      le[u] = (float)((log10(E) - l_ref) * inv_h);
    }
    const bool asc = energies_ascend<kBEPL>(le, lane);
    unsigned int n_valid = 0;  // evaluations per position by this lane
#pragma unroll
    for (int u = 0; u < kBEPL; ++u) n_valid += u * 64 + lane <= e_lim;
    const WaveEnergies<kBSets> we = wave_energies<kBEPL, kBSets>(le, lane);
    for (int b = 0; b < n_phase; ++b)
#pragma unroll
      for (int u = 0; u < kBEPL; ++u)
        if (u * 64 + lane <= e_lim) acc[(size_t)b * en + u * 64 + lane] = 0.0;

    for (int task = It.task0; task < It.task1; ++task) {
      const RingDesc* Dp = B.desc + task;
      if (!Dp->kept) continue;
      const RingTask T = B.tasks[task];
      const SpotParams& S = B.spots[T.spot];
      PT_MARK();
      __syncthreads();  // previous ring's LDS is no longer read
      const double* gph = G.ph + (size_t)(task - G.task0) * nf1;
      const float* gf = G.fine + (size_t)(task - G.task0) * 3 * nf1;
      if (FL)
        for (int k = tid; k < nf1; k += kBThreads) {
          s_ph[k] = gph[k];
          s_fI[k] = gf[k];
          s_zf[k] = gf[nf1 + k];
          s_cs[k] = gf[2 * nf1 + k];
        }
      int r_lo = 0, n_rows = 0;
      EAxis ax{0.f, 0.f, 0, 0, 1};
      if (ATM == PPM_ATM_NSX) {
        const double lkT = log10(S.kT_nsx);
        const double t_min = (l_ref + log10(Dp->zf_min) - lkT - N.logE_min) * inv_h;
        const double t_max = (log10(energies[e_hi]) + log10(Dp->zf_max) - lkT - N.logE_min) * inv_h;
        r_lo = clampi((int)floor(t_min - 0.01) - 1, 0, N.nE - 4);
        const int r_hi = clampi((int)floor(t_max + 0.01) - 1, 0, N.nE - 4) + 3;
        n_rows = r_hi - r_lo + 1;
        if (n_rows > rows_alloc) {
          if (tid == 0) {
            atomicOr(&B.desc[task].flags, 2);
            atomicOr(B.rows_flag, 2ull);
          }
          n_rows = rows_alloc;
        }
        ax = energy_axis<EM>(N.nE, r_lo, n_rows, rows_alloc);
        stage_nsx_rows(N, *Dp, r_lo, n_rows, s_S, tid, kBThreads);
      }
      __syncthreads();
      PT_ADD(kPhStage);
      const TxMap txm{(float)((l_ref - log10(S.kT_nsx) - N.logE_min) * inv_h - r_lo),
                      (float)(inv_h * 0.30102999566398120)};
      const float inv_kT = (float)S.inv_kT;
      const int mu_mode = nsx_mu_mode(S.nsx_interp);
      const RingTab R = FL ? RingTab{s_ph, s_fI, s_zf, s_cs, nf1} : RingTab{gph, gf, gf + nf1, gf + 2 * nf1, nf1};
      const double ph0 = R.ph[0];
      // read after the first class group's barrier
      const HuntIndex PH = build_hunt_index(R.ph, nf1, s_ph_idx, tid, kBThreads);
      const bool point = T.ring == 0;
      const bool ring_hidden = Dp->hidden != kHiddenNone;
      const bool gap_interval = Dp->hidden == kHiddenInterval;
      const double gap_lo = Dp->gap_lo, gap_hi = Dp->gap_hi;
      const int n = point ? 1 : Dp->n_pix;
      const double inv_n = 1.0 / n;
      const int n_runs = point ? 1 : Dp->n_runs;
      const ClassSet cset = G.csets[point ? 0 : n / 4];
      const double c_shift = point ? Dp->ps_shift : 1.0 - 0.25 * Dp->s_shift2;

      // Classes of the ring are swept together, each in a slot of a virtual
      // position axis (slot s holds positions s n_pad + k, k < n; n_pad is n
      // rounded up to 64, so a 64-position chunk lies in one slot).  A window
      // never leaves its slot, so one running prefix serves every class of a
      // group.  A group holds whole classes while their slots fit
      // kGroupPositions and their points grp_pts (rings of at most
      // kGroupRuns runs), or one class, or the points [g0, g1) of one class
      // too large for the event list (pt_cap).  The grouping depends on the
      // ring and the class table only.
      const int n_pad = (n + 63) & ~63;
      const int slot_cap = max(1, min(kMaxSlots, kGroupPositions / n_pad));
      const int pt_cap = max(1, B.ev_cap / (4 * n_runs));
      const int grp_pts = n_runs <= kGroupRuns ? B.ev_cap / (4 * kGroupRuns) : 0;
      const float inv_npad = 1.f / (float)n_pad;
      const int c_end = min(It.c1, cset.n_classes);
      for (int ci = It.c0, g0 = 0; ci < c_end;) {
        const ClassHdr C = G.chdr[cset.first_class + ci];
        int cj = ci + 1, g1 = C.n_pts, n_pts = C.n_pts - g0;
        if (n_pts > pt_cap) {
          g1 = g0 + pt_cap;
          n_pts = pt_cap;
        } else if (g0 == 0) {  // (the tail of a class swept in chunks stays alone)
          while (cj < c_end && cj - ci < slot_cap) {
            const int more = G.chdr[cset.first_class + cj].n_pts;
            if (n_pts + more > grp_pts) break;
            n_pts += more;
            ++cj;
          }
        }
        const int n_slots = cj - ci;
        const int pt0 = C.first_pt + g0;  // the classes' points are consecutive in the table
        const int nv = n_slots * n_pad;
        const int n_words = nv / 32;  // coverage words (64-position chunks)
        const int n_win = n_pts * n_runs;
        // rotational phase of position k of slot s, wrapped into [ph[0], ph[0] + 1)
        auto phase_of = [&](int k, int s) {
          const double v = (k + s_co[s]) * inv_n + 1.0;
          double pp = v - floor(v);
          if (pp < ph0) pp += 1.;
          return pp;
        };
        // ---- windows -> events and coverage
        PT_MARK();
        __syncthreads();
        for (int w = tid; w < n_words; w += kBThreads) s_cov[w] = 0u;
        if (tid == 0) s_nev = 0;
        for (int s = 0, off = 0; s < n_slots; ++s) {  // slot of each point, phase offset of each slot
          const ClassHdr Cs = G.chdr[cset.first_class + ci + s];
          const int cnt = s == 0 ? g1 - g0 : Cs.n_pts;
          for (int q = tid; q < cnt; q += kBThreads) s_pslot[off + q] = (uint8_t)s;
          if (tid == 0) s_co[s] = Cs.delta + c_shift;
          off += cnt;
        }
        __syncthreads();
        for (int wi = tid; wi < n_win; wi += kBThreads) {
          const int ql = wi / n_runs, r = wi - ql * n_runs;
          const ClassPt cp = G.cpts[pt0 + ql];
          const int vb = s_pslot[ql] * n_pad;  // the slot's first virtual position
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
            const int lo = vb + pieces[pc][0], hi = vb + pieces[pc][1];
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
        PT_ADD(kPhEvents);
        // ---- rank sort of the events (keys are unique) into records.  The
        // keys 2p, 2p + 1 are the two ends of one window piece: a thread
        // ranks both in one pass and keeps its first pair's ranks and
        // positions for the segment flags below
        int pr_r0 = -1, pr_r1 = -1, pr_p0 = 0, pr_p1 = 0;
        for (int p = tid; 2 * p < n_ev; p += kBThreads) {
          const uint64_t k0 = s_keys[2 * p], k1 = s_keys[2 * p + 1];
          int r0 = 0, r1 = 0;
          for (int j = 0; j < n_ev; ++j) {
            const uint64_t kj = s_keys[j];
            r0 += kj < k0;
            r1 += kj < k1;
          }
          const ClassPt cp = G.cpts[pt0 + (int)((k0 >> 12) & 0xfffff)];
          const int p0 = (int)(k0 >> 32), p1 = (int)(k1 >> 32);
          s_ev[r0] = EvRec{(uint32_t)p0 << 12 | (uint32_t)cp.bin, (float)-cp.w};  // window start: -w
          s_ev[r1] = EvRec{(uint32_t)p1 << 12 | (uint32_t)cp.bin, (float)cp.w};   // window end: +w
          if (p == tid) {
            pr_r0 = r0;
            pr_r1 = r1;
            pr_p0 = p0;
            pr_p1 = p1;
          }
        }
        PT_ADD(kPhSort);
        // ---- keep the covered positions whose patch is seen (flux/I != 0):
        // the others add nothing, so the split, the queue and the set-up
        // only count patches that contribute.  Each wavefront rewrites its
        // own 64-position chunks.  A ring without invisible fine nodes has
        // flux/I > 0 between all of them: nothing to drop.
        for (int base = wave * 64; ring_hidden && base < nv; base += kBThreads) {
          const int v = base + lane, s = base / n_pad, k = v - s * n_pad;
          bool keep = false;
          if (k < n && ((s_cov[v >> 5] >> (v & 31)) & 1u)) {
            const double pp = phase_of(k, s);
            if (gap_interval) {  // the set-up skips the rare kept patch whose flux/I rounds to 0
              keep = !(pp >= gap_lo && pp <= gap_hi);
            } else {
              int fi;
              double fa;
              keep = patch_flux_factor(pp, R, fi, fa, PH) != 0;
            }
          }
          const uint64_t m = __ballot(keep);
          if (lane == 0) s_cov[base >> 5] = (uint32_t)m;
          if (lane == 32) s_cov[(base >> 5) + 1] = (uint32_t)(m >> 32);
        }
        __syncthreads();

        PT_ADD(kPhVisible);
        // ---- sweep.  The covered positions are split in kBWP segments of
        // equal count, [seg_lo, seg_hi) for this wavefront; it takes the
        // events in its segment with a prefix that starts at zero, and the
        // sum of the earlier segments is added to those events afterwards.
        int seg_lo = 0, seg_hi = 0x7fffffff;
        if (kBWP > 1) {
          const int n_cov = covered_count(s_cov, n_words, lane);
          if (wp > 0) seg_lo = covered_position(s_cov, n_words, n_cov * wp / kBWP, nv, lane);
          if (wp < kBWP - 1) seg_hi = covered_position(s_cov, n_words, n_cov * (wp + 1) / kBWP, nv, lane);
          // each window piece's closing event records the segment of its
          // opening event: the segments in between owe it their totals
          // (the offset pass below)
          if (lane == 0) s_cut[wp] = seg_lo;
          __syncthreads();
          if (pr_r0 >= 0) {
            int s0 = 0;
#pragma unroll
            for (int w = 1; w < kBWP; ++w) s0 += pr_p0 >= s_cut[w];
            s_ev[pr_r1].pb |= (uint32_t)s0 << kEvSegShift;
          }
          __syncthreads();
        }
        const int ev_begin = wp > 0 ? events_before(s_ev, n_ev, seg_lo, lane) : 0;
        const int ev_end = wp < kBWP - 1 ? events_before(s_ev, n_ev, seg_hi, lane) : n_ev;
        int ev_stop = ev_end;  // events taken by load_ev: this segment's, then the later ones
        PT_ADD(kPhSplit);
        double P[kBEPL];
#pragma unroll
        for (int u = 0; u < kBEPL; ++u) P[u] = 0.0;
        // 64 events at a time are held one per lane; the next one is read
        // into scalar registers.  Contributions are posted as no-return
        // FP64 atomics to the item's private accumulator (one writer per
        // address: deterministic)
        int ev = ev_begin, win = ev_begin - 64;
        uint32_t v_pb = 0;
        float v_sw = 0.f;
        int ev_pos = 0x7fffffff, ev_bin = 0, ev_open = 0;
        double ev_sw = 0.0;
        auto load_ev = [&]() {
          if (ev < ev_stop) {
            if (ev - win >= 64) {
              win = ev;
              if (ev + lane < n_ev) {
                const EvRec r = s_ev[ev + lane];
                v_pb = r.pb;
                v_sw = r.sw;
              }
            }
            const uint32_t pb = (uint32_t)readlane_i((int)v_pb, ev - win);
            ev_pos = ev_position(pb);
            ev_bin = (int)(pb & 4095u);
            ev_open = ev_open_segment(pb);
            ev_sw = (double)readlane_f(v_sw, ev - win);
          } else {
            ev_pos = 0x7fffffff;
          }
        };
        auto fire = [&]() {
          PT_SCOPE_BEGIN(pt_f)
          double* const row = acc + lane + ev_bin * en;  // the energies' offsets are immediates
#pragma unroll
          for (int u = 0; u < kBEPL; ++u)
            if (u * 64 + lane <= e_lim) unsafeAtomicAdd(row + u * 64, ev_sw * P[u]);
          ++ev;
          load_ev();
          PT_SCOPE_END(pt_f, kPhEnds);
        };
        load_ev();
        // covered positions, in ascending order, 64 per batch: one patch
        // set-up per lane, then every lane evaluates each visible patch
        auto run_batch = [&](int cnt, int q_len) {
          PT_SCOPE_BEGIN(pt_s)
          const int k = lane < cnt ? s_q[lane] : 0;  // virtual position
          PatchVals pv;
          pv.im = -1;
          if (lane < cnt) {
            const int s = n_slots == 1 ? 0 : (int)(((float)k + 0.5f) * inv_npad);
            patch_setup<ATM>(phase_of(k - s * n_pad, s), 1.0, R, s_mu, s_mu + N.nm, N.nm, N.mu_min, mu_mode, txm, pv,
                             PH, MU);
          }
          // skipped patches add nothing to the prefix: the events between
          // two queued positions fire with the same prefix
          uint64_t todo = __ballot(lane < cnt && pv.im >= 0);
          n_done += n_valid * (unsigned int)__builtin_popcountll(todo);  // every visible patch is evaluated
          uint64_t due = __ballot(k >= ev_pos);
          PatchRows<kBSets> pr{};
          if (ATM == PPM_ATM_NSX && pv.im >= 0) pr = patch_rows(pv.tx, pv.im, we, ax);
          // patches evaluated without clamps, and those whose low row set
          // takes the clamped evaluation (the rest clamp the high set only:
          // energies beyond the end of the axis)
          static_assert(kBSets == 2, "the clamp classes assume two row sets");
          const uint64_t interior = __ballot(lane < cnt && pv.im >= 0 && pr.clamp == 0);
          const uint64_t clamp_lo = __ballot(lane < cnt && pv.im >= 0 && (pr.clamp & 1));
          // FP32 partial sums of at most one batch, folded into the FP64
          // prefix before an event fires and at the end of the batch
          float pf[kBEPL];
#pragma unroll
          for (int u = 0; u < kBEPL; ++u) pf[u] = 0.f;
          auto fold = [&]() {
#pragma unroll
            for (int u = 0; u < kBEPL; ++u) {
              P[u] += (double)pf[u];
              pf[u] = 0.f;
            }
          };
          // in position order: window ends at or before pos use the
          // prefix before F(pos); `due` marks the lanes whose position is
          // at or after the next window end
          auto accumulate = [&](int q, const float (&val)[kBEPL]) {
            if (__builtin_expect((due >> q) & 1, 0)) {
              const int pos = readlane_i(k, q);
              fold();
              while (ev_pos <= pos) fire();
              due = __ballot(k >= ev_pos);
            }
#pragma unroll
            for (int u = 0; u < kBEPL; ++u) pf[u] += val[u];
          };
          PT_SCOPE_END(pt_s, kPhSetup);
          // one patch at a time.  With the row-set path the slice rows of
          // the next patch are read while the current one is evaluated (two
          // row buffers, used in turn); the path is chosen once per batch
          if (ATM == PPM_ATM_NSX && we.fast) {
            auto step_rows = [&](const float (&p_cur)[kBSets][4], float (&p_next)[kBSets][4]) {
              const int q = __builtin_ctzll(todo);
              todo &= todo - 1;
              if (todo) nsx_load_rows(pr, __builtin_ctzll(todo), s_S, lane, p_next);
              float val[kBEPL];
              if (__builtin_expect((interior >> q) & 1, 1))
                nsx_eval_table<kBEPL, 0, kBSets, EM>(le, pv, pr, q, ax, p_cur, lane, s_C, val);
              else if ((clamp_lo >> q) & 1)
                nsx_eval_table<kBEPL, 3, kBSets, EM>(le, pv, pr, q, ax, p_cur, lane, s_C, val);
              else
                nsx_eval_table<kBEPL, 2, kBSets, EM>(le, pv, pr, q, ax, p_cur, lane, s_C, val);
              accumulate(q, val);
            };
            float p_a[kBSets][4], p_b[kBSets][4];
            if (todo) nsx_load_rows(pr, __builtin_ctzll(todo), s_S, lane, p_a);
            while (todo) {
              step_rows(p_a, p_b);
              if (!todo) break;
              step_rows(p_b, p_a);
            }
          } else {
            while (todo) {
              const int q = __builtin_ctzll(todo);
              todo &= todo - 1;
              float val[kBEPL];
              if (ATM == PPM_ATM_NSX) {
                float w[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) w[m] = readlane_f(pv.w[m], q);
                nsx_eval_edge<kBEPL, EM>(le, asc, readlane_f(pv.tx, q), readlane_i(pv.im, q), w, ax, s_S, s_C,
                                         lane, val);
              } else {
                const float c = readlane_f(pv.w[0], q), z = readlane_f(pv.zf, q);
#pragma unroll
                for (int u = 0; u < kBEPL; ++u) {
                  const float Ee = Ef[u] * z;
                  val[u] = c * (Ee * Ee * Ee / expm1f(Ee * inv_kT));
                }
              }
              accumulate(q, val);
            }
          }
          fold();
          // drop the batch from the queue
          const int rest = q_len - cnt;
          const int moved = lane < rest ? s_q[cnt + lane] : 0;
          __builtin_amdgcn_wave_barrier();
          if (lane < rest) s_q[lane] = moved;
          __builtin_amdgcn_wave_barrier();
          return rest;
        };
        int q_len = 0;
        const int sweep_end = min(nv, seg_hi);
        for (int base = seg_lo & ~63;;) {
          const bool more = base < sweep_end;
          if (more) {  // queue the covered positions of the next 64
            // (readfirstlane returns a signed int: widened through uint32_t,
            // so a low word with bit 31 set does not fill the high word)
            const uint32_t w_lo = (uint32_t)__builtin_amdgcn_readfirstlane(s_cov[base >> 5]);
            const uint32_t w_hi = (uint32_t)__builtin_amdgcn_readfirstlane(s_cov[(base >> 5) + 1]);
            uint64_t cov = (uint64_t)w_lo | ((uint64_t)w_hi << 32);
            if (seg_lo > base) cov &= ~0ull << (seg_lo - base);
            if (sweep_end - base < 64) cov &= (1ull << (sweep_end - base)) - 1;
            const uint32_t c_lo = (uint32_t)cov, c_hi = (uint32_t)(cov >> 32);
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(c_hi, __builtin_amdgcn_mbcnt_lo(c_lo, 0u));
            const bool mine = lane < 32 ? (c_lo >> lane) & 1u : (c_hi >> (lane - 32)) & 1u;
            if (mine) s_q[q_len + (int)below] = base + lane;
            __builtin_amdgcn_wave_barrier();
            q_len += __builtin_popcount(c_lo) + __builtin_popcount(c_hi);
            base += 64;
          }
          if (q_len >= 64 || (!more && q_len > 0))
            q_len = run_batch(min(q_len, 64), q_len);
          else if (!more)
            break;
        }
        while (ev < ev_end) fire();  // window ends past the last covered position
        PT_ADD(kPhSweep);
        if (kBWP > 1) {
          // segment offsets.  A window piece opening in segment a and
          // closing in segment b > a misses the totals of segments a..b-1
          // (each sweep's prefix starts at zero); P now holds this segment's
          // total, which this wavefront posts on the closing events of the
          // later segments whose window opened here or earlier, into its own
          // accumulator copy, with no wait for the other segments
          if (wp < kBWP - 1) {
            ev = ev_end;
            win = ev_end - 64;
            ev_stop = n_ev;
            load_ev();
            while (ev < n_ev) {
              if (ev_sw > 0.0 && ev_open <= wp) {
                fire();
              } else {
                ++ev;
                load_ev();
              }
            }
          }
          __syncthreads();  // the events are read before the next group writes them
          PT_ADD(kPhOffsets);
        }
        if (g1 < C.n_pts) {
          g0 = g1;
        } else {
          ci = cj;
          g0 = 0;
        }
      }
    }
  }
  PT_POST();
  // evaluations performed (statistics)
  unsigned long long tot = n_done;
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, 64);
  if (lane == 0) atomicAdd(B.n_evals, tot);
}

// flux[s - s0][b][e] = sum over the star's items whose energy chunk holds e
// of scale(spot) * (sum of the item's accumulator copies) / E, for the stars
// [s0, s1) of one grid.  One workgroup per (64 energies, bin, star) of
// n_waves = blockDim.x / 64 waves: wave w sums the items w, w + n_waves, ...
// of the star in item order and wave 0 adds the partial sums in wave order,
// so the result does not depend on timing.  The host takes 16 waves when the
// stars have many items (a star alone, its rings split by phase class: ~1500
// accumulators read in 0.5 ms against 1.0 with 4) and 4 otherwise.
constexpr int kReduceWavesMax = 16;
__global__ __launch_bounds__(64 * kReduceWavesMax) void k_reduce_batched(BatchLaunch B, int s0, int n_phase, int n_e,
                                                                         const double* energies, double* flux) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
  const int e = blockIdx.x * 64 + lane, b = blockIdx.y, s = s0 + (int)blockIdx.z;
  __shared__ double part[kReduceWavesMax][64];
  double total = 0.0;
  if (e < n_e) {
    for (int it = B.star_item0[s] + wave; it < B.star_item0[s + 1]; it += n_waves) {
      const BatchItem I = B.items[it];
      if (e < I.e0 || e >= I.e0 + I.en) continue;  // another chunk's item: it wrote none of e
      const size_t cp = (size_t)n_phase * I.en;
      const double* a = B.acc + I.acc + (size_t)b * I.en + (e - I.e0);
      double sum = 0.0;
#pragma unroll
      for (int c = 0; c < kBWP; ++c) sum += a[c * cp];
      total += B.spots[I.spot].scale * sum;
    }
  }
  part[wave][lane] = total;
  __syncthreads();
  if (wave == 0 && e < n_e) {
    double f = part[0][lane];
    for (int w = 1; w < n_waves; ++w) f += part[w][lane];
    flux[((size_t)(s - s0) * n_phase + b) * n_e + e] = f / energies[e];
  }
}

}  // namespace

size_t batch_lds_bytes(int n_fine, int atmosphere, int nm, int nsx_rows, int cov_words, int ev_cap, bool fine_lds) {
  const bool nsx = atmosphere == PPM_ATM_NSX;
  return batch_layout(fine_lds ? n_fine + 1 : 0, nsx ? nm : 0, nsx ? nsx_rows : 0, cov_words, ev_cap).total;
}

int batch_threads() { return kBThreads; }

bool batch_phase_timers(unsigned long long* cycles, int n, bool reset) {
#if PPM_PHASE_TIMERS
  unsigned long long h[kNumPhases + 1];
  if (hipMemcpyFromSymbol(h, HIP_SYMBOL(g_phase), sizeof(h)) != hipSuccess) return false;
  std::memcpy(cycles, h, sizeof(unsigned long long) * std::min(n, kNumPhases + 1));
  if (reset) {
    std::memset(h, 0, sizeof(h));
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_phase), h, sizeof(h)) != hipSuccess) return false;
  }
  return true;
#else
  (void)cycles;
  (void)n;
  (void)reset;
  return false;
#endif
}
int batch_acc_copies() { return kBWP; }
int batch_energies(int epl) { return 64 * epl; }

namespace {
template <int ATM, int EM, bool FL>
const void* batched_fn_epl(int epl) {
  switch (epl) {
    case 2: return (const void*)k_integral_batched<ATM, 2, EM, FL>;
    case 4: return (const void*)k_integral_batched<ATM, 4, EM, FL>;
    default: return (const void*)k_integral_batched<ATM, 6, EM, FL>;
  }
}
template <int ATM, int EM>
const void* batched_fn_fl(int epl, bool fine_lds) {
  return fine_lds ? batched_fn_epl<ATM, EM, true>(epl) : batched_fn_epl<ATM, EM, false>(epl);
}
// The kernel instance of a launch: atmosphere, energies per lane and the
// energy-axis mode of the NSX interpolant.
const void* batched_fn(int atmosphere, int epl, int emode, bool fine_lds) {
  if (atmosphere != PPM_ATM_NSX) return batched_fn_fl<PPM_ATM_BLACKBODY, kECubic>(epl, fine_lds);
  switch (emode) {
    case kECubicClip: return batched_fn_fl<PPM_ATM_NSX, kECubicClip>(epl, fine_lds);
    case kELinear: return batched_fn_fl<PPM_ATM_NSX, kELinear>(epl, fine_lds);
    default: return batched_fn_fl<PPM_ATM_NSX, kECubic>(epl, fine_lds);
  }
}
}  // namespace

// The dynamic LDS limit is set to the device's maximum (lds_max), the same
// value from every lane, so concurrent launches never lower it for another.
int batched_prepare(int atmosphere, int epl, int emode, bool fine_lds, int lds_max) {
  const void* fn = batched_fn(atmosphere, epl, emode, fine_lds);
  hipFuncAttributes fa{};
  const hipError_t e = hipFuncGetAttributes(&fa, fn);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max - (int)fa.sharedSizeBytes);
}

int launch_batched(const DevNsx& N, const BatchLaunch& B, int n_wg, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, B.queue, B.n_evals);
  const size_t lds =
      batch_lds_bytes(B.n_fine_max, B.atmosphere, N.nm, B.nsx_rows, B.cov_words, B.ev_cap, B.fine_lds);
  DevNsx n_arg = N;
  BatchLaunch b_arg = B;
  void* args[] = {&n_arg, &b_arg};
  const hipError_t e =
      hipLaunchKernel(batched_fn(B.atmosphere, B.epl, B.emode, B.fine_lds), dim3(n_wg), dim3(kBThreads), args, lds, st);
  return e != hipSuccess ? e : hipGetLastError();
}

int launch_zero_u64(unsigned long long* p, int n, void* stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, p, n);
  return hipGetLastError();
}

int launch_reduce_batched(const BatchLaunch& B, const BatchGrid& G, int s0, int s1, double* flux, void* stream) {
  if (s1 <= s0 || G.n_phase < 1 || G.n_e < 1) return hipSuccess;
  const int waves = B.n_items >= 64 * B.n_stars ? kReduceWavesMax : 4;
  constexpr int kMaxZ = 65535;  // grid z limit: stars per launch
  for (int a = s0; a < s1; a += kMaxZ) {
    const int n = min(kMaxZ, s1 - a);
    const dim3 grid((unsigned)((G.n_e + 63) / 64), (unsigned)G.n_phase, (unsigned)n);
    hipLaunchKernelGGL(k_reduce_batched, grid, dim3(64 * waves), 0, (hipStream_t)stream, B, a, G.n_phase,
                       G.n_e, G.energies, flux + (size_t)(a - s0) * G.n_phase * G.n_e);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Resident workgroups per CU: the runtime's figure, limited by the LDS a
// CU holds at the allocation granule measured on gfx950 (kLdsGranule): the
// runtime counts finer granules, and a launch sized to its figure got 2 of
// the 3 workgroups it reported (54064 bytes per workgroup ran at the speed
// of 2 per CU, 53552 at 3).
constexpr size_t kLdsGranule = 1280;
int batched_occupancy(int atmosphere, int epl, int emode, bool fine_lds, size_t lds, int* wg_per_cu) {
  const void* fn = batched_fn(atmosphere, epl, emode, fine_lds);
  int n = 0, dev = 0, lds_cu = 0;
  hipFuncAttributes fa{};
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kBThreads, lds);
  if (e == hipSuccess) e = hipFuncGetAttributes(&fa, fn);
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
  if (e == hipSuccess) {
    const size_t per_wg = (lds + fa.sharedSizeBytes + kLdsGranule - 1) / kLdsGranule * kLdsGranule;
    n = std::min(n, (int)((size_t)lds_cu / per_wg));
  }
  *wg_per_cu = n;
  return e;
}

}  // namespace ppm
