// Batched evaluation: many stars in one persistent launch, captured as a
// hipGraph (ppm_batch_kernels.hip holds the kernels).
#include "ppm_batch.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>

#include "ppm_batch_dev.hpp"

namespace ppm {

namespace {

// Phase points grouped by fractional offset for every pixel count of the
// grid: index 0 is the one-element point source, index t is n = 4t.
struct ClassTables {
  std::vector<ClassSet> sets;
  std::vector<ClassHdr> hdr;
  std::vector<ClassPt> pts;
};

void build_classes(int n_side, const std::vector<PhasePoint>& pp, ClassTables& T) {
  T.sets.assign(n_side + 1, ClassSet{0, 0});
  T.hdr.clear();
  T.pts.clear();
  const int np = (int)pp.size();
  std::vector<int> idx(np);
  std::vector<double> delta(np);
  std::vector<int> m(np);
  for (int t = 0; t <= n_side; ++t) {
    const int n = t == 0 ? 1 : 4 * t;
    for (int q = 0; q < np; ++q) {
      const double v = pp[q].p * n;
      double f = std::floor(v);
      double d = v - f;
      if (d > 1.0 - 1e-12) {  // an integer that rounded down
        d = 0.0;
        f += 1.0;
      }
      m[q] = (int)f % n;
      delta[q] = d;
    }
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
      if (std::fabs(delta[a] - delta[b]) > 1e-9) return delta[a] < delta[b];
      return m[a] < m[b];
    });
    T.sets[t].first_class = (int)T.hdr.size();
    int k = 0;
    while (k < np) {
      const double d0 = delta[idx[k]];
      ClassHdr h{d0, (int)T.pts.size(), 0};
      while (k < np && std::fabs(delta[idx[k]] - d0) <= 1e-9) {
        const int q = idx[k++];
        T.pts.push_back(ClassPt{m[q], pp[q].bin, pp[q].w});
        ++h.n_pts;
      }
      T.hdr.push_back(h);
    }
    T.sets[t].n_classes = (int)T.hdr.size() - T.sets[t].first_class;
  }
}

}  // namespace

// Class tables of one grid's phase points, on the host and the device.
struct ClassCache {
  ClassTables tab;
  DevBuf csets, chdr, cpts;
  uint64_t last_use = 0;
  int in_use = 0;  // launches in flight on some lane that read these tables
  void release() {
    csets.release();
    chdr.release();
    cpts.release();
  }
};

// The launch sequence is captured into a graph the second time the same
// configuration is launched in a row (a one-off configuration, as in
// serving, is launched directly) and replayed while it repeats.
struct GraphCache {
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned char> key, last_direct;
};

// Launch shape of the batched kernel: work-queue items per resident
// workgroup (how finely each spot's rings are split; more items balance the
// persistent workgroups better, fewer carry less per-item set-up) and
// resident workgroups per CU.
constexpr int kDefaultItemsPerSlot = 10;

// The last kTailChunks chunks of each spot's rings (of at least
// kTailChunks + 1) are cut kTailSplit times finer.  Items are taken
// largest-first, so the queue ends with small items and the persistent
// workgroups finish closer together.
constexpr int kTailChunks = 3;
constexpr int kTailSplit = 4;

// Caches shared by the lanes of a context (lanes run on several host
// threads, so each map has its mutex).
struct BatchState {
  std::mutex cls_mu;
  std::map<std::vector<int>, ClassCache> classes;  // per (n_side, n_phase, sampling, n_gl)
  uint64_t use_clock = 0;
  // shapes found by ppm_tune or set by the caller, per shape_key
  std::mutex shape_mu;
  std::map<std::vector<int>, ppm_launch_shape> shapes;
};

}  // namespace ppm

// Buffers and the graph of one lane's batched launches.
struct ppm::LaneBatch {
  ppm::GraphCache graph;
  ppm::DevBuf d_grids, d_order, d_star_item0, d_acc;
};

namespace ppm {

// The grid fields that shape the batched work, the atmosphere and the bit
// length of the star count.
static std::vector<int> shape_key(const ppm_grid& g, int n_e, int atm, int n) {
  int bits = 0;
  while (bits < 31 && (1 << bits) <= n) ++bits;
  return {g.n_side, g.n_phase, g.sampling, g.sampling == PPM_SAMPLE_GL ? g.n_gl : 0, g.n_fine, n_e, atm, bits};
}

static BatchState* state(ppm_ctx* c) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  if (!c->batch_state) c->batch_state = new BatchState();
  return c->batch_state;
}

static LaneBatch* lane_batch(Lane& L) {
  if (!L.lb) L.lb = new LaneBatch();
  return L.lb;
}

size_t lane_device_bytes(const Lane& L, size_t* fixed) {
  size_t b = 0;
  const Workspace& W = L.ws;
  for (const DevBuf* d : {&W.spots, &W.tasks, &W.desc, &W.ph, &W.fine, &W.pts, &W.energies, &W.partial, &W.red,
                          &W.scale, &W.out_flux, &W.out_counts, &W.fit_data, &W.fit_part, &W.fit_out})
    b += d->cap;
  const BatchWorkspace& B = L.batch;
  for (const DevBuf* d : {&B.spots, &B.tasks, &B.desc, &B.ph, &B.fine, &B.pts, &B.energies, &B.items, &B.scale,
                          &B.counter, &B.evals, &B.summary})
    b += d->cap;
  *fixed = 0;
  if (L.lb) {
    for (const DevBuf* d : {&L.lb->d_grids, &L.lb->d_order, &L.lb->d_star_item0, &L.lb->d_acc}) b += d->cap;
  }
  return b;
}

void release_lane_batch(Lane& L) {
  LaneBatch* B = L.lb;
  if (!B) return;
  if (B->graph.exec) hipGraphExecDestroy(B->graph.exec);
  for (DevBuf* b : {&B->d_grids, &B->d_order, &B->d_star_item0, &B->d_acc}) b->release();
  delete B;
  L.lb = nullptr;
}

// Fraction of a ring of colatitude th inside a disk (th1, rho), for cost
// estimates only.
// Fraction of a ring at z = cos(theta) within rho of a centre at th1, given
// sin th1, cos th1 and cos rho.
static double arc_fraction(double z, double st1, double ct1, double crho) {
  const double s = std::sqrt(std::max(0.0, 1.0 - z * z)) * st1;
  const double cth = crho - z * ct1;
  if (std::fabs(s) < 1e-15) return cth < 0 ? 1.0 : 0.0;
  const double x = cth / s;
  if (x <= -1) return 1.0;
  if (x >= 1) return 0.0;
  return std::acos(x) / kPi;
}

static double ring_z(int n_side, int i) {
  const bool north = i <= 2 * n_side;
  const int ii = north ? i : 4 * n_side - i;
  const double z = ii < n_side ? 1. - 1. * ii * ii / (3. * n_side * n_side) : 4. / 3. - 2. * ii / (3. * n_side);
  return north ? z : -z;
}

void release_batch_state(ppm_ctx* c) {
  BatchState* S = c->batch_state;
  if (!S) return;
  for (auto& kv : S->classes) kv.second.release();
  delete S;
  c->batch_state = nullptr;
}

static ppm_status eval_groups_rows(ppm_ctx* c, Lane& L, std::vector<BatchGroup>& groups, ppm_stats* stats,
                                   int rows_override);

ppm_status eval_batched(ppm_ctx* c, Lane& L, const ppm_star* stars, int n, const ppm_grid& g,
                        const std::vector<double>& energies, const std::vector<PhasePoint>& pts,
                        double* d_flux, double* d_counts, ppm_stats* stats) {
  // one launch per launch class present, each over the runs of consecutive
  // stars of its class
  std::vector<int> classes;
  for (int i = 0; i < n; ++i)
    if (std::find(classes.begin(), classes.end(), batch_class(stars[i])) == classes.end())
      classes.push_back(batch_class(stars[i]));
  const size_t fstride = (size_t)g.n_phase * energies.size();
  const size_t cstride = d_counts && g.response >= 0 ? (size_t)c->resp[g.response].n_ch * g.n_phase : 0;
  for (const int cl : classes) {
    std::vector<BatchGroup> groups;
    for (int i = 0; i < n;) {
      if (batch_class(stars[i]) != cl) {
        ++i;
        continue;
      }
      int j = i;
      while (j < n && batch_class(stars[j]) == cl) ++j;
      BatchGroup G;
      G.stars = stars + i;
      G.n = j - i;
      G.g = g;
      G.energies = energies;
      G.pts = pts;
      G.d_flux = d_flux + i * fstride;
      G.d_counts = d_counts ? d_counts + i * cstride : nullptr;
      G.fit_star0 = i;
      groups.push_back(std::move(G));
      i = j;
    }
    const ppm_status s = eval_groups_rows(c, L, groups, stats, 0);
    if (s) return s;
  }
  return PPM_OK;
}

ppm_status eval_batched_groups(ppm_ctx* c, Lane& L, std::vector<BatchGroup>& groups, ppm_stats* stats) {
  if (stats) std::memset(stats, 0, sizeof *stats);
  return eval_groups_rows(c, L, groups, stats, 0);
}

// Class tables of a grid's phase points, built and uploaded once per grid
// (kMaxClassGrids kept; the least recently used one not used by the current
// launch is dropped).
constexpr size_t kMaxClassGrids = 24;

// The caller holds BS->cls_mu; the entry returned is marked in use (the
// caller releases it when its launch has completed).
static ppm_status class_tables(ppm_ctx* c, BatchState* BS, const ppm_grid& g, const std::vector<PhasePoint>& pts,
                               ClassCache*& out) {
  const std::vector<int> key{g.n_side, g.n_phase, g.sampling, g.sampling == PPM_SAMPLE_GL ? g.n_gl : 0};
  auto it = BS->classes.find(key);
  if (it == BS->classes.end()) {
    if (BS->classes.size() >= kMaxClassGrids) {
      auto victim = BS->classes.end();
      for (auto v = BS->classes.begin(); v != BS->classes.end(); ++v)
        if (v->second.in_use == 0 &&
            (victim == BS->classes.end() || v->second.last_use < victim->second.last_use))
          victim = v;
      if (victim != BS->classes.end()) {
        victim->second.release();
        BS->classes.erase(victim);
      }
    }
    ClassCache& CC = BS->classes[key];
    build_classes(g.n_side, pts, CC.tab);
    const ClassTables& T = CC.tab;
    int e;
    if ((e = CC.csets.ensure(sizeof(ClassSet) * T.sets.size())) ||
        (e = CC.chdr.ensure(sizeof(ClassHdr) * std::max<size_t>(T.hdr.size(), 1))) ||
        (e = CC.cpts.ensure(sizeof(ClassPt) * std::max<size_t>(T.pts.size(), 1))) ||
        (e = hipMemcpy(CC.csets.p, T.sets.data(), sizeof(ClassSet) * T.sets.size(), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(CC.chdr.p, T.hdr.data(), sizeof(ClassHdr) * T.hdr.size(), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(CC.cpts.p, T.pts.data(), sizeof(ClassPt) * T.pts.size(), hipMemcpyHostToDevice))) {
      CC.release();
      BS->classes.erase(key);
      return hip_fail(c, e, "class tables");
    }
    it = BS->classes.find(key);
  }
  it->second.last_use = ++BS->use_clock;
  ++it->second.in_use;
  out = &it->second;
  return PPM_OK;
}

constexpr int kEvCap = 512;  // events per class group in LDS

// Class boundaries of k_integral_batched's sweep groups on a ring of npix
// pixels (its rule for rings of at most kGroupRuns runs: whole classes,
// starting from class 0, while their slots fit kGroupPositions and their
// points kEvCap / (4 kGroupRuns)): bounds = 0, ..., the ring's class count.
// A ring split into items by class is cut only there, so that its groups,
// and with them the order of every sum, are the ones of the whole ring.
static void class_groups(const ClassTables& T, int npix, std::vector<int>& bounds) {
  const ClassSet cs = T.sets[npix / 4];
  const int n_pad = (npix + 63) & ~63;
  const int slot_cap = std::max(1, std::min(kMaxSlots, kGroupPositions / n_pad));
  const int grp_pts = kEvCap / (4 * kGroupRuns);
  bounds.assign(1, 0);
  for (int ci = 0; ci < cs.n_classes;) {
    int cj = ci + 1, pts = T.hdr[cs.first_class + ci].n_pts;
    while (cj < cs.n_classes && cj - ci < slot_cap && pts + T.hdr[cs.first_class + cj].n_pts <= grp_pts)
      pts += T.hdr[cs.first_class + cj++].n_pts;
    bounds.push_back(ci = cj);
  }
}

// True when, at epl energies per lane, every row set of a work item (epl/2 x
// 64 consecutive energies of a 64 epl-energy chunk) spans at most the slice
// rows the table evaluation takes (58: WaveEnergies::fast; 57 here, below
// the device's FP32 rounding).  Beyond it a wavefront evaluates per energy.
static bool sets_fit(const ppm_ctx* c, const std::vector<BatchGroup>& groups, int epl) {
  const double h = (c->N.logE_max - c->N.logE_min) / std::max(c->N.nE - 1, 1);
  const int set = 32 * epl, chunk = 64 * epl;
  for (const BatchGroup& G : groups) {
    const int n = (int)G.energies.size();
    for (int a = 0; a < n; a += set) {
      const int b = std::min(a + set, std::min(n, (a / chunk + 1) * chunk)) - 1;
      if (std::log10(G.energies[b] / G.energies[a]) / h > 57.0) return false;
    }
  }
  return true;
}

// Energies per lane of a launch: the fewest of 2, 4 or 6 whose work item
// holds every grid's energies in one chunk, if its row sets fit (NSX); else
// the most whose row sets fit, in several chunks per spot; 2 at least.  A
// 256-energy grid over 0.2-12 keV takes 4 (45 rows per set; at 6, 67 rows
// per set would fall back to per-energy evaluation, 2.5x slower).
static int launch_epl(const ppm_ctx* c, const std::vector<BatchGroup>& groups, int atm) {
  size_t ne = 0;
  for (const BatchGroup& G : groups) ne = std::max(ne, G.energies.size());
  for (int epl : {2, 4, 6})
    if (ne <= (size_t)(64 * epl)) {
      if (atm != PPM_ATM_NSX || sets_fit(c, groups, epl)) return epl;
      break;
    }
  if (atm != PPM_ATM_NSX) return 6;
  for (int epl : {6, 4})
    if (sets_fit(c, groups, epl)) return epl;
  return 2;
}

// LDS layout of a launch over `groups`: the largest fine table, slice rows
// and coverage bitmap among them.  Returns the LDS bytes per workgroup with
// the fine tables staged in LDS (fine_lds) or read from global memory.
static size_t launch_geometry(const ppm_ctx* c, const std::vector<BatchGroup>& groups, int atm, int rows_override,
                              int epl, bool fine_lds, int& nsx_rows, int& nf_max, int& cov_words) {
  const int per_item = batch_energies(epl);
  nsx_rows = nf_max = cov_words = 0;
  const double h = (c->N.logE_max - c->N.logE_min) / std::max(c->N.nE - 1, 1);
  for (const BatchGroup& G : groups) {
    nf_max = std::max(nf_max, G.g.n_fine);
    cov_words = std::max(cov_words, std::max((4 * G.g.n_side + 63) / 64 * 2, kGroupPositions / 32) + 2);
    if (atm == PPM_ATM_NSX && !rows_override) {
      double span = 0;
      const int ne = (int)G.energies.size();
      for (int e0 = 0; e0 < ne; e0 += per_item) {
        const int e1 = std::min(ne, e0 + per_item) - 1;
        span = std::max(span, std::log10(G.energies[e1] / G.energies[e0]) / h);
      }
      nsx_rows = std::max(nsx_rows, std::min(c->N.nE, (int)std::ceil(span) + 16 + 8));
    }
  }
  if (atm == PPM_ATM_NSX) {
    if (rows_override) nsx_rows = rows_override;
    nsx_rows = std::max(nsx_rows, 64);  // a row set's 64 lanes read inside the allocation
  }
  return batch_lds_bytes(nf_max, atm, c->N.nm, nsx_rows, cov_words, kEvCap, fine_lds);
}

int batch_grid_epl(const ppm_ctx* c, const BatchGroup& G) {
  return launch_epl(c, {G}, G.stars[0].atmosphere);
}

bool batch_groups_fit(const ppm_ctx* c, const std::vector<BatchGroup>& groups) {
  if (groups.empty() || !groups[0].n) return false;
  int r, f, w;
  const int atm = groups[0].stars[0].atmosphere;
  return launch_geometry(c, groups, atm, 0, launch_epl(c, groups, atm), false, r, f, w) <=
         (size_t)c->max_lds;
}

static ppm_status eval_groups_rows(ppm_ctx* c, Lane& L, std::vector<BatchGroup>& groups, ppm_stats* stats,
                                   int rows_override) {
  BatchState* BS = state(c);
  LaneBatch* LB = lane_batch(L);
  BatchWorkspace& W = L.batch;
  const int n_groups = (int)groups.size();
  int n = 0;  // stars of the launch
  for (const BatchGroup& G : groups) n += G.n;
  if (!n_groups || n < 1) return fail(c, PPM_ERR_ARG, "no stars");
  const int atm = groups[0].stars[0].atmosphere;
  const int emode = batch_emode(groups[0].stars[0]);
  for (const BatchGroup& G : groups) {
    for (size_t e = 1; e < G.energies.size(); ++e)
      if (!(G.energies[e] > G.energies[e - 1]))
        return fail(c, PPM_ERR_ARG, "the batched path needs strictly increasing energies");
    for (int i = 0; i < G.n; ++i)
      if (G.stars[i].atmosphere != atm || batch_emode(G.stars[i]) != emode)
        return fail(c, PPM_ERR_ARG, "a launch needs one atmosphere and one energy-axis interpolant");
  }
  int e;
  // the launch shape: forced (autotuner), stored for a one-grid launch, or
  // the defaults
  ppm_launch_shape shape{0, 0, 0, 0};
  if (L.forced) {
    shape = *L.forced;
  } else if (n_groups == 1) {
    std::lock_guard<std::mutex> lk(BS->shape_mu);
    auto it = BS->shapes.find(shape_key(groups[0].g, (int)groups[0].energies.size(), atm, n));
    if (it != BS->shapes.end()) shape = it->second;
  }
  const int epl = shape.energies_per_lane == 2 || shape.energies_per_lane == 4 || shape.energies_per_lane == 6
                      ? shape.energies_per_lane
                      : launch_epl(c, groups, atm);
  const int per_item = batch_energies(epl);

  // ---- class tables (per grid), held in use until this launch has ended
  std::vector<ClassCache*> cls;
  struct Release {
    BatchState* BS;
    std::vector<ClassCache*>* cls;
    ~Release() {
      std::lock_guard<std::mutex> lk(BS->cls_mu);
      for (ClassCache* cc : *cls) --cc->in_use;
    }
  } release_cls{BS, &cls};
  {
    std::lock_guard<std::mutex> lk(BS->cls_mu);
    for (int k = 0; k < n_groups; ++k) {
      ClassCache* cc = nullptr;
      ppm_status s = class_tables(c, BS, groups[k].g, groups[k].pts, cc);
      if (s) return s;
      cls.push_back(cc);
    }
  }

  // ---- spots, ring tasks and cost estimates; the tasks of a grid are
  // contiguous, stars are numbered across the grids in order
  std::vector<SpotParams> sp;
  std::vector<RingTask> tasks;
  std::vector<double> cost;
  std::vector<int> n_cls;  // phase classes of each task's ring
  std::vector<int> spot_first_task, spot_star, spot_group;
  std::vector<int> group_task0(n_groups + 1), group_star0(n_groups + 1);
  std::vector<int> rings;
  for (int k = 0, sg = 0; k < n_groups; ++k) {
    const BatchGroup& G = groups[k];
    const ppm_grid& g = G.g;
    group_task0[k] = (int)tasks.size();
    group_star0[k] = sg;
    for (int s = 0; s < G.n; ++s, ++sg)
      for (int q = 0; q < G.stars[s].n_spots; ++q) {
        SpotParams P;
        spot_params(c, G.stars[s], G.stars[s].spot[q], g.n_side, P);
        const int si = (int)sp.size();
        sp.push_back(P);
        spot_first_task.push_back((int)tasks.size());
        spot_star.push_back(sg);
        spot_group.push_back(k);
        candidate_rings(P, G.stars[s].spot[q], g.n_side, rings);
        const ppm_spot& spt = G.stars[s].spot[q];
        const double st1 = std::sin(spt.theta1), ct1 = std::cos(spt.theta1), crho = std::cos(spt.angrad1);
        for (int r : rings) {
          tasks.push_back(RingTask{si, r});
          double est;
          if (r == 0) {
            est = (double)G.pts.size();
            n_cls.push_back(1);
          } else {
            const int npix = 4 * std::min({r, 4 * g.n_side - r, g.n_side});
            const ClassSet cs = cls[k]->tab.sets[npix / 4];
            const double a = arc_fraction(ring_z(g.n_side, r), st1, ct1, crho) * npix;
            const double per_cls = (double)G.pts.size() / std::max(cs.n_classes, 1);
            est = cs.n_classes * (std::min((double)npix, a * per_cls) + 64.0);
            n_cls.push_back(std::max(cs.n_classes, 1));
          }
          // an energy chunk of 384 energies costs about as much as the full
          // chunk; fewer energies cost proportionally less
          cost.push_back(est * std::min(1.0, (double)G.energies.size() / per_item));
        }
      }
  }
  group_task0[n_groups] = (int)tasks.size();
  group_star0[n_groups] = n;
  spot_first_task.push_back((int)tasks.size());
  const int n_spots = (int)sp.size();
  const int n_tasks = (int)tasks.size();

  // ---- launch geometry: the LDS layout of the largest grid
  // the fine tables go to LDS unless that lowers the occupancy
  int nsx_rows, nf_max, cov_words;
  const size_t lds_global = launch_geometry(c, groups, atm, rows_override, epl, false, nsx_rows, nf_max, cov_words);
  const size_t lds_staged = launch_geometry(c, groups, atm, rows_override, epl, true, nsx_rows, nf_max, cov_words);
  if (lds_global > (size_t)c->max_lds) return fail(c, PPM_ERR_ARG, "grid too large for the batched kernel's LDS");
  int occ_global = 0, occ_staged = 0;
  if ((e = batched_prepare(atm, epl, emode, false, c->max_lds))) return hip_fail(c, e, "kernel attributes");
  if (batched_occupancy(atm, epl, emode, false, lds_global, &occ_global) || occ_global < 1) occ_global = 1;
  if (lds_staged <= (size_t)c->max_lds) {
    if ((e = batched_prepare(atm, epl, emode, true, c->max_lds))) return hip_fail(c, e, "kernel attributes");
    if (batched_occupancy(atm, epl, emode, true, lds_staged, &occ_staged)) occ_staged = 0;
  }
  const bool fine_lds = occ_staged >= occ_global;
  const size_t lds = fine_lds ? lds_staged : lds_global;
  int wg_per_cu = fine_lds ? occ_staged : occ_global;
  L.occ_wg_max = wg_per_cu;
  if (shape.wg_per_cu > 0) wg_per_cu = std::min(wg_per_cu, shape.wg_per_cu);
  const int slots = wg_per_cu * c->n_cu;

  // ---- items: each spot's rings split into chunks of similar cost, times
  // the energy chunks of its grid; about per_slot items per resident
  // workgroup, fewer if their accumulators would pass the lane's cap
  // an item's accumulator copies cover its own energy chunk only
  std::vector<int> n_echunk(n_groups);
  for (int k = 0; k < n_groups; ++k) {
    const int ne = (int)groups[k].energies.size();
    n_echunk[k] = (ne + per_item - 1) / per_item;
  }
  double total_cost = 0;
  for (double v : cost) total_cost += v;
  const double per_slot = shape.items_per_slot > 0 ? shape.items_per_slot : kDefaultItemsPerSlot;
  std::vector<BatchItem> items;
  std::vector<double> item_cost;
  std::vector<int> star_item0(n + 1, 0);
  size_t acc_total = 0;
  std::vector<double> cuts;
  std::vector<int> bounds;
  // one item per energy chunk of tasks [t0, t1), classes [c0, c1) of each ring
  auto add_item = [&](int t0, int t1, int s, int si, int c0, int c1, double ic) {
    const int k_grid = spot_group[si];
    const int ne = (int)groups[k_grid].energies.size();
    for (int ec = 0; ec < n_echunk[k_grid]; ++ec) {
      const int e0 = ec * per_item, en = std::min(per_item, ne - e0);
      items.push_back(BatchItem{t0, t1, s, si, e0, k_grid, en, c0, c1, acc_total});
      acc_total += (size_t)groups[k_grid].g.n_phase * en * batch_acc_copies();
      item_cost.push_back(ic);
    }
  };
  auto build_items = [&](double target) {
    items.clear();
    item_cost.clear();
    acc_total = 0;
    for (int s = 0, si = 0; s < n; ++s) {
      star_item0[s] = (int)items.size();
      for (; si < n_spots && spot_star[si] == s; ++si) {
        const int t0 = spot_first_task[si], t1 = spot_first_task[si + 1];
        if (t0 == t1) continue;
        const int k_grid = spot_group[si];
        double cs = 0;
        for (int t = t0; t < t1; ++t) cs += cost[t];
        int k = (int)std::lround(total_cost > 0 ? target * cs / total_cost / n_echunk[k_grid] : 1.0);
        if (k > t1 - t0) {
          // more pieces wanted than the spot has rings (few stars): one item
          // per ring, and a ring that costs more than a third of a resident
          // workgroup's share of the launch is split by phase class (class
          // groups are independent sweeps; each item has its own
          // accumulator).  One N_side 1024 star: 8.7 -> 6.2 ms (a share: 6.5)
          const double unit = std::max(cs / k, total_cost / std::max(3 * slots, 1));
          for (int t = t0; t < t1; ++t) {
            int m = 1;
            if (n_cls[t] > 1 && cost[t] > 1.5 * unit)
              m = (int)std::min<double>(n_cls[t], std::max(1.0, std::round(cost[t] / unit)));
            if (m == 1) {
              add_item(t, t + 1, s, si, 0, kAllClasses, cost[t]);
              continue;
            }
            // m pieces cut at the sweep-group boundaries nearest to j n / m
            const int r = tasks[t].ring, ns = groups[k_grid].g.n_side, npix = 4 * std::min({r, 4 * ns - r, ns});
            class_groups(cls[k_grid]->tab, npix, bounds);
            int c0 = 0;
            for (int j = 1; j <= m; ++j) {
              const double target = (double)j * n_cls[t] / m;
              int c1 = n_cls[t];
              if (j < m) {
                const auto it = std::lower_bound(bounds.begin(), bounds.end(), (int)std::lround(target));
                c1 = it == bounds.end() ? n_cls[t] : *it;
                if (it != bounds.begin() && std::fabs(*(it - 1) - target) < std::fabs(c1 - target)) c1 = *(it - 1);
              }
              if (c1 <= c0) continue;
              add_item(t, t + 1, s, si, c0, c1, cost[t] * (c1 - c0) / n_cls[t]);
              c0 = c1;
            }
          }
          continue;
        }
        k = std::max(1, k);
        // cumulative cost at the end of each piece
        const int kt = k > kTailChunks ? kTailChunks : 0;
        cuts.clear();
        for (int j = 1; j <= k - kt; ++j) cuts.push_back(cs * j / k);
        for (int j = 1; j <= kt * kTailSplit; ++j) cuts.push_back(cs * (k - kt) / k + cs * j / ((double)k * kTailSplit));
        double acc = 0;
        int start = t0, made = 0;
        for (int t = t0; t < t1; ++t) {
          acc += cost[t];
          if (t == t1 - 1 || (made < (int)cuts.size() && acc >= cuts[made])) {
            double ic = 0;
            for (int u = start; u <= t; ++u) ic += cost[u];
            add_item(start, t + 1, s, si, 0, kAllClasses, ic);
            start = t + 1;
            ++made;
          }
        }
      }
    }
  };
  const double target = per_slot * slots;
  build_items(target);
  // fewer items until their accumulators fit the lane's cap (the rounding
  // per spot and the finer tail make the count only roughly proportional)
  for (double t = target; acc_total * sizeof(double) > L.acc_cap && t > 1.0;) {
    t = std::max(1.0, t * std::min(0.9, 0.95 * (double)L.acc_cap / (double)(acc_total * sizeof(double))));
    build_items(t);
  }
  if (acc_total * sizeof(double) > L.acc_cap)  // one item per spot and energy chunk is the least
    return fail(c, PPM_ERR_NOMEM, "the accumulators of this launch (" + std::to_string(acc_total * sizeof(double)) +
                                      " bytes at one item per spot) exceed the device-memory budget");
  star_item0[n] = (int)items.size();
  const int n_items = (int)items.size();
  std::vector<int> order(n_items);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return item_cost[a] > item_cost[b]; });

  // ---- per-grid offsets of energies and fine tables
  std::vector<size_t> e_off(n_groups), ph_off(n_groups), fine_off(n_groups);
  size_t n_energy_all = 0, n_ph_all = 0, n_fine_all = 0;
  for (int k = 0; k < n_groups; ++k) {
    const size_t nt = group_task0[k + 1] - group_task0[k], nf1 = groups[k].g.n_fine + 1;
    e_off[k] = n_energy_all;
    n_energy_all += groups[k].energies.size();
    ph_off[k] = n_ph_all;
    n_ph_all += nt * nf1;
    fine_off[k] = n_fine_all;
    n_fine_all += 3 * nt * nf1;
  }

  // ---- buffers and uploads
  if ((e = W.spots.ensure(sizeof(SpotParams) * std::max(n_spots, 1))) ||
      (e = W.tasks.ensure(sizeof(RingTask) * std::max(n_tasks, 1))) ||
      (e = W.desc.ensure(sizeof(RingDesc) * std::max(n_tasks, 1))) ||
      (e = W.ph.ensure(sizeof(double) * std::max<size_t>(n_ph_all, 1))) ||
      (e = W.fine.ensure(sizeof(float) * std::max<size_t>(n_fine_all, 1))) ||
      (e = W.energies.ensure(sizeof(double) * n_energy_all)) ||
      (e = W.items.ensure(sizeof(BatchItem) * std::max(n_items, 1))) ||
      (e = LB->d_grids.ensure(sizeof(BatchGrid) * n_groups)) ||
      (e = LB->d_order.ensure(sizeof(int) * std::max(n_items, 1))) ||
      (e = LB->d_star_item0.ensure(sizeof(int) * (n + 1))) || (e = W.counter.ensure(sizeof(int))) ||
      (e = W.evals.ensure(sizeof(unsigned long long))) ||
      (e = W.summary.ensure(sizeof(unsigned long long) * (4 * n_groups + 1))) ||
      (e = LB->d_acc.ensure(sizeof(double) * std::max<size_t>(acc_total, 1))) ||
      (e = W.scale.ensure(sizeof(double) * n)))
    return hip_fail(c, e, "device allocation");
  std::vector<BatchGrid> bg(n_groups);
  std::vector<double> energies_all(n_energy_all), fold_scale(n);
  for (int k = 0; k < n_groups; ++k) {
    const BatchGroup& G = groups[k];
    std::copy(G.energies.begin(), G.energies.end(), energies_all.begin() + e_off[k]);
    bg[k] = BatchGrid{(const double*)W.energies.p + e_off[k],
                      (const ClassSet*)cls[k]->csets.p,
                      (const ClassHdr*)cls[k]->chdr.p,
                      (const ClassPt*)cls[k]->cpts.p,
                      (const double*)W.ph.p + ph_off[k],
                      (const float*)W.fine.p + fine_off[k],
                      (int)G.energies.size(),
                      G.g.n_phase,
                      G.g.n_fine,
                      group_task0[k]};
    for (int s = 0; s < G.n; ++s)
      fold_scale[group_star0[k] + s] = G.stars[s].exposure_s * (1.0 / G.g.n_phase);
  }
  {
    Upload up;
    up.add(W.spots.p, sp.data(), sizeof(SpotParams) * n_spots);
    if (n_tasks) up.add(W.tasks.p, tasks.data(), sizeof(RingTask) * n_tasks);
    up.add(W.energies.p, energies_all.data(), sizeof(double) * n_energy_all);
    up.add(LB->d_grids.p, bg.data(), sizeof(BatchGrid) * n_groups);
    if (n_items) {
      up.add(W.items.p, items.data(), sizeof(BatchItem) * n_items);
      up.add(LB->d_order.p, order.data(), sizeof(int) * n_items);
    }
    up.add(LB->d_star_item0.p, star_item0.data(), sizeof(int) * (n + 1));
    up.add(W.scale.p, fold_scale.data(), sizeof(double) * n);
    if ((e = up.run(L))) return hip_fail(c, e, "upload");
  }

  // ring set-up per grid (its fine-table stride)
  std::vector<PlainLaunch> P(n_groups);
  for (int k = 0; k < n_groups; ++k) {
    PlainLaunch& Q = P[k];
    Q = PlainLaunch{};
    Q.d_spots = (const SpotParams*)W.spots.p;
    Q.d_tasks = (const RingTask*)W.tasks.p + group_task0[k];
    Q.d_desc = (RingDesc*)W.desc.p + group_task0[k];
    Q.d_ph = (double*)W.ph.p + ph_off[k];
    Q.d_fine = (float*)W.fine.p + fine_off[k];
    Q.n_tasks = group_task0[k + 1] - group_task0[k];
    Q.n_fine = groups[k].g.n_fine;
    Q.summary = (unsigned long long*)W.summary.p + 4 * k;
  }
  BatchLaunch B{};
  B.spots = (const SpotParams*)W.spots.p;
  B.tasks = (const RingTask*)W.tasks.p;
  B.desc = (RingDesc*)W.desc.p;
  B.grids = (const BatchGrid*)LB->d_grids.p;
  B.items = (const BatchItem*)W.items.p;
  B.order = (const int*)LB->d_order.p;
  B.star_item0 = (const int*)LB->d_star_item0.p;
  B.queue = (int*)W.counter.p;
  B.n_evals = (unsigned long long*)W.evals.p;
  B.rows_flag = (unsigned long long*)W.summary.p + 4 * n_groups;
  B.acc = (double*)LB->d_acc.p;
  B.n_items = n_items;
  B.n_stars = n;
  B.n_fine_max = nf_max;
  B.nsx_rows = nsx_rows;
  B.cov_words = cov_words;
  B.ev_cap = kEvCap;
  B.epl = epl;
  B.atmosphere = atm;
  B.emode = emode;
  B.fine_lds = fine_lds;
  B.persistent = !L.per_item;
  const int n_wg = L.per_item ? std::max(1, n_items) : std::max(1, std::min(slots, n_items));

  // ---- the launch sequence, captured once per configuration and replayed
  std::vector<unsigned char> key;
  auto put = [&](const void* p, size_t nb) {
    const unsigned char* b = (const unsigned char*)p;
    key.insert(key.end(), b, b + nb);
  };
  put(&B, sizeof B);
  put(&n_wg, sizeof n_wg);
  put(&lds, sizeof lds);
  {
    const FoldFit none{};
    put(L.fit ? L.fit : &none, sizeof(FoldFit));
  }
  for (int k = 0; k < n_groups; ++k) {
    put(&P[k], sizeof(PlainLaunch));
    put(&bg[k], sizeof(BatchGrid));
    put(&groups[k].d_flux, sizeof(void*));
    put(&groups[k].d_counts, sizeof(void*));
    put(&groups[k].g.response, sizeof(int));
    put(&group_star0[k], sizeof(int));
    put(&groups[k].fit_star0, sizeof(int));
  }
  // timed: per-kernel events around a direct launch; graph replays are timed
  // as a whole by events recorded around hipGraphLaunch (outside the graph)
  auto record = [&](bool timed) -> int {
    int r;
    if (timed && (r = hipEventRecord(L.ev0, L.stream))) return r;
    if ((r = launch_zero_u64((unsigned long long*)W.summary.p, 4 * n_groups + 1, L.stream))) return r;
    for (int k = 0; k < n_groups; ++k)
      if (P[k].n_tasks && (r = launch_ring_setup(c->L, c->N, P[k], L.stream))) return r;
    if (timed && (r = hipEventRecord(L.ev_i0, L.stream))) return r;
    if (n_items && (r = launch_batched(c->N, B, n_wg, L.stream))) return r;
    if (timed && (r = hipEventRecord(L.ev_i1, L.stream))) return r;
    for (int k = 0; k < n_groups; ++k)
      if ((r = launch_reduce_batched(B, bg[k], group_star0[k], group_star0[k + 1], groups[k].d_flux, L.stream)))
        return r;
    if (timed && (r = hipEventRecord(L.ev1, L.stream))) return r;
    for (int k = 0; k < n_groups; ++k) {
      const BatchGroup& G = groups[k];
      if (!G.d_counts || G.g.response < 0) continue;
      const Response& R = c->resp[G.g.response];
      if ((r = launch_fold(R.d_hi, R.d_lo, R.n_ch, R.n_e, G.d_flux, G.n, G.g.n_phase,
                           (const double*)W.scale.p + group_star0[k], G.d_counts, L.stream, L.fit,
                           G.fit_star0 >= 0 ? G.fit_star0 : group_star0[k])))
        return r;
    }
    if (timed && (r = hipEventRecord(L.ev2, L.stream))) return r;
    return hipSuccess;
  };
  GraphCache& GC = LB->graph;
  const bool use_graph = !L.no_graph && ((GC.exec && GC.key == key) || GC.last_direct == key);
  if (use_graph) {
    if (!GC.exec || GC.key != key) {
      if (GC.exec) {
        hipGraphExecDestroy(GC.exec);
        GC.exec = nullptr;
      }
      hipGraph_t graph = nullptr;
      if ((e = hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal))) return hip_fail(c, e, "capture");
      const int r = record(false);
      e = hipStreamEndCapture(L.stream, &graph);
      if (r) {
        if (graph) hipGraphDestroy(graph);
        return hip_fail(c, r, "batched launch");
      }
      if (e) return hip_fail(c, e, "capture");
      e = hipGraphInstantiate(&GC.exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (e) {
        GC.exec = nullptr;
        return hip_fail(c, e, "graph instantiation");
      }
      GC.key = key;
    }
    if ((e = hipEventRecord(L.ev0, L.stream))) return hip_fail(c, e, "event");
    if ((e = hipGraphLaunch(GC.exec, L.stream))) return hip_fail(c, e, "graph launch");
    if ((e = hipEventRecord(L.ev2, L.stream))) return hip_fail(c, e, "event");
  } else {
    if ((e = record(true))) return hip_fail(c, e, "batched launch");
    GC.last_direct = key;
  }
  if ((e = hipStreamSynchronize(L.stream))) return hip_fail(c, e, "execution");

  // the ring set-up's counters per grid and the integral's row-overflow flag
  std::vector<unsigned long long> sum(4 * n_groups + 1);
  if ((e = hipMemcpy(sum.data(), W.summary.p, sizeof(unsigned long long) * sum.size(), hipMemcpyDeviceToHost)))
    return hip_fail(c, e, "summary read-back");
  if ((sum[4 * n_groups] & 2) && nsx_rows < c->N.nE) return eval_groups_rows(c, L, groups, stats, c->N.nE);
  int64_t n_act = 0, n_eval_ref = 0, n_rings = 0;
  for (int k = 0; k < n_groups; ++k) {
    if (sum[4 * k + 3] & 1) return fail(c, PPM_ERR_UNSUPPORTED, "a ring has more active runs than supported");
    n_act += (int64_t)sum[4 * k];
    n_rings += (int64_t)sum[4 * k + 2];
    n_eval_ref += (int64_t)sum[4 * k + 1] * (int64_t)groups[k].pts.size() * (int64_t)groups[k].energies.size();
  }
  if (stats) {
    float t02 = 0;
    hipEventElapsedTime(&t02, L.ev0, L.ev2);
    stats->total_ms += t02;
    if (!use_graph) {  // per-kernel times exist only for direct launches
      float t01 = 0, t12 = 0, ti = 0;
      hipEventElapsedTime(&t01, L.ev0, L.ev1);
      hipEventElapsedTime(&t12, L.ev1, L.ev2);
      hipEventElapsedTime(&ti, L.ev_i0, L.ev_i1);
      stats->kernel_ms += t01;
      stats->integral_ms += ti;
      stats->fold_ms += t12;
    }
    stats->n_active += n_act;
    stats->n_eval_ref += n_eval_ref;
    stats->n_rings += n_rings;
    unsigned long long done = 0;
    if (hipMemcpy(&done, W.evals.p, sizeof done, hipMemcpyDeviceToHost) == hipSuccess)
      stats->n_eval_done += (int64_t)done;
  }
  return PPM_OK;
}

// ---- autotuner of the batched launch shape

namespace {

ppm_status shape_args(ppm_ctx* c, const ppm_grid* grid, int n_stars, int atmosphere, std::vector<int>& key) {
  if (!grid || n_stars < 1) return fail(c, PPM_ERR_ARG, "null grid or no stars");
  if (atmosphere != PPM_ATM_BLACKBODY && atmosphere != PPM_ATM_NSX) return fail(c, PPM_ERR_ARG, "unknown atmosphere");
  std::vector<double> energies;
  ppm_status s = resolve_energies(c, *grid, energies);
  if (s) return s;
  key = shape_key(*grid, (int)energies.size(), atmosphere, n_stars);
  return PPM_OK;
}

}  // namespace

}  // namespace ppm

using namespace ppm;

extern "C" {

ppm_status ppm_tune(ppm_ctx* c, const ppm_star* stars, int n, const ppm_grid* grid, int repeats, ppm_tune_result* out) {
  if (!c || !stars || !grid || n < 1 || !out) return fail(c, PPM_ERR_ARG, "null argument");
  if (repeats < 0 || repeats > 16) return fail(c, PPM_ERR_ARG, "repeats must be in 0..16");
  if (repeats == 0) repeats = 2;
  std::lock_guard<std::mutex> lk(c->exec_mu);
  hipSetDevice(c->device);
  std::vector<int> key;
  ppm_status s = shape_args(c, grid, n, stars[0].atmosphere, key);
  if (s) return s;
  // the grid's energies, given explicitly: the fold (which does not depend
  // on the shape) is not run
  std::vector<double> energies;
  if ((s = resolve_energies(c, *grid, energies))) return s;
  ppm_grid g = *grid;
  g.method = PPM_METHOD_BATCHED;
  g.energies_keV = energies.data();
  g.n_energy = (int)energies.size();
  g.response = -1;
  const size_t nf = (size_t)n * g.n_phase * key[5];
  DevBuf d_flux;
  int e = d_flux.ensure(sizeof(double) * nf);
  if (e) {
    d_flux.release();
    return hip_fail(c, e, "device allocation");
  }
  BatchState* BS = state(c);
  Lane& L = c->main;
  ppm_launch_shape cur{0, 0, 0, 0};
  // direct launches (the integral's own events) and the candidate shape,
  // restored on every exit
  struct Restore {
    Lane& L;
    DevBuf* buf;
    bool no_graph;
    ~Restore() {
      L.no_graph = no_graph;
      L.forced = nullptr;
      hipStreamSynchronize(L.stream);
      buf->release();
    }
  } restore{L, &d_flux, L.no_graph};
  L.no_graph = true;
  L.forced = &cur;
  int n_tried = 0;
  auto time_shape = [&](ppm_launch_shape sh, double& ms) -> ppm_status {
    cur = sh;
    ms = 1e300;
    for (int r = 0; r < repeats; ++r) {
      ppm_stats st{};
      const ppm_status q = eval_any(c, L, stars, n, &g, (double*)d_flux.p, nullptr, &st);
      if (q) return q;
      ms = std::min(ms, st.integral_ms);
    }
    ++n_tried;
    return PPM_OK;
  };
  // untimed warm-up (allocations, first launches); it also gives the
  // occupancy limit
  {
    cur = ppm_launch_shape{0, 0, 0, 0};
    ppm_stats st{};
    if ((s = eval_any(c, L, stars, n, &g, (double*)d_flux.p, nullptr, &st))) return s;
  }
  const int wg_max = std::max(1, L.occ_wg_max);
  ppm_launch_shape best{kDefaultItemsPerSlot, wg_max, 0, 0};
  double t_best, t_def;
  if ((s = time_shape(best, t_def))) return s;
  t_best = t_def;
  auto consider = [&](ppm_launch_shape sh) -> ppm_status {
    double t;
    const ppm_status q = time_shape(sh, t);
    if (q) return q;
    if (t < t_best) {
      t_best = t;
      best = sh;
    }
    return PPM_OK;
  };
  // items per workgroup: halve and double, then one more step the better way
  for (int f : {kDefaultItemsPerSlot / 2, kDefaultItemsPerSlot * 2})
    if ((s = consider(ppm_launch_shape{f, wg_max, 0, 0}))) return s;
  if (best.items_per_slot != kDefaultItemsPerSlot) {
    const int f = best.items_per_slot < kDefaultItemsPerSlot ? best.items_per_slot / 2 : best.items_per_slot * 2;
    if ((s = consider(ppm_launch_shape{std::max(f, 1), wg_max, 0, 0}))) return s;
  }
  // one resident workgroup fewer per CU
  if (wg_max > 1 && (s = consider(ppm_launch_shape{best.items_per_slot, wg_max - 1, 0, 0}))) return s;
  // more than 128 energies: the other energies-per-lane choices whose row
  // sets fit (fewer per lane: more energy chunks, each patch set up again;
  // more: idle lanes in the last chunk)
  if (key[5] > 128) {
    BatchGroup G;
    G.stars = stars;
    G.n = n;
    G.g = g;
    G.energies = energies;
    const int dflt = launch_epl(c, {G}, stars[0].atmosphere);
    const ppm_launch_shape base = best;
    for (int epl : {2, 4, 6})
      if (epl != dflt && (stars[0].atmosphere != PPM_ATM_NSX || sets_fit(c, {G}, epl)) &&
          (s = consider(ppm_launch_shape{base.items_per_slot, base.wg_per_cu, epl, 0})))
        return s;
  }
  {
    std::lock_guard<std::mutex> lk(BS->shape_mu);
    BS->shapes[key] = best;
  }
  out->best = best;
  out->best_ms = t_best;
  out->default_ms = t_def;
  out->n_tried = n_tried;
  out->reserved = 0;
  return PPM_OK;
}

ppm_status ppm_set_launch_shape(ppm_ctx* c, const ppm_grid* grid, int n_stars, int atmosphere,
                                const ppm_launch_shape* shape) {
  if (!c) return PPM_ERR_ARG;
  if (shape && (shape->items_per_slot < 0 || shape->items_per_slot > 4096 || shape->wg_per_cu < 0 ||
                shape->wg_per_cu > 64 ||
                (shape->energies_per_lane != 0 && shape->energies_per_lane != 2 && shape->energies_per_lane != 4 &&
                 shape->energies_per_lane != 6)))
    return fail(c, PPM_ERR_ARG,
                "launch shape out of range (items_per_slot 0..4096, wg_per_cu 0..64, energies_per_lane 0, 2, 4 or 6)");
  std::lock_guard<std::mutex> lk(c->exec_mu);
  std::vector<int> key;
  ppm_status s = shape_args(c, grid, n_stars, atmosphere, key);
  if (s) return s;
  BatchState* BS = state(c);
  std::lock_guard<std::mutex> lk2(BS->shape_mu);
  if (shape && (shape->items_per_slot || shape->wg_per_cu || shape->energies_per_lane)) {
    ppm_launch_shape sh = *shape;
    sh.reserved = 0;
    BS->shapes[key] = sh;
  } else {
    BS->shapes.erase(key);
  }
  return PPM_OK;
}

ppm_status ppm_get_launch_shape(ppm_ctx* c, const ppm_grid* grid, int n_stars, int atmosphere,
                                ppm_launch_shape* shape) {
  if (!c || !shape) return PPM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->exec_mu);
  std::vector<int> key;
  ppm_status s = shape_args(c, grid, n_stars, atmosphere, key);
  if (s) return s;
  BatchState* BS = state(c);
  std::lock_guard<std::mutex> lk2(BS->shape_mu);
  auto it = BS->shapes.find(key);
  *shape = it != BS->shapes.end() ? it->second : ppm_launch_shape{0, 0, 0, 0};
  return PPM_OK;
}

ppm_status ppm_phase_timers(ppm_ctx* c, uint64_t* cycles, int n, int reset) {
  if (!c || !cycles || n < 1 || n > PPM_PHASE_COUNT) return PPM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->exec_mu);
  hipSetDevice(c->device);
  if (hipDeviceSynchronize() != hipSuccess) return fail(c, PPM_ERR_DEVICE, "device synchronisation failed");
  static_assert(sizeof(uint64_t) == sizeof(unsigned long long), "cycle counters are 64-bit");
  if (!batch_phase_timers(reinterpret_cast<unsigned long long*>(cycles), n, reset != 0))
    return fail(c, PPM_ERR_UNSUPPORTED, "this library is built without phase timers (libppm_phase.so has them)");
  return PPM_OK;
}

}  // extern "C"
