// Batched evaluation: many stars in one persistent launch, captured as a
// hipGraph (ppm_batch_kernels.hip holds the kernels).
#include "ppm_batch.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
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

struct ClassCache {
  int n_side = -1, n_phase = -1, sampling = -1, n_gl = -1;
  ClassTables tab;
  bool uploaded = false;
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
constexpr int kDefaultItemsPerSlot = 12;

struct BatchState {
  ClassCache classes;
  GraphCache graph;
  DevBuf d_csets, d_chdr, d_cpts, d_order, d_star_item0, d_acc, d_seg;
  int n_cu = 0;
  // shapes found by ppm_tune or set by the caller, per shape_key
  std::map<std::vector<int>, ppm_launch_shape> shapes;
  const ppm_launch_shape* forced = nullptr;  // set while ppm_tune times a candidate
  int occ_wg_max = 0;                        // occupancy limit of the last launch
};

// The grid fields that shape the batched work, the atmosphere and the bit
// length of the star count.
static std::vector<int> shape_key(const ppm_grid& g, int n_e, int atm, int n) {
  int bits = 0;
  while (bits < 31 && (1 << bits) <= n) ++bits;
  return {g.n_side, g.n_phase, g.sampling, g.sampling == PPM_SAMPLE_GL ? g.n_gl : 0, g.n_fine, n_e, atm, bits};
}

static BatchState* state(ppm_ctx* c) {
  if (!c->batch_state) c->batch_state = new BatchState();
  return c->batch_state;
}

// Fraction of a ring of colatitude th inside a disk (th1, rho), for cost
// estimates only.
static double arc_fraction(double th, double th1, double rho) {
  const double s = std::sin(th) * std::sin(th1);
  const double cth = (std::cos(rho) - std::cos(th) * std::cos(th1));
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
  if (S->graph.exec) hipGraphExecDestroy(S->graph.exec);
  for (DevBuf* b : {&S->d_csets, &S->d_chdr, &S->d_cpts, &S->d_order, &S->d_star_item0, &S->d_acc, &S->d_seg})
    b->release();
  delete S;
  c->batch_state = nullptr;
}

ppm_status eval_batched_rows(ppm_ctx* c, const ppm_star* stars, int n, const ppm_grid& g,
                             const std::vector<double>& energies, const std::vector<PhasePoint>& pts,
                             double* d_flux, double* d_counts, ppm_stats* stats, int rows_override);

ppm_status eval_batched(ppm_ctx* c, const ppm_star* stars, int n, const ppm_grid& g,
                        const std::vector<double>& energies, const std::vector<PhasePoint>& pts,
                        double* d_flux, double* d_counts, ppm_stats* stats) {
  return eval_batched_rows(c, stars, n, g, energies, pts, d_flux, d_counts, stats, 0);
}

ppm_status eval_batched_rows(ppm_ctx* c, const ppm_star* stars, int n, const ppm_grid& g,
                             const std::vector<double>& energies, const std::vector<PhasePoint>& pts,
                             double* d_flux, double* d_counts, ppm_stats* stats, int rows_override) {
  BatchState* BS = state(c);
  BatchWorkspace& W = c->batch;
  const int n_e = (int)energies.size();
  const int nf1 = g.n_fine + 1;
  for (int e = 1; e < n_e; ++e)
    if (!(energies[e] > energies[e - 1]))
      return fail(c, PPM_ERR_ARG, "the batched path needs strictly increasing energies");
  int atm = stars[0].atmosphere;
  for (int i = 1; i < n; ++i)
    if (stars[i].atmosphere != atm) return fail(c, PPM_ERR_ARG, "a batch needs one atmosphere model");
  int e;
  if (!BS->n_cu && hipDeviceGetAttribute(&BS->n_cu, hipDeviceAttributeMultiprocessorCount, c->device))
    BS->n_cu = 256;

  // ---- class tables (per grid)
  ClassCache& CC = BS->classes;
  if (CC.n_side != g.n_side || CC.n_phase != g.n_phase || CC.sampling != g.sampling || CC.n_gl != g.n_gl) {
    build_classes(g.n_side, pts, CC.tab);
    CC.n_side = g.n_side;
    CC.n_phase = g.n_phase;
    CC.sampling = g.sampling;
    CC.n_gl = g.n_gl;
    CC.uploaded = false;
  }
  if (!CC.uploaded) {
    const ClassTables& T = CC.tab;
    if ((e = BS->d_csets.ensure(sizeof(ClassSet) * T.sets.size())) ||
        (e = BS->d_chdr.ensure(sizeof(ClassHdr) * T.hdr.size())) ||
        (e = BS->d_cpts.ensure(sizeof(ClassPt) * T.pts.size())) ||
        (e = hipMemcpy(BS->d_csets.p, T.sets.data(), sizeof(ClassSet) * T.sets.size(), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(BS->d_chdr.p, T.hdr.data(), sizeof(ClassHdr) * T.hdr.size(), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(BS->d_cpts.p, T.pts.data(), sizeof(ClassPt) * T.pts.size(), hipMemcpyHostToDevice)))
      return hip_fail(c, e, "class tables");
    CC.uploaded = true;
  }

  // ---- spots, ring tasks and cost estimates
  std::vector<SpotParams> sp;
  std::vector<RingTask> tasks;
  std::vector<double> cost;
  std::vector<int> spot_first_task, spot_star;
  std::vector<int> rings;
  for (int s = 0; s < n; ++s)
    for (int k = 0; k < stars[s].n_spots; ++k) {
      SpotParams P;
      spot_params(c, stars[s], stars[s].spot[k], g.n_side, P);
      P.pad = s;
      const int si = (int)sp.size();
      sp.push_back(P);
      spot_first_task.push_back((int)tasks.size());
      spot_star.push_back(s);
      candidate_rings(P, stars[s].spot[k], g.n_side, rings);
      const ppm_spot& spt = stars[s].spot[k];
      for (int r : rings) {
        tasks.push_back(RingTask{si, r});
        double est;
        if (r == 0) {
          est = (double)pts.size();
        } else {
          const int npix = 4 * std::min({r, 4 * g.n_side - r, g.n_side});
          const ClassSet cs = CC.tab.sets[npix / 4];
          const double a = arc_fraction(std::acos(ring_z(g.n_side, r)), spt.theta1, spt.angrad1) * npix;
          const double per_cls = (double)pts.size() / std::max(cs.n_classes, 1);
          est = cs.n_classes * (std::min((double)npix, a * per_cls) + 64.0);
        }
        cost.push_back(est);
      }
    }
  spot_first_task.push_back((int)tasks.size());
  const int n_spots = (int)sp.size();
  const int n_tasks = (int)tasks.size();

  // ---- launch geometry
  const int n_echunk = (n_e + batch_energies() - 1) / batch_energies();
  int nsx_rows = 0;
  if (atm == PPM_ATM_NSX) {
    if (rows_override) {
      nsx_rows = rows_override;
    } else {
      const double h = (c->N.logE_max - c->N.logE_min) / (c->N.nE - 1);
      double span = 0;
      for (int e0 = 0; e0 < n_e; e0 += batch_energies()) {
        const int e1 = std::min(n_e, e0 + batch_energies()) - 1;
        span = std::max(span, std::log10(energies[e1] / energies[e0]) / h);
      }
      nsx_rows = std::min(c->N.nE, (int)std::ceil(span) + 16 + 8);
    }
    nsx_rows = std::max(nsx_rows, 64);  // a row set's 64 lanes read inside the allocation
  }
  const int cov_words = (4 * g.n_side + 63) / 64 * 2 + 2;
  const int ev_cap = 512;
  const size_t lds = batch_lds_bytes(g.n_fine, atm, nsx_rows, cov_words, ev_cap);
  if (lds > (size_t)c->max_lds) return fail(c, PPM_ERR_ARG, "grid too large for the batched kernel's LDS");
  if ((e = batched_prepare(atm, lds))) return hip_fail(c, e, "kernel attributes");
  int wg_per_cu = 0;
  if (batched_occupancy(atm, lds, &wg_per_cu) || wg_per_cu < 1) wg_per_cu = 1;
  BS->occ_wg_max = wg_per_cu;
  ppm_launch_shape shape{0, 0};
  if (BS->forced) {
    shape = *BS->forced;
  } else {
    auto it = BS->shapes.find(shape_key(g, n_e, atm, n));
    if (it != BS->shapes.end()) shape = it->second;
  }
  if (shape.wg_per_cu > 0) wg_per_cu = std::min(wg_per_cu, shape.wg_per_cu);
  const int slots = wg_per_cu * BS->n_cu;

  // ---- items: each spot's rings split into chunks of similar cost
  double total_cost = 0;
  for (double v : cost) total_cost += v;
  const size_t acc_per_item = (size_t)g.n_phase * n_e * batch_acc_copies();
  const size_t cap_items = std::max<size_t>(1, ((size_t)3072 << 20) / (acc_per_item * sizeof(double)));
  const double per_slot = shape.items_per_slot > 0 ? shape.items_per_slot : kDefaultItemsPerSlot;
  const double target = std::min<double>((double)cap_items / n_echunk, per_slot * slots / n_echunk);
  std::vector<BatchItem> items;
  std::vector<double> item_cost;
  std::vector<int> star_item0(n + 1, 0);
  for (int s = 0, si = 0; s < n; ++s) {
    star_item0[s] = (int)items.size();
    for (; si < n_spots && spot_star[si] == s; ++si) {
      const int t0 = spot_first_task[si], t1 = spot_first_task[si + 1];
      if (t0 == t1) continue;
      double cs = 0;
      for (int t = t0; t < t1; ++t) cs += cost[t];
      int k = (int)std::lround(total_cost > 0 ? target * cs / total_cost : 1.0);
      k = std::max(1, std::min(k, t1 - t0));
      double acc = 0;
      int start = t0, made = 0;
      for (int t = t0; t < t1; ++t) {
        acc += cost[t];
        if (t == t1 - 1 || acc >= cs * (made + 1) / k) {
          double ic = 0;
          for (int u = start; u <= t; ++u) ic += cost[u];
          for (int ec = 0; ec < n_echunk; ++ec) {
            items.push_back(BatchItem{start, t + 1, s, si, ec * batch_energies(), 0});
            item_cost.push_back(ic);
          }
          start = t + 1;
          ++made;
        }
      }
    }
  }
  star_item0[n] = (int)items.size();
  const int n_items = (int)items.size();
  std::vector<int> order(n_items);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return item_cost[a] > item_cost[b]; });

  // ---- buffers and uploads
  if ((e = W.spots.ensure(sizeof(SpotParams) * std::max(n_spots, 1))) ||
      (e = W.tasks.ensure(sizeof(RingTask) * std::max(n_tasks, 1))) ||
      (e = W.desc.ensure(sizeof(RingDesc) * std::max(n_tasks, 1))) ||
      (e = W.ph.ensure(sizeof(double) * nf1 * std::max(n_tasks, 1))) ||
      (e = W.fine.ensure(sizeof(float) * 3 * nf1 * std::max(n_tasks, 1))) ||
      (e = W.energies.ensure(sizeof(double) * n_e)) || (e = W.items.ensure(sizeof(BatchItem) * std::max(n_items, 1))) ||
      (e = BS->d_order.ensure(sizeof(int) * std::max(n_items, 1))) ||
      (e = BS->d_star_item0.ensure(sizeof(int) * (n + 1))) || (e = W.counter.ensure(sizeof(int))) || (e = W.evals.ensure(sizeof(unsigned long long))) ||
      (e = BS->d_acc.ensure(sizeof(double) * acc_per_item * std::max(n_items, 1))) ||
      (e = BS->d_seg.ensure(sizeof(double) * batch_seg_doubles() * std::max(slots, 1))) ||
      (e = W.scale.ensure(sizeof(double) * n)))
    return hip_fail(c, e, "device allocation");
  std::vector<double> fold_scale(n);
  for (int s = 0; s < n; ++s) fold_scale[s] = stars[s].exposure_s * (1.0 / g.n_phase);
  {
    Upload up;
    up.add(W.spots.p, sp.data(), sizeof(SpotParams) * n_spots);
    if (n_tasks) up.add(W.tasks.p, tasks.data(), sizeof(RingTask) * n_tasks);
    up.add(W.energies.p, energies.data(), sizeof(double) * n_e);
    if (n_items) {
      up.add(W.items.p, items.data(), sizeof(BatchItem) * n_items);
      up.add(BS->d_order.p, order.data(), sizeof(int) * n_items);
    }
    up.add(BS->d_star_item0.p, star_item0.data(), sizeof(int) * (n + 1));
    up.add(W.scale.p, fold_scale.data(), sizeof(double) * n);
    if ((e = up.run(c))) return hip_fail(c, e, "upload");
  }

  PlainLaunch P{};
  P.d_spots = (const SpotParams*)W.spots.p;
  P.d_tasks = (const RingTask*)W.tasks.p;
  P.d_desc = (RingDesc*)W.desc.p;
  P.d_ph = (double*)W.ph.p;
  P.d_fine = (float*)W.fine.p;
  P.n_tasks = n_tasks;
  P.n_fine = g.n_fine;
  BatchLaunch B{};
  B.spots = P.d_spots;
  B.tasks = P.d_tasks;
  B.desc = P.d_desc;
  B.ph = P.d_ph;
  B.fine = P.d_fine;
  B.energies = (const double*)W.energies.p;
  B.csets = (const ClassSet*)BS->d_csets.p;
  B.chdr = (const ClassHdr*)BS->d_chdr.p;
  B.cpts = (const ClassPt*)BS->d_cpts.p;
  B.items = (const BatchItem*)W.items.p;
  B.order = (const int*)BS->d_order.p;
  B.star_item0 = (const int*)BS->d_star_item0.p;
  B.queue = (int*)W.counter.p;
  B.n_evals = (unsigned long long*)W.evals.p;
  B.acc = (double*)BS->d_acc.p;
  B.seg = (double*)BS->d_seg.p;
  B.flux = d_flux;
  B.n_items = n_items;
  B.n_stars = n;
  B.n_e = n_e;
  B.n_phase = g.n_phase;
  B.n_fine = g.n_fine;
  B.nsx_rows = nsx_rows;
  B.cov_words = cov_words;
  B.ev_cap = ev_cap;
  B.atmosphere = atm;
  const int n_wg = std::max(1, std::min(slots, n_items));
  const bool fold = d_counts && g.response >= 0;

  // ---- the launch sequence, captured once per configuration and replayed
  std::vector<unsigned char> key(sizeof(P) + sizeof(B) + 4 * sizeof(int) + sizeof(void*));
  {
    unsigned char* k = key.data();
    std::memcpy(k, &P, sizeof P);
    k += sizeof P;
    std::memcpy(k, &B, sizeof B);
    k += sizeof B;
    const int extra[4] = {n_wg, fold ? g.response : -1, (int)lds, 0};
    std::memcpy(k, extra, sizeof extra);
    k += sizeof extra;
    std::memcpy(k, &d_counts, sizeof(void*));
  }
  // timed: per-kernel events around a direct launch; graph replays are timed
  // as a whole by events recorded around hipGraphLaunch (outside the graph)
  auto record = [&](bool timed) -> int {
    int r;
    if (timed && (r = hipEventRecord(c->ev0, c->stream))) return r;
    if ((r = launch_ring_setup(c->L, c->N, P, c->stream))) return r;
    if (timed && (r = hipEventRecord(c->ev_i0, c->stream))) return r;
    if (n_items && (r = launch_batched(c->N, B, n_wg, c->stream))) return r;
    if (timed && (r = hipEventRecord(c->ev_i1, c->stream))) return r;
    if ((r = launch_reduce_batched(B, c->stream))) return r;
    if (timed && (r = hipEventRecord(c->ev1, c->stream))) return r;
    if (fold) {
      const Response& R = c->resp[g.response];
      if ((r = launch_fold(R.d_hi, R.d_lo, R.n_ch, R.n_e, d_flux, n, g.n_phase, (const double*)W.scale.p, d_counts,
                           c->stream)))
        return r;
    }
    if (timed && (r = hipEventRecord(c->ev2, c->stream))) return r;
    return hipSuccess;
  };
  GraphCache& G = BS->graph;
  const bool use_graph = !c->no_graph && ((G.exec && G.key == key) || G.last_direct == key);
  if (use_graph) {
    if (!G.exec || G.key != key) {
      if (G.exec) {
        hipGraphExecDestroy(G.exec);
        G.exec = nullptr;
      }
      hipGraph_t graph = nullptr;
      if ((e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal))) return hip_fail(c, e, "capture");
      const int r = record(false);
      e = hipStreamEndCapture(c->stream, &graph);
      if (r) {
        if (graph) hipGraphDestroy(graph);
        return hip_fail(c, r, "batched launch");
      }
      if (e) return hip_fail(c, e, "capture");
      e = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (e) {
        G.exec = nullptr;
        return hip_fail(c, e, "graph instantiation");
      }
      G.key = key;
    }
    if ((e = hipEventRecord(c->ev0, c->stream))) return hip_fail(c, e, "event");
    if ((e = hipGraphLaunch(G.exec, c->stream))) return hip_fail(c, e, "graph launch");
    if ((e = hipEventRecord(c->ev2, c->stream))) return hip_fail(c, e, "event");
  } else {
    if ((e = record(true))) return hip_fail(c, e, "batched launch");
    G.last_direct = key;
  }
  if ((e = hipStreamSynchronize(c->stream))) return hip_fail(c, e, "execution");

  std::vector<RingDesc> desc(n_tasks);
  if (n_tasks && (e = hipMemcpy(desc.data(), W.desc.p, sizeof(RingDesc) * n_tasks, hipMemcpyDeviceToHost)))
    return hip_fail(c, e, "descriptor read-back");
  bool rows_overflow = false;
  for (const RingDesc& d : desc) rows_overflow |= (d.flags & 2) != 0;
  if (rows_overflow && nsx_rows < c->N.nE)
    return eval_batched_rows(c, stars, n, g, energies, pts, d_flux, d_counts, stats, c->N.nE);
  int64_t n_act = 0, n_eval_ref = 0, n_rings = 0;
  for (const RingDesc& d : desc) {
    if (d.flags & 1) return fail(c, PPM_ERR_UNSUPPORTED, "a ring has more active runs than supported");
    n_act += d.n_act;
    if (d.n_act) ++n_rings;
    if (d.kept) n_eval_ref += (int64_t)d.n_act * (int64_t)pts.size() * n_e;
  }
  if (stats) {
    float t02 = 0;
    hipEventElapsedTime(&t02, c->ev0, c->ev2);
    stats->total_ms += t02;
    if (!use_graph) {  // per-kernel times exist only for direct launches
      float t01 = 0, t12 = 0, ti = 0;
      hipEventElapsedTime(&t01, c->ev0, c->ev1);
      hipEventElapsedTime(&t12, c->ev1, c->ev2);
      hipEventElapsedTime(&ti, c->ev_i0, c->ev_i1);
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
  ppm_launch_shape cur{0, 0};
  // direct launches (the integral's own events) and the candidate shape,
  // restored on every exit
  struct Restore {
    ppm_ctx* c;
    BatchState* BS;
    DevBuf* buf;
    bool no_graph;
    ~Restore() {
      c->no_graph = no_graph;
      BS->forced = nullptr;
      hipStreamSynchronize(c->stream);
      buf->release();
    }
  } restore{c, BS, &d_flux, c->no_graph};
  c->no_graph = true;
  BS->forced = &cur;
  int n_tried = 0;
  auto time_shape = [&](ppm_launch_shape sh, double& ms) -> ppm_status {
    cur = sh;
    ms = 1e300;
    for (int r = 0; r < repeats; ++r) {
      ppm_stats st{};
      const ppm_status q = eval_any(c, stars, n, &g, (double*)d_flux.p, nullptr, &st);
      if (q) return q;
      ms = std::min(ms, st.integral_ms);
    }
    ++n_tried;
    return PPM_OK;
  };
  // untimed warm-up (allocations, first launches); it also gives the
  // occupancy limit
  {
    cur = ppm_launch_shape{0, 0};
    ppm_stats st{};
    if ((s = eval_any(c, stars, n, &g, (double*)d_flux.p, nullptr, &st))) return s;
  }
  const int wg_max = std::max(1, BS->occ_wg_max);
  ppm_launch_shape best{kDefaultItemsPerSlot, wg_max};
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
    if ((s = consider(ppm_launch_shape{f, wg_max}))) return s;
  if (best.items_per_slot != kDefaultItemsPerSlot) {
    const int f = best.items_per_slot < kDefaultItemsPerSlot ? best.items_per_slot / 2 : best.items_per_slot * 2;
    if ((s = consider(ppm_launch_shape{std::max(f, 1), wg_max}))) return s;
  }
  // one resident workgroup fewer per CU
  if (wg_max > 1 && (s = consider(ppm_launch_shape{best.items_per_slot, wg_max - 1}))) return s;
  BS->shapes[key] = best;
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
                shape->wg_per_cu > 64))
    return fail(c, PPM_ERR_ARG, "launch shape out of range (items_per_slot 0..4096, wg_per_cu 0..64)");
  std::lock_guard<std::mutex> lk(c->exec_mu);
  std::vector<int> key;
  ppm_status s = shape_args(c, grid, n_stars, atmosphere, key);
  if (s) return s;
  BatchState* BS = state(c);
  if (shape && (shape->items_per_slot || shape->wg_per_cu))
    BS->shapes[key] = *shape;
  else
    BS->shapes.erase(key);
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
  auto it = BS->shapes.find(key);
  *shape = it != BS->shapes.end() ? it->second : ppm_launch_shape{0, 0};
  return PPM_OK;
}

}  // extern "C"
