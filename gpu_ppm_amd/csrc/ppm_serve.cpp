// Serving loop: continuous batching of single-profile requests, with work
// of requests that arrive while others run overlapping on the device.
//
// ppm_submit validates a request, copies its star, grid and energy grid,
// estimates its work and device memory, charges the memory against the
// context's budget and queues it.  Two workers, each with its own lane
// (stream and buffers), take queued requests whenever they are free:
//   - the bulk worker takes every queued request (the largest batches);
//   - the express worker takes light requests (estimated work below
//     kLightWork) while the bulk worker has a launch in flight.
// Serving launches run one workgroup per work item (not persistent
// workgroups), so workgroups retire every few milliseconds and the device
// dispatches an express launch's workgroups among a running bulk launch's:
// a light request that arrives during a heavy launch does not wait for it.
// A worker takes every queued request it may serve of the oldest such
// request's launch class (atmosphere and, for NSX, the interpolant's
// energy-axis evaluation), up to its lane's memory share, groups them by
// grid and evaluates them in persistent batched launches (several grids per
// launch, the cheapest launch first).  The results of a launch are copied to
// host memory of its requests as soon as it ends and handed out by ppm_poll
// in any order.  The reference has no serving layer; each request is one
// calc_one_star (gpu/devarshi24st_gpu.cu:431-511) without the text output.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <numeric>
#include <unordered_map>

#include "ppm_batch.hpp"
#include "ppm_batch_dev.hpp"
#include "ppm_ctx.hpp"

namespace ppm {

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kMaxBatch = 1024;          // requests per launch sequence
constexpr size_t kMaxQueued = 1 << 16;   // queued + running requests
constexpr double kLightWork = 1e11;      // reference-space evaluations of a light request
enum { kExpress = 0, kBulk = 1, kWorkers = 2 };

enum class ReqState { kQueued, kRunning, kDone, kFailed, kCancelled };

struct Request {
  ppm_star star{};
  ppm_grid grid{};
  std::vector<double> energies;  // the grid's energies, owned here
  int n_ch = 0;                  // channels of the fold, 0 without
  double work = 0;               // estimated evaluations in the reference iteration space
  size_t bytes = 0;              // device memory charged at admission (outputs and ring tables)
  size_t acc_bytes = 0;          // accumulators at one work item per spot and energy chunk
  std::vector<double> flux, counts;
  Clock::time_point t_submit, t_done;
  ReqState state = ReqState::kQueued;
  bool cancel_pending = false;   // cancelled while running
  ppm_status status = PPM_OK;
};

// Requests evaluate together when everything but the star is equal.
bool same_grid(const Request& a, const Request& b) {
  const ppm_grid &x = a.grid, &y = b.grid;
  return x.n_side == y.n_side && x.n_phase == y.n_phase && x.sampling == y.sampling &&
         (x.sampling != PPM_SAMPLE_GL || x.n_gl == y.n_gl) && x.n_fine == y.n_fine && x.response == y.response &&
         a.energies == b.energies && batch_class(a.star) == batch_class(b.star);
}

}  // namespace

struct ServeWorker {
  Lane lane;
  std::thread thread;
  bool busy = false;
  DevBuf d_flux, d_counts;
  HostBuf h_flux, h_counts;
  int64_t launches = 0, overlapped = 0;
  size_t held = 0, held_fixed = 0;  // device bytes of the lane's buffers after its last launch
};

struct ServeState {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> queue;  // queued tickets, arrival order
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_ticket = 1;
  size_t budget = 0, charged = 0;  // the context's hbm_budget_bytes
  size_t lane_share = 0;  // per worker: accumulator cap, and request bytes of one launch at most
  int active = 0;         // queued + running
  bool stop = false, started = false;
  ServeWorker workers[kWorkers];
};

namespace {

// Whether worker w may take request r now (S.mu held).
bool eligible(const ServeState& S, int w, const Request& r) {
  if (w == kBulk) return true;
  return S.workers[kBulk].busy && r.work < kLightWork;
}

void worker_loop(ppm_ctx* c, int w) {
  ServeState& S = *c->serve;
  ServeWorker& W = S.workers[w];
  Lane& L = W.lane;
  hipSetDevice(c->device);
  std::vector<Request*> batch;
  std::vector<uint64_t> tickets;
  std::vector<std::vector<int>> members;  // per grid group: indices into batch
  std::vector<std::vector<ppm_star>> stars;
  for (;;) {
    batch.clear();
    tickets.clear();
    {
      std::unique_lock<std::mutex> lk(S.mu);
      auto first = S.queue.end();
      S.cv_work.wait(lk, [&] {
        if (S.stop) return true;
        first = std::find_if(S.queue.begin(), S.queue.end(),
                             [&](uint64_t t) { return eligible(S, w, *S.reqs.at(t)); });
        return first != S.queue.end();
      });
      if (S.stop) return;
      // the oldest request this worker may serve and every queued one of
      // its launch class, within the lane's memory share
      const int cls = batch_class(S.reqs.at(*first)->star);
      size_t acc = 0, bytes = 0;
      for (auto it = first; it != S.queue.end() && (int)batch.size() < kMaxBatch;) {
        Request* r = S.reqs.at(*it).get();
        if (batch_class(r->star) == cls && eligible(S, w, *r) &&
            (batch.empty() || (acc + r->acc_bytes <= S.lane_share && bytes + r->bytes <= S.lane_share))) {
          acc += r->acc_bytes;
          bytes += r->bytes;
          r->state = ReqState::kRunning;
          batch.push_back(r);
          tickets.push_back(*it);
          it = S.queue.erase(it);
        } else {
          ++it;
        }
      }
      W.busy = true;
      if (w == kExpress) ++W.overlapped;  // express launches start only while a bulk launch runs
      ++W.launches;
    }
    const int n = (int)batch.size();
    // group by grid, groups in order of their first request
    members.clear();
    for (int k = 0; k < n; ++k) {
      size_t gi = 0;
      while (gi < members.size() && !same_grid(*batch[members[gi][0]], *batch[k])) ++gi;
      if (gi == members.size()) members.emplace_back();
      members[gi].push_back(k);
    }
    const int n_groups = (int)members.size();
    std::vector<size_t> f_off(n_groups + 1, 0), c_off(n_groups + 1, 0);
    for (int gi = 0; gi < n_groups; ++gi) {
      const Request& r0 = *batch[members[gi][0]];
      const size_t m = members[gi].size();
      f_off[gi + 1] = f_off[gi] + m * r0.grid.n_phase * r0.energies.size();
      c_off[gi + 1] = c_off[gi] + m * (size_t)r0.n_ch * r0.grid.n_phase;
    }
    std::vector<ppm_status> st(n_groups, PPM_OK);
    // hands the results of groups gis (already in host memory, or failed)
    // to their requests
    auto publish = [&](const std::vector<int>& gis) {
      const Clock::time_point now = Clock::now();
      {
        std::lock_guard<std::mutex> lk(S.mu);
        for (int gi : gis) {
          const size_t per_flux = (f_off[gi + 1] - f_off[gi]) / members[gi].size();
          const size_t per_counts = (c_off[gi + 1] - c_off[gi]) / members[gi].size();
          for (size_t j = 0; j < members[gi].size(); ++j) {
            const int k = members[gi][j];
            Request* r = batch[k];
            S.charged -= r->bytes + r->acc_bytes;
            --S.active;
            if (r->cancel_pending) {
              S.reqs.erase(tickets[k]);
              continue;
            }
            r->t_done = now;
            r->status = st[gi];
            if (st[gi] == PPM_OK) {
              const double* f = (const double*)W.h_flux.p + f_off[gi] + per_flux * j;
              r->flux.assign(f, f + per_flux);
              if (per_counts) {
                const double* q = (const double*)W.h_counts.p + c_off[gi] + per_counts * j;
                r->counts.assign(q, q + per_counts);
              }
              r->state = ReqState::kDone;
            } else {
              r->state = ReqState::kFailed;
            }
          }
        }
      }
      S.cv_done.notify_all();
    };
    // copies the results of groups gis to host memory, then publishes them
    auto finish = [&](const std::vector<int>& gis) {
      int e = 0;
      for (int gi : gis) {
        if (st[gi] != PPM_OK) continue;
        if (!e && f_off[gi + 1] > f_off[gi])
          e = hipMemcpyAsync((double*)W.h_flux.p + f_off[gi], (const double*)W.d_flux.p + f_off[gi],
                             sizeof(double) * (f_off[gi + 1] - f_off[gi]), hipMemcpyDeviceToHost, L.stream);
        if (!e && c_off[gi + 1] > c_off[gi])
          e = hipMemcpyAsync((double*)W.h_counts.p + c_off[gi], (const double*)W.d_counts.p + c_off[gi],
                             sizeof(double) * (c_off[gi + 1] - c_off[gi]), hipMemcpyDeviceToHost, L.stream);
      }
      if (!e) e = hipStreamSynchronize(L.stream);
      if (e) {
        const ppm_status q = hip_fail(c, e, "serving read-back");
        for (int gi : gis)
          if (st[gi] == PPM_OK) st[gi] = q;
      }
      publish(gis);
    };
    int e = 0;
    if ((e = W.d_flux.ensure(sizeof(double) * f_off[n_groups])) ||
        (e = W.h_flux.ensure(sizeof(double) * f_off[n_groups])) ||
        (c_off[n_groups] && ((e = W.d_counts.ensure(sizeof(double) * c_off[n_groups])) ||
                             (e = W.h_counts.ensure(sizeof(double) * c_off[n_groups]))))) {
      st.assign(n_groups, hip_fail(c, e, "serving buffers"));
      std::vector<int> all(n_groups);
      std::iota(all.begin(), all.end(), 0);
      publish(all);
    } else {
      std::vector<BatchGroup> groups(n_groups);
      stars.assign(n_groups, {});
      for (int gi = 0; gi < n_groups; ++gi) {
        const Request& r0 = *batch[members[gi][0]];
        for (int k : members[gi]) stars[gi].push_back(batch[k]->star);
        BatchGroup& G = groups[gi];
        G.stars = stars[gi].data();
        G.n = (int)stars[gi].size();
        G.g = r0.grid;
        G.g.energies_keV = nullptr;
        G.g.n_energy = (int)r0.energies.size();
        G.energies = r0.energies;
        phase_points(G.g, G.pts);
        G.d_flux = (double*)W.d_flux.p + f_off[gi];
        G.d_counts = r0.n_ch ? (double*)W.d_counts.p + c_off[gi] : nullptr;
      }
      // launches: one per class of grids (2, 4 or 6 energies per lane, fine
      // tables up to 384 nodes or more: the kernel instance and the LDS,
      // hence the occupancy, follow the largest grid of a launch),
      // groups packed in order while the LDS of one workgroup holds them;
      // a group the batched kernel cannot take (energies not strictly
      // increasing, or too large alone) runs on the plain path
      std::vector<std::vector<int>> launches;
      std::vector<int> plain;
      for (int cls = 0; cls < 6; ++cls) {  // (2, 4 or 6 energies per lane) x (n_fine <= 384 or more)
        std::vector<BatchGroup> cur;
        std::vector<int> cur_idx;
        for (int gi = 0; gi < n_groups; ++gi) {
          const BatchGroup& G = groups[gi];
          if ((batch_grid_epl(c, G) / 2 - 1) + 3 * (G.g.n_fine > 384) != cls) continue;
          bool ascending = true;
          for (size_t q = 1; q < G.energies.size(); ++q) ascending = ascending && G.energies[q] > G.energies[q - 1];
          if (!ascending || !batch_groups_fit(c, {G})) {
            plain.push_back(gi);
            continue;
          }
          cur.push_back(G);
          if (!batch_groups_fit(c, cur)) {
            cur.pop_back();
            launches.push_back(cur_idx);
            cur.assign(1, G);
            cur_idx.clear();
          }
          cur_idx.push_back(gi);
        }
        if (!cur_idx.empty()) launches.push_back(cur_idx);
      }
      // cheapest launch first, each launch's requests completed as soon as
      // it ends
      auto work = [&](const std::vector<int>& Lg) {
        double wsum = 0;
        for (int gi : Lg)
          for (int k : members[gi]) wsum += batch[k]->work;
        return wsum;
      };
      std::stable_sort(launches.begin(), launches.end(),
                       [&](const std::vector<int>& a, const std::vector<int>& b) { return work(a) < work(b); });
      for (const std::vector<int>& Lg : launches) {
        std::vector<BatchGroup> lg;
        for (int gi : Lg) lg.push_back(groups[gi]);
        ppm_stats stats{};
        const ppm_status q = eval_batched_groups(c, L, lg, &stats);
        for (int gi : Lg) st[gi] = q;
        finish(Lg);
      }
      for (int gi : plain) {
        ppm_grid g = groups[gi].g;
        g.energies_keV = groups[gi].energies.data();
        g.method = PPM_METHOD_PLAIN;
        ppm_stats stats{};
        st[gi] = eval_any(c, L, groups[gi].stars, groups[gi].n, &g, groups[gi].d_flux, groups[gi].d_counts, &stats);
        finish({gi});
      }
    }
    {
      size_t fixed = 0;
      const size_t held = lane_device_bytes(L, &fixed) + W.d_flux.cap + W.d_counts.cap;
      std::lock_guard<std::mutex> lk(S.mu);
      W.busy = false;
      W.held = held;
      W.held_fixed = fixed;
    }
    S.cv_work.notify_all();  // requests held back for this lane may go elsewhere now
  }
}

// Work and device memory of a request: evaluations in the reference
// iteration space (spot solid angle x pixels x phase points x energies), the
// outputs (device and its pinned host copy count once) and ring tables of its
// candidate rings, and its accumulators at one work item per spot and energy
// chunk (the least a batched launch allocates for it; more items only
// spread over the lane's fixed accumulator cap).
void request_cost(const ppm_ctx* c, Request& r) {
  const ppm_grid& g = r.grid;
  const size_t n_e = r.energies.size();
  const size_t n_pts = (size_t)g.n_phase * (g.sampling == PPM_SAMPLE_GL ? g.n_gl : 1);
  const size_t nf1 = (size_t)g.n_fine + 1;
  size_t n_rings = 0;
  double area = 0;
  std::vector<int> rings;
  for (int s = 0; s < r.star.n_spots; ++s) {
    SpotParams P;
    spot_params(c, r.star, r.star.spot[s], g.n_side, P);
    candidate_rings(P, r.star.spot[s], g.n_side, rings);
    n_rings += rings.size();
    area += P.point_source ? 1.0 / (12.0 * g.n_side * g.n_side) : 0.5 * (1.0 - std::cos(r.star.spot[s].angrad1));
  }
  r.work = area * 12.0 * g.n_side * g.n_side * (double)n_pts * (double)n_e;
  const size_t out = sizeof(double) * ((size_t)g.n_phase * n_e + (size_t)r.n_ch * g.n_phase);
  size_t tables = n_rings * (nf1 * (sizeof(double) + 3 * sizeof(float)) + sizeof(RingDesc) + sizeof(RingTask)) +
                  sizeof(SpotParams) * r.star.n_spots;
  bool ascending = true;
  for (size_t q = 1; q < n_e; ++q) ascending = ascending && r.energies[q] > r.energies[q - 1];
  if (!ascending)  // the plain path's per-ring partials
    tables += n_rings * (size_t)g.n_phase * n_e * sizeof(float) + 64 * (size_t)g.n_phase * n_e * sizeof(double);
  r.bytes = out + tables;
  r.acc_bytes = ascending ? sizeof(double) * (size_t)r.star.n_spots * batch_acc_copies() * g.n_phase * n_e : 0;
}

}  // namespace

void serve_init(ppm_ctx* c) {
  c->serve = new ServeState;
  if (c->budget) c->serve->budget = c->budget;
  // the budget splits evenly between the two workers; each worker's share
  // caps its launches' accumulators and, separately, the outputs and ring
  // tables of the requests of one launch (half the share each)
  c->serve->lane_share = std::min(kAccCapBytes, c->serve->budget / (2 * kWorkers));
}

void serve_shutdown(ppm_ctx* c) {
  ServeState* S = c->serve;
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (ServeWorker& W : S->workers) {
    if (W.thread.joinable()) W.thread.join();
    W.lane.release();
    W.d_flux.release();
    W.d_counts.release();
    W.h_flux.release();
    W.h_counts.release();
  }
  delete S;
  c->serve = nullptr;
}

}  // namespace ppm

using namespace ppm;

extern "C" {

ppm_status ppm_submit(ppm_ctx* c, const ppm_star* star, const ppm_grid* grid, uint64_t* ticket) {
  if (!c || !star || !grid || !ticket || !c->serve) return PPM_ERR_ARG;
  auto r = std::make_unique<Request>();
  r->star = *star;
  r->grid = *grid;
  ppm_status s = resolve_energies(c, *grid, r->energies);
  if (s) return s;
  if ((s = validate(c, r->star, r->grid, (int)r->energies.size()))) return s;
  r->grid.energies_keV = nullptr;  // the copy in r->energies is used
  r->grid.n_energy = (int)r->energies.size();
  r->n_ch = grid->response >= 0 ? c->resp[grid->response].n_ch : 0;
  request_cost(c, *r);
  ServeState& S = *c->serve;
  if (r->bytes > S.lane_share || r->acc_bytes > S.lane_share)
    return fail(c, PPM_ERR_NOMEM,
                "request needs " + std::to_string(r->bytes) + " bytes of tables and outputs and " +
                    std::to_string(r->acc_bytes) + " of accumulators; one serving launch may use " +
                    std::to_string(S.lane_share) + " of each (a quarter of the budget)");
  {
    std::lock_guard<std::mutex> lk(S.mu);
    if ((size_t)S.active >= kMaxQueued) return fail(c, PPM_ERR_BUSY, "serving queue full");
    const size_t need = r->bytes + r->acc_bytes;
    if (S.charged + need > S.budget) return fail(c, PPM_ERR_BUSY, "device-memory budget of the serving queue in use");
    if (!S.started) {
      for (int w = 0; w < kWorkers; ++w) {
        ServeWorker& W = S.workers[w];
        const int e = W.lane.init();
        if (e) return hip_fail(c, e, "serving lane");
        W.lane.no_graph = c->no_graph;
        W.lane.acc_cap = S.lane_share;
        W.lane.per_item = true;
      }
      for (int w = 0; w < kWorkers; ++w) S.workers[w].thread = std::thread(worker_loop, c, w);
      S.started = true;
    }
    S.charged += need;
    ++S.active;
    r->t_submit = Clock::now();
    *ticket = S.next_ticket++;
    S.queue.push_back(*ticket);
    S.reqs.emplace(*ticket, std::move(r));
  }
  S.cv_work.notify_all();
  return PPM_OK;
}

ppm_status ppm_poll(ppm_ctx* c, uint64_t ticket, double* flux, double* counts, double* latency_ms) {
  if (!c || !c->serve) return PPM_ERR_ARG;
  ServeState& S = *c->serve;
  std::lock_guard<std::mutex> lk(S.mu);
  auto it = S.reqs.find(ticket);
  if (it == S.reqs.end()) return PPM_ERR_NOT_FOUND;
  Request& r = *it->second;
  switch (r.state) {
    case ReqState::kQueued:
    case ReqState::kRunning:
      return PPM_PENDING;
    case ReqState::kCancelled:
      S.reqs.erase(it);
      return PPM_ERR_CANCELLED;
    case ReqState::kFailed: {
      const ppm_status st = r.status;
      S.reqs.erase(it);
      return st;
    }
    case ReqState::kDone:
      break;
  }
  if (flux) std::copy(r.flux.begin(), r.flux.end(), flux);
  if (counts && !r.counts.empty()) std::copy(r.counts.begin(), r.counts.end(), counts);
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
  S.reqs.erase(it);
  return PPM_OK;
}

ppm_status ppm_cancel(ppm_ctx* c, uint64_t ticket) {
  if (!c || !c->serve) return PPM_ERR_ARG;
  ServeState& S = *c->serve;
  {
    std::lock_guard<std::mutex> lk(S.mu);
    auto it = S.reqs.find(ticket);
    if (it == S.reqs.end()) return PPM_ERR_NOT_FOUND;
    Request& r = *it->second;
    if (r.state == ReqState::kQueued) {
      S.queue.erase(std::find(S.queue.begin(), S.queue.end(), ticket));
      S.charged -= r.bytes + r.acc_bytes;
      --S.active;
      r.state = ReqState::kCancelled;
    } else if (r.state == ReqState::kRunning) {
      r.cancel_pending = true;  // the result is dropped when the launch ends
    } else {
      S.reqs.erase(it);  // finished: the result is discarded
    }
  }
  S.cv_done.notify_all();
  return PPM_OK;
}

ppm_status ppm_serve_stats_get(ppm_ctx* c, ppm_serve_stats* out) {
  if (!c || !c->serve || !out) return PPM_ERR_ARG;
  ServeState& S = *c->serve;
  std::lock_guard<std::mutex> lk(S.mu);
  *out = ppm_serve_stats{};
  out->queued = (int64_t)S.queue.size();
  out->running = S.active - out->queued;
  for (int w = 0; w < kWorkers; ++w) {
    out->launches[w] = S.workers[w].launches;
    out->device_bytes += S.workers[w].held;
    out->fixed_bytes += S.workers[w].held_fixed;
  }
  out->overlapped = S.workers[kExpress].overlapped;
  out->charged_bytes = S.charged;
  out->budget_bytes = S.budget;
  return PPM_OK;
}

ppm_status ppm_drain(ppm_ctx* c) {
  if (!c || !c->serve) return PPM_ERR_ARG;
  ServeState& S = *c->serve;
  std::unique_lock<std::mutex> lk(S.mu);
  S.cv_done.wait(lk, [&] { return S.active == 0; });
  return PPM_OK;
}

}  // extern "C"
