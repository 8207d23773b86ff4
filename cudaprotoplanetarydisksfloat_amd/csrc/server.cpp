// server.cpp -- serving loop over the batch engine (fgx_server_*).
//
// The reference runs one disk per process from main() (src/Main.cu:87-401).
// Here independent disk runs ("requests") arrive over time and share one GPU
// through continuous batching:
// - at most max_live_batches batches are live at once (no more than the
//   device's hardware queues), each advanced one DT at a time by a worker
//   thread on the batch's own HIP stream, so batches overlap on the device;
// - a batch is created for a group of waiting requests that share the batch
//   mode flags, with spare slots beside them: frozen disks of the group's
//   grids, allocated with the batch;
// - a request that has done its DTs is frozen and its slot goes back to the
//   pool at once; at its batch's next DT boundary a waiting request of the
//   same grid and modes takes the slot (fgx_batch_reset_disk: nothing is
//   allocated, the disk evolves exactly as alone);
// - a batch whose slots are all free stays live for later refills, and is
//   released when a request needs its memory or its queue position, or when
//   no request is left to come;
// - snapshots are asynchronous copies into pinned host memory, queued behind
//   the DT's kernels (the reference writes its outputs every NINTERM steps,
//   src/Main.cu:348-371).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "fgx_internal.h"

struct fgx_server {
  struct Request {
    fgx_params p;
    fgx_planets pl;
    int n_dt = 0;
    double arrival = 0;
    fgx_request_info info{};
    size_t footprint = 0;
    std::vector<int> refused;   // batches whose slot set-up refused the request
    std::vector<double> final_fields[FGX_NFIELDS];
    std::string error;   // the library's message where status is an error
  };
  struct Slot {
    int req = -1;         // request held, -1 = free (frozen)
    bool fresh = false;   // assigned; its set-up is the worker's next task
    bool initial = false; // the batch was created with this request's disk
    int left = 0;         // DTs still to do
    int nrad = 0, nsec = 0;
  };
  struct Live {
    int id = 0;
    fgx_params P0{};            // the batch's mode flags
    std::vector<Slot> slots;
    size_t foot = 0;            // device memory of all its slots
    bool retire = false;        // take no more requests; end when idle
    bool dead = false;          // the worker has ended and returned the memory
    std::thread th;
  };
  fgx_server_config cfg{};
  std::vector<Request> reqs;
  std::mutex mu;
  std::condition_variable cv;
  std::vector<std::unique_ptr<Live>> live;
  std::vector<int> waiting;     // arrived, not yet in a slot; arrival order
  size_t in_use = 0;
  fgx_server_stats stats{};
  std::chrono::steady_clock::time_point t0;
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
  int max_live() const { return cfg.max_live_batches > 0 ? cfg.max_live_batches : 4; }
};

namespace {

using Server = fgx_server;
using Live = fgx_server::Live;

bool idle(const Live& L) {
  for (const auto& sl : L.slots)
    if (sl.req >= 0) return false;
  return true;
}

// a request back into the waiting queue at its arrival position
void requeue(Server* S, int rid) {
  auto it = std::upper_bound(S->waiting.begin(), S->waiting.end(), rid, [&](int a, int b) {
    return S->reqs[a].arrival < S->reqs[b].arrival || (S->reqs[a].arrival == S->reqs[b].arrival && a < b);
  });
  S->waiting.insert(it, rid);
}

void worker(Server* S, Live* L, std::vector<fgx_params> ps, std::vector<fgx_planets> pls) {
  const int n = (int)ps.size();
  fgx_batch* b = nullptr;
  int rc = fgx_batch_create(n, ps.data(), pls.data(), S->cfg.precision, &b);
  const int fields[4] = {FGX_DENS, FGX_VRAD, FGX_VTHETA, FGX_ENERGY};
  // record a request's end: final fields (device reads) first, then the
  // record under the lock; the slot is frozen and freed by the caller
  auto finish = [&](int k, int rid, int status) {
    const Server::Request& r = S->reqs[rid];
    fgx_disk_info di{};
    if (b && status == FGX_OK) status = fgx_batch_info(b, k, &di);
    std::vector<double> keep[4];
    if (b && status == FGX_OK && S->cfg.keep_final)
      for (int f = 0; f < 4 && status == FGX_OK; ++f) {
        if (fields[f] == FGX_ENERGY && !r.p.adiabatic) continue;
        keep[f].resize((size_t)(r.p.nrad + 1) * r.p.nsec);
        status = fgx_batch_read_field(b, k, fields[f], FGX_F64, keep[f].data(), keep[f].size());
      }
    // the message of this thread's last error (errors are kept per thread)
    const std::string msg = status == FGX_OK ? std::string() : std::string(fgx::last_error());
    std::lock_guard<std::mutex> g(S->mu);
    S->reqs[rid].error = msg;
    fgx_request_info& in = S->reqs[rid].info;
    in.finish = S->now();
    in.status = status;
    in.substeps = di.substeps;
    in.cell_updates = di.substeps * (long long)r.p.nrad * r.p.nsec;
    in.batch = L->id;
    in.batch_size = n;
    in.slot = k;
    for (int f = 0; f < 4; ++f) S->reqs[rid].final_fields[fields[f]] = std::move(keep[f]);
    L->slots[k].req = -1;
    L->slots[k].fresh = false;
    S->cv.notify_all();
  };
  if (rc) {
    for (int k = 0; k < n; ++k)
      if (L->slots[k].req >= 0) finish(k, L->slots[k].req, rc);
  } else {
    // the group's requests start as created; the spare slots wait frozen
    // (a request placed in one meanwhile is set up like any refill)
    std::vector<char> spare(n, 0);
    {
      std::lock_guard<std::mutex> g(S->mu);
      for (int k = 0; k < n; ++k) {
        Server::Slot& sl = L->slots[k];
        if (sl.initial) {
          sl.fresh = false;
          sl.left = S->reqs[sl.req].n_dt;
        } else {
          spare[k] = 1;
        }
        sl.initial = false;
      }
    }
    for (int k = 0; k < n; ++k)
      if (spare[k]) fgx_batch_freeze(b, k, 1);
    std::vector<int> todo, act;
    for (;;) {
      todo.clear();
      {
        std::unique_lock<std::mutex> lk(S->mu);
        bool any = false;
        for (int k = 0; k < n; ++k) {
          const Server::Slot& sl = L->slots[k];
          if (sl.req >= 0 && sl.fresh) todo.push_back(k);
          any |= sl.req >= 0;
        }
        if (!any) {
          if (L->retire) break;
          S->cv.wait_for(lk, std::chrono::milliseconds(20));   // idle in the pool
          continue;
        }
      }
      // refills: set up the assigned requests in their slots
      for (int k : todo) {
        const int rid = L->slots[k].req;
        const Server::Request& r = S->reqs[rid];
        rc = fgx_batch_reset_disk(b, k, &r.p, &r.pl);
        std::lock_guard<std::mutex> g(S->mu);
        Server::Slot& sl = L->slots[k];
        if (rc == FGX_OK) {
          sl.fresh = false;
          sl.left = r.n_dt;
          S->stats.refills++;
        } else if (rc == FGX_EPARAM) {
          // a pass the batch lacks: the request waits for another batch
          S->reqs[rid].refused.push_back(L->id);
          S->reqs[rid].info.admit = -1;
          sl.req = -1;
          sl.fresh = false;
          requeue(S, rid);
          S->cv.notify_all();
        } else {
          sl.fresh = false;
          S->reqs[rid].info.status = rc;
          S->reqs[rid].info.finish = S->now();
          sl.req = -1;
          S->cv.notify_all();
        }
      }
      act.clear();
      {
        std::lock_guard<std::mutex> g(S->mu);
        for (int k = 0; k < n; ++k)
          if (L->slots[k].req >= 0 && !L->slots[k].fresh) act.push_back(k);
      }
      if (act.empty()) continue;
      rc = fgx_batch_step(b, 1);
      if (rc == FGX_ENUMERIC) {
        // the disks found numerically unhealthy were frozen by the step;
        // their requests end with that status, the others went on
        std::vector<int> keep;
        for (int k : act) {
          fgx_disk_info di{};
          if (fgx_batch_info(b, k, &di) == FGX_OK && di.health) {
            int had = 0;
            fgx::batch_snapshot_wait(b, k, &had);
            finish(k, L->slots[k].req, FGX_ENUMERIC);
          } else {
            keep.push_back(k);
          }
        }
        act.swap(keep);
        rc = FGX_OK;
      }
      if (rc) {
        for (int k : act) {
          finish(k, L->slots[k].req, rc);
          fgx_batch_freeze(b, k, 1);
        }
        std::lock_guard<std::mutex> g(S->mu);
        L->retire = true;   // the batch is in an unknown state
        continue;
      }
      for (int k : act) {
        const int rid = L->slots[k].req;
        const Server::Request& r = S->reqs[rid];
        const int left = --L->slots[k].left;
        const int done = r.n_dt - left;
        int had = 0;
        if (left == 0) {
          // the last snapshot has landed long ago; the slot returns to the pool
          fgx::batch_snapshot_wait(b, k, &had);
          if (had) { std::lock_guard<std::mutex> g(S->mu); S->reqs[rid].info.snapshots++; }
          fgx_batch_freeze(b, k, 1);
          finish(k, rid, FGX_OK);
        } else if (S->cfg.snapshot_every > 0 && done % S->cfg.snapshot_every == 0) {
          fgx::batch_snapshot_wait(b, k, &had);
          if (had) { std::lock_guard<std::mutex> g(S->mu); S->reqs[rid].info.snapshots++; }
          if (fgx::batch_snapshot_queue(b, k) == FGX_OK) {}
        }
      }
    }
  }
  if (b) fgx_batch_destroy(b);
  std::lock_guard<std::mutex> g(S->mu);
  S->in_use -= L->foot;
  L->dead = true;
  S->cv.notify_all();
}

// place waiting requests into free slots of live batches; then form new
// batches while the live count and the budget allow.  Called with the lock.
void admit(Server* S, double t, bool more_to_come) {
  // 1. free slots of live batches
  for (size_t q = 0; q < S->waiting.size();) {
    const int rid = S->waiting[q];
    Server::Request& r = S->reqs[rid];
    bool placed = false;
    for (auto& Lp : S->live) {
      Live& L = *Lp;
      if (L.retire || L.dead || !fgx::batch_modes_equal(r.p, L.P0)) continue;
      if (std::find(r.refused.begin(), r.refused.end(), L.id) != r.refused.end()) continue;
      for (auto& sl : L.slots)
        if (sl.req < 0 && sl.nrad == r.p.nrad && sl.nsec == r.p.nsec) {
          sl.req = rid;
          sl.fresh = true;
          r.info.admit = t;
          r.info.batch = L.id;
          placed = true;
          break;
        }
      if (placed) break;
    }
    if (placed) {
      S->waiting.erase(S->waiting.begin() + q);
      S->cv.notify_all();
    } else {
      ++q;
    }
  }
  // 2. new batches
  while (!S->waiting.empty()) {
    int nlive = 0;
    for (auto& Lp : S->live) nlive += !Lp->retire && !Lp->dead;
    const int rid0 = S->waiting[0];
    Server::Request& first = S->reqs[rid0];
    if (S->cfg.batch_window > 0 && t < first.arrival + S->cfg.batch_window && nlive > 0) break;
    // the group: the first waiting request and the next ones of its modes
    std::vector<int> group{rid0};
    size_t foot = first.footprint;
    for (size_t q = 1; q < S->waiting.size() && (int)group.size() < S->cfg.max_batch_disks; ++q) {
      const Server::Request& r = S->reqs[S->waiting[q]];
      if (!fgx::batch_modes_equal(r.p, first.p)) continue;
      if (S->in_use + foot + r.footprint > S->cfg.hbm_budget) break;
      group.push_back(S->waiting[q]);
      foot += r.footprint;
    }
    const bool room = nlive < S->max_live() && S->in_use + foot <= S->cfg.hbm_budget;
    if (!room) {
      // release an idle batch for the memory or the queue position
      for (auto& Lp : S->live)
        if (!Lp->retire && !Lp->dead && idle(*Lp)) {
          Lp->retire = true;
          S->cv.notify_all();
          break;
        }
      break;
    }
    // spare slots: frozen copies of the group's grids, while the budget holds
    std::vector<int> spare;
    for (int q = 0; q < S->cfg.spare_slots && (int)(group.size() + spare.size()) < S->cfg.max_batch_disks; ++q) {
      const int src = group[q % group.size()];
      if (S->in_use + foot + S->reqs[src].footprint > S->cfg.hbm_budget) break;
      spare.push_back(src);
      foot += S->reqs[src].footprint;
    }
    auto L = std::make_unique<Live>();
    L->id = S->stats.batches++;
    L->P0 = first.p;
    L->foot = foot;
    std::vector<fgx_params> ps;
    std::vector<fgx_planets> pls;
    for (int rid : group) {
      Server::Request& r = S->reqs[rid];
      Server::Slot sl;
      sl.req = rid; sl.fresh = true; sl.initial = true; sl.nrad = r.p.nrad; sl.nsec = r.p.nsec;
      L->slots.push_back(sl);
      ps.push_back(r.p); pls.push_back(r.pl);
      r.info.admit = t;
      r.info.batch = L->id;
      S->waiting.erase(std::find(S->waiting.begin(), S->waiting.end(), rid));
    }
    for (int src : spare) {
      Server::Slot sl;
      sl.nrad = S->reqs[src].p.nrad; sl.nsec = S->reqs[src].p.nsec;
      L->slots.push_back(sl);
      ps.push_back(S->reqs[src].p); pls.push_back(S->reqs[src].pl);
    }
    S->in_use += foot;
    S->stats.peak_bytes = std::max(S->stats.peak_bytes, S->in_use);
    Live* raw = L.get();
    S->live.push_back(std::move(L));
    raw->th = std::thread(worker, S, raw, std::move(ps), std::move(pls));
    S->stats.peak_live = std::max(S->stats.peak_live, nlive + 1);
  }
  // no request left to come or wait: the idle batches are done
  if (!more_to_come && S->waiting.empty())
    for (auto& Lp : S->live)
      if (!Lp->retire) { Lp->retire = true; S->cv.notify_all(); }
}

}  // namespace

extern "C" {

int fgx_server_create(const fgx_server_config* cfg, fgx_server** out) {
  if (!cfg || !out) return fgx::set_error(FGX_EPARAM, "null argument");
  if (cfg->precision != FGX_F32 && cfg->precision != FGX_F64) return fgx::set_error(FGX_EPARAM, "bad precision");
  if (cfg->hbm_budget == 0 || cfg->max_batch_disks < 1) return fgx::set_error(FGX_EPARAM, "bad server budget");
  if (cfg->max_live_batches < 0 || cfg->spare_slots < 0) return fgx::set_error(FGX_EPARAM, "bad slot counts");
  auto* s = new fgx_server();
  s->cfg = *cfg;
  *out = s;
  return FGX_OK;
}

int fgx_server_submit(fgx_server* s, const fgx_params* p, const fgx_planets* pl, int n_dt, double arrival, int* id) {
  if (!s || !p || !pl || n_dt < 1 || arrival < 0) return fgx::set_error(FGX_EPARAM, "bad request");
  int rc = fgx::params_validate(p);
  if (rc) return rc;
  fgx_server::Request r;
  r.p = *p;
  r.pl = *pl;
  r.n_dt = n_dt;
  r.arrival = arrival;
  r.info.arrival = arrival;
  r.info.admit = r.info.finish = -1;
  r.info.n_dt = n_dt;
  r.info.batch = r.info.slot = -1;
  rc = fgx_footprint_bytes(p, s->cfg.precision, &r.footprint);
  if (rc) return rc;
  if (r.footprint > s->cfg.hbm_budget)
    return fgx::set_error(FGX_ENOMEM, "request needs %zu B, more than the server budget %zu B", r.footprint,
                          s->cfg.hbm_budget);
  std::lock_guard<std::mutex> g(s->mu);
  s->reqs.push_back(std::move(r));
  if (id) *id = (int)s->reqs.size() - 1;
  return FGX_OK;
}

int fgx_server_run(fgx_server* s) {
  if (!s) return fgx::set_error(FGX_EPARAM, "null handle");
  std::vector<int> order(s->reqs.size());
  for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return s->reqs[a].arrival < s->reqs[b].arrival; });
  s->t0 = std::chrono::steady_clock::now();
  std::unique_lock<std::mutex> lk(s->mu);
  size_t next = 0;   // first request in arrival order not yet arrived
  for (;;) {
    const double t = s->now();
    while (next < order.size() && s->reqs[order[next]].arrival <= t) s->waiting.push_back(order[next++]);
    admit(s, t, next < order.size());
    // join the batches that have ended
    for (size_t k = 0; k < s->live.size();) {
      if (s->live[k]->dead) {
        std::unique_ptr<Live> L = std::move(s->live[k]);
        s->live.erase(s->live.begin() + k);
        lk.unlock();
        L->th.join();
        lk.lock();
      } else {
        ++k;
      }
    }
    if (next >= order.size() && s->waiting.empty() && s->live.empty()) break;
    // sleep until the next arrival (or window end), or until a slot or a
    // batch frees up
    double wake = 1e30;
    if (next < order.size()) wake = s->reqs[order[next]].arrival;
    if (s->cfg.batch_window > 0 && !s->waiting.empty()) wake = std::min(wake, s->reqs[s->waiting[0]].arrival + s->cfg.batch_window);
    const double dt = std::max(0.0, std::min(wake - s->now(), 0.05));
    s->cv.wait_for(lk, std::chrono::duration<double>(dt));
  }
  lk.unlock();
  // the first failed request's status and message (set in its worker)
  for (size_t k = 0; k < s->reqs.size(); ++k)
    if (s->reqs[k].info.status != FGX_OK)
      return fgx::set_error(s->reqs[k].info.status, "request %zu: %s", k, s->reqs[k].error.c_str());
  return FGX_OK;
}

int fgx_server_stats_get(fgx_server* s, fgx_server_stats* out) {
  if (!s || !out) return fgx::set_error(FGX_EPARAM, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->stats;
  return FGX_OK;
}

int fgx_server_request(fgx_server* s, int id, fgx_request_info* out) {
  if (!s || !out) return fgx::set_error(FGX_EPARAM, "null argument");
  if (id < 0 || id >= (int)s->reqs.size()) return fgx::set_error(FGX_ERANGE, "request %d out of range", id);
  *out = s->reqs[id].info;
  return FGX_OK;
}

int fgx_server_read_field(fgx_server* s, int id, int field, int dtype, void* host, size_t count) {
  if (!s || !host) return fgx::set_error(FGX_EPARAM, "null argument");
  if (id < 0 || id >= (int)s->reqs.size()) return fgx::set_error(FGX_ERANGE, "request %d out of range", id);
  if (field < 0 || field >= FGX_NFIELDS) return fgx::set_error(FGX_ERANGE, "field %d out of range", field);
  const std::vector<double>& v = s->reqs[id].final_fields[field];
  if (v.empty()) return fgx::set_error(FGX_ERANGE, "no final field %d kept for request %d", field, id);
  const size_t n = std::min(count, v.size());
  if (dtype == FGX_F32) for (size_t k = 0; k < n; ++k) ((float*)host)[k] = (float)v[k];
  else std::memcpy(host, v.data(), n * sizeof(double));
  return FGX_OK;
}

int fgx_server_destroy(fgx_server* s) {
  delete s;
  return FGX_OK;
}

}  // extern "C"
