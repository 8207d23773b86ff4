// server.cpp -- serving loop over the batch engine (fgx_server_*).
//
// The reference runs one disk per process from main() (src/Main.cu:87-401).
// Here many independent disk runs ("requests") arrive over time and share
// one GPU: an admission thread hands each group of arrived requests to a
// worker that builds a batch (own HIP stream) and advances it one DT at a
// time, retiring requests as they finish and snapshotting fields on the
// reference's output cadence.  Workers of different batches run
// concurrently, so their kernels overlap on the device.
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
    bool admitted = false;
    std::vector<double> final_fields[FGX_NFIELDS];
  };
  fgx_server_config cfg{};
  std::vector<Request> reqs;
  std::mutex mu;
  std::condition_variable cv;
  size_t in_use = 0;
  int live = 0, batches = 0;
  std::chrono::steady_clock::time_point t0;
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

void run_batch(fgx_server* S, std::vector<int> ids) {
  const int n = (int)ids.size();
  std::vector<fgx_params> ps(n);
  std::vector<fgx_planets> pls(n);
  size_t foot = 0;
  int batch;
  {
    std::lock_guard<std::mutex> g(S->mu);
    for (int k = 0; k < n; ++k) { ps[k] = S->reqs[ids[k]].p; pls[k] = S->reqs[ids[k]].pl; foot += S->reqs[ids[k]].footprint; }
    batch = S->batches++;
  }
  fgx_batch* b = nullptr;
  int rc = fgx_batch_create(n, ps.data(), pls.data(), S->cfg.precision, &b);
  std::vector<int> left(n);
  std::vector<long long> cells(n);
  for (int k = 0; k < n; ++k) {
    left[k] = S->reqs[ids[k]].n_dt;
    cells[k] = (long long)ps[k].nrad * ps[k].nsec;
  }
  std::vector<double> snap;
  const int fields[4] = {FGX_DENS, FGX_VRAD, FGX_VTHETA, FGX_ENERGY};
  auto finish = [&](int k, int status) {
    fgx_disk_info di{};
    if (b && status == FGX_OK) status = fgx_batch_info(b, k, &di);
    std::vector<double> keep[4];
    if (b && status == FGX_OK && S->cfg.keep_final)
      for (int f = 0; f < 4; ++f) {
        if (fields[f] == FGX_ENERGY && !ps[k].adiabatic) continue;
        keep[f].resize((size_t)(ps[k].nrad + 1) * ps[k].nsec);
        if (fgx_batch_read_field(b, k, fields[f], FGX_F64, keep[f].data(), keep[f].size())) status = FGX_EDEVICE;
      }
    std::lock_guard<std::mutex> g(S->mu);
    fgx_request_info& in = S->reqs[ids[k]].info;
    in.finish = S->now();
    in.status = status;
    in.substeps = di.substeps;
    in.cell_updates = di.substeps * cells[k];
    in.batch = batch;
    in.batch_size = n;
    for (int f = 0; f < 4; ++f) S->reqs[ids[k]].final_fields[fields[f]] = std::move(keep[f]);
  };
  if (rc) {
    for (int k = 0; k < n; ++k) finish(k, rc);
  } else {
    for (int step = 1;; ++step) {
      bool any = false;
      for (int k = 0; k < n; ++k) any |= left[k] > 0;
      if (!any) break;
      rc = fgx_batch_step(b, 1);
      if (rc == FGX_ENUMERIC) {
        // the disks found numerically unhealthy were frozen by the step; their
        // requests end with that status, the others went on
        for (int k = 0; k < n; ++k) {
          fgx_disk_info di{};
          if (left[k] > 0 && fgx_batch_info(b, k, &di) == FGX_OK && di.health) {
            left[k] = 0;
            finish(k, FGX_ENUMERIC);
          }
        }
        rc = FGX_OK;
      }
      if (rc) {
        for (int k = 0; k < n; ++k)
          if (left[k] > 0) { left[k] = 0; finish(k, rc); }
        break;
      }
      for (int k = 0; k < n; ++k) {
        if (left[k] <= 0) continue;
        if (--left[k] == 0) {
          finish(k, FGX_OK);
          fgx_batch_freeze(b, k, 1);
        } else if (S->cfg.snapshot_every > 0 && step % S->cfg.snapshot_every == 0) {
          // snapshot of the running request (WriteDiskPolar cadence)
          snap.resize((size_t)(ps[k].nrad + 1) * ps[k].nsec);
          for (int f = 0; f < 3; ++f) fgx_batch_read_field(b, k, fields[f], FGX_F64, snap.data(), snap.size());
          std::lock_guard<std::mutex> g(S->mu);
          S->reqs[ids[k]].info.snapshots++;
        }
      }
    }
  }
  if (b) fgx_batch_destroy(b);
  std::lock_guard<std::mutex> g(S->mu);
  S->in_use -= foot;
  S->live--;
  S->cv.notify_all();
}

}  // namespace

extern "C" {

int fgx_server_create(const fgx_server_config* cfg, fgx_server** out) {
  if (!cfg || !out) return fgx::set_error(FGX_EPARAM, "null argument");
  if (cfg->precision != FGX_F32 && cfg->precision != FGX_F64) return fgx::set_error(FGX_EPARAM, "bad precision");
  if (cfg->hbm_budget == 0 || cfg->max_batch_disks < 1) return fgx::set_error(FGX_EPARAM, "bad server budget");
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
  r.info.batch = -1;
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
  std::vector<std::thread> workers;
  s->t0 = std::chrono::steady_clock::now();
  std::unique_lock<std::mutex> lk(s->mu);
  size_t next = 0;   // first request in arrival order not yet admitted
  for (;;) {
    const double t = s->now();
    // admission in arrival order: a group is the first waiting request and
    // the next arrived ones with the same modes, while the budget holds
    while (next < order.size() && s->reqs[order[next]].arrival <= t) {
      fgx_server::Request& first = s->reqs[order[next]];
      // wait a little for companions to batch with, unless the window is over
      if (s->cfg.batch_window > 0 && t < first.arrival + s->cfg.batch_window && s->live > 0) break;
      if (s->in_use + first.footprint > s->cfg.hbm_budget) break;
      std::vector<int> group{order[next]};
      size_t used = first.footprint;
      size_t k = next + 1;
      while (k < order.size() && (int)group.size() < s->cfg.max_batch_disks && s->reqs[order[k]].arrival <= t &&
             fgx::batch_modes_equal(s->reqs[order[k]].p, first.p) && s->in_use + used + s->reqs[order[k]].footprint <= s->cfg.hbm_budget) {
        group.push_back(order[k]);
        used += s->reqs[order[k]].footprint;
        ++k;
      }
      // requests skipped for a different mode stay in order behind this group
      std::vector<int> rest;
      for (size_t q = next + 1; q < k; ++q)
        if (std::find(group.begin(), group.end(), order[q]) == group.end()) rest.push_back(order[q]);
      for (int id : group) { s->reqs[id].admitted = true; s->reqs[id].info.admit = t; }
      std::copy(rest.begin(), rest.end(), order.begin() + next + group.size());
      std::copy(group.begin(), group.end(), order.begin() + next);
      next += group.size();
      s->in_use += used;
      s->live++;
      workers.emplace_back(run_batch, s, group);
    }
    if (next >= order.size() && s->live == 0) break;
    // sleep until the next arrival (or window end) or until a batch retires
    double wake = 1e30;
    if (next < order.size()) {
      wake = s->reqs[order[next]].arrival;
      if (s->cfg.batch_window > 0 && s->live > 0) wake += s->cfg.batch_window;
    }
    if (wake < 1e29) {
      const double dt = std::max(0.0, wake - s->now());
      s->cv.wait_for(lk, std::chrono::duration<double>(std::min(dt, 0.05)));
    } else {
      s->cv.wait_for(lk, std::chrono::milliseconds(50));
    }
  }
  lk.unlock();
  for (auto& w : workers) w.join();
  int rc = FGX_OK;
  for (auto& r : s->reqs)
    if (r.info.status != FGX_OK && rc == FGX_OK) rc = r.info.status;
  return rc;
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
