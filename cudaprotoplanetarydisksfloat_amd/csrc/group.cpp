// group.cpp -- one set of disks advanced as several batches at once
// (fgx_group_*, include/fgx.h).  No counterpart in the reference, which runs
// one disk per process (src/Main.cu:87-98).
//
// A batch's substep is a chain of launches on one stream, and the short
// per-disk launches in it (the substep's set-up, the CFL and force pass)
// leave most of the 256 CUs idle while they run.  Split over a few batches,
// each with its own stream and stepped by its own host thread, one batch's
// sweeps fill the CUs the other's short launches leave idle.  A disk's
// result does not depend on which disks share its batch, so the fields are
// those of the one big batch, bit for bit.
#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "fgx_internal.h"

struct fgx_group {
  std::vector<fgx_batch*> parts;
  std::vector<int> part_of, local;   // per disk: its batch and its index there
};

namespace {

void destroy(fgx_group* g) {
  for (fgx_batch* b : g->parts)
    if (b) fgx_batch_destroy(b);
  delete g;
}

}  // namespace

extern "C" {

int fgx_group_create(int n, const fgx_params* p, const fgx_planets* pl, int precision, int parts,
                     const fgx_tiling* t, fgx_group** out) {
  using fgx::set_error;
  if (n < 1 || !p || !pl || !out) return set_error(FGX_EPARAM, "bad group arguments");
  if (parts < 1 || parts > FGX_GROUP_MAX_PARTS)
    return set_error(FGX_EPARAM, "a group runs 1..%d batches (asked for %d)", FGX_GROUP_MAX_PARTS, parts);
  parts = std::min(parts, n);
  // largest disks first, each to the batch with the fewest cells so far
  // (ties to the lower batch); a batch keeps its disks in their given order
  std::vector<int> order(n);
  for (int d = 0; d < n; ++d) order[d] = d;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (long long)p[a].nrad * p[a].nsec > (long long)p[b].nrad * p[b].nsec;
  });
  std::vector<long long> cells(parts, 0);
  std::vector<int> part_of(n);
  for (int d : order) {
    const int k = (int)(std::min_element(cells.begin(), cells.end()) - cells.begin());
    part_of[d] = k;
    cells[k] += (long long)p[d].nrad * p[d].nsec;
  }
  fgx_group* g = new fgx_group;
  g->part_of = part_of;
  g->local.assign(n, 0);
  for (int k = 0; k < parts; ++k) {
    std::vector<fgx_params> ps;
    std::vector<fgx_planets> pls;
    for (int d = 0; d < n; ++d)
      if (part_of[d] == k) {
        g->local[d] = (int)ps.size();
        ps.push_back(p[d]);
        pls.push_back(pl[d]);
      }
    fgx_batch* b = nullptr;
    const int r = fgx_batch_create_tiled((int)ps.size(), ps.data(), pls.data(), precision, t, &b);
    if (r != FGX_OK) {
      destroy(g);
      return r;   // the batch's own message stays this thread's last error
    }
    g->parts.push_back(b);
  }
  *out = g;
  return FGX_OK;
}

// Every batch takes n_dt DTs, batch 0 on the calling thread and the others
// on threads of their own; the call returns when all have.  A batch's error
// is reported with its message; a device or parameter error ranks above
// FGX_ENUMERIC (a disk frozen for its numeric health, the others went on).
int fgx_group_step(fgx_group* g, int n_dt) {
  using fgx::set_error;
  if (!g) return set_error(FGX_EPARAM, "null handle");
  if (n_dt < 0) return set_error(FGX_EPARAM, "negative step count");
  const int np = (int)g->parts.size();
  if (np == 1) return fgx_batch_step(g->parts[0], n_dt);
  std::vector<int> rc(np, FGX_OK);
  std::vector<std::string> msg(np);
  auto run = [&](int k) {
    rc[k] = fgx_batch_step(g->parts[k], n_dt);
    if (rc[k] != FGX_OK) msg[k] = fgx::last_error();
  };
  std::vector<std::thread> th;
  th.reserve(np - 1);
  for (int k = 1; k < np; ++k) th.emplace_back(run, k);
  run(0);
  for (std::thread& x : th) x.join();
  int worst = -1;
  for (int k = 0; k < np; ++k)
    if (rc[k] != FGX_OK && (worst < 0 || (rc[worst] == FGX_ENUMERIC && rc[k] != FGX_ENUMERIC))) worst = k;
  if (worst < 0) return FGX_OK;
  return set_error(rc[worst], "%s", msg[worst].c_str());
}

int fgx_group_parts(fgx_group* g, int* n) {
  if (!g || !n) return fgx::set_error(FGX_EPARAM, "null argument");
  *n = (int)g->parts.size();
  return FGX_OK;
}

int fgx_group_batch(fgx_group* g, int part, fgx_batch** out) {
  if (!g || !out) return fgx::set_error(FGX_EPARAM, "null argument");
  if (part < 0 || part >= (int)g->parts.size())
    return fgx::set_error(FGX_ERANGE, "batch %d of a group of %d", part, (int)g->parts.size());
  *out = g->parts[part];
  return FGX_OK;
}

int fgx_group_disk(fgx_group* g, int disk, fgx_batch** batch, int* index) {
  if (!g || !batch || !index) return fgx::set_error(FGX_EPARAM, "null argument");
  if (disk < 0 || disk >= (int)g->part_of.size())
    return fgx::set_error(FGX_ERANGE, "disk index %d out of range", disk);
  *batch = g->parts[g->part_of[disk]];
  *index = g->local[disk];
  return FGX_OK;
}

int fgx_group_destroy(fgx_group* g) {
  if (g) destroy(g);
  return FGX_OK;
}

}  // extern "C"
