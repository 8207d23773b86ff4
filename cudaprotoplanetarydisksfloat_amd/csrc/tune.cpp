// tune.cpp -- the tile-shape autotuner (fgx_autotune, include/fgx.h).
//
// The batch kernels walk strips of rings.  Tall strips amortise the halo
// rings each strip re-reads and re-computes; short strips give more blocks,
// which a batch needs to fill 256 CUs when few of its disks are still active.
// Where the balance lies depends on the mix of disk sizes, on how unevenly
// their CFL substep counts spread, and on the device, so it is measured:
// scratch batches of the same disks are created with candidate tilings and
// timed.  Every tiling yields the same fields bit for bit, so the choice only
// changes speed.  The reference has no counterpart (one thread per cell,
// src/SourceEuler.cu:208-301).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "fgx_internal.h"

namespace fgx {
namespace {

uint64_t fnv1a(const std::string& s) {
  uint64_t h = 1469598103934665603ull;
  for (unsigned char c : s) { h ^= c; h *= 1099511628211ull; }
  return h;
}

// The workload key: precision, the device, the batch mode flags, and per disk
// (sorted, so the order of the disks does not matter) the grid and the
// parameters that set its CFL substep count and cost.
std::string signature(int n, const fgx_params* p, const fgx_planets* pl, int precision, std::string& text) {
  char dev[256] = "?";
  device_name(dev, sizeof dev);
  std::vector<std::string> disks;
  for (int k = 0; k < n; ++k) {
    char b[256];
    std::snprintf(b, sizeof b, "%dx%d h%.6g f%.6g a%.6g v%.6g s%.6g r%.6g-%.6g p%d", p[k].nrad, p[k].nsec,
                  (double)p[k].aspectratio, (double)p[k].flaringindex, (double)p[k].alphaviscosity,
                  (double)p[k].viscosity, (double)p[k].sigmaslope, (double)p[k].rmin, (double)p[k].rmax, pl[k].nb);
    disks.push_back(b);
  }
  std::sort(disks.begin(), disks.end());
  std::ostringstream o;
  o << (precision == FGX_F64 ? "f64" : "f32") << '|' << dev << "|ad" << p[0].adiabatic << " sg" << p[0].self_gravity
    << " ft" << p[0].fast_transport << '|';
  for (const std::string& d : disks) o << d << ';';
  text = o.str();
  char key[128];
  std::snprintf(key, sizeof key, "fgx-tile-%016llx-%dd", (unsigned long long)fnv1a(text), n);
  return key;
}

bool same(const fgx_tiling& a, const fgx_tiling& b) {
  return a.strip_rings == b.strip_rings && a.source_rings == b.source_rings &&
         a.fine_source_rings == b.fine_source_rings && a.fine_strip_rings == b.fine_strip_rings &&
         a.fine_cells == b.fine_cells;
}

// cell-updates per second of one scratch batch with tiling t over n_dt DTs
// after one warm-up DT; `used` receives the tiling as the engine resolved it,
// `secs` (if given) the timed wall time
int measure(int n, const fgx_params* p, const fgx_planets* pl, int precision, const fgx_tiling* t, int n_dt,
            fgx_tiling* used, double* rate, double* secs_out = nullptr) {
  fgx_batch* b = nullptr;
  if (int r = fgx_batch_create_tiled(n, p, pl, precision, t, &b)) return r;
  int r = fgx_batch_tiling(b, used);
  if (!r) r = fgx_batch_step(b, 1);
  std::vector<long long> s0(n, 0);
  for (int d = 0; !r && d < n; ++d) {
    fgx_disk_info in;
    r = fgx_batch_info(b, d, &in);
    s0[d] = in.substeps;
  }
  double secs = 0;
  if (!r) {
    const auto t0 = std::chrono::steady_clock::now();
    r = fgx_batch_step(b, n_dt);
    secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  double cu = 0;
  for (int d = 0; !r && d < n; ++d) {
    fgx_disk_info in;
    r = fgx_batch_info(b, d, &in);
    cu += (double)in.nrad * in.nsec * (double)(in.substeps - s0[d]);
  }
  fgx_batch_destroy(b);
  if (r) return r;
  *rate = secs > 0 ? cu / secs : 0;
  if (secs_out) *secs_out = secs;
  return FGX_OK;
}

// cache file: one record per line, "<key> strip source fine_source fine_strip
// fine_cells rate # workload text"
bool cache_find(const char* path, const std::string& key, fgx_tiling* t, double* rate) {
  if (!path) return false;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    std::string k;
    fgx_tiling v{};
    double r = 0;
    if (!(in >> k) || k != key) continue;
    if (in >> v.strip_rings >> v.source_rings >> v.fine_source_rings >> v.fine_strip_rings >> v.fine_cells >> r) {
      *t = v;
      *rate = r;
      return true;
    }
  }
  return false;
}

int cache_store(const char* path, const std::string& key, const fgx_tiling& t, double rate, const std::string& text) {
  if (!path) return FGX_OK;
  std::vector<std::string> keep;
  {
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
      std::istringstream in(line);
      std::string k;
      if ((in >> k) && k == key) continue;
      if (!line.empty()) keep.push_back(line);
    }
  }
  char rec[512];
  std::snprintf(rec, sizeof rec, "%s %d %d %d %d %.17g %.6e", key.c_str(), t.strip_rings, t.source_rings,
                t.fine_source_rings, t.fine_strip_rings, t.fine_cells, rate);
  keep.push_back(std::string(rec) + " # " + text.substr(0, 3000));
  const std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream f(tmp, std::ios::trunc);
    if (!f) return set_error(FGX_EIO, "cannot write the tiling cache %s", tmp.c_str());
    for (const std::string& l : keep) f << l << '\n';
    if (!f) return set_error(FGX_EIO, "cannot write the tiling cache %s", tmp.c_str());
  }
  if (std::rename(tmp.c_str(), path) != 0) return set_error(FGX_EIO, "cannot replace the tiling cache %s", path);
  return FGX_OK;
}

}  // namespace
}  // namespace fgx

using fgx::set_error;

extern "C" int fgx_autotune(int n, const fgx_params* p, const fgx_planets* pl, int precision, int n_dt,
                            const char* cache_path, fgx_tiling* out, fgx_tune_report* report) {
  if (n < 1 || !p || !pl || !out) return set_error(FGX_EPARAM, "bad autotune arguments");
  if (n_dt < 1) return set_error(FGX_EPARAM, "the autotuner needs at least one timed DT");
  const auto t0 = std::chrono::steady_clock::now();
  fgx_tune_report rep;
  std::memset(&rep, 0, sizeof rep);
  std::string text;
  const std::string key = fgx::signature(n, p, pl, precision, text);
  std::snprintf(rep.signature, sizeof rep.signature, "%s", key.c_str());
  fgx_tiling best{};
  double best_rate = 0;
  if (fgx::cache_find(cache_path, key, &best, &best_rate)) {
    rep.from_cache = 1;
    rep.best_rate = best_rate;
    *out = best;
    if (report) *report = rep;
    return FGX_OK;
  }
  // the engine's own choice first: it is the incumbent.  A workload whose
  // DTs are short is timed over more of them, at least about 40 ms per
  // candidate, so that the 1% rule below is not decided by timer noise
  const fgx_tiling none{};
  double secs = 0;
  if (int r = fgx::measure(n, p, pl, precision, &none, n_dt, &best, &best_rate, &secs)) return r;
  rep.candidates = 1;
  if (secs < 0.04) {
    n_dt = std::min(64, (int)std::ceil(n_dt * 0.04 / std::max(secs, 1e-4)));
    if (int r = fgx::measure(n, p, pl, precision, &none, n_dt, &best, &best_rate)) return r;
    rep.candidates++;
  }
  const fgx_tiling dflt = best;
  rep.default_rate = best_rate;
  std::vector<fgx_tiling> tried{best};
  // coordinate search: each knob in turn over its candidate values, the
  // others held at the incumbent; a candidate must beat it by more than 1%
  // (the scatter of repeated timings) to replace it
  const int src_v[] = {6, 8, 12, 16, 20, 24};
  const int strip_v[] = {4, 8, 16};
  const double fine_v[] = {-1.0, 0.5e6, 1e6, 2e6, 4e6};
  const int fsrc_v[] = {2, 4, 8};
  const int fstrip_v[] = {2, 4};
  auto try_one = [&](fgx_tiling c) -> int {
    for (const fgx_tiling& t : tried)
      if (fgx::same(t, c)) return FGX_OK;
    fgx_tiling used{};
    double rate = 0;
    if (int r = fgx::measure(n, p, pl, precision, &c, n_dt, &used, &rate)) return r;
    rep.candidates++;
    tried.push_back(c);
    if (!fgx::same(used, c)) tried.push_back(used);   // clamped to an equivalent tiling
    if (rate > best_rate * 1.01) { best = used; best_rate = rate; }
    return FGX_OK;
  };
  for (int v : src_v) { fgx_tiling c = best; c.source_rings = v; c.fine_source_rings = std::min(c.fine_source_rings, v); if (int r = try_one(c)) return r; }
  for (int v : strip_v) { fgx_tiling c = best; c.strip_rings = v; c.fine_strip_rings = std::min(c.fine_strip_rings, v); if (int r = try_one(c)) return r; }
  for (double v : fine_v) { fgx_tiling c = best; c.fine_cells = v; if (int r = try_one(c)) return r; }
  for (int v : fsrc_v) {
    if (v > best.source_rings) continue;
    fgx_tiling c = best; c.fine_source_rings = v; if (int r = try_one(c)) return r;
  }
  for (int v : fstrip_v) {
    if (v > best.strip_rings) continue;
    fgx_tiling c = best; c.fine_strip_rings = v; if (int r = try_one(c)) return r;
  }
  // a winner other than the engine's own choice is confirmed against it:
  // two more timings of each, alternating, and it stays only if its best
  // beats the default's best by more than 1%
  if (!fgx::same(best, dflt)) {
    double rb = best_rate, rd = rep.default_rate;
    for (int k = 0; k < 2; ++k) {
      fgx_tiling used{};
      double r1 = 0, r2 = 0;
      if (int r = fgx::measure(n, p, pl, precision, &dflt, n_dt, &used, &r1)) return r;
      if (int r = fgx::measure(n, p, pl, precision, &best, n_dt, &used, &r2)) return r;
      rd = std::max(rd, r1);
      rb = std::max(rb, r2);
      rep.candidates += 2;
    }
    rep.default_rate = rd;
    if (rb > rd * 1.01) {
      best_rate = rb;
    } else {
      best = dflt;
      best_rate = rd;
    }
  }
  rep.best_rate = best_rate;
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (int r = fgx::cache_store(cache_path, key, best, best_rate, text)) return r;
  *out = best;
  if (report) *report = rep;
  return FGX_OK;
}
