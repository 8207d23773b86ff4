// engine.hip -- device state of a batch of disks and the AlgoGas driver.
//
// A batch holds N >= 1 disks of possibly different sizes in one set of
// device arrays (one array per field, disks back to back).  One call of
// AlgoGas (SourceEuler.cu:208-301) for the batch is a loop of substeps; each
// substep is a fixed sequence of kernel launches that each cover every disk
// still inside its DT.  dt, the CFL count, planets and time live on the
// device, so the host only learns the substep count of the batch: it reads a
// two-int summary every few substeps instead of synchronising per kernel as
// the reference does.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <cstdint>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fgx_internal.h"
#include "kernels.hpp"
#include "transport.hpp"
#include "sg.hpp"
#include "source.hpp"

// Device scratch of fgx_disk_force_noalloc: the per-ring partials of the
// force kernel (44 doubles per ring, up to FGX_NOALLOC_FORCE_RINGS rings) and
// the 44 totals after them.  Each GPU the library runs on holds its own copy
// until the process ends.  In host code this name is only a stand-in: the
// device address comes from hipGetSymbolAddress (fgx::noalloc_scratch).
#define FGX_NOALLOC_FORCE_RINGS 4096
#define FGX_NOALLOC_SCRATCH (FGX_NOALLOC_FORCE_RINGS * 44 + 64)
static __device__ __align__(256) double scratch[FGX_NOALLOC_SCRATCH];

namespace fgx {

// the device address of `scratch` on device `dev` (which must be current),
// looked up once per device and kept
#define FGX_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[FGX_MAX_DEVICES] = {};
static int noalloc_scratch(int dev, double** out) {
  if (dev < 0 || dev >= FGX_MAX_DEVICES) return set_error(FGX_EDEVICE, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(FGX_EDEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return FGX_OK;
}

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return set_error(FGX_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                                 \
  } while (0)

// A kernel's dynamic-LDS limit (hipFuncAttributeMaxDynamicSharedMemorySize)
// is process-wide, and batches of different grids may be created at the same
// time (the serving loop's workers): it is only ever raised, under a lock, so
// that a batch setting a smaller limit cannot undercut one that needs more.
static std::mutex g_lds_attr_mu;
static int raise_dynamic_lds(const void* f, size_t bytes) {
  static std::unordered_map<const void*, size_t> limit;
  std::lock_guard<std::mutex> lk(g_lds_attr_mu);
  size_t& cur = limit[f];
  if (bytes <= cur || bytes <= 48 * 1024) return FGX_OK;
  HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  cur = bytes;
  return FGX_OK;
}

struct BatchBase {
  virtual ~BatchBase() {}
  virtual int step(int n_dt) = 0;
  virtual int read_field(int disk, int field, int dtype, void* host, size_t count) = 0;
  virtual int write_field(int disk, int field, int dtype, const void* host, size_t count) = 0;
  virtual int read_ring(int disk, int ring, int dtype, void* host, size_t count) = 0;
  virtual int info(int disk, fgx_disk_info* out) = 0;
  virtual int planets(int disk, fgx_planets* out) = 0;
  virtual int set_planets(int disk, const fgx_planets* in) = 0;
  virtual int force(int disk, double x, double y, double sm, double m, double* out44) = 0;
  virtual int force_noalloc(int disk, double x, double y, double sm, double m, double* out44) = 0;
  int device = 0;   // the HIP device the batch lives on
  virtual int mdcp(int disk, double* out) = 0;
  virtual int profile_enable(int on) = 0;
  virtual int profile_get(int k, const char** name, double* ms, long long* n, double* bytes) = 0;
  virtual int profile_reset() = 0;
  virtual const HostDisk& host(int d) const = 0;
  virtual int size() const = 0;
  virtual int freeze(int disk, int on) = 0;
  virtual int checkpoint(int disk, const char* path) = 0;
  virtual int restore(int disk, const char* path) = 0;
  virtual int reset_disk(int disk, const fgx_params& p, const fgx_planets& pl) = 0;
  virtual int snapshot_queue(int disk) = 0;
  virtual int snapshot_wait(int disk, int* had) = 0;
  virtual int tiling(fgx_tiling* out) const = 0;
};

// checkpoint file: header, per-disk state record, ring profiles carried from
// substep to substep, fields in the engine's representation, FNV-1a 64 sum
struct CkptHeader {
  char magic[8];             // "FGXCKPT3"
  int32_t real_bytes, nrad, nsec, adiabatic, nfields, nrings;
  uint32_t state_bytes, pad;
};
inline uint64_t fnv1a64(const char* p, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t k = 0; k < n; ++k) { h ^= (unsigned char)p[k]; h *= 1099511628211ull; }
  return h;
}

template <class real>
struct Batch : BatchBase {
  int nd = 0;
  std::vector<HostDisk> hd;
  std::vector<DiskConst> dch;
  hipStream_t stream = nullptr;
  Dev<real> D{};
  std::vector<void*> allocs;
  int n_ctiles = 0, n_rtiles = 0, max_nsec = 0, max_nrad = 0;
  // radial strips (d, first ring, sector block) and the azimuthal ring lists:
  // one per lane width K of the register-resident kernel, plus the rings
  // whose sector count fits none (shared-memory kernel)
  int n_stiles = 0, strip_rings = 16, n_sgctiles = 0;
  // fused source pass (k_source): 57-sector chunks of src_rings rings
  int n_srctiles = 0, src_rings = 12;
  bool fused_source = true;
  const CellTile* srctiles = nullptr;
  // two tilings of the strip kernels: the coarse one above for substeps in
  // which most of the batch is active, and a fine one (short strips, many
  // more blocks) for the latency-bound substeps in which only a few disks
  // still run; the host picks one per substep from the summary's forecast of
  // active cells.  Both give the same fields bit for bit.
  struct TileSet {
    const CellTile *st = nullptr, *src = nullptr;
    int n_st = 0, strip_rings = 0, n_src = 0, src_rings = 0;
  } coarse, fine;
  // fine tiling below this many active cells: coarse strips cannot fill the
  // chip with less (256 CUs x 4 SIMDs x a few waves of 64 x 12-16 rings)
  double fine_cells = 2.0e6;
  long long total_cells = 0;
  std::vector<char> fine_plan;   // tiling of the upcoming substeps
  int plan_pos = 0;
  void use_tiles(bool f) {
    const TileSet& t = f ? fine : coarse;
    stiles = t.st; n_stiles = t.n_st; strip_rings = t.strip_rings;
    srctiles = t.src; n_srctiles = t.n_src; src_rings = t.src_rings;
    D.stiles = t.st; D.strip_rings = t.strip_rings;
  }
  size_t sg_shm_row = 0, sg_shm_col = 0;
  const CellTile* stiles = nullptr;
  struct ThetaGroup { int k, w, n; const RingTile* tiles; };
  std::vector<ThetaGroup> theta_groups;
  ThetaMulti theta_multi{};   // the multi-arc groups, launched together
  // Tile lists narrowed to the disks still active in a DT.  A disk only
  // drops out within a DT, so the disks active at a summary cover every later
  // substep of the DT; the blocks of finished disks are then not launched at
  // all, instead of each loading its tile and its disk's state to exit.
  // Lists: 0 coarse strips, 1 coarse source, 2 fine strips, 3 fine source,
  // 4 rings, then the multi-arc ring groups and the other ring groups.
  struct TileList {
    const int* full = nullptr;   // the whole list (device, padded)
    int* part = nullptr;         // its narrowed copy (device)
    int elem = 0;                // ints per tile
    int n_full = 0;
    std::vector<int> off, cnt;   // per disk: its range in the whole list
    int n_part = 0;
  };
  std::vector<TileList> tlists;
  bool compact_on = true;        // FGX_COMPACT=0 launches every disk's tiles
  int covered = 0;               // disks the current lists cover
  long long covered_cells = 0;
  TileSet coarse_full, fine_full;
  ThetaMulti theta_multi_full{};
  std::vector<ThetaGroup> theta_groups_full;
  int theta_multi_n_full = 0, n_rtiles_full = 0;
  const RingTile* rtiles_full = nullptr;
  CompactRange* h_ranges = nullptr;   // pinned
  CompactRange* d_ranges = nullptr;
  int max_ranges = 0;
  template <class T>
  static void tile_ranges(const std::vector<T>& v, int nd, TileList& L) {
    L.off.assign(nd, 0); L.cnt.assign(nd, 0);
    for (size_t k = 0; k < v.size(); ++k) {
      const int d = v[k].d;
      if (d < 0) continue;
      if (L.cnt[d] == 0) L.off[d] = (int)k;
      L.cnt[d]++;
    }
  }
  template <class T>
  int add_list(const std::vector<T>& v, const T* dev, int& idx) {
    TileList L;
    L.full = reinterpret_cast<const int*>(dev);
    L.elem = (int)(sizeof(T) / sizeof(int));
    L.n_full = (int)v.size();
    tile_ranges(v, nd, L);
    T* part;
    if (int r = dalloc(&part, v.size())) return r;
    L.part = reinterpret_cast<int*>(part);
    idx = (int)tlists.size();
    tlists.push_back(std::move(L));
    return FGX_OK;
  }
  int list_coarse_st = -1, list_coarse_src = -1, list_fine_st = -1, list_fine_src = -1, list_rings = -1;
  int list_sgcols = -1, n_sgctiles_full = 0;
  const RingTile* sg_ctiles_full = nullptr;
  std::vector<char> frozen_h;   // disks frozen by the host (fgx_batch_freeze)
  std::vector<int> list_multi, list_groups;
  // the whole lists again (every DT starts and ends with them)
  void lists_full() {
    coarse = coarse_full; fine = fine_full;
    theta_multi = theta_multi_full; theta_multi_n = theta_multi_n_full; theta_groups = theta_groups_full;
    D.rtiles = rtiles_full; n_rtiles = n_rtiles_full;
    if (list_sgcols >= 0) { D.sg_ctiles = sg_ctiles_full; n_sgctiles = n_sgctiles_full; }
    covered = nd; covered_cells = total_cells;
  }
  // narrow every list to the disks with act[d] != 0
  int lists_narrow(const int* act) {
    int nr = 0;
    auto range = [&](const int* src, int* dst, int len, int e) {
      h_ranges[nr++] = CompactRange{src, dst, len, e};
    };
    for (TileList& L : tlists) {
      int n = 0;
      for (int d = 0; d < nd; ++d) {
        if (!act[d] || L.cnt[d] == 0) continue;
        range(L.full + (size_t)L.off[d] * L.elem, L.part + (size_t)n * L.elem, L.cnt[d], L.elem);
        n += L.cnt[d];
      }
      const int np = (n + FGX_TILE_PAD - 1) / FGX_TILE_PAD * FGX_TILE_PAD;
      if (np > n) range(nullptr, L.part + (size_t)n * L.elem, np - n, L.elem);
      L.n_part = np;
    }
    HIPCHK(hipMemcpyAsync(d_ranges, h_ranges, nr * sizeof(CompactRange), hipMemcpyHostToDevice, stream));
    if (int r = launch("compact", 0, k_compact, nr, FGX_TPB, 0, (const CompactRange*)d_ranges)) return r;
    // hipMemcpyAsync reads the pinned table when the copy runs; the table is
    // only refilled after a later summary has been waited for, which the
    // stream orders after this copy
    auto part = [&](int i) { return std::make_pair((const CellTile*)tlists[i].part, tlists[i].n_part); };
    coarse.st = part(list_coarse_st).first; coarse.n_st = part(list_coarse_st).second;
    coarse.src = part(list_coarse_src).first; coarse.n_src = part(list_coarse_src).second;
    fine.st = part(list_fine_st).first; fine.n_st = part(list_fine_st).second;
    fine.src = part(list_fine_src).first; fine.n_src = part(list_fine_src).second;
    D.rtiles = (const RingTile*)tlists[list_rings].part; n_rtiles = tlists[list_rings].n_part;
    if (list_sgcols >= 0) {
      D.sg_ctiles = (const RingTile*)tlists[list_sgcols].part;
      n_sgctiles = tlists[list_sgcols].n_part;
    }
    ThetaMulti M{};
    int first = 0;
    for (int g = 0; g < theta_multi_full.ng; ++g) {
      const TileList& L = tlists[list_multi[g]];
      if (L.n_part == 0) continue;
      M.k[M.ng] = theta_multi_full.k[g]; M.w[M.ng] = theta_multi_full.w[g];
      M.tiles[M.ng] = (const RingTile*)L.part;
      M.first[M.ng] = first;
      first += L.n_part;
      M.first[++M.ng] = first;
    }
    theta_multi = M; theta_multi_n = first;
    for (size_t g = 0; g < theta_groups_full.size(); ++g) {
      const TileList& L = tlists[list_groups[g]];
      theta_groups[g] = ThetaGroup{theta_groups_full[g].k, theta_groups_full[g].w, L.n_part, (const RingTile*)L.part};
    }
    covered = 0; covered_cells = 0;
    for (int d = 0; d < nd; ++d)
      if (act[d]) { covered++; covered_cells += (long long)hd[d].nrad * hd[d].nsec; }
    return FGX_OK;
  }
  bool theta_arcs = true;     // FGX_THETA_ARCS=0: one wave per ring wherever it fits
  int theta_multi_n = 0, theta_multi_block = 0;
  size_t F = 0, R = 0, S = 0;
  int* h_summary = nullptr;  // pinned
  int visc_dynamic = 0, alpha = 0, need_force = 1;
  // Launches of k_visc_profile that would rewrite what D.csmean and D.visc
  // already hold are skipped.  Without LambdaDoubling the profile is a
  // function of the ring means of the sound speed alone, and the sound speed
  // changes only where k_sound_speed runs.  visc_fresh: D.visc is the profile
  // of D.csmean; cs_dirty: D.csmean may differ from the ring means of D.cs;
  // visc_time: some disk's profile depends on time.  Every step() call starts
  // them over, so changes made between calls are always seen.
  bool visc_fresh = false, cs_dirty = true, visc_time = true;
  bool visc_skip = true;   // FGX_VISC_SKIP=0 launches every one
  bool accretion = false;       // some disk's planets accrete (StubbedPhysics)
  bool outer_source = false;    // some disk has the outer mass source (StubbedPhysics)
  size_t theta_shmem = 0;
  std::vector<std::vector<double>> vref_h;  // reference vtheta profile per disk (nrad+1)
  std::vector<int> health_h;                // FGX_HEALTH_* flags of each disk known to the host
  // profiling
  struct Prof { std::string name; double ms = 0; long long n = 0; double bytes = 0; };
  bool prof_on = false;
  std::vector<Prof> prof;
  struct Pending { int k; hipEvent_t a, b; double bytes; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> free_events;

  // asynchronous snapshots: per disk a pinned buffer of its stored fields and
  // the event that marks the copy's end
  std::vector<void*> snap_host;
  std::vector<hipEvent_t> snap_ev;
  std::vector<char> snap_pending;

  ~Batch() override {
    if (stream) hipStreamSynchronize(stream);
    graphs_clear();
    for (void* p : snap_host) if (p) hipHostFree(p);
    for (hipEvent_t e : snap_ev) if (e) hipEventDestroy(e);
    for (auto& p : pending) { hipEventDestroy(p.a); hipEventDestroy(p.b); }
    for (auto e : free_events) hipEventDestroy(e);
    for (void* p : allocs) hipFree(p);
    if (h_summary) hipHostFree(h_summary);
    if (h_ranges) hipHostFree(h_ranges);
    if (stream) hipStreamDestroy(stream);
  }
  int size() const override { return nd; }
  const HostDisk& host(int d) const override { return hd[d]; }

  // the arrays that hold a share per disk, by how that share is laid out:
  // 0 field ((nrad+1)*nsec at fo), 1 ring (nrad+1 at ro), 2 sector (nsec at
  // so), 3 self-gravity work (nrad*nsec complex at sgw_o), 4 force partials
  // (2*FGX_NFORCE_T per ring tile from rt0)
  struct Region { void* p; size_t elem; int kind; };
  std::vector<Region> regions;
  template <class T>
  int dalloc_share(T** p, size_t n, int kind) {
    if (int r = dalloc(p, n)) return r;
    regions.push_back(Region{(void*)*p, sizeof(T), kind});
    return FGX_OK;
  }
  // zero disk d's share of every per-disk array, as a fresh batch holds it
  int clear_share(int d) {
    const DiskConst& c = dch[d];
    for (const Region& g : regions) {
      size_t off = 0, n = 0;
      switch (g.kind) {
        case 0: off = (size_t)c.fo; n = (size_t)(c.nrad + 1) * c.nsec; break;
        case 1: off = (size_t)c.ro; n = (size_t)c.nrad + 1; break;
        case 2: off = (size_t)c.so; n = (size_t)c.nsec; break;
        case 3: off = (size_t)c.sgw_o; n = (size_t)c.nrad * c.nsec; break;
        default: off = (size_t)c.rt0 * 2 * FGX_NFORCE_T; n = (size_t)c.nrad * 2 * FGX_NFORCE_T; break;
      }
      HIPCHK(hipMemsetAsync((char*)g.p + off * g.elem, 0, n * g.elem, stream));
    }
    return FGX_OK;
  }
  template <class T>
  int dalloc(T** p, size_t n) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return set_error(FGX_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    allocs.push_back(q);
    HIPCHK(hipMemsetAsync(q, 0, bytes, stream));
    *p = (T*)q;
    return FGX_OK;
  }

  // ------------------------------------------------------------- profiling
  int prof_index(const char* name) {
    for (size_t k = 0; k < prof.size(); ++k)
      if (prof[k].name == name) return (int)k;
    prof.push_back(Prof{name});
    return (int)prof.size() - 1;
  }
  hipEvent_t get_event() {
    if (!free_events.empty()) { hipEvent_t e = free_events.back(); free_events.pop_back(); return e; }
    hipEvent_t e; hipEventCreate(&e); return e;
  }
  void prof_collect() {
    for (auto& p : pending) {
      float ms = 0;
      hipEventSynchronize(p.b);
      hipEventElapsedTime(&ms, p.a, p.b);
      prof[p.k].ms += ms; prof[p.k].n += 1; prof[p.k].bytes += p.bytes;
      free_events.push_back(p.a); free_events.push_back(p.b);
    }
    pending.clear();
  }

  // algorithmic bytes of one launch: bytes per cell of the kernel's distinct
  // full-field reads and writes, times the cells of the batch
  double cells() const {
    double c = 0;
    for (auto& h : hd) c += (double)h.nrad * h.nsec;
    return c;
  }
  template <class K, class... A>
  int launch(const char* name, double bytes_per_cell, K kern, int grid, int block, size_t shmem, A... args) {
    if (grid <= 0) return FGX_OK;
    if (launch_mode == 1) {
      // recording: the launch's kernel, shape and argument bytes go into the
      // substep's key, nothing is launched
      const void* kp = (const void*)kern;
      rec_mix(&kp, sizeof kp);
      rec_mix(&grid, sizeof grid);
      rec_mix(&block, sizeof block);
      rec_mix(&shmem, sizeof shmem);
      (rec_mix(&args, sizeof(args)), ...);
      return FGX_OK;
    }
    Pending pd{};
    const bool timed = prof_on && !in_group;
    if (timed) {
      // the events ride on the kernel's own dispatch, so they time its
      // execution alone, as a kernel trace does, not the launch gap too
      pd.k = prof_index(name); pd.a = get_event(); pd.b = get_event();
      pd.bytes = bytes_per_cell * sizeof(real) * cells();
      hipExtLaunchKernelGGL(kern, dim3(grid), dim3(block), shmem, stream, pd.a, pd.b, 0, args...);
    } else if (prof_on && in_group) {
      // a group spans from its first kernel's start to its last kernel's end
      hipExtLaunchKernelGGL(kern, dim3(grid), dim3(block), shmem, stream, grp_started ? nullptr : grp.a, grp.b, 0,
                            args...);
      grp_started = true;
    } else {
      hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shmem, stream, args...);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(FGX_EDEVICE, "launch of %s failed: %s", name, hipGetErrorString(e));
    if (timed) pending.push_back(pd);
    return FGX_OK;
  }
  // ---------------------------------------------------- substep graphs
  // With graphs on, a substep's launches run as one instantiated hipGraph.
  // The substep is first run with launches only recorded: its kernels,
  // shapes and argument bytes key a cache of graphs, and the host-side state
  // (velocity buffer parity, tiling, the plan position) advances exactly as
  // the launches would have advanced it.  On a miss the same substep is
  // captured from the same starting state and instantiated.  Two substeps
  // with the same key launch the same kernels with the same arguments, so
  // replaying the cached graph is the same work; the tile lists and disk
  // state the kernels read are device data, read when the graph runs.
  bool graphs_on = false;
  int launch_mode = 0;   // 0 launch, 1 record the key only
  uint64_t rec_hash = 0;
  std::unordered_map<uint64_t, hipGraphExec_t> graph_cache;
  long long graph_hits = 0, graph_builds = 0;
  void rec_mix(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t k = 0; k < n; ++k) { rec_hash ^= b[k]; rec_hash *= 1099511628211ull; }
  }
  void graphs_clear() {
    for (auto& kv : graph_cache) hipGraphExecDestroy(kv.second);
    graph_cache.clear();
  }
  struct HostTiles {
    const CellTile *st, *src;
    int n_st, strip_rings, n_src, src_rings;
  };
  HostTiles tiles_save() const { return HostTiles{stiles, srctiles, n_stiles, strip_rings, n_srctiles, src_rings}; }
  void tiles_restore(const HostTiles& t) {
    stiles = t.st; srctiles = t.src; n_stiles = t.n_st; strip_rings = t.strip_rings;
    n_srctiles = t.n_src; src_rings = t.src_rings;
  }
  int substep_graphed() {
    const Dev<real> D0 = D;
    const int pp0 = plan_pos;
    const bool vf0 = visc_fresh, cd0 = cs_dirty;
    const HostTiles t0 = tiles_save();
    launch_mode = 1;
    rec_hash = 1469598103934665603ull;
    int r = substep();
    launch_mode = 0;
    if (r) return r;
    auto it = graph_cache.find(rec_hash);
    if (it == graph_cache.end()) {
      const Dev<real> D1 = D;
      const int pp1 = plan_pos;
      const bool vf1 = visc_fresh, cd1 = cs_dirty;
      const HostTiles t1 = tiles_save();
      D = D0; plan_pos = pp0; visc_fresh = vf0; cs_dirty = cd0; tiles_restore(t0);
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      r = substep();
      const hipError_t e = hipStreamEndCapture(stream, &g);
      if (r) { if (g) hipGraphDestroy(g); return r; }
      if (e != hipSuccess) return set_error(FGX_EDEVICE, "substep capture failed: %s", hipGetErrorString(e));
      hipGraphExec_t x = nullptr;
      const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (ei != hipSuccess) return set_error(FGX_EDEVICE, "substep graph instantiation failed: %s", hipGetErrorString(ei));
      if (graph_cache.size() >= 512) graphs_clear();
      it = graph_cache.emplace(rec_hash, x).first;
      ++graph_builds;
      // the capture advanced the host state as the recording did
      if (plan_pos != pp1 || visc_fresh != vf1 || cs_dirty != cd1 || std::memcmp(&D, &D1, sizeof D) != 0 ||
          tiles_save().st != t1.st)
        return set_error(FGX_EDEVICE, "substep capture changed the host state differently from its recording");
    } else {
      ++graph_hits;
    }
    HIPCHK(hipGraphLaunch(it->second, stream));
    return FGX_OK;
  }

  // several launches timed as one kernel of the profile
  bool in_group = false, grp_started = false;
  Pending grp{};
  void group_begin(const char* name, double bytes_per_cell) {
    if (prof_on) grp = Pending{prof_index(name), get_event(), get_event(), bytes_per_cell * sizeof(real) * cells()};
    in_group = true;
    grp_started = false;
  }
  void group_end() {
    if (prof_on && grp_started) pending.push_back(grp);
    in_group = false;
  }
#define LCHK(x) do { int r_ = (x); if (r_) return r_; } while (0)

  // the constants of a disk that follow from its parameters and set-up (the
  // layout offsets are the batch's)
  static void fill_const(const HostDisk& h, DiskConst& c) {
    const fgx_params& p = h.P;
    c.accretes = 0;
    c.nb = h.sys.nb;
    for (int k = 0; k < c.nb; ++k) {
      c.feel_disk[k] = h.sys.feel_disk[k]; c.feel_others[k] = h.sys.feel_others[k]; c.acc[k] = h.sys.acc[k];
      if (p.ext_stubbed_physics && h.sys.acc[k] > 1e-10) c.accretes = 1;
    }
    c.outer_source = (p.ext_stubbed_physics && p.outer_source_mass) ? 1 : 0;
    c.outer_vr = p.imposeddiskdrift * std::pow(h.rinf[h.nrad - 1], -(double)p.sigmaslope);
    {
      const double r0 = h.rmed[0], rn = h.rmed[h.nrad - 1];
      c.pow_in_2f = std::pow(r0, 2.0 * (double)p.flaringindex);
      c.pow_out_2f = std::pow(rn, 2.0 * (double)p.flaringindex);
      c.ev_tin = 2.0 * kPI * std::pow(r0, 1.5);
      c.ev_tout = 2.0 * kPI * std::pow(rn, 1.5);
      c.ev_drmin = r0 * 1.25;
      c.ev_drmax = rn * 0.84;
    }
    c.DT = p.dt; c.masstaper = p.masstaper; c.sigmaslope = p.sigmaslope; c.flaringindex = p.flaringindex;
    c.aspectratio = p.aspectratio; c.adiabaticindex = p.adiabaticindex;
    c.thicknesssmoothing = p.thicknesssmoothing; c.rochesmoothing = p.rochesmoothing;
    c.releasedate = p.releasedate; c.releaseradius = p.releaseradius;
    c.transitionradius = p.transitionradius; c.transitionratio = p.transitionratio;
    c.transitionwidth = p.transitionwidth; c.lambdadoubling = p.lambdadoubling;
    c.cavityradius = p.cavityradius; c.cavityratio = p.cavityratio; c.cavitywidth = p.cavitywidth;
    c.viscosity = p.viscosity; c.alphaviscosity = p.alphaviscosity;
    c.hill_radius = h.hill_radius; c.sigma0 = p.sigma0; c.eccentricity = p.eccentricity;
    c.omega0 = h.omega_frame;
    c.substeps_forced = p.ext_substeps;
    c.max_substeps = p.ext_max_substeps > 0 ? p.ext_max_substeps : 0x7fffffff;
    // the viscosity profile changes only through the ring-mean sound speed
    // (adiabatic alpha viscosity) or the time-dependent transition scale
    c.visc_dynamic = ((p.viscosity_alpha && p.adiabatic) || p.lambdadoubling != 0.0f) ? 1 : 0;
    if (p.self_gravity) { c.sg_eps = h.sg_eps; c.sg_rstep = h.sg_rstep; c.sg_tstep = h.sg_tstep; }
  }

  // ------------------------------------------------------------------ create
  // the tiling the batch was created with (every field resolved)
  fgx_tiling tiling_used{};
  int tiling(fgx_tiling* out) const override { *out = tiling_used; return FGX_OK; }
  int create(const std::vector<fgx_params>& ps, const std::vector<fgx_planets>& pls, const fgx_tiling* tl) {
    nd = (int)ps.size();
    if (nd < 1) return set_error(FGX_EPARAM, "empty batch");
    const fgx_params& P0 = ps[0];
    for (int d = 0; d < nd; ++d) {
      const fgx_params& p = ps[d];
      LCHK(params_validate(&p));
      if (!batch_modes_equal(p, P0))
        return set_error(FGX_EPARAM, "disk %d of the batch differs from disk 0 in a mode flag", d);
      if (p.corotating && pls[d].nb == 0) return set_error(FGX_EPARAM, "a corotating frame needs a planet");
      if (!p.is_disk) return set_error(FGX_EPARAM, "DISK NO (planets only) is not supported");
    }
    HIPCHK(hipGetDevice(&device));
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipHostMalloc((void**)&h_summary, (2 + FGX_SUMMARY_HIST + nd) * sizeof(int)));
    hd.resize(nd);
    // a disk whose parameters and planets repeat an earlier disk's byte for
    // byte (the serving loop's spare slots) takes that disk's host set-up
    std::vector<int> same(nd, -1);
    for (int d = 1; d < nd; ++d)
      for (int e = 0; e < d && same[d] < 0; ++e)
        if (same[e] < 0 && !std::memcmp(&ps[d], &ps[e], sizeof(fgx_params)) &&
            !std::memcmp(&pls[d], &pls[e], sizeof(fgx_planets)))
          same[d] = e;
    for (int d = 0; d < nd; ++d)
      if (same[d] < 0) LCHK(setup_grid_and_density(ps[d], pls[d], hd[d]));
    // self-gravity set-up (FP64, host): kernel transforms, the initial radial
    // SG profile for the planets' and the gas's circular velocities
    std::vector<std::vector<double>> khr(nd), kht(nd);
    for (int d = 0; d < nd; ++d) {
      if (same[d] >= 0) continue;
      std::vector<double> axisg;
      if (hd[d].P.self_gravity) LCHK(sg_host_setup(hd[d], khr[d], kht[d], axisg));
      setup_velocities(hd[d], axisg);
    }
    for (int d = 0; d < nd; ++d)
      if (same[d] >= 0) { hd[d] = hd[same[d]]; khr[d] = khr[same[d]]; kht[d] = kht[same[d]]; }
    // layout
    dch.resize(nd);
    size_t SGW = 0, SGK = 0, STW = 0;
    std::vector<RingTile> sgct;
    std::vector<CellTile> ct;
    std::vector<RingTile> rt;
    for (int d = 0; d < nd; ++d) {
      const HostDisk& h = hd[d];
      DiskConst& c = dch[d];
      std::memset(&c, 0, sizeof c);
      c.nrad = h.nrad; c.nsec = h.nsec;
      F = (F + 63) & ~(size_t)63;   // every disk starts on a 256-byte boundary (vector loads)
      c.fo = (long long)F; c.ro = (int)R; c.so = (int)S;
      c.rt0 = (int)rt.size();
      F += (size_t)(h.nrad + 1) * h.nsec;
      R += (size_t)h.nrad + 1;
      S += (size_t)h.nsec;
      max_nsec = std::max(max_nsec, h.nsec);
      max_nrad = std::max(max_nrad, h.nrad);
      for (int i = 0; i < h.nrad; ++i) {
        rt.push_back(RingTile{d, i});
        for (int jb = 0; jb * FGX_TPB < h.nsec; ++jb) ct.push_back(CellTile{d, i, jb, 0});
      }
      fill_const(h, c);
      accretion |= c.accretes != 0;
      outer_source |= c.outer_source != 0;
      visc_dynamic |= c.visc_dynamic;
      const fgx_params& p = h.P;
      if (p.self_gravity) {
        c.sgw_o = (long long)SGW; c.sgk_o = (long long)SGK;
        c.tw_ns = (int)STW; c.tw_nr = (int)(STW + h.nsec);
        sg_factor(h.nsec, c.fac_ns, &c.nf_ns);
        sg_factor(2 * h.nrad, c.fac_nr, &c.nf_nr);
        SGW += (size_t)h.nrad * h.nsec;
        SGK += (size_t)2 * h.nrad * h.nsec;
        STW += (size_t)h.nsec + 2 * h.nrad;
        for (int k = 0; 2 * k <= h.nsec; ++k) sgct.push_back(RingTile{d, k});   // column pairs (k, nsec-k)
      }
    }
    alpha = P0.viscosity_alpha;
    // strip height: as tall as possible while the batch still yields about
    // 2048 strip tiles (32 waves per CU), between 2 and 16 rings
    {
      double tiles16 = 0;
      for (int d = 0; d < nd; ++d)
        tiles16 += std::ceil(hd[d].nrad / 16.0) * std::ceil(hd[d].nsec / (double)FGX_TPB);
      strip_rings = 16;
      while (strip_rings > 2 && tiles16 * 16.0 / strip_rings < 2048) strip_rings /= 2;
      if (tl && tl->strip_rings > 0) strip_rings = tl->strip_rings;
      else if (const char* e = std::getenv("FGX_STRIP_RINGS")) strip_rings = std::atoi(e);
      strip_rings = std::min(std::max(strip_rings, 2), 64);
    }
    std::vector<CellTile> st;
    std::vector<std::vector<RingTile>> tg(17 * 17);   // index K*17 + W (K = 0: shared-memory kernel)
    if (const char* e = std::getenv("FGX_THETA_ARCS")) theta_arcs = std::atoi(e) != 0;
    for (int d = 0; d < nd; ++d) {
      const HostDisk& h = hd[d];
      for (int i0 = 0; i0 < h.nrad; i0 += strip_rings)
        for (int jb = 0; jb * FGX_TPB < h.nsec; ++jb) st.push_back(CellTile{d, i0, jb, 0});
      int K = 0, W = 0;
      theta_split<real>(h.nsec, K, W, theta_arcs);
      for (int i = 0; i < h.nrad; ++i) tg[K * 17 + W].push_back(RingTile{d, i});
    }
    while (st.size() % FGX_TILE_PAD) st.push_back(CellTile{-1, 0, 0, 0});
    n_stiles = (int)st.size();
    // fused source pass: on by default, FGX_FUSED_SOURCE=0 keeps the
    // separate kernels; FGX_SOURCE_RINGS sets the strip height
    if (const char* e = std::getenv("FGX_FUSED_SOURCE")) fused_source = std::atoi(e) != 0;
    if (const char* e = std::getenv("FGX_VISC_SKIP")) visc_skip = std::atoi(e) != 0;
    if (tl && tl->source_rings > 0) src_rings = tl->source_rings;
    else if (const char* e = std::getenv("FGX_SOURCE_RINGS")) src_rings = std::atoi(e);
    src_rings = std::min(std::max(src_rings, 1), FGX_SRC_RINGS_MAX);   // k_source stages rs + 8 rings
    std::vector<CellTile> srt;
    for (int d = 0; d < nd; ++d)
      for (int i0 = 0; i0 < hd[d].nrad; i0 += src_rings)
        for (int jb = 0; jb * FGX_SRC_OUT < hd[d].nsec; ++jb) srt.push_back(CellTile{d, i0, jb, 0});
    while (srt.size() % FGX_TILE_PAD) srt.push_back(CellTile{-1, 0, 0, 0});
    n_srctiles = (int)srt.size();
    // the fine tiling: 2-ring strips, 4-ring source strips
    int fine_src = 4;
    if (tl && tl->fine_source_rings > 0) fine_src = tl->fine_source_rings;
    else if (const char* e = std::getenv("FGX_FINE_SOURCE_RINGS")) fine_src = std::atoi(e);
    fine_src = std::min(std::max(fine_src, 1), src_rings);
    int fine_strip = 2;
    if (tl && tl->fine_strip_rings > 0) fine_strip = tl->fine_strip_rings;
    fine_strip = std::min(std::max(fine_strip, 2), strip_rings);
    std::vector<CellTile> stf, srtf;
    for (int d = 0; d < nd; ++d) {
      total_cells += (long long)hd[d].nrad * hd[d].nsec;
      for (int i0 = 0; i0 < hd[d].nrad; i0 += fine_strip)
        for (int jb = 0; jb * FGX_TPB < hd[d].nsec; ++jb) stf.push_back(CellTile{d, i0, jb, 0});
      for (int i0 = 0; i0 < hd[d].nrad; i0 += fine_src)
        for (int jb = 0; jb * FGX_SRC_OUT < hd[d].nsec; ++jb) srtf.push_back(CellTile{d, i0, jb, 0});
    }
    while (stf.size() % FGX_TILE_PAD) stf.push_back(CellTile{-1, 0, 0, 0});
    while (srtf.size() % FGX_TILE_PAD) srtf.push_back(CellTile{-1, 0, 0, 0});
    if (tl && tl->fine_cells != 0.0) fine_cells = tl->fine_cells;
    else if (const char* e = std::getenv("FGX_FINE_CELLS")) fine_cells = std::atof(e);
    tiling_used = fgx_tiling{strip_rings, src_rings, fine_src, fine_strip, fine_cells};
    while (ct.size() % FGX_TILE_PAD) ct.push_back(CellTile{-1, 0, 0, 0});
    while (rt.size() % FGX_TILE_PAD) rt.push_back(RingTile{-1, 0});
    n_ctiles = (int)ct.size();
    n_rtiles = (int)rt.size();
    // device arrays
    const int ad = P0.adiabatic;
    auto f = [&](real** p, bool need) -> int { return need ? dalloc_share(p, F, 0) : dalloc(p, 1); };
    LCHK(f(&D.dens, true)); LCHK(f(&D.vrad, true)); LCHK(f(&D.vtheta, true)); LCHK(f(&D.energy, ad));
    LCHK(f(&D.cs, ad)); LCHK(f(&D.vradint, true)); LCHK(f(&D.vthetaint, true));
    LCHK(f(&D.taurr, true)); LCHK(f(&D.taupp, true)); LCHK(f(&D.taurp, true));
    LCHK(f(&D.divv, ad)); LCHK(f(&D.qplus, ad)); LCHK(f(&D.energyint, ad)); LCHK(f(&D.pressure, true));
    LCHK(f(&D.t_dens, true)); LCHK(f(&D.t_rmp, true)); LCHK(f(&D.t_rmm, true));
    LCHK(f(&D.t_tmp, true)); LCHK(f(&D.t_tmm, true)); LCHK(f(&D.t_e, ad)); LCHK(f(&D.work, true));
    const int lab = (P0.advect_label && P0.ext_stubbed_physics) ? 1 : 0;
    LCHK(f(&D.label, lab)); LCHK(f(&D.t_lab, lab));
    const int sg = P0.self_gravity;
    LCHK(f(&D.sg_accr, sg)); LCHK(f(&D.sg_acct, sg));
    LCHK(dalloc_share(&D.phi, F, 0));
    if (sg) {
      while (sgct.size() % FGX_TILE_PAD) sgct.push_back(RingTile{-1, 0});
      n_sgctiles = (int)sgct.size();
      cxd *w, *kr, *kt, *tw;
      RingTile* sgt;
      LCHK(dalloc_share(&w, SGW, 3)); LCHK(dalloc(&kr, SGK)); LCHK(dalloc(&kt, SGK)); LCHK(dalloc(&tw, STW));
      LCHK(dalloc_share(&D.axisg, R, 1)); LCHK(dalloc(&sgt, sgct.size()));
      D.sgw = w; D.sgkr = kr; D.sgkt = kt; D.sgtw = tw; D.sg_ctiles = sgt;
      HIPCHK(hipMemcpyAsync((void*)sgt, sgct.data(), sgct.size() * sizeof(RingTile), hipMemcpyHostToDevice, stream));
      LCHK(add_list(sgct, (const RingTile*)sgt, list_sgcols));
      sg_ctiles_full = sgt;
      n_sgctiles_full = n_sgctiles;
      std::vector<cxd> twh(STW);
      for (int d = 0; d < nd; ++d) {
        const DiskConst& c = dch[d];
        const int lens[2] = {c.nsec, 2 * c.nrad}, offs[2] = {c.tw_ns, c.tw_nr};
        for (int q = 0; q < 2; ++q)
          for (int m = 0; m < lens[q]; ++m) {
            const double a = 2.0 * kPI * (double)m / (double)lens[q];
            twh[offs[q] + m] = cxd{std::cos(a), -std::sin(a)};   // exp(-2 pi i m / n)
          }
        HIPCHK(hipMemcpy((void*)(kr + c.sgk_o), khr[d].data(), khr[d].size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void*)(kt + c.sgk_o), kht[d].data(), kht[d].size() * sizeof(double), hipMemcpyHostToDevice));
      }
      HIPCHK(hipMemcpy((void*)tw, twh.data(), twh.size() * sizeof(cxd), hipMemcpyHostToDevice));
    }
    real** rings[] = {&D.rinf, &D.rsup, &D.rmed, &D.surf, &D.invsurf, &D.invrmed, &D.invrinf, &D.invdiffrmed,
                      &D.invdiffrsup, &D.dxtheta, &D.supp_torque, &D.sigmamed, &D.energymed, &D.coolingtime,
                      &D.qplusmed, &D.visc, &D.cs_ring, &D.csmean, &D.vmed, &D.newdt};
    for (real** r : rings) LCHK(dalloc_share(r, R, 1));
    LCHK(dalloc_share(&D.nshift, R, 1));
    {
      double *vref, *acc0, *u0, *gq0, *rinf_d, *invrinf_d, *rmed_d, *invrmed_d, *invdiffrmed_d, *cosj_d, *sinj_d, *cs2_d;
      LCHK(dalloc_share(&vref, R, 1)); LCHK(dalloc_share(&acc0, R, 1)); LCHK(dalloc_share(&u0, R, 1));
      LCHK(dalloc_share(&gq0, R, 1)); LCHK(dalloc_share(&rinf_d, R, 1)); LCHK(dalloc_share(&invrinf_d, R, 1));
      LCHK(dalloc_share(&cs2_d, R, 1)); D.cs2_d = cs2_d;
      double *sigmamed_d, *energymed_d;
      LCHK(dalloc_share(&sigmamed_d, R, 1)); LCHK(dalloc_share(&energymed_d, R, 1));
      D.sigmamed_d = sigmamed_d; D.energymed_d = energymed_d;
      LCHK(dalloc_share(&rmed_d, R, 1)); LCHK(dalloc_share(&invrmed_d, R, 1)); LCHK(dalloc_share(&invdiffrmed_d, R, 1));
      LCHK(dalloc_share(&D.vmoy, R, 1)); LCHK(dalloc_share(&cosj_d, S, 2)); LCHK(dalloc_share(&sinj_d, S, 2));
      if (P0.evanescent) { LCHK(dalloc_share(&D.ev_d2, R, 1)); LCHK(dalloc_share(&D.ev_vt0s, R, 1)); }
      D.vref = vref; D.acc0 = acc0; D.u0 = u0; D.gq0 = gq0; D.rinf_d = rinf_d; D.invrinf_d = invrinf_d; D.rmed_d = rmed_d;
      D.invdiffrmed_d = invdiffrmed_d; D.cosj_d = cosj_d; D.sinj_d = sinj_d; D.invrmed_d = invrmed_d;
    }
    LCHK(dalloc_share(&D.cosj, S, 2)); LCHK(dalloc_share(&D.sinj, S, 2)); LCHK(dalloc_share(&D.lost, S, 2));
    DiskConst* dcd; DiskState* std_;
    LCHK(dalloc(&dcd, nd)); LCHK(dalloc(&std_, nd));
    D.dc = dcd; D.st = std_;
    LCHK(dalloc_share(&D.fpart, (size_t)n_rtiles * 2 * FGX_NFORCE_T, 4));
    LCHK(dalloc(&D.rpart, (size_t)std::max(max_nrad, 1) * 44));
    LCHK(dalloc(&D.summary, 2 + FGX_SUMMARY_HIST + nd));
    LCHK(dalloc(&D.health, nd));
    health_h.assign(nd, 0);
    frozen_h.assign(nd, 0);
    CellTile* ctd; RingTile* rtd;
    LCHK(dalloc(&ctd, ct.size())); LCHK(dalloc(&rtd, rt.size()));
    D.ctiles = ctd; D.rtiles = rtd;
    {
      CellTile* std_t;
      LCHK(dalloc(&std_t, st.size()));
      HIPCHK(hipMemcpyAsync((void*)std_t, st.data(), st.size() * sizeof(CellTile), hipMemcpyHostToDevice, stream));
      stiles = std_t;
      D.stiles = std_t;
      D.strip_rings = strip_rings;
      CellTile* srt_d;
      LCHK(dalloc(&srt_d, srt.size()));
      HIPCHK(hipMemcpyAsync((void*)srt_d, srt.data(), srt.size() * sizeof(CellTile), hipMemcpyHostToDevice, stream));
      srctiles = srt_d;
      coarse = TileSet{std_t, srt_d, n_stiles, strip_rings, n_srctiles, src_rings};
      CellTile *stf_d, *srtf_d;
      LCHK(dalloc(&stf_d, stf.size()));
      LCHK(dalloc(&srtf_d, srtf.size()));
      HIPCHK(hipMemcpyAsync((void*)stf_d, stf.data(), stf.size() * sizeof(CellTile), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync((void*)srtf_d, srtf.data(), srtf.size() * sizeof(CellTile), hipMemcpyHostToDevice, stream));
      fine = TileSet{stf_d, srtf_d, (int)stf.size(), fine_strip, (int)srtf.size(), fine_src};
      LCHK(add_list(st, (const CellTile*)std_t, list_coarse_st));
      LCHK(add_list(srt, (const CellTile*)srt_d, list_coarse_src));
      LCHK(add_list(stf, (const CellTile*)stf_d, list_fine_st));
      LCHK(add_list(srtf, (const CellTile*)srtf_d, list_fine_src));
      for (int kw = 0; kw < 17 * 17; ++kw) {
        std::vector<RingTile>& v = tg[kw];
        if (v.empty()) continue;
        while (v.size() % FGX_TILE_PAD) v.push_back(RingTile{-1, 0});
        RingTile* g;
        LCHK(dalloc(&g, v.size()));
        HIPCHK(hipMemcpyAsync((void*)g, v.data(), v.size() * sizeof(RingTile), hipMemcpyHostToDevice, stream));
        const int K = kw / 17, W = kw % 17;
        int li = -1;
        LCHK(add_list(v, (const RingTile*)g, li));
        if ((K == 2 || K == 3) && W >= 2 && theta_multi.ng < FGX_THETA_MULTI) {
          ThetaMulti& M = theta_multi;
          M.k[M.ng] = K; M.w[M.ng] = W; M.tiles[M.ng] = g;
          M.first[M.ng] = theta_multi_n;
          theta_multi_n += (int)v.size();
          M.first[++M.ng] = theta_multi_n;
          theta_multi_block = std::max(theta_multi_block, 64 * W);
          list_multi.push_back(li);
        } else {
          theta_groups.push_back(ThetaGroup{K, W, (int)v.size(), g});
          list_groups.push_back(li);
        }
      }
      HIPCHK(hipStreamSynchronize(stream));
    }
    D.max_nsec = max_nsec; D.ndisk = nd;
    D.adiabatic = ad; D.fast_transport = P0.fast_transport; D.open_inner = P0.open_inner;
    D.nonreflecting = P0.nonreflecting; D.evanescent = P0.evanescent; D.self_gravity = P0.self_gravity;
    D.cooling = P0.cooling; D.indirect = P0.indirect_term; D.exclude_hill = P0.exclude_hill;
    D.fix_taurp = P0.ext_fix_taurp_corner; D.fix_pp = P0.ext_fix_pp_distance; D.fix_sgwrap = P0.ext_fix_sg_wrap;
    D.sloppy = P0.sloppy_cfl; D.corotating = P0.corotating; D.guiding = P0.guiding_center;
    D.forced_circular = P0.forced_circular; D.roche = P0.roche_smoothing;
    D.write_pressure = (!ad && P0.write_temperature) ? 1 : 0;
    D.advect_label = lab;
    D.ignore_active = 0;
    D.only_disk = -1;
    int viscous = 0;
    for (int d = 0; d < nd; ++d)
      for (int i = 0; i < hd[d].nrad; ++i) viscous |= (hd[d].visc[i] != 0.0);
    D.viscous = viscous || visc_dynamic;
    // uploads
    HIPCHK(hipMemcpyAsync((void*)ctd, ct.data(), ct.size() * sizeof(CellTile), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync((void*)rtd, rt.data(), rt.size() * sizeof(RingTile), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync((void*)dcd, dch.data(), nd * sizeof(DiskConst), hipMemcpyHostToDevice, stream));
    // the whole lists, and room for their narrowed copies
    LCHK(add_list(rt, (const RingTile*)rtd, list_rings));
    coarse_full = coarse; fine_full = fine;
    theta_multi_full = theta_multi; theta_multi_n_full = theta_multi_n; theta_groups_full = theta_groups;
    rtiles_full = D.rtiles; n_rtiles_full = n_rtiles;
    covered = nd; covered_cells = total_cells;
    max_ranges = (int)tlists.size() * (nd + 1);
    HIPCHK(hipHostMalloc((void**)&h_ranges, max_ranges * sizeof(CompactRange)));
    LCHK(dalloc(&d_ranges, max_ranges));
    if (const char* e = std::getenv("FGX_COMPACT")) compact_on = std::atoi(e) != 0;
    if (const char* e = std::getenv("FGX_GRAPH")) graphs_on = std::atoi(e) != 0;
    std::vector<DiskState> sth(nd);
    for (int d = 0; d < nd; ++d) sth[d] = initial_state(d);
    HIPCHK(hipMemcpyAsync((void*)std_, sth.data(), nd * sizeof(DiskState), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    LCHK(upload_rings());
    for (int d = 0; d < nd; ++d) LCHK(upload_fields_disk(d));
    // initial pressure (InitEuler's ComputePressureField), used by the
    // isothermal temperature output
    D.ignore_active = 1;
    LCHK(launch("pressure_init", 2, k_pressure_init<real>, n_ctiles, FGX_TPB, 0, D));
    if (ad) LCHK(launch("sound_speed", 3, k_sound_speed<real>, n_stiles, FGX_TPB, 0, D));
    D.ignore_active = 0;
    // LDS of the azimuthal transport kernel
    int fb_nsec = 0;   // widest ring of the shared-memory azimuthal kernel
    for (int d = 0; d < nd; ++d) {
      int K = 0, W = 0;
      theta_split<real>(hd[d].nsec, K, W, theta_arcs);
      if (K == 0) fb_nsec = std::max(fb_nsec, hd[d].nsec);
    }
    theta_shmem = (size_t)(FGX_THETA_ARRAYS + lab) * fb_nsec * sizeof(real);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    size_t lds_max = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : prop.sharedMemPerBlock;
    if (theta_shmem + 1024 > lds_max)
      return set_error(FGX_EPARAM, "NSEC=%d needs %zu B of LDS per ring in this precision; the device has %zu", fb_nsec,
                       theta_shmem, lds_max);
    LCHK(raise_dynamic_lds((const void*)k_trans_theta<real>, theta_shmem));
    if (sg) {
      // self-gravity transforms: two rows of nsec, four columns of 2*nrad (FP64 complex,
      // one pad per eight elements: lds_len)
      sg_shm_row = 2 * (size_t)lds_len(max_nsec) * sizeof(cxd);
      sg_shm_col = 2 * (size_t)lds_len(4 * max_nrad) * sizeof(cxd);
      if (sg_shm_row + 1024 > lds_max || sg_shm_col + 1024 > lds_max)
        return set_error(FGX_EPARAM, "self-gravity transforms of NSEC=%d, 2*NRAD=%d exceed the %zu B of LDS", max_nsec,
                         2 * max_nrad, lds_max);
      const void* kf[3] = {(const void*)k_sg_rows_fwd<real>, (const void*)k_sg_rows_inv<real>,
                           (const void*)k_sg_cols<real>};
      const size_t sz[3] = {sg_shm_row, sg_shm_row, sg_shm_col};
      for (int q = 0; q < 3; ++q)
        LCHK(raise_dynamic_lds(kf[q], sz[q]));
      // the SG field of the initial density (read_field before the first step)
      D.ignore_active = 1;
      LCHK(self_gravity());
      D.ignore_active = 0;
    }
    HIPCHK(hipStreamSynchronize(stream));
    return FGX_OK;
  }

  // the state record of a disk at set-up
  DiskState initial_state(int d) const {
    DiskState s;
    std::memset(&s, 0, sizeof s);
    const HostDisk& h = hd[d];
    for (int k = 0; k < h.sys.nb; ++k) {
      s.x[k] = h.sys.x[k]; s.y[k] = h.sys.y[k]; s.vx[k] = h.sys.vx[k]; s.vy[k] = h.sys.vy[k];
      s.mass[k] = h.sys.mass[k];
    }
    s.omega_frame = h.omega_frame;
    s.active = 0;
    return s;
  }
  // the initial fields of a disk, in the engine's representation
  int upload_fields_disk(int d) {
    // the density is kept as a residual of S0, the field-precision SigmaMed
    {
      std::vector<double> sres(hd[d].dens);
      for (int i = 0; i < hd[d].nrad; ++i)
        for (int j = 0; j < hd[d].nsec; ++j) sres[(size_t)i * hd[d].nsec + j] -= field_base(d, FGX_DENS, i);
      LCHK(upload_field(d, D.dens, sres));
    }
    LCHK(upload_field(d, D.vrad, hd[d].vrad));
    std::vector<double> res(hd[d].vtheta);
    for (int i = 0; i < hd[d].nrad; ++i)
      for (int j = 0; j < hd[d].nsec; ++j) res[(size_t)i * hd[d].nsec + j] -= vref_h[d][i];
    LCHK(upload_field(d, D.vtheta, res));
    // the ghost ring (ring nrad) of the velocities into the other buffer of
    // each pair as well: no substep writes that ring, and a disk that ends
    // its DT before its batch does has its velocities copied across the pair
    // with it (k_vel_fixup), so with both copies equal the ring reads the
    // same whatever the batch's substep count, i.e. whatever shares the batch
    {
      const size_t ns = (size_t)hd[d].nsec, g0 = (size_t)hd[d].nrad * ns;
      LCHK(upload_vec(D.vradint + dch[d].fo + g0, std::vector<double>(hd[d].vrad.begin() + g0, hd[d].vrad.end()), ns));
      LCHK(upload_vec(D.vthetaint + dch[d].fo + g0, std::vector<double>(res.begin() + g0, res.end()), ns));
    }
    if (D.adiabatic) {
      // the energy is kept as a residual of its ring profile EnergyMed
      std::vector<double> eres(hd[d].energy);
      for (int i = 0; i < hd[d].nrad; ++i)
        for (int j = 0; j < hd[d].nsec; ++j) eres[(size_t)i * hd[d].nsec + j] -= hd[d].energymed[i];
      LCHK(upload_field(d, D.energy, eres));
    }
    if (D.advect_label) LCHK(upload_field(d, D.label, hd[d].label));
    return FGX_OK;
  }

  // The ring profile a field is stored as a residual of (vtheta: vref; the
  // energy: EnergyMed; the density: SigmaMed in the field precision), 0 for
  // the other fields and for the ghost ring
  double field_base(int d, int field, int i) const {
    const HostDisk& h = hd[d];
    if (i < 0 || i >= h.nrad) return 0.0;
    switch (field) {
      case FGX_VTHETA: return vref_h.empty() ? 0.0 : vref_h[d][i];
      case FGX_ENERGY: return h.P.adiabatic ? h.energymed[i] : 0.0;
      case FGX_DENS: return (double)(real)h.sigmamed[i];
      default: return 0.0;
    }
  }
  template <class V>
  int upload_vec(real* dst, const V& src, size_t n) {
    std::vector<real> tmp(n);
    for (size_t k = 0; k < n; ++k) tmp[k] = (real)src[k];
    HIPCHK(hipMemcpy(dst, tmp.data(), n * sizeof(real), hipMemcpyHostToDevice));
    return FGX_OK;
  }
  int upload_field(int d, real* base, const std::vector<double>& v) {
    return upload_vec(base + dch[d].fo, v, v.size());
  }
  int upload_d(const double* dst, const std::vector<double>& v) {
    HIPCHK(hipMemcpy((void*)dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    return FGX_OK;
  }
  int upload_rings() {
    vref_h.assign(nd, {});
    for (int d = 0; d < nd; ++d) LCHK(upload_rings_disk(d));
    return FGX_OK;
  }
  int upload_rings_disk(int d) {
    {
      const HostDisk& h = hd[d];
      const int o = dch[d].ro;
      const size_t n = (size_t)h.nrad + 1;
      // FP64 tables: reference vtheta (the initial, ring-uniform profile),
      // geometry, and Substep1's ring-profile terms
      std::vector<double> vref(n, 0.0), acc0(n, 0.0), u0(n, 0.0), gq0(n, 0.0), invrinf(n, 0.0), cs2(n, 0.0),
          invrmed(n, 0.0);
      for (int i = 0; i < h.nrad; ++i) cs2[i] = h.cs_ring[i] * h.cs_ring[i];
      LCHK(upload_d(D.cs2_d + o, cs2));
      LCHK(upload_d(D.sigmamed_d + o, h.sigmamed)); LCHK(upload_d(D.energymed_d + o, h.energymed));
      for (int i = 0; i < h.nrad; ++i) vref[i] = h.vtheta[(size_t)i * h.nsec];
      for (int i = 0; i < h.nrad; ++i) invrinf[i] = 1.0 / h.rinf[i];
      for (int i = 0; i < h.nrad; ++i) invrmed[i] = 1.0 / h.rmed[i];
      // Substep1's radial acceleration (Kernels.cu:52-68) of the ring profiles
      // themselves (density S0, the field-precision SigmaMed; energy EnergyMed;
      // vtheta vref; frame OmegaFrame at set-up; star potential -G/r), at each
      // interface i, in long double:
      //   acc0 = -(P0_i - P0_i-1) 2 / (S0_i + S0_i-1) / dRmed - (phi_i - phi_i-1) / dRmed
      //          + u0^2 / Rinf,   u0 = (vref_i + vref_i-1) / 2 + Rinf OmegaFrame
      // with the same FP64 geometry tables the kernels use.  The kernels add
      // the change that the fields' departures from the profiles make, so the
      // O(1) gravity and centrifugal terms, which cancel to O(h^2), never meet
      // in the field arithmetic.  gq0 = (P0_i - P0_i-1) / (S0_i + S0_i-1).
      {
        typedef long double W;
        const W g1 = (W)(h.P.adiabaticindex - 1.0);
        auto s0 = [&](int i) -> W { return (W)(real)h.sigmamed[i]; };
        auto p0 = [&](int i) -> W {
          return h.P.adiabatic ? g1 * (W)h.energymed[i] : s0(i) * (W)h.cs_ring[i] * (W)h.cs_ring[i];
        };
        for (int i = 1; i < h.nrad; ++i) {
          const W idr = (W)h.invdiffrmed[i];
          const W ss = s0(i) + s0(i - 1), dp = p0(i) - p0(i - 1);
          const W gradp = dp * (W)2 / ss * idr;
          const W gradphi = ((W)-1 / (W)h.rmed[i] + (W)1 / (W)h.rmed[i - 1]) * idr;
          const W u = ((W)2 * (W)vref[i] + (W)2 * (W)vref[i - 1]) / (W)4 + (W)h.rinf[i] * (W)h.omega_frame;
          acc0[i] = (double)(-gradp - gradphi + u * u * (W)invrinf[i]);
          u0[i] = (double)u;
          gq0[i] = (double)(dp / ss);
        }
      }
      vref_h[d] = vref;
      LCHK(upload_d(D.vref + o, vref)); LCHK(upload_d(D.acc0 + o, acc0)); LCHK(upload_d(D.u0 + o, u0));
      LCHK(upload_d(D.gq0 + o, gq0));
      LCHK(upload_d(D.rinf_d + o, h.rinf)); LCHK(upload_d(D.invrinf_d + o, invrinf));
      LCHK(upload_d(D.rmed_d + o, h.rmed)); LCHK(upload_d(D.invdiffrmed_d + o, h.invdiffrmed));
      LCHK(upload_d(D.invrmed_d + o, invrmed));
      LCHK(upload_d(D.cosj_d + dch[d].so, h.cosj)); LCHK(upload_d(D.sinj_d + dch[d].so, h.sinj));
      LCHK(upload_vec(D.rinf + o, h.rinf, n)); LCHK(upload_vec(D.rsup + o, h.rsup, n));
      LCHK(upload_vec(D.rmed + o, h.rmed, n)); LCHK(upload_vec(D.surf + o, h.surf, n));
      LCHK(upload_vec(D.invsurf + o, h.invsurf, n)); LCHK(upload_vec(D.invrmed + o, h.invrmed, n));
      LCHK(upload_vec(D.invrinf + o, h.invrinf, n)); LCHK(upload_vec(D.invdiffrmed + o, h.invdiffrmed, n));
      LCHK(upload_vec(D.invdiffrsup + o, h.invdiffrsup, n)); LCHK(upload_vec(D.dxtheta + o, h.dxtheta, n));
      LCHK(upload_vec(D.supp_torque + o, h.supp_torque, n)); LCHK(upload_vec(D.sigmamed + o, h.sigmamed, n));
      LCHK(upload_vec(D.energymed + o, h.energymed, n)); LCHK(upload_vec(D.coolingtime + o, h.coolingtime, n));
      LCHK(upload_vec(D.qplusmed + o, h.qplusmed, n)); LCHK(upload_vec(D.visc + o, h.visc, n));
      LCHK(upload_vec(D.cs_ring + o, h.cs_ring, n));
      // ring-mean sound speed as Make1Dprofile(1) leaves it after set-up
      std::vector<double> csm(n, 0.0);
      if (!h.P.adiabatic) csm = h.cs_ring;
      else
        for (int i = 0; i < h.nrad; ++i) {
          double s = 0;
          for (int j = 0; j < h.nsec; ++j) {
            size_t k = (size_t)i * h.nsec + j;
            s += (double)(real)std::sqrt((double)h.P.adiabaticindex * (h.P.adiabaticindex - 1.0) *
                                         (double)(real)h.energy[k] / (double)(real)h.dens[k]);
          }
          csm[i] = s / h.nsec;
        }
      LCHK(upload_vec(D.csmean + o, csm, n));
      LCHK(upload_vec(D.cosj + dch[d].so, h.cosj, h.nsec));
      LCHK(upload_vec(D.sinj + dch[d].so, h.sinj, h.nsec));
    }
    // the evanescent boundary's per-ring terms, from the geometry just uploaded
    if (D.evanescent) LCHK(launch("ev_tables", 0, k_ev_tables<real>, 1, FGX_TPB, 0, D, d));
    return FGX_OK;
  }

  // ------------------------------------------------------------------ step
  // the CFL and force pass: a ring of more than 768 sectors takes one wave per
  // 512 sectors (at most 8; cfl_waves), blocks sized for the batch's widest
  int cfl_force(int mode) {
    const double b = mode == 0 ? 3 : 1;
    switch (cfl_waves(max_nsec)) {
      case 1: return launch("cfl_force", b, k_cfl_force<real, 1>, n_rtiles, 64, 0, D, mode);
      case 2: return launch("cfl_force", b, k_cfl_force<real, 2>, n_rtiles, 128, 0, D, mode);
      case 3: return launch("cfl_force", b, k_cfl_force<real, 3>, n_rtiles, 192, 0, D, mode);
      case 4: return launch("cfl_force", b, k_cfl_force<real, 4>, n_rtiles, 256, 0, D, mode);
      case 5: return launch("cfl_force", b, k_cfl_force<real, 5>, n_rtiles, 320, 0, D, mode);
      case 6: return launch("cfl_force", b, k_cfl_force<real, 6>, n_rtiles, 384, 0, D, mode);
      case 7: return launch("cfl_force", b, k_cfl_force<real, 7>, n_rtiles, 448, 0, D, mode);
      default: return launch("cfl_force", b, k_cfl_force<real, 8>, n_rtiles, 512, 0, D, mode);
    }
  }
  // azimuthal transport: one launch per lane width K
  // g.w: 0 = one full wave per ring, 1 = one partial wave, >= 2 = arcs
  template <int K>
  int theta_k(const ThetaGroup& g) {
    if (D.advect_label) return D.adiabatic ? theta_kl<K, true, true>(g) : theta_kl<K, true, false>(g);
    return D.adiabatic ? theta_kl<K, false, true>(g) : theta_kl<K, false, false>(g);
  }
  template <int K, bool LAB, bool AD>
  int theta_kl(const ThetaGroup& g) {
    if constexpr (sizeof(real) == 8 && K > 8) {
      return set_error(FGX_EDEVICE, "no FP64 azimuthal kernel for K=%d", g.k);
    } else {
      if (g.w == 0) return launch("trans_theta", 0, k_trans_theta_arc<real, K, 1, LAB, AD>, g.n, 64, 0, D, g.tiles, 1);
      if (g.w == 1) return launch("trans_theta", 0, k_trans_theta_arc<real, K, 0, LAB, AD>, g.n, 64, 0, D, g.tiles, 1);
      if constexpr (K >= 2 && K <= 4) {
        return launch("trans_theta", 0, k_trans_theta_arc<real, K, 2, LAB, AD>, g.n, 64 * g.w, 0, D, g.tiles, g.w);
      } else {
        return set_error(FGX_EDEVICE, "no multi-wave azimuthal kernel for K=%d", g.k);
      }
    }
  }
  template <bool LAB, bool AD>
  int theta_multi_launch() {
    return launch("trans_theta", 0, k_trans_theta_multi<real, LAB, AD>, theta_multi_n, theta_multi_block, 0, D, theta_multi);
  }
  int trans_theta(double bytes_per_cell) {
    group_begin("trans_theta", bytes_per_cell);
    int r = FGX_OK;
    if (theta_multi.ng > 0) {
      if (D.advect_label) r = D.adiabatic ? theta_multi_launch<true, true>() : theta_multi_launch<true, false>();
      else r = D.adiabatic ? theta_multi_launch<false, true>() : theta_multi_launch<false, false>();
    }
    for (const ThetaGroup& g : theta_groups) {
      if (r) break;
      switch (g.k) {
        case 0: r = launch("trans_theta", 0, k_trans_theta<real>, g.n, FGX_TPB, theta_shmem, D, g.tiles); break;
        case 1: r = theta_k<1>(g); break;
        case 2: r = theta_k<2>(g); break;
        case 3: r = theta_k<3>(g); break;
        case 4: r = theta_k<4>(g); break;
        case 5: r = theta_k<5>(g); break;
        case 6: r = theta_k<6>(g); break;
        case 8: r = theta_k<8>(g); break;
        case 10: r = theta_k<10>(g); break;
        case 12: r = theta_k<12>(g); break;
        case 16: r = theta_k<16>(g); break;
        default: r = set_error(FGX_EDEVICE, "no azimuthal kernel for K=%d", g.k);
      }
      if (r) break;
    }
    group_end();
    return r;
  }
  // compute_selfgravity (SgMain.cu:22-47) on the current density: the three
  // transform passes, leaving SG_Accr, SG_Acct and their ring means
  int self_gravity() {
    const double r = 1.0 / sizeof(real);
    // as many threads as a stage has butterflies (at most 1024, at least 256)
    auto threads = [](int butterflies) { return std::min(1024, std::max(FGX_TPB, (butterflies + 63) / 64 * 64)); };
    const int trow = threads(max_nsec / 3), tcol = threads(max_nrad);
    LCHK(launch("sg_rows", 20 * r, k_sg_rows_fwd<real>, n_rtiles, trow, sg_shm_row, D));
    LCHK(launch("sg_cols", 96 * r, k_sg_cols<real>, n_sgctiles, tcol, sg_shm_col, D));
    LCHK(launch("sg_rows", 20 * r + 2, k_sg_rows_inv<real>, n_rtiles, trow, sg_shm_row, D));
    return FGX_OK;
  }
  int visc_profile(double bytes, int update_csmean) {
    if (visc_skip && !visc_time && visc_fresh && (!update_csmean || !cs_dirty)) return FGX_OK;
    LCHK(launch("visc_profile", bytes, k_visc_profile<real>, n_rtiles, FGX_TPB, 0, D, update_csmean, alpha));
    visc_fresh = true;
    if (update_csmean) cs_dirty = false;
    return FGX_OK;
  }
  int sound_speed() {
    cs_dirty = true;
    return launch("sound_speed", 3, k_sound_speed<real>, n_stiles, FGX_TPB, 0, D);
  }
  int boundary() {
    if (D.open_inner) LCHK(launch("open_boundary", 0, k_open_boundary<real>, nd, FGX_TPB, 0, D));
    if (D.nonreflecting) {
      if (D.adiabatic) LCHK(sound_speed());
      LCHK(launch("nonreflecting", 0, k_nonreflecting<real>, nd, FGX_TPB, 0, D));
    }
    if (D.evanescent) LCHK(launch("evanescent", D.adiabatic ? 8 : 6, k_evanescent<real>, n_stiles, FGX_TPB, 0, D));
    if (outer_source) LCHK(launch("outer_source", 0, k_outer_source<real>, nd, FGX_TPB, 0, D));
    return FGX_OK;
  }
  int substep() {
    const int ad = D.adiabatic;
    // the tiling the last summary's forecast chose for this substep
    use_tiles(plan_pos < (int)fine_plan.size() && fine_plan[plan_pos]);
    ++plan_pos;
    LCHK(cfl_force(0));
    if (accretion) {
      // AccreteOntoPlanets needs the substep's dt and comes before the forces
      LCHK(launch("substep_setup", 0, k_substep_setup<real>, nd, FGX_TPB, 0, D, 1));
      LCHK(launch("accrete", 0, k_accrete<real>, nd, FGX_TPB, 0, D));
      LCHK(cfl_force(1));
      LCHK(launch("substep_setup", 0, k_substep_setup<real>, nd, FGX_TPB, 0, D, 2));
    } else {
      LCHK(launch("substep_setup", 0, k_substep_setup<real>, nd, FGX_TPB, 0, D, 0));
    }
    if (D.corotating) LCHK(launch("correct_vtheta", 2, k_correct_vtheta<real>, n_stiles, FGX_TPB, 0, D));
    if (visc_dynamic) LCHK(visc_profile(0, 0));
    LCHK(boundary());
    if (fused_source) {
      // the viscosity profile only depends on the sound speed, which the
      // source stages do not change, so it can be formed first
      if (D.viscous && visc_dynamic) LCHK(visc_profile(ad ? 1 : 0, 1));
      if (D.self_gravity) {
        // the self-gravity field only depends on the density, which the
        // source stages read but do not change; the pass adds it after
        // Substep1, reading two more fields
        LCHK(self_gravity());
        LCHK(launch("source", ad ? 9 : 7, k_source<real, true>, n_srctiles, 64, 0, D, srctiles, src_rings));
      } else {
        LCHK(launch("source", ad ? 7 : 5, k_source<real, false>, n_srctiles, 64, 0, D, srctiles, src_rings));
      }
      // the new velocities were written to the VradInt / VthetaInt arrays
      std::swap(D.vrad, D.vradint);
      std::swap(D.vtheta, D.vthetaint);
      D.vpar ^= 1;
    } else {
      LCHK(launch("potential", 2, k_potential<real>, n_stiles, FGX_TPB, 0, D));
      LCHK(launch("substep1", ad ? 8 : 7, k_substep1<real>, n_stiles, FGX_TPB, 0, D));
      if (D.self_gravity) {
        LCHK(self_gravity());
        LCHK(launch("sg_update", 6, k_sg_update<real>, n_stiles, FGX_TPB, 0, D));
      }
      if (D.viscous) {
        if (visc_dynamic) LCHK(visc_profile(ad ? 1 : 0, 1));
        LCHK(launch("visc_tau", 6, k_visc_tau<real>, n_stiles, FGX_TPB, 0, D, (const real*)D.vradint,
                    (const real*)D.vthetaint, 0));
      }
      LCHK(launch("visc_update", 7, k_visc_update<real>, n_stiles, FGX_TPB, 0, D));
      LCHK(launch("substep2", ad ? 7 : 5, k_substep2<real>, n_stiles, FGX_TPB, 0, D));
    }
    LCHK(boundary());
    if (ad) {
      if (visc_dynamic) LCHK(visc_profile(1, 1));
      if (fused_source) {
        // the stresses stay in registers (bit-identical to the two kernels)
        LCHK(launch("substep3", 11, k_substep3_tau<real>, n_stiles, FGX_TPB, 0, D));
      } else {
        LCHK(launch("visc_tau", 7, k_visc_tau<real>, n_stiles, FGX_TPB, 0, D, (const real*)D.vrad,
                    (const real*)D.vtheta, 1));
        LCHK(launch("substep3", 9, k_substep3<real>, n_stiles, FGX_TPB, 0, D));
      }
    }
    if (D.advect_label)
      LCHK(launch("trans_rad", ad ? 12 : 10, k_trans_rad_strip<real, true>, n_stiles, FGX_TPB, 0, D, stiles, strip_rings));
    else
      LCHK(launch("trans_rad", ad ? 10 : 8, k_trans_rad_strip<real, false>, n_stiles, FGX_TPB, 0, D, stiles, strip_rings));
    LCHK(trans_theta(ad ? 11 : 9));
    LCHK(launch("trans_vr", 4, k_trans_vr<real>, n_stiles, FGX_TPB, 0, D));
    LCHK(boundary());
    use_tiles(false);
    return FGX_OK;
  }
  int step(int n_dt) override {
    if (covered < nd) lists_full();   // after a step that stopped on an error
    // the disks that take part in the DT: frozen ones (by the host, or for
    // their numeric health) are left out of every tile list from its start
    std::vector<int> act0(nd, 1);
    int n_act0 = 0;
    for (int d = 0; d < nd; ++d) {
      act0[d] = !frozen_h[d] && !health_h[d];
      n_act0 += act0[d];
    }
    visc_fresh = false;
    cs_dirty = true;
    visc_time = false;
    for (int d = 0; d < nd; ++d) visc_time |= dch[d].lambdadoubling != 0.0;
    for (int k = 0; k < n_dt; ++k) {
      LCHK(launch("begin_dt", 0, k_begin_dt<real>, (nd + 63) / 64, 64, 0, D));
      if (compact_on && n_act0 < nd) LCHK(lists_narrow(act0.data()));
      if (D.adiabatic) LCHK(sound_speed());
      int plan = 0;
      // the first substeps of the DT run every disk that is not frozen
      fine_plan.assign(1, (double)covered_cells < fine_cells);
      plan_pos = 0;
      for (int guard = 0;; ++guard) {
        if (guard > 10000000) return set_error(FGX_ENUMERIC, "substep loop did not terminate");
        LCHK((graphs_on && !prof_on) ? substep_graphed() : substep());
        if (plan > 1) { --plan; continue; }
        LCHK(launch("summary", 0, k_summary<real>, 1, FGX_TPB, 0, D));
        HIPCHK(hipMemcpyAsync(h_summary, D.summary, (2 + FGX_SUMMARY_HIST + nd) * sizeof(int), hipMemcpyDeviceToHost,
                              stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (prof_on) prof_collect();
        if (h_summary[0] == 0) break;
        // the disks still active cover the rest of the DT
        if (compact_on && h_summary[0] < covered) LCHK(lists_narrow(h_summary + 2 + FGX_SUMMARY_HIST));
        plan = std::max(1, h_summary[1]);
        // look again where the forecast halves the cells the lists cover (a
        // forecast that is off only moves that look)
        if (compact_on)
          for (int j = 2; j < plan && j < FGX_SUMMARY_HIST; ++j)
            if (2 * 64LL * h_summary[2 + j] < covered_cells) { plan = j; break; }
        // forecast: cells of the disks still active j substeps ahead (units
        // of 64 cells); short strips where that is a small share of the batch
        fine_plan.assign(plan, 0);
        plan_pos = 0;
        for (int j = 0; j < plan; ++j) {
          const long long act = 64LL * h_summary[2 + std::min(j, FGX_SUMMARY_HIST - 1)];
          fine_plan[j] = (double)act < fine_cells;
        }
      }
      // the end of every disk's last substep, then every later call sees
      // every disk again; then the disks that finished the DT before the
      // batch did
      LCHK(launch("finish_dt", 0, k_finish_dt<real>, nd, 64, 0, D));
      if (covered < nd) lists_full();
      use_tiles(false);
      if (fused_source) LCHK(launch("vel_fixup", 0, k_vel_fixup<real>, n_stiles, FGX_TPB, 0, D));
    }
    HIPCHK(hipStreamSynchronize(stream));
    if (prof_on) prof_collect();
    return check_health();
  }
  // disks found numerically unhealthy since the last look: the others have
  // completed the call; these stopped and were frozen where they were found
  int check_health() {
    std::vector<int> h(nd, 0);
    HIPCHK(hipMemcpy(h.data(), D.health, nd * sizeof(int), hipMemcpyDeviceToHost));
    int first = -1, nbad = 0;
    for (int d = 0; d < nd; ++d) {
      if (h[d] && !health_h[d]) { if (first < 0) first = d; ++nbad; }
      health_h[d] = h[d];
    }
    if (first >= 0)
      return set_error(FGX_ENUMERIC, "%d disk(s) numerically unhealthy and frozen, first disk %d (flags 0x%x)", nbad,
                       first, h[first]);
    return FGX_OK;
  }

  // ------------------------------------------------------------- accessors
  real* field_ptr(int field) {
    switch (field) {
      case FGX_DENS: return D.dens;
      case FGX_VRAD: return D.vrad;
      case FGX_VTHETA: return D.vtheta;
      case FGX_ENERGY: return D.adiabatic ? D.energy : nullptr;
      case FGX_SOUNDSPEED: return D.adiabatic ? D.cs : nullptr;
      case FGX_DIVV: return D.adiabatic ? D.divv : nullptr;
      case FGX_QPLUS: return D.adiabatic ? D.qplus : nullptr;
      case FGX_SG_ACCR: return D.self_gravity ? D.sg_accr : nullptr;
      case FGX_SG_ACCT: return D.self_gravity ? D.sg_acct : nullptr;
      case FGX_LABEL: return D.advect_label ? D.label : nullptr;
      default: return nullptr;
    }
  }
  int read_field(int d, int field, int dtype, void* host, size_t count) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    const HostDisk& h = hd[d];
    const size_t nfull = (size_t)(h.nrad + 1) * h.nsec, ncell = (size_t)h.nrad * h.nsec;
    if (count < ncell) return set_error(FGX_ERANGE, "buffer of %zu values < nrad*nsec = %zu", count, ncell);
    const size_t n = std::min(count, nfull);
    std::vector<real> tmp(nfull, (real)0);
    real* src = field_ptr(field);
    if (field == FGX_TEMPERATURE && !D.adiabatic && !D.write_pressure)
      // ComputeTemperatureField (Kernels.cu:211-223) divides the pressure of
      // the substep's start by the density of its end; that pressure is kept
      // only when the temperature is an output
      return set_error(FGX_EPARAM, "the isothermal temperature needs WRITETEMPERATURE YES");
    if (field == FGX_TEMPERATURE || field == FGX_POTENTIAL) {
      D.ignore_active = 1;
      int r = (field == FGX_TEMPERATURE) ? launch("temperature", 0, k_temperature<real>, n_ctiles, FGX_TPB, 0, D)
                                         : launch("potential", 0, k_potential_field<real>, n_ctiles, FGX_TPB, 0, D);
      D.ignore_active = 0;
      LCHK(r);
      src = D.work;
    } else if (field == FGX_LABEL && !src) {
      // without label advection the label keeps its initial values
      for (size_t k = 0; k < n; ++k) tmp[k] = (real)h.label[k];
      src = nullptr;
    } else if (field == FGX_SOUNDSPEED && !D.adiabatic) {
      for (int i = 0; i < h.nrad; ++i)
        for (int j = 0; j < h.nsec; ++j) tmp[(size_t)i * h.nsec + j] = (real)h.cs_ring[i];
    } else if (!src && field != FGX_ENERGY && field != FGX_DIVV && field != FGX_QPLUS) {
      return set_error(FGX_ERANGE, "unknown field %d", field);
    }
    if (src) {
      HIPCHK(hipMemcpyAsync(tmp.data(), src + dch[d].fo, nfull * sizeof(real), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    if (field == FGX_VTHETA || field == FGX_DENS || (field == FGX_ENERGY && D.adiabatic)) {
      // vtheta, the density and the energy are stored as residuals of ring profiles
      for (size_t k = 0; k < n; ++k) {
        const double v = (double)tmp[k] + field_base(d, field, (int)(k / h.nsec));
        if (dtype == FGX_F32) ((float*)host)[k] = (float)v; else ((double*)host)[k] = v;
      }
      return FGX_OK;
    }
    if (dtype == FGX_F32) for (size_t k = 0; k < n; ++k) ((float*)host)[k] = (float)tmp[k];
    else for (size_t k = 0; k < n; ++k) ((double*)host)[k] = (double)tmp[k];
    return FGX_OK;
  }
  int write_field(int d, int field, int dtype, const void* host, size_t count) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    if (covered < nd) { lists_full(); use_tiles(false); }
    const HostDisk& h = hd[d];
    const size_t nfull = (size_t)(h.nrad + 1) * h.nsec, ncell = (size_t)h.nrad * h.nsec;
    if (count < ncell) return set_error(FGX_ERANGE, "buffer of %zu values < nrad*nsec = %zu", count, ncell);
    real* dst = field_ptr(field);
    if (!dst || field == FGX_SOUNDSPEED || field == FGX_DIVV || field == FGX_QPLUS)
      return set_error(FGX_ERANGE, "field %d is not writable", field);
    const size_t n = std::min(count, nfull);
    std::vector<real> tmp(n);
    for (size_t k = 0; k < n; ++k) {
      double v = dtype == FGX_F32 ? (double)((const float*)host)[k] : ((const double*)host)[k];
      v -= field_base(d, field, (int)(k / h.nsec));
      tmp[k] = (real)v;
    }
    HIPCHK(hipMemcpyAsync(dst + dch[d].fo, tmp.data(), n * sizeof(real), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (field == FGX_DENS || field == FGX_ENERGY) {
      D.ignore_active = 1;
      int r = launch("pressure_init", 2, k_pressure_init<real>, n_ctiles, FGX_TPB, 0, D);
      if (!r && D.adiabatic) r = launch("sound_speed", 3, k_sound_speed<real>, n_stiles, FGX_TPB, 0, D);
      D.ignore_active = 0;
      LCHK(r);
      HIPCHK(hipStreamSynchronize(stream));
    }
    return FGX_OK;
  }
  int read_ring(int d, int ring, int dtype, void* host, size_t count) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    const HostDisk& h = hd[d];
    const size_t n = std::min(count, (size_t)h.nrad + 1);
    std::vector<double> out(n, 0.0);
    real* src = nullptr;
    switch (ring) {
      case FGX_R_RADII: for (size_t i = 0; i < n; ++i) out[i] = (real)h.radii[i]; break;
      case FGX_R_RMED: src = D.rmed; break;
      case FGX_R_SURF: src = D.surf; break;
      case FGX_R_SIGMAMED: src = D.sigmamed; break;
      case FGX_R_ENERGYMED: src = D.energymed; break;
      case FGX_R_VISCOSITY: src = D.visc; break;
      case FGX_R_VMED: src = D.vmed; break;
      case FGX_R_CS: src = D.cs_ring; break;
      case FGX_R_NEWDT: src = D.newdt; break;
      case FGX_R_INVRMED: src = D.invrmed; break;
      case FGX_R_NSHIFT: {
        std::vector<int> tmp(n);
        HIPCHK(hipMemcpyAsync(tmp.data(), D.nshift + dch[d].ro, n * sizeof(int), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (size_t i = 0; i < n; ++i) out[i] = tmp[i];
        break;
      }
      default: return set_error(FGX_ERANGE, "unknown ring profile %d", ring);
    }
    if (src) {
      std::vector<real> tmp(n);
      HIPCHK(hipMemcpyAsync(tmp.data(), src + dch[d].ro, n * sizeof(real), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      for (size_t i = 0; i < n; ++i) out[i] = (double)tmp[i];
    }
    if (dtype == FGX_F32) for (size_t i = 0; i < n; ++i) ((float*)host)[i] = (float)out[i];
    else for (size_t i = 0; i < n; ++i) ((double*)host)[i] = out[i];
    return FGX_OK;
  }
  int get_state(int d, DiskState& s) {
    HIPCHK(hipMemcpyAsync(&s, D.st + d, sizeof(DiskState), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return FGX_OK;
  }
  int info(int d, fgx_disk_info* o) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    DiskState s;
    LCHK(get_state(d, s));
    std::memset(o, 0, sizeof *o);
    o->nrad = hd[d].nrad; o->nsec = hd[d].nsec; o->precision = sizeof(real) == 8 ? FGX_F64 : FGX_F32;
    o->physical_time = s.time; o->omega_frame = s.omega_frame; o->lost_mass = s.lost_mass;
    o->last_dt = s.dt; o->last_ncfl = s.ncfl; o->substeps = s.substeps; o->dt_steps = s.dt_steps;
    o->health = s.health;
    double m = 0;
    if (s.substeps > 0) LCHK(mdcp(d, &m));
    o->mdcp1 = m; o->exces_mdcp = m;  // mdcp is never set by the reference (Main.cu:341)
    return FGX_OK;
  }
  int planets(int d, fgx_planets* out) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    DiskState s;
    LCHK(get_state(d, s));
    *out = hd[d].sys;
    for (int k = 0; k < out->nb; ++k) {
      out->x[k] = s.x[k]; out->y[k] = s.y[k]; out->vx[k] = s.vx[k]; out->vy[k] = s.vy[k]; out->mass[k] = s.mass[k];
    }
    return FGX_OK;
  }
  int set_planets(int d, const fgx_planets* in) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    if (in->nb != hd[d].sys.nb) return set_error(FGX_EPARAM, "planet count differs from the disk's");
    DiskState s;
    LCHK(get_state(d, s));
    for (int k = 0; k < in->nb; ++k) {
      s.x[k] = in->x[k]; s.y[k] = in->y[k]; s.vx[k] = in->vx[k]; s.vy[k] = in->vy[k]; s.mass[k] = in->mass[k];
    }
    HIPCHK(hipMemcpyAsync(D.st + d, &s, sizeof(DiskState), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return FGX_OK;
  }
  int force(int d, double x, double y, double sm, double m, double* out44) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    const int nr = hd[d].nrad;
    LCHK(launch("force_full", 1, k_force_full<real>, nr, FGX_TPB, 0, D, d, x, y, sm, m, D.rpart));
    std::vector<double> part((size_t)nr * 44);
    HIPCHK(hipMemcpyAsync(part.data(), D.rpart, part.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int q = 0; q < 44; ++q) {
      double s = 0;
      for (int i = 0; i < nr; ++i) s += part[(size_t)i * 44 + q];
      out44[q] = s;
    }
    return FGX_OK;
  }
  // the same 44 values with no allocation at all: the partials and their
  // ring-order sums live in the static device scratch, the totals are copied
  // straight into the caller's array.  Calls that overlap on one GPU share
  // that scratch and race.
  int force_noalloc(int d, double x, double y, double sm, double m, double* out44) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    const int nr = hd[d].nrad;
    if (nr > FGX_NOALLOC_FORCE_RINGS)
      return set_error(FGX_ERANGE, "fgx_disk_force_noalloc: NRAD=%d exceeds its scratch of %d rings", nr,
                       FGX_NOALLOC_FORCE_RINGS);
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != device) HIPCHK(hipSetDevice(device));
    double* sp = nullptr;
    int r = noalloc_scratch(device, &sp);
    if (!r) {
      // every partial the sum reads (rings 0..nr-1) is written by this call,
      // so nothing is cleared: the scratch keeps the last call's contents
      double* tot = sp + (size_t)FGX_NOALLOC_FORCE_RINGS * 44;
      hipLaunchKernelGGL(k_force_full<real>, dim3(nr), dim3(FGX_TPB), 0, stream, D, d, x, y, sm, m, sp);
      hipLaunchKernelGGL(k_force_sum, dim3(1), dim3(64), 0, stream, (const double*)sp, nr, tot);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(out44, tot, 44 * sizeof(double), hipMemcpyDeviceToHost, stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e != hipSuccess) r = set_error(FGX_EDEVICE, "fgx_disk_force_noalloc: %s", hipGetErrorString(e));
    }
    if (cur != device) hipSetDevice(cur);
    return r;
  }
  int mdcp(int d, double* out) override {
    if (hd[d].sys.nb == 0) { *out = 0; return FGX_OK; }
    DiskState s;
    LCHK(get_state(d, s));
    const int nr = hd[d].nrad;
    LCHK(launch("mdcp", 1, k_mdcp<real>, nr, FGX_TPB, 0, D, d, (double)(float)s.x[0], (double)(float)s.y[0],
                hd[d].hill_radius));
    std::vector<double> part(nr);
    HIPCHK(hipMemcpyAsync(part.data(), D.rpart, nr * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    double a = 0;
    for (double v : part) a += v;
    *out = a;
    return FGX_OK;
  }
  int profile_enable(int on) override {
    if (!on && prof_on) { hipStreamSynchronize(stream); prof_collect(); }
    prof_on = on != 0;
    return FGX_OK;
  }
  int profile_get(int k, const char** name, double* ms, long long* n, double* bytes) override {
    if (k < 0 || k >= (int)prof.size()) { if (name) *name = nullptr; return FGX_ERANGE; }
    if (name) *name = prof[k].name.c_str();
    if (ms) *ms = prof[k].ms;
    if (n) *n = prof[k].n;
    if (bytes) *bytes = prof[k].n ? prof[k].bytes / prof[k].n : 0;
    return FGX_OK;
  }
  int profile_reset() override { prof.clear(); return FGX_OK; }
  // ----------------------------------------------------------- checkpoints
  // A disk's evolving state in the engine's own representation (vtheta as
  // its residual, the FP64 planets and clocks of the state record, the ring
  // profiles that carry over between substeps), so that a restore into a disk
  // of the same parameters continues bit-identically.  Between DTs every
  // disk's velocities are in the current buffer pair.
  std::vector<real*> ckpt_fields() {
    std::vector<real*> f = {D.dens, D.vrad, D.vtheta};
    if (D.adiabatic) f.push_back(D.energy);
    if (D.advect_label) f.push_back(D.label);
    return f;
  }
  int checkpoint(int d, const char* path) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    if (!path) return set_error(FGX_EPARAM, "null path");
    const HostDisk& h = hd[d];
    const size_t nf = (size_t)(h.nrad + 1) * h.nsec, nrr = (size_t)h.nrad + 1;
    const std::vector<real*> fields = ckpt_fields();
    real* rings[2] = {D.csmean, D.visc};
    CkptHeader hdr{};
    std::memcpy(hdr.magic, "FGXCKPT3", 8);
    hdr.real_bytes = (int32_t)sizeof(real); hdr.nrad = h.nrad; hdr.nsec = h.nsec;
    hdr.adiabatic = D.adiabatic; hdr.nfields = (int32_t)fields.size(); hdr.nrings = 2;
    hdr.state_bytes = (uint32_t)sizeof(DiskState);
    std::vector<char> buf(sizeof hdr + sizeof(DiskState) + (2 * nrr + fields.size() * nf) * sizeof(real));
    char* q = buf.data();
    std::memcpy(q, &hdr, sizeof hdr); q += sizeof hdr;
    DiskState st;
    LCHK(get_state(d, st));
    std::memcpy(q, &st, sizeof st); q += sizeof st;
    for (real* r : rings) {
      HIPCHK(hipMemcpyAsync(q, r + dch[d].ro, nrr * sizeof(real), hipMemcpyDeviceToHost, stream));
      q += nrr * sizeof(real);
    }
    for (real* f : fields) {
      HIPCHK(hipMemcpyAsync(q, f + dch[d].fo, nf * sizeof(real), hipMemcpyDeviceToHost, stream));
      q += nf * sizeof(real);
    }
    HIPCHK(hipStreamSynchronize(stream));
    const uint64_t sum = fnv1a64(buf.data(), buf.size());
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return set_error(FGX_EIO, "cannot write checkpoint '%s'", path);
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size() && std::fwrite(&sum, 8, 1, fp) == 1;
    if (std::fclose(fp) != 0 || !ok) return set_error(FGX_EIO, "short write of checkpoint '%s'", path);
    return FGX_OK;
  }
  int restore(int d, const char* path) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    if (!path) return set_error(FGX_EPARAM, "null path");
    const HostDisk& h = hd[d];
    const size_t nf = (size_t)(h.nrad + 1) * h.nsec, nrr = (size_t)h.nrad + 1;
    const std::vector<real*> fields = ckpt_fields();
    real* rings[2] = {D.csmean, D.visc};
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return set_error(FGX_EIO, "cannot open checkpoint '%s'", path);
    std::vector<char> buf;
    char chunk[1 << 16];
    for (size_t n; (n = std::fread(chunk, 1, sizeof chunk, fp)) > 0;) buf.insert(buf.end(), chunk, chunk + n);
    std::fclose(fp);
    const size_t want = sizeof(CkptHeader) + sizeof(DiskState) + (2 * nrr + fields.size() * nf) * sizeof(real) + 8;
    CkptHeader hdr{};
    if (buf.size() >= sizeof hdr) std::memcpy(&hdr, buf.data(), sizeof hdr);
    if (buf.size() < sizeof hdr || std::memcmp(hdr.magic, "FGXCKPT3", 8) != 0)
      return set_error(FGX_EIO, "'%s' is not a checkpoint", path);
    if (hdr.real_bytes != (int32_t)sizeof(real) || hdr.nrad != h.nrad || hdr.nsec != h.nsec ||
        hdr.adiabatic != D.adiabatic || hdr.nfields != (int32_t)fields.size() || hdr.nrings != 2 ||
        hdr.state_bytes != (uint32_t)sizeof(DiskState))
      return set_error(FGX_EPARAM, "checkpoint '%s' (%d x %d, %d-byte fields, adiabatic %d) does not fit disk %d "
                       "(%d x %d, %d-byte fields, adiabatic %d)", path, hdr.nrad, hdr.nsec, hdr.real_bytes,
                       hdr.adiabatic, d, h.nrad, h.nsec, (int)sizeof(real), D.adiabatic);
    if (buf.size() != want) return set_error(FGX_EIO, "checkpoint '%s' has %zu bytes, expected %zu", path, buf.size(), want);
    uint64_t sum;
    std::memcpy(&sum, buf.data() + buf.size() - 8, 8);
    if (sum != fnv1a64(buf.data(), buf.size() - 8)) return set_error(FGX_EIO, "checkpoint '%s' fails its checksum", path);
    const char* q = buf.data() + sizeof hdr;
    DiskState st;
    std::memcpy(&st, q, sizeof st); q += sizeof st;
    st.vpar = D.vpar;   // the velocities go into this batch's current buffer pair
    HIPCHK(hipMemcpyAsync(D.st + d, &st, sizeof st, hipMemcpyHostToDevice, stream));
    // the health flags travel with the state: a disk saved unhealthy stays frozen
    HIPCHK(hipMemcpyAsync(D.health + d, &st.health, sizeof(int), hipMemcpyHostToDevice, stream));
    health_h[d] = st.health;
    frozen_h[d] = (char)(st.frozen != 0);
    for (real* r : rings) {
      HIPCHK(hipMemcpyAsync(r + dch[d].ro, q, nrr * sizeof(real), hipMemcpyHostToDevice, stream));
      q += nrr * sizeof(real);
    }
    for (real* f : fields) {
      HIPCHK(hipMemcpyAsync(f + dch[d].fo, q, nf * sizeof(real), hipMemcpyHostToDevice, stream));
      q += nf * sizeof(real);
    }
    HIPCHK(hipStreamSynchronize(stream));
    // fields derived from the state (pressure, and the adiabatic sound speed)
    LCHK(derive_fields(d, false));
    return FGX_OK;
  }
  // the fields derived from disk d's state alone (its pressure, adiabatic
  // sound speed and, with `sg`, self-gravity field); other disks untouched
  int derive_fields(int d, bool sg) {
    D.ignore_active = 1;
    D.only_disk = d;
    int r = launch("pressure_init", 2, k_pressure_init<real>, n_ctiles, FGX_TPB, 0, D);
    if (!r && D.adiabatic) r = launch("sound_speed", 3, k_sound_speed<real>, n_stiles, FGX_TPB, 0, D);
    if (!r && sg && D.self_gravity) r = self_gravity();
    D.ignore_active = 0;
    D.only_disk = -1;
    LCHK(r);
    HIPCHK(hipStreamSynchronize(stream));
    return FGX_OK;
  }
  // the fields a snapshot holds: Σ, vr, vθ and, adiabatic, e, as stored
  // (residuals of the ring profiles that fgx_batch_read_ring returns)
  std::vector<real*> snap_fields() {
    std::vector<real*> v{D.dens, D.vrad, D.vtheta};
    if (D.adiabatic) v.push_back(D.energy);
    return v;
  }
  // queue a copy of disk d's fields into its pinned buffer behind the work
  // already on the stream; the host does not wait
  int snapshot_queue(int d) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    if (snap_host.empty()) { snap_host.assign(nd, nullptr); snap_ev.assign(nd, nullptr); snap_pending.assign(nd, 0); }
    if (snap_pending[d]) HIPCHK(hipEventSynchronize(snap_ev[d]));   // the last copy, long done
    const std::vector<real*> fs = snap_fields();
    const size_t nf = (size_t)(hd[d].nrad + 1) * hd[d].nsec;
    if (!snap_host[d]) {
      HIPCHK(hipHostMalloc(&snap_host[d], fs.size() * nf * sizeof(real)));
      HIPCHK(hipEventCreateWithFlags(&snap_ev[d], hipEventDisableTiming));
    }
    for (size_t q = 0; q < fs.size(); ++q)
      HIPCHK(hipMemcpyAsync((real*)snap_host[d] + q * nf, fs[q] + dch[d].fo, nf * sizeof(real), hipMemcpyDeviceToHost,
                            stream));
    HIPCHK(hipEventRecord(snap_ev[d], stream));
    snap_pending[d] = 1;
    return FGX_OK;
  }
  // wait for disk d's queued copy (had = whether one was queued)
  int snapshot_wait(int d, int* had) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    *had = 0;
    if (snap_pending.empty() || !snap_pending[d]) return FGX_OK;
    HIPCHK(hipEventSynchronize(snap_ev[d]));
    snap_pending[d] = 0;
    *had = 1;
    return FGX_OK;
  }
  // slot d takes a new disk of the same grid and batch modes, set up as
  // fgx_batch_create would set it up; no allocation, the other disks keep
  // their state
  int reset_disk(int d, const fgx_params& p, const fgx_planets& pl) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    LCHK(params_validate(&p));
    if (p.nrad != hd[d].nrad || p.nsec != hd[d].nsec)
      return set_error(FGX_EPARAM, "slot %d holds a %d x %d grid, not %d x %d", d, hd[d].nrad, hd[d].nsec, p.nrad,
                       p.nsec);
    if (!batch_modes_equal(p, hd[0].P)) return set_error(FGX_EPARAM, "the disk differs from the batch in a mode flag");
    if (p.corotating && pl.nb == 0) return set_error(FGX_EPARAM, "a corotating frame needs a planet");
    HostDisk h;
    LCHK(setup_grid_and_density(p, pl, h));
    std::vector<double> khr, kht, axisg;
    if (p.self_gravity) LCHK(sg_host_setup(h, khr, kht, axisg));
    setup_velocities(h, axisg);
    const DiskConst& o = dch[d];
    DiskConst c;
    std::memset(&c, 0, sizeof c);
    c.nrad = o.nrad; c.nsec = o.nsec; c.fo = o.fo; c.ro = o.ro; c.so = o.so; c.rt0 = o.rt0;
    c.sgw_o = o.sgw_o; c.sgk_o = o.sgk_o; c.tw_ns = o.tw_ns; c.tw_nr = o.tw_nr; c.nf_ns = o.nf_ns; c.nf_nr = o.nf_nr;
    std::memcpy(c.fac_ns, o.fac_ns, sizeof c.fac_ns);
    std::memcpy(c.fac_nr, o.fac_nr, sizeof c.fac_nr);
    fill_const(h, c);
    // the launch sequence the batch was set up with must cover the disk
    bool viscous = false;
    for (int i = 0; i < h.nrad; ++i) viscous |= h.visc[i] != 0.0;
    if ((c.accretes && !accretion) || (c.outer_source && !outer_source) || (c.visc_dynamic && !visc_dynamic) ||
        ((viscous || c.visc_dynamic) && !D.viscous))
      return set_error(FGX_EPARAM, "the disk needs a pass (accretion, outer source or viscosity) the batch was set up "
                                   "without");
    hd[d] = std::move(h);
    dch[d] = c;
    HIPCHK(hipMemcpyAsync((void*)(D.dc + d), &dch[d], sizeof(DiskConst), hipMemcpyHostToDevice, stream));
    if (p.self_gravity) {
      HIPCHK(hipMemcpyAsync((void*)(D.sgkr + c.sgk_o), khr.data(), khr.size() * sizeof(double), hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipMemcpyAsync((void*)(D.sgkt + c.sgk_o), kht.data(), kht.size() * sizeof(double), hipMemcpyHostToDevice,
                            stream));
    }
    DiskState st = initial_state(d);
    st.vpar = D.vpar;   // the fields go into this batch's current buffer pair
    static const int zero = 0;
    HIPCHK(hipMemcpyAsync((void*)(D.st + d), &st, sizeof st, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(D.health + d, &zero, sizeof(int), hipMemcpyHostToDevice, stream));
    LCHK(clear_share(d));
    HIPCHK(hipStreamSynchronize(stream));
    health_h[d] = 0;
    frozen_h[d] = 0;
    if (!snap_pending.empty()) snap_pending[d] = 0;   // a copy of the slot's last disk has landed
    LCHK(upload_rings_disk(d));
    LCHK(upload_fields_disk(d));
    return derive_fields(d, true);
  }
  int freeze(int d, int on) override {
    if (d < 0 || d >= nd) return set_error(FGX_ERANGE, "disk index %d out of range", d);
    const int v = on ? 1 : 0;
    frozen_h[d] = (char)v;
    HIPCHK(hipMemcpyAsync((char*)D.st + (size_t)d * sizeof(DiskState) + offsetof(DiskState, frozen), &v, sizeof(int),
                          hipMemcpyHostToDevice, stream));
    if (!on) {
      // thawing also clears the disk's health flags (its fields may have been rewritten)
      static const int zero = 0;
      HIPCHK(hipMemcpyAsync((char*)D.st + (size_t)d * sizeof(DiskState) + offsetof(DiskState, health), &zero,
                            sizeof(int), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(D.health + d, &zero, sizeof(int), hipMemcpyHostToDevice, stream));
      health_h[d] = 0;
    }
    HIPCHK(hipStreamSynchronize(stream));
    return FGX_OK;
  }
};

}  // namespace fgx

// ============================================================== C ABI
struct fgx_batch {
  std::unique_ptr<fgx::BatchBase> impl;
};

using fgx::set_error;

static int make_batch(const std::vector<fgx_params>& ps, const std::vector<fgx_planets>& pls, int precision,
                      fgx_batch** out, const fgx_tiling* tl = nullptr) {
  if (!out) return set_error(FGX_EPARAM, "null output handle");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_error(FGX_EDEVICE, "no HIP device available");
  std::unique_ptr<fgx_batch> b(new fgx_batch());
  int r;
  if (precision == FGX_F64) {
    auto* impl = new fgx::Batch<double>();
    b->impl.reset(impl);
    r = impl->create(ps, pls, tl);
  } else if (precision == FGX_F32) {
    auto* impl = new fgx::Batch<float>();
    b->impl.reset(impl);
    r = impl->create(ps, pls, tl);
  } else {
    return set_error(FGX_EPARAM, "precision must be FGX_F32 or FGX_F64");
  }
  if (r) return r;
  *out = b.release();
  return FGX_OK;
}

extern "C" {

int fgx_disk_create(const fgx_params* p, const fgx_planets* pl, int precision, fgx_disk** out) {
  if (!p || !pl) return set_error(FGX_EPARAM, "null parameters");
  return make_batch({*p}, {*pl}, precision, out);
}
int fgx_batch_create(int n, const fgx_params* p, const fgx_planets* pl, int precision, fgx_batch** out) {
  if (n < 1 || !p || !pl) return set_error(FGX_EPARAM, "bad batch arguments");
  return make_batch(std::vector<fgx_params>(p, p + n), std::vector<fgx_planets>(pl, pl + n), precision, out);
}
int fgx_batch_create_tiled(int n, const fgx_params* p, const fgx_planets* pl, int precision, const fgx_tiling* t,
                           fgx_batch** out) {
  if (n < 1 || !p || !pl) return set_error(FGX_EPARAM, "bad batch arguments");
  return make_batch(std::vector<fgx_params>(p, p + n), std::vector<fgx_planets>(pl, pl + n), precision, out, t);
}
int fgx_batch_tiling(fgx_batch* b, fgx_tiling* out) {
  if (!b || !out) return set_error(FGX_EPARAM, "null argument");
  return b->impl->tiling(out);
}
int fgx_disk_step(fgx_disk* d, int n) {
  if (!d) return set_error(FGX_EPARAM, "null handle");
  if (n < 0) return set_error(FGX_EPARAM, "negative step count");
  return d->impl->step(n);
}
int fgx_batch_step(fgx_batch* b, int n) { return fgx_disk_step(b, n); }
int fgx_disk_destroy(fgx_disk* d) { delete d; return FGX_OK; }
int fgx_batch_destroy(fgx_batch* b) { delete b; return FGX_OK; }
int fgx_batch_freeze(fgx_batch* b, int disk, int frozen) {
  if (!b) return set_error(FGX_EPARAM, "null handle");
  return b->impl->freeze(disk, frozen);
}
int fgx_batch_checkpoint(fgx_batch* b, int disk, const char* path) {
  if (!b) return set_error(FGX_EPARAM, "null handle");
  return b->impl->checkpoint(disk, path);
}
int fgx_batch_restore(fgx_batch* b, int disk, const char* path) {
  if (!b) return set_error(FGX_EPARAM, "null handle");
  return b->impl->restore(disk, path);
}
int fgx_batch_reset_disk(fgx_batch* b, int disk, const fgx_params* p, const fgx_planets* pl) {
  if (!b || !p || !pl) return set_error(FGX_EPARAM, "null argument");
  return b->impl->reset_disk(disk, *p, *pl);
}
int fgx_batch_size(fgx_batch* b, int* n) {
  if (!b || !n) return set_error(FGX_EPARAM, "null argument");
  *n = b->impl->size();
  return FGX_OK;
}
int fgx_disk_info_get(fgx_disk* d, fgx_disk_info* info) {
  if (!d || !info) return set_error(FGX_EPARAM, "null argument");
  return d->impl->info(0, info);
}
int fgx_batch_info(fgx_batch* b, int disk, fgx_disk_info* info) {
  if (!b || !info) return set_error(FGX_EPARAM, "null argument");
  return b->impl->info(disk, info);
}
int fgx_batch_planets(fgx_batch* b, int disk, fgx_planets* out) {
  if (!b || !out) return set_error(FGX_EPARAM, "null argument");
  return b->impl->planets(disk, out);
}
int fgx_disk_read_field(fgx_disk* d, int field, int dtype, void* host, size_t count) {
  if (!d || !host) return set_error(FGX_EPARAM, "null argument");
  return d->impl->read_field(0, field, dtype, host, count);
}
int fgx_batch_read_field(fgx_batch* b, int disk, int field, int dtype, void* host, size_t count) {
  if (!b || !host) return set_error(FGX_EPARAM, "null argument");
  return b->impl->read_field(disk, field, dtype, host, count);
}
int fgx_disk_write_field(fgx_disk* d, int field, int dtype, const void* host, size_t count) {
  if (!d || !host) return set_error(FGX_EPARAM, "null argument");
  return d->impl->write_field(0, field, dtype, host, count);
}
int fgx_batch_write_field(fgx_batch* b, int disk, int field, int dtype, const void* host, size_t count) {
  if (!b || !host) return set_error(FGX_EPARAM, "null argument");
  return b->impl->write_field(disk, field, dtype, host, count);
}
int fgx_disk_read_ring(fgx_disk* d, int ring, int dtype, void* host, size_t count) {
  if (!d || !host) return set_error(FGX_EPARAM, "null argument");
  return d->impl->read_ring(0, ring, dtype, host, count);
}
int fgx_disk_planets(fgx_disk* d, fgx_planets* out) {
  if (!d || !out) return set_error(FGX_EPARAM, "null argument");
  return d->impl->planets(0, out);
}
int fgx_disk_set_planets(fgx_disk* d, const fgx_planets* pl) {
  if (!d || !pl) return set_error(FGX_EPARAM, "null argument");
  return d->impl->set_planets(0, pl);
}
int fgx_disk_force(fgx_disk* d, double x, double y, double sm, double m, double out[44]) {
  if (!d || !out) return set_error(FGX_EPARAM, "null argument");
  return d->impl->force(0, x, y, sm, m, out);
}
int fgx_disk_force_noalloc(fgx_disk* d, double x, double y, double sm, double m, double out[44]) {
  if (!d || !out) return set_error(FGX_EPARAM, "null argument");
  return d->impl->force_noalloc(0, x, y, sm, m, out);
}
int fgx_profile_enable(fgx_batch* b, int enable) {
  if (!b) return set_error(FGX_EPARAM, "null handle");
  return b->impl->profile_enable(enable);
}
int fgx_profile_get(fgx_batch* b, int k, const char** name, double* total_ms, long long* launches,
                    double* bytes_per_launch) {
  if (!b) return set_error(FGX_EPARAM, "null handle");
  return b->impl->profile_get(k, name, total_ms, launches, bytes_per_launch);
}
int fgx_profile_reset(fgx_batch* b) {
  if (!b) return set_error(FGX_EPARAM, "null handle");
  return b->impl->profile_reset();
}
int fgx_device_count(int* n) {
  if (!n) return set_error(FGX_EPARAM, "null argument");
  *n = 0;
  if (hipGetDeviceCount(n) != hipSuccess) *n = 0;
  return FGX_OK;
}
int fgx_footprint_bytes(const fgx_params* p, int precision, size_t* bytes) {
  if (!p || !bytes) return set_error(FGX_EPARAM, "null argument");
  const size_t F = (size_t)(p->nrad + 1) * p->nsec, R = (size_t)p->nrad + 1;
  const size_t rs = precision == FGX_F64 ? 8 : 4;
  // self-gravity: transform rows (nrad*nsec), two kernel transforms
  // ((2*nrad)*nsec), all complex FP64, and the two acceleration fields
  const size_t sgb = p->self_gravity ? (size_t)p->nrad * p->nsec * 16 * 5 + F * rs * 2 : 0;
  const int nfields = p->adiabatic ? 23 : 17;
  // fields, ring tables (working precision and the FP64 set), sector tables
  *bytes = F * rs * nfields + F * 8 /* FP64 potential */ + R * rs * 21 + R * 8 * 10 + (size_t)p->nsec * (rs * 3 + 8 * 2) +
           R * 8 * 2 * FGX_NFORCE_T + sgb + 4096;
  return FGX_OK;
}

}  // extern "C"

// internal accessors for the output writers (output.cpp)
namespace fgx {
const HostDisk* batch_host(fgx_batch* b, int d) { return &b->impl->host(d); }
int batch_mdcp(fgx_batch* b, int d, double* out) { return b->impl->mdcp(d, out); }
int batch_snapshot_queue(fgx_batch* b, int d) { return b->impl->snapshot_queue(d); }
int batch_snapshot_wait(fgx_batch* b, int d, int* had) { return b->impl->snapshot_wait(d, had); }
int device_name(char* out, size_t n) {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return set_error(FGX_EDEVICE, "no HIP device available");
  std::snprintf(out, n, "%s:%s:%d", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return FGX_OK;
}
}  // namespace fgx
