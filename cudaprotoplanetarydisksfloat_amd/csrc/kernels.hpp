// kernels.hpp -- HIP kernels of the AlgoGas step for gfx950 (CDNA4).
//
// Every kernel serves a whole batch of disks: the launch grid walks a tile
// table (CellTile = one ring x 256 sectors, RingTile = one ring) built on the
// host; the block -> tile mapping keeps consecutive tiles on one XCD so ring
// neighbours share that XCD's L2.  Fields are ring-major with sectors
// contiguous, so a wave64 reads 64 consecutive sectors of a ring: all global
// traffic is coalesced.  `real` is float (the reference's precision) or
// double (verification mode).  Per-disk scalars and planets are FP64.
//
// Each kernel names the reference kernels whose arithmetic it carries.
#pragma once
#include <hip/hip_runtime.h>

#include "fgx_internal.h"

#define FGX_TPB 256
#define FGX_SUMMARY_HIST 64   // substeps ahead the active-cell forecast covers
#define FGX_THETA_ARRAYS 8

namespace fgx {

template <class real>
struct Dev {
  // state and work fields, (nrad+1)*nsec per disk
  real *dens, *vrad, *vtheta, *energy, *cs;
  real *vradint, *vthetaint, *taurr, *taupp, *taurp, *divv, *qplus, *energyint, *pressure;
  real *t_dens, *t_rmp, *t_rmm, *t_tmp, *t_tmm, *t_e;
  real *label, *t_lab;   // advected label and its extensive form after the radial pass
  real *work, *sg_accr, *sg_acct;
  // ring tables, nrad+1 per disk
  real *rinf, *rsup, *rmed, *surf, *invsurf, *invrmed, *invrinf, *invdiffrmed, *invdiffrsup;
  real *dxtheta, *supp_torque, *sigmamed, *energymed, *coolingtime, *qplusmed, *visc;
  real *cs_ring, *csmean, *vmed, *newdt;
  int* nshift;
  // sector tables, nsec per disk
  real *cosj, *sinj, *lost;
  // FP64 ring/sector tables for the terms whose FP32 rounding would dominate:
  // the reference azimuthal velocity profile (vtheta and the angular momenta
  // are stored as FP32 residuals from it), the cell geometry of the
  // potential, and Substep1's ring-profile terms: acc0 the radial acceleration
  // of the ring profiles themselves at interface i, u0 their interface
  // azimuthal velocity, gq0 their pressure difference over the density sum
  const double *vref, *acc0, *u0, *gq0, *rinf_d, *invrinf_d, *rmed_d, *invrmed_d, *invdiffrmed_d, *cosj_d, *sinj_d, *cs2_d, *sigmamed_d,
      *energymed_d;
  // EvanescentBoundary's per-ring terms of a disk whose aspect ratio does not
  // change in time: the damping weight (damping^2 * 10) and the static part
  // of the Keplerian relaxation target of vtheta (evanescent batches only)
  double *ev_d2, *ev_vt0s;
  double* vmoy;  // ring-mean vtheta of the CFL condition
  double* phi;   // potential of the planets and the indirect term, FP64 field
  // self-gravity: transform work rows, kernel transforms, twiddles, ring
  // mean of the radial SG acceleration, column-pair tiles (d, k)
  cxd* sgw;
  const cxd *sgkr, *sgkt, *sgtw;
  double* axisg;
  const RingTile* sg_ctiles;
  const DiskConst* dc;
  DiskState* st;
  double* fpart;       // per ring tile: 2*FGX_NFORCE_T force partial sums
  double* rpart;       // per ring tile: generic partial sums (44 max)
  int* summary;        // [0] active disks, [1] max ncfl
  int* health;         // per disk: FGX_HEALTH_* flags raised by the CFL and force pass
  const CellTile* ctiles;
  const RingTile* rtiles;
  const CellTile* stiles;   // strips (d, first ring, sector block) of strip_rings rings
  int strip_rings;
  int max_nsec, ndisk;
  // batch-wide mode flags
  int adiabatic, fast_transport, open_inner, nonreflecting, evanescent, self_gravity;
  int cooling, indirect, exclude_hill, fix_taurp, fix_pp, fix_sgwrap, sloppy, corotating;
  int guiding, forced_circular, roche, write_pressure, ignore_active, viscous;
  int only_disk;      // with ignore_active: -1 every disk, else that disk alone
  int advect_label;   // LABELADVECTION with StubbedPhysics
  // which of the buffer pairs (vrad, vtheta) / (vradint, vthetaint) the
  // batch's pointers have been swapped to since set-up, modulo 2 (the fused
  // source pass writes the new velocities into the second pair)
  int vpar;
};

// ---- small helpers ------------------------------------------------------
// Data that no thread writes during a kernel (ring and sector tables, the
// per-disk constants, the per-substep disk state in the kernels that only read
// it) is read through the constant address space.  Uniform reads of it then
// use the scalar cache even after the kernel's own stores; through a plain
// pointer the compiler cannot rule out that a store changed it, issues vector
// loads instead, and has to drain every outstanding load and store to wait
// for them.
// (The host pass of hipcc also parses kernel bodies; there the qualifier is
// dropped.)
#if defined(__HIP_DEVICE_COMPILE__)
#define FGX_AS4 __attribute__((address_space(4)))
#else
#define FGX_AS4
#endif
template <class T>
__device__ __forceinline__ const FGX_AS4 T& cref(const T* p) { return *(const FGX_AS4 T*)p; }
template <class T>
__device__ __forceinline__ const FGX_AS4 T* cptr(const T* p) { return (const FGX_AS4 T*)p; }
#define RT(name) cptr(D.name)

// quotients of FP32 field values as products with the hardware reciprocal
// (1 ulp), exact division in FP64
template <class real>
__device__ __forceinline__ real rcp_of(real x) {
  if constexpr (sizeof(real) == 4) return __builtin_amdgcn_rcpf(x);
  else return (real)1 / x;
}
template <class real>
__device__ __forceinline__ real qdiv(real a, real b) {
  if constexpr (sizeof(real) == 4) return a * __builtin_amdgcn_rcpf(b);
  else return a / b;
}
// The density field holds Sigma - S0(ring), S0 the field-precision value of
// SigmaMed (the ring profile of InitGasDensity, Theo.cu:13-30): the density's
// rounding is then relative to its departure from the profile.  S0 is exactly
// representable in the field precision, so the whole value of a residual x is
// S0 + x with one rounding (or exactly, in FP64 arithmetic).
template <class real>
__device__ __forceinline__ real s_whole(const Dev<real>& D, int rb, int i, real x) {
  return RT(sigmamed)[rb + i] + x;
}
template <class real>
__device__ __forceinline__ double s_whole_d(const Dev<real>& D, int rb, int i, real x) {
  return (double)RT(sigmamed)[rb + i] + (double)x;
}
// The energy field holds e - EnergyMed(ring): a residual of the FP64 ring
// profile (Theo.cu:33-51), so that its rounding, and that of the
// intermediate energies, is relative to the departure from the profile and
// not to e itself.  The whole value of a residual x on ring i:
template <class real>
__device__ __forceinline__ real e_whole(const Dev<real>& D, int rb, int i, real x) {
  return (real)(RT(energymed_d)[rb + i] + (double)x);
}
// DivisePolarGrid's work of the density itself, density / (density + 1e-20)
// (Kernels.cu:1127-1134): exactly 1 wherever the 1e-20 is below half an ulp of
// the density (in FP32 above 1e-12, in FP64 above 3e-4), else the quotient
// (FP32 forms that rare quotient with the hardware reciprocal)
template <class real>
__device__ __forceinline__ real dens_work(real sg) {
  if constexpr (sizeof(real) == 4) return sg >= 1e-12f ? 1.0f : sg * __builtin_amdgcn_rcpf(sg + 1e-20f);
  else return sg >= 3e-4 ? 1.0 : sg / (sg + 1e-20);
}
// Substep1's radial velocity update (Kernels.cu:52-68): the acceleration
// -gradP/Sigma - gradPhi + vt^2/Rinf is the ring profile's own, acc0 (formed
// at set-up in extended precision, where the O(1) gravity and centrifugal
// terms cancel), plus the change that the departures from the profiles make:
//   dgp  = ((p_i - p_i-1) - gq0 (sr_i + sr_i-1)) 2 / (Sigma_i + Sigma_i-1) / dRmed
//   dphi = (phi_i - phi_i-1) / dRmed
//   sv   = the mean of the four vtheta residuals + Rinf (OmegaFrame - OmegaFrame at set-up)
//   acc  = acc0 - dgp - dphi + sv (2 u0 + sv) / Rinf
// with p the pressure departure, sr the density residual and gq0, u0 the
// profile's pressure-gradient coefficient and interface velocity.  Every term
// is a small departure, so the field precision carries it: FP32 forms it in
// FP32 (the ring constants rounded once from their FP64 tables, the quotient
// with a reciprocal), FP64 in FP64.  The potential difference is taken in
// FP64.
template <class real>
struct S1Ring { real cs2, idr, acc0, u0, gq0, invrinf, rinfdom; };
template <class real>
__device__ __forceinline__ S1Ring<real> s1_ring(const Dev<real>& D, int r, double dom) {
  return S1Ring<real>{(real)D.cs2_d[r], (real)D.invdiffrmed_d[r], (real)D.acc0[r], (real)D.u0[r], (real)D.gq0[r],
                      (real)D.invrinf_d[r], (real)(D.rinf_d[r] * dom)};
}
template <class real>
__device__ __forceinline__ real substep1_vr(real vr, real p, real pm, real sr, real srm, real sg, real sgm, double phi,
                                            double phim, real vsum, const S1Ring<real>& q, real dt) {
  const real dgp = qdiv((p - pm - q.gq0 * (sr + srm)) * (real)2, sg + sgm) * q.idr;
  const real dphi = (real)(phi - phim) * q.idr;
  const real sv = vsum / (real)4 + q.rinfdom;
  return vr + dt * (q.acc0 - dgp - dphi + sv * ((real)2 * q.u0 + sv) * q.invrinf);
}
// Update_sgvelocityKernel's two updates (Kernels.cu:1024-1041) in FP64: the
// radial acceleration interpolated to the inner interface of ring i, and the
// azimuthal one averaged over sectors j-1 and j.  Contraction is off, so that
// the separate update kernel and the fused source pass round them alike.
__device__ __forceinline__ double sg_vrad_new(double v, double dt, double ri, double rm1, double rm, double ar,
                                              double ar1, double idr) {
#pragma clang fp contract(off)
  return v + dt * ((ri - rm1) * ar + (rm - ri) * ar1) * idr;
}
__device__ __forceinline__ double sg_vtheta_new(double v, double dt, double at, double am) {
#pragma clang fp contract(off)
  return v + 0.5 * dt * (at + am);
}
// Substep1's azimuthal pressure gradient (times dx_theta): the difference of
// the pressure departures over the mean density (FP32 with a reciprocal)
template <class real>
__device__ __forceinline__ real azim_gradp(real p, real pl, real sig, real sigl) {
  return qdiv((p - pl) * (real)2, sig + sigl);
}
// block -> tile: the hardware hands blocks to the 8 XCDs round-robin (block
// b to XCD b & 7).  Groups of FGX_XCD_GROUP consecutive tiles (neighbouring
// strips of a disk, which share halo rings in L2) stay on one XCD, and the
// groups are dealt to the XCDs in turn, so the tiles of any one disk spread
// over the whole chip (when only a few disks are still active, a contiguous
// share per XCD would leave seven of the eight idle).  n is a multiple of
// 8 * FGX_XCD_GROUP.
#define FGX_XCD_GROUP 8
#define FGX_TILE_PAD (8 * FGX_XCD_GROUP)
__device__ __forceinline__ int xcd_tile(int b, int n) {
  (void)n;
  const int x = b & 7, slot = b >> 3;
  return ((slot / FGX_XCD_GROUP) * 8 + x) * FGX_XCD_GROUP + slot % FGX_XCD_GROUP;
}
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <class T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}
// block of FGX_TPB threads; every thread gets the result
template <class T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  T r = 0;
#pragma unroll
  for (int k = 0; k < FGX_TPB / 64; ++k) r += red[k];
  return r;
}
template <class T>
__device__ __forceinline__ T block_min(T v, T* red) {
  v = wave_min(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  T r = red[0];
#pragma unroll
  for (int k = 1; k < FGX_TPB / 64; ++k) r = red[k] < r ? red[k] : r;
  return r;
}
__device__ __forceinline__ double aspect_ratio_dev(const DiskConst& c, double r, double t) {
  double h = c.aspectratio;
  double rmin = c.transitionradius - c.transitionwidth * c.aspectratio;
  double rmax = c.transitionradius + c.transitionwidth * c.aspectratio;
  double scale = 1.0 + t * c.lambdadoubling;
  rmin *= scale; rmax *= scale;
  if (r < rmin) h *= c.transitionratio;
  if (r >= rmin && r <= rmax) h *= exp((rmax - r) / (rmax - rmin) * log(c.transitionratio));
  return h;
}
__device__ __forceinline__ double fviscosity_dev(const DiskConst& c, int alpha, double r, double csm, double t) {
  double nu = c.viscosity;
  if (alpha) nu = c.alphaviscosity * csm * csm * pow(r, 1.5);
  double rmin = c.cavityradius - c.cavitywidth * c.aspectratio;
  double rmax = c.cavityradius + c.cavitywidth * c.aspectratio;
  double scale = 1.0 + t * c.lambdadoubling;
  rmin *= scale; rmax *= scale;
  if (r < rmin) nu *= c.cavityratio;
  if (r >= rmin && r <= rmax) nu *= exp((rmax - r) / (rmax - rmin) * log(c.cavityratio));
  return nu;
}

// between two substeps (before k_substep_setup): a disk still taking
// substeps in this DT, i.e. active, and its pending substep (if any) has not
// reached the DT (that substep's end-of-substep update is applied where the
// next substep starts).  Within a substep, after k_substep_setup, the active
// flag alone says it.
__device__ __forceinline__ bool disk_running(const DiskConst& c, const DiskState& s) {
  return s.active && !(s.pending && !(s.dtemp < 0.999999999 * c.DT));
}

// the disks a kernel runs on: in a step the active ones; with ignore_active
// (set-up, derived reads) every disk, or the one only_disk names
template <class real>
__device__ __forceinline__ bool skip_disk(const Dev<real>& D, const DiskState& s, int d) {
  return D.ignore_active ? (D.only_disk >= 0 && d != D.only_disk) : !s.active;
}

#define FGX_CELL_PROLOGUE                                                   \
  const CellTile t = D.ctiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)];   \
  if (t.d < 0) return;                                                      \
  const FGX_AS4 DiskConst& c = cref(D.dc + t.d);                           \
  const FGX_AS4 DiskState& s = cref(D.st + t.d);                            \
  if (skip_disk(D, s, t.d)) return;                                \
  const int ns = c.nsec, nr = c.nrad;                                       \
  const int i = t.i;                                                        \
  const int j = t.jb * FGX_TPB + (int)threadIdx.x;                          \
  if (j >= ns) return;                                                      \
  const long long fb = c.fo;                                                \
  const int rb = c.ro;                                                      \
  (void)nr; (void)rb; (void)s;

#define FGX_RING_PROLOGUE                                                   \
  const RingTile t = D.rtiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)];   \
  if (t.d < 0) return;                                                      \
  const FGX_AS4 DiskConst& c = cref(D.dc + t.d);                           \
  const FGX_AS4 DiskState& s = cref(D.st + t.d);                            \
  if (skip_disk(D, s, t.d)) return;                                \
  const int ns = c.nsec, nr = c.nrad;                                       \
  const int i = t.i;                                                        \
  const long long fb = c.fo;                                                \
  const int rb = c.ro;                                                      \
  (void)nr; (void)rb; (void)fb;

// a strip tile: one thread per sector walks rings [t.i, t.i + strip_rings)
#define FGX_STRIP_BEGIN                                                     \
  const CellTile t = D.stiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)];   \
  if (t.d < 0) return;                                                      \
  const FGX_AS4 DiskConst& c = cref(D.dc + t.d);                           \
  const FGX_AS4 DiskState& s = cref(D.st + t.d);                            \
  if (skip_disk(D, s, t.d)) return;                                \
  const int ns = c.nsec, nr = c.nrad;                                       \
  const int j = t.jb * FGX_TPB + (int)threadIdx.x;                          \
  if (j >= ns) return;                                                      \
  const long long fb = c.fo;                                                \
  const int rb = c.ro;                                                      \
  const int i_end = (t.i + D.strip_rings < nr) ? t.i + D.strip_rings : nr;  \
  (void)rb; (void)s; (void)c;                                               \
  for (int i = t.i; i < i_end; ++i)

// Field element (ring ii, sector jj) of the tile's disk as a buffer access:
// the disk's field base goes into a buffer descriptor (scalar registers), the
// ring offset into the scalar offset and the sector offset into a 32-bit
// vector offset, so walking rings costs one scalar add and no vector address
// arithmetic.  Offsets are bytes; a disk's field is far below 2 GiB.
template <class T>
struct FieldRef {
  __amdgpu_buffer_rsrc_t r;
  int soff;
  unsigned voff;
  __device__ __forceinline__ T get() const {
    if constexpr (sizeof(T) == 4) {
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    } else {
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    }
  }
  __device__ __forceinline__ void put(T v) const {
    if constexpr (sizeof(T) == 4) {
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
    } else {
      using u2 = unsigned __attribute__((ext_vector_type(2)));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 0);
    }
  }
  __device__ __forceinline__ operator T() const { return get(); }
  __device__ __forceinline__ const FieldRef& operator=(T v) const { put(v); return *this; }
  __device__ __forceinline__ const FieldRef& operator=(const FieldRef& o) const { put(o.get()); return *this; }
  __device__ __forceinline__ const FieldRef& operator+=(T v) const { put(get() + v); return *this; }
  __device__ __forceinline__ const FieldRef& operator-=(T v) const { put(get() - v); return *this; }
};
// raw buffer descriptor word 3 for gfx9-family parts: 32-bit data format, no
// swizzle, no index stride
#define FGX_BUF_WORD3 0x00020000
template <class T>
__device__ __forceinline__ FieldRef<T> field_at(T* base, int soff, unsigned voff) {
  return FieldRef<T>{__builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, 0x7fffffff, FGX_BUF_WORD3), soff, voff};
}
#define AT(a, ii, jj) field_at((a) + fb, (int)((ii) * ns * (int)sizeof(*(a))), (unsigned)(jj) * (unsigned)sizeof(*(a)))

// potential of the planets and the indirect term at cell (i,j) in FP64:
// FillForcesArraysKernel (Kernels.cu:1322-1349) for all planets in one
// evaluation; the star's -1/r is handled per ring (its radial difference is
// part of the ring-profile acceleration acc0, its azimuthal difference
// vanishes)
// 1/sqrt(x) in FP64: the hardware estimate (good to ~2^-29) refined by
// Newton steps: two for FP64 fields (relative error below 1e-16), one for
// FP32 fields, whose potential differences need far less (below 2^-50)
template <class real>
__device__ __forceinline__ double rsqrt_d(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  if constexpr (sizeof(real) == 8) {
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
  }
  return fma(0.5 * y, e, y);
}

template <class real>
__device__ __forceinline__ double potential_planets_xy(const Dev<real>& D, const DiskConst& c, const DiskState& s,
                                                       double x, double y) {
  double pot = 0;
  for (int k = 0; k < c.nb; ++k) {
    const double xp = s.pot_x[k], yp = s.pot_y[k], mp = s.pot_m[k];
    const double dist = (x - xp) * (x - xp) + (y - yp) * (y - yp);
    double p = -mp * rsqrt_d<real>(dist + s.pot_sm2[k]);
    if (D.indirect) p += mp * s.pot_inv3[k] * (x * xp + y * yp);
    pot += p;
  }
  pot -= s.ind_x * x + s.ind_y * y;
  return pot;
}
template <class real>
__device__ __forceinline__ double potential_planets_at(const Dev<real>& D, const DiskConst& c, const DiskState& s,
                                                       double rmed, int j) {
  return potential_planets_xy(D, c, s, rmed * RT(cosj_d)[c.so + j], rmed * RT(sinj_d)[c.so + j]);
}

// ============================================================ CFL, forces
// ConditionCFLKernel1D + ConditionCFLKernel2D1 + per-ring min of
// ConditionCFLKernel2D2 (Kernels.cu:1363-1422) and ComputeForceKernel
// (Kernels.cu:804-839) for the substep's targets, both over the state at the
// start of the substep: one wave per ring, ring sums and minima by shuffles.
// A disk whose planets accrete needs its forces after the accretion, which
// needs the substep's dt: mode 0 skips its forces, mode 1 forms only those.
// Waves per ring of the CFL and force pass: a ring of more than 768 sectors
// takes one wave per 512 sectors (at most 8), each walking every wp-th group
// of 64 sectors, so a wide ring's passes run wp times shorter.  The count is
// the disk's own, so a disk's sums are the same alone and in any batch.
__device__ __host__ __forceinline__ int cfl_waves(int nsec) {
  return nsec <= 768 ? 1 : ((nsec + 511) / 512 < 8 ? (nsec + 511) / 512 : 8);
}
// WPR: the batch's largest count (a block of 64*WPR threads); the waves past
// the ring's own count end at once.  The waves' sums and extrema meet in LDS
// in wave order; with one wave every value is the one-wave pass's.
template <class real, int WPR>
__global__ void __launch_bounds__(64 * WPR) k_cfl_force(Dev<real> D, int mode) {
  const RingTile t = D.rtiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)];
  if (t.d < 0) return;
  const FGX_AS4 DiskConst& c = cref(D.dc + t.d);
  const FGX_AS4 DiskState& s = cref(D.st + t.d);
  // mode 0 runs before k_substep_setup, where the disk's previous substep may
  // still be pending; mode 1 runs after it, within the substep
  if (!D.ignore_active && !(mode == 0 ? disk_running(c, s) : s.active)) return;
  if (mode == 1 && !c.accretes) return;
  const int ns = c.nsec, i = t.i, rb = c.ro;
  const long long fb = c.fo;
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const int wp = WPR > 1 ? cfl_waves(ns) : 1;
  if (w >= wp) return;   // a barrier below waits only for the waves still running
  const int j0 = (int)threadIdx.x, js = 64 * wp;   // this thread's first sector and stride
  __shared__ double xred[WPR];
  // the block's sum (or max with MAX) of one value per wave, in wave order;
  // every thread receives it
  auto across = [&](double v, bool max) -> double {
    if constexpr (WPR == 1) {
      return v;
    } else {
      __syncthreads();   // the previous use of xred is over
      if (lane == 0) xred[w] = v;
      __syncthreads();
      double r = xred[0];
      for (int q = 1; q < wp; ++q) r = max ? (xred[q] > r ? xred[q] : r) : r + xred[q];
      return r;
    }
  };
  // disk forces on the targets (k = 0 or k = 10 Hill cut), per ring in FP64
  // numeric health of the density at the start of the substep (mode 0): a
  // value that is not positive, or not a number, flags the disk; the first
  // force target's sweep reads every cell anyway, else a sweep of its own
  const real s0 = RT(sigmamed)[rb + i];
  const int nt = s.ntarget;
  bool dbad = false;
  if (mode == 0 && (c.accretes || nt == 0)) {
#pragma unroll 4
    for (int j = j0; j < ns; j += js) dbad |= !(s0 + (real)AT(D.dens, i, j) > (real)0);
  }
  if (mode == 1 || !c.accretes) {
    const real rmed = RT(rmed)[rb + i], surf = RT(surf)[rb + i];
    double* out = D.fpart + (size_t)(c.rt0 + i) * (2 * FGX_NFORCE_T);
    for (int q = 0; q < nt; ++q) {
      const real tx = (real)s.tx[q], ty = (real)s.ty[q], sm2 = (real)s.tsm2[q], irh2 = (real)(1.0 / s.trh2[q]);
      const int kk = s.tk[q];
      double fx = 0, fy = 0;
#pragma unroll 4
      for (int j = j0; j < ns; j += js) {
        const real cm = surf * (s0 + (real)AT(D.dens, i, j));
        if (q == 0) dbad |= !(cm > (real)0);
        const real dx = rmed * RT(cosj)[c.so + j] - tx, dy = rmed * RT(sinj)[c.so + j] - ty;
        const real d2 = dx * dx + dy * dy;
        const real dist2 = d2 + sm2;
        // dist2^(-3/2) from the reciprocal square root
        const real rq = rsqrt(dist2);
        const real inv3 = rq * rq * rq;
        const real cut = kk ? (real)1 - exp(-d2 * irh2) : (real)1;
        fx += (double)(cm * dx * inv3 * cut);
        fy += (double)(cm * dy * inv3 * cut);
      }
      fx = across(wave_sum(fx), false);
      fy = across(wave_sum(fy), false);
      if (threadIdx.x == 0) { out[2 * q] = fx; out[2 * q + 1] = fy; }
    }
  }
  if (mode == 0 && __any(dbad) && lane == 0) atomicOr(D.health + t.d, FGX_HEALTH_DENSITY);
  if (mode == 1 || c.substeps_forced > 0 || (D.sloppy && s.sub_in_dt > 0)) return;
  const double vr0 = RT(vref)[rb + i];
  double acc = 0;
#pragma unroll 4
  for (int j = j0; j < ns; j += js) acc += (double)(real)AT(D.vtheta, i, j);
  const double mres = across(wave_sum(acc), false) / (double)ns;   // ring mean of the residual
  if (threadIdx.x == 0) D.vmoy[rb + i] = vr0 + mres;
  if (i == 0) return;
  const real dxrad = RT(rsup)[rb + i] - RT(rinf)[rb + i];
  const real dxt = RT(dxtheta)[rb + i];
  const real mind = dxrad < dxt ? dxrad : dxt;
  // the per-cell quotients by ring lengths as products with the ring's
  // reciprocals (the substep count only changes if DT/dt sits within an ulp
  // of an integer, and dt itself only depends on that count)
  const real imind = (real)1 / mind, idxr = (real)1 / dxrad, idxt = (real)1 / dxt;
  const real csr = D.adiabatic ? (real)0 : D.cs_ring[rb + i];
  // the per-cell limit CFLSECURITY / sqrt(x) only falls as x grows (sqrt and
  // the quotient are both correctly rounded, hence monotone), so the ring's
  // least limit is the limit of the ring's largest x, bit for bit; NaN x are
  // skipped by both forms, and the 1e30 start survives an empty maximum
  real xmax = 0;
  bool nbad = false;
#pragma unroll 4
  for (int j = j0; j < ns; j += js) {
    const real rs = AT(D.vtheta, i, j);
    const real vr = AT(D.vrad, i, j);
    const real cs = D.adiabatic ? (real)AT(D.cs, i, j) : csr;
    // Vresidual: vtheta - Vmoy (FARGO) or vtheta (standard)
    const real vres = D.fast_transport ? (real)((double)rs - mres) : (real)(vr0 + (double)rs);
    const real i1 = cs * imind;
    const real i2 = fabs(vr) * idxr;
    const real i3 = fabs(vres) * idxt;
    real dvr = (real)AT(D.vrad, i + 1, j) - vr;
    real dvt = (real)AT(D.vtheta, i, j + 1 == ns ? 0 : j + 1) - rs;
    dvr = (dvr >= 0) ? (real)1e-10 : -dvr;
    dvt = (dvt >= 0) ? (real)1e-10 : -dvt;
    real i4 = dvr * idxr;
    const real i4b = dvt * idxt;
    i4 = i4b > i4 ? i4b : i4;
    i4 *= (real)(4.0 * kCVNR * kCVNR);
    const real x = i1 * i1 + i2 * i2 + i3 * i3 + i4 * i4;
    nbad |= x != x;
    xmax = x > xmax ? x : xmax;
  }
  if (__any(nbad) && lane == 0) atomicOr(D.health + t.d, FGX_HEALTH_NAN);
  xmax = -wave_min(-xmax);
  xmax = (real)across((double)xmax, true);
  if (threadIdx.x == 0) {
    const real dt2 = (real)kCFLSECURITY / sqrt(xmax);
    D.newdt[rb + i] = dt2 < (real)1e30 ? dt2 : (real)1e30;
  }
}

// force targets of the next substep: the star (Pframeforce.cu via
// SourceEuler.cu:247) and the planets that feel the disk (Pframeforce.cu:68-92)
// at their current positions; set where a disk's substep starts its DT and
// after each of its substeps, so the CFL and force pass can run first
template <class real>
__device__ void set_force_targets(const Dev<real>& D, const DiskConst& c, DiskState& s) {
  s.tx[0] = 0; s.ty[0] = 0; s.tsm2[0] = 0; s.trh2[0] = 1e-30; s.tk[0] = 0;
  int nt = 1;
  const double tt = s.time - s.time0;
  for (int k = 0; k < c.nb; ++k) {
    if (!c.feel_disk[k]) continue;
    const double m = s.mass[k], x = s.x[k], y = s.y[k];
    const double r = sqrt(x * x + y * y);
    const double sm = D.roche ? r * pow(m / 3., 1. / 3.) * c.rochesmoothing
                              : c.thicknesssmoothing * aspect_ratio_dev(c, r, tt) * pow(r, 1.0 + c.flaringindex);
    const double rh = pow(m / 3., 1. / 3.) * r + 1e-15;
    s.tx[nt] = x; s.ty[nt] = y; s.tsm2[nt] = sm * sm; s.trh2[nt] = rh * rh;
    s.tk[nt] = D.exclude_hill ? 10 : 0;
    nt++;
  }
  s.ntarget = nt;
}

// The end of a substep (SourceEuler.cu:297-298) and the DT loop condition
// (SourceEuler.cu:230), LostMass accumulation (TransportEuler.cu:71).  It is
// applied where the disk's next substep starts (k_substep_setup) or at the
// end of the DT (k_finish_dt), not in a launch of its own: until then the
// substep is pending, and a pending disk whose dtemp has reached DT counts as
// done (disk_running).
template <class real>
__device__ void end_substep_state(const Dev<real>& D, const DiskConst& c, DiskState& s, double lost) {
  s.lost_mass += lost;
  s.time += s.dt;
  s.substeps++;
  s.sub_in_dt++;
  if (!(s.dtemp < 0.999999999 * c.DT)) {
    s.active = 0;
    s.dt_steps++;
    s.vpar = D.vpar;   // later swaps of the DT leave this disk's velocities behind
  }
  s.pending = 0;
}
// the lost mass of the pending substep: ring 0's outflow, summed by one wave
template <class real>
__device__ double lost_sum_wave(const Dev<real>& D, const DiskConst& c) {
  double lost = 0;
  if (D.open_inner) {
    for (int j = threadIdx.x; j < c.nsec; j += 64) lost += (double)D.lost[c.so + j];
    lost = wave_sum(lost);
  }
  return lost;
}

// ============================================================ substep begin
// MassTaper, the CFL count of ConditionCFL (SourceEuler.cu:494-536), dt,
// GetPsysInfo(MARK) and the force targets of the substep: one block per disk
template <class real>
__device__ double get_psys_info_dev(const Dev<real>& D, const DiskConst& c, DiskState& s, int action);

template <class real>
__device__ void substep_begin_block(const Dev<real>& D, int d, DiskState& s, real* redr) {
  const FGX_AS4 DiskConst& c = cref(D.dc + d);
  const int nr = c.nrad, rb = c.ro;
  const bool need_cfl = !(c.substeps_forced > 0) && !(D.sloppy && s.sub_in_dt > 0);
  real newdt = (real)1e30;
  if (need_cfl) {
    real mn = (real)1e30;
    for (int i = 1 + threadIdx.x; i < nr; i += FGX_TPB) {
      const real v = D.newdt[rb + i];
      mn = v < mn ? v : mn;
    }
    for (int i = threadIdx.x; i < nr - 1; i += FGX_TPB) {
      const double a = D.vmoy[rb + i] / RT(rmed_d)[rb + i] - D.vmoy[rb + i + 1] / RT(rmed_d)[rb + i + 1];
      const real v = (real)(2.0 * kPI * kCFLSECURITY / (double)c.nsec / fabs(a));
      mn = v < mn ? v : mn;
    }
    newdt = block_min(mn, redr);
  }
  if (threadIdx.x != 0) return;   // the rest is serial
  // MassTaper (SourceEuler.cu:231-232); the sine only while it ramps up
  const double mt = s.time / (c.masstaper * 2.0 * M_PI);
  if (mt > 1.0) {
    s.masstaper = 1.0;
  } else {
    const double sn = sin(mt * M_PI / 2.0);
    s.masstaper = sn * sn;
  }
  const double DT = c.DT;
  // numeric health: the flags of this substep's CFL and force pass, a time
  // step that is not positive (NaN included), a count above the cap.  An
  // unhealthy disk is frozen here, before its count or dt is used, and keeps
  // its fields as they are; its velocities stay in the current buffer pair.
  int h = D.health[d];
  if (need_cfl && !(newdt > (real)0)) h |= FGX_HEALTH_TIMESTEP;
  int n = 1;
  if (!h) {
    if (c.substeps_forced > 0) {
      n = D.sloppy ? c.substeps_forced : c.substeps_forced - s.sub_in_dt;
      if (n < 1) n = 1;
    } else if (need_cfl) {
      const real dT = (real)(DT - s.dtemp);
      real nd = newdt;
      if (dT < nd) nd = dT;
      const real q = ceil(dT / nd);
      if (!(q <= (real)c.max_substeps)) h |= FGX_HEALTH_SUBSTEPS;
      else n = (int)q;
    } else n = s.ncfl;
  }
  if (h) {
    s.health |= h;
    D.health[d] = s.health;
    s.active = 0;
    s.frozen = 1;
    s.ncfl = 0;
    s.vpar = D.vpar;
    return;
  }
  // SloppyCFL keeps the dt of the first substep (SourceEuler.cu:224-229)
  if (!(D.sloppy && s.sub_in_dt > 0)) s.dt = D.sloppy ? DT / (double)n : (DT - s.dtemp) / (double)n;
  s.ncfl = n;
  s.dtemp += s.dt;
  if (D.corotating) get_psys_info_dev(D, c, s, 1);
  const double tt = s.time - s.time0;
  // ApplySubKeplerianBoundary values (SideEuler.cu:148-166), FP64; the
  // powers of the edge radii are the disk's constants
  const double r0 = RT(rmed_d)[rb], rn = RT(rmed_d)[rb + nr - 1];
  const double fac = 1.0 + c.sigmaslope - 2.0 * c.flaringindex;
  const double h0 = aspect_ratio_dev(c, r0, tt), hn = aspect_ratio_dev(c, rn, tt);
  s.vkep_in = 1.0 / r0 * (1.0 - fac * h0 * h0 * c.pow_in_2f);
  s.vkep_out = 1.0 / rn * (1.0 - fac * hn * hn * c.pow_out_2f);
  // EvanescentBoundary constants (SideEuler.cu:125-130), formed on the host
  s.ev_tin = c.ev_tin;
  s.ev_tout = c.ev_tout;
  s.ev_drmin = c.ev_drmin;
  s.ev_drmax = c.ev_drmax;
}

// ================================================================= planets
template <class real>
__device__ double get_psys_info_dev(const Dev<real>& D, const DiskConst& c, DiskState& s, int action) {
  // GetPsysInfo (Psys.cu:162-232)
  double x = s.x[0], y = s.y[0], vx = s.vx[0], vy = s.vy[0];
  double xc = x, yc = y;
  const double m = s.mass[0] + 1.;
  const double h = x * vy - y * vx;
  const double dd = sqrt(x * x + y * y);
  const double Ax = x * vy * vy - y * vx * vy - m * x / dd;
  const double Ay = y * vx * vx - x * vx * vy - m * y / dd;
  const double e = sqrt(Ax * Ax + Ay * Ay) / m;
  const double a = h * h / m / (1. - e * e);
  const double arg = (e == 0.0) ? 1.0 : (1.0 - dd / a) / e;
  double E;
  if (fabs(arg) >= 1.0) E = kPI * (1. - arg / fabs(arg)) / 2.;
  else E = acos((1.0 - dd / a) / e);
  if ((x * y * (vy * vy - vx * vx) + vx * vy * (x * x - y * y)) < 0) E = -E;
  const double M = E - e * sin(E);
  const double omega = sqrt(m / a / a / a);
  const double pa = atan2(Ay, Ax);
  if (D.guiding) { xc = a * cos(M + pa); yc = a * sin(M + pa); }
  if (e < 1e-8) { xc = x; yc = y; }
  if (action == 1) { s.mark_x = xc; s.mark_y = yc; return 0.; }
  if (action == 0) {
    const double d2 = sqrt(xc * xc + yc * yc);
    const double d1 = sqrt(s.mark_x * s.mark_x + s.mark_y * s.mark_y);
    const double cross = s.mark_x * yc - xc * s.mark_y;
    s.mark_x = xc; s.mark_y = yc;
    return asin(cross / (d1 * d2));
  }
  return omega;
}

// DerivMotionRK5 (RungeKutta.cu:9-46)
template <class real>
__device__ __forceinline__ void deriv_rk5_dev(const Dev<real>& D, const DiskConst& c, const double* mass, const double* q, double* der,
                              double* dist, int nb, double dt) {
  const double *x = q, *y = q + nb, *vx = q + 2 * nb, *vy = q + 3 * nb;
  for (int i = 0; i < nb; ++i) dist[i] = sqrt(x[i] * x[i] + y[i] * y[i]);
  for (int i = 0; i < nb; ++i) {
    der[i] = vx[i];
    der[i + nb] = vy[i];
    double ax = -1.0 / dist[i] / dist[i] / dist[i] * x[i];
    double ay = -1.0 / dist[i] / dist[i] / dist[i] * y[i];
    for (int j = 0; j < nb; ++j) {
      if (D.indirect) {
        ax -= mass[j] / dist[j] / dist[j] / dist[j] * x[j];
        ay -= mass[j] / dist[j] / dist[j] / dist[j] * y[j];
      }
      if (j != i && c.feel_others[i]) {
        double d3;
        if (D.fix_pp) {
          const double dd = sqrt((x[i] - x[j]) * (x[i] - x[j]) + (y[i] - y[j]) * (y[i] - y[j]));
          d3 = dd * dd * dd;
        } else d3 = dist[j] * dist[j] * dist[j];
        ax += mass[j] / d3 * (x[j] - x[i]);
        ay += mass[j] / d3 * (y[j] - y[i]);
      }
    }
    der[i + 2 * nb] = ax;
    der[i + 3 * nb] = ay;
  }
  for (int i = 0; i < 4 * nb; ++i) der[i] *= dt;
}

// The Cash-Karp stages of AdvanceSystemRK5 (RungeKutta.cu:55-74): q1 = q0
// advanced by dt.  With a compile-time planet count (NB > 0) every stage
// array is a register array and the six derivative evaluations and stage
// sums are straight-line code, with no LDS round trip between dependent
// steps (one planet: 21.5 -> 16.4 us per setup launch of the 512x1536 disk);
// NB = 0 takes the count at run time, with the stages in the LDS scratch
// `rk`.  Both do the same operations in the same order.
template <class real, int NB>
__device__ __forceinline__ void rk5_stages(const Dev<real>& D, const DiskConst& c, const DiskState& s, int nb,
                                           double dt, double* rk, double* q1out) {
  constexpr int S = NB > 0 ? 4 * NB : 1;
  double q0r[S], q1r[S], k1r[S], k2r[S], k3r[S], k4r[S], k5r[S], k6r[S], distr[NB > 0 ? NB : 1];
  const int L = 4 * FGX_MAX_PLANETS;
  double *q0 = NB ? q0r : rk, *q1 = NB ? q1r : rk + L, *k1 = NB ? k1r : rk + 2 * L, *k2 = NB ? k2r : rk + 3 * L;
  double *k3 = NB ? k3r : rk + 4 * L, *k4 = NB ? k4r : rk + 5 * L, *k5 = NB ? k5r : rk + 6 * L;
  double *k6 = NB ? k6r : rk + 7 * L, *dist = NB ? distr : rk + 8 * L;
  if (NB) nb = NB;
  for (int k = 0; k < nb; ++k) {
    q0[k] = s.x[k]; q0[k + nb] = s.y[k]; q0[k + 2 * nb] = s.vx[k]; q0[k + 3 * nb] = s.vy[k];
  }
  const int n4 = 4 * nb;
  deriv_rk5_dev(D, c, s.mass, q0, k1, dist, nb, dt);
  for (int m = 0; m < n4; ++m) q1[m] = q0[m] + 0.2 * k1[m];
  deriv_rk5_dev(D, c, s.mass, q1, k2, dist, nb, dt);
  for (int m = 0; m < n4; ++m) q1[m] = q0[m] + 0.075 * k1[m] + 0.225 * k2[m];
  deriv_rk5_dev(D, c, s.mass, q1, k3, dist, nb, dt);
  for (int m = 0; m < n4; ++m) q1[m] = q0[m] + 0.3 * k1[m] - 0.9 * k2[m] + 1.2 * k3[m];
  deriv_rk5_dev(D, c, s.mass, q1, k4, dist, nb, dt);
  for (int m = 0; m < n4; ++m)
    q1[m] = q0[m] - 11.0 / 54.0 * k1[m] + 2.5 * k2[m] - 70.0 / 27.0 * k3[m] + 35.0 / 27.0 * k4[m];
  deriv_rk5_dev(D, c, s.mass, q1, k5, dist, nb, dt);
  for (int m = 0; m < n4; ++m)
    q1[m] = q0[m] + 1631.0 / 55296.0 * k1[m] + 175.0 / 512.0 * k2[m] + 575.0 / 13824.0 * k3[m] +
            44275.0 / 110592.0 * k4[m] + 253.0 / 4096.0 * k5[m];
  deriv_rk5_dev(D, c, s.mass, q1, k6, dist, nb, dt);
  for (int m = 0; m < n4; ++m)
    q1out[m] = q0[m] + 37.0 / 378.0 * k1[m] + 250.0 / 621.0 * k3[m] + 125.0 / 594.0 * k4[m] + 512.0 / 1771.0 * k6[m];
}

// reduction of the ring partial forces, FillForcesArrays' potential
// parameters, AdvanceSystemFromDisk, AdvanceSystemRK5 (Cash-Karp,
// RungeKutta.cu:55-74), corotation and RotatePsys: one block per disk
template <class real>
__device__ void planets_block(const Dev<real>& D, int d, DiskState& s, double* redd, double* tot, double* rk) {
  const FGX_AS4 DiskConst& c = cref(D.dc + d);
  const int nt = s.ntarget;
  for (int q = 0; q < 2 * nt; ++q) {
    double a = 0;
    for (int i = threadIdx.x; i < c.nrad; i += FGX_TPB) a += D.fpart[(size_t)(c.rt0 + i) * (2 * FGX_NFORCE_T) + q];
    a = block_sum(a, redd);
    if (threadIdx.x == 0) tot[q] = a;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double dt = s.dt;
  const double tt = s.time - s.time0;
  // ComputeIndirectTerm (Pframeforce.cu:27-35)
  s.ind_x = D.indirect ? -tot[0] : 0.0;
  s.ind_y = D.indirect ? -tot[1] : 0.0;
  const int nb = c.nb;
  // FillForcesArrays parameters (Pframeforce.cu:49-61) at the current positions
  for (int k = 0; k < nb; ++k) {
    const double xp = s.x[k], yp = s.y[k], mp = s.mass[k] * s.masstaper;
    const double pd = sqrt(xp * xp + yp * yp);
    double sm;
    if (D.roche) {
      sm = pd * pow((1.0 / 3.0 * mp), 1.0 / 3.0) * c.rochesmoothing;
    } else {
      // pd^(1 + flaring) is pd itself for a flat disk (pow(x, 1) == x)
      const double pw = c.flaringindex == 0.0 ? pd : pow(pd, 1.0 + c.flaringindex);
      sm = c.thicknesssmoothing * aspect_ratio_dev(c, pd, tt) * pw;
    }
    s.pot_x[k] = xp; s.pot_y[k] = yp; s.pot_m[k] = mp;
    s.pot_inv3[k] = 1.0 / pd / pd / pd;
    s.pot_sm2[k] = sm * sm;
  }
  // AdvanceSystemFromDisk (Pframeforce.cu:68-92)
  int q = 1;
  for (int k = 0; k < nb; ++k) {
    if (!c.feel_disk[k]) continue;
    s.vx[k] += dt * tot[2 * q]; s.vy[k] += dt * tot[2 * q + 1];
    s.vx[k] += dt * s.ind_x; s.vy[k] += dt * s.ind_y;
    q++;
  }
  // AdvanceSystemRK5 (Pframeforce.cu:94-157)
  if (nb > 0) {
    // the advanced state: one planet runs the stages in registers, more in
    // the LDS scratch (rk: 8 x 4*FGX_MAX_PLANETS doubles plus FGX_MAX_PLANETS
    // distances); register stages for two or three planets raised the
    // kernel's registers enough to slow the batch's one-planet launches
    double* q1 = rk + 4 * FGX_MAX_PLANETS;
    if (!D.forced_circular) {
      if (nb == 1) rk5_stages<real, 1>(D, c, s, nb, dt, rk, q1);
      else rk5_stages<real, 0>(D, c, s, nb, dt, rk, q1);
    }
    for (int k = 1 - (s.time >= c.releasedate ? 1 : 0); k < nb; ++k) {
      if (!D.forced_circular) {
        s.x[k] = q1[k]; s.y[k] = q1[k + nb]; s.vx[k] = q1[k + 2 * nb]; s.vy[k] = q1[k + 3 * nb];
      } else {
        const double x = s.x[k], y = s.y[k], th = atan2(y, x), vx = s.vx[k], vy = s.vy[k];
        const double r = sqrt(x * x + y * y);
        const double om = (-y * vx + x * vy) / r / r, dth = om * dt;
        double sn, cs;
        sincos(th + dth, &sn, &cs);
        s.x[k] = r * cs; s.y[k] = r * sn;
        s.vx[k] = vx * cs - vy * sn;
        s.vy[k] = vx * sn + vy * cs;
      }
    }
    if (s.time < c.releasedate) {
      const double x = s.x[0], y = s.y[0], r = sqrt(x * x + y * y), th = atan2(y, x);
      const double rdot = (c.releaseradius - r) / (c.releasedate - s.time);
      const double om = sqrt((1. + s.mass[0]) / r / r / r);
      const double nrr = r + rdot * dt, den = r - nrr;
      const double dth = (den != 0.0) ? 2. * dt * r * om / den * (sqrt(r / nrr) - 1.) : om * dt;
      const double vx = rdot, vy = nrr * sqrt((1. + s.mass[0]) / nrr / nrr / nrr);
      s.x[0] = nrr * cos(dth + th); s.y[0] = nrr * sin(dth + th);
      s.vx[0] = vx * cos(dth + th) - vy * sin(dth + th);
      s.vy[0] = vx * sin(dth + th) + vx * cos(dth + th);  // as Pframeforce.cu:154
    }
  }
  // corotating frame (SourceEuler.cu:257-263)
  s.domega = 0;
  if (D.corotating && nb > 0) {
    const double onew = get_psys_info_dev(D, c, s, 0) / dt;
    s.domega = onew - s.omega_frame;
    s.omega_frame = onew;
  }
  // RotatePsys (Psys.cu:234-251), by -OmegaFrame*dt; a zero angle is the
  // identity (cos 0 = 1, sin 0 = 0 leave every product exact)
  const double ang = s.omega_frame * dt;
  if (ang == 0.0) return;
  double sa, ca;
  sincos(ang, &sa, &ca);
  for (int k = 0; k < nb; ++k) {
    double xt = s.x[k], yt = s.y[k];
    s.x[k] = xt * ca + yt * sa; s.y[k] = -xt * sa + yt * ca;
    xt = s.vx[k]; yt = s.vy[k];
    s.vx[k] = xt * ca + yt * sa; s.vy[k] = -xt * sa + yt * ca;
  }
}

// the per-disk serial part of a substep's start: the CFL substep count and dt,
// MassTaper, boundary constants (substep_begin_block), then the disk forces
// summed over rings, the potential's planet parameters, AdvanceSystemFromDisk,
// the RK5 N-body step and the frame (planets_block): one block per disk.
// part 0 runs both halves, 1 the first, 2 the second (accretion sits between)
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_substep_setup(Dev<real> D, int part) {
  const int d = blockIdx.x;
  if (!D.st[d].active) return;
  __shared__ real redr[FGX_TPB / 64];
  __shared__ double redd[FGX_TPB / 64];
  __shared__ double tot[2 * FGX_NFORCE_T];
  __shared__ double rk[8 * 4 * FGX_MAX_PLANETS + FGX_MAX_PLANETS];
  // the disk's state is staged in LDS: the serial part reads and writes it
  // field by field, and an LDS round trip is far shorter than a global one
  __shared__ DiskState ls;
  constexpr int NW = (int)(sizeof(DiskState) / sizeof(int));
  for (int q = threadIdx.x; q < NW; q += FGX_TPB) reinterpret_cast<int*>(&ls)[q] = reinterpret_cast<const int*>(D.st + d)[q];
  __syncthreads();
  const FGX_AS4 DiskConst& c = cref(D.dc + d);
  bool run = true;
  if (part != 2 && ls.pending) {
    // the end of the disk's previous substep
    if (threadIdx.x < 64) {
      const double lost = lost_sum_wave(D, c);
      if (threadIdx.x == 0) end_substep_state(D, c, ls, lost);
    }
    __syncthreads();
    run = ls.active;
  }
  if (run) {
    if (part != 2) {
      substep_begin_block(D, d, ls, redr);
      __syncthreads();
      run = ls.active;   // not when the disk was found unhealthy
      if (run && threadIdx.x == 0) ls.pending = 1;
    }
    if (run && part != 1) {
      __syncthreads();
      planets_block(D, d, ls, redd, tot, rk);
      // the force targets of the disk's next substep, at the time that
      // substep starts
      if (threadIdx.x == 0) {
        const double t0 = ls.time;
        ls.time = t0 + ls.dt;
        set_force_targets(D, c, ls);
        ls.time = t0;
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < NW; q += FGX_TPB) reinterpret_cast<int*>(D.st + d)[q] = reinterpret_cast<const int*>(&ls)[q];
}

// CorrectVthetaKernel (Kernels.cu:1352-1359)
// AccreteOntoPlanets with FARGO's kernel (Planet.cu:7-47 computes the Roche
// box and leaves the launch commented out; it runs here with StubbedPhysics):
// for each accreting planet in turn, the cells of its Roche box closer than
// 0.75 R_Roche lose facc/3 of their density and those closer than 0.45
// R_Roche a further 2 facc/3 (facc = dt * acc, W. Kley's scheme); the planet
// takes that mass and its momentum, the cell velocity formed at the cell
// centre from the face values.  The angle of the box is real (the reference
// declares it int, Planet.cu:9) and a box wider than the ring is cut at nsec
// sectors so that no cell is visited twice.  Afterwards the force targets
// follow the new masses.  One block per disk; FP64 per cell, FP64 block sums.
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_accrete(Dev<real> D) {
  const int d = blockIdx.x;
  const DiskConst& c = D.dc[d];
  DiskState& s = D.st[d];
  if (!s.active || !c.accretes) return;
  __shared__ double redd[FGX_TPB / 64];
  __shared__ int box[4];
  const int ns = c.nsec, nr = c.nrad, rb = c.ro;
  const long long fb = c.fo;
  for (int k = 0; k < c.nb; ++k) {
    if (!(c.acc[k] > 1e-10)) continue;
    const double facc = s.dt * c.acc[k], facc1 = 1.0 / 3.0 * facc, facc2 = 2.0 / 3.0 * facc;
    const double xp = s.x[k], yp = s.y[k], mp = s.mass[k];
    const double rp = sqrt(xp * xp + yp * yp);
    const double rroche = pow(1.0 / 3.0 * mp, 1.0 / 3.0) * rp;
    // the Roche box is formed from these: a planet that is not finite (or
    // whose box angle would overflow) takes nothing
    if (!(rp > 0 && rp < 1e30 && rroche < rp)) continue;
    if (threadIdx.x == 0) {
      int imin = 0, imax = nr - 1;
      while (imin < nr && (double)RT(rsup)[rb + imin] < rp - rroche) imin++;
      while (imax > 0 && (double)RT(rinf)[rb + imax] > rp + rroche) imax--;
      const double angle = atan2(yp, xp);
      const int jmin = (int)((double)ns / 2.0 / kPI * (angle - 2.0 * rroche / rp));
      int jmax = (int)((double)ns / 2.0 / kPI * (angle + 2.0 * rroche / rp));
      if (jmax - jmin + 1 > ns) jmax = jmin + ns - 1;
      box[0] = imin; box[1] = imax; box[2] = jmin; box[3] = jmax;
    }
    __syncthreads();
    const int imin = box[0], jmin = box[2], nj = box[3] - box[2] + 1;
    const int ncell = (box[1] >= imin && nj > 0) ? (box[1] - imin + 1) * nj : 0;
    double dm = 0, dpx = 0, dpy = 0;
    for (int q = threadIdx.x; q < ncell; q += FGX_TPB) {
      const int i = imin + q / nj;
      int jf = (jmin + q % nj) % ns;
      if (jf < 0) jf += ns;
      const int jp = (jf + 1 == ns) ? 0 : jf + 1;
      const real rm = RT(rmed)[rb + i];
      const double xc = (double)(rm * RT(cosj)[c.so + jf]), yc = (double)(rm * RT(sinj)[c.so + jf]);
      const double dx = xp - xc, dy = yp - yc;
      const double dist = sqrt(dx * dx + dy * dy);
      if (!(dist < 0.75 * rroche)) continue;
      const double rmd = (double)rm;
      // Vtheta is held as a residual of the ring's reference velocity
      const double vt = RT(vref)[rb + i] + 0.5 * ((double)(real)AT(D.vtheta, i, jf) + (double)(real)AT(D.vtheta, i, jp)) +
                        rmd * s.omega_frame;
      const double vr = 0.5 * ((double)(real)AT(D.vrad, i, jf) + (double)(real)AT(D.vrad, i + 1, jf));
      const double vx = (vr * xc - vt * yc) / rmd, vy = (vr * yc + vt * xc) / rmd;
      const double surf = (double)RT(surf)[rb + i];
      // the density field is a residual of S0 (rounded as a residual)
      const double s0 = (double)RT(sigmamed)[rb + i];
      double den = s0 + (double)(real)AT(D.dens, i, jf);
      double m = facc1 * den * surf;
      den = s0 + (double)(real)(den * (1.0 - facc1) - s0);
      dpx += m * vx; dpy += m * vy; dm += m;
      if (dist < 0.45 * rroche) {
        m = facc2 * den * surf;
        den = s0 + (double)(real)(den * (1.0 - facc2) - s0);
        dpx += m * vx; dpy += m * vy; dm += m;
      }
      AT(D.dens, i, jf) = (real)(den - s0);
    }
    // the sums also order this planet's density updates before the next one's
    dm = block_sum(dm, redd);
    dpx = block_sum(dpx, redd);
    dpy = block_sum(dpy, redd);
    if (threadIdx.x == 0) {
      const double px = mp * s.vx[k] + dpx, py = mp * s.vy[k] + dpy, m = mp + dm;
      if (c.feel_disk[k]) { s.vx[k] = px / m; s.vy[k] = py / m; }
      s.mass[k] = m;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) set_force_targets(D, c, s);
}

template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_correct_vtheta(Dev<real> D) {
  FGX_STRIP_BEGIN {
    AT(D.vtheta, i, j) -= (real)s.domega * RT(rmed)[rb + i];
  }
}

// ============================================================ viscosity nu
// FViscosity (Viscosity.cu:19-37) per ring; with update_csmean the ring mean
// of the sound speed is refreshed first (Make1Dprofile(1), Kernels.cu:740-756)
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_visc_profile(Dev<real> D, int update_csmean, int alpha) {
  FGX_RING_PROLOGUE
  __shared__ double redd[FGX_TPB / 64];
  if (update_csmean) {
    double acc = 0;
    if (D.adiabatic) {
      for (int j = threadIdx.x; j < ns; j += FGX_TPB) acc += (double)AT(D.cs, i, j);
      acc = block_sum(acc, redd) / (double)ns;
    } else acc = (double)D.cs_ring[rb + i];
    if (threadIdx.x == 0) D.csmean[rb + i] = (real)acc;
  }
  if (threadIdx.x == 0) {
    const double nu = fviscosity_dev(c, alpha, RT(rmed_d)[rb + i], (double)D.csmean[rb + i], s.time - s.time0);
    D.visc[rb + i] = (real)nu;
  }
}

// ============================================================ sound speed
// ComputeSoundSpeedKernel, adiabatic branch (Kernels.cu:190-191)
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_sound_speed(Dev<real> D) {
  FGX_STRIP_BEGIN {
    const real g = (real)c.adiabaticindex;
    AT(D.cs, i, j) = sqrt(g * (g - (real)1) * e_whole(D, rb, i, (real)AT(D.energy, i, j)) /
                          s_whole(D, rb, i, (real)AT(D.dens, i, j)));
  }
}

// ================================================================ Substep1
// Substep1KernelVrad + Substep1KernelVtheta (Kernels.cu:52-88), pressure
// (ComputePressureFieldKernel, Kernels.cu:196-208) and potential evaluated in
// place instead of stored.  The radial acceleration is a near-cancellation of
// O(1) centrifugal and gravity terms, so it is formed in FP64 from the FP32
// fields and FP64 ring tables; vthetaint is stored as a residual of vref.
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_substep1(Dev<real> D) {
  FGX_STRIP_BEGIN {
    const double dt = s.dt;
    const int jp = (j + 1 == ns) ? 0 : j + 1, jm = (j == 0) ? ns - 1 : j - 1;
    // the pressure's departure from the ring profile's, in the field
    // precision from the residual fields: (Sigma - S0) cs^2 with the ring's
    // isothermal cs^2, or (gamma - 1)(e - EnergyMed)
    const double g1 = c.adiabaticindex - 1.0;
    const real g1r = (real)g1;
    auto press = [&](int ii, int jj) -> real {
      if (D.adiabatic) return g1r * AT(D.energy, ii, jj);
      return AT(D.dens, ii, jj) * (real)RT(cs2_d)[rb + ii];
    };
    const real sr = AT(D.dens, i, j);
    const real sig = s_whole(D, rb, i, sr);
    const real p = press(i, j);
    const double phi = AT(D.phi, i, j);
    if (D.write_pressure)
      AT(D.pressure, i, j) = (real)((double)RT(sigmamed)[rb + i] * RT(cs2_d)[rb + i] + (double)p);
    // radial: the ring profiles' own acceleration acc0 plus the change the
    // departures from the profiles make (substep1_vr)
    if (i > 0) {
      const real srm = AT(D.dens, i - 1, j);
      const real sigm = s_whole(D, rb, i - 1, srm);
      const real pm = press(i - 1, j);
      const double phim = AT(D.phi, i - 1, j);
      const real vsum = AT(D.vtheta, i, j) + AT(D.vtheta, i, jp) + AT(D.vtheta, i - 1, j) + AT(D.vtheta, i - 1, jp);
      const S1Ring<real> q = s1_ring(D, rb + i, s.omega_frame - c.omega0);
      AT(D.vradint, i, j) = substep1_vr((real)AT(D.vrad, i, j), p, pm, sr, srm, sig, sigm, phi, phim, vsum, q, (real)dt);
    } else {
      AT(D.vradint, i, j) = 0;
    }
    // azimuthal (the star's potential is the same at j and j-1)
    const real invdxt = (real)1 / RT(dxtheta)[rb + i];
    const real sigl = s_whole(D, rb, i, (real)AT(D.dens, i, jm));
    const real pl = press(i, jm);
    const double phil = AT(D.phi, i, jm);
    const real gradp = azim_gradp(p, pl, sig, sigl) * invdxt;
    const real gradphi = (real)(phi - phil) * invdxt;
    real v = AT(D.vtheta, i, j) - (real)dt * (gradp + gradphi);
    v += (real)dt * RT(supp_torque)[rb + i];
    AT(D.vthetaint, i, j) = v;
  }
}

// ============================================================ viscous terms
// ViscousTermsKernelDRP + ViscousTermsKernelTAURP (Kernels.cu:843-880): the
// strain rates stay in registers, only the three stresses (and the velocity
// divergence when an energy equation needs it) are stored
template <class real>
struct Stress { real rr, pp, rp, divv, sig; };
// a value the compiler must take as it is: no product or sum on either side of
// it is fused with one on the other (as if it had been stored and read back),
// so that a kernel that keeps the stresses in registers rounds exactly like
// the pair of kernels that passes them through memory
template <class T>
__device__ __forceinline__ T opaque(T x) {
  asm volatile("" : "+v"(x));
  return x;
}
template <class real>
__device__ __forceinline__ Stress<real> opaque(Stress<real> t) {
  t.rr = opaque(t.rr); t.pp = opaque(t.pp); t.rp = opaque(t.rp); t.divv = opaque(t.divv); t.sig = opaque(t.sig);
  return t;
}
// the stresses and the velocity divergence of cell (i, j) from the velocities
// vr, vt and the density (sig: the cell's whole density)
template <class real>
__device__ __forceinline__ Stress<real> stress_at(const Dev<real>& D, const DiskConst& c, int rb, long long fb,
                                                  const real* __restrict__ vr, const real* __restrict__ vt, int i,
                                                  int j) {
  const int ns = c.nsec;
  const int jp = (j + 1 == ns) ? 0 : j + 1, jm = (j == 0) ? ns - 1 : j - 1;
  const real invdphi = (real)((double)ns / (2.0 * kPI));
  const real irm = RT(invrmed)[rb + i];
  const real vr0 = AT(vr, i, j), vr1 = AT(vr, i + 1, j);
  const real vt0 = AT(vt, i, j), vtp = AT(vt, i, jp);
  const real drr = (vr1 - vr0) * RT(invdiffrsup)[rb + i];
  const real dpp = (vtp - vt0) * invdphi * irm + (real)0.5 * (vr1 + vr0) * irm;
  real divv = (vr1 * RT(rsup)[rb + i] - vr0 * RT(rinf)[rb + i]) * RT(invdiffrsup)[rb + i] * irm;
  divv += (vtp - vt0) * invdphi * irm;
  const real nu = RT(visc)[rb + i];
  const real sig = s_whole(D, rb, i, (real)AT(D.dens, i, j));
  const real third = (real)(1.0 / 3.0);
  Stress<real> t;
  t.rr = (real)2 * sig * nu * (drr - third * divv);
  t.pp = (real)2 * sig * nu * (dpp - third * divv);
  t.divv = divv;
  t.sig = sig;
  if (i > 0) {
    // the angular velocity difference between rings is formed in FP64 from
    // the residuals and the reference profile
    const double w0 = (RT(vref)[rb + i] + (double)vt0) * RT(invrmed_d)[rb + i];
    const double w1 = (RT(vref)[rb + i - 1] + (double)AT(vt, i - 1, j)) * RT(invrmed_d)[rb + i - 1];
    const real drp = (real)0.5 * ((real)(RT(rinf_d)[rb + i] * (w0 - w1) * RT(invdiffrmed_d)[rb + i]) +
                                  (vr0 - AT(vr, i, jm)) * invdphi * RT(invrinf)[rb + i]);
    // corner density: column 0 of ring i as the reference indexes it
    // (Kernels.cu:876), or column j-1 with the quirk fix
    const real sc = s_whole(D, rb, i, D.fix_taurp ? (real)AT(D.dens, i, jm) : (real)AT(D.dens, i, 0));
    const real sm_ = s_whole(D, rb, i - 1, (real)AT(D.dens, i - 1, j)), sml = s_whole(D, rb, i - 1, (real)AT(D.dens, i - 1, jm));
    t.rp = (real)(2.0 * 0.25) * (sig + sm_ + sc + sml) * nu * drp;
  } else {
    t.rp = 0;
  }
  return t;
}
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_visc_tau(Dev<real> D, const real* __restrict__ vr,
                                                      const real* __restrict__ vt, int write_divv) {
  FGX_STRIP_BEGIN {
    const Stress<real> t = stress_at(D, c, rb, fb, vr, vt, i, j);
    AT(D.taurr, i, j) = t.rr;
    AT(D.taupp, i, j) = t.pp;
    if (write_divv) AT(D.divv, i, j) = t.divv;
    AT(D.taurp, i, j) = t.rp;
  }
}

// UpdateVelocitiesKernel (Kernels.cu:142-160) + ApplySubKeplerianBoundary
// (SideEuler.cu:148-170, Kernels.cu:1490-1503)
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_visc_update(Dev<real> D) {
  FGX_STRIP_BEGIN {
    const real dt = (real)s.dt;
    const int jp = (j + 1 == ns) ? 0 : j + 1, jm = (j == 0) ? ns - 1 : j - 1;
    const real invdphi = (real)((double)ns / (2.0 * kPI));
    if (D.viscous && i > 0 && i < nr - 1) {
      const real trp0 = AT(D.taurp, i, j), trp1 = AT(D.taurp, i + 1, j);
      AT(D.vthetaint, i, j) +=
          qdiv(dt * RT(invrmed)[rb + i] *
                   ((RT(rsup)[rb + i] * trp1 - RT(rinf)[rb + i] * trp0) * RT(invdiffrsup)[rb + i] +
                    (AT(D.taupp, i, j) - AT(D.taupp, i, jm)) * invdphi + (real)0.5 * (trp0 + trp1)),
               (real)0.5 * (s_whole(D, rb, i, (real)AT(D.dens, i, j)) + s_whole(D, rb, i, (real)AT(D.dens, i, jm))));
    }
    if (D.viscous && i > 0) {
      AT(D.vradint, i, j) +=
          qdiv(dt * RT(invrinf)[rb + i] *
                   ((RT(rmed)[rb + i] * AT(D.taurr, i, j) - RT(rmed)[rb + i - 1] * AT(D.taurr, i - 1, j)) *
                        RT(invdiffrmed)[rb + i] +
                    (AT(D.taurp, i, jp) - AT(D.taurp, i, j)) * invdphi -
                    (real)0.5 * (AT(D.taupp, i, j) + AT(D.taupp, i - 1, j))),
               (real)0.5 * (s_whole(D, rb, i, (real)AT(D.dens, i, j)) + s_whole(D, rb, i - 1, (real)AT(D.dens, i - 1, j))));
    }
    if (!D.evanescent) {
      // with self-gravity the Keplerian values include -r * <SG_Accr>_ring
      // (ApplySubKeplerianBoundary, SideEuler.cu:158-166)
      if (i == 0) {
        const double v = s.vkep_in - (D.self_gravity ? RT(rmed_d)[rb] * D.axisg[rb] : 0.0);
        AT(D.vthetaint, i, j) = (real)(sqrt(v) - RT(rmed_d)[rb] * s.omega_frame - RT(vref)[rb]);
      }
      if (i == nr - 1) {
        const double v = s.vkep_out - (D.self_gravity ? RT(rmed_d)[rb + nr - 1] * D.axisg[rb + nr - 1] : 0.0);
        AT(D.vthetaint, i, j) = (real)(sqrt(v) - RT(rmed_d)[rb + nr - 1] * s.omega_frame - RT(vref)[rb + nr - 1]);
      }
    }
  }
}

// ================================================================ Substep2
// Substep2Kernel + Substep2Kernel2 (Kernels.cu:570-621): the artificial
// pressures are recomputed where needed instead of stored; writes the new
// velocities straight into the state (ActualiseGasVrad/Vtheta become no-ops)
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_substep2(Dev<real> D) {
  FGX_STRIP_BEGIN {
    const real dt = (real)s.dt;
    const int jp = (j + 1 == ns) ? 0 : j + 1, jm = (j == 0) ? ns - 1 : j - 1;
    const real cv2 = (real)(kCVNR * kCVNR);
    const real sig = s_whole(D, rb, i, (real)AT(D.dens, i, j));
    const real vri = AT(D.vradint, i, j), vrp = AT(D.vradint, i + 1, j);
    real dv = vrp - vri;
    const real qr = dv < 0 ? cv2 * sig * dv * dv : (real)0;
    const real vti = AT(D.vthetaint, i, j), vtp = AT(D.vthetaint, i, jp), vtm = AT(D.vthetaint, i, jm);
    dv = vtp - vti;
    const real qt = dv < 0 ? cv2 * sig * dv * dv : (real)0;
    const real sigl = s_whole(D, rb, i, (real)AT(D.dens, i, jm));
    dv = vti - vtm;
    const real qtl = dv < 0 ? cv2 * sigl * dv * dv : (real)0;
    const real invdxt = (real)1 / RT(dxtheta)[rb + i];
    if (i > 0) {
      const real sigm = s_whole(D, rb, i - 1, (real)AT(D.dens, i - 1, j));
      dv = vri - AT(D.vradint, i - 1, j);
      const real qrm = dv < 0 ? cv2 * sigm * dv * dv : (real)0;
      AT(D.vrad, i, j) = vri - qdiv(dt * (real)2, sig + sigm) * (qr - qrm) * RT(invdiffrmed)[rb + i];
    } else {
      AT(D.vrad, i, j) = 0;
    }
    AT(D.vtheta, i, j) = vti - qdiv(dt * (real)2, sig + sigl) * (qt - qtl) * invdxt;
    if (D.adiabatic)
      AT(D.energyint, i, j) = AT(D.energy, i, j) - dt * qr * (vrp - vri) * RT(invdiffrsup)[rb + i] -
                              dt * qt * (vtp - vti) * invdxt;
  }
}

// ================================================================ Substep3
// Substep3Kernel + Substep3Kernel2 (Kernels.cu:91-139)
template <class real>
__device__ __forceinline__ real qplus_of(real nu, real sig, real trr, real trp, real tpp, real dv) {
  real q = (real)0.5 / nu / sig * (trr * trr + trp * trp + tpp * tpp);
  q += (real)(2.0 / 9.0) * nu * sig * dv * dv;
  return q;
}
template <class real>
__device__ __forceinline__ real qplus_at(const Dev<real>& D, const DiskConst& c, int rb, long long fb, int ns, int i, int j) {
  const real nu = RT(visc)[rb + i];
  if (nu == (real)0) return 0;
  const real sig = s_whole(D, rb, i, (real)AT(D.dens, i, j));
  return qplus_of(nu, sig, (real)AT(D.taurr, i, j), (real)AT(D.taurp, i, j), (real)AT(D.taupp, i, j),
                  (real)AT(D.divv, i, j));
}
// the same from the stresses formed in place (stress_at on vrad, vtheta)
template <class real>
__device__ __forceinline__ real qplus_here(const Dev<real>& D, const DiskConst& c, int rb, long long fb, int i, int j) {
  const real nu = RT(visc)[rb + i];
  if (nu == (real)0) return 0;
  const Stress<real> t = opaque(stress_at(D, c, rb, fb, (const real*)D.vrad, (const real*)D.vtheta, i, j));
  return qplus_of(nu, t.sig, t.rr, t.rp, t.pp, t.divv);
}
// the new energy of cell (i, j) from its heating qp and divergence dvv
template <class real>
__device__ __forceinline__ real substep3_energy(const Dev<real>& D, const DiskConst& c, const DiskState& s, int rb,
                                                long long fb, int i, int j, real qp, real dvv) {
  const int ns = c.nsec;
  const real dt = (real)s.dt;
  // on residuals of the energy profile e0: e_new - e0 = (num - e0 * den) / den
  const real g1 = (real)c.adiabaticindex - (real)1;
  const real eint = AT(D.energyint, i, j);
  const real e0 = RT(energymed)[rb + i];
  real en;
  if (!D.cooling) {
    const real gd = g1 * dt * dvv;
    en = (dt * qp + eint - e0 * gd) / ((real)1 + gd);
  } else {
    // the density residual is Sigma - SigmaMed
    const real sres = AT(D.dens, i, j), sm = RT(sigmamed)[rb + i], sig = sm + sres;
    const real tc = RT(coolingtime)[rb + i];
    const real num = e0 * dt * sres / sm + tc * eint + dt * tc * (qp - RT(qplusmed)[rb + i] * sig / sm) -
                     e0 * g1 * dt * tc * dvv;
    const real den = dt + tc + g1 * dt * tc * dvv;
    en = num / den;
  }
  return en;
}
// ring 0's heating, extrapolated from rings 1 and 2 (Kernels.cu:91-139)
template <class real>
__device__ __forceinline__ real qplus_ring0(const Dev<real>& D, int rb, real q1, real q2) {
  return q1 * (real)exp((double)log(q1 / q2) * log(RT(rmed_d)[rb] / RT(rmed_d)[rb + 1]) / log(RT(rmed_d)[rb + 1] / RT(rmed_d)[rb + 2]));
}
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_substep3(Dev<real> D) {
  FGX_STRIP_BEGIN {
    real qp;
    if (i > 0) qp = qplus_at(D, c, rb, fb, ns, i, j);
    else if (RT(visc)[rb + nr - 1] != (real)0) {
      const real q1 = qplus_at(D, c, rb, fb, ns, 1, j), q2 = qplus_at(D, c, rb, fb, ns, 2, j);
      qp = qplus_ring0(D, rb, q1, q2);
    } else qp = 0;
    AT(D.qplus, i, j) = qp;
    AT(D.energy, i, j) = substep3_energy(D, c, s, rb, fb, i, j, qp, (real)AT(D.divv, i, j));
  }
}
// Substep3 with its viscous stresses formed in place (ViscousTerms on the
// new velocities + Substep3, Kernels.cu:91-139, 843-880): the operations of
// k_visc_tau and k_substep3, with the stresses kept in registers instead of
// stored and read back; ring 0 forms rings 1 and 2's for its extrapolation
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_substep3_tau(Dev<real> D) {
  FGX_STRIP_BEGIN {
    const Stress<real> t = opaque(stress_at(D, c, rb, fb, (const real*)D.vrad, (const real*)D.vtheta, i, j));
    AT(D.divv, i, j) = t.divv;
    real qp;
    if (i > 0) {
      const real nu = RT(visc)[rb + i];
      qp = nu == (real)0 ? (real)0 : qplus_of(nu, t.sig, t.rr, t.rp, t.pp, t.divv);
    } else if (RT(visc)[rb + nr - 1] != (real)0) {
      qp = qplus_ring0(D, rb, qplus_here(D, c, rb, fb, 1, j), qplus_here(D, c, rb, fb, 2, j));
    } else qp = 0;
    AT(D.qplus, i, j) = qp;
    AT(D.energy, i, j) = substep3_energy(D, c, s, rb, fb, i, j, qp, t.divv);
  }
}

// ============================================================== boundaries
// OpenBoundaryKernel (Kernels.cu:624-637): one block per disk
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_open_boundary(Dev<real> D) {
  const int d = blockIdx.x;
  const DiskConst& c = D.dc[d];
  if (!D.st[d].active) return;
  const int ns = c.nsec, rb = c.ro;
  const long long fb = c.fo;
  const real sm0 = RT(sigmamed)[rb];
  // energy residuals move between rings with the profile difference
  const double de01 = D.adiabatic ? RT(energymed_d)[rb + 1] - RT(energymed_d)[rb] : 0.0;
  const double ds01 = (double)RT(sigmamed)[rb + 1] - (double)RT(sigmamed)[rb];
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    const real s1 = AT(D.dens, 1, j);
    AT(D.dens, 0, j) = (real)(ds01 + (double)s1);
    if (D.adiabatic) AT(D.energy, 0, j) = (real)(de01 + (double)AT(D.energy, 1, j));
    if (AT(D.vrad, 2, j) > 0 || s_whole(D, rb, 1, s1) < sm0) AT(D.vrad, 1, j) = 0;
    else AT(D.vrad, 1, j) = AT(D.vrad, 2, j);
  }
}

// NonReflectingBoundary (SideEuler.cu:70-97) with ReduceCs, ReduceMean (99-120),
// NonReflectingBoundaryKernel and MinusMeanKernel (Kernels.cu:640-723): one
// block per disk
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_nonreflecting(Dev<real> D) {
  const int d = blockIdx.x;
  const DiskConst& c = D.dc[d];
  DiskState& s = D.st[d];
  if (!s.active) return;
  const int ns = c.nsec, nr = c.nrad, rb = c.ro;
  const long long fb = c.fo;
  __shared__ double redd[FGX_TPB / 64];
  __shared__ int ia_sh[2];
  auto csv = [&](int ii, int jj) -> real { return D.adiabatic ? AT(D.cs, ii, jj) : D.cs_ring[rb + ii]; };
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    a0 += (double)csv(0, j); a1 += (double)csv(1, j); a2 += (double)csv(nr - 1, j); a3 += (double)csv(nr - 2, j);
  }
  a0 = block_sum(a0, redd) / ns; a1 = block_sum(a1, redd) / ns;
  a2 = block_sum(a2, redd) / ns; a3 = block_sum(a3, redd) / ns;
  if (threadIdx.x == 0) {
    double da = (pow(RT(rinf_d)[rb + 1], -1.5) - 1.0) / (.5 * (a0 + a1));
    da *= (RT(rmed_d)[rb + 1] - RT(rmed_d)[rb]);
    ia_sh[0] = (int)(da / 2.0 / kPI * (double)ns + .5);
    double da2 = (pow(RT(rinf_d)[rb + nr - 2], -1.5) - 1.0) / (.5 * (a2 + a3));
    da2 *= (RT(rmed_d)[rb + nr - 1] - RT(rmed_d)[rb + nr - 2]);
    ia_sh[1] = (int)(da2 / 2.0 / kPI * (double)ns + .5);
    s.nr_iangle = ia_sh[0]; s.nr_iangle2 = ia_sh[1];
  }
  __syncthreads();
  // the shifts reduced into [0, nsec) (the same sectors; a non-finite sound
  // speed cannot make them index outside the ring)
  const int ia = ((ia_sh[0] % ns) + ns) % ns, ia2 = ((ia_sh[1] % ns) + ns) % ns;
  const real sm1 = RT(sigmamed)[rb + 1], sm2 = RT(sigmamed)[rb + nr - 2];
  // energy residuals move between rings with the profile difference
  const double de01 = D.adiabatic ? RT(energymed_d)[rb + 1] - RT(energymed_d)[rb] : 0.0;
  const double deo = D.adiabatic ? RT(energymed_d)[rb + nr - 2] - RT(energymed_d)[rb + nr - 1] : 0.0;
  const double ds01 = (double)RT(sigmamed)[rb + 1] - (double)RT(sigmamed)[rb];
  const double dso = (double)RT(sigmamed)[rb + nr - 2] - (double)RT(sigmamed)[rb + nr - 1];
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    int jp = j + ia;
    if (jp >= ns) jp -= ns;
    if (jp < 0) jp += ns;
    AT(D.dens, 0, jp) = (real)(ds01 + (double)AT(D.dens, 1, j));
    if (D.adiabatic) AT(D.energy, 0, jp) = (real)(de01 + (double)AT(D.energy, 1, j));
    const real vm = -csv(1, j) * AT(D.dens, 1, j) / sm1;   // (Sigma - SigmaMed) / SigmaMed
    AT(D.vrad, 1, j) = (real)2 * vm - AT(D.vrad, 2, j);
    const int io = nr - 1;
    int jo = j - ia2;
    if (jo >= ns) jo -= ns;
    if (jo < 0) jo += ns;
    AT(D.dens, io, j) = (real)(dso + (double)AT(D.dens, io - 1, jo));
    if (D.adiabatic) AT(D.energy, io, j) = (real)(deo + (double)AT(D.energy, io - 1, jo));
    const real vm2 = csv(io, j) * AT(D.dens, io - 1, j) / sm2;
    AT(D.vrad, io, j) = (real)2 * vm2 - AT(D.vrad, io - 1, j);
  }
  __syncthreads();
  double m0 = 0, m1 = 0, e0 = 0, e1 = 0;
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    m0 += (double)AT(D.dens, 0, j); m1 += (double)AT(D.dens, nr - 1, j);
    if (D.adiabatic) { e0 += (double)AT(D.energy, 0, j); e1 += (double)AT(D.energy, nr - 1, j); }
  }
  const double md = block_sum(m0, redd) / ns, md2 = block_sum(m1, redd) / ns;
  const double me = block_sum(e0, redd) / ns, me2 = block_sum(e1, redd) / ns;
  // ring means restored to the initial profile, in FP64 (md, md2: means of
  // the residuals of S0)
  const double cd = (RT(sigmamed_d)[rb] - (double)RT(sigmamed)[rb]) - md;
  const double cd2 = (RT(sigmamed_d)[rb + nr - 1] - (double)RT(sigmamed)[rb + nr - 1]) - md2;
  // (the energy field is a residual of the profile: its ring mean goes to 0)
  const double ce = D.adiabatic ? -me : 0.0;
  const double ce2 = D.adiabatic ? -me2 : 0.0;
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    AT(D.dens, 0, j) = (real)((double)AT(D.dens, 0, j) + cd);
    AT(D.dens, nr - 1, j) = (real)((double)AT(D.dens, nr - 1, j) + cd2);
    if (D.adiabatic) {
      AT(D.energy, 0, j) = (real)((double)AT(D.energy, 0, j) + ce);
      AT(D.energy, nr - 1, j) = (real)((double)AT(D.energy, nr - 1, j) + ce2);
    }
  }
}

// EvanescentBoundaryKernel (Kernels.cu:1076-1124)
// EvanescentBoundary's per-ring terms (SideEuler.cu:122-141, Kernels.cu:1076-1124)
// at time t: the damping weight d2 = damping^2 * 10 (lambda = d2 * dt / T of
// the zone) and the static part of the relaxation target of vtheta,
// sqrt(1/r (1 - (1 + sigma - 2f) h^2 r^2f)); zero outside the damping zones
template <class real>
__device__ __forceinline__ void ev_ring_terms(const Dev<real>& D, const DiskConst& c, int rb, int nr, int i, double t,
                                              double& d2, double& vt0s) {
  const double r = RT(rmed_d)[rb + i];
  d2 = 0;
  vt0s = 0;
  if (!(r < c.ev_drmin || r > c.ev_drmax)) return;
  double damping;
  if (r < c.ev_drmin) {
    damping = (r - c.ev_drmin) / (RT(rmed_d)[rb] - c.ev_drmin);
    d2 = damping * damping * 10.0;
  }
  if (r > c.ev_drmax) {
    damping = (r - c.ev_drmax) / (RT(rmed_d)[rb + nr - 1] - c.ev_drmax);
    d2 = damping * damping * 10.0;
  }
  const double h = aspect_ratio_dev(c, r, t);
  vt0s = sqrt(1.0 / r * (1.0 - (1.0 + c.sigmaslope - 2.0 * c.flaringindex) * h * h * pow(r, 2.0 * c.flaringindex)));
}
// the tables of disk d, once at set-up: its aspect ratio changes in time only
// through LambdaDoubling, so with LambdaDoubling 0 the terms are those of
// every substep; one block per disk
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_ev_tables(Dev<real> D, int d) {
  const FGX_AS4 DiskConst& c = cref(D.dc + d);
  for (int i = threadIdx.x; i < c.nrad; i += FGX_TPB) {
    double d2, vt0s;
    ev_ring_terms(D, c, c.ro, c.nrad, i, 0.0, d2, vt0s);
    D.ev_d2[c.ro + i] = d2;
    D.ev_vt0s[c.ro + i] = vt0s;
  }
}
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_evanescent(Dev<real> D) {
  FGX_STRIP_BEGIN {
    const double r = RT(rmed_d)[rb + i];
    if (!(r < s.ev_drmin || r > s.ev_drmax)) continue;
    const double dt = s.dt;
    // the ring's terms: from the set-up tables (a pow, a square root and
    // divisions per ring and launch less), or at this time with LambdaDoubling
    double d2, vt0;
    if (c.lambdadoubling == 0.0) {
      d2 = RT(ev_d2)[rb + i];
      vt0 = RT(ev_vt0s)[rb + i];
    } else {
      ev_ring_terms(D, c, rb, nr, i, s.time - s.time0, d2, vt0);
    }
    // the outer zone's time scale where the zones overlap, as the outer
    // assignment comes last in the reference
    const double lambda = d2 * dt / (r > s.ev_drmax ? s.ev_tout : s.ev_tin);
    vt0 -= r * s.omega_frame;
    const double vr0 = -3.0 * (double)RT(visc)[rb + i] / r * (-c.sigmaslope + .5);
    // relaxation towards the FP64 initial profiles, evaluated in FP64
    const double inv = 1.0 / (1.0 + lambda);
    AT(D.vrad, i, j) = (real)(((double)AT(D.vrad, i, j) + lambda * vr0) * inv);
    AT(D.vtheta, i, j) = (real)(((double)AT(D.vtheta, i, j) + lambda * (vt0 - RT(vref)[rb + i])) * inv);
    AT(D.dens, i, j) =
        (real)(((double)AT(D.dens, i, j) + lambda * (RT(sigmamed_d)[rb + i] - (double)RT(sigmamed)[rb + i])) * inv);
    // the energy relaxes to its profile EnergyMed, of which it is stored as a
    // residual: (e + lambda e0) / (1 + lambda) - e0 = (e - e0) / (1 + lambda)
    if (D.adiabatic) AT(D.energy, i, j) = (real)((double)AT(D.energy, i, j) * inv);
  }
}

// ApplyOuterSourceMass, FARGO's SideEuler.c (the reference leaves it
// commented out, SideEuler.cu:143-146,181; it runs here with StubbedPhysics):
// the last ring's mean density is reset to SigmaMed, keeping its azimuthal
// structure, and its radial velocity to IMPOSEDDISKDRIFT * Rinf^-SIGMASLOPE.
// One block per disk.
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_outer_source(Dev<real> D) {
  const int d = blockIdx.x;
  const DiskConst& c = D.dc[d];
  const DiskState& s = D.st[d];
  if (!c.outer_source || skip_disk(D, s, d)) return;
  __shared__ double redd[FGX_TPB / 64];
  const int ns = c.nsec, i = c.nrad - 1;
  const long long fb = c.fo;
  double a = 0;
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) a += (double)(real)AT(D.dens, i, j);
  a = block_sum(a, redd);
  // a: sum of the residuals of S0
  const double add = (RT(sigmamed_d)[c.ro + i] - (double)RT(sigmamed)[c.ro + i]) - a / (double)ns;
  const real vr = (real)c.outer_vr;
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    AT(D.dens, i, j) = (real)((double)(real)AT(D.dens, i, j) + add);
    AT(D.vrad, i, j) = vr;
  }
}

// ====================================================== azimuthal transport
// OneWindTheta (TransportEuler.cu:171-188): ComputeAverageThetaVelocities,
// ComputeResiduals, ComputeConstantResidual, both QuantitiesAdvection passes
// (StarTheta, DivisePolarGrid, VanLeerTheta) and AdvectSHIFT, plus the
// azimuthal half of ComputeVelocities (Kernels.cu:1177-1306).  One block per
// ring; the ring lives in LDS through both passes and the integer shift is a
// rotation of the LDS array on the way out.
template <class real>
__device__ __forceinline__ real th_slope(const real* w, int j, int ns, real invdxt) {
  const int jm = (j == 0) ? ns - 1 : j - 1, jp = (j + 1 == ns) ? 0 : j + 1;
  const real dqm = w[j] - w[jm], dqp = w[jp] - w[j];
  return (dqp * dqm > 0) ? dqp * dqm / (dqp + dqm) * invdxt : (real)0;
}
template <class real>
__device__ __forceinline__ real th_star(const real* w, int j, int ns, real v, real dt, real dxt, real invdxt) {
  const real ksi = v * dt;
  if (ksi > 0) {
    const int jm = (j == 0) ? ns - 1 : j - 1;
    return w[jm] + (dxt - ksi) * th_slope(w, jm, ns, invdxt);
  }
  return w[j] - (dxt + ksi) * th_slope(w, j, ns, invdxt);
}

// fargo integer shift with the reference's promotion order in FP32
// (ComputeConstantResidualKernel, Kernels.cu:1216-1221)
template <class real>
__device__ __forceinline__ int fargo_shift_dev(real vmed, real invrmed, double dt, int ns, real& ntilde, real& nround);
template <>
__device__ __forceinline__ int fargo_shift_dev<float>(float vmed, float invrmed, double dt, int ns, float& ntilde,
                                                      float& nround) {
  const float prod = vmed * invrmed * (float)dt * (float)ns;
  ntilde = (float)((double)prod / 2.0 / kPI);
  nround = (float)floor((double)ntilde + 0.5);
  return (int)(long)nround;
}
template <>
__device__ __forceinline__ int fargo_shift_dev<double>(double vmed, double invrmed, double dt, int ns, double& ntilde,
                                                       double& nround) {
  ntilde = vmed * invrmed * dt * (double)ns / 2.0 / kPI;
  nround = floor(ntilde + 0.5);
  return (int)(long)nround;
}

template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_trans_theta(Dev<real> D, const RingTile* __restrict__ tiles) {
  extern __shared__ unsigned char fgx_smem[];
  const RingTile t = tiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)];
  if (t.d < 0) return;
  const DiskConst& c = D.dc[t.d];
  DiskState& s = D.st[t.d];
  if (skip_disk(D, s, t.d)) return;
  const int ns = c.nsec, nr = c.nrad;
  const int i = t.i;
  const long long fb = c.fo;
  const int rb = c.ro;
  if (i >= nr) return;
  const int MS = D.max_nsec;
  // sD1, sD2 hold the density as the field does, a residual of S0 = s0
  real* sD1 = reinterpret_cast<real*>(fgx_smem);  // density after the radial pass
  real* sS1 = sD1 + MS;                            // DensStar, pass 1 (whole)
  real* sD2 = sD1 + 2 * MS;                        // density after pass 1
  real* sS2 = sD1 + 3 * MS;                        // DensStar, pass 2
  real* sV = sD1 + 4 * MS;                         // residual velocity
  real* sQ = sD1 + 5 * MS;                         // quantity in flight
  real* sA = sD1 + 6 * MS;                         // work = quantity / density
  real* sH = sD1 + 7 * MS;                         // held ThetaMomP
  real* sL = sD1 + 8 * MS;                         // held extensive label (label advection)
  __shared__ double redd[FGX_TPB / 64];
  const long long row = fb + (long long)i * ns;
  const double dtd = s.dt;
  const real dt = (real)dtd;
  // ring mean of vtheta (ComputeAverageThetaVelocitiesKernel), FP64 sum
  double acc = 0;
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) acc += (double)D.vtheta[row + j];
  acc = block_sum(acc, redd);
  const double mres = acc / (double)ns;             // ring mean of the residual
  const real vmed = (real)(RT(vref)[rb + i] + mres);  // VMed
  const real rmed = RT(rmed)[rb + i], irm = RT(invrmed)[rb + i];
  real ntilde, nround;
  const int nsh = fargo_shift_dev<real>(vmed, irm, dtd, ns, ntilde, nround);
  const real invdt = (real)1 / dt;
  const real dpinvns = (real)(2.0 * kPI / (double)ns);
  const real vconst = (ntilde - nround) * rmed * invdt * dpinvns;
  const real v2 = D.fast_transport ? vconst : (real)0;
  if (threadIdx.x == 0) { D.vmed[rb + i] = vmed; D.nshift[rb + i] = nsh; }
  const real dxt = RT(dxtheta)[rb + i];
  const real invdxt = (real)1 / dxt;
  const real dxrad = (RT(rsup)[rb + i] - RT(rinf)[rb + i]) * dt;
  const real invsurf = (real)1 / RT(surf)[rb + i];
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    sD1[j] = D.t_dens[row + j];
    real v = (real)((double)D.vtheta[row + j] - mres);   // VthetaRes
    if (!D.fast_transport) v = vconst + v;
    sV[j] = v;
  }
  __syncthreads();
  const real s0 = RT(sigmamed)[rb + i];
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) sS1[j] = s0 + th_star(sD1, j, ns, sV[j], dt, dxt, invdxt);
  // density, pass 1 -> sD2
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) sA[j] = dens_work(s0 + sD1[j]);
  __syncthreads();
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    const int jp = (j + 1 == ns) ? 0 : j + 1;
    real varq = dxrad * th_star(sA, j, ns, sV[j], dt, dxt, invdxt) * sS1[j] * sV[j];
    varq -= dxrad * th_star(sA, jp, ns, sV[jp], dt, dxt, invdxt) * sS1[jp] * sV[jp];
    sD2[j] = sD1[j] + varq * invsurf;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) sS2[j] = s0 + th_star(sD2, j, ns, v2, dt, dxt, invdxt);
  __syncthreads();
  // one pass of VanLeerTheta on sQ with divisor Dv, DensStar S, velocity v
  // xb: what the stored quantity lacks of the whole one (the energy profile
  // for the energy residual)
  auto pass = [&](const real* Dv, const real* S, const real* varr, real vu, real xb) {
    for (int j = threadIdx.x; j < ns; j += FGX_TPB) sA[j] = (sQ[j] + xb) / (s0 + Dv[j] + (real)1e-20);
    __syncthreads();
    for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
      const int jp = (j + 1 == ns) ? 0 : j + 1;
      const real vj = varr ? varr[j] : vu, vjp = varr ? varr[jp] : vu;
      real varq = dxrad * th_star(sA, j, ns, vj, dt, dxt, invdxt) * S[j] * vj;
      varq -= dxrad * th_star(sA, jp, ns, vjp, dt, dxt, invdxt) * S[jp] * vjp;
      sQ[j] += varq * invsurf;
    }
    __syncthreads();
  };
  int nsm = nsh % ns;
  if (nsm < 0) nsm += ns;
  // quantity order: RadMomP, RadMomM, [energy], [extensive label, held in
  // sL], ThetaMomP (held in sH), ThetaMomM (left in sQ); every quantity sees
  // the same divisors and interface densities, so the order is free here
  const int order[6] = {0, 1, 4, 5, 2, 3};
  for (int o = 0; o < 6; ++o) {
    const int q = order[o];
    if (q == 4 && !D.adiabatic) continue;
    if (q == 5 && !D.advect_label) continue;
    const real* src = q == 0 ? D.t_rmp : q == 1 ? D.t_rmm : q == 2 ? D.t_tmp : q == 3 ? D.t_tmm : q == 4 ? D.t_e : D.t_lab;
    for (int j = threadIdx.x; j < ns; j += FGX_TPB) sQ[j] = src[row + j];
    __syncthreads();
    const real xb = q == 4 ? RT(energymed)[rb + i] : (real)0;
    pass(sD1, sS1, sV, (real)0, xb);
    pass(sD2, sS2, nullptr, v2, xb);
    if (q == 2 || q == 5) {
      real* hold = q == 2 ? sH : sL;
      for (int j = threadIdx.x; j < ns; j += FGX_TPB) hold[j] = sQ[j];
      __syncthreads();
    } else if (q != 3) {
      // AdvectSHIFTKernel (Kernels.cu:1276-1289): out[j] = in[j - Nshift]
      real* dst = q == 0 ? D.t_rmp : q == 1 ? D.t_rmm : D.energy;
      for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
        int js = j - nsm;
        js += (js < 0) ? ns : 0;
        dst[row + j] = sQ[js];
      }
      __syncthreads();
    }
  }
  // density, pass 2 -> sD1 (the radial-pass density is no longer needed)
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) sA[j] = dens_work(s0 + sD2[j]);
  __syncthreads();
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    const int jp = (j + 1 == ns) ? 0 : j + 1;
    real varq = dxrad * th_star(sA, j, ns, v2, dt, dxt, invdxt) * sS2[j] * v2;
    varq -= dxrad * th_star(sA, jp, ns, v2, dt, dxt, invdxt) * sS2[jp] * v2;
    sD1[j] = sD2[j] + varq * invsurf;
  }
  __syncthreads();
  // shifted density and the azimuthal velocity of ComputeVelocitiesKernel
  // (Kernels.cu:1303-1304), evaluated in the unshifted frame and rotated
  // the angular momenta are residuals of density * l0 (k_trans_rad), so
  // (TMP + TMM)/(S + 1e-15)/R - R*Omega = vref + residual below, exactly
  const double rd = RT(rmed_d)[rb + i];
  const double l0 = (RT(vref)[rb + i] + rd * s.omega_frame) * rd;
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    int js = j - nsm;
    js += (js < 0) ? ns : 0;
    const int jsm = (js == 0) ? ns - 1 : js - 1;
    D.dens[row + j] = sD1[js];
    if (D.advect_label) D.label[row + j] = sL[js] / (s0 + sD1[js]);   // ComputeSpeQty (Kernels.cu:1309-1319)
    const double S = 2.0 * (double)s0 + (double)sD1[js] + (double)sD1[jsm];
    D.vtheta[row + j] = (real)(((double)sH[jsm] + (double)sQ[js] - l0 * 1e-15) / (S + 1e-15) / rd);
  }
}

// radial half of ComputeVelocitiesKernel (Kernels.cu:1299-1302)
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_trans_vr(Dev<real> D) {
  // vr(i) = (RadMomP(i-1) + RadMomM(i)) / (Sigma(i) + Sigma(i-1) + 1e-20),
  // vr(0) = 0: ring i-1's values are carried in registers and ring i+1's are
  // fetched before ring i's store
  const CellTile t = D.stiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)];
  if (t.d < 0) return;
  const FGX_AS4 DiskConst& c = cref(D.dc + t.d);
  const FGX_AS4 DiskState& s = cref(D.st + t.d);
  if (skip_disk(D, s, t.d)) return;
  const int ns = c.nsec, nr = c.nrad;
  const int j = t.jb * FGX_TPB + (int)threadIdx.x;
  if (j >= ns) return;
  const long long fb = c.fo;
  const int i0 = t.i, i_end = (t.i + D.strip_rings < nr) ? t.i + D.strip_rings : nr;
  const int rb = c.ro;
  real pm = 0, dm = 0;   // RadMomP and Sigma of ring i-1 (whole values)
  if (i0 > 0) { pm = AT(D.t_rmp, i0 - 1, j); dm = s_whole(D, rb, i0 - 1, (real)AT(D.dens, i0 - 1, j)); }
  real nm = AT(D.t_rmm, i0, j), nd = AT(D.dens, i0, j), np = AT(D.t_rmp, i0, j);
  for (int i = i0; i < i_end; ++i) {
    const real rmm = nm, den = s_whole(D, rb, i, nd), rmp = np;
    if (i + 1 < i_end) { nm = AT(D.t_rmm, i + 1, j); nd = AT(D.dens, i + 1, j); np = AT(D.t_rmp, i + 1, j); }
    AT(D.vrad, i, j) = (i == 0) ? (real)0 : (pm + rmm) / (den + dm + (real)1e-20);
    pm = rmp;
    dm = den;
  }
}

// ============================================================ bookkeeping
// start of an AlgoGas call (SourceEuler.cu:211-216)
template <class real>
__global__ void k_begin_dt(Dev<real> D) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= D.ndisk) return;
  DiskState& s = D.st[d];
  s.vpar = D.vpar;   // every disk's velocities are in the current buffers now
  if (s.frozen) { s.active = 0; return; }
  s.dtemp = 0;
  s.sub_in_dt = 0;
  s.pending = 0;
  s.active = 1;
  set_force_targets(D, D.dc[d], s);
}
// the end of each disk's last substep of the DT (end_substep_state): one
// wave per disk
template <class real>
__global__ void __launch_bounds__(64) k_finish_dt(Dev<real> D) {
  const int d = blockIdx.x;
  const DiskConst& c = D.dc[d];
  DiskState& s = D.st[d];
  if (!s.pending) return;
  const double lost = lost_sum_wave(D, c);
  if (threadIdx.x == 0) end_substep_state(D, c, s, lost);
}
// end of a DT with the fused source pass: a disk that stopped taking
// substeps before the batch did holds its velocities in the buffer pair that
// was current at its last substep; they are copied into the current pair
// (rings 0..nrad, the ghost ring with the last strip)
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_vel_fixup(Dev<real> D) {
  const CellTile t = D.stiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)];
  if (t.d < 0) return;
  if (D.st[t.d].vpar == D.vpar) return;
  const FGX_AS4 DiskConst& c = cref(D.dc + t.d);
  const int ns = c.nsec, nr = c.nrad;
  const int j = t.jb * FGX_TPB + (int)threadIdx.x;
  if (j >= ns) return;
  const long long fb = c.fo;
  const int i_end = (t.i + D.strip_rings < nr) ? t.i + D.strip_rings : nr + 1;
  for (int i = t.i; i < i_end; ++i) {
    AT(D.vrad, i, j) = (real)AT(D.vradint, i, j);
    AT(D.vtheta, i, j) = (real)AT(D.vthetaint, i, j);
  }
}
// batch summary for the host loop: active disks and the largest CFL count
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_summary(Dev<real> D) {
  static_assert(FGX_SUMMARY_HIST == 64, "the forecast is scanned by one wave");
  __shared__ int red[FGX_TPB / 64];
  __shared__ int hist[FGX_SUMMARY_HIST + 1];   // difference array of the forecast
  for (int j = threadIdx.x; j <= FGX_SUMMARY_HIST; j += FGX_TPB) hist[j] = 0;
  __syncthreads();
  int a = 0, m = 0;
  for (int d = threadIdx.x; d < D.ndisk; d += FGX_TPB) {
    const DiskState& st = D.st[d];
    const int act = disk_running(D.dc[d], st) ? 1 : 0;
    a += act;
    // substeps left after the pending one
    const int sub = st.sub_in_dt + st.pending;
    const int rem = act ? (D.sloppy ? st.ncfl - sub : st.ncfl - 1) : 0;
    m = rem > m ? rem : m;
    // forecast of the cells still active j substeps ahead, in units of 64:
    // the disk adds its cells to entries [0, rem)
    const int c64 = (D.dc[d].nrad * D.dc[d].nsec + 63) / 64;
    const int r = rem < FGX_SUMMARY_HIST ? rem : FGX_SUMMARY_HIST;
    if (r > 0) {
      atomicAdd(&hist[0], c64);
      atomicAdd(&hist[r], -c64);
    }
  }
  a = block_sum(a, red);
  m = -block_min(-m, red);
  if (threadIdx.x == 0) { D.summary[0] = a; D.summary[1] = m; }
  // which disks are active: the host narrows the tile lists to them
  for (int d = threadIdx.x; d < D.ndisk; d += FGX_TPB) D.summary[2 + FGX_SUMMARY_HIST + d] = disk_running(D.dc[d], D.st[d]);
  __syncthreads();
  if (threadIdx.x < 64) {
    // inclusive scan of the differences across the wave
    const int lane = (int)threadIdx.x;
    int v = hist[lane];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(v, o, 64);
      if (lane >= o) v += u;
    }
    D.summary[2 + lane] = v;
  }
}

// ============================================================ tile lists
// A tile list narrowed to the disks still active in a DT: range r copies len
// tiles of e ints from src to dst, or, with src null, writes len empty tiles
// (first int -1, the rest 0).  One block per range.
struct CompactRange {
  const int* src;
  int* dst;
  int len, e;
};
__global__ void __launch_bounds__(FGX_TPB) k_compact(const CompactRange* __restrict__ rg) {
  const CompactRange r = rg[blockIdx.x];
  const int n = r.len * r.e;
  for (int k = threadIdx.x; k < n; k += FGX_TPB) r.dst[k] = r.src ? r.src[k] : ((k % r.e) == 0 ? -1 : 0);
}

// ============================================================== diagnostics
// potential of the current substep into D.work (FillForcesArrays output)
// FillForcesArraysKernel (Kernels.cu:1322-1349) for the planets and the
// indirect term, once per cell and substep, in FP64
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_potential(Dev<real> D) {
  FGX_STRIP_BEGIN {
    AT(D.phi, i, j) = potential_planets_at(D, c, s, RT(rmed_d)[rb + i], j);
  }
}
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_potential_field(Dev<real> D) {
  FGX_CELL_PROLOGUE
  if (i >= nr) return;
  const double r = RT(rmed_d)[rb + i];
  AT(D.work, i, j) = (real)(potential_planets_at(D, c, s, r, j) - 1.0 / r);
}
// ComputeTemperatureFieldKernel (Kernels.cu:211-223) into D.work
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_temperature(Dev<real> D) {
  FGX_CELL_PROLOGUE
  if (i >= nr) return;
  const real sig = s_whole(D, rb, i, (real)AT(D.dens, i, j));
  if (!D.adiabatic) AT(D.work, i, j) = AT(D.pressure, i, j) / sig;
  else AT(D.work, i, j) = (real)(c.adiabaticindex - 1.0) * e_whole(D, rb, i, (real)AT(D.energy, i, j)) / sig;
}
// isothermal pressure at set-up (ComputePressureFieldKernel at InitEuler)
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_pressure_init(Dev<real> D) {
  FGX_CELL_PROLOGUE
  if (i >= nr) return;
  if (!D.adiabatic) {
    const real cs = D.cs_ring[rb + i];
    AT(D.pressure, i, j) = s_whole(D, rb, i, (real)AT(D.dens, i, j)) * cs * cs;
  } else {
    AT(D.pressure, i, j) = (real)(c.adiabaticindex - 1.0) * e_whole(D, rb, i, (real)AT(D.energy, i, j));
  }
}
// ComputeForce with all eleven Hill-cut variants (Force.cu:31-67,
// Kernels.cu:804-839) for one point, per-ring FP64 partials into part[44 i ..]
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_force_full(Dev<real> D, int disk, double x, double y, double sm,
                                                        double mass, double* __restrict__ part) {
  const int i = blockIdx.x;
  const DiskConst& c = D.dc[disk];
  if (i >= c.nrad) return;
  const int ns = c.nsec, rb = c.ro;
  const long long fb = c.fo;
  __shared__ double redd[FGX_TPB / 64];
  const double a = sqrt(x * x + y * y);
  const double rh = pow(mass / 3., 1. / 3.) * a + 1e-15;
  const real rmed = RT(rmed)[rb + i], surf = RT(surf)[rb + i];
  const bool inner = (double)rmed < (double)(float)a;
  double f[22];
  for (int k = 0; k < 22; ++k) f[k] = 0;
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    const real cm = surf * s_whole(D, rb, i, (real)AT(D.dens, i, j));
    const real dx = rmed * RT(cosj)[c.so + j] - (real)x, dy = rmed * RT(sinj)[c.so + j] - (real)y;
    const real d2 = dx * dx + dy * dy;
    const real dist2 = d2 + (real)(sm * sm);
    const real inv3 = (real)1 / dist2 / sqrt(dist2);
    for (int k = 0; k < 11; ++k) {
      real cut = 1;
      if (k) {
        const real rk = (real)(rh * (double)k / 10.0);
        cut = (real)1 - exp(-d2 / (rk * rk));
      }
      f[k] += (double)(cm * dx * inv3 * cut);
      f[11 + k] += (double)(cm * dy * inv3 * cut);
    }
  }
  double* out = part + (size_t)i * 44;
  for (int k = 0; k < 11; ++k) {
    const double fx = block_sum(f[k], redd), fy = block_sum(f[11 + k], redd);
    if (threadIdx.x == 0) {
      out[k] = inner ? fx : 0; out[11 + k] = inner ? 0 : fx;
      out[22 + k] = inner ? fy : 0; out[33 + k] = inner ? 0 : fy;
    }
  }
}
// the 44 force totals from k_force_full's per-ring partials, each summed in
// ring order (the order of the host sum of fgx_disk_force, so both agree bit
// for bit): one thread per component
__global__ void __launch_bounds__(64) k_force_sum(const double* __restrict__ part, int nr, double* __restrict__ out) {
  const int q = (int)threadIdx.x;
  if (q >= 44) return;
  double s = 0;
  for (int i = 0; i < nr; ++i) s += part[(size_t)i * 44 + q];
  out[q] = s;
}
// CircumPlanetaryMassKernel (Kernels.cu:245-261), per-ring FP64 partials
template <class real>
__global__ void __launch_bounds__(FGX_TPB) k_mdcp(Dev<real> D, int disk, double xpl, double ypl, double rhill) {
  const int i = blockIdx.x;
  const DiskConst& c = D.dc[disk];
  if (i >= c.nrad) return;
  const int ns = c.nsec, rb = c.ro;
  const long long fb = c.fo;
  __shared__ double redd[FGX_TPB / 64];
  const real rmed = RT(rmed)[rb + i], surf = RT(surf)[rb + i];
  double acc = 0;
  for (int j = threadIdx.x; j < ns; j += FGX_TPB) {
    const real dx = rmed * RT(cosj)[c.so + j] - (real)xpl, dy = rmed * RT(sinj)[c.so + j] - (real)ypl;
    if ((double)sqrt(dx * dx + dy * dy) < rhill) acc += (double)(surf * s_whole(D, rb, i, (real)AT(D.dens, i, j)));
  }
  acc = block_sum(acc, redd);
  if (threadIdx.x == 0) D.rpart[i] = acc;
}

}  // namespace fgx
