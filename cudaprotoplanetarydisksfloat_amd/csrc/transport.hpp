// transport.hpp -- the two transport sweeps of the FARGO step, register
// resident.
//
// Radial pass (OneWindRad, TransportEuler.cu:54-73 with LRMomentaKernel,
// StarRadKernel/2, DivisePolarGridKernel and VanLeerRadialKernel,
// Kernels.cu:883-947, 1127-1154): a thread owns one sector of a strip of
// rings and walks outwards with a three-ring window, so every ring is
// loaded once per strip (plus a four-ring halo), every van Leer slope and
// every interface flux is formed once and shared by the two cells it
// separates.
//
// Azimuthal pass (OneWindTheta, TransportEuler.cu:171-188, with the
// ComputeAverageThetaVelocities, ComputeResiduals, ComputeConstantResidual,
// StarTheta/2, VanLeerTheta, AdvectSHIFT and the azimuthal half of
// ComputeVelocities kernels, Kernels.cu:1177-1306): one wave64 per ring; lane
// l holds the K contiguous sectors [l*K, l*K+K) in registers through both
// advection passes of all quantities, chunk-edge neighbours come from the
// adjacent lanes by cross-lane reads, and global traffic goes through a
// padded LDS transpose so that loads and stores stay coalesced.  The integer
// shift is applied on the way out of LDS.
#pragma once
#include "kernels.hpp"


namespace fgx {

__device__ __forceinline__ float vl_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double vl_rcp(double x) { return 1.0 / x; }
// DivisePolarGrid's work x / (d + 1e-20): FP32 as a product with the
// reciprocal `inv` of d + 1e-20, FP64 as the quotient itself
__device__ __forceinline__ float work_of(float x, float d, float inv) { (void)d; return x * inv; }
__device__ __forceinline__ double work_of(double x, double d, double inv) { (void)inv; return x / (d + 1e-20); }
// a * b + c with one rounding
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

// whole-wave rotations by one lane (DPP wave_ror:1 / wave_rol:1): lane l
// receives lane l-1 (lane 0 receives lane 63), resp. lane l+1 (lane 63
// receives lane 0)
__device__ __forceinline__ int dpp_ror1(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x13C, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_rol1(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x134, 0xF, 0xF, false); }
__device__ __forceinline__ float lane_from_left(float x) { return __int_as_float(dpp_ror1(__float_as_int(x))); }
__device__ __forceinline__ float lane_from_right(float x) { return __int_as_float(dpp_rol1(__float_as_int(x))); }
__device__ __forceinline__ float lane_read(float x, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }
__device__ __forceinline__ double lane_from_left(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_ror1((int)(b & 0xffffffffLL)), hi = dpp_ror1((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_from_right(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_rol1((int)(b & 0xffffffffLL)), hi = dpp_rol1((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_read(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// ============================================================ radial sweep
template <class real, bool LAB = false>
struct RadRing {
  real q[LAB ? 8 : 7];   // 0 density (for DensStar), 1..6 work = quantity / (density + 1e-20):
                         // density, RadMomP, RadMomM, ThetaMomP, ThetaMomM, energy, [label]
  real sg, sr, vr0, vr1, e;   // sg the whole density, sr its residual of S0 (the field's value)
  real x;         // extensive label, density * label (label advection)
  double tp, tm;  // the angular momenta's departures from density * l0, in FP64:
                  // density * (vtheta residual) * R, l0 = (vref + R*Omega) * R
};

// a disk's field as a buffer resource (scalar registers): a ring's offset
// goes into the scalar offset and the sector's into a 32-bit vector offset
// fixed for the whole walk, so a ring's loads and stores cost no vector
// address arithmetic.  Offsets are bytes; a disk's field is far below 2 GiB.
template <class T>
struct FieldBuf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ T ld(int soff, unsigned voff) const { return FieldRef<T>{r, soff, voff}.get(); }
  __device__ __forceinline__ void st(int soff, unsigned voff, T v) const { FieldRef<T>{r, soff, voff}.put(v); }
};
template <class T>
__device__ __forceinline__ FieldBuf<T> field_buf(const T* base) {
  return FieldBuf<T>{__builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, 0x7fffffff, FGX_BUF_WORD3)};
}

// the disk's fields seen from its first element
template <class real>
struct RadFields {
  FieldBuf<real> dens, vrad, vtheta, energy, label;
  const FGX_AS4 double *rmed_d, *vref;
  const FGX_AS4 real* emed;   // ring profile the stored energy is a residual of
  const FGX_AS4 real* smed;   // S0, the ring profile the stored density is a residual of
};

// the raw values of ring k a thread needs (clamped rings only feed zero
// slopes / zero fluxes)
template <class real>
struct RadRaw {
  real sg, vr0, vr1, vtp, vtj, e, lab;
};
// vj, vjp: the byte offsets of sectors j and j+1 (cyclic) within a ring
template <class real, bool LAB = false>
__device__ __forceinline__ void rad_fetch(const RadFields<real>& f, int ns, int nr, int k, unsigned vj, unsigned vjp,
                                          bool ad, RadRaw<real>& w) {
  k = k < 0 ? 0 : (k > nr - 1 ? nr - 1 : k);
  const int o = k * ns * (int)sizeof(real), o1 = o + ns * (int)sizeof(real);
  w.sg = f.dens.ld(o, vj);
  w.vr0 = f.vrad.ld(o, vj);
  w.vr1 = f.vrad.ld(o1, vj);
  w.vtp = f.vtheta.ld(o, vjp);
  w.vtj = f.vtheta.ld(o, vj);
  w.e = ad ? f.energy.ld(o, vj) : (real)0;
  w.lab = LAB ? f.label.ld(o, vj) : (real)0;
}

template <class real, bool LAB = false>
__device__ __forceinline__ void rad_derive(const RadFields<real>& f, int nr, int k, double om, bool ad,
                                           const RadRaw<real>& w, RadRing<real, LAB>& r) {
  k = k < 0 ? 0 : (k > nr - 1 ? nr - 1 : k);
  const real s0 = f.smed[k];
  const real sg = s0 + w.sg;   // the density field holds the residual of S0
  const double rk = f.rmed_d[k];
  const double l0 = (f.vref[k] + rk * om) * rk;
  const double lp = l0 + (double)w.vtp * rk, lm = l0 + (double)w.vtj * rk;
  // DivisePolarGrid: quantity / (density + 1e-20); FP32 forms it as a
  // product with one reciprocal per cell.  The density's own work is exactly
  // 1 wherever the 1e-20 is below half an ulp of the density, as in the
  // quotient.
  const real inv = vl_rcp(sg + (real)1e-20);
  const real w1 = dens_work(sg);
  r.sg = sg; r.sr = w.sg; r.vr0 = w.vr0; r.vr1 = w.vr1;
  r.q[0] = sg;
  r.q[1] = w1;
  r.q[2] = work_of(sg * w.vr1, sg, inv);
  r.q[3] = work_of(sg * w.vr0, sg, inv);
  const double dd = (double)s0 + (double)w.sg;   // the whole density, exactly
  r.tp = dd * ((double)w.vtp * rk); r.tm = dd * ((double)w.vtj * rk);
  if constexpr (sizeof(real) == 4) {
    // (density * l) / (density + 1e-20): l times the density's work
    r.q[4] = (real)lp * w1; r.q[5] = (real)lm * w1;
  } else {
    r.q[4] = (real)(dd * lp / (dd + 1e-20)); r.q[5] = (real)(dd * lm / (dd + 1e-20));
  }
  // the energy is stored as a residual of its ring profile: its work is
  // formed from the whole value (its flux is small, so one rounding of the
  // sum costs nothing), its update stays a residual
  r.e = w.e;
  r.q[6] = ad ? work_of(f.emed[k] + w.e, sg, inv) : (real)0;
  if constexpr (LAB) {
    // ComputeExtQty (Kernels.cu:899-906): the label travels as density * label
    r.x = sg * w.lab;
    r.q[7] = r.x * inv;
  }
}

template <class real, bool LAB = false>
__device__ __forceinline__ void rad_load(const RadFields<real>& f, int ns, int nr, int k, unsigned vj, unsigned vjp,
                                         double om, bool ad, RadRing<real, LAB>& r) {
  RadRaw<real> w;
  rad_fetch<real, LAB>(f, ns, nr, k, vj, vjp, ad, w);
  rad_derive<real, LAB>(f, nr, k, om, ad, w, r);
}

// van Leer slope of ring k from rings k-1, k, k+1 (StarRadKernel); zero on
// the first and last ring
template <class real>
__device__ __forceinline__ real rad_vl(real wm, real w0, real wp, real idm, real idp) {
  const real dqm = (w0 - wm) * idm;
  const real dqp = (wp - w0) * idp;
  return (dqp * dqm > 0) ? (real)2 * dqp * dqm * vl_rcp(dqp + dqm) : (real)0;
}

// LAB: the label is advected as an eighth quantity (density * label).
// FP64 without the label is held to 168 VGPRs for three waves per SIMD (it
// fits without spilling, and the walk is latency-bound at two)
template <class real, bool LAB>
__global__ void __launch_bounds__(FGX_TPB, sizeof(real) == 8 && !LAB ? 3 : 1) k_trans_rad_strip(Dev<real> D, const CellTile* __restrict__ tiles, int rs) {
  constexpr int NQ = LAB ? 8 : 7;
  const CellTile t = tiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)];
  if (t.d < 0) return;
  const DiskConst& c = D.dc[t.d];
  const DiskState& s = D.st[t.d];
  if (skip_disk(D, s, t.d)) return;
  const int ns = c.nsec, nr = c.nrad;
  const int j = t.jb * FGX_TPB + (int)threadIdx.x;
  if (j >= ns) return;
  const long long fb = c.fo;
  const int rb = c.ro;
  const int i0 = t.i, i1 = (t.i + rs < nr) ? t.i + rs : nr;
  const int jp = (j + 1 == ns) ? 0 : j + 1;
  const bool ad = D.adiabatic;
  const double om = s.omega_frame;
  const real dt = (real)s.dt;
  const real dth = (real)(2.0 * kPI / (double)ns);
  const auto rmed = RT(rmed) + rb;
  const auto idr = RT(invdiffrmed) + rb;
  // quantities in use: 0..5 always, 6 (energy) if adiabatic, 7 (label) if LAB
  auto qon = [&](int q) { return q < 6 || q == 7 || ad; };
  const RadFields<real> F{field_buf(D.dens + fb), field_buf(D.vrad + fb), field_buf(D.vtheta + fb),
                          field_buf(D.energy + (ad ? fb : 0)), field_buf(LAB ? D.label + fb : D.dens + fb),
                          RT(rmed_d) + rb, RT(vref) + rb, RT(energymed) + rb, RT(sigmamed) + rb};
  const FieldBuf<real> o_dens = field_buf(D.t_dens + fb);
  const FieldBuf<real> o_rmp = field_buf(D.t_rmp + fb);
  const FieldBuf<real> o_rmm = field_buf(D.t_rmm + fb);
  const FieldBuf<real> o_tmp = field_buf(D.t_tmp + fb);
  const FieldBuf<real> o_tmm = field_buf(D.t_tmm + fb);
  const FieldBuf<real> o_e = field_buf(D.t_e + (ad ? fb : 0));
  const FieldBuf<real> o_lab = field_buf(LAB ? D.t_lab + fb : D.t_dens + fb);
  const unsigned vj = (unsigned)j * (unsigned)sizeof(real), vjp = (unsigned)jp * (unsigned)sizeof(real);

  using Ring = RadRing<real, LAB>;
  // a ring's slope coefficients, read once for all quantities (read inside
  // each quantity's edge test, they were fetched and waited for again per
  // quantity); the slope is zero on the first and last ring
  struct SlopeRing { real im, ip; bool inner; };
  auto slope_ring = [&](int k) -> SlopeRing {
    const int kc = k < 0 ? 0 : (k < nr - 1 ? k : (nr >= 2 ? nr - 2 : 0));
    return SlopeRing{idr[kc], idr[kc + 1], k > 0 && k < nr - 1};
  };
  auto slope_at = [&](const SlopeRing& g, const Ring& m, const Ring& z, const Ring& p, int q) -> real {
    const real sl = rad_vl(m.q[q], z.q[q], p.q[q], g.im, g.ip);
    return g.inner ? sl : (real)0;
  };
  // flux through the interface between ring lo and ring lo + 1, zero at the
  // disk's edges.  The ring spacings are formed before the quantity loop:
  // read inside the two upwind arms, the table loads had turned each
  // selection into a divergent branch.
  auto iface = [&](int lo, const Ring& L, const Ring& H, const real (&sL)[NQ], const real (&sH)[NQ],
                   real (&f)[NQ]) {
    const real v = H.vr0;
    const bool edge = lo < 0 || lo + 1 >= nr;
    const int l0 = lo < 0 ? 0 : lo, l1 = lo + 1 < 0 ? 0 : lo + 1, l2 = lo + 2 < 0 ? 0 : (lo + 2 <= nr ? lo + 2 : nr);
    const real hm = rmed[l1] - rmed[l0], hp = rmed[l2] - rmed[l1];
    const real a = edge ? (real)0 : dt * dth * RT(rinf)[rb + l1];
    const real cm = (hm - v * dt) * (real)0.5, cp = (hp + v * dt) * (real)0.5;
    const bool up = v > 0;
    real st0 = 0;
    f[0] = 0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (!qon(q)) { f[q] = 0; continue; }
      // both upwind values, then a selection (a conditional expression over
      // the two formulas compiles to a divergent branch per quantity)
      const real from_lo = L.q[q] + cm * sL[q], from_hi = H.q[q] - cp * sH[q];
      const real st = up ? from_lo : from_hi;
      if (q == 0) st0 = st;
      else f[q] = edge ? (real)0 : a * st * st0 * v;
    }
  };
  // The walk starts two rings before the strip: those two passes only form
  // the slopes of rings i0-1 and i0 and the flux through interface i0, with
  // the same instructions as every later ring, so no value depends on where
  // a strip begins (a separate prologue rounds them differently once the
  // compiler fuses its products into the flux arithmetic).
  Ring ra, rb1;
  rad_load<real, LAB>(F, ns, nr, i0 - 2, vj, vjp, om, ad, ra);
  rad_load<real, LAB>(F, ns, nr, i0 - 1, vj, vjp, om, ad, rb1);
  real slA[NQ], flux[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) slA[q] = flux[q] = (real)0;
  const auto isurf = RT(invsurf) + rb;
  // ring i+2 is fetched two iterations ahead, before the stores of rings
  // i-2 and i-1: the vector memory counter retires loads and stores in issue
  // order, and four rings are in flight when the walk starts
  RadRaw<real> nxt, nxt2;
  rad_fetch<real, LAB>(F, ns, nr, i0, vj, vjp, ad, nxt);
  rad_fetch<real, LAB>(F, ns, nr, i0 + 1, vj, vjp, ad, nxt2);
  for (int i = i0 - 2; i < i1; ++i) {
    const RadRaw<real> cur = nxt;
    nxt = nxt2;
    if (i + 2 < i1) rad_fetch<real, LAB>(F, ns, nr, i + 4, vj, vjp, ad, nxt2);
    Ring rc;
    rad_derive<real, LAB>(F, nr, i + 2, om, ad, cur, rc);
    // slope of ring i+1, flux through interface i+1 (zero at the outer edge)
    const SlopeRing gb = slope_ring(i + 1);
    real fnew[NQ], slB[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) slB[q] = qon(q) ? slope_at(gb, ra, rb1, rc, q) : (real)0;
    iface(i, ra, rb1, slA, slB, fnew);
    if (i >= i0 && i < i1) {
      const real is = isurf[i];
      const double rd = F.rmed_d[i], vref_i = F.vref[i];
      // cell i: quantity + (flux in - flux out) / Surf
      const real inc = (flux[1] - fnew[1]) * is;
      const real sg = ra.sg;
      const int oc = i * ns * (int)sizeof(real);
      o_dens.st(oc, vj, ra.sr + inc);   // a residual of S0, as the field
      if (i == 0 && D.open_inner) D.lost[c.so + j] = flux[1] - fnew[1];
      o_rmp.st(oc, vj, sg * ra.vr1 + (flux[2] - fnew[2]) * is);
      o_rmm.st(oc, vj, sg * ra.vr0 + (flux[3] - fnew[3]) * is);
      // angular momenta leave as residuals of density * l0(ring), with the
      // density taken before its field rounding, so the azimuthal velocity
      // rebuilt from them carries no trace of that rounding:
      //   (L + dL) - (Sigma + dSigma) l0 = Sigma (l - l0) + dL - dSigma l0,
      // formed from its small terms only (the O(1) products Sigma l and
      // Sigma l0 would each carry a rounding of the size of a vtheta ulp)
      const double l0i = (vref_i + rd * om) * rd;
      const double dsl = (double)inc * l0i;
      o_tmp.st(oc, vj, (real)(ra.tp + (double)((flux[4] - fnew[4]) * is) - dsl));
      o_tmm.st(oc, vj, (real)(ra.tm + (double)((flux[5] - fnew[5]) * is) - dsl));
      if (ad) o_e.st(oc, vj, ra.e + (flux[6] - fnew[6]) * is);
      if constexpr (LAB) o_lab.st(oc, vj, ra.x + (flux[7] - fnew[7]) * is);
    }
    // slide the window
#pragma unroll
    for (int q = 0; q < NQ; ++q) { flux[q] = fnew[q]; slA[q] = slB[q]; }
    ra = rb1;
    rb1 = rc;
  }
}

// ========================================================= azimuthal sweep
// K contiguous values at p (aligned to the vector width) into registers
template <class real, int K>
__device__ __forceinline__ void ld_chunk(const real* __restrict__ p, real (&v)[K]) {
  if constexpr (sizeof(real) == 4 && K % 4 == 0) {
#pragma unroll
    for (int q = 0; q < K / 4; ++q) {
      const float4 x = reinterpret_cast<const float4*>(p)[q];
      v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
    }
  } else if constexpr (sizeof(real) == 4 && K % 2 == 0) {
#pragma unroll
    for (int q = 0; q < K / 2; ++q) {
      const float2 x = reinterpret_cast<const float2*>(p)[q];
      v[2 * q] = x.x; v[2 * q + 1] = x.y;
    }
  } else if constexpr (sizeof(real) == 8 && K % 2 == 0) {
#pragma unroll
    for (int q = 0; q < K / 2; ++q) {
      const double2 x = reinterpret_cast<const double2*>(p)[q];
      v[2 * q] = x.x; v[2 * q + 1] = x.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = p[k];
  }
}

// A ring is split into W arcs of T*K sectors, one wave64 per arc (T = 64 when
// W > 1, T <= 64 when W == 1); lane l of wave w holds the K contiguous sectors
// starting at (w*64 + l)*K.  Neighbours inside a wave come from DPP lane
// rotations, across arcs through two LDS words per wave edge (double
// buffered, one barrier per exchange).  Every disk's field offset is a
// multiple of 64 elements, so each lane's chunk is aligned for vector loads.
// LAB: the extensive label is carried as one more quantity and the label
// (ComputeSpeQty, Kernels.cu:1309-1319) formed from the final density
template <class real, int K, int WM, bool LAB, bool AD>
__device__ __forceinline__ void theta_arc_ring(const Dev<real>& D, const RingTile t, int W) {
  // WM: 2 = W arcs of one full wave each, 1 = one full wave (nsec = 64*K),
  // 0 = one wave of T < 64 lanes
  constexpr bool MULTI = WM == 2;
  __shared__ real ex[2][12][16];
  __shared__ double redw[16];
  if (t.d < 0) return;
  const FGX_AS4 DiskConst& c = cref(D.dc + t.d);
  const FGX_AS4 DiskState& s = cref(D.st + t.d);
  if (skip_disk(D, s, t.d)) return;
  const int ns = c.nsec, i = t.i, rb = c.ro;
  const int w = MULTI ? (int)threadIdx.x >> 6 : 0, lane = (int)threadIdx.x & 63;
  const int T = (WM == 0) ? ns / K : 64;
  const bool on = (WM == 0) ? lane < T : true;
  const int wl = (w == 0) ? W - 1 : w - 1, wr = (w + 1 == W) ? 0 : w + 1;
  // lanes past the ring read its first chunk (never stored)
  const long long row = c.fo + (long long)i * ns;
  const int j0 = on ? (w * 64 + lane) * K : 0;   // first sector of this lane
  const long long my = row + j0;
  constexpr bool ad = AD;   // the energy equation, fixed per instantiation
  int xb = 0;

  // neighbour lanes along the ring; in(): values of this lane's left and
  // right neighbours for up to three exchanged quantities at once
  auto fix_single = [&](real& l1, real& l2, real& r1, real x1, real x2, real xr) {
    // T < 64: the wrap comes from lane T-1, resp. lane 0
    if constexpr (WM == 0) {
      const real e1 = lane_read(x1, T - 1), e2 = lane_read(x2, T - 1), er = lane_read(xr, 0);
      if (lane == 0) { l1 = e1; l2 = e2; }
      if (lane == T - 1) r1 = er;
    }
  };
  // left neighbours of x1 (this lane's X[K-1]) and x2 (X[K-2]), right
  // neighbour of xr (X[0])
  // Across arcs the wave's edge values are broadcast from lanes 63 and 0 and
  // written by every lane (one address, one value), and every lane reads its
  // neighbour's and selects: no divergent branch around the LDS accesses.
  auto halo3 = [&](real x1, real x2, real xr, real& l1, real& l2, real& r1) {
    l1 = lane_from_left(x1);
    l2 = lane_from_left(x2);
    r1 = lane_from_right(xr);
    if constexpr (MULTI) {
      const real e1 = lane_read(x1, 63), e2 = lane_read(x2, 63), er = lane_read(xr, 0);
      ex[xb][0][w] = e1; ex[xb][1][w] = e2; ex[xb][2][w] = er;
      __syncthreads();
      const real n1 = ex[xb][0][wl], n2 = ex[xb][1][wl], nr = ex[xb][2][wr];
      l1 = lane == 0 ? n1 : l1;
      l2 = lane == 0 ? n2 : l2;
      r1 = lane == 63 ? nr : r1;
      xb ^= 1;
    } else {
      fix_single(l1, l2, r1, x1, x2, xr);
    }
  };
  // left neighbour of a value that is already a left-neighbour value (K == 1,
  // W == 1)
  auto left_of = [&](real x) -> real {
    real r = lane_from_left(x);
    if constexpr (WM == 0) {
      const real e = lane_read(x, T - 1);
      if (lane == 0) r = e;
    }
    return r;
  };
  auto right1 = [&](real xr) -> real {
    real r = lane_from_right(xr);
    if constexpr (MULTI) {
      ex[xb][3][w] = lane_read(xr, 0);
      __syncthreads();
      const real nr = ex[xb][3][wr];
      r = lane == 63 ? nr : r;
      xb ^= 1;
    } else if constexpr (WM == 0) {
      const real e = lane_read(xr, 0);
      if (lane == T - 1) r = e;
    }
    return r;
  };
  auto left2 = [&](real x1, real x2, real& l1, real& l2) {
    l1 = lane_from_left(x1);
    l2 = lane_from_left(x2);
    if constexpr (MULTI) {
      const real e1 = lane_read(x1, 63), e2 = lane_read(x2, 63);
      ex[xb][0][w] = e1; ex[xb][1][w] = e2;
      __syncthreads();
      const real n1 = ex[xb][0][wl], n2 = ex[xb][1][wl];
      l1 = lane == 0 ? n1 : l1;
      l2 = lane == 0 ? n2 : l2;
      xb ^= 1;
    } else if constexpr (WM == 0) {
      const real e1 = lane_read(x1, T - 1), e2 = lane_read(x2, T - 1);
      if (lane == 0) { l1 = e1; l2 = e2; }
    }
  };

  // every global load of the ring is issued here, before any store: the
  // vector memory counter retires loads and stores in issue order, so a load
  // issued after a store could not be waited for without that store
  const int nq = ad ? 5 : 4, oh = nq - 2;   // oh: ThetaMomP
  real V[K], Dn[K], Qa[5][K];
  ld_chunk<real, K>(D.vtheta + my, V);
  ld_chunk<real, K>(D.t_dens + my, Dn);
  {
    const real* srcs[5] = {D.t_rmp, D.t_rmm, ad ? D.t_e : D.t_tmp, ad ? D.t_tmp : D.t_tmm, D.t_tmm};
#pragma unroll
    for (int o = 0; o < 5; ++o)
      if (o < nq) ld_chunk<real, K>(srcs[o] + my, Qa[o]);
  }
  real Lq[LAB ? K : 1];
  if constexpr (LAB) ld_chunk<real, K>(D.t_lab + my, Lq);

  // ring mean of vtheta (FP64 sum of the residuals)
  double acc = 0;
  if (on) {
#pragma unroll
    for (int k = 0; k < K; ++k) acc += (double)V[k];
  }
  acc = wave_sum(acc);
  if constexpr (MULTI) {
    if (lane == 0) redw[w] = acc;
    __syncthreads();
    acc = 0;
    for (int q = 0; q < W; ++q) acc += redw[q];
  }
  const double mres = acc / (double)ns;
  const real vmed = (real)(RT(vref)[rb + i] + mres);
  const real rmed = RT(rmed)[rb + i], irm = RT(invrmed)[rb + i];
  real ntilde, nround;
  const int nsh = fargo_shift_dev<real>(vmed, irm, s.dt, ns, ntilde, nround);
  const real dt = (real)s.dt;
  const real invdt = (real)1 / dt;
  const real dpinvns = (real)(2.0 * kPI / (double)ns);
  const real vconst = (ntilde - nround) * rmed * invdt * dpinvns;
  const real v2 = D.fast_transport ? vconst : (real)0;
  if (threadIdx.x == 0) { D.vmed[rb + i] = vmed; D.nshift[rb + i] = nsh; }
  int nsm = nsh % ns;
  if (nsm < 0) nsm += ns;
  const real dxt = RT(dxtheta)[rb + i];
  const real invdxt = (real)1 / dxt;
  const real dxrad = (RT(rsup)[rb + i] - RT(rinf)[rb + i]) * dt;
  const real invsurf = (real)1 / RT(surf)[rb + i];
  // AdvectSHIFTKernel: out[j] = in[j - Nshift], i.e. sector j lands at
  // j + Nshift
  const int o0 = [&] { int o = j0 + nsm; return o >= ns ? o - ns : o; }();
  auto store_shift = [&](real* dst, const real(&v)[K]) {
    if (!on) return;
    int o = o0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      dst[row + o] = v[k];
      o = (o + 1 == ns) ? 0 : o + 1;
    }
  };

  // van Leer slope without the 1/dxt factor (folded into the coefficients)
  auto slope = [&](real xm, real x0, real xp) -> real {
    const real dqm = x0 - xm, dqp = xp - x0;
    return (dqp * dqm > 0) ? dqp * dqm * vl_rcp(dqp + dqm) : (real)0;
  };
  // interface values (StarThetaKernel2) of X; per-sector upwind bits and
  // coefficients ((dxt - ksi) if ksi > 0, else (dxt + ksi)) / dxt
  auto star = [&](const real(&X)[K], unsigned pos, const real(&cf)[K], real(&out)[K]) {
    real xl1, xl2, xr1;
    halo3(X[K - 1], (K >= 2) ? X[K - 2] : X[0], X[0], xl1, xl2, xr1);
    if constexpr (K == 1) xl2 = left_of(xl1);   // W == 1 only
    real slp = slope(xl2, xl1, X[0]);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const real xm = (k == 0) ? xl1 : X[k - 1];
      const real xp = (k == K - 1) ? xr1 : X[k + 1];
      const real sl = slope(xm, X[k], xp);
      out[k] = ((pos >> k) & 1u) ? xm + cf[k] * slp : X[k] - cf[k] * sl;
      slp = sl;
    }
  };
  // the same with one velocity for the whole ring
  auto star_u = [&](const real(&X)[K], bool pos, real cf, real(&out)[K]) {
    real xl1, xl2, xr1;
    halo3(X[K - 1], (K >= 2) ? X[K - 2] : X[0], X[0], xl1, xl2, xr1);
    if constexpr (K == 1) xl2 = left_of(xl1);
    real slp = slope(xl2, xl1, X[0]);
    if (pos) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const real xm = (k == 0) ? xl1 : X[k - 1];
        const real xp = (k == K - 1) ? xr1 : X[k + 1];
        const real sl = slope(xm, X[k], xp);
        out[k] = xm + cf * slp;
        slp = sl;
      }
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const real xm = (k == 0) ? xl1 : X[k - 1];
        const real xp = (k == K - 1) ? xr1 : X[k + 1];
        out[k] = X[k] - cf * slope(xm, X[k], xp);
      }
    }
  };
  // flux divergence of one VanLeerTheta pass: F = star(work) * SV with SV =
  // dxrad * DensStar * v / Surf
  auto apply = [&](real(&X)[K], real(&F)[K], const real(&SV)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) F[k] *= SV[k];
    const real fr = right1(F[0]);
#pragma unroll
    for (int k = 0; k < K; ++k) X[k] += F[k] - ((k == K - 1) ? fr : F[k + 1]);
  };
  // the same with one coefficient for the whole ring
  auto apply_u = [&](real(&X)[K], real(&F)[K], real cu) {
#pragma unroll
    for (int k = 0; k < K; ++k) F[k] *= cu;
    const real fr = right1(F[0]);
#pragma unroll
    for (int k = 0; k < K; ++k) X[k] += F[k] - ((k == K - 1) ? fr : F[k + 1]);
  };
  // Pass 2 moves the ring by one uniform fraction of a cell, so its fluxes
  // are nearly equal from sector to sector and their differences, the
  // updates, are small.  Formed as products of whole values, each flux would
  // carry a rounding error of the size of the update itself.  They are
  // therefore formed against ring constants: the interface density as sbar
  // plus a residual S2r, a quantity's work as a ring value ab plus a residual
  // (ab = 1 for the density, the profile's specific energy for the energy,
  // 0 for the momenta), and the uniform part ab * sbar * v, which cancels
  // between neighbours, is left out:
  //   F = c2 * (ab * S2r + star(work - ab) * S2),  c2 = dxrad * v2 / Surf.
  // The arithmetic is the reference's to within the rounding of the terms.
  unsigned pos1 = 0;
  real cf1[K], SV1[K], invD1[K], S2r[K], SV2[K], invD2[K];
  const real ksi2 = v2 * dt;
  const bool pos2 = ksi2 > 0;
  const real cf2 = (pos2 ? dxt - ksi2 : dxt + ksi2) * invdxt;
  const real sbar = RT(sigmamed)[rb + i];
  const real c2 = dxrad * v2 * invsurf;
  // one quantity X through both passes: the momenta (ring value 0 of the
  // work: F = star(work) * S2 * c2 in pass 2)
  auto both = [&](real(&X)[K]) {
    real A[K], F[K];
#pragma unroll
    for (int k = 0; k < K; ++k) A[k] = X[k] * invD1[k];
    star(A, pos1, cf1, F);
    apply(X, F, SV1);
#pragma unroll
    for (int k = 0; k < K; ++k) A[k] = X[k] * invD2[k];
    star_u(A, pos2, cf2, F);
    apply(X, F, SV2);
  };
  // the energy, stored as a residual X of its ring profile e0: pass 1 takes
  // the work from the whole value (its fluxes are small), pass 2 splits it
  // as ab + (X + e0 - ab * density) / density
  auto both_energy = [&](real(&X)[K]) {
    const real e0 = RT(energymed)[rb + i];
    const real ab = e0 * vl_rcp(sbar);   // any value near the ring's e / density serves
    const real eb = fmadd(-ab, sbar, e0);   // e0 - ab * sbar, near 0
    real A[K], F[K];
#pragma unroll
    for (int k = 0; k < K; ++k) A[k] = (e0 + X[k]) * invD1[k];
    star(A, pos1, cf1, F);
    apply(X, F, SV1);
    // (X + e0 - ab * density) / density with density = sbar + Dn
#pragma unroll
    for (int k = 0; k < K; ++k) A[k] = (fmadd(-ab, Dn[k], eb) + X[k]) * invD2[k];
    star_u(A, pos2, cf2, F);
#pragma unroll
    for (int k = 0; k < K; ++k) F[k] = ab * S2r[k] + F[k] * (sbar + S2r[k]);
    apply_u(X, F, c2);
  };

  // FP32: the four momenta (RadMomP, RadMomM, ThetaMomP, ThetaMomM) go
  // through both passes in lockstep: each takes exactly `both`'s operations,
  // and each step's halo exchange carries all four, one barrier across arcs
  // where `both` takes one per quantity (16 -> 4 per ring).
  auto halo3x4 = [&](const real (&X)[4][K], real (&l1)[4], real (&l2)[4], real (&r1)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      l1[q] = lane_from_left(X[q][K - 1]);
      l2[q] = lane_from_left((K >= 2) ? X[q][K - 2] : X[q][0]);
      r1[q] = lane_from_right(X[q][0]);
    }
    if constexpr (MULTI) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const real e1 = lane_read(X[q][K - 1], 63), e2 = lane_read((K >= 2) ? X[q][K - 2] : X[q][0], 63);
        const real er = lane_read(X[q][0], 0);
        ex[xb][3 * q][w] = e1; ex[xb][3 * q + 1][w] = e2; ex[xb][3 * q + 2][w] = er;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const real n1 = ex[xb][3 * q][wl], n2 = ex[xb][3 * q + 1][wl], nr = ex[xb][3 * q + 2][wr];
        l1[q] = lane == 0 ? n1 : l1[q];
        l2[q] = lane == 0 ? n2 : l2[q];
        r1[q] = lane == 63 ? nr : r1[q];
      }
      xb ^= 1;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        fix_single(l1[q], l2[q], r1[q], X[q][K - 1], (K >= 2) ? X[q][K - 2] : X[q][0], X[q][0]);
    }
  };
  auto star4 = [&](const real (&X)[4][K], unsigned pos, const real (&cf)[K], real (&out)[4][K]) {
    real xl1[4], xl2[4], xr1[4];
    halo3x4(X, xl1, xl2, xr1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if constexpr (K == 1) xl2[q] = left_of(xl1[q]);
      real slp = slope(xl2[q], xl1[q], X[q][0]);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const real xm = (k == 0) ? xl1[q] : X[q][k - 1];
        const real xp = (k == K - 1) ? xr1[q] : X[q][k + 1];
        const real sl = slope(xm, X[q][k], xp);
        out[q][k] = ((pos >> k) & 1u) ? xm + cf[k] * slp : X[q][k] - cf[k] * sl;
        slp = sl;
      }
    }
  };
  auto star_u4 = [&](const real (&X)[4][K], bool pos, real cf, real (&out)[4][K]) {
    real xl1[4], xl2[4], xr1[4];
    halo3x4(X, xl1, xl2, xr1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if constexpr (K == 1) xl2[q] = left_of(xl1[q]);
      real slp = slope(xl2[q], xl1[q], X[q][0]);
      if (pos) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const real xm = (k == 0) ? xl1[q] : X[q][k - 1];
          const real xp = (k == K - 1) ? xr1[q] : X[q][k + 1];
          const real sl = slope(xm, X[q][k], xp);
          out[q][k] = xm + cf * slp;
          slp = sl;
        }
      } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const real xm = (k == 0) ? xl1[q] : X[q][k - 1];
          const real xp = (k == K - 1) ? xr1[q] : X[q][k + 1];
          out[q][k] = X[q][k] - cf * slope(xm, X[q][k], xp);
        }
      }
    }
  };
  auto apply4 = [&](real (&X)[4][K], real (&F)[4][K], const real (&SV)[K]) {
    real fr[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int k = 0; k < K; ++k) F[q][k] *= SV[k];
      fr[q] = lane_from_right(F[q][0]);
    }
    if constexpr (MULTI) {
#pragma unroll
      for (int q = 0; q < 4; ++q) ex[xb][q][w] = lane_read(F[q][0], 0);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const real nr = ex[xb][q][wr];
        fr[q] = lane == 63 ? nr : fr[q];
      }
      xb ^= 1;
    } else if constexpr (WM == 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const real e = lane_read(F[q][0], 0);
        if (lane == T - 1) fr[q] = e;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < K; ++k) X[q][k] += F[q][k] - ((k == K - 1) ? fr[q] : F[q][k + 1]);
  };
  auto both4 = [&](real (&M)[4][K]) {
    real A[4][K], F[4][K];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < K; ++k) A[q][k] = M[q][k] * invD1[k];
    star4(A, pos1, cf1, F);
    apply4(M, F, SV1);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < K; ++k) A[q][k] = M[q][k] * invD2[k];
    star_u4(A, pos2, cf2, F);
    apply4(M, F, SV2);
  };

  // pass-1 velocity (ComputeResiduals); STANDARD transport folds the
  // constant part in and the second pass moves nothing
#pragma unroll
  for (int k = 0; k < K; ++k) {
    real v = (real)((double)V[k] - mres);
    if (!D.fast_transport) v = vconst + v;
    V[k] = v;
    const real ksi = v * dt;
    pos1 |= (ksi > 0 ? 1u : 0u) << k;
    cf1[k] = ((ksi > 0) ? dxt - ksi : dxt + ksi) * invdxt;
  }
  // density: DensStar and pass 1 (the density's work is exactly 1 but for
  // densities below the 1e-20's resolution).  Dn holds the density as the
  // field does, a residual of sbar = S0: the van Leer slopes of the residual
  // are those of the density, and its star plus sbar is the density's.
  star(Dn, pos1, cf1, SV1);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    SV1[k] = dxrad * (sbar + SV1[k]) * V[k] * invsurf;
    invD1[k] = vl_rcp(sbar + Dn[k] + (real)1e-20);
  }
  {
    real A[K], F[K];
#pragma unroll
    for (int k = 0; k < K; ++k) A[k] = dens_work(sbar + Dn[k]);
    star(A, pos1, cf1, F);
    apply(Dn, F, SV1);
  }
  // pass-2 interface densities from the density after pass 1, as residuals
  // of sbar
  star_u(Dn, pos2, cf2, S2r);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    SV2[k] = (sbar + S2r[k]) * c2;
    invD2[k] = vl_rcp(sbar + Dn[k] + (real)1e-20);
  }
  // quantities in the order RadMomP, RadMomM, [energy], ThetaMomP (held),
  // ThetaMomM (held); the energy is stored as a residual of its ring profile
  real H[K], Qn[K];
  if constexpr (sizeof(real) == 4) {
    constexpr int m2 = AD ? 3 : 2, m3 = AD ? 4 : 3;
    real M[4][K];
#pragma unroll
    for (int k = 0; k < K; ++k) { M[0][k] = Qa[0][k]; M[1][k] = Qa[1][k]; M[2][k] = Qa[m2][k]; M[3][k] = Qa[m3][k]; }
    both4(M);
#pragma unroll
    for (int k = 0; k < K; ++k) { Qa[0][k] = M[0][k]; Qa[1][k] = M[1][k]; Qa[m2][k] = M[2][k]; Qa[m3][k] = M[3][k]; }
    if (ad) both_energy(Qa[2]);
    store_shift(D.t_rmp, Qa[0]);
    store_shift(D.t_rmm, Qa[1]);
    if (ad) store_shift(D.energy, Qa[2]);
  } else {
    // FP64 takes the quantities one at a time: four in lockstep spill
    real* dsts[3] = {D.t_rmp, D.t_rmm, D.energy};
#pragma unroll
    for (int o = 0; o < 5; ++o) {
      if (o >= nq) break;
      if (ad && o == 2) both_energy(Qa[o]);
      else both(Qa[o]);
      if (o < oh) store_shift(dsts[o], Qa[o]);
    }
  }
  if constexpr (LAB) both(Lq);
#pragma unroll
  for (int k = 0; k < K; ++k) { H[k] = ad ? Qa[3][k] : Qa[2][k]; Qn[k] = ad ? Qa[4][k] : Qa[3][k]; }
  // density, pass 2 (work 1 + residual)
  {
    real A[K], F[K];
#pragma unroll
    for (int k = 0; k < K; ++k) A[k] = dens_work(sbar + Dn[k]) - (real)1;
    star_u(A, pos2, cf2, F);
#pragma unroll
    for (int k = 0; k < K; ++k) F[k] = S2r[k] + F[k] * (sbar + S2r[k]);
    apply_u(Dn, F, c2);
  }
  store_shift(D.dens, Dn);
  if constexpr (LAB) {
#pragma unroll
    for (int k = 0; k < K; ++k) Lq[k] = Lq[k] / (sbar + Dn[k]);
    store_shift(D.label, Lq);
  }
  // azimuthal velocity of ComputeVelocitiesKernel at the face j-1/2:
  // (TMP[j-1] + TMM[j]) / (Sigma[j] + Sigma[j-1] + 1e-15) / R - R*Omega; with
  // the momenta stored as residuals of density * l0 this is vref + the
  // residual below, exactly
  const double rd = RT(rmed_d)[rb + i];
  const double l0 = (RT(vref)[rb + i] + rd * s.omega_frame) * rd;
  real hl, dl;
  left2(H[K - 1], Dn[K - 1], hl, dl);
  if constexpr (sizeof(real) == 4) {
    // the quotient is a small residual of vtheta: FP32 with a reciprocal
    // keeps it far inside the FP32 field's own rounding
    const real l0e = (real)(l0 * 1e-15);
    const real sbar2 = sbar + sbar;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const real hm = (k == 0) ? hl : H[k - 1];
      const real dm = (k == 0) ? dl : Dn[k - 1];
      V[k] = (hm + Qn[k] - l0e) * vl_rcp(sbar2 + (Dn[k] + dm) + (real)1e-15) * irm;
    }
  } else {
    const double ird = 1.0 / rd;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double hm = (double)((k == 0) ? hl : H[k - 1]);
      const double dm = (double)((k == 0) ? dl : Dn[k - 1]);
      V[k] = (real)((hm + (double)Qn[k] - l0 * 1e-15) / (2.0 * (double)sbar + (double)Dn[k] + dm + 1e-15) * ird);
    }
  }
  store_shift(D.vtheta, V);
}

template <class real, int K, int WM, bool LAB, bool AD>
__global__ void __launch_bounds__(WM == 2 ? 1024 : 64) k_trans_theta_arc(Dev<real> D, const RingTile* __restrict__ tiles,
                                                                        int W) {
  theta_arc_ring<real, K, WM, LAB, AD>(D, tiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)], W);
}

// The multi-arc ring groups of a batch (K = 2 or 3 sectors per lane, W = 2..16
// arcs) in one launch: group g owns blocks [first[g], first[g+1]), a multiple
// of 64 (so each block keeps its XCD under the grouped map), and the block
// has 64 * max W threads; the waves past a ring's W arcs end at once (a
// barrier waits only for the waves that have not ended).
#define FGX_THETA_MULTI 8
struct ThetaMulti {
  int ng;
  int first[FGX_THETA_MULTI + 1];
  int k[FGX_THETA_MULTI], w[FGX_THETA_MULTI];
  const RingTile* tiles[FGX_THETA_MULTI];
};
template <class real, bool LAB, bool AD>
__global__ void __launch_bounds__(1024) k_trans_theta_multi(Dev<real> D, ThetaMulti M) {
  const int b = (int)blockIdx.x;
  int g = 0;
  while (g + 1 < M.ng && b >= M.first[g + 1]) ++g;
  const int W = M.w[g];
  if ((int)threadIdx.x >= 64 * W) return;
  const RingTile t = M.tiles[g][xcd_tile(b - M.first[g], M.first[g + 1] - M.first[g])];
  if (M.k[g] == 3) theta_arc_ring<real, 3, 2, LAB, AD>(D, t, W);
  else theta_arc_ring<real, 2, 2, LAB, AD>(D, t, W);
}

// split of a ring of nsec sectors: W arcs of 64*K sectors (K in {3, 2, 4},
// 2 <= W <= 16), else one wave with nsec = T*K, W = 0 when T = 64 and W = 1
// when T < 64; K = 0 when nothing fits (shared-memory kernel)
template <class real>
inline void theta_split(int nsec, int& K, int& W, bool arcs = true) {
  static const int ka[] = {3, 2, 4};
  for (int k : ka)
    if (arcs && nsec % (64 * k) == 0 && nsec / (64 * k) >= 2 && nsec / (64 * k) <= 16) {
      K = k; W = nsec / (64 * k); return;
    }
  static const int kf[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 16};
  static const int kd[] = {1, 2, 3, 4, 5, 6, 8};
  const int* ks = sizeof(real) == 4 ? kf : kd;
  const int n = sizeof(real) == 4 ? 10 : 7;
  W = 1;
  for (int a = 0; a < n; ++a)
    if (nsec % ks[a] == 0 && nsec / ks[a] <= 64) { K = ks[a]; W = (nsec / ks[a] == 64) ? 0 : 1; return; }
  K = 0;
}

}  // namespace fgx
