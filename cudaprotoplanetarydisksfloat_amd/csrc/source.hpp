// source.hpp -- the source half of a substep in one pass over the state:
// the planets' potential (FillForcesArraysKernel, Kernels.cu:1322-1349),
// Substep1 (Kernels.cu:52-88), the viscous stresses (ViscousTermsKernelDRP /
// TAURP, Kernels.cu:843-880), the viscous velocity update with the
// sub-Keplerian boundary rows (Kernels.cu:142-160, 1490-1503) and Substep2
// (Kernels.cu:570-621).
//
// One wave owns a chunk of FGX_SRC_OUT sectors of a strip of rings plus four
// extra lanes on each side, and walks the strip outwards running four stages
// one ring apart:
//   A  ring k     potential, pressure, Substep1 -> VradInt, VthetaInt
//   B  ring k-1   stresses from VradInt, VthetaInt
//   C  ring k-2   viscous update of VradInt, VthetaInt; boundary rows
//   D  ring k-3   Substep2 -> new Vrad, Vtheta, [EnergyInt]   (stored)
// A stage's azimuthal neighbours come from the adjacent lanes (DPP), its
// radial neighbours from the registers of the previous iterations, so the
// intermediate fields (potential, VradInt, VthetaInt, the stresses) never
// reach memory.  Each stage needs one more lane on each side than the next,
// hence the four extra lanes.  Radially, the stored rings [i0, i1) need
// stage C's VradInt on rings i0-1 .. i1, which needs the stresses of rings
// i0-2 .. i1 (only their TAU_rr, TAU_pp below i0-1) and so stage A on rings
// i0-2 .. i1+1, with ring i0-3's pressure and potential for A's radial
// differences.  The arithmetic is that of the separate kernels, operation
// for operation.
#pragma once
#include "transport.hpp"
#ifndef FGX_SRC_ATTR
#define FGX_SRC_ATTR
#endif

namespace fgx {

#define FGX_SRC_OUT 57   // sectors stored per wave: lanes 4..60
#define FGX_SRC_STAGE 32  // rings of coefficients staged per strip: rs + 8
#define FGX_SRC_RINGS_MAX (FGX_SRC_STAGE - 8)

// SG: self-gravitating disks.  Their acceleration field (formed from the
// density before the launch) is added to VradInt and VthetaInt right after
// Substep1, as Update_sgvelocityKernel (Kernels.cu:1024-1041) does between
// Substep1 and the viscous stresses, in FP64.
template <class real, bool SG>
__global__ void __launch_bounds__(64) FGX_SRC_ATTR k_source(Dev<real> D, const CellTile* __restrict__ tiles, int rs) {
  const CellTile t = tiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)];
  if (t.d < 0) return;
  const FGX_AS4 DiskConst& c = cref(D.dc + t.d);
  const FGX_AS4 DiskState& s = cref(D.st + t.d);
  if (skip_disk(D, s, t.d)) return;
  const int ns = c.nsec, nr = c.nrad;
  const long long fb = c.fo;
  const int rb = c.ro;
  const int lane = (int)threadIdx.x;
  // this lane's sector; lanes 4..60 of the chunk are stored
  const int first = t.jb * FGX_SRC_OUT;
  int j = (first - 4 + lane) % ns;
  if (j < 0) j += ns;
  const int jp = (j + 1 == ns) ? 0 : j + 1, jm = (j == 0) ? ns - 1 : j - 1;
  const bool out = lane >= 4 && lane <= 60 && first + lane - 4 < ns;
  const int i0 = t.i, i1 = (t.i + rs < nr) ? t.i + rs : nr;
  const bool visc = D.viscous, ad = D.adiabatic;
  const double dtd = s.dt, dom = s.omega_frame - c.omega0;
  const real dt = (real)dtd;
  const double g1 = c.adiabaticindex - 1.0;
  const real invdphi = (real)((double)ns / (2.0 * kPI));
  const real third = (real)(1.0 / 3.0);
  const real cv2 = (real)(kCVNR * kCVNR);

  // the ring coefficients of rings [i0-4, i0+rs+3) staged in LDS, one lane
  // per ring: inside the loop they are read from LDS into vector registers,
  // instead of some fifty per-ring values and their table addresses being
  // held in scalar registers across the loop
  // acc0, u0, gq0: the ring profile's own radial acceleration, interface
  // velocity and pressure-gradient coefficient (Substep1 below); S0 the density
  // profile of which the density field is a residual
  // cr_: Substep1's departure terms are formed in the field precision, from
  // the FP64 tables rounded once here: cs^2, 1/(Rmed_i - Rmed_i-1), acc0, u0,
  // gq0, 1/Rinf and Rinf * (the frame's change since set-up)
  __shared__ double cd_[10][FGX_SRC_STAGE];  // cs2_d rmed_d invdiffrmed_d acc0 vref rinf_d invrinf_d invrmed_d u0 gq0
  __shared__ real cf_[11][FGX_SRC_STAGE];    // 1/dxtheta supp_torque invrmed invdiffrsup rsup rinf visc invrinf rmed invdiffrmed S0
  __shared__ real cr_[7][FGX_SRC_STAGE];     // cs2 idr acc0 u0 gq0 invrinf rinf*dom
  const int rlo = i0 - 4;
  {
    const int r = rlo + lane;
    if (r >= 0 && r < nr && lane < rs + 8 && lane < FGX_SRC_STAGE) {
      cd_[0][lane] = RT(cs2_d)[rb + r];
      cd_[1][lane] = RT(rmed_d)[rb + r];
      cd_[2][lane] = RT(invdiffrmed_d)[rb + r];
      cd_[3][lane] = RT(acc0)[rb + r];
      cd_[8][lane] = RT(u0)[rb + r];
      cd_[9][lane] = RT(gq0)[rb + r];
      cf_[10][lane] = RT(sigmamed)[rb + r];
      const S1Ring<real> q = s1_ring(D, rb + r, s.omega_frame - c.omega0);
      cr_[0][lane] = q.cs2; cr_[1][lane] = q.idr; cr_[2][lane] = q.acc0; cr_[3][lane] = q.u0;
      cr_[4][lane] = q.gq0; cr_[5][lane] = q.invrinf; cr_[6][lane] = q.rinfdom;
      cd_[4][lane] = RT(vref)[rb + r];
      cd_[5][lane] = RT(rinf_d)[rb + r];
      cd_[6][lane] = RT(invrinf_d)[rb + r];
      cd_[7][lane] = RT(invrmed_d)[rb + r];
      cf_[0][lane] = (real)1 / RT(dxtheta)[rb + r];
      cf_[1][lane] = RT(supp_torque)[rb + r];
      cf_[2][lane] = RT(invrmed)[rb + r];
      cf_[3][lane] = RT(invdiffrsup)[rb + r];
      cf_[4][lane] = RT(rsup)[rb + r];
      cf_[5][lane] = RT(rinf)[rb + r];
      cf_[6][lane] = RT(visc)[rb + r];
      cf_[7][lane] = RT(invrinf)[rb + r];
      cf_[8][lane] = RT(rmed)[rb + r];
      cf_[9][lane] = RT(invdiffrmed)[rb + r];
    }
  }
  __syncthreads();
#define FGX_CD(q, r) cd_[q][(r) - rlo]
#define FGX_CF(q, r) cf_[q][(r) - rlo]
#define FGX_CR(q, r) cr_[q][(r) - rlo]

  // ring k-1 inputs of stage A (sgA1 the whole density, srA1 its residual)
  real sgA1 = 0, srA1 = 0, vtA1 = 0, vtpA1 = 0, pA1 = 0;
  double phiA1 = 0;
  // density and energy of rings k-1 .. k-4 (index 1..4), A's outputs of
  // rings k-1, k-2, the stresses of rings k-2, k-3, C's outputs of rings
  // k-3, k-4
  real sg1 = 0, sg2 = 0, sg3 = 0, sg4 = 0, e1 = 0, e2 = 0, e3 = 0;
  real vri1 = 0, vri2 = 0, vti1 = 0, vti2 = 0;
  real trr1 = 0, trr2 = 0, tpp1 = 0, tpp2 = 0, trp1 = 0;
  real vrc1 = 0, vrc2 = 0, vtc1 = 0;

  // the sector's FP64 cosine and sine, loaded once: loaded inside the loop,
  // each ring would wait for them and so for every load in flight
  const double cj = RT(cosj_d)[c.so + j], sj = RT(sinj_d)[c.so + j];
  // consumed here, so that the loop does not wait for them (and, loads
  // retiring in order, for everything behind them) on every ring
  asm volatile("" ::"v"(cj), "v"(sj));
  // planet 0's potential constants and the indirect term, read once: in the
  // ring loop they would be fetched again, and waited for, on every ring.
  // The arithmetic is potential_planets_xy's, in the same order.
  const int nb = c.nb;
  const bool indirect = D.indirect;
  double px0 = s.pot_x[0], py0 = s.pot_y[0], pm0 = s.pot_m[0], ps0 = s.pot_sm2[0], pi0 = s.pot_inv3[0];
  double indx = s.ind_x, indy = s.ind_y;
  // held in scalar registers: left as plain loads, the compiler re-reads
  // some of them from memory inside the loop and waits for them on every ring
  asm("" : "+s"(px0), "+s"(py0), "+s"(pm0), "+s"(ps0), "+s"(pi0), "+s"(indx), "+s"(indy));
  auto potential = [&](double x, double y) -> double {
    double pot = 0;
    if (nb > 0) {
      const double dist = (x - px0) * (x - px0) + (y - py0) * (y - py0);
      double p = -pm0 * rsqrt_d<real>(dist + ps0);
      if (indirect) p += pm0 * pi0 * (x * px0 + y * py0);
      pot += p;
    }
    for (int q = 1; q < nb; ++q) {
      const double xp = s.pot_x[q], yp = s.pot_y[q], mp = s.pot_m[q];
      const double dist = (x - xp) * (x - xp) + (y - yp) * (y - yp);
      double p = -mp * rsqrt_d<real>(dist + s.pot_sm2[q]);
      if (indirect) p += mp * s.pot_inv3[q] * (x * xp + y * yp);
      pot += p;
    }
    pot -= indx * x + indy * y;
    return pot;
  };
  // the pressure's departure from the ring profile's: (Sigma - S0) cs^2, or
  // (gamma - 1)(e - EnergyMed), from the residual fields
  const real g1r = (real)g1;
  auto pressure = [&](real sr, real e, int k) -> real { return ad ? g1r * e : sr * FGX_CR(0, k); };
  const int ks = (i0 - 2 > 0) ? i0 - 2 : 0;
  if (ks > 0) {   // prime stage A with ring ks-1
    const int k = ks - 1;
    srA1 = AT(D.dens, k, j);
    sgA1 = FGX_CF(10, k) + srA1;
    vtA1 = AT(D.vtheta, k, j);
    vtpA1 = AT(D.vtheta, k, jp);
    pA1 = pressure(srA1, ad ? (real)AT(D.energy, k, j) : (real)0, k);
    phiA1 = potential(FGX_CD(1, k) * cj, FGX_CD(1, k) * sj);
  }
  // past the outer edge A and B are skipped and C yields VradInt(nrad) = 0
  const int ke = i1 + 2;
  // Stage A's inputs are fetched one ring ahead, so that the wave does not
  // wait for them at the top of every iteration.  The fetch also carries
  // the density of column 0 of the ring (the corner density of the
  // reference's TAURP, used by stage B one ring later): a load used in the
  // iteration that issues it would make the wave wait for every load in
  // flight, since loads retire in issue order.  Rings past the last one
  // stage A reads (i1+1, or the outer edge) load that ring again, which the
  // caches still hold (their values are never used): a fixed number of loads
  // per iteration lets the compiler wait for exactly the ring it needs,
  // without fetching rings no stage reads.
  real nsg, nvr, nvt, nvtp, ne, nc0, nar = 0, nat = 0;
  const int klast = (i1 + 1 < nr) ? i1 + 1 : nr - 1;
  auto fetch = [&](int k) {
    const int kk = k < klast ? k : klast;
    nsg = AT(D.dens, kk, j);
    nvr = AT(D.vrad, kk, j);
    nvt = AT(D.vtheta, kk, j);
    nvtp = AT(D.vtheta, kk, jp);
    ne = ad ? (real)AT(D.energy, kk, j) : (real)0;
    nc0 = AT(D.dens, kk, 0);
    if constexpr (SG) {
      nar = AT(D.sg_accr, kk, j);
      nat = AT(D.sg_acct, kk, j);
    }
  };
  // the self-gravity acceleration of ring k-1 at this lane's sector
  real arA1 = 0, atA1 = 0;
  if constexpr (SG) {
    if (ks > 0) {
      arA1 = AT(D.sg_accr, ks - 1, j);
      atA1 = AT(D.sg_acct, ks - 1, j);
    }
  }
  real c0A1 = 0;   // column-0 density of ring k-1 (whole)
  if (ks > 0) c0A1 = FGX_CF(10, ks - 1) + (real)AT(D.dens, ks - 1, 0);
  fetch(ks);
  auto ring_step = [&](const int k) {
    // ---------------------------------------------------------- stage A
    real sg0 = 0, e0 = 0, vri0 = 0, vti0 = 0;
    const real csg = nsg, cvr = nvr, cvt = nvt, cvtp = nvtp, ce = ne, cc0 = nc0, car = nar, cat = nat;
    fetch(k + 1);
    if (k < nr && k <= i1 + 1) {
      const real sr = csg;
      sg0 = FGX_CF(10, k) + sr;
      const real vr = cvr;
      const real vt = cvt, vtp = cvtp;
      if (ad) e0 = ce;
      const real p = pressure(sr, e0, k);
      const double phi = potential(FGX_CD(1, k) * cj, FGX_CD(1, k) * sj);
      if (D.write_pressure && out && k >= i0 && k < i1)
        AT(D.pressure, k, j) = (real)((double)FGX_CF(10, k) * FGX_CD(0, k) + (double)p);
      if (k > 0) {
        const S1Ring<real> q{FGX_CR(0, k), FGX_CR(1, k), FGX_CR(2, k), FGX_CR(3, k), FGX_CR(4, k), FGX_CR(5, k),
                             FGX_CR(6, k)};
        vri0 = substep1_vr(vr, p, pA1, sr, srA1, sg0, sgA1, phi, phiA1, vt + vtp + vtA1 + vtpA1, q, dt);
      }
      // azimuthal (the star's potential is the same at j and j-1)
      const real invdxt = FGX_CF(0, k);
      const real sigl = lane_from_left(sg0);
      const real pl = lane_from_left(p);
      const double phil = lane_from_left(phi);
      const real gradp = azim_gradp(p, pl, sg0, sigl) * invdxt;
      const real gradphi = (real)(phi - phil) * invdxt;
      real v = vt - (real)dtd * (gradp + gradphi);
      v += (real)dtd * FGX_CF(1, k);
      if constexpr (SG) {
        // the radial acceleration interpolated to the inner interface; the
        // azimuthal one averaged with sector j-1, which at j = 0 the
        // reference takes from ring k-1's last sector (a linear index of
        // -1 past the ring start) unless the quirk fix asks for ring k's
        if (k > 0)
          vri0 = (real)sg_vrad_new((double)vri0, dtd, FGX_CD(5, k), FGX_CD(1, k - 1), FGX_CD(1, k), (double)car,
                                   (double)arA1, FGX_CD(2, k));
        const real atl = lane_from_left(cat), atl1 = lane_from_left(atA1);
        const real am = (j == 0 && k > 0 && !D.fix_sgwrap) ? atl1 : atl;
        v = (real)sg_vtheta_new((double)v, dtd, (double)cat, (double)am);
        arA1 = car;
        atA1 = cat;
      }
      vti0 = v;
      sgA1 = sg0; srA1 = sr; vtA1 = vt; vtpA1 = vtp; pA1 = p; phiA1 = phi;
    }
    // ---------------------------------------------------------- stage B
    const int b = k - 1;
    real trr0 = 0, tpp0 = 0, trp0 = 0;
    // ring i0-2's TAU_rp reads stage A's ring i0-3, which is not formed:
    // it is never used (stage C's VthetaInt of ring i0-1 is not stored)
    if (visc && b >= 0 && b < nr && b >= i0 - 2 && b <= i1) {
      const real irm = FGX_CF(2, b);
      const real vr0 = vri1, vr1 = vri0;
      const real vt0 = vti1, vtp = lane_from_right(vti1);
      const real drr = (vr1 - vr0) * FGX_CF(3, b);
      const real dpp = (vtp - vt0) * invdphi * irm + (real)0.5 * (vr1 + vr0) * irm;
      real divv = (vr1 * FGX_CF(4, b) - vr0 * FGX_CF(5, b)) * FGX_CF(3, b) * irm;
      divv += (vtp - vt0) * invdphi * irm;
      const real nu = FGX_CF(6, b);
      const real sig = sg1;
      trr0 = (real)2 * sig * nu * (drr - third * divv);
      tpp0 = (real)2 * sig * nu * (dpp - third * divv);
      if (b > 0) {
        const double w0 = (FGX_CD(4, b) + (double)vt0) * FGX_CD(7, b);
        const double w1 = (FGX_CD(4, b - 1) + (double)vti2) * FGX_CD(7, b - 1);
        const real drp = (real)0.5 * ((real)(FGX_CD(5, b) * (w0 - w1) * FGX_CD(2, b)) +
                                      (vr0 - lane_from_left(vri1)) * invdphi * FGX_CF(7, b));
        // corner density: column 0 of ring b as the reference indexes it
        // (Kernels.cu:876), or column j-1 with the quirk fix
        const real sc = D.fix_taurp ? lane_from_left(sig) : c0A1;
        trp0 = (real)(2.0 * 0.25) * (sig + sg2 + sc + lane_from_left(sg2)) * nu * drp;
      }
    }
    // ---------------------------------------------------------- stage C
    const int cr = k - 2;
    real vrc0 = vri2, vtc0 = vti2;
    if (cr >= 0 && cr >= i0 - 1) {
      if (cr >= nr) {
        vrc0 = 0;
      } else {
        if (visc && cr > 0 && cr < nr - 1) {
          const real sgl = lane_from_left(sg2);
          vtc0 += qdiv(dt * FGX_CF(2, cr) *
                           ((FGX_CF(4, cr) * trp0 - FGX_CF(5, cr) * trp1) * FGX_CF(3, cr) +
                            (tpp1 - lane_from_left(tpp1)) * invdphi + (real)0.5 * (trp1 + trp0)),
                       (real)0.5 * (sg2 + sgl));
        }
        if (visc && cr > 0) {
          vrc0 += qdiv(dt * FGX_CF(7, cr) *
                           ((FGX_CF(8, cr) * trr1 - FGX_CF(8, cr - 1) * trr2) * FGX_CF(9, cr) +
                            (lane_from_right(trp1) - trp1) * invdphi - (real)0.5 * (tpp1 + tpp2)),
                       (real)0.5 * (sg2 + sg3));
        }
        if (!D.evanescent) {
          if (cr == 0) {
            const double v = s.vkep_in - (D.self_gravity ? RT(rmed_d)[rb] * D.axisg[rb] : 0.0);
            vtc0 = (real)(sqrt(v) - RT(rmed_d)[rb] * s.omega_frame - RT(vref)[rb]);
          }
          if (cr == nr - 1) {
            const double v = s.vkep_out - (D.self_gravity ? RT(rmed_d)[rb + nr - 1] * D.axisg[rb + nr - 1] : 0.0);
            vtc0 = (real)(sqrt(v) - RT(rmed_d)[rb + nr - 1] * s.omega_frame - RT(vref)[rb + nr - 1]);
          }
        }
      }
    }
    // ---------------------------------------------------------- stage D
    const int d = k - 3;
    if (d >= i0 && d < i1) {
      const real sig = sg3;
      const real vri = vrc1, vrp = vrc0;
      real dv = vrp - vri;
      const real qr = dv < 0 ? cv2 * sig * dv * dv : (real)0;
      const real vti = vtc1, vtp = lane_from_right(vtc1), vtm = lane_from_left(vtc1);
      dv = vtp - vti;
      const real qt = dv < 0 ? cv2 * sig * dv * dv : (real)0;
      const real sigl = lane_from_left(sig);
      dv = vti - vtm;
      const real qtl = dv < 0 ? cv2 * sigl * dv * dv : (real)0;
      const real invdxt = FGX_CF(0, d);
      real vrn = 0;
      if (d > 0) {
        const real sigm = sg4;
        dv = vri - vrc2;
        const real qrm = dv < 0 ? cv2 * sigm * dv * dv : (real)0;
        vrn = vri - qdiv(dt * (real)2, sig + sigm) * (qr - qrm) * FGX_CF(9, d);
      }
      const real vtn = vti - qdiv(dt * (real)2, sig + sigl) * (qt - qtl) * invdxt;
      if (out) {
        // into the spare VradInt / VthetaInt arrays, swapped with Vrad /
        // Vtheta after the launch: the neighbouring strips still read this
        // strip's old velocities as their halo
        AT(D.vradint, d, j) = vrn;
        AT(D.vthetaint, d, j) = vtn;
        if (ad)
          AT(D.energyint, d, j) = e3 - dt * qr * (vrp - vri) * FGX_CF(3, d) - dt * qt * (vtp - vti) * invdxt;
      }
    }
    // ---------------------------------------------------------- shift
    c0A1 = FGX_CF(10, k < nr ? k : nr - 1) + cc0;
    sg4 = sg3; sg3 = sg2; sg2 = sg1; sg1 = sg0;
    e3 = e2; e2 = e1; e1 = e0;
    vri2 = vri1; vri1 = vri0; vti2 = vti1; vti1 = vti0;
    trr2 = trr1; trr1 = trr0; tpp2 = tpp1; tpp1 = tpp0; trp1 = trp0;
    vrc2 = vrc1; vrc1 = vrc0; vtc1 = vtc0;
  };
  // FP32: two rings per trip, so that the register shifts of period two
  // cost no moves (an odd count runs one more ring past ke, which stores
  // nothing); FP64 keeps one ring per trip, where two would cost occupancy
  if constexpr (sizeof(real) == 4) {
    for (int k = ks; k <= ke; k += 2) {
      ring_step(k);
      ring_step(k + 1);
    }
  } else {
    for (int k = ks; k <= ke; ++k) ring_step(k);
  }
}
#undef FGX_CD
#undef FGX_CF
#undef FGX_CR

}  // namespace fgx
