// setup.cpp -- host-side FP64 set-up of one disk: ring geometry, analytic
// profiles and initial fields.  The values are rounded once to the disk's
// arithmetic type when uploaded.
#include "fgx_internal.h"

#include <cmath>
#include <cstdint>

namespace fgx {

// AspectRatioHost (Viscosity.cu:39-52) at time t (PhysicalTime-PhysicalTimeInitial)
double aspect_ratio(const fgx_params& P, double r, double t) {
  double h = P.aspectratio;
  double rmin = P.transitionradius - P.transitionwidth * P.aspectratio;
  double rmax = P.transitionradius + P.transitionwidth * P.aspectratio;
  double scale = 1.0 + t * P.lambdadoubling;
  rmin *= scale; rmax *= scale;
  if (r < rmin) h *= P.transitionratio;
  if (r >= rmin && r <= rmax) h *= std::exp((rmax - r) / (rmax - rmin) * std::log((double)P.transitionratio));
  return h;
}

// FViscosity (Viscosity.cu:19-37)
double fviscosity(const fgx_params& P, double r, double csmean, double t) {
  double nu = P.viscosity;
  if (P.viscosity_alpha) nu = P.alphaviscosity * csmean * csmean * std::pow(r, 1.5);
  double rmin = P.cavityradius - P.cavitywidth * P.aspectratio;
  double rmax = P.cavityradius + P.cavitywidth * P.aspectratio;
  double scale = 1.0 + t * P.lambdadoubling;
  rmin *= scale; rmax *= scale;
  if (r < rmin) nu *= P.cavityratio;
  if (r >= rmin && r <= rmax) nu *= std::exp((rmax - r) / (rmax - rmin) * std::log((double)P.cavityratio));
  return nu;
}

double psys_frequency(const fgx_params& P, const fgx_planets& s) {
  (void)P;
  if (s.nb == 0) return 0.0;
  double x = s.x[0], y = s.y[0], vx = s.vx[0], vy = s.vy[0];
  double m = s.mass[0] + 1.;
  double h = x * vy - y * vx;
  double d = std::sqrt(x * x + y * y);
  double Ax = x * vy * vy - y * vx * vy - kG * m * x / d;
  double Ay = y * vx * vx - x * vx * vy - kG * m * y / d;
  double e = std::sqrt(Ax * Ax + Ay * Ay) / m;
  double a = h * h / kG / m / (1. - e * e);
  return std::sqrt(m / a / a / a);
}

static uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int setup_grid_and_density(const fgx_params& P, const fgx_planets& pl, HostDisk& h) {
  h.P = P;
  h.sys = pl;
  const int NR = P.nrad, NS = P.nsec;
  h.nrad = NR; h.nsec = NS;
  for (auto* v : {&h.radii, &h.rinf, &h.rsup, &h.rmed, &h.surf, &h.invrmed, &h.invsurf, &h.invdiffrsup,
                  &h.invdiffrmed, &h.invrinf, &h.sigmamed, &h.sigmainf, &h.energymed, &h.coolingtime,
                  &h.qplusmed, &h.supp_torque, &h.dxtheta, &h.visc, &h.cs_ring})
    v->assign(NR + 1, 0.0);
  // FillPolar1DArrays (SourceEuler.cu:76-157)
  std::vector<double> R(NR + 1), Rm(NR + 1);
  float drrsep = (P.rmax - P.rmin) / (float)NR;
  for (int i = 0; i <= NR; ++i) {
    R[i] = P.log_grid ? (double)P.rmin * std::exp((double)i / (double)NR * std::log((double)P.rmax / (double)P.rmin))
                      : P.rmin + drrsep * (double)i;
    h.radii[i] = R[i];
  }
  for (int i = 0; i < NR; ++i) {
    h.rinf[i] = R[i];
    h.rsup[i] = R[i + 1];
    Rm[i] = 2.0 / 3.0 * (R[i + 1] * R[i + 1] * R[i + 1] - R[i] * R[i] * R[i]) / (R[i + 1] * R[i + 1] - R[i] * R[i]);
    h.rmed[i] = Rm[i];
    h.surf[i] = kPI * (R[i + 1] * R[i + 1] - R[i] * R[i]) / (double)NS;
    h.invrmed[i] = 1.0 / Rm[i];
    h.invsurf[i] = 1.0 / h.surf[i];
    h.invdiffrsup[i] = 1.0 / (R[i + 1] - R[i]);
    h.invrinf[i] = 1.0 / R[i];
  }
  h.rinf[NR] = R[NR];
  h.rsup[NR] = R[NR];
  h.rmed[NR] = R[NR];  // never set by the reference; any finite value (it multiplies a zero slope)
  for (int i = 1; i < NR; ++i) h.invdiffrmed[i] = 1.0 / (Rm[i] - Rm[i - 1]);
  // Main.cu:264-278
  for (int i = 0; i < NR; ++i) {
    h.supp_torque[i] = P.imposeddiskdrift * 0.5 * std::pow(h.rmed[i], -2.5 + P.sigmaslope);
    h.dxtheta[i] = 2.0 * kPI / (double)NS * h.rmed[i];
  }
  h.cosj.resize(NS); h.sinj.resize(NS);
  for (int j = 0; j < NS; ++j) {
    double a = 2.0 * kPI * (double)j / (double)NS;
    h.cosj[j] = std::cos(a); h.sinj[j] = std::sin(a);
  }
  h.hill_radius = pl.nb > 0 ? pl.x[0] * std::pow(pl.mass[0] / 3., 1. / 3.) : 0.0;
  // InitGasDensity / FillSigma / Sigma (Pframeforce.cu:173-182, Theo.cu:13-30)
  size_t NG = (size_t)(NR + 1) * NS;
  h.dens.assign(NG, 0.0); h.vrad.assign(NG, 0.0); h.vtheta.assign(NG, 0.0);
  h.energy.assign(NG, 0.0); h.label.assign(NG, 0.0);
  auto sigma = [&](double r) {
    double cav = (r < P.cavityradius) ? 1.0 / P.cavityratio : 1.0;
    return cav * P.scaling_factor * P.sigma0 * std::pow(r, -(double)P.sigmaslope);
  };
  for (int i = 0; i < NR; ++i) {
    h.sigmamed[i] = sigma(h.rmed[i]);
    h.sigmainf[i] = sigma(h.rinf[i]);
    for (int j = 0; j < NS; ++j) h.dens[(size_t)i * NS + j] = h.sigmamed[i];
  }
  // seeded density noise: Sigma *= 1 + A(2u-1), u = splitmix64 draw n for cell n
  if (P.ext_noise_amplitude != 0.0) {
    uint64_t s = P.ext_noise_seed + (uint64_t)(int64_t)P.ext_disk_id;
    for (size_t k = 0; k < (size_t)NR * NS; ++k) {
      double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
      h.dens[k] = h.dens[k] * (1.0 + P.ext_noise_amplitude * (2.0 * u - 1.0));
    }
  }
  // InitGasEnergy / Energy (Pframeforce.cu:184-192, Theo.cu:33-51)
  if (P.adiabatic) {
    for (int i = 0; i < NR; ++i) {
      h.energymed[i] = 1.0 / 1.0 / (P.adiabaticindex - 1.0) * P.sigma0 * std::pow((double)P.aspectratio, 2.0) *
                       std::pow(h.rmed[i], -(double)P.sigmaslope - 1.0 + 2.0 * P.flaringindex);
      for (int j = 0; j < NS; ++j) h.energy[(size_t)i * NS + j] = h.energymed[i];
    }
  }
  // isothermal sound speed (ComputeSoundSpeedKernel, Kernels.cu:185-188) at t=0
  for (int i = 0; i < NR; ++i)
    h.cs_ring[i] = aspect_ratio(P, h.rmed[i], 0.0) * std::sqrt(kG * 1.0 / h.rmed[i]) * std::pow(h.rmed[i], (double)P.flaringindex);
  h.omega_frame = P.omegaframe;
  // InitLabel (Init.cu:8-18)
  if (pl.nb > 0) {
    double xp = pl.x[0], yp = pl.y[0];
    double rhill = std::sqrt(xp * xp + yp * yp) * std::pow(pl.mass[0] / 3., 1. / 3);
    for (int i = 0; i < NR; ++i)
      for (int j = 0; j < NS; ++j) {
        double x = h.rmed[i] * h.cosj[j], y = h.rmed[i] * h.sinj[j];
        h.label[(size_t)i * NS + j] = std::sqrt((x - xp) * (x - xp) + (y - yp) * (y - yp)) < rhill ? 1.0 : 0.0;
      }
  }
  return FGX_OK;
}

// Exact centrifugal balance of -b with FARGO's InitGasVelocities formula (the
// reference's Pframeforce.cu:202-237 averages the sound speed instead of the
// pressure, drops the point-mass term and never applies the result,
// Kernels.cu:783): at each inner interface Radii[i],
//   v^2 / r = (P[i] - P[i-1]) / (Sigma-bar (Rmed[i] - Rmed[i-1])) + G / (Rmed[i] Rmed[i-1]) - g_sg
// from the ring-mean pressure, the analytic Sigma and the ring-mean radial SG
// acceleration interpolated to the interface, frame velocity removed; ring
// values u follow from (u[i] + u[i-1]) / 2 = v[i] with u[0] from one secant
// step on the summed second differences (ConstructSequence,
// SourceEuler.cu:416-425).  NRAD+1 values.
static std::vector<double> centrifugal_balance(const HostDisk& h, const std::vector<double>& sg_axi) {
  const fgx_params& P = h.P;
  const int NR = h.nrad, NS = h.nsec;
  auto sigma = [&](double r) {
    double cav = (r < P.cavityradius) ? 1.0 / P.cavityratio : 1.0;
    return cav * P.scaling_factor * P.sigma0 * std::pow(r, -(double)P.sigmaslope);
  };
  std::vector<double> gp(NR), v(NR + 1, 0.0), u(NR + 1, 0.0);
  for (int i = 0; i < NR; ++i) {
    double a = 0;
    for (int j = 0; j < NS; ++j) {
      const size_t k = (size_t)i * NS + j;
      // ComputePressureFieldKernel (Kernels.cu:196-208) at t = 0
      a += P.adiabatic ? (P.adiabaticindex - 1.0) * h.energy[k] : h.dens[k] * h.cs_ring[i] * h.cs_ring[i];
    }
    gp[i] = a / (double)NS;
  }
  for (int i = 1; i < NR; ++i) {
    const double r0 = h.rmed[i - 1], r1 = h.rmed[i];
    v[i] = (gp[i] - gp[i - 1]) / (.5 * (sigma(r1) + sigma(r0))) / (r1 - r0) + kG * (1.0 / r0 - 1.0 / r1) / (r1 - r0);
  }
  if (P.self_gravity && !sg_axi.empty())
    for (int i = 1; i < NR; ++i)
      v[i] -= ((h.radii[i] - h.rmed[i - 1]) * sg_axi[i] + (h.rmed[i] - h.radii[i]) * sg_axi[i - 1]) /
              (h.rmed[i] - h.rmed[i - 1]);
  for (int i = 1; i < NR; ++i) v[i] = std::sqrt(v[i] * h.radii[i]) - h.radii[i] * h.omega_frame;
  auto sequence = [&]() {
    for (int i = 1; i < NR; ++i) u[i] = 2.0 * v[i] - u[i - 1];
    double lapl = 0;
    for (int i = 1; i < NR - 1; ++i) lapl += std::fabs(u[i + 1] + u[i - 1] - 2.0 * u[i]);
    return lapl;
  };
  const double t1 = u[0] = v[1] + .75 * (v[1] - v[2]);
  const double r1 = sequence();
  u[0] += .25 * (v[1] - v[2]);
  const double t2 = u[0];
  const double r2 = sequence();
  u[0] = t1 - r1 / (r2 - r1) * (t2 - t1);
  sequence();
  u[NR] = u[NR - 1];
  return u;
}

void setup_velocities(HostDisk& h, const std::vector<double>& sg_axi) {
  const fgx_params& P = h.P;
  const int NR = h.nrad, NS = h.nsec;
  if (P.corotating) h.omega_frame = psys_frequency(P, h.sys);
  // ring-mean sound speed for alpha viscosity (Make1Dprofile(1)); in the
  // adiabatic case cs comes from the energy field
  std::vector<double> csm(NR + 1, 0.0);
  for (int i = 0; i < NR; ++i) {
    if (!P.adiabatic) csm[i] = h.cs_ring[i];
    else {
      double s = 0;
      for (int j = 0; j < NS; ++j) {
        size_t k = (size_t)i * NS + j;
        s += std::sqrt((double)P.adiabaticindex * (P.adiabaticindex - 1.0) * h.energy[k] / h.dens[k]);
      }
      csm[i] = s / NS;
    }
  }
  for (int i = 0; i < NR; ++i) h.visc[i] = fviscosity(P, h.rmed[i], csm[i], 0.0);
  h.visc[NR] = fviscosity(P, h.rmed[NR - 1], csm[NR - 1], 0.0);
  if (P.cooling)
    for (int i = 0; i < NR; ++i) {
      double r = h.rmed[i];
      h.coolingtime[i] = P.coolingtime0 * std::pow(r, 2.0 + 2.0 * P.flaringindex);
      h.qplusmed[i] = 2.25 * h.visc[i] * P.sigma0 * std::pow(r, -(double)P.sigmaslope - 3.0);
    }
  // Init_azimutalvelocity_withSG (SourceEuler.cu:428-437, Kernels.cu:1044-1059)
  if (!P.centrifugal_balance && P.self_gravity && !sg_axi.empty()) {
    for (int i = 0; i < NR; ++i) {
      double r = h.rmed[i], invr = 1. / r;
      double omk = std::sqrt(kG * invr * invr * invr);
      double om = std::sqrt(omk * omk * (1.0 - (1. + P.sigmaslope - 2.0 * P.flaringindex) *
                                                   std::pow((double)P.aspectratio, 2.0) *
                                                   std::pow(r, 2.0 * P.flaringindex)) - invr * sg_axi[i]);
      for (int j = 0; j < NS; ++j) h.vtheta[(size_t)i * NS + j] = r * om;
    }
  }
  std::vector<double> vt_cent;
  if (P.centrifugal_balance && P.ext_stubbed_physics) vt_cent = centrifugal_balance(h, sg_axi);
  // InitGasVelocitiesKernel (Kernels.cu:759-800)
  for (int i = 0; i <= NR; ++i) {
    double r = (i == NR) ? h.rmed[NR - 1] : h.rmed[i];
    double ri = (i == NR) ? h.rinf[NR - 1] : h.rinf[i];
    for (int j = 0; j < NS; ++j) {
      size_t k = (size_t)i * NS + j;
      if (!P.self_gravity) {
        double om = std::sqrt(kG * 1.0 / r / r / r);
        h.vtheta[k] = om * r * std::sqrt(1.0 - std::pow((double)P.aspectratio, 2.0) * std::pow(r, 2.0 * P.flaringindex) *
                                                   (1. + P.sigmaslope - 2.0 * P.flaringindex));
      }
      h.vtheta[k] -= h.omega_frame * r;
      if (!vt_cent.empty()) h.vtheta[k] = vt_cent[i];
      if (i == NR) h.vrad[k] = 0.0;
      else {
        double v = P.imposeddiskdrift * P.sigma0 / h.sigmainf[i] / ri;
        if (P.viscosity_alpha) v -= 3.0 * h.visc[i] / r * (-P.sigmaslope + 2.0 * P.flaringindex + 1.0);
        else v -= 3.0 * h.visc[i] / r * (-P.sigmaslope + .5);
        h.vrad[k] = v;
      }
      if (i == 0) { h.vrad[j] = 0.0; h.vrad[(size_t)NR * NS + j] = 0.0; }
    }
  }
}

}  // namespace fgx
