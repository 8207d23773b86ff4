// oracle.hpp -- CPU restatement of the reference's AlgoGas hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker the GPU engine is compared
// against; nothing in the product library links, loads or calls it.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// through liboracle.so (oracle_capi.cpp).
//
// It restates, loop for loop, the reference's kernels and host functions
// (file:line cited at each function) in C++ with OpenMP over rings, templated
// on the arithmetic type `real`.  real=double is the FP64 oracle and CPU
// baseline; real=float exists for cross-checks of integer results on
// identical FP32 inputs.  Deviations from the reference (each one is a point
// where the reference reads uninitialised memory or accumulates in FP32):
//   - ring means and full-grid sums accumulate in FP64 (the reference sums
//     FP32, Kernels.cu:271-557, 726-756, 1177-1188, 1363-1377);
//   - PhysicalTime is FP64 (Viscosity.cu:15 keeps it in float);
//   - Rmed[NRAD] is defined as Radii[NRAD] (never set in SourceEuler.cu:139-157);
//   - with zero planets the star potential is still applied
//     (Kernels.cu:1342 only adds it inside the planet loop);
//   - the SG azimuthal update at ring 0, sector 0 wraps periodically
//     (Kernels.cu:1039 reads index -1 there).
// Quirk switches in fgx_params (ext_fix_*) select the corrected forms.
// With ext_stubbed_physics it also runs what the reference parses but leaves
// stubbed, with FARGO's semantics: accretion (accrete_onto_planets), the
// outer mass source (outer_source_mass), exact centrifugal balance
// (centrifugal_balance) and label advection (advect_label); each is pinned
// by closed forms and invariants in tests/test_oracle_stubbed.py.
#pragma once
#include "../include/fgx.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

namespace orc {

constexpr double G = 1.0;
constexpr double PI = 3.14159265358979323844;   // include/Fondam.cuh:3
constexpr double MU = 1.0, RGAS = 1.0;          // include/Fondam.cuh:4-5
constexpr double CFLSECURITY = 0.5;             // include/Fondam.cuh:7
constexpr double CVNR = 1.41;                   // include/Fondam.cuh:10

// splitmix64 stream used by the seeded density noise (ext_noise_*): the n-th
// draw for cell n = i*NSEC+j, u = top 53 bits / 2^53.
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// ---- FFT (mixed radix, forward sign -1, unnormalised) for the SG oracle ----
// Written from the Cooley-Tukey decimation-in-time recursion; radices 2,3,5
// and a direct DFT for any other prime factor.
inline void fft_rec(std::complex<double>* a, int n, int stride, std::complex<double>* out,
                    int sign, std::vector<std::complex<double>>& tmp) {
  if (n == 1) { out[0] = a[0]; return; }
  int p = 0;
  for (int r : {2, 3, 5}) if (n % r == 0) { p = r; break; }
  if (!p) {
    for (int r = 7; r * r <= n; r += 2) if (n % r == 0) { p = r; break; }
    if (!p) p = n;
  }
  int m = n / p;
  // sub-transforms of the p decimated sequences
  std::vector<std::complex<double>> sub((size_t)n);
  for (int q = 0; q < p; ++q) fft_rec(a + q * stride, m, stride * p, sub.data() + q * m, sign, tmp);
  for (int k = 0; k < m; ++k) {
    for (int s = 0; s < p; ++s) {
      std::complex<double> acc = 0;
      int kk = k + s * m;
      for (int q = 0; q < p; ++q) {
        double ang = sign * 2.0 * PI * (double)((long long)q * kk % n) / (double)n;
        acc += sub[(size_t)q * m + k] * std::complex<double>(std::cos(ang), std::sin(ang));
      }
      out[kk] = acc;
    }
  }
}
inline void fft1d(std::complex<double>* a, int n, int stride, int sign) {
  std::vector<std::complex<double>> in((size_t)n), out((size_t)n), tmp;
  for (int k = 0; k < n; ++k) in[k] = a[(size_t)k * stride];
  fft_rec(in.data(), n, 1, out.data(), sign, tmp);
  for (int k = 0; k < n; ++k) a[(size_t)k * stride] = out[k];
}
// 2-D transform of a rows x cols row-major array (cuFFT plan2d(rows, cols) layout)
inline void fft2d(std::vector<std::complex<double>>& a, int rows, int cols, int sign) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < rows; ++r) fft1d(a.data() + (size_t)r * cols, cols, 1, sign);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < cols; ++c) fft1d(a.data() + c, rows, cols, sign);
}

template <class real>
struct Disk {
  // FP64 for the FP32 and FP64 instances, the extended type for the 80-bit
  // instance (the potential is formed in it)
  using wide = typename std::conditional<(sizeof(real) > sizeof(double)), real, double>::type;
  fgx_params P;
  fgx_planets sys;
  int NRAD = 0, NSEC = 0;
  size_t NG = 0;  // (NRAD+1)*NSEC

  // 1-D tables (NRAD+1): FillPolar1DArrays, SourceEuler.cu:76-157
  std::vector<real> Radii, Rinf, Rsup, Rmed, Surf, invRmed, invSurf, invdiffRsup,
      invdiffRmed, invRinf, powRmed, SigmaMed, SigmaInf, EnergyMed, CoolingTimeMed,
      QplusMed, viscosity_array, GLOBAL_bufarray, supp_torque, dxtheta, GLOBAL_AxiSGAccr;
  std::vector<real> cosj, sinj;  // per sector

  // fields (NRAD+1)*NSEC
  std::vector<real> Dens, Vrad, Vtheta, Energy, Label, SoundSpeed, Pressure, Temperature,
      Potential, VradInt, VthetaInt, VradNew, VthetaNew, EnergyInt, EnergyNew, DensInt,
      TemperInt, DensStar, QRStar, dq, Work, RadMomP, RadMomM, ThetaMomP, ThetaMomM,
      VthetaRes, TempShift, DRR, DPP, DRP, DivV, TAURR, TAUPP, TAURP, Qplus, CellAbscissa,
      CellOrdinate, SG_Accr, SG_Acct, Vresidual, DT2D, ExtLabel;
  std::vector<real> VMed, Vmoy, newDT, DT1D, LostByDisk;
  std::vector<int> Nshift, NoSplitAdvection;

  // self-gravity kernels in spectral space (complex, 2NRAD x NSEC)
  std::vector<std::complex<double>> Kr_hat, Kt_hat;
  double SGP_eps = 0, SGP_rstep = 0, SGP_tstep = 0;

  double PhysicalTime = 0, PhysicalTimeInitial = 0, OmegaFrame = 0, MassTaper = 0;
  double LostMass = 0, mdcp = 0, mdcp1 = 0, exces_mdcp = 0, HillRadius = 0;
  double DOPAx = 0, DOPAy = 0;    // DiskOnPrimaryAcceleration
  double Indx = 0, Indy = 0;      // IndirectTerm (Pframeforce.cu:23)
  double XplanetMark = 0, YplanetMark = 0;  // Psys.cu:5 statics
  double last_dt = 0;
  int last_ncfl = 0;
  long long substeps = 0;
  double GlobalForce[44];

  size_t id(int i, int j) const { return (size_t)i * NSEC + j; }
  int jp1(int j) const { return (j + 1) % NSEC; }
  int jm1(int j) const { return (j - 1 + NSEC) % NSEC; }

  // ---------------------------------------------------------------- helpers
  // AspectRatioHost / AspectRatioDevice (Viscosity.cu:39-52, Kernels.cu:1473-1488)
  double aspect_ratio(double r) const {
    double h = P.aspectratio;
    double rmin = P.transitionradius - P.transitionwidth * P.aspectratio;
    double rmax = P.transitionradius + P.transitionwidth * P.aspectratio;
    double scale = 1.0 + (PhysicalTime - PhysicalTimeInitial) * P.lambdadoubling;
    rmin *= scale; rmax *= scale;
    if (r < rmin) h *= P.transitionratio;
    if (r >= rmin && r <= rmax) h *= std::exp((rmax - r) / (rmax - rmin) * std::log((double)P.transitionratio));
    return h;
  }
  // FViscosity (Viscosity.cu:19-37)
  double fviscosity(double r) const {
    double nu = P.viscosity;
    if (P.viscosity_alpha) {
      int i = 0;
      while (i < NRAD - 1 && (double)Rmed[i] < r) i++;
      nu = P.alphaviscosity * (double)GLOBAL_bufarray[i] * (double)GLOBAL_bufarray[i] * std::pow(r, 1.5);
    }
    double rmin = P.cavityradius - P.cavitywidth * P.aspectratio;
    double rmax = P.cavityradius + P.cavitywidth * P.aspectratio;
    double scale = 1.0 + (PhysicalTime - PhysicalTimeInitial) * P.lambdadoubling;
    rmin *= scale; rmax *= scale;
    if (r < rmin) nu *= P.cavityratio;
    if (r >= rmin && r <= rmax) nu *= std::exp((rmax - r) / (rmax - rmin) * std::log((double)P.cavityratio));
    return nu;
  }
  // Sigma / Energy / CoolingTime / Qplusinit (Theo.cu:13-78)
  double sigma_profile(double r) const {
    double cavity = 1.0;
    if (r < P.cavityradius) cavity = 1.0 / P.cavityratio;
    return cavity * P.scaling_factor * P.sigma0 * std::pow(r, -(double)P.sigmaslope);
  }
  double energy_profile(double r) const {
    return RGAS / MU / (P.adiabaticindex - 1.0) * P.sigma0 * std::pow((double)P.aspectratio, 2.0) *
           std::pow(r, -(double)P.sigmaslope - 1.0 + 2.0 * P.flaringindex);
  }
  // Compute_smoothing (Force.cu:69-74)
  double compute_smoothing(double r) const {
    return P.thicknesssmoothing * aspect_ratio(r) * std::pow(r, 1.0 + P.flaringindex);
  }
  // ring mean, FP64 accumulation (Make1DprofileKernel, Kernels.cu:726-738)
  real ring_mean(const std::vector<real>& f, int i) const {
    wide s = 0;
    for (int j = 0; j < NSEC; ++j) s += (wide)f[id(i, j)];
    return (real)(s / (wide)NSEC);
  }

  // ------------------------------------------------------------------ setup
  void alloc() {
    NRAD = P.nrad; NSEC = P.nsec; NG = (size_t)(NRAD + 1) * NSEC;
    for (auto* v : {&Radii, &Rinf, &Rsup, &Rmed, &Surf, &invRmed, &invSurf, &invdiffRsup,
                    &invdiffRmed, &invRinf, &powRmed, &SigmaMed, &SigmaInf, &EnergyMed,
                    &CoolingTimeMed, &QplusMed, &viscosity_array, &GLOBAL_bufarray,
                    &supp_torque, &dxtheta, &GLOBAL_AxiSGAccr, &VMed, &Vmoy, &newDT, &DT1D})
      v->assign(NRAD + 1, real(0));
    cosj.assign(NSEC, 0); sinj.assign(NSEC, 0); LostByDisk.assign(NSEC, 0);
    Nshift.assign(NRAD + 1, 0); NoSplitAdvection.assign(NRAD + 1, 0);
    for (auto* v : {&Dens, &Vrad, &Vtheta, &Energy, &Label, &SoundSpeed, &Pressure, &Temperature,
                    &Potential, &VradInt, &VthetaInt, &VradNew, &VthetaNew, &EnergyInt, &EnergyNew,
                    &DensInt, &TemperInt, &DensStar, &QRStar, &dq, &Work, &RadMomP, &RadMomM,
                    &ThetaMomP, &ThetaMomM, &VthetaRes, &TempShift, &DRR, &DPP, &DRP, &DivV,
                    &TAURR, &TAUPP, &TAURP, &Qplus, &CellAbscissa, &CellOrdinate, &SG_Accr,
                    &SG_Acct, &Vresidual, &DT2D, &ExtLabel})
      v->assign(NG, real(0));
  }

  // FillPolar1DArrays (SourceEuler.cu:76-157), radii.dat input not supported
  void fill_polar_1d() {
    std::vector<double> R2(NRAD + 1), Rm2(NRAD + 1);
    float drrsep = (P.rmax - P.rmin) / (float)NRAD;
    for (int i = 0; i <= NRAD; ++i) {
      if (P.log_grid)
        R2[i] = (double)P.rmin * std::exp((double)i / (double)NRAD * std::log((double)P.rmax / (double)P.rmin));
      else
        R2[i] = P.rmin + drrsep * (double)i;
      Radii[i] = (real)R2[i];
    }
    for (int i = 0; i < NRAD; ++i) {
      Rinf[i] = (real)R2[i];
      Rsup[i] = (real)R2[i + 1];
      Rm2[i] = 2.0 / 3.0 * (R2[i + 1] * R2[i + 1] * R2[i + 1] - R2[i] * R2[i] * R2[i]);
      Rm2[i] = Rm2[i] / (R2[i + 1] * R2[i + 1] - R2[i] * R2[i]);
      Rmed[i] = (real)Rm2[i];
      Surf[i] = (real)(PI * (R2[i + 1] * R2[i + 1] - R2[i] * R2[i]) / (double)NSEC);
      invRmed[i] = (real)(1.0 / Rm2[i]);
      invSurf[i] = (real)(1.0 / (double)Surf[i]);
      invdiffRsup[i] = (real)(1.0 / (R2[i + 1] - R2[i]));
      invRinf[i] = (real)(1.0 / R2[i]);
    }
    Rinf[NRAD] = (real)R2[NRAD];
    Rsup[NRAD] = (real)R2[NRAD];
    Rmed[NRAD] = (real)R2[NRAD];   // Q3: defined (uninitialised in the reference)
    for (int i = 0; i < NRAD; ++i) {
      if (i > 0) invdiffRmed[i] = (real)(1.0 / (Rm2[i] - Rm2[i - 1]));
      powRmed[i] = (real)std::pow(Rm2[i], -2.5 + P.sigmaslope);
    }
  }

  // InitPlanetarySystem side effect HillRadius (Psys.cu:132)
  void set_hill_radius() {
    HillRadius = sys.nb > 0 ? sys.x[0] * std::pow(sys.mass[0] / 3., 1. / 3.) : 0.0;
  }

  // GetPsysInfo (Psys.cu:162-232); action 0=GET 1=MARK 2=FREQUENCY
  double get_psys_info(int action) {
    double x, y, vx, vy, xc, yc, vxc, vyc;
    xc = x = sys.x[0]; yc = y = sys.y[0]; vxc = vx = sys.vx[0]; vyc = vy = sys.vy[0];
    double m = sys.mass[0] + 1.;
    double h = x * vy - y * vx;
    double d = std::sqrt(x * x + y * y);
    double Ax = x * vy * vy - y * vx * vy - G * m * x / d;
    double Ay = y * vx * vx - x * vx * vy - G * m * y / d;
    double e = std::sqrt(Ax * Ax + Ay * Ay) / m;
    double a = h * h / G / m / (1. - e * e);
    double arg = (e == 0.0) ? 1.0 : (1.0 - d / a) / e;
    double E;
    if (std::fabs(arg) >= 1.0) E = PI * (1. - arg / std::fabs(arg)) / 2.;
    else E = std::acos((1.0 - d / a) / e);
    if ((x * y * (vy * vy - vx * vx) + vx * vy * (x * x - y * y)) < 0) E = -E;
    double M = E - e * std::sin(E);
    double omega = std::sqrt(m / a / a / a);
    double PerihelionPA = std::atan2(Ay, Ax);
    if (P.guiding_center) {
      xc = a * std::cos(M + PerihelionPA); yc = a * std::sin(M + PerihelionPA);
      vxc = -a * omega * std::sin(M + PerihelionPA); vyc = a * omega * std::cos(M + PerihelionPA);
    }
    if (e < 1e-8) { xc = x; yc = y; vxc = vx; vyc = vy; }
    (void)vxc; (void)vyc;
    if (action == 1) { XplanetMark = xc; YplanetMark = yc; return 0.; }
    if (action == 0) {
      double d2 = std::sqrt(xc * xc + yc * yc);
      double d1 = std::sqrt(XplanetMark * XplanetMark + YplanetMark * YplanetMark);
      double cross = XplanetMark * yc - xc * YplanetMark;
      XplanetMark = xc; YplanetMark = yc;
      return std::asin(cross / (d1 * d2));
    }
    return omega;
  }

  // ComputeSoundSpeedKernel (Kernels.cu:176-193)
  void compute_sound_speed() {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i) {
      double r = Rmed[i];
      for (int j = 0; j < NSEC; ++j) {
        if (!P.adiabatic)
          SoundSpeed[id(i, j)] = (real)(aspect_ratio(r) * std::sqrt(G * 1.0 / r) * std::pow(r, (double)P.flaringindex));
        else
          SoundSpeed[id(i, j)] = (real)std::sqrt((double)P.adiabaticindex * (P.adiabaticindex - 1.0) *
                                                 (double)Energy[id(i, j)] / (double)Dens[id(i, j)]);
      }
    }
  }
  // ComputePressureFieldKernel (Kernels.cu:196-208)
  void compute_pressure() {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        Pressure[k] = !P.adiabatic ? Dens[k] * SoundSpeed[k] * SoundSpeed[k]
                                   : (real)(P.adiabaticindex - 1.0) * Energy[k];
      }
  }
  // ComputeTemperatureFieldKernel (Kernels.cu:211-223)
  void compute_temperature() {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        Temperature[k] = !P.adiabatic ? (real)(MU / RGAS) * Pressure[k] / Dens[k]
                                      : (real)(MU / RGAS * (P.adiabaticindex - 1.0)) * Energy[k] / Dens[k];
      }
  }

  // ------------------------------------------------------------ self-gravity
  // compute_kernel (SgMain.cu:49-81) evaluated in FP64, then transformed once
  void sg_init_kernels() {
    double eps = SGP_eps;
    int NR2 = 2 * NRAD;
    std::vector<std::complex<double>> Kr((size_t)NR2 * NSEC), Kt((size_t)NR2 * NSEC);
    for (int i = 0; i < NR2; ++i) {
      double u = (i < NRAD) ? std::log((double)Radii[i] / (double)Radii[0])
                            : -std::log((double)Radii[NR2 - i] / (double)Radii[0]);
      for (int j = 0; j < NSEC; ++j) {
        double th = 2.0 * PI * (double)j / (double)NSEC;
        double kr = 1.0 + eps * eps - std::cos(th) * std::exp(-u);
        double base = eps * eps * std::exp(u) + 2.0 * (std::cosh(u) - std::cos(th));
        kr *= std::pow(base, -1.5);
        double kt = std::sin(th) * std::pow(base, -1.5);
        Kr[(size_t)i * NSEC + j] = kr;
        Kt[(size_t)i * NSEC + j] = kt;
      }
    }
    fft2d(Kr, NR2, NSEC, -1);
    fft2d(Kt, NR2, NSEC, -1);
    Kr_hat = std::move(Kr); Kt_hat = std::move(Kt);
  }
  // compute_FFT .. compute_sgacc (SgMain.cu:84-106, Kernels.cu:949-1021)
  void sg_accel() {
    int NR2 = 2 * NRAD;
    std::vector<std::complex<double>> Sr((size_t)NR2 * NSEC, 0.0), St((size_t)NR2 * NSEC, 0.0);
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        double var = (double)Dens[id(i, j)] * std::sqrt((double)Rmed[i] / (double)Rmed[0]);
        Sr[(size_t)i * NSEC + j] = var;
        St[(size_t)i * NSEC + j] = var * (double)Rmed[i] / (double)Rmed[0];
      }
    fft2d(Sr, NR2, NSEC, -1);
    fft2d(St, NR2, NSEC, -1);
    for (size_t k = 0; k < Sr.size(); ++k) {
      St[k] = -G * Kt_hat[k] * St[k];  // Gphi
      Sr[k] = -G * Kr_hat[k] * Sr[k];  // Gr
    }
    fft2d(St, NR2, NSEC, +1);
    fft2d(Sr, NR2, NSEC, +1);
    double norm = SGP_rstep * SGP_tstep / ((double)NR2 * (double)NSEC);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i) {
      double divR = (double)Rmed[i] / (double)Rmed[0];
      double nr = norm / std::sqrt(divR), nt = norm / (divR * std::sqrt(divR));
      for (int j = 0; j < NSEC; ++j) {
        SG_Acct[id(i, j)] = (real)(St[(size_t)i * NSEC + j].real() * nt);
        SG_Accr[id(i, j)] = (real)(Sr[(size_t)i * NSEC + j].real() * nr +
                                   G * (double)Dens[id(i, j)] * SGP_rstep * SGP_tstep / SGP_eps);
      }
    }
  }
  // Update_sgvelocityKernel (Kernels.cu:1024-1041) applied to VradInt/VthetaInt
  void sg_update_velocity(double dt) {
    std::vector<real> acct = SG_Acct;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        if (i > 0)
          VradInt[id(i, j)] += (real)(dt * (((double)Rinf[i] - (double)Rmed[i - 1]) * (double)SG_Accr[id(i, j)] +
                                            ((double)Rmed[i] - (double)Rinf[i]) * (double)SG_Accr[id(i - 1, j)]) *
                                      (double)invdiffRmed[i]);
        // (j-1)%nsec at j=0 is -1: the reference reads linear index i*nsec-1
        size_t km = (j > 0) ? id(i, j - 1)
                            : ((P.ext_fix_sg_wrap || i == 0) ? id(i, NSEC - 1) : id(i, 0) - 1);
        VthetaInt[id(i, j)] += (real)(0.5 * dt * ((double)acct[id(i, j)] + (double)acct[km]));
      }
  }
  void sg_setup() {
    double soft = P.ext_sg_softening > 0 ? P.ext_sg_softening : P.thicknesssmoothing;
    SGP_eps = soft * (double)P.aspectratio;
    SGP_rstep = std::log((double)Radii[NRAD] / (double)Radii[0]) / (double)NRAD;
    SGP_tstep = 2.0 * PI / (double)NSEC;
    sg_init_kernels();
    sg_accel();
    // Init_planetarysys_withSG (SgMain.cu:166-195)
    for (int i = 0; i < NRAD; ++i) GLOBAL_AxiSGAccr[i] = ring_mean(SG_Accr, i);
    for (int k = 0; k < sys.nb; ++k) {
      double r = sys.x[k];
      double dist = (double)(float)(r / (1. + P.eccentricity));
      int ipl = 0;
      while (ipl < NRAD - 2 && (double)Rmed[ipl] <= dist) ipl++;
      double ri = Rmed[ipl], rip1 = Rmed[ipl + 1], dr = rip1 - ri;
      double sgacc = ((dist - ri) * (double)GLOBAL_AxiSGAccr[ipl + 1] + (rip1 - dist) * (double)GLOBAL_AxiSGAccr[ipl]) / dr;
      sys.vy[k] *= std::sqrt(1. - dist * dist * sgacc);
    }
  }

  // ---------------------------------------------------------------- set-up
  // main() (Main.cu:198-344) minus file output
  void setup(const fgx_params& p, const fgx_planets& pl) {
    P = p; sys = pl;
    alloc();
    fill_polar_1d();
    for (int i = 0; i < NRAD; ++i) {  // Main.cu:264-268
      supp_torque[i] = (real)(P.imposeddiskdrift * 0.5 * std::pow((double)Rmed[i], -2.5 + P.sigmaslope));
      dxtheta[i] = (real)(2.0 * PI / (double)NSEC * (double)Rmed[i]);
    }
    for (int j = 0; j < NSEC; ++j) {
      double a = 2.0 * PI * (double)j / (double)NSEC;
      cosj[j] = (real)std::cos(a); sinj[j] = (real)std::sin(a);
    }
    set_hill_radius();
    // InitGasDensity + FillSigma (Pframeforce.cu:173-182, Theo.cu:23-30)
    for (int i = 0; i < NRAD; ++i) {
      SigmaMed[i] = (real)sigma_profile(Rmed[i]);
      SigmaInf[i] = (real)sigma_profile(Rinf[i]);
      for (int j = 0; j < NSEC; ++j) Dens[id(i, j)] = SigmaMed[i];
    }
    if (P.ext_noise_amplitude != 0.0) {
      SplitMix64 rng(P.ext_noise_seed + (uint64_t)(int64_t)P.ext_disk_id);
      for (int i = 0; i < NRAD; ++i)
        for (int j = 0; j < NSEC; ++j) {
          double u = rng.uniform();
          Dens[id(i, j)] = (real)((double)Dens[id(i, j)] * (1.0 + P.ext_noise_amplitude * (2.0 * u - 1.0)));
        }
    }
    if (P.adiabatic) {  // InitGasEnergy (Pframeforce.cu:184-192)
      for (int i = 0; i < NRAD; ++i) {
        EnergyMed[i] = (real)energy_profile(Rmed[i]);
        for (int j = 0; j < NSEC; ++j) Energy[id(i, j)] = EnergyMed[i];
      }
    }
    if (P.self_gravity) sg_setup();
    OmegaFrame = P.omegaframe;
    if (P.corotating) OmegaFrame = get_psys_info(2);
    // InitEuler (SourceEuler.cu:177-204)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        CellAbscissa[id(i, j)] = Rmed[i] * cosj[j];
        CellOrdinate[id(i, j)] = Rmed[i] * sinj[j];
      }
    compute_sound_speed();
    compute_pressure();
    compute_temperature();
    init_gas_velocities();
    // InitLabel (Init.cu:8-18)
    if (sys.nb > 0) {
      double xp = sys.x[0], yp = sys.y[0];
      double rhill = std::sqrt(xp * xp + yp * yp) * std::pow(sys.mass[0] / 3., 1. / 3);
      for (int i = 0; i < NRAD; ++i)
        for (int j = 0; j < NSEC; ++j) {
          double x = Rmed[i] * cosj[j], y = Rmed[i] * sinj[j];
          Label[id(i, j)] = std::sqrt((x - xp) * (x - xp) + (y - yp) * (y - yp)) < rhill ? 1 : 0;
        }
    }
    PhysicalTimeInitial = PhysicalTime;
  }

  // InitGasVelocities + InitVelocities (Pframeforce.cu:194-272, Kernels.cu:759-800)
  void init_gas_velocities() {
    std::vector<double> vt_cent;
    if (P.centrifugal_balance && P.ext_stubbed_physics) vt_cent = centrifugal_balance();
    if (!P.centrifugal_balance && P.self_gravity) {
      // Init_azimutalvelocity_withSG (SourceEuler.cu:428-437, Kernels.cu:1044-1059)
      for (int i = 0; i < NRAD; ++i) {
        double ax = ring_mean(SG_Accr, i);
        double r = Rmed[i], invr = 1. / r;
        double omk = std::sqrt(G * invr * invr * invr);
        double om = std::sqrt(omk * omk * (1.0 - (1. + P.sigmaslope - 2.0 * P.flaringindex) *
                                                     std::pow((double)P.aspectratio, 2.0) *
                                                     std::pow(r, 2.0 * P.flaringindex)) - invr * ax);
        for (int j = 0; j < NSEC; ++j) Vtheta[id(i, j)] = (real)(r * om);
      }
    }
    if (P.viscosity_alpha)
      for (int i = 0; i < NRAD; ++i) GLOBAL_bufarray[i] = ring_mean(SoundSpeed, i);
    if (P.cooling) {
      for (int i = 0; i < NRAD; ++i) {
        double r = Rmed[i];
        CoolingTimeMed[i] = (real)(P.coolingtime0 * std::pow(r, 2.0 + 2.0 * P.flaringindex));
        QplusMed[i] = (real)(2.25 * fviscosity(r) * P.sigma0 * std::pow(r, -(double)P.sigmaslope - 3.0));
      }
    }
    for (int i = 0; i < NRAD; ++i) viscosity_array[i] = (real)fviscosity(Rmed[i]);
    viscosity_array[NRAD] = (real)fviscosity(Rmed[NRAD - 1]);
    for (int i = 0; i <= NRAD; ++i) {
      double r = (i == NRAD) ? (double)Rmed[NRAD - 1] : (double)Rmed[i];
      double ri = (i == NRAD) ? (double)Rinf[NRAD - 1] : (double)Rinf[i];
      for (int j = 0; j < NSEC; ++j) {
        if (!P.self_gravity) {
          double om = std::sqrt(G * 1.0 / r / r / r);
          Vtheta[id(i, j)] = (real)(om * r * std::sqrt(1.0 - std::pow((double)P.aspectratio, 2.0) *
                                                                 std::pow(r, 2.0 * P.flaringindex) *
                                                                 (1. + P.sigmaslope - 2.0 * P.flaringindex)));
        }
        Vtheta[id(i, j)] = (real)((double)Vtheta[id(i, j)] - OmegaFrame * r);
        if (!vt_cent.empty()) Vtheta[id(i, j)] = (real)vt_cent[i];
        if (i == NRAD) Vrad[id(i, j)] = 0;
        else {
          double v = P.imposeddiskdrift * P.sigma0 / (double)SigmaInf[i] / ri;
          if (P.viscosity_alpha) v -= 3.0 * (double)viscosity_array[i] / r * (-P.sigmaslope + 2.0 * P.flaringindex + 1.0);
          else v -= 3.0 * (double)viscosity_array[i] / r * (-P.sigmaslope + .5);
          Vrad[id(i, j)] = (real)v;
        }
        if (i == 0) { Vrad[id(0, j)] = 0; Vrad[id(NRAD, j)] = 0; }
      }
    }
  }

  // Exact centrifugal balance of -b, FARGO's InitGasVelocities: at each inner
  // interface Radii[i], r*Omega^2 = (dP/dr)/Sigma + G/r^2 - g_sg, from the
  // ring-mean pressure, the analytic Sigma, the point-mass potential and the
  // ring-mean radial SG acceleration interpolated to the interface; ring
  // values u[i] then follow from (u[i] + u[i-1])/2 = v[i], with u[0] chosen by
  // one secant step on the summed second differences of u (ConstructSequence,
  // SourceEuler.cu:416-425).  The reference's version (Pframeforce.cu:202-237)
  // averages the sound speed instead of the pressure, drops the G/r^2 term
  // and is never applied (Kernels.cu:783), so it runs here only with
  // StubbedPhysics.  Returns NRAD+1 ring values, frame velocity removed.
  std::vector<double> centrifugal_balance() {
    std::vector<double> gp(NRAD), v(NRAD + 1, 0.0), u(NRAD + 1, 0.0);
    for (int i = 0; i < NRAD; ++i) {
      double a = 0;
      for (int j = 0; j < NSEC; ++j) a += (double)Pressure[id(i, j)];
      gp[i] = a / (double)NSEC;
    }
    for (int i = 1; i < NRAD; ++i) {
      const double r0 = Rmed[i - 1], r1 = Rmed[i];
      v[i] = (gp[i] - gp[i - 1]) / (.5 * (sigma_profile(r1) + sigma_profile(r0))) / (r1 - r0) +
             G * (1.0 / r0 - 1.0 / r1) / (r1 - r0);
    }
    if (P.self_gravity) {
      std::vector<double> ax(NRAD);
      for (int i = 0; i < NRAD; ++i) {
        double a = 0;
        for (int j = 0; j < NSEC; ++j) a += (double)SG_Accr[id(i, j)];
        ax[i] = a / (double)NSEC;
      }
      for (int i = 1; i < NRAD; ++i)
        v[i] -= (((double)Radii[i] - (double)Rmed[i - 1]) * ax[i] + ((double)Rmed[i] - (double)Radii[i]) * ax[i - 1]) /
                ((double)Rmed[i] - (double)Rmed[i - 1]);
    }
    for (int i = 1; i < NRAD; ++i) v[i] = std::sqrt(v[i] * (double)Radii[i]) - (double)Radii[i] * OmegaFrame;
    auto sequence = [&]() {
      for (int i = 1; i < NRAD; ++i) u[i] = 2.0 * v[i] - u[i - 1];
      double lapl = 0;
      for (int i = 1; i < NRAD - 1; ++i) lapl += std::fabs(u[i + 1] + u[i - 1] - 2.0 * u[i]);
      return lapl;
    };
    double t1 = u[0] = v[1] + .75 * (v[1] - v[2]);
    const double r1 = sequence();
    u[0] += .25 * (v[1] - v[2]);
    const double t2 = u[0];
    const double r2 = sequence();
    u[0] = t1 - r1 / (r2 - r1) * (t2 - t1);
    sequence();
    u[NRAD] = u[NRAD - 1];
    return u;
  }

  // ------------------------------------------------------------------ CFL
  // ConditionCFL (SourceEuler.cu:494-536) + kernels (Kernels.cu:1363-1429)
  int condition_cfl(double DeltaT) {
    for (int i = 0; i < NRAD; ++i) Vmoy[i] = ring_mean(Vtheta, i);
#pragma omp parallel for schedule(static)
    for (int i = 1; i < NRAD; ++i) {
      real dxrad = Rsup[i] - Rinf[i];
      real mn = (real)1e30;
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        Vresidual[k] = P.fast_transport ? Vtheta[k] - Vmoy[i] : Vtheta[k];
        real invdt1 = SoundSpeed[k] / std::min(dxrad, dxtheta[i]);
        real invdt2 = std::fabs(Vrad[k]) / dxrad;
        real invdt3 = std::fabs(Vresidual[k]) / dxtheta[i];
        real dvr = Vrad[id(i + 1, j)] - Vrad[k];
        real dvt = Vtheta[id(i, jp1(j))] - Vtheta[k];
        dvr = (dvr >= 0) ? (real)1e-10 : -dvr;
        dvt = (dvt >= 0) ? (real)1e-10 : -dvt;
        real invdt4 = std::max(dvr / dxrad, dvt / dxtheta[i]) * (real)(4.0 * CVNR * CVNR);
        DT2D[k] = (real)CFLSECURITY / std::sqrt(invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 + invdt4 * invdt4);
        if (DT2D[k] < mn) mn = DT2D[k];
      }
      newDT[i] = mn;
    }
    for (int i = 0; i < NRAD - 1; ++i)
      DT1D[i] = (real)(2.0 * PI * CFLSECURITY / (double)NSEC /
                       std::fabs((double)(Vmoy[i] * invRmed[i] - Vmoy[i + 1] * invRmed[i + 1])));
    real newdt = newDT[1];
    for (int i = 2; i < NRAD; ++i) if (newDT[i] < newdt) newdt = newDT[i];
    for (int i = 0; i < NRAD - 1; ++i) if (DT1D[i] < newdt) newdt = DT1D[i];
    real dT = (real)DeltaT;
    if (dT < newdt) newdt = dT;
    return (int)std::ceil(dT / newdt);
  }

  // ---------------------------------------------------------------- forces
  // ComputeForce (Force.cu:31-67) + ComputeForceKernel (Kernels.cu:804-839):
  // GlobalForce[44] = {fxi[11], fxo[11], fyi[11], fyo[11]}, FP64 sums
  void compute_force(double x, double y, double rsmoothing, double mass, double* gf) {
    const int dim = 11;
    double a = std::sqrt(x * x + y * y);
    double rh = std::pow(mass / 3., 1. / 3.) * a + 1e-15;
    std::vector<double> part((size_t)NRAD * 44, 0.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i) {
      double* pp = &part[(size_t)i * 44];
      bool inner = (double)Rmed[i] < a;
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        double cellmass = (double)Surf[i] * (double)Dens[k];
        double dx = (double)CellAbscissa[k] - x, dy = (double)CellOrdinate[k] - y;
        double d2 = dx * dx + dy * dy;
        double dist2 = d2 + rsmoothing * rsmoothing;
        double inv3 = 1.0 / dist2 / std::sqrt(dist2);
        for (int kk = 0; kk < dim; ++kk) {
          double cut = 1.0;
          if (kk != 0) {
            double rhk = rh * (double)kk / (double)(dim - 1);
            cut = 1. - std::exp(-d2 / (rhk * rhk));
          }
          double fx = G * cellmass * dx * inv3 * cut, fy = G * cellmass * dy * inv3 * cut;
          if (inner) { pp[kk] += fx; pp[kk + 2 * dim] += fy; }
          else { pp[kk + dim] += fx; pp[kk + 3 * dim] += fy; }
        }
      }
    }
    for (int q = 0; q < 44; ++q) {
      double s = 0;
      for (int i = 0; i < NRAD; ++i) s += part[(size_t)i * 44 + q];
      gf[q] = s;
    }
  }
  // ComputeAccel (SideEuler.cu:49-62)
  void compute_accel(double x, double y, double rsm, double mass, double& ax, double& ay) {
    compute_force(x, y, rsm, mass, GlobalForce);
    if (P.exclude_hill) { ax = GlobalForce[10] + GlobalForce[21]; ay = GlobalForce[32] + GlobalForce[43]; }
    else { ax = GlobalForce[0] + GlobalForce[11]; ay = GlobalForce[22] + GlobalForce[33]; }
  }

  // FillForcesArrays (Pframeforce.cu:37-66) + FillForcesArraysKernel (Kernels.cu:1322-1349)
  void fill_forces_arrays() {
    Indx = -DOPAx; Indy = -DOPAy;             // ComputeIndirectTerm (Pframeforce.cu:27-35)
    if (!P.indirect_term) { Indx = 0; Indy = 0; }
    std::fill(Potential.begin(), Potential.end(), real(0));
    int nb = sys.nb;
    for (int k = 0; k < std::max(nb, 1); ++k) {
      double xp = 0, yp = 0, mp = 0, inv3 = 0, smooth = 0;
      bool have_planet = k < nb;
      if (have_planet) {
        xp = sys.x[k]; yp = sys.y[k]; mp = sys.mass[k] * MassTaper;
        double d = std::sqrt(xp * xp + yp * yp);
        inv3 = 1.0 / d / d / d;
        double rroche = d * std::pow((1.0 / 3.0 * mp), 1.0 / 3.0);
        double sm = P.roche_smoothing ? rroche * P.rochesmoothing : compute_smoothing(d);
        smooth = sm * sm;
      }
#pragma omp parallel for schedule(static)
      for (int i = 0; i < NRAD; ++i) {
        wide r = Rmed[i];
        for (int j = 0; j < NSEC; ++j) {
          wide x = r * (wide)cosj[j], y = r * (wide)sinj[j];
          wide pot = 0;
          if (have_planet) {
            wide dist = (x - xp) * (x - xp) + (y - yp) * (y - yp);
            pot = -G * mp / std::sqrt(dist + smooth);
            if (P.indirect_term) pot += G * mp * inv3 * (x * xp + y * yp);
          }
          if (k == 0) pot += -G * (wide)1.0 / r - (Indx * x + Indy * y);
          Potential[id(i, j)] = (real)((wide)Potential[id(i, j)] + pot);
        }
      }
    }
  }

  // AdvanceSystemFromDisk (Pframeforce.cu:68-92)
  void advance_system_from_disk(double dt) {
    for (int k = 0; k < sys.nb; ++k) {
      if (!sys.feel_disk[k]) continue;
      double m = sys.mass[k], x = sys.x[k], y = sys.y[k];
      double r = std::sqrt(x * x + y * y);
      double sm = P.roche_smoothing ? r * std::pow(m / 3., 1. / 3.) * P.rochesmoothing : compute_smoothing(r);
      double gx, gy;
      compute_accel(x, y, sm, m, gx, gy);
      sys.vx[k] += dt * gx; sys.vy[k] += dt * gy;
      sys.vx[k] += dt * Indx; sys.vy[k] += dt * Indy;
    }
  }

  // DerivMotionRK5 / RungeKutta (RungeKutta.cu:9-74): Cash-Karp tableau
  void deriv_rk5(const double* q, const double* m, double* der, int nb, double dt) {
    const double *x = q, *y = q + nb, *vx = q + 2 * nb, *vy = q + 3 * nb;
    double *dx = der, *dy = der + nb, *dvx = der + 2 * nb, *dvy = der + 3 * nb;
    double Dist[FGX_MAX_PLANETS];
    for (int i = 0; i < nb; ++i) Dist[i] = std::sqrt(x[i] * x[i] + y[i] * y[i]);
    for (int i = 0; i < nb; ++i) {
      dx[i] = vx[i]; dy[i] = vy[i];
      dvx[i] = -G * 1.0 / Dist[i] / Dist[i] / Dist[i] * x[i];
      dvy[i] = -G * 1.0 / Dist[i] / Dist[i] / Dist[i] * y[i];
      for (int j = 0; j < nb; ++j) {
        if (P.indirect_term) {
          dvx[i] -= G * m[j] / Dist[j] / Dist[j] / Dist[j] * x[j];
          dvy[i] -= G * m[j] / Dist[j] / Dist[j] / Dist[j] * y[j];
        }
        if (j != i && sys.feel_others[i]) {
          double d3;
          if (P.ext_fix_pp_distance) {
            double dd = std::sqrt((x[i] - x[j]) * (x[i] - x[j]) + (y[i] - y[j]) * (y[i] - y[j]));
            d3 = dd * dd * dd;
          } else d3 = Dist[j] * Dist[j] * Dist[j];  // Q6
          dvx[i] += G * m[j] / d3 * (x[j] - x[i]);
          dvy[i] += G * m[j] / d3 * (y[j] - y[i]);
        }
      }
    }
    for (int i = 0; i < 4 * nb; ++i) der[i] *= dt;
  }
  void runge_kutta(const double* q0, double dt, const double* m, double* q1, int nb) {
    double k1[4 * FGX_MAX_PLANETS] = {0}, k2[4 * FGX_MAX_PLANETS] = {0}, k3[4 * FGX_MAX_PLANETS] = {0},
           k4[4 * FGX_MAX_PLANETS] = {0}, k5[4 * FGX_MAX_PLANETS] = {0}, k6[4 * FGX_MAX_PLANETS] = {0};
    auto tr = [&](double c1, double c2, double c3, double c4, double c5) {
      for (int i = 0; i < 4 * nb; ++i) q1[i] = q0[i] + c1 * k1[i] + c2 * k2[i] + c3 * k3[i] + c4 * k4[i] + c5 * k5[i];
    };
    deriv_rk5(q0, m, k1, nb, dt);
    tr(0.2, 0, 0, 0, 0);
    deriv_rk5(q1, m, k2, nb, dt);
    tr(0.075, 0.225, 0, 0, 0);
    deriv_rk5(q1, m, k3, nb, dt);
    tr(0.3, -0.9, 1.2, 0, 0);
    deriv_rk5(q1, m, k4, nb, dt);
    tr(-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0);
    deriv_rk5(q1, m, k5, nb, dt);
    tr(1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0);
    deriv_rk5(q1, m, k6, nb, dt);
    for (int i = 0; i < 4 * nb; ++i)
      q1[i] = q0[i] + 37.0 / 378.0 * k1[i] + 250.0 / 621.0 * k3[i] + 125.0 / 594.0 * k4[i] + 512.0 / 1771.0 * k6[i];
  }
  // AdvanceSystemRK5 (Pframeforce.cu:94-157)
  void advance_system_rk5(double dt) {
    int nb = sys.nb;
    if (nb == 0) return;
    double q0[4 * FGX_MAX_PLANETS], q1[4 * FGX_MAX_PLANETS], m[FGX_MAX_PLANETS];
    if (!P.forced_circular) {
      for (int k = 0; k < nb; ++k) {
        q0[k] = sys.x[k]; q0[k + nb] = sys.y[k]; q0[k + 2 * nb] = sys.vx[k]; q0[k + 3 * nb] = sys.vy[k];
        m[k] = sys.mass[k];
      }
      runge_kutta(q0, dt, m, q1, nb);
    }
    for (int i = 1 - (PhysicalTime >= P.releasedate ? 1 : 0); i < nb; ++i) {
      if (!P.forced_circular) {
        sys.x[i] = q1[i]; sys.y[i] = q1[i + nb]; sys.vx[i] = q1[i + 2 * nb]; sys.vy[i] = q1[i + 3 * nb];
      } else {
        double x = sys.x[i], y = sys.y[i], th = std::atan2(y, x), vx = sys.vx[i], vy = sys.vy[i];
        double r = std::sqrt(x * x + y * y);
        double om = (-y * vx + x * vy) / r / r, dth = om * dt;
        sys.x[i] = r * std::cos(th + dth); sys.y[i] = r * std::sin(th + dth);
        sys.vx[i] = vx * std::cos(dth + th) - vy * std::sin(dth + th);
        sys.vy[i] = vx * std::sin(dth + th) + vy * std::cos(dth + th);
      }
    }
    if (PhysicalTime < P.releasedate) {
      double x = sys.x[0], y = sys.y[0], r = std::sqrt(x * x + y * y), th = std::atan2(y, x);
      double rdot = (P.releaseradius - r) / (P.releasedate - PhysicalTime);
      double om = std::sqrt((1. + sys.mass[0]) / r / r / r);
      double nr = r + rdot * dt, den = r - nr, dth;
      if (den != 0.0) dth = 2. * dt * r * om / den * (std::sqrt(r / nr) - 1.);
      else dth = om * dt;
      double vx = rdot, vy = nr * std::sqrt((1. + sys.mass[0]) / nr / nr / nr);
      sys.x[0] = nr * std::cos(dth + th); sys.y[0] = nr * std::sin(dth + th);
      sys.vx[0] = vx * std::cos(dth + th) - vy * std::sin(dth + th);
      sys.vy[0] = vx * std::sin(dth + th) + vx * std::cos(dth + th);  // sic (Pframeforce.cu:154)
    }
  }
  // RotatePsys (Psys.cu:234-251): rotate by -angle
  void rotate_psys(double angle) {
    double s = std::sin(angle), c = std::cos(angle);
    for (int i = 0; i < sys.nb; ++i) {
      double xt = sys.x[i], yt = sys.y[i];
      sys.x[i] = xt * c + yt * s; sys.y[i] = -xt * s + yt * c;
      xt = sys.vx[i]; yt = sys.vy[i];
      sys.vx[i] = xt * c + yt * s; sys.vy[i] = -xt * s + yt * c;
    }
  }

  // ------------------------------------------------------------ boundaries
  // ApplyBoundaryCondition (SideEuler.cu:172-182)
  void apply_boundary(double dt) {
    if (P.open_inner) open_boundary();
    if (P.nonreflecting) {
      if (P.adiabatic) compute_sound_speed();
      nonreflecting_boundary();
    }
    if (P.evanescent) evanescent_boundary(dt);
    if (P.outer_source_mass && P.ext_stubbed_physics) outer_source_mass();
  }
  // ApplyOuterSourceMass, FARGO's SideEuler.c: the reference leaves it
  // commented out (SideEuler.cu:143-146,181).  The last ring's mean density
  // is reset to SigmaMed, keeping its azimuthal structure, and its radial
  // velocity to the imposed drift IMPOSEDDISKDRIFT * Rinf^-SIGMASLOPE.
  void outer_source_mass() {
    const int i = NRAD - 1;
    const real add = SigmaMed[i] - ring_mean(Dens, i);
    const real vr = (real)(P.imposeddiskdrift * std::pow((double)Rinf[i], -(double)P.sigmaslope));
    for (int j = 0; j < NSEC; ++j) {
      Dens[id(i, j)] += add;
      Vrad[id(i, j)] = vr;
    }
  }
  // OpenBoundaryKernel (Kernels.cu:624-637)
  void open_boundary() {
    for (int j = 0; j < NSEC; ++j) {
      Dens[id(0, j)] = Dens[id(1, j)];
      Energy[id(0, j)] = Energy[id(1, j)];
      if (Vrad[id(2, j)] > 0 || Dens[id(1, j)] < SigmaMed[0]) Vrad[id(1, j)] = 0;
      else Vrad[id(1, j)] = Vrad[id(2, j)];
    }
  }
  // NonReflectingBoundary (SideEuler.cu:70-120) + kernels (Kernels.cu:640-723)
  void nonreflecting_boundary() {
    double cs0 = ring_mean(SoundSpeed, 0), cs1 = ring_mean(SoundSpeed, 1);
    double csn1 = ring_mean(SoundSpeed, NRAD - 1), csn2 = ring_mean(SoundSpeed, NRAD - 2);
    double dangle = (std::pow((double)Rinf[1], -1.5) - 1.0) / (.5 * (cs0 + cs1));
    dangle *= ((double)Rmed[1] - (double)Rmed[0]);
    int ia = (int)(dangle / 2.0 / PI * (double)NSEC + .5);
    double dangle2 = (std::pow((double)Rinf[NRAD - 2], -1.5) - 1.0) / (.5 * (csn1 + csn2));
    dangle2 *= ((double)Rmed[NRAD - 1] - (double)Rmed[NRAD - 2]);
    int ia2 = (int)(dangle2 / 2.0 / PI * (double)NSEC + .5);
    real sm1 = SigmaMed[1], sm2 = SigmaMed[NRAD - 2];
    for (int j = 0; j < NSEC; ++j) {
      int jp = j + ia;
      if (jp >= NSEC) jp -= NSEC;
      if (jp < 0) jp += NSEC;
      Dens[id(0, jp)] = Dens[id(1, j)];
      Energy[id(0, jp)] = Energy[id(1, j)];
      real vmed = -SoundSpeed[id(1, j)] * (Dens[id(1, j)] - sm1) / sm1;
      Vrad[id(1, j)] = (real)2. * vmed - Vrad[id(2, j)];
    }
    int i = NRAD - 1;
    for (int j = 0; j < NSEC; ++j) {
      int jp = j - ia2;
      if (jp >= NSEC) jp -= NSEC;
      if (jp < 0) jp += NSEC;
      Dens[id(i, j)] = Dens[id(i - 1, jp)];
      Energy[id(i, j)] = Energy[id(i - 1, jp)];
      real vmed = SoundSpeed[id(i, j)] * (Dens[id(i - 1, j)] - sm2) / sm2;
      Vrad[id(i, j)] = (real)2. * vmed - Vrad[id(i - 1, j)];
    }
    // ReduceMean + MinusMeanKernel
    real md = ring_mean(Dens, 0), md2 = ring_mean(Dens, NRAD - 1);
    real me = ring_mean(Energy, 0), me2 = ring_mean(Energy, NRAD - 1);
    for (int j = 0; j < NSEC; ++j) {
      Dens[id(0, j)] += SigmaMed[0] - md;
      Energy[id(0, j)] += EnergyMed[0] - me;
      Dens[id(NRAD - 1, j)] += SigmaMed[NRAD - 1] - md2;
      Energy[id(NRAD - 1, j)] += EnergyMed[NRAD - 1] - me2;
    }
  }
  // EvanescentBoundary (SideEuler.cu:122-141) + kernel (Kernels.cu:1076-1124)
  void evanescent_boundary(double step) {
    double Tin = 2.0 * PI * std::pow((double)Rmed[0], 3. / 2);
    double Tout = 2.0 * PI * std::pow((double)Rmed[NRAD - 1], 3. / 2);
    double DRMIN = (double)Rmed[0] * 1.25, DRMAX = (double)Rmed[NRAD - 1] * 0.84;
    for (int i = 0; i < NRAD; ++i) viscosity_array[i] = (real)fviscosity(Rmed[i]);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i) {
      double r = Rmed[i];
      if (!(r < DRMIN || r > DRMAX)) continue;
      double lambda = 0, damping;
      if (r < DRMIN) { damping = (r - DRMIN) / ((double)Rmed[0] - DRMIN); lambda = damping * damping * 10.0 * step / Tin; }
      if (r > DRMAX) { damping = (r - DRMAX) / ((double)Rmed[NRAD - 1] - DRMAX); lambda = damping * damping * 10.0 * step / Tout; }
      const wide h = aspect_ratio(r), rw = Rmed[i];
      // Q9: the self-gravity profile of this kernel is never filled (zero)
      wide vt0 = std::sqrt(G * (wide)1.0 / rw * ((wide)1.0 - ((wide)1.0 + P.sigmaslope - (wide)2.0 * P.flaringindex) * h * h *
                                                     std::pow(rw, (wide)2.0 * P.flaringindex)));
      vt0 -= rw * OmegaFrame;
      const wide vr0 = (wide)-3.0 * (wide)viscosity_array[i] / rw * (-P.sigmaslope + (wide).5);  // Q12
      const wide d0 = SigmaMed[i], e0 = EnergyMed[i], lw = lambda;
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        Vrad[k] = (real)(((wide)Vrad[k] + lw * vr0) / ((wide)1.0 + lw));
        Vtheta[k] = (real)(((wide)Vtheta[k] + lw * vt0) / ((wide)1.0 + lw));
        Dens[k] = (real)(((wide)Dens[k] + lw * d0) / ((wide)1.0 + lw));
        if (P.adiabatic) Energy[k] = (real)(((wide)Energy[k] + lw * e0) / ((wide)1.0 + lw));
      }
    }
  }
  // ApplySubKeplerianBoundary (SideEuler.cu:148-170)
  void sub_keplerian_boundary() {
    double r0 = Rmed[0], rn = Rmed[NRAD - 1];
    double fac = (1.0 + P.sigmaslope - 2.0 * P.flaringindex);
    double vin = G * 1.0 / r0 * (1.0 - fac * std::pow(aspect_ratio(r0), 2.0) * std::pow(r0, 2.0 * P.flaringindex));
    double vout = G * 1.0 / rn * (1.0 - fac * std::pow(aspect_ratio(rn), 2.0) * std::pow(rn, 2.0 * P.flaringindex));
    if (P.self_gravity) {
      for (int i = 0; i < NRAD; ++i) GLOBAL_AxiSGAccr[i] = ring_mean(SG_Accr, i);
      vin -= r0 * (double)GLOBAL_AxiSGAccr[0];
      vout -= rn * (double)GLOBAL_AxiSGAccr[NRAD - 1];
    }
    vin = std::sqrt(vin); vout = std::sqrt(vout);
    for (int j = 0; j < NSEC; ++j) {
      VthetaInt[id(0, j)] = (real)(vin - r0 * OmegaFrame);
      VthetaInt[id(NRAD - 1, j)] = (real)(vout - rn * OmegaFrame);
    }
  }

  // --------------------------------------------------------------- source
  // Substep1KernelVrad / Substep1KernelVtheta (Kernels.cu:52-88)
  void substep1_kernels(double dt) {
    real rdt = (real)dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i) {
      real invdxt = (real)(1.0 / (2.0 * PI / (double)NSEC * (double)Rmed[i]));
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        if (i > 0) {
          size_t km = id(i - 1, j);
          real gradp = (Pressure[k] - Pressure[km]) * (real)2.0 / (Dens[k] + Dens[km]) * invdiffRmed[i];
          real gradphi = (Potential[k] - Potential[km]) * invdiffRmed[i];
          real vt2 = Vtheta[k] + Vtheta[id(i, jp1(j))] + Vtheta[km] + Vtheta[id(i - 1, jp1(j))];
          vt2 = vt2 / (real)4.0 + Rinf[i] * (real)OmegaFrame;
          vt2 = vt2 * vt2;
          VradInt[k] = Vrad[k] + rdt * (-gradp - gradphi + vt2 * invRinf[i]);
        }
        size_t kl = id(i, jm1(j));
        real gradp = (Pressure[k] - Pressure[kl]) * (real)2.0 / (Dens[k] + Dens[kl]) * invdxt;
        real gradphi = (Potential[k] - Potential[kl]) * invdxt;
        VthetaInt[k] = Vtheta[k] - rdt * (gradp + gradphi);
        VthetaInt[k] += rdt * supp_torque[i];
      }
    }
  }
  // ComputeViscousTerms (Viscosity.cu:77-95) + kernels (Kernels.cu:843-880)
  void compute_viscous_terms(const std::vector<real>& vr, const std::vector<real>& vt) {
    if (P.viscosity_alpha)
      for (int i = 0; i < NRAD; ++i) GLOBAL_bufarray[i] = ring_mean(SoundSpeed, i);
    for (int i = 0; i < NRAD; ++i) viscosity_array[i] = (real)fviscosity(Rmed[i]);
    real invdphi = (real)(1.0 / (2.0 * PI / (double)NSEC));
    real onethird = (real)(1.0 / 3.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j), ku = id(i + 1, j), kr = id(i, jp1(j));
        DRR[k] = (vr[ku] - vr[k]) * invdiffRsup[i];
        DPP[k] = (vt[kr] - vt[k]) * invdphi * invRmed[i] + (real)0.5 * (vr[ku] + vr[k]) * invRmed[i];
        DivV[k] = (vr[ku] * Rsup[i] - vr[k] * Rinf[i]) * invdiffRsup[i] * invRmed[i];
        DivV[k] += (vt[kr] - vt[k]) * invdphi * invRmed[i];
        if (i > 0)
          DRP[k] = (real)0.5 * (Rinf[i] * (vt[k] * invRmed[i] - vt[id(i - 1, j)] * invRmed[i - 1]) * invdiffRmed[i] +
                                (vr[k] - vr[id(i, jm1(j))]) * invdphi * invRinf[i]);
      }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        real nu = viscosity_array[i];
        TAURR[k] = (real)2.0 * Dens[k] * nu * (DRR[k] - onethird * DivV[k]);
        TAUPP[k] = (real)2.0 * Dens[k] * nu * (DPP[k] - onethird * DivV[k]);
        if (i > 0) {
          // Q1: ((j-1)*nsec)%nsec is always 0 -> column 0 of ring i
          size_t kc = P.ext_fix_taurp_corner ? id(i, jm1(j)) : id(i, 0);
          TAURP[k] = (real)(2.0 * 0.25) * (Dens[k] + Dens[id(i - 1, j)] + Dens[kc] + Dens[id(i - 1, jm1(j))]) * nu * DRP[k];
        }
      }
  }
  // UpdateVelocitiesKernel (Kernels.cu:142-160)
  void update_velocities_with_viscosity(double dt) {
    real rdt = (real)dt;
    real invdphi = (real)(1.0 / (2.0 * PI / (double)NSEC));
#pragma omp parallel for schedule(static)
    for (int i = 1; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        if (i < NRAD - 1) {
          size_t kl = id(i, jm1(j));
          VthetaInt[k] += rdt * invRmed[i] *
                          ((Rsup[i] * TAURP[id(i + 1, j)] - Rinf[i] * TAURP[k]) * invdiffRsup[i] +
                           (TAUPP[k] - TAUPP[kl]) * invdphi + (real)0.5 * (TAURP[k] + TAURP[id(i + 1, j)])) /
                          ((real)0.5 * (Dens[k] + Dens[kl]));
        }
        size_t km = id(i - 1, j);
        VradInt[k] += rdt * invRinf[i] *
                      ((Rmed[i] * TAURR[k] - Rmed[i - 1] * TAURR[km]) * invdiffRmed[i] +
                       (TAURP[id(i, jp1(j))] - TAURP[k]) * invdphi - (real)0.5 * (TAUPP[k] + TAUPP[km])) /
                      ((real)0.5 * (Dens[k] + Dens[km]));
      }
  }
  // Substep1 (SourceEuler.cu:305-335)
  void substep1(double dt) {
    substep1_kernels(dt);
    if (P.self_gravity) {
      sg_accel();
      sg_update_velocity(dt);
    }
    compute_viscous_terms(VradInt, VthetaInt);
    update_velocities_with_viscosity(dt);
    if (!P.evanescent) sub_keplerian_boundary();
  }
  // Substep2Kernel / Substep2Kernel2 (Kernels.cu:570-621)
  void substep2(double dt) {
    real rdt = (real)dt;
    real cv2 = (real)(CVNR * CVNR);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        real dv = VradInt[id(i + 1, j)] - VradInt[k];
        DensInt[k] = (dv < 0) ? cv2 * Dens[k] * dv * dv : (real)0;
        dv = VthetaInt[id(i, jp1(j))] - VthetaInt[k];
        TemperInt[k] = (dv < 0) ? cv2 * Dens[k] * dv * dv : (real)0;
      }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i) {
      real invdxt = (real)(1.0 / (2.0 * PI / (double)NSEC * (double)Rmed[i]));
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j), kl = id(i, jm1(j));
        if (i > 0) {
          size_t km = id(i - 1, j);
          VradNew[k] = VradInt[k] - rdt * (real)2.0 / (Dens[k] + Dens[km]) * (DensInt[k] - DensInt[km]) * invdiffRmed[i];
        }
        VthetaNew[k] = VthetaInt[k] - rdt * (real)2.0 / (Dens[k] + Dens[kl]) * (TemperInt[k] - TemperInt[kl]) * invdxt;
        if (P.adiabatic)
          EnergyInt[k] = Energy[k] - rdt * DensInt[k] * (VradInt[id(i + 1, j)] - VradInt[k]) * invdiffRsup[i] -
                         rdt * TemperInt[k] * (VthetaInt[id(i, jp1(j))] - VthetaInt[k]) * invdxt;
      }
    }
  }
  // Substep3 (SourceEuler.cu:348-362) + kernels (Kernels.cu:91-139)
  void substep3(double dt) {
    for (int i = 0; i < NRAD; ++i) viscosity_array[i] = (real)fviscosity(Rmed[i]);
    real rdt = (real)dt, gam = (real)P.adiabaticindex;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        real nu = viscosity_array[i];
        if (nu != 0) {
          Qplus[k] = (real)0.5 / nu / Dens[k] * (TAURR[k] * TAURR[k] + TAURP[k] * TAURP[k] + TAUPP[k] * TAUPP[k]);
          Qplus[k] += (real)(2.0 / 9.0) * nu * Dens[k] * DivV[k] * DivV[k];
        } else Qplus[k] = 0;
      }
    for (int j = 0; j < NSEC; ++j) {
      if (viscosity_array[NRAD - 1] != 0) {  // Q11
        real q1 = Qplus[id(1, j)], q2 = Qplus[id(2, j)];
        Qplus[id(0, j)] = (real)((double)q1 * std::exp(std::log((double)q1 / (double)q2) *
                                                       std::log((double)Rmed[0] / (double)Rmed[1]) /
                                                       std::log((double)Rmed[1] / (double)Rmed[2])));
      } else Qplus[id(0, j)] = 0;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        if (!P.cooling) {
          real num = rdt * Qplus[k] + EnergyInt[k];
          real den = (real)1.0 + (gam - (real)1.0) * rdt * DivV[k];
          EnergyNew[k] = num / den;
        } else {
          real num = EnergyMed[i] * rdt * Dens[k] / SigmaMed[i] + CoolingTimeMed[i] * EnergyInt[k] +
                     rdt * CoolingTimeMed[i] * (Qplus[k] - QplusMed[i] * Dens[k] / SigmaMed[i]);
          real den = rdt + CoolingTimeMed[i] + (gam - (real)1.0) * rdt * CoolingTimeMed[i] * DivV[k];
          EnergyNew[k] = num / den;
        }
      }
  }

  // -------------------------------------------------------------- transport
  // StarRadKernel + StarRadKernel2 (Kernels.cu:909-947): dq kept ring-major here
  void star_rad(const std::vector<real>& Q, std::vector<real>& QStar, double dt) {
    real rdt = (real)dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        if (i == 0 || i == NRAD - 1) dq[k] = 0;
        else {
          real dqm = (Q[k] - Q[id(i - 1, j)]) * invdiffRmed[i];
          real dqp = (Q[id(i + 1, j)] - Q[k]) * invdiffRmed[i + 1];
          dq[k] = (dqp * dqm > 0) ? (real)2.0 * dqp * dqm / (dqp + dqm) : (real)0;
        }
      }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        if (Vrad[k] > 0 && i > 0)
          QStar[k] = Q[id(i - 1, j)] + (Rmed[i] - Rmed[i - 1] - Vrad[k] * rdt) * (real)0.5 * dq[id(i - 1, j)];
        else
          QStar[k] = Q[k] - (Rmed[i + 1] - Rmed[i] + Vrad[k] * rdt) * (real)0.5 * dq[k];
      }
    for (int j = 0; j < NSEC; ++j) { QStar[id(0, j)] = 0; QStar[id(NRAD, j)] = 0; }
  }
  // DivisePolarGridKernel (Kernels.cu:1127-1134) over the cells it covers
  void divise(const std::vector<real>& Q) {
    size_t n = std::min(NG, (size_t)(NSEC + 1) * (size_t)NRAD);
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < (long long)n; ++k) Work[k] = Q[k] / (DensInt[k] + (real)1e-20);
  }
  // VanLeerRadial (TransportEuler.cu:100-150) + VanLeerRadialKernel (Kernels.cu:1137-1154)
  double van_leer_radial(std::vector<real>& Q, double dt, bool ret_lost) {
    divise(Q);
    star_rad(Work, QRStar, dt);
    real dtheta = (real)(2.0 * PI / (double)NSEC), rdt = (real)dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j), ku = id(i + 1, j);
        real varq = rdt * dtheta * Rinf[i] * QRStar[k] * DensStar[k] * Vrad[k];
        varq -= rdt * dtheta * Rsup[i] * QRStar[ku] * DensStar[ku] * Vrad[ku];
        Q[k] += varq * invSurf[i];
        if (i == 0 && P.open_inner) LostByDisk[j] = varq;
      }
    double lost = 0;
    if (ret_lost && P.open_inner)
      for (int j = 0; j < NSEC; ++j) lost += (double)LostByDisk[j];
    return lost;
  }
  // StarThetaKernel + StarThetaKernel2 (Kernels.cu:1235-1273)
  void star_theta(const std::vector<real>& Q, const std::vector<real>& V, std::vector<real>& QStar, double dt) {
    real rdt = (real)dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i) {
      real dxt = (real)(2.0 * PI / (double)NSEC * (double)Rmed[i]);
      real invdxt = (real)1.0 / dxt;
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        real dqm = Q[k] - Q[id(i, jm1(j))];
        real dqp = Q[id(i, jp1(j))] - Q[k];
        dq[k] = (dqp * dqm > 0) ? dqp * dqm / (dqp + dqm) * invdxt : (real)0;
      }
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        real ksi = V[k] * rdt;
        if (ksi > 0) QStar[k] = Q[id(i, jm1(j))] + (dxt - ksi) * dq[id(i, jm1(j))];
        else QStar[k] = Q[k] - (dxt + ksi) * dq[k];
      }
    }
  }
  // VanLeerTheta (TransportEuler.cu:243-251) + VanLeerThetaKernel (Kernels.cu:1157-1174)
  void van_leer_theta(const std::vector<real>& V, std::vector<real>& Q, double dt) {
    divise(Q);
    star_theta(Work, V, QRStar, dt);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i) {
      real invsurf = (real)1.0 / Surf[i];
      real dxrad = (Rsup[i] - Rinf[i]) * (real)dt;
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j), kr = id(i, jp1(j));
        real varq = dxrad * QRStar[k] * DensStar[k] * V[k];
        varq -= dxrad * QRStar[kr] * DensStar[kr] * V[kr];
        Q[k] += varq * invsurf;
      }
    }
  }
  // QuantitiesAdvection (TransportEuler.cu:223-239)
  void quantities_advection(const std::vector<real>& V, double dt) {
    star_theta(Dens, V, DensStar, dt);
    DensInt = Dens;
    van_leer_theta(V, RadMomP, dt);
    van_leer_theta(V, RadMomM, dt);
    van_leer_theta(V, ThetaMomP, dt);
    van_leer_theta(V, ThetaMomM, dt);
    if (P.adiabatic) van_leer_theta(V, Energy, dt);
    if (advect_label()) van_leer_theta(V, ExtLabel, dt);
    van_leer_theta(V, Dens, dt);
  }
  // Label advection (LABELADVECTION YES): the reference advects only the
  // never-filled ExtLabel radially and leaves the extensive/specific
  // conversions and the azimuthal pass commented out (TransportEuler.cu:28,
  // 34,68-69,186,235-236), so the label never moves; with StubbedPhysics it is
  // advected as FARGO does, as the density-weighted quantity Dens*Label
  bool advect_label() const { return P.advect_label && P.ext_stubbed_physics; }
  // ComputeConstantResidualKernel (Kernels.cu:1201-1232): integer FARGO shift
  static int fargo_shift(real vmed, real invrmed, double dt, int nsec, real& ntilde_out, real& nround_out) {
    if constexpr (sizeof(real) == 4) {
      // reference precision: float product, then /2.0/PI in double
      float prod = (float)vmed * (float)invrmed * (float)dt * (float)nsec;
      float ntilde = (float)((double)prod / 2.0 / PI);
      float nround = (float)std::floor((double)ntilde + 0.5);
      ntilde_out = ntilde; nround_out = nround;
      return (int)(long)nround;
    } else {
      double ntilde = (double)vmed * (double)invrmed * dt * (double)nsec / 2.0 / PI;
      double nround = std::floor(ntilde + 0.5);
      ntilde_out = (real)ntilde; nround_out = (real)nround;
      return (int)(long)nround;
    }
  }
  void advect_shift(std::vector<real>& Q) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        int ji = j - Nshift[i];
        while (ji < 0) ji += NSEC;
        while (ji >= NSEC) ji -= NSEC;
        TempShift[id(i, j)] = Q[id(i, ji)];
      }
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) Q[id(i, j)] = TempShift[id(i, j)];
  }
  // OneWindTheta (TransportEuler.cu:171-188)
  void one_wind_theta(double dt) {
    for (int i = 0; i < NRAD; ++i) VMed[i] = ring_mean(Vtheta, i);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i) {
      for (int j = 0; j < NSEC; ++j) VthetaRes[id(i, j)] = Vtheta[id(i, j)] - VMed[i];
      real ntilde, nround;
      Nshift[i] = fargo_shift(VMed[i], invRmed[i], dt, NSEC, ntilde, nround);
      real invdt = (real)(1.0 / dt);
      real dpinvns = (real)(2.0 * PI / (double)NSEC);
      real vconst = (ntilde - nround) * Rmed[i] * invdt * dpinvns;
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        Vtheta[k] = vconst;
        if (!P.fast_transport) {
          NoSplitAdvection[i] = 1;
          VthetaRes[k] = Vtheta[k] + VthetaRes[k];
          Vtheta[k] = 0;
        } else NoSplitAdvection[i] = 0;
      }
    }
    quantities_advection(VthetaRes, dt);
    quantities_advection(Vtheta, dt);
    advect_shift(RadMomP);
    advect_shift(RadMomM);
    advect_shift(ThetaMomP);
    advect_shift(ThetaMomM);
    if (P.adiabatic) advect_shift(Energy);
    if (advect_label()) advect_shift(ExtLabel);
    advect_shift(Dens);
  }
  // Transport (TransportEuler.cu:24-35)
  void transport(double dt) {
    // ComputeLRMomenta (LRMomentaKernel, Kernels.cu:883-896)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        RadMomP[k] = Dens[k] * Vrad[id(i + 1, j)];
        RadMomM[k] = Dens[k] * Vrad[k];
        ThetaMomP[k] = Dens[k] * (Vtheta[id(i, jp1(j))] + Rmed[i] * (real)OmegaFrame) * Rmed[i];
        ThetaMomM[k] = Dens[k] * (Vtheta[k] + Rmed[i] * (real)OmegaFrame) * Rmed[i];
      }
    // ComputeExtQty (ExtQtyKernel, Kernels.cu:899-906)
    if (advect_label())
      for (int i = 0; i < NRAD; ++i)
        for (int j = 0; j < NSEC; ++j) ExtLabel[id(i, j)] = Dens[id(i, j)] * Label[id(i, j)];
    // OneWindRad (TransportEuler.cu:54-73)
    star_rad(Dens, DensStar, dt);
    DensInt = Dens;
    van_leer_radial(RadMomP, dt, false);
    van_leer_radial(RadMomM, dt, false);
    van_leer_radial(ThetaMomP, dt, false);
    van_leer_radial(ThetaMomM, dt, false);
    if (P.adiabatic) van_leer_radial(Energy, dt, false);
    if (advect_label()) van_leer_radial(ExtLabel, dt, false);
    LostMass += van_leer_radial(Dens, dt, true);
    one_wind_theta(dt);
    // ComputeVelocitiesKernel (Kernels.cu:1292-1306)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NRAD; ++i)
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j), kl = id(i, jm1(j));
        if (i == 0) Vrad[k] = 0;
        else
          Vrad[k] = (RadMomP[id(i - 1, j)] + RadMomM[k]) / (Dens[k] + Dens[id(i - 1, j)] + (real)1e-20);
        Vtheta[k] = (ThetaMomP[kl] + ThetaMomM[k]) / (Dens[k] + Dens[kl] + (real)1e-15) / Rmed[i] -
                    Rmed[i] * (real)OmegaFrame;
      }
    // ComputeSpeQty (ComputeSpeQtyKernel, Kernels.cu:1309-1319)
    if (advect_label())
      for (int i = 0; i < NRAD; ++i)
        for (int j = 0; j < NSEC; ++j) Label[id(i, j)] = ExtLabel[id(i, j)] / Dens[id(i, j)];
  }

  // CircumPlanetaryMass (SourceEuler.cu:539-553, Kernels.cu:245-261)
  double circumplanetary_mass() {
    if (sys.nb == 0) return 0.0;
    double xpl = (float)sys.x[0], ypl = (float)sys.y[0];
    double s = 0;
    for (int i = 0; i < NRAD; ++i) {
      double rs = 0;
      for (int j = 0; j < NSEC; ++j) {
        size_t k = id(i, j);
        double dx = (double)CellAbscissa[k] - xpl, dy = (double)CellOrdinate[k] - ypl;
        if (std::sqrt(dx * dx + dy * dy) < HillRadius) rs += (double)Surf[i] * (double)Dens[k];
      }
      s += rs;
    }
    return s;
  }

  // AccreteOntoPlanets, FARGO's Planet.c (W. Kley's scheme): the reference
  // computes the Roche box but its kernel launch is commented out
  // (Planet.cu:7-47).  Gas within 0.75 R_Roche loses facc/3 of its density per
  // substep, gas within 0.45 R_Roche a further 2 facc/3, facc = dt * acc; the
  // planet gains that mass and its momentum (cell velocity at the cell
  // centre from the face-centred Vrad, Vtheta).  The angle is real as in
  // FARGO (the reference declares it int, Planet.cu:9).  A box wider than the
  // ring is cut at NSEC sectors so no cell is visited twice.
  void accrete_onto_planets(double dt) {
    for (int k = 0; k < sys.nb; ++k) {
      if (!(sys.acc[k] > 1e-10)) continue;
      double dM = 0, dPx = 0, dPy = 0;
      const double facc = dt * sys.acc[k], facc1 = 1.0 / 3.0 * facc, facc2 = 2.0 / 3.0 * facc;
      const double frac1 = 0.75, frac2 = 0.45;
      const double xp = sys.x[k], yp = sys.y[k], vxp = sys.vx[k], vyp = sys.vy[k], mp = sys.mass[k];
      const double rp = std::sqrt(xp * xp + yp * yp);
      const double rroche = std::pow(1.0 / 3.0 * mp, 1.0 / 3.0) * rp;
      int imin = 0, imax = NRAD - 1;
      while (imin < NRAD && (double)Rsup[imin] < rp - rroche) imin++;
      while (imax > 0 && (double)Rinf[imax] > rp + rroche) imax--;
      const double angle = std::atan2(yp, xp);
      const int jmin = (int)((double)NSEC / 2.0 / PI * (angle - 2.0 * rroche / rp));
      int jmax = (int)((double)NSEC / 2.0 / PI * (angle + 2.0 * rroche / rp));
      if (jmax - jmin + 1 > NSEC) jmax = jmin + NSEC - 1;
      for (int i = imin; i <= imax; ++i)
        for (int j = jmin; j <= jmax; ++j) {
          const int jf = ((j % NSEC) + NSEC) % NSEC;
          const size_t l = id(i, jf), lip = id(i + 1, jf), ljp = id(i, jp1(jf));
          const double xc = (double)(Rmed[i] * cosj[jf]), yc = (double)(Rmed[i] * sinj[jf]);
          const double dx = xp - xc, dy = yp - yc;
          const double dist = std::sqrt(dx * dx + dy * dy);
          const double vt = 0.5 * ((double)Vtheta[l] + (double)Vtheta[ljp]) + (double)Rmed[i] * OmegaFrame;
          const double vr = 0.5 * ((double)Vrad[l] + (double)Vrad[lip]);
          const double vx = (vr * xc - vt * yc) / (double)Rmed[i], vy = (vr * yc + vt * xc) / (double)Rmed[i];
          if (dist < frac1 * rroche) {
            const double m = facc1 * (double)Dens[l] * (double)Surf[i];
            Dens[l] = (real)((double)Dens[l] * (1.0 - facc1));
            dPx += m * vx; dPy += m * vy; dM += m;
          }
          if (dist < frac2 * rroche) {
            const double m = facc2 * (double)Dens[l] * (double)Surf[i];
            Dens[l] = (real)((double)Dens[l] * (1.0 - facc2));
            dPx += m * vx; dPy += m * vy; dM += m;
          }
        }
      const double px = mp * vxp + dPx, py = mp * vyp + dPy, m = mp + dM;
      if (sys.feel_disk[k]) { sys.vx[k] = px / m; sys.vy[k] = py / m; }
      sys.mass[k] = m;
    }
  }

  // ---------------------------------------------------------------- AlgoGas
  // AlgoGas (SourceEuler.cu:208-301); returns the number of substeps
  int algo_gas() {
    double DT = P.dt, dtemp = 0.0, dt;
    int n = 1;
    int nsub = 0;
    if (P.adiabatic) compute_sound_speed();
    auto cfl = [&](double dT) -> int {
      if (P.ext_substeps > 0) return std::max(1, P.ext_substeps - nsub);
      return condition_cfl(dT);
    };
    if (P.is_disk && P.sloppy_cfl) n = cfl(DT - dtemp);
    dt = DT / (double)n;
    while (dtemp < 0.999999999 * DT) {
      double mt = PhysicalTime / (P.masstaper * 2.0 * M_PI);
      MassTaper = (mt > 1.0) ? 1.0 : std::pow(std::sin(mt * M_PI / 2.0), 2.0);
      if (P.is_disk && !P.sloppy_cfl) {
        n = cfl(DT - dtemp);
        dt = (DT - dtemp) / (double)n;
      }
      if (P.is_disk && P.ext_stubbed_physics) accrete_onto_planets(dt);
      last_ncfl = n;
      dtemp += dt;
      DOPAx = DOPAy = 0.0;
      if (P.corotating) get_psys_info(1);
      if (P.is_disk) {
        compute_accel(0, 0, 0, 0, DOPAx, DOPAy);
        fill_forces_arrays();
        advance_system_from_disk(dt);
      }
      advance_system_rk5(dt);
      if (P.corotating) {
        double OmegaNew = get_psys_info(0) / dt;
        double domega = OmegaNew - OmegaFrame;
        if (P.is_disk) {  // CorrectVthetaKernel (Kernels.cu:1352-1359)
          for (int i = 0; i < NRAD; ++i)
            for (int j = 0; j < NSEC; ++j) Vtheta[id(i, j)] -= (real)domega * Rmed[i];
        }
        OmegaFrame = OmegaNew;
      }
      rotate_psys(OmegaFrame * dt);
      if (P.is_disk) {
        apply_boundary(dt);
        compute_pressure();
        substep1(dt);
        substep2(dt);
        std::copy(VradNew.begin(), VradNew.end(), Vrad.begin());
        std::copy(VthetaNew.begin(), VthetaNew.end(), Vtheta.begin());
        apply_boundary(dt);
        if (P.adiabatic) {
          compute_viscous_terms(Vrad, Vtheta);
          substep3(dt);
          std::copy(EnergyNew.begin(), EnergyNew.end(), Energy.begin());
        }
        transport(dt);
        apply_boundary(dt);
        compute_temperature();
        mdcp1 = circumplanetary_mass();
        exces_mdcp = mdcp1 - mdcp;
      }
      PhysicalTime += dt;
      last_dt = dt;
      ++nsub; ++substeps;
    }
    return nsub;
  }
};

}  // namespace orc
