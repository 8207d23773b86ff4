// oracle_capi.cpp -- C entry points of the CPU oracle (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker; never by the product library.
// See oracle.hpp for what is restated and where it deviates.
#include "oracle.hpp"

#include <cstdio>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <memory>
#include <string>

namespace {

struct Handle {
  int precision;  // 0 float, 1 double, 2 long double (x87 80-bit, 64-bit mantissa)
  std::unique_ptr<orc::Disk<float>> f;
  std::unique_ptr<orc::Disk<double>> d;
  std::unique_ptr<orc::Disk<long double>> q;
};

template <class D>
std::vector<typename std::remove_reference<decltype(std::declval<D>().Dens)>::type::value_type>* field_of(D& s, int f) {
  switch (f) {
    case FGX_DENS: return &s.Dens;
    case FGX_VRAD: return &s.Vrad;
    case FGX_VTHETA: return &s.Vtheta;
    case FGX_ENERGY: return &s.Energy;
    case FGX_TEMPERATURE: return &s.Temperature;
    case FGX_SOUNDSPEED: return &s.SoundSpeed;
    case FGX_DIVV: return &s.DivV;
    case FGX_QPLUS: return &s.Qplus;
    case FGX_LABEL: return &s.Label;
    case FGX_POTENTIAL: return &s.Potential;
    case FGX_SG_ACCR: return &s.SG_Accr;
    case FGX_SG_ACCT: return &s.SG_Acct;
    default: return nullptr;
  }
}

template <class D>
int get_field(D& s, int f, double* out, size_t n) {
  auto* v = field_of(s, f);
  if (!v || n > v->size()) return FGX_ERANGE;
  for (size_t k = 0; k < n; ++k) out[k] = (double)(*v)[k];
  return 0;
}
template <class D>
int set_field(D& s, int f, const double* in, size_t n) {
  auto* v = field_of(s, f);
  if (!v || n > v->size()) return FGX_ERANGE;
  using T = typename std::remove_reference<decltype(*v)>::type::value_type;
  for (size_t k = 0; k < n; ++k) (*v)[k] = (T)in[k];
  return 0;
}
template <class D>
int get_ring(D& s, int r, double* out, size_t n) {
  if (n > (size_t)s.NRAD + 1) return FGX_ERANGE;
  for (size_t i = 0; i < n; ++i) {
    double v = 0;
    switch (r) {
      case FGX_R_RADII: v = s.Radii[i]; break;
      case FGX_R_RMED: v = s.Rmed[i]; break;
      case FGX_R_SURF: v = s.Surf[i]; break;
      case FGX_R_SIGMAMED: v = s.SigmaMed[i]; break;
      case FGX_R_ENERGYMED: v = s.EnergyMed[i]; break;
      case FGX_R_VISCOSITY: v = s.viscosity_array[i]; break;
      case FGX_R_VMED: v = s.VMed[i]; break;
      case FGX_R_NSHIFT: v = s.Nshift[i]; break;
      case FGX_R_CS: v = (i < (size_t)s.NRAD) ? (double)s.SoundSpeed[i * s.NSEC] : 0.0; break;
      case FGX_R_NEWDT: v = s.newDT[i]; break;
      case FGX_R_INVRMED: v = s.invRmed[i]; break;
      default: return FGX_ERANGE;
    }
    out[i] = v;
  }
  return 0;
}
template <class D>
void info(D& s, double* o) {
  o[0] = s.PhysicalTime; o[1] = s.OmegaFrame; o[2] = s.LostMass; o[3] = s.mdcp1; o[4] = s.exces_mdcp;
  o[5] = s.last_dt; o[6] = s.last_ncfl; o[7] = (double)s.substeps; o[8] = s.MassTaper;
  o[9] = s.Indx; o[10] = s.Indy;
}

// Named arrays of the restatement, for the pins of single functions
// (tests/test_oracle_pins.py): the work fields, the stresses, the star values
template <class D>
std::vector<typename std::remove_reference<decltype(std::declval<D>().Dens)>::type::value_type>* array_of(D& s,
                                                                                                       const std::string& n) {
  if (n == "Dens") return &s.Dens;
  if (n == "Vrad") return &s.Vrad;
  if (n == "Vtheta") return &s.Vtheta;
  if (n == "Energy") return &s.Energy;
  if (n == "SoundSpeed") return &s.SoundSpeed;
  if (n == "Pressure") return &s.Pressure;
  if (n == "Potential") return &s.Potential;
  if (n == "VradInt") return &s.VradInt;
  if (n == "VthetaInt") return &s.VthetaInt;
  if (n == "VradNew") return &s.VradNew;
  if (n == "VthetaNew") return &s.VthetaNew;
  if (n == "EnergyInt") return &s.EnergyInt;
  if (n == "EnergyNew") return &s.EnergyNew;
  if (n == "DRR") return &s.DRR;
  if (n == "DPP") return &s.DPP;
  if (n == "DRP") return &s.DRP;
  if (n == "DivV") return &s.DivV;
  if (n == "TAURR") return &s.TAURR;
  if (n == "TAUPP") return &s.TAUPP;
  if (n == "TAURP") return &s.TAURP;
  if (n == "Qplus") return &s.Qplus;
  if (n == "Work") return &s.Work;
  if (n == "QRStar") return &s.QRStar;
  if (n == "dq") return &s.dq;
  if (n == "VthetaRes") return &s.VthetaRes;
  return nullptr;
}

}  // namespace

extern "C" {

void* orc_create(const fgx_params* p, const fgx_planets* pl, int precision) {
  auto* h = new Handle();
  h->precision = precision;
  try {
    if (precision == 0) { h->f.reset(new orc::Disk<float>()); h->f->setup(*p, *pl); }
    else if (precision == 2) { h->q.reset(new orc::Disk<long double>()); h->q->setup(*p, *pl); }
    else { h->d.reset(new orc::Disk<double>()); h->d->setup(*p, *pl); }
  } catch (...) { delete h; return nullptr; }
  return h;
}
void orc_destroy(void* hp) { delete (Handle*)hp; }

#define DISPATCH(hp, expr)                                \
  Handle* h = (Handle*)hp;                                \
  if (h->precision == 0) { auto& s = *h->f; expr; }       \
  else if (h->precision == 2) { auto& s = *h->q; expr; }  \
  else { auto& s = *h->d; expr; }

// n calls of AlgoGas; returns the total number of substeps
long long orc_step(void* hp, int n) {
  long long tot = 0;
  DISPATCH(hp, for (int k = 0; k < n; ++k) tot += s.algo_gas());
  return tot;
}
int orc_get_field(void* hp, int f, double* out, size_t n) { DISPATCH(hp, return get_field(s, f, out, n)); }
int orc_set_field(void* hp, int f, const double* in, size_t n) { DISPATCH(hp, return set_field(s, f, in, n)); }
int orc_get_ring(void* hp, int r, double* out, size_t n) { DISPATCH(hp, return get_ring(s, r, out, n)); }
void orc_get_planets(void* hp, fgx_planets* out) { DISPATCH(hp, *out = s.sys); }
void orc_set_planets(void* hp, const fgx_planets* in) { DISPATCH(hp, s.sys = *in); }
void orc_info(void* hp, double* out11) { DISPATCH(hp, info(s, out11)); }
void orc_force(void* hp, double x, double y, double sm, double m, double* out44) {
  DISPATCH(hp, s.compute_force(x, y, sm, m, out44));
}
// ConditionCFL on the current state (SourceEuler.cu:494-536)
int orc_condition_cfl(void* hp, double DeltaT) { DISPATCH(hp, return s.condition_cfl(DeltaT)); }
// one named stage of the substep, for per-stage parity checks
int orc_stage(void* hp, const char* name, double dt) {
  std::string n(name);
  DISPATCH(hp, {
    if (n == "pressure") s.compute_pressure();
    else if (n == "soundspeed") s.compute_sound_speed();
    else if (n == "substep1") s.substep1(dt);
    else if (n == "substep2") {
      s.substep2(dt);
      std::copy(s.VradNew.begin(), s.VradNew.end(), s.Vrad.begin());
      std::copy(s.VthetaNew.begin(), s.VthetaNew.end(), s.Vtheta.begin());
    } else if (n == "substep3") {
      s.compute_viscous_terms(s.Vrad, s.Vtheta);
      s.substep3(dt);
      std::copy(s.EnergyNew.begin(), s.EnergyNew.end(), s.Energy.begin());
    } else if (n == "transport") s.transport(dt);
    else if (n == "boundary") s.apply_boundary(dt);
    else if (n == "potential") s.fill_forces_arrays();
    else if (n == "accrete") s.accrete_onto_planets(dt);
    else return -1;
  });
  return 0;
}
// VradInt / VthetaInt after a substep1 stage
int orc_get_int_velocities(void* hp, double* vr, double* vt, size_t n) {
  DISPATCH(hp, {
    if (n > s.NG) return FGX_ERANGE;
    for (size_t k = 0; k < n; ++k) { vr[k] = s.VradInt[k]; vt[k] = s.VthetaInt[k]; }
  });
  return 0;
}
int orc_get_array(void* hp, const char* name, double* out, size_t n) {
  DISPATCH(hp, {
    auto* v = array_of(s, std::string(name));
    if (!v || n > v->size()) return FGX_ERANGE;
    for (size_t k = 0; k < n; ++k) out[k] = (double)(*v)[k];
  });
  return 0;
}
int orc_set_array(void* hp, const char* name, const double* in, size_t n) {
  DISPATCH(hp, {
    auto* v = array_of(s, std::string(name));
    if (!v || n > v->size()) return FGX_ERANGE;
    using T = typename std::remove_reference<decltype(*v)>::type::value_type;
    for (size_t k = 0; k < n; ++k) (*v)[k] = (T)in[k];
  });
  return 0;
}
// single functions of the restatement on the current arrays
int orc_call(void* hp, const char* name, double dt) {
  std::string n(name);
  DISPATCH(hp, {
    if (n == "viscous_terms") s.compute_viscous_terms(s.Vrad, s.Vtheta);
    else if (n == "visc_update") s.update_velocities_with_viscosity(dt);
    else if (n == "substep1_kernels") s.substep1_kernels(dt);
    else if (n == "substep2") s.substep2(dt);
    else if (n == "substep3") s.substep3(dt);
    else if (n == "star_rad") s.star_rad(s.Work, s.QRStar, dt);
    else if (n == "star_theta") s.star_theta(s.Work, s.VthetaRes, s.QRStar, dt);
    else if (n == "open_boundary") s.open_boundary();
    else if (n == "nonreflecting_boundary") s.nonreflecting_boundary();
    else if (n == "evanescent_boundary") s.evanescent_boundary(dt);
    else if (n == "mdcp") s.mdcp1 = s.circumplanetary_mass();
    else if (n == "fill_forces_arrays") s.fill_forces_arrays();
    else return -1;
  });
  return 0;
}

// FARGO integer shift with the reference's FP32 promotion order
// (ComputeConstantResidualKernel, Kernels.cu:1216-1221)
int orc_fargo_shift_f32(float vmed, float invrmed, double dt, int nsec) {
  float a, b;
  return orc::Disk<float>::fargo_shift(vmed, invrmed, dt, nsec, a, b);
}
int orc_fargo_shift_f64(double vmed, double invrmed, double dt, int nsec) {
  double a, b;
  return orc::Disk<double>::fargo_shift(vmed, invrmed, dt, nsec, a, b);
}
// 2-D complex FFT of the oracle (rows x cols, sign -1 forward), in place
void orc_fft2d(double* re, double* im, int rows, int cols, int sign) {
  std::vector<std::complex<double>> a((size_t)rows * cols);
  for (size_t k = 0; k < a.size(); ++k) a[k] = {re[k], im[k]};
  orc::fft2d(a, rows, cols, sign);
  for (size_t k = 0; k < a.size(); ++k) { re[k] = a[k].real(); im[k] = a[k].imag(); }
}
// self-gravity accelerations of the current density (FFT path) into
// accr/acct (nrad*nsec each)
int orc_sg_accel(void* hp, double* accr, double* acct) {
  DISPATCH(hp, {
    if (!s.P.self_gravity) return FGX_EPARAM;
    s.sg_accel();
    for (int i = 0; i < s.NRAD; ++i)
      for (int j = 0; j < s.NSEC; ++j) {
        accr[(size_t)i * s.NSEC + j] = s.SG_Accr[s.id(i, j)];
        acct[(size_t)i * s.NSEC + j] = s.SG_Acct[s.id(i, j)];
      }
  });
  return 0;
}
size_t orc_sizeof_params(void) { return sizeof(fgx_params); }
size_t orc_sizeof_planets(void) { return sizeof(fgx_planets); }
int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
