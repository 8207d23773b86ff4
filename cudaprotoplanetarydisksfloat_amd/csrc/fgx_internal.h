// fgx_internal.h -- declarations shared by the host set-up, the HIP engine
// and the C ABI.  Not installed; the public surface is include/fgx.h.
#pragma once
#include "../../include/fgx.h"

#include <string>
#include <vector>

namespace fgx {

// thread-local last error; returns `code`
int set_error(int code, const char* fmt, ...);
const char* last_error();

// ---- parameters / planets (params.cpp) ------------------------------------
void params_defaults(fgx_params* p);
int params_parse_text(const char* text, fgx_params* p);
int params_set_key(fgx_params* p, const char* key, const char* value);
int params_get_key(const fgx_params* p, const char* key, char* out, size_t n);
int params_validate(const fgx_params* p);
bool batch_modes_equal(const fgx_params& p, const fgx_params& P0);
int planets_parse_text(const char* text, const fgx_params* p, const double* rmed,
                       const double* radii, fgx_planets* out);
bool read_file(const char* path, std::string& out);

constexpr double kG = 1.0;
constexpr double kPI = 3.14159265358979323844;   // include/Fondam.cuh:3
constexpr double kCFLSECURITY = 0.5;             // include/Fondam.cuh:7
constexpr double kCVNR = 1.41;                   // include/Fondam.cuh:10

// ---- host-side set-up of one disk, FP64 (setup.cpp) --------------------------
// Everything main() computes before its loop that does not need the device:
// ring geometry, the analytic profiles, the initial fields.
struct HostDisk {
  fgx_params P;
  fgx_planets sys;
  int nrad = 0, nsec = 0;
  // ring tables, nrad+1 entries (FillPolar1DArrays, SourceEuler.cu:76-157)
  std::vector<double> radii, rinf, rsup, rmed, surf, invrmed, invsurf, invdiffrsup,
      invdiffrmed, invrinf, sigmamed, sigmainf, energymed, coolingtime, qplusmed,
      supp_torque, dxtheta, visc, cs_ring;
  std::vector<double> cosj, sinj;  // per sector
  // initial fields, (nrad+1)*nsec
  std::vector<double> dens, vrad, vtheta, energy, label;
  double omega_frame = 0, hill_radius = 0;
  double sg_eps = 0, sg_rstep = 0, sg_tstep = 0;   // SGP_eps, SGP_rstep, SGP_tstep
};

// grid, profiles, density (with noise) and energy; velocities are filled by
// setup_velocities once the self-gravity field (if any) is known
int setup_grid_and_density(const fgx_params& p, const fgx_planets& pl, HostDisk& h);
// InitGasVelocities / InitVelocities (Pframeforce.cu:194-272); sg_axi is the
// ring mean of the radial SG acceleration, or empty without self-gravity
void setup_velocities(HostDisk& h, const std::vector<double>& sg_axi);
// FViscosity at time t (Viscosity.cu:19-37); csmean = ring-mean sound speed
double fviscosity(const fgx_params& P, double r, double csmean, double t);
double aspect_ratio(const fgx_params& P, double r, double t);
// GetPsysInfo(FREQUENCY) (Psys.cu:162-232)
double psys_frequency(const fgx_params& P, const fgx_planets& s);
// self-gravity (sg.cpp): radix factorisation of a transform length (4, 2, 3,
// 5), and the per-disk set-up: FP64 kernel transforms (interleaved re/im,
// (2*nrad) x nsec stored by column, entry (i, j) at j*2*nrad + i), the ring mean of the initial radial SG acceleration, and
// the planets' SG-corrected circular velocities (Init_planetarysys_withSG)
#define FGX_SG_MAXFAC 24
int sg_factor(int n, int* fac, int* nf);
int sg_host_setup(HostDisk& h, std::vector<double>& khat_r, std::vector<double>& khat_t,
                  std::vector<double>& axisg);

// engine accessors used by the output writers (engine.hip)
const HostDisk* batch_host(fgx_batch* b, int d);
int batch_mdcp(fgx_batch* b, int d, double* out);
// serving: queue an asynchronous copy of a disk's stored fields into its
// pinned host buffer, and wait for the last one queued
int batch_snapshot_queue(fgx_batch* b, int d);
int batch_snapshot_wait(fgx_batch* b, int d, int* had);
// the name of the current HIP device (the autotuner keys its records on it)
int device_name(char* out, size_t n);

}  // namespace fgx

// ---- device-visible per-disk records ------------------------------------
#define FGX_NFORCE_T (1 + FGX_MAX_PLANETS)  // star + planets

// constant description of one disk of a batch
struct DiskConst {
  int nrad, nsec;
  long long fo;       // element offset of this disk in every field array
  int ro;             // offset in the ring arrays (nrad+1 entries per disk)
  int so;             // offset in the sector arrays (nsec entries per disk)
  int rt0;            // index of this disk's first ring tile
  int nb;             // planets
  int feel_disk[FGX_MAX_PLANETS], feel_others[FGX_MAX_PLANETS];
  double acc[FGX_MAX_PLANETS];   // accretion rates (PlanetarySystem acc, Psys.cu:126)
  int accretes;                  // StubbedPhysics with some rate above 1e-10: FARGO accretion
  int outer_source;              // StubbedPhysics + OUTERSOURCEMASS: FARGO's outer mass source
  double outer_vr;               // its radial velocity, IMPOSEDDISKDRIFT * Rinf[nrad-1]^-SIGMASLOPE
  // constants of the substep start: Rmed[0]^(2 FLARINGINDEX), Rmed[nrad-1]^(2
  // FLARINGINDEX) and the EvanescentBoundary constants (SideEuler.cu:125-130)
  double pow_in_2f, pow_out_2f, ev_tin, ev_tout, ev_drmin, ev_drmax;
  double DT, masstaper, sigmaslope, flaringindex, aspectratio, adiabaticindex;
  double thicknesssmoothing, rochesmoothing, releasedate, releaseradius;
  double transitionradius, transitionratio, transitionwidth, lambdadoubling;
  double cavityradius, cavityratio, cavitywidth, viscosity, alphaviscosity;
  double hill_radius, sigma0, eccentricity;
  double omega0;           // OmegaFrame at set-up: the frame of Substep1's profile tables
  int substeps_forced;     // SUBSTEPS key, 0 = CFL
  int max_substeps;        // MAXSUBSTEPS: a larger CFL count marks the disk unhealthy
  int visc_dynamic;        // recompute the viscosity profile each substep
  double sg_eps, sg_rstep, sg_tstep;
  // self-gravity transforms: offsets of this disk's work rows (nrad*nsec
  // complex), kernel transforms ((2*nrad)*nsec complex) and twiddle tables
  // (nsec and 2*nrad entries), radix factorisations of nsec and 2*nrad
  long long sgw_o, sgk_o;
  int tw_ns, tw_nr, nf_ns, nf_nr;
  int fac_ns[24], fac_nr[24];
};

// a complex FP64 value (self-gravity transforms)
struct cxd { double x, y; };

// mutable FP64 state of one disk (planets, time, frame, per-substep scalars)
struct DiskState {
  double time, time0, dtemp, dt, omega_frame, masstaper;
  double mass[FGX_MAX_PLANETS];         // planet masses (change only by accretion)
  double x[FGX_MAX_PLANETS], y[FGX_MAX_PLANETS], vx[FGX_MAX_PLANETS], vy[FGX_MAX_PLANETS];
  // potential of this substep (FillForcesArrays, Pframeforce.cu:37-66)
  double pot_x[FGX_MAX_PLANETS], pot_y[FGX_MAX_PLANETS], pot_m[FGX_MAX_PLANETS],
      pot_sm2[FGX_MAX_PLANETS], pot_inv3[FGX_MAX_PLANETS];
  double ind_x, ind_y;                  // IndirectTerm
  // force targets of this substep: star + feel-disk planets
  double tx[FGX_NFORCE_T], ty[FGX_NFORCE_T], tsm2[FGX_NFORCE_T], trh2[FGX_NFORCE_T];
  int tk[FGX_NFORCE_T];                 // Hill-cut variant (0 or 10)
  int ntarget;
  double mark_x, mark_y;                // GetPsysInfo(MARK) statics
  double domega;                        // corotation correction of this substep
  double vkep_in, vkep_out;             // ApplySubKeplerianBoundary values
  double lost_mass, mdcp1, exces_mdcp;
  double ev_tin, ev_tout, ev_drmin, ev_drmax;  // EvanescentBoundary constants
  long long substeps, dt_steps;
  int ncfl, active, sub_in_dt, health;
  int pending;                          // a substep ran and its end-of-substep update is still due
  int nr_iangle, nr_iangle2;            // non-reflecting shifts of this call
  int frozen;                           // skipped by every later DT (served request done)
  int vpar;                             // velocity buffer pair holding this disk's Vrad/Vtheta
};

// a field tile: ring i of disk d, sectors [jb*256, jb*256+256)
struct CellTile { int d, i, jb, pad; };
// a ring tile: ring i of disk d
struct RingTile { int d, i; };
