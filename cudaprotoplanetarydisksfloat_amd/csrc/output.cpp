// output.cpp -- the reference's output files and its executable's main loop.
//
// File names, layouts, number formats and cadence follow src/Output.cu,
// src/Force.cu:76-138, src/SourceEuler.cu:159-171 and src/Main.cu:87-401.
// Values the reference keeps in float globals are rounded to float before
// printing with %.18g, so the text matches digit for digit in form.
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fgx_internal.h"

using fgx::set_error;

namespace {

std::string outdir(const fgx_params& P) { return std::string(P.outputdir); }

int write_raw(const std::string& dir, const std::string& name, const std::vector<float>& v) {
  struct stat st;
  if (!dir.empty() && stat(dir.c_str(), &st) == -1) mkdir(dir.c_str(), 0700);
  FILE* f = std::fopen(name.c_str(), "wb");
  if (!f) return set_error(FGX_EIO, "Unable to open '%s'", name.c_str());
  std::fwrite(v.data(), sizeof(float), v.size(), f);
  std::fclose(f);
  return FGX_OK;
}

// WriteDiskPolar (Output.cu:104-135) and the FARGO gas<name><N>.dat form
int write_field(fgx_disk* d, const fgx_params& P, int field, const char* name, const char* gasname, int index) {
  const fgx::HostDisk* h = fgx::batch_host(d, 0);
  std::vector<float> v((size_t)h->nrad * h->nsec);
  int r = fgx_disk_read_field(d, field, FGX_F32, v.data(), v.size());
  if (r) return r;
  const std::string od = outdir(P);
  if (P.ext_gas_names == 0 || P.ext_gas_names == 2) {
    std::string dir = od + name;
    r = write_raw(dir, dir + "/" + name + std::to_string(index) + ".dat", v);
    if (r) return r;
  }
  if (P.ext_gas_names >= 1) {
    r = write_raw("", od + gasname + std::to_string(index) + ".dat", v);
    if (r) return r;
  }
  return FGX_OK;
}

FILE* open_append(const std::string& name) { return std::fopen(name.c_str(), "a"); }

}  // namespace

extern "C" {

int fgx_disk_write_outputs(fgx_disk* d, int index) {
  if (!d) return set_error(FGX_EPARAM, "null handle");
  const fgx_params& P = fgx::batch_host(d, 0)->P;
  int r = FGX_OK;
  // SendOutput (Output.cu:155-170)
  if (!P.is_disk) return FGX_OK;
  if (P.write_density && (r = write_field(d, P, FGX_DENS, "dens", "gasdens", index))) return r;
  if (P.write_velocity) {
    if ((r = write_field(d, P, FGX_VRAD, "vrad", "gasvrad", index))) return r;
    if ((r = write_field(d, P, FGX_VTHETA, "vtheta", "gasvtheta", index))) return r;
  }
  if (P.write_energy && P.adiabatic && (r = write_field(d, P, FGX_ENERGY, "energy", "gasenergy", index))) return r;
  if (P.write_temperature && (r = write_field(d, P, FGX_TEMPERATURE, "temperature", "gasTemperature", index)))
    return r;
  if (P.write_divv && P.adiabatic && (r = write_field(d, P, FGX_DIVV, "divergence", "gasdivergence", index))) return r;
  if (P.write_qplus && P.adiabatic && (r = write_field(d, P, FGX_QPLUS, "qplus", "gasqplus", index))) return r;
  if (P.advect_label && (r = write_field(d, P, FGX_LABEL, "label", "gaslabel", index))) return r;
  return FGX_OK;
}

int fgx_disk_write_logs(fgx_disk* d, int timestep) {
  if (!d) return set_error(FGX_EPARAM, "null handle");
  const fgx::HostDisk* h = fgx::batch_host(d, 0);
  const fgx_params& P = h->P;
  fgx_planets sys;
  int r = fgx_disk_planets(d, &sys);
  if (r) return r;
  fgx_disk_info info;
  if ((r = fgx_disk_info_get(d, &info))) return r;
  const std::string od = outdir(P);
  const float lost = (float)info.lost_mass, ptime = (float)info.physical_time, om = (float)info.omega_frame;
  const float mdcp1 = (float)info.mdcp1, exces = (float)info.exces_mdcp;
  // WriteBigPlanetSystemFile / WriteBigPlanetFile (Output.cu:70-102)
  for (int k = 0; k < sys.nb; ++k) {
    std::string name = od + "bigplanet" + std::to_string(k) + ".dat";
    FILE* f = open_append(name);
    if (!f) return set_error(FGX_EIO, "Can't write 'bigplanet.dat' file.");
    std::fprintf(f, "%d\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\n", timestep,
                 (double)(float)sys.x[k], (double)(float)sys.y[k], (double)(float)sys.vx[k], (double)(float)sys.vy[k],
                 (double)(float)sys.mass[k], (double)lost, (double)ptime, (double)om, (double)mdcp1, (double)exces);
    std::fclose(f);
  }
  // UpdateLog (Force.cu:76-138)
  for (int k = 0; k < sys.nb; ++k) {
    const double x = sys.x[k], y = sys.y[k], vx = sys.vx[k], vy = sys.vy[k], m = sys.mass[k];
    const double rr = std::sqrt(x * x + y * y);
    const double sm = P.roche_smoothing ? rr * std::pow(m / 3., 1. / 3.) * P.rochesmoothing
                                        : P.thicknesssmoothing * fgx::aspect_ratio(P, rr, info.physical_time) *
                                              std::pow(rr, 1.0 + P.flaringindex);
    double gf[44];
    if ((r = fgx_disk_force(d, x, y, sm, m, gf))) return r;
    // the Force struct holds floats (include/Types.cuh:17-29)
    const double fxi = (float)gf[0], fxei = (float)gf[10], fxo = (float)gf[11], fxeo = (float)gf[21];
    const double fyi = (float)gf[22], fyei = (float)gf[32], fyo = (float)gf[33], fyeo = (float)gf[43];
    std::string name = od + "tqwk" + std::to_string(k) + ".dat";
    FILE* f = open_append(name);
    if (!f) return set_error(FGX_EIO, "Can't open %s", name.c_str());
    std::fprintf(f, "%d\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\n", timestep,
                 x * fyi - y * fxi, x * fyo - y * fxo, x * fyei - y * fxei, x * fyeo - y * fxeo, vx * fxi + vy * fyi,
                 vx * fxo + vy * fyo, vx * fxei + vy * fyei, vx * fxeo + vy * fyeo, (double)ptime);
    std::fclose(f);
    if (!P.self_gravity) {
      for (int kk = 0; kk < 11; ++kk) {
        std::string n2 = od + "tqwk" + std::to_string(k) + "_" + std::to_string(kk) + ".dat";
        FILE* g = open_append(n2);
        if (!g) return set_error(FGX_EIO, "Can't open %s", n2.c_str());
        // GlobalForce holds the float reductions as doubles
        const double a = (float)gf[kk], b = (float)gf[11 + kk], c = (float)gf[22 + kk], e = (float)gf[33 + kk];
        std::fprintf(g, "%d\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\n", timestep, x * c - y * a, x * e - y * b,
                     vx * a + vy * c, vx * b + vy * e, (double)ptime);
        std::fclose(g);
      }
    }
  }
  return FGX_OK;
}

int fgx_disk_write_setup_files(fgx_disk* d) {
  if (!d) return set_error(FGX_EPARAM, "null handle");
  const fgx::HostDisk* h = fgx::batch_host(d, 0);
  const fgx_params& P = h->P;
  const std::string od = outdir(P);
  // used_rad.dat (SourceEuler.cu:159-171): FP64 radii
  {
    std::string name = od + "used_rad.dat";
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) return set_error(FGX_EIO, "Can't write %s.", name.c_str());
    for (int i = 0; i <= h->nrad; ++i) std::fprintf(f, "%.18g\n", h->radii[i]);
    std::fclose(f);
  }
  // WriteDim (Output.cu:137-153)
  {
    std::string name = od + "dims.dat";
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) return set_error(FGX_EIO, "Unable to open %s.", name.c_str());
    std::fprintf(f, "%d\t%d\t\t%d\t%d\t%f\t%d\t%d\t%d\n", 0, 0, 0, 0, (double)P.rmax, P.ntot / P.ninterm, P.nrad,
                 P.nsec);
    std::fclose(f);
  }
  // EmptyPlanetSystemFile (Output.cu:12-32)
  for (int k = 0; k < h->sys.nb; ++k) {
    std::string name = od + "planet" + std::to_string(k) + ".dat";
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) return set_error(FGX_EIO, "Can't write %s file.", name.c_str());
    std::fclose(f);
  }
  return FGX_OK;
}

// -v: TellEverything (Interpret.cu:235-306).  The reference keeps each figure
// in a float before printing it, and counts orbits at unit radius with G = 1.
static void tell_everything(const fgx_params& P) {
  const double pi = 3.14159265358979323846;
  const double rmin = P.rmin, rmax = P.rmax, h = P.aspectratio, s0 = P.sigma0, sl = P.sigmaslope;
  auto orbits = [&](float t) { return (float)(t / 2.0 / pi); };
  auto outputs = [&](float t) { return (float)(t / P.dt / P.ninterm); };
  float t;
  std::printf("\nDisc properties:\n----------------\n");
  std::printf("Inner Radius          : %g\n", rmin);
  std::printf("Outer Radius          : %g\n", rmax);
  std::printf("Aspect Ratio          : %g\n", h);
  std::printf("VKep at inner edge    : %.3g\n", std::sqrt(1.0 / rmin));
  std::printf("VKep at outer edge    : %.3g\n", std::sqrt(1.0 / rmax));
  // disk masses of Sigma0 r^-slope between the edges and unit radius
  t = (float)(2.0 * pi * s0 / (2.0 - sl) * (std::pow(rmax, 2.0 - sl) - std::pow(rmin, 2.0 - sl)));
  std::printf("Initial Disk Mass             : %g\n", (double)t);
  t = (float)(2.0 * pi * s0 / (2.0 - sl) * (1.0 - std::pow(rmin, 2.0 - sl)));
  std::printf("Initial Mass inner to r=1.0  : %g \n", (double)t);
  t = (float)(2.0 * pi * s0 / (2.0 - sl) * (std::pow(rmax, 2.0 - sl) - 1.0));
  std::printf("Initial Mass outer to r=1.0  : %g \n", (double)t);
  // sound crossing times, the integral of dr / (h r^-1/2)
  std::printf("Travelling time for acoustic density waves :\n");
  const double spans[3][2] = {{rmin, rmax}, {1.0, rmax}, {rmin, 1.0}};
  const char* labels[3] = {" * From Rmin to Rmax  ", " * From r=1.0 to Rmax", " * From r=1.0 to Rmin"};
  for (int k = 0; k < 3; ++k) {
    t = (float)(2.0 / 3.0 / h * (std::pow(spans[k][1], 1.5) - std::pow(spans[k][0], 1.5)));
    std::printf("%s: %.2g = %.2f orbits ~ %.1f outputs\n", labels[k], (double)t, (double)orbits(t), (double)outputs(t));
  }
  // r^3 is a float product there
  t = (float)(2.0 * pi * std::sqrt((double)(P.rmin * P.rmin * P.rmin)));
  std::printf("Orbital time at Rmin  : %.3g ~ %.2f outputs\n", (double)t, (double)outputs(t));
  t = (float)(2.0 * pi * std::sqrt((double)(P.rmax * P.rmax * P.rmax)));
  std::printf("Orbital time at Rmax  : %.3g ~ %.2f outputs\n", (double)t, (double)outputs(t));
  std::printf("Sound speed :\n");
  std::printf(" * At unit radius     : %.3g\n", h);
  std::printf(" * At outer edge      : %.3g\n", h * std::sqrt(1.0 / rmax));
  std::printf(" * At inner edge      : %.3g\n", h * std::sqrt(1.0 / rmin));
  std::printf("\nGrid properties:\n----------------\n");
  std::printf("Number of rings       : %d\n", P.nrad);
  std::printf("Number of sectors     : %d\n", P.nsec);
  std::printf("Total cells           : %d\n", P.nrad * P.nsec);
  std::printf("\nOutputs properties:\n-------------------\n");
  const float span = (float)(P.ninterm * P.dt);
  std::printf("Time increment between outputs : %.3f = %.3f orbits\n", (double)span, (double)orbits(span));
  std::printf("At each output #i, the following files are written:\n");
  const int bytes = (int)(P.nrad * P.nsec * sizeof(float));
  std::printf("gasdens[i].dat : %d bytes\n", bytes);
  std::printf("gasvrad[i].dat : %d bytes\n", bytes);
  std::printf("gasvtheta[i].dat : %d bytes\n", bytes);
  if (P.adiabatic) std::printf("gasTemperature[i].dat : %d bytes\n", bytes);
  if (P.advect_label) std::printf("gaslabel[i].dat : %d bytes\n", bytes);
  std::printf("There will be in total %d outputs\n", P.ntot / P.ninterm);
  const float total = (float)(P.ntot * P.dt);
  std::printf("(which correspond to an elapsed time = %.3f or to %.2f orbits)\n", (double)total, (double)orbits(total));
  const float nfiles = 3.0f + (P.adiabatic ? 1.0f : 0.0f) + (P.advect_label ? 1.0f : 0.0f);
  t = nfiles * P.nrad * P.nsec * sizeof(float);
  t *= (float)P.ntot / (float)P.ninterm;
  t /= 1024.0 * 1024.0;
  std::printf("So the code will produce ~%.2f Mbytes of data\n", (double)t);
  std::printf("Check (eg by issuing a 'df' command) that you have enough disk space,\n");
  std::printf("otherwise you will get a system full and the code will stop.\n");
  std::fflush(stdout);
}

static int usage(const char* exe) {
  std::printf("Usage : %s [-abcdeimnptvz] [-(0-9)] [-s number] [-f scaling] parameters file\n", exe);
  return 1;
}

// main() of the reference (Main.cu:87-401) over the C ABI.  FGX_PRECISION=64
// in the environment selects the FP64 fields.
int fgx_run_main(int argc, char** argv) {
  if (argc == 1) return usage(argv[0]);
  std::string parfile;
  int disable = 0, sloppy = 0, cb = 0, store = 0, restart = 0, override_out = 0, nb_restart = 0, verbose = 0;
  float scaling = 1.0f;
  std::string newout;
  for (int i = 1; i < argc; ++i) {
    if (argv[i][0] == '-') {
      if (std::strspn(argv[i], "-secndovtpfamzib0123456789") != std::strlen(argv[i])) return usage(argv[0]);
      if (std::strchr(argv[i], 'n')) disable = 1;
      if (std::strchr(argv[i], 'v')) verbose = 1;
      if (std::strchr(argv[i], 'c')) sloppy = 1;
      if (std::strchr(argv[i], 'b')) cb = 1;
      if (std::strchr(argv[i], 'i')) store = 1;
      if (std::strchr(argv[i], 's')) {
        restart = 1;
        if (++i >= argc) return usage(argv[0]);
        nb_restart = std::atoi(argv[i]);
        if (nb_restart < 0) {
          std::printf("Incorrect restart number\n");
          return usage(argv[0]);
        }
      }
      if (std::strchr(argv[i], 'o')) {
        override_out = 1;
        if (++i >= argc) return usage(argv[0]);
        newout = argv[i];
      } else if (std::strchr(argv[i], 'f')) {
        if (++i >= argc) return usage(argv[0]);
        scaling = (float)std::atof(argv[i]);
        std::printf("Scaling factor = %g\n", scaling);
        if (scaling <= 0) return usage(argv[0]);
      }
    } else parfile = argv[i];
  }
  if (store && !restart) {
    std::printf("You cannot use tabulated surface density\nor surface internal energy in a non-restart run.\nAborted\n");
    return 0;
  }
  if (parfile.empty()) return usage(argv[0]);
  fgx_params P;
  int r = fgx_params_load(parfile.c_str(), &P);
  if (r) { std::fprintf(stderr, "%s\n", fgx_last_error()); return 1; }
  if (override_out) std::snprintf(P.outputdir, sizeof P.outputdir, "%s", newout.c_str());
  P.sloppy_cfl = sloppy; P.centrifugal_balance = cb; P.scaling_factor = scaling;
  if (verbose) tell_everything(P);
  if (disable) return 0;
  fgx_planets pl;
  if ((r = fgx_planets_load(P.planetconfig, &P, &pl))) { std::fprintf(stderr, "%s\n", fgx_last_error()); return 1; }
  std::printf("%d planet(s) found.\n", pl.nb);
  const char* prec = std::getenv("FGX_PRECISION");
  const int precision = (prec && std::strcmp(prec, "64") == 0) ? FGX_F64 : FGX_F32;
  fgx_disk* d = nullptr;
  if ((r = fgx_disk_create(&P, &pl, precision, &d))) { std::fprintf(stderr, "%s\n", fgx_last_error()); return 1; }
  // -s n: the state at output n, written by an earlier run with CHECKPOINT
  // YES; the loop resumes at that output's iteration, whose log lines that run
  // has already written (the reference parses -s but restores nothing,
  // Main.cu:147-155, 175-180)
  int first = 0;
  if (restart) {
    const std::string ck = std::string(P.outputdir) + "restart" + std::to_string(nb_restart) + ".bin";
    if ((r = fgx_batch_restore(d, 0, ck.c_str()))) {
      std::fprintf(stderr, "%s\n", fgx_last_error());
      fgx_disk_destroy(d);
      return 1;
    }
    first = nb_restart * P.ninterm;
    std::printf("Restarting from output %d (iteration %d)\n", nb_restart, first);
  } else if ((r = fgx_disk_write_setup_files(d))) {
    std::fprintf(stderr, "%s\n", fgx_last_error());
    fgx_disk_destroy(d);
    return 1;
  }
  auto t0 = std::chrono::steady_clock::now();
  int timestep = restart ? nb_restart : 0;
  for (int i = first; i <= P.ntot; ++i) {
    if (!(restart && i == first) && (r = fgx_disk_write_logs(d, timestep))) break;
    if (P.ninterm * (timestep = (i / P.ninterm)) == i) {
      std::printf("\n*** OUTPUT %d ***\n", timestep);
      if ((r = fgx_disk_write_outputs(d, timestep))) break;
      if (P.ext_checkpoint) {
        const std::string ck = std::string(P.outputdir) + "restart" + std::to_string(timestep) + ".bin";
        if ((r = fgx_batch_checkpoint(d, 0, ck.c_str()))) break;
      }
    }
    if ((r = fgx_disk_step(d, 1))) break;
  }
  if (r) { std::fprintf(stderr, "%s\n", fgx_last_error()); fgx_disk_destroy(d); return 1; }
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("%g\n", secs);
  fgx_disk_destroy(d);
  return 0;
}

}  // extern "C"
