// params.cpp -- .par and planet .cfg parsing with the reference's semantics.
//
// ReadVariables / SearchVariable (src/Interpret.cu:35-208): keys are the first
// token of a line, upper-cased; the value starts after any run of "\t :=>_"
// following the key; numbers are read with %lf and stored as float; string
// values are kept raw (case-sensitive).  A line whose first token starts with
// '#' is a comment.  Unknown keys are ignored.  The derived flags are only
// ever switched away from their defaults, as in src/Interpret.cu:126-202.
#include "fgx_internal.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace fgx {

static void set_str(char* dst, size_t n, const char* src) {
  std::snprintf(dst, n, "%s", src);
}

void params_defaults(fgx_params* p) {
  std::memset(p, 0, sizeof(*p));
  // include/Param.cuh:9-62
  p->dt = 1.0f; p->sigma0 = 173.0f; p->ninterm = 10; p->ntot = 1501;
  set_str(p->outputdir, sizeof p->outputdir, "~masset");
  set_str(p->innerboundary, sizeof p->innerboundary, "WALL");
  set_str(p->labeladvection, sizeof p->labeladvection, "NO");
  set_str(p->transport, sizeof p->transport, "FAST");
  set_str(p->planetconfig, sizeof p->planetconfig, "Systems/SolarSystem.cfg");
  p->masstaper = 0.0000001f;
  set_str(p->radialspacing, sizeof p->radialspacing, "ARITHMETIC");
  p->nrad = 64; p->nsec = 64; p->rmin = 1.0f; p->rmax = 1.0f;
  p->thicknesssmoothing = 0.0f; p->rochesmoothing = 0.0f; p->aspectratio = 0.05f;
  p->viscosity = 0.0f; p->alphaviscosity = 0.0f; p->sigmaslope = 0.0f;
  p->releaseradius = 0.0f; p->releasedate = 0.0f; p->omegaframe = 0.0f;
  set_str(p->disk, sizeof p->disk, "YES");
  set_str(p->frame, sizeof p->frame, "FIXED");
  set_str(p->outersourcemass, sizeof p->outersourcemass, "NO");
  set_str(p->writedensity, sizeof p->writedensity, "YES");
  set_str(p->writevelocity, sizeof p->writevelocity, "YES");
  set_str(p->writeenergy, sizeof p->writeenergy, "NO");
  set_str(p->writetemperature, sizeof p->writetemperature, "NO");
  set_str(p->writedivv, sizeof p->writedivv, "NO");
  set_str(p->writeqplus, sizeof p->writeqplus, "NO");
  set_str(p->indirectterm, sizeof p->indirectterm, "YES");
  set_str(p->excludehill, sizeof p->excludehill, "NO");
  p->imposeddiskdrift = 0.0f; p->flaringindex = 0.0f; p->eccentricity = 0.0f;
  p->cavityradius = 0.0f; p->cavityratio = 1.0f; p->cavitywidth = 1.0f;
  p->transitionradius = 0.0f; p->transitionratio = 1.0f; p->transitionwidth = 1.0f;
  p->lambdadoubling = 0.0f;
  set_str(p->selfgravity, sizeof p->selfgravity, "NO");
  set_str(p->cicplanet, sizeof p->cicplanet, "NO");
  set_str(p->forcedcircular, sizeof p->forcedcircular, "NO");
  set_str(p->zmplus, sizeof p->zmplus, "NO");
  set_str(p->adiabatic_s, sizeof p->adiabatic_s, "NO");
  p->adiabaticindex = 1.4f;
  p->ext_max_substeps = 100000;
  set_str(p->cooling_s, sizeof p->cooling_s, "NO");
  p->coolingtime0 = 6.28f;
  // src/Interpret.cu:23-31 flag defaults
  p->fast_transport = 1; p->indirect_term = 1; p->is_disk = 1;
  p->write_density = 1; p->write_velocity = 1;
  p->scaling_factor = 1.0f;
  // extensions
  p->ext_noise_seed = 20261203ULL;
}

// SearchVariable (src/Interpret.cu:35-95).  The numeric value is parsed from
// the value string; when that fails the reference reads an uninitialised
// double, here 0 is used.
static void search_variable(fgx_params* p, const char* name, const char* sv) {
  double t = 0.0;
  if (std::sscanf(sv, "%lf", &t) != 1) t = 0.0;
  float f = (float)t;
  int ti = (int)t;
  struct FKey { const char* k; float* v; };
  const FKey fkeys[] = {
      {"DT", &p->dt}, {"SIGMA0", &p->sigma0}, {"MASSTAPER", &p->masstaper},
      {"RMIN", &p->rmin}, {"RMAX", &p->rmax}, {"THICKNESSSMOOTHING", &p->thicknesssmoothing},
      {"ROCHESMOOTHING", &p->rochesmoothing}, {"ASPECTRATIO", &p->aspectratio},
      {"VISCOSITY", &p->viscosity}, {"ALPHAVISCOSITY", &p->alphaviscosity},
      {"SIGMASLOPE", &p->sigmaslope}, {"RELEASERADIUS", &p->releaseradius},
      {"RELEASEDATE", &p->releasedate}, {"OMEGAFRAME", &p->omegaframe},
      {"IMPOSEDDISKDRIFT", &p->imposeddiskdrift}, {"FLARINGINDEX", &p->flaringindex},
      {"ECCENTRICITY", &p->eccentricity}, {"CAVITYRADIUS", &p->cavityradius},
      {"CAVITYRATIO", &p->cavityratio}, {"CAVITYWIDTH", &p->cavitywidth},
      {"TRANSITIONRADIUS", &p->transitionradius}, {"TRANSITIONRATIO", &p->transitionratio},
      {"TRANSITIONWIDTH", &p->transitionwidth}, {"LAMBDADOUBLING", &p->lambdadoubling},
      {"ADIABATICINDEX", &p->adiabaticindex}, {"COOLINGTIME0", &p->coolingtime0}};
  for (const auto& k : fkeys)
    if (std::strcmp(name, k.k) == 0) *k.v = f;
  if (std::strcmp(name, "NINTERM") == 0) p->ninterm = ti;
  if (std::strcmp(name, "NTOT") == 0) p->ntot = ti;
  if (std::strcmp(name, "NRAD") == 0) p->nrad = ti;
  if (std::strcmp(name, "NSEC") == 0) p->nsec = ti;
  struct SKey { const char* k; char* v; size_t n; };
  const SKey skeys[] = {
      {"OUTPUTDIR", p->outputdir, sizeof p->outputdir},
      {"INNERBOUNDARY", p->innerboundary, sizeof p->innerboundary},
      {"LABELADVECTION", p->labeladvection, sizeof p->labeladvection},
      {"TRANSPORT", p->transport, sizeof p->transport},
      {"PLANETCONFIG", p->planetconfig, sizeof p->planetconfig},
      {"RADIALSPACING", p->radialspacing, sizeof p->radialspacing},
      {"DISK", p->disk, sizeof p->disk}, {"FRAME", p->frame, sizeof p->frame},
      {"OUTERSOURCEMASS", p->outersourcemass, sizeof p->outersourcemass},
      {"WRITEDENSITY", p->writedensity, sizeof p->writedensity},
      {"WRITEVELOCITY", p->writevelocity, sizeof p->writevelocity},
      {"WRITEENERGY", p->writeenergy, sizeof p->writeenergy},
      {"WRITETEMPERATURE", p->writetemperature, sizeof p->writetemperature},
      {"WRITEQPLUS", p->writeqplus, sizeof p->writeqplus},
      {"INDIRECTTERM", p->indirectterm, sizeof p->indirectterm},
      {"EXCLUDEHILL", p->excludehill, sizeof p->excludehill},
      {"SELFGRAVITY", p->selfgravity, sizeof p->selfgravity},
      {"CICPLANET", p->cicplanet, sizeof p->cicplanet},
      {"FORCEDCIRCULAR", p->forcedcircular, sizeof p->forcedcircular},
      {"ZMPLUS", p->zmplus, sizeof p->zmplus},
      {"ADIABATIC", p->adiabatic_s, sizeof p->adiabatic_s},
      {"COOLING", p->cooling_s, sizeof p->cooling_s}};
  for (const auto& k : skeys)
    if (std::strcmp(name, k.k) == 0) set_str(k.v, k.n, sv);
  // WRITEDIVV = temp assigns a one-character string (src/Interpret.cu:73),
  // so the file can never turn the divergence output on.
  if (std::strcmp(name, "WRITEDIVV") == 0) {
    p->writedivv[0] = (char)(int)t;
    p->writedivv[1] = 0;
  }
  // keys the reference does not know
  if (std::strcmp(name, "NOISEAMPLITUDE") == 0) p->ext_noise_amplitude = t;
  if (std::strcmp(name, "NOISESEED") == 0) p->ext_noise_seed = std::strtoull(sv, nullptr, 10);
  if (std::strcmp(name, "DISKID") == 0) p->ext_disk_id = ti;
  if (std::strcmp(name, "SUBSTEPS") == 0) p->ext_substeps = ti;
  if (std::strcmp(name, "QUIRKFIXTAURP") == 0) p->ext_fix_taurp_corner = ti;
  if (std::strcmp(name, "QUIRKFIXPPDIST") == 0) p->ext_fix_pp_distance = ti;
  if (std::strcmp(name, "QUIRKFIXSGWRAP") == 0) p->ext_fix_sg_wrap = ti;
  if (std::strcmp(name, "SGSOFTENING") == 0) p->ext_sg_softening = f;
  if (std::strcmp(name, "GASNAMES") == 0) p->ext_gas_names = ti;
  if (std::strcmp(name, "CHECKPOINT") == 0) p->ext_checkpoint = ti != 0 || std::strcmp(sv, "YES") == 0;
  if (std::strcmp(name, "STUBBEDPHYSICS") == 0) p->ext_stubbed_physics = ti != 0 || std::strcmp(sv, "YES") == 0;
  if (std::strcmp(name, "MAXSUBSTEPS") == 0) p->ext_max_substeps = ti;
  // the executable's switches -c, -b, -f (Main.cu:122-170) as keys
  if (std::strcmp(name, "SLOPPYCFL") == 0) p->sloppy_cfl = ti;
  if (std::strcmp(name, "CENTRIFUGALBALANCE") == 0) p->centrifugal_balance = ti;
  if (std::strcmp(name, "SCALINGFACTOR") == 0) p->scaling_factor = f;
}

// derived flags and checks (src/Interpret.cu:126-201)
static int derive_flags(fgx_params* p) {
  auto eq = [](const char* a, const char* b) { return std::strcmp(a, b) == 0; };
  // flags are derived once from the final key values, starting from their
  // defaults (src/Interpret.cu:23-31), so a later key can undo an earlier one
  p->log_grid = p->open_inner = p->nonreflecting = p->evanescent = p->corotating = 0;
  p->guiding_center = p->self_gravity = p->sg_zero_mode = p->zmplus_f = p->adiabatic = 0;
  p->cooling = p->exclude_hill = p->cic_planet = p->forced_circular = p->viscosity_alpha = 0;
  p->roche_smoothing = p->advect_label = p->outer_source_mass = p->write_energy = 0;
  p->write_temperature = p->write_divv = p->write_qplus = 0;
  p->fast_transport = p->indirect_term = p->is_disk = p->write_density = p->write_velocity = 1;
  if (eq(p->labeladvection, "YES")) p->advect_label = 1;
  if (eq(p->outersourcemass, "YES")) p->outer_source_mass = 1;
  if (eq(p->transport, "STANDARD")) p->fast_transport = 0;
  if (eq(p->innerboundary, "OPEN")) p->open_inner = 1;
  if (eq(p->innerboundary, "NONREFLECTING")) p->nonreflecting = 1;
  if (eq(p->innerboundary, "EVANESCENT")) p->evanescent = 1;
  if (eq(p->radialspacing, "LOGARITHM")) p->log_grid = 1;
  if (eq(p->disk, "NO")) p->is_disk = 0;
  if (eq(p->frame, "COROTATING")) p->corotating = 1;
  if (eq(p->frame, "GUIDING-CENTER")) { p->guiding_center = 1; p->corotating = 1; }
  if (eq(p->writevelocity, "NO")) p->write_velocity = 0;
  if (eq(p->writedensity, "NO")) p->write_density = 0;
  if (eq(p->writeenergy, "YES")) p->write_energy = 1;
  if (eq(p->writetemperature, "YES")) p->write_temperature = 1;
  if (eq(p->writedivv, "YES")) p->write_divv = 1;
  if (eq(p->writeqplus, "YES")) p->write_qplus = 1;
  if (eq(p->indirectterm, "NO")) p->indirect_term = 0;
  if (eq(p->selfgravity, "YES")) p->self_gravity = 1;
  if (eq(p->selfgravity, "Z")) { p->self_gravity = 1; p->sg_zero_mode = 1; }
  if (eq(p->zmplus, "YES")) p->zmplus_f = 1;
  if (p->zmplus_f && !p->sg_zero_mode) p->zmplus_f = 0;
  if (eq(p->adiabatic_s, "YES")) { p->adiabatic = 1; p->write_temperature = 1; }
  if (eq(p->cooling_s, "YES")) p->cooling = 1;
  if (p->adiabatic && p->adiabaticindex == 1.0f) p->adiabatic = 0;
  if (eq(p->writeenergy, "NO")) p->write_energy = 0;
  if (eq(p->excludehill, "YES")) p->exclude_hill = 1;
  if (eq(p->cicplanet, "YES")) p->cic_planet = 1;
  if (eq(p->forcedcircular, "YES")) p->forced_circular = 1;
  if (p->alphaviscosity != 0.0f && p->viscosity != 0.0f)
    return set_error(FGX_EPARAM, "You cannot use at the same time VISCOSITY and ALPHAVISCOSITY.");
  if (p->alphaviscosity != 0.0f) p->viscosity_alpha = 1;
  if (p->thicknesssmoothing != 0.0f && p->rochesmoothing != 0.0f)
    return set_error(FGX_EPARAM, "You cannot use at the same time ThicknessSmoothing and RocheSmoothing.");
  if (p->thicknesssmoothing <= 0.0f && p->rochesmoothing <= 0.0f)
    return set_error(FGX_EPARAM, "A non-vanishing potential smoothing length is required.");
  if (p->rochesmoothing != 0.0f) p->roche_smoothing = 1;
  return FGX_OK;
}

// the flags that choose a batch's launch sequence, and the batch-wide
// switches: every disk of a batch shares them
bool batch_modes_equal(const fgx_params& p, const fgx_params& P0) {
  return !(p.adiabatic != P0.adiabatic || p.open_inner != P0.open_inner || p.nonreflecting != P0.nonreflecting ||
           p.evanescent != P0.evanescent || p.fast_transport != P0.fast_transport ||
           p.self_gravity != P0.self_gravity || p.cooling != P0.cooling || p.indirect_term != P0.indirect_term ||
           p.exclude_hill != P0.exclude_hill || p.sloppy_cfl != P0.sloppy_cfl || p.corotating != P0.corotating ||
           p.guiding_center != P0.guiding_center || p.forced_circular != P0.forced_circular ||
           p.roche_smoothing != P0.roche_smoothing || p.viscosity_alpha != P0.viscosity_alpha ||
           p.ext_fix_taurp_corner != P0.ext_fix_taurp_corner || p.ext_fix_pp_distance != P0.ext_fix_pp_distance ||
           p.is_disk != P0.is_disk || p.ext_fix_sg_wrap != P0.ext_fix_sg_wrap ||
           (p.advect_label && p.ext_stubbed_physics) != (P0.advect_label && P0.ext_stubbed_physics));
}

int params_validate(const fgx_params* p) {
  if (p->nrad < 4 || p->nsec < 4)
    return set_error(FGX_EPARAM, "NRAD and NSEC must be at least 4 (got %d x %d)", p->nrad, p->nsec);
  if ((long long)p->nrad * p->nsec > (1LL << 31) / 2)
    return set_error(FGX_EPARAM, "grid %d x %d too large", p->nrad, p->nsec);
  if (!(p->dt > 0.0f) || !std::isfinite(p->dt)) return set_error(FGX_EPARAM, "DT must be positive");
  if (!(p->rmax > p->rmin) || !(p->rmin > 0.0f))
    return set_error(FGX_EPARAM, "need 0 < RMIN < RMAX (got %g, %g)", p->rmin, p->rmax);
  if (p->self_gravity && !p->log_grid)
    return set_error(FGX_EPARAM, "A logarithmic grid is needed to compute self-gravity with polar method.");
  return FGX_OK;
}

// ReadVariables line loop (src/Interpret.cu:111-123)
int params_parse_text(const char* text, fgx_params* p) {
  params_defaults(p);
  char line[350], nm[300], sval[290];
  sval[0] = 0;
  const char* cur = text;
  while (*cur) {
    // fgets(s, 349, ...) semantics: up to 348 chars or through '\n'
    size_t n = 0;
    while (cur[n] && cur[n] != '\n' && n < 348) n++;
    if (cur[n] == '\n' && n < 348) n++;
    std::memcpy(line, cur, n);
    line[n] = 0;
    cur += n;
    if (std::sscanf(line, "%299s ", nm) == 1 && nm[0] != '#') {
      const char* s1 = line + std::strlen(nm);
      s1 += std::strspn(s1, "\t :=>_");
      char tmp[290];
      if (std::sscanf(s1, "%289s ", tmp) == 1) std::snprintf(sval, sizeof sval, "%s", tmp);
      for (char* c = nm; *c; ++c) *c = (char)std::toupper((unsigned char)*c);
      search_variable(p, nm, sval);
    }
  }
  return derive_flags(p);
}

int params_set_key(fgx_params* p, const char* key, const char* value) {
  char nm[300];
  std::snprintf(nm, sizeof nm, "%s", key);
  for (char* c = nm; *c; ++c) *c = (char)std::toupper((unsigned char)*c);
  search_variable(p, nm, value);
  return derive_flags(p);
}

int params_get_key(const fgx_params* p, const char* key, char* out, size_t n) {
  std::string k(key);
  for (auto& c : k) c = (char)std::toupper((unsigned char)c);
  struct F { const char* k; double v; };
  const F nums[] = {
      {"NRAD", (double)p->nrad}, {"NSEC", (double)p->nsec}, {"NINTERM", (double)p->ninterm},
      {"NTOT", (double)p->ntot}, {"DT", p->dt}, {"SIGMA0", p->sigma0}, {"MASSTAPER", p->masstaper},
      {"RMIN", p->rmin}, {"RMAX", p->rmax}, {"THICKNESSSMOOTHING", p->thicknesssmoothing},
      {"ROCHESMOOTHING", p->rochesmoothing}, {"ASPECTRATIO", p->aspectratio},
      {"VISCOSITY", p->viscosity}, {"ALPHAVISCOSITY", p->alphaviscosity},
      {"SIGMASLOPE", p->sigmaslope}, {"RELEASERADIUS", p->releaseradius},
      {"RELEASEDATE", p->releasedate}, {"OMEGAFRAME", p->omegaframe},
      {"IMPOSEDDISKDRIFT", p->imposeddiskdrift}, {"FLARINGINDEX", p->flaringindex},
      {"ECCENTRICITY", p->eccentricity}, {"CAVITYRADIUS", p->cavityradius},
      {"CAVITYRATIO", p->cavityratio}, {"CAVITYWIDTH", p->cavitywidth},
      {"TRANSITIONRADIUS", p->transitionradius}, {"TRANSITIONRATIO", p->transitionratio},
      {"TRANSITIONWIDTH", p->transitionwidth}, {"LAMBDADOUBLING", p->lambdadoubling},
      {"ADIABATICINDEX", p->adiabaticindex}, {"COOLINGTIME0", p->coolingtime0},
      {"NOISEAMPLITUDE", p->ext_noise_amplitude}, {"NOISESEED", (double)p->ext_noise_seed},
      {"DISKID", (double)p->ext_disk_id}, {"SUBSTEPS", (double)p->ext_substeps},
      {"SGSOFTENING", p->ext_sg_softening}, {"GASNAMES", (double)p->ext_gas_names},
      {"CHECKPOINT", (double)p->ext_checkpoint}, {"STUBBEDPHYSICS", (double)p->ext_stubbed_physics},
      {"MAXSUBSTEPS", (double)p->ext_max_substeps},
      // derived flags
      {"LOGGRID", (double)p->log_grid}, {"FASTTRANSPORT", (double)p->fast_transport},
      {"OPENINNER", (double)p->open_inner}, {"NONREFLECTING", (double)p->nonreflecting},
      {"EVANESCENT", (double)p->evanescent}, {"ISDISK", (double)p->is_disk},
      {"COROTATING", (double)p->corotating}, {"GUIDINGCENTER", (double)p->guiding_center},
      {"INDIRECT_TERM", (double)p->indirect_term}, {"SELFGRAVITY_F", (double)p->self_gravity},
      {"SGZEROMODE", (double)p->sg_zero_mode}, {"ADIABATIC_F", (double)p->adiabatic},
      {"COOLING_F", (double)p->cooling}, {"EXCLUDEHILL_F", (double)p->exclude_hill},
      {"CICPLANET_F", (double)p->cic_planet}, {"FORCEDCIRCULAR_F", (double)p->forced_circular},
      {"VISCOSITYALPHA", (double)p->viscosity_alpha}, {"ROCHESMOOTHING_F", (double)p->roche_smoothing},
      {"ADVECTLABEL", (double)p->advect_label}, {"WRITE_DENSITY", (double)p->write_density},
      {"WRITE_VELOCITY", (double)p->write_velocity}, {"WRITE_ENERGY", (double)p->write_energy},
      {"WRITE_TEMPERATURE", (double)p->write_temperature}, {"WRITE_DIVV", (double)p->write_divv},
      {"WRITE_QPLUS", (double)p->write_qplus}, {"SLOPPYCFL", (double)p->sloppy_cfl},
      {"CENTRIFUGALBALANCE", (double)p->centrifugal_balance}, {"SCALINGFACTOR", p->scaling_factor},
      {"QUIRKFIXTAURP", (double)p->ext_fix_taurp_corner}, {"QUIRKFIXPPDIST", (double)p->ext_fix_pp_distance},
      {"QUIRKFIXSGWRAP", (double)p->ext_fix_sg_wrap}};
  for (const auto& f : nums)
    if (k == f.k) { std::snprintf(out, n, "%.17g", f.v); return FGX_OK; }
  struct S { const char* k; const char* v; };
  const S strs[] = {{"OUTPUTDIR", p->outputdir}, {"INNERBOUNDARY", p->innerboundary},
                    {"LABELADVECTION", p->labeladvection}, {"TRANSPORT", p->transport},
                    {"PLANETCONFIG", p->planetconfig}, {"RADIALSPACING", p->radialspacing},
                    {"DISK", p->disk}, {"FRAME", p->frame}, {"OUTERSOURCEMASS", p->outersourcemass},
                    {"WRITEDENSITY", p->writedensity}, {"WRITEVELOCITY", p->writevelocity},
                    {"WRITEENERGY", p->writeenergy}, {"WRITETEMPERATURE", p->writetemperature},
                    {"WRITEQPLUS", p->writeqplus}, {"INDIRECTTERM", p->indirectterm},
                    {"EXCLUDEHILL", p->excludehill}, {"SELFGRAVITY", p->selfgravity},
                    {"CICPLANET", p->cicplanet}, {"FORCEDCIRCULAR", p->forcedcircular},
                    {"ZMPLUS", p->zmplus}, {"ADIABATIC", p->adiabatic_s}, {"COOLING", p->cooling_s}};
  for (const auto& s : strs)
    if (k == s.k) { std::snprintf(out, n, "%s", s.v); return FGX_OK; }
  return set_error(FGX_EPARAM, "unknown key '%s'", key);
}

// InitPlanetarySystem / FindNumberOfPlanets (src/Psys.cu:9-134).  CICPlanet
// needs the ring radii, passed in as rmed (nrad) and radii (nrad+1).
int planets_parse_text(const char* text, const fgx_params* p, const double* rmed,
                       const double* radii, fgx_planets* out) {
  std::memset(out, 0, sizeof(*out));
  char s[512], nm[512], t1[512] = "", t2[512] = "";
  const char* cur = text;
  int i = 0;
  while (*cur) {
    size_t n = 0;
    while (cur[n] && cur[n] != '\n' && n < 509) n++;
    if (cur[n] == '\n' && n < 509) n++;
    std::memcpy(s, cur, n);
    s[n] = 0;
    cur += n;
    if (std::sscanf(s, "%511s ", nm) != 1) nm[0] = 0;
    if (!std::isalpha((unsigned char)s[0])) continue;
    if (i >= FGX_MAX_PLANETS)
      return set_error(FGX_EPARAM, "more than %d planets in the configuration", FGX_MAX_PLANETS);
    const char* s1 = s + std::strlen(nm);
    double dist = 0, mass = 0, accret = 0;
    std::sscanf(s1 + std::strspn(s1, "\t :=>_"), "%lf %lf %lf %511s %511s", &dist, &mass, &accret, t1, t2);
    if (p->cic_planet && rmed && radii) {
      int j = 0;
      while (j < p->nrad - 1 && rmed[j] < dist) j++;
      dist = radii[j + 1];
    }
    double e = p->eccentricity;
    out->mass[i] = mass;
    int fd = 1, fo = 1;
    if (std::tolower((unsigned char)t1[0]) == 'n') fd = 0;
    if (std::tolower((unsigned char)t2[0]) == 'n') fo = 0;
    out->x[i] = dist * (1.0 + e);
    out->y[i] = 0.0;
    out->vy[i] = std::sqrt(1.0 * (1.0 + mass) / dist) * std::sqrt((1.0 - e) / (1.0 + e));
    out->vx[i] = -0.0000000001 * out->vy[i];
    out->acc[i] = accret;
    out->feel_disk[i] = fd;
    out->feel_others[i] = fo;
    std::snprintf(out->name[i], sizeof out->name[i], "%s", nm);
    i++;
  }
  out->nb = i;
  return FGX_OK;
}

bool read_file(const char* path, std::string& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  char buf[65536];
  size_t n;
  out.clear();
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
  std::fclose(f);
  return true;
}

}  // namespace fgx
