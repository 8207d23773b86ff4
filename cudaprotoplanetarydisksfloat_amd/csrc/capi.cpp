// capi.cpp -- C ABI for errors, parameters and planets (include/fgx.h).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "fgx_internal.h"

namespace fgx {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}
const char* last_error() { return g_last_error.c_str(); }

}  // namespace fgx

using fgx::set_error;

extern "C" {

int fgx_version(void) { return FGX_VERSION; }
const char* fgx_last_error(void) { return fgx::last_error(); }
size_t fgx_sizeof_params(void) { return sizeof(fgx_params); }
size_t fgx_sizeof_planets(void) { return sizeof(fgx_planets); }

int fgx_params_default(fgx_params* p) {
  if (!p) return set_error(FGX_EPARAM, "null parameters");
  fgx::params_defaults(p);
  return FGX_OK;
}
int fgx_params_parse(const char* text, fgx_params* p) {
  if (!text || !p) return set_error(FGX_EPARAM, "null argument");
  return fgx::params_parse_text(text, p);
}
int fgx_params_load(const char* path, fgx_params* p) {
  if (!path || !p) return set_error(FGX_EPARAM, "null argument");
  std::string text;
  if (!fgx::read_file(path, text)) return set_error(FGX_EIO, "Unable to read %s .", path);
  return fgx::params_parse_text(text.c_str(), p);
}
int fgx_params_set(fgx_params* p, const char* key, const char* value) {
  if (!p || !key || !value) return set_error(FGX_EPARAM, "null argument");
  return fgx::params_set_key(p, key, value);
}
int fgx_params_get(const fgx_params* p, const char* key, char* out, size_t n) {
  if (!p || !key || !out) return set_error(FGX_EPARAM, "null argument");
  return fgx::params_get_key(p, key, out, n);
}

// CICPlanet placement needs the ring radii; they are rebuilt here exactly as
// the disk set-up builds them
static int planets_from_text(const char* text, const fgx_params* p, fgx_planets* out) {
  if (p->cic_planet) {
    fgx::HostDisk h;
    fgx_planets none;
    std::memset(&none, 0, sizeof none);
    int r = fgx::setup_grid_and_density(*p, none, h);
    if (r) return r;
    std::vector<double> radii(h.radii.size());
    for (size_t i = 0; i < radii.size(); ++i) radii[i] = (double)(float)h.radii[i];
    std::vector<double> rmed(h.rmed.size());
    for (size_t i = 0; i < rmed.size(); ++i) rmed[i] = (double)(float)h.rmed[i];
    return fgx::planets_parse_text(text, p, rmed.data(), radii.data(), out);
  }
  return fgx::planets_parse_text(text, p, nullptr, nullptr, out);
}
int fgx_planets_parse(const char* text, const fgx_params* p, fgx_planets* out) {
  if (!text || !p || !out) return set_error(FGX_EPARAM, "null argument");
  return planets_from_text(text, p, out);
}
int fgx_planets_load(const char* path, const fgx_params* p, fgx_planets* out) {
  if (!path || !p || !out) return set_error(FGX_EPARAM, "null argument");
  std::string text;
  if (!fgx::read_file(path, text)) return set_error(FGX_EIO, "Error : can't find '%s'.", path);
  return planets_from_text(text.c_str(), p, out);
}

}  // extern "C"
