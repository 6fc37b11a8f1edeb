// Host-only parts of the C ABI (see host_abi.h).
#include "host_abi.h"

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <fstream>
#include <initializer_list>

namespace smaiso {

namespace {

thread_local std::string g_error = "no error";

std::string fmt(const char* f, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof(buf), f, ap);
  va_end(ap);
  return buf;
}

bool finite_all(std::initializer_list<double> v) {
  for (double x : v)
    if (!std::isfinite(x)) return false;
  return true;
}

const char kCkptMagic[8] = {'S', 'M', 'A', 'I', 'S', 'O', 'C', 'K'};

}  // namespace

int abi_fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

const char* abi_error() { return g_error.c_str(); }

int validate_params(const smaiso_params* p, std::string* why) {
  if (!p) { *why = "params is NULL"; return SMAISO_E_INVALID; }
  if (p->nx < 2 || p->ny < 2 || p->nz < 2) {
    *why = fmt("grid %dx%dx%d: every extent must be >= 2", p->nx, p->ny, p->nz);
    return SMAISO_E_INVALID;
  }
  if ((long long)p->nx * p->ny * p->nz > 2147483647LL / 2) {
    *why = "grid too large (Nx*Ny*Nz must stay below 2^30)";
    return SMAISO_E_INVALID;
  }
  if (!finite_all({p->dx, p->dy, p->dz, p->dt, p->alpha, p->beta, p->gamma, p->epsilon, p->q0,
                   p->nu, p->lambda, p->zeta, p->rho0, p->rho_m})) {
    *why = "non-finite parameter";
    return SMAISO_E_INVALID;
  }
  if (p->dx <= 0 || p->dy <= 0 || p->dz <= 0) { *why = "dx, dy, dz must be > 0"; return SMAISO_E_INVALID; }
  if (p->dt <= 0) { *why = "dt must be > 0"; return SMAISO_E_INVALID; }
  if (p->nu <= 0) { *why = "nu must be > 0"; return SMAISO_E_INVALID; }
  if (p->lambda + 2 * p->nu <= 0) { *why = "lambda + 2 nu must be > 0"; return SMAISO_E_INVALID; }
  const double disc = 9 * p->beta * p->beta + 40 * p->epsilon * p->gamma;
  if (disc < 0 || (3 * p->beta + std::sqrt(disc)) <= 0) {
    *why = "Amp = sqrt((3 beta + sqrt(9 beta^2 + 40 eps gamma))/5) is not a positive real";
    return SMAISO_E_INVALID;
  }
  if (p->precision != SMAISO_FP64 && p->precision != SMAISO_FP32_MIXED) {
    *why = fmt("unknown precision %d", p->precision);
    return SMAISO_E_INVALID;
  }
  if (p->divv_switch < 0) { *why = "divv_switch must be >= 0"; return SMAISO_E_INVALID; }
  if (p->slab_n0 < 0) { *why = "slab_n0 must be >= 0"; return SMAISO_E_INVALID; }
  return SMAISO_OK;
}

void ckpt_header_init(CkptHeader* h, int nx, int ny, int nz, int precision, int elem, int64_t nloop) {
  std::memset(h, 0, sizeof(*h));
  std::memcpy(h->magic, kCkptMagic, 8);
  h->version = kCkptVersion;
  h->nx = nx;
  h->ny = ny;
  h->nz = nz;
  h->precision = precision;
  h->elem = elem;
  h->nloop = nloop;
}

int ckpt_header_read(std::FILE* f, int nx, int ny, int nz, int precision, int elem, CkptHeader* out,
                     std::string* why) {
  CkptHeader hd{};
  if (std::fread(&hd, sizeof(hd), 1, f) != 1 || std::memcmp(hd.magic, kCkptMagic, 8) != 0) {
    *why = "not a smaiso checkpoint";
    return SMAISO_E_INVALID;
  }
  if (hd.version != kCkptVersion) {
    *why = fmt("checkpoint version %d (this library reads %d)", hd.version, kCkptVersion);
    return SMAISO_E_INVALID;
  }
  if (hd.nloop < 0) {
    *why = fmt("negative step count %lld", (long long)hd.nloop);
    return SMAISO_E_INVALID;
  }
  if (hd.nx != nx || hd.ny != ny || hd.nz != nz || hd.precision != precision || hd.elem != elem) {
    *why = fmt("checkpoint %dx%dx%d precision %d does not match the handle (%dx%dx%d precision %d)", hd.nx, hd.ny,
               hd.nz, hd.precision, nx, ny, nz, precision);
    return SMAISO_E_INVALID;
  }
  *out = hd;
  return SMAISO_OK;
}

}  // namespace smaiso

using smaiso::abi_fail;
using smaiso::fmt;

extern "C" {

int smaiso_abi_version(void) { return SMAISO_ABI_VERSION; }

void smaiso_params_default(smaiso_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->nx = p->ny = p->nz = 512;                     // S:218
  const double nw = 8, q0 = 1.0;                   // S:261, S:244
  p->dx = p->dy = p->dz = 2.0 * M_PI / (q0 * nw);  // L/N, S:267-272
  p->dt = 0.001;                                   // S:262
  p->gamma = 1.0; p->beta = 2.0; p->alpha = 1.0;   // S:239-241
  p->epsilon = -0.001 * 5;                         // ep_arg = 5, S:242-243
  p->q0 = q0;
  p->nu = 1.0;                                     // argv[1]
  p->lambda = 1.0; p->zeta = 1.0;                  // S:256-257
  p->rho0 = 0.01;                                  // S:251
  p->rho_m = 0.0;                                  // -> rho_s/2, S:255
  p->divv_switch = 50;                             // S:214
  p->load_mode = 0;
  p->precision = SMAISO_FP64;
  p->slab_n0 = 0;
}

int smaiso_make_interface_ic(const smaiso_params* p, const smaiso_ic* ic, const double* height,
                             double* psi, double* rho) {
  if (!p || !ic || !psi || !rho) return abi_fail(SMAISO_E_INVALID, "NULL argument");
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return abi_fail(SMAISO_E_INVALID, "bad grid");
  if (!(ic->width > 0)) return abi_fail(SMAISO_E_INVALID, "width must be > 0");
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const double lx = nx * p->dx, ly = ny * p->dy, lz = nz * p->dz;
  const double h0 = ic->h0 < 0 ? lz / 2 : ic->h0;
  const uint64_t gam = 0x9E3779B97F4A7C15ULL;
  uint64_t n = 0;
  for (int i = 0; i < nx; ++i) {
    const double x = i * p->dx;
    for (int j = 0; j < ny; ++j) {
      const double y = j * p->dy;
      const double h = height ? height[(size_t)i * ny + j]
                              : h0 + ic->undulation * std::cos(2 * M_PI * x / lx) * std::cos(2 * M_PI * y / ly);
      for (int k = 0; k < nz; ++k) {
        const double z = k * p->dz;
        const double prof = 0.5 * (1.0 - std::tanh((z - h) / ic->width));
        // splitmix64 (Steele, Lea, Flood 2014), element n uses seed + (n+1) gamma
        ++n;
        uint64_t zz = ic->seed + n * gam;
        zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ULL;
        zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBULL;
        zz = zz ^ (zz >> 31);
        const double u = (double)(zz >> 11) * (1.0 / 9007199254740992.0);
        const size_t idx = ((size_t)i * ny + j) * nz + k;
        psi[idx] = ic->amplitude * std::cos(p->q0 * z) * prof + ic->noise * (2.0 * u - 1.0);
        rho[idx] = ic->rho_iso + (ic->rho_sm - ic->rho_iso) * prof;
      }
    }
  }
  return SMAISO_OK;
}

int smaiso_interface_ic_vjp(const smaiso_params* p, const smaiso_ic* ic, const double* height,
                            const double* dpsi, const double* drho, double* dheight) {
  if (!p || !ic || !dpsi || !drho || !dheight) return abi_fail(SMAISO_E_INVALID, "NULL argument");
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return abi_fail(SMAISO_E_INVALID, "bad grid");
  if (!(ic->width > 0)) return abi_fail(SMAISO_E_INVALID, "width must be > 0");
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const double lx = nx * p->dx, ly = ny * p->dy, lz = nz * p->dz;
  const double h0 = ic->h0 < 0 ? lz / 2 : ic->h0;
  for (int i = 0; i < nx; ++i) {
    for (int j = 0; j < ny; ++j) {
      const double hh = height ? height[(size_t)i * ny + j]
                               : h0 + ic->undulation * std::cos(2 * M_PI * i * p->dx / lx) *
                                          std::cos(2 * M_PI * j * p->dy / ly);
      double acc = 0.0;
      for (int k = 0; k < nz; ++k) {
        const double z = k * p->dz;
        const double ch = std::cosh((z - hh) / ic->width);
        const double dprof = 0.5 / (ic->width * ch * ch);
        const size_t idx = ((size_t)i * ny + j) * nz + k;
        acc += (ic->amplitude * std::cos(p->q0 * z) * dpsi[idx] + (ic->rho_sm - ic->rho_iso) * drho[idx]) * dprof;
      }
      dheight[(size_t)i * ny + j] = acc;
    }
  }
  return SMAISO_OK;
}

int smaiso_read_psi_text(const char* const* paths, int32_t npaths, int64_t count, double* psi) {
  if (!paths || npaths < 1 || count < 0 || !psi) return abi_fail(SMAISO_E_INVALID, "bad argument");
  int64_t got = 0;
  for (int32_t r = 0; r < npaths && got < count; ++r) {
    if (!paths[r]) return abi_fail(SMAISO_E_INVALID, "path is NULL");
    std::ifstream in(paths[r]);
    if (!in.is_open()) return abi_fail(SMAISO_E_INVALID, fmt("cannot open %s", paths[r]));
    double v;
    while (got < count && (in >> v)) psi[got++] = v;  // S:739-748 reads with operator>>
  }
  if (got < count)
    return abi_fail(SMAISO_E_INVALID, fmt("the files hold %lld values, %lld needed", (long long)got, (long long)count));
  return SMAISO_OK;
}

}  // extern "C"
