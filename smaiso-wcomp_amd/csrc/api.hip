// C ABI (include/smaiso.h): handle, device buffers and the step schedule.
//
// One call of run_step<T>() is one iteration of the reference's loop body
// (smaiso-wcomp.cpp:1081-2198) on one GPU.  The schedule follows the
// reference's stage order; stages are kernels on the handle's stream.
#include "../../include/smaiso.h"
#include "common.h"
#include "r2r.h"
#include "stepper.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

using namespace smaiso;

namespace {

thread_local std::string g_create_error = "no error";

std::string fmt(const char* f, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof(buf), f, ap);
  va_end(ap);
  return buf;
}

const int kCT[3] = {REDFT10, REDFT10, REDFT10};
const int kICT[3] = {REDFT01, REDFT01, REDFT01};
const int kST[3][3] = {{RODFT10, REDFT10, REDFT10}, {REDFT10, RODFT10, REDFT10},
                       {REDFT10, REDFT10, RODFT10}};
const int kIST[3][3] = {{RODFT01, REDFT01, REDFT01}, {REDFT01, RODFT01, REDFT01},
                        {REDFT01, REDFT01, RODFT01}};

struct PendingTiming {
  int kc;
  hipEvent_t a, b;
  double bytes;
};

}  // namespace

struct smaiso_handle {
  smaiso_params p{};
  Geom g{};
  int device = 0;
  hipStream_t stream = nullptr;
  R2R r2r;
  std::string err = "no error";
  bool have_state = false;
  long long nloop = 0;
  bool mixed = false;
  size_t tsz = 8;  // bytes per work element

  // derived (S:250-255, S:278)
  double amp_eq = 0, kp = 0, rho_s = 0, rho_m = 0, scale = 0;

  // device memory
  std::vector<void*> allocs;
  double* vq[3] = {nullptr, nullptr, nullptr};
  double* vs[3] = {nullptr, nullptr, nullptr};
  // FP64 state
  double *psiq = nullptr, *psiq_old = nullptr, *rho = nullptr, *f_rho = nullptr;
  double *partials = nullptr, *diag_out = nullptr, *out64 = nullptr;
  // work fields (T = double or float)
  void *psi = nullptr, *psi_prev = nullptr, *nlq = nullptr, *nlq_old = nullptr;
  double* rhoq = nullptr;  // DCT of rho, FP64 in both precisions (rho is ~constant: its
                           // gradient is a small difference of large coefficients)
  void *gp[3] = {}, *rd[3] = {}, *S[3] = {}, *f[3] = {}, *vsolq[3] = {}, *virrq[3] = {}, *v[3] = {};
  void *amp = nullptr, *lpsi = nullptr, *laprl = nullptr, *mu = nullptr, *fgrad = nullptr,
       *erho = nullptr, *divv = nullptr, *divf = nullptr, *nl = nullptr, *tmp = nullptr;

  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev_free;
  std::vector<PendingTiming> pending;
  double kc_ms[SMAISO_KC_COUNT] = {};
  long long kc_n[SMAISO_KC_COUNT] = {};
  double kc_bytes[SMAISO_KC_COUNT] = {};

  ~smaiso_handle() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& pt : pending) { ev_free.push_back(pt.a); ev_free.push_back(pt.b); }
    for (auto e : ev_free) (void)hipEventDestroy(e);
    for (void* a : allocs) (void)hipFree(a);
    if (stream) (void)hipStreamDestroy(stream);
  }

  StepConsts consts() const {
    StepConsts c{};
    c.nx = g.nx; c.ny = g.ny; c.nz = g.nz;
    c.slab_n0 = p.slab_n0;
    c.scale = scale;
    c.dt = p.dt; c.alpha = p.alpha; c.beta = p.beta; c.gamma = p.gamma; c.eps = p.epsilon;
    c.q0 = p.q0; c.q02 = p.q0 * p.q0; c.nu = p.nu; c.lam = p.lambda; c.zeta = p.zeta;
    c.rho0 = p.rho0; c.kp = kp; c.amp_eq = amp_eq; c.rho_m = rho_m;
    for (int a = 0; a < 3; ++a) { c.vq[a] = vq[a]; c.vs[a] = vs[a]; }
    return c;
  }

  long long n3() const { return g.n3(); }

  // --- timing helpers ----------------------------------------------------
  hipEvent_t take_event() {
    if (!ev_free.empty()) { hipEvent_t e = ev_free.back(); ev_free.pop_back(); return e; }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  struct Scope {
    smaiso_handle* h;
    int kc;
    double bytes;
    hipEvent_t a = nullptr;
    Scope(smaiso_handle* hh, int k, double b) : h(hh), kc(k), bytes(b) {
      if (h->timing) { a = h->take_event(); (void)hipEventRecord(a, h->stream); }
    }
    ~Scope() {
      if (h->timing) {
        hipEvent_t b = h->take_event();
        (void)hipEventRecord(b, h->stream);
        h->pending.push_back({kc, a, b, bytes});
      }
    }
  };
  void collect_timing() {
    for (auto& pt : pending) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, pt.a, pt.b) == hipSuccess) {
        kc_ms[pt.kc] += ms;
        kc_n[pt.kc] += 1;
        kc_bytes[pt.kc] += pt.bytes;
      }
      ev_free.push_back(pt.a);
      ev_free.push_back(pt.b);
    }
    pending.clear();
  }

  // --- transforms ----------------------------------------------------------
  Fuse fuse_base() const {
    Fuse f;
    for (int a = 0; a < 3; ++a) { f.vq[a] = vq[a]; f.vs[a] = vs[a]; }
    f.q02 = p.q0 * p.q0;
    f.alpha = p.alpha;
    f.scale = scale;
    f.dt = p.dt;
    f.rho_m = rho_m;
    f.eps = p.epsilon;
    return f;
  }
  // Algorithmic bytes of one pass: its inputs and outputs, N^3 each.
  double pass_bytes(DType dt, const PassIO& io, int axis) const {
    const double es = dt == F64 ? 8 : 4, n = (double)n3();
    const int op = io.f.op;
    double in = es, out = es;
    if (op == OP_PLAIN) in = es * (io.in2 ? 2 : 1);
    else if (op == OP_LD_GRAD_D || op == OP_LD_SH || op == OP_LD_RELAXLAP) in = 8;
    else if (op == OP_LD_AMP) in = 4 * es;
    else if (op == OP_LD_PROD) in = 8 + es;
    else if (op == OP_LD_DIV) in = 3 * es;
    else if (op == OP_LD_PSI) { in = 16 + 2 * es; out = es + 8; }
    (void)axis;
    return n * (in + out);
  }
  hipError_t run_pass(DType dt, int kind, int axis, const PassIO& io) {
    Scope sc(this, axis == 2 ? SMAISO_KC_R2R_Z : axis == 1 ? SMAISO_KC_R2R_Y : SMAISO_KC_R2R_X,
             pass_bytes(dt, io, axis));
    return r2r.pass(dt, kind, axis, io, g, stream);
  }
  // 3-D transform z, y, x; f's load op acts in the z pass
  hipError_t xform(DType dt, const int kinds[3], const void* in, const void* in2, void* out,
                   const Fuse* f = nullptr) {
    PassIO io;
    io.in = in; io.in2 = in2; io.out = out;
    if (f) io.f = *f;
    hipError_t e = run_pass(dt, kinds[2], 2, io);
    if (e != hipSuccess) return e;
    io.in = out; io.in2 = nullptr; io.f.op = OP_PLAIN;
    if ((e = run_pass(dt, kinds[1], 1, io)) != hipSuccess) return e;
    return run_pass(dt, kinds[0], 0, io);
  }
  hipError_t xw(const int kinds[3], const void* in, void* out, const void* in2 = nullptr) {
    return xform(mixed ? F32 : F64, kinds, in, in2, out);
  }
  hipError_t xwf(const int kinds[3], const Fuse& f, void* out) {
    return xform(mixed ? F32 : F64, kinds, nullptr, nullptr, out, &f);
  }
  // DCT-II (z pass with f's load op), DCT-II along y, then one x pass doing
  // DCT-II, the diagonal operator mid_op and DCT-III, then DCT-III along y, z.
  hipError_t xmid(const Fuse& f, int mid_op, void* out) {
    const DType dt = mixed ? F32 : F64;
    PassIO io;
    io.out = out;
    io.f = f;
    hipError_t e = run_pass(dt, REDFT10, 2, io);
    if (e != hipSuccess) return e;
    io.in = out; io.f.op = OP_PLAIN;
    if ((e = run_pass(dt, REDFT10, 1, io)) != hipSuccess) return e;
    io.f.op = mid_op;
    if ((e = run_pass(dt, REDFT10, 0, io)) != hipSuccess) return e;
    io.f.op = OP_PLAIN;
    if ((e = run_pass(dt, REDFT01, 1, io)) != hipSuccess) return e;
    return run_pass(dt, REDFT01, 2, io);
  }
  double fb(int nfields) const { return (double)n3() * tsz * nfields; }
};

namespace {

int fail(smaiso_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  else g_create_error = msg;
  return code;
}

#define HIPTRY(h, expr)                                                                    \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      return fail((h), _e == hipErrorOutOfMemory ? SMAISO_E_NOMEM : SMAISO_E_HIP,         \
                  fmt("%s: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__)); \
  } while (0)

bool finite_all(std::initializer_list<double> v) {
  for (double x : v)
    if (!std::isfinite(x)) return false;
  return true;
}

int validate(const smaiso_params* p, std::string* why) {
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

template <typename T>
int run_step(smaiso_handle* h) {
  using K = StepKernels<T>;
  hipStream_t s = h->stream;
  const StepConsts c = h->consts();
  T* const* gp = reinterpret_cast<T* const*>(h->gp);
  T* const* rd = reinterpret_cast<T* const*>(h->rd);
  T* const* S = reinterpret_cast<T* const*>(h->S);
  T* const* f = reinterpret_cast<T* const*>(h->f);
  T* const* vsolq = reinterpret_cast<T* const*>(h->vsolq);
  T* const* virrq = reinterpret_cast<T* const*>(h->virrq);
  T* const* v = reinterpret_cast<T* const*>(h->v);
  const double fb1 = h->fb(1);
  const double db1 = (double)h->n3() * 8;
  const Fuse fb = h->fuse_base();

  h->nloop += 1;
  std::swap(h->nlq, h->nlq_old);  // Nl_old_local = Nl_local, S:1100

  // a6 (S:1104-1250): rho^ = DCT(rho) in FP64; the DCT->DST shift and -Vs
  // (times scale for rho) are applied on load by the z pass of each inverse
  // transform
  HIPTRY(h, h->xform(F64, kCT, h->rho, nullptr, h->rhoq));
  for (int a = 0; a < 3; ++a) {
    Fuse g = fb;
    g.axis = a;
    g.op = OP_LD_GRAD_D; g.src[0] = h->psiq; g.c0 = 1.0;
    HIPTRY(h, h->xwf(kIST[a], g, gp[a]));
    g.op = OP_LD_GRAD_D; g.src[0] = h->rhoq; g.c0 = h->scale;
    HIPTRY(h, h->xwf(kIST[a], g, rd[a]));
  }
  // a7 (S:1254-1289): |amp| computed on load, Gaussian applied between the
  // forward and inverse x passes
  {
    Fuse g = fb;
    g.op = OP_LD_AMP;
    g.src[0] = h->psi; g.src[1] = gp[0]; g.src[2] = gp[1]; g.src[3] = gp[2];
    HIPTRY(h, h->xmid(g, OP_MID_AMP, h->amp));
  }
  // a8 (S:1293-1308)
  {
    Fuse g = fb;
    g.op = OP_LD_SH; g.src[0] = h->psiq;
    HIPTRY(h, h->xwf(kICT, g, h->lpsi));
  }
  // a9 (S:1312-1380)
  if (h->nloop > h->p.divv_switch || h->p.load_mode == 1) {
    Fuse g = fb;
    g.op = OP_LD_PROD; g.src[0] = h->rho; g.src[1] = h->lpsi;
    HIPTRY(h, h->xmid(g, OP_MID_NEGLAP, h->laprl));
  } else {
    Fuse g = fb;
    g.op = OP_LD_RELAXLAP; g.src[0] = h->psiq;
    HIPTRY(h, h->xwf(kICT, g, h->laprl));
    smaiso_handle::Scope sc(h, SMAISO_KC_POINTWISE, fb1 * 3 + db1);
    K::relax(c, static_cast<T*>(h->amp), static_cast<T*>(h->laprl), h->rho, s);
  }
  // a10 (S:1385-1425)
  {
    smaiso_handle::Scope sc(h, SMAISO_KC_POINTWISE, fb1 * 19 + db1);
    K::forces(c, static_cast<T*>(h->lpsi), static_cast<T*>(h->laprl), h->rho,
              static_cast<T*>(h->amp), static_cast<T*>(h->psi), gp[0], gp[1], gp[2], rd[0],
              rd[1], rd[2], static_cast<T*>(h->fgrad), static_cast<T*>(h->erho),
              static_cast<T*>(h->mu), S[0], S[1], S[2], f[0], f[1], f[2], s);
  }
  // a11 (S:1427-1457)
  HIPTRY(h, h->xw(kCT, h->fgrad, h->fgrad));
  HIPTRY(h, h->xw(kCT, h->erho, h->erho));
  for (int a = 0; a < 3; ++a) {
    HIPTRY(h, h->xw(kST[a], S[a], S[a]));
    HIPTRY(h, h->xw(kST[a], f[a], f[a]));
  }
  // a11 tail + a12 (S:1479-1974)
  {
    smaiso_handle::Scope sc(h, SMAISO_KC_SPECTRAL, fb1 * 11);
    K::velocity(c, S[0], S[1], S[2], static_cast<T*>(h->fgrad), static_cast<T*>(h->erho),
                vsolq[0], vsolq[1], vsolq[2], virrq[0], virrq[1], virrq[2], s);
  }
  // a13 + a14 (S:1979-2117): the divergences are formed on load by the z
  // pass of their inverse transforms; v = iST(vsol^ + virr^) by linearity
  {
    Fuse g = fb;
    g.op = OP_LD_DIV; g.c0 = 1.0;
    g.src[0] = virrq[0]; g.src[1] = virrq[1]; g.src[2] = virrq[2];
    HIPTRY(h, h->xwf(kICT, g, h->divv));
    g.c0 = h->scale;
    g.src[0] = f[0]; g.src[1] = f[1]; g.src[2] = f[2];
    HIPTRY(h, h->xwf(kICT, g, h->divf));
  }
  for (int a = 0; a < 3; ++a) HIPTRY(h, h->xw(kIST[a], vsolq[a], v[a], virrq[a]));
  // a15 + a16 (S:2121-2162)
  {
    smaiso_handle::Scope sc(h, SMAISO_KC_POINTWISE, fb1 * 18 + db1 * 4);
    K::density_nl(c, v[0], v[1], v[2], rd[0], rd[1], rd[2], static_cast<T*>(h->divv), h->rho,
                  h->f_rho, static_cast<T*>(h->mu),
                  static_cast<T*>(h->amp), static_cast<T*>(h->psi), static_cast<T*>(h->divf),
                  gp[0], gp[1], gp[2], static_cast<T*>(h->nl), s);
  }
  // a16 tail (S:2166-2172)
  HIPTRY(h, h->xw(kCT, h->nl, h->nlq));
  // a17 (S:2177-2198): the z pass of psi = iDCT(psiq_new) computes the
  // update on load and stores psiq_new over psiq_old
  std::swap(h->psi, h->psi_prev);  // psi_old_local = psi_local, S:2192
  {
    Fuse g = fb;
    g.op = OP_LD_PSI;
    g.src[0] = h->nlq;
    g.psiq = h->psiq;
    g.psiq_new = h->psiq_old;
    g.nlq_old = h->nlq_old;
    g.first = h->nloop == 1 ? 1 : 0;  // quirk Q5
    HIPTRY(h, h->xwf(kICT, g, h->psi));
  }
  std::swap(h->psiq, h->psiq_old);
  return SMAISO_OK;
}

int do_set_state(smaiso_handle* h, const double* psi, const double* rho, hipMemcpyKind kind) {
  HIPTRY(h, hipSetDevice(h->device));
  const long long n3 = h->n3();
  const size_t db = (size_t)n3 * 8;
  hipStream_t s = h->stream;
  const StepConsts c = h->consts();
  // pre-loop S:965-1001 in FP64 for both precisions: psiq = scale DCT(psi),
  // psi = iDCT(psiq), psiq_old = psiq; f_rho and the history reset (Q5)
  HIPTRY(h, hipMemcpyAsync(h->f_rho, psi, db, kind, s));
  HIPTRY(h, h->xform(F64, kCT, h->f_rho, nullptr, h->psiq));
  StepKernels<double>::spec(c, SPEC_SCALE, h->psiq, h->psiq, s);
  HIPTRY(h, h->xform(F64, kICT, h->psiq, nullptr, h->f_rho));
  if (h->mixed) convert_d2f(h->f_rho, static_cast<float*>(h->psi), n3, s);
  else HIPTRY(h, hipMemcpyAsync(h->psi, h->f_rho, db, hipMemcpyDeviceToDevice, s));
  HIPTRY(h, hipMemcpyAsync(h->psiq_old, h->psiq, db, hipMemcpyDeviceToDevice, s));
  HIPTRY(h, hipMemcpyAsync(h->rho, rho, db, kind, s));
  fill_zero_d(h->f_rho, n3, s);
  HIPTRY(h, hipGetLastError());
  HIPTRY(h, hipStreamSynchronize(s));
  h->collect_timing();
  h->nloop = 0;
  h->have_state = true;
  return SMAISO_OK;
}

// Produce field `fid` as FP64 on the device; returns a device pointer
// (possibly a handle scratch buffer).
int field_device(smaiso_handle* h, int fid, const double** out) {
  const long long n3 = h->n3();
  hipStream_t s = h->stream;
  auto as64 = [&](void* work) -> const double* {
    if (!h->mixed) return static_cast<const double*>(work);
    convert_f2d(static_cast<const float*>(work), h->out64, n3, s);
    return h->out64;
  };
  const StepConsts c = h->consts();
  switch (fid) {
    case SMAISO_FIELD_PSI: *out = as64(h->psi); return SMAISO_OK;
    case SMAISO_FIELD_RHO: *out = h->rho; return SMAISO_OK;
    case SMAISO_FIELD_AMP: *out = as64(h->amp); return SMAISO_OK;
    case SMAISO_FIELD_VX: case SMAISO_FIELD_VY: case SMAISO_FIELD_VZ:
      *out = as64(h->v[fid - SMAISO_FIELD_VX]); return SMAISO_OK;
    case SMAISO_FIELD_VSOLX: case SMAISO_FIELD_VSOLY: case SMAISO_FIELD_VSOLZ: {
      const int a = fid - SMAISO_FIELD_VSOLX;
      HIPTRY(h, h->xw(kIST[a], h->vsolq[a], h->tmp));
      *out = as64(h->tmp); return SMAISO_OK;
    }
    case SMAISO_FIELD_VIRRX: case SMAISO_FIELD_VIRRY: case SMAISO_FIELD_VIRRZ: {
      const int a = fid - SMAISO_FIELD_VIRRX;
      HIPTRY(h, h->xw(kIST[a], h->virrq[a], h->tmp));
      *out = as64(h->tmp); return SMAISO_OK;
    }
    case SMAISO_FIELD_DIVV: *out = as64(h->divv); return SMAISO_OK;
    case SMAISO_FIELD_LPSI: *out = as64(h->lpsi); return SMAISO_OK;
    case SMAISO_FIELD_MU: *out = as64(h->mu); return SMAISO_OK;
    case SMAISO_FIELD_ENERGY:
      if (h->mixed)
        StepKernels<float>::energy_density(c, static_cast<float*>(h->psi), static_cast<float*>(h->lpsi),
                                           h->rho, static_cast<float*>(h->amp),
                                           static_cast<float*>(h->tmp), s);
      else
        StepKernels<double>::energy_density(c, static_cast<double*>(h->psi), static_cast<double*>(h->lpsi),
                                            h->rho, static_cast<double*>(h->amp),
                                            static_cast<double*>(h->tmp), s);
      *out = as64(h->tmp); return SMAISO_OK;
    case SMAISO_FIELD_PSIQ: *out = h->psiq; return SMAISO_OK;
    case SMAISO_FIELD_PSIQ_OLD: *out = h->psiq_old; return SMAISO_OK;
    case SMAISO_FIELD_NLQ: *out = as64(h->nlq); return SMAISO_OK;
    case SMAISO_FIELD_F_RHO: *out = h->f_rho; return SMAISO_OK;
    case SMAISO_FIELD_DIVF: *out = as64(h->divf); return SMAISO_OK;
    case SMAISO_FIELD_GPSIX: case SMAISO_FIELD_GPSIY: case SMAISO_FIELD_GPSIZ:
      *out = as64(h->gp[fid - SMAISO_FIELD_GPSIX]); return SMAISO_OK;
    case SMAISO_FIELD_GRHOX: case SMAISO_FIELD_GRHOY: case SMAISO_FIELD_GRHOZ:
      *out = as64(h->rd[fid - SMAISO_FIELD_GRHOX]); return SMAISO_OK;
    case SMAISO_FIELD_LAPRHOL: *out = as64(h->laprl); return SMAISO_OK;
    case SMAISO_FIELD_PSI_OLD: *out = as64(h->psi_prev); return SMAISO_OK;
    default: return fail(h, SMAISO_E_INVALID, fmt("unknown field id %d", fid));
  }
}

}  // namespace

// ===========================================================================
// exported
// ===========================================================================
extern "C" {

int smaiso_abi_version(void) { return SMAISO_ABI_VERSION; }

void smaiso_params_default(smaiso_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->nx = p->ny = p->nz = 512;                 // S:218
  const double nw = 8, q0 = 1.0;               // S:261, S:244
  p->dx = p->dy = p->dz = 2.0 * M_PI / (q0 * nw);  // L/N, S:267-272
  p->dt = 0.001;                               // S:262
  p->gamma = 1.0; p->beta = 2.0; p->alpha = 1.0;  // S:239-241
  p->epsilon = -0.001 * 5;                     // ep_arg = 5, S:242-243
  p->q0 = q0;
  p->nu = 1.0;                                 // argv[1]
  p->lambda = 1.0; p->zeta = 1.0;              // S:256-257
  p->rho0 = 0.01;                              // S:251
  p->rho_m = 0.0;                              // -> rho_s/2, S:255
  p->divv_switch = 50;                         // S:214
  p->load_mode = 0;
  p->precision = SMAISO_FP64;
  p->slab_n0 = 0;
}

int smaiso_create(const smaiso_params* p, smaiso_handle** out) {
  if (!out) return fail(nullptr, SMAISO_E_INVALID, "out is NULL");
  *out = nullptr;
  std::string why;
  int rc = validate(p, &why);
  if (rc != SMAISO_OK) return fail(nullptr, rc, why);
  std::unique_ptr<smaiso_handle> h(new smaiso_handle());
  h->p = *p;
  h->g = Geom{p->nx, p->ny, p->nz};
  h->mixed = p->precision == SMAISO_FP32_MIXED;
  h->tsz = h->mixed ? 4 : 8;
  h->amp_eq = std::sqrt((3 * p->beta + std::sqrt(9 * p->beta * p->beta + 40 * p->epsilon * p->gamma)) / 5);
  h->kp = (1 - p->rho0) / h->amp_eq;
  h->rho_s = h->kp * h->amp_eq + p->rho0;
  h->rho_m = p->rho_m > 0 ? p->rho_m : h->rho_s / 2;
  h->scale = 0.125 / ((double)p->nx * p->ny * p->nz);

  auto hipfail = [&](hipError_t e, const char* what) {
    return fail(nullptr, e == hipErrorOutOfMemory ? SMAISO_E_NOMEM : SMAISO_E_HIP,
                fmt("%s: %s", what, hipGetErrorString(e)));
  };
  hipError_t e;
  if ((e = hipGetDevice(&h->device)) != hipSuccess) return hipfail(e, "hipGetDevice");
  if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess)
    return hipfail(e, "hipStreamCreate");
  const long long n3 = h->g.n3();
  auto alloc = [&](size_t bytes, void** ptr) -> hipError_t {
    hipError_t r = hipMalloc(ptr, bytes);
    if (r == hipSuccess) h->allocs.push_back(*ptr);
    return r;
  };
  const size_t db = (size_t)n3 * 8, tb = (size_t)n3 * h->tsz;
  void* tmpp = nullptr;
  // wavenumbers, S:447-481 (same expressions as the reference)
  const int nn[3] = {p->nx, p->ny, p->nz};
  const double dd[3] = {p->dx, p->dy, p->dz};
  for (int a = 0; a < 3; ++a) {
    std::vector<double> q(nn[a]), w(nn[a]);
    for (int n = 0; n < nn[a]; ++n) {
      q[n] = M_PI * n / (dd[a] * nn[a]);
      w[n] = M_PI * (n + 1) / (dd[a] * nn[a]);
    }
    if ((e = alloc(nn[a] * 8, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
    h->vq[a] = static_cast<double*>(tmpp);
    if ((e = alloc(nn[a] * 8, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
    h->vs[a] = static_cast<double*>(tmpp);
    if ((e = hipMemcpy(h->vq[a], q.data(), nn[a] * 8, hipMemcpyHostToDevice)) != hipSuccess)
      return hipfail(e, "hipMemcpy");
    if ((e = hipMemcpy(h->vs[a], w.data(), nn[a] * 8, hipMemcpyHostToDevice)) != hipSuccess)
      return hipfail(e, "hipMemcpy");
  }
  double** dbufs[] = {&h->psiq, &h->psiq_old, &h->rho, &h->f_rho, &h->rhoq};
  for (double** b : dbufs) {
    if ((e = alloc(db, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc (state)");
    *b = static_cast<double*>(tmpp);
  }
  if ((e = alloc(8 * 8 * 1024, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
  h->partials = static_cast<double*>(tmpp);
  if ((e = alloc(8 * 4, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
  h->diag_out = static_cast<double*>(tmpp);
  std::vector<void**> tbufs = {&h->psi, &h->psi_prev, &h->nlq, &h->nlq_old, &h->amp,
                               &h->lpsi, &h->laprl, &h->mu, &h->fgrad, &h->erho, &h->divv,
                               &h->divf, &h->nl, &h->tmp};
  for (int a = 0; a < 3; ++a) {
    tbufs.push_back(&h->gp[a]); tbufs.push_back(&h->rd[a]); tbufs.push_back(&h->S[a]);
    tbufs.push_back(&h->f[a]); tbufs.push_back(&h->vsolq[a]); tbufs.push_back(&h->virrq[a]);
    tbufs.push_back(&h->v[a]);
  }
  for (void** b : tbufs)
    if ((e = alloc(tb, b)) != hipSuccess) return hipfail(e, "hipMalloc (work fields)");
  if (h->mixed) {
    if ((e = alloc(db, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
    h->out64 = static_cast<double*>(tmpp);
  }
  *out = h.release();
  return SMAISO_OK;
}

int smaiso_destroy(smaiso_handle* h) {
  delete h;
  return SMAISO_OK;
}

const char* smaiso_last_error(const smaiso_handle* h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}

int smaiso_set_state(smaiso_handle* h, const double* psi, const double* rho) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!psi || !rho) return fail(h, SMAISO_E_INVALID, "psi/rho is NULL");
  return do_set_state(h, psi, rho, hipMemcpyHostToDevice);
}

int smaiso_set_state_device(smaiso_handle* h, const void* psi, const void* rho) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!psi || !rho) return fail(h, SMAISO_E_INVALID, "psi/rho is NULL");
  return do_set_state(h, static_cast<const double*>(psi), static_cast<const double*>(rho),
                      hipMemcpyDeviceToDevice);
}

int smaiso_step_async(smaiso_handle* h, int32_t nsteps) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!h->have_state) return fail(h, SMAISO_E_STATE, "smaiso_step before smaiso_set_state");
  if (nsteps < 0) return fail(h, SMAISO_E_INVALID, "nsteps < 0");
  HIPTRY(h, hipSetDevice(h->device));
  for (int i = 0; i < nsteps; ++i) {
    int rc = h->mixed ? run_step<float>(h) : run_step<double>(h);
    if (rc != SMAISO_OK) return rc;
  }
  HIPTRY(h, hipGetLastError());
  return SMAISO_OK;
}

int smaiso_synchronize(smaiso_handle* h) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  HIPTRY(h, hipSetDevice(h->device));
  HIPTRY(h, hipStreamSynchronize(h->stream));
  h->collect_timing();
  return SMAISO_OK;
}

int smaiso_step(smaiso_handle* h, int32_t nsteps) {
  int rc = smaiso_step_async(h, nsteps);
  if (rc != SMAISO_OK) return rc;
  return smaiso_synchronize(h);
}

int64_t smaiso_step_count(const smaiso_handle* h) { return h ? h->nloop : -1; }

int smaiso_get_field_device(smaiso_handle* h, int32_t field, void* out_dev) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!out_dev) return fail(h, SMAISO_E_INVALID, "out is NULL");
  if (!h->have_state) return fail(h, SMAISO_E_STATE, "no state (call smaiso_set_state first)");
  HIPTRY(h, hipSetDevice(h->device));
  const double* src = nullptr;
  int rc = field_device(h, field, &src);
  if (rc != SMAISO_OK) return rc;
  HIPTRY(h, hipMemcpyAsync(out_dev, src, (size_t)h->n3() * 8, hipMemcpyDeviceToDevice, h->stream));
  HIPTRY(h, hipStreamSynchronize(h->stream));
  h->collect_timing();
  return SMAISO_OK;
}

int smaiso_get_field(smaiso_handle* h, int32_t field, double* out) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!out) return fail(h, SMAISO_E_INVALID, "out is NULL");
  if (!h->have_state) return fail(h, SMAISO_E_STATE, "no state (call smaiso_set_state first)");
  HIPTRY(h, hipSetDevice(h->device));
  const double* src = nullptr;
  int rc = field_device(h, field, &src);
  if (rc != SMAISO_OK) return rc;
  HIPTRY(h, hipMemcpyAsync(out, src, (size_t)h->n3() * 8, hipMemcpyDeviceToHost, h->stream));
  HIPTRY(h, hipStreamSynchronize(h->stream));
  h->collect_timing();
  return SMAISO_OK;
}

int smaiso_diagnostics(smaiso_handle* h, double* l1, double* mass, double* energy) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!h->have_state || h->nloop < 1)
    return fail(h, SMAISO_E_STATE, "diagnostics need at least one completed step");
  HIPTRY(h, hipSetDevice(h->device));
  const StepConsts c = h->consts();
  hipStream_t s = h->stream;
  int nb;
  {
    smaiso_handle::Scope sc(h, SMAISO_KC_REDUCE, h->fb(4) + (double)h->n3() * 8);
    if (h->mixed)
      nb = StepKernels<float>::diag_partial(c, static_cast<float*>(h->psi), static_cast<float*>(h->psi_prev),
                                            h->rho, static_cast<float*>(h->lpsi),
                                            static_cast<float*>(h->amp), h->partials, s);
    else
      nb = StepKernels<double>::diag_partial(c, static_cast<double*>(h->psi),
                                             static_cast<double*>(h->psi_prev), h->rho,
                                             static_cast<double*>(h->lpsi),
                                             static_cast<double*>(h->amp), h->partials, s);
    diag_final(h->partials, nb, h->diag_out, s);
  }
  double sums[4];
  HIPTRY(h, hipMemcpyAsync(sums, h->diag_out, sizeof(sums), hipMemcpyDeviceToHost, s));
  HIPTRY(h, hipStreamSynchronize(s));
  h->collect_timing();
  const double vol = h->p.dx * h->p.dy * h->p.dz;
  const double a = sums[0] / (h->p.dt * sums[1]);  // S:2246
  const double m = sums[2] * vol;                   // S:2252
  const double en = sums[3] * vol;                  // S:2258
  if (l1) *l1 = a;
  if (mass) *mass = m;
  if (energy) *energy = en;
  if (!std::isfinite(a) || !std::isfinite(m) || !std::isfinite(en))
    return fail(h, SMAISO_E_NONFINITE, "non-finite diagnostics (state blew up)");
  return SMAISO_OK;
}

namespace {
std::mutex g_r2r_mu;
std::map<int, std::unique_ptr<R2R>> g_r2r;  // per device

R2R* standalone_r2r(int dev) {
  std::lock_guard<std::mutex> lk(g_r2r_mu);
  auto& p = g_r2r[dev];
  if (!p) p.reset(new R2R());
  return p.get();
}

int check_r2r_args(const int32_t kinds[3], int32_t nx, int32_t ny, int32_t nz, int32_t dtype) {
  if (!kinds) return fail(nullptr, SMAISO_E_INVALID, "kinds is NULL");
  for (int a = 0; a < 3; ++a)
    if (kinds[a] < 0 || kinds[a] > 3) return fail(nullptr, SMAISO_E_INVALID, fmt("bad kind %d", kinds[a]));
  if (nx < 1 || ny < 1 || nz < 1) return fail(nullptr, SMAISO_E_INVALID, "extents must be >= 1");
  if ((long long)nx * ny * nz > 2147483647LL / 2) return fail(nullptr, SMAISO_E_INVALID, "grid too large");
  if (dtype != SMAISO_FP64 && dtype != SMAISO_FP32_MIXED) return fail(nullptr, SMAISO_E_INVALID, "bad dtype");
  return SMAISO_OK;
}

int run_r2r(const int32_t kinds[3], const Geom& g, const void* in, void* out, DType dt,
            hipStream_t s) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(nullptr, SMAISO_E_HIP, hipGetErrorString(e));
  R2R* r = standalone_r2r(dev);
  PassIO io;
  io.in = in;
  io.out = out;
  const int k[3] = {kinds[0], kinds[1], kinds[2]};
  e = r->transform3d(dt, k, io, g, s);
  if (e != hipSuccess) return fail(nullptr, SMAISO_E_HIP, fmt("r2r: %s", hipGetErrorString(e)));
  return SMAISO_OK;
}
}  // namespace

int smaiso_r2r3d_device(const int32_t kinds[3], int32_t nx, int32_t ny, int32_t nz,
                        const void* in_dev, void* out_dev, int32_t dtype, void* stream) {
  int rc = check_r2r_args(kinds, nx, ny, nz, dtype);
  if (rc != SMAISO_OK) return rc;
  if (!in_dev || !out_dev) return fail(nullptr, SMAISO_E_INVALID, "in/out is NULL");
  return run_r2r(kinds, Geom{nx, ny, nz}, in_dev, out_dev, dtype == SMAISO_FP64 ? F64 : F32,
                 static_cast<hipStream_t>(stream));
}

int smaiso_r2r3d(const int32_t kinds[3], int32_t nx, int32_t ny, int32_t nz, const double* in,
                 double* out, int32_t dtype) {
  int rc = check_r2r_args(kinds, nx, ny, nz, dtype);
  if (rc != SMAISO_OK) return rc;
  if (!in || !out) return fail(nullptr, SMAISO_E_INVALID, "in/out is NULL");
  const long long n3 = (long long)nx * ny * nz;
  const size_t es = dtype == SMAISO_FP64 ? 8 : 4;
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, (size_t)n3 * es);
  if (e != hipSuccess) return fail(nullptr, SMAISO_E_NOMEM, hipGetErrorString(e));
  std::vector<float> f;
  if (dtype == SMAISO_FP64) {
    e = hipMemcpy(d, in, (size_t)n3 * 8, hipMemcpyHostToDevice);
  } else {
    f.assign(in, in + n3);
    e = hipMemcpy(d, f.data(), (size_t)n3 * 4, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) {
    rc = run_r2r(kinds, Geom{nx, ny, nz}, d, d, dtype == SMAISO_FP64 ? F64 : F32, nullptr);
    if (rc == SMAISO_OK) {
      if (dtype == SMAISO_FP64) {
        e = hipMemcpy(out, d, (size_t)n3 * 8, hipMemcpyDeviceToHost);
      } else {
        e = hipMemcpy(f.data(), d, (size_t)n3 * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess)
          for (long long i = 0; i < n3; ++i) out[i] = f[i];
      }
    }
  }
  (void)hipFree(d);
  if (rc != SMAISO_OK) return rc;
  if (e != hipSuccess) return fail(nullptr, SMAISO_E_HIP, hipGetErrorString(e));
  return SMAISO_OK;
}

int smaiso_make_interface_ic(const smaiso_params* p, const smaiso_ic* ic, const double* height,
                             double* psi, double* rho) {
  if (!p || !ic || !psi || !rho) return fail(nullptr, SMAISO_E_INVALID, "NULL argument");
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(nullptr, SMAISO_E_INVALID, "bad grid");
  if (!(ic->width > 0)) return fail(nullptr, SMAISO_E_INVALID, "width must be > 0");
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

int smaiso_timing_enable(smaiso_handle* h, int32_t enable) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  HIPTRY(h, hipSetDevice(h->device));
  HIPTRY(h, hipStreamSynchronize(h->stream));
  h->collect_timing();
  h->timing = enable != 0;
  for (int k = 0; k < SMAISO_KC_COUNT; ++k) { h->kc_ms[k] = 0; h->kc_n[k] = 0; h->kc_bytes[k] = 0; }
  return SMAISO_OK;
}

int smaiso_timing_get(smaiso_handle* h, int32_t kc, double* total_ms, int64_t* launches,
                      double* bytes) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (kc < 0 || kc >= SMAISO_KC_COUNT) return fail(h, SMAISO_E_INVALID, "bad kernel class");
  if (total_ms) *total_ms = h->kc_ms[kc];
  if (launches) *launches = h->kc_n[kc];
  if (bytes) *bytes = h->kc_bytes[kc];
  return SMAISO_OK;
}

}  // extern "C"
