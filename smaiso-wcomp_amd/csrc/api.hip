// C ABI (include/smaiso.h): handles, device buffers, the step schedule,
// hipGraph replay and ensembles.
//
// A handle is a group of `nmem` simulations of one grid shape that step
// together: every field holds the members back to back and every kernel takes
// the member as blockIdx.y.  One call of run_step<T>() is one iteration of the
// reference's loop body (smaiso-wcomp.cpp:1081-2198) for all members.
#include "../../include/smaiso.h"
#include "host_abi.h"
#include "common.h"
#include "adjoint.h"
#include "fused.h"
#include "r2r.h"
#include "stepper.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

using namespace smaiso;

namespace {

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

// FP32-arithmetic stages of an FP32-mixed handle at creation: the stages
// whose FP32 rounding keeps psi within 3e-7 of the FP64 stepper at 256^3 and
// 512x512x256 after 100, 500 and 2000 steps: forces, div v, div f, v's
// transforms, Nl and the velocity solve (psi within 4.8e-8 with these alone,
// the FP32 storage level), the amplitude filter's Gaussian factors (psi
// within 2.8e-7 at step 500, 1.4e-7 at 2000) and grad rho's y and x passes
// (from rho - rho0: psi within 2.8e-7 at step 500, 1.5e-7 at 2000).  grad psi
// and lap(rho L psi) each reach 3e-5, L psi 1e-6.
constexpr uint32_t kFp32StagesDefault = SMAISO_STAGE_FORCES | SMAISO_STAGE_DIVV | SMAISO_STAGE_DIVF |
                                        SMAISO_STAGE_VELOCITY | SMAISO_STAGE_NL | SMAISO_STAGE_VELSOLVE |
                                        SMAISO_STAGE_AMP | SMAISO_STAGE_GRAD_RHO;

double derived_amp(const smaiso_params& p) {
  return std::sqrt((3 * p.beta + std::sqrt(9 * p.beta * p.beta + 40 * p.epsilon * p.gamma)) / 5);
}

}  // namespace

struct smaiso_handle {
  smaiso_params p{};               // parameters shared by the group (member 0's)
  std::vector<smaiso_params> mp;   // per member
  int nmem = 1;
  Geom g{};                        // g.nb = nmem
  int device = 0;
  hipStream_t stream = nullptr;
  R2R r2r;
  std::string err = "no error";
  std::vector<char> member_set;
  long long nloop = 0;  // steps since the last set_state (the reference's nLoop)
  long long epoch = 0;  // steps since creation: its parity names the buffer rotation
  // The work fields hold the intermediates of a completed step (amp, L psi,
  // v, mu, psi_prev, ...): false after set_state or a checkpoint load until
  // the next step, so getters and diagnostics never read fields no step wrote.
  bool derived = false;
  bool nlq_valid = false;  // N^l of the last step: written by a step or a checkpoint load
  // virrq holds the irrotational velocity of the step that ended at epoch
  // virr_epoch; keep_virr makes the step write it (the adjoint's recomputed
  // steps read it)
  long long virr_epoch = -1;
  bool keep_virr = false;
  // the adjoint stores the intermediates of each recomputed segment's steps
  // when they fit in free device memory (otherwise, or when off, it
  // recomputes each step before transposing it)
  bool adj_store = true;
  // divf holds the y and x terms of div f only: the density pass formed the z
  // term from f_z (still in f[2] until the next step)
  bool divf_partial = false;
  // laprl holds the z lines of lap(rho L psi) before its last inverse z pass
  // (the forces pass runs that pass in its tile); the read-out completes it
  bool laprl_zlines = false;
  // gpx_next holds d_x psi of the current psi, formed by the x pass of psi's
  // transform at the end of the last step (or after a state was set, loaded
  // or restored); the next step takes it as grad_x psi
  void* gpx_next = nullptr;
  bool gpx_valid = false;
  // lpsi_next holds the z lines of L psi of the current psi, formed by the z
  // pass of psi's transform at the end of the last step (or after a state
  // was set, loaded or restored); the next step runs their y and x passes
  void* lpsi_next = nullptr;
  bool lz_valid = false;
  // fgrad holds fgrad^ + G erho^ (formed by the x pass of the two forward
  // transforms) and erho is not transformed along x: the velocity solve
  // reads fgrad alone
  bool fg_combined = false;
  bool mixed = false;
  size_t tsz = 8;  // bytes per work element
  uint32_t fp32_stages = 0;  // SMAISO_STAGE_* computed in FP32 (FP32-mixed only)
  // FP32 arithmetic for the passes of stage `bit`
  int a32(uint32_t bit) const { return mixed && (fp32_stages & bit) ? 1 : 0; }

  // derived (S:250-255, S:278), shared by the group
  double amp_eq = 0, rho_m = 0, scale = 0;
  std::vector<MemberConsts> mc_host;
  MemberConsts* mc = nullptr;  // device

  // device memory
  std::vector<void*> allocs;
  double* vq[3] = {nullptr, nullptr, nullptr};
  double* vs[3] = {nullptr, nullptr, nullptr};
  double* gq[3] = {nullptr, nullptr, nullptr};  // exp(-1.57^2 2 vq^2) per axis
  double* gs[3] = {nullptr, nullptr, nullptr};  // exp(-1.57^2 2 vs^2) per axis
  // FP64 state
  double *psiq = nullptr, *psiq_old = nullptr, *rho = nullptr, *f_rho = nullptr;
  // grad rho: formed in FP64 from the FP64 rho (rho is ~constant, so its
  // gradient is a small difference of large values) and stored in the work
  // type, whose rounding is then relative to the gradient itself
  void* rd[3] = {nullptr, nullptr, nullptr};
  // d_z rho of the current rho between steps (work type), where the density
  // pass leaves it: a step swaps it into rd[2] first, so rd[2] keeps the
  // step's own grad rho for the read-outs and the adjoint
  void* rdz_next = nullptr;
  double *partials = nullptr, *diag_out = nullptr;
  // FP64 field of the group's size, FP32 storage only: the field getters'
  // widened copy, and within a step the intermediates of psi's inverse
  // transform (nothing is kept in it from one call to the next)
  double* out64 = nullptr;
  std::vector<double> diag_host;
  long long diag_epoch = -1;  // epoch of the cached diagnostics (-1: none)
  // work fields (T = double or float)
  void *psi = nullptr, *psi_prev = nullptr, *nlq = nullptr, *nlq_old = nullptr;
  // vtq: the total velocity vsol^ + virr^ of the last step (spectral, on the
  // staggered lattices); virrq: its irrotational part, written only when asked
  // for (virr_epoch, keep_virr): the step forms div v^ without it
  void *gp[3] = {}, *S[3] = {}, *f[3] = {}, *vtq[3] = {}, *virrq[3] = {}, *v[3] = {};
  void *amp = nullptr, *lpsi = nullptr, *laprl = nullptr, *mu = nullptr, *fgrad = nullptr,
       *erho = nullptr, *divv = nullptr, *divf = nullptr, *nl = nullptr, *tmp = nullptr;

  // hipGraph replay of two steps (the three buffer swaps of a step cancel
  // after two, so one graph serves every pair that starts at the same epoch
  // parity)
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_parity = -1;

  int* surf_dev = nullptr;  // surface tracker output (nx*ny), allocated on first use

  // adjoint work fields (FP64, allocated by the first smaiso_adjoint call)
  std::vector<double*> adj;
  double* adj_partials = nullptr;
  double* adj_acc = nullptr;

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
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& pt : pending) { ev_free.push_back(pt.a); ev_free.push_back(pt.b); }
    for (auto e : ev_free) (void)hipEventDestroy(e);
    for (void* a : allocs) (void)hipFree(a);
    if (stream) (void)hipStreamDestroy(stream);
  }

  StepConsts consts() const {
    StepConsts c{};
    c.nx = g.nx; c.ny = g.ny; c.nz = g.nz;
    c.nmem = nmem;
    c.slab_n0 = p.slab_n0;
    c.scale = scale;
    c.dt = p.dt; c.alpha = p.alpha; c.beta = p.beta; c.gamma = p.gamma; c.eps = p.epsilon;
    c.q0 = p.q0; c.q02 = p.q0 * p.q0; c.rho_m = rho_m;
    for (int a = 0; a < 3; ++a) {
      c.vq[a] = vq[a]; c.vs[a] = vs[a];
      c.qs[a] = M_PI / ((a == 0 ? p.dx : a == 1 ? p.dy : p.dz) * (a == 0 ? p.nx : a == 1 ? p.ny : p.nz));
      c.gq[a] = gq[a]; c.gs[a] = gs[a];
    }
    c.mc = mc;
    return c;
  }
  // one member as a batch of one (pointers are offset by the caller)
  StepConsts consts_one(int m) const {
    StepConsts c = consts();
    c.nmem = 1;
    c.mc = mc + m;
    return c;
  }
  Geom g_one() const { Geom q = g; q.nb = 1; return q; }

  long long n3() const { return g.n3(); }
  long long ntot() const { return g.n3() * nmem; }
  bool all_set() const {
    for (char c : member_set) if (!c) return false;
    return true;
  }
  // does step number n (1-based) take the steady branch with the full history?
  bool steady(long long n) const {
    const bool relax = !(n > p.divv_switch || p.load_mode == 1);
    return !relax && n > 1;
  }
  // the forces pass runs the last inverse z pass of lap(rho L psi) in steady
  // steps (except the adjoint's recomputed ones)
  bool laprl_in_forces() const { return forces_zpass_laprl(mixed ? F32 : F64, g.nz); }
  // psi's transform ends with the x pass that also forms d_x psi (length 256
  // along x; not when grad psi computes in FP32)
  bool gpx_in_psi_pass() const { return yx256_invgrad_supported(g) && !a32(SMAISO_STAGE_GRAD_PSI); }
  // ... and its z pass also forms the z lines of the next step's L psi (FP64
  // arithmetic, so not when L psi computes in FP32)
  // FP64 storage, 256 points along x: steady steps form fgrad^ + G erho^ in
  // one x pass
  bool fg_in_x_pass() const { return !mixed && yx256_fgx_supported(g) && velocity_zpass_supported(g.nz); }
  bool lz_in_psi_pass() const {
    return gpx_in_psi_pass() && !a32(SMAISO_STAGE_LPSI) && psi_sh_zpass_supported(g.nz);
  }

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
    const double dd[3] = {p.dx, p.dy, p.dz};
    const int nn[3] = {p.nx, p.ny, p.nz};
    for (int a = 0; a < 3; ++a) { f.vq[a] = vq[a]; f.vs[a] = vs[a]; f.qs[a] = M_PI / (dd[a] * nn[a]); }
    f.q02 = p.q0 * p.q0;
    f.alpha = p.alpha;
    f.scale = scale;
    f.dt = p.dt;
    f.rho_m = rho_m;
    f.eps = p.epsilon;
    return f;
  }
  // Algorithmic bytes of one pass: its inputs and outputs, N^3 each.
  double pass_bytes(DType dt, const PassIO& io, const Geom& gg) const {
    const double es = dt == F64 ? 8 : 4, n = (double)gg.n3() * gg.nb;
    const int op = io.f.op;
    double in = es, out = es;
    if (op == OP_PLAIN) in = es * (io.in2 ? 2 : 1);
    else if (op == OP_LD_GRAD_D || op == OP_LD_SH || op == OP_LD_RELAXLAP) in = 8;
    else if (op == OP_LD_AMP) in = 4 * es;
    else if (op == OP_LD_PROD) in = 8 + es;
    else if (op == OP_LD_PSI) { in = 16 + 2 * es; out = es + 8; }
    else if (op == OP_AX_GAUSS_AMP) in = 4 * es;
    else if (op == OP_AX_DIV && io.f.accumulate) in = 2 * es;
    return n * (in + out);
  }
  hipError_t run_pass(DType dt, int kind, int axis, const PassIO& io, const Geom& gg) {
    Scope sc(this, is_axis_op(io.f.op) ? SMAISO_KC_R2R_AXIS
                   : axis == 2 ? SMAISO_KC_R2R_Z : axis == 1 ? SMAISO_KC_R2R_Y : SMAISO_KC_R2R_X,
             pass_bytes(dt, io, gg));
    return r2r.pass(dt, kind, axis, io, gg, stream);
  }
  // one pass reading in_dt and storing out_dt (R2R::pass_cast)
  hipError_t run_pass_cast(DType in_dt, DType out_dt, int kind, int axis, const PassIO& io) {
    const double n = (double)g.n3() * g.nb;
    const double bytes = pass_bytes(in_dt, io, g) + n * ((out_dt == F64 ? 8 : 4) - (in_dt == F64 ? 8 : 4));
    Scope sc(this, is_axis_op(io.f.op) ? SMAISO_KC_R2R_AXIS
                   : axis == 2 ? SMAISO_KC_R2R_Z : axis == 1 ? SMAISO_KC_R2R_Y : SMAISO_KC_R2R_X,
             bytes);
    return r2r.pass_cast(in_dt, out_dt, kind, axis, io, g, stream);
  }
  // 3-D transform z, y, x of all members (or of gg's batch); f's load op acts
  // in the z pass
  hipError_t xform(DType dt, const int kinds[3], const void* in, const void* in2, void* out,
                   const Fuse* f = nullptr, const Geom* gg = nullptr) {
    const Geom& G = gg ? *gg : g;
    PassIO io;
    io.in = in; io.in2 = in2; io.out = out;
    if (f) io.f = *f;
    hipError_t e = run_pass(dt, kinds[2], 2, io, G);
    if (e != hipSuccess) return e;
    io.in = out; io.in2 = nullptr; io.f.op = OP_PLAIN;
    if ((e = run_pass(dt, kinds[1], 1, io, G)) != hipSuccess) return e;
    return run_pass(dt, kinds[0], 0, io, G);
  }
  hipError_t xw(const int kinds[3], const void* in, void* out, const void* in2 = nullptr, int arith32 = 0) {
    Fuse f;
    f.arith32 = arith32;
    return xform(mixed ? F32 : F64, kinds, in, in2, out, &f);
  }
  hipError_t xwf(const int kinds[3], const Fuse& f, void* out) {
    return xform(mixed ? F32 : F64, kinds, nullptr, nullptr, out, &f);
  }
  // As xwf for FP32 storage, with the outputs of the z and y passes kept in
  // FP64 (in out64): the z pass reads FP32 sources and stores FP64, the y
  // pass runs in FP64 and the x pass stores FP32.  Each pass computes in
  // FP64 either way; this removes the FP32 rounding of partly transformed
  // values, which each later pass spreads over whole lines of the output.
  hipError_t xwf_wide(const int kinds[3], const Fuse& f, void* out) {
    PassIO io;
    io.f = f;
    io.out = out64;
    hipError_t e = run_pass_cast(F32, F64, kinds[2], 2, io);
    if (e != hipSuccess) return e;
    io.in = out64;
    io.f.op = OP_PLAIN;
    if ((e = run_pass(F64, kinds[1], 1, io, g)) != hipSuccess) return e;
    io.out = out;
    return run_pass_cast(F64, F32, kinds[0], 0, io);
  }
  // The x pass of psi's inverse transform from its FP64 x-spectral lines
  // `wide` (members of gg from member m0), storing psi and d_x psi (into
  // gpx_next) in the storage type.
  hipError_t psi_dx_pass(const double* wide, const Geom& gg, long long m0) {
    const void *tw = nullptr, *mk = nullptr;
    hipError_t e = r2r.table_ptrs(gg.nx, &tw, &mk);
    if (e != hipSuccess) return e;
    const long long off = m0 * g.n3();
    char* const psi_o = static_cast<char*>(psi) + off * tsz;
    char* const gpx_o = static_cast<char*>(gpx_next) + off * tsz;
    Scope sc(this, SMAISO_KC_R2R_AXIS, (double)gg.n3() * gg.nb * (8 + 2 * tsz));
    return launch_yx256_invgrad(mixed ? F32 : F64, wide, psi_o, gpx_o, fuse_base().qs[0], gg, tw, mk, stream);
  }
  // The z pass of psi = ICT(psiq) that also stores the z lines of L psi
  // (into zl): with the update computed on load (f, as OP_LD_PSI), or from
  // psiq as it is (pq, update = false); psi's z lines (FP64) into zpsi.
  hipError_t psi_lz_pass(const Fuse& f, bool update, const double* pq, double* zpsi, void* zl,
                         const StepConsts& cc) {
    const double n = (double)cc.nx * cc.ny * cc.nz * cc.nmem;
    Scope sc(this, SMAISO_KC_R2R_Z, n * ((update ? 16 + 2 * (double)tsz + 8 : 8) + 8 + (double)tsz));
    if (mixed) {
      PsiShIO<float> io;
      io.f = f; io.psiq = pq; io.zpsi = zpsi; io.zlpsi = static_cast<float*>(zl);
      return psi_sh_zpass<float>(cc, io, update, r2r, stream);
    }
    PsiShIO<double> io;
    io.f = f; io.psiq = pq; io.zpsi = zpsi; io.zlpsi = static_cast<double*>(zl);
    return psi_sh_zpass<double>(cc, io, update, r2r, stream);
  }
  // psi = ICT(psiq_new) with the update computed on load by the z pass (f:
  // OP_LD_PSI), whose output and the y pass's stay FP64 (in psi for FP64
  // storage, out64 for FP32 storage); the y pass on its default tile; the x
  // pass also forms d_x psi for the next step.  With lz_valid the z pass is
  // psi_lz_pass, which also leaves the next step's L psi z lines in
  // lpsi_next.  psi_dx_refresh repeats the z and y passes from a stored psiq
  // with the same kernels and tiles, so it reproduces both bit for bit.
  hipError_t psi_transform_with_dx(const Fuse& f) {
    PassIO io;
    io.f = f;
    double* const wide = mixed ? out64 : static_cast<double*>(psi);
    io.out = wide;
    hipError_t e = lz_valid ? psi_lz_pass(f, true, nullptr, wide, lpsi_next, consts())
                   : mixed  ? run_pass_cast(F32, F64, REDFT01, 2, io)
                            : run_pass(F64, REDFT01, 2, io, g);
    if (e != hipSuccess) return e;
    io.in = wide;
    io.f.op = OP_PLAIN;
    io.f.untuned = 1;
    if ((e = run_pass(F64, REDFT01, 1, io, g)) != hipSuccess) return e;
    return psi_dx_pass(wide, g, 0);
  }
  // d_x psi of psi = ICT(psiq) (member m, or all) into gpx_next, and L psi's
  // z lines into lpsi_next, after a state was set, loaded or restored: z and
  // y passes of psiq as the step runs them (the z pass's plan is that of the
  // update-on-load pass, or psi_lz_pass without the update), then the fused
  // x pass, whose psi output (equal to the stored psi where it came from a
  // step) goes to scratch.
  hipError_t psi_dx_refresh(int m) {
    gpx_valid = gpx_in_psi_pass();
    lz_valid = lz_in_psi_pass();
    if (!gpx_valid) return hipSuccess;
    const Geom G = m < 0 ? g : g_one();
    const long long m0 = m < 0 ? 0 : m, off = m0 * g.n3();
    double* const wide = (mixed ? out64 : static_cast<double*>(tmp)) + off;
    PassIO io;
    io.in = psiq + off;
    io.out = wide;
    io.f.untuned = 1;
    hipError_t e = lz_valid ? psi_lz_pass(fuse_base(), false, psiq + off, wide,
                                          static_cast<char*>(lpsi_next) + off * tsz,
                                          m < 0 ? consts() : consts_one(m))
                            : run_pass(F64, REDFT01, 2, io, G);
    if (e != hipSuccess) return e;
    io.in = wide;
    if ((e = run_pass(F64, REDFT01, 1, io, G)) != hipSuccess) return e;
    return psi_dx_scratch_pass(wide, G, m0);
  }
  // as psi_dx_pass with psi stored back over `wide` (FP64) rather than into
  // the field; d_x psi into gpx_next
  hipError_t psi_dx_scratch_pass(double* wide, const Geom& gg, long long m0) {
    const void *tw = nullptr, *mk = nullptr;
    hipError_t e = r2r.table_ptrs(gg.nx, &tw, &mk);
    if (e != hipSuccess) return e;
    char* const gpx_o = static_cast<char*>(gpx_next) + m0 * g.n3() * tsz;
    if (mixed) {
      // FP32 outputs: psi into the FP32 scratch field, d_x psi into gpx_next
      char* const tmp_o = static_cast<char*>(tmp) + m0 * g.n3() * tsz;
      return launch_yx256_invgrad(F32, wide, tmp_o, gpx_o, fuse_base().qs[0], gg, tw, mk, stream);
    }
    return launch_yx256_invgrad(F64, wide, wide, gpx_o, fuse_base().qs[0], gg, tw, mk, stream);
  }
  // one y pass in place (the work type, FP64 arithmetic)
  hipError_t y_pass(int kind, void* buf) {
    PassIO io;
    io.in = buf;
    io.out = buf;
    return run_pass(mixed ? F32 : F64, kind, 1, io, g);
  }
  // the x pass of the forward transforms of fgrad and erho (their z and y
  // passes done) that stores fgrad^ + G erho^ over fgrad (fg_in_x_pass)
  hipError_t fg_x_pass() {
    const void *tw = nullptr, *mk = nullptr;
    hipError_t e = r2r.table_ptrs(g.nx, &tw, &mk);
    if (e != hipSuccess) return e;
    const StepConsts c = consts();
    Scope sc(this, SMAISO_KC_R2R_X, (double)g.n3() * g.nb * 24);
    return launch_yx256_fgx(static_cast<const double*>(fgrad), static_cast<const double*>(erho),
                            static_cast<double*>(fgrad), c.qs, g, tw, mk, stream);
  }
  // the y and x passes of a transform whose z pass has run, in place
  hipError_t xw_yx(const int kinds[3], void* buf, int arith32 = 0) {
    const DType dt = mixed ? F32 : F64;
    PassIO io;
    io.in = buf;
    io.out = buf;
    io.f.arith32 = arith32;
    hipError_t e = run_pass(dt, kinds[1], 1, io, g);
    if (e != hipSuccess) return e;
    return run_pass(dt, kinds[0], 0, io, g);
  }
  // DCT-II (z pass with f's load op), DCT-II along y, then one x pass doing
  // DCT-II, the diagonal operator mid_op and DCT-III, then DCT-III along y, z
  // (the last z pass only when last_z).
  hipError_t xmid(const Fuse& f, int mid_op, void* out, bool last_z = true) {
    return xmid_dt(mixed ? F32 : F64, nullptr, f, mid_op, out, last_z);
  }
  // as xmid in precision dt; `in` is read by the z pass when f.op is OP_PLAIN
  hipError_t xmid_dt(DType dt, const void* in, const Fuse& f, int mid_op, void* out, bool last_z = true) {
    PassIO io;
    io.in = in;
    io.out = out;
    io.f = f;
    hipError_t e = run_pass(dt, REDFT10, 2, io, g);
    if (e != hipSuccess) return e;
    io.in = out; io.f.op = OP_PLAIN;
    if ((e = run_pass(dt, REDFT10, 1, io, g)) != hipSuccess) return e;
    io.f.op = mid_op;
    if ((e = run_pass(dt, REDFT10, 0, io, g)) != hipSuccess) return e;
    io.f.op = OP_PLAIN;
    if ((e = run_pass(dt, REDFT01, 1, io, g)) != hipSuccess) return e;
    return last_z ? run_pass(dt, REDFT01, 2, io, g) : hipSuccess;
  }
  // one-axis operator pass (OP_AX_*) along `axis`
  hipError_t axpass(DType dt, int axis, const void* in, void* out, const Fuse& f) {
    PassIO io;
    io.in = in;
    io.out = out;
    io.f = f;
    return run_pass(dt, REDFT10, axis, io, g);
  }
  // as axpass, reading in_dt and storing out_dt
  hipError_t axpass_cast(DType in_dt, DType out_dt, int axis, const void* in, void* out, const Fuse& f) {
    PassIO io;
    io.in = in;
    io.out = out;
    io.f = f;
    return run_pass_cast(in_dt, out_dt, REDFT10, axis, io);
  }
  double fb(int nfields) const { return (double)ntot() * tsz * nfields; }
};

struct smaiso_ensemble {
  std::vector<std::unique_ptr<smaiso_handle>> groups;
  std::vector<std::pair<int, int>> where;  // member -> (group, index in group)
  int device = 0;
  hipStream_t master = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_parity = -1;
  bool use_graph = true;
  std::string err = "no error";
  ~smaiso_ensemble() {
    (void)hipSetDevice(device);
    if (master) (void)hipStreamSynchronize(master);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (fork) (void)hipEventDestroy(fork);
    for (auto e : joins) (void)hipEventDestroy(e);
    groups.clear();
    if (master) (void)hipStreamDestroy(master);
  }
};

namespace {

int fail(smaiso_handle* h, int code, const std::string& msg) {
  if (!h) return abi_fail(code, msg);
  h->err = msg;
  return code;
}

#define HIPTRY(h, expr)                                                                    \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      return fail((h), _e == hipErrorOutOfMemory ? SMAISO_E_NOMEM : SMAISO_E_HIP,         \
                  fmt("%s: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__)); \
  } while (0)

double rho_m_eff(const smaiso_params& p) {
  const double amp = derived_amp(p);
  const double kp = (1 - p.rho0) / amp;
  const double rho_s = kp * amp + p.rho0;  // S:253
  return p.rho_m > 0 ? p.rho_m : rho_s / 2;
}

// Parameters two members must share to step in one batch.
bool same_group(const smaiso_params& a, const smaiso_params& b) {
  return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dx == b.dx && a.dy == b.dy &&
         a.dz == b.dz && a.dt == b.dt && a.alpha == b.alpha && a.beta == b.beta &&
         a.gamma == b.gamma && a.epsilon == b.epsilon && a.q0 == b.q0 &&
         rho_m_eff(a) == rho_m_eff(b) && a.divv_switch == b.divv_switch &&
         a.load_mode == b.load_mode && a.precision == b.precision && a.slab_n0 == b.slab_n0;
}

// Create a group of n members (all valid, all same_group).
int create_group(const smaiso_params* ps, int n, smaiso_handle** out) {
  *out = nullptr;
  std::unique_ptr<smaiso_handle> h(new smaiso_handle());
  const smaiso_params& p = ps[0];
  h->p = p;
  h->mp.assign(ps, ps + n);
  h->nmem = n;
  h->member_set.assign(n, 0);
  h->g = Geom{p.nx, p.ny, p.nz, n};
  h->mixed = p.precision == SMAISO_FP32_MIXED;
  h->tsz = h->mixed ? 4 : 8;
  h->fp32_stages = kFp32StagesDefault;
  h->amp_eq = derived_amp(p);
  h->rho_m = rho_m_eff(p);
  h->scale = 0.125 / ((double)p.nx * p.ny * p.nz);
  for (int m = 0; m < n; ++m) {
    MemberConsts c;
    c.nu = ps[m].nu;
    c.lam = ps[m].lambda;
    c.zeta = ps[m].zeta;
    c.rho0 = ps[m].rho0;
    c.amp_eq = h->amp_eq;
    c.kp = (1 - ps[m].rho0) / h->amp_eq;  // S:252
    c.rnu = 1.0 / c.nu;
    c.rlam = 1.0 / (c.lam + 2 * c.nu);
    h->mc_host.push_back(c);
  }
  h->diag_host.assign(4 * n, 0.0);

  auto hipfail = [&](hipError_t e, const char* what) {
    return fail(nullptr, e == hipErrorOutOfMemory ? SMAISO_E_NOMEM : SMAISO_E_HIP,
                fmt("%s: %s", what, hipGetErrorString(e)));
  };
  hipError_t e;
  if ((e = hipGetDevice(&h->device)) != hipSuccess) return hipfail(e, "hipGetDevice");
  if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess)
    return hipfail(e, "hipStreamCreate");
  auto alloc = [&](size_t bytes, void** ptr) -> hipError_t {
    hipError_t r = hipMalloc(ptr, bytes);
    if (r == hipSuccess) h->allocs.push_back(*ptr);
    return r;
  };
  const long long nt = h->ntot();
  const size_t db = (size_t)nt * 8, tb = (size_t)nt * h->tsz;
  void* tmpp = nullptr;
  // wavenumbers, S:447-481 (same expressions as the reference)
  const int nn[3] = {p.nx, p.ny, p.nz};
  const double dd[3] = {p.dx, p.dy, p.dz};
  for (int a = 0; a < 3; ++a) {
    std::vector<double> q(nn[a]), w(nn[a]), eq(nn[a]), ew(nn[a]);
    for (int i = 0; i < nn[a]; ++i) {
      q[i] = M_PI * i / (dd[a] * nn[a]);
      w[i] = M_PI * (i + 1) / (dd[a] * nn[a]);
      eq[i] = std::exp(-1.57 * 1.57 * 2 * (q[i] * q[i]));
      ew[i] = std::exp(-1.57 * 1.57 * 2 * (w[i] * w[i]));
    }
    if ((e = alloc(nn[a] * 8, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
    h->vq[a] = static_cast<double*>(tmpp);
    if ((e = alloc(nn[a] * 8, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
    h->vs[a] = static_cast<double*>(tmpp);
    if ((e = hipMemcpy(h->vq[a], q.data(), nn[a] * 8, hipMemcpyHostToDevice)) != hipSuccess)
      return hipfail(e, "hipMemcpy");
    if ((e = hipMemcpy(h->vs[a], w.data(), nn[a] * 8, hipMemcpyHostToDevice)) != hipSuccess)
      return hipfail(e, "hipMemcpy");
    if ((e = alloc(nn[a] * 8, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
    h->gq[a] = static_cast<double*>(tmpp);
    if ((e = alloc(nn[a] * 8, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
    h->gs[a] = static_cast<double*>(tmpp);
    if ((e = hipMemcpy(h->gq[a], eq.data(), nn[a] * 8, hipMemcpyHostToDevice)) != hipSuccess)
      return hipfail(e, "hipMemcpy");
    if ((e = hipMemcpy(h->gs[a], ew.data(), nn[a] * 8, hipMemcpyHostToDevice)) != hipSuccess)
      return hipfail(e, "hipMemcpy");
  }
  if ((e = alloc(sizeof(MemberConsts) * n, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
  h->mc = static_cast<MemberConsts*>(tmpp);
  if ((e = hipMemcpy(h->mc, h->mc_host.data(), sizeof(MemberConsts) * n, hipMemcpyHostToDevice)) != hipSuccess)
    return hipfail(e, "hipMemcpy");
  double** dbufs[] = {&h->psiq, &h->psiq_old, &h->rho, &h->f_rho};
  for (double** b : dbufs) {
    if ((e = alloc(db, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc (state)");
    *b = static_cast<double*>(tmpp);
  }
  if ((e = alloc((size_t)8 * 8 * kDiagBlocks * n, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
  h->partials = static_cast<double*>(tmpp);
  if ((e = alloc((size_t)8 * 4 * n, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
  h->diag_out = static_cast<double*>(tmpp);
  std::vector<void**> tbufs = {&h->psi, &h->psi_prev, &h->nlq, &h->nlq_old, &h->amp,
                               &h->lpsi, &h->laprl, &h->mu, &h->fgrad, &h->erho, &h->divv,
                               &h->divf, &h->nl, &h->tmp, &h->rdz_next, &h->gpx_next,
                               &h->lpsi_next};
  for (int a = 0; a < 3; ++a) {
    tbufs.push_back(&h->gp[a]); tbufs.push_back(&h->rd[a]); tbufs.push_back(&h->S[a]);
    tbufs.push_back(&h->f[a]); tbufs.push_back(&h->vtq[a]); tbufs.push_back(&h->virrq[a]);
    tbufs.push_back(&h->v[a]);
  }
  for (void** b : tbufs)
    if ((e = alloc(tb, b)) != hipSuccess) return hipfail(e, "hipMalloc (work fields)");
  if (h->mixed) {
    if ((e = alloc(db, &tmpp)) != hipSuccess) return hipfail(e, "hipMalloc");
    h->out64 = static_cast<double*>(tmpp);
  }
  for (int a = 0; a < 3; ++a) {
    if ((e = h->r2r.prepare(nn[a], h->mixed)) != hipSuccess) return hipfail(e, "r2r tables");
  }
  *out = h.release();
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
  T* const* vtq = reinterpret_cast<T* const*>(h->vtq);
  T* const* virrq = reinterpret_cast<T* const*>(h->virrq);
  T* const* v = reinterpret_cast<T* const*>(h->v);
  const double fb1 = h->fb(1);
  const double db1 = (double)h->ntot() * 8;
  const Fuse fb = h->fuse_base();

  h->nloop += 1;
  h->epoch += 1;
  std::swap(h->nlq, h->nlq_old);  // Nl_old_local = Nl_local, S:1100

  // a6 (S:1104-1250): grad_a = IST_a(-Vs_a shift(scale CT(.))) acts on one
  // axis only (iDCT.DCT = 2N on the others), so each component is one pass
  // along its axis: from rho in FP64, and from psi (= ICT(psiq)) in the work
  // type (the passes compute in FP64, so the FP32 copy of psi contributes
  // only its storage rounding).  The passes that read one input run back to
  // back, so its later reads find it in the memory-side cache.
  // Along z it is there already where the density update runs in the z
  // pass of Nl's transform: that pass leaves d_z rho_new, and setting,
  // loading or restoring a state computes it with the same code.
  const bool rdz_kept = density_zpass_supported(h->g.nz);
  if (rdz_kept) std::swap(h->rd[2], h->rdz_next);
  for (int a = 0; a < (rdz_kept ? 2 : 3); ++a) {
    Fuse g = fb;
    g.axis = a;
    g.op = OP_AX_GRAD;
    if (h->mixed) {
      // FP32 arithmetic (the 256-point y and x passes) on rho - rho0: the
      // gradient ignores the constant, the FP32 tile rounds the difference
      g.arith32 = h->a32(SMAISO_STAGE_GRAD_RHO);
      g.ld_shift = h->mc_host[0].rho0;
      HIPTRY(h, h->axpass_cast(F64, F32, a, h->rho, rd[a], g));
    } else {
      HIPTRY(h, h->axpass(F64, a, h->rho, rd[a], g));
    }
  }
  // grad psi along x and y, then one z pass forms d_z psi, |amp| at the
  // same points and the z factor of a7's separable Gaussian filter (d_z psi
  // in FP64; FP32 for grad psi takes the separate passes)
  const bool afused = amp_zpass_supported(h->g.nz) && !h->a32(SMAISO_STAGE_GRAD_PSI);
  // along x it is there already where psi's transform left it
  const bool gx = h->gpx_valid && !h->a32(SMAISO_STAGE_GRAD_PSI);
  if (gx) std::swap(h->gp[0], h->gpx_next);
  for (int a = gx ? 1 : 0; a < (afused ? 2 : 3); ++a) {
    Fuse g = fb;
    g.axis = a;
    g.op = OP_AX_GRAD;
    g.arith32 = h->a32(SMAISO_STAGE_GRAD_PSI);
    HIPTRY(h, h->axpass(h->mixed ? F32 : F64, a, h->psi, gp[a], g));
  }
  // a7 (S:1254-1289): the Gaussian exp(-c |q|^2) is a product over the axes,
  // one pass per axis; |amp| is computed on load by the first
  {
    const DType dt = h->mixed ? F32 : F64;
    Fuse g = fb;
    if (afused) {
      AmpIO<T> io;
      io.psi = static_cast<const T*>(h->psi);
      io.gpx = gp[0];
      io.gpy = gp[1];
      io.gpz = gp[2];
      io.amp = static_cast<T*>(h->amp);
      io.q02 = fb.q02;
      io.qs = fb.qs[2];
      smaiso_handle::Scope sc(h, SMAISO_KC_R2R_AXIS, fb1 * 5);
      HIPTRY(h, amp_zpass<T>(c, io, h->r2r, h->a32(SMAISO_STAGE_AMP) != 0, s));
    } else {
      g.op = OP_AX_GAUSS_AMP;
      g.axis = 2;
      g.arith32 = h->a32(SMAISO_STAGE_AMP);
      g.src[0] = h->psi; g.src[1] = gp[0]; g.src[2] = gp[1]; g.src[3] = gp[2];
      HIPTRY(h, h->axpass(dt, 2, nullptr, h->amp, g));
    }
    g.op = OP_AX_GAUSS;
    g.arith32 = h->a32(SMAISO_STAGE_AMP);
    for (int a = 1; a >= 0; --a) {
      g.axis = a;
      HIPTRY(h, h->axpass(dt, a, h->amp, h->amp, g));
    }
  }
  // a8 (S:1293-1308); its z pass ran with psi's transform where it could
  if (h->lz_valid && !h->a32(SMAISO_STAGE_LPSI)) {
    std::swap(h->lpsi, h->lpsi_next);
    HIPTRY(h, h->xw_yx(kICT, h->lpsi));
  } else {
    Fuse g = fb;
    g.op = OP_LD_SH; g.src[0] = h->psiq;
    g.arith32 = h->a32(SMAISO_STAGE_LPSI);
    HIPTRY(h, h->xwf(kICT, g, h->lpsi));
  }
  // a9 (S:1312-1380); in steady steps of FP64 handles the forces pass runs
  // the last inverse z pass of lap(rho L psi) (not in the adjoint's
  // recomputed steps, whose transpose reads the field)
  const bool forces_z = forces_zpass_supported(h->mixed ? F32 : F64, h->g.nz);
  h->laprl_zlines = false;
  h->fg_combined = false;
  if (h->nloop > h->p.divv_switch || h->p.load_mode == 1) {
    Fuse g = fb;
    g.op = OP_LD_PROD; g.src[0] = h->rho; g.src[1] = h->lpsi;
    g.arith32 = h->a32(SMAISO_STAGE_LAPRHOL);
    const bool lapz = h->laprl_in_forces() && !h->keep_virr;
    HIPTRY(h, h->xmid(g, OP_MID_NEGLAP, h->laprl, !lapz));
    h->laprl_zlines = lapz;
  } else {
    Fuse g = fb;
    g.op = OP_LD_RELAXLAP; g.src[0] = h->psiq;
    HIPTRY(h, h->xwf(kICT, g, h->laprl));
    smaiso_handle::Scope sc(h, SMAISO_KC_POINTWISE, fb1 * 3 + db1);
    K::relax(c, static_cast<T*>(h->amp), static_cast<T*>(h->laprl), h->rho, s);
  }
  // a10 (S:1385-1425) and a11 (S:1427-1457): the z pass of the five
  // transforms computes the forces on load where the length allows
  if (forces_z) {
    ForcesIO<T> io;
    io.lpsi = static_cast<const T*>(h->lpsi); io.laprl = static_cast<const T*>(h->laprl);
    io.laprl_zlines = h->laprl_zlines;
    io.amp = static_cast<const T*>(h->amp); io.psi = static_cast<const T*>(h->psi);
    io.rho = h->rho;
    for (int a = 0; a < 3; ++a) { io.gp[a] = gp[a]; io.rd[a] = rd[a]; io.f[a] = f[a]; io.zout[2 + a] = S[a]; }
    io.mu = static_cast<T*>(h->mu);
    io.zout[0] = static_cast<T*>(h->fgrad);
    io.zout[1] = static_cast<T*>(h->erho);
    {
      smaiso_handle::Scope sc(h, SMAISO_KC_POINTWISE, fb1 * 19 + db1);
      HIPTRY(h, forces_zpass<T>(c, io, h->r2r, s));
    }
    const int fa = h->a32(SMAISO_STAGE_FORCES);
    if (!h->keep_virr && h->fg_in_x_pass()) {
      // the y passes of both, then one x pass storing fgrad^ + G erho^
      HIPTRY(h, h->y_pass(kCT[1], h->fgrad));
      HIPTRY(h, h->y_pass(kCT[1], h->erho));
      HIPTRY(h, h->fg_x_pass());
      h->fg_combined = true;
    } else {
      HIPTRY(h, h->xw_yx(kCT, h->fgrad, fa));
      HIPTRY(h, h->xw_yx(kCT, h->erho, fa));
    }
    for (int a = 0; a < 3; ++a) HIPTRY(h, h->xw_yx(kST[a], S[a], fa));
  } else {
    {
      smaiso_handle::Scope sc(h, SMAISO_KC_POINTWISE, fb1 * 19 + db1);
      K::forces(c, static_cast<T*>(h->lpsi), static_cast<T*>(h->laprl), h->rho,
                static_cast<T*>(h->amp), static_cast<T*>(h->psi), gp[0], gp[1], gp[2], rd[0],
                rd[1], rd[2], static_cast<T*>(h->fgrad), static_cast<T*>(h->erho),
                static_cast<T*>(h->mu), S[0], S[1], S[2], f[0], f[1], f[2], s);
    }
    HIPTRY(h, h->xw(kCT, h->fgrad, h->fgrad));
    HIPTRY(h, h->xw(kCT, h->erho, h->erho));
    for (int a = 0; a < 3; ++a) HIPTRY(h, h->xw(kST[a], S[a], S[a]));
  }
  // a11 tail + a12 (S:1479-1974), the spectral div v^ of a13 (S:2042-2063)
  // and the inverse transforms of div v and v = iST(vsol^ + virr^) (S:2085-
  // 2117, by linearity): the z pass of the four transforms solves the
  // velocity on load where the length allows; otherwise the solve stores v^
  // and div v^ and the transforms follow.  The adjoint's recomputed steps
  // also keep virr^ (the unfused path).
  const bool vfused = !h->keep_virr && velocity_zpass_supported(h->g.nz);
  if (vfused) {
    VelocityIO<T> io;
    for (int a = 0; a < 3; ++a) { io.S[a] = S[a]; io.zout[a] = v[a]; }
    io.fgq = static_cast<const T*>(h->fgrad);
    io.erq = h->fg_combined ? nullptr : static_cast<const T*>(h->erho);
    io.zout[3] = static_cast<T*>(h->divv);
    {
      smaiso_handle::Scope sc(h, SMAISO_KC_SPECTRAL, fb1 * (h->fg_combined ? 8 : 9));
      HIPTRY(h, velocity_zpass<T>(c, io, h->r2r, h->a32(SMAISO_STAGE_VELSOLVE) != 0,
                                  h->a32(SMAISO_STAGE_DIVV) && h->a32(SMAISO_STAGE_VELOCITY), s));
    }
    HIPTRY(h, h->xw_yx(kICT, h->divv, h->a32(SMAISO_STAGE_DIVV)));
  } else {
    smaiso_handle::Scope sc(h, SMAISO_KC_SPECTRAL, fb1 * (h->keep_virr ? 12 : 9));
    K::velocity(c, S[0], S[1], S[2], static_cast<T*>(h->fgrad),
                h->fg_combined ? nullptr : static_cast<T*>(h->erho),
                vtq[0], vtq[1], vtq[2], h->keep_virr ? virrq[0] : nullptr, virrq[1], virrq[2],
                static_cast<T*>(h->divv), h->a32(SMAISO_STAGE_VELSOLVE) != 0, s);
    if (h->keep_virr) h->virr_epoch = h->epoch;
  }
  if (!vfused) HIPTRY(h, h->xw(kICT, h->divv, h->divv, nullptr, h->a32(SMAISO_STAGE_DIVV)));
  // div f = iCT(scale sum_a Vq_a shift(ST_a f_a)) (S:2064-2092) is a sum of
  // one-axis operators, one accumulating pass per component; the z term is
  // formed by the density pass where it runs (divf then holds the y and x
  // terms)
  const bool dens_z = density_zpass_supported(h->g.nz);
  {
    Fuse d = fb;
    d.op = OP_AX_DIV;
    d.arith32 = h->a32(SMAISO_STAGE_DIVF);
    const int a0 = dens_z ? 1 : 2;
    for (int a = a0; a >= 0; --a) {
      d.axis = a;
      d.accumulate = a != a0;
      HIPTRY(h, h->axpass(h->mixed ? F32 : F64, a, f[a], h->divf, d));
    }
  }
  h->divf_partial = dens_z;
  for (int a = 0; a < 3; ++a) {
    if (vfused) HIPTRY(h, h->xw_yx(kIST[a], v[a], h->a32(SMAISO_STAGE_VELOCITY)));
    else HIPTRY(h, h->xw(kIST[a], vtq[a], v[a], nullptr, h->a32(SMAISO_STAGE_VELOCITY)));
  }
  // a15 + a16 (S:2121-2162) and the transform of Nl (S:2166-2168): the z
  // pass of the transform computes them on load where the length allows
  if (dens_z) {
    DensityIO<T> io;
    for (int a = 0; a < 3; ++a) { io.v[a] = v[a]; io.rd[a] = rd[a]; io.gp[a] = gp[a]; }
    io.divv = static_cast<const T*>(h->divv);
    io.rho = h->rho;
    io.f_rho = h->f_rho;
    io.mu = static_cast<const T*>(h->mu);
    io.amp = static_cast<const T*>(h->amp);
    io.psi = static_cast<const T*>(h->psi);
    io.divf = static_cast<const T*>(h->divf);
    io.nlz = static_cast<T*>(h->nlq);
    io.rdz_next = static_cast<T*>(h->rdz_next);
    io.fz = f[2];
    io.qs = fb.qs[2];
    {
      // reads 15 work fields (f_z among them), writes Nl's z lines and d_z rho
      smaiso_handle::Scope sc(h, SMAISO_KC_POINTWISE, fb1 * 17 + db1 * 4);
      HIPTRY(h, density_zpass<T>(c, io, h->r2r, s));
    }
    HIPTRY(h, h->xw_yx(kCT, h->nlq, h->a32(SMAISO_STAGE_NL)));
  } else {
    {
      smaiso_handle::Scope sc(h, SMAISO_KC_POINTWISE, fb1 * 15 + db1 * 4);
      K::density_nl(c, v[0], v[1], v[2], rd[0], rd[1], rd[2], static_cast<T*>(h->divv), h->rho,
                    h->f_rho, static_cast<T*>(h->mu), static_cast<T*>(h->amp),
                    static_cast<T*>(h->psi), static_cast<T*>(h->divf), gp[0], gp[1], gp[2],
                    static_cast<T*>(h->nl), s);
    }
    HIPTRY(h, h->xw(kCT, h->nl, h->nlq));
  }
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
    // FP32 storage: psi's transform keeps FP64 intermediates.  Their FP32
    // rounding, fed back through the nonlinear terms every step, was the
    // largest FP32 storage error on near-stationary states (about 20x the
    // rounding of the stored fields themselves on a flat interface).
    // Where it can, the x pass also forms d_x psi for the next step (the z
    // and y passes on their default tiles, which psi_dx_refresh repeats) and
    // the z pass the z lines of the next step's L psi
    h->gpx_valid = h->gpx_in_psi_pass();
    h->lz_valid = h->lz_in_psi_pass();
    if (h->gpx_valid) HIPTRY(h, h->psi_transform_with_dx(g));
    else if (h->mixed) HIPTRY(h, h->xwf_wide(kICT, g, h->psi));
    else HIPTRY(h, h->xwf(kICT, g, h->psi));
  }
  std::swap(h->psiq, h->psiq_old);
  h->derived = true;
  h->nlq_valid = true;
  return SMAISO_OK;
}

int run_step_any(smaiso_handle* h) { return h->mixed ? run_step<float>(h) : run_step<double>(h); }

// d_z rho of the current rho (all members, or member m) into rdz_next,
// where steps keep it from one to the next (the density pass leaves it)
hipError_t refresh_rdz(smaiso_handle* h, int m = -1) {
  if (!density_zpass_supported(h->g.nz)) return hipSuccess;
  const StepConsts c = m < 0 ? h->consts() : h->consts_one(m);
  const long long off = m < 0 ? 0 : (long long)m * h->n3();
  const double qs = h->fuse_base().qs[2];
  if (h->mixed) return rho_dz<float>(c, h->rho + off, static_cast<float*>(h->rdz_next) + off, qs, h->r2r, h->stream);
  return rho_dz<double>(c, h->rho + off, static_cast<double*>(h->rdz_next) + off, qs, h->r2r, h->stream);
}

// Pre-loop of member m (S:965-1001) in FP64 for both precisions:
// psiq = scale DCT(psi), psi = iDCT(psiq), psiq_old = psiq, f_rho = 0.
int do_set_state(smaiso_handle* h, int m, const double* psi, const double* rho, hipMemcpyKind kind) {
  HIPTRY(h, hipSetDevice(h->device));
  const long long n3 = h->n3(), off = (long long)m * n3;
  const size_t db = (size_t)n3 * 8;
  hipStream_t s = h->stream;
  const StepConsts c = h->consts_one(m);
  const Geom g1 = h->g_one();
  double* psiq = h->psiq + off;
  double* scratch = h->f_rho + off;
  HIPTRY(h, hipMemcpyAsync(scratch, psi, db, kind, s));
  HIPTRY(h, h->xform(F64, kCT, scratch, nullptr, psiq, nullptr, &g1));
  StepKernels<double>::spec(c, SPEC_SCALE, psiq, psiq, s);
  HIPTRY(h, h->xform(F64, kICT, psiq, nullptr, scratch, nullptr, &g1));
  if (h->mixed) convert_d2f(scratch, static_cast<float*>(h->psi) + off, n3, s);
  else HIPTRY(h, hipMemcpyAsync(static_cast<double*>(h->psi) + off, scratch, db, hipMemcpyDeviceToDevice, s));
  HIPTRY(h, hipMemcpyAsync(h->psiq_old + off, psiq, db, hipMemcpyDeviceToDevice, s));
  HIPTRY(h, hipMemcpyAsync(h->rho + off, rho, db, kind, s));
  HIPTRY(h, refresh_rdz(h, m));
  HIPTRY(h, h->psi_dx_refresh(m));
  fill_zero_d(scratch, n3, s);
  HIPTRY(h, hipGetLastError());
  HIPTRY(h, hipStreamSynchronize(s));
  h->collect_timing();
  // the whole group restarts its step count (quirk Q5): members of a group
  // that has stepped since its last set lose their history too, so they must
  // be set again before the group steps
  if (h->nloop > 0)
    for (int q = 0; q < h->nmem; ++q)
      if (q != m) h->member_set[q] = 0;
  h->nloop = 0;
  h->member_set[m] = 1;
  h->diag_epoch = -1;
  h->derived = false;
  h->nlq_valid = false;
  return SMAISO_OK;
}

// virr^ of the last step in virrq: the velocity solve run again from the
// step's spectral S, fgrad and erho (the step itself stores only the total
// and div v^); its div v^ goes to tmp, which the caller overwrites
template <typename T>
hipError_t solve_virr(smaiso_handle* h) {
  T* const* S = reinterpret_cast<T* const*>(h->S);
  T* const* vtq = reinterpret_cast<T* const*>(h->vtq);
  T* const* virrq = reinterpret_cast<T* const*>(h->virrq);
  StepKernels<T>::velocity(h->consts(), S[0], S[1], S[2], static_cast<T*>(h->fgrad),
                           h->fg_combined ? nullptr : static_cast<T*>(h->erho), vtq[0], vtq[1], vtq[2],
                           virrq[0], virrq[1],
                           virrq[2], static_cast<T*>(h->tmp), h->a32(SMAISO_STAGE_VELSOLVE) != 0,
                           h->stream);
  return hipGetLastError();
}
hipError_t ensure_virr(smaiso_handle* h) {
  if (h->virr_epoch == h->epoch) return hipSuccess;
  const hipError_t e = h->mixed ? solve_virr<float>(h) : solve_virr<double>(h);
  if (e == hipSuccess) h->virr_epoch = h->epoch;
  return e;
}

// Produce field `fid` of all members as FP64 on the device; returns the
// device pointer of member 0 (possibly a handle scratch buffer).
int field_device(smaiso_handle* h, int fid, const double** out) {
  const bool state = fid == SMAISO_FIELD_PSI || fid == SMAISO_FIELD_RHO || fid == SMAISO_FIELD_PSIQ ||
                     fid == SMAISO_FIELD_PSIQ_OLD || fid == SMAISO_FIELD_F_RHO;
  if (fid >= 0 && fid < SMAISO_FIELD_COUNT && !state &&
      !(fid == SMAISO_FIELD_NLQ ? h->nlq_valid : h->derived))
    return fail(h, SMAISO_E_STATE,
                fmt("field %d is produced by a step: none has run since the state was set or loaded", fid));
  const long long nt = h->ntot();
  hipStream_t s = h->stream;
  auto as64 = [&](void* work) -> const double* {
    if (!h->mixed) return static_cast<const double*>(work);
    convert_f2d(static_cast<const float*>(work), h->out64, nt, s);
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
      // vsol = v - iST(virr^): the step keeps the total; virr^ is solved again
      // from the step's spectral S, fgrad and erho, which stay until the next step
      const int a = fid - SMAISO_FIELD_VSOLX;
      HIPTRY(h, ensure_virr(h));
      HIPTRY(h, h->xw(kIST[a], h->virrq[a], h->tmp));
      if (h->mixed)
        sub_fields(static_cast<const float*>(h->v[a]), static_cast<const float*>(h->tmp),
                   static_cast<float*>(h->tmp), nt, s);
      else
        sub_fields(static_cast<const double*>(h->v[a]), static_cast<const double*>(h->tmp),
                   static_cast<double*>(h->tmp), nt, s);
      *out = as64(h->tmp); return SMAISO_OK;
    }
    case SMAISO_FIELD_VIRRX: case SMAISO_FIELD_VIRRY: case SMAISO_FIELD_VIRRZ: {
      const int a = fid - SMAISO_FIELD_VIRRX;
      HIPTRY(h, ensure_virr(h));
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
    case SMAISO_FIELD_DIVF:
      if (h->divf_partial) {
        // the stored y and x terms plus the z term from f_z
        HIPTRY(h, hipMemcpyAsync(h->tmp, h->divf, (size_t)nt * h->tsz, hipMemcpyDeviceToDevice, s));
        Fuse d = h->fuse_base();
        d.op = OP_AX_DIV;
        d.axis = 2;
        d.accumulate = 1;
        d.arith32 = h->a32(SMAISO_STAGE_DIVF);
        HIPTRY(h, h->axpass(h->mixed ? F32 : F64, 2, h->f[2], h->tmp, d));
        *out = as64(h->tmp);
        return SMAISO_OK;
      }
      *out = as64(h->divf);
      return SMAISO_OK;
    case SMAISO_FIELD_GPSIX: case SMAISO_FIELD_GPSIY: case SMAISO_FIELD_GPSIZ:
      *out = as64(h->gp[fid - SMAISO_FIELD_GPSIX]); return SMAISO_OK;
    case SMAISO_FIELD_GRHOX: case SMAISO_FIELD_GRHOY: case SMAISO_FIELD_GRHOZ:
      *out = as64(h->rd[fid - SMAISO_FIELD_GRHOX]); return SMAISO_OK;
    case SMAISO_FIELD_LAPRHOL:
      if (h->laprl_zlines) {
        // the last inverse z pass, which the step ran inside the forces pass
        PassIO io;
        io.in = h->laprl;
        io.out = h->tmp;
        io.f.arith32 = h->a32(SMAISO_STAGE_LAPRHOL);
        HIPTRY(h, h->run_pass(h->mixed ? F32 : F64, REDFT01, 2, io, h->g));
        *out = as64(h->tmp);
        return SMAISO_OK;
      }
      *out = as64(h->laprl);
      return SMAISO_OK;
    case SMAISO_FIELD_PSI_OLD: *out = as64(h->psi_prev); return SMAISO_OK;
    default: return fail(h, SMAISO_E_INVALID, fmt("unknown field id %d", fid));
  }
}

int get_field(smaiso_handle* h, int m, int field, void* out, hipMemcpyKind kind) {
  if (!out) return fail(h, SMAISO_E_INVALID, "out is NULL");
  if (!h->all_set()) return fail(h, SMAISO_E_STATE, "no state (call smaiso_set_state first)");
  HIPTRY(h, hipSetDevice(h->device));
  const double* src = nullptr;
  int rc = field_device(h, field, &src);
  if (rc != SMAISO_OK) return rc;
  HIPTRY(h, hipMemcpyAsync(out, src + (long long)m * h->n3(), (size_t)h->n3() * 8, kind, h->stream));
  HIPTRY(h, hipStreamSynchronize(h->stream));
  h->collect_timing();
  return SMAISO_OK;
}

int diagnostics(smaiso_handle* h, int m, double* l1, double* mass, double* energy) {
  if (!h->all_set() || h->nloop < 1 || !h->derived)
    return fail(h, SMAISO_E_STATE,
                "diagnostics need a completed step since the state was set or loaded");
  HIPTRY(h, hipSetDevice(h->device));
  if (h->diag_epoch != h->epoch) {
    const StepConsts c = h->consts();
    hipStream_t s = h->stream;
    {
      smaiso_handle::Scope sc(h, SMAISO_KC_REDUCE, h->fb(4) + (double)h->ntot() * 8);
      int nb;
      if (h->mixed)
        nb = StepKernels<float>::diag_partial(c, static_cast<float*>(h->psi), static_cast<float*>(h->psi_prev),
                                              h->rho, static_cast<float*>(h->lpsi),
                                              static_cast<float*>(h->amp), h->partials, s);
      else
        nb = StepKernels<double>::diag_partial(c, static_cast<double*>(h->psi),
                                               static_cast<double*>(h->psi_prev), h->rho,
                                               static_cast<double*>(h->lpsi),
                                               static_cast<double*>(h->amp), h->partials, s);
      diag_final(h->partials, nb, h->nmem, h->diag_out, s);
    }
    HIPTRY(h, hipMemcpyAsync(h->diag_host.data(), h->diag_out, sizeof(double) * 4 * h->nmem,
                             hipMemcpyDeviceToHost, s));
    HIPTRY(h, hipStreamSynchronize(s));
    h->collect_timing();
    h->diag_epoch = h->epoch;
  }
  const double* sums = h->diag_host.data() + 4 * m;
  const smaiso_params& p = h->mp[m];
  const double vol = p.dx * p.dy * p.dz;
  const double a = sums[0] / (p.dt * sums[1]);  // S:2246
  const double ms = sums[2] * vol;               // S:2252
  const double en = sums[3] * vol;               // S:2258
  if (l1) *l1 = a;
  if (mass) *mass = ms;
  if (energy) *energy = en;
  if (!std::isfinite(a) || !std::isfinite(ms) || !std::isfinite(en))
    return fail(h, SMAISO_E_NONFINITE, "non-finite diagnostics (state blew up)");
  return SMAISO_OK;
}

// Streams of a set of groups stepped together: each group works on its own
// stream; `master` orders them against graph launches.  For a single handle
// master is the handle's stream and fork/join are no-ops.
struct Lanes {
  std::vector<smaiso_handle*>& hs;
  hipStream_t master;
  hipEvent_t fork;
  std::vector<hipEvent_t>& joins;
  // group streams wait for master's work so far
  hipError_t fork_out() {
    bool any = false;
    for (smaiso_handle* h : hs) any |= h->stream != master;
    if (!any) return hipSuccess;
    hipError_t e = hipEventRecord(fork, master);
    for (size_t q = 0; q < hs.size() && e == hipSuccess; ++q)
      if (hs[q]->stream != master) e = hipStreamWaitEvent(hs[q]->stream, fork, 0);
    return e;
  }
  // master waits for every group stream's work so far
  hipError_t join_in() {
    hipError_t e = hipSuccess;
    for (size_t q = 0; q < hs.size() && e == hipSuccess; ++q) {
      if (hs[q]->stream == master) continue;
      e = hipEventRecord(joins[q], hs[q]->stream);
      if (e == hipSuccess) e = hipStreamWaitEvent(master, joins[q], 0);
    }
    return e;
  }
};

// Capture two steps of every group into one graph on lanes.master.
int capture_two_steps(smaiso_handle* owner, Lanes& L, hipGraph_t* graph, hipGraphExec_t* exec) {
  HIPTRY(owner, hipStreamBeginCapture(L.master, hipStreamCaptureModeRelaxed));
  int rc = SMAISO_OK;
  hipError_t e = L.fork_out();
  for (int k = 0; k < 2 && rc == SMAISO_OK && e == hipSuccess; ++k)
    for (size_t q = 0; q < L.hs.size() && rc == SMAISO_OK; ++q) {
      rc = run_step_any(L.hs[q]);
      if (rc != SMAISO_OK) owner->err = L.hs[q]->err;
    }
  if (rc == SMAISO_OK && e == hipSuccess) e = L.join_in();
  hipGraph_t gr = nullptr;
  hipError_t e2 = hipStreamEndCapture(L.master, &gr);
  if (rc != SMAISO_OK || e != hipSuccess || e2 != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    if (rc != SMAISO_OK) return rc;
    HIPTRY(owner, e);
    HIPTRY(owner, e2);
  }
  hipGraphExec_t ex = nullptr;
  hipError_t e3 = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
  if (e3 != hipSuccess) {
    (void)hipGraphDestroy(gr);
    HIPTRY(owner, e3);
  }
  *graph = gr;
  *exec = ex;
  return SMAISO_OK;
}

// Advance every group by nsteps: eager steps, except that pairs of steps that
// all take the steady branch replay one captured two-step graph when
// `use_graph` is set.  The work is ordered after master's earlier work and
// master is ordered after all of it on return.
int advance(smaiso_handle* owner, Lanes& L, bool use_graph, hipGraph_t* graph,
            hipGraphExec_t* gexec, int* parity, int nsteps) {
  bool need_fork = true, need_join = false;
  for (int i = 0; i < nsteps;) {
    bool pair_ok = use_graph && nsteps - i >= 2;
    for (smaiso_handle* h : L.hs) pair_ok = pair_ok && h->steady(h->nloop + 1) && h->steady(h->nloop + 2);
    const int par = (int)(L.hs[0]->epoch & 1);
    if (pair_ok && (!*gexec || par == *parity)) {
      if (need_join) { HIPTRY(owner, L.join_in()); need_join = false; }
      if (!*gexec) {
        // capturing records the two steps (and advances the counters)
        int rc = capture_two_steps(owner, L, graph, gexec);
        if (rc != SMAISO_OK) return rc;
        *parity = par;
      } else {
        // the host state the two captured (steady) steps leave
        for (smaiso_handle* h : L.hs) {
          h->nloop += 2;
          h->epoch += 2;
          h->laprl_zlines = h->laprl_in_forces();
          h->gpx_valid = h->gpx_in_psi_pass();
          h->lz_valid = h->lz_in_psi_pass();
          h->fg_combined = forces_zpass_supported(h->mixed ? F32 : F64, h->g.nz) && !h->keep_virr &&
                           h->fg_in_x_pass();
        }
      }
      HIPTRY(owner, hipGraphLaunch(*gexec, L.master));
      need_fork = true;
      i += 2;
      continue;
    }
    if (need_fork) { HIPTRY(owner, L.fork_out()); need_fork = false; }
    for (smaiso_handle* h : L.hs) {
      int rc = run_step_any(h);
      if (rc != SMAISO_OK) { owner->err = h->err; return rc; }
    }
    need_join = true;
    i += 1;
  }
  if (need_join) HIPTRY(owner, L.join_in());
  return SMAISO_OK;
}

}  // namespace

// ===========================================================================
// exported
// ===========================================================================
extern "C" {

int smaiso_create(const smaiso_params* p, smaiso_handle** out) {
  if (!out) return fail(nullptr, SMAISO_E_INVALID, "out is NULL");
  *out = nullptr;
  std::string why;
  int rc = validate_params(p, &why);
  if (rc != SMAISO_OK) return fail(nullptr, rc, why);
  return create_group(p, 1, out);
}

int smaiso_destroy(smaiso_handle* h) {
  delete h;
  return SMAISO_OK;
}

const char* smaiso_last_error(const smaiso_handle* h) {
  return h ? h->err.c_str() : abi_error();
}

int smaiso_set_state(smaiso_handle* h, const double* psi, const double* rho) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!psi || !rho) return fail(h, SMAISO_E_INVALID, "psi/rho is NULL");
  return do_set_state(h, 0, psi, rho, hipMemcpyHostToDevice);
}

int smaiso_set_state_device(smaiso_handle* h, const void* psi, const void* rho) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!psi || !rho) return fail(h, SMAISO_E_INVALID, "psi/rho is NULL");
  return do_set_state(h, 0, static_cast<const double*>(psi), static_cast<const double*>(rho),
                      hipMemcpyDeviceToDevice);
}

int smaiso_step_async(smaiso_handle* h, int32_t nsteps) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!h->all_set()) return fail(h, SMAISO_E_STATE, "smaiso_step before smaiso_set_state");
  if (nsteps < 0) return fail(h, SMAISO_E_INVALID, "nsteps < 0");
  HIPTRY(h, hipSetDevice(h->device));
  std::vector<smaiso_handle*> hs{h};
  std::vector<hipEvent_t> none;
  Lanes L{hs, h->stream, nullptr, none};
  int rc = advance(h, L, h->use_graph && !h->timing, &h->graph, &h->gexec, &h->graph_parity, nsteps);
  if (rc != SMAISO_OK) return rc;
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

int smaiso_set_graph(smaiso_handle* h, int32_t enable) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  h->use_graph = enable != 0;
  return SMAISO_OK;
}

int64_t smaiso_step_count(const smaiso_handle* h) { return h ? h->nloop : -1; }

uint32_t smaiso_fp32_stages_default(void) { return kFp32StagesDefault; }

int smaiso_set_fp32_stages(smaiso_handle* h, uint32_t mask) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (mask & ~SMAISO_STAGES_ALL) return fail(h, SMAISO_E_INVALID, fmt("unknown stage bits 0x%x", mask));
  HIPTRY(h, hipSetDevice(h->device));
  HIPTRY(h, hipStreamSynchronize(h->stream));
  h->fp32_stages = mask;
  // a captured step graph holds the old passes
  if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
  if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
  h->graph_parity = -1;
  return SMAISO_OK;
}

int smaiso_get_field_device(smaiso_handle* h, int32_t field, void* out_dev) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  return get_field(h, 0, field, out_dev, hipMemcpyDeviceToDevice);
}

int smaiso_get_field(smaiso_handle* h, int32_t field, double* out) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  return get_field(h, 0, field, out, hipMemcpyDeviceToHost);
}

int smaiso_diagnostics(smaiso_handle* h, double* l1, double* mass, double* energy) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  return diagnostics(h, 0, l1, mass, energy);
}

// --- ensembles --------------------------------------------------------------
namespace {
int efail(smaiso_ensemble* e, int code, const std::string& msg) {
  if (!e) return abi_fail(code, msg);
  e->err = msg;
  return code;
}
int member_of(smaiso_ensemble* e, int32_t member, smaiso_handle** h, int* m) {
  if (!e) return efail(nullptr, SMAISO_E_INVALID, "ensemble is NULL");
  if (member < 0 || member >= (int32_t)e->where.size())
    return efail(e, SMAISO_E_INVALID, fmt("member %d out of range", member));
  *h = e->groups[e->where[member].first].get();
  *m = e->where[member].second;
  return SMAISO_OK;
}
}  // namespace

int smaiso_ensemble_create(const smaiso_params* members, int32_t count, smaiso_ensemble** out) {
  if (!out) return efail(nullptr, SMAISO_E_INVALID, "out is NULL");
  *out = nullptr;
  if (!members || count < 1) return efail(nullptr, SMAISO_E_INVALID, "need at least one member");
  for (int i = 0; i < count; ++i) {
    std::string why;
    int rc = validate_params(&members[i], &why);
    if (rc != SMAISO_OK) return efail(nullptr, rc, fmt("member %d: %s", i, why.c_str()));
  }
  std::unique_ptr<smaiso_ensemble> e(new smaiso_ensemble());
  // group members that can step in one batch, in order of first appearance
  std::vector<std::vector<int>> groups;
  for (int i = 0; i < count; ++i) {
    size_t q = 0;
    for (; q < groups.size(); ++q)
      if (same_group(members[groups[q][0]], members[i])) break;
    if (q == groups.size()) groups.emplace_back();
    groups[q].push_back(i);
  }
  e->where.resize(count);
  for (size_t q = 0; q < groups.size(); ++q) {
    std::vector<smaiso_params> ps;
    for (size_t k = 0; k < groups[q].size(); ++k) {
      ps.push_back(members[groups[q][k]]);
      e->where[groups[q][k]] = {(int)q, (int)k};
    }
    smaiso_handle* h = nullptr;
    int rc = create_group(ps.data(), (int)ps.size(), &h);
    if (rc != SMAISO_OK) return efail(nullptr, rc, abi_error());
    e->groups.emplace_back(h);
  }
  hipError_t he = hipGetDevice(&e->device);
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&e->master, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
  e->joins.resize(e->groups.size(), nullptr);
  for (auto& j : e->joins)
    if (he == hipSuccess) he = hipEventCreateWithFlags(&j, hipEventDisableTiming);
  if (he != hipSuccess) return efail(nullptr, SMAISO_E_HIP, hipGetErrorString(he));
  *out = e.release();
  return SMAISO_OK;
}

int smaiso_ensemble_destroy(smaiso_ensemble* e) {
  delete e;
  return SMAISO_OK;
}

const char* smaiso_ensemble_last_error(const smaiso_ensemble* e) {
  return e ? e->err.c_str() : abi_error();
}

int smaiso_ensemble_groups(const smaiso_ensemble* e) { return e ? (int)e->groups.size() : -1; }

int smaiso_ensemble_set_state(smaiso_ensemble* e, int32_t member, const double* psi,
                              const double* rho) {
  smaiso_handle* h;
  int m;
  int rc = member_of(e, member, &h, &m);
  if (rc != SMAISO_OK) return rc;
  if (!psi || !rho) return efail(e, SMAISO_E_INVALID, "psi/rho is NULL");
  rc = do_set_state(h, m, psi, rho, hipMemcpyHostToDevice);
  if (rc != SMAISO_OK) return efail(e, rc, h->err);
  return SMAISO_OK;
}

int smaiso_ensemble_set_graph(smaiso_ensemble* e, int32_t enable) {
  if (!e) return efail(nullptr, SMAISO_E_INVALID, "ensemble is NULL");
  e->use_graph = enable != 0;
  return SMAISO_OK;
}

int smaiso_ensemble_step(smaiso_ensemble* e, int32_t nsteps) {
  if (!e) return efail(nullptr, SMAISO_E_INVALID, "ensemble is NULL");
  if (nsteps < 0) return efail(e, SMAISO_E_INVALID, "nsteps < 0");
  for (auto& g : e->groups)
    if (!g->all_set()) return efail(e, SMAISO_E_STATE, "every member needs smaiso_ensemble_set_state");
  hipError_t he = hipSetDevice(e->device);
  if (he != hipSuccess) return efail(e, SMAISO_E_HIP, hipGetErrorString(he));
  std::vector<smaiso_handle*> hs;
  for (auto& g : e->groups) hs.push_back(g.get());
  smaiso_handle* owner = hs[0];
  Lanes L{hs, e->master, e->fork, e->joins};
  int rc = advance(owner, L, e->use_graph, &e->graph, &e->gexec, &e->graph_parity, nsteps);
  if (rc != SMAISO_OK) return efail(e, rc, owner->err);
  he = hipStreamSynchronize(e->master);
  if (he != hipSuccess) return efail(e, SMAISO_E_HIP, hipGetErrorString(he));
  return SMAISO_OK;
}

int64_t smaiso_ensemble_step_count(const smaiso_ensemble* e, int32_t member) {
  if (!e || member < 0 || member >= (int32_t)e->where.size()) return -1;
  return e->groups[e->where[member].first]->nloop;
}

int smaiso_ensemble_get_field(smaiso_ensemble* e, int32_t member, int32_t field, double* out) {
  smaiso_handle* h;
  int m;
  int rc = member_of(e, member, &h, &m);
  if (rc != SMAISO_OK) return rc;
  rc = get_field(h, m, field, out, hipMemcpyDeviceToHost);
  if (rc != SMAISO_OK) return efail(e, rc, h->err);
  return SMAISO_OK;
}

int smaiso_ensemble_diagnostics(smaiso_ensemble* e, int32_t member, double* l1, double* mass,
                                double* energy) {
  smaiso_handle* h;
  int m;
  int rc = member_of(e, member, &h, &m);
  if (rc != SMAISO_OK) return rc;
  rc = diagnostics(h, m, l1, mass, energy);
  if (rc != SMAISO_OK) return efail(e, rc, h->err);
  return SMAISO_OK;
}

// --- standalone transforms ----------------------------------------------------
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


// --- autotuner -----------------------------------------------------------------
namespace {
void drop_graph(hipGraph_t* graph, hipGraphExec_t* gexec, int* parity) {
  if (*gexec) (void)hipGraphExecDestroy(*gexec);
  if (*graph) (void)hipGraphDestroy(*graph);
  *gexec = nullptr;
  *graph = nullptr;
  *parity = -1;
}

// tune the plain passes of the handle's batch geometry in FP64 (the rho
// transforms, and every transform of an FP64 handle) and, for FP32-mixed,
// in FP32
int tune_handle(smaiso_handle* h, float* ms, int32_t* chosen) {
  HIPTRY(h, hipSetDevice(h->device));
  HIPTRY(h, hipStreamSynchronize(h->stream));
  const DType dts[2] = {F64, F32};
  for (int d = 0; d < 2; ++d) {
    float* m = ms ? ms + d * 3 * R2R::kTileVariants : nullptr;
    int c[3] = {0, 0, 0};
    if (d == 1 && !h->mixed) {
      if (m) for (int k = 0; k < 3 * R2R::kTileVariants; ++k) m[k] = -1.0f;
    } else {
      HIPTRY(h, h->r2r.autotune(dts[d], h->g, h->stream, c, m));
    }
    if (chosen) for (int a = 0; a < 3; ++a) chosen[d * 3 + a] = c[a];
  }
  drop_graph(&h->graph, &h->gexec, &h->graph_parity);
  return SMAISO_OK;
}
}  // namespace

int smaiso_autotune(smaiso_handle* h, float* ms, int32_t* chosen) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  return tune_handle(h, ms, chosen);
}

int smaiso_ensemble_autotune(smaiso_ensemble* e) {
  if (!e) return efail(nullptr, SMAISO_E_INVALID, "ensemble is NULL");
  for (auto& g : e->groups) {
    int rc = tune_handle(g.get(), nullptr, nullptr);
    if (rc != SMAISO_OK) return efail(e, rc, g->err);
  }
  drop_graph(&e->graph, &e->gexec, &e->graph_parity);
  return SMAISO_OK;
}

int smaiso_r2r3d_tiles(const int32_t kinds[3], int32_t nx, int32_t ny, int32_t nz, const void* in_dev,
                       void* out_dev, int32_t dtype, int32_t variant, void* stream) {
  int rc = check_r2r_args(kinds, nx, ny, nz, dtype);
  if (rc != SMAISO_OK) return rc;
  if (!in_dev || !out_dev) return fail(nullptr, SMAISO_E_INVALID, "in/out is NULL");
  if (variant < 0 || variant >= R2R::kTileVariants) return fail(nullptr, SMAISO_E_INVALID, "bad tile variant");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(nullptr, SMAISO_E_HIP, hipGetErrorString(e));
  // a private transform object so that the shared one keeps its defaults
  R2R r;
  const Geom g{nx, ny, nz};
  const DType dt = dtype == SMAISO_FP64 ? F64 : F32;
  for (int a = 0; a < 3; ++a) r.set_variant(dt, a, g, variant);
  PassIO io;
  io.in = in_dev;
  io.out = out_dev;
  const int k[3] = {kinds[0], kinds[1], kinds[2]};
  e = r.transform3d(dt, k, io, g, static_cast<hipStream_t>(stream));
  if (e == hipErrorNotSupported) return fail(nullptr, SMAISO_E_INVALID, "tile variant not built for this length");
  if (e == hipSuccess) e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(nullptr, SMAISO_E_HIP, fmt("r2r: %s", hipGetErrorString(e)));
  return SMAISO_OK;
}

}  // extern "C"

// --- adjoint -------------------------------------------------------------------
namespace {

// One stored step input: psiq, psiq_old, rho, f_rho (FP64), the previous
// step's N^l and psi (work precision).
struct StateStore {
  std::vector<void*> mem;
  size_t n3 = 0, tsz = 8;
  int count = 0;
  ~StateStore() {
    for (void* m : mem) (void)hipFree(m);
  }
  hipError_t alloc(int n, long long points, size_t es) {
    count = n;
    n3 = (size_t)points;
    tsz = es;
    for (int k = 0; k < n; ++k) {
      void* m = nullptr;
      hipError_t e = hipMalloc(&m, n3 * (32 + 2 * tsz));
      if (e != hipSuccess) return e;
      mem.push_back(m);
    }
    return hipSuccess;
  }
  double* field(int k, int f) const {  // f: 0 psiq, 1 psiq_old, 2 rho, 3 f_rho
    return static_cast<double*>(mem[k]) + (size_t)f * n3;
  }
  void* nlq(int k) const { return static_cast<char*>(mem[k]) + 32 * n3; }
  void* psi(int k) const { return static_cast<char*>(mem[k]) + (32 + tsz) * n3; }
};

// The forward intermediates that the transposed step reads (AdjArgs), for
// `count` steps: 20 work fields of the storage type, then rho at the step's
// start and after its density update (FP64).
struct StepStore {
  static constexpr int kT = 20;  // psi, amp, lpsi, laprl, divv, gp[3], rd[3], v[3], vtq[3], virrq[3]
  std::vector<void*> mem;
  size_t n3 = 0, tsz = 8;
  int count = 0;
  ~StepStore() { release(); }
  void release() {
    for (void* m : mem) (void)hipFree(m);
    mem.clear();
    count = 0;
  }
  static size_t bytes_per_step(long long points, size_t es) { return (size_t)points * (kT * es + 16); }
  hipError_t alloc(int n, long long points, size_t es) {
    n3 = (size_t)points;
    tsz = es;
    for (int k = 0; k < n; ++k) {
      void* m = nullptr;
      hipError_t e = hipMalloc(&m, bytes_per_step(points, es));
      if (e != hipSuccess) return e;
      mem.push_back(m);
      count = k + 1;
    }
    return hipSuccess;
  }
  void* tf(int k, int f) const { return static_cast<char*>(mem[k]) + (size_t)f * n3 * tsz; }
  double* rho(int k, int which) const {  // which: 0 at the step's start, 1 after it
    return reinterpret_cast<double*>(static_cast<char*>(mem[k]) + kT * n3 * tsz) + (size_t)which * n3;
  }
};

// Copies the intermediates of the step the handle has just run, and its new
// rho, into slot k (the caller stores rho at the step's start).  The step
// ran as any steady step does, so v^ and virr^ are solved again from the
// S^, fgrad^ and erho^ it left (the same solve), and lap(rho L psi) gets its
// last inverse z pass where the forces pass ran it.
hipError_t save_step(smaiso_handle* h, const StepStore& st, int k) {
  const size_t tb = st.n3 * st.tsz, db = st.n3 * 8;
  hipStream_t s = h->stream;
  hipError_t e = h->mixed ? solve_virr<float>(h) : solve_virr<double>(h);
  if (e != hipSuccess) return e;
  h->virr_epoch = h->epoch;
  if (h->laprl_zlines) {
    PassIO io;
    io.in = h->laprl;
    io.out = st.tf(k, 3);
    io.f.arith32 = h->a32(SMAISO_STAGE_LAPRHOL);
    if ((e = h->run_pass(h->mixed ? F32 : F64, REDFT01, 2, io, h->g)) != hipSuccess) return e;
  }
  const void* src[StepStore::kT] = {h->psi_prev, h->amp, h->lpsi, h->laprl, h->divv};
  for (int a = 0; a < 3; ++a) {
    src[5 + a] = h->gp[a];
    src[8 + a] = h->rd[a];
    src[11 + a] = h->v[a];
    src[14 + a] = h->vtq[a];
    src[17 + a] = h->virrq[a];
  }
  for (int f = 0; f < StepStore::kT && e == hipSuccess; ++f)
    if (f != 3 || !h->laprl_zlines) e = hipMemcpyAsync(st.tf(k, f), src[f], tb, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(st.rho(k, 1), h->rho, db, hipMemcpyDeviceToDevice, s);
  return e;
}

hipError_t save_state(smaiso_handle* h, const StateStore& st, int k) {
  const size_t db = st.n3 * 8, tb = st.n3 * st.tsz;
  hipStream_t s = h->stream;
  hipError_t e = hipMemcpyAsync(st.field(k, 0), h->psiq, db, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(st.field(k, 1), h->psiq_old, db, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(st.field(k, 2), h->rho, db, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(st.field(k, 3), h->f_rho, db, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(st.nlq(k), h->nlq, tb, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(st.psi(k), h->psi, tb, hipMemcpyDeviceToDevice, s);
  return e;
}

hipError_t restore_state(smaiso_handle* h, const StateStore& st, int k, long long nloop) {
  const size_t db = st.n3 * 8, tb = st.n3 * st.tsz;
  hipStream_t s = h->stream;
  hipError_t e = hipMemcpyAsync(h->psiq, st.field(k, 0), db, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(h->psiq_old, st.field(k, 1), db, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(h->rho, st.field(k, 2), db, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(h->f_rho, st.field(k, 3), db, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(h->nlq, st.nlq(k), tb, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(h->psi, st.psi(k), tb, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = refresh_rdz(h);
  if (e == hipSuccess) e = h->psi_dx_refresh(-1);
  h->nloop = nloop;
  return e;
}

enum AdjField {
  AF_P, AF_PO, AF_R, AF_FR, AF_NO, AF_NL, AF_PSI, AF_AMP, AF_LAP, AF_LPSI, AF_YBAR, AF_A0BAR,
  AF_TMP, AF_DIVV, AF_DIVF, AF_WFG, AF_WGFG, AF_TARGET,
  AF_GP, AF_RD = AF_GP + 3, AF_V = AF_RD + 3, AF_DOT = AF_V + 3, AF_CFG = AF_DOT + 3,
  AF_DIR = AF_CFG + 3, AF_WFQ = AF_DIR + 3, AF_WSQ = AF_WFQ + 3, AF_COUNT = AF_WSQ + 3
};

int adj_alloc(smaiso_handle* h) {
  if (!h->adj.empty()) return SMAISO_OK;
  const size_t db = (size_t)h->n3() * 8;
  for (int f = 0; f < AF_COUNT; ++f) {
    void* m = nullptr;
    HIPTRY(h, hipMalloc(&m, db));
    h->allocs.push_back(m);
    h->adj.push_back(static_cast<double*>(m));
  }
  const StepConsts c = h->consts();
  void* m = nullptr;
  HIPTRY(h, hipMalloc(&m, sizeof(double) * kAdjSlots * adj_blocks(c)));
  h->allocs.push_back(m);
  h->adj_partials = static_cast<double*>(m);
  HIPTRY(h, hipMalloc(&m, sizeof(double) * 2 * kAdjSlots));
  h->allocs.push_back(m);
  h->adj_acc = static_cast<double*>(m);
  return SMAISO_OK;
}

// The forward intermediates come from the handle's work fields after it has
// recomputed the step (st null) or from slot k of a step store.
template <typename T>
AdjArgs<T> adj_args(smaiso_handle* h, const double* rho_n, bool first, const StepStore* st, int k) {
  AdjArgs<T> A{};
  auto tp = [](void* p) { return static_cast<const T*>(p); };
  if (st) {
    A.psi = tp(st->tf(k, 0));
    A.amp = tp(st->tf(k, 1));
    A.lpsi = tp(st->tf(k, 2));
    A.laprl = tp(st->tf(k, 3));
    A.divv = tp(st->tf(k, 4));
    for (int a = 0; a < 3; ++a) {
      A.gp[a] = tp(st->tf(k, 5 + a));
      A.rd[a] = tp(st->tf(k, 8 + a));
      A.v[a] = tp(st->tf(k, 11 + a));
      A.vtq[a] = tp(st->tf(k, 14 + a));
      A.virrq[a] = tp(st->tf(k, 17 + a));
    }
    A.rho = st->rho(k, 0);
    A.rho_new = st->rho(k, 1);
  } else {
    A.psi = tp(h->psi_prev);  // the step swapped psi and psi_prev
    A.amp = tp(h->amp);
    A.lpsi = tp(h->lpsi);
    A.laprl = tp(h->laprl);
    A.divv = tp(h->divv);
    for (int a = 0; a < 3; ++a) {
      A.gp[a] = tp(h->gp[a]);
      A.rd[a] = tp(h->rd[a]);
      A.v[a] = tp(h->v[a]);
      A.vtq[a] = tp(h->vtq[a]);
      A.virrq[a] = tp(h->virrq[a]);
    }
    A.rho = rho_n;
    A.rho_new = h->rho;
  }
  double* const* f = h->adj.data();
  A.bP = f[AF_P]; A.bPo = f[AF_PO]; A.bR = f[AF_R]; A.bFr = f[AF_FR]; A.bNo = f[AF_NO];
  A.bnl = f[AF_NL]; A.bpsi = f[AF_PSI]; A.bamp = f[AF_AMP]; A.blap = f[AF_LAP];
  A.blpsi = f[AF_LPSI]; A.ybar = f[AF_YBAR]; A.a0bar = f[AF_A0BAR]; A.tmp = f[AF_TMP];
  A.bdivv = f[AF_DIVV]; A.bdivf = f[AF_DIVF]; A.wfg = f[AF_WFG]; A.wGfg = f[AF_WGFG];
  for (int a = 0; a < 3; ++a) {
    A.bgp[a] = f[AF_GP + a]; A.brd[a] = f[AF_RD + a]; A.bv[a] = f[AF_V + a];
    A.dotb[a] = f[AF_DOT + a]; A.cfgb[a] = f[AF_CFG + a]; A.dirb[a] = f[AF_DIR + a];
    A.wfq[a] = f[AF_WFQ + a]; A.wSq[a] = f[AF_WSQ + a];
  }
  A.partials = h->adj_partials;
  A.first = first ? 1 : 0;
  return A;
}

// Transpose of the step that the handle has just recomputed from the stored
// input state (rho at its start is rho_n), or of the step in slot k of st:
// the adjoint of the step's outputs, held in the adjoint state fields,
// becomes that of its inputs.
template <typename T>
int adjoint_step(smaiso_handle* h, const double* rho_n, bool first, const StepStore* st, int k) {
  using K = AdjKernels<T>;
  const StepConsts c = h->consts();
  hipStream_t s = h->stream;
  const AdjArgs<T> A = adj_args<T>(h, rho_n, first, st, k);
  const int nb = (int)adj_blocks(c);
  Fuse fb = h->fuse_base();
  fb.op = OP_PLAIN;
  K::update(c, A, s);
  HIPTRY(h, h->xform(F64, kICT, A.tmp, nullptr, A.bnl));
  K::nl_density(c, A, s);
  for (int a = 0; a < 3; ++a) HIPTRY(h, h->xform(F64, kST[a], A.bv[a], nullptr, A.bv[a]));
  HIPTRY(h, h->xform(F64, kCT, A.bdivv, nullptr, A.bdivv));
  // div f = sum_a D_a(f_a) with one-axis D_a, so f_a bar = D_a^T(div f bar)
  {
    Fuse t = fb;
    t.op = OP_AX_DIV_T;
    for (int a = 0; a < 3; ++a) {
      t.axis = a;
      HIPTRY(h, h->axpass(F64, a, A.bdivf, A.wfq[a], t));
    }
  }
  K::velocity_a(c, A, s);
  K::velocity_b(c, A, s);
  for (int a = 0; a < 3; ++a) HIPTRY(h, h->xform(F64, kIST[a], A.wSq[a], nullptr, A.wSq[a]));
  HIPTRY(h, h->xform(F64, kICT, A.wfg, nullptr, A.wfg));
  HIPTRY(h, h->xform(F64, kICT, A.wGfg, nullptr, A.wGfg));
  K::forces(c, A, s);
  HIPTRY(h, h->xmid_dt(F64, A.blap, fb, OP_MID_NEGLAP, A.ybar));
  // the amplitude filter is symmetric and a product over the axes
  {
    Fuse t = fb;
    t.op = OP_AX_GAUSS;
    for (int a = 2; a >= 0; --a) {
      t.axis = a;
      HIPTRY(h, h->axpass(F64, a, a == 2 ? A.bamp : A.a0bar, A.a0bar, t));
    }
  }
  K::amp_lap(c, A, s);
  // grad_a = one-axis M_a of psi (= ICT(psiq)) and of rho: psi bar and rho bar
  // gain M_a^T(grad_a bar), accumulated by one-axis passes
  {
    Fuse t = fb;
    t.op = OP_AX_GRAD_T;
    t.accumulate = 1;
    for (int a = 0; a < 3; ++a) {
      t.axis = a;
      HIPTRY(h, h->axpass(F64, a, A.bgp[a], A.bpsi, t));
      HIPTRY(h, h->axpass(F64, a, A.brd[a], A.bR, t));
    }
  }
  HIPTRY(h, h->xform(F64, kCT, A.bpsi, nullptr, A.bpsi));
  HIPTRY(h, h->xform(F64, kCT, A.blpsi, nullptr, A.blpsi));
  K::grad(c, A, s);
  adj_reduce(h->adj_partials, 0, nb, h->adj_acc, 0, s);  // dzeta
  adj_reduce(h->adj_partials, 2, nb, h->adj_acc, 0, s);
  adj_reduce(h->adj_partials, 1, nb, h->adj_acc, 1, s);  // dnu
  HIPTRY(h, hipGetLastError());
  return SMAISO_OK;
}

int adjoint_step_any(smaiso_handle* h, const double* rho_n, bool first, const StepStore* st = nullptr,
                     int k = 0) {
  return h->mixed ? adjoint_step<float>(h, rho_n, first, st, k) : adjoint_step<double>(h, rho_n, first, st, k);
}

int run_adjoint(smaiso_handle* h, const double* psi0, const double* rho0, const double* target,
                int nsteps, int every, double* loss, double* dpsi0, double* drho0, double* dzeta,
                double* dnu) {
  if (h->nmem != 1) return fail(h, SMAISO_E_INVALID, "the adjoint runs on single handles");
  if (!(h->p.divv_switch == 0 || h->p.load_mode == 1))
    return fail(h, SMAISO_E_INVALID,
                "the adjoint covers the steady branch only: set divv_switch = 0 or load_mode = 1");
  if (nsteps < 0 || every < 1) return fail(h, SMAISO_E_INVALID, "need nsteps >= 0 and checkpoint_every >= 1");
  HIPTRY(h, hipSetDevice(h->device));
  int rc = adj_alloc(h);
  if (rc != SMAISO_OK) return rc;
  hipStream_t s = h->stream;
  const StepConsts c = h->consts();
  const long long n3 = h->n3();
  const size_t db = (size_t)n3 * 8;
  const int nb = (int)adj_blocks(c);
  double* const* F = h->adj.data();

  rc = do_set_state(h, 0, psi0, rho0, hipMemcpyHostToDevice);
  if (rc != SMAISO_OK) return rc;
  HIPTRY(h, hipMemcpyAsync(F[AF_TARGET], target, db, hipMemcpyHostToDevice, s));
  HIPTRY(h, hipMemsetAsync(h->adj_acc, 0, sizeof(double) * 2 * kAdjSlots, s));

  // forward sweep with a checkpoint every `every` steps
  const int ncp = nsteps > 0 ? (nsteps + every - 1) / every : 0;
  StateStore cps, seg;
  HIPTRY(h, cps.alloc(ncp, n3, h->tsz));
  for (int n = 0; n < nsteps; ++n) {
    if (n % every == 0) HIPTRY(h, save_state(h, cps, n / every));
    rc = run_step_any(h);
    if (rc != SMAISO_OK) return rc;
  }
  // loss = sum (psi_T - target)^2; psiq_T bar = ICT^T(2 (psi_T - target))
  if (h->mixed)
    AdjKernels<float>::loss(c, static_cast<const float*>(h->psi), F[AF_TARGET], F[AF_TMP], h->adj_partials, s);
  else
    AdjKernels<double>::loss(c, static_cast<const double*>(h->psi), F[AF_TARGET], F[AF_TMP], h->adj_partials, s);
  adj_reduce(h->adj_partials, 3, nb, h->adj_acc, 3, s);
  HIPTRY(h, h->xform(F64, kCT, F[AF_TMP], nullptr, F[AF_TMP]));
  adj_div_wc(c, F[AF_TMP], F[AF_P], s);
  for (int f : {AF_PO, AF_R, AF_FR, AF_NO}) fill_zero_d(F[f], n3, s);

  // reverse sweep: recompute each segment from its checkpoint, then
  // transpose its steps last to first.  With a step store the recomputation
  // keeps every step's intermediates (one forward step per transposed
  // step); without one each step is recomputed again from its stored state
  // before its transpose.
  const int seg_len = std::min(every, std::max(nsteps, 1));
  StepStore inter;
  if (h->adj_store && nsteps > 0) {
    size_t free_b = 0, total_b = 0;
    HIPTRY(h, hipMemGetInfo(&free_b, &total_b));
    const size_t need = StepStore::bytes_per_step(n3, h->tsz) * (size_t)seg_len;
    const size_t reserve = (size_t)1 << 30;
    if (need + reserve < free_b && inter.alloc(seg_len, n3, h->tsz) != hipSuccess) {
      (void)hipGetLastError();  // fall back to recomputation
      inter.release();
    }
  }
  const bool stored = inter.count == seg_len;
  if (!stored) HIPTRY(h, seg.alloc(seg_len, n3, h->tsz));
  for (int sg = ncp - 1; stored && sg >= 0; --sg) {
    const int n0 = sg * every, n1 = std::min(n0 + every, nsteps);
    HIPTRY(h, restore_state(h, cps, sg, n0));
    for (int n = n0; n < n1; ++n) {
      HIPTRY(h, hipMemcpyAsync(inter.rho(n - n0, 0), h->rho, db, hipMemcpyDeviceToDevice, s));
      rc = run_step_any(h);  // the forward sweep's step, so the states match it
      if (rc != SMAISO_OK) return rc;
      HIPTRY(h, save_step(h, inter, n - n0));
    }
    for (int n = n1 - 1; n >= n0; --n) {
      rc = adjoint_step_any(h, inter.rho(n - n0, 0), n == 0, &inter, n - n0);
      if (rc != SMAISO_OK) return rc;
    }
  }
  for (int sg = ncp - 1; !stored && sg >= 0; --sg) {
    const int n0 = sg * every, n1 = std::min(n0 + every, nsteps);
    HIPTRY(h, restore_state(h, cps, sg, n0));
    for (int n = n0; n < n1; ++n) {
      HIPTRY(h, save_state(h, seg, n - n0));
      if (n + 1 < n1) {
        rc = run_step_any(h);
        if (rc != SMAISO_OK) return rc;
      }
    }
    for (int n = n1 - 1; n >= n0; --n) {
      HIPTRY(h, restore_state(h, seg, n - n0, n));
      h->keep_virr = true;  // the transposed velocity solve reads virr^
      rc = run_step_any(h);
      h->keep_virr = false;
      if (rc != SMAISO_OK) return rc;
      rc = adjoint_step_any(h, seg.field(n - n0, 2), n == 0);
      if (rc != SMAISO_OK) return rc;
    }
  }
  // set_state: psiq = scale CT(psi0), psiq_old = psiq, rho = rho0
  adj_psi0(c, F[AF_P], F[AF_PO], F[AF_TMP], s);
  HIPTRY(h, h->xform(F64, kICT, F[AF_TMP], nullptr, F[AF_TMP]));
  double acc[2 * kAdjSlots];
  HIPTRY(h, hipMemcpyAsync(acc, h->adj_acc, sizeof(acc), hipMemcpyDeviceToHost, s));
  if (dpsi0) HIPTRY(h, hipMemcpyAsync(dpsi0, F[AF_TMP], db, hipMemcpyDeviceToHost, s));
  if (drho0) HIPTRY(h, hipMemcpyAsync(drho0, F[AF_R], db, hipMemcpyDeviceToHost, s));
  HIPTRY(h, hipStreamSynchronize(s));
  h->collect_timing();
  // the stepper holds an intermediate state now
  h->member_set[0] = 0;
  h->diag_epoch = -1;
  h->derived = false;
  h->nlq_valid = false;
  if (loss) *loss = acc[6] + acc[7];
  if (dzeta) *dzeta = acc[0] + acc[1];
  if (dnu) *dnu = acc[2] + acc[3];
  return SMAISO_OK;
}

}  // namespace

extern "C" {

int smaiso_adjoint(smaiso_handle* h, const double* psi0, const double* rho0, const double* target,
                   int32_t nsteps, int32_t checkpoint_every, double* loss, double* dpsi0,
                   double* drho0, double* dzeta, double* dnu) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!psi0 || !rho0 || !target) return fail(h, SMAISO_E_INVALID, "psi0/rho0/target is NULL");
  return run_adjoint(h, psi0, rho0, target, nsteps, checkpoint_every, loss, dpsi0, drho0, dzeta, dnu);
}

int smaiso_set_adjoint_store(smaiso_handle* h, int32_t enable) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  h->adj_store = enable != 0;
  return SMAISO_OK;
}


// --- output writers' data, checkpoints, the reference's psi files -------------
int smaiso_get_slice(smaiso_handle* h, int32_t field, int32_t j, double* out) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!out) return fail(h, SMAISO_E_INVALID, "out is NULL");
  if (j < 0 || j >= h->g.ny) return fail(h, SMAISO_E_INVALID, fmt("slice j=%d outside 0..%d", j, h->g.ny - 1));
  if (!h->all_set()) return fail(h, SMAISO_E_STATE, "no state (call smaiso_set_state first)");
  HIPTRY(h, hipSetDevice(h->device));
  const double* src = nullptr;
  int rc = field_device(h, field, &src);
  if (rc != SMAISO_OK) return rc;
  const size_t nz = (size_t)h->g.nz, ny = (size_t)h->g.ny;
  // saveMid (S:94-129): out[i*Nz + k] = field[(i*Ny + j)*Nz + k]
  HIPTRY(h, hipMemcpy2DAsync(out, nz * 8, src + (size_t)j * nz, ny * nz * 8, nz * 8, (size_t)h->g.nx,
                             hipMemcpyDeviceToHost, h->stream));
  HIPTRY(h, hipStreamSynchronize(h->stream));
  h->collect_timing();
  return SMAISO_OK;
}

int smaiso_surface_height(smaiso_handle* h, double threshold, double* surf) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!surf) return fail(h, SMAISO_E_INVALID, "surf is NULL");
  if (!h->all_set()) return fail(h, SMAISO_E_STATE, "no state (call smaiso_set_state first)");
  HIPTRY(h, hipSetDevice(h->device));
  const size_t cols = (size_t)h->g.nx * h->g.ny;
  if (!h->surf_dev) {
    void* m = nullptr;
    HIPTRY(h, hipMalloc(&m, cols * sizeof(int)));
    h->allocs.push_back(m);
    h->surf_dev = static_cast<int*>(m);
  }
  surface_height(h->rho, h->g.nx, h->g.ny, h->g.nz, threshold, h->surf_dev, h->stream);
  HIPTRY(h, hipGetLastError());
  std::vector<int> k(cols);
  HIPTRY(h, hipMemcpyAsync(k.data(), h->surf_dev, cols * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPTRY(h, hipStreamSynchronize(h->stream));
  // a column without a qualifying k keeps its previous value (S:2284-2293)
  for (size_t c = 0; c < cols; ++c)
    if (k[c] >= 0) surf[c] = (double)k[c];
  return SMAISO_OK;
}

int smaiso_save_checkpoint(smaiso_handle* h, const char* path) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!path) return fail(h, SMAISO_E_INVALID, "path is NULL");
  if (h->nmem != 1) return fail(h, SMAISO_E_INVALID, "checkpoints are per single handle");
  if (!h->all_set()) return fail(h, SMAISO_E_STATE, "no state (call smaiso_set_state first)");
  HIPTRY(h, hipSetDevice(h->device));
  HIPTRY(h, hipStreamSynchronize(h->stream));
  CkptHeader hd;
  ckpt_header_init(&hd, h->g.nx, h->g.ny, h->g.nz, h->p.precision, (int)h->tsz, h->nloop);
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(h, SMAISO_E_INVALID, fmt("cannot open %s for writing", path));
  const size_t n3 = (size_t)h->n3();
  std::vector<char> buf(n3 * 8);
  bool ok = std::fwrite(&hd, sizeof(hd), 1, f) == 1;
  const void* arrays[6] = {h->psiq, h->psiq_old, h->rho, h->f_rho, h->psi, h->nlq};
  for (int a = 0; a < 6 && ok; ++a) {
    const size_t bytes = n3 * (a < 4 ? 8 : h->tsz);
    if (hipMemcpy(buf.data(), arrays[a], bytes, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
    else ok = std::fwrite(buf.data(), 1, bytes, f) == bytes;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(h, SMAISO_E_HIP, fmt("writing checkpoint %s failed", path));
  return SMAISO_OK;
}

int smaiso_load_checkpoint(smaiso_handle* h, const char* path) {
  if (!h) return fail(nullptr, SMAISO_E_INVALID, "handle is NULL");
  if (!path) return fail(h, SMAISO_E_INVALID, "path is NULL");
  if (h->nmem != 1) return fail(h, SMAISO_E_INVALID, "checkpoints are per single handle");
  HIPTRY(h, hipSetDevice(h->device));
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(h, SMAISO_E_INVALID, fmt("cannot open %s", path));
  CkptHeader hd;
  std::string why;
  if (ckpt_header_read(f, h->g.nx, h->g.ny, h->g.nz, h->p.precision, (int)h->tsz, &hd, &why) != SMAISO_OK) {
    std::fclose(f);
    return fail(h, SMAISO_E_INVALID, fmt("%s: %s", path, why.c_str()));
  }
  HIPTRY(h, hipStreamSynchronize(h->stream));
  // from here on the handle's state is being overwritten: it counts as set
  // again only when the whole file has been read
  h->member_set[0] = 0;
  h->derived = false;
  h->nlq_valid = false;
  h->diag_epoch = -1;
  const size_t n3 = (size_t)h->n3();
  std::vector<char> buf(n3 * 8);
  void* arrays[6] = {h->psiq, h->psiq_old, h->rho, h->f_rho, h->psi, h->nlq};
  bool ok = true;
  for (int a = 0; a < 6 && ok; ++a) {
    const size_t bytes = n3 * (a < 4 ? 8 : h->tsz);
    ok = std::fread(buf.data(), 1, bytes, f) == bytes &&
         hipMemcpy(arrays[a], buf.data(), bytes, hipMemcpyHostToDevice) == hipSuccess;
  }
  std::fclose(f);
  if (!ok) return fail(h, SMAISO_E_INVALID, fmt("%s is truncated", path));
  HIPTRY(h, refresh_rdz(h));
  HIPTRY(h, h->psi_dx_refresh(-1));
  h->nloop = hd.nloop;
  h->member_set[0] = 1;
  h->nlq_valid = true;  // the checkpoint holds the last step's N^l
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
