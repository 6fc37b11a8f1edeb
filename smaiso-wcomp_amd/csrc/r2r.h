// Batched 1-D and 3-D real-to-real transforms (DCT-II/III, DST-II/III) on one
// MI355X, with FFTW's unnormalised REDFT10/REDFT01/RODFT10/RODFT01 semantics.
//
// These replace the reference's eight fftw_mpi_plan_r2r_3d plans
// (smaiso-wcomp.cpp:499-545).  A 3-D transform is three axis passes over HBM
// (z, then y, then x); each pass moves a tile of lines through LDS, so the
// x-slab MPI transposes of the reference become strided tile loads on one GPU.
#pragma once

#include "common.h"
#include <map>
#include <tuple>
#include <mutex>
#include <vector>

namespace smaiso {

enum DType : int { F64 = 0, F32 = 1 };

// Operations fused into an axis pass.  Load ops replace the element loads of
// the first (z) pass; the mid ops make an x pass compute forward REDFT10,
// multiply every mode by a diagonal operator and transform back with REDFT01;
// OP_LD_PSI makes the z pass of the inverse transform of psi compute the psi
// update on load and store the new coefficients on the way.
enum FuseOp : int {
  OP_PLAIN = 0,        // x = in (+ in2)
  OP_LD_GRAD_D = 1,    // x = -c0 Vs_a[n_a] src[n + e_a]   (src double, 0 at n_a = N_a-1)
  OP_LD_GRAD_T = 2,    // same with src of the work type
  OP_LD_AMP = 3,       // x = sqrt(q0^2 s0^2 + s1^2 + s2^2 + s3^2)
  OP_LD_SH = 4,        // x = alpha (q0^2 - |q|^2) psiq              (psiq double)
  OP_LD_RELAXLAP = 5,  // x = -|q|^2 alpha (q0^2 - |q|^2) psiq
  OP_LD_PROD = 6,      // x = s0 s1                                  (s0 double)
  OP_MID_AMP = 8,      // mode *= scale exp(-1.57*1.57*|q|^2*2)
  OP_MID_NEGLAP = 9,   // mode *= -scale |q|^2
  OP_LD_PSI = 10,      // x = new psiq from psiq, psiq_new (= old), Nl^ (src[0]), Nl^_old; stored to psiq_new
  // One-axis operators: forward 1-D transform along the pass axis, an operator
  // on the modes, inverse 1-D transform, in one pass; c = c0/(2N).  They are
  // the 3-D spectral operators of the reference restricted to the one axis
  // they act on (iDCT.DCT = 2N along every other axis).
  OP_AX_GRAD = 11,       // out = iDST(y), y[n] = -c Vs[n] DCT(in)[n+1] (0 at n = N-1)
  OP_AX_DIV = 12,        // out (+)= iDCT(y), y[n] = c Vq[n] DST(in)[n-1] (0 at n = 0)
  OP_AX_GAUSS = 13,      // out = iDCT(c exp(-1.57^2 2 Vq[n]^2) DCT(in))
  OP_AX_GAUSS_AMP = 14,  // OP_AX_GAUSS (z only) of the amplitude computed on load (OP_LD_AMP)
  // transposes (for the adjoint): REDFT10^T = REDFT01 diag(2,1,..,1) and
  // RODFT10^T = RODFT01 diag(1,..,1,2) make them one-axis operators too
  OP_AX_GRAD_T = 15,     // out (+)= iDCT(y), y[n] = -c Vs[n-1] DST(in)[n-1] (0 at n = 0)
  OP_AX_DIV_T = 16,      // out (+)= iDST(y), y[n] = c Vq[n+1] DCT(in)[n+1] (0 at n = N-1)
  OP_COUNT = 17,
};

inline bool is_axis_op(int op) {
  return op == OP_AX_GRAD || op == OP_AX_DIV || op == OP_AX_GAUSS || op == OP_AX_GAUSS_AMP ||
         op == OP_AX_GRAD_T || op == OP_AX_DIV_T;
}

struct Fuse {
  int op = OP_PLAIN;
  const void* src[4] = {nullptr, nullptr, nullptr, nullptr};
  int axis = 0;          // shift axis of OP_LD_GRAD_*
  double c0 = 1.0;       // multiplier of OP_LD_GRAD_*
  const double* vq[3] = {nullptr, nullptr, nullptr};  // DCT-lattice wavenumbers
  const double* vs[3] = {nullptr, nullptr, nullptr};  // DST-lattice wavenumbers
  // wavenumber spacing pi/(d_a N_a) per axis: vq_a[n] = pi n/(d_a N_a), vs_a[n]
  // = pi (n+1)/(d_a N_a) (S:447-481); when set the passes evaluate n qs_a
  // (within an ulp of the tables) instead of loading the tables (0: use the
  // tables)
  double qs[3] = {0, 0, 0};
  double q02 = 0, alpha = 0, scale = 0, dt = 0, rho_m = 0, eps = 0;
  // OP_LD_PSI
  const double* psiq = nullptr;
  double* psiq_new = nullptr;     // holds psiq_old on entry, overwritten in place
  const void* nlq_old = nullptr;
  int first = 0;                  // first step: Nl_old = Nl
  int accumulate = 0;             // OP_AX_DIV: add the result to out
  // FP32 storage only: compute the pass in FP32 (FP32 tile and tables)
  // where the length has such instances (r2r_arith32_supported); otherwise,
  // and for FP64 storage, the pass computes in FP64.  Also the gradient pass
  // from FP64 to FP32 storage along y or x at length 256 (R2R::pass_cast).
  int arith32 = 0;
  // OP_AX_GRAD from FP64 input in FP32 arithmetic: subtracted from the input
  // before it is rounded to the FP32 tile.  The operator never reads mode 0,
  // so a constant leaves the result unchanged in exact arithmetic while a
  // constant near the field's mean shrinks the rounding of the tile.
  double ld_shift = 0;
  // plain passes: the default tile even where the autotuner chose another
  // (passes whose results a later computation must reproduce bit for bit)
  int untuned = 0;
};

// Lengths with FP32-arithmetic instances of the transform passes.
inline bool r2r_arith32_supported(int n) { return n == 64 || n == 96 || n == 128 || n == 256 || n == 512; }

struct PassIO {
  const void* in = nullptr;
  const void* in2 = nullptr;   // added element-wise to `in` on load when set (OP_PLAIN)
  void* out = nullptr;
  Fuse f;
};

class R2R {
 public:
  R2R() = default;
  ~R2R();
  // One axis pass: out = kind_axis(in (+ in2)).  in may alias out.
  hipError_t pass(DType dt, int kind, int axis, const PassIO& io, const Geom& g,
                  hipStream_t s);
  // Full 3-D transform with kinds[a] on axis a; passes run z, y, x.  A load
  // op in io.f applies to the z pass.
  hipError_t transform3d(DType dt, const int kinds[3], const PassIO& io, const Geom& g,
                         hipStream_t s);
  // One axis pass that reads in_dt and stores out_dt: FP32 -> FP64 in a z
  // pass (OP_PLAIN or OP_LD_PSI, whose work-type sources are then FP32),
  // FP64 -> FP32 in a y or x pass (OP_PLAIN); hipErrorNotSupported for the
  // other directions.  With in_dt == out_dt it is pass().  A 3-D transform
  // between FP32 fields made of these keeps its intermediates in FP64.
  hipError_t pass_cast(DType in_dt, DType out_dt, int kind, int axis, const PassIO& io,
                       const Geom& g, hipStream_t s);
  // True when length n runs on the LDS FFT path (else the direct O(n^2) path).
  static bool fft_supported(int n);
  // Build the twiddle tables of length n now (they are otherwise built on
  // first use, with synchronous copies that a stream capture cannot hold):
  // FP64, and with fp32 also the FP32 ones of the FP32-arithmetic passes.
  hipError_t prepare(int n, bool fp32 = false) {
    Tables* t = nullptr;
    hipError_t e = tables(F64, n, &t);
    if (e == hipSuccess && fp32 && r2r_arith32_supported(n)) e = tables(F32, n, &t);
    return e;
  }
  // Device twiddle tables of length n (FP64 complex; the passes compute in
  // FP64 for both storage types): tw[m] = exp(-2 pi i m/n), mk[m] =
  // exp(-i pi m/(2n)); for kernels outside R2R that run an FFT pass of their
  // own.
  hipError_t table_ptrs(int n, const void** tw, const void** mk, DType dt = F64) {
    Tables* t = nullptr;
    hipError_t e = tables(dt, n, &t);
    if (e != hipSuccess) return e;
    *tw = t->tw;
    *mk = t->mk;
    return hipSuccess;
  }

  // Autotuner: for each axis of geometry g whose length runs on the FFT
  // path, time a forward+inverse plain pass with every tile variant
  // (kTileVariants, variant 0 = the built-in default) on scratch buffers and
  // keep the fastest; later plain passes of (dt, axis, g) use it.  ms
  // (optional, 3 x kTileVariants) receives the timings, -1 where a variant
  // is not built for the length; chosen (optional, 3) the variants kept.
  // Not capturable (allocates and synchronises).
  hipError_t autotune(DType dt, const Geom& g, hipStream_t s, int* chosen, float* ms);
  // Use tile variant v for plain passes of (dt, axis, g) (0 = default).
  void set_variant(DType dt, int axis, const Geom& g, int v);
  int variant(DType dt, int axis, const Geom& g) const;
  static constexpr int kTileVariants = 4;

 private:
  using TuneKey = std::tuple<int, int, int, int, int, int>;  // dtype, axis, nx, ny, nz, nb
  static TuneKey tune_key(DType dt, int axis, const Geom& g) {
    return TuneKey((int)dt, axis, g.nx, g.ny, g.nz, g.nb);
  }
  std::map<TuneKey, int> tuned_;
  struct Tables {
    void* tw = nullptr;      // n complex: exp(-2 pi i m / n)
    void* mk = nullptr;      // n complex: exp(-i pi m / (2n))
    void* costab = nullptr;  // 4n reals: cos(pi t / (2n))
  };
  hipError_t tables(DType dt, int n, Tables** out);
  std::map<std::pair<int, int>, Tables> cache_;
  std::mutex mu_;
};

// Host values of the transform tables of length n, as R2R uploads them:
// tw[2m], tw[2m+1] = cos, sin(-2 pi m/n) and mk likewise of -pi m/(2n) for
// m < n (2n doubles each); ct[t] = cos(pi t/(2n)) for t < 4n.
void r2r_table_values(int n, double* tw, double* mk, double* ct);

// One plain axis pass (io.f.op must be OP_PLAIN) with the default tiles and
// device tables the caller provides (the values of r2r_table_values for the
// pass axis' length): the same kernels as R2R::pass without R2R's table
// cache, so it allocates nothing.
hipError_t r2r_plain_pass_with_tables(DType dt, int kind, int axis, const PassIO& io, const Geom& g,
                                      const void* tw, const void* mk, const void* ct, hipStream_t s);

// The x pass of an inverse transform (DCT-III along x) of FP64 x-spectral
// lines `in` that stores the result to out and its x derivative (the
// gradient one-axis operator: DST-III of -Vs[n] in[n+1], qs = pi/(dx Nx)) to
// out2, both in out_dt, FP64 arithmetic; length 256 along x
// (yx256_invgrad_supported).  Defined in r2r_yx256.hip.
bool yx256_invgrad_supported(const Geom& g);
// The x pass of the forward DCTs of fgrad and e rho at length 256 (FP64
// fields, after their z and y passes) that stores fgrad^ + G e rho^ into out
// (may alias fg): G = exp(-1.57^2 2 |q|^2) with |q|^2 from the spacings qs.
bool yx256_fgx_supported(const Geom& g);
hipError_t launch_yx256_fgx(const double* fg, const double* er, double* out, const double qs[3], const Geom& g,
                            const void* tw, const void* mk, hipStream_t s);
hipError_t launch_yx256_invgrad(DType out_dt, const double* in, void* out, void* out2, double qs, const Geom& g,
                                const void* tw, const void* mk, hipStream_t s);

}  // namespace smaiso
