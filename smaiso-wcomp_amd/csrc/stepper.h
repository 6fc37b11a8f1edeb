// Per-mode (spectral) and pointwise (real-space) kernels of one time step
// that are not fused into a transform pass (the fused ones live in r2r.hip).
// Each launcher enqueues one kernel on stream s.  T is the storage type of the
// work fields (double for SMAISO_FP64, float for SMAISO_FP32_MIXED); the
// time-integrated state psiq, psiq_old, rho and f_rho is always double, and the
// pointwise kernels evaluate their expressions in FP64 registers.
//
// Every kernel works on a batch of `nmem` members of one grid shape: member m
// of a field lives at offset m * Nx*Ny*Nz and blockIdx.y is the member.
#pragma once

#include "common.h"

namespace smaiso {

// Parameters that may differ between the members of a batch.
struct MemberConsts {
  double nu, lam, zeta, rho0, kp, amp_eq;
  double rnu, rlam;  // 1/nu and 1/(lam + 2 nu), correctly rounded on the host
};

// Scalars and wavenumber tables shared by every kernel of a step.
struct StepConsts {
  int nx, ny, nz;
  int nmem;             // members in the batch
  int slab_n0;          // reference slab height for quirk Q3
  double scale;         // 1/(8 Nx Ny Nz), S:278
  double dt, alpha, beta, gamma, eps, q0, q02, rho_m;
  const double* vq[3];  // DCT-lattice wavenumbers pi n/(d N), S:447-463
  const double* vs[3];  // DST-lattice wavenumbers pi (n+1)/(d N), S:465-481
  double qs[3];         // their spacing pi/(d N): vq[n] ~ n qs, vs[n] ~ (n+1) qs (within an ulp)
  // one-axis factors of the Gaussian exp(-1.57*1.57*2 |q|^2) (S:1281 and the
  // velocity solve's exponentials): gq_a[n] = exp(-1.57*1.57*2 vq_a[n]^2),
  // gs_a[n] the same of vs_a[n]; the 3-D factor is the product over the axes
  // (the adjoint's kernels take it so; the forward velocity solve evaluates exp)
  const double* gq[3];
  const double* gs[3];
  const MemberConsts* mc;  // device array, nmem entries
};

enum SpecOp : int {
  SPEC_SCALE = 0,        // out = scale * in                         (pre-loop, S:982)
};

template <typename T>
struct StepKernels {
  static void spec(const StepConsts& c, int op, const T* in, T* out, hipStream_t s);
  // a9 relaxation branch: rho = rho0 + (1-rho0)/Amp amp; laprl *= rho
  static void relax(const StepConsts& c, const T* amp, T* laprl, double* rho, hipStream_t s);
  // a10 (FP64 arithmetic)
  static void forces(const StepConsts& c, const T* lpsi, const T* laprl, const double* rho,
                     const T* amp, const T* psi, const T* gpx, const T* gpy, const T* gpz,
                     const T* rdx, const T* rdy, const T* rdz, T* fgrad, T* erho, T* mu, T* sx,
                     T* sy, T* sz, T* fx, T* fy, T* fz, hipStream_t s);
  // a11 (fgrad += G erho) + a12 + the spectral divergence of a13: the total
  // velocity vsol^ + virr^ (the input of v's inverse transforms, summed before
  // the store), the irrotational part virr^ when virrx is not null, and
  // div v^ (the input of div v's inverse transform) when divq is not null;
  // FP64 arithmetic, FP32 for FP32 storage when arith32
  static void velocity(const StepConsts& c, const T* sxq, const T* syq, const T* szq,
                       const T* fgq, const T* erq, T* vx, T* vy, T* vz, T* virrx,
                       T* virry, T* virrz, T* divq, bool arith32, hipStream_t s);
  // a15 + a16 (FP64 arithmetic)
  static void density_nl(const StepConsts& c, const T* vx, const T* vy, const T* vz,
                         const T* rdx, const T* rdy, const T* rdz, const T* divv, double* rho,
                         double* f_rho, const T* mu, const T* amp, const T* psi,
                         const T* divf, const T* gpx, const T* gpy, const T* gpz, T* nl,
                         hipStream_t s);
  // free-energy density (S:2227-2233) for field readout
  static void energy_density(const StepConsts& c, const T* psi, const T* lpsi, const double* rho,
                             const T* amp, T* out, hipStream_t s);
  // a18: per member, block-partial compensated sums of the 4 diagnostics ->
  // partials[member][block][4][2]; returns the number of blocks per member
  static int diag_partial(const StepConsts& c, const T* psi, const T* psi_old, const double* rho,
                          const T* lpsi, const T* amp, double* partials, hipStream_t s);
};

// sum partials -> out[member][4] with compensated summation (one workgroup per member)
void diag_final(const double* partials, int nb, int nmem, double* out, hipStream_t s);
// element-wise conversions / fills over n elements
void convert_d2f(const double* in, float* out, long long n, hipStream_t s);
void convert_f2d(const float* in, double* out, long long n, hipStream_t s);
void fill_zero_d(double* out, long long n, hipStream_t s);
// out = a - b over n elements of type T (float or double)
template <typename T>
void sub_fields(const T* a, const T* b, T* out, long long n, hipStream_t s);
// Surface tracker of S:2278-2318 for one nx x ny x nz field rho: for each
// column (i, j), the largest k with rho > threshold, or -1 when there is none
// (the reference keeps the column's previous value then; the caller merges).
void surface_height(const double* rho, int nx, int ny, int nz, double threshold, int* out,
                    hipStream_t s);

// Largest number of diagnostic blocks per member.
constexpr int kDiagBlocks = 1024;

}  // namespace smaiso
