// Per-mode (spectral) and pointwise (real-space) kernels of one time step.
// Each launcher enqueues one kernel on stream s.  T is the storage type of the
// work fields (double for SMAISO_FP64, float for SMAISO_FP32_MIXED); the
// time-integrated state psiq, psiq_old, rho and f_rho is always double.
#pragma once

#include "common.h"

namespace smaiso {

// Scalars and wavenumber tables shared by every kernel of a step.
struct StepConsts {
  int nx, ny, nz;
  int slab_n0;        // reference slab height for quirk Q3
  double scale;       // 1/(8 Nx Ny Nz), S:278
  double dt, alpha, beta, gamma, eps, q0, q02, nu, lam, zeta, rho0, kp, amp_eq, rho_m;
  const double* vq[3];  // DCT-lattice wavenumbers pi n/(d N), S:447-463
  const double* vs[3];  // DST-lattice wavenumbers pi (n+1)/(d N), S:465-481
};

enum SpecOp : int {
  SPEC_SCALE = 0,        // out = scale * in                         (pre-loop, S:982)
  SPEC_AMP_SMOOTH = 1,   // out = scale * in * G(|q|^2)              (S:1279-1282)
  SPEC_NEGLAP = 2,       // out = -scale |q|^2 in                    (S:1334-1337)
};

enum PsiqOp : int {
  PSIQ_SH = 0,           // out = alpha (q0^2 - |q|^2) psiq          (S:1300-1301)
  PSIQ_RELAXLAP = 1,     // out = -|q|^2 alpha (q0^2 - |q|^2) psiq   (S:1356-1357)
};

template <typename T>
struct StepKernels {
  // a6: spectral gradients of psiq (normalised) and rhoq (unnormalised DCT of rho)
  static void grad(const StepConsts& c, const double* psiq, const T* rhoq, T* gpx, T* gpy, T* gpz,
                   T* rdx, T* rdy, T* rdz, hipStream_t s);
  // a7: |amp| = sqrt(q0^2 psi^2 + |grad psi|^2)
  static void amp(const StepConsts& c, const T* psi, const T* gx, const T* gy, const T* gz, T* amp,
                  hipStream_t s);
  static void spec(const StepConsts& c, int op, const T* in, T* out, hipStream_t s);
  static void psiq_op(const StepConsts& c, int op, const double* psiq, T* out, hipStream_t s);
  // w = rho L psi
  static void prod(const StepConsts& c, const T* a, const T* b, T* out, hipStream_t s);
  // a9 relaxation branch: rho = rho0 + (1-rho0)/Amp amp; laprl *= rho
  static void relax(const StepConsts& c, const T* amp, T* laprl, double* rho, T* rho_t,
                    hipStream_t s);
  // a10
  static void forces(const StepConsts& c, const T* lpsi, const T* laprl, const T* rho,
                     const T* amp, const T* psi, const T* gpx, const T* gpy, const T* gpz,
                     const T* rdx, const T* rdy, const T* rdz, T* fgrad, T* erho, T* mu, T* sx,
                     T* sy, T* sz, T* fx, T* fy, T* fz, hipStream_t s);
  // a11 (fgrad += G erho) + a12
  static void velocity(const StepConsts& c, const T* sxq, const T* syq, const T* szq,
                       const T* fgq, const T* erq, T* vsolx, T* vsoly, T* vsolz, T* virrx,
                       T* virry, T* virrz, hipStream_t s);
  // a13
  static void divergence(const StepConsts& c, const T* virrx, const T* virry, const T* virrz,
                         const T* fxq, const T* fyq, const T* fzq, T* divv, T* divf,
                         hipStream_t s);
  // a15 + a16
  static void density_nl(const StepConsts& c, const T* vx, const T* vy, const T* vz,
                         const T* rdx, const T* rdy, const T* rdz, const T* divv, double* rho,
                         double* f_rho, T* rho_t, const T* mu, const T* amp, const T* psi,
                         const T* divf, const T* gpx, const T* gpy, const T* gpz, T* nl,
                         hipStream_t s);
  // a17: psiq_new written over psiq_old; first != 0 uses nlq for nlq_old (quirk Q5)
  static void psi_update(const StepConsts& c, const double* psiq, double* psiq_old_to_new,
                         const T* nlq, const T* nlq_old, int first, T* psiq_t, hipStream_t s);
  // free-energy density (S:2227-2233) for field readout
  static void energy_density(const StepConsts& c, const T* psi, const T* lpsi, const T* rho,
                             const T* amp, T* out, hipStream_t s);
  // a18: 4 block-partial compensated sums -> partials[4*nblocks]; returns nblocks
  static int diag_partial(const StepConsts& c, const T* psi, const T* psi_old, const double* rho,
                          const T* lpsi, const T* amp, double* partials, hipStream_t s);
};

// sum partials[4*nb] -> out[4] with compensated summation (one workgroup)
void diag_final(const double* partials, int nb, double* out, hipStream_t s);
// element-wise conversions / fills
void convert_d2f(const double* in, float* out, long long n, hipStream_t s);
void convert_f2d(const float* in, double* out, long long n, hipStream_t s);
void fill_zero_d(double* out, long long n, hipStream_t s);

}  // namespace smaiso
