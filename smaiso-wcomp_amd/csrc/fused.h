// Time-step stages fused into the z pass of the transforms that consume
// them (kernels in fused.hip).  A pointwise stage that writes fields which
// are transformed next costs one write and one read of each such field; when
// the z pass computes the stage on load, those fields go from registers to
// LDS instead, and the pass writes their z-transformed lines in place of the
// real-space values.  The y and x passes of the transforms follow as usual.
#pragma once

#include "r2r.h"
#include "stepper.h"

namespace smaiso {

// a10 (smaiso-wcomp.cpp:1385-1425) fused with the z passes of a11
// (S:1427-1457): fgrad and erho (REDFT10 along z) and S_x, S_y (REDFT10),
// S_z (RODFT10).
template <typename T>
struct ForcesIO {
  const T *lpsi, *laprl, *amp, *psi;
  const T* gp[3];      // grad psi
  const double* rho;
  const T* rd[3];       // grad rho
  T* mu;                // chemical potential (real space)
  T* f[3];              // forces f_a (real space)
  T* zout[5];           // z pass of fgrad, erho, S_x, S_y, S_z
  // laprl holds the z lines of lap(rho L psi) before the last inverse z pass
  // of its transform (REDFT01), which the kernel then runs in its own tile
  // (FP64 storage: with FP32 storage the kernel is less memory-bound and the
  // extra phase measured slower than the separate pass)
  bool laprl_zlines = false;
};

// hipErrorNotSupported when nz has no fused kernel (a length off the FFT
// path, or an LDS tile above 64 KiB); the caller then runs the unfused stages.
template <typename T>
hipError_t forces_zpass(const StepConsts& c, const ForcesIO<T>& io, R2R& r2r, hipStream_t s);
bool forces_zpass_supported(DType dt, int nz);
// the forces pass runs lap(rho L psi)'s last inverse z pass (laprl_zlines)
bool forces_zpass_laprl(DType dt, int nz);

// a15 + a16 (S:2121-2162) fused with the z pass of the nonlinear term's
// transform (S:2166-2168, REDFT10 along z): the density update (rho and
// f_rho in place, FP64) and Nl evaluated on load, Nl's z-transformed lines
// stored to nlz.
template <typename T>
struct DensityIO {
  const T* v[3];        // total velocity (real space)
  const T* rd[3];       // grad rho at the step start
  const T* divv;
  double* rho;          // in: rho at the step start; out: rho_new
  double* f_rho;        // in: the last step's source; out: this step's
  const T *mu, *amp, *psi, *divf;
  const T* gp[3];       // grad psi
  T* nlz;               // z pass of Nl (REDFT10)
  // when set: d_z rho_new (the z component of the next step's grad rho, the
  // gradient one-axis operator of a6) is stored here; qs = pi/(dz Nz)
  T* rdz_next = nullptr;
  // when set: divf holds the y and x terms of div f only, and the pass forms
  // the z term from f_z (real space) itself
  const T* fz = nullptr;
  double qs = 0;
};

template <typename T>
hipError_t density_zpass(const StepConsts& c, const DensityIO<T>& io, R2R& r2r, hipStream_t s);
bool density_zpass_supported(int nz);

// d_z rho of the current rho into rdz with the density pass's tile and code,
// for a state that was set or loaded rather than left by a step
template <typename T>
hipError_t rho_dz(const StepConsts& c, const double* rho, T* rdz, double qs, R2R& r2r, hipStream_t s);

// a11's fgrad^ += G erho^, the velocity solve a12 (S:1479-1974) and the
// spectral div v^ of a13 (S:2042-2063) fused with the z passes of the inverse
// transforms that consume them (S:2085-2117): v_x, v_y and div v (REDFT01
// along z) and v_z (RODFT01).  The solve at a cosine mode gives the three
// components one index back along their own axis (velocity.h), so the z
// lines of v_x go one plane back and those of v_y one line back.  For FP32
// storage the solve computes in FP32 when arith32 and the transforms when
// tile32 (lengths with FP32 tables); FP64 otherwise.
template <typename T>
struct VelocityIO {
  const T* S[3];        // spectral S_a^ (after their forward transforms)
  const T *fgq, *erq;   // spectral fgrad^ and erho^
  T* zout[4];           // z pass of v_x, v_y, v_z (vsol^ + virr^) and div v
};

template <typename T>
hipError_t velocity_zpass(const StepConsts& c, const VelocityIO<T>& io, R2R& r2r, bool arith32, bool tile32,
                          hipStream_t s);
bool velocity_zpass_supported(int nz);

// The z component of grad psi (a6, S:1104-1250: a one-axis pass DCT-II,
// shift by -Vs, DST-III along z) and the amplitude a7 (S:1254-1289) in one z
// pass: the pass computes d_z psi, stores it, forms |amp| = sqrt(q0^2 psi^2
// + |grad psi|^2) at the same points and applies the z factor of the
// separable Gaussian filter (DCT-II, exp(-1.57^2 2 Vq^2)/(2N), DCT-III)
// before storing amp.  FP64 arithmetic, except the Gaussian factor's
// transforms in FP32 when gauss32 (FP32 storage, lengths with FP32 tables).
template <typename T>
struct AmpIO {
  const T *psi, *gpx, *gpy;
  T *gpz, *amp;
  double q02;  // q0^2
  double qs;   // wavenumber spacing pi/(dz Nz)
};

template <typename T>
hipError_t amp_zpass(const StepConsts& c, const AmpIO<T>& io, R2R& r2r, bool gauss32, hipStream_t s);
bool amp_zpass_supported(int nz);

// The z pass of psi's inverse transform (a17, S:2177-2198: the update of
// OP_LD_PSI computed on load, new psiq stored) together with the z pass of
// the next step's L psi (a8, S:1293-1308: iDCT of alpha (q0^2 - |q|^2) psiq
// from the same new psiq) in one tile: psi's z lines go to zpsi (FP64),
// L psi's to zlpsi, whose y and x passes the next step runs.  update =
// false transforms psiq as it is (the same two outputs, for a state that was
// set, loaded or restored).  FP64 arithmetic.
template <typename T>
struct PsiShIO {
  Fuse f;                        // OP_LD_PSI's fields, alpha, q02, qs
  const double* psiq = nullptr;  // update = false: the coefficients
  double* zpsi = nullptr;
  T* zlpsi = nullptr;
};

template <typename T>
hipError_t psi_sh_zpass(const StepConsts& c, const PsiShIO<T>& io, bool update, R2R& r2r, hipStream_t s);
bool psi_sh_zpass_supported(int nz);

}  // namespace smaiso
