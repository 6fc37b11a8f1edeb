// Kernels of the reverse-mode derivative (adjoint) of one time step.
//
// The reference has no adjoint; these kernels transpose, stage by stage, the
// steady branch of the step (smaiso-wcomp.cpp:1081-2198, with the first-step
// quirk Q5 and the velocity quirks Q1-Q3).  The forward intermediates of the
// step (T = double or float, as stored by the forward stepper) are read from
// the stepper's work fields after it has recomputed that step; every adjoint
// field is FP64.
//
// Transposed transforms reuse the forward r2r passes with diagonal weights
// fused into these kernels: CT^T y = ICT(wc y), ST_a^T y = IST_a(ws_a y),
// ICT^T x = CT(x)/wc, IST_a^T x = ST_a(x)/ws_a, where wc doubles every index-0
// plane and ws_a doubles the index-0 planes of the two DCT axes and the
// index N_a-1 plane of the DST axis a.  The one-axis operators of the forward
// step (gradients, div f, the Gaussian filter) transpose into one-axis
// passes (OP_AX_GRAD_T, OP_AX_DIV_T, OP_AX_GAUSS).
#pragma once

#include "stepper.h"

namespace smaiso {

template <typename T>
struct AdjArgs {
  // forward intermediates of the step being transposed
  const T* psi;       // psi at the step start
  const T* amp;       // Gaussian-filtered amplitude
  const T* lpsi;      // L psi
  const T* laprl;     // lap(rho L psi)
  const T* gp[3];     // grad psi
  const T* rd[3];     // grad rho
  const T* v[3];      // total velocity (real space)
  const T* divv;      // div v (real space)
  const T* vtq[3];    // total velocity vsol^ + virr^ on the staggered lattices (spectral)
  const T* virrq[3];  // irrotational velocity (spectral)
  const double* rho;      // rho at the step start
  const double* rho_new;  // rho after the density update
  // adjoint of the step's state (updated in place: outputs -> inputs)
  double *bP, *bPo, *bR, *bFr, *bNo;
  // adjoint work fields
  double *bnl, *bpsi, *bamp, *blap, *blpsi, *ybar, *a0bar, *tmp;
  double *bgp[3], *brd[3], *bv[3], *bdivv, *bdivf;
  double *dotb[3], *cfgb[3], *dirb[3], *wfq[3], *wSq[3], *wfg, *wGfg;
  double* partials;  // kAdjSlots x (blocks per field) block sums
  int first;         // the step is the first after set_state (quirk Q5)
};

constexpr int kAdjSlots = 4;  // 0: dzeta (a16), 1: dnu (a12), 2: dzeta (a10), 3: loss

template <typename T>
struct AdjKernels {
  // a17: psi update (and the N^l history), writes wc * bnlq to tmp
  static void update(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s);
  // a16 + a15 (real space), input bnl = CT^T(bnlq)
  static void nl_density(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s);
  // a14 + a13 + a12 part 1 (spectral): inputs ST_a(bv_a), CT(bdivv)
  static void velocity_a(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s);
  // a12 part 2 (spectral gathers) -> ws_a Sq^_a, wc fgq^, wc G fgq^
  static void velocity_b(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s);
  // a10 (real space), inputs IST'd wSq, ICT'd wfg, wGfg and f bar (wfq, from the
  // one-axis transposes of div f)
  static void forces(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s);
  // a9 + a7 pointwise tails, inputs ybar = op_lap(blap), a0bar = op_gauss(bamp)
  static void amp_lap(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s);
  // a8 + psi = ICT(psiq) (spectral), inputs CT(bpsi), CT(blpsi)
  static void grad(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s);
  // loss sum (psi - target)^2 into slot 3 partials, tmp = 2 (psi - target)
  static void loss(const StepConsts& c, const T* psi, const double* target, double* tmp,
                   double* partials, hipStream_t s);
};

// out = in / wc (ICT^T after a CT)
void adj_div_wc(const StepConsts& c, const double* in, double* out, hipStream_t s);
// out = scale * wc * (a + b)  (CT^T of set_state's psiq = scale CT(psi0), psiq_old = psiq)
void adj_psi0(const StepConsts& c, const double* a, const double* b, double* out, hipStream_t s);
// acc[2*slot_acc .. +1] += compensated sum of partials[slot * nb .. + nb]
void adj_reduce(const double* partials, int slot, int nb, double* acc, int slot_acc, hipStream_t s);
// number of blocks of a per-point adjoint kernel (and of partials per slot)
long long adj_blocks(const StepConsts& c);

}  // namespace smaiso
