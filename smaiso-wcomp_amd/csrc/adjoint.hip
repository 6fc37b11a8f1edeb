// Adjoint kernels of one time step (the transpose of the kernels in
// stepper.hip and of the fused passes in r2r.hip; see adjoint.h).
//
// Notation: a bar marks the adjoint of a forward quantity.  Every kernel is a
// flat grid over the Nx*Ny*Nz points of one simulation, layout
// (i*Ny + j)*Nz + k; "spectral" kernels index the DCT/DST lattices, whose
// wavenumbers are the forward tables (S:447-481).
#include "adjoint.h"

#include <cmath>

namespace smaiso {

namespace {

constexpr int kB = 256;

struct Ix {
  int i, j, k;
};

__device__ __forceinline__ Ix decomp(unsigned idx, int ny, int nz) {
  Ix r;
  const unsigned t = idx / (unsigned)nz;
  r.k = (int)(idx - t * (unsigned)nz);
  const unsigned i = t / (unsigned)ny;
  r.j = (int)(t - i * (unsigned)ny);
  r.i = (int)i;
  return r;
}

__device__ __forceinline__ double sqd(double x) { return x * x; }

// CT^T = ICT . wc: 2 on every index-0 plane
__device__ __forceinline__ double wc_at(int i, int j, int k) {
  return (i == 0 ? 2.0 : 1.0) * (j == 0 ? 2.0 : 1.0) * (k == 0 ? 2.0 : 1.0);
}
// ST_a^T = IST_a . ws_a: index N_a-1 on the DST axis, index 0 on the others
__device__ __forceinline__ double ws_at(const StepConsts& c, int a, int i, int j, int k) {
  const double wx = a == 0 ? (i == c.nx - 1 ? 2.0 : 1.0) : (i == 0 ? 2.0 : 1.0);
  const double wy = a == 1 ? (j == c.ny - 1 ? 2.0 : 1.0) : (j == 0 ? 2.0 : 1.0);
  const double wz = a == 2 ? (k == c.nz - 1 ? 2.0 : 1.0) : (k == 0 ? 2.0 : 1.0);
  return wx * wy * wz;
}

__device__ __forceinline__ double mq2c_at(const StepConsts& c, int i, int j, int k) {
  return sqd(c.vq[0][i]) + sqd(c.vq[1][j]) + sqd(c.vq[2][k]);
}

// deterministic block sum of v (every thread of the block must call it)
__device__ __forceinline__ void block_partial(double v, double* out) {
  __shared__ double sh[kB];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = kB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

__host__ __device__ inline unsigned nblocks(const StepConsts& c) {
  return (unsigned)(((long long)c.nx * c.ny * c.nz + kB - 1) / kB);
}

// ---------------------------------------------------------------------------
// a17 (S:2177-2190): psiq_new = (C1 psiq - psiq_old/2 + dt/2 scale (3 nlq - nlq_old))/C2
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kB) void k_adj_update(StepConsts c, AdjArgs<T> A) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned idx = blockIdx.x * kB + threadIdx.x;
  if (idx >= n3) return;
  const Ix p = decomp(idx, c.ny, c.nz);
  const double m2 = mq2c_at(c, p.i, p.j, p.k);
  const double alin = c.eps - c.alpha * sqd(c.q02 - m2);
  const double c1 = 2.0 - c.dt * c.rho_m * alin;
  const double c2 = 1.5 - c.dt * c.rho_m * alin;
  const double u = A.bP[idx] / c2;
  double bnlq = 1.5 * c.dt * c.scale * u + A.bNo[idx];
  double bno = -0.5 * c.dt * c.scale * u;
  if (A.first) {  // quirk Q5: nlq_old = nlq on the first step
    bnlq += bno;
    bno = 0.0;
  }
  A.bP[idx] = c1 * u + A.bPo[idx];  // psiq_old_new = psiq
  A.bPo[idx] = -0.5 * u;
  A.bNo[idx] = bno;
  A.tmp[idx] = wc_at(p.i, p.j, p.k) * bnlq;  // -> ICT: bnl = CT^T(bnlq)
}

// ---------------------------------------------------------------------------
// a16 (S:2144-2162) and a15 (S:2121-2139), real space
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kB) void k_adj_nl(StepConsts c, AdjArgs<T> A) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned idx = blockIdx.x * kB + threadIdx.x;
  double dz = 0.0;
  if (idx < n3) {
    const MemberConsts m = c.mc[0];
    const double kp = m.kp, zeta = m.zeta, rho0 = m.rho0;
    const double ps = A.psi[idx], am = A.amp[idx], R = A.rho[idx], Rn = A.rho_new[idx];
    const double bnl = A.bnl[idx];
    const double ia = 1.0 / (am + 0.001);
    const double cc = kp * zeta * ia;
    const double D = Rn - rho0 - kp * am;
    // nl = -mu + cc D psi - divf - v . grad psi
    const double bRn = A.bR[idx] + bnl * cc * ps;
    A.bamp[idx] = bnl * ps * (-kp * zeta * D * ia * ia - kp * cc);
    A.bpsi[idx] = bnl * cc * D;
    dz = bnl * kp * ia * D * ps;
    A.bdivf[idx] = -bnl;
    // rho_new = rho + dt (1.5 s - 0.5 f_rho), f_rho_new = s,
    // s = -v . grad rho - rho div v
    const double bs = 1.5 * c.dt * bRn + A.bFr[idx];
    A.bR[idx] = bRn - bs * (double)A.divv[idx];
    A.bFr[idx] = -0.5 * c.dt * bRn;
    A.bdivv[idx] = -bs * R;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double g = A.gp[a][idx], v = A.v[a][idx], r = A.rd[a][idx];
      A.bv[a][idx] = -bnl * g - bs * r;
      A.bgp[a][idx] = -bnl * v;
      A.brd[a][idx] = -bs * v;
    }
  }
  block_partial(dz, A.partials + 0 * (size_t)nblocks(c) + blockIdx.x);
}

// ---------------------------------------------------------------------------
// a14 + a13 (S:1979-2117) and the first half of the transposed velocity solve
// (S:1583-1974): per component a, from vsol^_a = ST_a(bv_a)/ws_a and
// virr^_a = vsol^_a + vq_a shift(CT(bdivv)/wc), the quantities that the
// gathers of velocity_b read: dot_a, cfg_a (the fgrad^ term) and dir_a (the
// S^_a diagonal); and dL/dnu.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kB) void k_adj_vel_a(StepConsts c, AdjArgs<T> A) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned idx = blockIdx.x * kB + threadIdx.x;
  double dn = 0.0;
  if (idx < n3) {
    const Ix p = decomp(idx, c.ny, c.nz);
    const MemberConsts m = c.mc[0];
    const double nu = m.nu, ilam = m.lam + 2 * m.nu;
    const unsigned sI = (unsigned)c.ny * c.nz, sJ = (unsigned)c.nz;
    const double qv[3] = {c.vq[0][p.i], c.vq[1][p.j], c.vq[2][p.k]};
    const double wv[3] = {c.vs[0][p.i], c.vs[1][p.j], c.vs[2][p.k]};
    const int nn[3] = {c.nx, c.ny, c.nz};
    const int ii[3] = {p.i, p.j, p.k};
    const unsigned st[3] = {sI, sJ, 1u};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double wsa = ws_at(c, a, p.i, p.j, p.k);
      double vs = A.bv[a][idx] / wsa;
      double vi = vs;
      if (ii[a] <= nn[a] - 2) {
        const unsigned nb = idx + st[a];
        const double q = c.vq[a][ii[a] + 1];
        const double wn = wc_at(p.i + (a == 0), p.j + (a == 1), p.k + (a == 2));
        vi += q * A.bdivv[nb] / wn;
      }
      // output masks, quirks Q1 and Q2
      const bool zero = a == 0 ? (p.i == c.nx - 1) : (p.j == c.ny - 1 || p.k == c.nz - 1);
      if (zero) vs = vi = 0.0;
      const double virr = (double)A.virrq[a][idx], vsol = (double)A.vtq[a][idx] - virr;
      dn += vs * (-vsol / nu) + vi * (-2.0 * virr / ilam);
      double mq2 = 0;
#pragma unroll
      for (int b = 0; b < 3; ++b) mq2 += sqd(b == a ? wv[b] : qv[b]);
      const double cm1 = c.scale / mq2, cm2 = wv[a] / mq2;
      // exp(gfac/CM1_a) as a product of the one-axis Gaussian tables
      const double gx = a == 0 ? c.gs[0][p.i] : c.gq[0][p.i];
      double gy = a == 1 ? c.gs[1][p.j] : c.gq[1][p.j];
      double gz = a == 2 ? c.gs[2][p.k] : c.gq[2][p.k];
      if (a == 2) {  // quirk Q3 (S:1932): CM1y in the exponent
        const int slab = c.slab_n0 > 0 ? c.slab_n0 : c.nx;
        if (p.k == 0 && p.j > 0 && (p.i % slab) != 0) { gy = c.gs[1][p.j]; gz = c.gq[2][p.k]; }
      }
      const double dot = -(cm1 / nu) * cm2 * vs + (cm1 / ilam) * cm2 * (gx * gy * gz) * vi;
      A.dotb[a][idx] = dot;
      A.cfgb[a][idx] = (cm1 / ilam) * wv[a] * vi;
      A.dirb[a][idx] = (cm1 / nu) * vs + wv[a] * dot;
    }
  }
  block_partial(dn, A.partials + 1 * (size_t)nblocks(c) + blockIdx.x);
}

// second half: gathers of the shifted reads of the forward solve
template <typename T>
__global__ __launch_bounds__(kB) void k_adj_vel_b(StepConsts c, AdjArgs<T> A) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned idx = blockIdx.x * kB + threadIdx.x;
  if (idx >= n3) return;
  const Ix p = decomp(idx, c.ny, c.nz);
  const unsigned sI = (unsigned)c.ny * c.nz, sJ = (unsigned)c.nz;
  const int i = p.i, j = p.j, k = p.k;
  double bsx = A.dirb[0][idx], bsy = A.dirb[1][idx], bsz = A.dirb[2][idx];
  // x solve read Sy(i+1, j-1, k), Sz(i+1, j, k-1), fgrad(i+1, j, k)
  // y solve read Sx(i-1, j+1, k), Sz(i, j+1, k-1), fgrad(i, j+1, k)
  // z solve read Sx(i-1, j, k+1), Sy(i, j-1, k+1), fgrad(i, j, k+1)
  if (i <= c.nx - 2) {
    const double q = c.vq[0][i + 1];
    if (j >= 1) bsx += q * A.dotb[1][idx + sI - sJ];
    if (k >= 1) bsx += q * A.dotb[2][idx + sI - 1];
  }
  if (j <= c.ny - 2) {
    const double q = c.vq[1][j + 1];
    if (i >= 1) bsy += q * A.dotb[0][idx - sI + sJ];
    if (k >= 1) bsy += q * A.dotb[2][idx + sJ - 1];
  }
  if (k <= c.nz - 2) {
    const double q = c.vq[2][k + 1];
    if (i >= 1) bsz += q * A.dotb[0][idx - sI + 1];
    if (j >= 1) bsz += q * A.dotb[1][idx - sJ + 1];
  }
  double bfg = 0.0;
  if (i >= 1) bfg += A.cfgb[0][idx - sI];
  if (j >= 1) bfg += A.cfgb[1][idx - sJ];
  if (k >= 1) bfg += A.cfgb[2][idx - 1];
  A.wSq[0][idx] = ws_at(c, 0, i, j, k) * bsx;  // -> IST_a: S_a bar = ST_a^T(Sq bar)
  A.wSq[1][idx] = ws_at(c, 1, i, j, k) * bsy;
  A.wSq[2][idx] = ws_at(c, 2, i, j, k) * bsz;
  // fgq^ = CT(fgrad) + G CT(erho), S:1479-1488
  const double w = wc_at(i, j, k) * bfg;
  A.wfg[idx] = w;
  A.wGfg[idx] = w * (c.gq[0][i] * c.gq[1][j] * c.gq[2][k]);
}

// ---------------------------------------------------------------------------
// a10 (S:1385-1425), real space
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kB) void k_adj_forces(StepConsts c, AdjArgs<T> A) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned idx = blockIdx.x * kB + threadIdx.x;
  double dz = 0.0;
  if (idx < n3) {
    const MemberConsts m = c.mc[0];
    const double kp = m.kp, zeta = m.zeta, rho0 = m.rho0;
    const double R = A.rho[idx], am = A.amp[idx], ps = A.psi[idx], L = A.lpsi[idx], lap = A.laprl[idx];
    const double bmu = -A.bnl[idx];
    const double bfgr = A.wfg[idx], ber = A.wGfg[idx];
    double gp[3], rd[3], bS[3], bf[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      gp[a] = A.gp[a][idx];
      rd[a] = A.rd[a][idx];
      bS[a] = A.wSq[a][idx];
      bf[a] = A.wfq[a][idx];
    }
    double bR = A.bR[idx], bam = A.bamp[idx], bps = A.bpsi[idx];
    const double ia = 1.0 / (am + 0.001);
    // f_a = fc grad_a psi, fc = kp zeta (rho - rho0 - kp amp)/(amp + 0.001)
    const double d0 = R - rho0 - kp * am;
    const double fc = kp * zeta * d0 * ia;
    const double bfc = bf[0] * gp[0] + bf[1] * gp[1] + bf[2] * gp[2];
    bR += bfc * kp * zeta * ia;
    bam += bfc * kp * zeta * (-kp * ia - d0 * ia * ia);
    dz += bfc * kp * d0 * ia;
    // erho = rho e, e = L^2/2
    const double e = 0.5 * L * L;
    bR += ber * e;
    double be = ber * R;
    // mu = lap + rho (L - eps psi - beta psi^3 + gamma psi^5)
    const double p2 = ps * ps;
    const double poly = L - c.eps * ps - c.beta * p2 * ps + c.gamma * p2 * p2 * ps;
    double blap = bmu;
    bR += bmu * poly;
    double blpsi = bmu * R;
    bps += bmu * R * (-c.eps - 3 * c.beta * p2 + 5 * c.gamma * p2 * p2);
    // S_a = mu1 grad_a psi + e grad_a rho, mu1 = lap + rho L
    const double mu1 = lap + R * L;
    const double bmu1 = bS[0] * gp[0] + bS[1] * gp[1] + bS[2] * gp[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      A.bgp[a][idx] += fc * bf[a] + bS[a] * mu1;
      be += bS[a] * rd[a];
      A.brd[a][idx] += bS[a] * e;
    }
    blap += bmu1;
    bR += bmu1 * L;
    blpsi += bmu1 * R;
    // fgrad = zeta/2 (rho^2 - (rho0 + kp amp)^2)
    const double rs = rho0 + kp * am;
    bR += bfgr * zeta * R;
    bam -= bfgr * zeta * rs * kp;
    dz += bfgr * 0.5 * (R * R - rs * rs);
    blpsi += be * L;
    A.bR[idx] = bR;
    A.bamp[idx] = bam;
    A.bpsi[idx] = bps;
    A.blap[idx] = blap;
    A.blpsi[idx] = blpsi;
  }
  block_partial(dz, A.partials + 2 * (size_t)nblocks(c) + blockIdx.x);
}

// ---------------------------------------------------------------------------
// a9 tail (lap_rl = op(rho L psi), S:1312-1380) and a7 tail (amp = op(a0),
// a0 = sqrt(q0^2 psi^2 + |grad psi|^2), S:1254-1289); both operators are
// symmetric, so ybar = op(blap) and a0bar = op(bamp) come from the forward
// passes.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kB) void k_adj_amp_lap(StepConsts c, AdjArgs<T> A) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned idx = blockIdx.x * kB + threadIdx.x;
  if (idx >= n3) return;
  const double R = A.rho[idx], L = A.lpsi[idx], ps = A.psi[idx];
  const double y = A.ybar[idx], a0b = A.a0bar[idx];
  A.bR[idx] += y * L;
  A.blpsi[idx] += y * R;
  double gp[3];
  double a2 = c.q02 * ps * ps;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    gp[a] = A.gp[a][idx];
    a2 += gp[a] * gp[a];
  }
  const double a0 = sqrt(a2);
  const double inv = a0 > 0 ? 1.0 / a0 : 0.0;
  A.bpsi[idx] += a0b * c.q02 * ps * inv;
#pragma unroll
  for (int a = 0; a < 3; ++a) A.bgp[a][idx] += a0b * gp[a] * inv;
}

// ---------------------------------------------------------------------------
// a8 (L psi = ICT(alpha (q0^2 - |q|^2) psiq)) and psi = ICT(psiq), spectral:
// psiq bar += (alpha (q0^2 - |q|^2) CT(Lpsi bar) + CT(psi bar)) / wc.  The
// gradients' transposes (a6) were added to psi bar and rho bar in real space
// by one-axis passes before.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kB) void k_adj_grad(StepConsts c, AdjArgs<T> A) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned idx = blockIdx.x * kB + threadIdx.x;
  if (idx >= n3) return;
  const Ix p = decomp(idx, c.ny, c.nz);
  const double lq = c.alpha * (c.q02 - mq2c_at(c, p.i, p.j, p.k));
  A.bP[idx] += (lq * A.blpsi[idx] + A.bpsi[idx]) / wc_at(p.i, p.j, p.k);
}

template <typename T>
__global__ __launch_bounds__(kB) void k_adj_loss(StepConsts c, const T* __restrict__ psi,
                                                 const double* __restrict__ target,
                                                 double* __restrict__ tmp, double* partials) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned idx = blockIdx.x * kB + threadIdx.x;
  double l = 0.0;
  if (idx < n3) {
    const double r = (double)psi[idx] - target[idx];
    tmp[idx] = 2.0 * r;
    l = r * r;
  }
  block_partial(l, partials + 3 * (size_t)nblocks(c) + blockIdx.x);
}

__global__ __launch_bounds__(kB) void k_adj_div_wc(StepConsts c, const double* __restrict__ in,
                                                   double* __restrict__ out) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned idx = blockIdx.x * kB + threadIdx.x;
  if (idx >= n3) return;
  const Ix p = decomp(idx, c.ny, c.nz);
  out[idx] = in[idx] / wc_at(p.i, p.j, p.k);
}

__global__ __launch_bounds__(kB) void k_adj_psi0(StepConsts c, const double* __restrict__ a,
                                                 const double* __restrict__ b,
                                                 double* __restrict__ out) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned idx = blockIdx.x * kB + threadIdx.x;
  if (idx >= n3) return;
  const Ix p = decomp(idx, c.ny, c.nz);
  out[idx] = c.scale * wc_at(p.i, p.j, p.k) * (a[idx] + b[idx]);
}

// one workgroup: acc[2 s], acc[2 s + 1] (sum, compensation) += sum of the slot's partials
__global__ __launch_bounds__(kB) void k_adj_reduce(const double* __restrict__ part, int nb,
                                                   double* __restrict__ acc, int s) {
  __shared__ double sh[kB], shc[kB];
  double sum = 0.0, comp = 0.0;
  for (int b = threadIdx.x; b < nb; b += kB) {
    const double y = part[b] - comp;
    const double t = sum + y;
    comp = (t - sum) - y;
    sum = t;
  }
  sh[threadIdx.x] = sum;
  shc[threadIdx.x] = -comp;
  __syncthreads();
  if (threadIdx.x == 0) {
    // Neumaier over the per-thread results, in thread order
    double s0 = acc[2 * s], c0 = acc[2 * s + 1];
    for (int t = 0; t < kB; ++t) {
      for (int q = 0; q < 2; ++q) {
        const double x = q == 0 ? sh[t] : shc[t];
        const double u = s0 + x;
        if (fabs(s0) >= fabs(x)) c0 += (s0 - u) + x;
        else c0 += (x - u) + s0;
        s0 = u;
      }
    }
    acc[2 * s] = s0;
    acc[2 * s + 1] = c0;
  }
}

}  // namespace

long long adj_blocks(const StepConsts& c) { return nblocks(c); }

#define ADJ_LAUNCH(kern) \
  hipLaunchKernelGGL((kern<T>), dim3(nblocks(c)), dim3(kB), 0, s, c, A)

template <typename T>
void AdjKernels<T>::update(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s) { ADJ_LAUNCH(k_adj_update); }
template <typename T>
void AdjKernels<T>::nl_density(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s) { ADJ_LAUNCH(k_adj_nl); }
template <typename T>
void AdjKernels<T>::velocity_a(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s) { ADJ_LAUNCH(k_adj_vel_a); }
template <typename T>
void AdjKernels<T>::velocity_b(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s) { ADJ_LAUNCH(k_adj_vel_b); }
template <typename T>
void AdjKernels<T>::forces(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s) { ADJ_LAUNCH(k_adj_forces); }
template <typename T>
void AdjKernels<T>::amp_lap(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s) { ADJ_LAUNCH(k_adj_amp_lap); }
template <typename T>
void AdjKernels<T>::grad(const StepConsts& c, const AdjArgs<T>& A, hipStream_t s) { ADJ_LAUNCH(k_adj_grad); }
template <typename T>
void AdjKernels<T>::loss(const StepConsts& c, const T* psi, const double* target, double* tmp,
                         double* partials, hipStream_t s) {
  hipLaunchKernelGGL((k_adj_loss<T>), dim3(nblocks(c)), dim3(kB), 0, s, c, psi, target, tmp, partials);
}

void adj_div_wc(const StepConsts& c, const double* in, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_div_wc, dim3(nblocks(c)), dim3(kB), 0, s, c, in, out);
}
void adj_psi0(const StepConsts& c, const double* a, const double* b, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_psi0, dim3(nblocks(c)), dim3(kB), 0, s, c, a, b, out);
}
void adj_reduce(const double* partials, int slot, int nb, double* acc, int slot_acc, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_reduce, dim3(1), dim3(kB), 0, s, partials + (size_t)slot * nb, nb, acc, slot_acc);
}

template struct AdjKernels<double>;
template struct AdjKernels<float>;

}  // namespace smaiso
