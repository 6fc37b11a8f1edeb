// Per-mode and pointwise kernels of the time step (smaiso-wcomp.cpp:1081-2263).
//
// Every kernel is a flat grid over the Nx*Ny*Nz points of one member (x
// dimension of the launch) times the members of the batch (y dimension), with
// the reference's layout (i*Ny + j)*Nz + k inside a member.  The reference's
// halo exchanges between x-slabs (S:1108-1164, S:1492-1579, S:1713-1753,
// S:1979-2035) are plain +-1 index shifts here; a shifted read outside the grid
// is 0, as on the first/last rank.  The operator tables of S:936-963 (mq2c,
// aLin, C1, C2, CM1a, CM2a) are evaluated in registers from the per-axis
// wavenumber vectors.
#include "stepper.h"
#include "velocity.h"

#include <cmath>

namespace smaiso {

constexpr int kBlock = 256;

static inline long long n3_of(const StepConsts& c) { return (long long)c.nx * c.ny * c.nz; }
static inline dim3 grid_of(const StepConsts& c) {
  return dim3((unsigned)((n3_of(c) + kBlock - 1) / kBlock), (unsigned)c.nmem);
}
static inline unsigned grid_for(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct Idx {
  int i, j, k;
};

__device__ __forceinline__ Idx decompose(unsigned idx, int ny, int nz) {
  Idx r;
  unsigned t = idx / (unsigned)nz;
  r.k = (int)(idx - t * (unsigned)nz);
  unsigned i = t / (unsigned)ny;
  r.j = (int)(t - i * (unsigned)ny);
  r.i = (int)i;
  return r;
}

__device__ __forceinline__ double sq(double x) { return x * x; }

// member offset of this workgroup
__device__ __forceinline__ long long member_offset(const StepConsts& c) {
  return (long long)blockIdx.y * ((long long)c.nx * c.ny * c.nz);
}

// ---------------------------------------------------------------------------
// pre-loop scaling (S:982)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_spec(StepConsts c, int op, const T* __restrict__ in,
                                                 T* out) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned li = blockIdx.x * kBlock + threadIdx.x;
  if (li >= n3) return;
  const long long idx = member_offset(c) + li;
  (void)op;  // SPEC_SCALE
  out[idx] = (T)(c.scale * (double)in[idx]);
}

// a9 relaxation branch (S:1364-1378): quirk Q4, rho is overwritten
template <typename T>
__global__ __launch_bounds__(kBlock) void k_relax(StepConsts c, const T* __restrict__ amp,
                                                  T* laprl, double* rho) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned li = blockIdx.x * kBlock + threadIdx.x;
  if (li >= n3) return;
  const long long idx = member_offset(c) + li;
  const MemberConsts m = c.mc[blockIdx.y];
  const double r = m.rho0 + ((1 - m.rho0) / m.amp_eq) * (double)amp[idx];
  rho[idx] = r;
  laprl[idx] = (T)(r * (double)laprl[idx]);
}

// ---------------------------------------------------------------------------
// a10: stresses, forces and chemical potential (S:1385-1425)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_forces(
    StepConsts c, const T* __restrict__ lpsi, const T* __restrict__ laprl,
    const double* __restrict__ rho, const T* __restrict__ amp, const T* __restrict__ psi,
    const T* __restrict__ gpx, const T* __restrict__ gpy, const T* __restrict__ gpz,
    const T* __restrict__ rdx, const T* __restrict__ rdy, const T* __restrict__ rdz,
    T* __restrict__ fgrad, T* __restrict__ erho, T* __restrict__ mu, T* __restrict__ sx,
    T* __restrict__ sy, T* __restrict__ sz, T* __restrict__ fx, T* __restrict__ fy,
    T* __restrict__ fz) {
  // FP64 arithmetic for both storage types: rho^2 - (rho0 + kp amp)^2 and
  // rho - rho0 - kp amp cancel to a few 1e-3 inside the smectic
  using A = double;
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned li = blockIdx.x * kBlock + threadIdx.x;
  if (li >= n3) return;
  const long long idx = member_offset(c) + li;
  const MemberConsts m = c.mc[blockIdx.y];
  const A L = lpsi[idx], lr = laprl[idx], r = rho[idx], am = amp[idx], ps = psi[idx];
  const A kp = m.kp, zeta = m.zeta, rho0 = m.rho0;
  const A e = A(0.5) * (L * L);
  const A ra = rho0 + kp * am;
  fgrad[idx] = (T)(A(0.5) * zeta * (r * r - ra * ra));
  const A mu1 = lr + r * L;
  const A gx = gpx[idx], gy = gpy[idx], gz = gpz[idx];
  sx[idx] = (T)(mu1 * gx + e * (A)rdx[idx]);
  sy[idx] = (T)(mu1 * gy + e * (A)rdy[idx]);
  sz[idx] = (T)(mu1 * gz + e * (A)rdz[idx]);
  const A p2 = ps * ps, p3 = p2 * ps, p5 = p3 * p2;
  mu[idx] = (T)(lr + r * (L - (A)c.eps * ps - (A)c.beta * p3 + (A)c.gamma * p5));
  erho[idx] = (T)(r * e);
  const A fc = kp * zeta * (r - rho0 - kp * am) / (am + A(0.001));
  fx[idx] = (T)(fc * gx);
  fy[idx] = (T)(fc * gy);
  fz[idx] = (T)(fc * gz);
}

// ---------------------------------------------------------------------------
// a11 (fgrad^ += erho^ G) fused with a12: velocity solve (S:1479-1974), and
// the spectral divergence of virr^ (S:2042-2063)
// ---------------------------------------------------------------------------
// Thread (i, j, k) solves cosine mode (i, j, k) (velocity_mode): the x
// component at (i-1, j, k), the y component at (i, j-1, k), the z component
// at (i, j, k-1) and div v^(i, j, k), so virr^ need not be stored.  Each thread
// reads five values and writes four.
//
// A: the arithmetic type (FP64; FP32 for FP32 storage where the velocity
// solve's FP32 stage is asked for).  VIRR: also store virr^ (for the vsol and
// virr read-outs and the adjoint), at the same points as the total.
template <typename T, typename A, int BS, bool VIRR>
__global__ __launch_bounds__(BS) void k_velocity(
    StepConsts c, const T* __restrict__ sxq_, const T* __restrict__ syq_,
    const T* __restrict__ szq_, const T* __restrict__ fgq_, const T* __restrict__ erq_,
    T* __restrict__ vtx, T* __restrict__ vty, T* __restrict__ vtz,
    T* __restrict__ virrx, T* __restrict__ virry, T* __restrict__ virrz, T* __restrict__ divq) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  // XCD-aware order: workgroups are dealt to the 8 XCDs round robin.  When
  // an i plane splits into 8 equal runs of whole workgroups, XCD x takes run
  // x of every plane (a band of j lines), plane after plane, so the j-1 line
  // and the i-1 plane a point reads are mostly its own band's, in its own 4
  // MiB L2.  Otherwise each XCD walks a contiguous range of lines.
  const unsigned nb = gridDim.x;
  const unsigned plane = (unsigned)c.ny * c.nz;
  unsigned b = blockIdx.x;
  if (plane % (8u * BS) == 0) {
    const unsigned bpp = plane / BS, bpx = bpp >> 3, s = b >> 3;
    b = (s / bpx) * bpp + (b & 7u) * bpx + (s % bpx);
  } else if ((nb & 7u) == 0) {
    b = (b & 7u) * (nb >> 3) + (b >> 3);
  }
  const unsigned idx = b * BS + threadIdx.x;
  if (idx >= n3) return;
  const long long mo = member_offset(c);
  const MemberConsts mm = c.mc[blockIdx.y];
  const Idx p = decompose(idx, c.ny, c.nz);
  const unsigned sI = (unsigned)c.ny * c.nz, sJ = (unsigned)c.nz;
  const bool hx = p.i > 0, hy = p.j > 0, hz = p.k > 0;
  // shifted reads outside the grid are 0, as on the first rank
  const A Sx = hx ? (A)sxq_[mo + idx - sI] : A(0);
  const A Sy = hy ? (A)syq_[mo + idx - sJ] : A(0);
  const A Sz = hz ? (A)szq_[mo + idx - 1] : A(0);
  const A er = erq_ ? (A)erq_[mo + idx] : A(0);
  const VelMode<A> r = velocity_mode<A>(c, mm, p.i, p.j, p.k, Sx, Sy, Sz, (A)fgq_[mo + idx], er);
  const long long o = mo + idx;
  const long long pox = hx ? o - sI : o + (long long)(c.nx - 1) * sI;
  const long long poy = hy ? o - sJ : o + (long long)(c.ny - 1) * sJ;
  const long long poz = hz ? o - 1 : o + (c.nz - 1);
  vtx[pox] = (T)r.tx;
  vty[poy] = (T)r.ty;
  vtz[poz] = (T)r.tz;
  if constexpr (VIRR) {
    virrx[pox] = (T)r.rx;
    virry[poy] = (T)r.ry;
    virrz[poz] = (T)r.rz;
  }
  if (divq) divq[o] = (T)r.div;
}

// ---------------------------------------------------------------------------
// a15 + a16: density AB2 (S:2121-2139) and the nonlinear term (S:2144-2162)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_density_nl(
    StepConsts c, const T* __restrict__ vx, const T* __restrict__ vy,
    const T* __restrict__ vz, const T* __restrict__ rdx, const T* __restrict__ rdy,
    const T* __restrict__ rdz, const T* __restrict__ divv, double* __restrict__ rho,
    double* __restrict__ f_rho, const T* __restrict__ mu,
    const T* __restrict__ amp, const T* __restrict__ psi, const T* __restrict__ divf,
    const T* __restrict__ gpx, const T* __restrict__ gpy, const T* __restrict__ gpz,
    T* __restrict__ nl) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned li = blockIdx.x * kBlock + threadIdx.x;
  if (li >= n3) return;
  const long long idx = member_offset(c) + li;
  const MemberConsts m = c.mc[blockIdx.y];
  // density in FP64 (both precisions)
  const double ux = vx[idx], uy = vy[idx], uz = vz[idx];
  const double r0 = rho[idx];
  const double src = -ux * (double)rdx[idx] - uy * (double)rdy[idx] - uz * (double)rdz[idx] -
                     r0 * (double)divv[idx];
  const double r1 = r0 + c.dt * (1.5 * src - 0.5 * f_rho[idx]);
  rho[idx] = r1;
  f_rho[idx] = src;
  // FP64 arithmetic (rho_new - rho0 - kp amp cancels inside the smectic)
  const double am = amp[idx];
  const double v = -(double)mu[idx] + ((m.kp * m.zeta) / (am + 0.001)) * (r1 - m.rho0 - m.kp * am) * (double)psi[idx] -
                   (double)divf[idx] - ux * (double)gpx[idx] - uy * (double)gpy[idx] - uz * (double)gpz[idx];
  nl[idx] = (T)v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_energy_density(StepConsts c, const T* __restrict__ psi,
                                                           const T* __restrict__ lpsi,
                                                           const double* __restrict__ rho,
                                                           const T* __restrict__ amp,
                                                           T* __restrict__ out) {
  const unsigned n3 = (unsigned)c.nx * c.ny * c.nz;
  const unsigned li = blockIdx.x * kBlock + threadIdx.x;
  if (li >= n3) return;
  const long long idx = member_offset(c) + li;
  const MemberConsts m = c.mc[blockIdx.y];
  const double ps = psi[idx], L = lpsi[idx], r = rho[idx], am = amp[idx];
  const double p2 = ps * ps;
  const double e = -0.5 * c.eps * p2 + 0.5 * (L * L) - 0.25 * c.beta * (p2 * p2) +
                   (c.gamma / 6) * (p2 * p2 * p2);
  const double d = r - m.rho0 - m.kp * am;
  out[idx] = (T)(r * e + 0.5 * m.zeta * (d * d));
}

// ---------------------------------------------------------------------------
// a18: diagnostics with compensated (two-sum) accumulation
// ---------------------------------------------------------------------------
struct CSum {
  double s, c;
};

__device__ __forceinline__ void csum_add(CSum& a, double x) {
  const double t = a.s + x;
  // Neumaier: carry the low part of whichever operand is larger
  if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
  else a.c += (x - t) + a.s;
  a.s = t;
}

__device__ __forceinline__ CSum csum_merge(CSum a, CSum b) {
  CSum r = a;
  csum_add(r, b.s);
  r.c += b.c;
  return r;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag(StepConsts c, const T* __restrict__ psi,
                                                 const T* __restrict__ psi_old,
                                                 const double* __restrict__ rho,
                                                 const T* __restrict__ lpsi,
                                                 const T* __restrict__ amp,
                                                 double* __restrict__ partials) {
  const long long n3 = (long long)c.nx * c.ny * c.nz;
  const long long mo = member_offset(c);
  const MemberConsts m = c.mc[blockIdx.y];
  CSum acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (long long li = (long long)blockIdx.x * kBlock + threadIdx.x; li < n3;
       li += (long long)gridDim.x * kBlock) {
    const long long idx = mo + li;
    const double ps = psi[idx], po = psi_old[idx], r = rho[idx], L = lpsi[idx], am = amp[idx];
    csum_add(acc[0], fabs(ps - po));
    csum_add(acc[1], fabs(ps));
    csum_add(acc[2], r);
    const double p2 = ps * ps;
    const double e = -0.5 * c.eps * p2 + 0.5 * (L * L) - 0.25 * c.beta * (p2 * p2) +
                     (c.gamma / 6) * (p2 * p2 * p2);
    const double d = r - m.rho0 - m.kp * am;
    csum_add(acc[3], r * e + 0.5 * m.zeta * (d * d));
  }
  __shared__ CSum sh[4][kBlock];
#pragma unroll
  for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = csum_merge(sh[q][threadIdx.x], sh[q][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) {
    const CSum r = sh[threadIdx.x][0];
    double* pm = partials + (size_t)blockIdx.y * kDiagBlocks * 8;
    pm[2 * (4 * blockIdx.x + threadIdx.x)] = r.s;
    pm[2 * (4 * blockIdx.x + threadIdx.x) + 1] = r.c;
  }
}

__global__ __launch_bounds__(kBlock) void k_diag_final(const double* __restrict__ partials, int nb,
                                                       double* __restrict__ out) {
  const double* pm = partials + (size_t)blockIdx.x * kDiagBlocks * 8;
  CSum acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (int b = threadIdx.x; b < nb; b += kBlock) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      acc[q] = csum_merge(acc[q], CSum{pm[2 * (4 * b + q)], pm[2 * (4 * b + q) + 1]});
  }
  __shared__ CSum sh[4][kBlock];
#pragma unroll
  for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = csum_merge(sh[q][threadIdx.x], sh[q][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) out[4 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0].s + sh[threadIdx.x][0].c;
}

__global__ void k_d2f(const double* __restrict__ in, float* __restrict__ out, long long n) {
  const long long idx = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (idx < n) out[idx] = (float)in[idx];
}
__global__ void k_f2d(const float* __restrict__ in, double* __restrict__ out, long long n) {
  const long long idx = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (idx < n) out[idx] = (double)in[idx];
}
__global__ void k_zero_d(double* __restrict__ out, long long n) {
  const long long idx = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (idx < n) out[idx] = 0.0;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T>
void StepKernels<T>::spec(const StepConsts& c, int op, const T* in, T* out, hipStream_t s) {
  hipLaunchKernelGGL(k_spec<T>, grid_of(c), dim3(kBlock), 0, s, c, op, in, out);
}
template <typename T>
void StepKernels<T>::relax(const StepConsts& c, const T* amp, T* laprl, double* rho, hipStream_t s) {
  hipLaunchKernelGGL(k_relax<T>, grid_of(c), dim3(kBlock), 0, s, c, amp, laprl, rho);
}
// 1024-point workgroups (four lines of 256) when they tile the field: the
// j-1 lines a point reads are then mostly read by its own workgroup, on the
// same CU at the same time
template <typename T, typename A, bool VIRR>
static void launch_velocity(const StepConsts& c, const T* sxq, const T* syq, const T* szq, const T* fgq,
                            const T* erq, T* vx, T* vy, T* vz, T* virrx, T* virry, T* virrz, T* divq,
                            hipStream_t s) {
  constexpr int BS = 1024;
  const long long n3 = n3_of(c);
  if (n3 % BS == 0) {
    hipLaunchKernelGGL((k_velocity<T, A, BS, VIRR>), dim3((unsigned)(n3 / BS), (unsigned)c.nmem), dim3(BS), 0, s,
                       c, sxq, syq, szq, fgq, erq, vx, vy, vz, virrx, virry, virrz, divq);
  } else {
    hipLaunchKernelGGL((k_velocity<T, A, kBlock, VIRR>), grid_of(c), dim3(kBlock), 0, s, c, sxq, syq, szq, fgq,
                       erq, vx, vy, vz, virrx, virry, virrz, divq);
  }
}
template <typename T, typename A>
static void launch_velocity_any(const StepConsts& c, const T* sxq, const T* syq, const T* szq, const T* fgq,
                                const T* erq, T* vx, T* vy, T* vz, T* virrx, T* virry, T* virrz, T* divq,
                                hipStream_t s) {
  if (virrx)
    launch_velocity<T, A, true>(c, sxq, syq, szq, fgq, erq, vx, vy, vz, virrx, virry, virrz, divq, s);
  else
    launch_velocity<T, A, false>(c, sxq, syq, szq, fgq, erq, vx, vy, vz, virrx, virry, virrz, divq, s);
}

template <typename T>
void StepKernels<T>::forces(const StepConsts& c, const T* lpsi, const T* laprl, const double* rho,
                            const T* amp, const T* psi, const T* gpx, const T* gpy, const T* gpz,
                            const T* rdx, const T* rdy, const T* rdz, T* fgrad, T* erho, T* mu,
                            T* sx, T* sy, T* sz, T* fx, T* fy, T* fz, hipStream_t s) {
  hipLaunchKernelGGL(k_forces<T>, grid_of(c), dim3(kBlock), 0, s, c, lpsi, laprl, rho, amp, psi,
                     gpx, gpy, gpz, rdx, rdy, rdz, fgrad, erho, mu, sx, sy, sz, fx, fy, fz);
}
template <typename T>
void StepKernels<T>::velocity(const StepConsts& c, const T* sxq, const T* syq, const T* szq,
                              const T* fgq, const T* erq, T* vx, T* vy, T* vz, T* virrx,
                              T* virry, T* virrz, T* divq, bool arith32, hipStream_t s) {
  if constexpr (sizeof(T) == 4) {
    if (arith32) {
      launch_velocity_any<T, float>(c, sxq, syq, szq, fgq, erq, vx, vy, vz, virrx, virry, virrz, divq, s);
      return;
    }
  }
  launch_velocity_any<T, double>(c, sxq, syq, szq, fgq, erq, vx, vy, vz, virrx, virry, virrz, divq, s);
}
template <typename T>
void StepKernels<T>::density_nl(const StepConsts& c, const T* vx, const T* vy, const T* vz,
                                const T* rdx, const T* rdy, const T* rdz, const T* divv,
                                double* rho, double* f_rho, const T* mu, const T* amp,
                                const T* psi, const T* divf, const T* gpx, const T* gpy,
                                const T* gpz, T* nl, hipStream_t s) {
  hipLaunchKernelGGL(k_density_nl<T>, grid_of(c), dim3(kBlock), 0, s, c, vx, vy, vz, rdx, rdy, rdz,
                     divv, rho, f_rho, mu, amp, psi, divf, gpx, gpy, gpz, nl);
}
template <typename T>
void StepKernels<T>::energy_density(const StepConsts& c, const T* psi, const T* lpsi, const double* rho,
                                    const T* amp, T* out, hipStream_t s) {
  hipLaunchKernelGGL(k_energy_density<T>, grid_of(c), dim3(kBlock), 0, s, c, psi, lpsi, rho, amp, out);
}
template <typename T>
int StepKernels<T>::diag_partial(const StepConsts& c, const T* psi, const T* psi_old,
                                 const double* rho, const T* lpsi, const T* amp, double* partials,
                                 hipStream_t s) {
  long long nb = grid_for(n3_of(c));
  if (nb > kDiagBlocks) nb = kDiagBlocks;
  hipLaunchKernelGGL(k_diag<T>, dim3((unsigned)nb, (unsigned)c.nmem), dim3(kBlock), 0, s, c, psi,
                     psi_old, rho, lpsi, amp, partials);
  return (int)nb;
}

void diag_final(const double* partials, int nb, int nmem, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_final, dim3((unsigned)nmem), dim3(kBlock), 0, s, partials, nb, out);
}
void convert_d2f(const double* in, float* out, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_d2f, dim3(grid_for(n)), dim3(kBlock), 0, s, in, out, n);
}
void convert_f2d(const float* in, double* out, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_f2d, dim3(grid_for(n)), dim3(kBlock), 0, s, in, out, n);
}
void fill_zero_d(double* out, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_d, dim3(grid_for(n)), dim3(kBlock), 0, s, out, n);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sub(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out,
                                                long long n) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
    out[i] = a[i] - b[i];
}
template <typename T>
void sub_fields(const T* a, const T* b, T* out, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_sub<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, a, b, out, n);
}
template void sub_fields<float>(const float*, const float*, float*, long long, hipStream_t);
template void sub_fields<double>(const double*, const double*, double*, long long, hipStream_t);

// One wavefront per column: chunks of 64 k values from the top, the first
// chunk with a qualifying value gives the highest such k (its lowest lane).
__global__ __launch_bounds__(kBlock) void k_surface(const double* __restrict__ rho, int nx, int ny,
                                                    int nz, double thr, int* __restrict__ out) {
  constexpr int kWave = 64;
  const int lane = threadIdx.x % kWave;
  const long long col = (long long)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (col >= (long long)nx * ny) return;  // whole wavefronts leave together
  const double* line = rho + col * nz;
  int found = -1;
  for (int top = nz - 1; top >= 0; top -= kWave) {
    const int k = top - lane;
    const bool hit = k >= 0 && line[k] > thr;
    const unsigned long long mask = __ballot(hit);
    if (mask) {
      found = top - (__ffsll(mask) - 1);
      break;
    }
  }
  if (lane == 0) out[col] = found;
}

void surface_height(const double* rho, int nx, int ny, int nz, double threshold, int* out,
                    hipStream_t s) {
  const long long cols = (long long)nx * ny;
  const unsigned blocks = (unsigned)((cols + (kBlock / 64) - 1) / (kBlock / 64));
  hipLaunchKernelGGL(k_surface, dim3(blocks), dim3(kBlock), 0, s, rho, nx, ny, nz, threshold, out);
}

template struct StepKernels<double>;
template struct StepKernels<float>;

}  // namespace smaiso
