// Stages of the time step fused into transform passes (see fused.h).
#include "fused.h"

#include "r2r_kernels.h"
#include "velocity.h"

#include <type_traits>

namespace smaiso {

namespace {

constexpr int kForceOut = 5;  // fgrad, erho, S_x, S_y, S_z

// Line-pair multiple of the z tiles for short lines: the largest power of two
// up to 256/N (1 from length 192 on), so that the threads of a tile keep
// several points each.
template <int N>
constexpr int short_line_scale() {
  int p = 256 / N, q = 1;
  while (q * 2 <= p) q *= 2;
  return q;
}

template <typename T, int N>
struct ForcesTile {
  static constexpr int P = N <= 256 ? 2 * short_line_scale<N>() : 1;  // line pairs per tile
  // 5 waves: the radix-8 stages of the 5P sequences (N <= 256: 10 x N/8
  // butterflies) give every thread one butterfly at N = 256 and 512
  static constexpr int NT = 320;
  static constexpr size_t lds = (size_t)kForceOut * P * seq_len(N, pick_maxr(N, kForceOut * P, NT)) * sizeof(cplx<double>);
  static constexpr bool fits = lds <= 65536;
};

// One workgroup: 2P consecutive z lines.  Load phase: every point of the
// tile evaluates a10 in FP64 registers (as k_forces), stores mu and f_a and
// puts the five transformed fields at their Makhoul positions of 5P complex
// sequences in LDS (field r, line pair c -> sequence r P + c).  Then one
// FFT over the 5P sequences and Makhoul's post-processing store the z pass
// outputs (S_z is a DST-II along z: odd inputs negated, outputs reversed).
// The tile and the FFT are FP64 for both storage types, as in the r2r passes.
// LZ: io.laprl holds the z lines of lap(rho L psi) before the last inverse z
// pass of its transform (REDFT01); the pass runs first, in the LDS sequences
// of the first output field (Makhoul's inverse: pre-twiddle, inverse FFT),
// and leaves x[m] of each line at position m/2 (m even) or N-1-(m-1)/2 (m
// odd): the slot where the load phase puts point m's value of that field,
// so each point reads its lap(rho L psi) from the slot it then overwrites.
template <typename T, int N, int P, int NT, bool LZ>
__global__ __launch_bounds__(NT) void forces_z_kernel(StepConsts c, ForcesIO<T> io,
                                                      const cplx<double>* __restrict__ tw,
                                                      const cplx<double>* __restrict__ mk) {
  constexpr int NF = kForceOut, TL = 2 * P, HN = N / 2;
  constexpr int MAXR = pick_maxr(N, NF * P, NT);
  constexpr int SEQ = seq_len(N, MAXR), SH = pad_shift(N, MAXR);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using A = double;
  cplx<A>* buf = reinterpret_cast<cplx<A>*>(smem_raw);
  A* rbuf = reinterpret_cast<A*>(smem_raw);
  const int tid = threadIdx.x;
  const long long nl = (long long)c.nx * c.ny;
  const long long first = (long long)blockIdx.x * TL;
  const int nvalid = (int)((nl - first) < TL ? (nl - first) : TL);
  const long long base = (long long)blockIdx.y * nl * N + first * N;
  const MemberConsts mc = c.mc[blockIdx.y];
  StageTw<A, N, NF * P, NT, MAXR> stw;
  stw.load(tw, tid);
  cplx<A>* const lbuf = buf;  // P sequences of lap(rho L psi), then field 0's
  const A* const lrb = rbuf;
  if constexpr (LZ) {
    constexpr int W = P * HN;
    constexpr int ITW = (W + NT - 1) / NT;
    MkSeq<A, N, P, NT> ms(mk, 2, tid);
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      const int w = tid + it * NT;
      if (W % NT == 0 || w < W) {
        const int cc = w / HN, k = w - cc * HN;
        const int kb = (k == 0) ? HN : N - k;
        cplx<A> rk = {A(0), A(0)}, rn = {A(0), A(0)};
        const int l0 = 2 * cc;
        if (l0 < nvalid) {
          const T* q = io.laprl + base + (long long)l0 * N;
          rk.x = (A)q[k];
          rn.x = (A)q[kb];
        }
        if (l0 + 1 < nvalid) {
          const T* q = io.laprl + base + (long long)(l0 + 1) * N;
          rk.y = (A)q[k];
          rn.y = (A)q[kb];
        }
        cplx<A> zk, zb;
        makhoul_pre<A, N>(mk, k, ms.get(mk, 2, k), rk, rn, zk, zb);
        lbuf[cc * SEQ + lp<SH>(k)] = zk;
        lbuf[cc * SEQ + lp<SH>(kb)] = zb;
      }
      ms.advance(2, it);
    }
    StageTw<A, N, P, NT, MAXR> stl;
#pragma unroll
    for (int q = 0; q < StageTw<A, N, P, NT, MAXR>::kStages; ++q) stl.w[q] = stw.w[q];
    __syncthreads();
    fft_stage<A, N, P, NT, MAXR, +1, 0>(lbuf, tw, stl);
  }

  constexpr int E = TL * N;
  constexpr int IT = (E + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + it * NT;
    if (E % NT == 0 || e < E) {
      const int l = e / N, m = e - l * N;
      A v[NF] = {A(0), A(0), A(0), A(0), A(0)};
      if (l < nvalid) {
        const long long idx = base + e;
        const int pl = (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1);
        const A lr = LZ ? lrb[2 * ((l >> 1) * SEQ + lp<SH>(pl)) + (l & 1)] : (A)io.laprl[idx];
        const A L = io.lpsi[idx], r = io.rho[idx], am = io.amp[idx], ps = io.psi[idx];
        const A kp = mc.kp, zeta = mc.zeta, rho0 = mc.rho0;
        const A en = A(0.5) * (L * L);
        const A ra = rho0 + kp * am;
        v[0] = A(0.5) * zeta * (r * r - ra * ra);
        const A mu1 = lr + r * L;
        const A gx = io.gp[0][idx], gy = io.gp[1][idx], gz = io.gp[2][idx];
        v[2] = mu1 * gx + en * io.rd[0][idx];
        v[3] = mu1 * gy + en * io.rd[1][idx];
        v[4] = mu1 * gz + en * io.rd[2][idx];
        const A p2 = ps * ps, p3 = p2 * ps, p5 = p3 * p2;
        io.mu[idx] = (T)(lr + r * (L - (A)c.eps * ps - (A)c.beta * p3 + (A)c.gamma * p5));
        v[1] = r * en;
        const A fc = kp * zeta * (r - rho0 - kp * am) / (am + A(0.001));
        io.f[0][idx] = (T)(fc * gx);
        io.f[1][idx] = (T)(fc * gy);
        io.f[2][idx] = (T)(fc * gz);
      }
      const int pos = (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1);
#pragma unroll
      for (int r = 0; r < NF; ++r) {
        A x = v[r];
        if (r == NF - 1 && (m & 1)) x = -x;
        rbuf[2 * ((r * P + (l >> 1)) * SEQ + lp<SH>(pos)) + (l & 1)] = x;
      }
    }
  }
  __syncthreads();
  fft_stage<A, N, NF * P, NT, MAXR, -1, 0>(buf, tw, stw);

  constexpr int W = NF * P * HN;
  constexpr int ITW = (W + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const int w = tid + it * NT;
    if (W % NT == 0 || w < W) {
      const int sq = w / HN, k = w - sq * HN;
      const int r = sq / P, cc = sq - r * P;
      const int kb = (k == 0) ? HN : N - k;
      A y0a, y1a, y0b, y1b;
      makhoul_post<A, N>(mk, k, kb, mk[k], buf[sq * SEQ + lp<SH>(k)], buf[sq * SEQ + lp<SH>(kb)], y0a, y1a, y0b, y1b);
      const bool sine = r == NF - 1;
      const int oa = sine ? N - 1 - k : k;
      const int ob = sine ? N - 1 - kb : kb;
      T* o = r == 0 ? io.zout[0] : r == 1 ? io.zout[1] : r == 2 ? io.zout[2] : r == 3 ? io.zout[3] : io.zout[4];
      const int l0 = 2 * cc;
      if (l0 < nvalid) {
        T* q = o + base + (long long)l0 * N;
        q[oa] = (T)y0a;
        q[ob] = (T)y0b;
      }
      if (l0 + 1 < nvalid) {
        T* q = o + base + (long long)(l0 + 1) * N;
        q[oa] = (T)y1a;
        q[ob] = (T)y1b;
      }
    }
  }
}

template <typename T, int N>
hipError_t launch_forces(const StepConsts& c, const ForcesIO<T>& io, const void* tw, const void* mk,
                         hipStream_t s) {
  using F = ForcesTile<T, N>;
  if constexpr (!F::fits) {
    return hipErrorNotSupported;
  } else {
    constexpr int TL = 2 * F::P;
    const long long blocks = ((long long)c.nx * c.ny + TL - 1) / TL;
    const dim3 grid((unsigned)blocks, (unsigned)c.nmem);
    const auto* t = static_cast<const cplx<double>*>(tw);
    const auto* m = static_cast<const cplx<double>*>(mk);
    if constexpr (sizeof(T) == 8) {
      if (io.laprl_zlines) {
        hipLaunchKernelGGL((forces_z_kernel<T, N, F::P, F::NT, true>), grid, dim3(F::NT), F::lds, s, c, io, t, m);
        return hipGetLastError();
      }
    } else {
      if (io.laprl_zlines) return hipErrorNotSupported;
    }
    hipLaunchKernelGGL((forces_z_kernel<T, N, F::P, F::NT, false>), grid, dim3(F::NT), F::lds, s, c, io, t, m);
    return hipGetLastError();
  }
}

template <typename T>
bool fits_for(int n) {
  switch (n) {
    case 16: return ForcesTile<T, 16>::fits;
    case 32: return ForcesTile<T, 32>::fits;
    case 48: return ForcesTile<T, 48>::fits;
    case 64: return ForcesTile<T, 64>::fits;
    case 96: return ForcesTile<T, 96>::fits;
    case 128: return ForcesTile<T, 128>::fits;
    case 192: return ForcesTile<T, 192>::fits;
    case 256: return ForcesTile<T, 256>::fits;
    case 384: return ForcesTile<T, 384>::fits;
    case 512: return ForcesTile<T, 512>::fits;
    case 1024: return ForcesTile<T, 1024>::fits;
    default: return false;
  }
}

}  // namespace

bool forces_zpass_supported(DType dt, int nz) {
  return dt == F64 ? fits_for<double>(nz) : fits_for<float>(nz);
}

bool forces_zpass_laprl(DType dt, int nz) { return dt == F64 && fits_for<double>(nz); }

template <typename T>
hipError_t forces_zpass(const StepConsts& c, const ForcesIO<T>& io, R2R& r2r, hipStream_t s) {
  const DType dt = sizeof(T) == 8 ? F64 : F32;
  if (!forces_zpass_supported(dt, c.nz)) return hipErrorNotSupported;
  const void *tw = nullptr, *mk = nullptr;
  hipError_t e = r2r.table_ptrs(c.nz, &tw, &mk);
  if (e != hipSuccess) return e;
  switch (c.nz) {
    case 16: return launch_forces<T, 16>(c, io, tw, mk, s);
    case 32: return launch_forces<T, 32>(c, io, tw, mk, s);
    case 48: return launch_forces<T, 48>(c, io, tw, mk, s);
    case 64: return launch_forces<T, 64>(c, io, tw, mk, s);
    case 96: return launch_forces<T, 96>(c, io, tw, mk, s);
    case 128: return launch_forces<T, 128>(c, io, tw, mk, s);
    case 192: return launch_forces<T, 192>(c, io, tw, mk, s);
    case 256: return launch_forces<T, 256>(c, io, tw, mk, s);
    case 384: return launch_forces<T, 384>(c, io, tw, mk, s);
    case 512: return launch_forces<T, 512>(c, io, tw, mk, s);
    case 1024: return launch_forces<T, 1024>(c, io, tw, mk, s);
    default: return hipErrorNotSupported;
  }
}

namespace {

// Operator phase of a one-axis operator along z on a tile of P sequences
// (line pairs) after the forward FFT: item k < N/2 of sequence c forms the
// forward outputs at k and kb (Makhoul post), applies the operator and
// Makhoul's inverse pre-twiddle, in place at k and kb (the sources of the
// modes the pre-twiddle reads are the item's own two forward outputs).
// kZGrad: DCT-II in, y[n] = -Vs[n] Y[n+1]/(2N), DST-III out; kZDiv: DST-II
// in (odd inputs negated, outputs reversed), y[n] = Vq[n] Y[n-1]/(2N),
// DCT-III out; kZGauss: the Gaussian factor exp(-1.57^2 2 Vq[n]^2)/(2N)
// between DCT-II and DCT-III.  Vq[i] = i qs.  The caller synchronizes before
// and after.
enum ZAxisOp { kZGrad = 0, kZDiv = 1, kZGauss = 2 };
template <int ZOP, int N, int P, int NT, int SEQ, int SH, typename A = double>
__device__ __forceinline__ void z_axis_op_phase(cplx<A>* buf, const cplx<A>* __restrict__ mk,
                                                const MkSeq<A, N, P, NT>& ms0, double qs) {
  // weights in FP64 (the one-axis passes' ax_weight), applied in the tile's type
  const double cw = 1.0 / (2.0 * N);
  const int tid = threadIdx.x;
  constexpr int HN = N / 2, W = P * HN, ITW = (W + NT - 1) / NT;
  auto vq = [&](int i) -> double { return (double)i * qs; };
  MkSeq<A, N, P, NT> ms = ms0;
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const int w = tid + it * NT;
    if (W % NT == 0 || w < W) {
      const int c = w / HN, k = w - c * HN;
      const int kb = (k == 0) ? HN : N - k;
      const cplx<A> mkk = ms.get(mk, 2, k);
      A y0a, y1a, y0b, y1b;
      makhoul_post<A, N>(mk, k, kb, mkk, buf[c * SEQ + lp<SH>(k)], buf[c * SEQ + lp<SH>(kb)], y0a, y1a, y0b, y1b);
      cplx<A> xk = {A(0), A(0)}, xb = {A(0), A(0)};
      if constexpr (ZOP == kZGrad) {
        // DST-III output: positions pa = N-1-k, pb = N-1-kb; source p+1
        const int pa = N - 1 - k, pb = N - 1 - kb;
        if (pa <= N - 2) {
          const double wgt = -cw * vq(pa + 1);
          const cplx<A> y = (pa + 1 == k) ? cplx<A>{y0a, y1a} : cplx<A>{y0b, y1b};
          xk = cplx<A>{(A)(wgt * y.x), (A)(wgt * y.y)};
        }
        if (pb <= N - 2) {
          const double wgt = -cw * vq(pb + 1);
          const cplx<A> y = (pb + 1 == k) ? cplx<A>{y0a, y1a} : cplx<A>{y0b, y1b};
          xb = cplx<A>{(A)(wgt * y.x), (A)(wgt * y.y)};
        }
      } else if constexpr (ZOP == kZDiv) {
        // DST-II input: forward output index N-1-k sits at Makhoul index k;
        // DCT-III output positions pa = k, pb = kb; source p-1
        const int fa = N - 1 - k;
        if (k >= 1) {
          const double wgt = cw * vq(k);
          const cplx<A> y = (k - 1 == fa) ? cplx<A>{y0a, y1a} : cplx<A>{y0b, y1b};
          xk = cplx<A>{(A)(wgt * y.x), (A)(wgt * y.y)};
        }
        {
          const double wgt = cw * vq(kb);
          const cplx<A> y = (kb - 1 == fa) ? cplx<A>{y0a, y1a} : cplx<A>{y0b, y1b};
          xb = cplx<A>{(A)(wgt * y.x), (A)(wgt * y.y)};
        }
      } else {
        const double qa = vq(k), qb = vq(kb);
        const double wa = cw * exp(-1.57 * 1.57 * qa * qa * 2), wb = cw * exp(-1.57 * 1.57 * qb * qb * 2);
        xk = cplx<A>{(A)(wa * y0a), (A)(wa * y1a)};
        xb = cplx<A>{(A)(wb * y0b), (A)(wb * y1b)};
      }
      cplx<A> zk, zb;
      makhoul_pre<A, N>(mk, k, mkk, xk, xb, zk, zb);
      buf[c * SEQ + lp<SH>(k)] = zk;
      buf[c * SEQ + lp<SH>(kb)] = zb;
    }
    ms.advance(2, it);
  }
}

// Line pairs per workgroup of the density pass and of rho_dz (one from
// length 192 on), so that short lines still give each of the 128 threads
// several points.
template <int N>
constexpr int density_pairs() {
  return short_line_scale<N>();
}

// d_z of P real z line pairs by the gradient one-axis operator (DCT-II along
// z, shift by -Vs, DST-III): the values of items e = tid + it NT (line e /
// N, element e mod N; the layout of the density pass) go into the LDS
// sequences in Makhoul's forward order (line pair c -> sequence c), and d_z
// lands in out (line l at out + l N; nvalid lines).  The caller has
// synchronized after its last use of buf.
template <typename T, int N, int P, int NT>
__device__ __forceinline__ void dz_lines(const double* x, cplx<double>* buf, T* out, int nvalid,
                                         const cplx<double>* __restrict__ tw, const cplx<double>* __restrict__ mk,
                                         double qs) {
  using A = double;
  constexpr int HN = N / 2;
  constexpr int MAXR = pick_maxr(N, P, NT);
  constexpr int SEQ = seq_len(N, MAXR), SH = pad_shift(N, MAXR);
  constexpr int E = 2 * P * N, IT = (E + NT - 1) / NT;
  A* rbuf = reinterpret_cast<A*>(buf);
  const int tid = threadIdx.x;
  StageTw<A, N, P, NT, MAXR> stw;
  stw.load(tw, tid);
  const MkSeq<A, N, P, NT> ms0(mk, 2, tid);
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + it * NT;
    if (E % NT == 0 || e < E) {
      const int l = e / N, m = e - l * N;
      const int pos = (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1);
      rbuf[2 * ((l >> 1) * SEQ + lp<SH>(pos)) + (l & 1)] = x[it];
    }
  }
  __syncthreads();
  fft_stage<A, N, P, NT, MAXR, -1, 0>(buf, tw, stw);
  z_axis_op_phase<kZGrad, N, P, NT, SEQ, SH>(buf, mk, ms0, qs);
  __syncthreads();
  fft_stage<A, N, P, NT, MAXR, +1, 0>(buf, tw, stw);
  constexpr int W = P * HN, ITW = (W + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const int w = tid + it * NT;
    if (W % NT == 0 || w < W) {
      const int cc = w / HN, n = w - cc * HN;
      const cplx<A> ua = buf[cc * SEQ + lp<SH>(n)];
      const cplx<A> ub = buf[cc * SEQ + lp<SH>(N - 1 - n)];
      // DST-III: y_{2n} = u_n, y_{2n+1} = -u_{N-1-n}
      const int l0 = 2 * cc;
      if (l0 < nvalid) { out[l0 * N + 2 * n] = (T)ua.x; out[l0 * N + 2 * n + 1] = (T)(-ub.x); }
      if (l0 + 1 < nvalid) { out[(l0 + 1) * N + 2 * n] = (T)ua.y; out[(l0 + 1) * N + 2 * n + 1] = (T)(-ub.y); }
    }
  }
}

// One workgroup: P = density_pairs<N>() z line pairs of 128 threads.  Load
// phase: every point evaluates a15 and a16 in FP64 registers as k_density_nl
// does (same expressions, same order), stores rho_new and the source f_rho,
// and puts Nl at its Makhoul position of its pair's complex sequence in
// LDS; then the
// FFT and Makhoul's post-processing store Nl's z pass (DCT-II).  DZF: the
// z term of div f first, from f_z in the same LDS sequence (the divergence
// one-axis operator: DST-II, y[n] = Vq[n] Y[n-1]/(2N), DCT-III); each point
// adds it to the stored y and x terms and overwrites its own slot with Nl.
template <typename T, int N, int NT, bool DZF>
__global__ __launch_bounds__(NT) void density_z_kernel(StepConsts c, DensityIO<T> io,
                                                       const cplx<double>* __restrict__ tw,
                                                       const cplx<double>* __restrict__ mk) {
  constexpr int P = density_pairs<N>(), TL = 2 * P, HN = N / 2;
  constexpr int MAXR = pick_maxr(N, P, NT);
  constexpr int SEQ = seq_len(N, MAXR), SH = pad_shift(N, MAXR);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using A = double;
  cplx<A>* buf = reinterpret_cast<cplx<A>*>(smem_raw);
  A* rbuf = reinterpret_cast<A*>(smem_raw);
  const int tid = threadIdx.x;
  const long long nl = (long long)c.nx * c.ny;
  const long long first = (long long)blockIdx.x * TL;
  const int nvalid = (int)((nl - first) < TL ? (nl - first) : TL);
  const long long base = (long long)blockIdx.y * nl * N + first * N;
  const MemberConsts mc = c.mc[blockIdx.y];
  StageTw<A, N, P, NT, MAXR> stw;
  stw.load(tw, tid);

  constexpr int E = TL * N;
  constexpr int IT = (E + NT - 1) / NT;
  if constexpr (DZF) {
    // f_z in DST-II input order: odd inputs negated at Makhoul positions
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      if (E % NT == 0 || e < E) {
        const int l = e / N, m = e - l * N;
        A x = l < nvalid ? (A)io.fz[base + e] : A(0);
        if (m & 1) x = -x;
        const int pos = (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1);
        rbuf[2 * ((l >> 1) * SEQ + lp<SH>(pos)) + (l & 1)] = x;
      }
    }
    __syncthreads();
    fft_stage<A, N, P, NT, MAXR, -1, 0>(buf, tw, stw);
    const MkSeq<A, N, P, NT> ms0(mk, 2, tid);
    z_axis_op_phase<kZDiv, N, P, NT, SEQ, SH>(buf, mk, ms0, io.qs);
    __syncthreads();
    // DCT-III output at position m sits in the slot of Makhoul position m
    fft_stage<A, N, P, NT, MAXR, +1, 0>(buf, tw, stw);
  }
  A rnew[IT];  // rho_new of the items, for d_z rho_new
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + it * NT;
    rnew[it] = A(0);
    if (E % NT == 0 || e < E) {
      const int l = e / N, m = e - l * N;
      A v = A(0);
      if (l < nvalid) {
        const long long idx = base + e;
        const A ux = io.v[0][idx], uy = io.v[1][idx], uz = io.v[2][idx];
        const A r0 = io.rho[idx];
        const A src = -ux * io.rd[0][idx] - uy * io.rd[1][idx] - uz * io.rd[2][idx] - r0 * (A)io.divv[idx];
        const A r1 = r0 + c.dt * (1.5 * src - 0.5 * io.f_rho[idx]);
        io.rho[idx] = r1;
        io.f_rho[idx] = src;
        rnew[it] = r1;
        const A am = io.amp[idx];
        const int pz = (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1);
        const A df = DZF ? (A)io.divf[idx] + rbuf[2 * ((l >> 1) * SEQ + lp<SH>(pz)) + (l & 1)] : (A)io.divf[idx];
        v = -(A)io.mu[idx] + ((mc.kp * mc.zeta) / (am + 0.001)) * (r1 - mc.rho0 - mc.kp * am) * (A)io.psi[idx] -
            df - ux * (A)io.gp[0][idx] - uy * (A)io.gp[1][idx] - uz * (A)io.gp[2][idx];
      }
      const int pos = (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1);
      rbuf[2 * ((l >> 1) * SEQ + lp<SH>(pos)) + (l & 1)] = v;
    }
  }
  __syncthreads();
  fft_stage<A, N, P, NT, MAXR, -1, 0>(buf, tw, stw);

  constexpr int W = P * HN;
  constexpr int ITW = (W + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const int w = tid + it * NT;
    if (W % NT == 0 || w < W) {
      const int cc = w / HN, k = w - cc * HN;
      const int kb = (k == 0) ? HN : N - k;
      A y0a, y1a, y0b, y1b;
      makhoul_post<A, N>(mk, k, kb, mk[k], buf[cc * SEQ + lp<SH>(k)], buf[cc * SEQ + lp<SH>(kb)], y0a, y1a, y0b,
                         y1b);
      const int l0 = 2 * cc;
      if (l0 < nvalid) {
        T* q = io.nlz + base + (long long)l0 * N;
        q[k] = (T)y0a;
        q[kb] = (T)y0b;
      }
      if (l0 + 1 < nvalid) {
        T* q = io.nlz + base + (long long)(l0 + 1) * N;
        q[k] = (T)y1a;
        q[kb] = (T)y1b;
      }
    }
  }
  // d_z rho_new: the next step's z component of grad rho
  if (io.rdz_next) {
    __syncthreads();
    dz_lines<T, N, P, NT>(rnew, buf, io.rdz_next + base, nvalid, tw, mk, io.qs);
  }
}

// d_z rho of the current rho, by the same tile and code as the density pass
// (so a state set or loaded gives the values a step would have left)
template <typename T, int N, int NT>
__global__ __launch_bounds__(NT) void rho_dz_kernel(StepConsts c, const double* __restrict__ rho, T* rdz,
                                                    double qs, const cplx<double>* __restrict__ tw,
                                                    const cplx<double>* __restrict__ mk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int P = density_pairs<N>(), TL = 2 * P, E = TL * N, IT = (E + NT - 1) / NT;
  const long long nl = (long long)c.nx * c.ny;
  const long long first = (long long)blockIdx.x * TL;
  const int nvalid = (int)((nl - first) < TL ? (nl - first) : TL);
  const long long base = (long long)blockIdx.y * nl * N + first * N;
  double x[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = threadIdx.x + it * NT;
    x[it] = 0.0;
    if ((E % NT == 0 || e < E) && e / N < nvalid) x[it] = rho[base + e];
  }
  dz_lines<T, N, P, NT>(x, reinterpret_cast<cplx<double>*>(smem_raw), rdz + base, nvalid, tw, mk, qs);
}

constexpr int kDensityNT = 128;

template <typename T, int N>
hipError_t launch_rho_dz(const StepConsts& c, const double* rho, T* rdz, double qs, const void* tw, const void* mk,
                         hipStream_t s) {
  constexpr int P = density_pairs<N>();
  const long long blocks = ((long long)c.nx * c.ny + 2 * P - 1) / (2 * P);
  const size_t lds = (size_t)P * seq_len(N, pick_maxr(N, P, kDensityNT)) * sizeof(cplx<double>);
  hipLaunchKernelGGL((rho_dz_kernel<T, N, kDensityNT>), dim3((unsigned)blocks, (unsigned)c.nmem),
                     dim3(kDensityNT), lds, s, c, rho, rdz, qs, static_cast<const cplx<double>*>(tw),
                     static_cast<const cplx<double>*>(mk));
  return hipGetLastError();
}

template <typename T, int N>
hipError_t launch_density(const StepConsts& c, const DensityIO<T>& io, const void* tw, const void* mk,
                          hipStream_t s) {
  constexpr int P = density_pairs<N>();
  const long long blocks = ((long long)c.nx * c.ny + 2 * P - 1) / (2 * P);
  const size_t lds = (size_t)P * seq_len(N, pick_maxr(N, P, kDensityNT)) * sizeof(cplx<double>);
  if (io.fz)
    hipLaunchKernelGGL((density_z_kernel<T, N, kDensityNT, true>), dim3((unsigned)blocks, (unsigned)c.nmem),
                       dim3(kDensityNT), lds, s, c, io, static_cast<const cplx<double>*>(tw),
                       static_cast<const cplx<double>*>(mk));
  else
    hipLaunchKernelGGL((density_z_kernel<T, N, kDensityNT, false>), dim3((unsigned)blocks, (unsigned)c.nmem),
                       dim3(kDensityNT), lds, s, c, io, static_cast<const cplx<double>*>(tw),
                       static_cast<const cplx<double>*>(mk));
  return hipGetLastError();
}

}  // namespace

bool density_zpass_supported(int nz) { return R2R::fft_supported(nz); }

template <typename T>
hipError_t density_zpass(const StepConsts& c, const DensityIO<T>& io, R2R& r2r, hipStream_t s) {
  if (!density_zpass_supported(c.nz)) return hipErrorNotSupported;
  const void *tw = nullptr, *mk = nullptr;
  hipError_t e = r2r.table_ptrs(c.nz, &tw, &mk);
  if (e != hipSuccess) return e;
  switch (c.nz) {
    case 16: return launch_density<T, 16>(c, io, tw, mk, s);
    case 32: return launch_density<T, 32>(c, io, tw, mk, s);
    case 48: return launch_density<T, 48>(c, io, tw, mk, s);
    case 64: return launch_density<T, 64>(c, io, tw, mk, s);
    case 96: return launch_density<T, 96>(c, io, tw, mk, s);
    case 128: return launch_density<T, 128>(c, io, tw, mk, s);
    case 192: return launch_density<T, 192>(c, io, tw, mk, s);
    case 256: return launch_density<T, 256>(c, io, tw, mk, s);
    case 384: return launch_density<T, 384>(c, io, tw, mk, s);
    case 512: return launch_density<T, 512>(c, io, tw, mk, s);
    case 1024: return launch_density<T, 1024>(c, io, tw, mk, s);
    default: return hipErrorNotSupported;
  }
}

namespace {

constexpr int kVelOut = 4;  // v_x, v_y, v_z, div v

template <int N, typename AT = double>
struct VelocityTile {
  static constexpr int P = N <= 256 ? 2 * short_line_scale<N>() : 1;  // line pairs per tile
  static constexpr int NT = 256;
  static constexpr size_t lds = (size_t)kVelOut * P * seq_len(N, pick_maxr(N, kVelOut * P, NT)) * sizeof(cplx<AT>);
  static constexpr bool fits = lds <= 65536;
};

// One workgroup: 2P consecutive z lines of cosine modes.  Compute phase:
// every mode of the tile is solved (velocity_mode, in AV arithmetic) from
// S_x one plane back, S_y one line back, S_z one mode back and fgrad^,
// erho^ at the mode; its four outputs go to LDS as the inputs of Makhoul's
// inverse (v_z at its reversed position: DST-III reads its input reversed).
// Then the pre-twiddle in place, one inverse FFT over the 4P sequences and
// the stores y_{2n} = u_n, y_{2n+1} = u_{N-1-n} (negated for v_z), each
// field to its own line: v_x of cosine line (i, j) is line (i-1, j) (cosine
// plane 0 gives the zero plane Nx-1), v_y is line (i, j-1) (cosine line j =
// 0 gives the zero line Ny-1).  AV: the solve's arithmetic; AT: the tile's
// and the tables' (FP64, or FP32 for FP32 storage where the transforms of
// div v and v compute in FP32).
template <typename T, typename AV, typename AT, int N, int P, int NT>
__global__ __launch_bounds__(NT) void velocity_z_kernel(StepConsts c, VelocityIO<T> io,
                                                        const cplx<AT>* __restrict__ tw,
                                                        const cplx<AT>* __restrict__ mk) {
  constexpr int NF = kVelOut, TL = 2 * P, HN = N / 2;
  constexpr int MAXR = pick_maxr(N, NF * P, NT);
  constexpr int SEQ = seq_len(N, MAXR), SH = pad_shift(N, MAXR);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using A = AT;
  cplx<A>* buf = reinterpret_cast<cplx<A>*>(smem_raw);
  A* rbuf = reinterpret_cast<A*>(smem_raw);
  const int tid = threadIdx.x;
  const long long nl = (long long)c.nx * c.ny;
  const long long first = (long long)blockIdx.x * TL;
  const int nvalid = (int)((nl - first) < TL ? (nl - first) : TL);
  const long long mo = (long long)blockIdx.y * nl * N;
  const long long base = mo + first * N;
  const MemberConsts mc = c.mc[blockIdx.y];
  StageTw<A, N, NF * P, NT, MAXR> stw;
  stw.load(tw, tid);

  constexpr int E = TL * N;
  constexpr int IT = (E + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + it * NT;
    if (E % NT == 0 || e < E) {
      const int l = e / N, k = e - l * N;
      A v[NF] = {A(0), A(0), A(0), A(0)};
      if (l < nvalid) {
        int i, j;
        line_ij(first + l, c.ny, i, j);
        const long long idx = base + e;
        // shifted reads outside the grid are 0, as on the first rank
        const AV Sx = i > 0 ? (AV)io.S[0][idx - (long long)c.ny * N] : AV(0);
        const AV Sy = j > 0 ? (AV)io.S[1][idx - N] : AV(0);
        const AV Sz = k > 0 ? (AV)io.S[2][idx - 1] : AV(0);
        const AV er = io.erq ? (AV)io.erq[idx] : AV(0);
        const VelMode<AV> r = velocity_mode<AV>(c, mc, i, j, k, Sx, Sy, Sz, (AV)io.fgq[idx], er);
        v[0] = (A)r.tx;
        v[1] = (A)r.ty;
        v[2] = (A)r.tz;
        v[3] = (A)r.div;
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        // v_z of mode k sits at k-1 (k = 0: the zero at N-1), reversed: N-k
        const int pos = f == 2 ? (k == 0 ? 0 : N - k) : k;
        rbuf[2 * ((f * P + (l >> 1)) * SEQ + lp<SH>(pos)) + (l & 1)] = v[f];
      }
    }
  }
  __syncthreads();
  // Makhoul's pre-twiddle, item k < N/2 in place at k and kb
  constexpr int W = NF * P * HN;
  constexpr int ITW = (W + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const int w = tid + it * NT;
    if (W % NT == 0 || w < W) {
      const int sq = w / HN, k = w - sq * HN;
      const int kb = (k == 0) ? HN : N - k;
      cplx<A> zk, zb;
      makhoul_pre<A, N>(mk, k, mk[k], buf[sq * SEQ + lp<SH>(k)], buf[sq * SEQ + lp<SH>(kb)], zk, zb);
      buf[sq * SEQ + lp<SH>(k)] = zk;
      buf[sq * SEQ + lp<SH>(kb)] = zb;
    }
  }
  __syncthreads();
  fft_stage<A, N, NF * P, NT, MAXR, +1, 0>(buf, tw, stw);
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const int w = tid + it * NT;
    if (W % NT == 0 || w < W) {
      const int sq = w / HN, n = w - sq * HN;
      const int f = sq / P, cc = sq - f * P;
      const cplx<A> ua = buf[sq * SEQ + lp<SH>(n)];
      const cplx<A> ub = buf[sq * SEQ + lp<SH>(N - 1 - n)];
      const A sg = f == 2 ? A(-1) : A(1);
      T* const o = f == 0 ? io.zout[0] : f == 1 ? io.zout[1] : f == 2 ? io.zout[2] : io.zout[3];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int l = 2 * cc + q;
        if (l < nvalid) {
          const long long g = first + l;
          int i, j;
          line_ij(g, c.ny, i, j);
          long long ol = g;
          if (f == 0) ol = i > 0 ? g - c.ny : g + (long long)(c.nx - 1) * c.ny;
          else if (f == 1) ol = j > 0 ? g - 1 : g + (c.ny - 1);
          T* const d = o + mo + ol * N + 2 * n;
          d[0] = (T)(q == 0 ? ua.x : ua.y);
          d[1] = (T)(sg * (q == 0 ? ub.x : ub.y));
        }
      }
    }
  }
}

template <typename T, typename AV, typename AT, int N>
hipError_t launch_velocity_z(const StepConsts& c, const VelocityIO<T>& io, const void* tw, const void* mk,
                             hipStream_t s) {
  using F = VelocityTile<N, AT>;
  if constexpr (!F::fits) {
    return hipErrorNotSupported;
  } else {
    constexpr int TL = 2 * F::P;
    const long long blocks = ((long long)c.nx * c.ny + TL - 1) / TL;
    hipLaunchKernelGGL((velocity_z_kernel<T, AV, AT, N, F::P, F::NT>), dim3((unsigned)blocks, (unsigned)c.nmem),
                       dim3(F::NT), F::lds, s, c, io, static_cast<const cplx<AT>*>(tw),
                       static_cast<const cplx<AT>*>(mk));
    return hipGetLastError();
  }
}

template <typename T, typename AV, typename AT>
hipError_t velocity_z_n(const StepConsts& c, const VelocityIO<T>& io, const void* tw, const void* mk,
                        hipStream_t s) {
  if constexpr (std::is_same<AT, float>::value) {
    // the lengths with FP32 table and tile instances (r2r_arith32_supported)
    switch (c.nz) {
      case 64: return launch_velocity_z<T, AV, AT, 64>(c, io, tw, mk, s);
      case 96: return launch_velocity_z<T, AV, AT, 96>(c, io, tw, mk, s);
      case 128: return launch_velocity_z<T, AV, AT, 128>(c, io, tw, mk, s);
      case 256: return launch_velocity_z<T, AV, AT, 256>(c, io, tw, mk, s);
      case 512: return launch_velocity_z<T, AV, AT, 512>(c, io, tw, mk, s);
      default: return hipErrorNotSupported;
    }
  } else {
    switch (c.nz) {
      case 16: return launch_velocity_z<T, AV, AT, 16>(c, io, tw, mk, s);
      case 32: return launch_velocity_z<T, AV, AT, 32>(c, io, tw, mk, s);
      case 48: return launch_velocity_z<T, AV, AT, 48>(c, io, tw, mk, s);
      case 64: return launch_velocity_z<T, AV, AT, 64>(c, io, tw, mk, s);
      case 96: return launch_velocity_z<T, AV, AT, 96>(c, io, tw, mk, s);
      case 128: return launch_velocity_z<T, AV, AT, 128>(c, io, tw, mk, s);
      case 192: return launch_velocity_z<T, AV, AT, 192>(c, io, tw, mk, s);
      case 256: return launch_velocity_z<T, AV, AT, 256>(c, io, tw, mk, s);
      case 384: return launch_velocity_z<T, AV, AT, 384>(c, io, tw, mk, s);
      case 512: return launch_velocity_z<T, AV, AT, 512>(c, io, tw, mk, s);
      default: return hipErrorNotSupported;
    }
  }
}

template <int N>
constexpr bool vel_fits() { return VelocityTile<N>::fits; }

}  // namespace

bool velocity_zpass_supported(int nz) {
  switch (nz) {
    case 16: return vel_fits<16>();
    case 32: return vel_fits<32>();
    case 48: return vel_fits<48>();
    case 64: return vel_fits<64>();
    case 96: return vel_fits<96>();
    case 128: return vel_fits<128>();
    case 192: return vel_fits<192>();
    case 256: return vel_fits<256>();
    case 384: return vel_fits<384>();
    case 512: return vel_fits<512>();
    default: return false;
  }
}

template <typename T>
hipError_t velocity_zpass(const StepConsts& c, const VelocityIO<T>& io, R2R& r2r, bool arith32, bool tile32,
                          hipStream_t s) {
  if (!velocity_zpass_supported(c.nz)) return hipErrorNotSupported;
  const bool t32 = sizeof(T) == 4 && tile32 && r2r_arith32_supported(c.nz);
  const void *tw = nullptr, *mk = nullptr;
  hipError_t e = r2r.table_ptrs(c.nz, &tw, &mk, t32 ? F32 : F64);
  if (e != hipSuccess) return e;
  if constexpr (sizeof(T) == 4) {
    if (t32) {
      if (arith32) return velocity_z_n<T, float, float>(c, io, tw, mk, s);
      return velocity_z_n<T, double, float>(c, io, tw, mk, s);
    }
    if (arith32) return velocity_z_n<T, float, double>(c, io, tw, mk, s);
  }
  return velocity_z_n<T, double, double>(c, io, tw, mk, s);
}

template hipError_t velocity_zpass<double>(const StepConsts&, const VelocityIO<double>&, R2R&, bool, bool, hipStream_t);
template hipError_t velocity_zpass<float>(const StepConsts&, const VelocityIO<float>&, R2R&, bool, bool, hipStream_t);

namespace {

// One workgroup: P consecutive z line pairs (the z tile of the one-axis
// passes).  Two one-axis operators on the same LDS tile, with a pointwise
// step between them:
//  1. psi in Makhoul's forward order, FFT, Makhoul post, the gradient
//     operator y[n] = -Vs[n] Y[n+1]/(2N) and the inverse pre-twiddle in
//     place, inverse FFT: item n holds d_z psi at 2n and 2n+1 of both lines
//     (DST-III: the odd output negated), stored to gpz;
//  2. the item reads psi, grad_x psi, grad_y psi at the same four points and
//     forms |amp| (the expression of the amplitude load of the z pass);
//     it writes the values back to its own two slots in Makhoul's forward
//     order (2n -> n, 2n+1 -> N-1-n), which it alone read;
//  3. FFT, Makhoul post, exp(-1.57^2 2 Vq[n]^2)/(2N) and the pre-twiddle,
//     inverse FFT, stores amp.
// AG: the arithmetic of step 3 (FP64, or FP32 with FP32 tables tw32, mk32
// for FP32 storage where the amplitude filter computes in FP32).
template <typename T, int N, int P, int NT, typename AG>
__global__ __launch_bounds__(NT) void amp_z_kernel(AmpIO<T> io, int nlines, const cplx<double>* __restrict__ tw,
                                                   const cplx<double>* __restrict__ mk,
                                                   const cplx<AG>* __restrict__ twg, const cplx<AG>* __restrict__ mkg) {
  using A = double;
  constexpr int TL = 2 * P, HN = N / 2;
  constexpr int MAXR = pick_maxr(N, P, NT);
  constexpr int SEQ = seq_len(N, MAXR), SH = pad_shift(N, MAXR);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<A>* buf = reinterpret_cast<cplx<A>*>(smem_raw);
  const int tid = threadIdx.x;
  const long long first = (long long)blockIdx.x * TL;
  const int nvalid = (int)((nlines - first) < TL ? (nlines - first) : TL);
  const long long base = (long long)blockIdx.y * nlines * N + first * N;
  StageTw<A, N, P, NT, MAXR> stw;
  stw.load(tw, tid);
  const MkSeq<A, N, P, NT> ms0(mk, 2, tid);

  // 1. psi -> LDS (DCT-II input order: v[m/2] = x[m] for even m, v[N-1-(m-1)/2] = x[m] for odd m)
  constexpr int E = P * N;
  constexpr int IT = (E + NT - 1) / NT;
  {
    cplx<A> v[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      const int c = e / N, m = e - c * N;
      v[it] = cplx<A>{A(0), A(0)};
      if (E % NT == 0 || e < E) {
        const int l0 = 2 * c;
        if (l0 < nvalid) v[it].x = (A)io.psi[base + (long long)l0 * N + m];
        if (l0 + 1 < nvalid) v[it].y = (A)io.psi[base + (long long)(l0 + 1) * N + m];
      }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      if (E % NT == 0 || e < E) {
        const int c = e / N, m = e - c * N;
        const int pos = (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1);
        buf[c * SEQ + lp<SH>(pos)] = v[it];
      }
    }
  }
  constexpr int W = P * HN;
  constexpr int ITW = (W + NT - 1) / NT;
  __syncthreads();
  fft_stage<A, N, P, NT, MAXR, -1, 0>(buf, tw, stw);
  z_axis_op_phase<kZGrad, N, P, NT, SEQ, SH>(buf, mk, ms0, io.qs);
  __syncthreads();
  fft_stage<A, N, P, NT, MAXR, +1, 0>(buf, tw, stw);

  // 2. item n: d_z psi at 2n, 2n+1 of both lines (DST-III: odd output
  // negated), stored; |amp| at the same points back into its two slots
  cplx<A> amv[std::is_same<AG, A>::value ? 1 : ITW][2];
  (void)amv;
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const int w = tid + it * NT;
    if (W % NT == 0 || w < W) {
      const int c = w / HN, n = w - c * HN;
      const cplx<A> ua = buf[c * SEQ + lp<SH>(n)];
      const cplx<A> ub = buf[c * SEQ + lp<SH>(N - 1 - n)];
      cplx<A> am0 = {A(0), A(0)}, am1 = {A(0), A(0)};  // (line 0, line 1) at 2n and 2n+1
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int l = 2 * c + q;
        if (l < nvalid) {
          const long long o = base + (long long)l * N + 2 * n;
          const A g0 = q == 0 ? ua.x : ua.y, g1 = -(q == 0 ? ub.x : ub.y);
          io.gpz[o] = (T)g0;
          io.gpz[o + 1] = (T)g1;
          // the amplitude load's expression, from the stored (rounded) d_z psi
          const A gz0 = (A)(T)g0, gz1 = (A)(T)g1;
          const A p0 = io.psi[o], p1 = io.psi[o + 1];
          const A x0 = io.gpx[o], x1 = io.gpx[o + 1], y0 = io.gpy[o], y1 = io.gpy[o + 1];
          const A a0 = sqrt(io.q02 * (p0 * p0) + x0 * x0 + y0 * y0 + gz0 * gz0);
          const A a1 = sqrt(io.q02 * (p1 * p1) + x1 * x1 + y1 * y1 + gz1 * gz1);
          if (q == 0) { am0.x = a0; am1.x = a1; } else { am0.y = a0; am1.y = a1; }
        }
      }
      if constexpr (std::is_same<AG, A>::value) {
        buf[c * SEQ + lp<SH>(n)] = am0;
        buf[c * SEQ + lp<SH>(N - 1 - n)] = am1;
      } else {
        amv[it][0] = am0;
        amv[it][1] = am1;
      }
    }
  }
  // 3. the z factor of the Gaussian filter on |amp|, in the AG tile (an FP32
  // view of the same LDS, written once every FP64 slot has been read)
  cplx<AG>* const bg = reinterpret_cast<cplx<AG>*>(smem_raw);
  if constexpr (!std::is_same<AG, A>::value) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      const int w = tid + it * NT;
      if (W % NT == 0 || w < W) {
        const int c = w / HN, n = w - c * HN;
        bg[c * SEQ + lp<SH>(n)] = cplx<AG>{(AG)amv[it][0].x, (AG)amv[it][0].y};
        bg[c * SEQ + lp<SH>(N - 1 - n)] = cplx<AG>{(AG)amv[it][1].x, (AG)amv[it][1].y};
      }
    }
  }
  __syncthreads();
  StageTw<AG, N, P, NT, MAXR> stg;
  stg.load(twg, tid);
  const MkSeq<AG, N, P, NT> msg(mkg, 2, tid);
  fft_stage<AG, N, P, NT, MAXR, -1, 0>(bg, twg, stg);
  z_axis_op_phase<kZGauss, N, P, NT, SEQ, SH, AG>(bg, mkg, msg, io.qs);
  __syncthreads();
  fft_stage<AG, N, P, NT, MAXR, +1, 0>(bg, twg, stg);
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const int w = tid + it * NT;
    if (W % NT == 0 || w < W) {
      const int c = w / HN, n = w - c * HN;
      const cplx<AG> ua = bg[c * SEQ + lp<SH>(n)];
      const cplx<AG> ub = bg[c * SEQ + lp<SH>(N - 1 - n)];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int l = 2 * c + q;
        if (l < nvalid) {
          const long long o = base + (long long)l * N + 2 * n;
          io.amp[o] = (T)(q == 0 ? ua.x : ua.y);
          io.amp[o + 1] = (T)(q == 0 ? ub.x : ub.y);
        }
      }
    }
  }
}

template <typename T, int N, typename AG>
hipError_t launch_amp(const StepConsts& c, const AmpIO<T>& io, const void* tw, const void* mk, const void* twg,
                      const void* mkg, hipStream_t s) {
  using C = TileChoice<T, N>;
  constexpr int P = C::PZ, NT = C::NTZ, TL = 2 * P;
  const long long nl = (long long)c.nx * c.ny;
  const long long blocks = (nl + TL - 1) / TL;
  const size_t lds = (size_t)P * seq_len(N, pick_maxr(N, P, NT)) * sizeof(cplx<double>);
  hipLaunchKernelGGL((amp_z_kernel<T, N, P, NT, AG>), dim3((unsigned)blocks, (unsigned)c.nmem), dim3(NT), lds, s,
                     io, (int)nl, static_cast<const cplx<double>*>(tw), static_cast<const cplx<double>*>(mk),
                     static_cast<const cplx<AG>*>(twg), static_cast<const cplx<AG>*>(mkg));
  return hipGetLastError();
}

template <typename T, typename AG>
hipError_t amp_z_n(const StepConsts& c, const AmpIO<T>& io, const void* tw, const void* mk, const void* twg,
                   const void* mkg, hipStream_t s) {
  if constexpr (std::is_same<AG, float>::value) {
    // the lengths with FP32 tables (r2r_arith32_supported)
    switch (c.nz) {
      case 64: return launch_amp<T, 64, AG>(c, io, tw, mk, twg, mkg, s);
      case 96: return launch_amp<T, 96, AG>(c, io, tw, mk, twg, mkg, s);
      case 128: return launch_amp<T, 128, AG>(c, io, tw, mk, twg, mkg, s);
      case 256: return launch_amp<T, 256, AG>(c, io, tw, mk, twg, mkg, s);
      case 512: return launch_amp<T, 512, AG>(c, io, tw, mk, twg, mkg, s);
      default: return hipErrorNotSupported;
    }
  } else {
    switch (c.nz) {
      case 16: return launch_amp<T, 16, AG>(c, io, tw, mk, twg, mkg, s);
      case 32: return launch_amp<T, 32, AG>(c, io, tw, mk, twg, mkg, s);
      case 48: return launch_amp<T, 48, AG>(c, io, tw, mk, twg, mkg, s);
      case 64: return launch_amp<T, 64, AG>(c, io, tw, mk, twg, mkg, s);
      case 96: return launch_amp<T, 96, AG>(c, io, tw, mk, twg, mkg, s);
      case 128: return launch_amp<T, 128, AG>(c, io, tw, mk, twg, mkg, s);
      case 192: return launch_amp<T, 192, AG>(c, io, tw, mk, twg, mkg, s);
      case 256: return launch_amp<T, 256, AG>(c, io, tw, mk, twg, mkg, s);
      case 384: return launch_amp<T, 384, AG>(c, io, tw, mk, twg, mkg, s);
      case 512: return launch_amp<T, 512, AG>(c, io, tw, mk, twg, mkg, s);
      case 1024: return launch_amp<T, 1024, AG>(c, io, tw, mk, twg, mkg, s);
      default: return hipErrorNotSupported;
    }
  }
}

}  // namespace

bool amp_zpass_supported(int nz) { return R2R::fft_supported(nz); }

template <typename T>
hipError_t amp_zpass(const StepConsts& c, const AmpIO<T>& io, R2R& r2r, bool gauss32, hipStream_t s) {
  if (!amp_zpass_supported(c.nz)) return hipErrorNotSupported;
  const void *tw = nullptr, *mk = nullptr;
  hipError_t e = r2r.table_ptrs(c.nz, &tw, &mk);
  if (e != hipSuccess) return e;
  if constexpr (sizeof(T) == 4) {
    if (gauss32 && r2r_arith32_supported(c.nz)) {
      const void *tw32 = nullptr, *mk32 = nullptr;
      if ((e = r2r.table_ptrs(c.nz, &tw32, &mk32, F32)) != hipSuccess) return e;
      return amp_z_n<T, float>(c, io, tw, mk, tw32, mk32, s);
    }
  }
  return amp_z_n<T, double>(c, io, tw, mk, tw, mk, s);
}

template hipError_t amp_zpass<double>(const StepConsts&, const AmpIO<double>&, R2R&, bool, hipStream_t);
template hipError_t amp_zpass<float>(const StepConsts&, const AmpIO<float>&, R2R&, bool, hipStream_t);

namespace {

// One workgroup: P z line pairs of 128 threads, 2P complex sequences in LDS
// (psi of pair c at sequence c, L psi of pair c at sequence P + c).  Item
// (c, k), k < N/2, takes modes k and kb = N - k (N/2 for k = 0) of both lines
// of pair c: UPD forms the new psiq there (the update of OP_LD_PSI) and
// stores it, else it reads psiq; L psi's mode is alpha (q0^2 - |q|^2) times
// it (the load of OP_LD_SH).  Makhoul's inverse pre-twiddle puts both into
// their sequences, one inverse FFT runs over all 2P, and y_2n = u_n,
// y_2n+1 = u_{N-1-n} are stored: psi's z lines in FP64, L psi's in T.
template <typename T, int N, int P, int NT, bool UPD>
__global__ __launch_bounds__(NT) void psi_sh_z_kernel(PsiShIO<T> io, int nlines, int ny,
                                                     const cplx<double>* __restrict__ tw,
                                                     const cplx<double>* __restrict__ mk) {
  using A = double;
  constexpr int NS = 2 * P, TL = 2 * P, HN = N / 2;
  constexpr int MAXR = pick_maxr(N, NS, NT);
  constexpr int SEQ = seq_len(N, MAXR), SH = pad_shift(N, MAXR);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<A>* buf = reinterpret_cast<cplx<A>*>(smem_raw);
  const int tid = threadIdx.x;
  const long long first = (long long)blockIdx.x * TL;
  const int nvalid = (int)((nlines - first) < TL ? (nlines - first) : TL);
  const long long base = (long long)blockIdx.y * nlines * N + first * N;
  const Fuse& f = io.f;
  StageTw<A, N, NS, NT, MAXR> stw;
  stw.load(tw, tid);

  constexpr int W = P * HN;
  constexpr int ITW = (W + NT - 1) / NT;
  cplx<A> pk[ITW], pb[ITW], sk[ITW], sb[ITW];  // psi and L psi at modes k, kb (line 0, line 1)
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const int w = tid + it * NT;
    pk[it] = pb[it] = sk[it] = sb[it] = cplx<A>{A(0), A(0)};
    if (W % NT == 0 || w < W) {
      const int c = w / HN, k = w - c * HN;
      const int kb = (k == 0) ? HN : N - k;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int l = 2 * c + q;
        if (l < nvalid) {
          int i, j;
          line_ij(first + l, ny, i, j);
          const long long o = base + (long long)l * N;
          A ra, rb;
          if constexpr (UPD) {
            ra = psi_update_value<T>(f, o + k, i, j, k);
            rb = psi_update_value<T>(f, o + kb, i, j, kb);
            f.psiq_new[o + k] = ra;
            f.psiq_new[o + kb] = rb;
          } else {
            ra = io.psiq[o + k];
            rb = io.psiq[o + kb];
          }
          const A la = f.alpha * (f.q02 - mq2c(f, i, j, k)) * ra;
          const A lb = f.alpha * (f.q02 - mq2c(f, i, j, kb)) * rb;
          if (q == 0) {
            pk[it].x = ra; pb[it].x = rb; sk[it].x = la; sb[it].x = lb;
          } else {
            pk[it].y = ra; pb[it].y = rb; sk[it].y = la; sb[it].y = lb;
          }
        }
      }
    }
  }
  MkSeq<A, N, P, NT> ms(mk, 2, tid);
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const int w = tid + it * NT;
    if (W % NT == 0 || w < W) {
      const int c = w / HN, k = w - c * HN;
      const int kb = (k == 0) ? HN : N - k;
      const cplx<A> mkk = ms.get(mk, 2, k);
      cplx<A> zk, zb;
      makhoul_pre<A, N>(mk, k, mkk, pk[it], pb[it], zk, zb);
      buf[c * SEQ + lp<SH>(k)] = zk;
      buf[c * SEQ + lp<SH>(kb)] = zb;
      makhoul_pre<A, N>(mk, k, mkk, sk[it], sb[it], zk, zb);
      buf[(P + c) * SEQ + lp<SH>(k)] = zk;
      buf[(P + c) * SEQ + lp<SH>(kb)] = zb;
    }
    ms.advance(2, it);
  }
  __syncthreads();
  fft_stage<A, N, NS, NT, MAXR, +1, 0>(buf, tw, stw);

  constexpr int WO = NS * HN;
  constexpr int ITO = (WO + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < ITO; ++it) {
    const int w = tid + it * NT;
    if (WO % NT == 0 || w < WO) {
      const int c = w / HN, n = w - c * HN;
      const cplx<A> ua = buf[c * SEQ + lp<SH>(n)];
      const cplx<A> ub = buf[c * SEQ + lp<SH>(N - 1 - n)];
      const bool lz = c >= P;
      const int pc = lz ? c - P : c;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int l = 2 * pc + q;
        if (l < nvalid) {
          const long long o = base + (long long)l * N + 2 * n;
          const A v0 = q == 0 ? ua.x : ua.y, v1 = q == 0 ? ub.x : ub.y;
          if (lz) *reinterpret_cast<cplx<T>*>(io.zlpsi + o) = cplx<T>{(T)v0, (T)v1};
          else *reinterpret_cast<cplx<double>*>(io.zpsi + o) = cplx<double>{v0, v1};
        }
      }
    }
  }
}

template <typename T, int N, bool UPD>
hipError_t launch_psi_sh(const StepConsts& c, const PsiShIO<T>& io, const void* tw, const void* mk,
                         hipStream_t s) {
  // the general z tile's line pairs with twice its sequences on 128 threads
  constexpr int P = TileChoice<double, N>::PZ, NT = 128, TL = 2 * P;
  const long long nl = (long long)c.nx * c.ny;
  const long long blocks = (nl + TL - 1) / TL;
  const size_t lds = (size_t)2 * P * seq_len(N, pick_maxr(N, 2 * P, NT)) * sizeof(cplx<double>);
  hipLaunchKernelGGL((psi_sh_z_kernel<T, N, P, NT, UPD>), dim3((unsigned)blocks, (unsigned)c.nmem), dim3(NT), lds,
                     s, io, (int)nl, c.ny, static_cast<const cplx<double>*>(tw),
                     static_cast<const cplx<double>*>(mk));
  return hipGetLastError();
}

template <typename T, bool UPD>
hipError_t psi_sh_n(const StepConsts& c, const PsiShIO<T>& io, const void* tw, const void* mk, hipStream_t s) {
  switch (c.nz) {
    case 16: return launch_psi_sh<T, 16, UPD>(c, io, tw, mk, s);
    case 32: return launch_psi_sh<T, 32, UPD>(c, io, tw, mk, s);
    case 48: return launch_psi_sh<T, 48, UPD>(c, io, tw, mk, s);
    case 64: return launch_psi_sh<T, 64, UPD>(c, io, tw, mk, s);
    case 96: return launch_psi_sh<T, 96, UPD>(c, io, tw, mk, s);
    case 128: return launch_psi_sh<T, 128, UPD>(c, io, tw, mk, s);
    case 192: return launch_psi_sh<T, 192, UPD>(c, io, tw, mk, s);
    case 256: return launch_psi_sh<T, 256, UPD>(c, io, tw, mk, s);
    case 384: return launch_psi_sh<T, 384, UPD>(c, io, tw, mk, s);
    case 512: return launch_psi_sh<T, 512, UPD>(c, io, tw, mk, s);
    case 1024: return launch_psi_sh<T, 1024, UPD>(c, io, tw, mk, s);
    default: return hipErrorNotSupported;
  }
}

}  // namespace

bool psi_sh_zpass_supported(int nz) { return R2R::fft_supported(nz); }

template <typename T>
hipError_t psi_sh_zpass(const StepConsts& c, const PsiShIO<T>& io, bool update, R2R& r2r, hipStream_t s) {
  if (!psi_sh_zpass_supported(c.nz)) return hipErrorNotSupported;
  const void *tw = nullptr, *mk = nullptr;
  hipError_t e = r2r.table_ptrs(c.nz, &tw, &mk);
  if (e != hipSuccess) return e;
  return update ? psi_sh_n<T, true>(c, io, tw, mk, s) : psi_sh_n<T, false>(c, io, tw, mk, s);
}

template hipError_t psi_sh_zpass<double>(const StepConsts&, const PsiShIO<double>&, bool, R2R&, hipStream_t);
template hipError_t psi_sh_zpass<float>(const StepConsts&, const PsiShIO<float>&, bool, R2R&, hipStream_t);

template <typename T>
hipError_t rho_dz(const StepConsts& c, const double* rho, T* rdz, double qs, R2R& r2r, hipStream_t s) {
  if (!density_zpass_supported(c.nz)) return hipErrorNotSupported;
  const void *tw = nullptr, *mk = nullptr;
  hipError_t e = r2r.table_ptrs(c.nz, &tw, &mk);
  if (e != hipSuccess) return e;
  switch (c.nz) {
    case 16: return launch_rho_dz<T, 16>(c, rho, rdz, qs, tw, mk, s);
    case 32: return launch_rho_dz<T, 32>(c, rho, rdz, qs, tw, mk, s);
    case 48: return launch_rho_dz<T, 48>(c, rho, rdz, qs, tw, mk, s);
    case 64: return launch_rho_dz<T, 64>(c, rho, rdz, qs, tw, mk, s);
    case 96: return launch_rho_dz<T, 96>(c, rho, rdz, qs, tw, mk, s);
    case 128: return launch_rho_dz<T, 128>(c, rho, rdz, qs, tw, mk, s);
    case 192: return launch_rho_dz<T, 192>(c, rho, rdz, qs, tw, mk, s);
    case 256: return launch_rho_dz<T, 256>(c, rho, rdz, qs, tw, mk, s);
    case 384: return launch_rho_dz<T, 384>(c, rho, rdz, qs, tw, mk, s);
    case 512: return launch_rho_dz<T, 512>(c, rho, rdz, qs, tw, mk, s);
    case 1024: return launch_rho_dz<T, 1024>(c, rho, rdz, qs, tw, mk, s);
    default: return hipErrorNotSupported;
  }
}

template hipError_t rho_dz<double>(const StepConsts&, const double*, double*, double, R2R&, hipStream_t);
template hipError_t rho_dz<float>(const StepConsts&, const double*, float*, double, R2R&, hipStream_t);

template hipError_t density_zpass<double>(const StepConsts&, const DensityIO<double>&, R2R&, hipStream_t);
template hipError_t density_zpass<float>(const StepConsts&, const DensityIO<float>&, R2R&, hipStream_t);

template hipError_t forces_zpass<double>(const StepConsts&, const ForcesIO<double>&, R2R&, hipStream_t);
template hipError_t forces_zpass<float>(const StepConsts&, const ForcesIO<float>&, R2R&, hipStream_t);

}  // namespace smaiso
