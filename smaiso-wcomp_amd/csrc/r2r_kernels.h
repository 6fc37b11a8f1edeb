// Device code of the r2r axis passes (included by r2r.hip and by the
// translation units of the autotuner's tile variants).
//
// LDS-staged batched r2r transforms for gfx950.
//
// Method (per axis pass):
//  * A workgroup owns a tile of TL = 2P lines.  For the contiguous z axis the
//    tile is TL consecutive lines; for the strided y/x axes it is TL
//    consecutive k columns at one outer index, so every global access is a
//    run of TL*sizeof(T) bytes.
//  * Two real lines form one complex sequence z = a + i b (two real FFTs for
//    the price of one).
//  * DCT-II (REDFT10) uses Makhoul's reordering (J. Makhoul, "A fast cosine
//    transform in one and two dimensions", IEEE TASSP 28(1), 1980):
//    v = (x0, x2, x4, ..., x5, x3, x1), Y_k = 2 Re(e^{-i pi k/2N} V_k) with
//    V = DFT_N(v).  DCT-III (REDFT01) is its transpose:
//    U_k = e^{+i pi k/2N}(Y_k - i Y_{N-k}), u = IDFT_N(U) (unnormalised),
//    y_{2n} = u_n, y_{2n+1} = u_{N-1-n}.
//  * DST-II (RODFT10)  = reverse(DCT-II((-1)^n x_n)) and
//    DST-III (RODFT01) = (-1)^k DCT-III(reverse(x))_k, folded into the load
//    and store index maps.
//  * The complex DFT_N is a self-sorting Stockham FFT in LDS with radix 2, 3,
//    4, 8 and 16 butterflies in registers and a per-N twiddle table.
//  * The tile, the twiddles and every operation on them are FP64 for both
//    storage types; a pass rounds to the storage type only when it stores.
//    An FP32 FFT's rounding is relative to the norm of each line and swamps
//    the small high modes that the spectral operators then amplify, while
//    rounding a stored value is relative to that value.  (Without packed
//    FP32 instructions, FP64 VALU arithmetic costs what FP32 costs.)
//  * Lengths outside the FFT list use a direct O(N^2) LDS kernel that
//    evaluates FFTW's defining sums with a cosine table.
#pragma once

#include "r2r.h"

#include <cmath>
#include <cstdint>

namespace smaiso {

// ---------------------------------------------------------------------------
// butterflies
// ---------------------------------------------------------------------------
template <typename T, int SIGN>
__device__ __forceinline__ cplx<T> mul_i(cplx<T> a) {
  // a * (SIGN * i)
  if constexpr (SIGN < 0) return {a.y, -a.x};
  else return {-a.y, a.x};
}

template <typename T, int SIGN>
__device__ __forceinline__ void dft2(cplx<T>& a0, cplx<T>& a1) {
  cplx<T> t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}

template <typename T, int SIGN>
__device__ __forceinline__ void dft4(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2, cplx<T>& a3) {
  cplx<T> t0 = cadd(a0, a2), t1 = csub(a0, a2);
  cplx<T> t2 = cadd(a1, a3), t3 = mul_i<T, SIGN>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <typename T, int SIGN>
__device__ __forceinline__ void dft3(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2) {
  const T h = T(0.8660254037844386);  // sqrt(3)/2
  cplx<T> t = cadd(a1, a2);
  cplx<T> d = csub(a1, a2);
  cplx<T> u = mul_i<T, SIGN>(cplx<T>{d.x * h, d.y * h});
  cplx<T> m = {a0.x - T(0.5) * t.x, a0.y - T(0.5) * t.y};
  a0 = cadd(a0, t);
  a1 = cadd(m, u);
  a2 = csub(m, u);
}

// w8^s = exp(SIGN i pi s / 4)
template <typename T, int SIGN, int S>
__device__ __forceinline__ cplx<T> tw8(cplx<T> a) {
  const T r = T(0.7071067811865476);
  if constexpr (S == 0) return a;
  else if constexpr (S == 1) return cmul(a, cplx<T>{r, T(SIGN) * r});
  else if constexpr (S == 2) return mul_i<T, SIGN>(a);
  else return cmul(a, cplx<T>{-r, T(SIGN) * r});
}

template <typename T, int SIGN>
__device__ __forceinline__ void dft8(cplx<T>* a) {
  cplx<T> e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  cplx<T> o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<T, SIGN>(e0, e1, e2, e3);
  dft4<T, SIGN>(o0, o1, o2, o3);
  o1 = tw8<T, SIGN, 1>(o1);
  o2 = tw8<T, SIGN, 2>(o2);
  o3 = tw8<T, SIGN, 3>(o3);
  a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
  a[1] = cadd(e1, o1); a[5] = csub(e1, o1);
  a[2] = cadd(e2, o2); a[6] = csub(e2, o2);
  a[3] = cadd(e3, o3); a[7] = csub(e3, o3);
}

template <typename T, int SIGN>
__device__ __forceinline__ void dft16(cplx<T>* a) {
  cplx<T> e[8], o[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) { e[q] = a[2 * q]; o[q] = a[2 * q + 1]; }
  dft8<T, SIGN>(e);
  dft8<T, SIGN>(o);
  // w16^s = exp(SIGN i pi s / 8)
  const T c1 = T(0.9238795325112867), s1 = T(0.3826834323650898), r = T(0.7071067811865476);
  const T sg = T(SIGN);
  o[1] = cmul(o[1], cplx<T>{c1, sg * s1});
  o[2] = cmul(o[2], cplx<T>{r, sg * r});
  o[3] = cmul(o[3], cplx<T>{s1, sg * c1});
  o[4] = mul_i<T, SIGN>(o[4]);
  o[5] = cmul(o[5], cplx<T>{-s1, sg * c1});
  o[6] = cmul(o[6], cplx<T>{-r, sg * r});
  o[7] = cmul(o[7], cplx<T>{-c1, sg * s1});
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    a[q] = cadd(e[q], o[q]);
    a[q + 8] = csub(e[q], o[q]);
  }
}

template <typename T, int R, int SIGN>
__device__ __forceinline__ void dft(cplx<T>* a) {
  if constexpr (R == 2) dft2<T, SIGN>(a[0], a[1]);
  else if constexpr (R == 3) dft3<T, SIGN>(a[0], a[1], a[2]);
  else if constexpr (R == 4) dft4<T, SIGN>(a[0], a[1], a[2], a[3]);
  else if constexpr (R == 8) dft8<T, SIGN>(a);
  else if constexpr (R == 16) dft16<T, SIGN>(a);
}

// ---------------------------------------------------------------------------
// compile-time FFT plans: radix of stage s for length n when no radix may
// exceed maxr (16, 8 or 4); 0 ends the plan
// ---------------------------------------------------------------------------
constexpr int fft_radix(int n, int maxr, int s) {
  if (maxr >= 16) {
    switch (n) {
      case 16: { const int r[] = {16, 0}; return r[s < 1 ? s : 1]; }
      case 32: { const int r[] = {16, 2, 0}; return r[s < 2 ? s : 2]; }
      case 48: { const int r[] = {16, 3, 0}; return r[s < 2 ? s : 2]; }
      case 64: { const int r[] = {16, 4, 0}; return r[s < 2 ? s : 2]; }
      case 96: { const int r[] = {16, 3, 2, 0}; return r[s < 3 ? s : 3]; }
      case 128: { const int r[] = {16, 8, 0}; return r[s < 2 ? s : 2]; }
      case 192: { const int r[] = {16, 4, 3, 0}; return r[s < 3 ? s : 3]; }
      case 256: { const int r[] = {16, 16, 0}; return r[s < 2 ? s : 2]; }
      case 384: { const int r[] = {16, 8, 3, 0}; return r[s < 3 ? s : 3]; }
      case 512: { const int r[] = {16, 8, 4, 0}; return r[s < 3 ? s : 3]; }
      case 1024: { const int r[] = {16, 16, 4, 0}; return r[s < 3 ? s : 3]; }
      default: return 0;
    }
  } else if (maxr >= 8) {
    switch (n) {
      case 16: { const int r[] = {8, 2, 0}; return r[s < 2 ? s : 2]; }
      case 32: { const int r[] = {8, 4, 0}; return r[s < 2 ? s : 2]; }
      case 48: { const int r[] = {8, 3, 2, 0}; return r[s < 3 ? s : 3]; }
      case 64: { const int r[] = {8, 8, 0}; return r[s < 2 ? s : 2]; }
      case 96: { const int r[] = {8, 4, 3, 0}; return r[s < 3 ? s : 3]; }
      case 128: { const int r[] = {8, 4, 4, 0}; return r[s < 3 ? s : 3]; }
      case 192: { const int r[] = {8, 8, 3, 0}; return r[s < 3 ? s : 3]; }
      case 256: { const int r[] = {8, 8, 4, 0}; return r[s < 3 ? s : 3]; }
      case 384: { const int r[] = {8, 8, 3, 2, 0}; return r[s < 4 ? s : 4]; }
      case 512: { const int r[] = {8, 8, 8, 0}; return r[s < 3 ? s : 3]; }
      case 1024: { const int r[] = {8, 8, 4, 4, 0}; return r[s < 4 ? s : 4]; }
      default: return 0;
    }
  } else {
    switch (n) {
      case 16: { const int r[] = {4, 4, 0}; return r[s < 2 ? s : 2]; }
      case 32: { const int r[] = {4, 4, 2, 0}; return r[s < 3 ? s : 3]; }
      case 48: { const int r[] = {4, 4, 3, 0}; return r[s < 3 ? s : 3]; }
      case 64: { const int r[] = {4, 4, 4, 0}; return r[s < 3 ? s : 3]; }
      case 96: { const int r[] = {4, 4, 3, 2, 0}; return r[s < 4 ? s : 4]; }
      case 128: { const int r[] = {4, 4, 4, 2, 0}; return r[s < 4 ? s : 4]; }
      case 192: { const int r[] = {4, 4, 4, 3, 0}; return r[s < 4 ? s : 4]; }
      case 256: { const int r[] = {4, 4, 4, 4, 0}; return r[s < 4 ? s : 4]; }
      case 384: { const int r[] = {4, 4, 4, 3, 2, 0}; return r[s < 5 ? s : 5]; }
      case 512: { const int r[] = {4, 4, 4, 4, 2, 0}; return r[s < 5 ? s : 5]; }
      case 1024: { const int r[] = {4, 4, 4, 4, 4, 0}; return r[s < 5 ? s : 5]; }
      default: return 0;
    }
  }
}

constexpr int ns_before(int n, int maxr, int s) {
  int p = 1;
  for (int i = 0; i < s; ++i) p *= fft_radix(n, maxr, i);
  return p;
}

// Largest radix that still gives every thread of the block a butterfly.
constexpr int pick_maxr(int n, int p, int nt) {
  return (p * n / 16 >= nt) ? 16 : (p * n / 8 >= nt) ? 8 : 4;
}

// LDS layout of a sequence: item x sits at x + (x >> SH), one spare slot
// after every 2^SH items, with SH = log2 of the plan's first radix R0.  The
// first stage writes with a stride of R0 items across lanes and the later
// stages in runs of NS; with the spare slots both map the lanes that LDS
// serves together onto distinct banks (a plain layout puts every lane of a
// stride-R0 write on the same few banks).  A sequence takes seq_len items,
// the last one spare so that neighbouring sequences start on other banks.
constexpr int pad_shift(int n, int maxr) {
  return fft_radix(n, maxr, 0) >= 16 ? 4 : fft_radix(n, maxr, 0) >= 8 ? 3 : 2;
}
constexpr int seq_len(int n, int maxr) { return n + (n >> pad_shift(n, maxr)) + 1; }
template <int SH>
__device__ __forceinline__ int lp(int x) { return x + (x >> SH); }
// lp(x0 + q STEP) for x0 >= 0, as lp(x0) plus a constant where STEP is a
// multiple of 2^SH, so that a thread's accesses for q = 0, 1, ... share one
// address register with immediate offsets (written as lp of the sum, the
// compiler keeps one register per access)
template <int SH, int STEP>
__device__ __forceinline__ int lp_step(int x0, int q) {
  if constexpr (STEP % (1 << SH) == 0) return lp<SH>(x0) + q * (STEP + (STEP >> SH));
  else return lp<SH>(x0 + q * STEP);
}

// T: storage type of the input and of the load ops' work-type sources;
// TO: storage type of the output (T unless a pass changes the storage type)
template <typename T, typename TO = T>
struct PassArgs {
  const T* in;
  const T* in2;
  TO* out;
  // twiddle and Makhoul tables of the tile's arithmetic type (cplx<A>: FP64
  // unless the pass computes in FP32)
  const void* tw;
  const void* mk;
  int nx, ny, nz;
  int axis;
  int kind;
  int vec;  // 16-byte pair accesses allowed (nz even, pointers 16-byte aligned)
  Fuse f;
};

struct TileGeo {
  long long base;
  long long line_step;
  long long stride_m;
  int nvalid;
  int outer;  // y/x passes: the fixed index (i for y, j for x); z pass: first line
  int k0;     // y/x passes: first column
};

template <int TL, int N>
__device__ __forceinline__ TileGeo tile_geometry(int axis, int nx, int ny, int nz) {
  TileGeo t;
  if (axis == 2) {
    long long first = (long long)blockIdx.x * TL;
    long long nl = (long long)nx * ny;
    t.nvalid = (int)((nl - first) < TL ? (nl - first) : TL);
    t.base = first * N;
    t.line_step = N;
    t.stride_m = 1;
    t.outer = (int)first;
    t.k0 = 0;
  } else {
    // y passes: the plane index i varies fastest over the workgroups, so the
    // tiles in flight spread over the whole field (as along x) rather than
    // covering a few consecutive planes
    const int tpo = (nz + TL - 1) / TL;
    int o, kc;
    if (axis == 1) { kc = blockIdx.x / nx; o = blockIdx.x - kc * nx; }
    else { o = blockIdx.x / tpo; kc = blockIdx.x - o * tpo; }
    const int k0 = kc * TL;
    t.nvalid = (nz - k0) < TL ? (nz - k0) : TL;
    t.line_step = 1;
    t.outer = o;
    t.k0 = k0;
    if (axis == 1) {
      t.base = (long long)o * ny * nz + k0;
      t.stride_m = nz;
    } else {
      t.base = (long long)o * nz + k0;
      t.stride_m = (long long)ny * nz;
    }
  }
  return t;
}

// A thread's twiddle base w = tw[k STRIDE] of each stage, loaded once before
// the tile's data: in a stage whose butterfly count per sequence J divides
// NT, every iteration of the thread takes butterfly j = tid mod J and so the
// same k = j mod NS.  Stages where J does not divide NT load w in the stage.
template <typename T, int N, int P, int NT, int MAXR>
struct StageTw {
  static constexpr int kStages = 6;
  cplx<T> w[kStages];
  template <int S>
  static constexpr bool fixed() {
    return fft_radix(N, MAXR, S) != 0 && ns_before(N, MAXR, S) > 1 && NT % (N / fft_radix(N, MAXR, S)) == 0;
  }
  __device__ __forceinline__ void load(const cplx<T>* __restrict__ tw, int tid) { load_s<0>(tw, tid); }
  template <int S>
  __device__ __forceinline__ void load_s(const cplx<T>* __restrict__ tw, int tid) {
    if constexpr (S < kStages && fft_radix(N, MAXR, S) != 0) {
      if constexpr (fixed<S>()) {
        constexpr int R = fft_radix(N, MAXR, S), NS = ns_before(N, MAXR, S), J = N / R;
        w[S] = tw[((tid % J) % NS) * (N / (NS * R))];
      }
      load_s<S + 1>(tw, tid);
    }
  }
};

// Sequence p of the tile lives at buf[p * SEQ], item x of it at
// buf[p * SEQ + lp<SH>(x)] (SEQ = seq_len(N, MAXR), SH = pad_shift(N, MAXR)).
template <typename T, int N, int P, int NT, int MAXR, int SIGN, int S, int SEQ = seq_len(N, MAXR),
          int SH = pad_shift(N, MAXR)>
__device__ __forceinline__ void fft_stage(cplx<T>* buf, const cplx<T>* __restrict__ tw,
                                          const StageTw<T, N, P, NT, MAXR>& pre) {
  constexpr int R = fft_radix(N, MAXR, S);
  if constexpr (R != 0) {
    constexpr int NS = ns_before(N, MAXR, S);
    constexpr int J = N / R;
    constexpr int TOT = P * J;
    constexpr int IT = (TOT + NT - 1) / NT;
    constexpr int STRIDE = N / (NS * R);
    cplx<T> v[IT][R];
    const int tid = threadIdx.x;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int w = tid + it * NT;
      if (TOT % NT == 0 || w < TOT) {
        const int p = w / J;
        const int j = w - p * J;
#pragma unroll
        for (int q = 0; q < R; ++q) v[it][q] = buf[p * SEQ + lp_step<SH, J>(j, q)];
        if constexpr (NS > 1) {
          // twiddles w^q, q = 1..R-1, of w = tw[k STRIDE]: one table load and
          // powers by products (a few ulp from the table's entries)
          cplx<T> w1;
          if constexpr (StageTw<T, N, P, NT, MAXR>::template fixed<S>()) w1 = pre.w[S];
          else w1 = tw[(j % NS) * STRIDE];
          if constexpr (SIGN > 0) w1.y = -w1.y;
          cplx<T> t = w1;
          v[it][1] = cmul(v[it][1], t);
#pragma unroll
          for (int q = 2; q < R; ++q) {
            t = cmul(t, w1);
            v[it][q] = cmul(v[it][q], t);
          }
        }
        dft<T, R, SIGN>(v[it]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int w = tid + it * NT;
      if (TOT % NT == 0 || w < TOT) {
        const int p = w / J;
        const int j = w - p * J;
        const int k = j % NS;
        const int idx = (j / NS) * NS * R + k;
        // the first stage (NS = 1) writes R = 2^SH consecutive items from a
        // multiple of R: one padding slot before them, none among them
        constexpr bool kRun = NS == 1 && R == (1 << SH);
#pragma unroll
        for (int q = 0; q < R; ++q)
          buf[p * SEQ + (kRun ? lp<SH>(idx) + q : lp_step<SH, NS>(idx, q))] = v[it][q];
      }
    }
    __syncthreads();
    fft_stage<T, N, P, NT, MAXR, SIGN, S + 1, SEQ, SH>(buf, tw, pre);
  }
}

// ---------------------------------------------------------------------------
// fused element operations
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sqd(double x) { return x * x; }

// Wavenumbers along axis a: from the spacing when set (every fused pass of
// the stepper), else from the tables
__device__ __forceinline__ double vq_at(const Fuse& f, int a, int n) {
  return f.qs[a] != 0.0 ? (double)n * f.qs[a] : f.vq[a][n];
}
__device__ __forceinline__ double vs_at(const Fuse& f, int a, int n) {
  return f.qs[a] != 0.0 ? (double)(n + 1) * f.qs[a] : f.vs[a][n];
}

// |q|^2 at mode (i, j, k)
__device__ __forceinline__ double mq2c(const Fuse& f, int i, int j, int k) {
  return sqd(vq_at(f, 0, i)) + sqd(vq_at(f, 1, j)) + sqd(vq_at(f, 2, k));
}

// psi update at one mode (S:2177-2190) with FP64 C1, C2: the new psiq from
// psiq, psiq_old and the transformed nonlinear terms of this and the last step.
template <typename T>
__device__ __forceinline__ double psi_update_value(const Fuse& f, long long idx, int i, int j, int k) {
  const double m2 = mq2c(f, i, j, k);
  const double alin = f.eps - f.alpha * sqd(f.q02 - m2);
  const double c1 = 2.0 - f.dt * f.rho_m * alin;
  const double c2 = 1.5 - f.dt * f.rho_m * alin;
  const double n1 = (double)static_cast<const T*>(f.src[0])[idx];
  const double nold = (double)static_cast<const T*>(f.nlq_old)[idx];  // read even on the first step
  const double n0 = f.first ? n1 : nold;
  return (c1 * f.psiq[idx] - 0.5 * f.psiq_new[idx] + (f.dt / 2) * f.scale * (3.0 * n1 - n0)) / c2;
}

// z-line index g = i*ny + j -> (i, j) with 32-bit arithmetic
__device__ __forceinline__ void line_ij(long long g, int ny, int& i, int& j) {
  const unsigned gu = (unsigned)g;
  const unsigned iu = gu / (unsigned)ny;
  i = (int)iu;
  j = (int)(gu - iu * (unsigned)ny);
}

// Element m of z-line g (= i*ny + j) of the first pass under load op OP;
// idx = g*nz + m.  The value is formed in FP64 and goes to the FP64 tile as
// it is: a pass rounds to the storage type T only when it stores.
template <typename T, int OP, typename TO>
__device__ __forceinline__ double load_elem(const PassArgs<T, TO>& a, int i, int j, int m, long long idx) {
  if constexpr (OP == OP_PLAIN) {
    double v = (double)a.in[idx];
    if (a.in2) v += (double)a.in2[idx];
    return v;
  } else if constexpr (OP == OP_LD_PROD) {
    // rho (FP64 state) times L psi
    return static_cast<const double*>(a.f.src[0])[idx] * (double)static_cast<const T*>(a.f.src[1])[idx];
  } else if constexpr (OP == OP_LD_AMP) {
    const double ps = (double)static_cast<const T*>(a.f.src[0])[idx];
    const double x = (double)static_cast<const T*>(a.f.src[1])[idx];
    const double y = (double)static_cast<const T*>(a.f.src[2])[idx];
    const double z = (double)static_cast<const T*>(a.f.src[3])[idx];
    return sqrt(a.f.q02 * (ps * ps) + x * x + y * y + z * z);
  } else if constexpr (OP == OP_LD_GRAD_D || OP == OP_LD_GRAD_T) {
    // branch-free: the load always happens (at a clamped address) so that the
    // compiler can keep every load of the tile in flight
    const int ax = a.f.axis;
    const int na = ax == 0 ? i : (ax == 1 ? j : m);
    const int nmax = ax == 0 ? a.nx : (ax == 1 ? a.ny : a.nz);
    const long long off = ax == 0 ? (long long)a.ny * a.nz : (ax == 1 ? (long long)a.nz : 1LL);
    const bool ok = na < nmax - 1;
    const double w = vs_at(a.f, ax, ok ? na : 0);
    const long long src_idx = ok ? idx + off : idx;
    double src;
    if constexpr (OP == OP_LD_GRAD_D) src = static_cast<const double*>(a.f.src[0])[src_idx];
    else src = (double)static_cast<const T*>(a.f.src[0])[src_idx];
    return ok ? (-a.f.c0 * w) * src : 0.0;
  } else if constexpr (OP == OP_LD_SH || OP == OP_LD_RELAXLAP) {
    const double m2 = mq2c(a.f, i, j, m);
    const double ps = static_cast<const double*>(a.f.src[0])[idx];
    if constexpr (OP == OP_LD_SH) return a.f.alpha * (a.f.q02 - m2) * ps;
    else return -m2 * a.f.alpha * (a.f.q02 - m2) * ps;
  } else if constexpr (OP == OP_LD_PSI) {
    // the new psiq is the transform's input and is also stored (over psiq_old)
    const double r = psi_update_value<T>(a.f, idx, i, j, m);
    a.f.psiq_new[idx] = r;
    return r;
  } else {
    return 0.0;
  }
}

template <typename T, typename TO>
__device__ double load_elem_rt(const PassArgs<T, TO>& a, long long g, int m, long long mo) {
  int i, j;
  line_ij(g, a.ny, i, j);
  const long long idx = mo + g * a.nz + m;
  switch (a.f.op) {
    case OP_LD_GRAD_D: return load_elem<T, OP_LD_GRAD_D>(a, i, j, m, idx);
    case OP_LD_GRAD_T: return load_elem<T, OP_LD_GRAD_T>(a, i, j, m, idx);
    case OP_LD_AMP: return load_elem<T, OP_LD_AMP>(a, i, j, m, idx);
    case OP_LD_SH: return load_elem<T, OP_LD_SH>(a, i, j, m, idx);
    case OP_LD_RELAXLAP: return load_elem<T, OP_LD_RELAXLAP>(a, i, j, m, idx);
    case OP_LD_PROD: return load_elem<T, OP_LD_PROD>(a, i, j, m, idx);
    case OP_LD_PSI: return load_elem<T, OP_LD_PSI>(a, i, j, m, idx);
    case OP_AX_GAUSS_AMP: return load_elem<T, OP_LD_AMP>(a, i, j, m, idx);
    default: return load_elem<T, OP_PLAIN>(a, i, j, m, idx);
  }
}

// Diagonal multiplier of the mid ops at mode (i, j, k); y is the forward
// output, evaluated in the reference's order: (scale y) G, (-scale |q|^2) y.
template <int OP>
__device__ __forceinline__ double mid_apply(const Fuse& f, int i, int j, int k, double y) {
  const double m2 = mq2c(f, i, j, k);
  if constexpr (OP == OP_MID_AMP) return f.scale * y * exp(-1.57 * 1.57 * m2 * 2);
  else return (-f.scale * m2) * y;
}

// Adjacent columns l0, l0+1 of a strided tile at row m (zero outside the
// grid), widened to FP64.
template <typename T, typename TO>
__device__ __forceinline__ cplx<double> load_pair(const PassArgs<T, TO>& a, const TileGeo& g, int l0, int m) {
  const long long addr = g.base + l0 + (long long)m * g.stride_m;
  cplx<double> v = {0.0, 0.0};
  if (a.vec && l0 + 1 < g.nvalid) {
    const cplx<T> u = *reinterpret_cast<const cplx<T>*>(a.in + addr);
    v = cplx<double>{(double)u.x, (double)u.y};
    if (a.in2) {
      const cplx<T> w = *reinterpret_cast<const cplx<T>*>(a.in2 + addr);
      v.x += (double)w.x;
      v.y += (double)w.y;
    }
  } else {
    if (l0 < g.nvalid) { v.x = (double)a.in[addr]; if (a.in2) v.x += (double)a.in2[addr]; }
    if (l0 + 1 < g.nvalid) { v.y = (double)a.in[addr + 1]; if (a.in2) v.y += (double)a.in2[addr + 1]; }
  }
  return v;
}

// Store (y0, y1) at columns l0, l0+1 of row m, rounded to the storage type
// (added to what is there when acc).
template <typename T, typename TO>
__device__ __forceinline__ void store_pair(const PassArgs<T, TO>& a, const TileGeo& g, int l0, int m, double y0,
                                           double y1, bool acc = false) {
  TO* o = a.out + g.base + l0 + (long long)m * g.stride_m;
  if (a.vec && l0 + 1 < g.nvalid) {
    if (acc) {
      const cplx<TO> old = *reinterpret_cast<const cplx<TO>*>(o);
      y0 += (double)old.x;
      y1 += (double)old.y;
    }
    *reinterpret_cast<cplx<TO>*>(o) = cplx<TO>{(TO)y0, (TO)y1};
  } else {
    if (l0 < g.nvalid) o[0] = (TO)(acc ? (double)o[0] + y0 : y0);
    if (l0 + 1 < g.nvalid) o[1] = (TO)(acc ? (double)o[1] + y1 : y1);
  }
}

// Mode m of the one-axis operators' output from the forward outputs y of one
// line (a stride-1 view of the sequence): see FuseOp.  For both lines of a
// pair, y is the complex packing (line 0, line 1).
template <typename T, int OP>
__device__ __forceinline__ double ax_weight(const Fuse& f, int m, int n) {
  const double c = f.c0 / (2.0 * n);
  // DCT-lattice wavenumber of index i along the pass axis, evaluated from the
  // spacing rather than loaded from the table: the load would stall the
  // operator step behind an L2 round trip
  const double qs = f.qs[f.axis];
  auto vq = [&](int i) { return qs != 0.0 ? (double)i * qs : f.vq[f.axis][i]; };
  if constexpr (OP == OP_AX_GRAD) return -c * vq(m + 1);          // Vs[m] = Vq[m+1]
  else if constexpr (OP == OP_AX_DIV) return c * vq(m);
  else if constexpr (OP == OP_AX_GRAD_T) return -c * vq(m);       // Vs[m-1]
  else if constexpr (OP == OP_AX_DIV_T) return c * vq(m + 1);
  else {
    const double q = vq(m);
    return c * exp(-1.57 * 1.57 * q * q * 2);
  }
}
// source mode of output mode m, or -1 for a zero output
template <int OP>
__device__ __forceinline__ int ax_source(int m, int n) {
  if constexpr (OP == OP_AX_GRAD || OP == OP_AX_DIV_T) return m <= n - 2 ? m + 1 : -1;
  else if constexpr (OP == OP_AX_DIV || OP == OP_AX_GRAD_T) return m >= 1 ? m - 1 : -1;
  else return m;
}

// mk[N-k] from mk[k] = exp(-i pi k/2N): exp(-i pi/2) conj(mk[k]) = (-mk.y,
// -mk.x), exact; item 0's partner N/2 takes the table entry (a uniform
// address).
template <typename T, int N>
__device__ __forceinline__ cplx<T> mk_mirror(const cplx<T>* __restrict__ mk, int k, cplx<T> mkk) {
  return k == 0 ? mk[N / 2] : cplx<T>{-mkk.y, -mkk.x};
}

// Makhoul post-processing of item k of one sequence: Z at k and kb (kb = N-k,
// or N/2 for k = 0) -> the forward outputs of both lines at k and at kb.
// mkk = mk[k].
template <typename T, int N>
__device__ __forceinline__ void makhoul_post(const cplx<T>* __restrict__ mk, int k, int kb, cplx<T> mkk,
                                             cplx<T> za, cplx<T> zb, T& y0a, T& y1a, T& y0b, T& y1b) {
  // partner of k is N-k: Z_k pairs with conj Z_{N-k}; item 0 pairs 0 with 0 and N/2 with N/2
  const cplx<T> pa = (k == 0) ? za : zb;
  const cplx<T> qa = (k == 0) ? zb : za;
  {
    const cplx<T> m = mkk;
    const cplx<T> S = {za.x + pa.x, za.y - pa.y};
    const cplx<T> D = {za.x - pa.x, za.y + pa.y};
    y0a = m.x * S.x - m.y * S.y;
    y1a = m.x * D.y + m.y * D.x;
  }
  {
    const cplx<T> m = mk_mirror<T, N>(mk, k, mkk);
    const cplx<T> S = {zb.x + qa.x, zb.y - qa.y};
    const cplx<T> D = {zb.x - qa.x, zb.y + qa.y};
    y0b = m.x * S.x - m.y * S.y;
    y1b = m.x * D.y + m.y * D.x;
  }
}

// Makhoul pre-processing of the inverse: (Y0, Y1) at k (rk) and at kb (rn) ->
// z_k, z_kb of U0 + i U1 with U_k = e^{+i pi k/2N} (Y_k - i Y_{N-k}); mkk =
// mk[k].
template <typename T, int N>
__device__ __forceinline__ void makhoul_pre(const cplx<T>* __restrict__ mk, int k, cplx<T> mkk, cplx<T> rk,
                                            cplx<T> rn, cplx<T>& zk, cplx<T>& zb) {
  constexpr int HN = N / 2;
  if (k == 0) {
    zk = rk;
    cplx<T> cc = mk[HN];
    cc.y = -cc.y;
    zb = cmul(cc, cplx<T>{rn.x + rn.y, rn.y - rn.x});
  } else {
    cplx<T> cc = mkk;
    cc.y = -cc.y;
    zk = cmul(cc, cplx<T>{rk.x + rn.y, rk.y - rn.x});
    cc = mk_mirror<T, N>(mk, k, mkk);
    cc.y = -cc.y;
    zb = cmul(cc, cplx<T>{rn.x + rk.y, rn.y - rk.x});
  }
}

// mk[k] for the items k a thread visits in the loops over it (w = tid + it
// NT), from one table load and products with a uniform step (a few ulp):
// - along y and x, k = w / P advances by NT/P per iteration (at most 7
//   products);
// - along z, k = w mod N/2: when NT is a multiple of N/2, k = tid mod N/2 in
//   every iteration; when N/2 is a multiple of NT, k = tid + (it mod
//   (N/2)/NT) NT, restarting at tid with each new sequence; other lengths
//   load mk[k] per item.
template <typename A, int N, int P, int NT>
struct MkSeq {
  static constexpr int HN = N / 2;
  static constexpr bool ZREP = NT % HN == 0;
  static constexpr bool ZCYC = !ZREP && HN % NT == 0;
  static constexpr int ZQ = ZCYC ? HN / NT : 1;
  // a thread whose first item is past the table has no valid item (k0 >=
  // N/2), and when NT/P >= N no thread has a valid item after the first
  // iteration: both indexes are clamped into the N-entry table
  static constexpr int SI = NT / P < N ? NT / P : 0;
  cplx<A> base, cur, step;
  __device__ __forceinline__ MkSeq(const cplx<A>* __restrict__ mk, int axis, int tid) {
    base = cur = step = cplx<A>{A(0), A(0)};
    if (axis != 2) {
      const int k0 = tid / P;
      cur = mk[k0 < N ? k0 : 0];
      step = mk[SI];
    } else if constexpr (ZREP) {
      cur = mk[tid % HN];
    } else if constexpr (ZCYC) {
      base = cur = mk[tid];  // tid < NT <= N/2
      step = mk[NT];
    }
  }
  // mk[k] of this iteration's item
  __device__ __forceinline__ cplx<A> get(const cplx<A>* __restrict__ mk, int axis, int k) const {
    if constexpr (!ZREP && !ZCYC) {
      if (axis == 2) return mk[k];
    }
    return cur;
  }
  // move to iteration it + 1
  __device__ __forceinline__ void advance(int axis, int it) {
    if (axis != 2) {
      cur = cmul(cur, step);
    } else if constexpr (ZCYC) {
      cur = ((it + 1) % ZQ == 0) ? base : cmul(cur, step);
    }
  }
};

// A pair loaded in FP64 as the tile's arithmetic type.
template <typename A>
__device__ __forceinline__ cplx<A> to_arith(cplx<double> v) {
  return cplx<A>{(A)v.x, (A)v.y};
}

// Minimum waves per SIMD a pass is compiled for (its register cap), for
// tiles whose threads hold at most 8 values (2 P N / NT): the 256-thread y/x
// tiles are held to 4 by their LDS anyway, so 128 VGPRs cost nothing there;
// one-wave z tiles fit 6 waves (5 for the one-axis operators, which would
// spill at 6).  Larger per-thread tiles and other block sizes keep the
// compiler's choice.
constexpr int pass_min_waves(int nt, int op, int per_thread) {
  return per_thread > 8 ? 1 : nt >= 256 ? 4 : nt == 64 ? (op >= OP_AX_GRAD ? 5 : 6) : 1;
}

// One axis pass over a tile of TL = 2P lines; lines 2c and 2c+1 of the tile
// form complex sequence c.  OP selects a fused operation (FuseOp); the
// output is stored as TO.
template <typename T, int N, int P, int NT, int OP, typename TO = T, typename A = double, int AXC = -1>
__global__ __launch_bounds__(NT, pass_min_waves(NT, OP, 2 * P * N / NT)) void fft_pass_kernel(PassArgs<T, TO> a) {
  // the pass axis: known at compile time for z passes (AXC = 2); y or x
  // (AXC = 3) is taken as a.axis & 1, so every test for the z axis folds
  const int axis = AXC == 2 ? 2 : AXC == 3 ? (a.axis & 1) : a.axis;
  constexpr int TL = 2 * P;
  constexpr int MAXR = pick_maxr(N, P, NT);
  constexpr int SEQ = seq_len(N, MAXR);
  constexpr int SH = pad_shift(N, MAXR);
  constexpr int HN = N / 2;
  constexpr bool MID = (OP == OP_MID_AMP || OP == OP_MID_NEGLAP);
  constexpr bool AX = (OP == OP_AX_GRAD || OP == OP_AX_DIV || OP == OP_AX_GAUSS || OP == OP_AX_GAUSS_AMP ||
                       OP == OP_AX_GRAD_T || OP == OP_AX_DIV_T);
  // load op of the first (z) pass
  constexpr int LD_OP = OP == OP_AX_GAUSS_AMP ? OP_LD_AMP : ((MID || AX) ? OP_PLAIN : OP);
  constexpr int AXOP = OP == OP_AX_GAUSS_AMP ? OP_AX_GAUSS : OP;
  // the tile and every operation on it are of type A: FP64 (both storage
  // types), or FP32 for FP32 storage where FP32 arithmetic is asked for; the
  // tables are of type A too
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<A>* buf = reinterpret_cast<cplx<A>*>(smem_raw);
  const int tid = threadIdx.x;
  const cplx<A>* const tw = reinterpret_cast<const cplx<A>*>(a.tw);
  const cplx<A>* const mk = reinterpret_cast<const cplx<A>*>(a.mk);
  TileGeo g = tile_geometry<TL, N>(axis, a.nx, a.ny, a.nz);
  const long long mo = (long long)blockIdx.y * ((long long)a.nx * a.ny * a.nz);  // member offset
  g.base += mo;
  const bool fwd = MID || AX || (a.kind == REDFT10 || a.kind == RODFT10);
  // sine: DST on the load side (forward DST-II, or inverse DST-III input);
  // sine_out: DST on the store side
  const bool sine = AX ? (OP == OP_AX_DIV || OP == OP_AX_GRAD_T) : (!MID && (a.kind == RODFT10 || a.kind == RODFT01));
  const bool sine_out = AX ? (OP == OP_AX_GRAD || OP == OP_AX_DIV_T) : sine;
  StageTw<A, N, P, NT, MAXR> stw;
  stw.load(tw, tid);
  // the Makhoul factors' first loads, also ahead of the tile's data
  const MkSeq<A, N, P, NT> ms0(mk, axis, tid);

  if (fwd) {
    // ---- load: every global load of the tile is issued before the first LDS
    // write, so each thread keeps all of its loads in flight ---------------
    if (axis == 2) {
      // item e = (line pair c, element m): lanes walk m along the two lines
      // of a pair, and each item stores one complex value to LDS
      constexpr int E = P * N;
      constexpr int IT = (E + NT - 1) / NT;
      cplx<A> v[IT];
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int e = tid + it * NT;
        const int c = e / N, m = e - (e / N) * N;
        v[it] = cplx<A>{A(0), A(0)};
        if (E % NT == 0 || e < E) {
          const int l0 = 2 * c;
          const long long gl = (long long)g.outer + l0;
          int i, j;
          line_ij(gl, a.ny, i, j);
          if (l0 < g.nvalid) v[it].x = load_elem<T, LD_OP>(a, i, j, m, mo + gl * N + m);
          if (l0 + 1 < g.nvalid) {
            int i1 = i, j1 = j + 1;
            if (j1 == a.ny) { j1 = 0; ++i1; }
            v[it].y = load_elem<T, LD_OP>(a, i1, j1, m, mo + (gl + 1) * N + m);
          }
        }
      }
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int e = tid + it * NT;
        if (E % NT == 0 || e < E) {
          const int c = e / N, m = e - (e / N) * N;
          cplx<A> x = v[it];
          if (sine && (m & 1)) { x.x = -x.x; x.y = -x.y; }
          const int pos = (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1);
          buf[c * SEQ + lp<SH>(pos)] = x;
        }
      }
    } else {
      constexpr int E = P * N;  // pairs of adjacent columns
      constexpr int IT = (E + NT - 1) / NT;
      cplx<A> v[IT];
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int e = tid + it * NT;
        v[it] = (E % NT == 0 || e < E) ? to_arith<A>(load_pair(a, g, 2 * (e % P), e / P)) : cplx<A>{A(0), A(0)};
      }
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int e = tid + it * NT;
        if (E % NT == 0 || e < E) {
          const int c = e % P, m = e / P;
          cplx<A> x = v[it];
          if (sine && (m & 1)) { x.x = -x.x; x.y = -x.y; }
          const int pos = (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1);
          buf[c * SEQ + lp<SH>(pos)] = x;
        }
      }
    }
  } else {
    // ---- load straight into Makhoul's pre-twiddled sequence: item k < N/2
    // loads Y_k and its mirror Y_{N-k} (item 0 takes Y_0 and Y_{N/2}) for both
    // lines of a pair.  DST-III reads the input reversed, Y_k = x_{N-1-k}.
    constexpr int W = P * HN;
    constexpr int ITW = (W + NT - 1) / NT;
    cplx<A> rk[ITW], rn[ITW];
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      const int w = tid + it * NT;
      rk[it] = rn[it] = cplx<A>{A(0), A(0)};
      if (W % NT == 0 || w < W) {
        int c, k;
        if (axis == 2) { c = w / HN; k = w - c * HN; }
        else { k = w / P; c = w - k * P; }
        const int kb = (k == 0) ? HN : N - k;
        const int pa = sine ? N - 1 - k : k;
        const int pb = sine ? N - 1 - kb : kb;
        if (axis == 2) {
          const int l0 = 2 * c;
          const long long g0 = (long long)g.outer + l0;
          int i, j;
          line_ij(g0, a.ny, i, j);
          if (l0 < g.nvalid) {
            rk[it].x = load_elem<T, OP>(a, i, j, pa, mo + g0 * N + pa);
            rn[it].x = load_elem<T, OP>(a, i, j, pb, mo + g0 * N + pb);
          }
          if (l0 + 1 < g.nvalid) {
            int i1 = i, j1 = j + 1;
            if (j1 == a.ny) { j1 = 0; ++i1; }
            rk[it].y = load_elem<T, OP>(a, i1, j1, pa, mo + (g0 + 1) * N + pa);
            rn[it].y = load_elem<T, OP>(a, i1, j1, pb, mo + (g0 + 1) * N + pb);
          }
        } else {
          rk[it] = to_arith<A>(load_pair(a, g, 2 * c, pa));
          rn[it] = to_arith<A>(load_pair(a, g, 2 * c, pb));
        }
      }
    }
    MkSeq<A, N, P, NT> ms = ms0;
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      const int w = tid + it * NT;
      if (W % NT == 0 || w < W) {
        int c, k;
        if (axis == 2) { c = w / HN; k = w - c * HN; }
        else { k = w / P; c = w - k * P; }
        const int kb = (k == 0) ? HN : N - k;
        cplx<A> zk, zb;
        makhoul_pre<A, N>(mk, k, ms.get(mk, axis, k), rk[it], rn[it], zk, zb);
        buf[c * SEQ + lp<SH>(k)] = zk;
        buf[c * SEQ + lp<SH>(kb)] = zb;
      }
      ms.advance(axis, it);
    }
  }
  __syncthreads();

  if (fwd) {
    fft_stage<A, N, P, NT, MAXR, -1, 0>(buf, tw, stw);
    constexpr int W = P * HN;
    constexpr int ITW = (W + NT - 1) / NT;
    if constexpr (MID) {
      // forward outputs -> diagonal operator -> inverse pre-twiddle, in place:
      // item k owns positions k and kb of its sequence, so no barrier is needed
      MkSeq<A, N, P, NT> ms = ms0;
#pragma unroll
      for (int it = 0; it < ITW; ++it) {
        const int w = tid + it * NT;
        if (W % NT == 0 || w < W) {
          int c, k;
          if (axis == 2) { c = w / HN; k = w - c * HN; }
          else { k = w / P; c = w - k * P; }
          const int kb = (k == 0) ? HN : N - k;
          const cplx<A> mkk = ms.get(mk, axis, k);
          A y0a, y1a, y0b, y1b;
          makhoul_post<A, N>(mk, k, kb, mkk, buf[c * SEQ + lp<SH>(k)], buf[c * SEQ + lp<SH>(kb)], y0a, y1a, y0b, y1b);
          // x pass (axis 0): mode index along x, column j = outer, k = k0 + line
          const int jj = g.outer;
          const int kc0 = g.k0 + 2 * c < a.nz ? g.k0 + 2 * c : a.nz - 1;
          const int kc1 = g.k0 + 2 * c + 1 < a.nz ? g.k0 + 2 * c + 1 : a.nz - 1;
          const cplx<A> rk = {(A)mid_apply<OP>(a.f, k, jj, kc0, y0a), (A)mid_apply<OP>(a.f, k, jj, kc1, y1a)};
          const cplx<A> rn = {(A)mid_apply<OP>(a.f, kb, jj, kc0, y0b), (A)mid_apply<OP>(a.f, kb, jj, kc1, y1b)};
          cplx<A> zk, zb;
          makhoul_pre<A, N>(mk, k, mkk, rk, rn, zk, zb);
          buf[c * SEQ + lp<SH>(k)] = zk;
          buf[c * SEQ + lp<SH>(kb)] = zb;
        }
        ms.advance(axis, it);
      }
      __syncthreads();
    } else if constexpr (AX) {
      // forward outputs -> operator -> inverse pre-twiddle, in registers and
      // in place: item k holds the forward outputs at fa, fb (k and kb, or
      // N-1-k and N-1-kb for a forward DST), and the modes its pre-twiddle
      // reads, the operator's sources of output indices pa, pb (k and kb, or
      // N-1-k and N-1-kb for an inverse DST), are always among them or zero:
      // the one-mode shifts compose with the DST index reversal.  So no
      // exchange between items is needed, and item k rewrites only the
      // positions k, kb it read
      MkSeq<A, N, P, NT> ms = ms0;
#pragma unroll
      for (int it = 0; it < ITW; ++it) {
        const int w = tid + it * NT;
        if (W % NT == 0 || w < W) {
          int c, k;
          if (axis == 2) { c = w / HN; k = w - c * HN; }
          else { k = w / P; c = w - k * P; }
          const int kb = (k == 0) ? HN : N - k;
          const cplx<A> mkk = ms.get(mk, axis, k);
          A y0a, y1a, y0b, y1b;
          makhoul_post<A, N>(mk, k, kb, mkk, buf[c * SEQ + lp<SH>(k)], buf[c * SEQ + lp<SH>(kb)], y0a, y1a, y0b, y1b);
          const int fa = sine ? N - 1 - k : k;
          const int pa = sine_out ? N - 1 - k : k;
          const int pb = sine_out ? N - 1 - kb : kb;
          const int sa = ax_source<AXOP>(pa, N), sb = ax_source<AXOP>(pb, N);
          cplx<A> xk = {A(0), A(0)}, xb = {A(0), A(0)};
          if (sa >= 0) {
            const double wgt = ax_weight<T, AXOP>(a.f, pa, N);
            const cplx<A> y = sa == fa ? cplx<A>{y0a, y1a} : cplx<A>{y0b, y1b};
            xk = cplx<A>{(A)(wgt * y.x), (A)(wgt * y.y)};
          }
          if (sb >= 0) {
            const double wgt = ax_weight<T, AXOP>(a.f, pb, N);
            const cplx<A> y = sb == fa ? cplx<A>{y0a, y1a} : cplx<A>{y0b, y1b};
            xb = cplx<A>{(A)(wgt * y.x), (A)(wgt * y.y)};
          }
          cplx<A> zk, zb;
          makhoul_pre<A, N>(mk, k, mkk, xk, xb, zk, zb);
          buf[c * SEQ + lp<SH>(k)] = zk;
          buf[c * SEQ + lp<SH>(kb)] = zb;
        }
        ms.advance(axis, it);
      }
      __syncthreads();
    } else {
      // ---- post: item k < N/2 writes outputs k and N-k (item 0: 0 and N/2)
      MkSeq<A, N, P, NT> ms = ms0;
#pragma unroll
      for (int it = 0; it < ITW; ++it) {
        const int w = tid + it * NT;
        if (W % NT == 0 || w < W) {
          int c, k;
          if (axis == 2) { c = w / HN; k = w - c * HN; }
          else { k = w / P; c = w - k * P; }
          const int kb = (k == 0) ? HN : N - k;
          A y0a, y1a, y0b, y1b;
          makhoul_post<A, N>(mk, k, kb, ms.get(mk, axis, k), buf[c * SEQ + lp<SH>(k)], buf[c * SEQ + lp<SH>(kb)],
                             y0a, y1a, y0b, y1b);
          const int oa = sine ? (N - 1 - k) : k;
          const int ob = sine ? (N - 1 - kb) : kb;
          const int l0 = 2 * c;
          if (axis == 2) {
            if (l0 < g.nvalid) {
              TO* o = a.out + g.base + (long long)l0 * N;
              o[oa] = (TO)y0a;
              o[ob] = (TO)y0b;
            }
            if (l0 + 1 < g.nvalid) {
              TO* o = a.out + g.base + (long long)(l0 + 1) * N;
              o[oa] = (TO)y1a;
              o[ob] = (TO)y1b;
            }
          } else {
            store_pair(a, g, l0, oa, y0a, y1a);
            store_pair(a, g, l0, ob, y0b, y1b);
          }
        }
        ms.advance(axis, it);
      }
      return;
    }
  }

  // ---- inverse FFT and store: y_{2n} = u_n, y_{2n+1} = u_{N-1-n}
  fft_stage<A, N, P, NT, MAXR, +1, 0>(buf, tw, stw);
  constexpr int WO = P * HN;
  constexpr int ITO = (WO + NT - 1) / NT;
  const A s = sine_out ? A(-1) : A(1);
  const bool acc = AX && a.f.accumulate;
#pragma unroll
  for (int it = 0; it < ITO; ++it) {
    const int w = tid + it * NT;
    if (WO % NT == 0 || w < WO) {
      int c, n;
      if (axis == 2) { c = w / HN; n = w - c * HN; }
      else { n = w / P; c = w - n * P; }
      const cplx<A> ua = buf[c * SEQ + lp<SH>(n)];
      const cplx<A> ub = buf[c * SEQ + lp<SH>(N - 1 - n)];
      const int l0 = 2 * c;
      if (axis == 2) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if (l0 + q < g.nvalid) {
            TO* o = a.out + g.base + (long long)(l0 + q) * N + 2 * n;
            A v0 = q == 0 ? ua.x : ua.y;
            A v1 = s * (q == 0 ? ub.x : ub.y);
            if (a.vec) {
              if (acc) {
                const cplx<TO> old = *reinterpret_cast<const cplx<TO>*>(o);
                v0 += (A)old.x;
                v1 += (A)old.y;
              }
              *reinterpret_cast<cplx<TO>*>(o) = cplx<TO>{(TO)v0, (TO)v1};
            } else {
              if (acc) { v0 += (A)o[0]; v1 += (A)o[1]; }
              o[0] = (TO)v0;
              o[1] = (TO)v1;
            }
          }
        }
      } else {
        store_pair(a, g, l0, 2 * n, ua.x, ua.y, acc);
        store_pair(a, g, l0, 2 * n + 1, s * ub.x, s * ub.y, acc);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// direct O(N^2) fallback for any length, with the same fused operations
// ---------------------------------------------------------------------------
constexpr int kThreads = 256;

template <typename T, typename TO = T>
struct DirectArgs {
  PassArgs<T, TO> p;
  const double* costab;  // 4n entries: cos(pi t/(2n))
  int n;
  int tl;
};

// FFTW's defining sum of `kind` at output k for one line xl of length n; the
// angle index t (units of pi/(2n)) is advanced incrementally mod 4n.
__device__ inline double direct_sum(int kind, const double* xl, int n, int k, const double* __restrict__ ct) {
  using T = double;
  const int n4 = 4 * n;
  T acc = T(0);
  if (kind == REDFT10) {            // 2 sum x_m cos(pi (2m+1) k / 2n)
    int t = k % n4, step = (2 * k) % n4;
    for (int m = 0; m < n; ++m) { acc += xl[m] * ct[t]; t += step; if (t >= n4) t -= n4; }
    return T(2) * acc;
  } else if (kind == REDFT01) {     // x_0 + 2 sum_{m>=1} x_m cos(pi m (2k+1)/2n)
    int step = (2 * k + 1) % n4, t = step;
    for (int m = 1; m < n; ++m) { acc += xl[m] * ct[t]; t += step; if (t >= n4) t -= n4; }
    return xl[0] + T(2) * acc;
  } else if (kind == RODFT10) {     // 2 sum x_m sin(pi (2m+1)(k+1)/2n)
    int step = (2 * (k + 1)) % n4;
    int t = ((k + 1) - n) % n4; if (t < 0) t += n4;   // sin(u) = cos(u - pi/2)
    for (int m = 0; m < n; ++m) { acc += xl[m] * ct[t]; t += step; if (t >= n4) t -= n4; }
    return T(2) * acc;
  } else {                          // (-1)^k x_{n-1} + 2 sum_{m<n-1} x_m sin(pi (m+1)(2k+1)/2n)
    int step = (2 * k + 1) % n4;
    int t = (step - n) % n4; if (t < 0) t += n4;
    for (int m = 0; m < n - 1; ++m) { acc += xl[m] * ct[t]; t += step; if (t >= n4) t -= n4; }
    return ((k & 1) ? -xl[n - 1] : xl[n - 1]) + T(2) * acc;
  }
}

template <typename T, typename TO = T>
__global__ __launch_bounds__(kThreads) void direct_pass_kernel(DirectArgs<T, TO> d) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const PassArgs<T, TO>& a = d.p;
  const int n = d.n, tl = d.tl;
  // FP64 tile for both storage types
  double* x = reinterpret_cast<double*>(smem_raw);
  double* y = x + tl * n;  // second buffer of the mid ops
  const bool mid = a.f.op == OP_MID_AMP || a.f.op == OP_MID_NEGLAP;
  const int op = a.f.op;
  const bool ax = op == OP_AX_GRAD || op == OP_AX_DIV || op == OP_AX_GAUSS || op == OP_AX_GAUSS_AMP ||
                  op == OP_AX_GRAD_T || op == OP_AX_DIV_T;
  const bool ax_fsine = op == OP_AX_DIV || op == OP_AX_GRAD_T;   // forward DST-II
  const bool ax_isine = op == OP_AX_GRAD || op == OP_AX_DIV_T;   // inverse DST-III
  TileGeo g;
  if (a.axis == 2) {
    long long first = (long long)blockIdx.x * tl;
    long long nl = (long long)a.nx * a.ny;
    g.nvalid = (int)((nl - first) < tl ? (nl - first) : tl);
    g.base = first * n;
    g.line_step = n;
    g.stride_m = 1;
    g.outer = (int)first;
    g.k0 = 0;
  } else {
    int tpo = (a.nz + tl - 1) / tl;
    int o = blockIdx.x / tpo;
    int k0 = (blockIdx.x - o * tpo) * tl;
    g.nvalid = (a.nz - k0) < tl ? (a.nz - k0) : tl;
    g.line_step = 1;
    g.outer = o;
    g.k0 = k0;
    if (a.axis == 1) { g.base = (long long)o * a.ny * a.nz + k0; g.stride_m = a.nz; }
    else { g.base = (long long)o * a.nz + k0; g.stride_m = (long long)a.ny * a.nz; }
  }
  const long long mo = (long long)blockIdx.y * ((long long)a.nx * a.ny * a.nz);  // member offset
  g.base += mo;
  const int E = tl * n;
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int l, m;
    if (a.axis == 2) { l = e / n; m = e - l * n; }
    else { m = e / tl; l = e - m * tl; }
    double v = 0.0;
    if (l < g.nvalid) {
      if (a.axis == 2) {
        v = load_elem_rt(a, (long long)g.outer + l, m, mo);
      } else {
        const long long addr = g.base + l * g.line_step + m * g.stride_m;
        v = (double)a.in[addr];
        if (a.in2) v += (double)a.in2[addr];
      }
    }
    x[l * n + m] = v;
  }
  __syncthreads();
  if (ax) {
    // forward 1-D transform into y, the operator back into x
    const int fk = ax_fsine ? RODFT10 : REDFT10;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
      const int k = e / tl, l = e - (e / tl) * tl;
      if (l < g.nvalid) y[l * n + k] = direct_sum(fk, x + l * n, n, k, d.costab);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
      const int m = e / tl, l = e - (e / tl) * tl;
      if (l >= g.nvalid) continue;
      double v = 0.0;
      if (op == OP_AX_GRAD) {
        if (ax_source<OP_AX_GRAD>(m, n) >= 0) v = (ax_weight<T, OP_AX_GRAD>(a.f, m, n) * y[l * n + m + 1]);
      } else if (op == OP_AX_DIV) {
        if (ax_source<OP_AX_DIV>(m, n) >= 0) v = (ax_weight<T, OP_AX_DIV>(a.f, m, n) * y[l * n + m - 1]);
      } else if (op == OP_AX_GRAD_T) {
        if (ax_source<OP_AX_GRAD_T>(m, n) >= 0) v = (ax_weight<T, OP_AX_GRAD_T>(a.f, m, n) * y[l * n + m - 1]);
      } else if (op == OP_AX_DIV_T) {
        if (ax_source<OP_AX_DIV_T>(m, n) >= 0) v = (ax_weight<T, OP_AX_DIV_T>(a.f, m, n) * y[l * n + m + 1]);
      } else {
        v = (ax_weight<T, OP_AX_GAUSS>(a.f, m, n) * y[l * n + m]);
      }
      x[l * n + m] = v;
    }
    __syncthreads();
  } else if (mid) {
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
      const int k = e / tl, l = e - (e / tl) * tl;
      if (l >= g.nvalid) continue;
      const double v = direct_sum(REDFT10, x + l * n, n, k, d.costab);
      const int kc = g.k0 + l;
      y[l * n + k] = (a.f.op == OP_MID_AMP) ? mid_apply<OP_MID_AMP>(a.f, k, g.outer, kc, v)
                                            : mid_apply<OP_MID_NEGLAP>(a.f, k, g.outer, kc, v);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int l, k;
    if (a.axis == 2) { l = e / n; k = e - l * n; }
    else { k = e / tl; l = e - k * tl; }
    if (l >= g.nvalid) continue;
    double r;
    if (ax) r = direct_sum(ax_isine ? RODFT01 : REDFT01, x + l * n, n, k, d.costab);
    else if (mid) r = direct_sum(REDFT01, y + l * n, n, k, d.costab);
    else r = direct_sum(a.kind, x + l * n, n, k, d.costab);
    const long long oi = g.base + l * g.line_step + k * g.stride_m;
    a.out[oi] = (TO)((ax && a.f.accumulate) ? (double)a.out[oi] + r : r);
  }
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

template <typename T, typename TO = T>
PassArgs<T, TO> make_args(const PassIO& io, void* tw, void* mk, int kind, int axis, const Geom& g) {
  PassArgs<T, TO> a;
  a.in = static_cast<const T*>(io.in);
  a.in2 = static_cast<const T*>(io.in2);
  a.out = static_cast<TO*>(io.out);
  a.tw = tw;
  a.mk = mk;
  a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
  a.axis = axis;
  a.kind = kind;
  a.vec = (g.nz % 2 == 0) && aligned16(io.in) && aligned16(io.out) &&
          (io.in2 == nullptr || aligned16(io.in2));
  a.f = io.f;
  return a;
}

template <typename T, int N, int P, int NT, int OP, typename TO = T, typename A = double, int AXC = -1>
hipError_t launch_fft_pn(const PassIO& io, void* tw, void* mk, int kind, int axis, const Geom& g,
                         hipStream_t s) {
  constexpr int TL = 2 * P;
  if ((AXC == 2 && axis != 2) || (AXC == 3 && axis != 0 && axis != 1)) return hipErrorInvalidValue;
  PassArgs<T, TO> a = make_args<T, TO>(io, tw, mk, kind, axis, g);
  long long blocks;
  if (axis == 2) blocks = ((long long)g.nx * g.ny + TL - 1) / TL;
  else blocks = (long long)(axis == 1 ? g.nx : g.ny) * ((g.nz + TL - 1) / TL);
  const size_t lds = (size_t)P * seq_len(N, pick_maxr(N, P, NT)) * sizeof(cplx<A>);
  hipLaunchKernelGGL((fft_pass_kernel<T, N, P, NT, OP, TO, A, AXC>), dim3((unsigned)blocks, (unsigned)g.nb), dim3(NT), lds, s, a);
  return hipGetLastError();
}


// Tile and block size per (length, axis).  Measured at 256^3 on MI355X: one
// wavefront per line pair along z (half the LDS of two pairs, so more
// wavefronts per CU hide the loads), i.e. 256 elements per wavefront; along y
// and x 8 pairs x 256 threads, whose rows are 16 columns: 128-byte segments
// for FP64 storage.  FP32 storage takes 16 pairs from length 64 on (128-byte
// rows of 32 columns) with P N / 8 threads, so that radix-8 stages give
// every thread a butterfly (16 elements per thread, as before).
template <typename T, int N>
struct TileChoice {
  static constexpr int cap(int p) { return p < 1 ? 1 : p; }
  static constexpr int budget(int b) {
    int p = b / N, q = 1;
    while (q * 2 <= p) q *= 2;
    return q;
  }
  static constexpr bool kWide = sizeof(T) == 4 && N >= 64;
  // z: about 256 elements per wavefront (one line pair at N = 256, more for
  // shorter lines so every thread keeps a butterfly)
  static constexpr int PZ = cap(budget(256));
  static constexpr int PS = kWide ? cap(budget(4096) < 16 ? budget(4096) : 16) : cap(budget(2048) < 8 ? budget(2048) : 8);
  static constexpr int NTZ = 64;
  static constexpr int NTS = kWide ? (PS * N / 8 < 512 ? PS * N / 8 : 512) : 256;
};

// A plain y or x pass on the 256-point kernels of r2r_yx256.h when they
// apply (length 256 along a strided axis, full tiles of 16 FP64 or 32 FP32
// columns, aligned pointers, element offsets within 32 bits, no fused op);
// hipErrorNotSupported otherwise, with nothing launched.  Storage in_dt ->
// out_dt: equal, or FP64 -> FP32.  Defined in r2r_yx256.hip.
hipError_t launch_yx256(DType in_dt, DType out_dt, int kind, int axis, const PassIO& io, const Geom& g,
                        const void* tw, const void* mk, hipStream_t s, bool arith32 = false);

// One-axis operator pass (OP_AX_*) of FFT length n with the default tiles;
// defined in r2r_axis_ops.hip.
template <typename T, typename A = double>
hipError_t launch_axis_op(int n, const PassIO& io, void* tw, void* mk, int axis, const Geom& g,
                          hipStream_t s);

// ---------------------------------------------------------------------------
// tile variants that R2R::autotune times for plain passes (variant 0 is the
// default TileChoice of r2r.hip): lines per tile 2P and threads NT
// ---------------------------------------------------------------------------
template <typename T, bool Z, int V>
struct TileVariant;
template <typename T> struct TileVariant<T, true, 1> { static constexpr int P = 2, NT = 64; };
template <typename T> struct TileVariant<T, true, 2> { static constexpr int P = 2, NT = 128; };
template <typename T> struct TileVariant<T, true, 3> { static constexpr int P = 4, NT = 128; };
template <typename T> struct TileVariant<T, false, 1> { static constexpr int P = 4, NT = 128; };
template <typename T> struct TileVariant<T, false, 2> { static constexpr int P = 8, NT = 512; };
template <typename T> struct TileVariant<T, false, 3> { static constexpr int P = 8, NT = 128; };

// LDS a workgroup may take on gfx950 (160 KiB per CU, all of it available to
// one workgroup).
constexpr long long kMaxLds = 160 * 1024;

// A variant is built for length N when its LDS tile fits kMaxLds and each
// thread holds at most 256 bytes of the tile in registers.
template <typename T, int N, int P, int NT>
constexpr bool tile_fits() {
  return (long long)P * seq_len(N, pick_maxr(N, P, NT)) * (long long)sizeof(cplx<double>) <= kMaxLds &&
         (2LL * P * N / NT) * 8LL <= 256 && NT % 64 == 0 && NT <= 1024;
}

// Plain pass of length n with tile variant v (1..kTileVariants-1);
// hipErrorNotSupported when the variant is not built for n.  Defined in
// r2r_tune_f64.hip / r2r_tune_f32.hip.
template <typename T>
hipError_t launch_plain_variant(int n, int v, const PassIO& io, void* tw, void* mk, int kind, int axis,
                                const Geom& g, hipStream_t s);

}  // namespace smaiso
