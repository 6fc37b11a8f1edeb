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
//  * Lengths outside the FFT list use a direct O(N^2) LDS kernel that
//    evaluates FFTW's defining sums with a cosine table.
#include "r2r.h"

#include <cmath>
#include <cstdio>

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

template <typename T>
struct PassArgs {
  const T* in;
  const T* in2;
  T* out;
  const cplx<T>* tw;
  const cplx<T>* mk;
  int nx, ny, nz;
  int axis;
  int kind;
  int vec;  // 16-byte pair accesses allowed (nz even, pointers 16-byte aligned)
};

struct TileGeo {
  long long base;
  long long line_step;
  long long stride_m;
  int nvalid;
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
  } else {
    int tpo = (nz + TL - 1) / TL;
    int o = blockIdx.x / tpo;
    int k0 = (blockIdx.x - o * tpo) * TL;
    t.nvalid = (nz - k0) < TL ? (nz - k0) : TL;
    t.line_step = 1;
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

template <typename T, int N, int P, int NT, int MAXR, int SIGN, int S>
__device__ __forceinline__ void fft_stage(cplx<T>* buf, const cplx<T>* __restrict__ tw) {
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
        for (int q = 0; q < R; ++q) v[it][q] = buf[p * N + j + q * J];
        if constexpr (NS > 1) {
          const int k = j % NS;
#pragma unroll
          for (int q = 1; q < R; ++q) {
            cplx<T> t = tw[q * k * STRIDE];
            if constexpr (SIGN > 0) t.y = -t.y;
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
#pragma unroll
        for (int q = 0; q < R; ++q) buf[p * N + idx + q * NS] = v[it][q];
      }
    }
    __syncthreads();
    fft_stage<T, N, P, NT, MAXR, SIGN, S + 1>(buf, tw);
  }
}

// One axis pass over a tile of TL = 2P lines; lines 2c and 2c+1 of the tile
// form complex sequence c.
template <typename T, int N, int P, int NT>
__global__ __launch_bounds__(NT) void fft_pass_kernel(PassArgs<T> a) {
  constexpr int TL = 2 * P;
  constexpr int MAXR = pick_maxr(N, P, NT);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  T* rbuf = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x;
  const TileGeo g = tile_geometry<TL, N>(a.axis, a.nx, a.ny, a.nz);
  const bool fwd = (a.kind == REDFT10 || a.kind == RODFT10);
  const bool sine = (a.kind == RODFT10 || a.kind == RODFT01);

  // ---- load: every global load of the tile is issued before the first LDS
  // write, so each thread keeps all of its loads in flight -----------------
  if (a.axis == 2) {
    constexpr int E = TL * N;
    constexpr int IT = (E + NT - 1) / NT;
    T v[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      const int l = e / N, m = e - (e / N) * N;
      v[it] = T(0);
      if ((E % NT == 0 || e < E) && l < g.nvalid) {
        const long long addr = g.base + (long long)l * N + m;
        v[it] = a.in[addr];
        if (a.in2) v[it] += a.in2[addr];
      }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      if (E % NT == 0 || e < E) {
        const int l = e / N, m = e - (e / N) * N;
        int pos;
        T x = v[it];
        if (fwd) {
          if (sine && (m & 1)) x = -x;
          pos = (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1);
        } else {
          pos = sine ? (N - 1 - m) : m;
        }
        rbuf[2 * ((l >> 1) * N + pos) + (l & 1)] = x;
      }
    }
  } else {
    constexpr int E = P * N;  // pairs of adjacent columns
    constexpr int IT = (E + NT - 1) / NT;
    cplx<T> v[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      const int c = e % P, m = e / P;
      v[it] = cplx<T>{T(0), T(0)};
      if (E % NT == 0 || e < E) {
        const int l0 = 2 * c;
        const long long addr = g.base + l0 + (long long)m * g.stride_m;
        if (a.vec && l0 + 1 < g.nvalid) {
          v[it] = *reinterpret_cast<const cplx<T>*>(a.in + addr);
          if (a.in2) {
            const cplx<T> w = *reinterpret_cast<const cplx<T>*>(a.in2 + addr);
            v[it].x += w.x;
            v[it].y += w.y;
          }
        } else {
          if (l0 < g.nvalid) { v[it].x = a.in[addr]; if (a.in2) v[it].x += a.in2[addr]; }
          if (l0 + 1 < g.nvalid) { v[it].y = a.in[addr + 1]; if (a.in2) v[it].y += a.in2[addr + 1]; }
        }
      }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      if (E % NT == 0 || e < E) {
        const int c = e % P, m = e / P;
        int pos;
        cplx<T> x = v[it];
        if (fwd) {
          if (sine && (m & 1)) { x.x = -x.x; x.y = -x.y; }
          pos = (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1);
        } else {
          pos = sine ? (N - 1 - m) : m;
        }
        buf[c * N + pos] = x;
      }
    }
  }
  __syncthreads();

  if (!fwd) {
    // U_k = e^{+i pi k/2N} (Y_k - i Y_{N-k}) for the two lines of a pair,
    // packed as z_k = U0_k + i U1_k = c_k (Y0_k + Y1_{N-k}, Y1_k - Y0_{N-k}).
    constexpr int H = N / 2 + 1;
    constexpr int W = P * H;
    constexpr int ITW = (W + NT - 1) / NT;
    cplx<T> rk[ITW], rn[ITW];
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      const int w = tid + it * NT;
      if (w < W) {
        const int c = w / H;
        const int k = w - c * H;
        rk[it] = buf[c * N + (k < N ? k : 0)];
        rn[it] = (k > 0) ? buf[c * N + N - k] : cplx<T>{T(0), T(0)};
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      const int w = tid + it * NT;
      if (w < W) {
        const int c = w / H;
        const int k = w - c * H;
        if (k < N) {
          cplx<T> cc = a.mk[k];
          cc.y = -cc.y;
          buf[c * N + k] = cmul(cc, cplx<T>{rk[it].x + rn[it].y, rk[it].y - rn[it].x});
        }
        if (k > 0 && k < N - k) {
          cplx<T> cc = a.mk[N - k];
          cc.y = -cc.y;
          buf[c * N + N - k] = cmul(cc, cplx<T>{rn[it].x + rk[it].y, rn[it].y - rk[it].x});
        }
      }
    }
    __syncthreads();
    fft_stage<T, N, P, NT, MAXR, +1, 0>(buf, a.tw);
    // ---- store: y_{2n} = u_n, y_{2n+1} = u_{N-1-n}, line 2c from Re, 2c+1 from Im
    constexpr int HN = N / 2;
    constexpr int WO = P * HN;
    constexpr int ITO = (WO + NT - 1) / NT;
    const T s = sine ? T(-1) : T(1);
#pragma unroll
    for (int it = 0; it < ITO; ++it) {
      const int w = tid + it * NT;
      if (WO % NT == 0 || w < WO) {
        int c, n;
        if (a.axis == 2) { c = w / HN; n = w - c * HN; }
        else { n = w / P; c = w - n * P; }
        const cplx<T> ua = buf[c * N + n];
        const cplx<T> ub = buf[c * N + N - 1 - n];
        const int l0 = 2 * c;
        if (a.axis == 2) {
          if (l0 < g.nvalid) {
            T* o = a.out + g.base + (long long)l0 * N + 2 * n;
            if (a.vec) *reinterpret_cast<cplx<T>*>(o) = cplx<T>{ua.x, s * ub.x};
            else { o[0] = ua.x; o[1] = s * ub.x; }
          }
          if (l0 + 1 < g.nvalid) {
            T* o = a.out + g.base + (long long)(l0 + 1) * N + 2 * n;
            if (a.vec) *reinterpret_cast<cplx<T>*>(o) = cplx<T>{ua.y, s * ub.y};
            else { o[0] = ua.y; o[1] = s * ub.y; }
          }
        } else {
          T* o0 = a.out + g.base + l0 + (long long)(2 * n) * g.stride_m;
          T* o1 = a.out + g.base + l0 + (long long)(2 * n + 1) * g.stride_m;
          if (a.vec && l0 + 1 < g.nvalid) {
            *reinterpret_cast<cplx<T>*>(o0) = ua;
            *reinterpret_cast<cplx<T>*>(o1) = cplx<T>{s * ub.x, s * ub.y};
          } else {
            if (l0 < g.nvalid) { o0[0] = ua.x; o1[0] = s * ub.x; }
            if (l0 + 1 < g.nvalid) { o0[1] = ua.y; o1[1] = s * ub.y; }
          }
        }
      }
    }
  } else {
    fft_stage<T, N, P, NT, MAXR, -1, 0>(buf, a.tw);
    // ---- post: Y0_k = Re(mk_k (Z_k + conj Z_{N-k})), Y1_k = Im(mk_k (Z_k - conj Z_{N-k}))
    constexpr int H = N / 2 + 1;
    constexpr int W = P * H;
    constexpr int ITW = (W + NT - 1) / NT;
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      const int w = tid + it * NT;
      if (w < W) {
        int c, k;
        if (a.axis == 2) { c = w / H; k = w - c * H; }
        else { k = w / P; c = w - k * P; }
        const int kr = (k == 0) ? 0 : N - k;
        const cplx<T> zk = buf[c * N + k];
        const cplx<T> zr = buf[c * N + kr];
        T y0k, y1k, y0r, y1r;
        {
          const cplx<T> m = a.mk[k];
          const cplx<T> S = {zk.x + zr.x, zk.y - zr.y};
          const cplx<T> D = {zk.x - zr.x, zk.y + zr.y};
          y0k = m.x * S.x - m.y * S.y;
          y1k = m.x * D.y + m.y * D.x;
        }
        {
          const cplx<T> m = a.mk[kr];
          const cplx<T> S = {zr.x + zk.x, zr.y - zk.y};
          const cplx<T> D = {zr.x - zk.x, zr.y + zk.y};
          y0r = m.x * S.x - m.y * S.y;
          y1r = m.x * D.y + m.y * D.x;
        }
        const int ok = sine ? (N - 1 - k) : k;
        const int orr = sine ? (N - 1 - kr) : kr;
        const bool has_r = (k > 0 && k < N - k);
        const int l0 = 2 * c;
        if (a.axis == 2) {
          if (l0 < g.nvalid) {
            T* o = a.out + g.base + (long long)l0 * N;
            o[ok] = y0k;
            if (has_r) o[orr] = y0r;
          }
          if (l0 + 1 < g.nvalid) {
            T* o = a.out + g.base + (long long)(l0 + 1) * N;
            o[ok] = y1k;
            if (has_r) o[orr] = y1r;
          }
        } else {
          T* ok_p = a.out + g.base + l0 + (long long)ok * g.stride_m;
          T* or_p = a.out + g.base + l0 + (long long)orr * g.stride_m;
          if (a.vec && l0 + 1 < g.nvalid) {
            *reinterpret_cast<cplx<T>*>(ok_p) = cplx<T>{y0k, y1k};
            if (has_r) *reinterpret_cast<cplx<T>*>(or_p) = cplx<T>{y0r, y1r};
          } else {
            if (l0 < g.nvalid) { ok_p[0] = y0k; if (has_r) or_p[0] = y0r; }
            if (l0 + 1 < g.nvalid) { ok_p[1] = y1k; if (has_r) or_p[1] = y1r; }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// direct O(N^2) fallback for any length
// ---------------------------------------------------------------------------
constexpr int kThreads = 256;

template <typename T>
struct DirectArgs {
  const T* in;
  const T* in2;
  T* out;
  const T* costab;  // 4n entries: cos(pi t/(2n))
  int nx, ny, nz;
  int axis;
  int kind;
  int n;
  int tl;
};

template <typename T>
__global__ __launch_bounds__(kThreads) void direct_pass_kernel(DirectArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* x = reinterpret_cast<T*>(smem_raw);
  const int n = a.n, tl = a.tl, n4 = 4 * n;
  TileGeo g;
  if (a.axis == 2) {
    long long first = (long long)blockIdx.x * tl;
    long long nl = (long long)a.nx * a.ny;
    g.nvalid = (int)((nl - first) < tl ? (nl - first) : tl);
    g.base = first * n;
    g.line_step = n;
    g.stride_m = 1;
  } else {
    int tpo = (a.nz + tl - 1) / tl;
    int o = blockIdx.x / tpo;
    int k0 = (blockIdx.x - o * tpo) * tl;
    g.nvalid = (a.nz - k0) < tl ? (a.nz - k0) : tl;
    g.line_step = 1;
    if (a.axis == 1) { g.base = (long long)o * a.ny * a.nz + k0; g.stride_m = a.nz; }
    else { g.base = (long long)o * a.nz + k0; g.stride_m = (long long)a.ny * a.nz; }
  }
  const int E = tl * n;
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int l, m;
    if (a.axis == 2) { l = e / n; m = e - l * n; }
    else { m = e / tl; l = e - m * tl; }
    T v = T(0);
    if (l < g.nvalid) {
      const long long addr = g.base + l * g.line_step + m * g.stride_m;
      v = a.in[addr];
      if (a.in2) v += a.in2[addr];
    }
    x[l * n + m] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int l, k;
    if (a.axis == 2) { l = e / n; k = e - l * n; }
    else { k = e / tl; l = e - k * tl; }
    if (l >= g.nvalid) continue;
    const T* xl = x + l * n;
    T acc = T(0);
    // angle index t (units of pi/(2n)) is advanced incrementally mod 4n
    if (a.kind == REDFT10) {            // 2 sum x_m cos(pi (2m+1) k / 2n)
      int t = k % n4, step = (2 * k) % n4;
      for (int m = 0; m < n; ++m) { acc += xl[m] * a.costab[t]; t += step; if (t >= n4) t -= n4; }
      acc *= T(2);
    } else if (a.kind == REDFT01) {     // x_0 + 2 sum_{m>=1} x_m cos(pi m (2k+1)/2n)
      int step = (2 * k + 1) % n4, t = step;
      for (int m = 1; m < n; ++m) { acc += xl[m] * a.costab[t]; t += step; if (t >= n4) t -= n4; }
      acc = xl[0] + T(2) * acc;
    } else if (a.kind == RODFT10) {     // 2 sum x_m sin(pi (2m+1)(k+1)/2n)
      int step = (2 * (k + 1)) % n4;
      int t = ((k + 1) - n) % n4; if (t < 0) t += n4;   // sin(u) = cos(u - pi/2)
      for (int m = 0; m < n; ++m) { acc += xl[m] * a.costab[t]; t += step; if (t >= n4) t -= n4; }
      acc *= T(2);
    } else {                            // (-1)^k x_{n-1} + 2 sum_{m<n-1} x_m sin(pi (m+1)(2k+1)/2n)
      int step = (2 * k + 1) % n4;
      int t = (step - n) % n4; if (t < 0) t += n4;
      for (int m = 0; m < n - 1; ++m) { acc += xl[m] * a.costab[t]; t += step; if (t >= n4) t -= n4; }
      acc = ((k & 1) ? -xl[n - 1] : xl[n - 1]) + T(2) * acc;
    }
    a.out[g.base + l * g.line_step + k * g.stride_m] = acc;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool R2R::fft_supported(int n) { return fft_radix(n, 16, 0) != 0; }

R2R::~R2R() {
  for (auto& kv : cache_) {
    if (kv.second.tw) (void)hipFree(kv.second.tw);
    if (kv.second.mk) (void)hipFree(kv.second.mk);
    if (kv.second.costab) (void)hipFree(kv.second.costab);
  }
}

hipError_t R2R::tables(DType dt, int n, Tables** out) {
  std::lock_guard<std::mutex> lk(mu_);
  auto key = std::make_pair((int)dt, n);
  auto it = cache_.find(key);
  if (it != cache_.end()) { *out = &it->second; return hipSuccess; }
  const size_t es = dt == F64 ? 8 : 4;
  std::vector<double> tw(2 * n), mk(2 * n), ct(4 * n);
  for (int m = 0; m < n; ++m) {
    // angles reduced exactly in integers before the libm call
    const double a = -2.0 * M_PI * (double)m / (double)n;
    tw[2 * m] = std::cos(a);
    tw[2 * m + 1] = std::sin(a);
    const double b = -M_PI * (double)m / (2.0 * n);
    mk[2 * m] = std::cos(b);
    mk[2 * m + 1] = std::sin(b);
  }
  for (int t = 0; t < 4 * n; ++t) ct[t] = std::cos(M_PI * (double)t / (2.0 * n));
  Tables tb;
  hipError_t e;
  auto upload = [&](const std::vector<double>& v, void** d) -> hipError_t {
    hipError_t r = hipMalloc(d, v.size() * es);
    if (r != hipSuccess) return r;
    if (dt == F64) return hipMemcpy(*d, v.data(), v.size() * es, hipMemcpyHostToDevice);
    std::vector<float> f(v.begin(), v.end());
    return hipMemcpy(*d, f.data(), f.size() * es, hipMemcpyHostToDevice);
  };
  if ((e = upload(tw, &tb.tw)) != hipSuccess) return e;
  if ((e = upload(mk, &tb.mk)) != hipSuccess) return e;
  if ((e = upload(ct, &tb.costab)) != hipSuccess) return e;
  auto res = cache_.emplace(key, tb);
  *out = &res.first->second;
  return hipSuccess;
}

static inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

template <typename T, int N, int P, int NT>
hipError_t launch_fft_pn(const PassIO& io, void* tw, void* mk, int kind, int axis, const Geom& g,
                         hipStream_t s) {
  constexpr int TL = 2 * P;
  PassArgs<T> a;
  a.in = static_cast<const T*>(io.in);
  a.in2 = static_cast<const T*>(io.in2);
  a.out = static_cast<T*>(io.out);
  a.tw = static_cast<const cplx<T>*>(tw);
  a.mk = static_cast<const cplx<T>*>(mk);
  a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
  a.axis = axis;
  a.kind = kind;
  a.vec = (g.nz % 2 == 0) && aligned16(io.in) && aligned16(io.out) &&
          (io.in2 == nullptr || aligned16(io.in2));
  long long blocks;
  if (axis == 2) blocks = ((long long)g.nx * g.ny + TL - 1) / TL;
  else blocks = (long long)(axis == 1 ? g.nx : g.ny) * ((g.nz + TL - 1) / TL);
  const size_t lds = (size_t)P * N * sizeof(cplx<T>);
  hipLaunchKernelGGL((fft_pass_kernel<T, N, P, NT>), dim3((unsigned)blocks), dim3(NT), lds, s, a);
  return hipGetLastError();
}

// Tile and block size per (type, length, axis).  The strided y/x passes keep
// TL*sizeof(T) >= 128 bytes so that every row segment is a whole cache line.
template <typename T, int N>
struct TileChoice {
  static constexpr int cap(int p) { return p < 1 ? 1 : p; }
  static constexpr int budget(int b) {
    int p = b / N, q = 1;
    while (q * 2 <= p) q *= 2;
    return q;
  }
  static constexpr int PZ = cap(budget(2048) < 8 ? budget(2048) : 8);
  static constexpr int PS = cap(budget(4096) < (int)(64 / sizeof(T)) ? budget(4096) : (int)(64 / sizeof(T)));
  static constexpr int NTZ = 256;
  static constexpr int NTS = 256;
};

template <typename T, int N>
static hipError_t launch_fft(const PassIO& io, void* tw, void* mk, int kind, int axis,
                             const Geom& g, hipStream_t s) {
  using C = TileChoice<T, N>;
  if (axis == 2) return launch_fft_pn<T, N, C::PZ, C::NTZ>(io, tw, mk, kind, axis, g, s);
  return launch_fft_pn<T, N, C::PS, C::NTS>(io, tw, mk, kind, axis, g, s);
}

template <typename T>
static hipError_t launch_direct(const PassIO& io, void* costab, int kind, int axis, int n,
                                const Geom& g, hipStream_t s) {
  DirectArgs<T> a;
  a.in = static_cast<const T*>(io.in);
  a.in2 = static_cast<const T*>(io.in2);
  a.out = static_cast<T*>(io.out);
  a.costab = static_cast<const T*>(costab);
  a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
  a.axis = axis;
  a.kind = kind;
  a.n = n;
  int tl = 16;
  while (tl > 1 && (size_t)tl * n * sizeof(T) > 48 * 1024) tl /= 2;
  a.tl = tl;
  long long blocks;
  if (axis == 2) blocks = ((long long)g.nx * g.ny + tl - 1) / tl;
  else blocks = (long long)(axis == 1 ? g.nx : g.ny) * ((g.nz + tl - 1) / tl);
  const size_t lds = (size_t)tl * n * sizeof(T);
  hipLaunchKernelGGL((direct_pass_kernel<T>), dim3((unsigned)blocks), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

template <typename T>
static hipError_t dispatch(int n, const PassIO& io, void* tw, void* mk, void* ct, int kind,
                           int axis, const Geom& g, hipStream_t s) {
  switch (n) {
    case 16: return launch_fft<T, 16>(io, tw, mk, kind, axis, g, s);
    case 32: return launch_fft<T, 32>(io, tw, mk, kind, axis, g, s);
    case 48: return launch_fft<T, 48>(io, tw, mk, kind, axis, g, s);
    case 64: return launch_fft<T, 64>(io, tw, mk, kind, axis, g, s);
    case 96: return launch_fft<T, 96>(io, tw, mk, kind, axis, g, s);
    case 128: return launch_fft<T, 128>(io, tw, mk, kind, axis, g, s);
    case 192: return launch_fft<T, 192>(io, tw, mk, kind, axis, g, s);
    case 256: return launch_fft<T, 256>(io, tw, mk, kind, axis, g, s);
    case 384: return launch_fft<T, 384>(io, tw, mk, kind, axis, g, s);
    case 512: return launch_fft<T, 512>(io, tw, mk, kind, axis, g, s);
    case 1024: return launch_fft<T, 1024>(io, tw, mk, kind, axis, g, s);
    default: return launch_direct<T>(io, ct, kind, axis, n, g, s);
  }
}

hipError_t R2R::pass(DType dt, int kind, int axis, const PassIO& io, const Geom& g,
                     hipStream_t s) {
  const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  Tables* tb = nullptr;
  hipError_t e = tables(dt, n, &tb);
  if (e != hipSuccess) return e;
  if (dt == F64) return dispatch<double>(n, io, tb->tw, tb->mk, tb->costab, kind, axis, g, s);
  return dispatch<float>(n, io, tb->tw, tb->mk, tb->costab, kind, axis, g, s);
}

hipError_t R2R::transform3d(DType dt, const int kinds[3], const PassIO& io, const Geom& g,
                            hipStream_t s) {
  PassIO p = io;
  hipError_t e = pass(dt, kinds[2], 2, p, g, s);
  if (e != hipSuccess) return e;
  p.in = io.out;
  p.in2 = nullptr;
  if ((e = pass(dt, kinds[1], 1, p, g, s)) != hipSuccess) return e;
  return pass(dt, kinds[0], 0, p, g, s);
}

}  // namespace smaiso
