// Axis passes of length 256 along y and x (the strided axes), FP64
// arithmetic, for FP64 or FP32 storage: plain forward (DCT-II/DST-II) and
// inverse (DCT-III/DST-III) passes, and the forward -> operator -> inverse
// passes (one-axis operators OP_AX_*, the x-pass Laplacian OP_MID_*).
//
// Layout: a workgroup holds P complex sequences (column pairs 2c, 2c+1 of a
// tile of 2P consecutive k columns) and 16 P threads; thread (c, j) owns
// radix-16 butterfly j of sequence c in both FFT stages (256 = 16 x 16), so
// a transform takes two LDS exchanges instead of one per stage:
//  * forward: the thread loads its 16 elements straight from HBM in
//    Makhoul's order (v[j + 16 q] = x[2(j + 16 q)] or x[511 - 2(j + 16 q)]),
//    runs the first radix-16 butterfly in registers, exchanges through LDS,
//    runs the second (twiddles w^q, w = exp(-2 pi i j/256), by products) and
//    holds Z[j + 16 q]; Makhoul's post-twiddle pairs Z[k] with Z[256 - k],
//    which is read from LDS;
//  * inverse: Y_k and Y_{256-k} (through LDS) give Makhoul's pre-twiddled
//    z_k at the positions k = j + 16 q that the thread's first butterfly
//    needs, and the outputs u_n go to one row each (y_{2n} = u_n,
//    y_{2n+1} = u_{255-n});
//  * forward -> operator -> inverse: the post-twiddle gives the forward
//    outputs at k and 256 - k, which hold every source of the operator's
//    outputs at those two positions (the one-mode shifts compose with the DST
//    index reversal), so the pre-twiddle follows in registers.
// Workgroups are persistent: each walks tiles t = blockIdx.x + i gridDim.x
// and loads the next tile into registers before transforming the current
// one, so its loads overlap its own arithmetic and stores.  The threads hold
// 32 complex values (no register cap: one wavefront per SIMD).
#pragma once

#include "r2r_kernels.h"

namespace smaiso {
namespace yx256 {

constexpr int kN = 256;
// 16-byte LDS slots per sequence: one spare after every 16 items (the
// first-stage writes of a thread are 16 consecutive items) and one at the end
constexpr int kSeq = kN + kN / 16 + 1;
__device__ __forceinline__ int slot(int x) { return x + (x >> 4); }

// pass structure
constexpr int kFwd = 0;    // DCT-II / DST-II (kind)
constexpr int kInv = 1;    // DCT-III / DST-III (kind)
constexpr int kFwdOp = 2;  // forward, operator OP, inverse

template <typename T, typename TO>
struct Args {
  const T* in;
  TO* out;
  const void* tw;  // exp(-2 pi i m/256), cplx<A>
  const void* mk;  // exp(-i pi m/512), cplx<A>
  int nx, ny, nz;
  int kind;
  int ntiles;  // tiles per member
  int tpo;     // tiles per outer index (x passes)
  Fuse f;
};

template <typename A, typename T>
__device__ __forceinline__ cplx<A> ld2(const T* p) {
  const cplx<T> u = *reinterpret_cast<const cplx<T>*>(p);
  return cplx<A>{(A)u.x, (A)u.y};
}

template <typename TO>
__device__ __forceinline__ void st2(TO* p, double a, double b, bool acc) {
  if (acc) {
    const cplx<TO> o = *reinterpret_cast<const cplx<TO>*>(p);
    a += (double)o.x;
    b += (double)o.y;
  }
  *reinterpret_cast<cplx<TO>*>(p) = cplx<TO>{(TO)a, (TO)b};
}

// v[q] *= w^q, q = 1..15, powers of w by a product tree of depth 3
template <typename A>
__device__ __forceinline__ void stage_twiddle(cplx<A>* v, cplx<A> w1) {
  using C = cplx<A>;
  const C w2 = cmul(w1, w1), w3 = cmul(w2, w1), w4 = cmul(w2, w2);
  const C w8 = cmul(w4, w4), w12 = cmul(w8, w4);
  const C lo[4] = {C{A(1), A(0)}, w1, w2, w3};
  const C hi[4] = {C{A(1), A(0)}, w4, w8, w12};
#pragma unroll
  for (int q = 1; q < 16; ++q) {
    const int u = q >> 2, l = q & 3;
    const C w = u == 0 ? lo[l] : (l == 0 ? hi[u] : cmul(hi[u], lo[l]));
    v[q] = cmul(v[q], w);
  }
}

template <int OP>
constexpr bool ax_op() {
  return OP == OP_AX_GRAD || OP == OP_AX_DIV || OP == OP_AX_GAUSS || OP == OP_AX_GRAD_T || OP == OP_AX_DIV_T;
}

// Register cap, as waves per SIMD, for FP64 storage.  The FP64 tile's LDS
// (8 sequences, 35 KiB, 128 threads) allows four workgroups per CU, two
// waves per SIMD; the forward passes fit that without spilling, the inverse
// and operator passes (one more LDS exchange, or two transforms) keep the
// full register file.  FP32 tiles (16 sequences, 256 threads) measured
// fastest with the full register file and the tables held across tiles.
template <typename T, int MODE>
constexpr int min_waves() {
  return sizeof(T) == 8 && MODE == kFwd ? 2 : 1;
}

// AXIS: 1 = y pass, 0 = x pass.  P sequences (2P columns) per tile.  A:
// the arithmetic type of the tile and the tables (FP64, or FP32 for FP32
// storage where FP32 arithmetic is asked for).
template <typename T, typename TO, int AXIS, int MODE, int OP, int P, typename A = double>
__global__ __launch_bounds__(16 * P, (min_waves<T, MODE>())) void pass_kernel(Args<T, TO> a) {
  using C = cplx<A>;
  constexpr int N = kN;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* const buf = reinterpret_cast<C*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid % P, j = tid / P;
  C* const s = buf + c * kSeq;
  const long long mo = (long long)blockIdx.y * ((long long)a.nx * a.ny * a.nz);
  const int stride = AXIS == 1 ? a.nz : a.ny * a.nz;
  // sine on the load side (forward DST-II input, inverse DST-III input
  // reversal) and on the store side (forward DST output reversal, inverse
  // DST-III odd-output sign)
  constexpr bool AX = ax_op<OP>();
  const bool sine_in = MODE == kFwd ? a.kind == RODFT10
                       : MODE == kInv ? a.kind == RODFT01
                                      : (OP == OP_AX_DIV || OP == OP_AX_GRAD_T);
  const bool sine_out = MODE == kFwd ? sine_in
                        : MODE == kInv ? sine_in
                                       : (OP == OP_AX_GRAD || OP == OP_AX_DIV_T);
  const bool acc = AX && a.f.accumulate;
  const C* const tw = reinterpret_cast<const C*>(a.tw);
  const C* const mkt = reinterpret_cast<const C*>(a.mk);
  const C wf = tw[j];                         // stage-2 twiddle base, forward
  const C wi = C{wf.x, -wf.y};                // and inverse
  const C mk0 = mkt[j], mks = mkt[16];        // Makhoul factors mk[j + 16 q] by products

  auto tile = [&](int t, int& outer, int& k0) -> long long {
    int o, kc;
    if (AXIS == 1) { kc = t / a.nx; o = t - kc * a.nx; }  // plane index fastest
    else { o = t / a.tpo; kc = t - o * a.tpo; }
    outer = o;
    k0 = kc * 2 * P;
    return mo + (AXIS == 1 ? (long long)o * a.ny * a.nz : (long long)o * a.nz) + k0;
  };
  // the 16 elements a thread loads for tile t
  auto load = [&](C* v, int t) {
    int outer, k0;
    const T* p = a.in + tile(t, outer, k0) + 2 * c;
    if constexpr (MODE == kInv) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = j + 16 * q;
        v[q] = ld2<A>(p + (sine_in ? N - 1 - k : k) * stride);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int row = q < 8 ? 2 * j + 32 * q : 2 * N - 1 - 2 * j - 32 * q;
        v[q] = ld2<A>(p + row * stride);
      }
    }
  };
  // first inverse butterfly on z_{j+16q}, LDS exchange, second butterfly,
  // store y_{2n} = u_n, y_{2n+1} = u_{N-1-n} (times -1 for DST-III)
  auto inverse_tail = [&](C* v, TO* o, C wi_t, int j_t) {
    dft<A, 16, +1>(v);
    __syncthreads();  // every read of this tile's slots is done
#pragma unroll
    for (int q = 0; q < 16; ++q) s[slot(16 * j + q)] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = s[slot(j + 16 * q)];
    stage_twiddle(v, wi_t);
    dft<A, 16, +1>(v);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int n = j_t + 16 * q;
      if (q < 8) st2(o + (2 * n) * stride, v[q].x, v[q].y, acc);
      else if (sine_out) st2(o + (2 * N - 1 - 2 * n) * stride, -v[q].x, -v[q].y, acc);
      else st2(o + (2 * N - 1 - 2 * n) * stride, v[q].x, v[q].y, acc);
    }
  };

  // the plain passes load the next tile before transforming the current one;
  // the operator passes (two transforms per tile) load it afterwards, which
  // keeps their registers within the file
  constexpr bool kPrefetch = MODE != kFwdOp;
  C v[16], nxt[16];
  int t = blockIdx.x;
  if (t < a.ntiles) load(v, t);
  for (; t < a.ntiles; t += gridDim.x) {
    const int tn = t + gridDim.x;
    if (kPrefetch && tn < a.ntiles) load(nxt, tn);
    // The twiddle powers, Makhoul factors and operator weights of a thread
    // are the same for every tile; held across the loop in FP64 they would
    // take about half the register file.  Passing their seeds through an
    // empty asm statement each tile makes them values the compiler cannot
    // hoist, so they are recomputed per tile (a few dozen ALU operations).
    C wf_t = wf, wi_t = wi, mk0_t = mk0, mks_t = mks;
    int j_t = j;
    if constexpr (sizeof(T) == 8) {
      asm volatile("" : "+v"(j_t), "+v"(wf_t.x), "+v"(wf_t.y), "+v"(wi_t.x), "+v"(wi_t.y));
      asm volatile("" : "+v"(mk0_t.x), "+v"(mk0_t.y), "+v"(mks_t.x), "+v"(mks_t.y));
    }
    int outer, k0;
    TO* const o = a.out + tile(t, outer, k0) + 2 * c;
    if constexpr (MODE == kInv) {
      // pre-twiddle: z_k = conj(mk[k]) (Y0_k + Y1_{N-k} + i (Y1_k - Y0_{N-k})), Y_N = 0
#pragma unroll
      for (int q = 0; q < 16; ++q) s[slot(j + 16 * q)] = v[q];
      __syncthreads();
      C m = mk0_t;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = j_t + 16 * q;
        // FP64: both from LDS, so the loaded values need no registers past
        // the barrier
        const C rk = sizeof(T) == 8 ? s[slot(k)] : v[q];
        const C rn = k == 0 ? C{A(0), A(0)} : s[slot((N - k) & (N - 1))];
        v[q] = cmul(C{m.x, -m.y}, C{rk.x + rn.y, rk.y - rn.x});
        m = cmul(m, mks_t);
      }
      inverse_tail(v, o, wi_t, j_t);
    } else {
      if (sine_in) {
#pragma unroll
        for (int q = 8; q < 16; ++q) { v[q].x = -v[q].x; v[q].y = -v[q].y; }  // odd rows
      }
      dft<A, 16, -1>(v);
#pragma unroll
      for (int q = 0; q < 16; ++q) s[slot(16 * j + q)] = v[q];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = s[slot(j + 16 * q)];
      stage_twiddle(v, wf_t);
      dft<A, 16, -1>(v);
      // v[q] = Z[j + 16 q]; the slots written are the ones this thread read
#pragma unroll
      for (int q = 0; q < 16; ++q) s[slot(j + 16 * q)] = v[q];
      __syncthreads();
      C m = mk0_t;
      if constexpr (MODE == kFwd) {
        // Y_k = (Re(mk[k] (Z_k + conj Z_{N-k})), Im(mk[k] (Z_k - conj Z_{N-k})))
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int k = j_t + 16 * q;
          const C za = v[q], pa = s[slot((N - k) & (N - 1))];
          const A y0 = m.x * (za.x + pa.x) - m.y * (za.y - pa.y);
          const A y1 = m.x * (za.y + pa.y) + m.y * (za.x - pa.x);
          st2(o + (sine_out ? N - 1 - k : k) * stride, y0, y1, false);
          m = cmul(m, mks_t);
        }
      } else {
        // y and z parts of |q|^2 of the pair's two columns (mid op)
        C qyz = {A(0), A(0)};
        if constexpr (!AX) {
          const double qy = vq_at(a.f, 1, outer);
          const double q0 = vq_at(a.f, 2, k0 + 2 * c), q1 = vq_at(a.f, 2, k0 + 2 * c + 1);
          qyz = C{(A)(qy * qy + q0 * q0), (A)(qy * qy + q1 * q1)};
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int k = j_t + 16 * q;
          const int kb = (N - k) & (N - 1);
          const C za = sizeof(T) == 8 ? s[slot(k)] : v[q], zb = s[slot(kb)];
          const C mb = k == 0 ? m : C{-m.y, -m.x};  // mk[N - k], exact
          // forward outputs of both lines at positions k and kb
          const C ya = {m.x * (za.x + zb.x) - m.y * (za.y - zb.y), m.x * (za.y + zb.y) + m.y * (za.x - zb.x)};
          const C yb = {mb.x * (zb.x + za.x) - mb.y * (zb.y - za.y), mb.x * (zb.y + za.y) + mb.y * (zb.x - za.x)};
          C xk, xb;
          if constexpr (AX) {
            const int fa = sine_in ? N - 1 - k : k;
            const int pa = sine_out ? N - 1 - k : k, pb = sine_out ? N - 1 - kb : kb;
            const int sa = ax_source<OP>(pa, N), sb = ax_source<OP>(pb, N);
            xk = C{A(0), A(0)};
            xb = C{A(0), A(0)};
            if (sa >= 0) {
              const A w = (A)ax_weight<T, OP>(a.f, pa, N);
              const C y = sa == fa ? ya : yb;
              xk = C{w * y.x, w * y.y};
            }
            if (sb >= 0) {
              const A w = (A)ax_weight<T, OP>(a.f, pb, N);
              const C y = sb == fa ? ya : yb;
              xb = C{w * y.x, w * y.y};
            }
          } else {
            // x pass of -scale |q|^2 (OP_MID_NEGLAP): mode index k along x,
            // y index = outer, z columns k0 + 2c and k0 + 2c + 1, in the
            // reference's order (-scale |q|^2) y
            const A qk = (A)vq_at(a.f, 0, k), qb = (A)vq_at(a.f, 0, kb), sc = (A)a.f.scale;
            xk = C{(-sc * (qk * qk + qyz.x)) * ya.x, (-sc * (qk * qk + qyz.y)) * ya.y};
            xb = C{(-sc * (qb * qb + qyz.x)) * yb.x, (-sc * (qb * qb + qyz.y)) * yb.y};
          }
          const C rn = k == 0 ? C{A(0), A(0)} : xb;
          v[q] = cmul(C{m.x, -m.y}, C{xk.x + rn.y, xk.y - rn.x});
          m = cmul(m, mks_t);
          if ((q & 1) == 1) __builtin_amdgcn_sched_barrier(0);  // bounds the registers in flight
        }
        inverse_tail(v, o, wi_t, j_t);
      }
    }
    __syncthreads();  // the next tile's first writes reuse slots other threads read last
    if (tn < a.ntiles) {
      if (kPrefetch) {
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = nxt[q];
      } else {
        load(v, tn);
      }
    }
  }
}

// Sequences per tile: 16 columns of FP64 or 32 of FP32 (128-byte row
// segments).
template <typename T>
constexpr int tile_p() { return sizeof(T) == 8 ? 8 : 16; }

}  // namespace yx256

}  // namespace smaiso
