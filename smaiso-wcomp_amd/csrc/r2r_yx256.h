// Plain axis passes of length 256 along y and x (the strided axes): forward
// (DCT-II/DST-II) and inverse (DCT-III/DST-III), FP64 or FP32 storage, FP64
// arithmetic (FP32 where FP32 arithmetic is asked for with FP32 storage).
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
//    y_{2n+1} = u_{255-n}).
// Workgroups are persistent: each walks tiles t = blockIdx.x + i gridDim.x
// and loads the next tile into registers before transforming the current
// one, so its loads overlap its own arithmetic and stores.  The threads hold
// 32 complex values (register caps in min_waves).
// The one-axis operator passes (axop_kernel below) run forward transform,
// operator and inverse transform per tile on the same layout, with the
// operator applied in registers between the two.
#pragma once

#include "r2r_kernels.h"

namespace smaiso {
namespace yx256 {

constexpr int kN = 256;
// 16-byte LDS slots per sequence: one spare after every 16 items (the
// first-stage writes of a thread are 16 consecutive items) and one at the end
constexpr int kSeq = kN + kN / 16 + 1;
__device__ __forceinline__ int slot(int x) { return x + (x >> 4); }
// slot(16 j + q) and slot(j + 16 q) for 0 <= j, q < 16, written so that the
// q-dependence is a constant: the 16 accesses of a thread then share one
// address register with immediate offsets (as slot(), the compiler kept one
// register per access)
__device__ __forceinline__ int row_slot(int j, int q) { return 17 * j + q; }
__device__ __forceinline__ int col_slot(int j, int q) { return j + 17 * q; }
// slot of the partner 256 - k of k = j + 16 q (k = 0 pairs with itself):
// 271 - j - 17 q for j > 0, 272 - 17 q for j = 0 < q
__device__ __forceinline__ int partner_slot(int j, int q) {
  return (j == 0 && q == 0) ? 0 : (j == 0 ? 272 : 271 - j) - 17 * q;
}

// pass structure
constexpr int kFwd = 0;  // DCT-II / DST-II (kind)
constexpr int kInv = 1;  // DCT-III / DST-III (kind)

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

// Register cap, as waves per SIMD, and whether the pass loads the next tile
// while it transforms the current one (measured at 256^3 on MI355X).  An FP64
// tile of 8 sequences (35 KiB LDS, 128 threads) allows four workgroups per
// CU, two waves per SIMD; an FP64 tile of 16 sequences over FP32 storage (70
// KiB, 256 threads) two workgroups, two waves per SIMD.  FP64 arithmetic: the
// forward pass fits two waves with the prefetch (FP32 storage: y pass 44 ->
// 36 us against one wave); the inverse pass (one more LDS exchange) runs at
// two waves without the prefetch (FP64 storage: 244 VGPRs, same-box FP64 step
// 279.2 -> 280.5 steps/s against one wave with the prefetch, whose LDS
// addresses once took a register each; FP32 storage: 43 -> 37-42 us).  FP32
// arithmetic (16 sequences, 256 threads) measured fastest with the full
// register file and the tables held across tiles.
template <typename T, int MODE, typename A>
constexpr int min_waves() {
  return sizeof(A) == 8 ? 2 : 1;
}
template <typename T, int MODE, typename A>
constexpr bool pass_prefetch() {
  return !(MODE == kInv && sizeof(A) == 8);
}

// AXIS: 1 = y pass, 0 = x pass.  MODE: kFwd or kInv.  P sequences (2P
// columns) per tile.  A: the arithmetic type of the tile and the tables
// (FP64, or FP32 for FP32 storage where FP32 arithmetic is asked for).
template <typename T, typename TO, int AXIS, int MODE, int P, typename A = double>
__global__ __launch_bounds__(16 * P, (min_waves<T, MODE, A>())) void pass_kernel(Args<T, TO> a) {
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
  const bool sine = MODE == kFwd ? a.kind == RODFT10 : a.kind == RODFT01;
  const C* const tw = reinterpret_cast<const C*>(a.tw);
  const C* const mkt = reinterpret_cast<const C*>(a.mk);
  const C wf = tw[j];                         // stage-2 twiddle base, forward
  const C wi = C{wf.x, -wf.y};                // and inverse
  const C mk0 = mkt[j], mks = mkt[16];        // Makhoul factors mk[j + 16 q] by products

  // first element of tile t (y passes: the plane index varies fastest)
  auto tile = [&](int t) -> long long {
    int o, kc;
    if (AXIS == 1) { kc = t / a.nx; o = t - kc * a.nx; }
    else { o = t / a.tpo; kc = t - o * a.tpo; }
    return mo + (AXIS == 1 ? (long long)o * a.ny * a.nz : (long long)o * a.nz) + kc * 2 * P;
  };
  // the 16 elements a thread loads for tile t, in the storage type
  auto load = [&](cplx<T>* v, int t) {
    const T* p = a.in + tile(t) + 2 * c;
    if constexpr (MODE == kInv) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = j + 16 * q;
        v[q] = *reinterpret_cast<const cplx<T>*>(p + (sine ? N - 1 - k : k) * stride);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int row = q < 8 ? 2 * j + 32 * q : 2 * N - 1 - 2 * j - 32 * q;
        v[q] = *reinterpret_cast<const cplx<T>*>(p + row * stride);
      }
    }
  };

  // the next tile is held in the storage type until it is transformed; the
  // FP64-arithmetic inverse pass loads each tile at its start instead (two
  // waves per SIMD without spilling the prefetch)
  constexpr bool kPrefetch = pass_prefetch<T, MODE, A>();
  C v[16];
  cplx<T> nxt[16];
  int t = blockIdx.x;
  if (kPrefetch && t < a.ntiles) load(nxt, t);
  for (; t < a.ntiles; t += gridDim.x) {
    if (!kPrefetch) load(nxt, t);
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = C{(A)nxt[q].x, (A)nxt[q].y};
    const int tn = t + gridDim.x;
    if (kPrefetch && tn < a.ntiles) load(nxt, tn);
    // The twiddle powers and Makhoul factors of a thread are the same for
    // every tile; held across the loop in FP64 they would take about half
    // the register file.  Passing their seeds through an empty asm statement
    // each tile makes them values the compiler cannot hoist, so they are
    // recomputed per tile (a few dozen ALU operations).
    C wf_t = wf, wi_t = wi, mk0_t = mk0, mks_t = mks;
    int j_t = j;
    if constexpr (sizeof(A) == 8) {
      asm volatile("" : "+v"(j_t), "+v"(wf_t.x), "+v"(wf_t.y), "+v"(wi_t.x), "+v"(wi_t.y));
      asm volatile("" : "+v"(mk0_t.x), "+v"(mk0_t.y), "+v"(mks_t.x), "+v"(mks_t.y));
    }
    TO* const o = a.out + tile(t) + 2 * c;
    if constexpr (MODE == kInv) {
      // pre-twiddle: z_k = conj(mk[k]) (Y0_k + Y1_{N-k} + i (Y1_k - Y0_{N-k})), Y_N = 0
#pragma unroll
      for (int q = 0; q < 16; ++q) s[col_slot(j, q)] = v[q];
      __syncthreads();
      C m = mk0_t;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = j_t + 16 * q;
        // FP64: both from LDS, so the loaded values need no registers past
        // the barrier
        const C rk = sizeof(A) == 8 ? s[col_slot(j_t, q)] : v[q];
        const C rn = k == 0 ? C{A(0), A(0)} : s[partner_slot(j_t, q)];
        v[q] = cmul(C{m.x, -m.y}, C{rk.x + rn.y, rk.y - rn.x});
        m = cmul(m, mks_t);
      }
      // first butterfly on z_{j+16q}, LDS exchange, second butterfly, store
      // y_{2n} = u_n, y_{2n+1} = u_{N-1-n} (times -1 for DST-III)
      dft<A, 16, +1>(v);
      __syncthreads();  // every read of this tile's slots is done
#pragma unroll
      for (int q = 0; q < 16; ++q) s[row_slot(j, q)] = v[q];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = s[col_slot(j, q)];
      stage_twiddle(v, wi_t);
      dft<A, 16, +1>(v);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int n = j_t + 16 * q;
        if (q < 8) st2(o + (2 * n) * stride, v[q].x, v[q].y, false);
        else if (sine) st2(o + (2 * N - 1 - 2 * n) * stride, -v[q].x, -v[q].y, false);
        else st2(o + (2 * N - 1 - 2 * n) * stride, v[q].x, v[q].y, false);
      }
    } else {
      if (sine) {
#pragma unroll
        for (int q = 8; q < 16; ++q) { v[q].x = -v[q].x; v[q].y = -v[q].y; }  // odd rows
      }
      dft<A, 16, -1>(v);
#pragma unroll
      for (int q = 0; q < 16; ++q) s[row_slot(j, q)] = v[q];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = s[col_slot(j, q)];
      stage_twiddle(v, wf_t);
      dft<A, 16, -1>(v);
      // v[q] = Z[j + 16 q]; the slots written are the ones this thread read
#pragma unroll
      for (int q = 0; q < 16; ++q) s[col_slot(j, q)] = v[q];
      __syncthreads();
      // Y_k = (Re(mk[k] (Z_k + conj Z_{N-k})), Im(mk[k] (Z_k - conj Z_{N-k})))
      C m = mk0_t;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = j_t + 16 * q;
        const C za = v[q], pa = s[partner_slot(j_t, q)];
        const A y0 = m.x * (za.x + pa.x) - m.y * (za.y - pa.y);
        const A y1 = m.x * (za.y + pa.y) + m.y * (za.x - pa.x);
        st2(o + (sine ? N - 1 - k : k) * stride, y0, y1, false);
        m = cmul(m, mks_t);
      }
    }
    __syncthreads();  // the next tile's first writes reuse slots other threads read last
  }
}

// The x pass of the forward transforms of fgrad and e rho (DCT-II along x)
// that also forms fgrad^ + G e rho^ (S:1479-1488, G = exp(-1.57^2 2 |q|^2)
// at mode (i, j, k), with the velocity solve's own expression), the one
// combination of the two the solve reads.  A sequence packs the fgrad line
// (real part) and the e rho line (imaginary part) at one (y, z) position, so
// Makhoul's post-processing yields both transforms at mode i and the pass
// stores one real value per point: P z-adjacent lines per tile of 16 P
// threads (128-byte row segments per field at P = 16).  G is taken as the
// product of the x factor exp(-c qx^2) (a table of the 256 modes in LDS,
// made once per workgroup) and exp(-c (qy^2 + qz^2)) (once per line), within
// a few ulp of the exponential of the sum.  FP64 storage and arithmetic;
// the next tile is loaded while the current one is transformed.
template <int P>
__global__ __launch_bounds__(16 * P, 2) void fgx_kernel(const double* fg, const double* __restrict__ er, double* out,
                                                      const void* twv, const void* mkv,
                                                      int ny, int nz, int ntiles, int tpo, double qs0, double qs1,
                                                      double qs2) {
  using A = double;
  using C = cplx<A>;
  constexpr int N = kN;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* const buf = reinterpret_cast<C*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid % P, j = tid / P;
  C* const s = buf + c * kSeq;
  const long long mo = (long long)blockIdx.y * ((long long)N * ny * nz);
  const int stride = ny * nz;
  const C* const tw = reinterpret_cast<const C*>(twv);
  const C* const mkt = reinterpret_cast<const C*>(mkv);
  const C wf = tw[j];
  const C mk0 = mkt[j], mks = mkt[16];
  constexpr A kG = A(-1.57 * 1.57 * 2);
  A* const gx = reinterpret_cast<A*>(buf + P * kSeq);  // exp(-c qx^2) of the N modes
  for (int i = tid; i < N; i += 16 * P) {
    const A qx = i * qs0;
    gx[i] = exp(kG * (qx * qx));
  }
  __syncthreads();
  // line (y index o, z index kz) of tile t and of this thread
  auto line = [&](int t, int& o, int& kz) {
    o = t / tpo;
    kz = (t - o * tpo) * P + c;
  };
  auto load = [&](C* v, int t) {
    int o, kz;
    line(t, o, kz);
    const long long b = mo + (long long)o * nz + kz;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int row = q < 8 ? 2 * j + 32 * q : 2 * N - 1 - 2 * j - 32 * q;
      v[q] = C{fg[b + (long long)row * stride], er[b + (long long)row * stride]};
    }
  };
  C v[16], nxt[16];
  int t = blockIdx.x;
  if (t < ntiles) load(nxt, t);
  for (; t < ntiles; t += gridDim.x) {
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = nxt[q];
    const int tn = t + gridDim.x;
    if (tn < ntiles) load(nxt, tn);
    C wf_t = wf, mk0_t = mk0, mks_t = mks;
    int j_t = j;
    asm volatile("" : "+v"(j_t), "+v"(wf_t.x), "+v"(wf_t.y));
    asm volatile("" : "+v"(mk0_t.x), "+v"(mk0_t.y), "+v"(mks_t.x), "+v"(mks_t.y));
    dft<A, 16, -1>(v);
#pragma unroll
    for (int q = 0; q < 16; ++q) s[row_slot(j, q)] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = s[col_slot(j, q)];
    stage_twiddle(v, wf_t);
    dft<A, 16, -1>(v);
#pragma unroll
    for (int q = 0; q < 16; ++q) s[col_slot(j, q)] = v[q];
    __syncthreads();
    int o, kz;
    line(t, o, kz);
    const A qy = o * qs1, qz = kz * qs2;
    const A gyz = exp(kG * (qy * qy + qz * qz));
    double* const po = out + mo + (long long)o * nz + kz;
    // Y_k: fgrad^ (y0) and e rho^ (y1) at mode k along x
    C m = mk0_t;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = j_t + 16 * q;
      const C za = v[q], pa = s[partner_slot(j_t, q)];
      const A y0 = m.x * (za.x + pa.x) - m.y * (za.y - pa.y);
      const A y1 = m.x * (za.y + pa.y) + m.y * (za.x - pa.x);
      po[(long long)k * stride] = y0 + y1 * (gx[k] * gyz);
      m = cmul(m, mks_t);
    }
    __syncthreads();  // the next tile's first writes reuse slots other threads read last
  }
}

// One-axis operator passes (OP_AX_GRAD, OP_AX_DIV, OP_AX_GAUSS, OP_AX_GRAD_T,
// OP_AX_DIV_T) of length 256 along y and x on the same layout: forward
// transform, operator, inverse transform per tile with three LDS exchanges.
//  * forward: Makhoul-ordered loads, first butterfly, exchange, second
//    butterfly: the thread holds Z[k], k = j + 16 q;
//  * Z[k] goes back to its own slots and Z[256 - k] comes from the partner's:
//    the thread forms the forward outputs at k and 256 - k, and since the
//    one-mode shifts of the operators compose with the DST index reversal,
//    the operator outputs that Makhoul's inverse pre-twiddle of z_k needs
//    (its inputs at k and 256 - k) take their sources among those two modes:
//    the operator and the pre-twiddle run in registers, for the k of the
//    thread's first inverse butterfly;
//  * inverse: first butterfly, exchange, second butterfly, stores y_{2n} =
//    u_n, y_{2n+1} = u_{255-n} (added to the output for OP_AX_DIV with
//    accumulate).
// Persistent workgroups load the next tile, in the storage type, while
// transforming the current one.
template <typename T, typename A>
constexpr int ax_min_waves() {
  return 2;
}

template <typename T, typename TO, int AXIS, int OP, int P, typename A = double>
__global__ __launch_bounds__(16 * P, (ax_min_waves<T, A>())) void axop_kernel(Args<T, TO> a, Fuse f) {
  using C = cplx<A>;
  constexpr int N = kN;
  // forward DST-II (sine on the load side) and inverse DST-III (sine on the
  // store side)
  constexpr bool FS = OP == OP_AX_DIV || OP == OP_AX_GRAD_T;
  constexpr bool IS = OP == OP_AX_GRAD || OP == OP_AX_DIV_T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* const buf = reinterpret_cast<C*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid % P, j = tid / P;
  C* const s = buf + c * kSeq;
  const long long mo = (long long)blockIdx.y * ((long long)a.nx * a.ny * a.nz);
  const int stride = AXIS == 1 ? a.nz : a.ny * a.nz;
  const C* const tw = reinterpret_cast<const C*>(a.tw);
  const C* const mkt = reinterpret_cast<const C*>(a.mk);
  const C wf = tw[j];
  const C wi = C{wf.x, -wf.y};
  const C mk0 = mkt[j], mks = mkt[16];
  // the operators that may add to their output
  constexpr bool kAcc = OP == OP_AX_DIV || OP == OP_AX_GRAD_T || OP == OP_AX_DIV_T;
  const bool acc = kAcc && f.accumulate != 0;
  // operator weights: c0/(2N) and the wavenumber spacing along this axis
  // (set: the launcher sends passes without it to the general kernel)
  const double cw = f.c0 / (2.0 * N), qs = f.qs[AXIS];

  auto tile = [&](int t) -> long long {
    int o, kc;
    if (AXIS == 1) { kc = t / a.nx; o = t - kc * a.nx; }
    else { o = t / a.tpo; kc = t - o * a.tpo; }
    return mo + (AXIS == 1 ? (long long)o * a.ny * a.nz : (long long)o * a.nz) + kc * 2 * P;
  };
  // Makhoul-ordered loads of the forward transform, in the storage type
  auto load = [&](cplx<T>* v, int t, int jj) {
    const T* p = a.in + tile(t) + 2 * c;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int row = q < 8 ? 2 * jj + 32 * q : 2 * N - 1 - 2 * jj - 32 * q;
      v[q] = *reinterpret_cast<const cplx<T>*>(p + row * stride);
    }
  };

  for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    C wf_t = wf, wi_t = wi, mk0_t = mk0, mks_t = mks;
    int j_t = j;
    // per-tile recomputation of the seeds' powers, the operator weights and
    // the row offsets: hoisted out of the tile loop they would hold a large
    // part of the register file
    asm volatile("" : "+v"(j_t), "+v"(wf_t.x), "+v"(wf_t.y), "+v"(wi_t.x), "+v"(wi_t.y));
    asm volatile("" : "+v"(mk0_t.x), "+v"(mk0_t.y), "+v"(mks_t.x), "+v"(mks_t.y));
    C v[16];
    {
      cplx<T> ld[16];
      load(ld, t, j_t);
      if constexpr (OP == OP_AX_GRAD && sizeof(T) > sizeof(A)) {
        // FP64 input rounded to the FP32 tile after the constant shift
        const T sh = (T)f.ld_shift;
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = C{(A)(ld[q].x - sh), (A)(ld[q].y - sh)};
      } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = C{(A)ld[q].x, (A)ld[q].y};
      }
    }
    // ---- forward: DST-II negates the odd rows (q >= 8)
    if constexpr (FS) {
#pragma unroll
      for (int q = 8; q < 16; ++q) { v[q].x = -v[q].x; v[q].y = -v[q].y; }
    }
    dft<A, 16, -1>(v);
#pragma unroll
    for (int q = 0; q < 16; ++q) s[row_slot(j, q)] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = s[col_slot(j, q)];
    stage_twiddle(v, wf_t);
    dft<A, 16, -1>(v);
    // v[q] = Z[j + 16 q], into the slots this thread read
#pragma unroll
    for (int q = 0; q < 16; ++q) s[col_slot(j, q)] = v[q];
    __syncthreads();
    // ---- forward outputs at k and N - k, operator, inverse pre-twiddle
    C m = mk0_t;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = j_t + 16 * q;
      // partner N - k (k = 0 and k = 128 pair with themselves)
      const C za = v[q], pa = s[partner_slot(j_t, q)];
      // Makhoul outputs (both lines of the pair): Y at k with mk[k], and at
      // N - k with mk[N - k] = (-mk[k].y, -mk[k].x)
      const C mb = k == 0 ? m : C{-m.y, -m.x};
      const C ya = {m.x * (za.x + pa.x) - m.y * (za.y - pa.y), m.x * (za.y + pa.y) + m.y * (za.x - pa.x)};
      const C yb = {mb.x * (pa.x + za.x) - mb.y * (pa.y - za.y), mb.x * (pa.y + za.y) + mb.y * (pa.x - za.x)};
      // operator outputs at inverse-input positions k (rk) and N - k (rn;
      // Y_N = 0 for k = 0).  Position r is mode p = r (N - 1 - r after the
      // DST-III reversal); its source mode p + 1 (GRAD, DIV_T), p - 1 (DIV,
      // GRAD_T) or p (GAUSS) sits at forward Makhoul index N - k for r = k
      // and at k for r = N - k in the shifting operators (the shift composes
      // with one DST reversal), at r itself for GAUSS.  Weights c0/(2N) Vq[.]
      // with Vq[i] = i qs; mode k = 0 of a shifting operator has no source.
      C rk, rn;
      {
        const double dk = (double)k * qs, dn = (double)(N - k) * qs;
        double wk, wn;
        if constexpr (OP == OP_AX_GRAD) { wk = -cw * dn; wn = -cw * dk; }
        else if constexpr (OP == OP_AX_DIV) { wk = cw * dk; wn = cw * dn; }
        else if constexpr (OP == OP_AX_GRAD_T) { wk = -cw * dk; wn = -cw * dn; }
        else if constexpr (OP == OP_AX_DIV_T) { wk = cw * dn; wn = cw * dk; }
        else {
          wk = cw * exp(-1.57 * 1.57 * dk * dk * 2);
          wn = cw * exp(-1.57 * 1.57 * dn * dn * 2);
        }
        constexpr bool kShift = OP != OP_AX_GAUSS;
        const C sk = kShift ? yb : ya, sn = kShift ? ya : yb;
        rk = C{(A)(wk * (double)sk.x), (A)(wk * (double)sk.y)};
        rn = C{(A)(wn * (double)sn.x), (A)(wn * (double)sn.y)};
        if (k == 0) {
          if constexpr (kShift) rk = C{A(0), A(0)};
          rn = C{A(0), A(0)};
        }
      }
      // z_k = conj(mk[k]) (Y0_k + Y1_{N-k} + i (Y1_k - Y0_{N-k}))
      v[q] = cmul(C{m.x, -m.y}, C{rk.x + rn.y, rk.y - rn.x});
      m = cmul(m, mks_t);
      __builtin_amdgcn_sched_barrier(0);
    }
    // ---- inverse: first butterfly, exchange, second butterfly, store
    dft<A, 16, +1>(v);
    __syncthreads();  // every partner read of this tile is done
#pragma unroll
    for (int q = 0; q < 16; ++q) s[row_slot(j, q)] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = s[col_slot(j, q)];
    stage_twiddle(v, wi_t);
    dft<A, 16, +1>(v);
    // u_n goes to row 2n (q < 8) or 2N - 1 - 2n (q >= 8), negated there for
    // DST-III; an accumulating pass reads the old outputs in two batches of
    // eight, each batch's loads issued together
    TO* const o = a.out + tile(t) + 2 * c;
    auto row = [&](int q) { const int n = j_t + 16 * q; return q < 8 ? 2 * n : 2 * N - 1 - 2 * n; };
    if (kAcc && acc) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        cplx<TO> old[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) old[q] = *reinterpret_cast<const cplx<TO>*>(o + row(8 * h + q) * stride);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const C u = v[8 * h + q];
          const double sg = (IS && h == 1) ? -1.0 : 1.0;
          *reinterpret_cast<cplx<TO>*>(o + row(8 * h + q) * stride) =
              cplx<TO>{(TO)((double)old[q].x + sg * (double)u.x), (TO)((double)old[q].y + sg * (double)u.y)};
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const C u = (IS && q >= 8) ? C{-v[q].x, -v[q].y} : v[q];
        *reinterpret_cast<cplx<TO>*>(o + row(q) * stride) = cplx<TO>{(TO)u.x, (TO)u.y};
      }
    }
    __syncthreads();  // the next tile's first writes reuse slots other threads read last
  }
}

// The x pass of psi's inverse transform (DCT-III along x of the x-spectral
// lines Y) that also forms d_x psi from the same lines: d_x psi is the
// gradient one-axis operator applied to psi, whose forward DCT-II along x is
// 2N Y, so it is DST-III(-Vs[n] Y[n+1]) with no other factor.  The
// workgroup's first 16 P threads form psi and the other 16 P threads d_x psi
// (whole wavefronts each), with an LDS tile each: the first group loads the
// tile into its LDS sequences, both take their Makhoul pre-twiddle from them
// (d_x psi's as in axop_kernel: input position k takes -(N-k) qs Y[N-k],
// position N-k takes -k qs Y[k], none at k = 0), and the two inverse
// transforms run side by side, psi stored to out, d_x psi to out2.
template <typename T, typename TO, int P, typename A = double>
__global__ __launch_bounds__(32 * P) void invgrad_kernel(Args<T, TO> a, TO* __restrict__ out2, double qs) {
  using C = cplx<A>;
  constexpr int N = kN;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* const buf = reinterpret_cast<C*>(smem_raw);
  const int tid = threadIdx.x;
  const bool dx = tid >= 16 * P;  // wavefront-uniform
  const int lt = dx ? tid - 16 * P : tid;
  const int c = lt % P, j = lt / P;
  C* const sa = buf + c * kSeq;        // the loaded tile, then psi's exchange
  C* const sx = dx ? buf + (P + c) * kSeq : sa;  // this group's exchange
  const long long mo = (long long)blockIdx.y * ((long long)a.nx * a.ny * a.nz);
  const int stride = a.ny * a.nz;
  const C* const tw = reinterpret_cast<const C*>(a.tw);
  const C* const mkt = reinterpret_cast<const C*>(a.mk);
  const C wf = tw[j];
  const C wi = C{wf.x, -wf.y};
  const C mk0 = mkt[j], mks = mkt[16];
  auto tile = [&](int t) -> long long {
    const int o = t / a.tpo, kc = t - o * a.tpo;
    return mo + (long long)o * a.nz + kc * 2 * P;
  };
  for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    C wi_t = wi, mk0_t = mk0, mks_t = mks;
    int j_t = j;
    asm volatile("" : "+v"(j_t), "+v"(wi_t.x), "+v"(wi_t.y));
    asm volatile("" : "+v"(mk0_t.x), "+v"(mk0_t.y), "+v"(mks_t.x), "+v"(mks_t.y));
    C v[16];
    if (!dx) {
      const T* p = a.in + tile(t) + 2 * c;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const cplx<T> u = *reinterpret_cast<const cplx<T>*>(p + (j + 16 * q) * stride);
        sa[col_slot(j, q)] = C{(A)u.x, (A)u.y};
      }
    }
    __syncthreads();
    // z_k = conj(mk[k]) (R0_k + R1_{N-k} + i (R1_k - R0_{N-k})) of R = Y (psi)
    // or R = the gradient's inputs (d_x psi)
    C m = mk0_t;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = j_t + 16 * q;
      const C rk = sa[col_slot(j_t, q)];
      const C rn = k == 0 ? C{A(0), A(0)} : sa[partner_slot(j_t, q)];
      C r0 = rk, r1 = rn;
      if (dx) {
        const double dk = (double)k * qs, dn = (double)(N - k) * qs;
        r0 = k == 0 ? C{A(0), A(0)} : C{(A)(-dn * (double)rn.x), (A)(-dn * (double)rn.y)};
        r1 = k == 0 ? C{A(0), A(0)} : C{(A)(-dk * (double)rk.x), (A)(-dk * (double)rk.y)};
      }
      v[q] = cmul(C{m.x, -m.y}, C{r0.x + r1.y, r0.y - r1.x});
      m = cmul(m, mks_t);
    }
    // first butterfly, exchange, second butterfly; y_{2n} = u_n, y_{2n+1} =
    // u_{N-1-n} (negated for d_x psi's DST-III)
    dft<A, 16, +1>(v);
    __syncthreads();  // every pre-twiddle read of the loaded tile is done
#pragma unroll
    for (int q = 0; q < 16; ++q) sx[row_slot(j, q)] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = sx[col_slot(j, q)];
    stage_twiddle(v, wi_t);
    dft<A, 16, +1>(v);
    TO* const o = (dx ? out2 : a.out) + tile(t) + 2 * c;
    const A sg = dx ? A(-1) : A(1);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int n = j_t + 16 * q;
      const int row = q < 8 ? 2 * n : 2 * N - 1 - 2 * n;
      const C u = q >= 8 ? C{sg * v[q].x, sg * v[q].y} : v[q];
      *reinterpret_cast<cplx<TO>*>(o + row * stride) = cplx<TO>{(TO)u.x, (TO)u.y};
    }
    __syncthreads();  // the next tile's loads reuse slots other threads read last
  }
}

// Sequences per tile: 16 columns of FP64 or 32 of FP32 (128-byte row
// segments).
template <typename T>
constexpr int tile_p() { return sizeof(T) == 8 ? 8 : 16; }

}  // namespace yx256

}  // namespace smaiso
