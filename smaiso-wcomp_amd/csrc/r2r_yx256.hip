// Launchers of the 256-point y/x passes (r2r_yx256.h); a translation unit of
// its own so that these instantiations compile in parallel with the rest.
#include "r2r_yx256.h"

#include <algorithm>
#include <cstdint>
#include <type_traits>

namespace smaiso {

namespace {

template <typename T, typename TO, int AXIS, int MODE, typename A>
hipError_t launch(const PassIO& io, const Geom& g, int kind, const void* tw, const void* mk, hipStream_t s) {
  constexpr int P = yx256::tile_p<T>();
  yx256::Args<T, TO> a;
  a.in = static_cast<const T*>(io.in);
  a.out = static_cast<TO*>(io.out);
  a.tw = tw;
  a.mk = mk;
  a.nx = g.nx;
  a.ny = g.ny;
  a.nz = g.nz;
  a.kind = kind;
  a.tpo = g.nz / (2 * P);
  a.ntiles = (AXIS == 1 ? g.nx : g.ny) * a.tpo;
  // two resident workgroups per CU for FP64 tiles (128 threads), one for
  // FP32 tiles (256 threads); twice that many keep every CU's queue full
  const int grid = std::min(a.ntiles, 256 * (P == 8 ? 2 : 1) * 2);
  const size_t lds = (size_t)P * yx256::kSeq * sizeof(cplx<A>);
  hipLaunchKernelGGL((yx256::pass_kernel<T, TO, AXIS, MODE, P, A>), dim3((unsigned)grid, (unsigned)g.nb),
                     dim3(16 * P), lds, s, a);
  return hipGetLastError();
}

template <typename T, typename TO, int AXIS, int OP, typename A>
hipError_t launch_op(const PassIO& io, const Geom& g, const void* tw, const void* mk, hipStream_t s) {
  constexpr int P = yx256::tile_p<T>();
  yx256::Args<T, TO> a;
  a.in = static_cast<const T*>(io.in);
  a.out = static_cast<TO*>(io.out);
  a.tw = tw;
  a.mk = mk;
  a.nx = g.nx;
  a.ny = g.ny;
  a.nz = g.nz;
  a.kind = REDFT10;  // unused by the operators
  a.tpo = g.nz / (2 * P);
  a.ntiles = (AXIS == 1 ? g.nx : g.ny) * a.tpo;
  const size_t lds = (size_t)P * yx256::kSeq * sizeof(cplx<A>);
  // resident workgroups per CU (LDS bound), twice that many in the grid
  const int per_cu = (int)std::max<size_t>(1, (160 * 1024) / lds);
  const int grid = std::min(a.ntiles, 256 * per_cu * 2);
  hipLaunchKernelGGL((yx256::axop_kernel<T, TO, AXIS, OP, P, A>), dim3((unsigned)grid, (unsigned)g.nb),
                     dim3(16 * P), lds, s, a, io.f);
  return hipGetLastError();
}

// plain passes and the one-axis operators (the x pass's mid operators stay
// on the general pass kernel); FP32 arithmetic and the narrowing (FP64 ->
// FP32) pass have the operators the step runs in them
template <typename T, typename TO, int AXIS, typename A>
hipError_t launch_axis(int kind, const PassIO& io, const Geom& g, const void* tw, const void* mk, hipStream_t s) {
  const int op = io.f.op;
  if constexpr (sizeof(T) > sizeof(A)) {
    // FP64 input in FP32 arithmetic: the gradient only
    if (op != OP_AX_GRAD || io.f.axis != AXIS || io.f.qs[AXIS] == 0.0) return hipErrorNotSupported;
    return launch_op<T, TO, AXIS, OP_AX_GRAD, A>(io, g, tw, mk, s);
  }
  if (op == OP_PLAIN) {
    if (kind == REDFT10 || kind == RODFT10) return launch<T, TO, AXIS, yx256::kFwd, A>(io, g, kind, tw, mk, s);
    return launch<T, TO, AXIS, yx256::kInv, A>(io, g, kind, tw, mk, s);
  }
  if (io.f.axis != AXIS || io.f.qs[AXIS] == 0.0) return hipErrorNotSupported;
  constexpr bool kSame = std::is_same<T, TO>::value;
  constexpr bool kFull = kSame && std::is_same<T, double>::value;
  if (op == OP_AX_GRAD) return launch_op<T, TO, AXIS, OP_AX_GRAD, A>(io, g, tw, mk, s);
  if constexpr (kSame) {
    if (op == OP_AX_DIV) return launch_op<T, TO, AXIS, OP_AX_DIV, A>(io, g, tw, mk, s);
    if (op == OP_AX_GAUSS) return launch_op<T, TO, AXIS, OP_AX_GAUSS, A>(io, g, tw, mk, s);
  }
  if constexpr (kFull) {
    if (op == OP_AX_GRAD_T) return launch_op<T, TO, AXIS, OP_AX_GRAD_T, A>(io, g, tw, mk, s);
    if (op == OP_AX_DIV_T) return launch_op<T, TO, AXIS, OP_AX_DIV_T, A>(io, g, tw, mk, s);
  }
  return hipErrorNotSupported;
}

template <typename T, typename TO, typename A = double>
hipError_t launch_typed(int kind, int axis, const PassIO& io, const Geom& g, const void* tw, const void* mk,
                        hipStream_t s) {
  constexpr int P = yx256::tile_p<T>();
  const int n = axis == 1 ? g.ny : g.nx;
  const long long stride = axis == 1 ? (long long)g.nz : (long long)g.ny * g.nz;
  const bool aligned = ((uintptr_t)io.in % (2 * sizeof(T))) == 0 && ((uintptr_t)io.out % (2 * sizeof(TO))) == 0 &&
                       (g.n3() % 2) == 0;
  if (n != yx256::kN || g.nz % (2 * P) != 0 || !aligned || io.in2 != nullptr ||
      (yx256::kN - 1) * stride + g.nz > 2147483647LL)
    return hipErrorNotSupported;
  return axis == 1 ? launch_axis<T, TO, 1, A>(kind, io, g, tw, mk, s) : launch_axis<T, TO, 0, A>(kind, io, g, tw, mk, s);
}

template <typename TO>
hipError_t launch_invgrad_t(const double* in, TO* out, TO* out2, double qs, const Geom& g, const void* tw,
                            const void* mk, hipStream_t s) {
  constexpr int P = yx256::tile_p<double>();
  yx256::Args<double, TO> a;
  a.in = in;
  a.out = out;
  a.tw = tw;
  a.mk = mk;
  a.nx = g.nx;
  a.ny = g.ny;
  a.nz = g.nz;
  a.kind = REDFT01;
  a.tpo = g.nz / (2 * P);
  a.ntiles = g.ny * a.tpo;
  // two LDS tiles (the loaded lines and psi's exchange, d_x psi's exchange)
  const size_t lds = (size_t)2 * P * yx256::kSeq * sizeof(cplx<double>);
  const int per_cu = (int)std::max<size_t>(1, (160 * 1024) / lds);
  const int grid = std::min(a.ntiles, 256 * per_cu * 2);
  hipLaunchKernelGGL((yx256::invgrad_kernel<double, TO, P, double>), dim3((unsigned)grid, (unsigned)g.nb),
                     dim3(32 * P), lds, s, a, out2, qs);
  return hipGetLastError();
}

}  // namespace

bool yx256_fgx_supported(const Geom& g) {
  return g.nx == yx256::kN && g.nz % 16 == 0 && (long long)(yx256::kN - 1) * g.ny * g.nz + g.nz <= 2147483647LL;
}

hipError_t launch_yx256_fgx(const double* fg, const double* er, double* out, const double qs[3], const Geom& g,
                            const void* tw, const void* mk, hipStream_t s) {
  constexpr int P = 16;
  if (!yx256_fgx_supported(g)) return hipErrorNotSupported;
  const int tpo = g.nz / P;
  const int ntiles = g.ny * tpo;
  const size_t lds = (size_t)P * yx256::kSeq * sizeof(cplx<double>) + yx256::kN * sizeof(double);
  const int per_cu = (int)std::max<size_t>(1, (160 * 1024) / lds);
  const int grid = std::min(ntiles, 256 * per_cu * 2);
  hipLaunchKernelGGL((yx256::fgx_kernel<P>), dim3((unsigned)grid, (unsigned)g.nb), dim3(16 * P), lds, s, fg, er,
                     out, tw, mk, g.ny, g.nz, ntiles, tpo, qs[0], qs[1], qs[2]);
  return hipGetLastError();
}

bool yx256_invgrad_supported(const Geom& g) {
  constexpr int P = yx256::tile_p<double>();
  return g.nx == yx256::kN && g.nz % (2 * P) == 0 && (g.n3() % 2) == 0 &&
         (long long)(yx256::kN - 1) * g.ny * g.nz + g.nz <= 2147483647LL;
}

hipError_t launch_yx256_invgrad(DType out_dt, const double* in, void* out, void* out2, double qs, const Geom& g,
                                const void* tw, const void* mk, hipStream_t s) {
  if (!yx256_invgrad_supported(g) || qs == 0.0) return hipErrorNotSupported;
  if (out_dt == F64)
    return launch_invgrad_t<double>(in, static_cast<double*>(out), static_cast<double*>(out2), qs, g, tw, mk, s);
  return launch_invgrad_t<float>(in, static_cast<float*>(out), static_cast<float*>(out2), qs, g, tw, mk, s);
}

hipError_t launch_yx256(DType in_dt, DType out_dt, int kind, int axis, const PassIO& io, const Geom& g,
                        const void* tw, const void* mk, hipStream_t s, bool arith32) {
  if (axis != 0 && axis != 1) return hipErrorNotSupported;
  if (arith32) {
    if (in_dt == F64 && out_dt == F32) return launch_typed<double, float, float>(kind, axis, io, g, tw, mk, s);
    if (in_dt != F32 || out_dt != F32) return hipErrorNotSupported;
    return launch_typed<float, float, float>(kind, axis, io, g, tw, mk, s);
  }
  if (in_dt == F64 && out_dt == F64) return launch_typed<double, double>(kind, axis, io, g, tw, mk, s);
  if (in_dt == F32 && out_dt == F32) return launch_typed<float, float>(kind, axis, io, g, tw, mk, s);
  if (in_dt == F64 && out_dt == F32) return launch_typed<double, float>(kind, axis, io, g, tw, mk, s);
  return hipErrorNotSupported;
}

}  // namespace smaiso
