// Axis passes whose output storage type differs from the input's: FP32
// sources into an FP64 intermediate (z pass, plain or with the psi update on
// load) and an FP64 intermediate into FP32 output (y or x pass, plain).  A 3-D
// transform between FP32 fields then keeps its partly transformed values in
// FP64.  Also the one-axis gradient of an FP64 field stored as FP32 (any
// axis: grad rho of FP32 storage).  A translation unit of its own so that
// these instantiations compile in parallel with the rest of the library.
#include "r2r_kernels.h"

namespace smaiso {

namespace {

template <typename T, typename TO, int N>
hipError_t launch_cast_n(const PassIO& io, void* tw, void* mk, int kind, int axis, const Geom& g,
                         hipStream_t s) {
  using C = TileChoice<T, N>;
  constexpr bool kNarrowing = sizeof(T) == 8 && sizeof(TO) == 4;
  if (axis == 2) {
    if (io.f.op == OP_LD_PSI) return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_LD_PSI, TO, double, 2>(io, tw, mk, kind, axis, g, s);
    if (io.f.op == OP_PLAIN) return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_PLAIN, TO, double, 2>(io, tw, mk, kind, axis, g, s);
    if constexpr (kNarrowing) {
      if (io.f.op == OP_AX_GRAD) return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_GRAD, TO, double, 2>(io, tw, mk, kind, axis, g, s);
    }
    return hipErrorInvalidValue;
  }
  if (io.f.op == OP_PLAIN) return launch_fft_pn<T, N, C::PS, C::NTS, OP_PLAIN, TO, double, 3>(io, tw, mk, kind, axis, g, s);
  if constexpr (kNarrowing) {
    if (io.f.op == OP_AX_GRAD) return launch_fft_pn<T, N, C::PS, C::NTS, OP_AX_GRAD, TO, double, 3>(io, tw, mk, kind, axis, g, s);
  }
  return hipErrorInvalidValue;
}

template <typename T, typename TO>
hipError_t launch_cast_direct(const PassIO& io, void* costab, int kind, int axis, int n, const Geom& g,
                              hipStream_t s) {
  DirectArgs<T, TO> d;
  d.p = make_args<T, TO>(io, nullptr, nullptr, kind, axis, g);
  d.costab = static_cast<const double*>(costab);
  d.n = n;
  // the one-axis operators keep the forward outputs beside the tile
  const int mul = is_axis_op(io.f.op) ? 2 : 1;
  int tl = 16;  // lines per tile; the LDS tile is FP64
  while (tl > 1 && (size_t)tl * n * sizeof(double) * mul > 48 * 1024) tl /= 2;
  d.tl = tl;
  long long blocks;
  if (axis == 2) blocks = ((long long)g.nx * g.ny + tl - 1) / tl;
  else blocks = (long long)(axis == 1 ? g.nx : g.ny) * ((g.nz + tl - 1) / tl);
  const size_t lds = (size_t)tl * n * sizeof(double) * mul;
  hipLaunchKernelGGL((direct_pass_kernel<T, TO>), dim3((unsigned)blocks, (unsigned)g.nb), dim3(kThreads), lds, s, d);
  return hipGetLastError();
}

template <typename T, typename TO>
hipError_t dispatch_cast(int n, const PassIO& io, void* tw, void* mk, void* ct, int kind, int axis,
                         const Geom& g, hipStream_t s) {
  if (n == 256 && axis != 2) {
    const hipError_t e = launch_yx256(sizeof(T) == 8 ? F64 : F32, sizeof(TO) == 8 ? F64 : F32, kind, axis, io, g,
                                      tw, mk, s);
    if (e != hipErrorNotSupported) return e;
  }
  switch (n) {
    case 16: return launch_cast_n<T, TO, 16>(io, tw, mk, kind, axis, g, s);
    case 32: return launch_cast_n<T, TO, 32>(io, tw, mk, kind, axis, g, s);
    case 48: return launch_cast_n<T, TO, 48>(io, tw, mk, kind, axis, g, s);
    case 64: return launch_cast_n<T, TO, 64>(io, tw, mk, kind, axis, g, s);
    case 96: return launch_cast_n<T, TO, 96>(io, tw, mk, kind, axis, g, s);
    case 128: return launch_cast_n<T, TO, 128>(io, tw, mk, kind, axis, g, s);
    case 192: return launch_cast_n<T, TO, 192>(io, tw, mk, kind, axis, g, s);
    case 256: return launch_cast_n<T, TO, 256>(io, tw, mk, kind, axis, g, s);
    case 384: return launch_cast_n<T, TO, 384>(io, tw, mk, kind, axis, g, s);
    case 512: return launch_cast_n<T, TO, 512>(io, tw, mk, kind, axis, g, s);
    case 1024: return launch_cast_n<T, TO, 1024>(io, tw, mk, kind, axis, g, s);
    default: return launch_cast_direct<T, TO>(io, ct, kind, axis, n, g, s);
  }
}

}  // namespace

hipError_t R2R::pass_cast(DType in_dt, DType out_dt, int kind, int axis, const PassIO& io, const Geom& g,
                          hipStream_t s) {
  if (in_dt == out_dt) return pass(in_dt, kind, axis, io, g, s);
  const int op = io.f.op;
  if (!(op == OP_PLAIN || (op == OP_LD_PSI && axis == 2) || op == OP_AX_GRAD)) return hipErrorInvalidValue;
  if (axis < 0 || axis > 2) return hipErrorInvalidValue;
  if (op == OP_AX_GRAD) {
    // one pass: FP64 -> FP32 along any axis
    if (in_dt != F64) return hipErrorNotSupported;
  } else if ((in_dt == F32) != (axis == 2)) {
    // FP32 -> FP64 in the first (z) pass, FP64 -> FP32 in a later pass
    return hipErrorNotSupported;
  }
  const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  Tables* tb = nullptr;
  hipError_t e;
  if (op == OP_AX_GRAD && io.f.arith32 && n == 256 && axis != 2) {
    // FP32 arithmetic over the FP64 input (256-point y/x passes only; other
    // lengths compute in FP64)
    if ((e = tables(F32, n, &tb)) != hipSuccess) return e;
    e = launch_yx256(F64, F32, kind, axis, io, g, tb->tw, tb->mk, s, true);
    if (e != hipErrorNotSupported) return e;
  }
  e = tables(F64, n, &tb);
  if (e != hipSuccess) return e;
  if (in_dt == F32) return dispatch_cast<float, double>(n, io, tb->tw, tb->mk, tb->costab, kind, axis, g, s);
  return dispatch_cast<double, float>(n, io, tb->tw, tb->mk, tb->costab, kind, axis, g, s);
}

}  // namespace smaiso
