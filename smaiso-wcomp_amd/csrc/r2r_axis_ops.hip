// One-axis operator passes (OP_AX_*, see r2r.h) for FFT lengths, with the
// default tiles; a translation unit of its own so that these instantiations
// compile in parallel with the rest of the library.
#include "r2r_kernels.h"

#include <type_traits>

namespace smaiso {

namespace {

template <typename T, int N, typename A>
hipError_t launch_ax(const PassIO& io, void* tw, void* mk, int axis, const Geom& g, hipStream_t s) {
  using C = TileChoice<T, N>;
  const int kind = REDFT10;  // unused by the one-axis operators
  if (axis == 2) {
    switch (io.f.op) {
      case OP_AX_GRAD: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_GRAD, T, A, 2>(io, tw, mk, kind, axis, g, s);
      case OP_AX_DIV: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_DIV, T, A, 2>(io, tw, mk, kind, axis, g, s);
      case OP_AX_GAUSS: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_GAUSS, T, A, 2>(io, tw, mk, kind, axis, g, s);
      case OP_AX_GAUSS_AMP: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_GAUSS_AMP, T, A, 2>(io, tw, mk, kind, axis, g, s);
      case OP_AX_GRAD_T: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_GRAD_T, T, A, 2>(io, tw, mk, kind, axis, g, s);
      case OP_AX_DIV_T: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_DIV_T, T, A, 2>(io, tw, mk, kind, axis, g, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (io.f.op) {
    case OP_AX_GRAD: return launch_fft_pn<T, N, C::PS, C::NTS, OP_AX_GRAD, T, A, 3>(io, tw, mk, kind, axis, g, s);
    case OP_AX_DIV: return launch_fft_pn<T, N, C::PS, C::NTS, OP_AX_DIV, T, A, 3>(io, tw, mk, kind, axis, g, s);
    case OP_AX_GAUSS: return launch_fft_pn<T, N, C::PS, C::NTS, OP_AX_GAUSS, T, A, 3>(io, tw, mk, kind, axis, g, s);
    case OP_AX_GRAD_T: return launch_fft_pn<T, N, C::PS, C::NTS, OP_AX_GRAD_T, T, A, 3>(io, tw, mk, kind, axis, g, s);
    case OP_AX_DIV_T: return launch_fft_pn<T, N, C::PS, C::NTS, OP_AX_DIV_T, T, A, 3>(io, tw, mk, kind, axis, g, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T, typename A>
hipError_t launch_axis_any(int n, const PassIO& io, void* tw, void* mk, int axis, const Geom& g,
                           hipStream_t s) {
  if constexpr (std::is_same<A, float>::value) {
    switch (n) {
      case 64: return launch_ax<T, 64, A>(io, tw, mk, axis, g, s);
      case 96: return launch_ax<T, 96, A>(io, tw, mk, axis, g, s);
      case 128: return launch_ax<T, 128, A>(io, tw, mk, axis, g, s);
      case 256: return launch_ax<T, 256, A>(io, tw, mk, axis, g, s);
      case 512: return launch_ax<T, 512, A>(io, tw, mk, axis, g, s);
      default: return hipErrorInvalidValue;
    }
  } else {
  switch (n) {
    case 16: return launch_ax<T, 16, A>(io, tw, mk, axis, g, s);
    case 32: return launch_ax<T, 32, A>(io, tw, mk, axis, g, s);
    case 48: return launch_ax<T, 48, A>(io, tw, mk, axis, g, s);
    case 64: return launch_ax<T, 64, A>(io, tw, mk, axis, g, s);
    case 96: return launch_ax<T, 96, A>(io, tw, mk, axis, g, s);
    case 128: return launch_ax<T, 128, A>(io, tw, mk, axis, g, s);
    case 192: return launch_ax<T, 192, A>(io, tw, mk, axis, g, s);
    case 256: return launch_ax<T, 256, A>(io, tw, mk, axis, g, s);
    case 384: return launch_ax<T, 384, A>(io, tw, mk, axis, g, s);
    case 512: return launch_ax<T, 512, A>(io, tw, mk, axis, g, s);
    case 1024: return launch_ax<T, 1024, A>(io, tw, mk, axis, g, s);
    default: return hipErrorInvalidValue;
  }
  }
}

}  // namespace

template <>
hipError_t launch_axis_op<double, double>(int n, const PassIO& io, void* tw, void* mk, int axis, const Geom& g,
                                          hipStream_t s) {
  return launch_axis_any<double, double>(n, io, tw, mk, axis, g, s);
}
template <>
hipError_t launch_axis_op<float, double>(int n, const PassIO& io, void* tw, void* mk, int axis, const Geom& g,
                                         hipStream_t s) {
  return launch_axis_any<float, double>(n, io, tw, mk, axis, g, s);
}
template <>
hipError_t launch_axis_op<float, float>(int n, const PassIO& io, void* tw, void* mk, int axis, const Geom& g,
                                        hipStream_t s) {
  return launch_axis_any<float, float>(n, io, tw, mk, axis, g, s);
}

}  // namespace smaiso
