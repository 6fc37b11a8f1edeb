// One-axis operator passes (OP_AX_*, see r2r.h) for FFT lengths, with the
// default tiles; a translation unit of its own so that these instantiations
// compile in parallel with the rest of the library.
#include "r2r_kernels.h"

namespace smaiso {

namespace {

template <typename T, int N>
hipError_t launch_ax(const PassIO& io, void* tw, void* mk, int axis, const Geom& g, hipStream_t s) {
  using C = TileChoice<T, N>;
  const int kind = REDFT10;  // unused by the one-axis operators
  if (axis == 2) {
    switch (io.f.op) {
      case OP_AX_GRAD: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_GRAD>(io, tw, mk, kind, axis, g, s);
      case OP_AX_DIV: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_DIV>(io, tw, mk, kind, axis, g, s);
      case OP_AX_GAUSS: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_GAUSS>(io, tw, mk, kind, axis, g, s);
      case OP_AX_GAUSS_AMP: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_GAUSS_AMP>(io, tw, mk, kind, axis, g, s);
      case OP_AX_GRAD_T: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_GRAD_T>(io, tw, mk, kind, axis, g, s);
      case OP_AX_DIV_T: return launch_fft_pn<T, N, C::PZ, C::NTZ, OP_AX_DIV_T>(io, tw, mk, kind, axis, g, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (io.f.op) {
    case OP_AX_GRAD: return launch_fft_pn<T, N, C::PS, C::NTS, OP_AX_GRAD>(io, tw, mk, kind, axis, g, s);
    case OP_AX_DIV: return launch_fft_pn<T, N, C::PS, C::NTS, OP_AX_DIV>(io, tw, mk, kind, axis, g, s);
    case OP_AX_GAUSS: return launch_fft_pn<T, N, C::PS, C::NTS, OP_AX_GAUSS>(io, tw, mk, kind, axis, g, s);
    case OP_AX_GRAD_T: return launch_fft_pn<T, N, C::PS, C::NTS, OP_AX_GRAD_T>(io, tw, mk, kind, axis, g, s);
    case OP_AX_DIV_T: return launch_fft_pn<T, N, C::PS, C::NTS, OP_AX_DIV_T>(io, tw, mk, kind, axis, g, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_axis_any(int n, const PassIO& io, void* tw, void* mk, int axis, const Geom& g,
                           hipStream_t s) {
  switch (n) {
    case 16: return launch_ax<T, 16>(io, tw, mk, axis, g, s);
    case 32: return launch_ax<T, 32>(io, tw, mk, axis, g, s);
    case 48: return launch_ax<T, 48>(io, tw, mk, axis, g, s);
    case 64: return launch_ax<T, 64>(io, tw, mk, axis, g, s);
    case 96: return launch_ax<T, 96>(io, tw, mk, axis, g, s);
    case 128: return launch_ax<T, 128>(io, tw, mk, axis, g, s);
    case 192: return launch_ax<T, 192>(io, tw, mk, axis, g, s);
    case 256: return launch_ax<T, 256>(io, tw, mk, axis, g, s);
    case 384: return launch_ax<T, 384>(io, tw, mk, axis, g, s);
    case 512: return launch_ax<T, 512>(io, tw, mk, axis, g, s);
    case 1024: return launch_ax<T, 1024>(io, tw, mk, axis, g, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

template <>
hipError_t launch_axis_op<double>(int n, const PassIO& io, void* tw, void* mk, int axis, const Geom& g,
                                  hipStream_t s) {
  return launch_axis_any<double>(n, io, tw, mk, axis, g, s);
}
template <>
hipError_t launch_axis_op<float>(int n, const PassIO& io, void* tw, void* mk, int axis, const Geom& g,
                                 hipStream_t s) {
  return launch_axis_any<float>(n, io, tw, mk, axis, g, s);
}

}  // namespace smaiso
