// Tile variants of the plain r2r pass for T = double (see TileVariant in
// r2r_kernels.h); a translation unit of its own so that the variants compile
// in parallel with the rest of the library.
#include "r2r_kernels.h"

namespace smaiso {

namespace {

template <typename T, int N, bool Z, int V>
hipError_t launch_tv(const PassIO& io, void* tw, void* mk, int kind, int axis, const Geom& g,
                     hipStream_t s) {
  using TV = TileVariant<T, Z, V>;
  if constexpr (tile_fits<T, N, TV::P, TV::NT>())
    return launch_fft_pn<T, N, TV::P, TV::NT, OP_PLAIN, T, double, Z ? 2 : 3>(io, tw, mk, kind, axis, g, s);
  else
    return hipErrorNotSupported;
}

template <typename T, int N>
hipError_t launch_n(int v, const PassIO& io, void* tw, void* mk, int kind, int axis, const Geom& g,
                    hipStream_t s) {
  const bool z = axis == 2;
  switch (v) {
    case 1: return z ? launch_tv<T, N, true, 1>(io, tw, mk, kind, axis, g, s)
                     : launch_tv<T, N, false, 1>(io, tw, mk, kind, axis, g, s);
    case 2: return z ? launch_tv<T, N, true, 2>(io, tw, mk, kind, axis, g, s)
                     : launch_tv<T, N, false, 2>(io, tw, mk, kind, axis, g, s);
    case 3: return z ? launch_tv<T, N, true, 3>(io, tw, mk, kind, axis, g, s)
                     : launch_tv<T, N, false, 3>(io, tw, mk, kind, axis, g, s);
    default: return hipErrorNotSupported;
  }
}

}  // namespace

template <>
hipError_t launch_plain_variant<double>(int n, int v, const PassIO& io, void* tw, void* mk, int kind,
                                     int axis, const Geom& g, hipStream_t s) {
  using T = double;
  switch (n) {
    case 16: return launch_n<T, 16>(v, io, tw, mk, kind, axis, g, s);
    case 32: return launch_n<T, 32>(v, io, tw, mk, kind, axis, g, s);
    case 48: return launch_n<T, 48>(v, io, tw, mk, kind, axis, g, s);
    case 64: return launch_n<T, 64>(v, io, tw, mk, kind, axis, g, s);
    case 96: return launch_n<T, 96>(v, io, tw, mk, kind, axis, g, s);
    case 128: return launch_n<T, 128>(v, io, tw, mk, kind, axis, g, s);
    case 192: return launch_n<T, 192>(v, io, tw, mk, kind, axis, g, s);
    case 256: return launch_n<T, 256>(v, io, tw, mk, kind, axis, g, s);
    case 384: return launch_n<T, 384>(v, io, tw, mk, kind, axis, g, s);
    case 512: return launch_n<T, 512>(v, io, tw, mk, kind, axis, g, s);
    case 1024: return launch_n<T, 1024>(v, io, tw, mk, kind, axis, g, s);
    default: return hipErrorNotSupported;
  }
}

}  // namespace smaiso
