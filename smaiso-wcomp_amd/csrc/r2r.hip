// LDS-staged batched r2r transforms for gfx950: host side (tables, tile
// choice, dispatch).  The method and the kernels are in r2r_kernels.h.
#include "r2r_kernels.h"

#include <cstdio>
#include <type_traits>

namespace smaiso {

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool R2R::fft_supported(int n) { return fft_radix(n, 16, 0) != 0; }

void r2r_table_values(int n, double* tw, double* mk, double* ct) {
  for (int m = 0; m < n; ++m) {
    const double a = -2.0 * M_PI * (double)m / (double)n;
    tw[2 * m] = std::cos(a);
    tw[2 * m + 1] = std::sin(a);
    const double b = -M_PI * (double)m / (2.0 * n);
    mk[2 * m] = std::cos(b);
    mk[2 * m + 1] = std::sin(b);
  }
  for (int t = 0; t < 4 * n; ++t) ct[t] = std::cos(M_PI * (double)t / (2.0 * n));
}

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
  r2r_table_values(n, tw.data(), mk.data(), ct.data());
  Tables tb;
  hipError_t e;
  auto upload = [&](const std::vector<double>& v, void** d) -> hipError_t {
    hipError_t r = hipMalloc(d, v.size() * es);
    if (r != hipSuccess) return r;
    if (dt == F64) return hipMemcpy(*d, v.data(), v.size() * es, hipMemcpyHostToDevice);
    std::vector<float> f(v.begin(), v.end());
    return hipMemcpy(*d, f.data(), f.size() * es, hipMemcpyHostToDevice);
  };
  if ((e = upload(tw, &tb.tw)) != hipSuccess || (e = upload(mk, &tb.mk)) != hipSuccess ||
      (e = upload(ct, &tb.costab)) != hipSuccess) {
    for (void* d : {tb.tw, tb.mk, tb.costab})
      if (d) (void)hipFree(d);
    return e;
  }
  auto res = cache_.emplace(key, tb);
  *out = &res.first->second;
  return hipSuccess;
}

template <typename T, int N, typename A = double>
static hipError_t launch_fft(const PassIO& io, void* tw, void* mk, int kind, int axis,
                             const Geom& g, hipStream_t s) {
  using C = TileChoice<T, N>;
  constexpr int PZ = C::PZ, NTZ = C::NTZ, PS = C::PS, NTS = C::NTS;
  if (axis == 2) {
    switch (io.f.op) {
      case OP_LD_SH: return launch_fft_pn<T, N, PZ, NTZ, OP_LD_SH, T, A, 2>(io, tw, mk, kind, axis, g, s);
      case OP_LD_RELAXLAP: return launch_fft_pn<T, N, PZ, NTZ, OP_LD_RELAXLAP, T, A, 2>(io, tw, mk, kind, axis, g, s);
      case OP_LD_PROD: return launch_fft_pn<T, N, PZ, NTZ, OP_LD_PROD, T, A, 2>(io, tw, mk, kind, axis, g, s);
      case OP_PLAIN: return launch_fft_pn<T, N, PZ, NTZ, OP_PLAIN, T, A, 2>(io, tw, mk, kind, axis, g, s);
      default: break;
    }
    // load ops the step runs in FP64 arithmetic only
    if constexpr (std::is_same<A, double>::value) {
      switch (io.f.op) {
        case OP_LD_GRAD_D: return launch_fft_pn<T, N, PZ, NTZ, OP_LD_GRAD_D, T, double, 2>(io, tw, mk, kind, axis, g, s);
        case OP_LD_GRAD_T: return launch_fft_pn<T, N, PZ, NTZ, OP_LD_GRAD_T, T, double, 2>(io, tw, mk, kind, axis, g, s);
        case OP_LD_AMP: return launch_fft_pn<T, N, PZ, NTZ, OP_LD_AMP, T, double, 2>(io, tw, mk, kind, axis, g, s);
        case OP_LD_PSI: return launch_fft_pn<T, N, PZ, NTZ, OP_LD_PSI, T, double, 2>(io, tw, mk, kind, axis, g, s);
        default: break;
      }
    }
    return hipErrorInvalidValue;
  }
  switch (io.f.op) {
    case OP_MID_NEGLAP: return launch_fft_pn<T, N, PS, NTS, OP_MID_NEGLAP, T, A, 3>(io, tw, mk, kind, axis, g, s);
    case OP_PLAIN: return launch_fft_pn<T, N, PS, NTS, OP_PLAIN, T, A, 3>(io, tw, mk, kind, axis, g, s);
    default: break;
  }
  if constexpr (std::is_same<A, double>::value) {
    if (io.f.op == OP_MID_AMP) return launch_fft_pn<T, N, PS, NTS, OP_MID_AMP, T, double, 3>(io, tw, mk, kind, axis, g, s);
  }
  return hipErrorInvalidValue;
}

template <typename T>
static hipError_t launch_direct(const PassIO& io, void* costab, int kind, int axis, int n,
                                const Geom& g, hipStream_t s) {
  DirectArgs<T> d;
  d.p = make_args<T>(io, nullptr, nullptr, kind, axis, g);
  d.costab = static_cast<const double*>(costab);
  d.n = n;
  const bool mid = io.f.op == OP_MID_AMP || io.f.op == OP_MID_NEGLAP || is_axis_op(io.f.op);
  int tl = 16;  // lines per tile; the LDS tile is FP64
  while (tl > 1 && (size_t)tl * n * sizeof(double) * (mid ? 2 : 1) > 48 * 1024) tl /= 2;
  d.tl = tl;
  long long blocks;
  if (axis == 2) blocks = ((long long)g.nx * g.ny + tl - 1) / tl;
  else blocks = (long long)(axis == 1 ? g.nx : g.ny) * ((g.nz + tl - 1) / tl);
  const size_t lds = (size_t)tl * n * sizeof(double) * (mid ? 2 : 1);
  hipLaunchKernelGGL((direct_pass_kernel<T>), dim3((unsigned)blocks, (unsigned)g.nb), dim3(kThreads), lds, s, d);
  return hipGetLastError();
}

// FP32 arithmetic (A = float, FP32 tables) exists for T = float and the
// lengths of r2r_arith32_supported only; the caller checks both.
template <typename T, typename A = double>
static hipError_t dispatch(int n, const PassIO& io, void* tw, void* mk, void* ct, int kind,
                           int axis, const Geom& g, hipStream_t s) {
  constexpr bool a32 = std::is_same<A, float>::value;
  if (n == 256 && axis != 2) {
    const DType dt = sizeof(T) == 8 ? F64 : F32;
    const hipError_t e = launch_yx256(dt, dt, kind, axis, io, g, tw, mk, s, a32);
    if (e != hipErrorNotSupported) return e;
  }
  if (is_axis_op(io.f.op) && R2R::fft_supported(n)) return launch_axis_op<T, A>(n, io, tw, mk, axis, g, s);
  if constexpr (a32) {
    switch (n) {
      case 64: return launch_fft<T, 64, A>(io, tw, mk, kind, axis, g, s);
      case 96: return launch_fft<T, 96, A>(io, tw, mk, kind, axis, g, s);
      case 128: return launch_fft<T, 128, A>(io, tw, mk, kind, axis, g, s);
      case 256: return launch_fft<T, 256, A>(io, tw, mk, kind, axis, g, s);
      case 512: return launch_fft<T, 512, A>(io, tw, mk, kind, axis, g, s);
      default: return hipErrorInvalidValue;
    }
  }
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

static bool op_valid(int op, int axis) {
  if (op < 0 || op >= OP_COUNT) return false;
  if (op == OP_PLAIN) return true;
  if (op == OP_AX_GAUSS_AMP) return axis == 2;
  if (is_axis_op(op)) return true;
  if (op == OP_MID_AMP || op == OP_MID_NEGLAP) return axis == 0;
  return axis == 2;  // load ops act in the first (z) pass
}

hipError_t R2R::pass(DType dt, int kind, int axis, const PassIO& io, const Geom& g,
                     hipStream_t s) {
  if (!op_valid(io.f.op, axis)) return hipErrorInvalidValue;
  const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  // FP64 arithmetic and tables for both storage types unless FP32
  // arithmetic is asked for (FP32 storage, a length with FP32 instances)
  const bool a32 = dt == F32 && io.f.arith32 && r2r_arith32_supported(n);
  Tables* tb = nullptr;
  hipError_t e = tables(a32 ? F32 : F64, n, &tb);
  if (e != hipSuccess) return e;
  if (a32) return dispatch<float, float>(n, io, tb->tw, tb->mk, tb->costab, kind, axis, g, s);
  if (io.f.op == OP_PLAIN && !io.f.untuned && !tuned_.empty()) {
    const auto it = tuned_.find(tune_key(dt, axis, g));
    if (it != tuned_.end() && it->second > 0) {
      if (dt == F64) return launch_plain_variant<double>(n, it->second, io, tb->tw, tb->mk, kind, axis, g, s);
      return launch_plain_variant<float>(n, it->second, io, tb->tw, tb->mk, kind, axis, g, s);
    }
  }
  if (dt == F64) return dispatch<double>(n, io, tb->tw, tb->mk, tb->costab, kind, axis, g, s);
  return dispatch<float>(n, io, tb->tw, tb->mk, tb->costab, kind, axis, g, s);
}

hipError_t r2r_plain_pass_with_tables(DType dt, int kind, int axis, const PassIO& io, const Geom& g,
                                      const void* tw, const void* mk, const void* ct, hipStream_t s) {
  if (io.f.op != OP_PLAIN) return hipErrorInvalidValue;
  const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  void* t = const_cast<void*>(tw);
  void* m = const_cast<void*>(mk);
  void* c = const_cast<void*>(ct);
  if (dt == F64) return dispatch<double>(n, io, t, m, c, kind, axis, g, s);
  return dispatch<float>(n, io, t, m, c, kind, axis, g, s);
}

void R2R::set_variant(DType dt, int axis, const Geom& g, int v) {
  if (v <= 0) tuned_.erase(tune_key(dt, axis, g));
  else tuned_[tune_key(dt, axis, g)] = v;
}

int R2R::variant(DType dt, int axis, const Geom& g) const {
  const auto it = tuned_.find(tune_key(dt, axis, g));
  return it == tuned_.end() ? 0 : it->second;
}

hipError_t R2R::autotune(DType dt, const Geom& g, hipStream_t s, int* chosen, float* ms) {
  const size_t bytes = (size_t)g.n3() * g.nb * (dt == F64 ? 8 : 4);
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, bytes);
  if (e == hipSuccess) e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, bytes, s);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  constexpr int kReps = 5;
  for (int axis = 0; axis < 3 && e == hipSuccess; ++axis) {
    const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
    int best = 0;
    float best_ms = 0;
    for (int v = 0; v < kTileVariants; ++v) {
      if (ms) ms[axis * kTileVariants + v] = -1.0f;
      if (!fft_supported(n) && v > 0) continue;
      set_variant(dt, axis, g, v);
      PassIO fwd, inv;
      fwd.in = a; fwd.out = b;
      inv.in = b; inv.out = a;
      hipError_t r = pass(dt, REDFT10, axis, fwd, g, s);
      if (r == hipErrorNotSupported) {
        (void)hipGetLastError();
        continue;
      }
      if (r == hipSuccess) r = pass(dt, REDFT01, axis, inv, g, s);
      if (r == hipSuccess) r = hipEventRecord(e0, s);
      for (int k = 0; k < kReps && r == hipSuccess; ++k) {
        r = pass(dt, REDFT10, axis, fwd, g, s);
        if (r == hipSuccess) r = pass(dt, REDFT01, axis, inv, g, s);
      }
      if (r == hipSuccess) r = hipEventRecord(e1, s);
      if (r == hipSuccess) r = hipEventSynchronize(e1);
      float t = 0;
      if (r == hipSuccess) r = hipEventElapsedTime(&t, e0, e1);
      if (r != hipSuccess) { e = r; break; }
      t /= kReps;
      if (ms) ms[axis * kTileVariants + v] = t;
      if (v == 0 || t < best_ms) { best = v; best_ms = t; }
    }
    set_variant(dt, axis, g, best);
    if (chosen) chosen[axis] = best;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return e;
}

hipError_t R2R::transform3d(DType dt, const int kinds[3], const PassIO& io, const Geom& g,
                            hipStream_t s) {
  PassIO p = io;
  hipError_t e = pass(dt, kinds[2], 2, p, g, s);
  if (e != hipSuccess) return e;
  p.in = io.out;
  p.in2 = nullptr;
  p.f.op = OP_PLAIN;
  if ((e = pass(dt, kinds[1], 1, p, g, s)) != hipSuccess) return e;
  return pass(dt, kinds[0], 0, p, g, s);
}

}  // namespace smaiso
