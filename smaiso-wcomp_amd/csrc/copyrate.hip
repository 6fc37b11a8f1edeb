// smaiso_hbm_copy_rate: the HBM rate a plain copy kernel reaches on the
// current device, the measured counterpart of the 8 TB/s specification that
// the roofline figures are quoted against.
#include "../../include/smaiso.h"
#include "host_abi.h"

#include <hip/hip_runtime.h>
#include <string>

namespace {

// Four 16-byte elements per lane, one pass over the buffer (no grid-stride
// loop): all four loads are in flight before the stores.  Of the shapes
// tried on MI355X (grid-stride, 4 or 8 per lane) this one copies fastest
// over buffers beyond the memory-side cache.
constexpr int kPerLane = 4;
__global__ __launch_bounds__(256) void k_copy16(const double2* __restrict__ a, double2* __restrict__ b,
                                                long long n) {
  const long long base = (long long)blockIdx.x * kPerLane * 256 + threadIdx.x;
  double2 v[kPerLane];
#pragma unroll
  for (int u = 0; u < kPerLane; ++u) {
    const long long i = base + u * 256;
    v[u] = i < n ? a[i] : double2{0.0, 0.0};
  }
#pragma unroll
  for (int u = 0; u < kPerLane; ++u) {
    const long long i = base + u * 256;
    if (i < n) b[i] = v[u];
  }
}

}  // namespace

extern "C" int smaiso_hbm_copy_rate(int64_t bytes, int32_t reps, double* gbs) {
  using smaiso::abi_fail;
  if (!gbs || reps < 1 || bytes < (1 << 20) || bytes % 16 != 0)
    return abi_fail(SMAISO_E_INVALID, "smaiso_hbm_copy_rate: need gbs, reps >= 1 and a multiple of 16 bytes >= 1 MiB");
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, (size_t)bytes);
  if (e == hipSuccess) e = hipMalloc(&b, (size_t)bytes);
  if (e == hipSuccess) e = hipMemset(a, 0, (size_t)bytes);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const long long n = bytes / 16;
  const unsigned blocks = (unsigned)((n + kPerLane * 256 - 1) / (kPerLane * 256));
  float ms = 0;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, nullptr, static_cast<const double2*>(a),
                       static_cast<double2*>(b), n);
    e = hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
      hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, nullptr, static_cast<const double2*>(a),
                         static_cast<double2*>(b), n);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e != hipSuccess)
    return abi_fail(e == hipErrorOutOfMemory ? SMAISO_E_NOMEM : SMAISO_E_HIP,
                    std::string("smaiso_hbm_copy_rate: ") + hipGetErrorString(e));
  *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
  return SMAISO_OK;
}
