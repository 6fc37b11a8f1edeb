// smaiso_r2r3d_noalloc: the standalone 3-D r2r transform of smaiso_r2r3d for
// callers that cannot allocate device memory.  smaiso_r2r3d stages the data
// in a hipMalloc'd buffer on every call and builds its twiddle tables in
// hipMalloc'd memory on first use of a length; this variant keeps both in
// `scratch`, a file-scope device array each GPU holds until the process
// ends, and runs the same pass kernels with the same table values, so its
// results are bit-identical to smaiso_r2r3d's.
#include "../../include/smaiso.h"
#include "host_abi.h"
#include "r2r.h"
#include "stepper.h"

#include <cstdio>
#include <map>
#include <mutex>
#include <string>

namespace {

using namespace smaiso;

// Largest field (points) and largest extent along one axis.
constexpr long long kMaxPoints = SMAISO_NOALLOC_MAX_POINTS;
constexpr int kMaxLen = SMAISO_NOALLOC_MAX_LEN;
// Layout of scratch: the FP64 staging of the field, the FP32 work copy
// (FP32 dtype only), then per axis the tables tw (2n), mk (2n), ct (4n).
constexpr size_t kStageOff = 0;
constexpr size_t kWorkOff = kStageOff + (size_t)kMaxPoints * 8;
constexpr size_t kTabOff = kWorkOff + (size_t)kMaxPoints * 4;
constexpr size_t kTabDoubles = 8 * (size_t)kMaxLen;
constexpr size_t kScratchBytes = kTabOff + 3 * kTabDoubles * 8;

}  // namespace

// About 24 MiB per GPU, reserved when the library's code object is loaded on
// that GPU.  Host code never uses this name as an address (on the host it is
// only a stand-in for the symbol): it uses the per-device pointer that
// hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mu;                          // serialises the calls in this process
std::map<int, unsigned char*> g_scratch;  // device -> address of its scratch

// Host staging of the table values (static: nothing is allocated per call).
double g_tab[3][8 * kMaxLen];

int scratch_ptr(int dev, unsigned char** out) {
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) { *out = it->second; return SMAISO_OK; }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    return abi_fail(SMAISO_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[dev] = static_cast<unsigned char*>(p);
  *out = static_cast<unsigned char*>(p);
  return SMAISO_OK;
}

}  // namespace

extern "C" int smaiso_r2r3d_noalloc(const int32_t kinds[3], int32_t nx, int32_t ny, int32_t nz,
                                    const double* in, double* out, int32_t dtype) {
  if (!kinds) return abi_fail(SMAISO_E_INVALID, "kinds is NULL");
  for (int a = 0; a < 3; ++a)
    if (kinds[a] < 0 || kinds[a] > 3) return abi_fail(SMAISO_E_INVALID, "bad kind " + std::to_string(kinds[a]));
  if (!in || !out) return abi_fail(SMAISO_E_INVALID, "in/out is NULL");
  if (dtype != SMAISO_FP64 && dtype != SMAISO_FP32_MIXED) return abi_fail(SMAISO_E_INVALID, "bad dtype");
  if (nx < 1 || ny < 1 || nz < 1) return abi_fail(SMAISO_E_INVALID, "extents must be >= 1");
  const long long n3 = (long long)nx * ny * nz;
  if (nx > kMaxLen || ny > kMaxLen || nz > kMaxLen || n3 > kMaxPoints) {
    char msg[200];
    std::snprintf(msg, sizeof msg,
                  "smaiso_r2r3d_noalloc: %dx%dx%d exceeds its fixed scratch (at most %lld points and %d per "
                  "axis); use smaiso_r2r3d",
                  nx, ny, nz, (long long)kMaxPoints, kMaxLen);
    return abi_fail(SMAISO_E_INVALID, msg);
  }
  std::lock_guard<std::mutex> lk(g_mu);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return abi_fail(SMAISO_E_HIP, hipGetErrorString(e));
  unsigned char* base = nullptr;
  int rc = scratch_ptr(dev, &base);
  if (rc != SMAISO_OK) return rc;
  double* stage = reinterpret_cast<double*>(base + kStageOff);
  float* work = reinterpret_cast<float*>(base + kWorkOff);
  const int len[3] = {nx, ny, nz};
  const void* tw[3];
  const void* mk[3];
  const void* ct[3];
  const hipStream_t s = nullptr;  // the null stream, synchronous like smaiso_r2r3d
  // every byte the passes read is written here first: nothing needs clearing
  for (int a = 0; a < 3 && e == hipSuccess; ++a) {
    const int n = len[a];
    double* h = g_tab[a];
    r2r_table_values(n, h, h + 2 * n, h + 4 * n);
    double* d = reinterpret_cast<double*>(base + kTabOff) + a * kTabDoubles;
    e = hipMemcpyAsync(d, h, (size_t)8 * n * sizeof(double), hipMemcpyHostToDevice, s);
    tw[a] = d;
    mk[a] = d + 2 * n;
    ct[a] = d + 4 * n;
  }
  if (e == hipSuccess) e = hipMemcpyAsync(stage, in, (size_t)n3 * 8, hipMemcpyHostToDevice, s);
  const Geom g{nx, ny, nz};
  const DType dt = dtype == SMAISO_FP64 ? F64 : F32;
  void* buf = stage;
  if (e == hipSuccess && dt == F32) {
    convert_d2f(stage, work, n3, s);
    buf = work;
    e = hipGetLastError();
  }
  PassIO io;
  io.in = buf;
  io.out = buf;
  for (int a = 2; a >= 0 && e == hipSuccess; --a)
    e = r2r_plain_pass_with_tables(dt, kinds[a], a, io, g, tw[a], mk[a], ct[a], s);
  if (e == hipSuccess && dt == F32) {
    convert_f2d(work, stage, n3, s);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, stage, (size_t)n3 * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return abi_fail(SMAISO_E_HIP, std::string("smaiso_r2r3d_noalloc: ") + hipGetErrorString(e));
  return SMAISO_OK;
}
