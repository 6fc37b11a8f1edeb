// Batched 1-D and 3-D real-to-real transforms (DCT-II/III, DST-II/III) on one
// MI355X, with FFTW's unnormalised REDFT10/REDFT01/RODFT10/RODFT01 semantics.
//
// These replace the reference's eight fftw_mpi_plan_r2r_3d plans
// (smaiso-wcomp.cpp:499-545).  A 3-D transform is three axis passes over HBM
// (z, then y, then x); each pass moves a tile of lines through LDS, so the
// x-slab MPI transposes of the reference become strided tile loads on one GPU.
#pragma once

#include "common.h"
#include <map>
#include <mutex>
#include <vector>

namespace smaiso {

enum DType : int { F64 = 0, F32 = 1 };

// Optional fused operations on the load side of an axis pass:
//   x = in[a] (+ in2[a])
struct PassIO {
  const void* in = nullptr;
  const void* in2 = nullptr;   // added element-wise to `in` on load when set
  void* out = nullptr;
};

class R2R {
 public:
  R2R() = default;
  ~R2R();
  // One axis pass: out = kind_axis(in (+ in2)).  in may alias out.
  hipError_t pass(DType dt, int kind, int axis, const PassIO& io, const Geom& g,
                  hipStream_t s);
  // Full 3-D transform with kinds[a] on axis a; passes run z, y, x.
  hipError_t transform3d(DType dt, const int kinds[3], const PassIO& io, const Geom& g,
                         hipStream_t s);
  // True when length n runs on the LDS FFT path (else the direct O(n^2) path).
  static bool fft_supported(int n);

 private:
  struct Tables {
    void* tw = nullptr;      // n complex: exp(-2 pi i m / n)
    void* mk = nullptr;      // n complex: exp(-i pi m / (2n))
    void* costab = nullptr;  // 4n reals: cos(pi t / (2n))
  };
  hipError_t tables(DType dt, int n, Tables** out);
  std::map<std::pair<int, int>, Tables> cache_;
  std::mutex mu_;
};

}  // namespace smaiso
