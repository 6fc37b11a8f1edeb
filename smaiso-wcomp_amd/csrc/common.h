// Shared device/host helpers for the smaiso MI355X stepper.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace smaiso {

// r2r kinds, numbered as FFTW_REDFT10, FFTW_REDFT01, FFTW_RODFT10, FFTW_RODFT01
// are used by the reference's plans (smaiso-wcomp.cpp:499-545).
enum R2RKind : int { REDFT10 = 0, REDFT01 = 1, RODFT10 = 2, RODFT01 = 3 };

template <typename T> struct cplx { T x, y; };

template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}

// Grid geometry of one field, row-major (i*ny + j)*nz + k with k fastest,
// the reference's layout (smaiso-wcomp.cpp:943); nb fields of this shape lie
// back to back (a batch of ensemble members).
struct Geom {
  int nx, ny, nz;
  int nb = 1;
  long long n3() const { return (long long)nx * ny * nz; }
};

}  // namespace smaiso
