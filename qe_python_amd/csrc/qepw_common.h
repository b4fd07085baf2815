// Shared device helpers for the gfx950 plane-wave h_psi / Davidson kernels.
//
// Complex numbers are interleaved (re, im) pairs, the layout of numpy's
// complex128/complex64 and of torch's complex tensors, held in HIP's
// double2 / float2 vector types.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace qepw {

template <typename T> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };
template <typename T> using cplx = typename cplx_t<T>::type;

template <typename T>
__device__ __forceinline__ cplx<T> mk(T r, T i) { cplx<T> z; z.x = r; z.y = i; return z; }
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { C z; z.x = a.x + b.x; z.y = a.y + b.y; return z; }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { C z; z.x = a.x - b.x; z.y = a.y - b.y; return z; }
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C z; z.x = a.x * b.x - a.y * b.y; z.y = a.x * b.y + a.y * b.x; return z;
}
// a * conj(b)
template <typename C>
__device__ __forceinline__ C cmulc(C a, C b) {
  C z; z.x = a.x * b.x + a.y * b.y; z.y = a.y * b.x - a.x * b.y; return z;
}
template <typename C, typename T>
__device__ __forceinline__ C cscale(C a, T s) { C z; z.x = a.x * s; z.y = a.y * s; return z; }
// multiply by +i (sign>0) or -i (sign<0)
template <typename C>
__device__ __forceinline__ C cmuli(C a, int sign) {
  C z; if (sign > 0) { z.x = -a.y; z.y = a.x; } else { z.x = a.y; z.y = -a.x; } return z;
}

// Host-visible limits of the LDS FFT machinery.
constexpr int kMaxLine = 256;        // longest 1-D FFT along any axis
constexpr int kMaxStages = 12;       // radix stages per line
constexpr int kPlaneThreads = 512;   // threads of the xy-plane kernel
constexpr int kPlanePoints = 16;     // complex points held per thread per stage
constexpr int kStickThreads = 256;   // threads of the z-stick kernels

// Radix plan of one line length, passed by value to kernels.
struct LinePlan {
  int n;
  int nstages;
  int radix[kMaxStages];
};

}  // namespace qepw
