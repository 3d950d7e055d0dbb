// Shared types and helpers for the gfx950 dense complex LU library.
//
// Internal matrices are ROW-MAJOR: element (i, j) of a matrix with leading
// dimension ld lives at p[i * ld + j].  Row interchanges are then contiguous
// (coalesced) and a panel of w columns is a set of contiguous w-element row
// segments.  The C ABI accepts column-major (Fortran) storage and transposes
// on the way in and out.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

#define HPR_CHECK(call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      fprintf(stderr, "hprmat_hip: HIP error %s at %s:%d\n",                   \
              hipGetErrorString(e_), __FILE__, __LINE__);                      \
      return -1;                                                               \
    }                                                                          \
  } while (0)

// Complex element of precision T (x = real, y = imaginary), 8- or 16-byte
// aligned so one element is a single vector load.
template <class T> struct cplx_t;
template <> struct cplx_t<double> { typedef double2 type; };
template <> struct cplx_t<float> { typedef float2 type; };

template <class T> using cx = typename cplx_t<T>::type;

template <class T> __host__ __device__ inline cx<T> mkc(T r, T i) {
  cx<T> z; z.x = r; z.y = i; return z;
}
template <class T> __device__ inline cx<T> cmul(cx<T> a, cx<T> b) {
  return mkc<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class T> __device__ inline cx<T> csub(cx<T> a, cx<T> b) {
  return mkc<T>(a.x - b.x, a.y - b.y);
}
// a - b*c
template <class T> __device__ inline cx<T> cfms(cx<T> a, cx<T> b, cx<T> c) {
  return mkc<T>(a.x - (b.x * c.x - b.y * c.y), a.y - (b.x * c.y + b.y * c.x));
}
// |Re| + |Im|, the pivot measure of LAPACK i?amax for complex data.
template <class T> __device__ inline T cabs1(cx<T> a) {
  return fabs(a.x) + fabs(a.y);
}
// 1/z by Smith's scaling, avoids overflow of |z|^2.
template <class T> __device__ inline cx<T> crecip(cx<T> z) {
  if (fabs(z.x) >= fabs(z.y)) {
    T r = z.y / z.x, d = z.x + z.y * r;
    return mkc<T>(T(1) / d, -r / d);
  } else {
    T r = z.x / z.y, d = z.y + z.x * r;
    return mkc<T>(r / d, T(-1) / d);
  }
}
// a / b by Smith's method.
template <class T> __device__ inline cx<T> cdiv(cx<T> a, cx<T> b) {
  if (fabs(b.x) >= fabs(b.y)) {
    T r = b.y / b.x, d = b.x + b.y * r;
    return mkc<T>((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    T r = b.x / b.y, d = b.y + b.x * r;
    return mkc<T>((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}

// One 16x16x4 MFMA step.  Lane l supplies a = A[l % 16][l / 16] and
// b = B[l / 16][l % 16].  Accumulator element r of lane l holds D[i][l % 16]
// with i = 4 * (l / 16) + r for f32, but i = 4 * r + l / 16 for f64 (measured
// on gfx950); mfma_row gives i for either.
__device__ inline double4_t mfma16x4(double a, double b, double4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline float4_t mfma16x4(float a, float b, float4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <class T> __device__ inline int mfma_row(int lk, int r);
template <> __device__ inline int mfma_row<double>(int lk, int r) { return 4 * r + lk; }
template <> __device__ inline int mfma_row<float>(int lk, int r) { return 4 * lk + r; }
template <class T> struct acc_t;
template <> struct acc_t<double> { typedef double4_t type; };
template <> struct acc_t<float> { typedef float4_t type; };

static inline int64_t cdiv_i64(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Raise the issue priority of the calling wave within its SIMD (s_setprio 3)
// when `on`: kernels on the latency-bound panel chain of the factorization
// use it so that their instructions win arbitration against the MFMA-bound
// trailing-update waves that share the SIMDs with them.
__device__ __forceinline__ void raise_wave_priority(int on) {
  if (on) __builtin_amdgcn_s_setprio(3);
}
