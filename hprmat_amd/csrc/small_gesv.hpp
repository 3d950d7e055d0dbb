// Small dense complex solves inside one workgroup, in LDS and FP64: the
// nch x nch (nch <= 64) steps of the R-matrix pipeline (S = Z_O^-1 Z_I, the
// propagation of the R-matrix over intervals) that are too small for the
// blocked LU but sit on the device between larger launches.
#pragma once

#include <hip/hip_runtime.h>

namespace hpr {

constexpr int SMALL_GESV_MAX_N = 64;

__device__ inline double2 zmul_d(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline double2 zdiv_d(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// Z X = W for Z (n x n, n <= 64) and W (n x m), both row-major in LDS, by
// Gaussian elimination with partial pivoting (largest |Re| + |Im|, the first
// row on ties, as LAPACK's izamax) and back substitution; X overwrites W, Z
// is destroyed.  Called by every thread of the workgroup (any block size
// that is a multiple of 64); lcol: n LDS entries; piv: one LDS int.  A column
// that is exactly zero is skipped, so a singular Z leaves inf / nan in X.
__device__ inline void lds_gesv(double2* Z, double2* W, int n, int m, double2* lcol, int* piv) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int k = 0; k < n; ++k) {
    if (tid < 64) {   // pivot search in wave 0: rows k .. n-1 (n <= 64)
      const int i = k + tid;
      double v = -1.0;
      int vi = 0x7fffffff;
      if (i < n) {
        const double2 z = Z[i * n + k];
        v = fabs(z.x) + fabs(z.y);
        vi = i;
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(vi, off);
        if (ov > v || (ov == v && oi < vi)) { v = ov; vi = oi; }
      }
      if (tid == 0) *piv = (v > 0.0) ? vi : -1;
    }
    __syncthreads();
    const int p = *piv;
    __syncthreads();       // every thread has read piv before the next column rewrites it
    if (p < 0) continue;
    if (p != k) {
      for (int j = tid; j < n + m; j += nt) {
        const bool z = j < n;
        double2* M = z ? Z : W;
        const int c = z ? j : j - n, ld = z ? n : m;
        const double2 t = M[k * ld + c];
        M[k * ld + c] = M[p * ld + c];
        M[p * ld + c] = t;
      }
      __syncthreads();
    }
    for (int i = k + 1 + tid; i < n; i += nt) lcol[i] = zdiv_d(Z[i * n + k], Z[k * n + k]);
    __syncthreads();
    const int zc = n - k - 1, cols = zc + m;   // Z columns > k, then all of W
    for (int e = tid; e < zc * cols; e += nt) {
      const int i = k + 1 + e / cols, j = e % cols;
      const bool z = j < zc;
      double2* M = z ? Z : W;
      const int c = z ? k + 1 + j : j - zc, ld = z ? n : m;
      const double2 lu = zmul_d(lcol[i], M[k * ld + c]), v = M[i * ld + c];
      M[i * ld + c] = make_double2(v.x - lu.x, v.y - lu.y);
    }
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    const double2 d = Z[k * n + k];
    for (int c = tid; c < m; c += nt) W[k * m + c] = zdiv_d(W[k * m + c], d);
    __syncthreads();
    for (int e = tid; e < k * m; e += nt) {
      const int i = e / m, c = e % m;
      const double2 lu = zmul_d(Z[i * n + k], W[k * m + c]), v = W[i * m + c];
      W[i * m + c] = make_double2(v.x - lu.x, v.y - lu.y);
    }
    __syncthreads();
  }
}

}  // namespace hpr
