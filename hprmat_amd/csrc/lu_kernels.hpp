// Kernels of the blocked right-looking LU with partial pivoting (GEPP) and of
// the triangular solves, for row-major complex matrices on gfx950.
//
// Algorithm: LAPACK getrf/getf2/laswp/getrs (Anderson et al., LAPACK Users'
// Guide 3rd ed. 1999; Golub & Van Loan 4th ed. 3.4).  The pivot of column j
// is the first row holding the largest |Re| + |Im| (the i?amax rule), so
// pivot sequences match LAPACK on matrices without near ties.
#pragma once
#include "common.hpp"

// ---------------------------------------------------------------------------
// Grid-wide barrier for the cooperative panel kernel.  The counter is zeroed
// before the launch and counts arrivals monotonically; the spin is bounded so
// that a barrier that can never complete (workgroups not co-resident) ends
// with *err set instead of hanging.  Every waiting workgroup also watches
// *err so one timeout releases all of them quickly.
// ---------------------------------------------------------------------------
__device__ inline void grid_barrier(unsigned* counter, unsigned target, int* err) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    atomicAdd(counter, 1u);
    unsigned spins = 0;
    while (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if ((++spins & 1023u) == 0) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (spins > (1u << 22)) { atomicExch(err, 1); break; }
      }
    }
    __threadfence();
  }
  __syncthreads();
}

// Exchange area of one panel-leaf launch (device memory, per leaf).
template <class T> struct LeafXchg {
  T* val;        // [2][G]   candidate |Re|+|Im|  (negative: no candidate)
  int* idx;      // [2][G]   candidate panel-local row
  cx<T>* row;    // [2][G][W] candidate row (all W leaf columns)
  cx<T>* rowj;   // [2][W]   current row j (published by its owner)
};

// ---------------------------------------------------------------------------
// Cooperative GEPP of an m x w leaf (w <= W) at P (row-major, ld).  G
// workgroups each keep a slab of rpw consecutive rows in LDS for the whole
// leaf.  Per column j: every workgroup publishes its best candidate (value,
// row and the full row) and the owner of row j publishes row j; one grid
// barrier; then every workgroup picks the same winner (largest value, lowest
// row on ties), performs the interchange on the rows it owns, scales the
// multipliers and applies the rank-1 update to its slab.  One grid barrier
// per column.  ipiv receives global 0-based row indices (row_base + local).
// ---------------------------------------------------------------------------
template <class T, int W>
__global__ void __launch_bounds__(256)
leaf_getf2_kernel(cx<T>* __restrict__ P, int64_t ld, int m, int w, int rpw, int row_base,
                  int* __restrict__ ipiv, int* __restrict__ info, LeafXchg<T> x,
                  unsigned* __restrict__ counter, int* __restrict__ err) {
  extern __shared__ unsigned char smem_raw[];
  cx<T>* slab = reinterpret_cast<cx<T>*>(smem_raw);  // [rpw][W]
  __shared__ cx<T> prow[W];
  __shared__ cx<T> oldj[W];
  __shared__ T red_v[4];
  __shared__ int red_i[4];
  __shared__ int s_win, s_p;
  __shared__ T s_best;

  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  const int lo = g * rpw, hi = min(m, lo + rpw), nrows = max(0, hi - lo);

  for (int e = tid; e < nrows * W; e += 256) {
    int i = e / W, c = e % W;
    slab[e] = (c < w) ? P[(int64_t)(lo + i) * ld + c] : mkc<T>(0, 0);
  }
  __syncthreads();

  for (int j = 0; j < w; ++j) {
    const int par = j & 1;
    // -- local candidate among owned rows >= j
    T bv = T(-1);
    int bi = 0x7fffffff;
    for (int i = max(lo, j) + tid; i < hi; i += 256) {
      T v = cabs1<T>(slab[(i - lo) * W + j]);
      if (v > bv) { bv = v; bi = i; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      T ov = __shfl_xor(bv, off);
      int oi = __shfl_xor(bi, off);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { red_v[tid >> 6] = bv; red_i[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
      for (int q = 1; q < 4; ++q)
        if (red_v[q] > bv || (red_v[q] == bv && red_i[q] < bi)) { bv = red_v[q]; bi = red_i[q]; }
      x.val[par * G + g] = bv;
      x.idx[par * G + g] = bi;
      s_best = bv; s_p = bi;
    }
    __syncthreads();
    if (s_best >= T(0) && tid < W)
      x.row[((int64_t)par * G + g) * W + tid] = slab[(s_p - lo) * W + tid];
    if (j >= lo && j < hi && tid < W) x.rowj[par * W + tid] = slab[(j - lo) * W + tid];

    grid_barrier(counter, (unsigned)(G * (j + 1)), err);

    // -- global winner: every workgroup reduces the same G candidates; the
    //    winning row's slab owner is row / rpw.
    bv = T(-1); bi = 0x7fffffff;
    for (int q = tid; q < G; q += 256) {
      T v = x.val[par * G + q];
      int i = x.idx[par * G + q];
      if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      T ov = __shfl_xor(bv, off);
      int oi = __shfl_xor(bi, off);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { red_v[tid >> 6] = bv; red_i[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
      for (int q = 1; q < 4; ++q)
        if (red_v[q] > bv || (red_v[q] == bv && red_i[q] < bi)) { bv = red_v[q]; bi = red_i[q]; }
      s_best = bv;
      s_p = (bv >= T(0)) ? bi : j;
      s_win = (bv >= T(0)) ? (bi / rpw) : 0;
    }
    __syncthreads();
    const int p = s_p;
    const T best = s_best;
    if (tid < W) {
      prow[tid] = (best >= T(0)) ? x.row[((int64_t)par * G + s_win) * W + tid] : mkc<T>(0, 0);
      oldj[tid] = x.rowj[par * W + tid];
    }
    __syncthreads();
    if (g == 0 && tid == 0) {
      ipiv[j] = row_base + p;
      if (best == T(0) && *info == 0) *info = row_base + j + 1;
    }
    // -- interchange rows j and p inside the owned slabs
    if (p != j && tid < w) {
      if (j >= lo && j < hi) slab[(j - lo) * W + tid] = prow[tid];
      if (p >= lo && p < hi) slab[(p - lo) * W + tid] = oldj[tid];
    }
    __syncthreads();
    // -- multipliers and rank-1 update of the rows below j
    if (best > T(0)) {
      const cx<T> rp = crecip<T>(prow[j]);
      const int r0 = max(lo, j + 1);
      const int cols = w - j - 1;
      if (cols > 0) {
        for (int e = tid; e < (hi - r0) * cols; e += 256) {
          int i = r0 + e / cols, c = j + 1 + e % cols;
          cx<T> l = cmul<T>(slab[(i - lo) * W + j], rp);
          slab[(i - lo) * W + c] = cfms<T>(slab[(i - lo) * W + c], l, prow[c]);
        }
        __syncthreads();
      }
      for (int i = r0 + tid; i < hi; i += 256)
        slab[(i - lo) * W + j] = cmul<T>(slab[(i - lo) * W + j], rp);
    }
    __syncthreads();
  }

  for (int e = tid; e < nrows * W; e += 256) {
    int i = e / W, c = e % W;
    if (c < w) P[(int64_t)(lo + i) * ld + c] = slab[e];
  }
}

// ---------------------------------------------------------------------------
// Row interchanges k1 .. k1+nk-1 (global 0-based ipiv, ipiv[k] >= k) applied
// to the columns [c0, c1) of a row-major matrix.  Each workgroup takes a
// 128-byte column strip, stages every row the sequence touches in LDS (the nk
// rows k1.. and each distinct pivot row below them), replays the swaps in
// order there and writes the rows back: two coalesced passes instead of a
// chain of dependent global swaps.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256)
laswp_kernel(cx<T>* __restrict__ A, int64_t ld, int c0, int c1, int k1, int nk,
             const int* __restrict__ ipiv, const int* __restrict__ skip) {
  if (skip && *skip) return;
  constexpr int CW = 128 / sizeof(cx<T>);
  extern __shared__ unsigned char smem_raw[];
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem_raw);  // [2 * nk][CW]
  int* slot = reinterpret_cast<int*>(buf + 2 * nk * CW);  // [nk]
  int* srcrow = slot + nk;                                 // [2 * nk]
  const int tid = threadIdx.x;
  const int cb = c0 + blockIdx.x * CW;
  const int ncol = min(CW, c1 - cb);
  if (ncol <= 0) return;

  for (int t = tid; t < nk; t += 256) {
    int p = ipiv[k1 + t];
    int s;
    if (p < k1 + nk) {
      s = p - k1;
    } else {
      int first = t;
      for (int q = 0; q < t; ++q)
        if (ipiv[k1 + q] == p) { first = q; break; }
      s = nk + first;
    }
    slot[t] = s;
    srcrow[t] = k1 + t;
    srcrow[nk + t] = (s == nk + t) ? p : -1;
  }
  __syncthreads();
  for (int e = tid; e < 2 * nk * CW; e += 256) {
    int r = e / CW, c = e % CW;
    int gr = srcrow[r];
    if (gr >= 0 && c < ncol) buf[e] = A[(int64_t)gr * ld + cb + c];
  }
  __syncthreads();
  if (tid < ncol) {
    for (int t = 0; t < nk; ++t) {
      int s = slot[t];
      if (s != t) {
        cx<T> a = buf[t * CW + tid];
        buf[t * CW + tid] = buf[s * CW + tid];
        buf[s * CW + tid] = a;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < 2 * nk * CW; e += 256) {
    int r = e / CW, c = e % CW;
    int gr = srcrow[r];
    if (gr >= 0 && c < ncol) A[(int64_t)gr * ld + cb + c] = buf[e];
  }
}

// ---------------------------------------------------------------------------
// In-place triangular solve with one diagonal block D (nb x nb, nb <= 64) of a
// row-major factor: lower unit (forward) or upper non-unit (backward), on the
// nb x ncols block row S of a row-major right-hand side.  Each workgroup
// stages D and a 16-column strip of S in LDS and substitutes there.
// ---------------------------------------------------------------------------
template <class T, bool UPPER>
__global__ void __launch_bounds__(256)
trsm_diag_kernel(const cx<T>* __restrict__ D, int64_t ldd, int nb, cx<T>* __restrict__ S,
                 int64_t lds, int ncols, const int* __restrict__ skip) {
  if (skip && *skip) return;
  constexpr int CW = 16, NBM = 64;
  __shared__ cx<T> d[NBM][NBM + 1];
  __shared__ cx<T> s[NBM][CW];
  __shared__ cx<T> rdiag[NBM];
  const int tid = threadIdx.x;
  const int cb = blockIdx.x * CW;
  const int ncol = min(CW, ncols - cb);
  for (int e = tid; e < nb * nb; e += 256) {
    int i = e / nb, k = e % nb;
    bool need = UPPER ? (k >= i) : (k < i);
    d[i][k] = need ? D[(int64_t)i * ldd + k] : mkc<T>(0, 0);
  }
  for (int e = tid; e < nb * CW; e += 256) {
    int i = e / CW, c = e % CW;
    s[i][c] = (c < ncol) ? S[(int64_t)i * lds + cb + c] : mkc<T>(0, 0);
  }
  __syncthreads();
  if (UPPER) {
    if (tid < nb) rdiag[tid] = crecip<T>(d[tid][tid]);
    __syncthreads();
    for (int i = nb - 1; i >= 0; --i) {
      if (tid < CW) s[i][tid] = cmul<T>(s[i][tid], rdiag[i]);
      __syncthreads();
      for (int e = tid; e < i * CW; e += 256) {
        int r = e / CW, c = e % CW;
        s[r][c] = cfms<T>(s[r][c], d[r][i], s[i][c]);
      }
      __syncthreads();
    }
  } else {
    for (int i = 0; i < nb - 1; ++i) {
      int rows = nb - 1 - i;
      for (int e = tid; e < rows * CW; e += 256) {
        int r = i + 1 + e / CW, c = e % CW;
        s[r][c] = cfms<T>(s[r][c], d[r][i], s[i][c]);
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < nb * CW; e += 256) {
    int i = e / CW, c = e % CW;
    if (c < ncol) S[(int64_t)i * lds + cb + c] = s[i][c];
  }
}
