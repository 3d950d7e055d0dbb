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
// Exchange area of the cooperative panel leaves (device memory), made of
// tagged 64-bit words: high 32 bits = column tag (unique per column within a
// factorization, the area is zeroed at its start), low 32 bits = payload.  A
// reader accepts a word only when its tag is the current column's, so data
// and "arrival" travel together: no flag, no fence, no store-completion wait.
// Words are written and read with relaxed agent-scope atomics, which compile
// to single coherent (sc1) global loads/stores; an agent-scope fence would
// instead emit buffer_wbl2/buffer_inv and stall on (and evict) the dirty
// tiles of the trailing-update GEMM that runs concurrently on the side stream.
// Slots are double-buffered by column parity: a workgroup can only write
// parity p for column j+2 after all workgroups published column j+1, which
// each does only after it finished reading column j.
// ---------------------------------------------------------------------------
typedef unsigned long long u64;
template <class T> struct words_per;                 // 32-bit payload words per real
template <> struct words_per<float> { static constexpr int n = 1; };
template <> struct words_per<double> { static constexpr int n = 2; };

struct LeafXchg {
  u64* val;    // [2][Gmax]            candidate |Re|+|Im| (payload words: words_per)
  u64* idx;    // [2][Gmax]            candidate panel-local row
  u64* row;    // [2][Gmax][W*2*wpr]   candidate row (all W leaf columns)
  u64* rowj;   // [2][W*2*wpr]         current row j (published by its owner)
  int gstride; // Gmax
};

__device__ inline void put_word(u64* p, unsigned tag, unsigned payload) {
  __hip_atomic_store(p, ((u64)tag << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline u64 get_word(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool tag_ok(u64 w, unsigned tag) { return (unsigned)(w >> 32) == tag; }

// Real value <-> tagged words.
__device__ inline void put_real(u64* p, unsigned tag, float v) { put_word(p, tag, __float_as_uint(v)); }
__device__ inline void put_real(u64* p, unsigned tag, double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  put_word(p, tag, (unsigned)b);
  put_word(p + 1, tag, (unsigned)(b >> 32));
}
// Spin (bounded, watching *err) until every word of a real carries the tag.
template <class T> __device__ inline T get_real(const u64* p, unsigned tag, int* err);
template <> __device__ inline float get_real<float>(const u64* p, unsigned tag, int* err) {
  unsigned spins = 0;
  u64 w = get_word(p);
  while (!tag_ok(w, tag)) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 255u) == 0) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      if (spins > (1u << 22)) { atomicExch(err, 1); break; }
    }
    w = get_word(p);
  }
  return __uint_as_float((unsigned)w);
}
template <> __device__ inline double get_real<double>(const u64* p, unsigned tag, int* err) {
  unsigned spins = 0;
  u64 lo = get_word(p), hi = get_word(p + 1);
  while (!tag_ok(lo, tag) || !tag_ok(hi, tag)) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 255u) == 0) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      if (spins > (1u << 22)) { atomicExch(err, 1); break; }
    }
    lo = get_word(p);
    hi = get_word(p + 1);
  }
  return __longlong_as_double((long long)((lo & 0xffffffffull) | ((hi & 0xffffffffull) << 32)));
}
__device__ inline int get_int(const u64* p, unsigned tag, int* err) {
  unsigned spins = 0;
  u64 w = get_word(p);
  while (!tag_ok(w, tag)) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 255u) == 0) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      if (spins > (1u << 22)) { atomicExch(err, 1); break; }
    }
    w = get_word(p);
  }
  return (int)(unsigned)w;
}

template <class T>
__device__ inline void best_of(T& bv, int& bi, T ov, int oi) {
  if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
}

// Workgroup argmax of (value, index): largest value, lowest index on ties
// (red_v / red_i hold one slot per wavefront).
template <class T>
__device__ inline void wg_argmax(T& bv, int& bi, T* red_v, int* red_i) {
  for (int off = 32; off > 0; off >>= 1) {
    T ov = __shfl_xor(bv, off);
    int oi = __shfl_xor(bi, off);
    best_of(bv, bi, ov, oi);
  }
  if ((threadIdx.x & 63) == 0) { red_v[threadIdx.x >> 6] = bv; red_i[threadIdx.x >> 6] = bi; }
  __syncthreads();
  bv = red_v[0]; bi = red_i[0];
  for (int q = 1; q < (int)(blockDim.x >> 6); ++q) best_of(bv, bi, red_v[q], red_i[q]);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Cooperative GEPP of an m x w leaf (w <= W) at P (row-major, ld).  G
// workgroups of NT threads own NT * RPT consecutive rows each; every thread
// keeps its RPT rows of the leaf in registers for the whole leaf (the column
// loop is fully unrolled, so row updates are register FMAs against the pivot
// row broadcast from LDS).  Per column j: each workgroup finds its best
// candidate (wavefront shuffles + LDS), publishes value, row index and the
// candidate row as tagged words (and the owner of row j publishes row j);
// every workgroup polls the G (value, index) pairs until all carry column j's
// tag, picks the same winner (largest |Re|+|Im|, lowest row on ties),
// fetches the winning row, performs the interchange on the rows it owns and
// applies the scaling and rank-1 update: two global round trips per column,
// none when G = 1.  ipiv receives global 0-based rows.
// ---------------------------------------------------------------------------
template <class T, int W, int NT, int RPT>
__global__ void __launch_bounds__(NT)
leaf_getf2_kernel(cx<T>* __restrict__ P, int64_t ld, int m, int w, int row_base,
                  int* __restrict__ ipiv, int* __restrict__ info, LeafXchg x, unsigned seq_base,
                  int* __restrict__ err) {
  constexpr int RW = 2 * words_per<T>::n;   // words per complex element
  constexpr int RPW = NT * RPT;             // rows per workgroup
  constexpr int STAGE_ROWS = 128;           // rows per LDS staging chunk
  __shared__ cx<T> prow[W];                 // pivot row (rotated order, see below)
  __shared__ cx<T> oldj[W];                 // row j before the interchange
  __shared__ T red_v[NT / 64];
  __shared__ int red_i[NT / 64];
  const int tid = threadIdx.x;
  const int G = gridDim.x, g = blockIdx.x;
  const int lo = g * RPW, hi = min(m, lo + RPW);
  const int gs = x.gstride;

  // Rows live in registers in ROTATED order: before step j, register c holds
  // column (c + j) mod W, so the step always works on register 0 and every
  // index is a compile-time constant; each step rotates the row left by one
  // and appends the finished multiplier.  Registers W-j .. W-1 hold finished
  // L entries, which the update must not touch (the pivot row is zeroed there).
  // coalesced load through LDS, NT rows at a time (a thread's row is W
  // consecutive elements, so direct per-thread loads would touch one cache
  // line per lane)
  __shared__ cx<T> stage[STAGE_ROWS][W + 1];
  cx<T> a[RPT][W];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    for (int s0 = 0; s0 < NT; s0 += STAGE_ROWS) {
      __syncthreads();
      for (int e = tid; e < STAGE_ROWS * W; e += NT) {
        const int i = e / W, c = e % W, row = lo + r * NT + s0 + i;
        stage[i][c] = (row < hi && c < w) ? P[(int64_t)row * ld + c] : mkc<T>(0, 0);
      }
      __syncthreads();
      if (tid >= s0 && tid < s0 + STAGE_ROWS) {
#pragma unroll
        for (int c = 0; c < W; ++c) a[r][c] = stage[tid - s0][c];
      }
    }
  }

  for (int j = 0; j < w; ++j) {
    const int par = j & 1;
    const unsigned tag = seq_base + (unsigned)j + 1u;
    // -- local candidate among owned rows >= j (rows ascend with r)
    T bv = T(-1);
    int bi = 0x7fffffff;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int row = lo + tid + r * NT;
      if (row < hi && row >= j) {
        const T v = cabs1<T>(a[r][0]);
        if (v > bv) { bv = v; bi = row; }
      }
    }
    wg_argmax(bv, bi, red_v, red_i);
    // -- stage this workgroup's candidate row and row j in LDS
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int row = lo + tid + r * NT;
      if (row == bi && bv >= T(0)) {
#pragma unroll
        for (int c = 0; c < W; ++c) prow[c] = a[r][c];
      }
      if (row == j && row < hi) {
#pragma unroll
        for (int c = 0; c < W; ++c) oldj[c] = a[r][c];
      }
    }
    __syncthreads();
    int p;
    T best;
    if (G > 1) {
      // -- publish (fire and forget): value, index, candidate row, row j
      if (tid == 0) {
        put_real(x.val + ((size_t)par * gs + g) * words_per<T>::n, tag, bv);
        put_word(x.idx + (size_t)par * gs + g, tag, (unsigned)bi);
      }
      if (tid < W * RW) {
        const int c = tid / RW, part = tid % RW;
        const int comp = part / words_per<T>::n;
        if (words_per<T>::n == 1 || (part % 2) == 0) {
          u64* dst = x.row + (((size_t)par * gs + g) * W + c) * RW + comp * words_per<T>::n;
          put_real(dst, tag, comp ? prow[c].y : prow[c].x);
          if (j >= lo && j < hi) {
            u64* dj = x.rowj + ((size_t)par * W + c) * RW + comp * words_per<T>::n;
            put_real(dj, tag, comp ? oldj[c].y : oldj[c].x);
          }
        }
      }
      // -- gather the G candidates (polling their tags) and reduce
      bv = T(-1); bi = 0x7fffffff;
      for (int q = tid; q < G; q += NT)
        best_of(bv, bi, get_real<T>(x.val + ((size_t)par * gs + q) * words_per<T>::n, tag, err),
                get_int(x.idx + (size_t)par * gs + q, tag, err));
      wg_argmax(bv, bi, red_v, red_i);   // its barriers also order the prow/oldj reads above
      best = bv;
      p = (bv >= T(0)) ? bi : j;
      const int win = (bv >= T(0)) ? (bi / RPW) : 0;
      // -- winning row and row j (each word's tag checked)
      if (tid < 2 * W) {
        const int c = tid % W;
        const bool want_j = tid >= W;
        const u64* src = want_j ? x.rowj + ((size_t)par * W + c) * RW
                                : x.row + (((size_t)par * gs + win) * W + c) * RW;
        cx<T> v = mkc<T>(0, 0);
        if (want_j || bv >= T(0))
          v = mkc<T>(get_real<T>(src, tag, err), get_real<T>(src + words_per<T>::n, tag, err));
        if (want_j) oldj[c] = v; else prow[c] = v;
      }
      __syncthreads();
    } else {
      best = bv;
      p = (bv >= T(0)) ? bi : j;
    }
    if (g == 0 && tid == 0) {
      ipiv[j] = row_base + p;
      if (best == T(0) && *info == 0) *info = row_base + j + 1;
    }
    // -- interchange, multipliers, rank-1 update and rotation, in registers
    const bool piv_ok = best > T(0);
    const cx<T> rp = piv_ok ? crecip<T>(prow[0]) : mkc<T>(0, 0);
    const int live = W - j;   // registers 1 .. live-1 still hold unfinished columns
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int row = lo + tid + r * NT;
      if (p != j) {
        if (row == j) {
#pragma unroll
          for (int c = 0; c < W; ++c) a[r][c] = prow[c];
        } else if (row == p) {
#pragma unroll
          for (int c = 0; c < W; ++c) a[r][c] = oldj[c];
        }
      }
      cx<T> head = a[r][0];
      if (piv_ok && row > j && row < hi) {
        const cx<T> l = cmul<T>(head, rp);
        head = l;
#pragma unroll
        for (int c = 1; c < W; ++c)
          if (c < live) a[r][c] = cfms<T>(a[r][c], l, prow[c]);
      }
#pragma unroll
      for (int c = 0; c < W - 1; ++c) a[r][c] = a[r][c + 1];
      a[r][W - 1] = head;
    }
    __syncthreads();   // prow / oldj are rewritten by the next column
  }

  // undo the rotation (register c now holds column (c + w) mod W) and store
  // coalesced through LDS
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    for (int s0 = 0; s0 < NT; s0 += STAGE_ROWS) {
      __syncthreads();
      if (tid >= s0 && tid < s0 + STAGE_ROWS) {
#pragma unroll
        for (int c = 0; c < W; ++c) {
          int col = c + w;
          if (col >= W) col -= W;
          stage[tid - s0][col] = a[r][c];
        }
      }
      __syncthreads();
      for (int e = tid; e < STAGE_ROWS * W; e += NT) {
        const int i = e / W, c = e % W, row = lo + r * NT + s0 + i;
        if (row < hi && c < w) P[(int64_t)row * ld + c] = stage[i][c];
      }
    }
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
  int* piv = srcrow + 2 * nk;                              // [nk]
  const int tid = threadIdx.x;
  const int cb = c0 + blockIdx.x * CW;
  const int ncol = min(CW, c1 - cb);
  if (ncol <= 0) return;

  for (int t = tid; t < nk; t += 256) piv[t] = ipiv[k1 + t];
  __syncthreads();
  for (int t = tid; t < nk; t += 256) {
    int p = piv[t];
    int s;
    if (p < k1 + nk) {
      s = p - k1;
    } else {
      int first = t;
      for (int q = 0; q < t; ++q)
        if (piv[q] == p) { first = q; break; }
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

// ---------------------------------------------------------------------------
// Final row permutation of a factorization: perm[i] = source row of row i of
// P A, from the interchange sequence ipiv (LAPACK order).  One workgroup;
// the array lives in LDS when n fits, ipiv is staged in chunks so the single
// sequential swap loop never waits on global memory.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
perm_from_ipiv_kernel(const int* __restrict__ ipiv, int n, int* __restrict__ perm, int in_lds) {
  extern __shared__ int sh_perm[];
  __shared__ int chunk[1024];
  int* p = in_lds ? sh_perm : perm;
  for (int i = threadIdx.x; i < n; i += 256) p[i] = i;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += 1024) {
    const int cn = min(1024, n - c0);
    for (int i = threadIdx.x; i < cn; i += 256) chunk[i] = ipiv[c0 + i];
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 0; i < cn; ++i) {
        int k = chunk[i], a = c0 + i;
        if (k != a) { int t = p[a]; p[a] = p[k]; p[k] = t; }
      }
    }
    __syncthreads();
  }
  if (in_lds)
    for (int i = threadIdx.x; i < n; i += 256) perm[i] = sh_perm[i];
}

// out[i][:] = (Tout) (s[perm[i]] in[perm[i]][:]) for row-major n x nrhs arrays
// (s = null: no scaling).
template <class Tin, class Tout>
__global__ void __launch_bounds__(256)
permute_cast_kernel(int n, int nrhs, const int* __restrict__ perm, const cx<Tin>* __restrict__ in,
                    cx<Tout>* __restrict__ out, const double* __restrict__ scale,
                    const int* __restrict__ skip) {
  if (skip && *skip) return;
  const size_t total = (size_t)n * nrhs;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const int i = (int)(e / nrhs), c = (int)(e % nrhs);
    const int src = perm[i];
    cx<Tin> v = in[(size_t)src * nrhs + c];
    const double f = scale ? scale[src] : 1.0;
    out[e] = mkc<Tout>((Tout)(v.x * f), (Tout)(v.y * f));
  }
}

// ---------------------------------------------------------------------------
// Inverses of the 64 x 64 diagonal blocks of a row-major LU factor, in place
// in LDS (LAPACK ztrti2 order): blockIdx.y = 0 -> inverse of the unit lower
// triangle, 1 -> inverse of the upper triangle.  Output: dinv[(b*2+y)*4096 +
// i*64 + k], zero outside the triangle (and unit diagonal for L).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256)
diag_inv_kernel(const cx<T>* __restrict__ LU, int64_t ld, int n, int blk0,
                cx<T>* __restrict__ dinv) {
  __shared__ cx<T> a[64][65];
  __shared__ cx<T> tmp[64];
  const int bi = blk0 + blockIdx.x, upper = blockIdx.y, tid = threadIdx.x;
  const int r0 = bi * 64, bs = min(64, n - r0);
  for (int e = tid; e < 64 * 64; e += 256) {
    int i = e / 64, k = e % 64;
    bool in = i < bs && k < bs && (upper ? (k >= i) : (k < i));
    a[i][k] = in ? LU[(int64_t)(r0 + i) * ld + r0 + k] : mkc<T>(0, 0);
  }
  __syncthreads();
  if (upper) {
    for (int j = 0; j < bs; ++j) {
      if (tid == 0) a[j][j] = crecip<T>(a[j][j]);
      if (tid < j) tmp[tid] = a[tid][j];
      __syncthreads();
      cx<T> s = mkc<T>(0, 0);
      if (tid < j)
        for (int k = tid; k < j; ++k) s = mkc<T>(s.x + a[tid][k].x * tmp[k].x - a[tid][k].y * tmp[k].y,
                                                 s.y + a[tid][k].x * tmp[k].y + a[tid][k].y * tmp[k].x);
      __syncthreads();
      if (tid < j) {
        cx<T> ajj = a[j][j];
        a[tid][j] = mkc<T>(-(ajj.x * s.x - ajj.y * s.y), -(ajj.x * s.y + ajj.y * s.x));
      }
      __syncthreads();
    }
  } else {
    for (int j = bs - 1; j >= 0; --j) {
      if (tid > j && tid < bs) tmp[tid] = a[tid][j];
      __syncthreads();
      cx<T> s = mkc<T>(0, 0);
      if (tid > j && tid < bs) {
        s = tmp[tid];
        for (int k = j + 1; k < tid; ++k) s = mkc<T>(s.x + a[tid][k].x * tmp[k].x - a[tid][k].y * tmp[k].y,
                                                     s.y + a[tid][k].x * tmp[k].y + a[tid][k].y * tmp[k].x);
      }
      __syncthreads();
      if (tid > j && tid < bs) a[tid][j] = mkc<T>(-s.x, -s.y);
      __syncthreads();
    }
    if (tid < bs) a[tid][tid] = mkc<T>(1, 0);
    __syncthreads();
  }
  cx<T>* out = dinv + ((int64_t)bi * 2 + upper) * 4096;
  for (int e = tid; e < 64 * 64; e += 256) out[e] = a[e / 64][e % 64];
}

// ---------------------------------------------------------------------------
// 64 x 64 x 64 complex tile product on MFMA for one workgroup of 4 waves:
// acc[t] += A(rows 16w .. 16w+15, :) x B(:, cols 16t .. 16t+15), with A and B
// held in LDS as real/imaginary planes of row stride 65 (conflict-free
// operand reads).  Accumulator element r of lane l is row mfma_row(l/16, r),
// column l % 16 of its 16 x 16 tile.
// ---------------------------------------------------------------------------
template <class T>
__device__ inline void prod64(const T (*ar)[65], const T (*ai)[65], const T (*br)[65],
                              const T (*bi)[65], typename acc_t<T>::type* re,
                              typename acc_t<T>::type* im) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int k0 = 0; k0 < 64; k0 += 4) {
    const T a_r = ar[16 * w + li][k0 + lk], a_i = ai[16 * w + li][k0 + lk];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const T b_r = br[k0 + lk][16 * t + li], b_i = bi[k0 + lk][16 * t + li];
      re[t] = mfma16x4(a_r, b_r, re[t]);
      re[t] = mfma16x4(-a_i, b_i, re[t]);
      im[t] = mfma16x4(a_r, b_i, im[t]);
      im[t] = mfma16x4(a_i, b_r, im[t]);
    }
  }
}

// ---------------------------------------------------------------------------
// Inverse of one 256 x 256 diagonal block of a row-major LU factor, one 64-
// column block column per workgroup (grid: blocks x {L, U} x 4), from the
// 64 x 64 inverses of its diagonal sub-blocks (block forward / backward
// substitution on MFMA):
//   L: X_jj = inv(L_jj),  X_ij = -inv(L_ii) sum_{k=j}^{i-1} L_ik X_kj  (i > j)
//   U: X_jj = inv(U_jj),  X_ij = -inv(U_ii) sum_{k=i+1}^{j} U_ik X_kj  (i < j)
// Output dinv256[(b*2+u)*65536 + r*256 + c], zero outside the triangle and
// outside a trailing partial block.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256)
blk_inv_kernel(const cx<T>* __restrict__ LU, int64_t ld, int n, int bb0,
               const cx<T>* __restrict__ dinv64, cx<T>* __restrict__ dinv256) {
  typedef typename acc_t<T>::type acc;
  __shared__ T ar[64][65], ai[64][65], br[64][65], bi[64][65];
  const int bb = bb0 + blockIdx.x, upper = blockIdx.y, jb = blockIdx.z;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
  const int r0 = bb * 256, kb = min(256, n - r0), nsb = (kb + 63) / 64;
  cx<T>* X = dinv256 + ((int64_t)bb * 2 + upper) * 65536;
  for (int ib = 0; ib < 4; ++ib) {
    const bool zero = jb >= nsb || ib >= nsb || (upper ? ib > jb : ib < jb);
    if (!zero) continue;
    for (int e = tid; e < 4096; e += 256) X[(ib * 64 + e / 64) * 256 + jb * 64 + e % 64] = mkc<T>(0, 0);
  }
  if (jb >= nsb) return;
  auto inv64 = [&](int sb) { return dinv64 + ((int64_t)(bb * 4 + sb) * 2 + upper) * 4096; };
  {
    const cx<T>* src = inv64(jb);
    for (int e = tid; e < 4096; e += 256) X[(jb * 64 + e / 64) * 256 + jb * 64 + e % 64] = src[e];
  }
  __threadfence();
  __syncthreads();
  auto load = [&](const cx<T>* src, int64_t lds, int rows, int cols, T (*dr)[65], T (*di)[65]) {
    for (int e = tid; e < 4096; e += 256) {
      const int i = e / 64, k = e % 64;
      cx<T> v = (i < rows && k < cols) ? src[(int64_t)i * lds + k] : mkc<T>(0, 0);
      dr[i][k] = v.x;
      di[i][k] = v.y;
    }
  };
  const int step = upper ? -1 : 1;
  for (int ib = jb + step; ib >= 0 && ib < nsb; ib += step) {
    acc re[4], im[4];
    for (int t = 0; t < 4; ++t) { re[t] = acc{}; im[t] = acc{}; }
    const int k_lo = upper ? ib + 1 : jb, k_hi = upper ? jb : ib - 1;
    for (int kk = k_lo; kk <= k_hi; ++kk) {
      __syncthreads();
      load(LU + (int64_t)(r0 + 64 * ib) * ld + r0 + 64 * kk, ld, min(64, kb - 64 * ib),
           min(64, kb - 64 * kk), ar, ai);
      load(X + (int64_t)(64 * kk) * 256 + 64 * jb, 256, 64, 64, br, bi);
      __syncthreads();
      prod64<T>(ar, ai, br, bi, re, im);
    }
    __syncthreads();
    for (int t = 0; t < 4; ++t)
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * w + mfma_row<T>(lk, r), col = 16 * t + li;
        br[row][col] = re[t][r];
        bi[row][col] = im[t][r];
      }
    load(inv64(ib), 64, 64, 64, ar, ai);
    __syncthreads();
    for (int t = 0; t < 4; ++t) { re[t] = acc{}; im[t] = acc{}; }
    prod64<T>(ar, ai, br, bi, re, im);
    for (int t = 0; t < 4; ++t)
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * w + mfma_row<T>(lk, r), col = 16 * t + li;
        X[(int64_t)(64 * ib + row) * 256 + 64 * jb + col] = mkc<T>(-re[t][r], -im[t][r]);
      }
    __threadfence();
    __syncthreads();
  }
}

// dst[r * ldd + c] = src[r * lds + c] for r < rows, c < cols.
template <class T>
__global__ void __launch_bounds__(256)
copy2d_kernel(int rows, int cols, const cx<T>* __restrict__ src, int64_t lds, cx<T>* __restrict__ dst,
              int64_t ldd) {
  const size_t total = (size_t)rows * cols;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const int r = (int)(e / cols), c = (int)(e % cols);
    dst[(int64_t)r * ldd + c] = src[(int64_t)r * lds + c];
  }
}
