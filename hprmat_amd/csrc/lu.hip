// Blocked right-looking LU with partial pivoting, triangular solves and
// mixed-precision iterative refinement for complex matrices on gfx950.
//
// getrf follows LAPACK zgetrf (src/cusolver_interface.cu calls the closed
// cusolverDn?getrf at :383 and :779 for this step):
//   for each panel of nb columns:
//     factor the panel in 32-column leaves (cooperative GEPP kernel), each
//       followed by the leaf's interchanges on the other panel columns, a
//       32-row triangular solve and a K = 32 MFMA update of the panel;
//     apply the panel's interchanges to the columns left and right of it;
//     U12 = L11^-1 A12 (64-row diagonal solves + MFMA updates);
//     A22 -= L21 U12 (MFMA complex GEMM, the O(N^3) hot loop).
// getrs follows LAPACK zgetrs: row interchanges of B, forward substitution
// with unit L, backward substitution with U, blocked by rs_nb rows with MFMA
// GEMM updates (cusolverDn?getrs, src/cusolver_interface.cu:393, 787).
#include "lu.hpp"
#include "gemm.hpp"
#include "lu_kernels.hpp"

namespace hpr {

namespace {

constexpr int LEAF_W = 32;
constexpr int MAX_LEAF_G = 480;

template <class T> constexpr int leaf_rpw_max() { return (64 * 1024) / (LEAF_W * (int)sizeof(cx<T>)); }

// Dynamic LDS above 64 KiB needs an explicit opt-in per kernel; the return
// code is checked so a refusal does not linger as the sticky last error.
template <class K> void allow_big_lds(K kernel) {
  hipError_t e = hipFuncSetAttribute((const void*)kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
  if (e != hipSuccess) {
    fprintf(stderr, "hprmat_hip: hipFuncSetAttribute failed: %s\n", hipGetErrorString(e));
    (void)hipGetLastError();
  }
}

__global__ void fill_u32_kernel(unsigned* p, size_t n, unsigned v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

template <class Tin, class Tout>
__global__ void __launch_bounds__(256)
transpose_cast_kernel(int rows, int cols, const cx<Tin>* __restrict__ in, int64_t ldi,
                      cx<Tout>* __restrict__ out, int64_t ldo) {
  __shared__ cx<Tin> tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int q = ty; q < 32; q += 8) {  // read: in[c * ldi + r], r fastest
    int r = r0 + tx, c = c0 + q;
    if (r < rows && c < cols) tile[q][tx] = in[(int64_t)c * ldi + r];
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {  // write: out[r * ldo + c], c fastest
    int r = r0 + q, c = c0 + tx;
    if (r < rows && c < cols) {
      cx<Tin> v = tile[tx][q];
      out[(int64_t)r * ldo + c] = mkc<Tout>((Tout)v.x, (Tout)v.y);
    }
  }
}

template <class Tin, class Tout>
__global__ void cast_kernel(size_t n, const cx<Tin>* __restrict__ in, cx<Tout>* __restrict__ out,
                            const int* __restrict__ skip) {
  if (skip && *skip) return;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
    cx<Tin> v = in[i];
    out[i] = mkc<Tout>((Tout)v.x, (Tout)v.y);
  }
}

template <class TF>
__global__ void add_correction_kernel(size_t n, double2* __restrict__ X, const cx<TF>* __restrict__ D,
                                      RefineState* __restrict__ st) {
  if (st->done) return;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) st->sweeps += 1;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
    cx<TF> d = D[i];
    double2 x = X[i];
    x.x += (double)d.x;
    x.y += (double)d.y;
    X[i] = x;
  }
}

// Residual statistics, stage 1: per-block (max |Re|+|Im|, its column-major
// flat index, its modulus, sum of |r|^2) over a row-major n x nrhs array.
constexpr int NORM_BLOCKS = 1024;
__global__ void __launch_bounds__(256)
norm_partial_kernel(int n, int nrhs, const double2* __restrict__ R, double* __restrict__ part,
                    const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sv[256], sm[256], ss[256], sa[256];
  __shared__ long long si[256];
  const size_t total = (size_t)n * nrhs;
  double bv = -1.0, bm = 0.0, sq = 0.0, am = 0.0;
  long long bi = 0x7fffffffffffffffLL;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    double2 r = R[e];
    double v = fabs(r.x) + fabs(r.y);
    long long flat = (long long)(e % nrhs) * n + (long long)(e / nrhs);
    double mod2 = r.x * r.x + r.y * r.y;
    if (v > bv || (v == bv && flat < bi)) { bv = v; bi = flat; bm = sqrt(mod2); }
    sq += mod2;
    am = fmax(am, mod2);
  }
  sv[threadIdx.x] = bv; si[threadIdx.x] = bi; sm[threadIdx.x] = bm; ss[threadIdx.x] = sq;
  sa[threadIdx.x] = am;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      int o = threadIdx.x + s;
      if (sv[o] > sv[threadIdx.x] || (sv[o] == sv[threadIdx.x] && si[o] < si[threadIdx.x])) {
        sv[threadIdx.x] = sv[o]; si[threadIdx.x] = si[o]; sm[threadIdx.x] = sm[o];
      }
      ss[threadIdx.x] += ss[o];
      sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x * 5 + 0] = sv[0];
    part[blockIdx.x * 5 + 1] = (double)si[0];
    part[blockIdx.x * 5 + 2] = sm[0];
    part[blockIdx.x * 5 + 3] = ss[0];
    part[blockIdx.x * 5 + 4] = sa[0];
  }
}

// Stage 2 (one block): combine the partials in a fixed order; mode 0 stores
// the RHS norms, mode 1 the residual norms and the convergence decision.
__global__ void __launch_bounds__(256)
norm_final_kernel(int nblocks, const double* __restrict__ part, RefineState* __restrict__ st,
                  int mode, int crit, double tol) {
  if (mode == 1 && st->done) return;
  __shared__ double sv[256], sm[256], ss[256], si[256], sa[256];
  double bv = -1.0, bi = 1e300, bm = 0.0, sq = 0.0, am = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += 256) {
    double v = part[b * 5], i = part[b * 5 + 1];
    if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; bm = part[b * 5 + 2]; }
    sq += part[b * 5 + 3];
    am = fmax(am, part[b * 5 + 4]);
  }
  sv[threadIdx.x] = bv; si[threadIdx.x] = bi; sm[threadIdx.x] = bm; ss[threadIdx.x] = sq;
  sa[threadIdx.x] = am;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      int o = threadIdx.x + s;
      if (sv[o] > sv[threadIdx.x] || (sv[o] == sv[threadIdx.x] && si[o] < si[threadIdx.x])) {
        sv[threadIdx.x] = sv[o]; si[threadIdx.x] = si[o]; sm[threadIdx.x] = sm[o];
      }
      ss[threadIdx.x] += ss[o];
      sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double fro = sqrt(ss[0]);
    if (mode == 0) {
      st->b_fro = fro;
      st->b_max = sqrt(sa[0]);
      st->done = 0; st->sweeps = 0; st->checks = 0;
      st->res_max = 0.0; st->res_fro = 0.0; st->res_absmax = 0.0;
    } else {
      st->res_max = sm[0];
      st->res_absmax = sqrt(sa[0]);
      st->res_fro = fro;
      double bf = st->b_fro != 0.0 ? st->b_fro : 1.0;
      if (st->checks < 32) st->hist[st->checks] = fro / bf;
      st->checks += 1;
      bool conv = (crit == CRIT_ABS_MAX) ? (sm[0] < tol) : (fro / bf <= tol);
      if (conv) st->done = 1;
    }
  }
}

template <class T> int launch_leaf(cx<T>* P, int64_t ld, int m, int w, int row_base, int* ipiv,
                                   int* info, const Scratch& s, int leaf, hipStream_t st) {
  const int rpw_max = leaf_rpw_max<T>();
  int G = (m + rpw_max - 1) / rpw_max;
  if (G < 1) G = 1;
  if (G > MAX_LEAF_G) return -1;
  const int rpw = (m + G - 1) / G;
  LeafXchg<T> x;
  char* base = (char*)s.xchg;
  x.val = (T*)base;
  x.idx = (int*)(base + 2 * MAX_LEAF_G * sizeof(double));
  x.rowj = (cx<T>*)(base + 2 * MAX_LEAF_G * (sizeof(double) + sizeof(int)));
  x.row = x.rowj + 2 * LEAF_W;
  size_t lds = (size_t)rpw * LEAF_W * sizeof(cx<T>);
  static bool attr = false;
  if (!attr) { allow_big_lds(leaf_getf2_kernel<T, LEAF_W>); attr = true; }
  leaf_getf2_kernel<T, LEAF_W><<<G, 256, lds, st>>>(P, ld, m, w, rpw, row_base, ipiv, info, x,
                                                    s.counters + leaf, s.err);
  return 0;
}

template <class T> void launch_laswp(cx<T>* A, int64_t ld, int c0, int c1, int k1, int nk,
                                     const int* ipiv, const int* skip, hipStream_t st) {
  if (c1 <= c0 || nk <= 0) return;
  constexpr int CW = 128 / sizeof(cx<T>);
  size_t lds = (size_t)2 * nk * 128 + (size_t)3 * nk * sizeof(int);
  static bool attr = false;
  if (!attr) { allow_big_lds(laswp_kernel<T>); attr = true; }
  int grid = (c1 - c0 + CW - 1) / CW;
  laswp_kernel<T><<<grid, 256, lds, st>>>(A, ld, c0, c1, k1, nk, ipiv, skip);
}

// Triangular solve with the kb x kb diagonal block at D (lower unit or upper
// non-unit) of the kb x ncols block row S, by 64-row diagonal solves and MFMA
// updates of the remaining rows of the block.
template <class T, bool UPPER>
void trsm_rows(const cx<T>* D, int64_t ldd, int kb, cx<T>* S, int64_t lds, int ncols, int tri,
               bool g3m, const int* skip, hipStream_t st) {
  if (ncols <= 0 || kb <= 0) return;
  const int grid = (ncols + 15) / 16;
  if (!UPPER) {
    for (int b = 0; b < kb; b += tri) {
      int bs = min(tri, kb - b);
      trsm_diag_kernel<T, false><<<grid, 256, 0, st>>>(D + (int64_t)b * ldd + b, ldd, bs,
                                                       S + (int64_t)b * lds, lds, ncols, skip);
      if (b + bs < kb)
        gemm<T>(kb - b - bs, ncols, bs, D + (int64_t)(b + bs) * ldd + b, ldd, false,
                S + (int64_t)b * lds, lds, S + (int64_t)(b + bs) * lds, lds, T(-1), 1, g3m, skip, st);
    }
  } else {
    for (int b = ((kb - 1) / tri) * tri; b >= 0; b -= tri) {
      int bs = min(tri, kb - b);
      trsm_diag_kernel<T, true><<<grid, 256, 0, st>>>(D + (int64_t)b * ldd + b, ldd, bs,
                                                      S + (int64_t)b * lds, lds, ncols, skip);
      if (b > 0)
        gemm<T>(b, ncols, bs, D + b, ldd, false, S + (int64_t)b * lds, lds, S, lds, T(-1), 1, g3m,
                skip, st);
    }
  }
}

}  // namespace

size_t scratch_bytes(int n) {
  size_t leaves = (size_t)n / LEAF_W + 8;
  size_t x = 2 * MAX_LEAF_G * (sizeof(double) + sizeof(int)) +
             (2 * LEAF_W + (size_t)2 * MAX_LEAF_G * LEAF_W) * sizeof(double2);
  x = (x + 255) & ~(size_t)255;
  return x + leaves * sizeof(unsigned) + 256;
}

int scratch_bind(Scratch& s, void* base, int n) {
  size_t leaves = (size_t)n / LEAF_W + 8;
  size_t x = 2 * MAX_LEAF_G * (sizeof(double) + sizeof(int)) +
             (2 * LEAF_W + (size_t)2 * MAX_LEAF_G * LEAF_W) * sizeof(double2);
  x = (x + 255) & ~(size_t)255;
  s.xchg = base;
  s.xchg_bytes = x;
  s.counters = (unsigned*)((char*)base + x);
  s.ncounters = leaves;
  s.err = (int*)((char*)base + x + leaves * sizeof(unsigned));
  return 0;
}

template <class T>
int gemm(int M, int N, int K, const cx<T>* A, int64_t lda, bool a_col, const cx<T>* B,
         int64_t ldb, cx<T>* C, int64_t ldc, T alpha, int beta, bool g3m, const int* skip,
         hipStream_t st) {
  if (M <= 0 || N <= 0) return 0;
  if (K <= 0) return 0;
  constexpr int BM = 64, BN = 64, BK = 16, WM = 32, WN = 32;
  int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  if (a_col) {
    if (g3m) zgemm_kernel<T, BM, BN, BK, WM, WN, true, true><<<tiles, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip);
    else zgemm_kernel<T, BM, BN, BK, WM, WN, true, false><<<tiles, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip);
  } else {
    if (g3m) zgemm_kernel<T, BM, BN, BK, WM, WN, false, true><<<tiles, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip);
    else zgemm_kernel<T, BM, BN, BK, WM, WN, false, false><<<tiles, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip);
  }
  return 0;
}

template <class T>
int getrf(cx<T>* A, int64_t ld, int n, int* ipiv, int* info, const Scratch& s,
          const LuPlan& plan, bool g3m, hipStream_t st) {
  if (n <= 0) return 0;
  const int nb = plan.nb;
  size_t nleaves = 0;
  for (int k = 0; k < n; k += nb) nleaves += (min(nb, n - k) + LEAF_W - 1) / LEAF_W;
  if (nleaves > s.ncounters) return -1;
  fill_u32_kernel<<<64, 256, 0, st>>>(s.counters, s.ncounters, 0u);
  fill_u32_kernel<<<1, 64, 0, st>>>((unsigned*)s.err, 1, 0u);
  fill_u32_kernel<<<1, 64, 0, st>>>((unsigned*)info, 1, 0u);
  int leaf = 0;
  for (int k = 0; k < n; k += nb) {
    const int kb = min(nb, n - k), m = n - k;
    for (int t = 0; t < kb; t += LEAF_W) {
      const int wt = min(LEAF_W, kb - t);
      cx<T>* P = A + (int64_t)(k + t) * ld + (k + t);
      if (launch_leaf<T>(P, ld, m - t, wt, k + t, ipiv + k + t, info, s, leaf++, st)) return -1;
      if (t > 0) launch_laswp<T>(A, ld, k, k + t, k + t, wt, ipiv, nullptr, st);
      if (t + wt < kb) {
        launch_laswp<T>(A, ld, k + t + wt, k + kb, k + t, wt, ipiv, nullptr, st);
        trsm_rows<T, false>(P, ld, wt, P + wt, ld, kb - t - wt, plan.tri, g3m, nullptr, st);
        gemm<T>(m - t - wt, kb - t - wt, wt, P + (int64_t)wt * ld, ld, false, P + wt, ld,
                P + (int64_t)wt * ld + wt, ld, T(-1), 1, g3m, nullptr, st);
      }
    }
    if (k > 0) launch_laswp<T>(A, ld, 0, k, k, kb, ipiv, nullptr, st);
    if (k + kb < n) {
      cx<T>* Akk = A + (int64_t)k * ld + k;
      launch_laswp<T>(A, ld, k + kb, n, k, kb, ipiv, nullptr, st);
      trsm_rows<T, false>(Akk, ld, kb, Akk + kb, ld, n - k - kb, plan.tri, g3m, nullptr, st);
      gemm<T>(n - k - kb, n - k - kb, kb, Akk + (int64_t)kb * ld, ld, false, Akk + kb, ld,
              Akk + (int64_t)kb * ld + kb, ld, T(-1), 1, g3m, nullptr, st);
    }
  }
  HPR_CHECK(hipGetLastError());
  return 0;
}

template <class T>
int getrs(const cx<T>* LU, int64_t ld, int n, const int* ipiv, cx<T>* B, int64_t ldb, int nrhs,
          const int* skip, const LuPlan& plan, bool g3m, hipStream_t st) {
  if (n <= 0 || nrhs <= 0) return 0;
  const int rs = plan.rs_nb;
  for (int k = 0; k < n; k += 256) launch_laswp<T>(B, ldb, 0, nrhs, k, min(256, n - k), ipiv, skip, st);
  for (int b = 0; b < n; b += rs) {
    int bs = min(rs, n - b);
    trsm_rows<T, false>(LU + (int64_t)b * ld + b, ld, bs, B + (int64_t)b * ldb, ldb, nrhs, plan.tri,
                        g3m, skip, st);
    if (b + bs < n)
      gemm<T>(n - b - bs, nrhs, bs, LU + (int64_t)(b + bs) * ld + b, ld, false, B + (int64_t)b * ldb,
              ldb, B + (int64_t)(b + bs) * ldb, ldb, T(-1), 1, g3m, skip, st);
  }
  for (int b = ((n - 1) / rs) * rs; b >= 0; b -= rs) {
    int bs = min(rs, n - b);
    trsm_rows<T, true>(LU + (int64_t)b * ld + b, ld, bs, B + (int64_t)b * ldb, ldb, nrhs, plan.tri,
                       g3m, skip, st);
    if (b > 0)
      gemm<T>(b, nrhs, bs, LU + b, ld, false, B + (int64_t)b * ldb, ldb, B, ldb, T(-1), 1, g3m,
              skip, st);
  }
  HPR_CHECK(hipGetLastError());
  return 0;
}

template <class Tin, class Tout>
int transpose_cast(int rows, int cols, const cx<Tin>* in, int64_t ldi, cx<Tout>* out, int64_t ldo,
                   hipStream_t st) {
  if (rows <= 0 || cols <= 0) return 0;
  dim3 grid((rows + 31) / 32, (cols + 31) / 32);
  transpose_cast_kernel<Tin, Tout><<<grid, 256, 0, st>>>(rows, cols, in, ldi, out, ldo);
  HPR_CHECK(hipGetLastError());
  return 0;
}

static int launch_norms(int n, int nrhs, const double2* R, double* part, RefineState* state,
                        int mode, int crit, double tol, const int* skip, hipStream_t st) {
  size_t total = (size_t)n * nrhs;
  int blocks = (int)min((size_t)NORM_BLOCKS, (total + 255) / 256);
  if (blocks < 1) blocks = 1;
  norm_partial_kernel<<<blocks, 256, 0, st>>>(n, nrhs, R, part, skip);
  norm_final_kernel<<<1, 256, 0, st>>>(blocks, part, state, mode, crit, tol);
  return 0;
}

template <class TF>
int refine_solve(const cx<TF>* LU, int64_t ld, int n, const int* ipiv, const double2* A64,
                 int64_t lda, const double2* B, double2* X, int nrhs, int max_refine, double tol,
                 int crit, RefineState* state, double2* R, cx<TF>* RF, double* part,
                 const LuPlan& plan, bool g3m, hipStream_t st) {
  const size_t cnt = (size_t)n * nrhs;
  const int eg = (int)min((size_t)4096, (cnt + 255) / 256);
  const int* done = &state->done;
  launch_norms(n, nrhs, B, part, state, 0, crit, tol, nullptr, st);
  cast_kernel<double, TF><<<eg, 256, 0, st>>>(cnt, B, RF, nullptr);
  if (getrs<TF>(LU, ld, n, ipiv, RF, nrhs, nrhs, nullptr, plan, g3m, st)) return -1;
  cast_kernel<TF, double><<<eg, 256, 0, st>>>(cnt, RF, X, nullptr);
  for (int it = 0; it < max_refine; ++it) {
    cast_kernel<double, double><<<eg, 256, 0, st>>>(cnt, B, R, done);
    gemm<double>(n, nrhs, n, A64, lda, true, X, nrhs, R, nrhs, -1.0, 1, false, done, st);
    launch_norms(n, nrhs, R, part, state, 1, crit, tol, done, st);
    cast_kernel<double, TF><<<eg, 256, 0, st>>>(cnt, R, RF, done);
    if (getrs<TF>(LU, ld, n, ipiv, RF, nrhs, nrhs, done, plan, g3m, st)) return -1;
    add_correction_kernel<TF><<<eg, 256, 0, st>>>(cnt, X, RF, state);
  }
  HPR_CHECK(hipGetLastError());
  return 0;
}

// Residual check without correction (final verification of a refined X).
int residual_check(int n, int nrhs, const double2* A64, int64_t lda, const double2* B,
                   const double2* X, double2* R, double* part, RefineState* state, int crit,
                   double tol, hipStream_t st) {
  const size_t cnt = (size_t)n * nrhs;
  const int eg = (int)min((size_t)4096, (cnt + 255) / 256);
  const int* done = &state->done;
  cast_kernel<double, double><<<eg, 256, 0, st>>>(cnt, B, R, done);
  gemm<double>(n, nrhs, n, A64, lda, true, X, nrhs, R, nrhs, -1.0, 1, false, done, st);
  launch_norms(n, nrhs, R, part, state, 1, crit, tol, done, st);
  HPR_CHECK(hipGetLastError());
  return 0;
}

#define INST(T)                                                                               \
  template int getrf<T>(cx<T>*, int64_t, int, int*, int*, const Scratch&, const LuPlan&, bool, \
                        hipStream_t);                                                          \
  template int getrs<T>(const cx<T>*, int64_t, int, const int*, cx<T>*, int64_t, int,          \
                        const int*, const LuPlan&, bool, hipStream_t);                         \
  template int gemm<T>(int, int, int, const cx<T>*, int64_t, bool, const cx<T>*, int64_t,      \
                       cx<T>*, int64_t, T, int, bool, const int*, hipStream_t);                \
  template int refine_solve<T>(const cx<T>*, int64_t, int, const int*, const double2*, int64_t, \
                               const double2*, double2*, int, int, double, int, RefineState*,   \
                               double2*, cx<T>*, double*, const LuPlan&, bool, hipStream_t);
INST(double)
INST(float)
template int transpose_cast<double, double>(int, int, const cx<double>*, int64_t, cx<double>*, int64_t, hipStream_t);
template int transpose_cast<double, float>(int, int, const cx<double>*, int64_t, cx<float>*, int64_t, hipStream_t);
template int transpose_cast<float, double>(int, int, const cx<float>*, int64_t, cx<double>*, int64_t, hipStream_t);

}  // namespace hpr
