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
//   after each panel, the inverses of its 64 x 64 diagonal blocks of L and U
//   are formed (diag_inv_kernel) and kept for the triangular solves;
//   the final row permutation is composed once at the end.
// getrs follows LAPACK zgetrs: the permutation is applied while casting the
// right-hand side, then forward (unit L) and backward (U) substitution by
// 256-row block rows: one MFMA block-solve kernel per block row using the
// stored 64 x 64 inverses, and one MFMA GEMM update of the remaining rows
// (cusolverDn?getrs, src/cusolver_interface.cu:393, 787).
#include <mutex>
#include <cmath>
#include <vector>

#include <type_traits>

#include "lu.hpp"
#include "split_gemm.hpp"
#include "gemm.hpp"
#include "lu_kernels.hpp"

namespace hpr {

namespace {

constexpr int LEAF_W = 32;
constexpr int MAX_LEAF_G = 240;               // co-resident: one leaf workgroup per CU

// Dynamic LDS above 64 KiB needs an explicit opt-in per kernel; the return
// code is checked so a refusal does not linger as the sticky last error.
template <class K> void allow_big_lds(K kernel) {
  hipError_t e = hipFuncSetAttribute((const void*)kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 152 * 1024);
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
                      cx<Tout>* __restrict__ out, int64_t ldo, const double* __restrict__ rs,
                      const double* __restrict__ cs, int64_t sIn = 0, int64_t sOut = 0,
                      int sScale = 0) {
  in += blockIdx.z * sIn;   // system of a batch
  out += blockIdx.z * sOut;
  if (rs) {
    rs += (int64_t)blockIdx.z * sScale;
    cs += (int64_t)blockIdx.z * sScale;
  }
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
      if (rs) {
        const double f = rs[r] * cs[c];
        out[(int64_t)r * ldo + c] = mkc<Tout>((Tout)(v.x * f), (Tout)(v.y * f));
      } else {
        out[(int64_t)r * ldo + c] = mkc<Tout>((Tout)v.x, (Tout)v.y);
      }
    }
  }
}

template <class Tin, class Tout>
__global__ void cast_kernel(size_t n, const cx<Tin>* __restrict__ in, cx<Tout>* __restrict__ out,
                            const int* __restrict__ skip, int64_t sV = 0, int sSkip = 0) {
  if (skip && skip[(int64_t)blockIdx.y * sSkip]) return;
  in += blockIdx.y * sV;
  out += blockIdx.y * sV;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
    cx<Tin> v = in[i];
    out[i] = mkc<Tout>((Tout)v.x, (Tout)v.y);
  }
}

// X (+)= diag(s) D for row-major n x nrhs arrays (s = null: no scaling);
// accumulate = 0 stores the first solution, 1 adds a correction (and counts
// the sweep); both are skipped once the refinement has converged.
template <class TF>
__global__ void add_correction_kernel(int nrhs, size_t cnt, double2* __restrict__ X,
                                      const cx<TF>* __restrict__ D, const double* __restrict__ s,
                                      int accumulate, RefineState* __restrict__ st,
                                      int64_t sV = 0, int sS = 0) {
  {
    const int b = blockIdx.y;   // system of a batch
    X += b * sV;
    D += b * sV;
    if (s) s += (int64_t)b * sS;
    st += b;
  }
  if (accumulate && st->done) return;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && accumulate) st->sweeps += 1;
  for (; i < cnt; i += (size_t)gridDim.x * blockDim.x) {
    cx<TF> d = D[i];
    const double f = s ? s[i / nrhs] : 1.0;
    double2 x = accumulate ? X[i] : make_double2(0.0, 0.0);
    x.x += (double)d.x * f;
    x.y += (double)d.y * f;
    X[i] = x;
  }
}

// Takes back the correction last added by add_correction_kernel (still in D)
// from a system whose refinement stopped at residual evaluation `check`
// because the residual grew (RefineState::diverged == check): X returns to
// the previous, better iterate, and the sweep count with it.
template <class TF>
__global__ void undo_correction_kernel(int nrhs, size_t cnt, double2* __restrict__ X,
                                       const cx<TF>* __restrict__ D, const double* __restrict__ s,
                                       RefineState* __restrict__ st, int check, int64_t sV = 0,
                                       int sS = 0) {
  {
    const int b = blockIdx.y;   // system of a batch
    X += b * sV;
    D += b * sV;
    if (s) s += (int64_t)b * sS;
    st += b;
  }
  if (st->diverged != check) return;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) st->sweeps -= 1;
  for (; i < cnt; i += (size_t)gridDim.x * blockDim.x) {
    cx<TF> d = D[i];
    const double f = s ? s[i / nrhs] : 1.0;
    double2 x = X[i];
    x.x -= (double)d.x * f;
    x.y -= (double)d.y * f;
    X[i] = x;
  }
}

// Symmetric equilibration factors s_i = ||row i of A||_2^(-1/2) (1 for a zero
// row) of a column-major FP64 matrix: the factored matrix is diag(s) A diag(s).
// One workgroup per 64 rows; its four waves sum every fourth column (each
// wave reads 64 consecutive rows of a column, 1 KB), the four partial sums are
// added in a fixed order, so the factors are reproducible bit for bit.
__global__ void __launch_bounds__(256)
row_scale_kernel(int n, const double2* __restrict__ A, int64_t lda, double* __restrict__ s,
                 int64_t sA = 0, int sS = 0) {
  A += blockIdx.y * sA;   // system of a batch
  s += (int64_t)blockIdx.y * sS;
  __shared__ double part[4][64];
  const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + r;
  double acc = 0.0;
  if (i < n)
    for (int j = q; j < n; j += 4) {
      const double2 a = A[(int64_t)j * lda + i];
      acc += a.x * a.x + a.y * a.y;
    }
  part[q][r] = acc;
  __syncthreads();
  if (q == 0 && i < n) {
    const double t = ((part[0][r] + part[1][r]) + part[2][r]) + part[3][r];
    s[i] = t > 0.0 ? 1.0 / sqrt(sqrt(t)) : 1.0;
  }
}

// Residual statistics, stage 1: per-block (max |Re|+|Im|, its column-major
// flat index, its modulus, sum of |r|^2) over a row-major n x nrhs array.
constexpr int NORM_BLOCKS = 1024;
__global__ void __launch_bounds__(256)
norm_partial_kernel(int n, int nrhs, const double2* __restrict__ R, double* __restrict__ part,
                    const int* __restrict__ skip, int64_t sR = 0, int sSkip = 0) {
  if (skip && skip[(int64_t)blockIdx.y * sSkip]) return;
  R += blockIdx.y * sR;   // system of a batch
  part += (size_t)blockIdx.y * gridDim.x * 5;
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
// With `guard` (classic refinement), a residual whose Frobenius norm exceeds
// the previous evaluation's stops the refinement: `diverged` records the
// evaluation, `done` is set, and the stored norms stay those of the previous
// (better) iterate, which undo_correction_kernel restores.
__global__ void __launch_bounds__(256)
norm_final_kernel(int nblocks, const double* __restrict__ part, RefineState* __restrict__ st,
                  int mode, int crit, double tol, int guard) {
  part += (size_t)blockIdx.x * nblocks * 5;   // one workgroup per system of a batch
  st += blockIdx.x;
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
      st->gm_halt = 0; st->gm_steps = 0; st->diverged = 0;
      st->res_max = 0.0; st->res_fro = 0.0; st->res_absmax = 0.0;
    } else {
      double bf = st->b_fro != 0.0 ? st->b_fro : 1.0;
      if (st->checks < 32) st->hist[st->checks] = fro / bf;
      const bool grew = guard && st->checks > 0 && fro > st->res_fro;
      st->checks += 1;
      if (grew) {
        st->diverged = st->checks;
        st->done = 1;
      } else {
        st->res_max = sm[0];
        st->res_absmax = sqrt(sa[0]);
        st->res_fro = fro;
        bool conv = (crit == CRIT_ABS_MAX) ? (sm[0] < tol) : (fro / bf <= tol);
        if (conv) st->done = 1;
      }
    }
  }
}


// Exchange area layout (u64 words): val [2][MAX_LEAF_G][2], idx [2][MAX_LEAF_G],
// row [2][MAX_LEAF_G][LEAF_W][4], rowj [2][LEAF_W][4] (sized for FP64).
constexpr size_t XW_VAL = 2 * MAX_LEAF_G * 2, XW_IDX = 2 * MAX_LEAF_G,
                 XW_ROW = (size_t)2 * MAX_LEAF_G * LEAF_W * 4, XW_ROWJ = 2 * LEAF_W * 4;
constexpr size_t XW_TOTAL = XW_VAL + XW_IDX + XW_ROW + XW_ROWJ;

template <class T> int launch_leaf(cx<T>* P, int64_t ld, int m, int w, int row_base, int* ipiv,
                                   int* info, const Scratch& s, int leaf, hipStream_t st,
                                   int prio = 0, int pivot = 1) {
  constexpr int NT = 256;   // one row per thread
  const int G = (m + NT - 1) / NT;
  if (G < 1 || G > MAX_LEAF_G) return -1;   // all leaf workgroups must be co-resident
  LeafXchg x;
  u64* base = (u64*)s.xchg;
  x.val = base;
  x.idx = x.val + XW_VAL;
  x.row = x.idx + XW_IDX;
  x.rowj = x.row + XW_ROW;
  x.gstride = MAX_LEAF_G;
  if constexpr (std::is_same<T, float>::value) {
    if (pivot == 0) {
      // tournament pivoting; its exchange words alias the candidate rows
      static_assert((size_t)(CALU_LEVELS * MAX_LEAF_G + 4 * LEAF_W) * LEAF_W <= XW_ROW,
                    "CALU exchange fits the leaf exchange area");
      CaluXchg cx_;
      cx_.sets = x.row;
      cx_.lu = cx_.sets + (size_t)CALU_LEVELS * MAX_LEAF_G * LEAF_W;
      cx_.disp = cx_.lu + (size_t)LEAF_W * LEAF_W * 2;
      cx_.gstride = MAX_LEAF_G;
      leaf_calu_kernel<LEAF_W><<<G, NT, 0, st>>>(P, ld, m, w, row_base, ipiv, info, cx_,
                                                 (unsigned)leaf * LEAF_W, s.err, prio);
      return 0;
    }
  }
  leaf_getf2_kernel<T, LEAF_W><<<G, NT, 0, st>>>(P, ld, m, w, row_base, ipiv, info, x,
                                                  (unsigned)leaf * LEAF_W, s.err, 0, 0, 0, prio);
  return 0;
}

// Interchanges k1 .. k1+nk-1 over the columns [c0, c1) and [c2, c3) (either
// may be empty) in one launch.
template <class T> void launch_laswp(cx<T>* A, int64_t ld, int c0, int c1, int k1, int nk,
                                     const int* ipiv, const int* skip, hipStream_t st,
                                     int prio = 0, int c2 = 0, int c3 = 0) {
  if (c1 < c0) c1 = c0;
  if (c3 < c2) c3 = c2;
  if ((c1 == c0 && c3 == c2) || nk <= 0) return;
  constexpr int CW = 128 / sizeof(cx<T>);
  size_t lds = (size_t)2 * nk * 128 + (size_t)4 * nk * sizeof(int);
  static bool attr = false;
  if (!attr) { allow_big_lds(laswp_kernel<T>); attr = true; }
  const int grid = (c1 - c0 + CW - 1) / CW + (c3 - c2 + CW - 1) / CW;
  laswp_kernel<T><<<grid, 256, lds, st>>>(A, ld, c0, c1, k1, nk, ipiv, skip, 0, 0, 0, prio, c2,
                                          c3);
}

// The panel's interchanges (k1, nk <= 256) as row moves, then their
// application to column ranges (laswp_map_kernel, laswp_apply_kernel).
void launch_laswp_map(const int* ipiv, int k1, int nk, int* map, hipStream_t st) {
  if (nk <= 0) return;
  laswp_map_kernel<<<1, 256, 0, st>>>(ipiv, k1, nk, map);
}
template <class T> void launch_laswp_apply(cx<T>* A, int64_t ld, int c0, int c1, const int* map,
                                           int nk, hipStream_t st, int prio = 0) {
  if (c1 <= c0 || nk <= 0) return;
  laswp_apply_kernel<T><<<(c1 - c0 + 15) / 16, 256, 0, st>>>(A, ld, c0, c1, map, 2 * nk, prio);
}

// ---- optional HIP-event timing of the trailing-update GEMM ------------------
// When enabled, each trailing update of getrf is bracketed by two events on
// the launch stream; prof_read sums the elapsed times and the algorithmic
// flops (8 M N K per complex update).
struct ProfRec {
  hipEvent_t a, b;
  double flops;
};
bool g_prof_on = false;
std::vector<ProfRec> g_prof;
size_t g_prof_used = 0;

ProfRec* prof_slot() {
  if (g_prof_used == g_prof.size()) {
    ProfRec r;
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return nullptr;
    g_prof.push_back(r);
  }
  return &g_prof[g_prof_used++];
}

size_t xchg_bytes() { return (XW_TOTAL * sizeof(u64) + 255) & ~(size_t)255; }

}  // namespace

int prof_enable(int on) {
  if (on) g_prof_used = 0;
  g_prof_on = on != 0;
  return 0;
}

int prof_read(double* ms, double* flops, int* count) {
  double t = 0.0, f = 0.0;
  for (size_t i = 0; i < g_prof_used; ++i) {
    float e = 0.f;
    HPR_CHECK(hipEventSynchronize(g_prof[i].b));
    HPR_CHECK(hipEventElapsedTime(&e, g_prof[i].a, g_prof[i].b));
    t += e;
    f += g_prof[i].flops;
  }
  *ms = t;
  *flops = f;
  *count = (int)g_prof_used;
  g_prof_used = 0;
  return 0;
}

size_t scratch_bytes() { return xchg_bytes() + 256 + 3 * 1024 * sizeof(int); }

void scratch_bind(Scratch& s, void* base) {
  s.xchg = base;
  s.xchg_bytes = xchg_bytes();
  s.err = (int*)((char*)base + s.xchg_bytes);
  s.lmap = (int*)((char*)base + s.xchg_bytes + 256);
}

template <class T>
int gemm(int M, int N, int K, const cx<T>* A, int64_t lda, bool a_col, const cx<T>* B,
         int64_t ldb, cx<T>* C, int64_t ldc, T alpha, int beta, bool g3m, const int* skip,
         hipStream_t st, size_t lds_extra, int prio) {
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  constexpr int BM = 64, BN = 64, BK = 16, WM = 32, WN = 32;
  int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  if (!a_col && lds_extra == 0 && tiles < 672) {
    // a grid of a few rounds over the CUs: 32 x 32 tiles (four times the
    // workgroups, each element summed in the same order)
    const int t32 = ((M + 31) / 32) * ((N + 31) / 32);
    if (g3m)
      zgemm_kernel<T, 32, 32, BK, 16, 16, false, true><<<t32, 256, 0, st>>>(
          M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, 0, 0, 0, 0, prio);
    else
      zgemm_kernel<T, 32, 32, BK, 16, 16, false, false><<<t32, 256, 0, st>>>(
          M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, 0, 0, 0, 0, prio);
    return 0;
  }
  // lds_extra reserves unused LDS per workgroup, which caps the number of
  // workgroups per CU and so leaves room for the concurrently running panel
  if (a_col) {
    if (g3m) zgemm_kernel<T, BM, BN, BK, WM, WN, true, true><<<tiles, 256, lds_extra, st>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, 0, 0, 0, 0, prio);
    else zgemm_kernel<T, BM, BN, BK, WM, WN, true, false><<<tiles, 256, lds_extra, st>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, 0, 0, 0, 0, prio);
  } else {
    if (g3m) zgemm_kernel<T, BM, BN, BK, WM, WN, false, true><<<tiles, 256, lds_extra, st>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, 0, 0, 0, 0, prio);
    else zgemm_kernel<T, BM, BN, BK, WM, WN, false, false><<<tiles, 256, lds_extra, st>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, 0, 0, 0, 0, prio);
  }
  return 0;
}

// Wave priority of the panel chain while the look-ahead stream updates the
// trailing matrix (HPRMAT_HIP_PANEL_PRIO): 0 none, 1 the panel kernels (leaf,
// interchanges, triangular solves, inverses), 2 also the main-stream GEMMs
// of the panel (in-panel updates, U12, the next panel's columns).
int panel_prio_level() {
  static int v = [] {
    const char* e = getenv("HPRMAT_HIP_PANEL_PRIO");
    return e ? atoi(e) : 2;
  }();
  return v;
}

// Initial value of the panel leaves' error flag.  0 in normal use;
// HPRMAT_HIP_FORCE_BARRIER_TIMEOUT=1 starts a factorization with the flag
// already raised, so every cross-workgroup wait that is not satisfied at
// once gives up (each leaf column then takes no pivot and no update, and
// every kernel ends): the host-side way to exercise the timeout report
// (stats.info = -3, batch status HPRMAT_STATUS_TIMEOUT).  Read per call.
unsigned barrier_err_init() {
  const char* e = getenv("HPRMAT_HIP_FORCE_BARRIER_TIMEOUT");
  return (e && e[0] == '1') ? 1u : 0u;
}

// Extra LDS of the look-ahead trailing update (bytes, HPRMAT_HIP_TRAIL_LDS;
// default: FP32 updates at three workgroups per CU; FP64 updates, built for
// three (150 registers), held at two, so that every CU keeps the registers
// and LDS of a panel-leaf workgroup free (176 registers, 17 KB) while the
// update runs: N = 25600 FP64 solve 965 -> 902 ms, the update at 60 instead
// of 55 TFLOP/s; three per CU starve the panel chain, 1002 ms).
size_t trail_lds_extra(bool fp32) {
  static long v = [] {
    const char* e = getenv("HPRMAT_HIP_TRAIL_LDS");
    return e ? atol(e) : -1L;
  }();
  if (v >= 0) return (size_t)v;
  return fp32 ? 32768 : 24576;
}

// ---- split-K FP64 refinement product ---------------------------------------
namespace {
constexpr int SK_BM = 64, SK_BN = 64, SK_BK = 16;

// Slice s = blockIdx.y of K: P_s = A[:, k0:k0+kn] X[k0:k0+kn, :] (A column-major).
// Built for three workgroups per CU (150 registers instead of 226 with the
// accumulators in AGPRs): when two FGMRES lanes share the GPU, the other
// lane's substitution launches then find registers free beside these
// long-running workgroups instead of waiting for one to retire (N = 25600,
// two lanes of 64 columns: f16x3_gmres 458 -> 449 ms, same iterates).
__global__ void __launch_bounds__(256, 3)
zgemm_splitk_kernel(int M, int N, int K, int Kc, const double2* __restrict__ A, int64_t lda,
                    const double2* __restrict__ B, int64_t ldb, double2* __restrict__ P,
                    int64_t sP, const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int k0 = blockIdx.y * Kc;
  const int kn = min(Kc, K - k0);
  zgemm_tile<double, SK_BM, SK_BN, SK_BK, 32, 32, true, false>(
      M, N, kn, A + (int64_t)k0 * lda, lda, B + (int64_t)k0 * ldb, ldb, P + blockIdx.y * sP, N,
      1.0, 0, blockIdx.x);
}

// C = beta C + alpha (P_0 + ... + P_{S-1}), summed in slice order.
__global__ void __launch_bounds__(256)
splitk_reduce_kernel(size_t cnt, int S, const double2* __restrict__ P, size_t sP,
                     double2* __restrict__ C, double alpha, int beta,
                     const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
       i += (size_t)gridDim.x * blockDim.x) {
    double2 acc = P[i];
    for (int s = 1; s < S; ++s) {
      const double2 v = P[(size_t)s * sP + i];
      acc.x += v.x;
      acc.y += v.y;
    }
    const double2 c = beta ? C[i] : make_double2(0.0, 0.0);
    C[i] = make_double2(c.x + alpha * acc.x, c.y + alpha * acc.y);
  }
}
}  // namespace

// The slice count fills rounds of two workgroups per CU.  It also fixes the
// summation order of the products, which moves the refinement on
// ill-conditioned systems: kept at the rule it was tuned with (rounds of
// three per CU, the kernel's residency, would take the N = 25600 benchmark
// system's Gauss-factor FGMRES from 6 to 8 products).
int resid_splits(int n, int nrhs) {
  static int slots = 0;
  if (!slots) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      return 1;
    slots = cus * 2;
  }
  if (n <= 0 || nrhs <= 0) return 1;
  const double tiles = (double)((n + SK_BM - 1) / SK_BM) * ((nrhs + SK_BN - 1) / SK_BN);
  if (tiles >= 4.0 * slots) return 1;
  // fraction of the last round of workgroups that is busy, less a small
  // charge per slice for the partial-sum traffic
  auto score = [&](int s) {
    const double rounds = tiles * s / slots;
    return rounds / std::ceil(rounds) - 0.01 * s;
  };
  int best = 1;
  for (int s = 2; s <= RESID_MAX_SPLITS; ++s) {
    if (n / s < 512) break;
    if (score(s) > score(best)) best = s;
  }
  return best;
}

int resid_gemm(int n, int nrhs, const double2* A64, int64_t lda, const double2* X, double2* R,
               double alpha, int beta, const int* skip, const SplitWork& sw, hipStream_t st) {
  if (n <= 0 || nrhs <= 0) return 0;
  if (sw.mesh) {
    if ((int64_t)sw.mesh->nch * sw.mesh->nr != n) return -1;
    return mesh_apply(*sw.mesh, nrhs, X, R, alpha, beta, skip, st);
  }
  if (sw.splits <= 1 || !sw.part)
    return gemm<double>(n, nrhs, n, A64, lda, true, X, nrhs, R, nrhs, alpha, beta, false, skip,
                        st);
  const int S = sw.splits;
  const int Kc = ((n + S - 1) / S + SK_BK - 1) / SK_BK * SK_BK;
  const size_t cnt = (size_t)n * nrhs;
  const int tiles = ((n + SK_BM - 1) / SK_BM) * ((nrhs + SK_BN - 1) / SK_BN);
  zgemm_splitk_kernel<<<dim3(tiles, S), 256, 0, st>>>(n, nrhs, n, Kc, A64, lda, X, nrhs, sw.part,
                                                      (int64_t)cnt, skip);
  splitk_reduce_kernel<<<(int)min((size_t)4096, (cnt + 255) / 256), 256, 0, st>>>(
      cnt, S, sw.part, cnt, R, alpha, beta, skip);
  HPR_CHECK(hipGetLastError());
  return 0;
}

// Split operands of one trailing update in `split` (split_gemm.hpp): L21
// (nc x kb at L, row stride ld) as packed rows, U12^T (U12: kb x nc at U, row
// stride ldu) as packed rows after them, then the scale exponents of both
// (FP16 pieces).
struct SplitView {
  half_bits* a;
  half_bits* b;
  int* ka;            // scale exponents of the rows of a / b (FP16 pieces)
  int* kb;
  unsigned* colmax;   // K scales: column maxima of L21 (256, FP16 pieces)
};
inline int split_row_halves_of(int kind) {
  return kind == SPLIT_BF16X3 ? split_row_halves<SPLIT_BF16X3>() : split_row_halves<SPLIT_F16X2>();
}
inline SplitView split_view(unsigned short* split, int nc, int kind) {
  const size_t rh = split_row_halves_of(kind);
  SplitView v;
  v.a = split;
  v.b = split + (size_t)nc * rh;
  v.ka = reinterpret_cast<int*>(split + 2 * (size_t)nc * rh);   // 4 KB aligned
  v.kb = v.ka + nc;
  v.colmax = reinterpret_cast<unsigned*>(v.kb + nc);
  return v;
}

template <int KIND>
void launch_split(int nc, int kb, const cx<float>* L, int64_t ld, const cx<float>* U, int64_t ldu,
                  const SplitView& v, hipStream_t st, bool split_l, int c0, int cols) {
  if (split_l) {
    if constexpr (SplitTraits<KIND>::SCALED) {
      // the K scales: column maxima of L21 (read by every split of this
      // step), atomic maxima over a zeroed array
      fill_u32_kernel<<<1, 256, 0, st>>>(v.colmax, 256, 0u);
      split_colmax_kernel<<<(nc + 63) / 64, 1024, 0, st>>>(nc, kb, L, ld, v.colmax);
    }
    split_rows_kernel<KIND><<<(nc + 3) / 4, 256, 0, st>>>(nc, kb, L, ld, v.a, v.ka, v.colmax);
  }
  if (cols > 0)
    split_cols_kernel<KIND><<<(cols + 31) / 32, 256, 0, st>>>(
        cols, kb, U + c0, ldu, v.b + (size_t)c0 * split_row_halves<KIND>(), v.kb + c0, v.colmax);
}

// With split_l false only U12's columns [c0, c0 + cols) are split (the packed
// rows keep the positions of all nc columns).
template <class T>
void split_operands(int nc, int kb, const cx<T>* L, int64_t ld, const cx<T>* U, int64_t ldu,
                    unsigned short* split, hipStream_t st, bool split_l, int c0, int cols,
                    int kind) {
  if constexpr (std::is_same<T, float>::value) {
    if (cols < 0) cols = nc - c0;
    const SplitView v = split_view(split, nc, kind);
    if (kind == SPLIT_F16X2)
      launch_split<SPLIT_F16X2>(nc, kb, L, ld, U, ldu, v, st, split_l, c0, cols);
    else if (kind == SPLIT_BF16X3)
      launch_split<SPLIT_BF16X3>(nc, kb, L, ld, U, ldu, v, st, split_l, c0, cols);
    else
      launch_split<SPLIT_BF16X2>(nc, kb, L, ld, U, ldu, v, st, split_l, c0, cols);
  }
}

// C[:, 0:N] -= L21 U12[:, c0:c0+N] over the split operands (M = nc rows).
// Tiles (measured at M = N = 25344): 128 x 128 with 32-K slices, two
// register sets and eight waves of 32 x 64 (two per SIMD) for the two-piece
// splits (4.9 ms against 5.2 ms with four waves of 64 x 64); 128 x 64 with
// 16-K slices and two workgroups per CU for the three-piece split (its six
// planes).
template <int KIND>
void launch_split_gemm(int M, int N, const SplitView& v, int c0, cx<float>* C, int64_t ldc,
                       hipStream_t st) {
  constexpr bool three = KIND == SPLIT_BF16X3;
  constexpr int BM = 128, BN = three ? 64 : 128, WM = three ? 64 : 32, WN = three ? 32 : 64;
  constexpr int BK = three ? 16 : 32, NSET = three ? 1 : 2, MINB = three ? 2 : 1;
  constexpr int NTH = three ? 256 : 512;
  static bool attr = false;
  if (!attr) {
    allow_big_lds(zgemm_split_kernel<KIND, BM, BN, WM, WN, BK, NSET, MINB, NTH>);
    attr = true;
  }
  const int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  zgemm_split_kernel<KIND, BM, BN, WM, WN, BK, NSET, MINB, NTH>
      <<<tiles, NTH, zgemm_split_lds_bytes<KIND, BM, BN, BK>(), st>>>(
          M, N, v.a, v.ka, v.b + (size_t)c0 * split_row_halves<KIND>(), v.kb + c0, C, ldc, -1.f);
}

// Trailing update C[:, 0:ncols] -= L21 U12[:, c0:c0+ncols] (C is the block
// at column c0 of A22), with the FP32/FP64 complex GEMM or the split-BF16
// three-product GEMM over the planes made by split_operands.
template <class T, int BM, int BN, int BK, int WM, int WN>
void launch_gemm2(int M, int N, int K, const cx<T>* A, int64_t lda, const cx<T>* B, int64_t ldb,
                  cx<T>* C, int64_t ldc, T alpha, int beta, hipStream_t st) {
  static bool attr = false;
  if (!attr) { allow_big_lds(zgemm2_kernel<T, BM, BN, BK, WM, WN, false, false>); attr = true; }
  const int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  zgemm2_kernel<T, BM, BN, BK, WM, WN, false, false>
      <<<tiles, 256, zgemm2_lds_bytes<BK, BM, BN, false, T>(), st>>>(M, N, K, A, lda, B, ldb, C,
                                                                      ldc, alpha, beta, nullptr);
}

// FP32 trailing update on 32x32x2 MFMA (zgemm32_kernel); lds_extra as in gemm().
template <int BM, int BN, int BK, int WM, int WN>
void launch_gemm32(int M, int N, int K, const cx<float>* A, int64_t lda, const cx<float>* B,
                   int64_t ldb, cx<float>* C, int64_t ldc, float alpha, int beta, hipStream_t st,
                   size_t lds_extra) {
  static bool attr = false;
  if (!attr) { allow_big_lds(zgemm32_kernel<BM, BN, BK, WM, WN, false>); attr = true; }
  const int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  const size_t lds = zgemm32_lds_bytes<BK, BM, BN>() + lds_extra;
  zgemm32_kernel<BM, BN, BK, WM, WN, false><<<tiles, 256, lds, st>>>(M, N, K, A, lda, B, ldb, C,
                                                                      ldc, alpha, beta, nullptr);
}

template <class T>
void trailing(bool bfx, int nc, int ncols, int c0, int kb, const cx<T>* L, int64_t ld,
              const cx<T>* U, int64_t ldu, cx<T>* C, int64_t ldc, bool g3m,
              const unsigned short* split, hipStream_t st, size_t lds_extra, int variant,
              int prio = 0, bool timed = false, int skind = SPLIT_BF16X2) {
  if (nc <= 0 || ncols <= 0) return;
  if constexpr (std::is_same<T, float>::value) {
    if (!g3m && !bfx && variant >= 3) {
      const cx<float>* B = U + c0;
      switch (variant) {
        case 3: launch_gemm32<64, 64, 16, 32, 32>(nc, ncols, kb, L, ld, B, ldu, C, ldc, -1.f, 1, st, lds_extra); return;
        case 4: launch_gemm32<128, 128, 16, 64, 64>(nc, ncols, kb, L, ld, B, ldu, C, ldc, -1.f, 1, st, lds_extra); return;
        case 5: launch_gemm32<128, 64, 16, 64, 32>(nc, ncols, kb, L, ld, B, ldu, C, ldc, -1.f, 1, st, lds_extra); return;
        case 6: launch_gemm32<64, 128, 16, 32, 64>(nc, ncols, kb, L, ld, B, ldu, C, ldc, -1.f, 1, st, lds_extra); return;
        default: break;
      }
    }
  }
  if (!g3m && !bfx && variant == 1) {
    launch_gemm2<T, 64, 128, 16, 32, 64>(nc, ncols, kb, L, ld, U + c0, ldu, C, ldc, T(-1), 1, st);
    return;
  }
  if (!g3m && !bfx && variant == 2) {
    launch_gemm2<T, 128, 64, 16, 64, 32>(nc, ncols, kb, L, ld, U + c0, ldu, C, ldc, T(-1), 1, st);
    return;
  }
  if constexpr (std::is_same<T, float>::value) {
    if (bfx) {
      const SplitView v = split_view(const_cast<unsigned short*>(split), nc, skind);
      if (skind == SPLIT_F16X2)
        launch_split_gemm<SPLIT_F16X2>(nc, ncols, v, c0, C, ldc, st);
      else if (skind == SPLIT_BF16X3)
        launch_split_gemm<SPLIT_BF16X3>(nc, ncols, v, c0, C, ldc, st);
      else
        launch_split_gemm<SPLIT_BF16X2>(nc, ncols, v, c0, C, ldc, st);
      return;
    }
  }
  if (timed) {
    // the same tile as gemm() under the trailing_update_kernel name
    constexpr int BM = 64, BN = 64, BK = 16, WM = 32, WN = 32;
    const int tiles = ((nc + BM - 1) / BM) * ((ncols + BN - 1) / BN);
    if (g3m)
      trailing_update_kernel<T, BM, BN, BK, WM, WN, true><<<tiles, 256, lds_extra, st>>>(
          nc, ncols, kb, L, ld, U + c0, ldu, C, ldc, T(-1));
    else
      trailing_update_kernel<T, BM, BN, BK, WM, WN, false><<<tiles, 256, lds_extra, st>>>(
          nc, ncols, kb, L, ld, U + c0, ldu, C, ldc, T(-1));
    return;
  }
  gemm<T>(nc, ncols, kb, L, ld, false, U + c0, ldu, C, ldc, T(-1), 1, g3m, nullptr, st, lds_extra,
          prio);
}

// Tile of the substitutions' block updates (M x N, K = 256): 0 = 64 x 64,
// 1 = 32 x 64, 2 = 32 x 32 (HPRMAT_HIP_SUBST_TILE forces one; -1 = by size).
// A launch takes about one tile time per round of workgroups over the 256
// CUs plus a fixed ~12 us, so small tiles win until the 64 x 64 grid holds
// several rounds: at N = 25600, nrhs = 128 the 32 x 32 tiles take an update
// of M = 8960 rows from 45.7 to 32.3 us and M = 768 from 29.7 to 15.3 us,
// while 64 x 64 stays 2-4 us ahead above about 670 of its tiles (M > 21500).
int subst_tile(int M, int N) {
  static int force = [] {
    const char* e = getenv("HPRMAT_HIP_SUBST_TILE");
    return e ? atoi(e) : -1;
  }();
  if (force >= 0 && force <= 2) return force;
  const int64_t tiles64 = (int64_t)((M + 63) / 64) * ((N + 63) / 64);
  return tiles64 >= 672 ? 0 : 2;
}

template <class T, bool LOWER, int BM2, int BN2, int WM2, int WN2>
void tri_dual_launch(const GemmProb<T>& p1, const GemmProb<T>& p2, int t1, const int* skip,
                     hipStream_t st) {
  const int t2 = (p2.M > 0 && p2.N > 0 && p2.K > 0)
                     ? ((p2.M + BM2 - 1) / BM2) * ((p2.N + BN2 - 1) / BN2) : 0;
  tri_dual_kernel<T, LOWER, BM2, BN2, WM2, WN2><<<t1 + t2, 256, 0, st>>>(p1, p2, t1, skip);
}

// D X into p1 (D a 256 x 256 triangular block inverse, p1.M = p1.K <= 256)
// and the independent product p2 (skipped when empty) in one launch.
template <class T>
int tri_dual(bool lower, const GemmProb<T>& p1, const GemmProb<T>& p2, const int* skip,
             hipStream_t st) {
  if (p1.M <= 0 || p1.N <= 0) return 0;
  if (p1.K != p1.M || p1.M > 256) return -1;
  const int t1 = ((p1.M + 31) / 32) * ((p1.N + 31) / 32);
  const int tile = subst_tile(p2.M, p2.N);
  if (lower) {
    if (tile == 2) tri_dual_launch<T, true, 32, 32, 16, 16>(p1, p2, t1, skip, st);
    else if (tile == 1) tri_dual_launch<T, true, 32, 64, 16, 32>(p1, p2, t1, skip, st);
    else tri_dual_launch<T, true, 64, 64, 32, 32>(p1, p2, t1, skip, st);
  } else {
    if (tile == 2) tri_dual_launch<T, false, 32, 32, 16, 16>(p1, p2, t1, skip, st);
    else if (tile == 1) tri_dual_launch<T, false, 32, 64, 16, 32>(p1, p2, t1, skip, st);
    else tri_dual_launch<T, false, 64, 64, 32, 32>(p1, p2, t1, skip, st);
  }
  return 0;
}

// C = beta C + alpha A B (row-major A) with the substitution update tile.
template <class T>
void subst_gemm(int M, int N, int K, const cx<T>* A, int64_t lda, const cx<T>* B, int64_t ldb,
                cx<T>* C, int64_t ldc, T alpha, int beta, const int* skip, hipStream_t st) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  const int tile = subst_tile(M, N);
  if (tile == 2)
    zgemm_kernel<T, 32, 32, 16, 16, 16, false, false><<<((M + 31) / 32) * ((N + 31) / 32), 256,
                                                        0, st>>>(M, N, K, A, lda, B, ldb, C, ldc,
                                                                 alpha, beta, skip);
  else if (tile == 1)
    zgemm_kernel<T, 32, 64, 16, 16, 32, false, false><<<((M + 31) / 32) * ((N + 63) / 64), 256,
                                                        0, st>>>(M, N, K, A, lda, B, ldb, C, ldc,
                                                                 alpha, beta, skip);
  else
    gemm<T>(M, N, K, A, lda, false, B, ldb, C, ldc, alpha, beta, false, skip, st);
}

template <class T>
GemmProb<T> prob(int M, int N, int K, const cx<T>* A, int64_t lda, const cx<T>* B, int64_t ldb,
                 cx<T>* C, int64_t ldc, T alpha, int beta) {
  GemmProb<T> p;
  p.M = M; p.N = N; p.K = K; p.A = A; p.lda = lda; p.B = B; p.ldb = ldb; p.C = C; p.ldc = ldc;
  p.alpha = alpha; p.beta = beta;
  return p;
}

// Fold the inverses of the unit-lower diagonal blocks into L once per
// factorization: for every 256-column block b, L[>b, b] <- L[>b, b] inv(L_bb),
// so that each block step of the forward substitution is one two-problem
// launch (inv(L_bb) B_b and B_>b -= W B_b read the same B_b).  L_bb has a
// unit diagonal and |l_ij| <= 1, so inv(L_bb) is well conditioned; U is not
// folded (its diagonal spans many orders of magnitude on Lagrange-mesh
// matrices, and U[<b, b] inv(U_bb) in FP32 would cost refinement sweeps).
// Row-wise transform, staged through `tmp` (256 x n elements), copied back.
// The fold is a right multiplication of whole rows, so it commutes with the
// row interchanges of later steps: the look-ahead stream folds finished block
// columns during the factorization (fold_block), and fold_inverses does the
// rest, from block column b_first on, after the last step.
template <class T>
void fold_block(cx<T>* A, int64_t ld, int n, int b0, const cx<T>* dinv256, cx<T>* tmp,
                hipStream_t st) {
  const int bs = min(256, n - b0), ml = n - b0 - bs;
  if (ml <= 0) return;
  const cx<T>* iL = dinv256 + (size_t)(b0 / 256) * 2 * 65536;
  cx<T>* Lb = A + (int64_t)(b0 + bs) * ld + b0;
  gemm<T>(ml, bs, bs, Lb, ld, false, iL, 256, tmp, bs, T(1), 0, false, nullptr, st);
  const size_t cnt = (size_t)ml * bs;
  copy2d_kernel<T><<<(int)min((size_t)4096, (cnt + 255) / 256), 256, 0, st>>>(ml, bs, tmp, bs,
                                                                               Lb, ld);
}

template <class T>
void fold_inverses(cx<T>* A, int64_t ld, int n, const cx<T>* dinv256, cx<T>* tmp, hipStream_t st,
                   int b_first = 0) {
  for (int b0 = b_first; b0 < n; b0 += 256) fold_block<T>(A, ld, n, b0, dinv256, tmp, st);
}

template <class T>
int getrf(cx<T>* A, int64_t ld, int n, int* ipiv, int* info, const FactorWork<T>& w,
          const Scratch& s, const LuPlan& plan, int gmode, hipStream_t st, const SideStream* side) {
  if (n <= 0) return 0;
  const int nb = plan.nb;
  const bool g3m = gmode == GEMM_3M;
  const bool bfx = (gmode == GEMM_BF16X3 || gmode == GEMM_BF16X6 || gmode == GEMM_F16X3) &&
                   std::is_same<T, float>::value;
  const int skind = gmode == GEMM_BF16X6  ? SPLIT_BF16X3
                    : gmode == GEMM_F16X3 ? SPLIT_F16X2
                                          : SPLIT_BF16X2;   // split_gemm.hpp kind
  if (bfx && !w.split) return -1;
  if (nb != 256) return -1;
  // pivot rule of the panel leaves (LuPlan::pivot)
  const bool tournament =
      std::is_same<T, float>::value && (plan.pivot == 2 || (plan.pivot == 0 && !g3m));
  const bool la = side && side->stream && n > 2 * nb;  // look-ahead
  const int pk = la && panel_prio_level() >= 1;   // priority of the panel kernels
  const int pg = la && panel_prio_level() >= 2;   // ... and of the main-stream GEMMs
  (void)hipGetLastError();  // start from a clean error state (checked at the end)
  fill_u32_kernel<<<64, 256, 0, st>>>((unsigned*)s.xchg, s.xchg_bytes / sizeof(unsigned), 0u);
  fill_u32_kernel<<<1, 64, 0, st>>>((unsigned*)s.err, 1, barrier_err_init());
  fill_u32_kernel<<<1, 64, 0, st>>>((unsigned*)info, 1, 0u);
  int leaf = 0;
  bool rest_pending = false;
  int folded = 0;   // block columns [0, folded) already hold L[>b, b] inv(L_bb)
  for (int k = 0; k < n; k += nb) {
    const int kb = min(nb, n - k), m = n - k;
    // -- panel: 32-column cooperative leaves, right-looking inside the panel.
    //    With look-ahead this overlaps the side stream's update of the
    //    columns beyond the panel from the previous step.
    for (int t = 0; t < kb; t += LEAF_W) {
      const int wt = min(LEAF_W, kb - t);
      cx<T>* P = A + (int64_t)(k + t) * ld + (k + t);
      if (launch_leaf<T>(P, ld, m - t, wt, k + t, ipiv + k + t, info, s, leaf++, st, pk,
                         tournament ? 0 : 1))
        return -1;
      // the leaf's interchanges on the panel's columns left and right of it
      launch_laswp<T>(A, ld, k, k + t, k + t, wt, ipiv, nullptr, st, pk, k + t + wt, k + kb);
      if (t + wt < kb) {
        trsm_diag_kernel<T, false><<<(kb - t - wt + 15) / 16, 256, 0, st>>>(
            P, ld, wt, P + wt, ld, kb - t - wt, nullptr, 0, pk);
        gemm<T>(m - t - wt, kb - t - wt, wt, P + (int64_t)wt * ld, ld, false, P + wt, ld,
                P + (int64_t)wt * ld + wt, ld, T(-1), 1, g3m, nullptr, st, 0, pg);
      }
    }
    const int nc = n - k - kb;
    cx<T>* Akk = A + (int64_t)k * ld + k;
    const cx<T>* iL = w.dinv256 + (size_t)(k / 256) * 2 * 65536;
    // the panel's interchanges as row moves, for every column range outside
    // the panel.  Three maps in turn: the last reader of a map, this step's
    // left interchanges on the look-ahead stream, precedes that stream's next
    // bulk update, whose `rest` event the main stream waits for before it
    // writes the map again two steps later.
    int* lmap = s.lmap + (size_t)((k / nb) % 3) * 1024;
    launch_laswp_map(ipiv, k, kb, lmap, st);
    // Inverses of the panel's diagonal blocks (64 x 64, then 256 x 256):
    // inv(L11) now, it is needed for U12; inv(U11) only by the solves, so
    // those are formed for every block in one pass after the last step
    // (U11 is final once its panel is factored: later interchanges only
    // touch rows below it), off the update chain.
    diag_inv_kernel<T><<<dim3((kb + 63) / 64, 1), 64, 0, st>>>(A, ld, n, k / 64, w.dinv64, 0, 0,
                                                                pk, 0);
    blk_inv_kernel<T><<<dim3(1, 1, 16), 256, 0, st>>>(A, ld, n, k / 256, w.dinv64, w.dinv256, 0, 0,
                                                     0, pk, 0);
    // U12 = inv(L11) A12 for the trailing columns [c0, c0 + cols) of the
    // step, staged in ustage (row stride nc), then C -= L21 U12 over them
    // and U12 copied back into A's rows k .. k + kb.
    auto u12 = [&](int c0, int cols, hipStream_t s_, int prio_) {
      gemm<T>(kb, cols, kb, iL, 256, false, Akk + kb + c0, ld, w.ustage + c0, nc, T(1), 0, g3m,
              nullptr, s_, 0, prio_);
    };
    auto copy_back = [&](int c0, int cols, hipStream_t s_) {
      const size_t cnt = (size_t)kb * cols;
      copy2d_kernel<T><<<(int)min((size_t)4096, (cnt + 255) / 256), 256, 0, s_>>>(
          kb, cols, w.ustage + c0, nc, Akk + kb + c0, ld);
    };
    if (la && nc > nb) {
      // Look-ahead step.  Main stream: the next panel's columns only (swap,
      // U12, update, copy back), so that the next panel starts as early as
      // possible.  Side stream: the swap, U12 and update of every column
      // beyond the next panel.
      // (split-BF16: the side stream also reads the L21 planes made below, so
      // it starts after them)
      if (!bfx) HPR_CHECK(hipEventRecord(side->ready, st));
      // the next panel's columns were last updated on the side stream, and
      // ustage / the split planes were last read there
      if (rest_pending) {
        HPR_CHECK(hipStreamWaitEvent(st, side->rest, 0));
        rest_pending = false;
      }
      launch_laswp_apply<T>(A, ld, k + kb, k + 2 * kb, lmap, kb, st, pk);
      u12(0, nb, st, pg);
      if (bfx) {
        split_operands<T>(nc, kb, Akk + (int64_t)kb * ld, ld, w.ustage, nc, w.split, st, true, 0,
                          nb, skind);
        HPR_CHECK(hipEventRecord(side->ready, st));
      }
      trailing<T>(bfx, nc, nb, 0, kb, Akk + (int64_t)kb * ld, ld, w.ustage, nc,
                  Akk + (int64_t)kb * ld + kb, ld, g3m, w.split, st, 0, plan.trail, pg, false,
                  skind);
      copy_back(0, nb, st);
      hipStream_t cs = side->stream;
      HPR_CHECK(hipStreamWaitEvent(cs, side->ready, 0));
      launch_laswp_apply<T>(A, ld, k + 2 * kb, n, lmap, kb, cs);
      u12(nb, nc - nb, cs, 0);
      if (bfx)
        split_operands<T>(nc, kb, Akk + (int64_t)kb * ld, ld, w.ustage, nc, w.split, cs, false,
                          nb, nc - nb, skind);
      ProfRec* pr2 = g_prof_on ? prof_slot() : nullptr;
      if (pr2) (void)hipEventRecord(pr2->a, cs);
      trailing<T>(bfx, nc, nc - nb, nb, kb, Akk + (int64_t)kb * ld, ld, w.ustage, nc,
                  Akk + (int64_t)kb * ld + kb + nb, ld, g3m, w.split, cs,
                  trail_lds_extra(sizeof(T) == 4), plan.trail, 0, true, skind);
      if (pr2) {
        (void)hipEventRecord(pr2->b, cs);
        pr2->flops = 8.0 * (double)nc * (double)(nc - nb) * (double)kb;
      }
      copy_back(nb, nc - nb, cs);
      HPR_CHECK(hipEventRecord(side->rest, cs));
      rest_pending = true;
      // the finished L columns left of the panel are not read again until
      // the solves: their interchanges follow the update on the side stream,
      // after the event the main stream waits for (run concurrently with the
      // update they would be starved by its workgroups)
      if (k > 0) launch_laswp_apply<T>(A, ld, 0, k, lmap, kb, cs);
      // in the second half, where the panel chain outlasts the update, this
      // stream has slack: fold inv(L_bb) into up to two finished block
      // columns b0 < k per step (the current panel's column is still read
      // by this step's updates)
      if (w.ftmp && 2 * nc <= n)
        for (int f = 0; f < 2 && folded + 256 <= k; ++f, folded += 256)
          fold_block<T>(A, ld, n, folded, w.dinv256, w.ftmp, cs);
    } else {
      // no look-ahead (or the last steps): everything on the main stream,
      // after all side-stream work (the left interchanges included)
      if (rest_pending) {
        HPR_CHECK(hipEventRecord(side->rest, side->stream));
        HPR_CHECK(hipStreamWaitEvent(st, side->rest, 0));
        rest_pending = false;
      }
      if (k > 0) launch_laswp_apply<T>(A, ld, 0, k, lmap, kb, st, pk);
      if (nc > 0) {
        launch_laswp_apply<T>(A, ld, k + kb, n, lmap, kb, st, pk);
        u12(0, nc, st, pg);
        ProfRec* pr = g_prof_on ? prof_slot() : nullptr;
        if (bfx)
          split_operands<T>(nc, kb, Akk + (int64_t)kb * ld, ld, w.ustage, nc, w.split, st, true,
                            0, -1, skind);
        if (pr) (void)hipEventRecord(pr->a, st);
        trailing<T>(bfx, nc, nc, 0, kb, Akk + (int64_t)kb * ld, ld, w.ustage, nc,
                    Akk + (int64_t)kb * ld + kb, ld, g3m, w.split, st, 0, plan.trail, 0, true,
                    skind);
        if (pr) {
          (void)hipEventRecord(pr->b, st);
          pr->flops = 8.0 * (double)nc * (double)nc * (double)kb;
        }
        copy_back(0, nc, st);
      }
    }
  }
  if (rest_pending) {   // all side-stream work, the left interchanges included
    HPR_CHECK(hipEventRecord(side->rest, side->stream));
    HPR_CHECK(hipStreamWaitEvent(st, side->rest, 0));
  }
  // the final row permutation needs only ipiv (one workgroup replaying the
  // interchanges, 1.8 ms at n = 25600): on the look-ahead stream, beside the
  // inverses and folds below
  hipStream_t ps = st;
  if (side) {
    HPR_CHECK(hipEventRecord(side->ready, st));
    HPR_CHECK(hipStreamWaitEvent(side->stream, side->ready, 0));
    ps = side->stream;
  }
  const size_t perm_lds = (size_t)n * sizeof(int);
  const int in_lds = perm_lds <= 140 * 1024;
  static bool attr = false;
  if (!attr) { allow_big_lds(perm_from_ipiv_kernel); attr = true; }
  perm_from_ipiv_kernel<<<1, 256, in_lds ? perm_lds : 0, ps>>>(ipiv, n, w.perm, in_lds);
  // inv(U_bb) of every diagonal block: 64 x 64 blocks, then 256 x 256
  diag_inv_kernel<T><<<dim3((n + 63) / 64, 1), 64, 0, st>>>(A, ld, n, 0, w.dinv64, 0, 0, 0, 1);
  blk_inv_kernel<T><<<dim3((n + 255) / 256, 1, 16), 256, 0, st>>>(A, ld, n, 0, w.dinv64, w.dinv256,
                                                                  0, 0, 0, 0, 1);
  fold_inverses<T>(A, ld, n, w.dinv256, w.ustage, st, folded);
  if (side) {
    HPR_CHECK(hipEventRecord(side->rest, side->stream));
    HPR_CHECK(hipStreamWaitEvent(st, side->rest, 0));
  }
  HPR_CHECK(hipGetLastError());
  return 0;
}

template <class T>
int getrs(const cx<T>* LU, int64_t ld, int n, const cx<T>* dinv256, cx<T>* B, int64_t ldb,
          int nrhs, cx<T>* Y, const int* skip, const LuPlan& plan, bool g3m, hipStream_t st) {
  if (n <= 0 || nrhs <= 0) return 0;
  const int rs = 256;
  // L holds W = L[>b, b] inv(L_bb) (see fold_inverses).  Forward:
  // Y_b = inv(L_bb) B_b and B_>b -= W B_b, one launch per block; backward:
  // B_b = inv(U_bb) Y_b, then Y_<b -= U[<b, b] B_b.
  for (int b = 0; b < n; b += rs) {
    const int bs = min(rs, n - b);
    const cx<T>* iL = dinv256 + (size_t)(b / 256) * 2 * 65536;
    const GemmProb<T> p1 = prob<T>(bs, nrhs, bs, iL, 256, B + (int64_t)b * ldb, ldb,
                                   Y + (int64_t)b * ldb, ldb, T(1), 0);
    const GemmProb<T> p2 = prob<T>(n - b - bs, nrhs, bs, LU + (int64_t)(b + bs) * ld + b, ld,
                                   B + (int64_t)b * ldb, ldb, B + (int64_t)(b + bs) * ldb, ldb,
                                   T(-1), 1);
    if (g3m) {
      gemm<T>(p1.M, p1.N, p1.K, p1.A, p1.lda, false, p1.B, p1.ldb, p1.C, p1.ldc, p1.alpha,
              p1.beta, true, skip, st);
      gemm<T>(p2.M, p2.N, p2.K, p2.A, p2.lda, false, p2.B, p2.ldb, p2.C, p2.ldc, p2.alpha,
              p2.beta, true, skip, st);
    } else {
      if (tri_dual<T>(true, p1, p2, skip, st)) return -1;
    }
  }
  for (int b = ((n - 1) / rs) * rs; b >= 0; b -= rs) {
    const int bs = min(rs, n - b);
    const cx<T>* iU = dinv256 + ((size_t)(b / 256) * 2 + 1) * 65536;
    if (g3m)
      gemm<T>(bs, nrhs, bs, iU, 256, false, Y + (int64_t)b * ldb, ldb, B + (int64_t)b * ldb, ldb,
              T(1), 0, true, skip, st);
    else if (tri_dual<T>(false, prob<T>(bs, nrhs, bs, iU, 256, Y + (int64_t)b * ldb, ldb,
                                        B + (int64_t)b * ldb, ldb, T(1), 0),
                         prob<T>(0, 0, 0, iU, 256, Y, ldb, B, ldb, T(1), 0), skip, st))
      return -1;
    if (b > 0) {
      if (g3m)
        gemm<T>(b, nrhs, bs, LU + b, ld, false, B + (int64_t)b * ldb, ldb, Y, ldb, T(-1), 1, true,
                skip, st);
      else
        subst_gemm<T>(b, nrhs, bs, LU + b, ld, B + (int64_t)b * ldb, ldb, Y, ldb, T(-1), 1, skip,
                      st);
    }
  }
  HPR_CHECK(hipGetLastError());
  return 0;
}

void subst_graph_release(SubstGraph& g) {
  if (g.exec) {
    (void)hipStreamSynchronize(g.last);
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGetLastError();
  }
  g = SubstGraph();
}

// One non-blocking stream per device, used only to record substitution
// graphs (nothing is ever executed on it).  Created on first use, kept until
// the process ends.
static hipStream_t capture_stream() {
  static std::mutex mu;
  static hipStream_t streams[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!streams[dev] && hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    streams[dev] = nullptr;
  }
  return streams[dev];
}

template <class T>
int getrs_replay(SubstGraph* g, const cx<T>* LU, int64_t ld, int n, const cx<T>* dinv256,
                 cx<T>* B, int64_t ldb, int nrhs, cx<T>* Y, const int* skip, const LuPlan& plan,
                 bool g3m, hipStream_t st) {
  const char* env = getenv("HPRMAT_HIP_SUBST_GRAPH");   // read per call (tests toggle it)
  const bool enabled = !(env && env[0] == '0');
  if (!g || !enabled || n <= 0 || nrhs <= 0)
    return getrs<T>(LU, ld, n, dinv256, B, ldb, nrhs, Y, skip, plan, g3m, st);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (st && (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)) {
    (void)hipGetLastError();   // the caller is capturing: its graph takes the launches
    return getrs<T>(LU, ld, n, dinv256, B, ldb, nrhs, Y, skip, plan, g3m, st);
  }
  if (!g->exec || g->LU != LU || g->dinv != dinv256 || g->B != B || g->Y != Y ||
      g->skip != skip || g->n != n || g->nrhs != nrhs || g->ld != ld || g->ldb != ldb ||
      g->g3m != g3m || g->elem != (int)sizeof(cx<T>)) {
    subst_graph_release(*g);
    // The launches are recorded on a private stream of this device (the
    // caller's stream may be the null stream, which cannot be captured); the
    // instantiated graph is then launched on the caller's stream.  Host
    // threads (the FGMRES lanes) take turns at the recording.
    static std::mutex cap_mu;
    std::lock_guard<std::mutex> cap_lock(cap_mu);
    hipStream_t cap = capture_stream();
    hipGraph_t graph = nullptr;
    if (!cap || hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      return getrs<T>(LU, ld, n, dinv256, B, ldb, nrhs, Y, skip, plan, g3m, st);
    }
    const int rc = getrs<T>(LU, ld, n, dinv256, B, ldb, nrhs, Y, skip, plan, g3m, cap);
    const hipError_t e = hipStreamEndCapture(cap, &graph);
    if (rc || e != hipSuccess || !graph) {
      if (graph) (void)hipGraphDestroy(graph);
      (void)hipGetLastError();
      return getrs<T>(LU, ld, n, dinv256, B, ldb, nrhs, Y, skip, plan, g3m, st);
    }
    const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {   // launch the kernels directly instead
      g->exec = nullptr;
      (void)hipGetLastError();
      return getrs<T>(LU, ld, n, dinv256, B, ldb, nrhs, Y, skip, plan, g3m, st);
    }
    g->LU = LU;
    g->dinv = dinv256;
    g->B = B;
    g->Y = Y;
    g->skip = skip;
    g->n = n;
    g->nrhs = nrhs;
    g->ld = ld;
    g->ldb = ldb;
    g->g3m = g3m;
    g->elem = (int)sizeof(cx<T>);
  }
  HPR_CHECK(hipGraphLaunch(g->exec, st));
  g->last = st;
  return 0;
}

template <class Tin, class Tout>
int permute_cast(int n, int nrhs, const int* perm, const cx<Tin>* in, cx<Tout>* out,
                 const int* skip, hipStream_t st, const double* scale) {
  const size_t cnt = (size_t)n * nrhs;
  if (cnt == 0) return 0;
  const int grid = (int)min((size_t)4096, (cnt + 255) / 256);
  permute_cast_kernel<Tin, Tout><<<grid, 256, 0, st>>>(n, nrhs, perm, in, out, scale, skip);
  return 0;
}

int equil_scale(int n, const double2* A, int64_t lda, double* s, hipStream_t st) {
  if (n <= 0) return 0;
  row_scale_kernel<<<(n + 63) / 64, 256, 0, st>>>(n, A, lda, s);
  HPR_CHECK(hipGetLastError());
  return 0;
}

template <class Tin, class Tout>
int transpose_cast(int rows, int cols, const cx<Tin>* in, int64_t ldi, cx<Tout>* out, int64_t ldo,
                   hipStream_t st, const double* rs, const double* cs) {
  if (rows <= 0 || cols <= 0) return 0;
  dim3 grid((rows + 31) / 32, (cols + 31) / 32);
  transpose_cast_kernel<Tin, Tout><<<grid, 256, 0, st>>>(rows, cols, in, ldi, out, ldo, rs, cs);
  HPR_CHECK(hipGetLastError());
  return 0;
}

int residual_norms(int n, int nrhs, const double2* R, double* part, RefineState* state, int mode,
                   int crit, double tol, const int* skip, hipStream_t st, int guard) {
  size_t total = (size_t)n * nrhs;
  int blocks = (int)min((size_t)NORM_BLOCKS, (total + 255) / 256);
  if (blocks < 1) blocks = 1;
  norm_partial_kernel<<<blocks, 256, 0, st>>>(n, nrhs, R, part, skip);
  norm_final_kernel<<<1, 256, 0, st>>>(blocks, part, state, mode, crit, tol, guard);
  return 0;
}

FlagPoll flag_poll(hipStream_t st, bool allowed) {
  FlagPoll p;
  if (!allowed) return p;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return p;
  }
  p.on = cs == hipStreamCaptureStatusNone;
  return p;
}

// The flag after the work queued so far (0 when polling is off or the copy
// fails: the device flags still gate the remaining launches).
int FlagPoll::read(const RefineState* st, hipStream_t s, int RefineState::*field) const {
  if (!on) return 0;
  int v = 0;
  if (hipMemcpyAsync(&v, &(st->*field), sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return v;
}

template <class TF>
int refine_solve(const cx<TF>* LU, int64_t ld, int n, const cx<TF>* dinv256, const int* perm,
                 const double2* A64, int64_t lda, const double2* B, double2* X, int nrhs,
                 int max_refine, double tol, int crit, RefineState* state, double2* R, cx<TF>* RF,
                 cx<TF>* RY, double* part, const LuPlan& plan, bool g3m, hipStream_t st,
                 const double* scale, const SplitWork& sw, bool poll, bool final_check) {
  const size_t cnt = (size_t)n * nrhs;
  const int eg = (int)min((size_t)4096, (cnt + 255) / 256);
  const int* done = &state->done;
  residual_norms(n, nrhs, B, part, state, 0, crit, tol, nullptr, st);
  permute_cast<double, TF>(n, nrhs, perm, B, RF, nullptr, st, scale);
  if (getrs<TF>(LU, ld, n, dinv256, RF, nrhs, nrhs, RY, nullptr, plan, g3m, st)) return -1;
  add_correction_kernel<TF><<<eg, 256, 0, st>>>(nrhs, cnt, X, RF, scale, 0, state);
  const FlagPoll fp = flag_poll(st, poll);
  // Evaluation `it` checks the iterate after `it` corrections; the last one
  // (it == max_refine) runs only with final_check.  A residual that grows
  // stops the sweeps and takes the last correction back
  // (RefineState::diverged): the solve returns its best iterate.
  for (int it = 0; it <= max_refine; ++it) {
    if (it == max_refine && !final_check) break;
    cast_kernel<double, double><<<eg, 256, 0, st>>>(cnt, B, R, done);
    if (resid_gemm(n, nrhs, A64, lda, X, R, -1.0, 1, done, sw, st)) return -1;
    residual_norms(n, nrhs, R, part, state, 1, crit, tol, done, st, 1);
    if (it > 0)
      undo_correction_kernel<TF><<<eg, 256, 0, st>>>(nrhs, cnt, X, RF, scale, state, it + 1);
    if (it == max_refine) break;
    if (fp.read(state, st, &RefineState::done)) break;
    permute_cast<double, TF>(n, nrhs, perm, R, RF, done, st, scale);
    if (getrs<TF>(LU, ld, n, dinv256, RF, nrhs, nrhs, RY, done, plan, g3m, st)) return -1;
    add_correction_kernel<TF><<<eg, 256, 0, st>>>(nrhs, cnt, X, RF, scale, 1, state);
  }
  HPR_CHECK(hipGetLastError());
  return 0;
}

// ===========================================================================
// Uniform batches: `count` systems of one order n, factored and solved in
// lockstep.  Every launch carries the system index in a grid dimension, so
// one chain of kernels serves the whole batch: a system of a few hundred to
// a few thousand unknowns is a chain of 10^2 - 10^3 small dependent kernels
// whose dispatch latency, not their work, dominates when systems run one at
// a time.  No look-ahead (the batch fills the GPU).  The cooperative panel
// leaf of every system (G = ceil(m / 256) workgroups each) is launched in
// one grid of up to BATCH_LEAF_WG workgroups: the workgroups of one system
// are consecutive in dispatch order and a system waits only on its own
// workgroups, so the earliest unfinished system always has all of its
// workgroups resident (workgroups of later systems queue behind it), as long
// as G leaf workgroups of each concurrently running batch fit on the GPU at
// once (G <= 25 for n <= 6400; about 512 FP64 leaf workgroups are resident).
// ===========================================================================
namespace {
constexpr int BATCH_LEAF_WG = 4096;

inline int64_t up32(int64_t x) { return (x + 31) & ~(int64_t)31; }
inline size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

// exchange words of one system's leaf: val [2][G][2], idx [2][G],
// row [2][G][LEAF_W][4], rowj [2][LEAF_W][4] (G = leaf workgroups)
inline int batch_leaf_g(int n) { return n > 256 ? (n + 255) / 256 : 1; }
inline int64_t batch_xchg_words(int n) {
  const int64_t g = batch_leaf_g(n);
  return up32(4 * g + 2 * g + 2 * g * LEAF_W * 4 + 2 * LEAF_W * 4);
}

// Tiles of a batched product (row-major A): 0 = 64 x 64 (waves of 32 x 32),
// 1 = 32 x 32 (waves of 16 x 16: four times the workgroups, for small
// systems), 2 = 128 x 64 (waves of 64 x 32).
template <class T, bool G3M>
void gemm_b_tiled(int tile, int M, int N, int K, const cx<T>* A, int64_t lda, const cx<T>* B,
                  int64_t ldb, cx<T>* C, int64_t ldc, T alpha, int beta, const int* skip,
                  int sSkip, int count, int64_t sA, int64_t sB, int64_t sC, hipStream_t st) {
  if (tile == 1) {
    const dim3 g(((M + 31) / 32) * ((N + 31) / 32), count);
    zgemm_kernel<T, 32, 32, 16, 16, 16, false, G3M><<<g, 256, 0, st>>>(
        M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, sA, sB, sC, sSkip);
  } else if (tile == 2) {
    const dim3 g(((M + 127) / 128) * ((N + 63) / 64), count);
    zgemm_kernel<T, 128, 64, 16, 64, 32, false, G3M><<<g, 256, 0, st>>>(
        M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, sA, sB, sC, sSkip);
  } else {
    const dim3 g(((M + 63) / 64) * ((N + 63) / 64), count);
    zgemm_kernel<T, 64, 64, 16, 32, 32, false, G3M><<<g, 256, 0, st>>>(
        M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, sA, sB, sC, sSkip);
  }
}

template <class T>
void gemm_b(int M, int N, int K, const cx<T>* A, int64_t lda, bool a_col, const cx<T>* B,
            int64_t ldb, cx<T>* C, int64_t ldc, T alpha, int beta, const int* skip, int sSkip,
            int count, int64_t sA, int64_t sB, int64_t sC, hipStream_t st, bool g3m = false,
            int tile = 0) {
  if (M <= 0 || N <= 0 || K <= 0 || count <= 0) return;
  if (tile != 0 && !a_col) {
    if (g3m)
      gemm_b_tiled<T, true>(tile, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, sSkip,
                            count, sA, sB, sC, st);
    else
      gemm_b_tiled<T, false>(tile, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, sSkip,
                             count, sA, sB, sC, st);
    return;
  }
  constexpr int BM = 64, BN = 64, BK = 16, WM = 32, WN = 32;
  if (N <= 32 && !g3m) {
    // narrow products (a batch's right-hand sides): 64 x 32 tiles, no idle
    // half tile, each element summed in the same order
    const dim3 g32(((M + 63) / 64) * ((N + 31) / 32), count);
    if (a_col)
      zgemm_kernel<T, 64, 32, BK, 32, 16, true, false><<<g32, 256, 0, st>>>(
          M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, sA, sB, sC, sSkip);
    else
      zgemm_kernel<T, 64, 32, BK, 32, 16, false, false><<<g32, 256, 0, st>>>(
          M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, sA, sB, sC, sSkip);
    return;
  }
  const dim3 grid(((M + BM - 1) / BM) * ((N + BN - 1) / BN), count);
  if (a_col)
    zgemm_kernel<T, BM, BN, BK, WM, WN, true, false><<<grid, 256, 0, st>>>(
        M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, sA, sB, sC, sSkip);
  else if (g3m)
    zgemm_kernel<T, BM, BN, BK, WM, WN, false, true><<<grid, 256, 0, st>>>(
        M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, sA, sB, sC, sSkip);
  else
    zgemm_kernel<T, BM, BN, BK, WM, WN, false, false><<<grid, 256, 0, st>>>(
        M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, skip, sA, sB, sC, sSkip);
}

template <class T>
void laswp_b(const BatchArgs<T>& b, int c0, int c1, int k1, int nk, hipStream_t st) {
  if (c1 <= c0 || nk <= 0) return;
  constexpr int CW = 128 / sizeof(cx<T>);
  const size_t lds = (size_t)2 * nk * 128 + (size_t)4 * nk * sizeof(int);
  static bool attr = false;
  if (!attr) { allow_big_lds(laswp_kernel<T>); attr = true; }
  const dim3 grid((c1 - c0 + CW - 1) / CW, b.count);
  laswp_kernel<T><<<grid, 256, lds, st>>>(b.F, b.ld, c0, c1, k1, nk, b.ipiv, nullptr, b.sF,
                                          b.sIv, 0);
}

inline dim3 copy_grid(size_t cnt, int count) {
  return dim3((unsigned)min((size_t)1024, (cnt + 255) / 256), (unsigned)count);
}

void norms_b(int n, int nrhs, const double2* R, int64_t sR, double* part, RefineState* state,
             int mode, int crit, double tol, const int* skip, int sSkip, int count,
             hipStream_t st, int guard = 0) {
  const size_t total = (size_t)n * nrhs;
  int blocks = (int)min((size_t)BATCH_NORM_BLOCKS, (total + 255) / 256);
  if (blocks < 1) blocks = 1;
  norm_partial_kernel<<<dim3(blocks, count), 256, 0, st>>>(n, nrhs, R, part, skip, sR, sSkip);
  norm_final_kernel<<<count, 256, 0, st>>>(blocks, part, state, mode, crit, tol, guard);
}

// per-system outcome: relative residual of the last evaluation and
// 1 converged / 0 not converged / -info (exactly singular pivot), or
// STATUS_TIMEOUT for every system when a panel leaf's bounded wait gave up
// (the factors are then invalid)
__global__ void batch_result_kernel(int count, const RefineState* __restrict__ st,
                                    const int* __restrict__ info, const int* __restrict__ err,
                                    double* __restrict__ res, int* __restrict__ conv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double bf = st[i].b_fro != 0.0 ? st[i].b_fro : 1.0;
  if (res) res[i] = st[i].res_fro / bf;
  if (conv)
    conv[i] = *err != 0     ? STATUS_TIMEOUT
              : info[i] != 0 ? -info[i]
                             : (st[i].done && !st[i].diverged ? 1 : 0);
}
}  // namespace

size_t batch_work_bytes(int n, int nrhs, int count, size_t elem) {
  if (n <= 0 || nrhs <= 0 || count <= 0) return 0;
  const int64_t ld = (((int64_t)n + 15) / 16) * 16 + 16;
  const int64_t sIv = (n + 1 + 63) & ~63;
  const int64_t sV = up32((int64_t)n * nrhs);
  size_t per = 0;
  per += (size_t)up32(ld * n) * elem;                                   // factors
  per += (dinv64_elems(n) + dinv256_elems(n)) * elem;                  // inverses
  per += (size_t)up32((int64_t)ustage_elems(n)) * elem;                 // U12 staging
  per += (size_t)sIv * (2 * sizeof(int) + sizeof(double));             // ipiv, perm, scale
  per += (size_t)batch_xchg_words(n) * sizeof(u64);                    // leaf exchange
  per += sizeof(RefineState) + sizeof(int) + BATCH_NORM_BLOCKS * 5 * sizeof(double);
  per += (size_t)sV * (3 * sizeof(double2) + 2 * elem);                // Brm, Xrm, R, RF, RY
  return per * (size_t)count + 16 * 256;                               // + alignment, err
}

template <class T>
int batch_bind(BatchArgs<T>& b, void* base, size_t bytes, int n, int nrhs, int count, bool equil) {
  const size_t e = sizeof(cx<T>);
  if (!base || n <= 0 || nrhs <= 0 || count <= 0) return -1;
  if (batch_work_bytes(n, nrhs, count, e) > bytes) return -2;
  char* p = (char*)base;
  auto take = [&](size_t total) {
    p = (char*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
    char* q = p;
    p += total;
    return q;
  };
  b.n = n;
  b.count = count;
  b.ld = (((int64_t)n + 15) / 16) * 16 + 16;
  b.sF = up32(b.ld * n);
  b.s64 = (int64_t)dinv64_elems(n);
  b.s256 = (int64_t)dinv256_elems(n);
  b.sU = up32((int64_t)ustage_elems(n));
  b.sIv = (n + 1 + 63) & ~63;
  b.sX = batch_xchg_words(n);
  b.gstride = batch_leaf_g(n);
  b.sV = up32((int64_t)n * nrhs);
  b.F = (cx<T>*)take((size_t)b.sF * count * e);
  b.dinv64 = (cx<T>*)take((size_t)b.s64 * count * e);
  b.dinv256 = (cx<T>*)take((size_t)b.s256 * count * e);
  b.ustage = (cx<T>*)take((size_t)b.sU * count * e);
  b.ipiv = (int*)take((size_t)b.sIv * count * sizeof(int));
  b.perm = (int*)take((size_t)b.sIv * count * sizeof(int));
  double* sc = (double*)take((size_t)b.sIv * count * sizeof(double));
  b.scale = equil ? sc : nullptr;
  b.xchg = (u64*)take((size_t)b.sX * count * sizeof(u64));
  b.state = (RefineState*)take(sizeof(RefineState) * count);
  b.info = (int*)take(sizeof(int) * count);
  b.part = (double*)take((size_t)BATCH_NORM_BLOCKS * 5 * count * sizeof(double));
  b.Brm = (double2*)take((size_t)b.sV * count * sizeof(double2));
  b.Xrm = (double2*)take((size_t)b.sV * count * sizeof(double2));
  b.R = (double2*)take((size_t)b.sV * count * sizeof(double2));
  b.RF = (cx<T>*)take((size_t)b.sV * count * e);
  b.RY = (cx<T>*)take((size_t)b.sV * count * e);
  b.err = (int*)take(sizeof(int));
  return (size_t)(p - (char*)base) <= bytes ? 0 : -2;
}

template <class T>
int getrf_batched(const BatchArgs<T>& b, hipStream_t st) {
  const int n = b.n, count = b.count;
  if (n <= 0 || count <= 0) return 0;
  const int64_t ld = b.ld, sF = b.sF;
  (void)hipGetLastError();   // start from a clean error state (checked at the end)
  fill_u32_kernel<<<1024, 256, 0, st>>>((unsigned*)b.xchg, (size_t)b.sX * 2 * count, 0u);
  fill_u32_kernel<<<1, 64, 0, st>>>((unsigned*)b.err, 1, barrier_err_init());
  fill_u32_kernel<<<(count + 255) / 256, 256, 0, st>>>((unsigned*)b.info, (size_t)count, 0u);
  int leaf = 0;
  for (int k = 0; k < n; k += 256) {
    const int kb = min(256, n - k), m = n - k;
    for (int t = 0; t < kb; t += LEAF_W) {
      const int wt = min(LEAF_W, kb - t);
      cx<T>* P = b.F + (int64_t)(k + t) * ld + (k + t);
      const int G = (m - t + 255) / 256;
      if (G > b.gstride) return -1;
      const int chunk = max(1, BATCH_LEAF_WG / G);
      for (int c0 = 0; c0 < count; c0 += chunk) {
        const int cn = min(chunk, count - c0);
        LeafXchg x;
        x.val = b.xchg + (int64_t)c0 * b.sX;
        x.idx = x.val + 4 * (int64_t)b.gstride;
        x.row = x.idx + 2 * (int64_t)b.gstride;
        x.rowj = x.row + 2 * (int64_t)b.gstride * LEAF_W * 4;
        x.gstride = b.gstride;
        leaf_getf2_kernel<T, LEAF_W><<<dim3(G, cn), 256, 0, st>>>(
            P + (int64_t)c0 * sF, ld, m - t, wt, k + t, b.ipiv + (int64_t)c0 * b.sIv + k + t,
            b.info + c0, x, (unsigned)leaf * LEAF_W, b.err, sF, b.sIv, b.sX);
      }
      ++leaf;
      if (t > 0) laswp_b<T>(b, k, k + t, k + t, wt, st);
      if (t + wt < kb) {
        laswp_b<T>(b, k + t + wt, k + kb, k + t, wt, st);
        trsm_diag_kernel<T, false><<<dim3((kb - t - wt + 15) / 16, count), 256, 0, st>>>(
            P, ld, wt, P + wt, ld, kb - t - wt, nullptr, sF);
        gemm_b<T>(m - t - wt, kb - t - wt, wt, P + (int64_t)wt * ld, ld, false, P + wt, ld,
                  P + (int64_t)wt * ld + wt, ld, T(-1), 1, nullptr, 0, count, sF, sF, sF, st);
      }
    }
    diag_inv_kernel<T><<<dim3((kb + 63) / 64, 2, count), 64, 0, st>>>(b.F, ld, n, k / 64,
                                                                       b.dinv64, sF, b.s64);
    // 16 workgroups per system in the grid's z extent (<= 65535)
    for (int c0 = 0; c0 < count; c0 += 4095) {
      const int cn = min(4095, count - c0);
      blk_inv_kernel<T><<<dim3(1, 2, 16 * cn), 256, 0, st>>>(
          b.F + c0 * sF, ld, n, k / 256, b.dinv64 + c0 * b.s64, b.dinv256 + c0 * b.s256, sF,
          b.s64, b.s256);
    }
    if (k > 0) laswp_b<T>(b, 0, k, k, kb, st);
    if (k + kb < n) {
      const int nc = n - k - kb;
      cx<T>* Akk = b.F + (int64_t)k * ld + k;
      laswp_b<T>(b, k + kb, n, k, kb, st);
      // U12 = inv(L11) A12 (staged, row stride nc), A22 -= L21 U12, U12 back
      gemm_b<T>(kb, nc, kb, b.dinv256 + (size_t)(k / 256) * 2 * 65536, 256, false, Akk + kb, ld,
                b.ustage, nc, T(1), 0, nullptr, 0, count, b.s256, sF, b.sU, st, false,
                b.trail_tile);
      gemm_b<T>(nc, nc, kb, Akk + (int64_t)kb * ld, ld, false, b.ustage, nc,
                Akk + (int64_t)kb * ld + kb, ld, T(-1), 1, nullptr, 0, count, sF, b.sU, sF, st,
                b.g3m, b.trail_tile);
      copy2d_kernel<T><<<copy_grid((size_t)kb * nc, count), 256, 0, st>>>(
          kb, nc, b.ustage, nc, Akk + kb, ld, b.sU, sF);
    }
  }
  // fold inv(L_bb) into the blocks of L below the diagonal (as fold_inverses)
  for (int b0 = 0; b0 < n; b0 += 256) {
    const int bs = min(256, n - b0), ml = n - b0 - bs;
    if (ml <= 0) break;
    const cx<T>* iL = b.dinv256 + (size_t)(b0 / 256) * 2 * 65536;
    cx<T>* Lb = b.F + (int64_t)(b0 + bs) * ld + b0;
    gemm_b<T>(ml, bs, bs, Lb, ld, false, iL, 256, b.ustage, bs, T(1), 0, nullptr, 0, count, sF,
              b.s256, b.sU, st);
    copy2d_kernel<T><<<copy_grid((size_t)ml * bs, count), 256, 0, st>>>(ml, bs, b.ustage, bs, Lb,
                                                                         ld, b.sU, sF);
  }
  const size_t perm_lds = (size_t)n * sizeof(int);
  const int in_lds = perm_lds <= 140 * 1024;
  static bool attr = false;
  if (!attr) { allow_big_lds(perm_from_ipiv_kernel); attr = true; }
  perm_from_ipiv_kernel<<<count, 256, in_lds ? perm_lds : 0, st>>>(b.ipiv, n, b.perm, in_lds,
                                                                    b.sIv);
  HPR_CHECK(hipGetLastError());
  return 0;
}

template <class T>
int getrs_batched(const BatchArgs<T>& b, cx<T>* B, cx<T>* Y, int nrhs, const int* skip, int sSkip,
                  hipStream_t st) {
  const int n = b.n, count = b.count;
  if (n <= 0 || nrhs <= 0 || count <= 0) return 0;
  const int64_t ld = b.ld, sV = b.sV;
  constexpr int BM = 64, BN = 64;
  auto tiles = [](const GemmProb<T>& p) {
    return (p.M > 0 && p.N > 0 && p.K > 0) ? ((p.M + BM - 1) / BM) * ((p.N + BN - 1) / BN) : 0;
  };
  // forward (L holds W = L[>b, b] inv(L_bb)): Y_b = inv(L_bb) B_b and
  // B_>b -= W B_b in one launch per block
  for (int r = 0; r < n; r += 256) {
    const int bs = min(256, n - r);
    GemmProb<T> p1 = prob<T>(bs, nrhs, bs, b.dinv256 + (size_t)(r / 256) * 2 * 65536, 256,
                             B + (int64_t)r * nrhs, nrhs, Y + (int64_t)r * nrhs, nrhs, T(1), 0);
    GemmProb<T> p2 = prob<T>(n - r - bs, nrhs, bs, b.F + (int64_t)(r + bs) * ld + r, ld,
                             B + (int64_t)r * nrhs, nrhs, B + (int64_t)(r + bs) * nrhs, nrhs,
                             T(-1), 1);
    p1.sA = b.s256; p1.sB = sV; p1.sC = sV;
    p2.sA = b.sF; p2.sB = sV; p2.sC = sV;
    // the product with inv(L_bb) on 32 x 32 tiles over its nonzero triangle
    // (tri_dual_kernel), the update on 64 x 32 tiles for narrow batches
    const int t1 = ((bs + 31) / 32) * ((nrhs + 31) / 32);
    if (b.subst_tile == 1) {   // 32 x 32 update tiles
      const int t2 = p2.M > 0 ? ((p2.M + 31) / 32) * ((nrhs + 31) / 32) : 0;
      tri_dual_kernel<T, true, 32, 32, 16, 16><<<dim3(t1 + t2, count), 256, 0, st>>>(
          p1, p2, t1, skip, sSkip);
    } else if (nrhs <= 32) {
      const int t2 = p2.M > 0 ? ((p2.M + 63) / 64) * ((nrhs + 31) / 32) : 0;
      tri_dual_kernel<T, true, 64, 32, 32, 16><<<dim3(t1 + t2, count), 256, 0, st>>>(
          p1, p2, t1, skip, sSkip);
    } else {
      const int t2 = tiles(p2);
      tri_dual_kernel<T, true, 64, 64, 32, 32><<<dim3(t1 + t2, count), 256, 0, st>>>(
          p1, p2, t1, skip, sSkip);
    }
  }
  // backward: B_b = inv(U_bb) Y_b, then Y_<b -= U[<b, b] B_b
  for (int r = ((n - 1) / 256) * 256; r >= 0; r -= 256) {
    const int bs = min(256, n - r);
    GemmProb<T> pu = prob<T>(bs, nrhs, bs, b.dinv256 + ((size_t)(r / 256) * 2 + 1) * 65536, 256,
                             Y + (int64_t)r * nrhs, nrhs, B + (int64_t)r * nrhs, nrhs, T(1), 0);
    pu.sA = b.s256; pu.sB = sV; pu.sC = sV;
    const GemmProb<T> none = prob<T>(0, 0, 0, pu.A, 256, Y, nrhs, B, nrhs, T(1), 0);
    const int tu = ((bs + 31) / 32) * ((nrhs + 31) / 32);
    tri_dual_kernel<T, false, 64, 64, 32, 32><<<dim3(tu, count), 256, 0, st>>>(pu, none, tu, skip,
                                                                               sSkip);
    if (r > 0)
      gemm_b<T>(r, nrhs, bs, b.F + r, ld, false, B + (int64_t)r * nrhs, nrhs, Y, nrhs, T(-1), 1,
                skip, sSkip, count, b.sF, sV, sV, st, false, b.subst_tile == 1 ? 1 : 0);
  }
  HPR_CHECK(hipGetLastError());
  return 0;
}

template <class TF>
int solve_batched(const BatchArgs<TF>& b, const double2* A64, int64_t lda, int64_t sA,
                  double2* dB, int64_t ldb, int64_t sB, int nrhs, int max_refine, double tol,
                  int crit, double* res_out, int* conv_out, hipStream_t st,
                  const MeshOperator* mesh) {
  const int n = b.n, count = b.count;
  if (n <= 0 || count <= 0) return 0;
  if (mesh && (int64_t)mesh->nch * mesh->nr != n) return -1;
  if (nrhs <= 0 || lda < n || ldb < n || max_refine < 0) return -1;
  const int64_t sV = b.sV;
  const size_t cnt = (size_t)n * nrhs;
  const dim3 eg = copy_grid(cnt, count);
  const int* done = &b.state->done;
  constexpr int sDone = (int)(sizeof(RefineState) / sizeof(int));
  static_assert(sizeof(RefineState) % sizeof(int) == 0, "state stride in ints");
  // factors of the (equilibrated) matrices, row-major, factor precision
  if (b.scale)
    row_scale_kernel<<<dim3((n + 63) / 64, count), 256, 0, st>>>(n, A64, lda, b.scale, sA,
                                                                   b.sIv);
  transpose_cast_kernel<double, TF><<<dim3((n + 31) / 32, (n + 31) / 32, count), 256, 0, st>>>(
      n, n, A64, lda, b.F, b.ld, b.scale, b.scale, sA, b.sF, b.sIv);
  if (getrf_batched<TF>(b, st)) return -1;
  // right-hand sides, row-major FP64
  transpose_cast_kernel<double, double>
      <<<dim3((n + 31) / 32, (nrhs + 31) / 32, count), 256, 0, st>>>(
          n, nrhs, dB, ldb, b.Brm, nrhs, nullptr, nullptr, sB, sV, 0);
  norms_b(n, nrhs, b.Brm, sV, b.part, b.state, 0, crit, tol, nullptr, 0, count, st);
  permute_cast_kernel<double, TF><<<eg, 256, 0, st>>>(n, nrhs, b.perm, b.Brm, b.RF, b.scale,
                                                      nullptr, b.sIv, sV, 0);
  if (getrs_batched<TF>(b, b.RF, b.RY, nrhs, nullptr, 0, st)) return -1;
  add_correction_kernel<TF><<<eg, 256, 0, st>>>(nrhs, cnt, b.Xrm, b.RF, b.scale, 0, b.state, sV,
                                                b.sIv);
  // classic refinement: residual in FP64, correction through the factors;
  // a converged system skips every later launch; the last evaluation only checks
  for (int it = 0; it <= max_refine; ++it) {
    cast_kernel<double, double><<<eg, 256, 0, st>>>(cnt, b.Brm, b.R, done, sV, sDone);
    if (mesh) {
      if (mesh_apply_batched(*mesh, count, nrhs, b.Xrm, sV, b.R, sV, -1.0, 1, done, sDone, st))
        return -1;
    } else {
      gemm_b<double>(n, nrhs, n, A64, lda, true, b.Xrm, nrhs, b.R, nrhs, -1.0, 1, done, sDone,
                     count, sA, sV, sV, st);
    }
    norms_b(n, nrhs, b.R, sV, b.part, b.state, 1, crit, tol, done, sDone, count, st, 1);
    if (it > 0)   // a system whose residual grew returns its previous iterate
      undo_correction_kernel<TF><<<eg, 256, 0, st>>>(nrhs, cnt, b.Xrm, b.RF, b.scale, b.state,
                                                     it + 1, sV, b.sIv);
    if (it == max_refine) break;
    permute_cast_kernel<double, TF><<<eg, 256, 0, st>>>(n, nrhs, b.perm, b.R, b.RF, b.scale, done,
                                                        b.sIv, sV, sDone);
    if (getrs_batched<TF>(b, b.RF, b.RY, nrhs, done, sDone, st)) return -1;
    add_correction_kernel<TF><<<eg, 256, 0, st>>>(nrhs, cnt, b.Xrm, b.RF, b.scale, 1, b.state, sV,
                                                  b.sIv);
  }
  transpose_cast_kernel<double, double>
      <<<dim3((nrhs + 31) / 32, (n + 31) / 32, count), 256, 0, st>>>(
          nrhs, n, b.Xrm, nrhs, dB, ldb, nullptr, nullptr, sV, sB, 0);
  batch_result_kernel<<<(count + 255) / 256, 256, 0, st>>>(count, b.state, b.info, b.err,
                                                           res_out, conv_out);
  HPR_CHECK(hipGetLastError());
  return 0;
}

#define INST(T)                                                                               \
  template int getrf<T>(cx<T>*, int64_t, int, int*, int*, const FactorWork<T>&, const Scratch&, \
                        const LuPlan&, int, hipStream_t, const SideStream*);                   \
  template int getrs<T>(const cx<T>*, int64_t, int, const cx<T>*, cx<T>*, int64_t, int, cx<T>*, \
                        const int*, const LuPlan&, bool, hipStream_t);                         \
  template int getrs_replay<T>(SubstGraph*, const cx<T>*, int64_t, int, const cx<T>*, cx<T>*,  \
                               int64_t, int, cx<T>*, const int*, const LuPlan&, bool,          \
                               hipStream_t);                                                   \
  template int gemm<T>(int, int, int, const cx<T>*, int64_t, bool, const cx<T>*, int64_t,      \
                       cx<T>*, int64_t, T, int, bool, const int*, hipStream_t, size_t, int);   \
  template int refine_solve<T>(const cx<T>*, int64_t, int, const cx<T>*, const int*,          \
                               const double2*, int64_t, const double2*, double2*, int, int,    \
                               double, int, RefineState*, double2*, cx<T>*, cx<T>*, double*,   \
                               const LuPlan&, bool, hipStream_t, const double*,               \
                               const SplitWork&, bool, bool);                                  \
  template int batch_bind<T>(BatchArgs<T>&, void*, size_t, int, int, int, bool);              \
  template int getrf_batched<T>(const BatchArgs<T>&, hipStream_t);                             \
  template int getrs_batched<T>(const BatchArgs<T>&, cx<T>*, cx<T>*, int, const int*, int,     \
                                hipStream_t);                                                  \
  template int solve_batched<T>(const BatchArgs<T>&, const double2*, int64_t, int64_t,         \
                                double2*, int64_t, int64_t, int, int, double, int, double*,    \
                                int*, hipStream_t, const MeshOperator*);
INST(double)
INST(float)
template int transpose_cast<double, double>(int, int, const cx<double>*, int64_t, cx<double>*, int64_t, hipStream_t, const double*, const double*);
template int transpose_cast<double, float>(int, int, const cx<double>*, int64_t, cx<float>*, int64_t, hipStream_t, const double*, const double*);
template int transpose_cast<float, double>(int, int, const cx<float>*, int64_t, cx<double>*, int64_t, hipStream_t, const double*, const double*);
template int permute_cast<double, double>(int, int, const int*, const cx<double>*, cx<double>*, const int*, hipStream_t, const double*);
template int permute_cast<double, float>(int, int, const int*, const cx<double>*, cx<float>*, const int*, hipStream_t, const double*);

}  // namespace hpr
