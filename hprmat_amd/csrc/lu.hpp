// Host-side drivers of the dense complex LU + solve + refinement path.
#pragma once
#include "common.hpp"

namespace hpr {

// Tunables of the blocked factorization.
struct LuPlan {
  int nb = 256;     // outer panel width and getrs block height (<= 256, multiple of 64)
  // trailing-update GEMM: 0 64x64x16, 1 64x128x16, 2 128x64x16 (16x16x4 MFMA);
  // FP32 only, 32x32x2 MFMA: 3 64x64x16, 4 128x128x16, 5 128x64x16, 6 64x128x16
  // (FP64 uses variant 0 for 3..6)
  int trail = 0;
  // panel leaves of the FP32 factors: 1 partial pivoting (leaf_getf2_kernel,
  // LAPACK's pivots), 2 tournament pivoting (leaf_calu_kernel), 0 automatic:
  // tournament except for GEMM_3M; FP64 always partial
  int pivot = 0;
};

// Device scratch used by getrf: exchange area of the cooperative panel
// leaves (tagged words, zeroed at the start of each factorization) and the
// error flag of their bounded waits.
struct Scratch {
  void* xchg = nullptr;
  size_t xchg_bytes = 0;
  int* err = nullptr;
  int* lmap = nullptr;   // interchange maps of the last three panels (3 x 1024 ints)
};

// Refinement bookkeeping kept on the device so no host round trip is needed
// between sweeps (graph-capturable).
struct RefineState {
  int done;          // convergence reached: later sweeps become no-ops
  int sweeps;        // corrections applied
  int checks;        // residual evaluations
  int pad;
  double res_max;    // modulus of the residual entry with the largest |Re|+|Im|
  double res_fro;    // ||B - A X||_F
  double b_fro;      // ||B||_F
  double b_max;      // largest |b_ij|
  double res_absmax; // largest |r_ij|
  double hist[32];   // ||r||_F / ||b||_F after each residual evaluation
  int gm_halt;       // FGMRES: no further Arnoldi step in this cycle (or done)
  int gm_steps;      // FGMRES: Arnoldi steps taken in this cycle
  int diverged;      // classic refinement: the residual evaluation (1-based
                     // count) at which ||r||_F grew; 0 = never.  `done` is
                     // set with it, and the last correction is undone.
  int pad2;
};

enum Criterion { CRIT_ABS_MAX = 0, CRIT_REL_FRO = 1 };

// Batch status of a system whose factorization's panel leaf gave up a wait
// (HPRMAT_STATUS_TIMEOUT of the C ABI).
constexpr int STATUS_TIMEOUT = -0x40000000;

// Host view of the refinement flags.  When the caller allows it (a solve
// that synchronises anyway) and the stream is not being captured, the
// refinement loops read `done` / `gm_halt` back after each sweep or Arnoldi
// step and stop launching once they are set: one synchronisation per step,
// against milliseconds of FP64 work per step at the sizes that refine, and
// no trail of no-op launches up to the sweep budget.  Otherwise the device
// flags alone turn the remaining launches into no-ops.
struct FlagPoll {
  bool on = false;
  int read(const RefineState* st, hipStream_t s, int RefineState::*field) const;
};
FlagPoll flag_poll(hipStream_t st, bool allowed);

size_t scratch_bytes();
void scratch_bind(Scratch& s, void* base);

// Device work areas of one factorization (element counts for an n x n
// factor): inverses of the 64 x 64 and 256 x 256 diagonal blocks of L and U,
// the 256-row staging block of U12 and the final row permutation.
inline size_t dinv64_elems(int n) { return (size_t)((n + 63) / 64) * 2 * 4096; }
inline size_t dinv256_elems(int n) { return (size_t)((n + 255) / 256) * 2 * 65536; }
inline size_t ustage_elems(int n) { return (size_t)256 * (n > 0 ? n : 1); }

// Split operands of one trailing update (16-bit elements): packed L21 and
// U12^T (n rows of 256 K x up to six planes each, split_gemm.hpp) followed
// by their int32 scale exponents (n each) and the 256 column maxima of L21
// (FP16 pieces only).
inline size_t split_elems(int n) {
  const size_t r = n > 0 ? n : 1;
  return (size_t)2 * 6 * 256 * r + 4 * r + 1024;
}

// Arithmetic of the FP32 trailing update A22 -= L21 U12: FP32 products
// (four or Gauss's three per complex product) or split 16-bit pieces on
// the 32x32x16 matrix cores (split_gemm.hpp): two BF16 pieces (3 products),
// three BF16 pieces (6 products), two scaled FP16 pieces (3 products).
enum GemmMode { GEMM_4M = 0, GEMM_3M = 1, GEMM_BF16X3 = 2, GEMM_BF16X6 = 3, GEMM_F16X3 = 4 };

template <class T> struct FactorWork {
  cx<T>* dinv64 = nullptr;
  cx<T>* dinv256 = nullptr;
  cx<T>* ustage = nullptr;
  int* perm = nullptr;
  unsigned short* split = nullptr;   // GEMM_BF16X3 / GEMM_BF16X6 only
  // staging of the look-ahead stream's folds of inv(L_bb) into finished
  // block columns (ustage_elems(n)); null: every fold after the last step
  cx<T>* ftmp = nullptr;
};

// Second stream of the look-ahead: the trailing update of all but the next
// panel's columns runs there while the main stream factors the next panel.
// Both streams keep the default priority.
struct SideStream {
  hipStream_t stream = nullptr;
  hipEvent_t ready = nullptr;   // main -> side: U12 staged, next panel updated
  hipEvent_t rest = nullptr;    // side -> main: rest of the trailing update done
  hipEvent_t join = nullptr;    // FGMRES lanes: the second lane's work done
};

// Device factorization P A = L U (row-major A, in place).  Also produces the
// inverses of the 256 x 256 diagonal blocks of L and U (w.dinv256, used as
// GEMM operands by the triangular solves) and the final row permutation
// w.perm (perm[i] = source row of row i of P A) used by getrs.  gmode
// (GemmMode) selects the arithmetic of the trailing update; the panel and
// U12 = L11^-1 A12 always use the plain complex product of T (or Gauss's
// three products for GEMM_3M).
template <class T>
int getrf(cx<T>* A, int64_t ld, int n, int* ipiv, int* info, const FactorWork<T>& w,
          const Scratch& s, const LuPlan& plan, int gmode, hipStream_t st,
          const SideStream* side = nullptr);

// Solve L U X = B in place for a right-hand side already permuted by perm
// (row-major n x nrhs, leading dimension ldb); Y is an n x nrhs work array
// with the same leading dimension.
template <class T>
int getrs(const cx<T>* LU, int64_t ld, int n, const cx<T>* dinv256, cx<T>* B, int64_t ldb,
          int nrhs, cx<T>* Y, const int* skip, const LuPlan& plan, bool gauss3m, hipStream_t st);

// A substitution chain (getrs) captured once into a HIP graph and then
// replayed: one graph launch instead of about 3 n / 256 kernel launches,
// which at n = 25600 cost the host about 6 ms per substitution -- more than
// the GPU needs for it when two FGMRES lanes share the host thread.  The
// graph bakes in the factors, the right-hand-side and work buffers and the
// skip flag; a call with any of them (or n, nrhs, ld, ldb) different
// captures anew.  Under stream capture, or with HPRMAT_HIP_SUBST_GRAPH=0,
// the kernels are launched directly.
struct SubstGraph {
  hipGraphExec_t exec = nullptr;
  hipStream_t last = nullptr;   // stream of the latest launch (drained before release)
  const void *LU = nullptr, *dinv = nullptr, *B = nullptr, *Y = nullptr;
  const int* skip = nullptr;
  int n = 0, nrhs = 0, elem = 0;
  int64_t ld = 0, ldb = 0;
  bool g3m = false;
};
void subst_graph_release(SubstGraph& g);
template <class T>
int getrs_replay(SubstGraph* g, const cx<T>* LU, int64_t ld, int n, const cx<T>* dinv256,
                 cx<T>* B, int64_t ldb, int nrhs, cx<T>* Y, const int* skip, const LuPlan& plan,
                 bool gauss3m, hipStream_t st);

// out[i][:] = (Tout) (scale[perm[i]] in[perm[i]][:]) (row-major n x nrhs;
// scale = null: no scaling).
template <class Tin, class Tout>
int permute_cast(int n, int nrhs, const int* perm, const cx<Tin>* in, cx<Tout>* out,
                 const int* skip, hipStream_t st, const double* scale = nullptr);

// out[r * ldo + c] = (Tout) (rs[r] cs[c] in[c * ldi + r]), r < rows, c < cols
// (rs = null: no scaling).
template <class Tin, class Tout>
int transpose_cast(int rows, int cols, const cx<Tin>* in, int64_t ldi, cx<Tout>* out,
                   int64_t ldo, hipStream_t st, const double* rs = nullptr,
                   const double* cs = nullptr);

// Symmetric equilibration factors s[i] = ||row i of A||_2^(-1/2) (1 for a
// zero row) of a column-major FP64 matrix: the factored matrix is
// diag(s) A diag(s) (still complex symmetric when A is).
int equil_scale(int n, const double2* A, int64_t lda, double* s, hipStream_t st);

// Complex GEMM launcher: C = beta C + alpha A B (row-major; A column-major when a_col).
template <class T>
int gemm(int M, int N, int K, const cx<T>* A, int64_t lda, bool a_col, const cx<T>* B,
         int64_t ldb, cx<T>* C, int64_t ldc, T alpha, int beta, bool gauss3m,
         const int* skip, hipStream_t st, size_t lds_extra = 0, int prio = 0);

// Split-K partial sums of the FP64 refinement product R = beta R + alpha A X
// (A column-major n x n, X and R row-major n x nrhs): `splits` slices of K
// are multiplied by separate workgroups into `part` (splits x n x nrhs) and
// summed in a fixed order, so a product that has too few output tiles to
// fill the GPU still runs at full occupancy and results stay reproducible.
// The Lagrange-mesh matrix described by its parts (device arrays, as for
// hprmat_hip_assemble): channel blocks from tc, blo0, r, rmax0, lval, qk,
// xc1 and the couplings cpot[i][c][c'] (nr x nch x nch complex).
struct MeshOperator {
  int nch = 0, nr = 0;
  const double* tc = nullptr;
  const double* blo0 = nullptr;
  const double* r = nullptr;
  double rmax0 = 1.0;
  const int* lval = nullptr;
  const double* qk = nullptr;
  const double* xc1 = nullptr;
  const double2* cpot = nullptr;
  int64_t sq = 0;   // batches: stride of the per-system rows of qk and xc1
};
// R = beta R + alpha C X (X, R row-major N x nrhs, N = nch nr) from the parts
// of C, skipped when *skip is set.
int mesh_apply(const MeshOperator& op, int nrhs, const double2* X, double2* R, double alpha,
               int beta, const int* skip, hipStream_t st);
// The same for `count` systems of one partial wave (X, R at strides sX, sR;
// qk / xc1 rows at stride op.sq; system e skipped when skip[e sSkip] is set).
int mesh_apply_batched(const MeshOperator& op, int count, int nrhs, const double2* X,
                       int64_t sX, double2* R, int64_t sR, double alpha, int beta,
                       const int* skip, int sSkip, hipStream_t st);

// How the FP64 refinement product is formed: split-K partial sums, or, with
// `mesh`, the Lagrange-mesh operator instead of the dense matrix.
struct SplitWork {
  double2* part = nullptr;
  int splits = 1;
  const MeshOperator* mesh = nullptr;
};
// Number of K slices for the product of an n x n matrix with nrhs columns
// (1: no split); resid_splits(n, nrhs) * n * nrhs complex doubles of `part`.
int resid_splits(int n, int nrhs);
constexpr int RESID_MAX_SPLITS = 8;
// R = beta R + alpha A64 X (skipped when *skip is set), split-K when sw has
// more than one slice and a partial buffer; through sw.mesh when it is set.
int resid_gemm(int n, int nrhs, const double2* A64, int64_t lda, const double2* X, double2* R,
               double alpha, int beta, const int* skip, const SplitWork& sw, hipStream_t st);

// Mixed/plain refinement around resident factors.  A64 is the FP64 matrix in
// column-major order (lda); B and X are row-major n x nrhs (ld = nrhs).
// Work buffers: R (n x nrhs FP64), RF and RY (n x nrhs in factor precision),
// part (>= 5120 doubles).  With scale non-null the factors are those of
// diag(s) A diag(s): right-hand sides are scaled by s before the solves and
// the solves' results by s after them.  At most max_refine corrections, each
// after a residual evaluation; final_check evaluates the last iterate too.
// An evaluation whose ||r||_F exceeds the previous one stops the refinement
// and undoes the last correction (RefineState::diverged; not converged).
template <class TF>
int refine_solve(const cx<TF>* LU, int64_t ld, int n, const cx<TF>* dinv256, const int* perm,
                 const double2* A64, int64_t lda, const double2* B, double2* X, int nrhs,
                 int max_refine, double tol, int crit, RefineState* state, double2* R,
                 cx<TF>* RF, cx<TF>* RY, double* part, const LuPlan& plan, bool gauss3m,
                 hipStream_t st, const double* scale, const SplitWork& sw = SplitWork(),
                 bool poll = false, bool final_check = false);

// Norms of an n x nrhs FP64 block R (row-major): mode 0 records ||B|| and
// resets the state, mode 1 records a residual evaluation and the stopping
// rule (skipped when *skip is set; mode 1 also when state->done).  guard:
// stop (RefineState::diverged) when ||R||_F exceeds the previous evaluation.
int residual_norms(int n, int nrhs, const double2* R, double* part, RefineState* state, int mode,
                   int crit, double tol, const int* skip, hipStream_t st, int guard = 0);

// Workspace of gmres_solve for n x nrhs blocks and cycles of m <= GMRES_MAX_M steps.
struct GmresWork {
  double2* V = nullptr;     // (m + 1) Krylov blocks, n x nrhs each
  double2* Z = nullptr;     // m preconditioned blocks
  double2* W = nullptr;     // n x nrhs
  double2* H = nullptr;     // (m + 1) x m x nrhs rotated Hessenberg columns
  double2* g = nullptr;     // (m + 1) x nrhs
  double2* sn = nullptr;    // m x nrhs
  double* cs = nullptr;     // m x nrhs
  double2* h = nullptr;     // 2 x (m + 1) x nrhs projections (two CGS passes)
  double2* y = nullptr;     // m x nrhs
  double* inv = nullptr;    // nrhs
  double2* part = nullptr;  // GMRES_DOT_BLOCKS x (m + 1) x nrhs partial dots
};
constexpr int GMRES_MAX_M = 24;
constexpr int GMRES_DEFAULT_CYCLE = 16;
constexpr int GMRES_DOT_BLOCKS = 256;
size_t gmres_work_bytes(int n, int nrhs, int m);
void gmres_work_bind(GmresWork& w, void* base, int n, int nrhs, int m);

// Flexible GMRES iterative refinement (FGMRES-IR, right preconditioning by
// the resident LU factors in their own precision, FP64 matrix-vector
// products and orthogonalization).  Arguments as refine_solve; max_steps
// bounds the preconditioned matrix-vector products (cycles of m steps, each
// cycle starting from the true residual B - A X).  X ends with a true
// residual evaluation recorded in state.  Nothing synchronises with the host
// unless `poll` is set (see FlagPoll).
template <class TF>
int gmres_solve(const cx<TF>* LU, int64_t ld, int n, const cx<TF>* dinv256, const int* perm,
                const double2* A64, int64_t lda, const double2* B, double2* X, int nrhs,
                int max_steps, int m, double tol, int crit, RefineState* state, double2* R,
                cx<TF>* RF, cx<TF>* RY, double* part, const GmresWork& gw, const LuPlan& plan,
                bool gauss3m, hipStream_t st, const double* scale,
                const SplitWork& sw = SplitWork(), bool poll = false);

// One lane of a multi-lane FGMRES solve: a subset of the right-hand sides
// (row-major n x nrhs blocks of their own) with its own refinement state,
// work buffers and stream.
template <class TF> struct GmresLane {
  const double2* B = nullptr;
  double2* X = nullptr;
  int nrhs = 0;
  RefineState* state = nullptr;
  double2* R = nullptr;
  cx<TF>* RF = nullptr;
  cx<TF>* RY = nullptr;
  double* part = nullptr;
  GmresWork gw;
  SplitWork sw;
  hipStream_t st = nullptr;
  SubstGraph* graph = nullptr;   // replayed substitution of the Arnoldi steps (null: launches)
};
constexpr int GMRES_MAX_LANES = 4;
// gmres_solve for nl lanes at once, each lane on its own stream with its
// own stopping rule (a lane that meets the rule stops; the union meets it
// too).  The lanes start together and, doing the same work, keep in step:
// their latency-bound substitutions run beside each other, and so do their
// FP64 products (N = 25600, two lanes of 64 columns: 180 ms per solve phase
// against 185 ms with the second lane started one substitution later, whose
// substitutions then ran beside the other lane's products and took twice as
// long).  The caller joins the lanes' streams.
template <class TF>
int gmres_solve_lanes(const cx<TF>* LU, int64_t ld, int n, const cx<TF>* dinv256, const int* perm,
                      const double2* A64, int64_t lda, int max_steps, int m, double tol, int crit,
                      const LuPlan& plan, bool gauss3m, const double* scale, int nl,
                      GmresLane<TF>* lanes, bool poll);

// ---- uniform batches ------------------------------------------------------
// `count` independent systems of one order n, each at a fixed element
// stride from the previous one in every buffer; one chain of launches (the
// system index in a grid dimension) factors and solves all of them.
template <class T> struct BatchArgs {
  int n = 0, count = 0;
  int64_t ld = 0;                                   // row stride of the factors
  cx<T>* F = nullptr;        int64_t sF = 0;        // row-major factors
  cx<T>* dinv64 = nullptr;   int64_t s64 = 0;       // 64 x 64 diagonal-block inverses
  cx<T>* dinv256 = nullptr;  int64_t s256 = 0;      // 256 x 256 diagonal-block inverses
  cx<T>* ustage = nullptr;   int64_t sU = 0;        // U12 staging
  int* ipiv = nullptr;
  int* perm = nullptr;       int sIv = 0;           // also the stride of `scale`
  int* info = nullptr;                              // count entries
  double* scale = nullptr;                          // equilibration (null: none)
  unsigned long long* xchg = nullptr; int64_t sX = 0; int gstride = 0;
  int* err = nullptr;
  RefineState* state = nullptr;                     // count entries
  double* part = nullptr;                           // count x BATCH_NORM_BLOCKS x 5
  double2* Brm = nullptr;
  double2* Xrm = nullptr;
  double2* R = nullptr;
  cx<T>* RF = nullptr;
  cx<T>* RY = nullptr;
  int64_t sV = 0;                                   // stride of the n x nrhs blocks
  bool g3m = false;                                 // Gauss 3-product trailing update
  int trail_tile = 0;    // tiles of U12 and the trailing update (gemm_b_tiled: 0, 1, 2)
  int subst_tile = 0;    // substitution update tiles: 0 by width, 1 32 x 32
};
constexpr int BATCH_NORM_BLOCKS = 64;

size_t batch_work_bytes(int n, int nrhs, int count, size_t elem);
template <class T>
int batch_bind(BatchArgs<T>& b, void* base, size_t bytes, int n, int nrhs, int count, bool equil);
// Blocked GEPP of every system (no look-ahead: the batch fills the GPU).
template <class T> int getrf_batched(const BatchArgs<T>& b, hipStream_t st);
// Substitution with the factors of getrf_batched on row-major n x nrhs
// blocks at stride b.sV (B overwritten with the solution, Y workspace);
// systems whose skip flag (stride sSkip ints) is set are left untouched.
template <class T>
int getrs_batched(const BatchArgs<T>& b, cx<T>* B, cx<T>* Y, int nrhs, const int* skip, int sSkip,
                  hipStream_t st);
// Factor + classic FP64 refinement of every system: A64 column-major (lda,
// stride sA), dB column-major right-hand sides (ldb, stride sB) overwritten
// with the solutions; per system the relative residual ||B - A X||_F/||B||_F
// (res_out) and 1 converged / 0 not converged / -info singular (conv_out).
// With mesh, the FP64 products are formed from the Lagrange-mesh parts of
// the matrices (mesh_apply_batched) instead of A64.
template <class TF>
int solve_batched(const BatchArgs<TF>& b, const double2* A64, int64_t lda, int64_t sA,
                  double2* dB, int64_t ldb, int64_t sB, int nrhs, int max_refine, double tol,
                  int crit, double* res_out, int* conv_out, hipStream_t st,
                  const MeshOperator* mesh = nullptr);

// HIP-event timing of getrf's trailing-update GEMM launches (for roofline
// reporting): enable, run, then read total ms / algorithmic flops / count.
int prof_enable(int on);
int prof_read(double* ms, double* flops, int* count);

}  // namespace hpr
