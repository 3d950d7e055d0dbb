/*
 * hprmat_hip.h - C ABI of libhprmat_hip, the gfx950 dense complex LU + solve
 * library behind HPRMAT's GPU solver types.
 *
 * Part 1 reproduces, name for name and argument for argument, the 15
 * extern "C" entry points that HPRMAT's Fortran module
 * src/gpu_solver_interface.F90:42-162 binds with bind(C, name=...): a build
 * of HPRMAT that links libhprmat_hip instead of its CUDA object needs no
 * source change.  Conventions kept from the reference:
 *   - arguments by reference (Fortran style), n and nrhs are 32-bit;
 *   - A and B are complex(c_double_complex) = interleaved (re, im) doubles,
 *     column-major, lda = ldb = n;
 *   - A is read only; B holds the RHS on entry and the solution on success
 *     and is left untouched on failure;
 *   - return 0 on success, -1 on failure; *info = 0 ok, -1 init failed,
 *     -2 allocation failed, > 0 zero pivot (1-based), 99 refinement did not
 *     converge (multi_mixed guard).
 * Internal indexing is 64-bit (the reference's int32 n*n overflows at
 * N > 46340); device buffers are library-owned and grow-only.
 *
 * Part 2 is the native device-pointer interface used by the Python/PyTorch
 * binding: stream-ordered, no hidden host synchronisation unless statistics
 * are requested, LU handles whose factors are reused for adjoint solves.
 */
#ifndef HPRMAT_HIP_H
#define HPRMAT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Part 1: reference-compatible ABI --------------------- */

/* src/cusolver_interface.cu:195-219 (gpu_solver_init_): select the device,
   print its name; idempotent. */
int gpu_solver_init_(int *device_id);
/* src/cusolver_interface.cu:221-256: release all device buffers. */
int gpu_solver_finalize_(void);
/* src/cusolver_interface.cu:262-269: drop the cached page-locked host
   registration.  This library never page-locks the caller's buffer, so the
   call only returns 0 (kept for link compatibility). */
int gpu_host_unregister_(void);
/* src/cusolver_interface.cu:321-443: complex-FP32 LU on the GPU, FP32 solve,
   then up to max_refine FP64 refinement sweeps stopping when the residual
   entry of largest |Re|+|Im| has modulus < tol. */
int gpu_solve_mixed_(double *A, double *B, int *n, int *nrhs, int *max_refine,
                     double *tol, int *info);
/* src/cusolver_interface.cu:456-592: same factorization; the reference forms
   the FP64 residual on the host, here it is formed on the GPU (MI355X FP64 is
   not throttled), with identical stopping rule and results. */
int gpu_solve_hybrid_(double *A, double *B, int *n, int *nrhs, int *max_refine,
                      double *tol, int *info);
/* src/cusolver_interface.cu:600-734: reduced-precision factorization with
   refinement (default 5 sweeps when max_refine <= 0).  gfx950 has no TF32
   matrix path; this entry point uses the split-BF16 three-product
   factorization (TF32-like accuracy). */
int gpu_solve_tf32_(double *A, double *B, int *n, int *nrhs, int *max_refine,
                    double *tol, int *info);
/* src/cusolver_interface.cu:1600-1639: memory-fit warning, then mixed. */
int gpu_solve_auto_(double *A, double *B, int *n, int *nrhs, int *max_refine,
                    double *tol, int *info);
/* src/cusolver_interface.cu:737-795: plain FP64 LU + solve.  The reference
   applies no refinement (LAPACK-quality substitution); here the substitution
   multiplies by explicit diagonal-block inverses, so at most two FP64
   refinement sweeps follow until ||B - A X||_F / ||B||_F <= 1e-12, which
   gives LAPACK zgesv's accuracy (or better) on kappa ~ 1e9 systems.  Not
   reaching the gate is not an error. */
int gpu_solve_double_(double *A, double *B, int *n, int *nrhs, int *info);
/* src/cusolver_interface.cu:797-804 */
int gpu_is_available_(void);
/* src/cusolver_interface.cu:807-814 */
int gpu_get_device_count_(void);
/* src/cusolver_interface.cu:816-825: "name (xx.x GB)", NUL-terminated. */
void gpu_get_info_(char *info_str, int *len);
/* src/cusolver_interface.cu:832-912: one-GPU build; *ngpu is set to the
   number of devices used (1). */
int gpu_multi_init_(int *ngpu);
/* src/cusolver_interface.cu:915-1019 */
int gpu_multi_finalize_(void);
/* src/cusolver_interface.cu:1023-1331: FP64 LU + solve (one GPU). */
int gpu_solve_multi_(double *A, double *B, int *n, int *nrhs, int *info);
/* src/cusolver_interface.cu:1345-1597: FP32 LU + FP64 refinement, then the
   reference's convergence guard ||r||_max > 1e-6 (||b||_max + 1) -> info 99
   (disabled by HPRMAT_MGPU_NO_FALLBACK). */
int gpu_solve_multi_mixed_(double *A, double *B, int *n, int *nrhs,
                           int *max_refine, double *tol, int *info);

/* ---------------- Part 2: native device-pointer interface -------------- */

/* Factorization precision of the LU (all use FP64 residuals/updates). */
enum {
  HPRMAT_PLAIN_FP64 = 0,     /* complex FP64 LU on FP64 MFMA                */
  HPRMAT_MIXED_FP32 = 1,     /* complex FP32 LU, 4 real MFMA products       */
  HPRMAT_MIXED_FP32_3M = 2,  /* complex FP32 LU, Gauss 3-product MFMA       */
  HPRMAT_SPLIT_BF16 = 3,     /* FP32-stored LU, split-BF16 3-term products  */
  HPRMAT_SPLIT_BF16X6 = 4,   /* FP32-stored LU, three-piece split-BF16 with
                                6 products per real product: FP32 accuracy  */
  HPRMAT_SPLIT_F16 = 5,      /* FP32-stored LU, two-piece FP16 split with
                                power-of-two row / column scales, 3 products
                                per real product (about 2^-22 relative)     */
  /* flag OR-ed into the method: factor diag(s) A diag(s), s_i = ||row i of
     A||_2^-1/2 (symmetric equilibration; keeps A^T = A), refinement
     residuals still on A */
  HPRMAT_EQUILIBRATE = 0x100
};
/* Stopping rule of the refinement. */
enum {
  HPRMAT_CRIT_ABS_MAX = 0,   /* |r_ij| of the largest |Re|+|Im| entry < tol  */
  HPRMAT_CRIT_REL_FRO = 1,   /* ||B - A X||_F / ||B||_F <= tol               */
  /* flag OR-ed into the rule: flexible GMRES refinement (FP64 Arnoldi per
     right-hand side, the resident factors as right preconditioner, cycles of
     min(max_refine, 24) steps); max_refine then bounds the preconditioned
     matrix-vector products, `sweeps` counts them */
  HPRMAT_FGMRES = 0x10
};

typedef struct hprmat_hip_stats {
  int info;        /* 0 ok, > 0 zero pivot (1-based), -3 a panel leaf's
                      bounded cross-workgroup wait gave up (the factors and
                      the solution are invalid)                            */
  int sweeps;      /* corrections applied                                */
  int checks;      /* residual evaluations                               */
  int converged;   /* stopping rule met at the last evaluation           */
  double res_max;  /* modulus of the largest-|Re|+|Im| residual entry    */
  double res_rel;  /* ||B - A X||_F / ||B||_F at the last evaluation     */
  double b_fro;    /* ||B||_F                                            */
  double hist[32]; /* relative residual after each evaluation            */
  int diverged;    /* classic refinement stopped because ||B - A X||_F
                      grew; the last correction was undone, so X is the
                      best iterate (res_rel / res_max are its residual)   */
  int pad_;
} hprmat_hip_stats;

/* Batch status of every system when a panel leaf's bounded wait gave up
   (hprmat_hip_batch_solve*; cannot collide with a zero-pivot -info). */
enum { HPRMAT_STATUS_TIMEOUT = -0x40000000 };

typedef struct hprmat_hip_lu *hprmat_hip_lu_t;

/* No reference counterpart (the reference keeps one process-global factor).
   Create / destroy an LU handle owning device factors and work buffers. */
int hprmat_hip_lu_create(hprmat_hip_lu_t *h);
int hprmat_hip_lu_destroy(hprmat_hip_lu_t h);

/* No reference counterpart (cuSOLVER chooses internally): factorization
   plan of a handle, set by an autotuner or by hand before factoring.
     HPRMAT_PLAN_TRAIL_TILE  trailing-update GEMM tile of the FP32/FP64
                             4-product path: 0 = 64x64x16 (default),
                             1 = 64x128x16, 2 = 128x64x16 (double-buffered LDS)
     HPRMAT_PLAN_LOOKAHEAD   1 = overlap the next panel with the trailing
                             update on a second stream (default), 0 = off
     HPRMAT_PLAN_GMRES_CYCLE steps per FGMRES cycle, 1..24 (0 = default 16);
                             the solve's step budget spans several cycles
     HPRMAT_PLAN_SOLVE_LANES FGMRES solves: 1 = all columns in one lane,
                             2 = two lanes of columns on two streams, the
                             second staggered by one substitution (one
                             lane's substitutions overlap the other's FP64
                             products; each lane stops on its own rule), 0 =
                             automatic (two for n >= 8192 and >= 64 columns,
                             one under stream capture)
     HPRMAT_PLAN_PIVOT       pivot rule of the FP32 factors: 1 = partial
                             pivoting (LAPACK's pivots, one exchange per
                             column), 2 = tournament pivoting (CALU: a few
                             exchanges per 32 columns), 0 = automatic
                             (default): tournament, except partial for
                             HPRMAT_MIXED_FP32_3M, whose Gauss form amplifies
                             the larger multipliers tournament pivots allow;
                             the FP64 factors always use partial pivoting
   Returns -1 for an unknown key or value. */
enum { HPRMAT_PLAN_TRAIL_TILE = 1, HPRMAT_PLAN_LOOKAHEAD = 2, HPRMAT_PLAN_GMRES_CYCLE = 3,
       HPRMAT_PLAN_SOLVE_LANES = 4, HPRMAT_PLAN_PIVOT = 5 };
int hprmat_hip_lu_set_plan(hprmat_hip_lu_t h, int key, int value);

/* No reference counterpart: describe the factored matrix by its
   Lagrange-mesh parts, the device inputs of hprmat_hip_assemble (one
   interval; no non-local potential).  The refinement's FP64 products C X
   are then formed from the channel blocks and the couplings (8 N (nr + nch)
   nrhs flops) instead of the dense matrix (8 N^2 nrhs); the factors, the
   substitutions and the stopping rule are unchanged.  Applies to solves
   with nch * nr = n; the arrays must stay valid and describe the matrix
   that was factored.  tc = NULL returns to the dense matrix.  Returns -1 on
   invalid arguments. */
int hprmat_hip_lu_set_operator(hprmat_hip_lu_t h, int nch, int nr, const double *tc,
                               const double *blo0, const double *r, double rmax0,
                               const int *lval, const double *qk, const double *xc1,
                               const void *cpot);

/* Factor the n x n complex FP64 matrix at device pointer dA (column-major,
   lda) in the given precision.  dA is not modified and must stay valid while
   the handle is used for refined solves (the FP64 residual reads it).
   Stream-ordered, no host synchronisation; the zero-pivot flag is read with
   hprmat_hip_lu_info.  Replaces cusolverDn?getrf at
   src/cusolver_interface.cu:383 and :779. */
int hprmat_hip_lu_factor(hprmat_hip_lu_t h, int method, const void *dA,
                         int64_t lda, int n, void *stream);
/* Synchronise and return the factorization's zero-pivot index (0 = none). */
int hprmat_hip_lu_info(hprmat_hip_lu_t h, int *info);

/* Synchronise and copy the row interchanges to host ipiv[n], 1-based as
   LAPACK returns them (row i was interchanged with row ipiv[i]).  The
   reference discards cusolverDn?getrf's d_ipiv (src/cusolver_interface.cu:111). */
int hprmat_hip_lu_pivots(hprmat_hip_lu_t h, int *ipiv);

/* Solve A X = B in place on the device (dB column-major, ldb, nrhs columns):
   factor-precision solve, then up to max_refine FP64 refinement sweeps with
   the given stopping rule; final_check adds one residual evaluation after
   the last correction.  stats may be NULL (then nothing synchronises).
   Replaces cusolverDn?getrs + the cublasZgemm / cublasIzamax refinement loop
   at src/cusolver_interface.cu:393-434. */
int hprmat_hip_lu_solve(hprmat_hip_lu_t h, void *dB, int64_t ldb, int nrhs,
                        int max_refine, double tol, int criterion,
                        int final_check, void *stream, hprmat_hip_stats *stats);

/* One-shot factor + solve with a library-global handle (device pointers). */
int hprmat_hip_solve(int method, const void *dA, int64_t lda, void *dB,
                     int64_t ldb, int n, int nrhs, int max_refine, double tol,
                     int criterion, int final_check, void *stream,
                     hprmat_hip_stats *stats);

/* Assemble the complex-symmetric Lagrange-mesh matrix of one interval
   (ns = 1) into column-major device storage dA (lda >= nch*nr):
     C[(c,i),(c',j)] = delta_cc' [(T+L)_ij + delta_ij (l_c(l_c+1)/r_i^2 -+ k_c^2)
                                  - [qk_c < 0] xc1_c blo0_ij / rmax0]
                       + delta_ij cpot[i][c][c'].
   All pointers are device arrays: tc, blo0 (nr x nr row-major, scaled as
   rmat_ini leaves them), r (nr radii), lval, qk, xc1 (nch), cpot (nr x nch x
   nch complex).  Replaces the host assembly at src/rmatrix_hp.F90:246-293. */
int hprmat_hip_assemble(int nch, int nr, const double *tc, const double *blo0,
                        const double *r, double rmax0, const int *lval,
                        const double *qk, const double *xc1, const void *cpot,
                        void *dA, int64_t lda, void *stream);

/* Add a non-local potential to a matrix assembled by hprmat_hip_assemble
   (one interval): C[(c,i),(c',j)] += sqrt(wle_i wle_j) rmax0 K[i][j][c][c'],
   device arrays, wle the (0, 1) Gauss weights, K nr x nr x nch x nch
   complex.  Replaces the cpnl loop of src/rmatrix_hp.F90:283-289. */
int hprmat_hip_add_nonlocal(int nch, int nr, const double *wle, double rmax0, const void *K,
                            void *dA, int64_t lda, void *stream);

/* R = normfac q2^T X (nch x nch) and S = Z_O^-1 Z_I (nopen x nopen), both
   column-major FP64 on the device.  dX is the column-major N x nch solution
   (ldx), q2 the nr boundary values, qk the nch wave numbers, open the
   0-based open-channel indices, co = G + iF and cop = rho O' per open
   channel (device arrays).  Replaces src/rmat_solvers.F90:597-609 and
   src/rmatrix_hp.F90:336-368. */
int hprmat_hip_r_to_s(int nch, int nr, const void *dX, int64_t ldx,
                      const double *q2, double normfac, double rmax,
                      const double *qk, int nopen, const int *open,
                      const void *co, const void *cop, void *dR, void *dS,
                      void *stream);
/* No reference counterpart: hprmat_hip_r_to_s with a caller-owned device
   workspace (hprmat_hip_r_to_s_work_bytes(nch) bytes) instead of the
   library-global one, for concurrent calls on several streams and for
   stream capture (no allocation inside).  Returns -2 when work_bytes is too
   small. */
size_t hprmat_hip_r_to_s_work_bytes(int nch);
int hprmat_hip_r_to_s_ws(int nch, int nr, const void *dX, int64_t ldx,
                         const double *q2, double normfac, double rmax,
                         const double *qk, int nopen, const int *open,
                         const void *co, const void *cop, void *dR, void *dS,
                         void *work, size_t work_bytes, void *stream);

/* No reference counterpart (the reference assembles one energy per call):
   hprmat_hip_assemble for `count` energies of one system at once; qk and
   xc1 are count x nch (row e for system e), dA receives count column-major
   matrices stride_a elements apart.  One launch. */
int hprmat_hip_assemble_batch(int nch, int nr, int count, const double *tc,
                              const double *blo0, const double *r, double rmax0,
                              const int *lval, const double *qk, const double *xc1,
                              const void *cpot, void *dA, int64_t lda, int64_t stride_a,
                              void *stream);

/* No reference counterpart: hprmat_hip_r_to_s for `count` solutions of one
   system (nch <= 64) in one launch, one workgroup per system (S by Gaussian
   elimination with partial pivoting in FP64).  dX: count column-major N x
   nch solutions (ldx, stride_x apart); qk, open, co, cop: count x nch
   (row e for system e; open, co, cop use the first nopen[e] entries);
   nopen: count device ints; dR, dS: count blocks of nch*nch complex
   (R column-major nch x nch, S column-major nopen x nopen at the block
   start).  Returns -1 for nch > 64. */
int hprmat_hip_r_to_s_batch(int nch, int nr, int count, const void *dX, int64_t ldx,
                            int64_t stride_x, const double *q2, double normfac, double rmax,
                            const double *qk, const int *nopen, const int *open,
                            const void *co, const void *cop, void *dR, void *dS,
                            void *stream);

/* No reference counterpart (the reference has no derivatives): adjoint
   contraction for gradients with respect to a local coupling potential.
   With X = A^-1 B and M = A^-T G (for complex-symmetric A the same factors
   solve it), a real loss with torch-convention gradient G = dL/dconj(X)
   has dL/dconj(V_cc'(r_i)) = -conj(K[i][c][c']) where
     K[i][c][c'] = sum_k M[(c nr + i), k] X[(c' nr + i), k].
   dM, dX: column-major N x nrhs (ldm, ldx), N = nch nr; dK: nr x nch x nch
   complex (row-major), all on the device. */
int hprmat_hip_adjoint_contract(int nch, int nr, int nrhs, const void *dM, int64_t ldm,
                                const void *dX, int64_t ldx, void *dK, void *stream);

/* No reference counterpart: HIP-event timing of the LU's trailing-update
   GEMM (the dominant kernel) for roofline reports.  Enable, factor, then
   read the summed elapsed milliseconds, the algorithmic flops (8 M N K per
   complex update) and the number of timed launches (read synchronises and
   resets). */
int hprmat_hip_prof_enable(int on);
int hprmat_hip_prof_read(double *ms, double *flops, int *count);

/* No reference counterpart: the library's complex GEMM on the device,
   C = beta C + alpha A B with alpha = alpha_sign (+1 or -1), beta in {0, 1},
   row-major C (ldc) and B (ldb), A row-major or, with a_col, column-major
   (lda); precision 64 (complex128 data) or 32 (complex64 data).  The kernel
   of the LU trailing update and of the FP64 refinement residual. */
int hprmat_hip_zgemm(int precision, int M, int N, int K, const void *A, int64_t lda, int a_col,
                     const void *B, int64_t ldb, void *C, int64_t ldc, int alpha_sign, int beta,
                     void *stream);

/* No reference counterpart (the reference loops rmatrix over energies, one
   zgesv per system): factor + solve + classic FP64 refinement of `count`
   independent systems of one order n in lockstep, one chain of launches for
   the whole batch.  dA: count column-major n x n matrices (lda, stride_a
   elements apart); dB: count column-major n x nrhs right-hand sides (ldb,
   stride_b apart), overwritten with the solutions.  method_flags:
   HPRMAT_PLAIN_FP64, HPRMAT_MIXED_FP32 or HPRMAT_MIXED_FP32_3M (Gauss's
   three-product trailing update), optionally | HPRMAT_EQUILIBRATE and a
   tile plan | HPRMAT_BATCH_TRAIL(t) | HPRMAT_BATCH_SUBST(s) (see below);
   criterion: HPRMAT_CRIT_* (no FGMRES).  max_refine correction sweeps at
   most, each system stopping once it meets tol; res_rel[i] receives the
   final ||B - A X||_F / ||B||_F and status[i] 1 (converged), 0 (not
   converged) or -info (exactly zero pivot at column info), both device
   arrays (may be NULL).  work: device workspace of
   hprmat_hip_batch_work_bytes(method, n, nrhs, count) bytes (256-byte
   aligned); nothing is allocated and nothing synchronises with the host,
   so the call can be captured into a HIP graph.  dA must not be modified
   until the call completes on `stream`.  Batches running concurrently on
   several streams need streams * ceil(n / 256) <= 512: the panel leaf of
   one system of each batch keeps ceil(n / 256) workgroups resident. */
/* Tile plan of a batched solve (OR-ed into method_flags; the scan's
   autotuner picks it per system size): HPRMAT_BATCH_TRAIL(t) sets the tiles
   of U12 and of the trailing update (0: 64 x 64, 1: 32 x 32 -- four times
   the workgroups, for small systems --, 2: 128 x 64), HPRMAT_BATCH_SUBST(s)
   those of the substitutions' block updates (0: 64 x 64, or 64 x 32 for at
   most 32 right-hand sides; 1: 32 x 32).  Any other bit is an error. */
#define HPRMAT_BATCH_TRAIL(t) (((t) & 0xf) << 16)
#define HPRMAT_BATCH_SUBST(s) (((s) & 0xf) << 20)
size_t hprmat_hip_batch_work_bytes(int method, int n, int nrhs, int count);
int hprmat_hip_batch_solve(int method_flags, int n, int nrhs, int count, const void *dA,
                           int64_t lda, int64_t stride_a, void *dB, int64_t ldb,
                           int64_t stride_b, int max_refine, double tol, int criterion,
                           double *res_rel, int *status, void *work, size_t work_bytes,
                           void *stream);

/* No reference counterpart: hprmat_hip_batch_solve for `count` systems of
   one partial wave (one energy each) whose refinement products are formed
   from their Lagrange-mesh parts (hprmat_hip_lu_set_operator's form): the
   mesh, angular momenta and couplings are shared, qk and xc1 have one row
   of nch per system, stride_q apart.  n = nch * nr. */
typedef struct {
  int nch, nr;
  const double *tc, *blo0, *r;
  double rmax0;
  const int *lval;
  const double *qk, *xc1;
  int64_t stride_q;
  const void *cpot;
} hprmat_hip_mesh_parts;
int hprmat_hip_batch_solve_mesh(int method_flags, int n, int nrhs, int count, const void *dA,
                                int64_t lda, int64_t stride_a, void *dB, int64_t ldb,
                                int64_t stride_b, int max_refine, double tol, int criterion,
                                double *res_rel, int *status, void *work, size_t work_bytes,
                                const hprmat_hip_mesh_parts *parts, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): hprmat_hip_batch_solve with the workspace taken from a
   scratch array of HPRMAT_NOALLOC_SCRATCH_BYTES built into the library (one
   per GPU, held from the first call on that GPU until the process ends)
   instead of a caller buffer.  Same arguments and results as
   hprmat_hip_batch_solve without work / work_bytes; calls no hipMalloc,
   hipFree or allocating routine and does not synchronise.  Returns -2 (with
   a message) when hprmat_hip_batch_work_bytes(method, n, nrhs, count)
   exceeds the array, e.g. more than one FP32 system of order 2000.  Every
   call on one GPU uses the same array: calls that overlap on one GPU (on
   different streams, or from several threads) race and give wrong results;
   order them on one stream. */
#define HPRMAT_NOALLOC_SCRATCH_BYTES (64u << 20)
int hprmat_hip_batch_solve_noalloc(int method_flags, int n, int nrhs, int count, const void *dA,
                                   int64_t lda, int64_t stride_a, void *dB, int64_t ldb,
                                   int64_t stride_b, int max_refine, double tol, int criterion,
                                   double *res_rel, int *status, void *stream);

/* Library version string. */
const char *hprmat_hip_version(void);

/* No reference counterpart as one call (the reference runs it on the host,
   src/rmatrix_hp.F90:319-431): R-matrix propagation over ns intervals, S
   and, when nc > 0, the wave functions, on the device.  dX holds the ns
   interval solutions C_s^-1 [q1_s blocks | q2_s blocks] (column-major N x
   2 nch, N = nch nr, ldx, stride_x apart); q1, q2 (ns x nr), qk (nch),
   open / co / cop (nopen: open channels and G + iF, rho O' at rmax = ns
   rmax0), nvc (nc, 0-based among the open channels), wle (nr), closed /
   xc2 (nclo closed channels and W(|k| rmax)) are device arrays.  Outputs:
   dR (nch x nch, the propagated R-matrix times rmax, optional), dS (nopen x
   nopen), dCf (ld1 x ld2 x nc wave functions at the mesh points), dAmp
   (nclo x nc bound-state amplitudes).  nch <= 64.  work: a device buffer of
   hprmat_hip_propagate_work_bytes(nch, ns, nc) bytes.  Nothing is
   allocated or synchronised. */
size_t hprmat_hip_propagate_work_bytes(int nch, int ns, int nc);
int hprmat_hip_propagate(int nch, int nr, int ns, const void *dX, int64_t ldx, int64_t stride_x,
                         const double *q1, const double *q2, double normfac, double rmax0,
                         const double *qk, int nopen, const int *open, const void *co,
                         const void *cop, void *dR, void *dS, int nc, const int *nvc,
                         const double *wle, int nclo, const int *closed, const double *xc2,
                         void *dCf, int64_t ld1, int64_t ld2, void *dAmp, void *work,
                         size_t work_bytes, void *stream);

/* ---------------- Part 3: host-pointer physics interface --------------- */
/* bindings/c/hprmat.h:20-24: solver type constants of the C binding. */
#define HPRMAT_SOLVER_DENSE     1   /* FP64 LU (the reference's LAPACK ZGESV type)   */
#define HPRMAT_SOLVER_MIXED     2   /* FP32 LU + two FP64 sweeps (1e-10)             */
#define HPRMAT_SOLVER_WOODBURY  3   /* the reference's CPU Woodbury: FP64 result here */
#define HPRMAT_SOLVER_GPU       4   /* FP32 LU + FGMRES refinement to 1e-12          */
/* The reference's C binding (bindings/c/hprmat.h:30-166, implemented in
   bindings/c/hprmat_c.F90:25-261), same names and argument meaning; complex
   arrays are double _Complex (interleaved re, im), column-major, and
   channel indices (nvc, iv, nom) are 1-based as in the reference.  Every
   interval's system is assembled, factored and solved on the GPU, and R, S,
   the propagation over intervals and the wave functions are formed there.
   Provided for eta = 0; several intervals (ns > 1) or
   wave functions need nch <= 64.  Other inputs print a message and return
   nopen = -1 (hprmat_solve, hprmat_solve_full) or a zero R-matrix
   (hprmat_solve_linear) where the reference stops the program.  Solver
   types: 1 dense FP64 (+ at most two FP64 sweeps to ||r||_F/||b||_F <=
   1e-12), 2 FP32 LU + two FP64 sweeps (1e-10), 3 as 1 (the reference's
   CPU Woodbury solver gives the FP64 result), 4 FP32 LU + FGMRES to 1e-12;
   0 = the type set by hprmat_set_solver (default 1). */
/* bindings/c/hprmat.h:30 (rmat_ini): the nr*ns mesh abscissas. */
void hprmat_init(int nr, int ns, double rmax, double *zrma);
/* bindings/c/hprmat.h:37-45 */
void hprmat_set_solver(int solver);
int hprmat_get_solver(void);
/* bindings/c/hprmat.h:78-83 (rmatrix, src/rmatrix_hp.F90:175-368):
   cpot[ir + c*cpot_dim1 + c'*cpot_dim1*nch] = V_cc'(r_ir); cu (nch x nch)
   receives S in its leading nopen x nopen block (open channels in input
   order), zeros elsewhere. */
void hprmat_solve(int nch, int *lval, double *qk, double *eta, double rmax, int nr, int ns,
                  void *cpot, int cpot_dim1, void *cu, int *nopen, int solver);
/* bindings/c/hprmat.h:108-114 (rmatrix with twf = .true.,
   src/rmatrix_hp.F90:370-431): as hprmat_solve, plus the internal wave
   functions cf[r + c cf_dim1 + ic cf_dim1 cf_dim2] at the nr*ns mesh points
   for the nc_entrance entrance channels nvc (1-based among the open
   channels) and, for each entrance column ic, the bound-state amplitudes of
   the closed channels in cu rows nopen.. .  cpnl (cpnl_dim1 >= nr*nr rows,
   row ir*nr + ir' for the mesh pair (ir, ir'), then nch x nch) is the
   optional non-local potential, one interval only (as in the reference).
   The wave functions are those of Descouvemont's code and its golden
   outputs (C^-1 applied to the boundary terms, rmat_pierre/rmatrix.f:296). */
void hprmat_solve_full(int nch, int *lval, double *qk, double *eta, double rmax, int nr, int ns,
                       void *cpot, int cpot_dim1, void *cu, void *cf, int cf_dim1, int cf_dim2,
                       int nc_entrance, int *nvc, void *cpnl, int cpnl_dim1, int *nopen,
                       int solver);
/* bindings/c/hprmat.h:142-147 (wf_print, src/rmatrix_hp.F90:678-734): the
   wave function of channel iv for entrance column nom at r = h, 2h, ...,
   npoin h: natural cubic spline of cf inside rmax, the asymptotic form from
   cu outside.  Host work only. */
void hprmat_wavefunction(int nch, int *lval, double *qk, double *eta, double rmax, int nr, int ns,
                         void *cu, int nopen, void *cf, int cf_dim1, int cf_dim2, double *zrma,
                         int iv, int nom, int npoin, double h, void *cwftab);
/* bindings/c/hprmat.h:162-164 (solve_rmatrix, src/rmat_solvers.F90:49-90):
   C X = B with column c of B = B_vector in the rows of channel c, then
   R(i,j) = normfac sum_r B_vector(r) X(i*nlag + r, j), Rmat column-major. */
void hprmat_solve_linear(void *cmat, void *B_vector, int nch, int nlag, double normfac,
                         void *Rmat, int solver);
/* No reference counterpart (the reference keeps them in module arrays,
   src/rmatrix_hp.F90:61-159): the mesh matrices of ns intervals of (0, rmax)
   for callers that assemble on the device with hprmat_hip_assemble: per
   interval s, tc and bloch (ns x nr x nr, row-major blocks; bloch is the
   closed-channel Bloch matrix with its radius factors, so the assembly takes
   rmax0 = 1), r, q1, q2 (ns x nr) and the (0, 1) weights wle (nr). */
int hprmat_hip_mesh(int nr, int ns, double rmax, double *tc, double *bloch, double *r,
                    double *q1, double *q2, double *wle);
/* bindings/c/hprmat_c.F90:273-287 and :295: version and solver description. */
void hprmat_version(char *version, int maxlen);
void hprmat_solver_info(int solver, char *info, int maxlen);
/* bindings/c/hprmat.h:187-197: the reference releases its cached page-locked
   registration of the caller's matrix.  This library never page-locks
   caller memory; the call releases the interface's cached device buffers,
   LU handle and stream (all recreated on the next call).  Safe to call at
   any time from the thread that calls the interface, also repeatedly. */
void hprmat_gpu_finalize(void);

#ifdef __cplusplus
}
#endif
#endif /* HPRMAT_HIP_H */
