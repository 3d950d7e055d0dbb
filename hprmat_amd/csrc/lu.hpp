// Host-side drivers of the dense complex LU + solve + refinement path.
#pragma once
#include "common.hpp"

namespace hpr {

// Tunables of the blocked factorization (chosen per precision by the tuner).
struct LuPlan {
  int nb = 256;     // outer panel width
  int tri = 64;     // diagonal block of the triangular solves (<= 64)
  int rs_nb = 256;  // block row height of the getrs sweeps
};

// Device scratch used by getrf (exchange area of the cooperative panel
// leaves, grid-barrier counters and the barrier error flag).
struct Scratch {
  void* xchg = nullptr;
  size_t xchg_bytes = 0;
  unsigned* counters = nullptr;
  size_t ncounters = 0;
  int* err = nullptr;
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
};

enum Criterion { CRIT_ABS_MAX = 0, CRIT_REL_FRO = 1 };

// Factor precision / trailing-update arithmetic.
enum Method { M_FP64 = 0, M_FP32 = 1, M_FP32_3M = 2, M_BF16X3 = 3 };

size_t scratch_bytes(int n);
int scratch_bind(Scratch& s, void* base, int n);

template <class T>
int getrf(cx<T>* A, int64_t ld, int n, int* ipiv, int* info, const Scratch& s,
          const LuPlan& plan, bool gauss3m, hipStream_t st);

template <class T>
int getrs(const cx<T>* LU, int64_t ld, int n, const int* ipiv, cx<T>* B, int64_t ldb,
          int nrhs, const int* skip, const LuPlan& plan, bool gauss3m, hipStream_t st);

// out[r * ldo + c] = (Tout) in[c * ldi + r], r < rows, c < cols
template <class Tin, class Tout>
int transpose_cast(int rows, int cols, const cx<Tin>* in, int64_t ldi, cx<Tout>* out,
                   int64_t ldo, hipStream_t st);

// Complex GEMM launcher: C = beta C + alpha A B (row-major; A column-major when a_col).
template <class T>
int gemm(int M, int N, int K, const cx<T>* A, int64_t lda, bool a_col, const cx<T>* B,
         int64_t ldb, cx<T>* C, int64_t ldc, T alpha, int beta, bool gauss3m,
         const int* skip, hipStream_t st);

// Mixed/plain refinement around resident factors.  A64 is the FP64 matrix in
// column-major order (lda); B and X are row-major n x nrhs (ld = nrhs).
// Work buffers: R (n x nrhs FP64), RF (n x nrhs in factor precision),
// part (>= 4096 doubles).
template <class TF>
int refine_solve(const cx<TF>* LU, int64_t ld, int n, const int* ipiv, const double2* A64,
                 int64_t lda, const double2* B, double2* X, int nrhs, int max_refine,
                 double tol, int crit, RefineState* state, double2* R, cx<TF>* RF,
                 double* part, const LuPlan& plan, bool gauss3m, hipStream_t st);

// One residual evaluation of a refined X (skipped when state->done is set):
// R = B - A64 X, norms and the stopping rule recorded in state.
int residual_check(int n, int nrhs, const double2* A64, int64_t lda, const double2* B,
                   const double2* X, double2* R, double* part, RefineState* state, int crit,
                   double tol, hipStream_t st);

}  // namespace hpr
