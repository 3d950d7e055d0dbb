// Lagrange-mesh system assembly and the R -> S step on the device.
//
// Assembly restates src/rmatrix_hp.F90:246-293 for one interval (ns = 1):
//   C[(c,i),(c',j)] = delta_cc' [ (T+L)_ij + delta_ij (l_c(l_c+1)/r_i^2 -+ k_c^2)
//                                 - [closed c] xc1_c blo0_ij / a ]
//                     + delta_ij V_cc'(r_i)
// written straight into column-major device storage, so the N x N matrix
// never crosses PCIe (the inputs are O(nr^2 + nr nch^2)).
//
// R and S restate src/rmat_solvers.F90:597-609 (R = normfac q2^T X) and
// src/rmatrix_hp.F90:336-368 (Z_O, Z_I, S = Z_O^-1 Z_I, FP64), with the
// nopen x nopen solve done by this library's own getrf/getrs.
#include "hprmat_hip.h"
#include "lu.hpp"
#include "small_gesv.hpp"

namespace {

__global__ void __launch_bounds__(256)
assemble_kernel(int nch, int nr, const double* __restrict__ tc, const double* __restrict__ blo0,
                const double* __restrict__ r, double rmax0, const int* __restrict__ lval,
                const double* __restrict__ qk, const double* __restrict__ xc1,
                const double2* __restrict__ cpot, double2* __restrict__ A, int64_t lda,
                int64_t sA = 0) {
  // batch over energies: blockIdx.z selects qk / xc1 (nch each) and A
  qk += (int64_t)blockIdx.z * nch;
  xc1 += (int64_t)blockIdx.z * nch;
  A += blockIdx.z * sA;
  const int n = nch * nr;
  const int row = blockIdx.x * 256 + threadIdx.x;
  const int col = blockIdx.y;
  if (row >= n) return;
  const int c = row / nr, i = row % nr;
  const int c2 = col / nr, j = col % nr;
  double vr = 0.0, vi = 0.0;
  if (c == c2) {
    vr = tc[i * nr + j];
    const double q = qk[c];
    if (i == j) {
      const double l = (double)lval[c];
      vr += l * (l + 1.0) / (r[i] * r[i]);
      if (q > 0.0) vr -= q * q;
      if (q < 0.0) vr += q * q;
    }
    if (q < 0.0) vr -= xc1[c] * blo0[i * nr + j] / rmax0;
  }
  if (i == j) {
    double2 v = cpot[((int64_t)i * nch + c) * nch + c2];
    vr += v.x;
    vi += v.y;
  }
  A[(int64_t)col * lda + row] = make_double2(vr, vi);
}

// Non-local potential (src/rmatrix_hp.F90:283-289): C[(c,i),(c',j)] +=
// sqrt(w_i w_j) a0 K_cc'(r_i, r_j), K as [i][j][c][c'] complex.
__global__ void __launch_bounds__(256)
nonlocal_kernel(int nch, int nr, const double* __restrict__ wle, double rmax0,
                const double2* __restrict__ K, double2* __restrict__ A, int64_t lda) {
  const int n = nch * nr;
  const int row = blockIdx.x * 256 + threadIdx.x;
  const int col = blockIdx.y;
  if (row >= n) return;
  const int c = row / nr, i = row % nr;
  const int c2 = col / nr, j = col % nr;
  const double f = sqrt(wle[i] * wle[j]) * rmax0;
  const double2 k = K[(((int64_t)i * nr + j) * nch + c) * nch + c2];
  double2* a = A + (int64_t)col * lda + row;
  *a = make_double2(a->x + f * k.x, a->y + f * k.y);
}

// R(a, b) = normfac * sum_r q2[r] X[a*nr + r, b]  (X column-major, R column-major)
__global__ void rmatrix_kernel(int nch, int nr, const double2* __restrict__ X, int64_t ldx,
                               const double* __restrict__ q2, double normfac,
                               double2* __restrict__ R) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nch * nch) return;
  const int a = t % nch, b = t / nch;
  double sr = 0.0, si = 0.0;
  const double2* x = X + (int64_t)b * ldx + (int64_t)a * nr;
  for (int r = 0; r < nr; ++r) {
    sr += q2[r] * x[r].x;
    si += q2[r] * x[r].y;
  }
  R[(int64_t)b * nch + a] = make_double2(sr * normfac, si * normfac);
}

__device__ inline double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Z_O (row-major, the matrix to factor) and Z_I (row-major RHS) over the
// open channels: crma2(m1, m2) = R(o1, o2)/a * sqrt(k_o1 / k_o2),
// Z_O(:, m) = -crma2(:, m) cop(m) + delta co(m), Z_I likewise with conjugates.
__global__ void zmatrices_kernel(int nch, int nopen, const double2* __restrict__ R, double rmax,
                                 const double* __restrict__ qk, const int* __restrict__ open,
                                 const double2* __restrict__ co, const double2* __restrict__ cop,
                                 double2* __restrict__ ZO, double2* __restrict__ ZI) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nopen * nopen) return;
  const int m1 = t / nopen, m2 = t % nopen;  // row m1, column m2
  const int o1 = open[m1], o2 = open[m2];
  double2 rr = R[(int64_t)o2 * nch + o1];
  const double s = sqrt(qk[o1] / qk[o2]) / rmax;
  double2 c2 = make_double2(rr.x * s, rr.y * s);
  double2 zo = zmul(c2, cop[m2]);
  double2 cc = make_double2(cop[m2].x, -cop[m2].y);
  double2 zi = zmul(c2, cc);
  zo = make_double2(-zo.x, -zo.y);
  zi = make_double2(-zi.x, -zi.y);
  if (m1 == m2) {
    zo.x += co[m1].x; zo.y += co[m1].y;
    zi.x += co[m1].x; zi.y -= co[m1].y;
  }
  ZO[(int64_t)m1 * nopen + m2] = zo;
  ZI[(int64_t)m1 * nopen + m2] = zi;
}

// S column-major from the row-major solution.
__global__ void store_colmajor_kernel(int n, const double2* __restrict__ Xr, double2* __restrict__ S) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const int i = t / n, j = t % n;
  S[(int64_t)j * n + i] = Xr[t];
}

// Batched R -> S, one workgroup per system (nch <= R2S_BATCH_MAX_NCH):
// R = normfac q2^T X to global memory; Z_O and Z_I over the open channels in
// LDS (row-major, nopen x nopen each); Z_O S = Z_I by Gaussian elimination
// with partial pivoting (largest |Re|+|Im|, first on ties) and back
// substitution, all in FP64; S stored column-major.  Per-system inputs are
// rows of count x nch arrays (qk, open, co, cop) and nopen[count].
constexpr int R2S_BATCH_MAX_NCH = hpr::SMALL_GESV_MAX_N;

__global__ void __launch_bounds__(256)
r_to_s_batch_kernel(int nch, int nr, const double2* __restrict__ X, int64_t ldx, int64_t sX,
                    const double* __restrict__ q2, double normfac, double rmax,
                    const double* __restrict__ qk, const int* __restrict__ nopen,
                    const int* __restrict__ open, const double2* __restrict__ co,
                    const double2* __restrict__ cop, double2* __restrict__ R,
                    double2* __restrict__ S) {
  const int b = blockIdx.x, tid = threadIdx.x;
  X += b * sX;
  qk += (int64_t)b * nch;
  open += (int64_t)b * nch;
  co += (int64_t)b * nch;
  cop += (int64_t)b * nch;
  R += (int64_t)b * nch * nch;
  S += (int64_t)b * nch * nch;
  const int no = nopen[b];
  extern __shared__ double2 zsh[];
  double2* ZO = zsh;               // [no][no]
  double2* ZI = zsh + no * no;     // [no][no]
  __shared__ double2 lcol[R2S_BATCH_MAX_NCH];
  __shared__ int piv;
  auto rdot = [&](int a, int c) {  // R(a, c) = normfac sum_r q2[r] X[a nr + r, c]
    const double2* x = X + (int64_t)c * ldx + (int64_t)a * nr;
    double sr = 0.0, si = 0.0;
    for (int r = 0; r < nr; ++r) {
      sr += q2[r] * x[r].x;
      si += q2[r] * x[r].y;
    }
    return make_double2(sr * normfac, si * normfac);
  };
  for (int t = tid; t < nch * nch; t += 256) R[t] = rdot(t % nch, t / nch);
  for (int t = tid; t < no * no; t += 256) {
    const int m1 = t / no, m2 = t % no, o1 = open[m1], o2 = open[m2];
    const double2 rr = rdot(o1, o2);
    const double f = sqrt(qk[o1] / qk[o2]) / rmax;
    const double2 c2 = make_double2(rr.x * f, rr.y * f);
    const double2 cp = cop[m2], cpc = make_double2(cp.x, -cp.y);
    double2 zo = zmul(c2, cp), zi = zmul(c2, cpc);
    zo = make_double2(-zo.x, -zo.y);
    zi = make_double2(-zi.x, -zi.y);
    if (m1 == m2) {
      zo.x += co[m1].x; zo.y += co[m1].y;
      zi.x += co[m1].x; zi.y -= co[m1].y;
    }
    ZO[t] = zo;
    ZI[t] = zi;
  }
  __syncthreads();
  hpr::lds_gesv(ZO, ZI, no, no, lcol, &piv);
  for (int t = tid; t < no * no; t += 256) S[(t % no) * no + t / no] = ZI[t];
}

struct SmallWork {
  void* buf = nullptr;
  size_t cap = 0;
} g_small;

}  // namespace

// K[i][c][c'] = sum_k M[(c nr + i) + k ldm] X[(c' nr + i) + k ldx]: the
// contraction of the adjoint and forward solutions over the right-hand sides
// at every mesh point, one workgroup per mesh point i.
__global__ void __launch_bounds__(256)
adjoint_contract_kernel(int nch, int nr, int nrhs, const double2* __restrict__ M, int64_t ldm,
                        const double2* __restrict__ X, int64_t ldx, double2* __restrict__ K) {
  const int i = blockIdx.x;
  for (int q = threadIdx.x; q < nch * nch; q += blockDim.x) {
    const int c = q / nch, d = q % nch;
    const double2* m = M + (int64_t)c * nr + i;
    const double2* x = X + (int64_t)d * nr + i;
    double sr = 0.0, si = 0.0;
    for (int k = 0; k < nrhs; ++k) {
      const double2 a = m[(int64_t)k * ldm], b = x[(int64_t)k * ldx];
      sr += a.x * b.x - a.y * b.y;
      si += a.x * b.y + a.y * b.x;
    }
    K[((int64_t)i * nch + c) * nch + d] = make_double2(sr, si);
  }
}

namespace {
int r_to_s_impl(int nch, int nr, const void* dX, int64_t ldx, const double* q2, double normfac,
                double rmax, const double* qk, int nopen, const int* open, const void* co,
                const void* cop, void* dR, void* dS, void* stream, void* work, size_t work_cap) {
  hipStream_t st = (hipStream_t)stream;
  if (nch <= 0 || nr <= 0 || nopen < 0 || nopen > nch) return -1;
  rmatrix_kernel<<<(nch * nch + 255) / 256, 256, 0, st>>>(nch, nr, (const double2*)dX, ldx, q2,
                                                         normfac, (double2*)dR);
  if (nopen == 0) return hipGetLastError() == hipSuccess ? 0 : -1;
  const int64_t ld = nopen;
  const size_t zbytes = (size_t)nopen * nopen * sizeof(double2);
  const size_t ibytes = ((size_t)(nopen + 64) * sizeof(int) + 255) & ~(size_t)255;
  const size_t d64 = (hpr::dinv64_elems(nopen) * sizeof(double2) + 255) & ~(size_t)255;
  const size_t d256 = (hpr::dinv256_elems(nopen) * sizeof(double2) + 255) & ~(size_t)255;
  const size_t dst = (hpr::ustage_elems(nopen) * sizeof(double2) + 255) & ~(size_t)255;
  const size_t need = 4 * zbytes + 3 * ibytes + d64 + d256 + dst + hpr::scratch_bytes() + 2048;
  if (!work) {   // library-global workspace (one caller at a time)
    if (need > g_small.cap) {
      if (g_small.buf) (void)hipFree(g_small.buf);
      g_small.buf = nullptr;
      g_small.cap = 0;
      if (hipMalloc(&g_small.buf, need) != hipSuccess) return -2;
      g_small.cap = need;
    }
    work = g_small.buf;
  } else if (need > work_cap) {
    return -2;
  }
  char* p = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  double2* ZO = (double2*)p;
  double2* ZI = (double2*)(p + zbytes);
  double2* ZP = (double2*)(p + 2 * zbytes);
  double2* ZY = (double2*)(p + 3 * zbytes);
  p += ((4 * zbytes + 255) & ~(size_t)255);
  int* ipiv = (int*)p;
  int* perm = (int*)(p + ibytes);
  int* info = (int*)(p + 2 * ibytes);
  p += 3 * ibytes;
  hpr::FactorWork<double> w;
  w.dinv64 = (double2*)p;
  w.dinv256 = (double2*)(p + d64);
  w.ustage = (double2*)(p + d64 + d256);
  w.perm = perm;
  p += d64 + d256 + dst;
  hpr::Scratch sc;
  hpr::scratch_bind(sc, p);
  zmatrices_kernel<<<(nopen * nopen + 255) / 256, 256, 0, st>>>(
      nch, nopen, (const double2*)dR, rmax, qk, open, (const double2*)co, (const double2*)cop, ZO, ZI);
  hpr::LuPlan plan;
  if (hpr::getrf<double>(ZO, ld, nopen, ipiv, info, w, sc, plan, hpr::GEMM_4M, st)) return -1;
  hpr::permute_cast<double, double>(nopen, nopen, perm, ZI, ZP, nullptr, st);
  if (hpr::getrs<double>(ZO, ld, nopen, w.dinv256, ZP, ld, nopen, ZY, nullptr, plan, false, st))
    return -1;
  store_colmajor_kernel<<<(nopen * nopen + 255) / 256, 256, 0, st>>>(nopen, ZP, (double2*)dS);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
}  // namespace

extern "C" {

/* Assemble the complex-symmetric Lagrange-mesh matrix of one interval into
   column-major device storage dA (lda >= nch*nr).  Inputs are device arrays:
   tc, blo0 (nr x nr, row-major, already scaled by the interval length as
   rmat_ini does), r (nr mesh radii), lval, qk, xc1 (nch; qk < 0 marks a
   closed channel, xc1 its Whittaker log-derivative a W'/W), cpot (nr x nch x
   nch complex, index [i][c][c']).  Replaces the host loops at
   src/rmatrix_hp.F90:246-293. */
int hprmat_hip_assemble(int nch, int nr, const double* tc, const double* blo0, const double* r,
                        double rmax0, const int* lval, const double* qk, const double* xc1,
                        const void* cpot, void* dA, int64_t lda, void* stream) {
  const int n = nch * nr;
  if (nch <= 0 || nr <= 0 || lda < n || n > 65535 * 4) return -1;
  dim3 grid((n + 255) / 256, n);
  assemble_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(nch, nr, tc, blo0, r, rmax0, lval, qk, xc1,
                                                        (const double2*)cpot, (double2*)dA, lda);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* Add a non-local potential to an assembled one-interval matrix:
   C[(c,i),(c',j)] += sqrt(wle_i wle_j) rmax0 K[i][j][c][c'] (device arrays;
   wle the (0, 1) weights, K nr x nr x nch x nch complex).  Replaces the
   cpnl loop of src/rmatrix_hp.F90:283-289. */
int hprmat_hip_add_nonlocal(int nch, int nr, const double* wle, double rmax0, const void* K,
                            void* dA, int64_t lda, void* stream) {
  const int n = nch * nr;
  if (nch <= 0 || nr <= 0 || lda < n || !wle || !K || !dA) return -1;
  dim3 grid((n + 255) / 256, n);
  nonlocal_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(nch, nr, wle, rmax0, (const double2*)K,
                                                         (double2*)dA, lda);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* R = normfac q2^T X (nch x nch, column-major) and, when nopen > 0,
   S = Z_O^-1 Z_I (nopen x nopen, column-major) in FP64 on the device.
   X is the column-major N x nch solution of the Lagrange-mesh system; open
   lists the open channels (0-based), co = G + iF and cop = rho O' per open
   channel.  Replaces src/rmat_solvers.F90:597-609 and
   src/rmatrix_hp.F90:336-368 (ZGESV(nopen) there). */
int hprmat_hip_r_to_s(int nch, int nr, const void* dX, int64_t ldx, const double* q2,
                      double normfac, double rmax, const double* qk, int nopen, const int* open,
                      const void* co, const void* cop, void* dR, void* dS, void* stream) {
  return r_to_s_impl(nch, nr, dX, ldx, q2, normfac, rmax, qk, nopen, open, co, cop, dR, dS,
                     stream, nullptr, 0);
}

size_t hprmat_hip_r_to_s_work_bytes(int nch) {
  const int k = nch > 0 ? nch : 1;
  const size_t zbytes = (size_t)k * k * sizeof(double2);
  const size_t ibytes = ((size_t)(k + 64) * sizeof(int) + 255) & ~(size_t)255;
  const size_t d64 = (hpr::dinv64_elems(k) * sizeof(double2) + 255) & ~(size_t)255;
  const size_t d256 = (hpr::dinv256_elems(k) * sizeof(double2) + 255) & ~(size_t)255;
  const size_t dst = (hpr::ustage_elems(k) * sizeof(double2) + 255) & ~(size_t)255;
  return 4 * zbytes + 3 * ibytes + d64 + d256 + dst + hpr::scratch_bytes() + 2048 + 256;
}

int hprmat_hip_r_to_s_ws(int nch, int nr, const void* dX, int64_t ldx, const double* q2,
                         double normfac, double rmax, const double* qk, int nopen, const int* open,
                         const void* co, const void* cop, void* dR, void* dS, void* work,
                         size_t work_bytes, void* stream) {
  if (!work) return -1;
  return r_to_s_impl(nch, nr, dX, ldx, q2, normfac, rmax, qk, nopen, open, co, cop, dR, dS,
                     stream, work, work_bytes);
}

int hprmat_hip_assemble_batch(int nch, int nr, int count, const double* tc, const double* blo0,
                              const double* r, double rmax0, const int* lval, const double* qk,
                              const double* xc1, const void* cpot, void* dA, int64_t lda,
                              int64_t stride_a, void* stream) {
  const int n = nch * nr;
  if (count == 0) return 0;
  if (nch <= 0 || nr <= 0 || count < 0 || count > 65535 || lda < n || n > 65535 ||
      (count > 1 && stride_a < lda * n))
    return -1;
  dim3 grid((n + 255) / 256, n, count);
  assemble_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(nch, nr, tc, blo0, r, rmax0, lval, qk,
                                                        xc1, (const double2*)cpot, (double2*)dA,
                                                        lda, stride_a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int hprmat_hip_r_to_s_batch(int nch, int nr, int count, const void* dX, int64_t ldx,
                            int64_t stride_x, const double* q2, double normfac, double rmax,
                            const double* qk, const int* nopen, const int* open, const void* co,
                            const void* cop, void* dR, void* dS, void* stream) {
  if (count == 0) return 0;
  if (nch <= 0 || nr <= 0 || count < 0 || nch > R2S_BATCH_MAX_NCH ||
      ldx < (int64_t)nch * nr || (count > 1 && stride_x < ldx * nch))
    return -1;
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute((const void*)r_to_s_batch_kernel,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            2 * R2S_BATCH_MAX_NCH * R2S_BATCH_MAX_NCH * 16) != hipSuccess) {
      (void)hipGetLastError();
      return -1;
    }
    attr = true;
  }
  const size_t lds = (size_t)2 * nch * nch * sizeof(double2);   // nopen <= nch
  r_to_s_batch_kernel<<<count, 256, lds, (hipStream_t)stream>>>(
      nch, nr, (const double2*)dX, ldx, stride_x, q2, normfac, rmax, qk, nopen, open,
      (const double2*)co, (const double2*)cop, (double2*)dR, (double2*)dS);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int hprmat_hip_adjoint_contract(int nch, int nr, int nrhs, const void* dM, int64_t ldm,
                                const void* dX, int64_t ldx, void* dK, void* stream) {
  if (nch <= 0 || nr <= 0 || nrhs < 0 || ldm < (int64_t)nch * nr || ldx < (int64_t)nch * nr)
    return -1;
  hipStream_t st = (hipStream_t)stream;
  adjoint_contract_kernel<<<nr, 256, 0, st>>>(nch, nr, nrhs, (const double2*)dM, ldm,
                                              (const double2*)dX, ldx, (double2*)dK);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
