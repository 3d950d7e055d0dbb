// The Lagrange-mesh matrix as an operator: R = beta R + alpha C X in FP64
// without the dense N x N matrix.
//
// C (src/rmatrix_hp.F90:246-293, as assembled by hprmat_hip_assemble) has
// two kinds of nonzeros: the channel blocks
//   D_c[i][j] = (T+L)_ij + delta_ij (l_c(l_c+1)/r_i^2 -+ k_c^2)
//               - [closed c] xc1_c blo0_ij / a           (real, nr x nr)
// and the couplings V_cc'(r_i) between channels at the same mesh point
// (complex, diagonal in the mesh index).  A row has nr + nch - 1 nonzeros, so
// C X costs 8 N (nr + nch) nrhs flops instead of 8 N^2 nrhs: 8.6 GFLOP
// against 671 GFLOP for N = 25600 (nch = 128, nr = 200) and 128 columns.
// The refinement's products A Z are these products when the caller
// describes C this way (hprmat_hip_lu_set_operator); the result is the same
// matrix times the same vectors, summed in another order.
//
// Two LDS-tiled FP64 kernels (VALU; the operands are real x complex or
// small complex blocks), 64 x 64 outputs per workgroup, 4 x 4 per thread:
//   block:  R_c = beta R_c + alpha D_c X_c          grid (nrhs/64, nr/64, nch x systems)
//   couple: R(i) += alpha V(i) X(i), X(i)[c'] = X[c' nr + i]
//                                                    grid (nrhs/64, nch/64, nr x systems)
// X and R are row-major N x nrhs with row index c nr + i.  A batch of
// systems shares the mesh, the angular momenta and the couplings (one
// partial wave at several energies); qk and xc1 have one row per system.
#include "lu.hpp"

namespace hpr {

namespace {

constexpr int MT = 64, KT = 16;

__global__ void __launch_bounds__(256)
mesh_block_kernel(int nch, int nr, int nrhs, const double* __restrict__ tc,
                  const double* __restrict__ blo0, const double* __restrict__ r, double rmax0,
                  const int* __restrict__ lval, const double* __restrict__ qk,
                  const double* __restrict__ xc1, int64_t sq, const double2* __restrict__ X,
                  int64_t ldx, int64_t sX, double2* __restrict__ R, int64_t ldr, int64_t sR,
                  double alpha, int beta, const int* __restrict__ skip, int sSkip, int e0) {
  // blockIdx.z = c + nch * (system - e0)
  const int c = blockIdx.z % nch, sys = e0 + blockIdx.z / nch;
  if (skip && skip[(int64_t)sys * sSkip]) return;
  qk += sys * sq;
  xc1 += sys * sq;
  X += sys * sX;
  R += sys * sR;
  __shared__ double dt[KT][MT + 1];     // D_c[i0 + ii][j0 + jj] as dt[jj][ii]
  __shared__ double2 xt[KT][MT];        // X_c[j0 + jj][k0 + kk]
  const int i0 = blockIdx.y * MT, k0 = blockIdx.x * MT, tid = threadIdx.x;
  const int ti = (tid / 16) * 4, tk = (tid % 16) * 4;
  const double q = qk[c];
  const double l = (double)lval[c];
  const double sb = q < 0.0 ? xc1[c] / rmax0 : 0.0;
  const double2* xc = X + (int64_t)c * nr * ldx;
  double2 acc[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) acc[a][b] = make_double2(0.0, 0.0);
  for (int j0 = 0; j0 < nr; j0 += KT) {
    for (int e = tid; e < KT * MT; e += 256) {
      const int jj = e / MT, ii = e % MT, i = i0 + ii, j = j0 + jj;
      double v = 0.0;
      if (i < nr && j < nr) {
        v = tc[(int64_t)i * nr + j];
        if (i == j) {
          v += l * (l + 1.0) / (r[i] * r[i]);
          if (q > 0.0) v -= q * q;
          if (q < 0.0) v += q * q;
        }
        if (q < 0.0) v -= sb * blo0[(int64_t)i * nr + j];
      }
      dt[jj][ii] = v;
      const int kk = ii, k = k0 + kk;
      xt[jj][kk] = (j < nr && k < nrhs) ? xc[(int64_t)j * ldx + k] : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 4
    for (int jj = 0; jj < KT; ++jj) {
      double d[4];
      double2 x[4];
      for (int a = 0; a < 4; ++a) d[a] = dt[jj][ti + a];
      for (int b = 0; b < 4; ++b) x[b] = xt[jj][tk + b];
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
          acc[a][b].x = fma(d[a], x[b].x, acc[a][b].x);
          acc[a][b].y = fma(d[a], x[b].y, acc[a][b].y);
        }
    }
    __syncthreads();
  }
  double2* rc = R + (int64_t)c * nr * ldr;
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + ti + a;
    if (i >= nr) continue;
    for (int b = 0; b < 4; ++b) {
      const int k = k0 + tk + b;
      if (k >= nrhs) continue;
      double2 o = make_double2(alpha * acc[a][b].x, alpha * acc[a][b].y);
      if (beta) {
        const double2 p = rc[(int64_t)i * ldr + k];
        o.x += p.x;
        o.y += p.y;
      }
      rc[(int64_t)i * ldr + k] = o;
    }
  }
}

__global__ void __launch_bounds__(256)
mesh_couple_kernel(int nch, int nr, int nrhs, const double2* __restrict__ cpot,
                   const double2* __restrict__ X, int64_t ldx, int64_t sX,
                   double2* __restrict__ R, int64_t ldr, int64_t sR, double alpha,
                   const int* __restrict__ skip, int sSkip, int e0) {
  // blockIdx.z = i + nr * (system - e0)
  const int i = blockIdx.z % nr, sys = e0 + blockIdx.z / nr;
  if (skip && skip[(int64_t)sys * sSkip]) return;
  X += sys * sX;
  R += sys * sR;
  __shared__ double2 vt[KT][MT + 1];    // V(i)[c0 + cc][d0 + dd] as vt[dd][cc]
  __shared__ double2 xt[KT][MT];        // X[(d0 + dd) nr + i][k0 + kk]
  const int c0 = blockIdx.y * MT, k0 = blockIdx.x * MT, tid = threadIdx.x;
  const int tc_ = (tid / 16) * 4, tk = (tid % 16) * 4;
  const double2* v = cpot + (int64_t)i * nch * nch;
  double2 acc[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) acc[a][b] = make_double2(0.0, 0.0);
  for (int d0 = 0; d0 < nch; d0 += KT) {
    for (int e = tid; e < KT * MT; e += 256) {
      const int dd = e / MT, cc = e % MT, c = c0 + cc, d = d0 + dd;
      vt[dd][cc] = (c < nch && d < nch) ? v[(int64_t)c * nch + d] : make_double2(0.0, 0.0);
      const int k = k0 + cc;
      xt[dd][cc] = (d < nch && k < nrhs) ? X[((int64_t)d * nr + i) * ldx + k]
                                         : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 4
    for (int dd = 0; dd < KT; ++dd) {
      double2 p[4], x[4];
      for (int a = 0; a < 4; ++a) p[a] = vt[dd][tc_ + a];
      for (int b = 0; b < 4; ++b) x[b] = xt[dd][tk + b];
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
          acc[a][b].x = fma(p[a].x, x[b].x, fma(-p[a].y, x[b].y, acc[a][b].x));
          acc[a][b].y = fma(p[a].x, x[b].y, fma(p[a].y, x[b].x, acc[a][b].y));
        }
    }
    __syncthreads();
  }
  for (int a = 0; a < 4; ++a) {
    const int c = c0 + tc_ + a;
    if (c >= nch) continue;
    double2* rr = R + ((int64_t)c * nr + i) * ldr;
    for (int b = 0; b < 4; ++b) {
      const int k = k0 + tk + b;
      if (k >= nrhs) continue;
      const double2 p = rr[k];
      rr[k] = make_double2(p.x + alpha * acc[a][b].x, p.y + alpha * acc[a][b].y);
    }
  }
}

}  // namespace

int mesh_apply_batched(const MeshOperator& op, int count, int nrhs, const double2* X,
                       int64_t sX, double2* R, int64_t sR, double alpha, int beta,
                       const int* skip, int sSkip, hipStream_t st) {
  if (op.nch <= 0 || op.nr <= 0 || nrhs <= 0 || count < 0) return -1;
  if (op.nr > 65535 || op.nch > 65535) return -1;
  // systems per launch: the grid's z extent (nch or nr blocks per system) stays <= 65535
  const int per = max(1, 65535 / max(op.nch, op.nr));
  for (int e0 = 0; e0 < count; e0 += per) {
    const int ce = min(per, count - e0);
    const dim3 gb((nrhs + MT - 1) / MT, (op.nr + MT - 1) / MT, op.nch * ce);
    mesh_block_kernel<<<gb, 256, 0, st>>>(op.nch, op.nr, nrhs, op.tc, op.blo0, op.r, op.rmax0,
                                          op.lval, op.qk, op.xc1, op.sq, X, nrhs, sX, R, nrhs, sR,
                                          alpha, beta, skip, sSkip, e0);
    const dim3 gc((nrhs + MT - 1) / MT, (op.nch + MT - 1) / MT, op.nr * ce);
    mesh_couple_kernel<<<gc, 256, 0, st>>>(op.nch, op.nr, nrhs, op.cpot, X, nrhs, sX, R, nrhs, sR,
                                           alpha, skip, sSkip, e0);
  }
  HPR_CHECK(hipGetLastError());
  return 0;
}

int mesh_apply(const MeshOperator& op, int nrhs, const double2* X, double2* R, double alpha,
               int beta, const int* skip, hipStream_t st) {
  return mesh_apply_batched(op, 1, nrhs, X, 0, R, 0, alpha, beta, skip, 0, st);
}

}  // namespace hpr
