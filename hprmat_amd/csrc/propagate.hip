// Multi-interval R-matrix propagation and wave functions on the device.
//
// Restates, for the interval solutions X_s = C_s^-1 [q1_s | q2_s] (N x 2 nch
// each, N = nch nr; the ns systems are independent and are solved by this
// library's LU before this step):
//   * the boundary blocks of solve_rmatrix_hp_4 (src/rmatrix_hp.F90:590-632):
//     R00 = q1^T C^-1 q1, R01 = q1^T C^-1 q2, R11 = q2^T C^-1 q2 per interval;
//   * the propagation chain (src/rmatrix_hp.F90:335-345)
//     crma0_1 = R11_1 / a_1,
//     crma0_s = (R11_s - R01_s^T (R00_s + a_{s-1} crma0_{s-1})^-1 R01_s) / a_s
//     with a_s = s a / ns;
//   * S = Z_O^-1 Z_I from crma0_ns (src/rmatrix_hp.F90:348-368);
//   * the wave functions (src/rmatrix_hp.F90:370-431): the amplitudes cfp at
//     the interval boundaries by the backward recursion
//     cfp_{s-1} = (R00_s + a_{s-1} crma0_{s-1})^-1 R01_s cfp_s,
//     the internal wave functions
//     cf_s = C_s^-1 (q2_s cfp_s - q1_s cfp_{s-1}) / sqrt(a0 w)
//          = (X2_s cfp_s - X1_s cfp_{s-1}) / sqrt(a0 w)
//     (the inverse C is what Descouvemont's code multiplies by,
//     rmat_pierre/rmatrix.f:223,296, and what its golden wave functions
//     hold), and the bound-state amplitudes of the closed channels.
// Launches: one grid for the R blocks (3 ns nch^2 dot products of length
// nr), one workgroup for the sequential O(ns nch^3) chain in LDS (nch <= 64),
// one grid for the wave functions (ns N nc outputs, 2 nch terms each) and a
// small one for the bound-state amplitudes.
#include "hprmat_hip.h"
#include "small_gesv.hpp"

namespace {

using hpr::zmul_d;

constexpr int PROP_MAX_NCH = hpr::SMALL_GESV_MAX_N;

__device__ inline double2 zaxpy_d(double2 y, double a, double2 x) {
  return make_double2(y.x + a * x.x, y.y + a * x.y);
}

// Rb[s][t][a][b], t = 0: R00, 1: R01, 2: R11 (row-major nch x nch blocks).
__global__ void __launch_bounds__(256)
rblocks_kernel(int nch, int nr, const double2* __restrict__ X, int64_t ldx, int64_t sX,
               const double* __restrict__ q1, const double* __restrict__ q2, double normfac,
               double2* __restrict__ Rb) {
  const int s = blockIdx.y;
  const int n2 = nch * nch;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * n2) return;
  const int t = e / n2, a = (e % n2) / nch, b = e % nch;
  const double* q = (t == 2 ? q2 : q1) + (int64_t)s * nr;
  const int col = t == 0 ? b : nch + b;
  const double2* x = X + s * sX + (int64_t)col * ldx + (int64_t)a * nr;
  double sr = 0.0, si = 0.0;
  for (int r = 0; r < nr; ++r) {
    sr += q[r] * x[r].x;
    si += q[r] * x[r].y;
  }
  Rb[(int64_t)s * 3 * n2 + e] = make_double2(sr * normfac, si * normfac);
}

// The sequential part, one workgroup: the chain, S, the boundary amplitudes.
// crma0 and Zs (= (R00_s + a_{s-1} crma0_{s-1})^-1 R01_s, s >= 1) are kept
// in global memory per interval for the backward recursion; S in Srow
// (row-major) and dS (column-major).  cfp[b][c][ic], b = 0 .. ns.
__global__ void __launch_bounds__(256)
propagate_kernel(int nch, int ns, double rmax0, const double2* __restrict__ Rb,
                 const double* __restrict__ qk, int nopen, const int* __restrict__ open,
                 const double2* __restrict__ co, const double2* __restrict__ cop,
                 double2* __restrict__ crma0, double2* __restrict__ Zs, double2* __restrict__ Srow,
                 double2* __restrict__ dS, double2* __restrict__ dR, int nc,
                 const int* __restrict__ nvc, double2* __restrict__ cfp) {
  const int tid = threadIdx.x, nt = blockDim.x, n2 = nch * nch;
  extern __shared__ double2 zsh[];
  double2* A = zsh;
  double2* B = zsh + n2;
  __shared__ double2 lcol[PROP_MAX_NCH];
  __shared__ int piv;
  for (int e = tid; e < n2; e += nt) {
    const double2 v = Rb[2 * n2 + e];
    crma0[e] = make_double2(v.x / rmax0, v.y / rmax0);
  }
  __syncthreads();
  for (int s = 1; s < ns; ++s) {
    const double r1 = s * rmax0, r2 = (s + 1) * rmax0;
    const double2* r00 = Rb + (int64_t)s * 3 * n2;
    const double2* r01 = r00 + n2;
    const double2* r11 = r00 + 2 * n2;
    const double2* prev = crma0 + (int64_t)(s - 1) * n2;
    for (int e = tid; e < n2; e += nt) {
      A[e] = zaxpy_d(r00[e], r1, prev[e]);
      B[e] = r01[e];
    }
    __syncthreads();
    hpr::lds_gesv(A, B, nch, nch, lcol, &piv);
    double2* z = Zs + (int64_t)s * n2;
    double2* cur = crma0 + (int64_t)s * n2;
    for (int e = tid; e < n2; e += nt) {
      const int a = e / nch, b = e % nch;
      double sr = 0.0, si = 0.0;
      for (int k = 0; k < nch; ++k) {
        const double2 p = zmul_d(r01[k * nch + a], B[k * nch + b]);
        sr += p.x;
        si += p.y;
      }
      z[e] = B[e];
      cur[e] = make_double2((r11[e].x - sr) / r2, (r11[e].y - si) / r2);
    }
    __syncthreads();
  }
  const double2* last = crma0 + (int64_t)(ns - 1) * n2;
  const double rmax = ns * rmax0;
  if (dR)
    for (int e = tid; e < n2; e += nt)
      dR[(e % nch) * nch + e / nch] = make_double2(last[e].x * rmax, last[e].y * rmax);
  const int no = nopen;
  if (no > 0) {
    for (int e = tid; e < no * no; e += nt) {
      const int m1 = e / no, m2 = e % no, o1 = open[m1], o2 = open[m2];
      const double f = sqrt(qk[o1] / qk[o2]);
      const double2 c2 = make_double2(last[o1 * nch + o2].x * f, last[o1 * nch + o2].y * f);
      const double2 cp = cop[m2], cpc = make_double2(cp.x, -cp.y);
      double2 zo = zmul_d(c2, cp), zi = zmul_d(c2, cpc);
      zo = make_double2(-zo.x, -zo.y);
      zi = make_double2(-zi.x, -zi.y);
      if (m1 == m2) {
        zo.x += co[m1].x; zo.y += co[m1].y;
        zi.x += co[m1].x; zi.y -= co[m1].y;
      }
      A[e] = zo;
      B[e] = zi;
    }
    __syncthreads();
    hpr::lds_gesv(A, B, no, no, lcol, &piv);
    for (int e = tid; e < no * no; e += nt) {
      dS[(e % no) * no + e / no] = B[e];
      Srow[e] = B[e];
    }
    __syncthreads();
  }
  if (nc <= 0 || no <= 0) return;
  const int nb = nch * nc;
  for (int e = tid; e < (ns + 1) * nb; e += nt) cfp[e] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int e = tid; e < no * nc; e += nt) {
    const int m1 = e / nc, ic = e % nc, nvex = nvc[ic];
    if (nvex < 0 || nvex >= no) continue;   // not an open channel: left zero
    const int i1 = open[m1], i2 = open[nvex];
    double2 v = zmul_d(Srow[m1 * no + nvex], cop[m1]);
    const double f = -sqrt(qk[i2] / qk[i1]);
    v = make_double2(v.x * f, v.y * f);
    if (i1 == i2) {
      v.x += cop[m1].x;
      v.y -= cop[m1].y;
    }
    cfp[(int64_t)ns * nb + i1 * nc + ic] = make_double2(v.x / rmax, v.y / rmax);
  }
  __syncthreads();
  for (int b = ns - 1; b >= 1; --b) {
    const double2* z = Zs + (int64_t)b * n2;
    const double2* nxt = cfp + (int64_t)(b + 1) * nb;
    for (int e = tid; e < nb; e += nt) {
      const int c = e / nc, ic = e % nc;
      double sr = 0.0, si = 0.0;
      for (int k = 0; k < nch; ++k) {
        const double2 p = zmul_d(z[c * nch + k], nxt[k * nc + ic]);
        sr += p.x;
        si += p.y;
      }
      cfp[(int64_t)b * nb + e] = make_double2(sr, si);
    }
    __syncthreads();
  }
}

// cf[(s nr + r) + c ld1 + ic ld1 ld2] = (X2_s cfp_{s+1} - X1_s cfp_s)[c nr + r] / sqrt(a0 w_r)
__global__ void __launch_bounds__(256)
wave_kernel(int nch, int nr, int ns, const double2* __restrict__ X, int64_t ldx, int64_t sX,
            const double2* __restrict__ cfp, int nc, const double* __restrict__ wle,
            double rmax0, double2* __restrict__ cf, int64_t ld1, int64_t ld2) {
  const int n = nch * nr;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  const int ic = blockIdx.y;
  if (row >= ns * n) return;
  const int s = row / n, cr = row % n, c = cr / nr, r = cr % nr;
  const double2* x = X + s * sX + cr;
  const int nb = nch * nc;
  const double2* fa = cfp + (int64_t)(s + 1) * nb + ic;
  const double2* fb = cfp + (int64_t)s * nb + ic;
  double sr = 0.0, si = 0.0;
  for (int k = 0; k < nch; ++k) {
    const double2 p = zmul_d(x[(int64_t)(nch + k) * ldx], fa[k * nc]);
    const double2 q = zmul_d(x[(int64_t)k * ldx], fb[k * nc]);
    sr += p.x - q.x;
    si += p.y - q.y;
  }
  const double w = 1.0 / sqrt(rmax0 * wle[r]);
  cf[(int64_t)s * nr + r + (int64_t)c * ld1 + (int64_t)ic * ld1 * ld2] = make_double2(sr * w, si * w);
}

// amp[m + ic nclo] = sum_r cf_ns(r) sqrt(a0 w_r) q2_ns(r) / W_c(a) for the
// m-th closed channel c (src/rmatrix_hp.F90:418-423).
__global__ void bound_kernel(int nr, int ns, const double2* __restrict__ cf, int64_t ld1,
                             int64_t ld2, int nc, const double* __restrict__ q2_last,
                             const double* __restrict__ wle, double rmax0, int nclo,
                             const int* __restrict__ closed, const double* __restrict__ xc2,
                             double2* __restrict__ amp) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nclo * nc) return;
  const int m = e % nclo, ic = e / nclo, c = closed[m];
  const double2* f = cf + (int64_t)(ns - 1) * nr + (int64_t)c * ld1 + (int64_t)ic * ld1 * ld2;
  double sr = 0.0, si = 0.0;
  for (int r = 0; r < nr; ++r) {
    const double g = sqrt(rmax0 * wle[r]) * q2_last[r];
    sr += g * f[r].x;
    si += g * f[r].y;
  }
  amp[e] = make_double2(sr / xc2[c], si / xc2[c]);
}

struct PropLayout {
  size_t rb, crma0, zs, srow, cfp, total;
};

PropLayout prop_layout(int nch, int ns, int nc) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t n2 = (size_t)nch * nch * sizeof(double2);
  PropLayout l;
  l.rb = 0;
  l.crma0 = l.rb + al(3 * (size_t)ns * n2);
  l.zs = l.crma0 + al((size_t)ns * n2);
  l.srow = l.zs + al((size_t)ns * n2);
  l.cfp = l.srow + al(n2);
  l.total = l.cfp + al((size_t)(ns + 1) * nch * (nc > 0 ? nc : 0) * sizeof(double2)) + 256;
  return l;
}

}  // namespace

extern "C" {

size_t hprmat_hip_propagate_work_bytes(int nch, int ns, int nc) {
  if (nch <= 0 || nch > PROP_MAX_NCH || ns <= 0 || nc < 0) return 0;
  return prop_layout(nch, ns, nc).total;
}

int hprmat_hip_propagate(int nch, int nr, int ns, const void* dX, int64_t ldx, int64_t stride_x,
                         const double* q1, const double* q2, double normfac, double rmax0,
                         const double* qk, int nopen, const int* open, const void* co,
                         const void* cop, void* dR, void* dS, int nc, const int* nvc,
                         const double* wle, int nclo, const int* closed, const double* xc2,
                         void* dCf, int64_t ld1, int64_t ld2, void* dAmp, void* work,
                         size_t work_bytes, void* stream) {
  if (nch <= 0 || nch > PROP_MAX_NCH || nr <= 0 || ns <= 0 || !dX || !q1 || !q2 || !qk ||
      !(rmax0 > 0.0) || ldx < (int64_t)nch * nr || (ns > 1 && stride_x < ldx * 2 * nch) ||
      nopen < 0 || nopen > nch || (nopen > 0 && (!open || !co || !cop || !dS)) || nc < 0 ||
      nclo < 0 || nopen + nclo > nch || !work)
    return -1;
  if (nc > 0 && (nopen == 0 || !nvc || !wle || !dCf || ld1 < (int64_t)nr * ns || ld2 < nch ||
                 (nclo > 0 && (!closed || !xc2 || !dAmp))))
    return -1;
  const PropLayout l = prop_layout(nch, ns, nc);
  if (work_bytes < l.total) return -2;
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute((const void*)propagate_kernel,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            2 * PROP_MAX_NCH * PROP_MAX_NCH * (int)sizeof(double2)) != hipSuccess) {
      (void)hipGetLastError();
      return -1;
    }
    attr = true;
  }
  hipStream_t st = (hipStream_t)stream;
  char* base = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  double2* Rb = (double2*)(base + l.rb);
  double2* crma0 = (double2*)(base + l.crma0);
  double2* Zs = (double2*)(base + l.zs);
  double2* Srow = (double2*)(base + l.srow);
  double2* cfp = (double2*)(base + l.cfp);
  const int n2 = nch * nch;
  rblocks_kernel<<<dim3((3 * n2 + 255) / 256, ns), 256, 0, st>>>(
      nch, nr, (const double2*)dX, ldx, stride_x, q1, q2, normfac, Rb);
  propagate_kernel<<<1, 256, (size_t)2 * n2 * sizeof(double2), st>>>(
      nch, ns, rmax0, Rb, qk, nopen, open, (const double2*)co, (const double2*)cop, crma0, Zs,
      Srow, (double2*)dS, (double2*)dR, nc, nvc, cfp);
  if (nc > 0) {
    const int rows = ns * nch * nr;
    wave_kernel<<<dim3((rows + 255) / 256, nc), 256, 0, st>>>(
        nch, nr, ns, (const double2*)dX, ldx, stride_x, cfp, nc, wle, rmax0, (double2*)dCf, ld1,
        ld2);
    if (nclo > 0)
      bound_kernel<<<(nclo * nc + 255) / 256, 256, 0, st>>>(
          nr, ns, (const double2*)dCf, ld1, ld2, nc, q2 + (int64_t)(ns - 1) * nr, wle, rmax0,
          nclo, closed, xc2, (double2*)dAmp);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
