// Flexible GMRES iterative refinement (FGMRES-IR) for A X = B with the LU
// factors of A (possibly in FP32, possibly of the equilibrated matrix) as a
// right preconditioner M^-1 (Saad, SIAM J. Sci. Comput. 14 (1993) 461;
// GMRES-based refinement: Carson & Higham, SIAM J. Sci. Comput. 40 (2018)
// A817).  Every right-hand side runs its own Arnoldi process; all columns
// advance in lockstep so each step is one block operation:
//
//   cycle:  R = B - A X (FP64), stop if the rule holds;  beta = ||R_c||,
//           V_0 = R / beta,  g = beta e_1
//   step j: Z_j = M^-1 V_j (factor precision, recorded as computed);
//           W = A Z_j (FP64);  two classical Gram-Schmidt passes against
//           V_0..V_j give column j of the Hessenberg matrix; V_j+1 = W / h;
//           complex Givens rotations keep it triangular and update g, whose
//           last entry is the residual of the correction for that column;
//           the cycle ends when sqrt(sum_c |g_j+1|^2) / ||B||_F <= margin tol
//           (margin 1/2 unless HPRMAT_HIP_GMRES_MARGIN says otherwise) or
//           after m steps.
//   update: X += Z y with R y = g (per column).
// Because the Z_j are kept (flexible variant), A Z = V H holds in FP64 and
// the residual estimate stays accurate although M^-1 is applied in FP32.
// Device flags in RefineState replace host synchronisation: `done` (the
// true residual met the rule) and `gm_halt` (the current cycle stopped)
// turn later kernels into no-ops.
#include <thread>

#include "lu.hpp"

namespace hpr {

namespace {

__device__ inline double2 cmulc(double2 a, double2 b) {   // conj(a) b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ inline double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// part[b][i][c] = sum over the rows of block b of conj(V_i[r][c]) W[r][c].
__global__ void __launch_bounds__(256)
dots_partial_kernel(int n, int nrhs, int nv, const double2* __restrict__ V, size_t vstride,
                    const double2* __restrict__ W, double2* __restrict__ part,
                    const int* __restrict__ halt) {
  if (*halt) return;
  const int rows = (n + gridDim.x - 1) / gridDim.x;
  const int r0 = blockIdx.x * rows, r1 = min(n, r0 + rows);
  for (int q = threadIdx.x; q < nv * nrhs; q += blockDim.x) {
    const int i = q / nrhs, c = q % nrhs;
    const double2* v = V + i * vstride;
    double sr = 0.0, si = 0.0;
    for (int r = r0; r < r1; ++r) {
      const double2 d = cmulc(v[(size_t)r * nrhs + c], W[(size_t)r * nrhs + c]);
      sr += d.x;
      si += d.y;
    }
    part[((size_t)blockIdx.x * nv + i) * nrhs + c] = make_double2(sr, si);
  }
}

// out[i][c] = (acc ? out : 0) + sum_b part[b][i][c], blocks in fixed order.
__global__ void dots_final_kernel(int nb, int nv, int nrhs, const double2* __restrict__ part,
                                  double2* __restrict__ out, int acc,
                                  const int* __restrict__ halt) {
  if (*halt) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nv * nrhs) return;
  double sr = 0.0, si = 0.0;
  for (int b = 0; b < nb; ++b) {
    const double2 p = part[(size_t)b * nv * nrhs + q];
    sr += p.x;
    si += p.y;
  }
  if (acc) { sr += out[q].x; si += out[q].y; }
  out[q] = make_double2(sr, si);
}

// W -= sum_i V_i h_i (per column coefficients).
__global__ void axpy_kernel(size_t cnt, int nrhs, int nv, const double2* __restrict__ V,
                            size_t vstride, const double2* __restrict__ h,
                            double2* __restrict__ W, const int* __restrict__ halt) {
  if (*halt) return;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt;
       e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % nrhs);
    double2 w = W[e];
    for (int i = 0; i < nv; ++i) {
      const double2 t = cmul2(V[i * vstride + e], h[i * nrhs + c]);
      w.x -= t.x;
      w.y -= t.y;
    }
    W[e] = w;
  }
}

// out = in (skipped when *halt).
__global__ void copy_kernel(size_t cnt, const double2* __restrict__ in, double2* __restrict__ out,
                            const int* __restrict__ halt) {
  if (*halt) return;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt;
       e += (size_t)gridDim.x * blockDim.x)
    out[e] = in[e];
}

// out = in * f[c].
__global__ void colscale_kernel(size_t cnt, int nrhs, const double2* __restrict__ in,
                                const double* __restrict__ f, double2* __restrict__ out,
                                const int* __restrict__ halt) {
  if (*halt) return;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt;
       e += (size_t)gridDim.x * blockDim.x) {
    const double s = f[e % nrhs];
    out[e] = make_double2(in[e].x * s, in[e].y * s);
  }
}

// Z = s[row] * D (factor precision -> FP64; s = null: no scaling).
template <class TF>
__global__ void widen_kernel(size_t cnt, int nrhs, const cx<TF>* __restrict__ D,
                             const double* __restrict__ s, double2* __restrict__ Z,
                             const int* __restrict__ halt) {
  if (*halt) return;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt;
       e += (size_t)gridDim.x * blockDim.x) {
    const double f = s ? s[e / nrhs] : 1.0;
    Z[e] = make_double2((double)D[e].x * f, (double)D[e].y * f);
  }
}

// Cycle start (one block): beta_c = ||R_c|| from the dots R^H R, g = beta e_1,
// inv = 1 / beta; gm_halt follows done.
__global__ void cycle_begin_kernel(int nrhs, int m, const double2* __restrict__ rr,
                                   double2* __restrict__ g, double* __restrict__ inv,
                                   RefineState* __restrict__ st) {
  const bool done = st->done != 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    st->gm_halt = done ? 1 : 0;
    st->gm_steps = 0;
  }
  if (done) return;
  for (int c = threadIdx.x; c < nrhs; c += blockDim.x) {
    const double beta = sqrt(fmax(rr[c].x, 0.0));
    g[c] = make_double2(beta, 0.0);
    for (int i = 1; i <= m; ++i) g[(size_t)i * nrhs + c] = make_double2(0.0, 0.0);
    inv[c] = beta > 0.0 ? 1.0 / beta : 0.0;
  }
}

// Step j (one block): column j of the Hessenberg matrix from the two CGS
// passes (h1 + h2) and ||w||^2, previous rotations, a new complex Givens
// rotation, g, 1/||w||, and the stopping decision of the cycle.
__global__ void __launch_bounds__(256)
arnoldi_step_kernel(int nrhs, int m, int j, const double2* __restrict__ h,
                    const double2* __restrict__ ww, double2* __restrict__ H,
                    double* __restrict__ cs, double2* __restrict__ sn, double2* __restrict__ g,
                    double* __restrict__ inv, RefineState* __restrict__ st, double thresh) {
  __shared__ double red[256];
  const bool halt = st->gm_halt != 0;
  __syncthreads();
  if (halt) return;
  double e2 = 0.0;
  for (int c = threadIdx.x; c < nrhs; c += blockDim.x) {
    double2 col[GMRES_MAX_M + 1];
    for (int i = 0; i <= j; ++i) {
      const double2 a = h[(size_t)i * nrhs + c], b = h[(size_t)(m + 1 + i) * nrhs + c];
      col[i] = make_double2(a.x + b.x, a.y + b.y);
    }
    const double hn = sqrt(fmax(ww[c].x, 0.0));
    col[j + 1] = make_double2(hn, 0.0);
    for (int i = 0; i < j; ++i) {
      const double ci = cs[(size_t)i * nrhs + c];
      const double2 si = sn[(size_t)i * nrhs + c], a = col[i], b = col[i + 1];
      const double2 sb = cmul2(si, b);
      col[i] = make_double2(ci * a.x + sb.x, ci * a.y + sb.y);
      const double2 sa = cmul2(make_double2(si.x, -si.y), a);
      col[i + 1] = make_double2(ci * b.x - sa.x, ci * b.y - sa.y);
    }
    // rotation zeroing col[j+1] (real, >= 0) against col[j]
    const double2 a = col[j];
    const double b = col[j + 1].x, na = sqrt(a.x * a.x + a.y * a.y);
    double c0;
    double2 s0, r0;
    if (b == 0.0) {
      c0 = 1.0; s0 = make_double2(0.0, 0.0); r0 = a;
    } else if (na == 0.0) {
      c0 = 0.0; s0 = make_double2(1.0, 0.0); r0 = make_double2(b, 0.0);
    } else {
      const double t = sqrt(na * na + b * b);
      const double2 ph = make_double2(a.x / na, a.y / na);
      c0 = na / t;
      s0 = make_double2(ph.x * b / t, ph.y * b / t);
      r0 = make_double2(ph.x * t, ph.y * t);
    }
    col[j] = r0;
    cs[(size_t)j * nrhs + c] = c0;
    sn[(size_t)j * nrhs + c] = s0;
    for (int i = 0; i <= j; ++i) H[((size_t)i * m + j) * nrhs + c] = col[i];
    const double2 gj = g[(size_t)j * nrhs + c];
    g[(size_t)j * nrhs + c] = make_double2(c0 * gj.x, c0 * gj.y);
    const double2 gn = cmul2(make_double2(-s0.x, s0.y), gj);   // -conj(s) g_j
    g[(size_t)(j + 1) * nrhs + c] = gn;
    e2 += gn.x * gn.x + gn.y * gn.y;
    inv[c] = hn > 0.0 ? 1.0 / hn : 0.0;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    st->sweeps += 1;
    st->gm_steps = j + 1;
    const double bf = st->b_fro != 0.0 ? st->b_fro : 1.0;
    if (sqrt(red[0]) / bf <= thresh || j + 1 == m) st->gm_halt = 1;
  }
}

// y = R^-1 g per column (back substitution over the steps taken).
__global__ void solve_y_kernel(int nrhs, int m, const double2* __restrict__ H,
                               const double2* __restrict__ g, double2* __restrict__ y,
                               const RefineState* __restrict__ st) {
  if (st->done) return;
  const int k = st->gm_steps;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nrhs; c += gridDim.x * blockDim.x) {
    double2 yy[GMRES_MAX_M];
    for (int i = k - 1; i >= 0; --i) {
      double2 t = g[(size_t)i * nrhs + c];
      for (int l = i + 1; l < k; ++l) {
        const double2 p = cmul2(H[((size_t)i * m + l) * nrhs + c], yy[l]);
        t.x -= p.x;
        t.y -= p.y;
      }
      const double2 d = H[((size_t)i * m + i) * nrhs + c];
      const double dd = d.x * d.x + d.y * d.y;
      yy[i] = dd > 0.0 ? make_double2((t.x * d.x + t.y * d.y) / dd, (t.y * d.x - t.x * d.y) / dd)
                       : make_double2(0.0, 0.0);
      y[(size_t)i * nrhs + c] = yy[i];
    }
  }
}

// X += sum_{i < steps} Z_i y_i.
__global__ void update_x_kernel(size_t cnt, int nrhs, const double2* __restrict__ Z,
                                size_t zstride, const double2* __restrict__ y,
                                double2* __restrict__ X, const RefineState* __restrict__ st) {
  if (st->done) return;
  const int k = st->gm_steps;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt;
       e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % nrhs);
    double2 x = X[e];
    for (int i = 0; i < k; ++i) {
      const double2 t = cmul2(Z[i * zstride + e], y[(size_t)i * nrhs + c]);
      x.x += t.x;
      x.y += t.y;
    }
    X[e] = x;
  }
}

void dots(int n, int nrhs, int nv, const double2* V, size_t vstride, const double2* W,
          double2* out, int acc, const GmresWork& gw, const int* halt, hipStream_t st) {
  const int nb = min(GMRES_DOT_BLOCKS, n);
  dots_partial_kernel<<<nb, 256, 0, st>>>(n, nrhs, nv, V, vstride, W, gw.part, halt);
  const int q = nv * nrhs;
  dots_final_kernel<<<(q + 255) / 256, 256, 0, st>>>(nb, nv, nrhs, gw.part, out, acc, halt);
}

}  // namespace

size_t gmres_work_bytes(int n, int nrhs, int m) {
  const size_t blk = (size_t)n * nrhs * sizeof(double2);
  const size_t small = ((size_t)(m + 1) * m + (m + 1) + m + 2 * (m + 1) + m) * nrhs *
                           sizeof(double2) +
                       ((size_t)m + 1) * nrhs * sizeof(double) +
                       (size_t)GMRES_DOT_BLOCKS * (m + 1) * nrhs * sizeof(double2);
  return (2 * (size_t)m + 2) * blk + small + 16 * 256;
}

void gmres_work_bind(GmresWork& w, void* base, int n, int nrhs, int m) {
  char* p = (char*)(((uintptr_t)base + 255) & ~(uintptr_t)255);
  auto take = [&](size_t bytes) {
    char* q = p;
    p += (bytes + 255) & ~(size_t)255;
    return q;
  };
  const size_t blk = (size_t)n * nrhs * sizeof(double2);
  w.V = (double2*)take((size_t)(m + 1) * blk);
  w.Z = (double2*)take((size_t)m * blk);
  w.W = (double2*)take(blk);
  w.H = (double2*)take((size_t)(m + 1) * m * nrhs * sizeof(double2));
  w.g = (double2*)take((size_t)(m + 1) * nrhs * sizeof(double2));
  w.sn = (double2*)take((size_t)m * nrhs * sizeof(double2));
  w.cs = (double*)take((size_t)m * nrhs * sizeof(double));
  w.h = (double2*)take((size_t)2 * (m + 1) * nrhs * sizeof(double2));
  w.y = (double2*)take((size_t)m * nrhs * sizeof(double2));
  w.inv = (double*)take((size_t)nrhs * sizeof(double));
  w.part = (double2*)take((size_t)GMRES_DOT_BLOCKS * (m + 1) * nrhs * sizeof(double2));
}

namespace {

// Fraction of the tolerance the Arnoldi residual estimate must reach to end a
// cycle (the true residual is then checked against the tolerance itself).
double gmres_margin() {
  const char* e = getenv("HPRMAT_HIP_GMRES_MARGIN");
  const double v = e ? atof(e) : 0.5;
  return v > 0.0 && v <= 1.0 ? v : 0.5;
}

// Host-side progress of one lane through the algorithm above.  advance()
// enqueues work up to the next point where the host may read a device flag
// (after a cycle's true residual: `done`; after an Arnoldi step that is not
// the cycle's last: `gm_halt`) and then returns; lane_drive() alternates
// advance() and the flag reads of one lane.
template <class TF> struct LaneRun {
  enum Stage { CYCLE, AFTER_CHECK, STEP, AFTER_STEP, FINAL, DONE };
  GmresLane<TF>* L = nullptr;
  int left = 0, mc = 0, j = 0;
  Stage stage = CYCLE;
  bool converged = false;
  int RefineState::*pending = nullptr;   // flag to read before the next advance
  FlagPoll fp;
};

template <class TF>
int lane_advance(LaneRun<TF>& r, const cx<TF>* LU, int64_t ld, int n, const cx<TF>* dinv256,
                 const int* perm, const double2* A64, int64_t lda, int m, double tol, int crit,
                 const LuPlan& plan, bool g3m, const double* scale) {
  GmresLane<TF>& L = *r.L;
  const int nrhs = L.nrhs;
  const size_t cnt = (size_t)n * nrhs;
  const int eg = (int)min((size_t)4096, (cnt + 255) / 256);
  const hipStream_t st = L.st;
  const GmresWork& gw = L.gw;
  RefineState* state = L.state;
  const int* done = &state->done;
  const int* halt = &state->gm_halt;
  r.pending = nullptr;
  for (;;) {
    switch (r.stage) {
      case LaneRun<TF>::CYCLE:
        if (r.left <= 0) { r.stage = LaneRun<TF>::FINAL; break; }
        r.mc = min(m, r.left);
        r.left -= r.mc;
        // true residual and the stopping rule
        copy_kernel<<<eg, 256, 0, st>>>(cnt, L.B, L.R, done);
        if (resid_gemm(n, nrhs, A64, lda, L.X, L.R, -1.0, 1, done, L.sw, st)) return -1;
        residual_norms(n, nrhs, L.R, L.part, state, 1, crit, tol, done, st);
        r.stage = LaneRun<TF>::AFTER_CHECK;
        r.pending = &RefineState::done;
        return 0;
      case LaneRun<TF>::AFTER_CHECK:   // not converged: start a cycle
        dots(n, nrhs, 1, L.R, 0, L.R, gw.h, 0, gw, done, st);
        cycle_begin_kernel<<<1, 256, 0, st>>>(nrhs, r.mc, gw.h, gw.g, gw.inv, state);
        colscale_kernel<<<eg, 256, 0, st>>>(cnt, nrhs, L.R, gw.inv, gw.V, halt);
        r.j = 0;
        r.stage = LaneRun<TF>::STEP;
        break;
      case LaneRun<TF>::STEP: {
        const int j = r.j, mc = r.mc;
        double2* Vj = gw.V + (size_t)j * cnt;
        double2* Zj = gw.Z + (size_t)j * cnt;
        // Z_j = M^-1 V_j
        permute_cast<double, TF>(n, nrhs, perm, Vj, L.RF, halt, st, scale);
        if (getrs_replay<TF>(L.graph, LU, ld, n, dinv256, L.RF, nrhs, nrhs, L.RY, halt, plan,
                             g3m, st))
          return -1;
        widen_kernel<TF><<<eg, 256, 0, st>>>(cnt, nrhs, L.RF, scale, Zj, halt);
        // W = A Z_j, two CGS passes, ||W||^2
        if (resid_gemm(n, nrhs, A64, lda, Zj, gw.W, 1.0, 0, halt, L.sw, st)) return -1;
        double2* h1 = gw.h;
        double2* h2 = gw.h + (size_t)(mc + 1) * nrhs;
        dots(n, nrhs, j + 1, gw.V, cnt, gw.W, h1, 0, gw, halt, st);
        axpy_kernel<<<eg, 256, 0, st>>>(cnt, nrhs, j + 1, gw.V, cnt, h1, gw.W, halt);
        dots(n, nrhs, j + 1, gw.V, cnt, gw.W, h2, 0, gw, halt, st);
        axpy_kernel<<<eg, 256, 0, st>>>(cnt, nrhs, j + 1, gw.V, cnt, h2, gw.W, halt);
        dots(n, nrhs, 1, gw.W, 0, gw.W, gw.y, 0, gw, halt, st);   // ||W||^2 into y (scratch)
        arnoldi_step_kernel<<<1, 256, 0, st>>>(nrhs, mc, j, gw.h, gw.y, gw.H, gw.cs, gw.sn, gw.g,
                                                gw.inv, state, gmres_margin() * tol);
        if (j + 1 < mc) {
          r.stage = LaneRun<TF>::AFTER_STEP;
          r.pending = &RefineState::gm_halt;
          return 0;
        }
        solve_y_kernel<<<(nrhs + 255) / 256, 256, 0, st>>>(nrhs, mc, gw.H, gw.g, gw.y, state);
        update_x_kernel<<<eg, 256, 0, st>>>(cnt, nrhs, gw.Z, cnt, gw.y, L.X, state);
        r.stage = LaneRun<TF>::CYCLE;
        break;
      }
      case LaneRun<TF>::AFTER_STEP:   // the cycle goes on
        colscale_kernel<<<eg, 256, 0, st>>>(cnt, nrhs, gw.W, gw.inv,
                                            gw.V + (size_t)(r.j + 1) * cnt, halt);
        r.j += 1;
        r.stage = LaneRun<TF>::STEP;
        break;
      case LaneRun<TF>::FINAL:   // final true residual
        if (!r.converged) {
          copy_kernel<<<eg, 256, 0, st>>>(cnt, L.B, L.R, done);
          if (resid_gemm(n, nrhs, A64, lda, L.X, L.R, -1.0, 1, done, L.sw, st)) return -1;
          residual_norms(n, nrhs, L.R, L.part, state, 1, crit, tol, done, st);
        }
        r.stage = LaneRun<TF>::DONE;
        return 0;
      case LaneRun<TF>::DONE:
        return 0;
    }
  }
}

// One lane from its first cycle to the end: enqueue up to the next flag,
// read the flag (in poll mode), go on.
template <class TF>
int lane_drive(LaneRun<TF>& r, const cx<TF>* LU, int64_t ld, int n, const cx<TF>* dinv256,
               const int* perm, const double2* A64, int64_t lda, int m, double tol, int crit,
               const LuPlan& plan, bool g3m, const double* scale) {
  if (lane_advance(r, LU, ld, n, dinv256, perm, A64, lda, m, tol, crit, plan, g3m, scale))
    return -1;
  while (r.stage != LaneRun<TF>::DONE) {
    if (r.pending && r.fp.read(r.L->state, r.L->st, r.pending)) {
      if (r.stage == LaneRun<TF>::AFTER_CHECK) {   // converged
        r.converged = true;
        r.stage = LaneRun<TF>::DONE;
        break;
      }
      // the cycle stopped early (its estimate met the rule): update X; the
      // steps it did not take return to the budget
      r.left += r.mc - (r.j + 1);
      const GmresWork& gw = r.L->gw;
      const size_t cnt = (size_t)n * r.L->nrhs;
      const int eg = (int)min((size_t)4096, (cnt + 255) / 256);
      solve_y_kernel<<<(r.L->nrhs + 255) / 256, 256, 0, r.L->st>>>(r.L->nrhs, r.mc, gw.H, gw.g,
                                                                    gw.y, r.L->state);
      update_x_kernel<<<eg, 256, 0, r.L->st>>>(cnt, r.L->nrhs, gw.Z, cnt, gw.y, r.L->X,
                                               r.L->state);
      r.stage = LaneRun<TF>::CYCLE;
    }
    if (lane_advance(r, LU, ld, n, dinv256, perm, A64, lda, m, tol, crit, plan, g3m, scale))
      return -1;
  }
  return 0;
}

}  // namespace

template <class TF>
int gmres_solve_lanes(const cx<TF>* LU, int64_t ld, int n, const cx<TF>* dinv256, const int* perm,
                      const double2* A64, int64_t lda, int max_steps, int m, double tol, int crit,
                      const LuPlan& plan, bool g3m, const double* scale, int nl,
                      GmresLane<TF>* lanes, bool poll) {
  if (m < 1 || m > GMRES_MAX_M || nl < 1 || nl > GMRES_MAX_LANES) return -1;
  LaneRun<TF> run[GMRES_MAX_LANES];
  // initial solutions X = M^-1 B
  for (int l = 0; l < nl; ++l) {
    GmresLane<TF>& L = lanes[l];
    if (L.nrhs < 1 || L.nrhs > 1024) return -1;
    const size_t cnt = (size_t)n * L.nrhs;
    const int eg = (int)min((size_t)4096, (cnt + 255) / 256);
    residual_norms(n, L.nrhs, L.B, L.part, L.state, 0, crit, tol, nullptr, L.st);
    permute_cast<double, TF>(n, L.nrhs, perm, L.B, L.RF, nullptr, L.st, scale);
    if (getrs<TF>(LU, ld, n, dinv256, L.RF, L.nrhs, L.nrhs, L.RY, nullptr, plan, g3m, L.st))
      return -1;
    widen_kernel<TF><<<eg, 256, 0, L.st>>>(cnt, L.nrhs, L.RF, scale, L.X, &L.state->gm_halt);
    run[l].L = &L;
    run[l].left = max_steps;
    run[l].fp = flag_poll(L.st, poll);
  }
  // Each lane after the first is driven by a host thread of its own: the
  // host blocks in a lane's flag reads, and a single host thread serving the
  // lanes in turn would hold a lane that is ready behind the wait for
  // another (under a kernel tracer, where a graph launch also blocks until
  // the graph has run, 8 ms per lane and step at N = 25600 with two lanes
  // of 64 columns).
  int dev = 0;
  HPR_CHECK(hipGetDevice(&dev));
  int rc[GMRES_MAX_LANES] = {};
  std::thread th[GMRES_MAX_LANES];
  bool threaded[GMRES_MAX_LANES] = {};
  for (int l = 1; l < nl; ++l) {
    try {
      th[l] = std::thread([&, l] {
        rc[l] = hipSetDevice(dev) == hipSuccess
                    ? lane_drive(run[l], LU, ld, n, dinv256, perm, A64, lda, m, tol, crit, plan,
                                 g3m, scale)
                    : -1;
      });
      threaded[l] = true;
    } catch (...) {   // no thread: the lane runs after lane 0 on this one
    }
  }
  rc[0] = lane_drive(run[0], LU, ld, n, dinv256, perm, A64, lda, m, tol, crit, plan, g3m, scale);
  for (int l = 1; l < nl; ++l) {
    if (threaded[l])
      th[l].join();
    else
      rc[l] = lane_drive(run[l], LU, ld, n, dinv256, perm, A64, lda, m, tol, crit, plan, g3m,
                         scale);
  }
  for (int l = 0; l < nl; ++l)
    if (rc[l]) return rc[l];
  HPR_CHECK(hipGetLastError());
  return 0;
}

template <class TF>
int gmres_solve(const cx<TF>* LU, int64_t ld, int n, const cx<TF>* dinv256, const int* perm,
                const double2* A64, int64_t lda, const double2* B, double2* X, int nrhs,
                int max_steps, int m, double tol, int crit, RefineState* state, double2* R,
                cx<TF>* RF, cx<TF>* RY, double* part, const GmresWork& gw, const LuPlan& plan,
                bool g3m, hipStream_t st, const double* scale, const SplitWork& sw, bool poll) {
  GmresLane<TF> lane;
  lane.B = B;
  lane.X = X;
  lane.nrhs = nrhs;
  lane.state = state;
  lane.R = R;
  lane.RF = RF;
  lane.RY = RY;
  lane.part = part;
  lane.gw = gw;
  lane.sw = sw;
  lane.st = st;
  return gmres_solve_lanes<TF>(LU, ld, n, dinv256, perm, A64, lda, max_steps, m, tol, crit, plan,
                               g3m, scale, 1, &lane, poll);
}

template int gmres_solve<double>(const cx<double>*, int64_t, int, const cx<double>*, const int*,
                                 const double2*, int64_t, const double2*, double2*, int, int, int,
                                 double, int, RefineState*, double2*, cx<double>*, cx<double>*,
                                 double*, const GmresWork&, const LuPlan&, bool, hipStream_t,
                                 const double*, const SplitWork&, bool);
template int gmres_solve<float>(const cx<float>*, int64_t, int, const cx<float>*, const int*,
                                const double2*, int64_t, const double2*, double2*, int, int, int,
                                double, int, RefineState*, double2*, cx<float>*, cx<float>*,
                                double*, const GmresWork&, const LuPlan&, bool, hipStream_t,
                                const double*, const SplitWork&, bool);
template int gmres_solve_lanes<double>(const cx<double>*, int64_t, int, const cx<double>*,
                                       const int*, const double2*, int64_t, int, int, double, int,
                                       const LuPlan&, bool, const double*, int,
                                       GmresLane<double>*, bool);
template int gmres_solve_lanes<float>(const cx<float>*, int64_t, int, const cx<float>*,
                                      const int*, const double2*, int64_t, int, int, double, int,
                                      const LuPlan&, bool, const double*, int, GmresLane<float>*,
                                      bool);

}  // namespace hpr
