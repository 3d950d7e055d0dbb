// Host-pointer physics interface with the reference's C binding names
// (bindings/c/hprmat.h:30-166, implemented in bindings/c/hprmat_c.F90:25-261):
// hprmat_init, hprmat_set_solver / hprmat_get_solver, hprmat_solve (collision
// matrix of a coupled-channel problem), hprmat_solve_linear (R-matrix of a
// given Lagrange-mesh system), hprmat_version, hprmat_solver_info.
//
// The N x N system is assembled, factored, solved and reduced to R and S on
// the GPU through this library's device entry points (hprmat_hip_assemble,
// hprmat_hip_lu_*, hprmat_hip_r_to_s); the host does the O(nr^2 + nch) set-up:
// the Lagrange-Legendre mesh of one interval (src/rmatrix_hp.F90:61-159) and
// the channel functions at the channel radius for eta = 0 (the Riccati-Bessel
// closed forms that COULFG / WHIT reduce to, src/rmatrix_hp.F90:224-244).
//
// Scope: one interval (ns = 1) and eta = 0, like the rest of this path; other
// inputs are reported on stderr and answered with nopen = -1 (hprmat_solve)
// or a zero R-matrix (hprmat_solve_linear).  The reference stops the program
// on errors instead.  hprmat_solve_full / hprmat_wavefunction (wave functions,
// src/rmatrix_hp.F90:370-431) are not provided.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hprmat_hip.h"

namespace {

typedef std::complex<double> zd;
constexpr double kPi = 3.14159265358979323846;

int g_solver = 1;                 // default solver type (reference default: 1)
hipStream_t g_stream = nullptr;   // the interface's own stream

int ensure_stream() {
  if (g_stream) return 0;
  if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) {
    fprintf(stderr, "hprmat: stream creation failed\n");
    g_stream = nullptr;
    return -1;
  }
  return 0;
}

// Gauss-Legendre nodes (ascending) and weights of n points on (0, 1):
// Newton iteration on P_n from t_i = cos(pi (i + 3/4) / (n + 1/2)), with P_n
// and P_n' from the three-term recurrence.
void legendre01(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double t = std::cos(kPi * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      const double dt = p1 / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-16) break;
    }
    // roots come out descending: store ascending on (0, 1)
    x[n - 1 - i] = 0.5 * (1.0 + t);
    w[n - 1 - i] = 1.0 / ((1.0 - t * t) * dp * dp);
  }
}

// One interval (0, a) of the regularised Lagrange-Legendre mesh: T + L(0)
// (scaled by 1/a^2), the Bloch block (1/a), the basis values at r = a
// (1/sqrt(a)) and the radii, all row-major nr x nr where matrices.
struct Mesh {
  int nr = 0;
  double a = 0.0;
  std::vector<double> r, tc, blo0, q2;
};

void mesh_one_interval(int nr, double a, Mesh& m) {
  std::vector<double> x, w;
  legendre01(nr, x, w);
  m.nr = nr;
  m.a = a;
  m.r.resize(nr);
  m.q2.resize(nr);
  m.tc.assign((size_t)nr * nr, 0.0);
  m.blo0.assign((size_t)nr * nr, 0.0);
  const double nn = (double)nr * (nr + 1);
  for (int i = 0; i < nr; ++i) {
    const double xi = x[i], x2i = xi * (1.0 - xi);
    m.r[i] = a * xi;
    m.q2[i] = (((nr + i + 1) % 2) ? -1.0 : 1.0) / std::sqrt(x2i) / std::sqrt(a);
    m.tc[(size_t)i * nr + i] = (4.0 * nn + 3.0 + (1.0 - 6.0 * xi) / x2i) / (3.0 * x2i) / (a * a);
    m.blo0[(size_t)i * nr + i] = 1.0 / x2i / a;
    for (int j = 0; j < i; ++j) {   // lower triangle, mirrored
      const double xj = x[j], x2j = xj * (1.0 - xj);
      const double sg = ((i + j) % 2) ? -1.0 : 1.0;
      const double d = xi - xj;
      const double off = (nn + 1.0 + (xi + xj - 2.0 * xi * xj) / (d * d) - 1.0 / (1.0 - xi) -
                          1.0 / (1.0 - xj)) / std::sqrt(x2i * x2j);
      m.tc[(size_t)i * nr + j] = m.tc[(size_t)j * nr + i] = sg * off / (a * a);
      m.blo0[(size_t)i * nr + j] = m.blo0[(size_t)j * nr + i] = sg / std::sqrt(x2i * x2j) / a;
    }
  }
}

// Spherical Bessel functions j_l, y_l (l = 0 .. L + 1) at rho > 0: y by
// upward recurrence (stable), j by downward (Miller) recurrence from far
// above L, rescaled on the way to stay in range and normalised with j_0 or
// j_1 (whichever is larger in magnitude; they never vanish together).
void sph_bessel(int L, double rho, std::vector<double>& j, std::vector<double>& y) {
  j.assign(L + 2, 0.0);
  y.assign(L + 2, 0.0);
  const double s = std::sin(rho), c = std::cos(rho);
  y[0] = -c / rho;
  y[1] = -c / (rho * rho) - s / rho;
  for (int l = 1; l <= L; ++l) y[l + 1] = (2.0 * l + 1.0) / rho * y[l] - y[l - 1];
  const int top = L + 2 + (int)rho + 30 + (int)std::sqrt(40.0 * (L + 2 + rho));
  std::vector<double> f(top + 2, 0.0);
  f[top] = 1e-200;
  for (int k = top; k >= 1; --k) {
    f[k - 1] = (2.0 * k + 1.0) / rho * f[k] - f[k + 1];
    if (std::fabs(f[k - 1]) > 1e200)
      for (int q = k - 1; q <= top; ++q) f[q] *= 1e-200;
  }
  const double j0 = s / rho, j1 = s / (rho * rho) - c / rho;
  const double scale = std::fabs(j0) >= std::fabs(j1) ? j0 / f[0] : j1 / f[1];
  for (int l = 0; l <= L + 1; ++l) j[l] = scale * f[l];
}

// Open channel at rho = k a: O = G + iF and rho O' with F = rho j_l,
// G = -rho y_l (eta = 0).
void open_channel(int l, double rho, zd& co, zd& cop) {
  std::vector<double> j, y;
  sph_bessel(l + 1, rho, j, y);
  const double jd = l == 0 ? -j[1] : j[l - 1] - (l + 1.0) / rho * j[l];
  const double yd = l == 0 ? -y[1] : y[l - 1] - (l + 1.0) / rho * y[l];
  co = zd(-rho * y[l], rho * j[l]);
  cop = zd(-(y[l] + rho * yd), j[l] + rho * jd) * rho;
}

// Closed channel at rho = |k| a: rho W'/W for the decaying solution W ~
// rho k_l(rho) (eta = 0); k_l scaled by e^rho (the factor cancels).
double closed_channel(int l, double rho) {
  std::vector<double> k(l + 2, 0.0);
  k[0] = 1.0 / rho;
  k[1] = 1.0 / rho + 1.0 / (rho * rho);
  for (int q = 1; q <= l; ++q) k[q + 1] = k[q - 1] + (2.0 * q + 1.0) / rho * k[q];
  const double kd = l == 0 ? -k[1] : -k[l - 1] - (l + 1.0) / rho * k[l];
  return (k[l] + rho * kd) / k[l];
}

// solver type -> factorization method flags, refinement sweeps / products,
// tolerance and criterion (the mapping of hprmat_amd.rmatrix.SOLVER_TYPES)
struct SolverPlan {
  int method, max_refine, criterion;
  double tol;
};

bool plan_for(int solver, SolverPlan& p) {
  switch (solver) {
    case 1:   // dense ZGESV: FP64 LU + <= 2 sweeps to LAPACK's accuracy
    case 3:   // Woodbury (CPU, exact in FP64): the same FP64 result
      p = {HPRMAT_PLAIN_FP64, 2, HPRMAT_CRIT_REL_FRO, 1e-12};
      return true;
    case 2:   // mixed: FP32 LU + FP64 sweeps, ||r||_F/||b||_F < 1e-10, two at most
      p = {HPRMAT_MIXED_FP32, 2, HPRMAT_CRIT_REL_FRO, 1e-10};
      return true;
    case 4:   // GPU: FP32 LU + FGMRES refinement to 1e-12 (equilibrated)
      p = {HPRMAT_MIXED_FP32 | HPRMAT_EQUILIBRATE, 24, HPRMAT_CRIT_REL_FRO | HPRMAT_FGMRES,
           1e-12};
      return true;
    default:
      return false;
  }
}

// Device buffers of one call, freed on scope exit (after the stream drains:
// an error path may return with work still queued).
struct DevBufs {
  std::vector<void*> ptrs;
  ~DevBufs() {
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T> T* alloc(size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, count * sizeof(T) + 16) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return (T*)p;
  }
};

template <class T> bool upload(T* d, const T* h, size_t count) {
  return hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, g_stream) == hipSuccess;
}

// Factor the column-major N x N dA and solve for the N x nrhs dX in place.
int factor_solve(const SolverPlan& p, const void* dA, int n, void* dX, int nrhs) {
  hprmat_hip_lu_t h = nullptr;
  if (hprmat_hip_lu_create(&h) != 0) return -1;
  int rc = hprmat_hip_lu_factor(h, p.method, dA, n, n, g_stream);
  hprmat_hip_stats st;
  if (!rc) rc = hprmat_hip_lu_solve(h, dX, n, nrhs, p.max_refine, p.tol, p.criterion, 1, g_stream,
                                    &st);
  if (!rc && st.info != 0) {
    fprintf(stderr, "hprmat: zero pivot at column %d\n", st.info);
    rc = -1;
  }
  hprmat_hip_lu_destroy(h);
  return rc;
}

}  // namespace

extern "C" {

void hprmat_init(int nr, int ns, double rmax, double* zrma) {
  if (nr <= 0 || ns <= 0 || !zrma) {
    fprintf(stderr, "hprmat_init: invalid arguments\n");
    return;
  }
  std::vector<double> x, w;
  legendre01(nr, x, w);
  const double h = rmax / ns;
  for (int s = 0; s < ns; ++s)
    for (int i = 0; i < nr; ++i) zrma[(size_t)s * nr + i] = (s + x[i]) * h;
}

void hprmat_set_solver(int solver) { g_solver = solver; }

int hprmat_get_solver(void) { return g_solver; }

void hprmat_solve(int nch, int* lval, double* qk, double* eta, double rmax, int nr, int ns,
                  void* cpot, int cpot_dim1, void* cu, int* nopen, int solver) {
  if (nopen) *nopen = -1;
  if (cu && nch > 0) memset(cu, 0, sizeof(zd) * (size_t)nch * nch);
  SolverPlan plan;
  if (nch <= 0 || nr <= 0 || !lval || !qk || !cpot || !cu || !nopen || cpot_dim1 < nr * ns) {
    fprintf(stderr, "hprmat_solve: invalid arguments\n");
    return;
  }
  if (ns != 1) {
    fprintf(stderr, "hprmat_solve: only one interval (ns = 1) is provided\n");
    return;
  }
  for (int c = 0; c < nch; ++c)
    if (eta && eta[c] != 0.0) {
      fprintf(stderr, "hprmat_solve: only eta = 0 (no Coulomb interaction) is provided\n");
      return;
    }
  if (!plan_for(solver > 0 ? solver : g_solver, plan)) {
    fprintf(stderr, "hprmat_solve: solver type %d is not available\n", solver);
    return;
  }
  if (ensure_stream()) return;
  Mesh m;
  mesh_one_interval(nr, rmax, m);
  // channel functions (open channels in their input order)
  std::vector<int> open;
  std::vector<zd> co, cop;
  std::vector<double> xc1(nch, 0.0);
  for (int c = 0; c < nch; ++c) {
    const double rho = std::fabs(qk[c]) * rmax;
    if (qk[c] > 0) {
      zd o, op;
      open_channel(lval[c], rho, o, op);
      open.push_back(c);
      co.push_back(o);
      cop.push_back(op);
    } else if (qk[c] < 0) {
      xc1[c] = closed_channel(lval[c], rho);
    }
  }
  const int no = (int)open.size();
  const int n = nch * nr;
  // potential: cpot[ir + c dim1 + c' dim1 nch] -> [ir][c][c']
  const zd* cp = (const zd*)cpot;
  std::vector<zd> pot((size_t)nr * nch * nch);
  for (int i = 0; i < nr; ++i)
    for (int c = 0; c < nch; ++c)
      for (int d = 0; d < nch; ++d)
        pot[((size_t)i * nch + c) * nch + d] =
            cp[i + (size_t)c * cpot_dim1 + (size_t)d * cpot_dim1 * nch];
  // right-hand sides: column c holds q2 in the rows of channel c
  std::vector<zd> rhs((size_t)n * nch, zd(0.0, 0.0));
  for (int c = 0; c < nch; ++c)
    for (int i = 0; i < nr; ++i) rhs[(size_t)c * n + (size_t)c * nr + i] = m.q2[i];

  DevBufs d;
  double* dtc = d.alloc<double>((size_t)nr * nr);
  double* dblo = d.alloc<double>((size_t)nr * nr);
  double* dr = d.alloc<double>(nr);
  double* dq2 = d.alloc<double>(nr);
  int* dl = d.alloc<int>(nch);
  double* dqk = d.alloc<double>(nch);
  double* dxc = d.alloc<double>(nch);
  zd* dpot = d.alloc<zd>(pot.size());
  zd* dA = d.alloc<zd>((size_t)n * n);
  zd* dX = d.alloc<zd>((size_t)n * nch);
  int* dopen = d.alloc<int>(nch);
  zd* dco = d.alloc<zd>(nch);
  zd* dcop = d.alloc<zd>(nch);
  zd* dR = d.alloc<zd>((size_t)nch * nch);
  zd* dS = d.alloc<zd>((size_t)nch * nch);
  if (!dtc || !dblo || !dr || !dq2 || !dl || !dqk || !dxc || !dpot || !dA || !dX || !dopen ||
      !dco || !dcop || !dR || !dS) {
    fprintf(stderr, "hprmat_solve: device allocation failed\n");
    return;
  }
  bool ok = upload(dtc, m.tc.data(), m.tc.size()) && upload(dblo, m.blo0.data(), m.blo0.size()) &&
            upload(dr, m.r.data(), (size_t)nr) && upload(dq2, m.q2.data(), (size_t)nr) &&
            upload(dl, lval, (size_t)nch) && upload(dqk, qk, (size_t)nch) &&
            upload(dxc, xc1.data(), (size_t)nch) && upload(dpot, pot.data(), pot.size()) &&
            upload(dX, rhs.data(), rhs.size());
  if (no > 0)
    ok = ok && upload(dopen, open.data(), (size_t)no) && upload(dco, co.data(), (size_t)no) &&
         upload(dcop, cop.data(), (size_t)no);
  if (!ok ||
      hprmat_hip_assemble(nch, nr, dtc, dblo, dr, rmax, dl, dqk, dxc, dpot, dA, n, g_stream) ||
      factor_solve(plan, dA, n, dX, nch) ||
      hprmat_hip_r_to_s(nch, nr, dX, n, dq2, 1.0, rmax, dqk, no, dopen, dco, dcop, dR, dS,
                        g_stream)) {
    fprintf(stderr, "hprmat_solve: device solve failed\n");
    (void)hipStreamSynchronize(g_stream);
    return;
  }
  std::vector<zd> s((size_t)no * no + 1);
  if (no > 0 &&
      hipMemcpyAsync(s.data(), dS, sizeof(zd) * (size_t)no * no, hipMemcpyDeviceToHost,
                     g_stream) != hipSuccess) {
    fprintf(stderr, "hprmat_solve: copy of S failed\n");
    return;
  }
  if (hipStreamSynchronize(g_stream) != hipSuccess) {
    fprintf(stderr, "hprmat_solve: device solve failed\n");
    return;
  }
  // S (column-major nopen x nopen) into the leading block of cu (nch x nch)
  zd* out = (zd*)cu;
  for (int j = 0; j < no; ++j)
    for (int i = 0; i < no; ++i) out[i + (size_t)j * nch] = s[i + (size_t)j * no];
  *nopen = no;
}

void hprmat_solve_linear(void* cmat, void* B_vector, int nch, int nlag, double normfac,
                         void* Rmat, int solver) {
  if (Rmat && nch > 0) memset(Rmat, 0, sizeof(zd) * (size_t)nch * nch);
  SolverPlan plan;
  if (!cmat || !B_vector || !Rmat || nch <= 0 || nlag <= 0) {
    fprintf(stderr, "hprmat_solve_linear: invalid arguments\n");
    return;
  }
  if (!plan_for(solver > 0 ? solver : g_solver, plan)) {
    fprintf(stderr, "hprmat_solve_linear: solver type %d is not available\n", solver);
    return;
  }
  if (ensure_stream()) return;
  const int n = nch * nlag;
  const zd* b = (const zd*)B_vector;
  std::vector<zd> rhs((size_t)n * nch, zd(0.0, 0.0));
  for (int c = 0; c < nch; ++c)
    for (int r = 0; r < nlag; ++r) rhs[(size_t)c * n + (size_t)c * nlag + r] = b[r];
  DevBufs d;
  zd* dA = d.alloc<zd>((size_t)n * n);
  zd* dX = d.alloc<zd>((size_t)n * nch);
  if (!dA || !dX) {
    fprintf(stderr, "hprmat_solve_linear: device allocation failed\n");
    return;
  }
  if (!upload(dA, (const zd*)cmat, (size_t)n * n) || !upload(dX, rhs.data(), rhs.size()) ||
      factor_solve(plan, dA, n, dX, nch) ||
      hipMemcpyAsync(rhs.data(), dX, sizeof(zd) * rhs.size(), hipMemcpyDeviceToHost, g_stream) !=
          hipSuccess ||
      hipStreamSynchronize(g_stream) != hipSuccess) {
    fprintf(stderr, "hprmat_solve_linear: device solve failed\n");
    return;
  }
  // R(i, j) = normfac sum_r B(r) X((i nlag + r), j), column-major nch x nch
  zd* R = (zd*)Rmat;
  for (int j = 0; j < nch; ++j)
    for (int i = 0; i < nch; ++i) {
      zd acc(0.0, 0.0);
      for (int r = 0; r < nlag; ++r) acc += b[r] * rhs[(size_t)j * n + (size_t)i * nlag + r];
      R[i + (size_t)j * nch] = normfac * acc;
    }
}

void hprmat_version(char* version, int maxlen) {
  if (!version || maxlen <= 0) return;
  snprintf(version, (size_t)maxlen, "%s", hprmat_hip_version());
}

void hprmat_solver_info(int solver, char* info, int maxlen) {
  if (!info || maxlen <= 0) return;
  const char* s;
  switch (solver) {
    case 1: s = "Dense FP64 LU on the GPU (+ FP64 refinement to LAPACK accuracy)"; break;
    case 2: s = "Mixed precision: FP32 LU on the GPU + FP64 refinement (rel. 1e-10)"; break;
    case 3: s = "Woodbury (CPU) served by the dense FP64 GPU solve"; break;
    case 4: s = "GPU: FP32 LU on MFMA + FP64 FGMRES refinement (rel. 1e-12)"; break;
    default: s = "Unknown solver"; break;
  }
  snprintf(info, (size_t)maxlen, "%s", s);
}

}  // extern "C"
