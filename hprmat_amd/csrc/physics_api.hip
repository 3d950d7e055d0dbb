// Host-pointer physics interface with the reference's C binding names
// (bindings/c/hprmat.h:30-166, implemented in bindings/c/hprmat_c.F90:25-261):
// hprmat_init, hprmat_set_solver / hprmat_get_solver, hprmat_solve (collision
// matrix of a coupled-channel problem), hprmat_solve_full (with the wave
// functions), hprmat_wavefunction (wave function on a uniform mesh),
// hprmat_solve_linear (R-matrix of a given Lagrange-mesh system),
// hprmat_version, hprmat_solver_info.
//
// Each interval's N x N system is assembled, factored and solved on the GPU
// through this library's device entry points (hprmat_hip_assemble,
// hprmat_hip_lu_*); R, S, the propagation over intervals and the wave
// functions follow on the device (hprmat_hip_r_to_s, hprmat_hip_propagate).
// The host does the O(ns nr^2 + nch) set-up: the Lagrange-Legendre mesh
// (src/rmatrix_hp.F90:61-159) and the channel functions at the channel
// radius for eta = 0 (the Riccati-Bessel closed forms that COULFG / WHIT
// reduce to, src/rmatrix_hp.F90:224-244), and the spline of
// hprmat_wavefunction (O(nr ns + npoin)).
//
// Scope: eta = 0; several intervals or wave functions need nch <= 64, and a
// non-local potential one interval (as in the reference).  Other inputs are reported on stderr and answered with
// nopen = -1 (hprmat_solve, hprmat_solve_full) or a zero R-matrix
// (hprmat_solve_linear); the reference stops the program on errors instead.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <algorithm>
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

// The regularised Lagrange-Legendre mesh of ns intervals of (0, a)
// (src/rmatrix_hp.F90:61-159), per interval s (row-major nr x nr matrices,
// interval length h = a / ns): T + L (scaled by 1/h^2; interval 0 has the
// regularised form, the others the shifted-Legendre form), the closed-channel
// Bloch matrix that the assembly multiplies by x W'/W and subtracts (blo0 / h
// on interval 0, blo2 / ((s+1) h) - blo1 / (s h) after it; blo* scaled by
// 1/h), the radii (s + x_i) h and the basis values at the interval ends, q1
// (left, zero on interval 0) and q2 (right), scaled by 1/sqrt(h).  The odd
// (i + j) entries of blo0 change sign once per interval, as the reference's
// loop flips them on every pass (src/rmatrix_hp.F90:126-131): for even ns
// they keep the unflipped sign.
struct Mesh {
  int nr = 0, ns = 0;
  double a = 0.0, h = 0.0;
  std::vector<double> x, w;                // (0, 1) nodes and weights
  std::vector<double> tc, bloch, r, q1, q2;  // ns * nr * nr, ns * nr * nr, ns * nr x 3
};

void mesh_intervals(int nr, int ns, double a, Mesh& m) {
  legendre01(nr, m.x, m.w);
  const std::vector<double>& x = m.x;
  m.nr = nr;
  m.ns = ns;
  m.a = a;
  const double h = a / ns;
  m.h = h;
  const size_t nn2 = (size_t)nr * nr;
  m.tc.assign(ns * nn2, 0.0);
  m.bloch.assign(ns * nn2, 0.0);
  m.r.assign((size_t)ns * nr, 0.0);
  m.q1.assign((size_t)ns * nr, 0.0);
  m.q2.assign((size_t)ns * nr, 0.0);
  const double nn = (double)nr * (nr + 1);
  const double flip0 = (ns % 2) ? -1.0 : 1.0;   // sign of the odd entries of blo0
  std::vector<double> blo1(nn2), blo2(nn2);
  for (int s = 0; s < ns; ++s) {
    double* tc = &m.tc[s * nn2];
    double* bl = &m.bloch[s * nn2];
    for (int i = 0; i < nr; ++i) {
      const double xi = x[i], x2i = xi * (1.0 - xi);
      m.r[(size_t)s * nr + i] = (s + xi) * h;
      const double q2s = ((nr + i + 1) % 2) ? -1.0 : 1.0;   // odd nr, then odd 1-based i
      const double q1s = (i % 2) ? 1.0 : -1.0;
      if (s == 0) {
        m.q2[i] = q2s / std::sqrt(x2i) / std::sqrt(h);
        tc[(size_t)i * nr + i] = (4.0 * nn + 3.0 + (1.0 - 6.0 * xi) / x2i) / (3.0 * x2i) / (h * h);
        bl[(size_t)i * nr + i] = 1.0 / x2i / h / h;
      } else {
        const double q = std::sqrt(xi / (1.0 - xi));
        m.q2[(size_t)s * nr + i] = q2s * q / std::sqrt(h);
        m.q1[(size_t)s * nr + i] = q1s * (-1.0 / q) / std::sqrt(h);
        const double xlb = xi / (1.0 - xi) * (nn - 1.0 / (1.0 - xi));
        const double xla = (1.0 - xi) / xi * (-nn + 1.0 / xi);
        tc[(size_t)i * nr + i] = ((nn + 6.0 - 2.0 / x2i) / (3.0 * x2i) + xlb - xla) / (h * h);
        blo1[(size_t)i * nr + i] = (1.0 - xi) / xi / h;
        blo2[(size_t)i * nr + i] = xi / (1.0 - xi) / h;
      }
      for (int j = 0; j < i; ++j) {   // lower triangle (x_i > x_j), mirrored
        const double xj = x[j], x2j = xj * (1.0 - xj);
        const double sg = ((i + j) % 2) ? -1.0 : 1.0;
        const double d = xi - xj;
        double t;
        if (s == 0) {
          t = (nn + 1.0 + (xi + xj - 2.0 * xi * xj) / (d * d) - 1.0 / (1.0 - xi) -
               1.0 / (1.0 - xj)) / std::sqrt(x2i * x2j);
          const double b0 = (((i + j) % 2) ? flip0 : 1.0) / std::sqrt(x2i * x2j) / h / h;
          bl[(size_t)i * nr + j] = bl[(size_t)j * nr + i] = b0;
        } else {
          const double yy = std::sqrt(x2j / (x2i * x2i * x2i)) *
                            (2.0 * xi * xj + 3.0 * xi - xj - 4.0 * xi * xi) / (d * d);
          const double xlb = std::sqrt(xi * xj / (1.0 - xi) / (1.0 - xj)) * (nn - 1.0 / (1.0 - xj));
          const double xla = std::sqrt((1.0 - xi) * (1.0 - xj) / xi / xj) * (-nn + 1.0 / xj);
          t = yy + xlb - xla;
          blo1[(size_t)i * nr + j] = blo1[(size_t)j * nr + i] =
              sg * std::sqrt((1.0 - xi) * (1.0 - xj) / xi / xj) / h;
          blo2[(size_t)i * nr + j] = blo2[(size_t)j * nr + i] =
              sg * std::sqrt(xi * xj / (1.0 - xi) / (1.0 - xj)) / h;
        }
        tc[(size_t)i * nr + j] = tc[(size_t)j * nr + i] = sg * t / (h * h);
      }
    }
    if (s > 0) {
      const double r1 = s * h, r2 = (s + 1) * h;
      for (size_t e = 0; e < nn2; ++e) bl[e] = blo2[e] / r2 - blo1[e] / r1;
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

// Closed channel at rho = |k| a: rho W'/W and W for the decaying solution
// W(rho) = (2/pi) rho k_l(rho) (eta = 0; W_{0,l+1/2}(2 rho), normalised to
// e^-rho for l = 0); k_l by upward recurrence, scaled by e^rho.
void closed_channel(int l, double rho, double& logd, double& w) {
  std::vector<double> k(l + 2, 0.0);
  k[0] = 1.0 / rho;
  k[1] = 1.0 / rho + 1.0 / (rho * rho);
  for (int q = 1; q <= l; ++q) k[q + 1] = k[q - 1] + (2.0 * q + 1.0) / rho * k[q];
  const double kd = l == 0 ? -k[1] : -k[l - 1] - (l + 1.0) / rho * k[l];
  logd = (k[l] + rho * kd) / k[l];
  w = rho * k[l] * std::exp(-rho);
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

// Device buffers of the interface, kept between calls and grown when a call
// needs more, like the reference's persistent GPU buffers
// (src/cusolver_interface.cu:104-130): the k-th buffer of a call reuses slot
// k, so repeated calls (an energy scan through hprmat_solve) allocate
// nothing after the largest one.  Every call drains the stream before it
// returns (also on error paths, in the destructor), so a slot is never
// reallocated under work in flight.  One caller at a time, as for the
// reference's module state.
struct Slot {
  void* p = nullptr;
  size_t cap = 0;
};
std::vector<Slot> g_slots;

struct DevBufs {
  size_t next = 0;
  ~DevBufs() {
    if (g_stream) (void)hipStreamSynchronize(g_stream);
  }
  template <class T> T* alloc(size_t count) {
    const size_t bytes = count * sizeof(T) + 16;
    if (next >= g_slots.size()) g_slots.resize(next + 1);
    Slot& sl = g_slots[next++];
    if (sl.cap < bytes) {
      if (sl.p) (void)hipFree(sl.p);
      sl.p = nullptr;
      sl.cap = 0;
      if (hipMalloc(&sl.p, bytes) != hipSuccess) {
        sl.p = nullptr;
        return nullptr;
      }
      sl.cap = bytes;
    }
    return (T*)sl.p;
  }
};

template <class T> bool upload(T* d, const T* h, size_t count) {
  return hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, g_stream) == hipSuccess;
}

// The Lagrange-mesh parts of one interval's matrix (device arrays): the
// refinement's FP64 products are formed from them instead of the dense
// matrix (hprmat_hip_lu_set_operator).
struct MeshParts {
  int nch = 0, nr = 0;
  const double *tc = nullptr, *bloch = nullptr, *r = nullptr;
  const int* lval = nullptr;
  const double *qk = nullptr, *xc1 = nullptr;
  const void* cpot = nullptr;
};

// Intervals of at most this order are solved as one uniform batch.
constexpr int kBatchIntervalMaxN = 4096;

// The interface's LU handle, kept between calls (its buffers grow only).
hprmat_hip_lu_t g_lu = nullptr;

// Factor the column-major N x N dA and solve for the N x nrhs dX in place.
int factor_solve(const SolverPlan& p, const void* dA, int n, void* dX, int nrhs,
                 const MeshParts* mp = nullptr) {
  if (!g_lu && hprmat_hip_lu_create(&g_lu) != 0) {
    g_lu = nullptr;
    return -1;
  }
  hprmat_hip_lu_t h = g_lu;
  int rc;
  if (mp)   // the Bloch matrix carries its radius factors: unit scale
    rc = hprmat_hip_lu_set_operator(h, mp->nch, mp->nr, mp->tc, mp->bloch, mp->r, 1.0, mp->lval,
                                    mp->qk, mp->xc1, mp->cpot);
  else
    rc = hprmat_hip_lu_set_operator(h, 0, 0, nullptr, nullptr, nullptr, 1.0, nullptr, nullptr,
                                    nullptr, nullptr);
  if (!rc) rc = hprmat_hip_lu_factor(h, p.method, dA, n, n, g_stream);
  hprmat_hip_stats st;
  if (!rc) rc = hprmat_hip_lu_solve(h, dX, n, nrhs, p.max_refine, p.tol, p.criterion, 1, g_stream,
                                    &st);
  if (!rc && st.info != 0) {
    fprintf(stderr, "hprmat: zero pivot at column %d\n", st.info);
    rc = -1;
  }
  return rc;
}

// Natural cubic spline through (x, y): second derivatives (the tridiagonal
// recurrence of src/rmatrix_hp.F90:786-821 with both end conditions free)
// and the value at t (bisection for the bracketing knots, :823-841).
void spline_natural(const double* x, const double* y, int n, std::vector<double>& y2) {
  y2.assign(n, 0.0);
  std::vector<double> u(n, 0.0);
  for (int i = 1; i < n - 1; ++i) {
    const double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    const double p = sig * y2[i - 1] + 2.0;
    y2[i] = (sig - 1.0) / p;
    u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                (x[i + 1] - x[i - 1]) -
            sig * u[i - 1]) /
           p;
  }
  y2[n - 1] = 0.0;
  for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
}

double spline_at(const double* x, const double* y, const std::vector<double>& y2, int n, double t) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int k = (hi + lo) / 2;
    if (x[k] > t) hi = k;
    else lo = k;
  }
  const double h = x[hi] - x[lo];
  const double a = (x[hi] - t) / h, b = (t - x[lo]) / h;
  return a * y[lo] + b * y[hi] + ((a * a * a - a) * y2[lo] + (b * b * b - b) * y2[hi]) * h * h / 6.0;
}

size_t nn2_of(int nr) { return (size_t)nr * nr; }

// Wave-function request of hprmat_solve_full: nc entrance channels nvc
// (1-based indices among the open channels), cf (dim1 x dim2 x nc).
struct WaveOut {
  int nc = 0;
  const int* nvc = nullptr;
  zd* cf = nullptr;
  int dim1 = 0, dim2 = 0;
};

// rmatrix (src/rmatrix_hp.F90:175-431) for eta = 0 on the device: returns
// false (after a message) when the inputs are outside what is provided.
// One interval without wave functions keeps the nch-column solve and the
// library's R -> S; otherwise every interval is solved for its 2 nch
// boundary columns and hprmat_hip_propagate forms S (and the wave functions).
bool physics_solve(const char* who, int nch, const int* lval, const double* qk, const double* eta,
                   double rmax, int nr, int ns, const void* cpot, int cpot_dim1, void* cu,
                   int* nopen, int solver, const WaveOut& wv, const void* cpnl = nullptr,
                   int cpnl_dim1 = 0) {
  if (nopen) *nopen = -1;
  if (cu && nch > 0) memset(cu, 0, sizeof(zd) * (size_t)nch * nch);
  SolverPlan plan;
  if (nch <= 0 || nr <= 0 || ns <= 0 || !(rmax > 0.0) || !lval || !qk || !cpot || !cu ||
      !nopen || cpot_dim1 < nr * ns) {
    fprintf(stderr, "%s: invalid arguments\n", who);
    return false;
  }
  for (int c = 0; c < nch; ++c)
    if (eta && eta[c] != 0.0) {
      fprintf(stderr, "%s: only eta = 0 (no Coulomb interaction) is provided\n", who);
      return false;
    }
  if (!plan_for(solver > 0 ? solver : g_solver, plan)) {
    fprintf(stderr, "%s: solver type %d is not available\n", who, solver);
    return false;
  }
  const bool nonlocal = cpnl_dim1 > 0;
  if (nonlocal && (!cpnl || cpnl_dim1 < nr * nr)) {
    fprintf(stderr, "%s: invalid non-local potential arguments\n", who);
    return false;
  }
  if (nonlocal && ns != 1) {
    fprintf(stderr, "%s: ns must be 1 for non-local potentials\n", who);
    return false;
  }
  const bool chain = ns > 1 || wv.nc > 0;
  if (chain && nch > 64) {
    fprintf(stderr, "%s: more than one interval or wave functions need nch <= 64\n", who);
    return false;
  }
  Mesh m;
  mesh_intervals(nr, ns, rmax, m);
  // channel functions at the channel radius (open channels in input order)
  std::vector<int> open, closed;
  std::vector<zd> co, cop;
  std::vector<double> xc1(nch, 0.0), xc2(nch, 0.0);
  for (int c = 0; c < nch; ++c) {
    const double rho = std::fabs(qk[c]) * rmax;
    if (qk[c] > 0) {
      zd o, op;
      open_channel(lval[c], rho, o, op);
      open.push_back(c);
      co.push_back(o);
      cop.push_back(op);
    } else {
      closed.push_back(c);
      if (qk[c] < 0) closed_channel(lval[c], rho, xc1[c], xc2[c]);
    }
  }
  const int no = (int)open.size(), ncl = (int)closed.size();
  if (wv.nc > 0) {
    if (!wv.nvc || !wv.cf || wv.dim1 < nr * ns || wv.dim2 < nch || wv.nc > no) {
      fprintf(stderr, "%s: invalid wave-function arguments\n", who);
      return false;
    }
    for (int ic = 0; ic < wv.nc; ++ic)
      if (wv.nvc[ic] < 1 || wv.nvc[ic] > no) {
        fprintf(stderr, "%s: entrance channel %d is not an open channel\n", who, wv.nvc[ic]);
        return false;
      }
    for (int c : closed)
      if (qk[c] == 0.0) {
        fprintf(stderr, "%s: a channel at threshold (k = 0) has no bound-state amplitude\n", who);
        return false;
      }
  }
  const int n = nch * nr, nrhs = chain ? 2 * nch : nch;
  // potential: cpot[ir + c dim1 + c' dim1 nch] -> [ir][c][c'] over all intervals
  const zd* cp = (const zd*)cpot;
  const int rows = nr * ns;
  std::vector<zd> pot((size_t)rows * nch * nch);
  for (int i = 0; i < rows; ++i)
    for (int c = 0; c < nch; ++c)
      for (int d = 0; d < nch; ++d)
        pot[((size_t)i * nch + c) * nch + d] =
            cp[i + (size_t)c * cpot_dim1 + (size_t)d * cpot_dim1 * nch];
  // non-local kernel: cpnl[(ir nr + ir') + c dim1 + c' dim1 nch] -> [ir][ir'][c][c']
  std::vector<zd> knl(nonlocal ? nn2_of(nr) * nch * nch : 1);
  if (nonlocal) {
    const zd* kp = (const zd*)cpnl;
    for (size_t p = 0; p < nn2_of(nr); ++p)
      for (int c = 0; c < nch; ++c)
        for (int d = 0; d < nch; ++d)
          knl[(p * nch + c) * nch + d] =
              kp[p + (size_t)c * cpnl_dim1 + (size_t)d * cpnl_dim1 * nch];
  }
  // right-hand sides per interval: [q1 blocks | q2 blocks] or the q2 blocks
  std::vector<zd> rhs((size_t)ns * n * nrhs, zd(0.0, 0.0));
  for (int s = 0; s < ns; ++s)
    for (int c = 0; c < nch; ++c)
      for (int i = 0; i < nr; ++i) {
        zd* x = &rhs[(size_t)s * n * nrhs];
        if (chain) {
          x[(size_t)c * n + (size_t)c * nr + i] = m.q1[(size_t)s * nr + i];
          x[(size_t)(nch + c) * n + (size_t)c * nr + i] = m.q2[(size_t)s * nr + i];
        } else {
          x[(size_t)c * n + (size_t)c * nr + i] = m.q2[i];
        }
      }
  std::vector<int> nvc0(wv.nc > 0 ? wv.nc : 1, 0);
  for (int ic = 0; ic < wv.nc; ++ic) nvc0[ic] = wv.nvc[ic] - 1;

  if (ensure_stream()) return false;
  const size_t nn2 = (size_t)nr * nr;
  const size_t pw = chain ? hprmat_hip_propagate_work_bytes(nch, ns, wv.nc) : 0;
  const size_t cfn = wv.nc > 0 ? (size_t)wv.dim1 * wv.dim2 * wv.nc : 1;
  DevBufs d;
  double* dtc = d.alloc<double>(ns * nn2);
  double* dbl = d.alloc<double>(ns * nn2);
  double* dr = d.alloc<double>((size_t)ns * nr);
  double* dq1 = d.alloc<double>((size_t)ns * nr);
  double* dq2 = d.alloc<double>((size_t)ns * nr);
  double* dw = d.alloc<double>(nr);
  int* dl = d.alloc<int>(nch);
  double* dqk = d.alloc<double>(nch);
  double* dxc = d.alloc<double>(nch);
  double* dxc2 = d.alloc<double>(nch);
  zd* dpot = d.alloc<zd>(pot.size());
  zd* dA = d.alloc<zd>((size_t)n * n);
  zd* dX = d.alloc<zd>(rhs.size());
  int* dopen = d.alloc<int>(nch);
  int* dclosed = d.alloc<int>(nch);
  int* dnvc = d.alloc<int>(nvc0.size());
  zd* dco = d.alloc<zd>(nch);
  zd* dcop = d.alloc<zd>(nch);
  zd* dR = d.alloc<zd>((size_t)nch * nch);
  zd* dS = d.alloc<zd>((size_t)nch * nch);
  zd* dcf = d.alloc<zd>(cfn);
  zd* damp = d.alloc<zd>((size_t)nch * (wv.nc > 0 ? wv.nc : 1));
  zd* dknl = d.alloc<zd>(knl.size());
  void* dwork = chain ? d.alloc<char>(pw) : nullptr;
  if (!dtc || !dbl || !dr || !dq1 || !dq2 || !dw || !dl || !dqk || !dxc || !dxc2 || !dpot ||
      !dA || !dX || !dopen || !dclosed || !dnvc || !dco || !dcop || !dR || !dS || !dcf || !damp ||
      !dknl || (chain && !dwork)) {
    fprintf(stderr, "%s: device allocation failed\n", who);
    return false;
  }
  bool ok = upload(dtc, m.tc.data(), m.tc.size()) && upload(dbl, m.bloch.data(), m.bloch.size()) &&
            upload(dr, m.r.data(), m.r.size()) && upload(dq1, m.q1.data(), m.q1.size()) &&
            upload(dq2, m.q2.data(), m.q2.size()) && upload(dw, m.w.data(), (size_t)nr) &&
            upload(dl, lval, (size_t)nch) && upload(dqk, qk, (size_t)nch) &&
            upload(dxc, xc1.data(), (size_t)nch) && upload(dxc2, xc2.data(), (size_t)nch) &&
            upload(dpot, pot.data(), pot.size()) && upload(dX, rhs.data(), rhs.size()) &&
            upload(dnvc, nvc0.data(), nvc0.size()) && (!nonlocal || upload(dknl, knl.data(), knl.size()));
  if (no > 0)
    ok = ok && upload(dopen, open.data(), (size_t)no) && upload(dco, co.data(), (size_t)no) &&
         upload(dcop, cop.data(), (size_t)no);
  if (ncl > 0) ok = ok && upload(dclosed, closed.data(), (size_t)ncl);
  if (wv.nc > 0) ok = ok && hipMemsetAsync(dcf, 0, sizeof(zd) * cfn, g_stream) == hipSuccess;
  // each interval: assemble (the Bloch matrix already carries its radius
  // factors, hence the unit scale) and solve; the intervals are independent
  // small intervals of one order: one uniform batch (one launch chain for
  // all ns systems) with the solver type's precision and gate; systems the
  // batch leaves unconverged are re-solved below with the FP64 plan
  const bool batched = ns > 1 && !nonlocal && n <= kBatchIntervalMaxN;
  std::vector<int> redo;
  if (ok && batched) {
    const int code = plan.method == HPRMAT_PLAIN_FP64
                         ? (HPRMAT_PLAIN_FP64 | HPRMAT_EQUILIBRATE)
                         : ((plan.criterion & HPRMAT_FGMRES) ? HPRMAT_MIXED_FP32_3M
                                                             : HPRMAT_MIXED_FP32) |
                               HPRMAT_EQUILIBRATE;
    const size_t wb = hprmat_hip_batch_work_bytes(code, n, nrhs, ns);
    zd* dAall = d.alloc<zd>((size_t)ns * n * n);
    double* dres = d.alloc<double>(ns);
    int* dst = d.alloc<int>(ns);
    void* dbw = wb ? d.alloc<char>(wb) : nullptr;
    ok = dAall && dres && dst && dbw;
    for (int s = 0; ok && s < ns; ++s)
      ok = !hprmat_hip_assemble(nch, nr, dtc + s * nn2, dbl + s * nn2, dr + (size_t)s * nr, 1.0, dl,
                                dqk, dxc, dpot + (size_t)s * nr * nch * nch,
                                dAall + (size_t)s * n * n, n, g_stream);
    ok = ok && !hprmat_hip_batch_solve(code, n, nrhs, ns, dAall, n, (int64_t)n * n, dX, n,
                                       (int64_t)n * nrhs, plan.max_refine, plan.tol,
                                       HPRMAT_CRIT_REL_FRO, dres, dst, dbw, wb, g_stream);
    std::vector<int> status(ns, 0);
    ok = ok && hipMemcpyAsync(status.data(), dst, sizeof(int) * ns, hipMemcpyDeviceToHost,
                              g_stream) == hipSuccess &&
         hipStreamSynchronize(g_stream) == hipSuccess;
    for (int s = 0; ok && s < ns; ++s)
      if (status[s] != 1) redo.push_back(s);
    SolverPlan p64;
    plan_for(1, p64);
    for (int s : redo)
      ok = ok && upload(dX + (size_t)s * n * nrhs, &rhs[(size_t)s * n * nrhs], (size_t)n * nrhs) &&
           !factor_solve(p64, dAall + (size_t)s * n * n, n, dX + (size_t)s * n * nrhs, nrhs);
  }
  for (int s = 0; ok && !batched && s < ns; ++s) {
    MeshParts mp;
    mp.nch = nch;
    mp.nr = nr;
    mp.tc = dtc + s * nn2;
    mp.bloch = dbl + s * nn2;
    mp.r = dr + (size_t)s * nr;
    mp.lval = dl;
    mp.qk = dqk;
    mp.xc1 = dxc;
    mp.cpot = dpot + (size_t)s * nr * nch * nch;
    ok = !hprmat_hip_assemble(nch, nr, mp.tc, mp.bloch, mp.r, 1.0, dl, dqk, dxc,
                              (const zd*)mp.cpot, dA, n, g_stream) &&
         (!nonlocal || !hprmat_hip_add_nonlocal(nch, nr, dw, m.h, dknl, dA, n, g_stream)) &&
         !factor_solve(plan, dA, n, dX + (size_t)s * n * nrhs, nrhs, nonlocal ? nullptr : &mp);
  }
  if (ok) {
    if (chain)
      ok = !hprmat_hip_propagate(nch, nr, ns, dX, n, (int64_t)n * nrhs, dq1, dq2, 1.0, m.h, dqk, no,
                                 dopen, dco, dcop, dR, dS, wv.nc, dnvc, dw, ncl, dclosed, dxc2, dcf,
                                 wv.dim1, wv.dim2, damp, dwork, pw, g_stream);
    else
      ok = !hprmat_hip_r_to_s(nch, nr, dX, n, dq2, 1.0, rmax, dqk, no, dopen, dco, dcop, dR, dS,
                              g_stream);
  }
  if (!ok) {
    fprintf(stderr, "%s: device solve failed\n", who);
    (void)hipStreamSynchronize(g_stream);
    return false;
  }
  std::vector<zd> sh((size_t)no * no + 1), amp((size_t)ncl * (wv.nc > 0 ? wv.nc : 0) + 1);
  if (no > 0)
    ok = hipMemcpyAsync(sh.data(), dS, sizeof(zd) * (size_t)no * no, hipMemcpyDeviceToHost,
                        g_stream) == hipSuccess;
  if (ok && wv.nc > 0)
    ok = hipMemcpyAsync(wv.cf, dcf, sizeof(zd) * cfn, hipMemcpyDeviceToHost, g_stream) ==
             hipSuccess &&
         (ncl == 0 || hipMemcpyAsync(amp.data(), damp, sizeof(zd) * (size_t)ncl * wv.nc,
                                     hipMemcpyDeviceToHost, g_stream) == hipSuccess);
  if (!ok || hipStreamSynchronize(g_stream) != hipSuccess) {
    fprintf(stderr, "%s: copy of the results failed\n", who);
    return false;
  }
  // S (column-major nopen x nopen) into the leading block of cu (nch x nch),
  // the bound-state amplitudes below it in the entrance columns
  zd* out = (zd*)cu;
  for (int j = 0; j < no; ++j)
    for (int i = 0; i < no; ++i) out[i + (size_t)j * nch] = sh[i + (size_t)j * no];
  for (int ic = 0; ic < wv.nc; ++ic)
    for (int k = 0; k < ncl; ++k) out[no + k + (size_t)ic * nch] = amp[k + (size_t)ic * ncl];
  *nopen = no;
  return true;
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

int hprmat_hip_mesh(int nr, int ns, double rmax, double* tc, double* bloch, double* r,
                    double* q1, double* q2, double* wle) {
  if (nr <= 0 || ns <= 0 || !(rmax > 0.0) || !tc || !bloch || !r || !q1 || !q2 || !wle) return -1;
  Mesh m;
  mesh_intervals(nr, ns, rmax, m);
  std::copy(m.tc.begin(), m.tc.end(), tc);
  std::copy(m.bloch.begin(), m.bloch.end(), bloch);
  std::copy(m.r.begin(), m.r.end(), r);
  std::copy(m.q1.begin(), m.q1.end(), q1);
  std::copy(m.q2.begin(), m.q2.end(), q2);
  std::copy(m.w.begin(), m.w.end(), wle);
  return 0;
}

void hprmat_set_solver(int solver) { g_solver = solver; }

int hprmat_get_solver(void) { return g_solver; }

void hprmat_solve(int nch, int* lval, double* qk, double* eta, double rmax, int nr, int ns,
                  void* cpot, int cpot_dim1, void* cu, int* nopen, int solver) {
  physics_solve("hprmat_solve", nch, lval, qk, eta, rmax, nr, ns, cpot, cpot_dim1, cu, nopen,
                solver, WaveOut());
}

void hprmat_solve_full(int nch, int* lval, double* qk, double* eta, double rmax, int nr, int ns,
                       void* cpot, int cpot_dim1, void* cu, void* cf, int cf_dim1, int cf_dim2,
                       int nc_entrance, int* nvc, void* cpnl, int cpnl_dim1, int* nopen,
                       int solver) {
  WaveOut wv;
  wv.nc = nc_entrance;
  wv.nvc = nvc;
  wv.cf = (zd*)cf;
  wv.dim1 = cf_dim1;
  wv.dim2 = cf_dim2;
  if (nc_entrance <= 0 || !cf || !nvc) {
    if (nopen) *nopen = -1;
    if (cu && nch > 0) memset(cu, 0, sizeof(zd) * (size_t)nch * nch);
    fprintf(stderr, "hprmat_solve_full: invalid wave-function arguments\n");
    return;
  }
  physics_solve("hprmat_solve_full", nch, lval, qk, eta, rmax, nr, ns, cpot, cpot_dim1, cu, nopen,
                solver, wv, cpnl, cpnl_dim1);
}

void hprmat_wavefunction(int nch, int* lval, double* qk, double* eta, double rmax, int nr, int ns,
                         void* cu, int nopen, void* cf, int cf_dim1, int cf_dim2, double* zrma,
                         int iv, int nom, int npoin, double h, void* cwftab) {
  if (!lval || !qk || !cu || !cf || !zrma || !cwftab || nch <= 0 || nr <= 0 || ns <= 0 ||
      iv < 1 || iv > nch || nom < 1 || nom > nch || npoin < 0 || cf_dim1 < nr * ns ||
      cf_dim2 < nch || nopen < 0 || nopen > nch || nr * ns < 2) {
    fprintf(stderr, "hprmat_wavefunction: invalid arguments\n");
    return;
  }
  if (eta)
    for (int c = 0; c < nch; ++c)
      if (eta[c] != 0.0) {
        fprintf(stderr, "hprmat_wavefunction: only eta = 0 (no Coulomb interaction) is provided\n");
        return;
      }
  // position of channel iv among the open (nop) or closed (nclo) channels
  int nop = 0, nclo = nopen;
  for (int w = 0; w < iv; ++w) {
    if (qk[w] > 0) ++nop;
    if (qk[w] < 0) ++nclo;
  }
  const int nsr = nr * ns, c = iv - 1, e = nom - 1;
  const zd* f = (const zd*)cf + (size_t)c * cf_dim1 + (size_t)e * cf_dim1 * cf_dim2;
  std::vector<double> yr(nsr), yi(nsr), y2r, y2i;
  for (int i = 0; i < nsr; ++i) {
    yr[i] = f[i].real();
    yi[i] = f[i].imag();
  }
  spline_natural(zrma, yr.data(), nsr, y2r);
  spline_natural(zrma, yi.data(), nsr, y2i);
  const zd* u = (const zd*)cu;
  zd* out = (zd*)cwftab;
  for (int i = 1; i <= npoin; ++i) {
    const double r = i * h;
    if (r <= rmax) {
      out[i - 1] = zd(spline_at(zrma, yr.data(), y2r, nsr, r), spline_at(zrma, yi.data(), y2i, nsr, r));
    } else if (qk[c] > 0) {
      zd o, op;
      open_channel(lval[c], qk[c] * r, o, op);
      o *= std::sqrt(qk[e] / qk[c]);
      zd v = -u[(nop - 1) + (size_t)e * nch] * o;
      if (c == e) v += std::conj(o);
      out[i - 1] = v;
    } else {
      double logd = 0.0, w = 0.0;
      if (qk[c] < 0) closed_channel(lval[c], -qk[c] * r, logd, w);
      out[i - 1] = u[(nclo - 1) + (size_t)e * nch] * w;
    }
  }
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

// bindings/c/hprmat.h:187-197.  The reference drops the page-locked
// registration of the caller's matrix here; this interface never page-locks
// caller memory, so it releases what it does keep between calls: the device
// buffers, the LU handle and the stream (recreated by the next call).
void hprmat_gpu_finalize(void) {
  if (g_stream) (void)hipStreamSynchronize(g_stream);
  if (g_lu) {
    (void)hprmat_hip_lu_destroy(g_lu);
    g_lu = nullptr;
  }
  for (Slot& s : g_slots)
    if (s.p) (void)hipFree(s.p);
  g_slots.clear();
  if (g_stream) {
    (void)hipStreamDestroy(g_stream);
    g_stream = nullptr;
  }
}

}  // extern "C"
