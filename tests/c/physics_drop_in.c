/*
 * A C program written against HPRMAT's C binding (bindings/c/hprmat.h): it
 * includes "hprmat.h" and uses only that header's names (HPRMAT_SOLVER_*,
 * hprmat_init, hprmat_set_solver, hprmat_get_solver, hprmat_solver_info,
 * hprmat_version, hprmat_solve, hprmat_gpu_finalize), built with
 *   gcc -Iinclude physics_drop_in.c -Lhprmat_amd -lhprmat_hip -lm
 *
 * Problem: three coupled channels (l = 0, 1, 2; thresholds 0, 0.7, 3.1;
 * E = 2, so channel 3 is closed), complex Woods-Saxon couplings, 60
 * Lagrange-Legendre points on (0, 15), one interval, eta = 0.
 *
 * Output lines (parsed by tests/test_c_fortran_drop_in.py):
 *   qk c value                 wave numbers (negative: closed)
 *   V ir c c' re im            the potential at the mesh points
 *   S solver nopen i j re im   collision matrix of each solver type
 * "--info" runs the host-only calls (no GPU needed) and exits.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hprmat.h"

#define NCH 3
#define NR 60

static double complex woods_saxon(double r, double v0, double w0, double rr, double d) {
  return (v0 + I * w0) / (1.0 + exp((r - rr) / d));
}

static int print_s(const char *tag, int solver, const double complex *cu, int nopen) {
  if (nopen <= 0) {
    printf("%s solver %d failed (nopen %d)\n", tag, solver, nopen);
    return 1;
  }
  for (int j = 0; j < nopen; ++j)
    for (int i = 0; i < nopen; ++i)
      printf("%s %d %d %d %d %.17e %.17e\n", tag, solver, nopen, i, j,
             creal(cu[i + j * NCH]), cimag(cu[i + j * NCH]));
  return 0;
}

int main(int argc, char **argv) {
  const int info_only = argc > 1 && strcmp(argv[1], "--info") == 0;
  const double rmax = 15.0, energy = 2.0;
  int lval[NCH] = {0, 1, 2};
  const double eps[NCH] = {0.0, 0.7, 3.1};
  double qk[NCH], eta[NCH] = {0.0, 0.0, 0.0};
  double zrma[NR];
  char buf[256];

  hprmat_version(buf, (int)sizeof buf);
  printf("version %s\n", buf);
  const int types[4] = {HPRMAT_SOLVER_DENSE, HPRMAT_SOLVER_MIXED, HPRMAT_SOLVER_WOODBURY,
                        HPRMAT_SOLVER_GPU};
  for (int t = 0; t < 4; ++t) {
    hprmat_solver_info(types[t], buf, (int)sizeof buf);
    printf("solver %d: %s\n", types[t], buf);
  }
  hprmat_set_solver(HPRMAT_SOLVER_DENSE);
  if (hprmat_get_solver() != HPRMAT_SOLVER_DENSE) {
    printf("set/get solver mismatch\n");
    return 1;
  }
  hprmat_init(NR, 1, rmax, zrma);
  printf("mesh %.17e %.17e\n", zrma[0], zrma[NR - 1]);
  if (info_only) {
    hprmat_gpu_finalize();
    printf("info ok\n");
    return 0;
  }

  for (int c = 0; c < NCH; ++c) {
    const double d = energy - eps[c];
    qk[c] = d > 0 ? sqrt(d) : -sqrt(-d);
    printf("qk %d %.17e\n", c, qk[c]);
  }
  /* pair (c, c'): depth, absorption, radius, diffuseness; off-diagonal x 0.2 */
  const double par[NCH][NCH][4] = {
      {{-42.0, -5.0, 4.1, 0.62}, {-35.0, -3.0, 3.6, 0.55}, {-47.0, -8.0, 4.6, 0.68}},
      {{-35.0, -3.0, 3.6, 0.55}, {-38.5, -6.5, 3.9, 0.60}, {-31.0, -2.5, 4.3, 0.51}},
      {{-47.0, -8.0, 4.6, 0.68}, {-31.0, -2.5, 4.3, 0.51}, {-44.0, -4.0, 3.2, 0.66}}};
  static double complex cpot[NR * NCH * NCH];
  for (int ir = 0; ir < NR; ++ir)
    for (int c = 0; c < NCH; ++c)
      for (int cp = 0; cp < NCH; ++cp) {
        const double *p = par[c][cp];
        const double s = c == cp ? 1.0 : 0.2;
        const double complex v = s * woods_saxon(zrma[ir], p[0], p[1], p[2], p[3]);
        cpot[ir + c * NR + cp * NR * NCH] = v;
        printf("V %d %d %d %.17e %.17e\n", ir, c, cp, creal(v), cimag(v));
      }

  double complex cu[NCH * NCH];
  int nopen = 0, bad = 0;
  for (int t = 0; t < 4; ++t) {
    memset(cu, 0, sizeof cu);
    hprmat_solve(NCH, lval, qk, eta, rmax, NR, 1, cpot, NR, cu, &nopen, types[t]);
    bad |= print_s("S", types[t], cu, nopen);
  }
  /* release the interface's device state, then solve again with the default
     solver type (0): everything is recreated on demand */
  hprmat_gpu_finalize();
  memset(cu, 0, sizeof cu);
  hprmat_solve(NCH, lval, qk, eta, rmax, NR, 1, cpot, NR, cu, &nopen, 0);
  bad |= print_s("S", 0, cu, nopen);
  hprmat_gpu_finalize();
  printf(bad ? "failed\n" : "done\n");
  return bad;
}
