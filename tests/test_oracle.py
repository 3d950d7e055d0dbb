"""CPU checks that pin the oracle (test infrastructure) independently of its own formulas.

* LEGZO restatement vs numpy's Gauss-Legendre rule;
* eta = 0 channel functions vs mpmath Coulomb / Whittaker functions and the
  Wronskian; the restated WHIT (asymptotic branch exact, Numerov branch
  within its method error) and the change of S it causes;
* rmat_ini's T + L(0) and boundary values vs brute-force Galerkin integrals
  of the regularised Lagrange-Legendre basis;
* the whole ns = 1 pipeline vs the golden phase shifts printed by the
  reference's examples (rmat_pierre/Ex2 Reid NN, rmat_pierre/Ex5 Yamaguchi);
* free motion (V = 0) gives S = 1; a real potential with all channels open
  gives a unitary, symmetric S;
* the unblocked GEPP restatement reproduces LAPACK zgetrf (pivots and
  factors), and the type-2 / GPU refinement rules behave as the reference's.
"""
import json
import math
import os

import mpmath
import numpy as np
import pytest
from scipy.linalg import lapack

from oracle import channels, lagrange, rmatrix, solve, synth

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden', 'rmat_pierre.json')


@pytest.mark.parametrize('n', [1, 2, 3, 10, 15, 60, 200])
def test_legzo_matches_numpy(n):
    x, w = lagrange.legzo(n)
    t, ww = np.polynomial.legendre.leggauss(n)
    assert np.allclose(x, (1 + t) / 2, atol=1e-14, rtol=0)
    assert np.allclose(w, ww / 2, atol=1e-14, rtol=0)


def _lagrange_basis(n, x):
    """The regularised Lagrange-Legendre functions on (0, 1),
    f_i(t) = (-1)^(n-i) sqrt(x_i (1 - x_i)) t P_n(2t - 1) / (x_i (t - x_i))
    (i = 1..n), with first and second derivatives, as callables of t: each
    f_i is a polynomial of degree n, and f_i(x_j) = lambda_i^-1/2 delta_ij."""
    from numpy.polynomial import legendre as leg
    cn = np.zeros(n + 1)
    cn[n] = 1.0
    d1, d2 = leg.legder(cn, 1), leg.legder(cn, 2)

    def f(i, t):
        s = 2 * t - 1
        p, p1, p2 = leg.legval(s, cn), 2 * leg.legval(s, d1), 4 * leg.legval(s, d2)
        c = (-1.0) ** (n - (i + 1)) * math.sqrt(x[i] * (1 - x[i])) / x[i]
        u, u1, u2 = t * p, p + t * p1, 2 * p1 + t * p2
        v = t - x[i]
        return c * u / v, c * (u1 / v - u / v ** 2), c * (u2 / v - 2 * u1 / v ** 2 + 2 * u / v ** 3)
    return f


@pytest.mark.parametrize('n', [8, 12, 20])
def test_rmat_ini_kinetic_bloch_matches_galerkin_integrals(n):
    # T + L(0) of rmat_ini against the brute-force Galerkin matrix elements
    # <f_i| -d2/dt2 |f_j> + f_i(1) f_j'(1) of the basis itself, integrated by
    # an independent 2n-point Gauss rule (exact: the integrand has degree
    # 2n - 2), and q2 against the basis values f_i(1)
    m = lagrange.rmat_ini(n, 1, 1.0)
    x = m['xle']
    f = _lagrange_basis(n, x)
    t, w = np.polynomial.legendre.leggauss(2 * n)
    t, w = (t + 1) / 2, w / 2
    one = np.array([1.0])
    g = np.zeros((n, n))
    for i in range(n):
        fi = f(i, t)[0]
        for j in range(n):
            g[i, j] = np.sum(w * fi * -f(j, t)[2]) + f(i, one)[0][0] * f(j, one)[1][0]
    tc = m['tc'][0]
    assert np.abs(g - tc).max() <= 1e-12 * np.abs(tc).max()
    q2 = np.array([f(i, one)[0][0] for i in range(n)])
    assert np.allclose(m['q2'][0], q2, rtol=1e-12, atol=0)


def test_mesh_symmetry_and_rank_one_bloch():
    m = lagrange.rmat_ini(40, 1, 10.0)
    assert np.allclose(m['tc'][0], m['tc'][0].T)
    # the Bloch block of the first interval is the rank-one q2 q2^T
    q2 = m['q2'][0]
    assert np.allclose(m['blo0'], np.outer(q2, q2), rtol=1e-12, atol=0)


@pytest.mark.parametrize('l', [0, 1, 2, 3])
@pytest.mark.parametrize('rho', [0.3, 2.0, 7.5, 31.0])
def test_open_channel_functions_match_mpmath(l, rho):
    co, cop = channels.open_channel(l, rho, 1.0)
    f = float(mpmath.coulombf(l, 0, rho))
    g = float(mpmath.coulombg(l, 0, rho))
    fd = float(mpmath.diff(lambda r: mpmath.coulombf(l, 0, r), rho))
    gd = float(mpmath.diff(lambda r: mpmath.coulombg(l, 0, r), rho))
    assert abs(co - complex(g, f)) <= 1e-12 * max(1.0, abs(co))
    assert abs(cop - complex(gd, fd) * rho) <= 1e-11 * max(1.0, abs(cop))
    assert abs(fd * g - f * gd - 1.0) < 1e-12          # Wronskian


@pytest.mark.parametrize('l', [0, 1, 2, 3])
@pytest.mark.parametrize('rho', [0.2, 1.5, 6.0, 25.0])
def test_closed_channel_log_derivative_matches_whittaker(l, rho):
    xc1, w = channels.closed_channel(l, rho, 1.0)
    wm = lambda z: mpmath.whitw(0, l + 0.5, 2 * z)
    ld = float(mpmath.diff(wm, rho) / wm(rho)) * rho
    assert abs(xc1 - ld) <= 1e-10 * max(1.0, abs(ld))
    assert abs(w / float(wm(rho)) - 1.0) < 1e-10


def _whit_exact(l, rho):
    wm = lambda z: mpmath.whitw(0, l + 0.5, 2 * z)
    return float(wm(rho)), float(mpmath.diff(wm, rho))


@pytest.mark.parametrize('rho', [10.0, 12.5, 25.0])
def test_whit_asymptotic_branch_is_the_closed_form(rho):
    # rho >= 10 (eta + 1): WHIT sums its series at rho; at eta = 0 every term
    # after the first vanishes and W = exp(-rho), exact; l > 0 by recurrence
    f, fd, ie = channels.whit(0.0, rho, 1.0, 3, 0)
    assert ie == 0
    for l in range(4):
        w, wd = _whit_exact(l, rho)
        assert abs(f[l] / w - 1.0) <= 1e-13 and abs(fd[l] / wd - 1.0) <= 1e-13


@pytest.mark.parametrize('rho', [0.3, 1.0, 4.64, 7.5, 9.99])
def test_whit_numerov_branch_has_the_method_error(rho):
    # rho < 10: Numerov from two series values beyond rho, the 7-point
    # difference for W', the l-recurrence.  The restatement lands within its
    # method's error of the exact function (about 1e-8: h ~ 1/30 and the
    # single-precision literals of the Fortran source), not at rounding level
    # and nowhere near O(1), which a wrong sign, index or step would give
    f, fd, _ = channels.whit(0.0, rho, 1.0, 3, 0)
    errs = []
    for l in range(4):
        w, wd = _whit_exact(l, rho)
        errs += [abs(f[l] / w - 1.0), abs(fd[l] / wd - 1.0)]
    assert max(errs) <= 1e-7
    assert max(errs) >= 1e-10


@pytest.mark.parametrize('nch,seed', [(8, 0), (8, 6), (8, 9)])
def test_whit_closed_channels_change_of_s(nch, seed):
    # the reference's closed-channel data (WHIT) against the closed form on
    # N = 480 systems of the configuration recipe (thresholds in [0, 3],
    # E = 2) with closed channels at kappa a = 2.3 .. 9.95: the
    # log-derivatives differ by up to about 1e-8 relative, and S by the
    # printed amount (measured 1.7e-14 .. 1.1e-12, below the 1e-10 S
    # tolerance; the product follows the closed form)
    lval, qk, cpot, _ = synth.system(nch, 60, 15.0, 3.0, 2.0, seed)
    mesh = lagrange.rmat_ini(60, 1, 15.0)
    npo, co, cop, xc1 = channels.channel_functions(lval, qk, 15.0)
    xw = xc1.copy()
    closed = [c for c in range(nch) if qk[c] < 0]
    assert closed and min(-qk[c] * 15.0 for c in closed) < 10.0
    for c in closed:
        xw[c] = channels.closed_channel_whit(int(lval[c]), -qk[c], 15.0)[0]
        assert abs(xw[c] / xc1[c] - 1.0) <= 1e-7
    b = rmatrix.rhs(mesh['q2'][0], nch)
    s = []
    for x1 in (xc1, xw):
        c = rmatrix.assemble(mesh, lval, qk, cpot, x1)
        x, info = solve.zgesv(c, b)
        r = rmatrix.extract_r(x, mesh['q2'][0], nch)
        s.append(rmatrix.r_to_s(r, 15.0, qk, npo, co, cop))
    ds = np.abs(s[0] - s[1]).max()
    print('N=%d seed %d: |S(WHIT) - S(closed form)| = %.2e' % (nch * 60, seed, ds))
    assert 0.0 < ds <= 1e-10


def _reid(j, nr, a):
    hm = 20.736 / 0.5
    m = lagrange.rmat_ini(nr, 1, a)
    cp = np.zeros((nr, 2, 2), complex)
    for i, r in enumerate(m['zrma']):
        xx = r * 0.7
        zz = math.exp(-xx)
        vcen = (-10.463 * zz + 105.468 * zz ** 2 - 3187.8 * zz ** 4 + 9924.3 * zz ** 6) / xx
        vtens = (-10.463 * ((1 + 3 / xx + 3 / xx ** 2) * zz - (12 / xx + 3 / xx ** 2) * zz ** 4) / xx
                 + 351.77 * zz ** 4 / xx - 1673.5 * zz ** 6 / xx)
        vls = 708.91 * zz ** 4 / xx - 2713.1 * zz ** 6 / xx
        cp[i, 0, 0] = vcen - 2 * (j - 1) * vtens / (2 * j + 1) + (j - 1) * vls
        cp[i, 0, 1] = cp[i, 1, 0] = 6 * vtens * math.sqrt(j * (j + 1.0)) / (2 * j + 1)
        cp[i, 1, 1] = vcen - 2 * (j + 2) * vtens / (2 * j + 1) - (j + 2) * vls
    return m, cp / hm, hm


def test_golden_ex2_reid():
    gold = json.load(open(GOLDEN))['ex2_reid_nn']
    for g in gold:
        m, cp, hm = _reid(g['j'], g['nr'], g['rmax'])
        qk = np.array([math.sqrt(g['e'] / hm)] * 2)
        s = rmatrix.rmatrix([g['j'] - 1, g['j'] + 1], qk, g['rmax'], g['nr'], cp, mesh=m)['S']
        d1 = math.atan2(s[0, 0].imag, s[0, 0].real) / 2
        d2 = math.atan2(s[1, 1].imag, s[1, 1].real) / 2
        for got, ref in [(d1, g['delta1_rad']), (d2, g['delta2_rad']), (abs(s[0, 1]), g['eta12'])]:
            assert abs(got - ref) <= 5e-5 * abs(ref) + 1e-9, (g, got, ref)


def _yamaguchi(nr, a):
    al, be = 0.2316053, 1.3918324
    m = lagrange.rmat_ini(nr, 1, a)
    z = m['zrma']
    cpnl = np.zeros((nr, nr, 1, 1), complex)
    cpnl[:, :, 0, 0] = -2 * be * (al + be) ** 2 * np.exp(-be * (z[:, None] + z[None, :]))
    return m, cpnl


def test_golden_ex5_yamaguchi():
    hm = 41.472
    for g in json.load(open(GOLDEN))['ex5_yamaguchi']:
        m, cpnl = _yamaguchi(g['nr'], g['rmax'])
        qk = np.array([math.sqrt(g['e'] / hm)])
        s = rmatrix.rmatrix([0], qk, g['rmax'], g['nr'], np.zeros((g['nr'], 1, 1), complex),
                            cpnl=cpnl, mesh=m)['S']
        delta = math.atan2(s[0, 0].imag, s[0, 0].real) / 2 * 180 / math.pi
        assert abs(delta - g['delta_deg']) <= 5e-5 * abs(g['delta_deg'])
        assert abs(abs(s[0, 0]) - g['modulus']) < 1e-4


@pytest.mark.parametrize('l', [0, 1, 2, 3])
def test_free_motion_gives_unit_s(l):
    nr, a = 60, 15.0
    m = lagrange.rmat_ini(nr, 1, a)
    s = rmatrix.rmatrix([l], np.array([1.3]), a, nr, np.zeros((nr, 1, 1), complex), mesh=m)['S']
    assert abs(s[0, 0] - 1.0) < 1e-8


def test_real_potential_all_open_unitary_symmetric():
    lval, qk, cpot, _ = synth.system(4, 60, 15.0, 3.0, 6.0, 3, w0_zero=True)
    out = rmatrix.rmatrix(lval, qk, 15.0, 60, cpot)
    s = out['S']
    assert out['nopen'] == 4
    assert np.allclose(s, s.T, atol=1e-11)
    assert np.allclose(s.conj().T @ s, np.eye(4), atol=1e-10)


def test_absorption_bounds_s():
    lval, qk, cpot, _ = synth.system(4, 60, 15.0, 3.0, 2.0, 0)
    s = rmatrix.rmatrix(lval, qk, 15.0, 60, cpot)['S']
    assert np.linalg.norm(s, 2) <= 1.0 + 1e-10
    assert np.allclose(s, s.T, atol=1e-11)


@pytest.mark.parametrize('n', [1, 5, 17, 40])
def test_getf2_matches_lapack_zgetrf(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    lu, piv, info = solve.getf2(a)
    lu2, piv2, info2 = lapack.zgetrf(a)
    assert info == info2 == 0
    assert np.array_equal(piv, piv2)
    assert np.allclose(lu, lu2, rtol=1e-12, atol=1e-12)


def test_getf2_zero_pivot_info_matches_lapack():
    rng = np.random.default_rng(5)
    a = rng.standard_normal((6, 6)) + 0j
    a[:, 2] = 0
    _, _, info = solve.getf2(a)
    _, _, info2 = lapack.zgetrf(a)
    assert info == info2 == 3


def test_manufactured_solution_and_exact_condition():
    # X = 1 (examples/Ex0/benchmark_mgpu.f90:71-81) on a diagonally dominant matrix
    rng = np.random.default_rng(1)
    n = 300
    a = rng.random((n, n)) + 1j * rng.random((n, n)) + 2 * n * np.eye(n)
    x, info = solve.zgesv(a, a @ np.ones((n, 2)))
    assert info == 0 and np.abs(x - 1).max() < 1e-12
    # Householder H D H with cond = kappa exactly (examples/Ex0/benchmark_mgpu_illcond.f90:41-78)
    v = rng.standard_normal(n)
    h = np.eye(n) - 2 * np.outer(v, v) / (v @ v)
    kappa = 1e6
    d = np.geomspace(1.0, 1.0 / kappa, n)
    a = h @ np.diag(d) @ h
    assert abs(np.linalg.cond(a) / kappa - 1) < 1e-6
    x, info = solve.zgesv(a, a @ np.ones(n))
    assert np.abs(x - 1).max() < 1e-16 * kappa * 100


def test_refinement_rules():
    lval, qk, cpot, _ = synth.system(4, 60, 15.0, 3.0, 2.0, 0)
    out = rmatrix.rmatrix(lval, qk, 15.0, 60, cpot)
    a, b = out['C'], out['B']
    x, used, hist = solve.mixed_cpu(a, b)
    assert used <= 2 and len(hist) >= 1
    # the type-2 rule stops once ||r||_F / ||b||_F < 1e-10
    assert hist[-1] < 1e-10 or used == 2
    x2, used2 = solve.mixed_gpu_semantics(a, b, 5, 1e-10)
    assert solve.izamax_modulus(b - a @ x2) < 1e-10 or used2 == 5
    xf, _ = solve.fp64_refined(a, b)
    assert solve.rel_residual(a, xf, b) <= 1e-12


# ---- adjoint gradient of S-matrix losses (oracle/adjoint.py) ----------------

def _fit_problem_small(nch=4, nr=16, energies=(1.5, 3.0, 6.0), seed=1):
    from oracle import adjoint, synth
    prob = adjoint.problem(nch, nr, 15.0, 3.0, seed, energies)
    _, u = synth.draw(nch, 3.0, seed, 'term')
    th_true = np.array(synth.params_from_u(u)).T
    return prob, th_true, adjoint.s_matrices(prob, th_true)


def test_adjoint_gradient_matches_central_differences():
    # pins every factor and sign of the hand-derived adjoint chain
    from oracle import adjoint
    prob, th_true, tg = _fit_problem_small()
    th = th_true * 1.05
    _, g = adjoint.loss_and_grad(prob, th, tg)
    fd = np.zeros((4, 4))
    for t in range(4):
        for p in range(4):
            h = 1e-6 * max(1.0, abs(th[t, p]))
            tp, tm = th.copy(), th.copy()
            tp[t, p] += h
            tm[t, p] -= h
            fd[t, p] = (adjoint.loss(prob, tp, tg) - adjoint.loss(prob, tm, tg)) / (2 * h)
    assert np.abs(g - fd).max() <= 1e-7 * np.abs(fd).max()


def test_adjoint_loss_vanishes_at_target():
    from oracle import adjoint
    prob, th_true, tg = _fit_problem_small()
    loss, g = adjoint.loss_and_grad(prob, th_true, tg)
    assert loss == 0.0 and np.abs(g).max() == 0.0
