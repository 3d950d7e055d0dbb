"""Pins of the multi-interval propagation / wave-function oracle
(oracle/propagation.py) against the golden outputs of the reference's
examples (rmat_pierre/Ex2: phase shifts for ns = 2, 3 and the internal wave
functions of wf2.txt), the one-interval oracle, the independence of S from
the split of (0, a) into intervals, continuity of the internal and
asymptotic wave functions, and scipy's natural cubic spline."""
import json
import math
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import lagrange, propagation as P, rmatrix as ormat

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden', 'rmat_pierre.json')
HM = 20.736 / 0.5


def reid(j, r):
    """Reid soft-core NN coupled-channel potential of the reference's example 2
    (divided by hbar^2 / 2mu)."""
    cp = np.zeros((len(r), 2, 2), complex)
    for i, ri in enumerate(r):
        xx = ri * 0.7
        zz = math.exp(-xx)
        vcen = (-10.463 * zz + 105.468 * zz ** 2 - 3187.8 * zz ** 4 + 9924.3 * zz ** 6) / xx
        vtens = (-10.463 * ((1 + 3 / xx + 3 / xx ** 2) * zz - (12 / xx + 3 / xx ** 2) * zz ** 4) / xx
                 + 351.77 * zz ** 4 / xx - 1673.5 * zz ** 6 / xx)
        vls = 708.91 * zz ** 4 / xx - 2713.1 * zz ** 6 / xx
        cp[i, 0, 0] = vcen - 2 * (j - 1) * vtens / (2 * j + 1) + (j - 1) * vls
        cp[i, 0, 1] = cp[i, 1, 0] = 6 * vtens * math.sqrt(j * (j + 1.0)) / (2 * j + 1)
        cp[i, 1, 1] = vcen - 2 * (j + 2) * vtens / (2 * j + 1) - (j + 2) * vls
    return cp / HM


def coupled_ws(r, nch, seed=0):
    """Complex Woods-Saxon couplings (symmetric in the channels)."""
    rng = np.random.default_rng(seed)
    v0 = rng.uniform(-60, -20, (nch, nch))
    w0 = rng.uniform(-8, -1, (nch, nch))
    v0, w0 = (v0 + v0.T) / 2, (w0 + w0.T) / 2
    off = np.where(np.eye(nch) > 0, 1.0, 0.2)
    f = 1 / (1 + np.exp((np.asarray(r) - 4.0) / 0.6))
    return off[None] * (v0 + 1j * w0)[None] * f[:, None, None] / 20.0


def test_golden_ex2_phase_shifts_multi_interval():
    gold = json.load(open(GOLDEN))['ex2_reid_nn_intervals']
    assert {g['ns'] for g in gold} == {2, 3}
    for g in gold:
        m = lagrange.rmat_ini(g['nr'], g['ns'], g['rmax'])
        qk = np.array([math.sqrt(g['e'] / HM)] * 2)
        s = P.rmatrix_ns([g['j'] - 1, g['j'] + 1], qk, g['rmax'], g['nr'], g['ns'],
                         reid(g['j'], m['zrma']))['S']
        d1 = math.atan2(s[0, 0].imag, s[0, 0].real) / 2
        d2 = math.atan2(s[1, 1].imag, s[1, 1].real) / 2
        for got, want in [(d1, g['delta1_rad']), (d2, g['delta2_rad']), (abs(s[0, 1]), g['eta12'])]:
            assert abs(got - want) <= 5e-5 * abs(want) + 1e-9, (g, got, want)


def test_golden_ex2_internal_wave_functions():
    # printed with es12.4: relative rounding <= 5e-5
    for g in json.load(open(GOLDEN))['ex2_wave_functions']:
        m = lagrange.rmat_ini(g['nr'], g['ns'], g['rmax'])
        assert np.allclose(m['zrma'], g['r'], atol=5e-4)
        qk = np.array([math.sqrt(g['e'] / HM)] * 2)
        out = P.rmatrix_ns([0, 2], qk, g['rmax'], g['nr'], g['ns'], reid(1, m['zrma']), nvc=[0])
        cf = out['cf'][:, :, 0]
        got = np.stack([cf[:, 0].real, cf[:, 0].imag, cf[:, 1].real, cf[:, 1].imag], 1)
        want = np.array(g['wf'])
        assert np.all(np.abs(got - want) <= 5.01e-5 * np.abs(want)), g['e']


def test_one_interval_matches_single_interval_oracle():
    m = lagrange.rmat_ini(60, 1, 12.0)
    cp = coupled_ws(m['zrma'], 3)
    lval, qk = np.array([0, 2, 1]), np.array([1.3, 0.9, -0.7])
    a = P.rmatrix_ns(lval, qk, 12.0, 60, 1, cp)
    b = ormat.rmatrix(lval, qk, 12.0, 60, cp)
    assert np.abs(a['S'] - b['S']).max() < 1e-11
    assert np.abs(a['crma'][0][2] - b['R']).max() < 1e-11 * np.abs(b['R']).max()


@pytest.mark.parametrize('qk,splits', [
    ([1.3, 0.9, 0.5], [(60, 2), (40, 3), (30, 4)]),
    # a closed channel: the reference flips the odd entries of the first
    # interval's Bloch matrix once per interval (src/rmatrix_hp.F90:126-131),
    # so only an odd number of intervals keeps its sign; ns = 3 and 5 here
    ([1.3, 0.9, -0.7], [(40, 3), (24, 5)]),
])
def test_s_independent_of_interval_split(qk, splits):
    lval, a = [0, 2, 1], 12.0
    m = lagrange.rmat_ini(120, 1, a)
    ref = P.rmatrix_ns(lval, qk, a, 120, 1, coupled_ws(m['zrma'], 3))['S']
    for nr, ns in splits:
        m = lagrange.rmat_ini(nr, ns, a)
        s = P.rmatrix_ns(lval, qk, a, nr, ns, coupled_ws(m['zrma'], 3))['S']
        assert np.abs(s - ref).max() < 1e-7, (nr, ns)


def test_even_split_with_closed_channel_follows_reference_sign():
    # the same system with ns = 2: the first interval's closed-channel Bloch
    # term has the sign pattern of the reference's second flip, so S moves
    lval, qk, a = [0, 2, 1], [1.3, 0.9, -0.7], 12.0
    m = lagrange.rmat_ini(120, 1, a)
    ref = P.rmatrix_ns(lval, qk, a, 120, 1, coupled_ws(m['zrma'], 3))['S']
    m2 = lagrange.rmat_ini(60, 2, a)
    s = P.rmatrix_ns(lval, qk, a, 60, 2, coupled_ws(m2['zrma'], 3))['S']
    assert np.abs(s - ref).max() > 1e-3
    i, j = np.triu_indices(60, 1)
    odd = (i + j) % 2 == 1
    b1 = lagrange.rmat_ini(60, 1, a)['blo0']
    assert np.allclose(m2['blo0'][i[odd], j[odd]], -2 * b1[i[odd], j[odd]])


def test_internal_wave_matches_asymptotic_form_at_the_radius():
    # Reid at 8 fm is negligible: the wave function at the last mesh point
    # (7.994 fm) equals the asymptotic form built from S (rmat_pierre/Ex2/
    # wf2.txt prints both, equal to five digits)
    m = lagrange.rmat_ini(25, 3, 8.0)
    qk = np.array([math.sqrt(12.0 / HM)] * 2)
    out = P.rmatrix_ns([0, 2], qk, 8.0, 25, 3, reid(1, m['zrma']), nvc=[0])
    r = m['zrma'][-1]
    for iv in range(2):
        ext = P.external_wave([0, 2], qk, out['cu'], 2, iv, 0, r)
        assert abs(out['cf'][-1, iv, 0] - ext) < 2e-4 * max(abs(ext), 0.05)


def test_spline_matches_scipy_natural_cubic():
    rng = np.random.default_rng(3)
    x = np.sort(rng.uniform(0, 10, 40))
    y = np.sin(x) + 0.1 * rng.standard_normal(40)
    cs = CubicSpline(x, y, bc_type='natural')
    y2 = P.spline_coeffs(x, y)
    assert np.allclose(y2, cs(x, 2), atol=1e-10)
    for t in np.linspace(x[0], x[-1], 57):
        assert abs(P.splint(x, y, y2, t) - cs(t)) < 1e-12


def test_wf_print_interpolates_inside_and_continues_outside():
    m = lagrange.rmat_ini(25, 3, 8.0)
    qk = np.array([math.sqrt(24.0 / HM)] * 2)
    out = P.rmatrix_ns([0, 2], qk, 8.0, 25, 3, reid(1, m['zrma']), nvc=[0])
    tab = P.wf_print([0, 2], qk, 8.0, out['cu'], 2, out['cf'], m['zrma'], 0, 0, 50, 0.2)
    # inside: close to the mesh values nearby; outside: the asymptotic form
    k = np.argmin(np.abs(m['zrma'] - 4.0))
    assert abs(tab[19] - out['cf'][k, 0, 0]) < 0.05 * abs(out['cf'][k, 0, 0]) + 0.05
    assert tab[45] == P.external_wave([0, 2], qk, out['cu'], 2, 0, 0, 46 * 0.2)
