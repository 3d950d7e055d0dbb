"""GPU parity of the physics C interface (hprmat_solve, hprmat_solve_linear;
the reference's bindings/c/hprmat.h:78-166) against the FP64 oracle: the
collision matrix and its layout, high angular momenta and closed channels
(the host's Riccati-Bessel / modified-Bessel channel functions), the golden
Reid NN phase shifts, and the R-matrix of a given system."""
import json
import math
import os

import numpy as np
import pytest

from hprmat_amd import physics
from oracle import rmatrix as ormat, synth as osynth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden', 'rmat_pierre.json')


@pytest.mark.parametrize('solver,tol', [(1, 1e-10), (3, 1e-10), (4, 1e-10), (2, 1e-6)])
def test_solve_matches_oracle_small_config(solver, tol):
    # N = 240 (4 channels x 60 points, thresholds in [0, 3], E = 2)
    lval, qk, cpot, _ = osynth.system(4, 60, 15.0, 3.0, 2.0, 0)
    ref = ormat.rmatrix(lval, qk, 15.0, 60, cpot)
    print('hprmat_solve solver %d nopen %d' % (solver, ref['nopen']), flush=True)
    cu, nopen = physics.solve(lval, qk, None, 15.0, 60, cpot, solver=solver)
    no = ref['nopen']
    assert nopen == no
    assert np.abs(cu[:no, :no] - ref['S']).max() <= tol
    assert not cu[no:, :].any() and not cu[:, no:].any()


def test_solve_high_l_and_closed_channels():
    # angular momenta up to 11, rho = |k| a from 4.5 to 30, three closed
    # channels: the channel functions at the radius (Miller recurrence for
    # j_l, upward recurrences for y_l and k_l) against the oracle's closed forms
    nch, nr, a = 6, 80, 15.0
    _, _, cpot, _ = osynth.system(nch, nr, a, 3.0, 2.0, 3)
    lval = np.array([0, 2, 5, 8, 3, 11])
    qk = np.array([1.2, 0.8, -0.6, 0.3, -1.5, 2.0])
    ref = ormat.rmatrix(lval, qk, a, nr, cpot)
    cu, nopen = physics.solve(lval, qk, np.zeros(nch), a, nr, cpot, solver=1)
    assert nopen == ref['nopen'] == 4
    assert np.abs(cu[:4, :4] - ref['S']).max() <= 1e-10


def test_golden_ex2_reid_through_c_interface():
    # Reid soft-core NN phase shifts and mixing of the reference's example 2
    gold = json.load(open(GOLDEN))['ex2_reid_nn']
    hm = 20.736 / 0.5
    for g in gold:
        j, nr, a = g['j'], g['nr'], g['rmax']
        r = physics.init(nr, 1, a)
        cp = np.zeros((nr, 2, 2), complex)
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
        cp /= hm
        qk = np.array([math.sqrt(g['e'] / hm)] * 2)
        for solver in (1, 4):
            print('ex2 j=%d solver %d' % (j, solver), flush=True)
            cu, nopen = physics.solve([j - 1, j + 1], qk, None, a, nr, cp, solver=solver)
            assert nopen == 2
            d1 = math.atan2(cu[0, 0].imag, cu[0, 0].real) / 2
            d2 = math.atan2(cu[1, 1].imag, cu[1, 1].real) / 2
            for got, want in [(d1, g['delta1_rad']), (d2, g['delta2_rad']), (abs(cu[0, 1]), g['eta12'])]:
                assert abs(got - want) <= 5e-5 * abs(want) + 1e-9, (g, solver, got, want)


@pytest.mark.parametrize('solver', [1, 4])
def test_solve_linear_matches_oracle_r(solver):
    lval, qk, cpot, _ = osynth.system(4, 60, 15.0, 3.0, 2.0, 1)
    ref = ormat.rmatrix(lval, qk, 15.0, 60, cpot)
    q2 = ref['mesh']['q2'][0]
    r = physics.solve_linear(ref['C'], q2.astype(np.complex128), 4, 60, 1.0, solver)
    assert np.abs(r - ref['R']).max() <= 1e-10 * np.abs(ref['R']).max()


def test_repeated_calls_reuse_buffers_across_sizes():
    # the interface keeps its device buffers and LU handle between calls:
    # alternate sizes (grow, shrink, grow) and entry points, each checked
    import time
    cases = [(4, 60, 0), (8, 80, 1), (2, 40, 2), (8, 80, 3), (4, 60, 4)]
    times = []
    for nch, nr, seed in cases:
        lval, qk, cpot, _ = osynth.system(nch, nr, 15.0, 3.0, 2.0, seed)
        ref = ormat.rmatrix(lval, qk, 15.0, nr, cpot)
        no = ref['nopen']
        t0 = time.perf_counter()
        cu, nopen = physics.solve(lval, qk, None, 15.0, nr, cpot, solver=1)
        times.append(time.perf_counter() - t0)
        assert nopen == no and np.abs(cu[:no, :no] - ref['S']).max() <= 1e-10, (nch, nr)
        r = physics.solve_linear(ref['C'], ref['mesh']['q2'][0].astype(np.complex128), nch, nr)
        assert np.abs(r - ref['R']).max() <= 1e-10 * np.abs(ref['R']).max()
    print('hprmat_solve call times (s):', ' '.join('%.4f' % t for t in times), flush=True)
