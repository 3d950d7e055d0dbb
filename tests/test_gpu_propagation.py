"""GPU parity of the multi-interval propagation and the wave functions
(hprmat_hip_propagate; rmatrix.rmatrix_intervals; hprmat_solve /
hprmat_solve_full with ns > 1) against the oracle (oracle/propagation.py)
and the golden outputs of the reference's example 2 (phase shifts for
ns = 2, 3; internal wave functions of rmat_pierre/Ex2/wf2.txt)."""
import json
import math
import os

import numpy as np
import pytest
import torch

from hprmat_amd import physics, rmatrix
from oracle import lagrange as olag, propagation as oprop
from tests.test_oracle_propagation import coupled_ws, reid

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden', 'rmat_pierre.json')
HM = 20.736 / 0.5


def _np(t):
    return t.cpu().numpy()


@pytest.mark.parametrize('solver_type', [1, 'mixed'])
@pytest.mark.parametrize('nr,ns', [(60, 1), (40, 3), (30, 4)])
def test_intervals_match_oracle(solver_type, nr, ns):
    # three channels, the third closed; entrance channels 0 and 1
    lval, qk, a = [0, 2, 1], [1.3, 0.9, -0.7], 12.0
    o = olag.rmat_ini(nr, ns, a)
    cp = coupled_ws(o['zrma'], 3)
    print('intervals nr=%d ns=%d solver %s' % (nr, ns, solver_type), flush=True)
    ref = oprop.rmatrix_ns(lval, qk, a, nr, ns, cp, nvc=[0, 1])
    out = rmatrix.rmatrix_intervals(lval, qk, a, nr, ns, cp, solver_type=solver_type, nvc=[0, 1])
    assert out['nopen'] == 2
    assert np.abs(_np(out['S']) - ref['S']).max() < 1e-10
    cf = _np(out['cf'])
    assert np.abs(cf - ref['cf']).max() < 1e-9 * np.abs(ref['cf']).max()
    assert np.abs(_np(out['amps']) - ref['cu'][2:, :2]).max() < 1e-9 * np.abs(ref['cu'][2:, :2]).max()
    r_ref = ref['crma0'][-1] * a
    assert np.abs(_np(out['R']) - r_ref).max() < 1e-10 * np.abs(r_ref).max()


def test_golden_ex2_intervals_on_gpu():
    d = json.load(open(GOLDEN))
    for g in d['ex2_reid_nn_intervals']:
        o = olag.rmat_ini(g['nr'], g['ns'], g['rmax'])
        qk = np.array([math.sqrt(g['e'] / HM)] * 2)
        out = rmatrix.rmatrix_intervals([0, 2], qk, g['rmax'], g['nr'], g['ns'],
                                        reid(g['j'], o['zrma']))
        s = _np(out['S'])
        d1 = math.atan2(s[0, 0].imag, s[0, 0].real) / 2
        d2 = math.atan2(s[1, 1].imag, s[1, 1].real) / 2
        for got, want in [(d1, g['delta1_rad']), (d2, g['delta2_rad']), (abs(s[0, 1]), g['eta12'])]:
            assert abs(got - want) <= 5e-5 * abs(want) + 1e-9, (g, got, want)
    for g in d['ex2_wave_functions']:
        o = olag.rmat_ini(g['nr'], g['ns'], g['rmax'])
        qk = np.array([math.sqrt(g['e'] / HM)] * 2)
        out = rmatrix.rmatrix_intervals([0, 2], qk, g['rmax'], g['nr'], g['ns'],
                                        reid(1, o['zrma']), nvc=[0])
        cf = _np(out['cf'])[:, :, 0]
        got = np.stack([cf[:, 0].real, cf[:, 0].imag, cf[:, 1].real, cf[:, 1].imag], 1)
        want = np.array(g['wf'])
        assert np.all(np.abs(got - want) <= 5.01e-5 * np.abs(want)), g['e']


def test_s_independent_of_interval_split_on_gpu():
    lval, qk, a = [0, 2, 1], [1.3, 0.9, 0.5], 12.0
    o = olag.rmat_ini(120, 1, a)
    ref = _np(rmatrix.rmatrix_intervals(lval, qk, a, 120, 1, coupled_ws(o['zrma'], 3))['S'])
    for nr, ns in [(60, 2), (40, 3), (24, 5)]:
        o = olag.rmat_ini(nr, ns, a)
        s = _np(rmatrix.rmatrix_intervals(lval, qk, a, nr, ns, coupled_ws(o['zrma'], 3))['S'])
        assert np.abs(s - ref).max() < 1e-7, (nr, ns)


def test_sixty_four_channels_two_intervals():
    # the largest channel count of the one-workgroup chain (LDS: 2 x 64^2 complex)
    nch, nr, ns, a = 64, 10, 2, 10.0
    o = olag.rmat_ini(nr, ns, a)
    rng = np.random.default_rng(11)
    lval = rng.integers(0, 6, nch)
    qk = np.where(np.arange(nch) % 5 == 4, -0.6, 0.4 + 0.02 * np.arange(nch))
    cp = coupled_ws(o['zrma'], nch, seed=5) * 0.05
    ref = oprop.rmatrix_ns(lval, qk, a, nr, ns, cp, nvc=[0, 7])
    out = rmatrix.rmatrix_intervals(lval, qk, a, nr, ns, cp, solver_type=1, nvc=[0, 7])
    assert out['nopen'] == ref['nopen']
    assert np.abs(_np(out['S']) - ref['S']).max() < 1e-9
    assert np.abs(_np(out['cf']) - ref['cf']).max() < 1e-8 * np.abs(ref['cf']).max()


@pytest.mark.parametrize('solver', [1, 4])
def test_c_interface_intervals_and_wave_functions(solver):
    lval, qk, a, nr, ns = [0, 2, 1], [1.3, 0.9, -0.7], 12.0, 40, 3
    o = olag.rmat_ini(nr, ns, a)
    cp = coupled_ws(o['zrma'], 3)
    ref = oprop.rmatrix_ns(lval, qk, a, nr, ns, cp, nvc=[0, 1])
    cu, nopen = physics.solve(lval, qk, None, a, nr, cp, ns=ns, solver=solver)
    assert nopen == 2
    assert np.abs(cu[:2, :2] - ref['S']).max() < 1e-10
    assert not cu[2:].any() and not cu[:, 2:].any()
    cu, cf, nopen = physics.solve_full(lval, qk, None, a, nr, cp, [1, 2], ns=ns, solver=solver)
    assert nopen == 2
    assert np.abs(cu - ref['cu']).max() < 1e-9 * np.abs(ref['cu']).max()
    assert np.abs(cf - ref['cf']).max() < 1e-9 * np.abs(ref['cf']).max()
    # and the uniform-mesh wave function built from them
    tab = physics.wavefunction(lval, qk, None, a, nr, ns, cu, nopen, cf, o['zrma'], 2, 1, 80, 0.2)
    want = oprop.wf_print(lval, qk, a, ref['cu'], 2, ref['cf'], o['zrma'], 1, 0, 80, 0.2)
    assert np.abs(tab - want).max() < 1e-8 * np.abs(want).max()


def test_c_interface_golden_ex2_wave_functions():
    for g in json.load(open(GOLDEN))['ex2_wave_functions']:
        z = physics.init(g['nr'], g['ns'], g['rmax'])
        qk = np.array([math.sqrt(g['e'] / HM)] * 2)
        cu, cf, nopen = physics.solve_full([0, 2], qk, None, g['rmax'], g['nr'], reid(1, z), [1],
                                           ns=g['ns'], solver=4)
        assert nopen == 2
        got = np.stack([cf[:, 0, 0].real, cf[:, 0, 0].imag, cf[:, 1, 0].real, cf[:, 1, 0].imag], 1)
        assert np.all(np.abs(got - np.array(g['wf'])) <= 5.01e-5 * np.abs(np.array(g['wf']))), g['e']


def _yamaguchi(nr, a):
    al, be = 0.2316053, 1.3918324
    z = olag.rmat_ini(nr, 1, a)['zrma']
    k = np.zeros((nr, nr, 1, 1), complex)
    k[:, :, 0, 0] = -2 * be * (al + be) ** 2 * np.exp(-be * (z[:, None] + z[None, :]))
    return k


def test_golden_ex5_nonlocal_on_gpu():
    # Yamaguchi separable non-local potential (the reference's example 5):
    # through rmatrix (device assembly + hprmat_hip_add_nonlocal), the
    # interval pipeline and the C entry point hprmat_solve_full
    from oracle import rmatrix as ormat
    hm = 41.472
    for g in json.load(open(GOLDEN))['ex5_yamaguchi']:
        nr, a = g['nr'], g['rmax']
        k = _yamaguchi(nr, a)
        qk = np.array([math.sqrt(g['e'] / hm)])
        cp = np.zeros((nr, 1, 1), complex)
        ref = ormat.rmatrix([0], qk, a, nr, cp, cpnl=k)['S']
        s1 = _np(rmatrix.rmatrix([0], qk, a, nr, cp, cpnl=k)['S'])
        s2 = _np(rmatrix.rmatrix_intervals([0], qk, a, nr, 1, cp, cpnl=k)['S'])
        cu, cf, nopen = physics.solve_full([0], qk, None, a, nr, cp, [1], solver=1, cpnl=k)
        assert nopen == 1
        for s in (s1, s2, cu[:1, :1]):
            assert abs(s[0, 0] - ref[0, 0]) < 1e-10
            delta = math.atan2(s[0, 0].imag, s[0, 0].real) / 2 * 180 / math.pi
            assert abs(delta - g['delta_deg']) <= 5e-5 * abs(g['delta_deg'])
            assert abs(abs(s[0, 0]) - g['modulus']) < 1e-4
