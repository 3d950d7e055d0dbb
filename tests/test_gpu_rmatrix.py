"""GPU parity of the Lagrange-mesh front end: device assembly, R and S on the
device, and the solver_type dispatch, against the FP64 oracle and the
reference's golden phase shifts."""
import json
import math
import os

import numpy as np
import pytest
import torch

from hprmat_amd import rmatrix, solver, synth
from oracle import channels as och, lagrange as olag, rmatrix as ormat, synth as osynth

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden', 'rmat_pierre.json')


def oracle_case(nch, nr, emax, energy, seed, w0_zero=False):
    lval, qk, cpot, _ = osynth.system(nch, nr, 15.0, emax, energy, seed, w0_zero=w0_zero)
    return lval, qk, cpot, ormat.rmatrix(lval, qk, 15.0, nr, cpot)


def test_device_assembly_matches_oracle():
    s = synth.config('small')
    lval, qk, cpot, ref = oracle_case(4, 60, 3.0, 2.0, 0)
    dm = rmatrix.DeviceMesh(s.mesh)
    _, _, _, xc1 = rmatrix.channel_data(s.lval, s.qk, s.a)
    a_cm = rmatrix.assemble(dm, s.lval, s.qk, xc1, s.cpot)
    a = a_cm.cpu().numpy().T            # column-major storage -> matrix
    c = ref['C']
    assert np.abs(a - c).max() <= 1e-12 * np.abs(c).max()
    assert np.abs(a - a.T).max() == 0.0  # complex symmetric by construction


# Types 1 and 5 reach FP64 accuracy: |dS| <= 1e-10 against ZGESV.  Types 2, 4
# and 6 stop by the reference's own rules (2 sweeps or a 1e-10 residual:
# src/rmat_solvers.F90:176-177,521,645 and src/cusolver_interface.cu:414-424),
# so their answers are only that close to ZGESV: what is left after the
# second sweep follows the rounding of the FP32 factors (the pivot rule and
# fused products move it by about 2x on this kappa ~ 1e7 system), so the
# bound there is 1e-7.
@pytest.mark.parametrize('solver_type,max_refine,tol', [(1, 0, 1e-10), (2, 0, 1e-7), (4, 3, 1e-7),
                                                        (5, 0, 1e-10), (6, 0, 1e-7)])
def test_small_config_s_matrix(solver_type, max_refine, tol):
    # N = 240 (4 channels x 60 points), E = 2
    lval, qk, cpot, ref = oracle_case(4, 60, 3.0, 2.0, 0)
    out = rmatrix.rmatrix(lval, qk, 15.0, 60, cpot, solver_type=solver_type, max_refine=max_refine)
    s = out['S'].cpu().numpy()
    assert out['nopen'] == ref['nopen']
    ds = float(np.abs(s - ref['S']).max())
    assert ds <= tol, (solver_type, ds)
    r = out['R'].cpu().numpy()
    dr = float(np.abs(r - ref['R']).max())
    assert dr <= tol * max(1.0, float(np.abs(ref['R']).max())), (solver_type, dr)


def test_solution_matches_zgesv_small():
    lval, qk, cpot, ref = oracle_case(4, 60, 3.0, 2.0, 0)
    s = synth.config('small')
    dm = rmatrix.DeviceMesh(s.mesh)
    _, _, _, xc1 = rmatrix.channel_data(s.lval, s.qk, s.a)
    a_cm = rmatrix.assemble(dm, s.lval, s.qk, xc1, s.cpot)
    b = rmatrix.rhs_colmajor(dm, 4)
    st = solver.LU().factor(a_cm, 'fp64').solve_(b, 2, 1e-12, 'rel_fro', True)
    x = b.cpu().numpy().T
    assert np.abs(x - ref['X']).max() <= 1e-9 * np.abs(ref['X']).max()
    assert st['res_rel'] <= 1e-12


def test_mid_size_s_matrix_mixed_vs_oracle():
    # N = 3200 (16 channels x 200 points, E = 5, closed channels present)
    lval, qk, cpot, ref = oracle_case(16, 200, 8.0, 5.0, 2)
    for kind in ('plain', 'mixed'):
        out = rmatrix.rmatrix(lval, qk, 15.0, 200, cpot, solver_type=kind)
        s = out['S'].cpu().numpy()
        assert out['stats']['res_rel'] <= 1e-12
        assert np.abs(s - ref['S']).max() <= 1e-10, kind


def test_unitarity_real_potential():
    lval, qk, cpot, ref = oracle_case(8, 60, 3.0, 6.0, 5, w0_zero=True)
    out = rmatrix.rmatrix(lval, qk, 15.0, 60, cpot, solver_type=1)
    s = out['S'].cpu().numpy()
    assert out['nopen'] == 8
    assert np.abs(s.conj().T @ s - np.eye(8)).max() <= 1e-10
    assert np.abs(s - s.T).max() <= 1e-11


def test_golden_ex2_reid_on_gpu():
    gold = json.load(open(GOLDEN))['ex2_reid_nn']
    hm = 20.736 / 0.5
    for g in gold:
        j, nr, a = g['j'], g['nr'], g['rmax']
        mesh = rmatrix.rmat_ini(nr, a)
        cp = np.zeros((nr, 2, 2), complex)
        for i, r in enumerate(mesh.r):
            xx = r * 0.7
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
        out = rmatrix.rmatrix([j - 1, j + 1], qk, a, nr, cp, solver_type=1,
                              dmesh=rmatrix.DeviceMesh(mesh))
        s = out['S'].cpu().numpy()
        d1 = math.atan2(s[0, 0].imag, s[0, 0].real) / 2
        d2 = math.atan2(s[1, 1].imag, s[1, 1].real) / 2
        for got, want in [(d1, g['delta1_rad']), (d2, g['delta2_rad']), (abs(s[0, 1]), g['eta12'])]:
            assert abs(got - want) <= 5e-5 * abs(want) + 1e-9, (g, got, want)


def test_plain_fp64_matches_zgesv_on_ill_conditioned_scan_system():
    # the N = 6400 partial wave (64 channels x 100 points, thresholds in
    # [0, 3], seed 7919) of the energy scan at E = 2.288 (kappa ~ 1e9): the
    # reference's plain FP64 path (solver_type 1, gpu_solve_double_) must give
    # LAPACK zgesv's S to 1e-10; the unrefined substitution through explicit
    # diagonal-block inverses alone missed that (5e-10)
    from hprmat_amd import synth as hsynth
    e = float(np.linspace(0.5, 10.0, 256)[48])
    sysm = hsynth.System(64, 100, 15.0, 3.0, e, 7919)
    ref = ormat.rmatrix(sysm.lval, sysm.qk, 15.0, 100, sysm.cpot)
    print('plain fp64 N=6400 E=%.4f nopen=%d' % (e, ref['nopen']), flush=True)
    out = rmatrix.rmatrix(sysm.lval, sysm.qk, 15.0, 100, sysm.cpot, solver_type=1)
    assert out['nopen'] == ref['nopen'] == 50
    assert out['stats']['res_rel'] <= 1e-12
    assert np.abs(out['S'].cpu().numpy() - ref['S']).max() <= 1e-10
    # the reference ABI entry point on host arrays: same parity
    import ctypes
    from hprmat_amd import _lib
    lib = _lib.load()
    c = np.asfortranarray(ref['C'])
    b = np.asfortranarray(ref['B'].copy())
    info = ctypes.c_int(7)
    rc = lib.gpu_solve_double_(c.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p),
                               ctypes.byref(ctypes.c_int(6400)), ctypes.byref(ctypes.c_int(64)),
                               ctypes.byref(info))
    assert rc == 0 and info.value == 0
    assert np.linalg.norm(ref['B'] - ref['C'] @ b) / np.linalg.norm(ref['B']) <= 1e-12
    r = ormat.extract_r(b, ref['mesh']['q2'][0], 64)
    s = ormat.r_to_s(r, 15.0, sysm.qk, ref['npo'], ref['co'], ref['cop'])
    assert np.abs(s - ref['S']).max() <= 1e-10
