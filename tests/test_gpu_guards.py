"""Failure reporting on the GPU path.

* Classic refinement that makes the residual grow stops, takes the last
  correction back and reports "not converged, diverged" (the FP32 factors of
  an exact-condition kappa = 1e12 Householder matrix: on the CPU, LAPACK
  cgetrf + FP64 refinement grows the relative residual from 1.9e-7 to
  2.3e-7 to 1.1e-6 ...).
* gpu_solve_tf32_ (split-BF16 factors, FGMRES refinement, the reference's
  ||r||_max > 1e-6 (||b||_max + 1) guard of src/cusolver_interface.cu:
  1558-1591) on the kappa ~ 1e9 N = 6400 scan system either converges or
  returns info 99 with B untouched, and returns 99 on the kappa = 1e12 matrix.
* A panel leaf whose cross-workgroup wait gives up (forced from the host with
  HPRMAT_HIP_FORCE_BARRIER_TIMEOUT=1: the error flag starts raised, so waits
  end after their first poll interval and every kernel finishes) is reported
  as stats.info = -3, hprmat_hip_lu_info = -3, gpu_solve_* info -3 with B
  untouched, and batch status HPRMAT_STATUS_TIMEOUT; without the variable the
  next factorization is correct again.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

from hprmat_amd import _lib, solver
from oracle import rmatrix as ormat

pytestmark = pytest.mark.gpu

I = ctypes.c_int
D = ctypes.c_double


def householder(n=200, kappa=1e12, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n)
    h = np.eye(n) - 2 * np.outer(v, v) / (v @ v)
    return (h @ np.diag(np.geomspace(1, 1 / kappa, n)) @ h).astype(complex)


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def host_call(name, a, b, max_refine=None, tol=None):
    lib = _lib.load()
    a = np.asfortranarray(a)
    bb = np.asfortranarray(b.copy())
    info = I(-99)
    n, r = I(a.shape[0]), I(b.shape[1])
    if max_refine is None:
        rc = getattr(lib, name)(ptr(a), ptr(bb), ctypes.byref(n), ctypes.byref(r), ctypes.byref(info))
    else:
        rc = getattr(lib, name)(ptr(a), ptr(bb), ctypes.byref(n), ctypes.byref(r),
                                ctypes.byref(I(max_refine)), ctypes.byref(D(tol)), ctypes.byref(info))
    return rc, info.value, bb


def test_classic_refinement_stops_when_the_residual_grows():
    a = householder()
    b = a @ np.ones((a.shape[0], 2))
    print('FP32 IR on kappa=1e12 Householder n=200', flush=True)
    lu = solver.LU().factor(torch.from_numpy(a.T.copy()).cuda(), 'fp32')
    b_cm = torch.from_numpy(b.T.copy()).cuda()
    st = lu.solve_(b_cm, 10, 1e-12, 'rel_fro', True)
    h = st['hist']
    assert st['diverged'] and not st['converged'], st
    # stopped at the first evaluation that grew; the state keeps the
    # previous (better) evaluation, and X is that iterate again
    assert h[-1] > h[-2] and st['res_rel'] == h[-2] == min(h)
    assert all(h[i + 1] <= h[i] for i in range(len(h) - 2))
    assert st['sweeps'] == st['checks'] - 2
    x = b_cm.cpu().numpy().T
    rel = np.linalg.norm(b - a @ x) / np.linalg.norm(b)
    assert abs(rel - st['res_rel']) <= 0.05 * st['res_rel'], (rel, st['res_rel'])


def test_tf32_entry_converges_or_reports_99_on_ill_conditioned_scan_system():
    from hprmat_amd import synth as hsynth
    e = float(np.linspace(0.5, 10.0, 256)[48])
    sysm = hsynth.System(64, 100, 15.0, 3.0, e, 7919)
    ref = ormat.rmatrix(sysm.lval, sysm.qk, 15.0, 100, sysm.cpot)
    c, b = ref['C'], ref['B']
    bmax = np.abs(b).max()
    for refine in (5, 12):
        print('gpu_solve_tf32_ N=6400 kappa~1e9, max_refine %d' % refine, flush=True)
        rc, info, x = host_call('gpu_solve_tf32_', c, b, refine, 1e-10)
        if info == 99:
            assert rc == -1 and np.array_equal(x, b)
            continue
        assert rc == 0 and info == 0
        r = b - c @ x
        assert np.abs(r).max() <= 1e-6 * (bmax + 1.0)
        # no worse than the unrefined residual of the split-BF16 factors
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-2
    rc, info, x = host_call('gpu_solve_tf32_', c, b, 12, 1e-10)
    assert rc == 0 and info == 0   # 12 products pass the guard on this system


def test_tf32_entry_reports_99_and_leaves_b_on_hopeless_system():
    # Wilkinson's growth matrix: the reduced-precision factors carry no
    # accuracy (pivot growth 2^99), two FGMRES products cannot fix that
    from test_gpu_abi import wilkinson
    a = wilkinson()
    b = a @ np.ones((a.shape[0], 1))
    rc, info, x = host_call('gpu_solve_tf32_', a, b, 2, 1e-12)
    assert rc == -1 and info == 99
    assert np.array_equal(x, b)


@pytest.fixture
def forced_timeout():
    os.environ['HPRMAT_HIP_FORCE_BARRIER_TIMEOUT'] = '1'
    try:
        yield
    finally:
        del os.environ['HPRMAT_HIP_FORCE_BARRIER_TIMEOUT']


def _system(n, seed=1):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return a, rng.standard_normal((n, 3)) + 0j


def test_barrier_timeout_is_reported(forced_timeout):
    # n = 700: panel leaves of three workgroups (rows > 256), so the pivot
    # search waits on other workgroups
    a, b = _system(700)
    print('forced barrier timeout: LU handle n=700', flush=True)
    lu = solver.LU().factor(torch.from_numpy(a.T.copy()).cuda(), 'fp32')
    b_cm = torch.from_numpy(b.T.copy()).cuda()
    st = lu.solve_(b_cm, 2, 1e-12, 'rel_fro', True)
    assert st['info'] == -3 and not st['converged']
    v = ctypes.c_int()
    assert _lib.load().hprmat_hip_lu_info(lu._h, ctypes.byref(v)) == -3
    print('forced barrier timeout: gpu_solve_double_ n=700', flush=True)
    rc, info, x = host_call('gpu_solve_double_', a, b)
    assert rc == -1 and info == -3 and np.array_equal(x, b)
    print('forced barrier timeout: batch of 2, n=700', flush=True)
    bt = solver.Batch(700, 3, 2, 'fp32')
    a_cm = torch.from_numpy(np.stack([a.T.copy(), a.T.copy()])).cuda()
    b_cm2 = torch.from_numpy(np.stack([b.T.copy(), b.T.copy()])).cuda()
    bt.solve_(a_cm, b_cm2, 2, 1e-12)
    torch.cuda.synchronize()
    assert bt.status.cpu().tolist() == [_lib.STATUS_TIMEOUT] * 2


def test_factorization_is_correct_after_a_forced_timeout():
    a, b = _system(700)
    lu = solver.LU().factor(torch.from_numpy(a.T.copy()).cuda(), 'fp64')
    b_cm = torch.from_numpy(b.T.copy()).cuda()
    st = lu.solve_(b_cm, 2, 1e-12, 'rel_fro', True)
    assert st['info'] == 0 and st['converged']
    x = b_cm.cpu().numpy().T
    assert np.linalg.norm(b - a @ x) / np.linalg.norm(b) <= 1e-12
