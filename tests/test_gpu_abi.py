"""The reference-compatible gpu_* C ABI (host pointers, Fortran conventions),
called through ctypes exactly as src/gpu_solver_interface.F90 binds it."""
import ctypes

import numpy as np
import pytest

from hprmat_amd import _lib
from oracle import rmatrix as ormat, solve as osolve, synth as osynth

pytestmark = pytest.mark.gpu

I = ctypes.c_int
D = ctypes.c_double


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lagrange_case(nch=4, nr=60, seed=0):
    lval, qk, cpot, _ = osynth.system(nch, nr, 15.0, 3.0, 2.0, seed)
    out = ormat.rmatrix(lval, qk, 15.0, nr, cpot)
    return np.asfortranarray(out['C']), np.asfortranarray(out['B']), out['X']


def call_refined(name, a, b, max_refine, tol):
    lib = _lib.load()
    bb = np.asfortranarray(b.copy())
    info = I(-99)
    rc = getattr(lib, name)(ptr(a), ptr(bb), ctypes.byref(I(a.shape[0])), ctypes.byref(I(b.shape[1])),
                            ctypes.byref(I(max_refine)), ctypes.byref(D(tol)), ctypes.byref(info))
    return rc, info.value, bb


def test_device_queries():
    lib = _lib.load()
    assert lib.gpu_is_available_() == 1
    assert lib.gpu_get_device_count_() >= 1
    buf = ctypes.create_string_buffer(128)
    lib.gpu_get_info_(buf, ctypes.byref(I(128)))
    assert b'GB' in buf.value
    assert lib.gpu_solver_init_(ctypes.byref(I(0))) == 0
    assert lib.gpu_solver_init_(ctypes.byref(I(0))) == 0      # idempotent
    ng = I(0)
    assert lib.gpu_multi_init_(ctypes.byref(ng)) == 0 and ng.value == 1
    assert lib.gpu_host_unregister_() == 0


def test_gpu_solve_double_matches_zgesv():
    a, b, x = lagrange_case()
    lib = _lib.load()
    bb = np.asfortranarray(b.copy())
    info = I(-99)
    rc = lib.gpu_solve_double_(ptr(a), ptr(bb), ctypes.byref(I(a.shape[0])), ctypes.byref(I(b.shape[1])),
                               ctypes.byref(info))
    assert rc == 0 and info.value == 0
    assert np.abs(bb - x).max() <= 1e-9 * np.abs(x).max()
    rc = lib.gpu_solve_multi_(ptr(a), ptr(bb := np.asfortranarray(b.copy())), ctypes.byref(I(a.shape[0])),
                              ctypes.byref(I(b.shape[1])), ctypes.byref(info))
    assert rc == 0 and np.abs(bb - x).max() <= 1e-9 * np.abs(x).max()


@pytest.mark.parametrize('name', ['gpu_solve_mixed_', 'gpu_solve_hybrid_', 'gpu_solve_auto_',
                                  'gpu_solve_tf32_', 'gpu_solve_multi_mixed_'])
def test_refined_entry_points_reach_tolerance(name):
    a, b, x = lagrange_case()
    rc, info, xb = call_refined(name, a, b, 8, 1e-13)
    assert rc == 0 and info == 0
    # the reference's stopping rule: modulus of the largest-|Re|+|Im| residual entry
    assert osolve.izamax_modulus(b - a @ xb) < 1e-10
    assert np.abs(xb - x).max() <= 1e-8 * np.abs(x).max()


def test_mixed_without_refinement_is_fp32_accurate():
    a, b, x = lagrange_case()
    rc, info, xb = call_refined('gpu_solve_mixed_', a, b, 0, 1e-10)
    assert rc == 0 and info == 0
    err = np.abs(xb - x).max() / np.abs(x).max()
    assert 1e-12 < err < 1e-2        # an FP32 solve: not FP64-exact, not garbage


def test_singular_matrix_reports_pivot_and_leaves_b():
    rng = np.random.default_rng(3)
    n = 64
    a = np.asfortranarray(rng.random((n, n)) + 1j * rng.random((n, n)))
    a[:, 10] = 0
    b = np.asfortranarray(rng.random((n, 2)) + 0j)
    lib = _lib.load()
    bb = b.copy(order='F')
    info = I(-99)
    rc = lib.gpu_solve_double_(ptr(a), ptr(bb), ctypes.byref(I(n)), ctypes.byref(I(2)), ctypes.byref(info))
    assert rc == -1 and info.value == 11
    assert np.array_equal(bb, b)


def wilkinson(n=100):
    """Wilkinson's growth matrix (1 on the diagonal, -1 below, 1 in the last
    column): partial pivoting takes no interchange and the last column grows
    like 2^(n-1), so an FP32 LU is no usable factor (LAPACK cgetrf of the
    equilibrated n = 100 matrix + FP64 refinement on the CPU: ||r||_max /
    (||b||_max + 1) = 2e20, 1.3e13 after one sweep, 2e7 after two)."""
    a = np.eye(n) - np.tril(np.ones((n, n)), -1)
    a[:, -1] = 1.0
    return np.asfortranarray(a.astype(complex))


def test_multi_mixed_guard_on_hopeless_refinement():
    # one sweep on a factor without accuracy -> info 99 and B untouched.
    # (Backward-stable LU leaves a small residual however large kappa is, so
    # the guard needs pivot growth, not just ill-conditioning, to fire.)
    a = wilkinson()
    b = np.asfortranarray(a @ np.ones((a.shape[0], 1)))
    rc, info, xb = call_refined('gpu_solve_multi_mixed_', a, b, 1, 1e-12)
    assert rc == -1 and info == 99
    assert np.array_equal(xb, b)
