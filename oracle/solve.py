"""Dense complex solves of the reference's CPU and GPU solver types (oracle).

* ``zgesv``           - solver type 1, LAPACK ZGESV (src/rmatrix_hp.F90:441-477,
                        src/rmat_solvers.F90:98-148).  SciPy's LAPACK
                        (OpenBLAS) is the textbook routine the reference links.
* ``mixed_cpu``       - solver type 2: CGETRF/CGETRS + FP64 refinement, stop
                        when ||r||_F/||b||_F < 1e-10, at most 2 sweeps
                        (src/rmat_solvers.F90:157-296).
* ``mixed_gpu_semantics`` - the device path of gpu_solve_mixed_: FP32 LU,
                        FP32 solve, then up to max_refine FP64 sweeps that stop
                        when the IZAMAX element of the residual (largest
                        |Re|+|Im|) has modulus < tol (absolute),
                        src/cusolver_interface.cu:321-443.
* ``getrf_nopiv_blocked`` / ``getf2`` - unblocked GEPP restatement with the
                        IZAMAX tie rule (first index of max |Re|+|Im|) used
                        to pin pivot sequences on tiny inputs.

Pinning: LAPACK itself (brute-force small systems, manufactured solutions
X = 1 as examples/Ex0/benchmark_mgpu.f90:71-81, and exact-condition
Householder matrices as examples/Ex0/benchmark_mgpu_illcond.f90:41-78).
"""
import numpy as np
from scipy.linalg import lapack


def zgesv(a, b):
    """X = A^-1 B by LAPACK ZGESV on copies; returns (X, info)."""
    a = np.array(a, dtype=np.complex128, order='F', copy=True)
    bb = np.array(b, dtype=np.complex128, order='F', copy=True)
    squeeze = bb.ndim == 1
    if squeeze:
        bb = bb[:, None]
    _, _, x, info = lapack.zgesv(a, bb, overwrite_a=True, overwrite_b=True)
    return (x[:, 0] if squeeze else x), int(info)


def zgetrf(a):
    """LAPACK ZGETRF on a copy: (LU, ipiv 0-based, info)."""
    lu, piv, info = lapack.zgetrf(np.array(a, dtype=np.complex128, order='F', copy=True))
    return lu, piv, int(info)


def getf2(a):
    """Unblocked right-looking GEPP, pivot = first index of max |Re|+|Im|.

    Restates LAPACK zgetf2 (the algorithm cuSOLVER/LAPACK getrf reduce to)
    for tiny matrices; returns (LU, ipiv 0-based, info 1-based or 0).
    """
    a = np.array(a, dtype=np.complex128, copy=True)
    m, n = a.shape
    ipiv = np.zeros(min(m, n), dtype=np.int64)
    info = 0
    for j in range(min(m, n)):
        best, p = -1.0, j
        for i in range(j, m):
            v = abs(a[i, j].real) + abs(a[i, j].imag)
            if v > best:
                best, p = v, i
        ipiv[j] = p
        if a[p, j] != 0:
            if p != j:
                a[[j, p], :] = a[[p, j], :]
            for i in range(j + 1, m):
                a[i, j] = a[i, j] / a[j, j]
        elif info == 0:
            info = j + 1
        for i in range(j + 1, m):
            for k in range(j + 1, n):
                a[i, k] -= a[i, j] * a[j, k]
    return a, ipiv, info


def rel_residual(a, x, b):
    """||B - A X||_F / ||B||_F, the relative rule of src/rmat_solvers.F90:253-262."""
    r = b - a @ x
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(r) / (nb if nb != 0 else 1.0))


def backward_error(a, x, b):
    """Normwise backward error ||r||_inf / (||A||_inf ||X||_inf + ||B||_inf)."""
    r = b - a @ x
    na = np.abs(a).sum(axis=1).max()
    nx = np.abs(x).sum(axis=1).max() if x.ndim == 2 else np.abs(x).max()
    nbb = np.abs(b).sum(axis=1).max() if b.ndim == 2 else np.abs(b).max()
    return float(np.abs(r).sum(axis=1).max() / (na * nx + nbb))


def izamax_modulus(r):
    """Modulus of the element with the largest |Re|+|Im| (first index wins)."""
    flat = np.asarray(r).reshape(-1, order='F')
    idx = int(np.argmax(np.abs(flat.real) + np.abs(flat.imag)))
    return float(abs(flat[idx]))


def mixed_cpu(a, b, max_iter=2, tol=1.0e-10):
    """Solver type 2 semantics; returns (X, sweeps_used, rel_res_history)."""
    a32 = np.array(a, dtype=np.complex64, order='F')
    lu, piv, info = lapack.cgetrf(a32)
    if info != 0:
        x, _ = zgesv(a, b)
        return x, 0, []
    bnorm = np.linalg.norm(b)
    bnorm = bnorm if bnorm != 0 else 1.0
    x32, info = lapack.cgetrs(lu, piv, np.array(b, dtype=np.complex64, order='F'))
    x = x32.astype(np.complex128)
    hist = []
    used = 0
    for _ in range(max_iter):
        r = b - a @ x
        rel = np.linalg.norm(r) / bnorm
        hist.append(float(rel))
        if rel < tol:
            break
        d32, info = lapack.cgetrs(lu, piv, r.astype(np.complex64))
        x = x + d32.astype(np.complex128)
        used += 1
    return x, used, hist


def mixed_gpu_semantics(a, b, max_refine, tol):
    """gpu_solve_mixed_ refinement rule; returns (X, corrections_applied)."""
    a32 = np.array(a, dtype=np.complex64, order='F')
    lu, piv, info = lapack.cgetrf(a32)
    if info != 0:
        raise RuntimeError('CGETRF failed with info = %d' % info)
    x32, _ = lapack.cgetrs(lu, piv, np.array(b, dtype=np.complex64, order='F'))
    x = x32.astype(np.complex128)
    used = 0
    for _ in range(max_refine):
        r = b - a @ x
        if izamax_modulus(r) < tol:
            break
        d32, _ = lapack.cgetrs(lu, piv, r.astype(np.complex64))
        x = x + d32.astype(np.complex128)
        used += 1
    return x, used


def fp64_refined(a, b, max_refine=5, rel_tol=1.0e-12):
    """FP64 LU + FP64 refinement until ||r||_F/||b||_F <= rel_tol."""
    lu, piv, info = lapack.zgetrf(np.array(a, dtype=np.complex128, order='F'))
    x, _ = lapack.zgetrs(lu, piv, np.array(b, dtype=np.complex128, order='F'))
    bnorm = np.linalg.norm(b) or 1.0
    used = 0
    for _ in range(max_refine):
        r = b - a @ x
        if np.linalg.norm(r) / bnorm <= rel_tol:
            break
        d, _ = lapack.zgetrs(lu, piv, r)
        x = x + d
        used += 1
    return x, used
