"""NumPy-facing mirror of the reference's C physics binding (bindings/c/
hprmat.h:30-166; src/rmatrix_hp.F90:175-368 behind it), over libhprmat_hip's
host-pointer entry points of the same names: the system is assembled,
factored, solved and reduced to R / S on the GPU.

``init(nr, ns, rmax)`` returns the mesh abscissas, ``solve(...)`` the
collision matrix (nch x nch, S in the leading nopen x nopen block) and nopen,
``solve_linear(cmat, b_vector, nch, nlag, normfac, solver)`` the R-matrix.
One interval and eta = 0 only (``solve`` returns nopen = -1 otherwise).
"""
import ctypes

import numpy as np

from . import _lib

SOLVER_DENSE, SOLVER_MIXED, SOLVER_WOODBURY, SOLVER_GPU = 1, 2, 3, 4


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def init(nr, ns, rmax):
    """Mesh abscissas (nr*ns,) of hprmat_init."""
    z = np.zeros(nr * ns)
    _lib.load().hprmat_init(int(nr), int(ns), float(rmax), _ptr(z))
    return z


def set_solver(solver):
    _lib.load().hprmat_set_solver(int(solver))


def get_solver():
    return int(_lib.load().hprmat_get_solver())


def solve(lval, qk, eta, rmax, nr, cpot, ns=1, solver=0):
    """(cu, nopen) of hprmat_solve.  cpot: (nr*ns, nch, nch) complex with
    cpot[ir, c, c'] = V_cc'(r_ir) (passed in the reference's column-major
    layout); cu: (nch, nch) complex, S in cu[:nopen, :nopen]."""
    lval = np.ascontiguousarray(lval, dtype=np.int32)
    nch = len(lval)
    qk = np.ascontiguousarray(qk, dtype=np.float64)
    eta = np.ascontiguousarray(np.zeros(nch) if eta is None else eta, dtype=np.float64)
    cp = np.asfortranarray(np.asarray(cpot, dtype=np.complex128))
    cu = np.zeros((nch, nch), dtype=np.complex128, order='F')
    nopen = ctypes.c_int(-1)
    _lib.load().hprmat_solve(nch, _ptr(lval), _ptr(qk), _ptr(eta), float(rmax), int(nr), int(ns),
                             _ptr(cp), cp.shape[0], _ptr(cu), ctypes.byref(nopen), int(solver))
    return cu, nopen.value


def solve_linear(cmat, b_vector, nch, nlag, normfac=1.0, solver=0):
    """R-matrix (nch x nch) of hprmat_solve_linear for the (nch nlag)^2 matrix cmat."""
    c = np.asfortranarray(np.asarray(cmat, dtype=np.complex128))
    b = np.ascontiguousarray(b_vector, dtype=np.complex128)
    r = np.zeros((nch, nch), dtype=np.complex128, order='F')
    _lib.load().hprmat_solve_linear(_ptr(c), _ptr(b), int(nch), int(nlag), float(normfac),
                                    _ptr(r), int(solver))
    return r


def version():
    buf = ctypes.create_string_buffer(64)
    _lib.load().hprmat_version(buf, 64)
    return buf.value.decode()


def solver_info(solver):
    buf = ctypes.create_string_buffer(128)
    _lib.load().hprmat_solver_info(int(solver), buf, 128)
    return buf.value.decode()
