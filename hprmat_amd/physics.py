"""NumPy-facing mirror of the reference's C physics binding (bindings/c/
hprmat.h:30-166; src/rmatrix_hp.F90:175-368 behind it), over libhprmat_hip's
host-pointer entry points of the same names: the system is assembled,
factored, solved and reduced to R / S on the GPU.

``init(nr, ns, rmax)`` returns the mesh abscissas, ``solve(...)`` the
collision matrix (nch x nch, S in the leading nopen x nopen block) and nopen,
``solve_full(...)`` also the internal wave functions (src/rmatrix_hp.F90:
370-431), ``wavefunction(...)`` one of them on a uniform mesh (wf_print,
:678-734), ``solve_linear(cmat, b_vector, nch, nlag, normfac, solver)`` the
R-matrix of a given system, ``mesh(nr, ns, rmax)`` the interval matrices.
eta = 0 and local potentials only; several intervals or wave functions need
nch <= 64 (``solve`` / ``solve_full`` return nopen = -1 otherwise).
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


def solve_full(lval, qk, eta, rmax, nr, cpot, nvc, ns=1, solver=0, cpnl=None):
    """(cu, cf, nopen) of hprmat_solve_full.  nvc: entrance channels,
    1-based among the open channels; cf: (nr*ns, nch, len(nvc)) complex
    wave functions at the mesh points; cu: S in cu[:nopen, :nopen], the
    bound-state amplitudes in cu[nopen:, :len(nvc)].  cpnl: optional
    non-local potential (nr, nr, nch, nch), K_cc'(r_i, r_j) (ns = 1)."""
    lval = np.ascontiguousarray(lval, dtype=np.int32)
    nch = len(lval)
    qk = np.ascontiguousarray(qk, dtype=np.float64)
    eta = np.ascontiguousarray(np.zeros(nch) if eta is None else eta, dtype=np.float64)
    cp = np.asfortranarray(np.asarray(cpot, dtype=np.complex128))
    nv = np.ascontiguousarray(nvc, dtype=np.int32)
    cu = np.zeros((nch, nch), dtype=np.complex128, order='F')
    cf = np.zeros((nr * ns, nch, len(nv)), dtype=np.complex128, order='F')
    nopen = ctypes.c_int(-1)
    kp, kdim = None, 0
    if cpnl is not None:
        k = np.asarray(cpnl, dtype=np.complex128)
        kf = np.asfortranarray(k.reshape(nr * nr, nch, nch))   # row ir*nr + ir'
        kp, kdim = _ptr(kf), nr * nr
    _lib.load().hprmat_solve_full(nch, _ptr(lval), _ptr(qk), _ptr(eta), float(rmax), int(nr),
                                  int(ns), _ptr(cp), cp.shape[0], _ptr(cu), _ptr(cf), nr * ns, nch,
                                  len(nv), _ptr(nv), kp, kdim, ctypes.byref(nopen), int(solver))
    return cu, cf, nopen.value


def wavefunction(lval, qk, eta, rmax, nr, ns, cu, nopen, cf, zrma, iv, nom, npoin, h):
    """hprmat_wavefunction: channel iv (1-based) of entrance column nom
    (1-based) at r = h .. npoin h."""
    lval = np.ascontiguousarray(lval, dtype=np.int32)
    nch = len(lval)
    qk = np.ascontiguousarray(qk, dtype=np.float64)
    eta = np.ascontiguousarray(np.zeros(nch) if eta is None else eta, dtype=np.float64)
    cu = np.asfortranarray(cu, dtype=np.complex128)
    cf = np.asfortranarray(cf, dtype=np.complex128)
    z = np.ascontiguousarray(zrma, dtype=np.float64)
    out = np.zeros(int(npoin), dtype=np.complex128)
    _lib.load().hprmat_wavefunction(nch, _ptr(lval), _ptr(qk), _ptr(eta), float(rmax), int(nr),
                                    int(ns), _ptr(cu), int(nopen), _ptr(cf), cf.shape[0],
                                    cf.shape[1], _ptr(z), int(iv), int(nom), int(npoin), float(h),
                                    _ptr(out))
    return out


def mesh(nr, ns, rmax):
    """Interval matrices of hprmat_hip_mesh: dict of tc, bloch (ns, nr, nr),
    r, q1, q2 (ns, nr), wle (nr,)."""
    tc = np.zeros((ns, nr, nr))
    bloch = np.zeros((ns, nr, nr))
    r, q1, q2 = np.zeros((ns, nr)), np.zeros((ns, nr)), np.zeros((ns, nr))
    wle = np.zeros(nr)
    rc = _lib.load().hprmat_hip_mesh(int(nr), int(ns), float(rmax), _ptr(tc), _ptr(bloch), _ptr(r),
                                     _ptr(q1), _ptr(q2), _ptr(wle))
    _lib.check(rc, 'hprmat_hip_mesh')
    return dict(tc=tc, bloch=bloch, r=r, q1=q1, q2=q2, wle=wle)


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
