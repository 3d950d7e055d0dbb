"""Lagrange-mesh R-matrix front end of the GPU solve path.

Mirrors the reference's physics entry points (one interval unless noted):
* ``rmat_ini``       - src/rmatrix_hp.F90:61-159 (mesh, T+L, Bloch, q2)
* ``channel_data``   - src/rmatrix_hp.F90:224-244 at eta = 0 (closed forms)
* ``assemble``       - src/rmatrix_hp.F90:246-293, on the device
* ``solve_rmatrix``  - src/rmat_solvers.F90:49-90 solver_type dispatch,
                       every type served by libhprmat_hip on the GPU
* ``r_to_s``         - src/rmat_solvers.F90:597-609 + src/rmatrix_hp.F90:336-368
* ``rmatrix``        - the whole ns = 1 pipeline returning (S, nopen)
* ``rmatrix_intervals`` - src/rmatrix_hp.F90:175-431 for ns >= 1 intervals:
                       propagation, S and the wave functions on the device

Host work is O(nr^2 + nr nch^2); the N x N matrix is built, factored and
solved on the device.
"""
import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch
from scipy import special

from . import _lib, solver


@dataclass
class Mesh:
    nr: int
    a: float
    x: np.ndarray      # Gauss-Legendre nodes on (0, 1)
    w: np.ndarray      # weights on (0, 1)
    r: np.ndarray      # radii a x
    tc: np.ndarray     # T + L(0), nr x nr, scaled by 1/a^2
    blo0: np.ndarray   # Bloch block, nr x nr, scaled by 1/a
    q2: np.ndarray     # basis values at r = a, scaled by 1/sqrt(a)


def rmat_ini(nr, a):
    """Regularised Lagrange-Legendre mesh of one interval (0, a)."""
    t, w = np.polynomial.legendre.leggauss(nr)
    x = (1.0 + t) / 2.0
    w = w / 2.0
    x2 = x * (1.0 - x)
    i = np.arange(nr)
    sign = np.where((i[:, None] + i[None, :]) % 2 == 1, -1.0, 1.0)
    xi, xj = x[:, None], x[None, :]
    with np.errstate(divide='ignore', invalid='ignore'):
        off = (nr * (nr + 1) + 1 + (xi + xj - 2 * xi * xj) / (xi - xj) ** 2
               - 1 / (1 - xi) - 1 / (1 - xj)) / np.sqrt(x2[:, None] * x2[None, :])
    tc = sign * off
    np.fill_diagonal(tc, (4 * nr * (nr + 1) + 3 + (1 - 6 * x) / x2) / (3 * x2))
    tc = np.tril(tc) + np.tril(tc, -1).T           # exactly symmetric, lower triangle mirrored
    blo0 = sign / np.sqrt(x2[:, None] * x2[None, :])
    blo0 = np.tril(blo0) + np.tril(blo0, -1).T
    q2 = np.where((nr + i + 1) % 2 == 1, -1.0, 1.0) / np.sqrt(x2)
    return Mesh(nr, a, x, w, x * a, tc / a ** 2, blo0 / a, q2 / math.sqrt(a))


def channel_data(lval, qk, a):
    """Open-channel (co, cop) and closed-channel xc1 at eta = 0.

    co = G + iF, cop = rho (G' + iF') with F = rho j_l, G = -rho y_l, and
    xc1 = rho W'/W with W = (2/pi) rho k_l(rho), rho = |k| a.
    """
    lval = np.asarray(lval)
    qk = np.asarray(qk, dtype=np.float64)
    rho = np.abs(qk) * a
    open_idx = np.nonzero(qk > 0)[0]
    co = np.zeros(len(open_idx), dtype=np.complex128)
    cop = np.zeros(len(open_idx), dtype=np.complex128)
    xc1 = np.zeros(len(qk))
    for m, c in enumerate(open_idx):
        l, p = int(lval[c]), rho[c]
        j, jd = special.spherical_jn(l, p), special.spherical_jn(l, p, derivative=True)
        y, yd = special.spherical_yn(l, p), special.spherical_yn(l, p, derivative=True)
        co[m] = complex(-p * y, p * j)
        cop[m] = complex(-(y + p * yd), j + p * jd) * p
    for c in np.nonzero(qk < 0)[0]:
        l, p = int(lval[c]), rho[c]
        k, kd = special.spherical_kn(l, p), special.spherical_kn(l, p, derivative=True)
        xc1[c] = p * (k + p * kd) / (p * k)
    return open_idx.astype(np.int32), co, cop, xc1


def _dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype).cuda()


class DeviceMesh:
    """Mesh arrays resident on the device (shared by many assemblies)."""

    def __init__(self, mesh):
        self.mesh = mesh
        self.tc = _dev(mesh.tc, torch.float64)
        self.blo0 = _dev(mesh.blo0, torch.float64)
        self.r = _dev(mesh.r, torch.float64)
        self.q2 = _dev(mesh.q2, torch.float64)


def assemble(dmesh, lval, qk, xc1, cpot, out=None):
    """Column-major storage of C (returned as an N x N tensor, see solver.colmajor)."""
    lib = _lib.load()
    nr = dmesh.mesh.nr
    nch = len(lval)
    n = nch * nr
    if out is None:
        out = torch.empty((n, n), dtype=torch.complex128, device='cuda')
    lv = _dev(np.asarray(lval, dtype=np.int32), torch.int32)
    q = _dev(qk, torch.float64)
    xc = _dev(xc1, torch.float64)
    cp = cpot if isinstance(cpot, torch.Tensor) else _dev(cpot, torch.complex128)
    rc = lib.hprmat_hip_assemble(
        nch, nr, ctypes.c_void_p(dmesh.tc.data_ptr()), ctypes.c_void_p(dmesh.blo0.data_ptr()),
        ctypes.c_void_p(dmesh.r.data_ptr()), ctypes.c_double(dmesh.mesh.a),
        ctypes.c_void_p(lv.data_ptr()), ctypes.c_void_p(q.data_ptr()),
        ctypes.c_void_p(xc.data_ptr()), ctypes.c_void_p(cp.data_ptr()),
        ctypes.c_void_p(out.data_ptr()), ctypes.c_int64(n), solver._stream_ptr())
    _lib.check(rc, 'hprmat_hip_assemble')
    return out


def rhs_colmajor(dmesh, nch):
    """Column-major storage of B (nch x N tensor): column c holds q2 in block c."""
    nr = dmesh.mesh.nr
    b = torch.zeros((nch, nch * nr), dtype=torch.complex128, device='cuda')
    for c in range(nch):
        b[c, c * nr:(c + 1) * nr] = dmesh.q2.to(torch.complex128)
    return b


# solver_type -> (method, max_refine, tol, criterion, equilibrate, refinement).
# Types 1/2 keep the semantics of the reference's CPU types (plain FP64; FP32
# + FP64 refinement with ||r||_F/||b||_F < 1e-10, two sweeps) but run on the
# GPU; type 1 adds at most two FP64 sweeps to ||r||_F/||b||_F <= 1e-12 so that
# it matches LAPACK zgesv's accuracy (the substitution multiplies by explicit
# diagonal-block inverses); type 3 (Woodbury, CPU-only) is not provided.  'plain' and 'mixed' are
# this library's refined methods: FP64 factors with classic refinement, or
# FP32 factors with flexible-GMRES refinement, both of the equilibrated
# matrix, until ||r||_F/||b||_F <= 1e-12.
SOLVER_TYPES = {
    1: ('fp64', 2, 1.0e-12, 'rel_fro', False, 'ir'),
    2: ('fp32', 2, 1.0e-10, 'rel_fro', False, 'ir'),
    4: ('fp32', None, 1.0e-10, 'abs_max', False, 'ir'),
    5: ('bf16x3', 5, 1.0e-12, 'abs_max', False, 'ir'),
    6: ('fp32', 2, 1.0e-12, 'abs_max', False, 'ir'),
    'plain': ('fp64', 5, 1.0e-12, 'rel_fro', True, 'ir'),
    'mixed': ('fp32', 24, 1.0e-12, 'rel_fro', True, 'fgmres'),
}


def solve_colmajor(a_cm, b_cm, solver_type=1, max_refine=0, operator=None):
    """Solve C X = B in place (column-major storages) for a solver_type.
    operator: optional Lagrange-mesh parts of C, the arguments of
    solver.LU.set_operator_parts, for the refinement's FP64 products."""
    if solver_type not in SOLVER_TYPES:
        raise NotImplementedError('solver_type %r is not provided by libhprmat_hip' % (solver_type,))
    method, mr, tol, crit, eq, refinement = SOLVER_TYPES[solver_type]
    if mr is None:
        mr = max_refine
    lu = solver.LU()
    if operator is not None:
        lu.set_operator_parts(*operator)
    lu.factor(a_cm, method, eq)
    st = lu.solve_(b_cm, mr, tol, crit, final_check=eq, refinement=refinement)
    if st['info'] != 0:
        raise RuntimeError('zero pivot at column %d' % st['info'])
    return st


def r_to_s(dmesh, x_cm, nch, qk, open_idx, co, cop, normfac=1.0):
    """(R, S) on the device from the column-major solution storage x_cm (nch x N)."""
    lib = _lib.load()
    nopen = len(open_idx)
    r = torch.empty((nch, nch), dtype=torch.complex128, device='cuda')
    s = torch.empty((max(nopen, 1), max(nopen, 1)), dtype=torch.complex128, device='cuda')
    q = _dev(qk, torch.float64)
    o = _dev(np.asarray(open_idx, dtype=np.int32) if nopen else np.zeros(1, np.int32), torch.int32)
    c0 = _dev(co if nopen else np.zeros(1, complex), torch.complex128)
    c1 = _dev(cop if nopen else np.zeros(1, complex), torch.complex128)
    rc = lib.hprmat_hip_r_to_s(nch, dmesh.mesh.nr, ctypes.c_void_p(x_cm.data_ptr()),
                               ctypes.c_int64(x_cm.shape[1]), ctypes.c_void_p(dmesh.q2.data_ptr()),
                               ctypes.c_double(normfac), ctypes.c_double(dmesh.mesh.a),
                               ctypes.c_void_p(q.data_ptr()), nopen, ctypes.c_void_p(o.data_ptr()),
                               ctypes.c_void_p(c0.data_ptr()), ctypes.c_void_p(c1.data_ptr()),
                               ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(s.data_ptr()),
                               solver._stream_ptr())
    _lib.check(rc, 'hprmat_hip_r_to_s')
    # column-major storages -> matrices
    return r.transpose(0, 1), s[:nopen, :nopen].transpose(0, 1)


def assemble_batch(dmesh, lval_dev, qk, xc1, cpot, out, count=None):
    """Column-major storages of C at several energies in one launch: qk and
    xc1 are (count, nch) float64 device tensors (row e for energy e), lval_dev
    the (nch,) int32 device angular momenta, cpot the (nr, nch, nch) device
    potential; out is a (>= count, N, N) complex128 device tensor."""
    lib = _lib.load()
    nr = dmesh.mesh.nr
    nch = qk.shape[1]
    n = nch * nr
    c = qk.shape[0] if count is None else int(count)
    if out.dim() != 3 or out.shape[1:] != (n, n) or out.shape[0] < c or xc1.shape[0] < c:
        raise ValueError('out must be (count, N, N) and xc1 (count, nch)')
    rc = lib.hprmat_hip_assemble_batch(
        nch, nr, c, ctypes.c_void_p(dmesh.tc.data_ptr()), ctypes.c_void_p(dmesh.blo0.data_ptr()),
        ctypes.c_void_p(dmesh.r.data_ptr()), ctypes.c_double(dmesh.mesh.a),
        ctypes.c_void_p(lval_dev.data_ptr()), ctypes.c_void_p(qk.data_ptr()),
        ctypes.c_void_p(xc1.data_ptr()), ctypes.c_void_p(cpot.data_ptr()),
        ctypes.c_void_p(out.data_ptr()), ctypes.c_int64(n), ctypes.c_int64(n * n),
        solver._stream_ptr())
    _lib.check(rc, 'hprmat_hip_assemble_batch')
    return out


def r_to_s_batch(dmesh, x_cm, qk, nopen, open_idx, co, cop, r_out, s_out, count=None,
                 normfac=1.0):
    """R and S of several solutions of one system in one launch (nch <= 64).
    x_cm: (count, nch, N) column-major solution storages; qk (count, nch)
    float64, nopen (count,) int32, open_idx (count, nch) int32, co / cop
    (count, nch) complex128 (first nopen[e] entries used), all on the device.
    r_out / s_out: (count, nch, nch) complex128; block e receives R
    (column-major nch x nch) and S (column-major nopen x nopen at its start)."""
    lib = _lib.load()
    nch = qk.shape[1]
    c = x_cm.shape[0] if count is None else int(count)
    n = nch * dmesh.mesh.nr
    if x_cm.dim() != 3 or x_cm.shape[1:] != (nch, n) or x_cm.shape[0] < c:
        raise ValueError('x_cm must be (count, nch, N)')
    rc = lib.hprmat_hip_r_to_s_batch(
        nch, dmesh.mesh.nr, c, ctypes.c_void_p(x_cm.data_ptr()), ctypes.c_int64(n),
        ctypes.c_int64(n * nch), ctypes.c_void_p(dmesh.q2.data_ptr()), ctypes.c_double(normfac),
        ctypes.c_double(dmesh.mesh.a), ctypes.c_void_p(qk.data_ptr()),
        ctypes.c_void_p(nopen.data_ptr()), ctypes.c_void_p(open_idx.data_ptr()),
        ctypes.c_void_p(co.data_ptr()), ctypes.c_void_p(cop.data_ptr()),
        ctypes.c_void_p(r_out.data_ptr()), ctypes.c_void_p(s_out.data_ptr()),
        solver._stream_ptr())
    _lib.check(rc, 'hprmat_hip_r_to_s_batch')
    return r_out, s_out


def add_nonlocal(dmesh, cpnl, a_cm):
    """C += sqrt(w_i w_j) a K_cc'(r_i, r_j) on the device (src/rmatrix_hp.F90:
    283-289); cpnl: (nr, nr, nch, nch) complex."""
    return _add_nonlocal(dmesh.mesh.nr, dmesh.mesh.w, dmesh.mesh.a, cpnl, a_cm)


def _add_nonlocal(nr, w, a0, cpnl, a_cm):
    k = cpnl if isinstance(cpnl, torch.Tensor) else _dev(np.ascontiguousarray(cpnl), torch.complex128)
    nch = k.shape[2]
    if k.shape != (nr, nr, nch, nch) or not k.is_contiguous() or a_cm.shape != (nch * nr, nch * nr):
        raise ValueError('cpnl must be a contiguous (nr, nr, nch, nch) array for an (nch nr)^2 matrix')
    wd = _dev(np.asarray(w, dtype=np.float64), torch.float64)
    rc = _lib.load().hprmat_hip_add_nonlocal(
        nch, nr, ctypes.c_void_p(wd.data_ptr()), ctypes.c_double(a0),
        ctypes.c_void_p(k.data_ptr()), ctypes.c_void_p(a_cm.data_ptr()),
        ctypes.c_int64(nch * nr), solver._stream_ptr())
    _lib.check(rc, 'hprmat_hip_add_nonlocal')
    return a_cm


def _operator_parts(tc, blo0, r, rmax0, lval, qk, xc1, cpot):
    """Device tensors of the Lagrange-mesh operator (solver.LU.set_operator_parts)."""
    def dev(v, dt, npdt):
        if isinstance(v, torch.Tensor) and v.is_cuda:
            return v.to(dt).contiguous()
        return _dev(np.asarray(v, dtype=npdt), dt)
    return (tc, blo0, r, float(rmax0), dev(lval, torch.int32, np.int32),
            dev(qk, torch.float64, np.float64), dev(xc1, torch.float64, np.float64),
            dev(cpot, torch.complex128, np.complex128))


def rmatrix(lval, qk, a, nr, cpot, solver_type=1, max_refine=0, dmesh=None, cpnl=None,
            structured=True):
    """ns = 1 pipeline on the GPU: returns dict(S, R, nopen, stats).
    cpnl: optional non-local potential (nr, nr, nch, nch).  structured: the
    refinement's FP64 products use the Lagrange-mesh parts of C instead of
    the dense matrix (local potentials only)."""
    if dmesh is None:
        dmesh = DeviceMesh(rmat_ini(nr, a))
    nch = len(lval)
    open_idx, co, cop, xc1 = channel_data(lval, qk, a)
    cp = cpot if isinstance(cpot, torch.Tensor) else _dev(np.ascontiguousarray(cpot), torch.complex128)
    a_cm = assemble(dmesh, lval, qk, xc1, cp)
    if cpnl is not None:
        add_nonlocal(dmesh, cpnl, a_cm)
    op = None
    if structured and cpnl is None:
        op = _operator_parts(dmesh.tc, dmesh.blo0, dmesh.r, dmesh.mesh.a, lval, qk, xc1, cp)
    b_cm = rhs_colmajor(dmesh, nch)
    st = solve_colmajor(a_cm, b_cm, solver_type, max_refine, operator=op)
    r, s = r_to_s(dmesh, b_cm, nch, qk, open_idx, co, cop)
    return dict(S=s, R=r, nopen=len(open_idx), stats=st, open=open_idx)


def closed_w(lval, qk, a):
    """W(|k| a) = (2/pi) rho k_l(rho) per channel (zero for open channels):
    the closed-channel Whittaker function at eta = 0, src/rmatrix_hp.F90:240."""
    w = np.zeros(len(qk))
    for c, (l, k) in enumerate(zip(lval, qk)):
        if k < 0:
            rho = -k * a
            w[c] = (2.0 / math.pi) * rho * special.spherical_kn(int(l), rho)
    return w


# Intervals of at most this order are solved as one uniform batch; larger
# ones one at a time (each then fills the GPU, with look-ahead and FGMRES).
BATCH_INTERVAL_MAX_N = 4096
# solver_type -> (batched method, tolerance, sweep budget) of the batch: the
# single-system plans' precision and gate with classic sweeps
_BATCH_PLANS = {1: ('fp64', 1.0e-12, 2), 2: ('fp32', 1.0e-10, 2), 4: ('fp32_3m', 1.0e-12, 24),
                'plain': ('fp64', 1.0e-12, 5), 'mixed': ('fp32_3m', 1.0e-12, 24)}


def rmatrix_intervals(lval, qk, a, nr, ns, cpot, solver_type=1, nvc=None, max_refine=0,
                      cpnl=None, structured=True):
    """rmatrix (src/rmatrix_hp.F90:175-431) for ns >= 1 intervals on the GPU.

    Every interval's system C_s is assembled on the device and solved for
    its 2 nch boundary columns [q1 | q2] (solve_rmatrix_hp_4,
    src/rmatrix_hp.F90:590-632, with the solver_type's precision): as one
    uniform batch when N <= BATCH_INTERVAL_MAX_N (local potentials), else
    one system at a time; the R blocks,
    the propagation chain, S and, for the entrance channels ``nvc`` (0-based
    among the open channels), the internal wave functions and bound-state
    amplitudes follow in hprmat_hip_propagate.  cpot: (nr ns, nch, nch);
    cpnl: optional non-local potential (nr, nr, nch, nch), ns = 1 only.
    structured: the refinement's FP64 products use the Lagrange-mesh parts
    of each C_s instead of the dense matrix (local potentials only).
    nch <= 64.  Returns dict(S, R, nopen, cf (nr ns, nch, len(nvc)) or None,
    amps (nclo, len(nvc)) or None, X (ns, 2 nch, N) column-major solutions).
    """
    from . import physics
    lib = _lib.load()
    lval = np.asarray(lval)
    qk = np.asarray(qk, dtype=np.float64)
    nch = len(lval)
    n = nch * nr
    h = a / ns
    if cpnl is not None and ns != 1:
        raise ValueError('ns must be 1 for non-local potentials')
    m = physics.mesh(nr, ns, a)
    tc, bl = _dev(m['tc'], torch.float64), _dev(m['bloch'], torch.float64)
    rr, q1, q2 = _dev(m['r'], torch.float64), _dev(m['q1'], torch.float64), _dev(m['q2'], torch.float64)
    wle = _dev(m['wle'], torch.float64)
    open_idx, co, cop, xc1 = channel_data(lval, qk, a)
    closed = [c for c in range(nch) if qk[c] <= 0]
    nopen, nclo = len(open_idx), len(closed)
    lv = _dev(lval.astype(np.int32), torch.int32)
    qd, xd = _dev(qk, torch.float64), _dev(xc1, torch.float64)
    cp = cpot if isinstance(cpot, torch.Tensor) else _dev(np.ascontiguousarray(cpot), torch.complex128)
    if cp.shape != (nr * ns, nch, nch) or not cp.is_contiguous():
        raise ValueError('cpot must be a contiguous (nr*ns, nch, nch) array')
    a_cm = torch.empty((n, n), dtype=torch.complex128, device='cuda')
    x = torch.zeros((ns, 2 * nch, n), dtype=torch.complex128, device='cuda')
    for c in range(nch):
        x[:, c, c * nr:(c + 1) * nr] = q1.to(torch.complex128)
        x[:, nch + c, c * nr:(c + 1) * nr] = q2.to(torch.complex128)
    x0 = x.clone() if ns > 1 else x
    def assemble_interval(s, out):
        rc = lib.hprmat_hip_assemble(
            nch, nr, ctypes.c_void_p(tc[s].data_ptr()), ctypes.c_void_p(bl[s].data_ptr()),
            ctypes.c_void_p(rr[s].data_ptr()), ctypes.c_double(1.0),
            ctypes.c_void_p(lv.data_ptr()), ctypes.c_void_p(qd.data_ptr()),
            ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(cp[s * nr].data_ptr()),
            ctypes.c_void_p(out.data_ptr()), ctypes.c_int64(n), solver._stream_ptr())
        _lib.check(rc, 'hprmat_hip_assemble')

    st = []
    batched = (ns > 1 and n <= BATCH_INTERVAL_MAX_N and cpnl is None
               and solver_type in _BATCH_PLANS)
    if batched:
        # the ns interval systems are independent and of one order: one
        # uniform batch (one chain of launches for all of them)
        method, tol, budget = _BATCH_PLANS[solver_type]
        a_all = torch.empty((ns, n, n), dtype=torch.complex128, device='cuda')
        for s in range(ns):
            assemble_interval(s, a_all[s])
        bt = solver.Batch(n, 2 * nch, ns, method, equilibrate=True)
        bt.solve_(a_all, x, max_refine=budget, tol=tol)
        status = bt.status.cpu().numpy()
        for s in np.nonzero(status != 1)[0]:   # stragglers: the single-system FP64 path
            x[s].copy_(x0[s])
            st.append(solve_colmajor(a_all[s], x[s], 1, 0))
    else:
        for s in range(ns):
            assemble_interval(s, a_cm)
            op = None
            if cpnl is not None:
                _add_nonlocal(nr, m['wle'], h, cpnl, a_cm)
            elif structured:
                op = _operator_parts(tc[s], bl[s], rr[s], 1.0, lv, qd, xd, cp[s * nr:(s + 1) * nr])
            st.append(solve_colmajor(a_cm, x[s], solver_type, max_refine, operator=op))
    nc = 0 if nvc is None else len(nvc)
    if nc and nopen == 0:
        raise ValueError('wave functions need an open entrance channel')
    if nc and any(not 0 <= v < nopen for v in nvc):
        raise ValueError('entrance channels are 0-based indices among the open channels')
    if nc and np.any(qk == 0):
        raise ValueError('a channel at threshold (k = 0) has no bound-state amplitude')
    wb = int(lib.hprmat_hip_propagate_work_bytes(nch, ns, nc))
    if wb == 0:
        raise ValueError('propagation needs 1 <= nch <= 64')
    work = torch.empty(wb, dtype=torch.uint8, device='cuda')
    r = torch.empty((nch, nch), dtype=torch.complex128, device='cuda')
    s_cm = torch.empty((max(nopen, 1), max(nopen, 1)), dtype=torch.complex128, device='cuda')
    ivec = lambda v: _dev(np.asarray(v if len(v) else [0], dtype=np.int32), torch.int32)
    o, cl, nv = ivec(list(open_idx)), ivec(closed), ivec(list(nvc) if nc else [])
    c0 = _dev(co if nopen else np.zeros(1, complex), torch.complex128)
    c1 = _dev(cop if nopen else np.zeros(1, complex), torch.complex128)
    x2 = _dev(closed_w(lval, qk, a), torch.float64)
    cf = torch.zeros((max(nc, 1), nch, nr * ns), dtype=torch.complex128, device='cuda')
    amp = torch.zeros((max(nc, 1), max(nclo, 1)), dtype=torch.complex128, device='cuda')
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    rc = lib.hprmat_hip_propagate(
        nch, nr, ns, p(x), n, 2 * nch * n, p(q1), p(q2), 1.0, h, p(qd), nopen, p(o), p(c0), p(c1),
        p(r), p(s_cm), nc, p(nv), p(wle), nclo if nc else 0, p(cl), p(x2), p(cf), nr * ns, nch,
        p(amp), p(work), wb, solver._stream_ptr())
    _lib.check(rc, 'hprmat_hip_propagate')
    out = dict(S=s_cm[:nopen, :nopen].transpose(0, 1), R=r.transpose(0, 1), nopen=nopen, cf=None,
               amps=None, X=x, stats=st, open=open_idx)
    if nc:
        out['cf'] = cf.permute(2, 1, 0)
        out['amps'] = amp[:, :nclo].transpose(0, 1)
    return out
