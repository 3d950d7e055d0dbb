"""PyTorch-facing dense complex solver over libhprmat_hip's native C ABI.

Tensors are complex128 on the GPU.  The library works on column-major
(Fortran) storage, like the reference's gpu_* entry points
(src/cusolver_interface.cu); a torch tensor is row-major, so ``colmajor(A)``
returns the contiguous storage of A in column-major order (a tensor shaped
like A^T).  For complex-symmetric matrices (A^T = A, the Lagrange-mesh case)
the row-major storage already is the column-major one and no copy is made.
All work is ordered on torch's current stream.
"""
import ctypes

import torch

from . import _lib

METHODS = {'fp64': _lib.PLAIN_FP64, 'fp32': _lib.MIXED_FP32,
           'fp32_3m': _lib.MIXED_FP32_3M, 'bf16x3': _lib.SPLIT_BF16,
           'bf16x6': _lib.SPLIT_BF16X6, 'f16x3': _lib.SPLIT_F16}
CRITERIA = {'abs_max': _lib.CRIT_ABS_MAX, 'rel_fro': _lib.CRIT_REL_FRO}


def _stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check_tensor(t, name):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.complex128 or not t.is_cuda:
        raise TypeError('%s must be a complex128 CUDA tensor' % name)
    if not t.is_contiguous():
        raise ValueError('%s must be contiguous' % name)


def colmajor(a, symmetric=False):
    """Column-major storage of matrix ``a`` (returned as a contiguous tensor)."""
    return a if symmetric else a.transpose(-2, -1).contiguous()


class LU:
    """LU factors of one n x n complex matrix, resident on the device.

    ``factor`` keeps a reference to the FP64 column-major matrix because the
    refinement residual reads it; ``solve`` reuses the factors (also for the
    adjoint solve of a complex-symmetric matrix, A^T = A).
    """

    def __init__(self):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.hprmat_hip_lu_create(ctypes.byref(h)), 'hprmat_hip_lu_create')
        self._h = h
        self.n = 0
        self.method = None
        self._a_cm = None

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            self._lib.hprmat_hip_lu_destroy(h)
            self._h = None

    def set_plan(self, trail_tile=None, lookahead=None, gmres_cycle=None, solve_lanes=None,
                 pivot=None):
        """Factorization plan: trailing-update GEMM tile (0 64x64x16,
        1 64x128x16, 2 128x64x16 on 16x16x4 MFMA; FP32 factors only, on
        32x32x2 MFMA: 3 64x64x16, 4 128x128x16, 5 128x64x16, 6 64x128x16),
        look-ahead on/off, the FGMRES cycle length (1..24 steps; 0 =
        default 16), the FGMRES lanes (0 automatic, 1, 2) and the pivot rule
        of the FP32 factors ('auto', the default: tournament except for
        fp32_3m; 'tournament'; 'partial': LAPACK's pivots)."""
        if pivot is not None:
            code = {'auto': 0, 'partial': 1, 'tournament': 2}[pivot]
            _lib.check(self._lib.hprmat_hip_lu_set_plan(self._h, _lib.PLAN_PIVOT, code),
                       'set_plan')
        if solve_lanes is not None:
            _lib.check(self._lib.hprmat_hip_lu_set_plan(self._h, _lib.PLAN_SOLVE_LANES,
                                                         int(solve_lanes)), 'set_plan')
        if gmres_cycle is not None:
            _lib.check(self._lib.hprmat_hip_lu_set_plan(self._h, _lib.PLAN_GMRES_CYCLE,
                                                         int(gmres_cycle)), 'set_plan')
        if trail_tile is not None:
            _lib.check(self._lib.hprmat_hip_lu_set_plan(self._h, _lib.PLAN_TRAIL_TILE,
                                                         int(trail_tile)), 'set_plan')
        if lookahead is not None:
            _lib.check(self._lib.hprmat_hip_lu_set_plan(self._h, _lib.PLAN_LOOKAHEAD,
                                                         int(bool(lookahead))), 'set_plan')
        return self

    def set_operator(self, dmesh, lval, qk, xc1, cpot):
        """Form the refinement's FP64 products from the Lagrange-mesh parts of
        the factored matrix (hprmat_hip_lu_set_operator): dmesh a
        rmatrix.DeviceMesh, lval (int32), qk, xc1 (float64) and cpot
        ((nr, nch, nch) complex128) device tensors; None clears.  The tensors
        are kept alive by the handle."""
        if dmesh is None:
            _lib.check(self._lib.hprmat_hip_lu_set_operator(self._h, 0, 0, None, None, None, 1.0,
                                                            None, None, None, None),
                       'hprmat_hip_lu_set_operator')
            self._op = None
            return self
        return self.set_operator_parts(dmesh.tc, dmesh.blo0, dmesh.r, dmesh.mesh.a, lval, qk, xc1,
                                       cpot)

    def set_operator_parts(self, tc, blo0, r, rmax0, lval, qk, xc1, cpot):
        """set_operator from the raw parts of one interval (the arguments of
        hprmat_hip_assemble as float64 / int32 / complex128 CUDA tensors)."""
        nch, nr = int(lval.numel()), int(r.numel())
        for t, dt in ((tc, torch.float64), (blo0, torch.float64), (r, torch.float64),
                      (lval, torch.int32), (qk, torch.float64), (xc1, torch.float64),
                      (cpot, torch.complex128)):
            if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != dt or not t.is_contiguous():
                raise TypeError('operator parts must be contiguous CUDA tensors of the stated types')
        if (tc.numel() != nr * nr or blo0.numel() != nr * nr or cpot.shape != (nr, nch, nch)
                or qk.numel() != nch or xc1.numel() != nch):
            raise ValueError('operator parts do not match (nr, nch)')
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        rc = self._lib.hprmat_hip_lu_set_operator(self._h, nch, nr, p(tc), p(blo0), p(r),
                                                  float(rmax0), p(lval), p(qk), p(xc1), p(cpot))
        _lib.check(rc, 'hprmat_hip_lu_set_operator')
        self._op = (tc, blo0, r, lval, qk, xc1, cpot)
        return self

    def factor(self, a_cm, method='fp64', equilibrate=False):
        """Factor the matrix whose column-major storage is ``a_cm`` (n x n).

        ``equilibrate`` factors diag(s) A diag(s) with s_i = ||row i||_2^-1/2
        (symmetric, so A^T = A is kept), which lowers the effective condition
        number of Lagrange-mesh matrices and so the refinement work of the
        reduced-precision factors (the refinement residual still uses A).
        """
        _check_tensor(a_cm, 'a_cm')
        n = a_cm.shape[0]
        if a_cm.shape != (n, n):
            raise ValueError('matrix must be square')
        code = METHODS[method] | (_lib.EQUILIBRATE if equilibrate else 0)
        rc = self._lib.hprmat_hip_lu_factor(self._h, code, ctypes.c_void_p(a_cm.data_ptr()),
                                            n, n, _stream_ptr())
        _lib.check(rc, 'hprmat_hip_lu_factor')
        self.n, self.method, self._a_cm = n, method, a_cm
        return self

    def info(self):
        v = ctypes.c_int()
        _lib.check(self._lib.hprmat_hip_lu_info(self._h, ctypes.byref(v)), 'hprmat_hip_lu_info')
        return v.value

    def pivots(self):
        """Row interchanges as LAPACK returns them (1-based numpy int32 array)."""
        import numpy as np
        out = np.zeros(self.n, dtype=np.int32)
        rc = self._lib.hprmat_hip_lu_pivots(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        _lib.check(rc, 'hprmat_hip_lu_pivots')
        return out

    def solve_(self, b_cm, max_refine=0, tol=0.0, criterion='rel_fro', final_check=False,
               stats=True, refinement='ir'):
        """Solve in place; ``b_cm`` is the column-major storage of B (nrhs x n tensor).

        ``refinement``: 'ir' (classic iterative refinement, up to max_refine
        corrections) or 'fgmres' (flexible GMRES with the factors as
        preconditioner, up to max_refine preconditioned products).
        """
        _check_tensor(b_cm, 'b_cm')
        if b_cm.dim() != 2 or b_cm.shape[1] != self.n:
            raise ValueError('b_cm must have shape (nrhs, n)')
        if refinement not in ('ir', 'fgmres'):
            raise ValueError('refinement must be ir or fgmres')
        crit = CRITERIA[criterion] | (_lib.FGMRES if refinement == 'fgmres' else 0)
        st = _lib.Stats() if stats else None
        rc = self._lib.hprmat_hip_lu_solve(self._h, ctypes.c_void_p(b_cm.data_ptr()), self.n,
                                           b_cm.shape[0], max_refine, tol, crit,
                                           int(final_check), _stream_ptr(),
                                           ctypes.byref(st) if st is not None else None)
        _lib.check(rc, 'hprmat_hip_lu_solve')
        return st.as_dict() if st is not None else None


def solve(a, b, method='fp64', max_refine=0, tol=1e-12, criterion='rel_fro', final_check=True,
          symmetric=False, equilibrate=False, refinement='ir', pivot=None):
    """X = A^-1 B for complex128 CUDA tensors a (n x n) and b (n x nrhs).

    ``pivot``: pivot rule of FP32 factors ('tournament' or 'partial'; None:
    the library default).  Returns (X, stats) with X an (n x nrhs) tensor.
    """
    _check_tensor(a, 'a')
    a_cm = colmajor(a, symmetric)
    b_cm = b.transpose(0, 1).contiguous()
    lu = LU()
    if pivot is not None:
        lu.set_plan(pivot=pivot)
    lu.factor(a_cm, method, equilibrate)
    st = lu.solve_(b_cm, max_refine, tol, criterion, final_check, refinement=refinement)
    if st['info'] != 0:
        raise RuntimeError('zero pivot at column %d' % st['info'])
    return b_cm.transpose(0, 1), st


BATCH_TRAIL_TILES = (0, 1, 2)   # 64 x 64, 32 x 32, 128 x 64 (HPRMAT_BATCH_TRAIL)
BATCH_SUBST_TILES = (0, 1)      # by width, 32 x 32 (HPRMAT_BATCH_SUBST)


def batch_code(method, equilibrate=True, trail=0, subst=0):
    """Method flags of hprmat_hip_batch_solve ('fp64', 'fp32' or 'fp32_3m': FP32
    factors whose trailing updates use Gauss's three-product complex GEMM),
    with the tile plan: ``trail`` the U12 / trailing-update tiles, ``subst``
    the substitution-update tiles (HPRMAT_BATCH_TRAIL / HPRMAT_BATCH_SUBST)."""
    if method not in ('fp64', 'fp32', 'fp32_3m'):
        raise ValueError('batched method must be fp64, fp32 or fp32_3m')
    if trail not in BATCH_TRAIL_TILES or subst not in BATCH_SUBST_TILES:
        raise ValueError('unknown batched tile plan (%r, %r)' % (trail, subst))
    return (METHODS[method] | (_lib.EQUILIBRATE if equilibrate else 0)
            | (int(trail) << 16) | (int(subst) << 20))


def batch_work_bytes(method, n, nrhs, count, equilibrate=True):
    """Device workspace (bytes) of a uniform batch of ``count`` systems."""
    return int(_lib.load().hprmat_hip_batch_work_bytes(batch_code(method, equilibrate), int(n),
                                                       int(nrhs), int(count)))


def mesh_parts(dmesh, lval, qk, xc1, cpot):
    """hprmat_hip_mesh_parts of a batch of one partial wave: dmesh a
    rmatrix.DeviceMesh, lval (nch,) int32, qk and xc1 (count, nch) float64,
    cpot (nr, nch, nch) complex128, all contiguous device tensors (the
    caller keeps them alive while the batch runs)."""
    nch, nr = int(lval.numel()), int(dmesh.mesh.nr)
    if qk.dim() != 2 or qk.shape[1] != nch or xc1.shape != qk.shape or cpot.shape != (nr, nch, nch):
        raise ValueError('mesh parts do not match (nr, nch)')
    p = lambda t: t.data_ptr()
    return _lib.MeshParts(nch, nr, p(dmesh.tc), p(dmesh.blo0), p(dmesh.r), float(dmesh.mesh.a),
                          p(lval), p(qk), p(xc1), int(qk.stride(0)), p(cpot))


def batch_solve(code, n, nrhs, count, a_ptr, b_ptr, max_refine, tol, criterion, res_ptr,
                status_ptr, work, work_bytes=None, mesh=None):
    """Raw hprmat_hip_batch_solve on the current stream: ``count`` column-major
    n x n matrices at ``a_ptr`` (stride n*n) and n x nrhs right-hand sides at
    ``b_ptr`` (stride n*nrhs, overwritten with the solutions); ``work`` a
    uint8 device tensor of at least the batch's workspace size (shared by
    calls that are ordered on one stream).  With ``mesh`` (mesh_parts) the
    refinement's FP64 products use the Lagrange-mesh parts of the matrices
    (hprmat_hip_batch_solve_mesh).  Allocates nothing, does not
    synchronise."""
    wb = work.numel() if work_bytes is None else int(work_bytes)
    args = (int(code), int(n), int(nrhs), int(count), ctypes.c_void_p(a_ptr), int(n),
            int(n) * int(n), ctypes.c_void_p(b_ptr), int(n), int(n) * int(nrhs), int(max_refine),
            float(tol), CRITERIA[criterion], ctypes.c_void_p(res_ptr),
            ctypes.c_void_p(status_ptr), ctypes.c_void_p(work.data_ptr()), wb)
    lib = _lib.load()
    if mesh is None:
        rc = lib.hprmat_hip_batch_solve(*args, _stream_ptr())
        _lib.check(rc, 'hprmat_hip_batch_solve')
    else:
        rc = lib.hprmat_hip_batch_solve_mesh(*args, ctypes.byref(mesh), _stream_ptr())
        _lib.check(rc, 'hprmat_hip_batch_solve_mesh')


class Batch:
    """Factor + solve + classic FP64 refinement of ``count`` systems of one
    order ``n`` in lockstep (libhprmat_hip's hprmat_hip_batch_solve: one chain
    of launches serves the whole batch).

    The device workspace is allocated once here; ``solve_`` allocates nothing
    and does not synchronise, so it can be captured into a CUDA/HIP graph.
    Methods: 'fp64' (plain) and 'fp32' (FP32 factors, FP64 refinement).
    """

    def __init__(self, n, nrhs, count, method='fp32', equilibrate=True):
        self._lib = _lib.load()
        self.n, self.nrhs, self.count, self.method = int(n), int(nrhs), int(count), method
        self.code = batch_code(method, equilibrate)
        self.work_bytes = int(self._lib.hprmat_hip_batch_work_bytes(self.code, self.n, self.nrhs,
                                                                    self.count))
        if self.work_bytes == 0:
            raise ValueError('invalid batch shape')
        self.work = torch.empty(self.work_bytes, dtype=torch.uint8, device='cuda')
        self.res = torch.zeros(self.count, dtype=torch.float64, device='cuda')
        self.status = torch.zeros(self.count, dtype=torch.int32, device='cuda')

    def solve_(self, a_cm, b_cm, max_refine=10, tol=1e-12, criterion='rel_fro', count=None):
        """Solve in place.  ``a_cm``: (count, n, n) column-major storages of the
        matrices; ``b_cm``: (count, nrhs, n) column-major storages of the
        right-hand sides, overwritten with the solutions.  ``count`` (default:
        the batch size) solves only the first systems.  Per-system relative
        residuals and statuses (1 converged, 0 not, -info singular,
        ``_lib.STATUS_TIMEOUT`` when a panel leaf gave up a wait) are left in
        ``self.res`` / ``self.status`` on the device."""
        _check_tensor(a_cm, 'a_cm')
        _check_tensor(b_cm, 'b_cm')
        c = self.count if count is None else int(count)
        n, r = self.n, self.nrhs
        if not 0 <= c <= self.count:
            raise ValueError('count exceeds the batch size')
        if a_cm.dim() != 3 or a_cm.shape[1:] != (n, n) or a_cm.shape[0] < c:
            raise ValueError('a_cm must have shape (count, n, n)')
        if b_cm.dim() != 3 or b_cm.shape[1:] != (r, n) or b_cm.shape[0] < c:
            raise ValueError('b_cm must have shape (count, nrhs, n)')
        batch_solve(self.code, n, r, c, a_cm.data_ptr(), b_cm.data_ptr(), max_refine, tol,
                    criterion, self.res.data_ptr(), self.status.data_ptr(), self.work,
                    self.work_bytes)
        return self
