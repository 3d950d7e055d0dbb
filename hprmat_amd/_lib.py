"""ctypes binding of libhprmat_hip.so (the C ABI declared in include/hprmat_hip.h).

The shared library is built in-tree by ``python -m hprmat_amd.build``.  There
is no fallback: if the library is missing or fails to load, every entry point
raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libhprmat_hip.so')

# Methods and stopping rules (include/hprmat_hip.h).
PLAIN_FP64 = 0
MIXED_FP32 = 1
MIXED_FP32_3M = 2
SPLIT_BF16 = 3
SPLIT_BF16X6 = 4
SPLIT_F16 = 5
EQUILIBRATE = 0x100
CRIT_ABS_MAX = 0
CRIT_REL_FRO = 1
PLAN_TRAIL_TILE = 1      # hprmat_hip_lu_set_plan keys
PLAN_LOOKAHEAD = 2
PLAN_GMRES_CYCLE = 3
PLAN_SOLVE_LANES = 4
PLAN_PIVOT = 5
FGMRES = 0x10            # flag OR-ed into the criterion: flexible GMRES refinement
STATUS_TIMEOUT = -0x40000000   # batch status: a panel leaf's bounded wait gave up

REFERENCE_SYMBOLS = (
    'gpu_solver_init_', 'gpu_solver_finalize_', 'gpu_host_unregister_',
    'gpu_solve_mixed_', 'gpu_solve_hybrid_', 'gpu_solve_tf32_', 'gpu_solve_auto_',
    'gpu_solve_double_', 'gpu_is_available_', 'gpu_get_device_count_', 'gpu_get_info_',
    'gpu_multi_init_', 'gpu_multi_finalize_', 'gpu_solve_multi_', 'gpu_solve_multi_mixed_',
)
NATIVE_SYMBOLS = (
    'hprmat_hip_lu_create', 'hprmat_hip_lu_destroy', 'hprmat_hip_lu_set_plan',
    'hprmat_hip_lu_factor',
    'hprmat_hip_lu_info', 'hprmat_hip_lu_pivots', 'hprmat_hip_lu_solve', 'hprmat_hip_solve',
    'hprmat_hip_assemble', 'hprmat_hip_r_to_s', 'hprmat_hip_r_to_s_work_bytes',
    'hprmat_hip_r_to_s_ws', 'hprmat_hip_adjoint_contract',
    'hprmat_hip_zgemm', 'hprmat_hip_prof_enable', 'hprmat_hip_prof_read', 'hprmat_hip_version',
    'hprmat_hip_batch_work_bytes', 'hprmat_hip_batch_solve', 'hprmat_hip_assemble_batch',
    'hprmat_hip_r_to_s_batch', 'hprmat_hip_propagate_work_bytes', 'hprmat_hip_propagate',
    'hprmat_hip_mesh', 'hprmat_hip_add_nonlocal', 'hprmat_hip_lu_set_operator',
    'hprmat_hip_batch_solve_mesh',
)
PHYSICS_SYMBOLS = (
    'hprmat_init', 'hprmat_set_solver', 'hprmat_get_solver', 'hprmat_solve',
    'hprmat_solve_linear', 'hprmat_version', 'hprmat_solver_info', 'hprmat_solve_full',
    'hprmat_wavefunction', 'hprmat_gpu_finalize',
)


class MeshParts(ctypes.Structure):
    """hprmat_hip_mesh_parts: the Lagrange-mesh parts of a batch of matrices
    of one partial wave (device pointers)."""
    _fields_ = [('nch', ctypes.c_int), ('nr', ctypes.c_int), ('tc', ctypes.c_void_p),
                ('blo0', ctypes.c_void_p), ('r', ctypes.c_void_p), ('rmax0', ctypes.c_double),
                ('lval', ctypes.c_void_p), ('qk', ctypes.c_void_p), ('xc1', ctypes.c_void_p),
                ('stride_q', ctypes.c_int64), ('cpot', ctypes.c_void_p)]


class Stats(ctypes.Structure):
    _fields_ = [('info', ctypes.c_int), ('sweeps', ctypes.c_int), ('checks', ctypes.c_int),
                ('converged', ctypes.c_int), ('res_max', ctypes.c_double),
                ('res_rel', ctypes.c_double), ('b_fro', ctypes.c_double),
                ('hist', ctypes.c_double * 32), ('diverged', ctypes.c_int),
                ('pad_', ctypes.c_int)]

    def as_dict(self):
        return dict(info=self.info, sweeps=self.sweeps, checks=self.checks,
                    converged=bool(self.converged), res_max=self.res_max,
                    res_rel=self.res_rel, b_fro=self.b_fro, diverged=bool(self.diverged),
                    hist=[self.hist[i] for i in range(min(self.checks, 32))])


_lib = None


def load():
    """Load the library once; raises RuntimeError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError('libhprmat_hip.so is not built (run: python -m hprmat_amd.build)')
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    L = ctypes.c_int64
    D = ctypes.c_double
    IP = ctypes.POINTER(ctypes.c_int)
    DP = ctypes.POINTER(ctypes.c_double)
    for name in ('gpu_solve_mixed_', 'gpu_solve_hybrid_', 'gpu_solve_tf32_', 'gpu_solve_auto_',
                 'gpu_solve_multi_mixed_'):
        f = getattr(lib, name)
        f.argtypes = [P, P, IP, IP, IP, DP, IP]
        f.restype = I
    for name in ('gpu_solve_double_', 'gpu_solve_multi_'):
        f = getattr(lib, name)
        f.argtypes = [P, P, IP, IP, IP]
        f.restype = I
    lib.gpu_solver_init_.argtypes = [IP]
    lib.gpu_multi_init_.argtypes = [IP]
    for name in ('gpu_solver_finalize_', 'gpu_host_unregister_', 'gpu_is_available_',
                 'gpu_get_device_count_', 'gpu_multi_finalize_'):
        getattr(lib, name).argtypes = []
        getattr(lib, name).restype = I
    lib.gpu_get_info_.argtypes = [ctypes.c_char_p, IP]
    lib.gpu_get_info_.restype = None
    lib.hprmat_hip_lu_create.argtypes = [ctypes.POINTER(P)]
    lib.hprmat_hip_lu_destroy.argtypes = [P]
    lib.hprmat_hip_lu_set_plan.argtypes = [P, I, I]
    lib.hprmat_hip_lu_factor.argtypes = [P, I, P, L, I, P]
    lib.hprmat_hip_lu_info.argtypes = [P, IP]
    lib.hprmat_hip_lu_pivots.argtypes = [P, IP]
    DPtr = ctypes.c_void_p
    lib.hprmat_hip_assemble.argtypes = [I, I, DPtr, DPtr, DPtr, D, DPtr, DPtr, DPtr, DPtr, DPtr, L, P]
    lib.hprmat_hip_r_to_s.argtypes = [I, I, DPtr, L, DPtr, D, D, DPtr, I, DPtr, DPtr, DPtr, DPtr, DPtr, P]
    lib.hprmat_hip_adjoint_contract.argtypes = [I, I, I, DPtr, L, DPtr, L, DPtr, P]
    lib.hprmat_hip_r_to_s_work_bytes.argtypes = [I]
    lib.hprmat_hip_r_to_s_work_bytes.restype = ctypes.c_size_t
    lib.hprmat_hip_r_to_s_ws.argtypes = [I, I, DPtr, L, DPtr, D, D, DPtr, I, DPtr, DPtr, DPtr, DPtr,
                                         DPtr, DPtr, ctypes.c_size_t, P]
    lib.hprmat_hip_lu_solve.argtypes = [P, P, L, I, I, D, I, I, P, ctypes.POINTER(Stats)]
    lib.hprmat_hip_solve.argtypes = [I, P, L, P, L, I, I, I, D, I, I, P, ctypes.POINTER(Stats)]
    lib.hprmat_hip_zgemm.argtypes = [I, I, I, I, DPtr, L, I, DPtr, L, DPtr, L, I, I, P]
    lib.hprmat_hip_batch_work_bytes.argtypes = [I, I, I, I]
    lib.hprmat_hip_batch_work_bytes.restype = ctypes.c_size_t
    lib.hprmat_hip_batch_solve.argtypes = [I, I, I, I, DPtr, L, L, DPtr, L, L, I, D, I, DPtr,
                                           DPtr, DPtr, ctypes.c_size_t, P]
    lib.hprmat_hip_batch_solve_mesh.argtypes = [I, I, I, I, DPtr, L, L, DPtr, L, L, I, D, I,
                                                DPtr, DPtr, DPtr, ctypes.c_size_t,
                                                ctypes.POINTER(MeshParts), P]
    lib.hprmat_hip_assemble_batch.argtypes = [I, I, I, DPtr, DPtr, DPtr, D, DPtr, DPtr, DPtr,
                                              DPtr, DPtr, L, L, P]
    lib.hprmat_hip_r_to_s_batch.argtypes = [I, I, I, DPtr, L, L, DPtr, D, D, DPtr, DPtr, DPtr,
                                            DPtr, DPtr, DPtr, DPtr, P]
    lib.hprmat_hip_propagate_work_bytes.argtypes = [I, I, I]
    lib.hprmat_hip_propagate_work_bytes.restype = ctypes.c_size_t
    lib.hprmat_hip_propagate.argtypes = [I, I, I, DPtr, L, L, DPtr, DPtr, D, D, DPtr, I, DPtr,
                                         DPtr, DPtr, DPtr, DPtr, I, DPtr, DPtr, I, DPtr, DPtr,
                                         DPtr, L, L, DPtr, DPtr, ctypes.c_size_t, P]
    lib.hprmat_hip_mesh.argtypes = [I, I, D, P, P, P, P, P, P]
    lib.hprmat_hip_lu_set_operator.argtypes = [P, I, I, DPtr, DPtr, DPtr, D, DPtr, DPtr, DPtr,
                                               DPtr]
    lib.hprmat_hip_add_nonlocal.argtypes = [I, I, DPtr, D, DPtr, DPtr, L, P]
    lib.hprmat_hip_version.restype = ctypes.c_char_p
    lib.hprmat_hip_prof_enable.argtypes = [I]
    lib.hprmat_hip_prof_read.argtypes = [DP, DP, IP]
    lib.hprmat_init.argtypes = [I, I, D, P]
    lib.hprmat_init.restype = None
    lib.hprmat_set_solver.argtypes = [I]
    lib.hprmat_set_solver.restype = None
    lib.hprmat_get_solver.argtypes = []
    lib.hprmat_get_solver.restype = I
    lib.hprmat_solve.argtypes = [I, P, P, P, D, I, I, P, I, P, IP, I]
    lib.hprmat_solve.restype = None
    lib.hprmat_solve_linear.argtypes = [P, P, I, I, D, P, I]
    lib.hprmat_solve_full.argtypes = [I, P, P, P, D, I, I, P, I, P, P, I, I, I, P, P, I, IP, I]
    lib.hprmat_solve_full.restype = None
    lib.hprmat_wavefunction.argtypes = [I, P, P, P, D, I, I, P, I, P, I, I, P, I, I, I, D, P]
    lib.hprmat_wavefunction.restype = None
    lib.hprmat_solve_linear.restype = None
    lib.hprmat_version.argtypes = [ctypes.c_char_p, I]
    lib.hprmat_version.restype = None
    lib.hprmat_solver_info.argtypes = [I, ctypes.c_char_p, I]
    lib.hprmat_solver_info.restype = None
    lib.hprmat_gpu_finalize.argtypes = []
    lib.hprmat_gpu_finalize.restype = None
    for name in REFERENCE_SYMBOLS + NATIVE_SYMBOLS + PHYSICS_SYMBOLS:
        if getattr(lib, name, None) is None:
            raise RuntimeError('libhprmat_hip.so lacks symbol ' + name)
    _lib = lib
    return lib


def check(rc, what):
    if rc != 0:
        raise RuntimeError('%s failed with status %d' % (what, rc))
