"""hprmat_hip_batch_solve_noalloc (the allocation-free batched solve, its
workspace in the library's per-GPU scratch array) against
hprmat_hip_batch_solve with a caller workspace: the same kernels on the same
inputs, so solutions, residuals and statuses agree bit for bit.  Two calls in
a row, the second smaller (the scratch array keeps the first call's
contents), then a request larger than the array, which is refused before any
device work.  Called through the C ABI with device pointers."""
import ctypes

import numpy as np
import pytest
import torch

from hprmat_amd import _lib

pytestmark = pytest.mark.gpu

I, L, D, P = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
SCRATCH_BYTES = 64 << 20   # HPRMAT_NOALLOC_SCRATCH_BYTES


def _lib_noalloc():
    lib = _lib.load()
    f = lib.hprmat_hip_batch_solve_noalloc
    f.argtypes = [I, I, I, I, P, L, L, P, L, L, I, D, I, P, P, P]
    f.restype = I
    return lib


def _systems(n, nrhs, count, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((count, n, n)) + 1j * rng.standard_normal((count, n, n))
    a += 4 * np.sqrt(n) * np.eye(n)[None]          # well conditioned
    b = rng.standard_normal((count, n, nrhs)) + 1j * rng.standard_normal((count, n, nrhs))
    # column-major storages: (count, n, n) tensors holding A^T row-major
    a_cm = torch.from_numpy(np.ascontiguousarray(a.transpose(0, 2, 1))).cuda()
    b_cm = torch.from_numpy(np.ascontiguousarray(b.transpose(0, 2, 1))).cuda()
    return a, b, a_cm, b_cm


def _run(lib, code, n, nrhs, count, a_cm, b_cm, noalloc):
    x = b_cm.clone()
    res = torch.full((count,), -1.0, dtype=torch.float64, device='cuda')
    status = torch.full((count,), 7, dtype=torch.int32, device='cuda')
    st = P(torch.cuda.current_stream().cuda_stream)
    args = (code, n, nrhs, count, P(a_cm.data_ptr()), n, n * n, P(x.data_ptr()), n, n * nrhs,
            8, 1e-12, _lib.CRIT_REL_FRO, P(res.data_ptr()), P(status.data_ptr()))
    if noalloc:
        rc = lib.hprmat_hip_batch_solve_noalloc(*args, st)
    else:
        wb = int(lib.hprmat_hip_batch_work_bytes(code, n, nrhs, count))
        work = torch.empty(wb, dtype=torch.uint8, device='cuda')
        rc = lib.hprmat_hip_batch_solve(*args, P(work.data_ptr()), wb, st)
    torch.cuda.synchronize()
    assert rc == 0
    return x, res, status


@pytest.mark.parametrize('method', ['fp32', 'fp64'])
def test_noalloc_matches_workspace_entry_point(method):
    lib = _lib_noalloc()
    code = (_lib.MIXED_FP32 if method == 'fp32' else _lib.PLAIN_FP64) | _lib.EQUILIBRATE
    # the second call is smaller and reuses the scratch array after the first
    for n, nrhs, count, seed in ((520, 4, 3, 11), (300, 2, 2, 12)):
        print('noalloc %s: n=%d nrhs=%d count=%d' % (method, n, nrhs, count), flush=True)
        assert lib.hprmat_hip_batch_work_bytes(code, n, nrhs, count) <= SCRATCH_BYTES
        a, b, a_cm, b_cm = _systems(n, nrhs, count, seed)
        x0, r0, s0 = _run(lib, code, n, nrhs, count, a_cm, b_cm, noalloc=True)
        x1, r1, s1 = _run(lib, code, n, nrhs, count, a_cm, b_cm, noalloc=False)
        assert torch.equal(x0, x1) and torch.equal(r0, r1) and torch.equal(s0, s1)
        assert s0.cpu().tolist() == [1] * count
        for k in range(count):
            xk = x0[k].cpu().numpy().T
            rel = np.linalg.norm(b[k] - a[k] @ xk) / np.linalg.norm(b[k])
            assert rel <= 1e-12 and abs(rel - float(r0[k])) <= 0.5 * rel + 1e-16


def test_noalloc_refuses_requests_larger_than_the_scratch_array():
    lib = _lib_noalloc()
    code = _lib.MIXED_FP32 | _lib.EQUILIBRATE
    n, nrhs, count = 2000, 8, 2     # about 91 MiB of workspace
    assert lib.hprmat_hip_batch_work_bytes(code, n, nrhs, count) > SCRATCH_BYTES
    a = torch.zeros((count, n, n), dtype=torch.complex128, device='cuda')
    b = torch.zeros((count, nrhs, n), dtype=torch.complex128, device='cuda')
    res = torch.zeros(count, dtype=torch.float64, device='cuda')
    status = torch.full((count,), 7, dtype=torch.int32, device='cuda')
    rc = lib.hprmat_hip_batch_solve_noalloc(code, n, nrhs, count, P(a.data_ptr()), n, n * n,
                                            P(b.data_ptr()), n, n * nrhs, 8, 1e-12,
                                            _lib.CRIT_REL_FRO, P(res.data_ptr()),
                                            P(status.data_ptr()),
                                            P(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert rc == -2
    assert status.cpu().tolist() == [7] * count     # nothing ran
