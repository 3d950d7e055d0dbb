"""CPU tests of the host side: C-ABI exports, product mesh / channel data /
synthetic generator against the oracle, and argument checking."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from hprmat_amd import _lib, rmatrix, synth
from oracle import channels as ochannels, lagrange as olagrange, synth as osynth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'hprmat_hip.h')


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'^\s*(?:int|void|size_t|const char)\s*\*?\s*(\w+)\s*\(', txt,
                                 flags=re.M)))


def test_header_declares_reference_abi():
    names = header_functions()
    for n in _lib.REFERENCE_SYMBOLS:
        assert n in names, n
    assert len([n for n in names if n.startswith('gpu_')]) == 15


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), 'libhprmat_hip.so not built'
    out = subprocess.run(['nm', '-D', '--defined-only', _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if ' T ' in l)
    missing = [n for n in header_functions() if n not in exported]
    assert not missing, missing


def test_library_loads_and_binds():
    # loading binds every symbol; no device call is made here
    code = ('import sys; sys.path.insert(0, %r); from hprmat_amd import _lib; '
            'lib = _lib.load(); print(lib.hprmat_hip_version().decode())' % ROOT)
    p = subprocess.run([sys.executable, '-c', code], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert 'hprmat_hip' in p.stdout


def test_batch_workspace_and_argument_checks():
    # host-only paths of the batch entry points: workspace arithmetic and the
    # argument checks that return before any device call
    code = '''
import sys, ctypes
sys.path.insert(0, %r)
from hprmat_amd import _lib
lib = _lib.load()
wb = lib.hprmat_hip_batch_work_bytes
one, four = wb(1, 1000, 8, 1), wb(1, 1000, 8, 4)
assert one > 1000 * 1000 * 8 and 3 * one < four <= 4 * one + 4096, (one, four)
assert wb(0, 1000, 8, 1) > one            # FP64 factors are twice as large
assert wb(2, 1000, 8, 1) == one            # Gauss 3-product FP32: the FP32 workspace
assert wb(3, 1000, 8, 1) == 0 and wb(1, 0, 8, 1) == 0 and wb(1, 10, 0, 1) == 0
nul = ctypes.c_void_p(0)
assert lib.hprmat_hip_batch_solve(3, 10, 1, 1, nul, 10, 100, nul, 10, 10, 0, 0.0, 1, nul, nul,
                                  nul, 0, nul) == -1     # split-BF16: not batched
assert lib.hprmat_hip_batch_solve(1, 10, 1, 1, nul, 10, 100, nul, 10, 10, 0, 0.0, 0x11, nul,
                                  nul, nul, 0, nul) == -1  # FGMRES: not batched
assert lib.hprmat_hip_batch_solve(1, 10, 1, 2, ctypes.c_void_p(16), 10, 50, ctypes.c_void_p(16),
                                  10, 10, 0, 0.0, 1, nul, nul, ctypes.c_void_p(16), 1 << 20,
                                  nul) == -1               # overlapping matrices
assert lib.hprmat_hip_batch_solve(1, 0, 1, 5, nul, 1, 0, nul, 1, 0, 0, 0.0, 1, nul, nul, nul,
                                  0, nul) == 0             # empty systems: nothing to do
assert lib.hprmat_hip_r_to_s_batch(65, 10, 1, nul, 650, 650 * 65, nul, 1.0, 1.0, nul, nul, nul,
                                   nul, nul, nul, nul, nul) == -1   # nch > 64
for bad in (1 | (3 << 16), 1 | (2 << 20), 1 | (1 << 12), 1 | (1 << 24)):   # unknown plan bits
    assert lib.hprmat_hip_batch_solve(bad, 10, 1, 1, nul, 10, 100, nul, 10, 10, 0, 0.0, 1, nul,
                                      nul, nul, 0, nul) == -1, hex(bad)
from hprmat_amd import solver
assert solver.batch_code('fp32_3m', trail=2, subst=1) == 2 | 0x100 | (2 << 16) | (1 << 20)
assert solver.batch_code('fp64', equilibrate=False) == 0
for kw in ({'trail': 3}, {'subst': 2}):
    try:
        solver.batch_code('fp32', **kw)
    except ValueError:
        pass
    else:
        raise AssertionError(kw)
print('ok')
''' % ROOT
    p = subprocess.run([sys.executable, '-c', code], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert 'ok' in p.stdout


@pytest.mark.parametrize('nr,a', [(10, 8.0), (60, 7.0), (200, 15.0)])
def test_product_mesh_matches_oracle(nr, a):
    m = olagrange.rmat_ini(nr, 1, a)
    p = rmatrix.rmat_ini(nr, a)
    assert np.allclose(p.tc, m['tc'][0], rtol=1e-11, atol=1e-11 * np.abs(p.tc).max())
    assert np.allclose(p.blo0, m['blo0'], rtol=1e-11, atol=1e-11 * np.abs(p.blo0).max())
    assert np.allclose(p.q2, m['q2'][0], rtol=1e-12)
    assert np.allclose(p.r, m['zrma'], rtol=1e-14)


def test_product_channel_data_matches_oracle():
    lval = [0, 1, 2, 3, 0, 1, 2]
    qk = np.array([0.5, -0.3, 1.2, -2.0, 0.01, 2.5, -0.05])
    npo, co, cop, xc1 = ochannels.channel_functions(lval, qk, 15.0)
    o, c0, c1, x1 = rmatrix.channel_data(lval, qk, 15.0)
    assert list(o) == npo
    assert np.allclose(c0, co, rtol=1e-13) and np.allclose(c1, cop, rtol=1e-13)
    assert np.allclose(x1, xc1, rtol=1e-13)


@pytest.mark.parametrize('mode', ['pair', 'term'])
def test_product_generator_matches_oracle_bitwise(mode):
    s = synth.System(6, 30, 15.0, 3.0, 2.0, 11, mode=mode)
    lval, qk, cpot, eps = osynth.system(6, 30, 15.0, 3.0, 2.0, 11, mode=mode)
    assert np.array_equal(s.lval, lval)
    assert np.array_equal(s.qk, qk)
    assert np.array_equal(s.eps, eps)
    # same draws, same arithmetic up to the exp() evaluation order
    assert np.allclose(s.cpot, cpot, rtol=1e-14, atol=1e-14)
    assert np.array_equal(s.cpot, np.transpose(s.cpot, (0, 2, 1)))


def test_named_workloads():
    s = synth.config('small')
    assert s.n == 240 and s.nch == 4
    assert 0 < int((s.qk > 0).sum()) <= 4


def test_solver_rejects_non_device_tensors():
    import torch
    from hprmat_amd import solver
    with pytest.raises((TypeError, RuntimeError)):
        solver.solve(torch.zeros((4, 4), dtype=torch.complex128), torch.zeros((4, 1), dtype=torch.complex128))


def test_solver_type_table_covers_gpu_types():
    for t in (1, 2, 4, 5, 6):
        assert t in rmatrix.SOLVER_TYPES
    with pytest.raises(NotImplementedError):
        rmatrix.solve_colmajor(None, None, solver_type=3)
