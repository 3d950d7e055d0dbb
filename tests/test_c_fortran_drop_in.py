"""The drop-in boundary from the reference's host languages.

C: tests/c/physics_drop_in.c is written against the reference's C binding
(bindings/c/hprmat.h: HPRMAT_SOLVER_*, hprmat_init, hprmat_solve,
hprmat_gpu_finalize, ...), includes "hprmat.h" and links to libhprmat_hip.
Fortran: hprmat_amd/fortran/hprmat_hip_solver.f90 declares the 15 bind(C)
names of src/gpu_solver_interface.F90:42-162 and wraps them with that
module's procedure names; tests/fortran/drop_in_solve.f90 calls every
solver through it.  Both are built with the image's compilers (gcc,
amdflang) into a temporary directory.

CPU tests: build, link and run the host-only paths; where the reference tree
is present in this container, its own bindings/c/test_ex1.c compiles and
links unchanged against include/, and the undefined gpu_*_ symbols of its
Fortran module src/gpu_solver_interface.F90 (compiled, not linked or run)
are all exported by the library.  GPU tests: the C program's collision
matrices against the oracle (same potential, printed by the program), and
the Fortran program's solutions of the N = 240 system against LAPACK zgesv.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from hprmat_amd import _lib
from oracle import lagrange as olagrange, rmatrix as ormat, synth as osynth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, 'include')
LIBDIR = os.path.dirname(_lib.LIB_PATH)
C_SRC = os.path.join(ROOT, 'tests', 'c', 'physics_drop_in.c')
F_MOD = os.path.join(ROOT, 'hprmat_amd', 'fortran', 'hprmat_hip_solver.f90')
F_PROG = os.path.join(ROOT, 'tests', 'fortran', 'drop_in_solve.f90')
FLANG = '/opt/rocm/bin/amdflang'
REF = '/root/reference'
LINK = ['-L' + LIBDIR, '-lhprmat_hip', '-lm', '-Wl,-rpath,' + LIBDIR]


def _run(cmd, timeout=120, **kw):
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, **kw)
    assert p.returncode == 0, '%s\n%s\n%s' % (' '.join(cmd), p.stdout[-3000:], p.stderr[-3000:])
    return p


def build_c(tmp):
    assert os.path.exists(_lib.LIB_PATH), 'libhprmat_hip.so not built'
    exe = os.path.join(str(tmp), 'physics_drop_in')
    _run(['gcc', '-O2', '-Wall', '-Werror', '-std=c11', '-I' + INCLUDE, C_SRC, '-o', exe] + LINK)
    return exe


def build_fortran(tmp):
    assert os.path.exists(_lib.LIB_PATH), 'libhprmat_hip.so not built'
    if not os.path.exists(FLANG):
        pytest.skip('amdflang is not installed')
    tmp = str(tmp)
    exe = os.path.join(tmp, 'drop_in_solve')
    _run([FLANG, '-O2', '-c', F_MOD, '-module-dir', tmp, '-o', os.path.join(tmp, 'mod.o')])
    _run([FLANG, '-O2', '-I', tmp, F_PROG, os.path.join(tmp, 'mod.o'), '-o', exe] + LINK)
    return exe


def test_c_program_builds_and_runs_host_calls(tmp_path):
    exe = build_c(tmp_path)
    p = _run([exe, '--info'])
    out = p.stdout
    assert 'info ok' in out and 'version hprmat_hip' in out
    for t in (1, 2, 3, 4):
        assert 'solver %d: ' % t in out
    # hprmat_init's mesh: the oracle's Lagrange-Legendre points on (0, 15)
    first, last = [float(v) for v in out.split('mesh ')[1].split()[:2]]
    mesh = olagrange.rmat_ini(60, 1, 15.0)
    r = mesh['xle'] * 15.0
    assert abs(first - r[0]) <= 1e-13 and abs(last - r[-1]) <= 1e-13


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, 'bindings', 'c')),
                    reason='reference tree not present')
def test_reference_c_example_compiles_unchanged(tmp_path):
    # bindings/c/test_ex1.c includes "hprmat.h" and uses HPRMAT_SOLVER_*:
    # compile and link it against include/ and the library (it is not run)
    exe = os.path.join(str(tmp_path), 'test_ex1')
    _run(['gcc', '-O2', '-Wall', '-I' + INCLUDE, os.path.join(REF, 'bindings', 'c', 'test_ex1.c'),
          '-o', exe] + LINK)
    assert os.path.exists(exe)


def test_fortran_module_builds_and_links(tmp_path):
    exe = build_fortran(tmp_path)
    if _lib.load().gpu_is_available_():
        return   # the GPU run is test_fortran_drop_in_n240
    p = _run([exe, os.devnull, os.devnull])
    assert 'devices 0' in p.stdout and 'no GPU' in p.stdout


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, 'src', 'gpu_solver_interface.F90')),
                    reason='reference tree not present')
def test_reference_fortran_module_symbols_are_exported(tmp_path):
    # the reference's own module (with its precision module) compiled with
    # amdflang: every gpu_*_ symbol it leaves undefined is defined by the library
    if not os.path.exists(FLANG):
        pytest.skip('amdflang is not installed')
    tmp = str(tmp_path)
    for src in ('precision.F90', 'gpu_solver_interface.F90'):
        _run([FLANG, '-c', os.path.join(REF, 'src', src), '-module-dir', tmp,
              '-o', os.path.join(tmp, src + '.o')], cwd=tmp)
    und = _run(['nm', '-u', os.path.join(tmp, 'gpu_solver_interface.F90.o')]).stdout.split()
    wanted = sorted(s for s in und if s.startswith('gpu_') and s.endswith('_'))
    assert wanted == sorted(_lib.REFERENCE_SYMBOLS)
    exported = _run(['nm', '-D', '--defined-only', _lib.LIB_PATH]).stdout
    have = set(l.split()[-1] for l in exported.splitlines() if ' T ' in l)
    assert not [s for s in wanted if s not in have]


def _parse_c_output(out):
    nch, nr = 3, 60
    cpot = np.zeros((nr, nch, nch), complex)
    qk = np.zeros(nch)
    smat = {}
    for line in out.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == 'qk':
            qk[int(f[1])] = float(f[2])
        elif f[0] == 'V':
            cpot[int(f[1]), int(f[2]), int(f[3])] = float(f[4]) + 1j * float(f[5])
        elif f[0] == 'S':
            solver, nopen, i, j = (int(v) for v in f[1:5])
            s = smat.setdefault(solver, np.zeros((nopen, nopen), complex))
            s[i, j] = float(f[5]) + 1j * float(f[6])
    return qk, cpot, smat


@pytest.mark.gpu
def test_c_program_collision_matrices_match_oracle(tmp_path):
    exe = build_c(tmp_path)
    print('C drop-in program: 3 channels x 60 points, solver types 1-4 and default', flush=True)
    p = _run([exe], timeout=300)
    assert 'done' in p.stdout
    qk, cpot, smat = _parse_c_output(p.stdout)
    ref = ormat.rmatrix(np.array([0, 1, 2]), qk, 15.0, 60, cpot)
    assert ref['nopen'] == 2 and qk[2] < 0
    # solver 2 stops at ||r||_F/||b||_F < 1e-10 (the reference's mixed rule)
    for solver, tol in ((1, 1e-10), (3, 1e-10), (4, 1e-10), (2, 1e-6), (0, 1e-10)):
        s = smat[solver]
        assert s.shape == (2, 2)
        err = np.abs(s - ref['S']).max()
        assert err <= tol, (solver, err)
    assert np.abs(smat[0] - smat[1]).max() == 0.0   # default type 1, bit-identical after finalize


@pytest.mark.gpu
def test_fortran_drop_in_n240(tmp_path):
    exe = build_fortran(tmp_path)
    lval, qk, cpot, _ = osynth.system(4, 60, 15.0, 3.0, 2.0, 0)
    mesh = olagrange.rmat_ini(60, 1, 15.0)
    from oracle import channels
    _, _, _, xc1 = channels.channel_functions(lval, qk, 15.0)
    a = ormat.assemble(mesh, lval, qk, cpot, xc1)
    b = ormat.rhs(mesh['q2'][0], 4)
    n, nrhs = a.shape[0], b.shape[1]
    fin, fout = os.path.join(str(tmp_path), 'in.bin'), os.path.join(str(tmp_path), 'out.bin')
    with open(fin, 'wb') as f:
        np.array([n, nrhs], np.int32).tofile(f)
        np.asfortranarray(a).ravel(order='F').astype(np.complex128).tofile(f)
        np.asfortranarray(b).ravel(order='F').astype(np.complex128).tofile(f)
    print('Fortran drop-in program: N=%d nrhs=%d, 7 solvers' % (n, nrhs), flush=True)
    p = _run([exe, fin, fout], timeout=300)
    assert 'finalize 0' in p.stdout, p.stdout
    x_ref = np.linalg.solve(a, b)
    raw = open(fout, 'rb').read()
    rec = 4 + 16 * n * nrhs
    assert len(raw) == 7 * rec
    names = ('double', 'mixed', 'hybrid', 'tf32', 'auto', 'multi', 'multi_mixed')
    for k, name in enumerate(names):
        ierr = int(np.frombuffer(raw[k * rec:k * rec + 4], np.int32)[0])
        x = np.frombuffer(raw[k * rec + 4:(k + 1) * rec], np.complex128).reshape((n, nrhs), order='F')
        assert ierr == 0, (name, ierr)
        rel = np.linalg.norm(b - a @ x) / np.linalg.norm(b)
        err = np.abs(x - x_ref).max() / np.abs(x_ref).max()
        print('%s: rel resid %.2e, rel err vs zgesv %.2e' % (name, rel, err), flush=True)
        # every solver refines in FP64 (tol 1e-13 on the largest residual entry,
        # at most 10 sweeps or products)
        assert rel <= 1e-11 and err <= 1e-8, (name, rel, err)
