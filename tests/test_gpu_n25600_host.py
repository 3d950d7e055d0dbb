"""The benchmark system (N = 25600 complex: 128 channels x 200 mesh points,
128 right-hand sides) against the FP64 host oracle.

The oracle is the committed fixture tests/golden/lagrange_n25600.npz
(written by tests/golden/make_lagrange_n25600.py from the oracle's host
assembly): the reference's CPU path, solver type 1 (LAPACK ZGESV + R + S,
src/rmatrix_hp.F90:441-477,336-368), and the same LU factors refined by two
FP64 sweeps.  The system has kappa ~ 1e9: ZGESV's relative residual is
5.3e-12 and its S lies 2.9e-10 from the refined S (residual 7e-13, the FP64
floor), so S parity is gated against the refined host solution, and the
distance to ZGESV's S is reported.

Two matrices per method:
* the product path, assembled on the device (hprmat_hip_assemble).  Its
  entries differ from the oracle's host assembly by up to 1.5e-16 of
  max |C| (summation order and mesh formulas at rounding level), which on
  this system moves the residual measured against the other matrix to
  3.4e-11; the residual is therefore recomputed on the host against the
  device's own matrix (copied back);
* the oracle's matrix itself, uploaded: solver parity on identical bits,
  the host residual against the oracle matrix.
Methods: the benchmark's head (FP32 LU, Gauss three-product update,
equilibrated, FGMRES) and plain FP64 (LU + classic FP64 refinement), both
to ||B - A X||_F / ||B||_F <= 1e-12 by the library's own evaluation.
Gates: host-recomputed residual <= 1.2e-12 (the host's summation order
re-evaluates a residual at the 7e-13 floor to within about 1e-13),
|dS| <= 1e-10 against the refined oracle, max |dX| / max |X| <= 1e-9 on
the sampled rows.
"""
import os

import numpy as np
import pytest
import torch

from hprmat_amd import rmatrix, solver, synth
from oracle import channels, lagrange, rmatrix as ormat, synth as osynth

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'lagrange_n25600.npz')


def fingerprint(c):
    return np.array([c.sum(), np.abs(c).sum(), (c * c.conj()).real.sum(), c.trace()])


def same_matrix(fp, ref, rtol):
    """Fingerprints agree up to summation order: the complex sums to rtol of
    the sum of moduli, the sums of moduli and squares to rtol of themselves."""
    scale = abs(ref[1])
    return (abs(fp[0] - ref[0]) <= rtol * scale and abs(fp[3] - ref[3]) <= rtol * scale and
            abs(fp[1] - ref[1]) <= rtol * abs(ref[1]) and abs(fp[2] - ref[2]) <= rtol * abs(ref[2]))


@pytest.fixture(scope='module')
def host_system():
    gold = np.load(GOLD)
    nch, nr, a, emax, e, seed = gold['config']
    nch, nr, seed = int(nch), int(nr), int(seed)
    print('host oracle system: N=%d, assembling on the host' % (nch * nr), flush=True)
    lval, qk, cpot, _ = osynth.system(nch, nr, a, emax, e, seed)
    mesh = lagrange.rmat_ini(nr, 1, a)
    npo, co, cop, xc1 = channels.channel_functions(lval, qk, a)
    c = ormat.assemble(mesh, lval, qk, cpot, xc1)
    assert same_matrix(fingerprint(c), gold['fingerprint'], 1e-13)
    b = ormat.rhs(mesh['q2'][0], nch)
    return dict(c=c, b=b, gold=gold, q2=mesh['q2'][0], qk=qk, npo=npo, co=co, cop=cop,
                nch=nch, nr=nr, a=a)


@pytest.mark.parametrize('matrix', ['device', 'oracle'])
@pytest.mark.parametrize('method,refinement,budget', [('fp32_3m', 'fgmres', 24),
                                                      ('fp64', 'ir', 10)])
def test_benchmark_system_matches_host_oracle(host_system, method, refinement, budget, matrix):
    h = host_system
    gold = h['gold']
    s = synth.System(h['nch'], h['nr'], 15.0, 8.0, 5.0, 0)
    dm = rmatrix.DeviceMesh(s.mesh)
    o, co, cop, xc1 = rmatrix.channel_data(s.lval, s.qk, s.a)
    if matrix == 'device':
        a_cm = rmatrix.assemble(dm, s.lval, s.qk, xc1, s.cpot)
        # the device-assembled matrix is the oracle's up to rounding (A^T = A:
        # the row-major storage of the column-major matrix holds the same numbers)
        dev_fp = np.array([a_cm.sum().item(), a_cm.abs().sum().item(),
                           (a_cm.real ** 2 + a_cm.imag ** 2).sum().item(),
                           torch.diagonal(a_cm).sum().item()])
        assert same_matrix(dev_fp, gold['fingerprint'], 1e-12)
    else:
        a_cm = torch.from_numpy(h['c']).cuda()
    b = rmatrix.rhs_colmajor(dm, s.nch)
    print('N=25600 %s + %s on the %s matrix' % (method, refinement, matrix), flush=True)
    lu = solver.LU().factor(a_cm, method, equilibrate=True)
    st = lu.solve_(b, budget, 1e-12, 'rel_fro', True, refinement=refinement)
    assert st['info'] == 0 and st['converged'] and st['res_rel'] <= 1e-12, st
    _, s_dev = rmatrix.r_to_s(dm, b, s.nch, s.qk, o, co, cop)
    x = b.cpu().numpy().T                      # (N, nrhs)
    c = a_cm.cpu().numpy() if matrix == 'device' else h['c']
    del lu, a_cm, b
    torch.cuda.empty_cache()
    res = np.linalg.norm(h['b'] - c @ x) / np.linalg.norm(h['b'])
    del c
    r = ormat.extract_r(x, h['q2'], h['nch'])
    s_host = ormat.r_to_s(r, h['a'], h['qk'], h['npo'], h['co'], h['cop'])
    ds = np.abs(s_host - gold['S']).max()
    ds_dev = np.abs(s_dev.cpu().numpy() - gold['S']).max()
    ds_zgesv = np.abs(s_host - gold['S_zgesv']).max()
    xr = gold['X_rows']
    dx = np.abs(x[::int(gold['row_step'])] - xr).max() / np.abs(xr).max()
    print('%s on the %s matrix: host residual %.2e (library %.2e), |dS| vs refined oracle %.2e '
          '(device S %.2e), vs plain zgesv %.2e, dX %.2e, products %d'
          % (method, matrix, res, st['res_rel'], ds, ds_dev, ds_zgesv, dx, st['sweeps']),
          flush=True)
    assert res <= 1.2e-12
    assert ds <= 1e-10 and ds_dev <= 1e-10
    assert dx <= 1e-9


def test_reference_whit_channel_data_change_of_s_at_full_size():
    # the reference computes the closed-channel log-derivatives with WHIT,
    # which integrates by Numerov below kappa a = 10 (6 of the 55 closed
    # channels here, xc1 off by 3e-9 .. 2.4e-8 relative); the product uses
    # the closed form.  Both variants of the device-assembled matrix are
    # solved in FP64 (+ refinement) and the change of S is reported; it stays
    # far below the S tolerance of the parity gates
    s = synth.System(128, 200, 15.0, 8.0, 5.0, 0)
    dm = rmatrix.DeviceMesh(s.mesh)
    o, co, cop, xc1 = rmatrix.channel_data(s.lval, s.qk, s.a)
    xw = xc1.copy()
    low = 0
    for c in range(len(s.qk)):
        if s.qk[c] < 0:
            xw[c] = channels.closed_channel_whit(int(s.lval[c]), -s.qk[c], s.a)[0]
            low += -s.qk[c] * s.a < 10.0
    assert low == 6
    smat = []
    for x1 in (xc1, xw):
        a_cm = rmatrix.assemble(dm, s.lval, s.qk, x1, s.cpot)
        b = rmatrix.rhs_colmajor(dm, s.nch)
        lu = solver.LU().factor(a_cm, 'fp64', equilibrate=True)
        st = lu.solve_(b, 10, 1e-12, 'rel_fro', True)
        assert st['converged'], st
        smat.append(rmatrix.r_to_s(dm, b, s.nch, s.qk, o, co, cop)[1].cpu().numpy())
        del lu, a_cm, b
        torch.cuda.empty_cache()
    ds = np.abs(smat[0] - smat[1]).max()
    print('N=25600: |S(WHIT channel data) - S(closed form)| = %.2e' % ds, flush=True)
    assert 0.0 < ds <= 1e-8
