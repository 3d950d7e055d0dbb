"""Writes tests/golden/lagrange_n25600.npz: the FP64 host oracle of the
N = 25600 Lagrange-mesh system (128 channels x 200 mesh points on (0, 15),
thresholds U[0, 8], E = 5, seed 0; 128 right-hand sides, one per channel),
assembled by the oracle and solved on the host in FP64.

Two solutions are stored:
* the reference's CPU path, solver type 1: LAPACK ZGESV (here ZGETRF +
  ZGETRS, which is what ZGESV does; src/rmatrix_hp.F90:441-477), its R
  (src/rmat_solvers.F90:135-144) and S = Z_O^-1 Z_I (src/rmatrix_hp.F90:
  336-368) -> S_zgesv;
* the same factors plus two FP64 refinement sweeps X += LU^-1 (B - C X) ->
  S, R, X_rows.  On this system (kappa ~ 1e9) ZGESV's relative residual is
  5e-12 and its S is about 3e-10 away from the refined one, whose residual
  is at the FP64 floor (7e-13): the refined solution is the accurate FP64
  oracle, ZGESV's is the reference's own output.
Also stored: every 400th row of the refined X, the residual of each iterate, and
fingerprints of the matrix (so a test can tell that it assembled the same
system).  About 2.5 minutes and 21 GB of host memory with 8 threads.

Usage: python tests/golden/make_lagrange_n25600.py
"""
import os
import sys
import time

import numpy as np
from scipy.linalg import lapack

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import channels, lagrange, rmatrix as ormat, synth as osynth  # noqa: E402

NCH, NR, A, EMAX, E, SEED = 128, 200, 15.0, 8.0, 5.0, 0
ROW_STEP = 400
SWEEPS = 2


def fingerprint(c):
    """Sums of the matrix that pin it to 1e-15 relative."""
    return np.array([c.sum(), np.abs(c).sum(), (c * c.conj()).real.sum(), c.trace()])


def main():
    t0 = time.time()
    lval, qk, cpot, _ = osynth.system(NCH, NR, A, EMAX, E, SEED)
    mesh = lagrange.rmat_ini(NR, 1, A)
    npo, co, cop, xc1 = channels.channel_functions(lval, qk, A)
    c = ormat.assemble(mesh, lval, qk, cpot, xc1)
    q2 = mesh['q2'][0]
    b = ormat.rhs(q2, NCH)
    print('assembled N=%d in %.1f s' % (c.shape[0], time.time() - t0), flush=True)
    fp = fingerprint(c)
    t1 = time.time()
    # C^T = C exactly: the row-major array is the column-major matrix
    lu, piv, info = lapack.zgetrf(np.array(c.T, order='F', copy=True), overwrite_a=True)
    assert info == 0
    x, info = lapack.zgetrs(lu, piv, b)
    assert info == 0
    print('zgetrf + zgetrs %.1f s' % (time.time() - t1), flush=True)
    out = {}
    res = []
    for sweep in range(SWEEPS + 1):
        r = b - c @ x
        res.append(np.linalg.norm(r) / np.linalg.norm(b))
        rm = ormat.extract_r(x, q2, NCH)
        sm = ormat.r_to_s(rm, A, qk, npo, co, cop)
        if sweep == 0:
            out.update(S_zgesv=sm)
        print('iterate %d: ||B - C X||_F / ||B||_F = %.3e' % (sweep, res[-1]), flush=True)
        if sweep == SWEEPS:
            break
        d, info = lapack.zgetrs(lu, piv, r)
        assert info == 0
        x = x + d
    out.update(S=sm, R=rm, X_rows=x[::ROW_STEP].copy())
    print('|S - S_zgesv| = %.3e' % np.abs(out['S'] - out['S_zgesv']).max(), flush=True)
    path = os.path.join(HERE, 'lagrange_n25600.npz')
    np.savez_compressed(path, row_step=ROW_STEP, res_rel=np.array(res), fingerprint=fp, qk=qk,
                        lval=lval, npo=np.array(npo),
                        config=np.array([NCH, NR, A, EMAX, E, SEED]), **out)
    print('wrote %s: nopen %d, %.1f s' % (path, len(npo), time.time() - t0))


if __name__ == '__main__':
    main()
