"""Assembly of C = H - E + Bloch, R-matrix extraction and R -> S (oracle).

Follows src/rmatrix_hp.F90:246-293 (assembly, one interval ns = 1),
src/rmat_solvers.F90:120-144 (RHS layout and R extraction) and
src/rmatrix_hp.F90:336-368 (Z_O, Z_I and S = Z_O^-1 Z_I).

Pinning: golden phase shifts of rmat_pierre/Ex2 (Reid NN, two coupled
channels) and rmat_pierre/Ex5 (non-local Yamaguchi, with its analytic phase
shift, Phys. Rev. 95 (1954) 1628), and the invariants S^T = S, S^dagger S = I
for real potentials with all channels open, |S| <= 1 with absorption
(tests/test_oracle.py).
"""
import numpy as np

from . import channels
from .solve import zgesv


def assemble(mesh, lval, qk, cpot, xc1, cpnl=None):
    """Dense C for interval 1 (ns = 1), src/rmatrix_hp.F90:246-293.

    cpot: (nr, nch, nch) complex local potential divided by hbar^2/2mu.
    cpnl: optional (nr, nr, nch, nch) non-local kernel.
    """
    nr = mesh['nr']
    nch = len(lval)
    ntot = nch * nr
    xle, wle = mesh['xle'], mesh['wle']
    rmax0 = mesh['rmax0']
    rmax2 = rmax0
    tc = mesh['tc'][0]
    blo0 = mesh['blo0']
    ch = np.zeros((ntot, ntot), dtype=complex)
    m1 = 0
    for i1 in range(nch):
        ch[m1:m1 + nr, m1:m1 + nr] = tc
        fac = lval[i1] * (lval[i1] + 1)
        for ir in range(nr):
            xx = xle[ir] * rmax0
            ch[m1 + ir, m1 + ir] += fac / xx ** 2
            if qk[i1] > 0:
                ch[m1 + ir, m1 + ir] -= qk[i1] ** 2
            if qk[i1] < 0:
                ch[m1 + ir, m1 + ir] += qk[i1] ** 2
        if qk[i1] < 0:
            ch[m1:m1 + nr, m1:m1 + nr] -= xc1[i1] * blo0 / rmax2
        m2 = 0
        for i2 in range(nch):
            for ir in range(nr):
                ch[m1 + ir, m2 + ir] += cpot[ir, i1, i2]
                if cpnl is not None:
                    for irp in range(nr):
                        ch[m1 + ir, m2 + irp] += np.sqrt(wle[ir] * wle[irp]) * rmax0 * cpnl[ir, irp, i1, i2]
            m2 += nr
        m1 += nr
    return ch


def rhs(q2, nch):
    """B(:, c) = q2 in the rows of block c, src/rmat_solvers.F90:120-124."""
    nr = len(q2)
    b = np.zeros((nch * nr, nch), dtype=complex)
    for ich in range(nch):
        for ir in range(nr):
            b[ich * nr + ir, ich] = q2[ir]
    return b


def extract_r(x, q2, nch, normfac=1.0):
    """R(i, j) = normfac * sum_r q2(r) X((i-1) nr + r, j), src/rmat_solvers.F90:135-144."""
    nr = len(q2)
    r = np.zeros((nch, nch), dtype=complex)
    for jch in range(nch):
        for ich in range(nch):
            acc = 0j
            for ir in range(nr):
                acc += q2[ir] * x[ich * nr + ir, jch]
            r[ich, jch] = acc
    return r * normfac


def r_to_s(rmat, rmax, qk, npo, co, cop):
    """S = Z_O^-1 Z_I over the open channels, src/rmatrix_hp.F90:336-368."""
    nopen = len(npo)
    crma0 = rmat / rmax
    crma2 = np.zeros((nopen, nopen), dtype=complex)
    for m1 in range(nopen):
        for m2 in range(nopen):
            crma2[m1, m2] = crma0[npo[m1], npo[m2]] * np.sqrt(qk[npo[m1]] / qk[npo[m2]])
    cz = np.zeros((nopen, nopen), dtype=complex)
    cu = np.zeros((nopen, nopen), dtype=complex)
    for m1 in range(nopen):
        cz[:, m1] = -crma2[:, m1] * cop[m1]
        cu[:, m1] = -crma2[:, m1] * np.conj(cop[m1])
        cz[m1, m1] += co[m1]
        cu[m1, m1] += np.conj(co[m1])
    s, info = zgesv(cz, cu)
    if info != 0:
        raise RuntimeError('Error in ZGESV for collision matrix')
    return s


def rmatrix(lval, qk, rmax, nr, cpot, cpnl=None, mesh=None):
    """Type-1 pipeline for ns = 1: assemble, ZGESV, R, S.

    Returns a dict with C, B, X, R, S, nopen and the channel data.
    """
    from .lagrange import rmat_ini
    if mesh is None:
        mesh = rmat_ini(nr, 1, rmax)
    nch = len(lval)
    npo, co, cop, xc1 = channels.channel_functions(lval, qk, rmax)
    c = assemble(mesh, lval, qk, cpot, xc1, cpnl)
    q2 = mesh['q2'][0]
    b = rhs(q2, nch)
    x, info = zgesv(c, b)
    if info != 0:
        raise RuntimeError('Error in solve_rmatrix_hp ZGESV')
    r = extract_r(x, q2, nch)
    s = r_to_s(r, rmax, qk, npo, co, cop) if npo else np.zeros((0, 0), dtype=complex)
    return dict(C=c, B=b, X=x, R=r, S=s, nopen=len(npo), npo=npo, co=co, cop=cop,
                xc1=xc1, mesh=mesh)
