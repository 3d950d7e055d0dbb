"""FP64 CPU reference for the gradient of S-matrix losses (oracle).

TEST INFRASTRUCTURE ONLY (see the package docstring).

The reference has no derivatives; this module restates its forward pipeline
(assembly src/rmatrix_hp.F90:246-293, ZGESV solve src/rmatrix_hp.F90:441-477,
R extraction src/rmat_solvers.F90:135-144, R -> S src/rmatrix_hp.F90:336-368)
through the other oracle modules and differentiates it by hand, as a chain
of adjoint steps written out here independently of the product's torch path:

  L = sum_E sum |S - T|^2,  S = Z_O^-1 Z_I,
  Z_O = diag(co) - C2 diag(cop),  Z_I = diag(conj co) - C2 diag(conj cop),
  C2[a, b] = R[o_a, o_b] / rmax sqrt(k_a / k_b),
  dL = 2 Re tr((S - T)^H dS),  dS = Z_O^-1 (dZ_I - dZ_O S)
     = 2 Re tr(Q dC2),  Q = (diag(cop) S - diag(conj cop)) P^H,  P = Z_O^-H (S - T)
     = 2 Re sum F . dX,  F[(i, r), j] = q2[r] Q[b, a] sqrt(k_a / k_b) / rmax
                          for i = o_a, j = o_b (zero for closed channels)
     = -2 Re tr(M^T dC X),  M = C^-T F = C^-1 F   (C complex symmetric)
     = -2 Re sum_{c, c', i} dV_cc'(r_i) K[i, c, c'],  K[i, c, c'] = sum_j M[(c, i), j] X[(c', i), j],
and the Woods-Saxon derivatives (V = s (V0 + i W0) f, f = 1 / (1 + e),
e = exp((r - R) / d)):
  dV/dV0 = s f,  dV/dW0 = i s f,  dV/dR = s (V0 + i W0) e / (d (1 + e)^2),
  dV/dd = s (V0 + i W0) e (r - R) / (d^2 (1 + e)^2).

Pinned by central finite differences of ``loss`` on small systems (tests),
which fix every factor and sign of the chain above.
"""
import numpy as np

from . import channels, lagrange, rmatrix, synth


def _potential_and_derivs(r, nch, theta):
    """cpot (nr, nch, nch) and dcpot (16, nr, nch, nch): d cpot / d theta[t, p]."""
    theta = np.asarray(theta, dtype=np.float64)
    nr = len(r)
    cpot = np.zeros((nr, nch, nch), dtype=complex)
    dc = np.zeros((4, 4, nr, nch, nch), dtype=complex)
    for c in range(nch):
        for c2 in range(nch):
            t = min(abs(c - c2), 3)
            s = 1.0 if t == 0 else 0.2
            v0, w0, rr, d = theta[t]
            for ir in range(nr):
                e = np.exp((r[ir] - rr) / d)
                f = 1.0 / (1.0 + e)
                amp = s * complex(v0, w0)
                cpot[ir, c, c2] = amp * f
                dc[t, 0, ir, c, c2] = s * f
                dc[t, 1, ir, c, c2] = 1j * s * f
                dc[t, 2, ir, c, c2] = amp * e / (d * (1.0 + e) ** 2)
                dc[t, 3, ir, c, c2] = amp * e * (r[ir] - rr) / (d * d * (1.0 + e) ** 2)
    return cpot, dc.reshape(16, nr, nch, nch)


def _forward(lval, qk, rmax, mesh, cpot):
    npo, co, cop, xc1 = channels.channel_functions(lval, qk, rmax)
    c = rmatrix.assemble(mesh, lval, qk, cpot, xc1)
    q2 = mesh['q2'][0]
    b = rmatrix.rhs(q2, len(lval))
    x = np.linalg.solve(c, b)
    r = rmatrix.extract_r(x, q2, len(lval))
    s = rmatrix.r_to_s(r, rmax, qk, npo, co, cop)
    return dict(C=c, X=x, R=r, S=s, npo=npo, co=np.asarray(co), cop=np.asarray(cop), q2=q2)


def problem(nch, nr, a, emax, seed, energies):
    """Channel data of the fit problem (term mode): lval, eps, mesh, energies."""
    lval, _, _, eps = synth.system(nch, nr, a, emax, 4.0, seed, mode='term')
    mesh = lagrange.rmat_ini(nr, 1, a)
    return dict(lval=np.asarray(lval), eps=np.asarray(eps), mesh=mesh, a=a, nch=nch, nr=nr,
                energies=np.asarray(energies, dtype=np.float64))


def _qk(eps, e):
    k2 = e - eps
    return np.where(k2 > 0, np.sqrt(np.abs(k2)), -np.sqrt(np.abs(k2)))


def s_matrices(prob, theta):
    r = prob['mesh']['xle'] * prob['a']
    cpot, _ = _potential_and_derivs(r, prob['nch'], theta)
    return [_forward(prob['lval'], _qk(prob['eps'], e), prob['a'], prob['mesh'], cpot)['S']
            for e in prob['energies']]


def loss(prob, theta, targets):
    return float(sum(np.sum(np.abs(s - t) ** 2) for s, t in zip(s_matrices(prob, theta), targets)))


def loss_and_grad(prob, theta, targets):
    """(L, dL/dtheta as a (4, 4) array) by the adjoint chain above."""
    nch, nr, a = prob['nch'], prob['nr'], prob['a']
    r = prob['mesh']['xle'] * a
    cpot, dcpot = _potential_and_derivs(r, nch, theta)
    total = 0.0
    grad = np.zeros(16)
    for e, tgt in zip(prob['energies'], targets):
        qk = _qk(prob['eps'], e)
        f = _forward(prob['lval'], qk, a, prob['mesh'], cpot)
        s, npo, co, cop = f['S'], f['npo'], f['co'], f['cop']
        total += float(np.sum(np.abs(s - tgt) ** 2))
        nopen = len(npo)
        c2 = np.zeros((nopen, nopen), dtype=complex)
        for m1 in range(nopen):
            for m2 in range(nopen):
                c2[m1, m2] = f['R'][npo[m1], npo[m2]] / a * np.sqrt(qk[npo[m1]] / qk[npo[m2]])
        zo = np.diag(co) - c2 * cop[None, :]
        p = np.linalg.solve(zo.conj().T, s - tgt)
        q = (np.diag(cop) @ s - np.diag(np.conj(cop))) @ p.conj().T
        ff = np.zeros((nch * nr, nch), dtype=complex)
        for ia, i in enumerate(npo):
            for jb, j in enumerate(npo):
                w = np.sqrt(qk[i] / qk[j]) / a
                ff[i * nr:(i + 1) * nr, j] = f['q2'] * q[jb, ia] * w
        m = np.linalg.solve(f['C'], ff)
        x = f['X']
        k = np.einsum('cij,dij->icd', m.reshape(nch, nr, nch), x.reshape(nch, nr, nch))
        grad += -2.0 * np.real(np.einsum('picd,icd->p', dcpot, k))
    return total, grad.reshape(4, 4)
