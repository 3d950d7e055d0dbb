"""Seeded synthetic Lagrange-mesh systems of the benchmark workloads (oracle).

Recipe (units hbar^2/2mu = 1, eta = 0, one interval):
* channel c = 0..nch-1 has l_c = c mod 4 and threshold eps_c ~ U[0, emax];
  k_c^2 = E - eps_c; qk_c = +sqrt(k^2) when open, -sqrt(-k^2) when closed
  (the reference's sign convention for closed channels, rmat_pierre/README).
* for every unordered pair c <= c' (c-major order) four uniforms u are
  drawn as one row of ``rng.random((npairs, 4))`` after the thresholds:
  V0 = -50 + 20 u0, W0 = -10 + 8 u1, R = 3 + 2 u2, d = 0.5 + 0.2 u3;
  V_cc'(r) = s (V0 + i W0) / (1 + exp((r - R)/d)), s = 1 if c = c' else 0.2,
  and V is symmetric.  ``w0_zero`` keeps the draws and sets W0 = 0.
* ``term`` mode (gradient workload): 4 coupling terms t = min(|c - c'|, 3)
  with parameters theta[t] = (V0, W0, R, d) drawn as ``rng.random((4, 4))``
  after the thresholds and mapped as above; s_t = 1 for t = 0, else 0.2.

The pair loop restates the config text literally, which is what pins it:
the product generator is checked against this one bit for bit.
"""
import math

import numpy as np


def draw(nch, emax, seed, mode='pair'):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(0.0, emax, nch)
    if mode == 'pair':
        npairs = nch * (nch + 1) // 2
        u = rng.random((npairs, 4))
    else:
        u = rng.random((4, 4))
    return eps, u


def params_from_u(u):
    return (-50.0 + 20.0 * u[..., 0], -10.0 + 8.0 * u[..., 1],
            3.0 + 2.0 * u[..., 2], 0.5 + 0.2 * u[..., 3])


def woods_saxon(r, v0, w0, rr, d):
    return complex(v0, w0) / (1.0 + math.exp((r - rr) / d))


def system(nch, nr, a, emax, energy, seed, w0_zero=False, mode='pair', theta=None):
    """Returns (lval, qk, cpot (nr, nch, nch), zrma-independent params).

    ``theta`` overrides the term parameters in ``term`` mode, as a (4, 4)
    array of (V0, W0, R, d) rows.
    """
    from .lagrange import legzo
    eps, u = draw(nch, emax, seed, mode)
    xle, _ = legzo(nr)
    r = xle * a
    lval = np.array([c % 4 for c in range(nch)])
    qk = np.zeros(nch)
    for c in range(nch):
        k2 = energy - eps[c]
        qk[c] = math.sqrt(k2) if k2 > 0 else -math.sqrt(-k2)
    cpot = np.zeros((nr, nch, nch), dtype=complex)
    if mode == 'pair':
        v0, w0, rr, d = params_from_u(u)
        p = 0
        for c in range(nch):
            for c2 in range(c, nch):
                s = 1.0 if c == c2 else 0.2
                wv = 0.0 if w0_zero else w0[p]
                for ir in range(nr):
                    v = s * woods_saxon(r[ir], v0[p], wv, rr[p], d[p])
                    cpot[ir, c, c2] = v
                    cpot[ir, c2, c] = v
                p += 1
    else:
        th = np.array(params_from_u(u)).T if theta is None else np.asarray(theta)
        for c in range(nch):
            for c2 in range(nch):
                t = min(abs(c - c2), 3)
                s = 1.0 if t == 0 else 0.2
                v0, w0, rr, d = th[t]
                if w0_zero:
                    w0 = 0.0
                for ir in range(nr):
                    cpot[ir, c, c2] = s * woods_saxon(r[ir], v0, w0, rr, d)
    return lval, qk, cpot, eps
