"""Lagrange-Legendre mesh of the R-matrix method (oracle restatement).

Follows src/rmatrix_hp.F90:61-159 (``rmat_ini``) and :741-783 (``LEGZO``).

Pinning:
* ``legzo`` against ``numpy.polynomial.legendre.leggauss`` (textbook routine).
* ``rmat_ini`` T+L against a brute-force Galerkin evaluation of
  <f_i| -d2/dr2 + L(0) |f_j> with the regularised Lagrange-Legendre basis,
  integrated by an independent high-order Gauss rule (tests/test_oracle.py),
  and through the full pipeline against the golden phase shifts of
  rmat_pierre/Ex2 and Ex5.
"""
import math

import numpy as np


def legzo(n):
    """Gauss-Legendre nodes/weights mapped to (0, 1), ascending.

    Newton iteration with deflation as in src/rmatrix_hp.F90:741-783.
    """
    x = [0.0] * n
    w = [0.0] * n
    pi = 3.1415926535898
    n0 = (n + 1) // 2
    for nr in range(1, n0 + 1):
        z = math.cos(pi * (nr - 0.25) / (n + 0.5))
        while True:
            z0 = z
            p = 1.0
            for i in range(nr - 1):
                p *= (z - x[i])
            f0 = 1.0
            if nr == n0 and n % 2 == 1:
                z = 0.0
            f1 = z
            pf = z
            pd = 1.0
            for k in range(2, n + 1):
                pf = (2.0 - 1.0 / k) * z * f1 - (1.0 - 1.0 / k) * f0
                pd = k * (f1 - z * pf) / (1.0 - z * z)
                f0 = f1
                f1 = pf
            if n == 1:
                pf, pd = z, 1.0
            if z == 0.0:
                break
            fd = pf / p
            q = 0.0
            for i in range(nr - 1):
                wp = 1.0
                for j in range(nr - 1):
                    if j != i:
                        wp *= (z - x[j])
                q += wp
            gd = (pd - q * fd) / p
            z = z - fd / gd
            if abs(z - z0) <= abs(z) * 1.0e-15:
                break
        x[nr - 1] = z
        x[n - nr] = -z
        w[nr - 1] = 2.0 / ((1.0 - z * z) * pd * pd)
        w[n - nr] = w[nr - 1]
    xs = np.array([(1.0 + x[n - 1 - i]) / 2.0 for i in range(n)])
    ws = np.array([w[n - 1 - i] / 2.0 for i in range(n)])
    return xs, ws


def rmat_ini(nr, ns, rmax):
    """Mesh matrices of ``rmat_ini`` (src/rmatrix_hp.F90:61-159).

    Returns a dict with xle, wle (nr), tc (ns, nr, nr), blo0/blo1/blo2
    (nr, nr), q1/q2 (ns, nr) and zrma (ns*nr), all scaled by the interval
    length rmax/ns exactly as the reference does.
    """
    xle, wle = legzo(nr)
    rmax0 = rmax / ns
    tc = np.zeros((ns, nr, nr))
    blo0 = np.zeros((nr, nr))
    blo1 = np.zeros((nr, nr))
    blo2 = np.zeros((nr, nr))
    q1 = np.zeros((ns, nr))
    q2 = np.zeros((ns, nr))
    for is_ in range(1, ns + 1):
        for i1 in range(1, nr + 1):
            xi = xle[i1 - 1]
            xi2 = xi * (1 - xi)
            if is_ == 1:
                xx = 4 * nr * (nr + 1) + 3 + (1 - 6 * xi) / xi2
                tc[0, i1 - 1, i1 - 1] = xx / (3 * xi2)
                blo0[i1 - 1, i1 - 1] = 1 / xi2
            else:
                xlb = xi / (1 - xi) * (nr * (nr + 1) - 1 / (1 - xi))
                xla = (1 - xi) / xi * (-nr * (nr + 1) + 1 / xi)
                tc[is_ - 1, i1 - 1, i1 - 1] = (nr * nr + nr + 6 - 2 / xi2) / (3 * xi2) + xlb - xla
                blo1[i1 - 1, i1 - 1] = (1 - xi) / xi
                blo2[i1 - 1, i1 - 1] = xi / (1 - xi)
            for i2 in range(1, i1):
                xj = xle[i2 - 1]
                xj2 = xj * (1 - xj)
                if is_ == 1:
                    xx = (nr * (nr + 1) + 1 + (xi + xj - 2 * xi * xj) / (xi - xj) ** 2
                          - 1 / (1 - xi) - 1 / (1 - xj))
                    t = xx / math.sqrt(xi2 * xj2)
                    b0 = 1 / math.sqrt(xi2 * xj2)
                    b1 = b2 = 0.0
                else:
                    yy = math.sqrt(xj2 / xi2 ** 3) * (2 * xi * xj + 3 * xi - xj - 4 * xi ** 2) / (xj - xi) ** 2
                    xlb = math.sqrt(xi * xj / (1 - xi) / (1 - xj)) * (nr * (nr + 1) - 1 / (1 - xj))
                    xla = math.sqrt((1 - xi) * (1 - xj) / xi / xj) * (-nr * (nr + 1) + 1 / xj)
                    t = yy + xlb - xla
                    b0 = blo0[i1 - 1, i2 - 1]
                    b1 = math.sqrt((1 - xi) * (1 - xj) / xi / xj)
                    b2 = math.sqrt(xi * xj / (1 - xi) / (1 - xj))
                # The reference flips the odd (i1+i2) entries of all Bloch blocks
                # on every interval, so blo0 is re-flipped for each is > 1
                # (src/rmatrix_hp.F90:126-131); restated as is.
                if (i1 + i2) % 2 == 1:
                    t, b0, b1, b2 = -t, -b0, -b1, -b2
                tc[is_ - 1, i1 - 1, i2 - 1] = t
                tc[is_ - 1, i2 - 1, i1 - 1] = t
                blo0[i1 - 1, i2 - 1] = blo0[i2 - 1, i1 - 1] = b0
                if is_ > 1:
                    blo1[i1 - 1, i2 - 1] = blo1[i2 - 1, i1 - 1] = b1
                    blo2[i1 - 1, i2 - 1] = blo2[i2 - 1, i1 - 1] = b2
        if is_ == 1:
            q2[0] = 1 / np.sqrt(xle * (1 - xle))
            q1[0] = 0
        else:
            q2[is_ - 1] = np.sqrt(xle / (1 - xle))
            q1[is_ - 1] = -1 / q2[is_ - 1]
        if nr % 2 == 1:
            q2[is_ - 1] = -q2[is_ - 1]
        q1[is_ - 1, 0::2] = -q1[is_ - 1, 0::2]
        q2[is_ - 1, 0::2] = -q2[is_ - 1, 0::2]
    tc = tc / rmax0 ** 2
    blo0 = blo0 / rmax0
    blo1 = blo1 / rmax0
    blo2 = blo2 / rmax0
    q1 = q1 / math.sqrt(rmax0)
    q2 = q2 / math.sqrt(rmax0)
    zrma = np.concatenate([((is_ - 1) + xle) * rmax0 for is_ in range(1, ns + 1)])
    return dict(nr=nr, ns=ns, rmax=rmax, rmax0=rmax0, xle=xle, wle=wle, tc=tc,
                blo0=blo0, blo1=blo1, blo2=blo2, q1=q1, q2=q2, zrma=zrma)
