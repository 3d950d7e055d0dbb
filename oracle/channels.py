"""Channel (boundary) functions at eta = 0 (oracle restatement).

The reference evaluates them in src/rmatrix_hp.F90:224-244 with COULFG
(Steed's method, src/special_functions.f:1-276) for open channels and WHIT
(src/special_functions.f:309-416) for closed channels.  With eta = 0 both
reduce to textbook functions:

* open:   F_l(rho) = rho j_l(rho),  G_l(rho) = -rho y_l(rho);
          co  = G + iF,  cop = (G' + iF') * k * a   (derivatives in rho)
* closed: W_{0,l+1/2}(2 rho) = (2/pi) rho k_l(rho) (normalised to e^-rho),
          xc1 = a * (kappa W') / W = rho W'(rho) / W(rho),  xc2 = W(rho)

Pinning: mpmath.coulombf / coulombg (eta = 0) and mpmath.whitw, plus the
Wronskian F'G - FG' = 1 (tests/test_oracle.py).  ``whit`` restates WHIT
itself: for kappa*a >= 10 it is the closed form to rounding, below it
integrates numerically (Numerov) and differs from the exact W by about
1e-8 relative; ``closed_channel_whit`` gives the reference's xc1 from it.
The product uses the closed form; the S-matrix change this causes is
measured in tests/test_oracle.py and tests/test_gpu_n25600_host.py.
"""
import math

import numpy as np
from scipy import special


def open_channel(l, k, a):
    """(co, cop) for an open channel, src/rmatrix_hp.F90:228-233."""
    rho = k * a
    j = special.spherical_jn(l, rho)
    jd = special.spherical_jn(l, rho, derivative=True)
    y = special.spherical_yn(l, rho)
    yd = special.spherical_yn(l, rho, derivative=True)
    f = rho * j
    fd = j + rho * jd
    g = -rho * y
    gd = -(y + rho * yd)
    co = complex(g, f)
    cop = complex(gd, fd) * k * a
    return co, cop


def closed_channel(l, kappa, a):
    """(xc1, xc2) for a closed channel, src/rmatrix_hp.F90:234-242."""
    rho = kappa * a
    kl = special.spherical_kn(l, rho)
    kld = special.spherical_kn(l, rho, derivative=True)
    w = (2.0 / math.pi) * rho * kl
    wd = (2.0 / math.pi) * (kl + rho * kld)
    return rho * wd / w, w


def _sp(num, den=1.0):
    """A default-REAL (single precision) Fortran literal expression such as
    10.0/12.0, folded in single precision and promoted to double: the
    reference is compiled without -fdefault-real-8 (Makefile:21)."""
    return float(np.float32(num) / np.float32(den))


def whit(heta, r, xk, ll, ie=0):
    """Restatement of WHIT (src/special_functions.f:309-416): the Whittaker
    function W_{-eta, l+1/2}(2 k r) normalised to exp(-(k r + eta log 2kr))
    for l = 0 .. ll, and its derivative in r.  Returns (F, FD, IE), F and FD
    scaled by e^IE.

    At rho = k r >= 10 (eta + 1) the 26-term asymptotic series is summed at
    rho itself (:376-398).  Below, the series is summed at two grid points
    beyond rho and Numerov's method integrates W'' = (1 - 2 eta/rho) W
    inwards with step h = rho / max(int(30 rho + 1), 4) (:358-375), and W'
    comes from the 7-point central difference (:374); higher l follow by the
    upward recurrence (:404-409).  Constants the source writes as default
    REAL literals (10.0/12.0, 1.0/6.0, 1.0/12.0, 1.0/60.0, 0.15) are used at
    their single-precision values, as the reference's build computes them.

    Parity: at eta = 0 and rho >= 10 the series is exact (every term after
    the first vanishes) and F = exp(-rho); below, the result differs from the
    exact W by the Numerov and difference-formula errors, of order 1e-8
    relative (tests/test_oracle.py pins both regimes against mpmath.whitw).
    """
    fpmax = 1.0e290
    nl = ll + 1
    ee = -1.0
    ak = xk
    eta = heta
    rho = ak * r
    s = [0.0] * 8                      # S(1..7)
    t = [0.0] * 4                      # T(1..3)
    h = float(max(int(30.0 * rho + 1.0), 4))
    h = rho / h
    rhoa = 10.0 * (eta + 1.0)
    if rhoa - rho <= 0.0:
        ifeql = 1
        rhoa = rho
    else:
        ifeql = 0
    rhoa = h * int(rhoa / h + 0.5)
    if ifeql <= 0 and rhoa - rho - 1.5 * h < 0.0:
        rhoa = rho + 2.0 * h
    f = [0.0] * (nl + 1)
    fd = [0.0] * (nl + 1)
    slot = 7
    while True:                        # label 55: the asymptotic series at rhoa
        c = 1.0 / rhoa
        a = 1.0
        b = 1.0 - c * eta
        f1, fd1 = a, b
        for m in range(1, 27):
            d = 0.5 * (eta + float(m - 1)) * (eta + float(m)) * c / float(m)
            a = -a * d
            b = -b * d - a * c
            f1 += a
            fd1 += b
        a = -eta * math.log(2.0 * rhoa) - rhoa
        fpminl = -math.log(fpmax)
        if ie == 0 and a < fpminl:
            ie = int(fpminl - a)
        a = math.exp(a + ie)
        f1 = a * f1
        fd1 = a * fd1 * (-1.0 - 2.0 * eta / rhoa)
        if ifeql > 0:
            f[1], fd[1] = f1, fd1
            break
        s[slot] = f1                   # label 57
        if slot == 7:
            slot = 6
            rhoa = rhoa + h
            continue
        # label 27: Numerov inwards from rhoa to rho - 3h
        ca = 2.0 - _sp(10.0, 12.0) * h * h * ee
        cb = _sp(1.0, 6.0) * h * eta
        cc = 1.0 + _sp(1.0, 12.0) * h * h * ee
        m1 = int(rhoa / h - 0.5)
        m2 = int(rho / h - 1.5)
        t[2] = cb / float(m1 + 1)
        t[3] = cb / float(m1)
        js = m1
        for _ in range(m2, m1 + 1):
            s[1:7] = s[2:8]
            t[1] = t[2]
            t[2] = t[3]
            t[3] = cb / float(js - 1)
            s[7] = ((ca + 10.0 * t[2]) * s[6] - (cc - t[1]) * s[5]) / (cc - t[3])
            js -= 1
            if abs(s[7]) > fpmax:
                for i in range(2, 8):
                    s[i] = s[i] / fpmax
        f[1] = s[4]
        fd[1] = (_sp(1.0, 60.0) * (s[1] - s[7]) + _sp(0.15) * (s[6] - s[2]) +
                 0.75 * (s[3] - s[5])) / h
        break
    c = 1.0 / rho                      # label 61: upward recurrence in l
    for m in range(1, nl):
        a = eta / float(m)
        b = a + c * float(m)
        f[m + 1] = (b * f[m] - fd[m]) / (a + 1.0)
        fd[m + 1] = (a - 1.0) * f[m] - b * f[m + 1]
    for m in range(1, nl + 1):
        fd[m] = ak * fd[m]
    return np.array(f[1:]), np.array(fd[1:]), ie


def closed_channel_whit(l, kappa, a):
    """(xc1, xc2) as the reference computes them, through WHIT at the
    channel radius (src/rmatrix_hp.F90:234-242): xc1 = a W'/W, xc2 = W."""
    f, fd, _ = whit(0.0, a, kappa, l, 0)
    return a * fd[l] / f[l], f[l]


def channel_functions(lval, qk, a):
    """Per-channel boundary data in the order the reference builds it.

    Returns (npo_open, co, cop, xc1) where npo_open lists open channel
    indices (0-based), co/cop are per open channel and xc1 is per channel
    (zero for open ones).
    """
    nch = len(lval)
    npo = []
    co = []
    cop = []
    xc1 = np.zeros(nch)
    for i in range(nch):
        if qk[i] > 0:
            npo.append(i)
            c, cp = open_channel(int(lval[i]), qk[i], a)
            co.append(c)
            cop.append(cp)
        else:
            xc1[i] = closed_channel(int(lval[i]), -qk[i], a)[0]
    return npo, np.array(co, dtype=complex), np.array(cop, dtype=complex), xc1
