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
Wronskian F'G - FG' = 1 (tests/test_oracle.py).  WHIT integrates
numerically (Numerov) when kappa*a < 10; its value differs from the exact
W by the integration error, which is why the channel functions enter the
device path as inputs and the S-matrix parity tests compare like with like.
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
