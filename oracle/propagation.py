"""Multi-interval R-matrix propagation and wave functions (oracle).

Test infrastructure only: the CPU statement that the device path is checked
against.  Restates, for eta = 0:

* the per-interval assembly of src/rmatrix_hp.F90:251-293 (interval radii,
  interval T+L, the closed-channel Bloch term blo0 / blo2 - blo1, the
  potential rows of the interval);
* the 2 nch right-hand-side solve of solve_rmatrix_hp_4
  (src/rmatrix_hp.F90:590-632): R00 = q1^T C^-1 q1, R01 = q1^T C^-1 q2,
  R11 = q2^T C^-1 q2 per interval, ZGESV here;
* the propagation chain src/rmatrix_hp.F90:335-345
  crma0_1 = R11_1 / a_1,
  crma0_s = (R11_s - R01_s^T (R00_s + a_{s-1} crma0_{s-1})^-1 R01_s) / a_s;
* S from crma0_ns (src/rmatrix_hp.F90:348-368);
* the wave functions src/rmatrix_hp.F90:370-431: amplitudes cfp at the
  interval boundaries (backward recursion), internal wave functions
  cf_s = C_s^-1 (q2_s cfp_s - q1_s cfp_{s-1}) / sqrt(a0 w), bound-state
  amplitudes.  C_s^-1 is what Descouvemont's original code multiplies by
  (rmat_pierre/rmatrix.f:223 inverts C in place, :296 uses it); the
  HPRMAT driver keeps C un-inverted at that point (src/rmatrix_hp.F90:415)
  and so does not reproduce its own golden wave functions
  (rmat_pierre/Ex2/wf2.txt).  This restatement follows the golden output;
* wf_print (src/rmatrix_hp.F90:678-734) with the natural cubic spline of
  spline / splint (:786-841): the internal wave function on a uniform mesh
  inside the channel radius, the asymptotic form outside.

Pinning (tests/test_oracle_propagation.py): the golden phase shifts and
mixing of rmat_pierre/Ex2 for ns = 2 and 3, the golden internal wave
functions of rmat_pierre/Ex2/wf2.txt (ns = 3), independence of S from the
split of (0, a) into intervals (ns = 1 with many points against ns = 3),
the ns = 1 case against the one-interval oracle, continuity of the internal
and asymptotic wave functions at the channel radius, and the spline against
scipy's natural cubic spline.
"""
import math

import numpy as np
from scipy import special

from . import channels
from .lagrange import rmat_ini
from .solve import zgesv


def interval_radii(mesh, s):
    """Mesh radii of interval s (0-based), src/rmatrix_hp.F90:258."""
    return s * mesh['rmax0'] + mesh['xle'] * mesh['rmax0']


def bloch_closed(mesh, s):
    """Closed-channel Bloch matrix of interval s (multiplied by xc1 and
    subtracted), src/rmatrix_hp.F90:264-271."""
    rmax0 = mesh['rmax0']
    rmax1, rmax2 = s * rmax0, (s + 1) * rmax0
    if s == 0:
        return mesh['blo0'] / rmax2
    return mesh['blo2'] / rmax2 - mesh['blo1'] / rmax1


def assemble_interval(mesh, s, lval, qk, cpot, xc1):
    """Dense C of interval s (0-based), src/rmatrix_hp.F90:251-293.
    cpot: (nr*ns, nch, nch), rows s*nr .. s*nr+nr-1 belong to interval s."""
    nr = mesh['nr']
    nch = len(lval)
    r = interval_radii(mesh, s)
    tc = mesh['tc'][s]
    bl = bloch_closed(mesh, s)
    ch = np.zeros((nch * nr, nch * nr), dtype=complex)
    for i1 in range(nch):
        b1 = slice(i1 * nr, (i1 + 1) * nr)
        blk = tc.astype(complex)
        d = lval[i1] * (lval[i1] + 1) / r ** 2
        if qk[i1] > 0:
            d = d - qk[i1] ** 2
        elif qk[i1] < 0:
            d = d + qk[i1] ** 2
        blk = blk + np.diag(d)
        if qk[i1] < 0:
            blk = blk - xc1[i1] * bl
        ch[b1, b1] = blk
        for i2 in range(nch):
            b2 = slice(i2 * nr, (i2 + 1) * nr)
            ch[b1, b2] += np.diag(cpot[s * nr:(s + 1) * nr, i1, i2])
    return ch


def boundary_rhs(mesh, s, nch):
    """[q1 blocks | q2 blocks] (nch nr x 2 nch), src/rmatrix_hp.F90:608-614."""
    nr = mesh['nr']
    b = np.zeros((nch * nr, 2 * nch), dtype=complex)
    for c in range(nch):
        b[c * nr:(c + 1) * nr, c] = mesh['q1'][s]
        b[c * nr:(c + 1) * nr, nch + c] = mesh['q2'][s]
    return b


def r_blocks(x, mesh, s, nch):
    """(R00, R01, R11) of interval s from X = C^-1 [q1 | q2],
    src/rmatrix_hp.F90:619-629."""
    nr = mesh['nr']
    q1, q2 = mesh['q1'][s], mesh['q2'][s]
    r00 = np.zeros((nch, nch), dtype=complex)
    r01 = np.zeros((nch, nch), dtype=complex)
    r11 = np.zeros((nch, nch), dtype=complex)
    for a in range(nch):
        rows = slice(a * nr, (a + 1) * nr)
        r00[a] = q1 @ x[rows, :nch]
        r01[a] = q1 @ x[rows, nch:]
        r11[a] = q2 @ x[rows, nch:]
    return r00, r01, r11


def propagate(crma, rmax0):
    """crma0 per interval (src/rmatrix_hp.F90:335-345); crma: list of
    (R00, R01, R11)."""
    out = []
    for s, (r00, r01, r11) in enumerate(crma):
        rmax1, rmax2 = s * rmax0, (s + 1) * rmax0
        if s == 0:
            out.append(r11 / rmax2)
            continue
        cz = r00 + out[-1] * rmax1
        czi, info = zgesv(cz, np.eye(len(cz), dtype=complex))
        if info != 0:
            raise RuntimeError('singular propagation matrix')
        out.append((r11 - r01.T @ czi @ r01) / rmax2)
    return out


def s_from_crma0(crma0, qk, npo, co, cop):
    """S = Z_O^-1 Z_I from the propagated R-matrix, src/rmatrix_hp.F90:348-368."""
    nopen = len(npo)
    c2 = np.zeros((nopen, nopen), dtype=complex)
    for m1 in range(nopen):
        for m2 in range(nopen):
            c2[m1, m2] = crma0[npo[m1], npo[m2]] * math.sqrt(qk[npo[m1]] / qk[npo[m2]])
    cz = -c2 * cop[None, :]
    cu = -c2 * np.conj(cop)[None, :]
    cz[np.diag_indices(nopen)] += co
    cu[np.diag_indices(nopen)] += np.conj(co)
    s, info = zgesv(cz, cu)
    if info != 0:
        raise RuntimeError('Error in ZGESV for collision matrix')
    return s


def rmatrix_ns(lval, qk, rmax, nr, ns, cpot, nvc=None):
    """rmatrix (src/rmatrix_hp.F90:175-431) for eta = 0 and any ns.

    cpot: (nr*ns, nch, nch).  nvc: entrance channels (0-based indices among
    the open channels) for which the wave functions are built.  Returns a
    dict with S (nopen x nopen), cu (nch x nch: S, then the bound-state
    amplitudes in rows nopen.. of the entrance columns), nopen, crma (per
    interval (R00, R01, R11)), crma0, X (per interval C^-1 [q1 | q2]), cf
    ((nr ns, nch, len(nvc)) or None), mesh.
    """
    lval = np.asarray(lval)
    qk = np.asarray(qk, dtype=float)
    nch = len(lval)
    mesh = rmat_ini(nr, ns, rmax)
    rmax0 = mesh['rmax0']
    npo, co, cop, xc1 = channels.channel_functions(lval, qk, rmax)
    ncl = [i for i in range(nch) if qk[i] <= 0]
    xc2 = np.zeros(nch)
    for i in ncl:
        xc2[i] = channels.closed_channel(int(lval[i]), -qk[i], rmax)[1]
    cmats, xs, crma = [], [], []
    for s in range(ns):
        c = assemble_interval(mesh, s, lval, qk, cpot, xc1)
        x, info = zgesv(c, boundary_rhs(mesh, s, nch))
        if info != 0:
            raise RuntimeError('Error in solve_rmatrix_hp_4')
        cmats.append(c)
        xs.append(x)
        crma.append(r_blocks(x, mesh, s, nch))
    crma0 = propagate(crma, rmax0)
    nopen = len(npo)
    cu = np.zeros((nch, nch), dtype=complex)
    s_mat = s_from_crma0(crma0[-1], qk, npo, co, cop) if nopen else np.zeros((0, 0), complex)
    cu[:nopen, :nopen] = s_mat
    cf = None
    if nvc is not None:
        cf = np.zeros((nr * ns, nch, len(nvc)), dtype=complex)
        for ic, nvex in enumerate(nvc):
            cf[:, :, ic], amps = wave_functions(xs, crma, crma0, mesh, qk, npo, cop, s_mat,
                                                nvex, ncl, xc2)
            cu[nopen:nopen + len(ncl), ic] = amps
    return dict(S=s_mat, cu=cu, nopen=nopen, crma=crma, crma0=crma0, X=xs, cf=cf, mesh=mesh,
                npo=npo, co=co, cop=cop, xc1=xc1, xc2=xc2)


def boundary_amplitudes(crma, crma0, rmax0, qk, npo, cop, s_mat, nvex, rmax, nch):
    """cfp (ns + 1, nch): amplitudes at the interval boundaries for entrance
    channel nvex, cfp[0] = 0 (src/rmatrix_hp.F90:380-396)."""
    ns = len(crma)
    cfp = np.zeros((ns + 1, nch), dtype=complex)
    i2 = npo[nvex]
    for m1, i1 in enumerate(npo):
        v = -s_mat[m1, nvex] * cop[m1] * math.sqrt(qk[i2] / qk[i1])
        if i1 == i2:
            v += np.conj(cop[m1])
        cfp[ns, i1] = v / rmax
    for s in range(ns - 1, 0, -1):      # reference is = ns-1 .. 1
        rin = s * rmax0
        cz = crma[s][0] + rin * crma0[s - 1]
        z, info = zgesv(cz, crma[s][1])
        if info != 0:
            raise RuntimeError('singular propagation matrix')
        cfp[s] = z @ cfp[s + 1]
    return cfp


def wave_functions(xs, crma, crma0, mesh, qk, npo, cop, s_mat, nvex, ncl, xc2):
    """Internal wave functions (nr ns, nch) at the mesh points and bound-state
    amplitudes of entrance channel nvex (src/rmatrix_hp.F90:398-430)."""
    nr, ns = mesh['nr'], len(xs)
    nch = xs[0].shape[1] // 2
    rmax0 = mesh['rmax0']
    cfp = boundary_amplitudes(crma, crma0, rmax0, qk, npo, cop, s_mat, nvex, rmax0 * ns, nch)
    cf = np.zeros((nr * ns, nch), dtype=complex)
    for s in range(ns):
        # C_s^-1 (q2_s cfp_s - q1_s cfp_{s-1}) = X2_s cfp_s - X1_s cfp_{s-1}
        y = xs[s][:, nch:] @ cfp[s + 1] - xs[s][:, :nch] @ cfp[s]
        for c in range(nch):
            cf[s * nr:(s + 1) * nr, c] = y[c * nr:(c + 1) * nr]
    amps = np.array([np.sum(cf[(ns - 1) * nr:, i1] * mesh['q2'][ns - 1]) / xc2[i1] for i1 in ncl],
                    dtype=complex)
    norm = np.tile(1.0 / np.sqrt(rmax0 * mesh['wle']), ns)
    return cf * norm[:, None], amps


def spline_coeffs(x, y):
    """Second derivatives of the natural cubic spline through (x, y)
    (spline with yp1 = ypn = 1e30, src/rmatrix_hp.F90:786-821)."""
    n = len(x)
    y2 = np.zeros(n)
    u = np.zeros(n)
    for i in range(1, n - 1):
        sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1])
        p = sig * y2[i - 1] + 2
        y2[i] = (sig - 1) / p
        u[i] = (6 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]))
                / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p
    y2[n - 1] = 0.0
    for k in range(n - 2, -1, -1):
        y2[k] = y2[k] * y2[k + 1] + u[k]
    return y2


def splint(x, y, y2, t):
    """Cubic-spline value at t (splint, src/rmatrix_hp.F90:823-841)."""
    lo, hi = 0, len(x) - 1
    while hi - lo > 1:
        k = (hi + lo) // 2
        if x[k] > t:
            hi = k
        else:
            lo = k
    h = x[hi] - x[lo]
    a = (x[hi] - t) / h
    b = (t - x[lo]) / h
    return a * y[lo] + b * y[hi] + ((a ** 3 - a) * y2[lo] + (b ** 3 - b) * y2[hi]) * h * h / 6


def external_wave(lval, qk, cu, nopen, iv, nom, r):
    """Asymptotic wave function of channel iv (0-based) at r > a for entrance
    channel nom (0-based channel index, open): src/rmatrix_hp.F90:711-729."""
    nop = sum(1 for w in range(iv + 1) if qk[w] > 0) - 1
    ncl = nopen + sum(1 for w in range(iv + 1) if qk[w] < 0) - 1
    l = int(lval[iv])
    if qk[iv] > 0:
        rho = qk[iv] * r
        f = rho * special.spherical_jn(l, rho)
        g = -rho * special.spherical_yn(l, rho)
        co = complex(g, f) * math.sqrt(qk[nom] / qk[iv])
        v = -cu[nop, nom] * co
        if iv == nom:
            v += np.conj(co)
        return v
    w = channels.closed_channel(l, -qk[iv], r)[1]
    return cu[ncl, nom] * w


def wf_print(lval, qk, rmax, cu, nopen, cf, zrma, iv, nom, npoin, h):
    """Wave function of channel iv, entrance column nom, at r = h, 2h, ...
    (npoin points): spline inside rmax, asymptotic form outside
    (src/rmatrix_hp.F90:678-734).  cf: (nr ns, nch, nc)."""
    yr = spline_coeffs(zrma, cf[:, iv, nom].real)
    yi = spline_coeffs(zrma, cf[:, iv, nom].imag)
    out = np.zeros(npoin, dtype=complex)
    for i in range(1, npoin + 1):
        r = i * h
        if r <= rmax:
            out[i - 1] = complex(splint(zrma, cf[:, iv, nom].real, yr, r),
                                 splint(zrma, cf[:, iv, nom].imag, yi, r))
        else:
            out[i - 1] = external_wave(lval, qk, cu, nopen, iv, nom, r)
    return out
