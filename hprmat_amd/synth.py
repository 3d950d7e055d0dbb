"""Seeded synthetic coupled-channel systems of the benchmark workloads.

Units hbar^2/2mu = 1, Sommerfeld parameter eta = 0, one interval of radius a,
Lagrange-Legendre mesh of nr points:

* channel c = 0..nch-1: l_c = c mod 4, threshold eps_c ~ U[0, emax];
  k_c^2 = E - eps_c, qk_c = +sqrt when open and -sqrt(eps_c - E) when closed
  (the reference's closed-channel sign convention);
* coupling of the unordered pair c <= c' (c-major order): one row
  u = rng.random((npairs, 4)) after the thresholds gives V0 = -50 + 20 u0,
  W0 = -10 + 8 u1, R = 3 + 2 u2, d = 0.5 + 0.2 u3 and
  V_cc'(r) = s (V0 + i W0) / (1 + exp((r - R)/d)), s = 1 (c = c') or 0.2;
  ``w0_zero`` keeps the draws and sets W0 = 0 (real potential);
* ``mode='term'``: four coupling terms t = min(|c - c'|, 3) with parameters
  theta[t] = (V0, W0, R, d) from rng.random((4, 4)), s_t = 1 (t = 0) or 0.2.

RHS: column c of B is q2 in the rows of block c (nrhs = nch), as
src/rmat_solvers.F90:120-124 builds it.
"""
import numpy as np

from . import rmatrix


def draw(nch, emax, seed, mode='pair'):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(0.0, emax, nch)
    u = rng.random((nch * (nch + 1) // 2, 4)) if mode == 'pair' else rng.random((4, 4))
    return eps, u


def theta_from_u(u):
    """(V0, W0, R, d) columns from uniforms."""
    u = np.asarray(u)
    return np.stack([-50.0 + 20.0 * u[..., 0], -10.0 + 8.0 * u[..., 1],
                     3.0 + 2.0 * u[..., 2], 0.5 + 0.2 * u[..., 3]], axis=-1)


def wave_numbers(eps, energy):
    k2 = energy - np.asarray(eps)
    return np.where(k2 > 0, np.sqrt(np.abs(k2)), -np.sqrt(np.abs(k2)))


def woods_saxon_potential(r, nch, theta, mode='pair', w0_zero=False):
    """cpot[i, c, c'] (nr x nch x nch complex) for the pair or term recipe."""
    r = np.asarray(r, dtype=np.float64)
    th = np.array(theta, dtype=np.float64, copy=True)
    if w0_zero:
        th[..., 1] = 0.0
    cpot = np.zeros((len(r), nch, nch), dtype=np.complex128)
    if mode == 'pair':
        iu, ju = np.triu_indices(nch)
        s = np.where(iu == ju, 1.0, 0.2)
        f = 1.0 / (1.0 + np.exp((r[:, None] - th[None, :, 2]) / th[None, :, 3]))
        v = (s * (th[:, 0] + 1j * th[:, 1]))[None, :] * f
        cpot[:, iu, ju] = v
        cpot[:, ju, iu] = v
    else:
        c = np.arange(nch)
        t = np.minimum(np.abs(c[:, None] - c[None, :]), 3)
        s = np.where(t == 0, 1.0, 0.2)
        for k in range(4):
            f = 1.0 / (1.0 + np.exp((r - th[k, 2]) / th[k, 3]))
            mask = (t == k)
            cpot[:, mask] = (s[mask] * (th[k, 0] + 1j * th[k, 1]))[None, :] * f[:, None]
    return cpot


class System:
    """One synthetic system: mesh, channel data and potential (host arrays)."""

    def __init__(self, nch, nr, a, emax, energy, seed, mode='pair', w0_zero=False, theta=None):
        self.nch, self.nr, self.a, self.energy, self.seed = nch, nr, a, energy, seed
        self.mesh = rmatrix.rmat_ini(nr, a)
        self.eps, u = draw(nch, emax, seed, mode)
        self.lval = np.arange(nch) % 4
        self.qk = wave_numbers(self.eps, energy)
        self.mode = mode
        self.theta = theta_from_u(u) if theta is None else np.asarray(theta, dtype=np.float64)
        self.cpot = woods_saxon_potential(self.mesh.r, nch, self.theta, mode, w0_zero)

    @property
    def n(self):
        return self.nch * self.nr

    def at_energy(self, energy):
        """Same potential at another energy (new wave numbers)."""
        s = object.__new__(System)
        s.__dict__.update(self.__dict__)
        s.energy = energy
        s.qk = wave_numbers(self.eps, energy)
        return s


def config(name, seed=0):
    """Systems of the named workloads (sizes of the benchmark configurations)."""
    if name == 'small':      # N = 240, CPU-runnable
        return System(4, 60, 15.0, 3.0, 2.0, seed)
    if name == 'single':     # N = 25600, nrhs = 128
        return System(128, 200, 15.0, 8.0, 5.0, seed)
    if name == 'memory':     # N = 51200, nrhs = 256
        return System(256, 200, 15.0, 3.0, 6.0, seed)
    if name == 'grad':       # N = 6400, 16 coupling parameters
        return System(32, 200, 15.0, 3.0, 4.0, seed, mode='term')
    raise KeyError(name)
