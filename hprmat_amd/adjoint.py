"""Gradients of S-matrix losses with respect to coupling-potential parameters.

Workload (energy-scan fit): for energies E_1..E_ne, the S-matrix S(E; theta)
of a Lagrange-mesh system whose local coupling potential is
    V_cc'(r) = s_t (V0_t + i W0_t) / (1 + exp((r - R_t) / d_t)),
    t = min(|c - c'|, 3),  s_0 = 1,  s_1 = s_2 = s_3 = 0.2,
with theta = (V0, W0, R, d) for each of the 4 terms (16 real parameters), and
the loss  L = sum_E sum_cc' |S_cc'(E) - S_target_cc'(E)|^2.

Forward, per energy: the matrix C(E) = A0(E) + P(V) (P puts V_cc'(r_i) on
the (c, i), (c', i) entries) is assembled on the device and X = C^-1 B is
solved by libhprmat_hip (FP32 LU + FGMRES refinement to 1e-12 by default);
R = q2^T X and S follow with differentiable torch operations in FP64 on the
device.

Backward (``_Solve.backward``): with G the torch gradient of X (the
conjugate-Wirtinger convention, G = 2 dL/dconj(X)), dL = Re <G, dX> and
dX = -C^-1 dC X give
    dL = -Re sum_cc'i dV_cc'(r_i) K_cc'(i),   K_cc'(i) = sum_k M_(c,i)k X_(c',i)k,
    M = C^-T conj(G) = C^-1 conj(G)      (C is complex symmetric),
so the adjoint solve reuses the forward LU factors (no refactorization) and
the gradient of V is -conj(K), formed by the library's adjoint contraction
kernel.  torch then carries it to theta through ``coupling_potential``.
"""
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, rmatrix, solver, synth

TERM_SCALE = (1.0, 0.2, 0.2, 0.2)


def term_index(nch):
    """t(c, c') = min(|c - c'|, 3) as an (nch, nch) long tensor."""
    c = torch.arange(nch)
    return torch.clamp((c[:, None] - c[None, :]).abs(), max=3)


def coupling_potential(theta, r, nch):
    """cpot[i, c, c'] (nr x nch x nch complex128) from theta (4 x 4 real:
    rows t = 0..3, columns V0, W0, R, d); differentiable in theta."""
    v0, w0, rr, d = theta.unbind(-1)
    s = torch.tensor(TERM_SCALE, dtype=theta.dtype, device=theta.device)
    f = 1.0 / (1.0 + torch.exp((r[None, :] - rr[:, None]) / d[:, None]))       # (4, nr)
    v = torch.complex(s * v0, s * w0)[:, None] * f                              # (4, nr)
    t = term_index(nch).to(theta.device)
    return v[t].permute(2, 0, 1).contiguous()                                   # (nr, nch, nch)


@dataclass
class SolveOptions:
    """Forward/adjoint solve settings.  tol is the gate on ||B - C X||_F /
    ||B||_F.  A solve that stops above it is still accepted when its residual
    is within ``floor_factor`` of the FP64 floor u ||(|C| |X|)||_F / ||B||_F
    (u = 2^-53): no FP64 X can do better (the N = 6400 fit system reaches
    that floor, about 2e-12, at its lowest energies)."""
    method: str = 'fp32'          # factorization: fp64, fp32, fp32_3m, bf16x3
    refinement: str = 'fgmres'    # ir or fgmres
    max_steps: int = 24
    tol: float = 1e-12
    floor_factor: float = 4.0


def fp64_floor(a_cm, x_cm, b_cm):
    """u ||(|C| |X|)||_F / ||B||_F for column-major storages a_cm (n x n) and
    x_cm, b_cm (nrhs x n) on the device."""
    ax = x_cm.abs() @ a_cm.abs()          # row k: (|C| |x_k|)^T, column-major storages
    return float(2.0 ** -53 * torch.linalg.vector_norm(ax) / torch.linalg.vector_norm(b_cm))


def _accept(st, a_cm, x_cm, b_cm, opts, what, energy):
    """Stats of a converged solve, or of one that ended at the FP64 floor
    (then st['at_floor'] is set); raises otherwise."""
    if st['info'] == 0 and st['converged']:
        return st
    if st['info'] == 0 and not st['diverged']:
        fl = fp64_floor(a_cm, x_cm, b_cm)
        if st['res_rel'] <= opts.floor_factor * fl:
            st['at_floor'] = fl
            return st
    raise RuntimeError('%s solve at E = %g did not converge: %r' % (what, energy, st))


class EnergyPoint:
    """Energy-dependent data of one system at one energy (device resident)."""

    def __init__(self, sysm, energy, dmesh):
        self.nch, self.nr, self.a = sysm.nch, sysm.nr, sysm.a
        self.lval = sysm.lval
        self.energy = float(energy)
        self.qk = synth.wave_numbers(sysm.eps, energy)
        self.open_idx, co, cop, self.xc1 = rmatrix.channel_data(self.lval, self.qk, self.a)
        self.co = torch.as_tensor(co)
        self.cop = torch.as_tensor(cop)
        self.dmesh = dmesh


class _Solve(torch.autograd.Function):
    """X = C(cpot)^-1 B on the device; X is returned as its column-major
    storage (nch x N tensor: row k holds column k of X)."""

    @staticmethod
    def forward(ctx, cpot, pt, opts):
        a_cm = rmatrix.assemble(pt.dmesh, pt.lval, pt.qk, pt.xc1, cpot.detach())
        x = rmatrix.rhs_colmajor(pt.dmesh, pt.nch)
        b = x.clone()
        lu = solver.LU().factor(a_cm, opts.method, equilibrate=True)
        st = lu.solve_(x, opts.max_steps, opts.tol, 'rel_fro', True, refinement=opts.refinement)
        _accept(st, a_cm, x, b, opts, 'forward', pt.energy)
        ctx.lu, ctx.pt, ctx.opts, ctx.a_cm = lu, pt, opts, a_cm
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, gx):
        (x,) = ctx.saved_tensors
        pt, opts, lu = ctx.pt, ctx.opts, ctx.lu
        # conj_physical: a lazy conj() view would hand unconjugated memory to
        # the library through data_ptr()
        m = torch.conj_physical(gx).contiguous()
        b = m.clone()
        st = lu.solve_(m, opts.max_steps, opts.tol, 'rel_fro', True, refinement=opts.refinement)
        _accept(st, ctx.a_cm, m, b, opts, 'adjoint', pt.energy)
        k = adjoint_contract(m, x, pt.nch, pt.nr)
        ctx.lu = ctx.a_cm = None
        return -k.conj(), None, None


def adjoint_contract(m, x, nch, nr):
    """K[i, c, c'] = sum_k M[(c, i), k] X[(c', i), k] for column-major storages
    m, x (nrhs x N tensors), on the device (libhprmat_hip kernel)."""
    import ctypes
    lib = _lib.load()
    n = nch * nr
    if m.shape != x.shape or m.shape[1] != n:
        raise ValueError('m and x must be (nrhs, nch*nr) storages')
    k = torch.empty((nr, nch, nch), dtype=torch.complex128, device=x.device)
    rc = lib.hprmat_hip_adjoint_contract(nch, nr, m.shape[0], ctypes.c_void_p(m.data_ptr()),
                                         ctypes.c_int64(n), ctypes.c_void_p(x.data_ptr()),
                                         ctypes.c_int64(n), ctypes.c_void_p(k.data_ptr()),
                                         solver._stream_ptr())
    _lib.check(rc, 'hprmat_hip_adjoint_contract')
    return k


def r_matrix(x, q2, nch, nr, normfac=1.0):
    """R[i, j] = normfac sum_r q2[r] X[(i, r), j] from the storage x (nch x N)."""
    return normfac * (x.view(nch, nch, nr) @ q2.to(x.dtype)).transpose(0, 1)


def s_matrix(rm, a, qk, open_idx, co, cop):
    """S = Z_O^-1 Z_I over the open channels (src/rmatrix_hp.F90:336-368),
    differentiable (complex128, on the device of rm)."""
    dev = rm.device
    o = torch.as_tensor(np.asarray(open_idx, dtype=np.int64), device=dev)
    k = torch.as_tensor(np.asarray(qk, dtype=np.float64)[np.asarray(open_idx, dtype=np.int64)],
                        device=dev)
    co = co.to(dev)
    cop = cop.to(dev)
    c2 = rm[o][:, o] / a * torch.sqrt(k[:, None] / k[None, :]).to(rm.dtype)
    zo = torch.diag(co) - c2 * cop[None, :]
    zi = torch.diag(co.conj()) - c2 * cop.conj()[None, :]
    return torch.linalg.solve(zo, zi)


def s_of_energy(cpot, pt, opts, q2):
    """S(E) on the device (R and the nopen x nopen solve in FP64)."""
    x = _Solve.apply(cpot, pt, opts)
    rm = r_matrix(x, q2, pt.nch, pt.nr)
    return s_matrix(rm, pt.a, pt.qk, pt.open_idx, pt.co, pt.cop)


def scan_loss(theta, points, targets, r, q2, opts=None):
    """L(theta) = sum_E sum |S(E; theta) - S_target(E)|^2 (theta on the device)."""
    opts = opts or SolveOptions()
    cpot = coupling_potential(theta, r, points[0].nch)
    loss = torch.zeros((), dtype=torch.float64, device=theta.device)
    for pt, tgt in zip(points, targets):
        s = s_of_energy(cpot, pt, opts, q2)
        loss = loss + (s - tgt).abs().pow(2).sum()
    return loss


class FitProblem:
    """The gradient workload: nch x nr system in 'term' mode, energies on a
    uniform grid, target S-matrices from the seeded parameters theta_true,
    evaluation at theta = (1 + perturb) theta_true."""

    def __init__(self, nch=32, nr=200, a=15.0, emax=3.0, energies=None, seed=0, perturb=0.05,
                 opts=None):
        self.sysm = synth.System(nch, nr, a, emax, 4.0, seed, mode='term')
        self.energies = np.linspace(1.0, 8.0, 32) if energies is None else np.asarray(energies)
        self.opts = opts or SolveOptions()
        self.dmesh = rmatrix.DeviceMesh(self.sysm.mesh)
        self.r = self.dmesh.r
        self.q2 = self.dmesh.q2
        self.points = [EnergyPoint(self.sysm, e, self.dmesh) for e in self.energies]
        self.theta_true = torch.as_tensor(self.sysm.theta, device='cuda')
        self.theta0 = self.theta_true * (1.0 + perturb)
        with torch.no_grad():
            cpot = coupling_potential(self.theta_true, self.r, nch)
            self.targets = [s_of_energy(cpot, pt, self.opts, self.q2).detach()
                            for pt in self.points]

    def loss(self, theta, opts=None):
        return scan_loss(theta, self.points, self.targets, self.r, self.q2, opts or self.opts)

    def loss_and_grad(self, theta=None, opts=None):
        th = (self.theta0 if theta is None else theta).detach().clone().requires_grad_(True)
        loss = self.loss(th, opts)
        loss.backward()
        return float(loss.detach()), th.grad.detach()
