"""GPU parity of the S-matrix gradient path (hprmat_amd.adjoint): the coupling
potential, the adjoint contraction kernel, the differentiable R -> S, and the
full gradient against the FP64 CPU adjoint oracle and central differences."""
import numpy as np
import pytest
import torch

from hprmat_amd import adjoint, rmatrix, synth
from oracle import adjoint as oadj, synth as osynth

pytestmark = pytest.mark.gpu


def test_coupling_potential_matches_term_recipe():
    s = synth.System(9, 30, 15.0, 3.0, 4.0, 2, mode='term')
    th = torch.as_tensor(s.theta, device='cuda')
    r = torch.as_tensor(s.mesh.r, device='cuda')
    cp = adjoint.coupling_potential(th, r, 9).cpu().numpy()
    assert np.abs(cp - s.cpot).max() <= 1e-13 * np.abs(s.cpot).max()


def test_adjoint_contract_kernel_matches_einsum():
    g = torch.Generator(device='cuda').manual_seed(5)
    nch, nr, nrhs = 7, 23, 5
    n = nch * nr
    m = torch.complex(torch.randn(nrhs, n, generator=g, device='cuda', dtype=torch.float64),
                      torch.randn(nrhs, n, generator=g, device='cuda', dtype=torch.float64))
    x = torch.complex(torch.randn(nrhs, n, generator=g, device='cuda', dtype=torch.float64),
                      torch.randn(nrhs, n, generator=g, device='cuda', dtype=torch.float64))
    k = adjoint.adjoint_contract(m, x, nch, nr).cpu().numpy()
    mm = m.cpu().numpy().reshape(nrhs, nch, nr)
    xx = x.cpu().numpy().reshape(nrhs, nch, nr)
    ref = np.einsum('kci,kdi->icd', mm, xx)
    assert np.abs(k - ref).max() <= 1e-13 * np.abs(ref).max()


def test_torch_s_matrix_matches_device_r_to_s():
    s = synth.System(6, 40, 15.0, 3.0, 2.5, 4, mode='term')
    dm = rmatrix.DeviceMesh(s.mesh)
    pt = adjoint.EnergyPoint(s, 2.5, dm)
    cpot = adjoint.coupling_potential(torch.as_tensor(s.theta, device='cuda'), dm.r, 6)
    opts = adjoint.SolveOptions(method='fp64', refinement='ir', max_steps=3)
    x = adjoint._Solve.apply(cpot, pt, opts)
    rm = adjoint.r_matrix(x, dm.q2, 6, 40)
    st = adjoint.s_matrix(rm.cpu(), pt.a, pt.qk, pt.open_idx, pt.co, pt.cop).numpy()
    r_dev, s_dev = rmatrix.r_to_s(dm, x, 6, pt.qk, pt.open_idx, pt.co.numpy(), pt.cop.numpy())
    assert np.abs(rm.cpu().numpy() - r_dev.cpu().numpy()).max() <= 1e-13 * np.abs(rm.cpu().numpy()).max()
    assert np.abs(st - s_dev.cpu().numpy()).max() <= 1e-12


def _small_fit(method, refinement, energies=(1.5, 3.0, 6.0), nch=4, nr=40, seed=1):
    opts = adjoint.SolveOptions(method=method, refinement=refinement, max_steps=16)
    fp = adjoint.FitProblem(nch=nch, nr=nr, energies=energies, seed=seed, opts=opts)
    prob = oadj.problem(nch, nr, 15.0, 3.0, seed, energies)
    th_true = fp.theta_true.cpu().numpy()
    tg = oadj.s_matrices(prob, th_true)
    return fp, prob, th_true, tg


@pytest.mark.parametrize('method,refinement', [('fp64', 'ir'), ('fp32', 'fgmres')])
def test_gradient_matches_cpu_adjoint_and_differences(method, refinement):
    fp, prob, th_true, tg = _small_fit(method, refinement)
    # targets from the device path agree with the oracle's
    for a, b in zip(fp.targets, tg):
        assert np.abs(a.cpu().numpy() - b).max() <= 1e-10
    loss, g = fp.loss_and_grad()
    th = fp.theta0.cpu().numpy()
    loss_ref, g_ref = oadj.loss_and_grad(prob, th, tg)
    assert abs(loss - loss_ref) <= 1e-10 * loss_ref
    g = g.cpu().numpy()
    assert np.abs(g - g_ref).max() <= 1e-8 * np.abs(g_ref).max(), (g, g_ref)
    # central differences of the FP64 oracle loss (h = 1e-6 relative)
    fd = np.zeros((4, 4))
    for t in range(4):
        for p in range(4):
            h = 1e-6 * max(1.0, abs(th[t, p]))
            tp, tm = th.copy(), th.copy()
            tp[t, p] += h
            tm[t, p] -= h
            fd[t, p] = (oadj.loss(prob, tp, tg) - oadj.loss(prob, tm, tg)) / (2 * h)
    assert np.abs(g - fd).max() <= 1e-6 * np.abs(fd).max()


def test_gradient_differences_at_larger_size():
    # central differences of the FP64 oracle loss at N = 640 (nch = 8 x nr =
    # 80), four energies, mixed-precision forward and adjoint solves to 1e-12
    fp, prob, th_true, tg = _small_fit('fp32', 'fgmres', energies=(1.2, 2.9, 5.1, 7.6), nch=8,
                                       nr=80, seed=3)
    loss, g = fp.loss_and_grad()
    th = fp.theta0.cpu().numpy()
    g = g.cpu().numpy()
    loss_ref, g_ref = oadj.loss_and_grad(prob, th, tg)
    assert np.abs(g - g_ref).max() <= 1e-8 * np.abs(g_ref).max()
    fd = np.zeros((4, 4))
    for t in range(4):
        for p in range(4):
            h = 1e-6 * max(1.0, abs(th[t, p]))
            tp, tm = th.copy(), th.copy()
            tp[t, p] += h
            tm[t, p] -= h
            fd[t, p] = (oadj.loss(prob, tp, tg) - oadj.loss(prob, tm, tg)) / (2 * h)
    assert np.abs(g - fd).max() <= 1e-6 * np.abs(fd).max()


def test_floor_acceptance_uses_the_fp64_floor():
    # a solve that stops above the gate is accepted only within a few times
    # u ||(|C| |X|)||_F / ||B||_F, and the floor matches a host evaluation
    s = synth.System(6, 40, 15.0, 3.0, 2.5, 4, mode='term')
    dm = rmatrix.DeviceMesh(s.mesh)
    pt = adjoint.EnergyPoint(s, 2.5, dm)
    cpot = adjoint.coupling_potential(torch.as_tensor(s.theta, device='cuda'), dm.r, 6)
    a_cm = rmatrix.assemble(dm, pt.lval, pt.qk, pt.xc1, cpot)
    b = rmatrix.rhs_colmajor(dm, 6)
    x = b.clone()
    solver_lu = __import__('hprmat_amd.solver', fromlist=['LU']).LU().factor(a_cm, 'fp64')
    st = solver_lu.solve_(x, 2, 1e-12, 'rel_fro', True)
    fl = adjoint.fp64_floor(a_cm, x, b)
    a = a_cm.cpu().numpy().T
    xh, bh = x.cpu().numpy().T, b.cpu().numpy().T
    ref = 2.0 ** -53 * np.linalg.norm(np.abs(a) @ np.abs(xh)) / np.linalg.norm(bh)
    assert abs(fl - ref) <= 1e-12 * ref
    opts = adjoint.SolveOptions()
    fake = dict(st, converged=False, res_rel=3.0 * fl)
    assert adjoint._accept(fake, a_cm, x, b, opts, 'test', 2.5)['at_floor'] == fl
    with pytest.raises(RuntimeError):
        adjoint._accept(dict(fake, res_rel=10.0 * fl), a_cm, x, b, opts, 'test', 2.5)
    with pytest.raises(RuntimeError):
        adjoint._accept(dict(fake, diverged=True), a_cm, x, b, opts, 'test', 2.5)


def test_gradient_full_size_against_cpu_adjoint():
    # the gradient workload's system (nch = 32 x nr = 200, N = 6400), four of
    # its energies (both ends of the grid among them), mixed-precision forward
    # and adjoint solves to 1e-12 (or the FP64 floor)
    energies = (1.0, 3.2580645161290325, 6.5, 8.0)
    opts = adjoint.SolveOptions(method='fp32', refinement='fgmres', max_steps=20)
    fp = adjoint.FitProblem(nch=32, nr=200, energies=energies, seed=0, opts=opts)
    loss, g = fp.loss_and_grad()
    prob = oadj.problem(32, 200, 15.0, 3.0, 0, energies)
    th = fp.theta0.cpu().numpy()
    tg = [t.cpu().numpy() for t in fp.targets]
    loss_ref, g_ref = oadj.loss_and_grad(prob, th, tg)
    assert abs(loss - loss_ref) <= 1e-8 * loss_ref
    g = g.cpu().numpy()
    assert np.abs(g - g_ref).max() <= 1e-8 * np.abs(g_ref).max()
