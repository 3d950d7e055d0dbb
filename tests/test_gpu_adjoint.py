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
        assert np.abs(a.numpy() - b).max() <= 1e-10
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


def test_gradient_full_size_against_cpu_adjoint():
    # the gradient workload's system (nch = 32 x nr = 200, N = 6400), two of
    # its energies, mixed-precision forward and adjoint solves
    energies = (2.0, 6.5)
    opts = adjoint.SolveOptions(method='fp32', refinement='fgmres', max_steps=20)
    fp = adjoint.FitProblem(nch=32, nr=200, energies=energies, seed=0, opts=opts)
    loss, g = fp.loss_and_grad()
    prob = oadj.problem(32, 200, 15.0, 3.0, 0, energies)
    th = fp.theta0.cpu().numpy()
    tg = [t.numpy() for t in fp.targets]
    loss_ref, g_ref = oadj.loss_and_grad(prob, th, tg)
    assert abs(loss - loss_ref) <= 1e-8 * loss_ref
    g = g.cpu().numpy()
    assert np.abs(g - g_ref).max() <= 1e-8 * np.abs(g_ref).max()
