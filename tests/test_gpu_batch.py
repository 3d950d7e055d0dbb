"""GPU parity of the uniform-batch path (hprmat_hip_batch_solve and the
batched assembly / R->S): against LAPACK solves of random systems, bit for
bit against the single-system path on Lagrange-mesh systems, and S against
the FP64 oracle."""
import numpy as np
import pytest
import torch

from hprmat_amd import rmatrix, scan, solver
from oracle import rmatrix as ormat

pytestmark = pytest.mark.gpu


def _random_systems(n, nrhs, count, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((count, n, n)) + 1j * rng.standard_normal((count, n, n))
    b = rng.standard_normal((count, n, nrhs)) + 1j * rng.standard_normal((count, n, nrhs))
    return a, b


def _cm(x):
    """(count, rows, cols) matrices -> device column-major storages (count, cols, rows)."""
    return torch.as_tensor(np.ascontiguousarray(x.transpose(0, 2, 1))).cuda()


@pytest.mark.parametrize('method', ['fp64', 'fp32', 'fp32_3m'])
@pytest.mark.parametrize('n,count', [(80, 1), (300, 5), (520, 2), (1100, 3)])
def test_batch_solve_matches_lapack(method, n, count):
    # n = 300 and 520: the panel leaf spans two and three workgroups and the
    # last 256-column block is ragged; 1100: five blocks, 20 leaf workgroups
    nrhs = 3
    print('batch vs LAPACK: method %s n %d count %d' % (method, n, count), flush=True)
    a, b = _random_systems(n, nrhs, count, seed=7 * n + count)
    x_ref = np.linalg.solve(a, b)
    bt = solver.Batch(n, nrhs, count, method, equilibrate=False)
    b_cm = _cm(b)
    bt.solve_(_cm(a), b_cm, max_refine=10, tol=1e-12)
    x = b_cm.cpu().numpy().transpose(0, 2, 1)
    status, res = bt.status.cpu().numpy(), bt.res.cpu().numpy()
    assert (status == 1).all(), (status, res)
    assert (res <= 1e-12).all(), res
    err = np.abs(x - x_ref).max(axis=(1, 2)) / np.abs(x_ref).max(axis=(1, 2))
    assert (err <= 1e-9).all(), err


def test_batch_singular_member_and_independent_stopping():
    # system 1 has an exactly zero column (zero pivot at column 8, reported as
    # -8); system 2 is badly scaled so its FP32 refinement takes more sweeps;
    # the other systems must be unaffected by either
    n, nrhs, count = 200, 2, 4
    a, b = _random_systems(n, nrhs, count, seed=11)
    a[1][:, 7] = 0.0
    a[2] = a[2] + np.diag(np.full(n, 1e-3)) - np.diag(np.diag(a[2]))
    a[2][:, :20] *= 1e3
    bt = solver.Batch(n, nrhs, count, 'fp32', equilibrate=False)
    b_cm = _cm(b)
    bt.solve_(_cm(a), b_cm, max_refine=30, tol=1e-12)
    x = b_cm.cpu().numpy().transpose(0, 2, 1)
    status, res = bt.status.cpu().numpy(), bt.res.cpu().numpy()
    assert status[1] == -8, status
    for e in (0, 2, 3):
        assert status[e] == 1 and res[e] <= 1e-12, (e, status, res)
        ref = np.linalg.solve(a[e], b[e])
        assert np.abs(x[e] - ref).max() <= 1e-8 * np.abs(ref).max()


def _wave(nch, nr, energies, index=0):
    return scan.Wave(index, nch, nr, np.asarray(energies, dtype=np.float64))


@pytest.mark.parametrize('method', ['fp64', 'fp32'])
@pytest.mark.parametrize('nch,nr', [(4, 60), (8, 100)])
def test_batch_bitwise_matches_single_path(method, nch, nr):
    # Same kernels in the same order: the batched factors and first solution
    # must equal the single-system path's bit for bit (no refinement: the
    # residual norms are reduced in a different order, which may move a
    # stopping decision but not the arithmetic of a sweep)
    w = _wave(nch, nr, (0.9, 2.5, 4.0, 7.5))
    ne, n = len(w.energies), w.n
    print('batch vs single: method %s n %d' % (method, n), flush=True)
    a = torch.empty((ne, n, n), dtype=torch.complex128, device='cuda')
    rmatrix.assemble_batch(w.dmesh, w.lv, w.qk, w.xc1, w.cpot, a)
    b = w.b.unsqueeze(0).repeat(ne, 1, 1).contiguous()
    bt = solver.Batch(n, nch, ne, method, equilibrate=True)
    bt.solve_(a, b, max_refine=0, tol=1e-12)
    for k in range(ne):
        a1 = rmatrix.assemble(w.dmesh, w.sysm.lval, w.qk_host[k], w.xc1[k].cpu().numpy(),
                              w.cpot)
        assert torch.equal(a1, a[k]), k
        x1 = w.b.clone()
        # (the batched leaves use partial pivoting: the single path's rule)
        lu = solver.LU().set_plan(lookahead=0, pivot='partial').factor(a1, method, True)
        lu.solve_(x1, 0, 1e-12, 'rel_fro', True)
        assert torch.equal(x1, b[k]), (k, (x1 - b[k]).abs().max().item())


def test_r_to_s_batch_matches_single_and_oracle():
    w = _wave(8, 40, (0.2, 1.5, 4.0, 9.0))
    ne, n = len(w.energies), w.n
    a = torch.empty((ne, n, n), dtype=torch.complex128, device='cuda')
    rmatrix.assemble_batch(w.dmesh, w.lv, w.qk, w.xc1, w.cpot, a)
    x = w.b.unsqueeze(0).repeat(ne, 1, 1).contiguous()
    bt = solver.Batch(n, w.nch, ne, 'fp64', equilibrate=True)
    bt.solve_(a, x, max_refine=5, tol=1e-12)
    assert (bt.status.cpu().numpy() == 1).all()
    nopen = torch.as_tensor(w.nopen.astype(np.int32)).cuda()
    r = torch.zeros((ne, w.nch, w.nch), dtype=torch.complex128, device='cuda')
    s = torch.zeros_like(r)
    rmatrix.r_to_s_batch(w.dmesh, x, w.qk, nopen, w.open, w.co, w.cop, r, s)
    assert len(set(w.nopen.tolist())) > 1      # energies with different open sets
    for k in range(ne):
        no = int(w.nopen[k])
        o, co, cop, _ = rmatrix.channel_data(w.sysm.lval, w.qk_host[k], w.a)
        r1, s1 = rmatrix.r_to_s(w.dmesh, x[k].contiguous(), w.nch, w.qk_host[k], o, co, cop)
        assert torch.allclose(r[k].transpose(0, 1), r1, rtol=0, atol=1e-14 * r1.abs().max().item())
        sk = s[k].reshape(-1)[:no * no].reshape(no, no).transpose(0, 1)
        if no:
            assert (sk - s1).abs().max().item() <= 1e-13
            ref = ormat.rmatrix(w.sysm.lval, w.qk_host[k], w.a, w.nr, w.sysm.cpot)['S']
            assert np.abs(sk.cpu().numpy() - ref).max() <= 1e-10


def test_batch_graph_replays_match_eager():
    # assembly + batched solve + R->S captured once, replayed three times
    # with eager work and synchronisations in between (no memset nodes)
    w = _wave(4, 60, (1.1, 3.3, 6.6))
    ne, n = len(w.energies), w.n
    a = torch.empty((ne, n, n), dtype=torch.complex128, device='cuda')
    x = torch.empty((ne, w.nch, n), dtype=torch.complex128, device='cuda')
    nopen = torch.as_tensor(w.nopen.astype(np.int32)).cuda()
    r = torch.zeros((ne, w.nch, w.nch), dtype=torch.complex128, device='cuda')
    s = torch.zeros_like(r)
    bt = solver.Batch(n, w.nch, ne, 'fp32', equilibrate=True)

    def run():
        rmatrix.assemble_batch(w.dmesh, w.lv, w.qk, w.xc1, w.cpot, a)
        x.copy_(w.b.unsqueeze(0).expand(ne, -1, -1))
        bt.solve_(a, x, max_refine=10, tol=1e-12)
        rmatrix.r_to_s_batch(w.dmesh, x, w.qk, nopen, w.open, w.co, w.cop, r, s)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    eager = s.clone()
    assert (bt.status.cpu().numpy() == 1).all()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run()
    for rep in range(3):
        s.zero_()
        x.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(s, eager), rep
