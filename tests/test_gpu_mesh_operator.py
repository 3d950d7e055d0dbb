"""The refinement's FP64 products formed from the Lagrange-mesh parts of the
matrix (hprmat_hip_lu_set_operator): the solves reach the gate, their true
residual recomputed with the dense matrix agrees, and the solutions match
those refined with the dense matrix."""
import numpy as np
import pytest
import torch

from hprmat_amd import rmatrix, solver, synth

pytestmark = pytest.mark.gpu


def _system(nch, nr, emax, energy, seed, closed_fix=False):
    s = synth.System(nch, nr, 15.0, emax, energy, seed)
    dm = rmatrix.DeviceMesh(s.mesh)
    _, _, _, xc1 = rmatrix.channel_data(s.lval, s.qk, s.a)
    a_cm = rmatrix.assemble(dm, s.lval, s.qk, xc1, s.cpot)
    b = rmatrix.rhs_colmajor(dm, s.nch)
    parts = (torch.as_tensor(s.lval, dtype=torch.int32, device='cuda'),
             torch.as_tensor(s.qk, dtype=torch.float64, device='cuda'),
             torch.as_tensor(xc1, dtype=torch.float64, device='cuda'),
             torch.as_tensor(s.cpot, dtype=torch.complex128, device='cuda').contiguous())
    return s, dm, a_cm, b, parts


def _dense_resid(a_cm, x_cm, b_cm):
    # a_cm is the column-major storage of C (C = a_cm^T), x_cm / b_cm the
    # column-major storages of X and B (nrhs x N)
    r = b_cm.transpose(0, 1) - a_cm.transpose(0, 1) @ x_cm.transpose(0, 1)
    return float(torch.linalg.norm(r) / torch.linalg.norm(b_cm))


@pytest.mark.parametrize('nch,nr,emax,energy,method,refinement', [
    (4, 60, 3.0, 2.0, 'fp64', 'ir'),          # N = 240, one closed channel or more
    (4, 60, 3.0, 2.0, 'fp32', 'fgmres'),
    (16, 100, 8.0, 5.0, 'fp32', 'fgmres'),    # N = 1600, open and closed channels
    (32, 200, 8.0, 5.0, 'fp32_3m', 'fgmres'), # N = 6400
    (32, 200, 8.0, 5.0, 'fp64', 'ir'),
])
def test_operator_refinement_matches_dense(nch, nr, emax, energy, method, refinement):
    s, dm, a_cm, b0, parts = _system(nch, nr, emax, energy, 3)
    assert (s.qk < 0).any() and (s.qk > 0).any()
    print('operator %s %s N=%d' % (method, refinement, nch * nr), flush=True)
    mr = 24 if refinement == 'fgmres' else 10
    xd = b0.clone()
    lu = solver.LU().factor(a_cm, method, equilibrate=True)
    std = lu.solve_(xd, mr, 1e-12, 'rel_fro', True, refinement=refinement)
    xo = b0.clone()
    lu2 = solver.LU().set_operator(dm, *parts).factor(a_cm, method, equilibrate=True)
    sto = lu2.solve_(xo, mr, 1e-12, 'rel_fro', True, refinement=refinement)
    assert std['converged'] and sto['converged'], (std, sto)
    assert sto['res_rel'] <= 1e-12
    # the residual the operator reported is the dense matrix's
    rd = _dense_resid(a_cm, xo, b0)
    assert rd <= 1.5e-12, rd
    assert abs(rd - sto['res_rel']) <= 0.05 * rd + 1e-15, (rd, sto['res_rel'])
    assert float((xo - xd).abs().max() / xd.abs().max()) <= 1e-10


def test_operator_of_another_size_is_not_used():
    s, dm, a_cm, b0, parts = _system(4, 60, 3.0, 2.0, 1)
    s2, dm2, a2, b2, parts2 = _system(8, 60, 3.0, 2.0, 1)
    lu = solver.LU().set_operator(dm2, *parts2).factor(a_cm, 'fp32', equilibrate=True)
    x = b0.clone()
    st = lu.solve_(x, 10, 1e-12, 'rel_fro', True)
    assert st['converged'] and _dense_resid(a_cm, x, b0) <= 1.5e-12
    # and clearing it returns to the dense matrix
    lu.set_operator(None, None, None, None, None)
    x = b0.clone()
    st = lu.solve_(x, 10, 1e-12, 'rel_fro', True)
    assert st['converged']


@pytest.mark.parametrize('method', ['fp32', 'fp64'])
def test_batched_operator_matches_dense_batch(method):
    # one partial wave at several energies (open and closed channels), solved
    # as a uniform batch with the mesh operator and with the dense matrices
    from hprmat_amd import scan
    w = scan.Wave(0, 16, 60, np.linspace(0.5, 10.0, 12), emax=3.0, seed=5)
    n, nch, c = w.n, w.nch, 12
    a = torch.empty((c, n, n), dtype=torch.complex128, device='cuda')
    rmatrix.assemble_batch(w.dmesh, w.lv, w.qk[:c], w.xc1[:c], w.cpot, a)
    code = solver.batch_code(method)
    work = torch.empty(solver.batch_work_bytes(method, n, nch, c), dtype=torch.uint8, device='cuda')
    out = {}
    for tag, mesh in (('dense', None), ('mesh', solver.mesh_parts(w.dmesh, w.lv, w.qk[:c],
                                                                  w.xc1[:c], w.cpot))):
        x = w.b.unsqueeze(0).expand(c, -1, -1).contiguous()
        res = torch.zeros(c, dtype=torch.float64, device='cuda')
        status = torch.zeros(c, dtype=torch.int32, device='cuda')
        solver.batch_solve(code, n, nch, c, a.data_ptr(), x.data_ptr(), 20, 1e-12, 'rel_fro',
                           res.data_ptr(), status.data_ptr(), work, mesh=mesh)
        torch.cuda.synchronize()
        out[tag] = (x, res, status)
    x, res, status = out['mesh']
    assert bool((status == 1).all()), status
    assert float(res.max()) <= 1e-12
    xd = out['dense'][0]
    assert float((x - xd).abs().max() / xd.abs().max()) <= 1e-10
    # the dense residual of every system agrees with the one the operator
    # reported (to 5 %, or to 5e-14 ~ 200 eps where the residual itself sits
    # at the rounding floor and the two summation orders differ by that much)
    b = w.b
    for e in range(c):
        rd = _dense_resid(a[e], x[e], b)
        assert rd <= 1.5e-12 and abs(rd - float(res[e])) <= 0.05 * rd + 5e-14, (e, rd, float(res[e]))
