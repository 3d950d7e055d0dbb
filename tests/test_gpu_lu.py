"""GPU parity of the dense complex LU + solve path (libhprmat_hip) against the
FP64 oracle (LAPACK).  Needs an MI355X; the library must be built in-tree."""
import numpy as np
import pytest
import torch
from scipy.linalg import lapack

from hprmat_amd import rmatrix, solver, synth
from oracle import solve as osolve

pytestmark = pytest.mark.gpu


def rand_system(n, nrhs, seed, dd=0.02):
    rng = np.random.default_rng(seed)
    a = (rng.random((n, n)) - 0.5) + 1j * (rng.random((n, n)) - 0.5) + 2 * n * dd * np.eye(n)
    b = rng.random((n, nrhs)) + 1j * rng.random((n, nrhs))
    return a, b


def gpu_solve(a, b, method, max_refine=0, tol=1e-13, equilibrate=False, pivot=None):
    x, st = solver.solve(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), method=method,
                         max_refine=max_refine, tol=tol, equilibrate=equilibrate, pivot=pivot)
    return x.cpu().numpy(), st


@pytest.mark.parametrize('n', [1, 7, 32, 33, 64, 255, 256, 257, 500, 1100])
@pytest.mark.parametrize('method', ['fp64', 'fp32', 'fp32_3m', 'bf16x3', 'bf16x6', 'f16x3'])
def test_solve_matches_lapack(n, method):
    a, b = rand_system(n, 5, n)
    x_ref, info = osolve.zgesv(a, b)
    assert info == 0
    x, st = gpu_solve(a, b, method, max_refine=0 if method == 'fp64' else 12)
    err = np.abs(x - x_ref).max() / np.abs(x_ref).max()
    assert st['info'] == 0
    assert err <= 1e-12, (n, method, err, st)


@pytest.mark.parametrize('n', [8, 40, 100, 300, 700, 1300])
def test_pivots_match_lapack_zgetrf(n):
    # index output: bit-exact
    a, _ = rand_system(n, 1, 1000 + n, dd=0.0)
    lu = solver.LU().factor(solver.colmajor(torch.from_numpy(a).cuda()), 'fp64')
    piv = lu.pivots()
    _, piv_ref, info = lapack.zgetrf(a)
    assert info == 0
    assert np.array_equal(piv, piv_ref + 1)


@pytest.mark.parametrize('n', [700, 1500])
def test_split_bf16_trailing_update_precision(n):
    # A split-BF16 operand keeps about 17 significant bits (hi and lo carry 8
    # each; the dropped lo*lo term is below 2^-17 relative), FP32 24: with
    # the panel in FP32 and only the trailing update in BF16 products, the
    # unrefined bf16x3 error lies clearly above the FP32 one but within two
    # orders of magnitude, which also shows that the BF16 kernel ran.  No
    # diagonal shift: the multipliers are O(1), so the update is as large as A22.
    a, b = rand_system(n, 4, 31 + n, dd=0.0)
    x_ref, _ = osolve.zgesv(a, b)
    errs = {}
    for method in ('fp32', 'bf16x3'):
        x, st = gpu_solve(a, b, method, max_refine=0, pivot='partial')
        assert st['info'] == 0
        errs[method] = np.abs(x - x_ref).max() / np.abs(x_ref).max()
    assert 2 * errs['fp32'] < errs['bf16x3'] < 200 * errs['fp32'], errs


@pytest.mark.parametrize('n', [700, 1500, 3000])
def test_split_bf16x6_trailing_update_precision(n):
    # Three BF16 pieces (hi, mid, lo) carry 24 significant bits and the six
    # kept piece products drop only terms below 2^-24 relative, so the
    # unrefined bf16x6 factorization is as accurate as the FP32 one (within a
    # small factor either way) and far more accurate than bf16x3; its result
    # differs from FP32's in the last bits, which shows its kernel ran.
    a, b = rand_system(n, 4, 57 + n, dd=0.0)
    x_ref, _ = osolve.zgesv(a, b)
    errs, xs = {}, {}
    for method in ('fp32', 'bf16x3', 'bf16x6'):
        x, st = gpu_solve(a, b, method, max_refine=0, pivot='partial')
        assert st['info'] == 0
        xs[method] = x
        errs[method] = np.abs(x - x_ref).max() / np.abs(x_ref).max()
    print('n %d errors %s' % (n, errs), flush=True)
    assert errs['bf16x6'] < 4 * errs['fp32'], errs
    assert errs['bf16x6'] * 3 < errs['bf16x3'], errs
    assert not np.array_equal(xs['bf16x6'], xs['fp32'])


@pytest.mark.parametrize('n', [700, 1500, 3000])
def test_split_f16_trailing_update_precision(n):
    # (partial pivoting for all: the same pivot rule as the other arithmetic)
    # Two FP16 pieces carry 22 significant bits; with every row of L21 and
    # every column of U12 scaled by a power of two into FP16's range, the
    # three kept products (hh, hl, lh) drop only terms below about 2^-22
    # relative.  The unrefined f16x3 factorization is within a small factor
    # of the FP32 one and far more accurate than bf16x3 (about 2^-16); its
    # result differs from FP32's, which shows its kernel ran.
    a, b = rand_system(n, 4, 83 + n, dd=0.0)
    x_ref, _ = osolve.zgesv(a, b)
    errs, xs = {}, {}
    for method in ('fp32', 'bf16x3', 'f16x3'):
        x, st = gpu_solve(a, b, method, max_refine=0, pivot='partial')
        assert st['info'] == 0
        xs[method] = x
        errs[method] = np.abs(x - x_ref).max() / np.abs(x_ref).max()
    print('n %d errors %s' % (n, errs), flush=True)
    assert errs['f16x3'] < 8 * errs['fp32'], errs
    assert errs['f16x3'] * 4 < errs['bf16x3'], errs
    assert not np.array_equal(xs['f16x3'], xs['fp32'])


@pytest.mark.parametrize('n', [600, 1300])
def test_split_f16_scaled_rows_and_columns(n):
    # Rows and columns of A scaled by powers of two spanning 2^-60 .. 2^60:
    # far outside FP16's range (6e-5 .. 65504), so the factorization only
    # works through the power-of-two scales of the split.  The row scaling
    # D A puts D_k into column k of L (inversely) and row k of U, so without
    # the K-index scales of L21's columns / U12's rows this fails (errors of
    # order 1e2); with them the unrefined error stays within a small factor
    # of the FP32 factorization's, and the refined one too.
    rng = np.random.default_rng(n)
    a, b = rand_system(n, 3, 7 + n, dd=0.0)
    dr = np.exp2(rng.integers(-60, 61, n).astype(float))
    dc = np.exp2(rng.integers(-60, 61, n).astype(float))
    a = dr[:, None] * a * dc[None, :]
    b = dr[:, None] * b
    x_ref, info = osolve.zgesv(a, b)
    assert info == 0
    errs = {}
    for method in ('fp32', 'f16x3'):
        x, st = gpu_solve(a, b, method, max_refine=0, pivot='partial')
        assert st['info'] == 0
        # error of the unscaled solution dc * x
        errs[method] = np.abs((x - x_ref) * dc[:, None]).max() / np.abs(x_ref * dc[:, None]).max()
    print('n %d scaled errors %s' % (n, errs), flush=True)
    assert errs['f16x3'] < 8 * errs['fp32'], errs
    # refined: as close to LAPACK as the refined FP32 factorization gets
    # (this system's scaling leaves a floor of a few 1e-12 for both)
    ref = {}
    for method in ('fp32', 'f16x3'):
        x, st = gpu_solve(a, b, method, max_refine=12, tol=1e-13, pivot='partial')
        assert st['info'] == 0
        ref[method] = np.abs((x - x_ref) * dc[:, None]).max() / np.abs(x_ref * dc[:, None]).max()
    print('n %d refined errors %s' % (n, ref), flush=True)
    assert ref['f16x3'] <= max(4 * ref['fp32'], 1e-12), ref


@pytest.mark.parametrize('n', [16500, 20000])
@pytest.mark.parametrize('method', ['fp64', 'fp32'])
def test_first_panel_pivots_tall(n, method):
    # The pivots of the first 256-column panel depend only on A[:, :256], so a
    # tall-skinny LAPACK getrf pins them at sizes where the panel leaf runs on
    # 65 .. 79 cooperating workgroups (index output: bit-exact).  FP32 with
    # the partial-pivoting plan (its default is tournament pivoting).
    g = torch.Generator(device='cuda').manual_seed(n)
    a = torch.complex(torch.rand(n, n, generator=g, device='cuda', dtype=torch.float64) - 0.5,
                      torch.rand(n, n, generator=g, device='cuda', dtype=torch.float64) - 0.5)
    lu = solver.LU().set_plan(pivot='partial').factor(solver.colmajor(a), method)
    piv = lu.pivots()[:256]
    panel = a[:, :256].cpu().numpy()
    if method == 'fp64':
        _, piv_ref, info = lapack.zgetrf(panel)
    else:
        _, piv_ref, info = lapack.cgetrf(panel.astype(np.complex64))
    assert info == 0
    mism = np.nonzero(piv != piv_ref[:256] + 1)[0]
    assert mism.size == 0, (n, method, mism[:10])


def tournament_pivots(panel, w=32, rows_per_wg=256, fan_in=8):
    """Restatement of the FP32 tournament-pivoting leaf (leaf_calu_kernel):
    GEPP (first largest |Re|+|Im|, complex64 arithmetic) on every block of
    256 rows picks w candidates; groups of 8 candidate sets are reduced by
    GEPP on their original rows until one set is left.  Returns the pivot
    rows in order."""
    def gepp_select(rows):
        a = rows.astype(np.complex64).copy()
        live = np.ones(len(a), bool)
        sel = []
        for j in range(w):
            v = np.where(live, np.abs(a[:, j].real) + np.abs(a[:, j].imag), -1.0)
            if v.max() < 0:
                break
            i = int(np.argmax(v))
            sel.append(i)
            live[i] = False
            if v[i] > 0:
                l = a[live, j] / a[i, j]
                a[live, j + 1:] -= np.outer(l, a[i, j + 1:])
        return sel
    m = panel.shape[0]
    sets = [[lo + i for i in gepp_select(panel[lo:min(m, lo + rows_per_wg)])]
            for lo in range(0, m, rows_per_wg)]
    while len(sets) > 1:
        new = []
        for q in range(0, len(sets), fan_in):
            cand = [r for st in sets[q:q + fan_in] for r in st]
            new.append([cand[i] for i in gepp_select(panel[cand])])
        sets = new
    return sets[0]


def ipiv_from_sequence(piv):
    """LAPACK interchanges (row i <-> ipiv[i], 0-based) moving the rows of a
    pivot sequence to the top in order."""
    pos = {}   # current position of a row that moved
    at = {}    # current row at a position that changed
    out = []
    for i, r in enumerate(piv):
        p = pos.get(r, r)
        out.append(p)
        x = at.get(i, i)
        at[p], pos[x] = x, p
        at[i], pos[r] = r, i
    return np.array(out)


@pytest.mark.parametrize('n', [16500, 20000])
def test_first_leaf_tournament_pivots(n):
    # The FP32 factors' default panel leaf chooses pivots by a tournament:
    # 65 .. 79 row blocks of 256, three reduction levels.  The first leaf's
    # 32 pivots depend only on A[:, :32]: they equal the restatement's
    # (index output: bit-exact), and the solve matches LAPACK after
    # refinement (test_solve_matches_lapack covers smaller sizes).
    g = torch.Generator(device='cuda').manual_seed(n)
    a = torch.complex(torch.rand(n, n, generator=g, device='cuda', dtype=torch.float64) - 0.5,
                      torch.rand(n, n, generator=g, device='cuda', dtype=torch.float64) - 0.5)
    lu = solver.LU().factor(solver.colmajor(a), 'fp32')
    assert lu.info() == 0
    piv = lu.pivots()[:32] - 1
    panel = a[:, :32].cpu().numpy().astype(np.complex64)
    ref = ipiv_from_sequence(tournament_pivots(panel))
    mism = np.nonzero(piv != ref)[0]
    assert mism.size == 0, (n, mism[:10], piv[:8], ref[:8])


def test_gauss_3m_factors_keep_partial_pivoting():
    # the automatic pivot rule keeps LAPACK's partial pivoting for the Gauss
    # three-product factorization: its first leaf's 32 pivots (no in-panel
    # update before them) equal cgetrf's on a tall matrix (73 workgroups)
    n = 18500
    g = torch.Generator(device='cuda').manual_seed(n)
    a = torch.complex(torch.rand(n, n, generator=g, device='cuda', dtype=torch.float64) - 0.5,
                      torch.rand(n, n, generator=g, device='cuda', dtype=torch.float64) - 0.5)
    lu = solver.LU().factor(solver.colmajor(a), 'fp32_3m')
    piv = lu.pivots()[:32]
    _, piv_ref, info = lapack.cgetrf(a[:, :32].cpu().numpy().astype(np.complex64))
    assert info == 0
    assert np.array_equal(piv, piv_ref[:32] + 1)


def test_tournament_interchange_replay():
    # the interchange replay of the leaf on sequences that move rows < 32
    # several times (pure index logic, checked against applying the swaps)
    rng = np.random.default_rng(5)
    for _ in range(200):
        m = int(rng.integers(32, 90))
        piv = list(rng.choice(m, 32, replace=False))
        ip = ipiv_from_sequence(piv)
        order = np.arange(m)
        for i, p in enumerate(ip):
            assert p >= i
            order[[i, p]] = order[[p, i]]
        assert list(order[:32]) == piv


@pytest.mark.parametrize('n', [700, 3000, 9000])
def test_tournament_vs_partial_pivoting_accuracy(n):
    # unrefined FP32 solves: tournament pivots within a small factor of
    # partial pivoting's error (both against LAPACK's FP64 solution)
    a, b = rand_system(n, 4, 123 + n, dd=0.0)
    x_ref, _ = osolve.zgesv(a, b)
    errs = {}
    for piv in ('partial', 'tournament'):
        lu = solver.LU().set_plan(pivot=piv).factor(solver.colmajor(torch.from_numpy(a).cuda()),
                                                    'fp32')
        assert lu.info() == 0
        bb = torch.from_numpy(b.T.copy()).cuda()
        lu.solve_(bb, 0, 1e-13, 'rel_fro', False)
        x = bb.cpu().numpy().T
        errs[piv] = np.abs(x - x_ref).max() / np.abs(x_ref).max()
    print('n %d errors %s' % (n, errs), flush=True)
    assert errs['tournament'] < 10 * errs['partial'], errs


@pytest.mark.parametrize('col', [0, 5, 63])
def test_zero_pivot_info_matches_lapack(col):
    n = 96
    a, _ = rand_system(n, 1, 7, dd=0.0)
    a[:, col] = 0
    lu = solver.LU().factor(solver.colmajor(torch.from_numpy(a).cuda()), 'fp64')
    _, _, info_ref = lapack.zgetrf(a)
    assert lu.info() == info_ref == col + 1


@pytest.mark.parametrize('method,col', [('fp32', 5), ('fp32', 300), ('f16x3', 300)])
def test_zero_pivot_info_tournament(method, col):
    # FP32 factors pivot by tournament (three leaf workgroups at n = 700):
    # an exactly zero column stays zero under every elimination, so the
    # first zero pivot is still at that column, as LAPACK reports it
    n = 700
    a, _ = rand_system(n, 1, 17, dd=0.0)
    a[:, col] = 0
    lu = solver.LU().factor(solver.colmajor(torch.from_numpy(a).cuda()), method)
    _, _, info_ref = lapack.zgetrf(a)
    assert lu.info() == info_ref == col + 1


@pytest.mark.parametrize('method', ['fp64', 'fp32'])
def test_equilibrated_factor_same_solution(method):
    s = synth.config('small')
    dm = rmatrix.DeviceMesh(s.mesh)
    _, _, _, xc1 = rmatrix.channel_data(s.lval, s.qk, s.a)
    a_cm = rmatrix.assemble(dm, s.lval, s.qk, xc1, s.cpot)
    b0 = rmatrix.rhs_colmajor(dm, s.nch)
    xs = []
    for eq in (False, True):
        b = b0.clone()
        st = solver.LU().factor(a_cm, method, eq).solve_(b, 10, 1e-13, 'rel_fro', True)
        assert st['res_rel'] <= 1e-12
        xs.append(b.cpu().numpy())
    assert np.abs(xs[0] - xs[1]).max() <= 1e-11 * np.abs(xs[0]).max()


def test_manufactured_solution_large():
    # X = 1 (examples/Ex0/benchmark_mgpu.f90:71-81), diagonally dominant, N = 6000
    n = 6000
    g = torch.Generator(device='cuda').manual_seed(3)
    a = torch.complex(torch.rand(n, n, generator=g, device='cuda', dtype=torch.float64),
                      torch.rand(n, n, generator=g, device='cuda', dtype=torch.float64))
    a += 2 * n * torch.eye(n, device='cuda', dtype=torch.complex128)
    b = a @ torch.ones(n, 3, device='cuda', dtype=torch.complex128)
    for m, r in [('fp64', 0), ('fp32', 5)]:
        x, st = solver.solve(a, b, method=m, max_refine=r, tol=1e-15)
        assert (x - 1).abs().max().item() < 1e-12, (m, st)


def test_refinement_history_is_monotone_until_floor():
    s = synth.System(16, 200, 15.0, 8.0, 5.0, 0)
    dm = rmatrix.DeviceMesh(s.mesh)
    _, _, _, xc1 = rmatrix.channel_data(s.lval, s.qk, s.a)
    a_cm = rmatrix.assemble(dm, s.lval, s.qk, xc1, s.cpot)
    b = rmatrix.rhs_colmajor(dm, s.nch)
    st = solver.LU().factor(a_cm, 'fp32', True).solve_(b, 12, 1e-12, 'rel_fro', True)
    h = st['hist']
    assert st['converged'] and h[-1] <= 1e-12
    assert all(h[i + 1] < h[i] for i in range(len(h) - 1))


@pytest.mark.parametrize('n', [300, 1100, 3000])
@pytest.mark.parametrize('method', ['fp64', 'fp32', 'bf16x3', 'f16x3'])
def test_fgmres_refinement_matches_lapack(n, method):
    a, b = rand_system(n, 6, 500 + n, dd=0.002)
    x_ref, info = osolve.zgesv(a, b)
    assert info == 0
    x, st = gpu_solve_fgmres(a, b, method)
    assert st['info'] == 0 and st['converged'], st
    assert st['res_rel'] <= 1e-12, st
    err = np.abs(x - x_ref).max() / np.abs(x_ref).max()
    assert err <= 1e-10, (n, method, err)
    # the residual of the returned X, recomputed on the host
    r = np.linalg.norm(b - a @ x) / np.linalg.norm(b)
    assert r <= 2e-12, r


def gpu_solve_fgmres(a, b, method, max_steps=16):
    x, st = solver.solve(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), method=method,
                         max_refine=max_steps, tol=1e-12, refinement='fgmres')
    return x.cpu().numpy(), st


def test_fgmres_zero_rhs_and_exact_start():
    # B = 0 gives X = 0 and no Arnoldi step (the first residual meets the rule)
    a, _ = rand_system(200, 3, 9)
    b = np.zeros((200, 3), complex)
    x, st = gpu_solve_fgmres(a, b, 'fp32')
    assert np.all(x == 0) and st['sweeps'] == 0, st


@pytest.mark.parametrize('method,nrhs', [('fp32', 128), ('fp32_3m', 100), ('fp64', 70)])
def test_fgmres_two_lanes_match_one_lane(method, nrhs):
    # FGMRES with the columns in two staggered lanes on two streams: every
    # lane reaches the rule on its own, so the union does; the solution
    # agrees with the one-lane solve (same products per column up to where
    # each lane stops) and with LAPACK; the combined statistics are the
    # norms of the union (ragged split: 64 + 36 and 64 + 6 columns)
    from hprmat_amd import synth as hsynth
    sysm = hsynth.System(16, 160, 15.0, 8.0, 5.0, 3)
    dm = rmatrix.DeviceMesh(sysm.mesh)
    _, _, _, xc1 = rmatrix.channel_data(sysm.lval, sysm.qk, sysm.a)
    a_cm = rmatrix.assemble(dm, sysm.lval, sysm.qk, xc1, sysm.cpot)
    n = a_cm.shape[0]
    # random combinations of the physical right-hand sides (q2 in one channel
    # block each), whose solutions the FP64 floor lets reach 1e-12
    rng = np.random.default_rng(nrhs)
    b = rmatrix.rhs_colmajor(dm, 16).cpu().numpy().T @ (
        rng.standard_normal((16, nrhs)) + 1j * rng.standard_normal((16, nrhs)))
    out = {}
    for lanes in (1, 2):
        print('fgmres lanes %d method %s n %d nrhs %d' % (lanes, method, n, nrhs), flush=True)
        lu = solver.LU().set_plan(solve_lanes=lanes).factor(a_cm, method, equilibrate=True)
        b_cm = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
        st = lu.solve_(b_cm, 24, 1e-12, 'rel_fro', True, refinement='fgmres')
        assert st['info'] == 0 and st['converged'] and st['res_rel'] <= 1e-12, (lanes, st)
        out[lanes] = (b_cm.cpu().numpy().T, st)
    a = a_cm.cpu().numpy()          # A^T = A: the storage is the matrix
    x_ref, info = osolve.zgesv(a, b)
    for lanes, (x, st) in out.items():
        rel = np.linalg.norm(b - a @ x) / np.linalg.norm(b)
        assert rel <= 1.5e-12 and abs(rel - st['res_rel']) <= 0.2 * rel, (lanes, rel, st['res_rel'])
        assert np.abs(x - x_ref).max() / np.abs(x_ref).max() <= 1e-8
    x1, x2 = out[1][0], out[2][0]
    assert np.abs(x1 - x2).max() / np.abs(x1).max() <= 1e-9


def test_full_size_mixed_solve_gate():
    # N = 25600 complex, 128 right-hand sides: the benchmark system.  Checked by
    # properties that do not need a CPU solve of that size: the FP64 relative
    # residual gate and the complex symmetry of R = q2^T X.
    s = synth.config('single')
    dm = rmatrix.DeviceMesh(s.mesh)
    o, co, cop, xc1 = rmatrix.channel_data(s.lval, s.qk, s.a)
    a_cm = rmatrix.assemble(dm, s.lval, s.qk, xc1, s.cpot)
    b = rmatrix.rhs_colmajor(dm, s.nch)
    b0 = b.clone()
    lu = solver.LU().factor(a_cm, 'fp32', True)
    st = lu.solve_(b, 24, 1e-12, 'rel_fro', True, refinement='fgmres')
    assert st['converged'] and st['res_rel'] <= 1e-12, st
    r, _ = rmatrix.r_to_s(dm, b, s.nch, s.qk, o, co, cop)
    r = r.cpu().numpy()
    assert np.abs(r - r.T).max() <= 1e-9 * np.abs(r).max()
    # classic refinement on the same factors reaches the gate too, with at
    # least as many corrections as FGMRES needs preconditioned products
    st_ir = lu.solve_(b0, 30, 1e-12, 'rel_fro', True)
    assert st_ir['converged'] and st_ir['res_rel'] <= 1e-12, st_ir
    assert st['sweeps'] <= st_ir['sweeps'], (st['sweeps'], st_ir['sweeps'])


def _memory_system(w0_zero):
    s = synth.System(256, 200, 15.0, 3.0, 6.0, 0, w0_zero=w0_zero)
    dm = rmatrix.DeviceMesh(s.mesh)
    o, co, cop, xc1 = rmatrix.channel_data(s.lval, s.qk, s.a)
    assert len(o) == 256                       # all channels open at E = 6
    a_cm = rmatrix.assemble(dm, s.lval, s.qk, xc1, s.cpot)
    b = rmatrix.rhs_colmajor(dm, s.nch)
    return s, dm, o, co, cop, a_cm, b


def test_memory_scale_mixed_solve():
    # N = 51200 (nch = 256 x nr = 200, about 42 GB as complex128 in HBM),
    # 256 right-hand sides: FP32 factorization, FGMRES refinement.  The FP64
    # attainable relative residual of this system is 1.35e-12 (FP64 LU + FP64
    # refinement stalls there too), so the gate is 2e-12 (8 products with the
    # row-2-norm equilibration); S (256 x 256, FP64 on the device) is
    # symmetric and, with absorption, |S_ij| <= 1.
    # The gate's floor is shown, not assumed: the plain FP64 LU with FP64
    # refinement on the same matrix stops above 1e-12 as well, and the mixed
    # solve lands within 20 % of it; u ||(|C| |X|)||_F / ||B||_F is printed.
    from hprmat_amd import adjoint
    s, dm, o, co, cop, a_cm, b = _memory_system(False)
    b0 = b.clone()
    lu = solver.LU().set_plan(gmres_cycle=24).factor(a_cm, 'fp32', True)
    st = lu.solve_(b, 40, 2e-12, 'rel_fro', True, refinement='fgmres')
    floor = adjoint.fp64_floor(a_cm, b, b0)
    print('memory-scale mixed: res_rel %.2e products %d hist %s; u|C||X| floor %.2e'
          % (st['res_rel'], st['sweeps'], ' '.join('%.1e' % h for h in st['hist']), floor))
    assert st['converged'] and st['res_rel'] <= 2e-12, st
    del lu
    b64 = b0.clone()
    lu64 = solver.LU().factor(a_cm, 'fp64', True)
    st64 = lu64.solve_(b64, 4, 1e-13, 'rel_fro', True)
    print('memory-scale FP64 LU + FP64 sweeps: res_rel %.2e after %d sweeps, hist %s'
          % (st64['res_rel'], st64['sweeps'], ' '.join('%.1e' % h for h in st64['hist'])))
    assert st64['res_rel'] > 1e-12              # FP64 cannot reach the nominal gate either
    assert st['res_rel'] <= 1.2 * st64['res_rel']
    assert st['res_rel'] <= 4.0 * floor
    del lu64, b64
    _, sm = rmatrix.r_to_s(dm, b, s.nch, s.qk, o, co, cop)
    sm = sm.cpu().numpy()
    assert np.abs(sm - sm.T).max() <= 1e-10
    assert np.abs(sm).max() <= 1.0 + 1e-10     # absorption: |S_ij| <= 1


def test_memory_scale_split_f16_solve():
    # The same N = 51200 system with the scaled split-FP16 trailing update
    # and tournament-pivoted leaves (the benchmark head): FGMRES to the
    # system's FP64 floor (1.35e-12, shown by the test above), and S
    # symmetric with |S_ij| <= 1.
    s, dm, o, co, cop, a_cm, b = _memory_system(False)
    lu = solver.LU().set_plan(gmres_cycle=24).factor(a_cm, 'f16x3', True)
    st = lu.solve_(b, 40, 2e-12, 'rel_fro', True, refinement='fgmres')
    print('memory-scale f16x3: res_rel %.2e products %d hist %s'
          % (st['res_rel'], st['sweeps'], ' '.join('%.1e' % h for h in st['hist'])))
    assert st['converged'] and st['res_rel'] <= 2e-12, st
    del lu
    _, sm = rmatrix.r_to_s(dm, b, s.nch, s.qk, o, co, cop)
    sm = sm.cpu().numpy()
    assert np.abs(sm - sm.T).max() <= 1e-10
    assert np.abs(sm).max() <= 1.0 + 1e-10


def test_memory_scale_real_potential_unitarity():
    # W0 = 0 (real potential): C is real symmetric indefinite with E inside
    # its spectrum, far too ill-conditioned for FP32 factors to precondition;
    # solved with the plain FP64 method.  S must be unitary and symmetric.
    s, dm, o, co, cop, a_cm, b = _memory_system(True)
    lu = solver.LU().factor(a_cm, 'fp64', True)
    st = lu.solve_(b, 5, 1e-12, 'rel_fro', True)
    _, sm = rmatrix.r_to_s(dm, b, s.nch, s.qk, o, co, cop)
    sm = sm.cpu().numpy()
    defect = np.abs(sm.conj().T @ sm - np.eye(256)).max()
    print('memory-scale W0=0: res_rel %.2e sweeps %d unitarity defect %.2e'
          % (st['res_rel'], st['sweeps'], defect))
    assert st['res_rel'] <= 1e-10, st
    assert defect <= 1e-8, defect
    assert np.abs(sm - sm.T).max() <= 1e-8


@pytest.mark.parametrize('tile', [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize('method', ['fp64', 'fp32'])
def test_trailing_tile_variants_match_lapack(tile, method):
    # every trailing-update kernel variant (16x16x4 and 32x32x2 MFMA tiles;
    # FP64 maps 3..6 to the default tile), with look-ahead, ragged edges
    n = 1337
    print('trail tile %d method %s n %d' % (tile, method, n), flush=True)
    a, b = rand_system(n, 4, 4242 + tile)
    x_ref, info = osolve.zgesv(a, b)
    assert info == 0
    lu = solver.LU().set_plan(trail_tile=tile, lookahead=1)
    lu.factor(solver.colmajor(torch.from_numpy(a).cuda()), method)
    b_cm = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    st = lu.solve_(b_cm, 0 if method == 'fp64' else 12, 1e-13, 'rel_fro', True)
    x = b_cm.cpu().numpy().T
    assert st['info'] == 0
    assert np.abs(x - x_ref).max() / np.abs(x_ref).max() <= 1e-12, st
    if method == 'fp32':
        # unrefined FP32 factors: first residual at FP32 level
        assert st['hist'][0] <= 1e-4, st['hist'][:2]


@pytest.mark.parametrize('method', ['fp64', 'fp32'])
def test_substitution_tiles_wide_rhs(method):
    # the block updates of the substitutions take 64 x 64 tiles while their
    # grid holds >= 672 of them and 32 x 32 tiles below: n = 3000 with 1024
    # right-hand sides uses both within one solve (ragged last block of 184)
    n, nrhs = 3000, 1024
    print('substitution tiles method %s n %d nrhs %d' % (method, n, nrhs), flush=True)
    a, b = rand_system(n, nrhs, 77)
    x_ref, info = osolve.zgesv(a, b)
    assert info == 0
    x, st = gpu_solve(a, b, method, max_refine=0 if method == 'fp64' else 12)
    assert st['info'] == 0
    assert np.abs(x - x_ref).max() / np.abs(x_ref).max() <= 1e-12, st


def test_replayed_substitution_graph_matches_direct_launches(monkeypatch):
    # FGMRES replays each lane's substitution chain from a captured HIP graph;
    # with HPRMAT_HIP_SUBST_GRAPH=0 the same kernels are launched directly.
    # Three solves through the graphs, with other device work and host
    # synchronisation between them, give the direct launches' solution and
    # statistics bit for bit
    s = synth.System(8, 200, 15.0, 8.0, 5.0, 1)
    dm = rmatrix.DeviceMesh(s.mesh)
    _, _, _, xc1 = rmatrix.channel_data(s.lval, s.qk, s.a)
    a_cm = rmatrix.assemble(dm, s.lval, s.qk, xc1, s.cpot)
    b0 = rmatrix.rhs_colmajor(dm, s.nch)
    lu = solver.LU().set_plan(solve_lanes=2).factor(a_cm, 'fp32', equilibrate=True)

    def run():
        b = b0.clone()
        st = lu.solve_(b, 24, 1e-12, 'rel_fro', True, refinement='fgmres')
        return b, st
    monkeypatch.setenv('HPRMAT_HIP_SUBST_GRAPH', '0')
    x_direct, st_direct = run()
    assert st_direct['converged'], st_direct
    monkeypatch.setenv('HPRMAT_HIP_SUBST_GRAPH', '1')
    for rep in range(3):
        torch.randn(1000, 1000, device='cuda').sum().item()   # other work in between
        torch.cuda.synchronize()
        x, st = run()
        assert torch.equal(x, x_direct), rep
        assert st == st_direct, rep
