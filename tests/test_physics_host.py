"""Host-side parts of the physics C interface (hprmat_init, solver state,
argument checks) through libhprmat_hip, on the CPU: nothing here touches the
GPU.  The mesh abscissas are pinned by the oracle's mesh (itself pinned to
numpy's Gauss-Legendre rule) and by the reference's multi-interval rule
zrma = (s + x_i) rmax / ns (src/rmatrix_hp.F90:153-156)."""
import numpy as np
import pytest

from hprmat_amd import physics
from oracle import lagrange as olag


@pytest.mark.parametrize('nr,a', [(1, 3.0), (7, 8.0), (60, 15.0), (200, 15.0)])
def test_init_matches_oracle_mesh(nr, a):
    z = physics.init(nr, 1, a)
    mesh = olag.rmat_ini(nr, 1, a)
    assert np.abs(z - mesh['xle'] * a).max() <= 1e-13 * a


def test_init_multi_interval_rule():
    nr, ns, a = 12, 3, 9.0
    z = physics.init(nr, ns, a)
    x = physics.init(nr, 1, 1.0)
    for s in range(ns):
        assert np.abs(z[s * nr:(s + 1) * nr] - (s + x) * a / ns).max() <= 1e-14


def test_solver_state_and_strings():
    old = physics.get_solver()
    try:
        physics.set_solver(4)
        assert physics.get_solver() == 4
    finally:
        physics.set_solver(old)
    assert physics.version().startswith('hprmat_hip')
    assert 'FP64' in physics.solver_info(1) and 'FGMRES' in physics.solver_info(4)
    assert physics.solver_info(9) == 'Unknown solver'


@pytest.mark.parametrize('kwargs', [dict(eta=[0.3, 0.0]), dict(ns=2), dict(solver=7)])
def test_solve_rejects_unsupported_inputs(kwargs):
    # eta != 0, several intervals and unknown solver types are refused before
    # any device work: nopen = -1 and a zero collision matrix
    nr = 10
    args = dict(lval=[0, 1], qk=[1.0, -0.5], eta=None, rmax=8.0, nr=nr,
                cpot=np.zeros((nr * kwargs.get('ns', 1), 2, 2)), ns=1, solver=0)
    args.update(kwargs)
    cu, nopen = physics.solve(**args)
    assert nopen == -1 and not cu.any()
