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


# ---- several intervals and wave functions (host parts) ----

from oracle import lagrange as olag, propagation as oprop  # noqa: E402


@pytest.mark.parametrize('nr,ns', [(7, 1), (8, 2), (9, 3), (10, 4)])
def test_interval_mesh_matches_oracle(nr, ns):
    a = 11.0
    m = physics.mesh(nr, ns, a)
    o = olag.rmat_ini(nr, ns, a)
    for s in range(ns):
        assert np.allclose(m['tc'][s], o['tc'][s], rtol=1e-12, atol=1e-12 * np.abs(o['tc'][s]).max())
        assert np.allclose(m['bloch'][s], oprop.bloch_closed(o, s), rtol=1e-12, atol=1e-14)
        assert np.allclose(m['r'][s], oprop.interval_radii(o, s), rtol=1e-14)
        assert np.allclose(m['q1'][s], o['q1'][s], rtol=1e-12, atol=1e-15)
        assert np.allclose(m['q2'][s], o['q2'][s], rtol=1e-12)
    assert np.allclose(m['wle'], o['wle'], rtol=1e-12)


def test_wavefunction_spline_and_asymptotics_match_oracle():
    # hprmat_wavefunction is host work: interpolate the oracle's wave
    # functions and continue them with the oracle's S
    hm = 20.736 / 0.5
    o = olag.rmat_ini(25, 3, 8.0)
    from tests.test_oracle_propagation import reid
    qk = np.array([np.sqrt(24.0 / hm)] * 2)
    out = oprop.rmatrix_ns([0, 2], qk, 8.0, 25, 3, reid(1, o['zrma']), nvc=[0])
    for iv in (1, 2):
        got = physics.wavefunction([0, 2], qk, None, 8.0, 25, 3, out['cu'], 2, out['cf'],
                                   o['zrma'], iv, 1, 60, 0.2)
        want = oprop.wf_print([0, 2], qk, 8.0, out['cu'], 2, out['cf'], o['zrma'], iv - 1, 0, 60, 0.2)
        assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()


def test_wavefunction_closed_channel_asymptotics():
    # a closed channel outside the radius: amplitude x W(|k| r) = e^-|k|r (l = 0)
    nch, nr, ns = 2, 6, 2
    qk = np.array([0.8, -0.5])
    cu = np.zeros((nch, nch), complex)
    cu[1, 0] = 0.3 - 0.2j
    cf = np.zeros((nr * ns, nch, 1), complex)
    z = physics.init(nr, ns, 4.0)
    out = physics.wavefunction([0, 0], qk, None, 4.0, nr, ns, cu, 1, cf, z, 2, 1, 30, 0.2)
    r = 0.2 * np.arange(1, 31)
    sel = r > 4.0
    assert np.allclose(out[sel], cu[1, 0] * np.exp(-0.5 * r[sel]), rtol=1e-12)
    assert not out[~sel].any()


def test_solve_full_rejects_unsupported_inputs():
    nch, nr = 2, 10
    cp = np.zeros((nr * 2, nch, nch), complex)
    lib = physics._lib.load()
    cu = np.ones((nch, nch), complex, order='F')
    cf = np.zeros((nr * 2, nch, 1), complex, order='F')
    nopen = physics.ctypes.c_int(7)
    lval = np.zeros(nch, np.int32)
    qk = np.array([0.5, 0.7])
    eta = np.zeros(nch)
    nvc = np.array([1], np.int32)
    cpnl = np.zeros((4, nch, nch), complex, order='F')
    lib.hprmat_solve_full(nch, physics._ptr(lval), physics._ptr(qk), physics._ptr(eta), 8.0, nr,
                          2, physics._ptr(np.asfortranarray(cp)), nr * 2, physics._ptr(cu),
                          physics._ptr(cf), nr * 2, nch, 1, physics._ptr(nvc),
                          physics._ptr(cpnl), 4, physics.ctypes.byref(nopen), 1)
    assert nopen.value == -1 and not cu.any()
    # entrance channel not open; nch > 64 with several intervals
    cu, cf2, no = physics.solve_full([0, 0], [0.5, -0.7], None, 8.0, nr, cp, [2], ns=2)
    assert no == -1
    big = np.zeros((nr * 2, 65, 65), complex)
    cu, no = physics.solve(np.zeros(65, int), np.full(65, 0.5), None, 8.0, nr, big, ns=2)
    assert no == -1


def test_solve_full_rejects_nonlocal_with_several_intervals():
    nr = 8
    cp = np.zeros((nr * 2, 1, 1), complex)
    k = np.zeros((nr, nr, 1, 1), complex)
    cu, cf, no = physics.solve_full([0], [0.5], None, 8.0, nr, cp, [1], ns=2, cpnl=k)
    assert no == -1
