"""GPU parity of the batched energy scan (hprmat_amd.scan): every system's S
against the FP64 oracle, graph replays against the eager run, the FP64
serial loop, and the autotuner's plans."""
import numpy as np
import pytest
import torch

from hprmat_amd import scan
from oracle import rmatrix as ormat

pytestmark = pytest.mark.gpu

SIZES = [(2, 40), (4, 60), (8, 40), (16, 40), (4, 100)]


def _waves(energies=(0.7, 2.1, 4.4, 6.0, 9.5)):
    return [scan.Wave(i, nc, nr, np.asarray(energies)) for i, (nc, nr) in enumerate(SIZES)]


def _oracle_s(w, k):
    out = ormat.rmatrix(w.sysm.lval, w.qk_host[k], w.a, w.nr, w.sysm.cpot)
    return out['S']


def _oracle_s_refined(w, k):
    """S of the host FP64 solve refined to ||r||_F/||b||_F <= 1e-12 (the
    reference's ZGESV result, then FP64 sweeps with the same factors), and S
    of the plain ZGESV.  On the kappa ~ 1e9 waves ZGESV's residual is a few
    1e-12 and its S up to 2e-10 from the refined one."""
    from oracle import solve as osolve
    out = ormat.rmatrix(w.sysm.lval, w.qk_host[k], w.a, w.nr, w.sysm.cpot)
    x, _ = osolve.fp64_refined(out['C'], out['B'])
    r = ormat.extract_r(x, out['mesh']['q2'][0], w.nch)
    s = (ormat.r_to_s(r, w.a, w.qk_host[k], out['npo'], out['co'], out['cop'])
         if out['nopen'] else out['S'])
    return s, out['S']


def test_scan_matches_oracle():
    waves = _waves()
    sc = scan.EnergyScan(waves, lanes=3, graph_max=300)   # some systems stay uncaptured
    sc.capture()
    assert any(l.graphs for l in sc.lanes)
    sc.run()
    torch.cuda.synchronize()
    for w in waves:
        for k in range(len(w.energies)):
            ref = _oracle_s(w, k)
            got = w.s_of(k).cpu().numpy()
            assert got.shape == ref.shape
            if ref.size:
                assert np.abs(got - ref).max() <= 1e-10, (w.n, w.energies[k])


def test_graph_replays_match_eager_run():
    waves = _waves((1.0, 3.0, 8.0))
    sc = scan.EnergyScan(waves, lanes=2, graph_max=10 ** 6)
    sc.run()                                   # eager (no graph yet)
    torch.cuda.synchronize()
    eager = [w.S.clone() for w in waves]
    sc.capture()
    for rep in range(3):
        for w in waves:
            w.S.zero_()
        torch.cuda.synchronize()
        sc.run()                               # graph replay
        torch.cuda.synchronize()
        for w, e in zip(waves, eager):
            assert torch.equal(w.S, e), rep


def test_serial_fp64_loop_matches_oracle():
    waves = _waves((2.5, 7.0))
    pairs = [(w, k) for w in waves for k in range(len(w.energies))]
    scan.serial_fp64(waves, pairs)
    torch.cuda.synchronize()
    for w, k in pairs:
        ref = _oracle_s(w, k)
        if ref.size:
            assert np.abs(w.s_of(k).cpu().numpy() - ref).max() <= 1e-10


def test_autotune_plans_and_cache(tmp_path):
    waves = [scan.Wave(i, nc, nr, np.linspace(0.5, 10, 4)) for i, (nc, nr) in
             enumerate([(2, 40), (8, 80)])]
    cache = str(tmp_path / 'tune.json')
    plans, lanes = scan.autotune(waves, sample=2, cache=cache, lane_options=(1, 2))
    assert set(plans) == {(80, 2), (640, 8)} and lanes in (1, 2)
    for p in plans.values():
        assert p['method'] in ('fp32', 'bf16x3') and p['refinement'] in ('ir', 'fgmres')
    plans2, lanes2 = scan.autotune(waves, sample=2, cache=cache)
    assert plans2 == plans and lanes2 == lanes
    sc = scan.EnergyScan(waves, lanes=lanes, plans=plans)
    sc.run()
    torch.cuda.synchronize()
    for w in waves:
        for k in range(4):
            ref = _oracle_s(w, k)
            if ref.size:
                assert np.abs(w.s_of(k).cpu().numpy() - ref).max() <= 1e-10


def test_batch_scan_matches_oracle_and_replays():
    # uniform-batch scan: one chain per chunk of energies; a small memory
    # budget forces several chunks per wave, three lanes, then graph replays
    waves = _waves((0.7, 2.1, 4.4, 6.0, 9.5))
    sc = scan.BatchScan(waves, method='fp32', lanes=3, mem_bytes=3 * 2 * 400 * 400 * 16 * 2)
    assert any(sc.chunk[id(w)] < len(w.energies) for w in waves)
    print('batch scan: eager', flush=True)
    sc.run()
    assert sc.finish() == 0
    for w in waves:
        assert (w.status.cpu().numpy() == 1).all()
        assert (w.res.cpu().numpy() <= 1e-12).all()
        for k in range(len(w.energies)):
            ref = _oracle_s(w, k)
            got = w.s_of(k).cpu().numpy()
            assert got.shape == ref.shape
            if ref.size:
                assert np.abs(got - ref).max() <= 1e-10, (w.n, w.energies[k])
    eager = [w.S.clone() for w in waves]
    print('batch scan: capture', flush=True)
    sc.capture()
    assert sc.graph is not None          # the whole scan is one graph
    for rep in range(3):
        for w in waves:
            w.S.zero_()
            w.status.zero_()
        torch.cuda.synchronize()
        sc.run()
        assert sc.finish() == 0
        for w, e in zip(waves, eager):
            assert torch.equal(w.S, e), rep


def test_batch_scan_bench_configuration_at_measured_sizes():
    # the benchmark's scan configuration (Gauss three-product FP32 factors,
    # refinement products from the Lagrange-mesh parts, 15 sweeps, two lanes,
    # captured graphs, FP64 re-solve of stragglers) on two of its own waves:
    # wave 0 (64 channels x 100 points, N = 6400, batched leaves of 25
    # workgroups) and wave 21 (32 x 80, N = 2560), at four energies of its
    # grid, among them E = 2.288 where kappa ~ 1e9; S against the host FP64
    # solve refined to the residual gate (plain ZGESV printed beside it)
    grid = np.linspace(0.5, 10.0, 256)
    energies = grid[[0, 48, 128, 255]]
    sizes = scan.draw_waves(64, 0)
    assert sizes[0] == (64, 100) and sizes[21] == (32, 80)
    waves = [scan.Wave(i, *sizes[i], energies, emax=3.0, seed=0) for i in (0, 21)]
    sc = scan.BatchScan(waves, method='fp32_3m', lanes=2, max_refine=15)
    assert sc.mesh_operator
    print('bench-configuration scan: capture', flush=True)
    sc.capture()
    print('bench-configuration scan: replay', flush=True)
    sc.run()
    sc.finish()
    for w in waves:
        res = w.res.cpu().numpy()
        for k in range(len(w.energies)):
            ref, ref_zgesv = _oracle_s_refined(w, k)
            got = w.s_of(k).cpu().numpy()
            ds = np.abs(got - ref).max() if ref.size else 0.0
            dz = np.abs(got - ref_zgesv).max() if ref.size else 0.0
            print('wave %d N=%d E=%.3f res %.2e |dS| %.2e (vs plain zgesv %.2e)'
                  % (w.index, w.n, w.energies[k], res[k], ds, dz), flush=True)
            assert got.shape == ref.shape
            assert ds <= 1e-10, (w.n, w.energies[k], ds)
            assert res[k] <= 3e-12


def test_batch_scan_lane_graphs_match_single_graph():
    # the per-lane chain graphs (capture(single=False)) give the single
    # graph's S bit for bit
    waves = _waves((1.1, 5.3, 9.1))
    sc1 = scan.BatchScan(waves, method='fp32_3m', lanes=3)
    sc1.capture()
    sc1.run()
    assert sc1.finish() == 0
    one = [w.S.clone() for w in waves]
    sc2 = scan.BatchScan(waves, method='fp32_3m', lanes=3)
    sc2.capture(single=False)
    assert sc2.graph is None and any(l.graphs for l in sc2.lanes)
    for w in waves:
        w.S.zero_()
    sc2.run()
    assert sc2.finish() == 0
    for w, e in zip(waves, one):
        assert torch.equal(w.S, e)


def test_batch_scan_fallback_resolves_unconverged():
    # with no refinement sweep allowed the FP32 factors alone miss 1e-12:
    # finish() must re-solve those systems in FP64 and reach the oracle's S
    waves = _waves((1.3, 5.5))[:3]
    sc = scan.BatchScan(waves, method='fp32', lanes=2, max_refine=0)
    sc.run()
    n_redo = sc.finish()
    assert n_redo > 0
    for w in waves:
        assert (w.status.cpu().numpy() == 1).all()
        for k in range(len(w.energies)):
            ref = _oracle_s(w, k)
            if ref.size:
                assert np.abs(w.s_of(k).cpu().numpy() - ref).max() <= 1e-10


@pytest.mark.parametrize('method', ['fp32', 'fp32_3m'])
@pytest.mark.parametrize('trail,subst', [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)])
def test_batch_tile_plans_match_oracle(method, trail, subst):
    # every tile plan the batched autotuner chooses from gives the oracle's S
    # on waves of one block (N = 80, 240), several blocks (N = 400, 960) and a
    # ragged last block (N = 640)
    sizes = [(2, 40), (4, 60), (4, 100), (16, 60), (8, 80)]
    waves = [scan.Wave(i, nc, nr, np.asarray([0.9, 4.7, 8.8])) for i, (nc, nr) in
             enumerate(sizes)]
    plans = {(w.n, w.nch): {'method': method, 'trail': trail, 'subst': subst, 'max_refine': 20}
             for w in waves}
    sc = scan.BatchScan(waves, lanes=2, plans=plans)
    assert all(sc.plan_of(w)['trail'] == trail for w in waves)
    sc.run()
    assert sc.finish() == 0
    for w in waves:
        assert (w.status.cpu().numpy() == 1).all()
        for k in range(len(w.energies)):
            ref = _oracle_s(w, k)
            if ref.size:
                assert np.abs(w.s_of(k).cpu().numpy() - ref).max() <= 1e-10, (w.n, k)


def test_batch_autotune_plans_and_cache(tmp_path):
    waves = [scan.Wave(i, nc, nr, np.linspace(0.5, 10, 6)) for i, (nc, nr) in
             enumerate([(2, 40), (8, 80), (16, 40)])]
    cache = str(tmp_path / 'batch_tune.json')
    plans, lanes = scan.autotune_batch(waves, sample=3, cache=cache, lane_options=(1, 2))
    assert set(plans) == {(80, 2), (640, 8), (640, 16)} and lanes in (1, 2)
    for p in plans.values():
        assert p['method'] in ('fp32', 'fp32_3m')
        assert p['trail'] in (0, 1, 2) and p['subst'] in (0, 1)
        assert p['max_refine'] in range(3, 21)
    plans2, lanes2 = scan.autotune_batch(waves, sample=3, cache=cache)
    assert plans2 == plans and lanes2 == lanes
    sc = scan.BatchScan(waves, lanes=lanes, plans=plans)
    sc.capture()
    sc.run()
    sc.finish()
    for w in waves:
        for k in range(len(w.energies)):
            ref = _oracle_s(w, k)
            if ref.size:
                assert np.abs(w.s_of(k).cpu().numpy() - ref).max() <= 1e-10
