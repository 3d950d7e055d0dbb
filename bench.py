#!/usr/bin/env python
"""Benchmark of the dense complex LU + solve path on one MI355X.

Workload: one Lagrange-mesh system of nch = 128 channels x nr = 200 mesh
points (N = 25600 complex, complex-symmetric), 128 right-hand sides (one per
channel), synthetic potentials from hprmat_amd.synth (seed 0).  A step is one
full solve with the matrix already resident in HBM: factor (LU with partial
pivoting) + triangular solves + FP64 iterative refinement until
||B - A X||_F / ||B||_F <= 1e-12 (the last residual evaluation included),
reporting its convergence statistics to the host (the solve stops launching
refinement work once converged).

Methods timed back to back (name: factorization, refinement):
  fp64      plain FP64 LU, FP64 residual check / refinement to the gate;
  fp32      complex-FP32 LU on MFMA, FP64 refinement to the gate (<= 30 sweeps);
  fp32_ir5  the same factorization with at most 5 FP64 sweeps;
  bf16x3    FP32 LU whose trailing update uses split-BF16 three-product
            MFMA arithmetic, FP64 refinement to the gate (<= 30 sweeps);
  fp32_gmres  the fp32 factorization with flexible-GMRES refinement (FP64
            Arnoldi, the FP32 factors as preconditioner, <= 24 products);
  bf16x3_gmres  the bf16x3 factorization with flexible-GMRES refinement;
  fp32_3m   complex-FP32 LU with Gauss's three-product complex GEMM;
  fp32_3m_gmres  the fp32_3m factorization with flexible-GMRES refinement;
  bf16x6    FP32 LU whose trailing update uses three-piece split-BF16 (six
            products per real product, FP32 accuracy), classic FP64 IR;
  bf16x6_gmres  the bf16x6 factorization with flexible-GMRES refinement;
  fp32_3m_gmres_op, fp64_op  fp32_3m_gmres and fp64 with the refinement's
            FP64 products formed from the Lagrange-mesh parts of the matrix
            (channel blocks and couplings, hprmat_hip_lu_set_operator)
            instead of the dense matrix; reported in their own block, not
            as the headline, which keeps the dense drop-in operator.
Each method's solution is compared with the fp64 method's (max relative
error of X and max |dS| of the S-matrix formed from it on the device).
``value`` is the solves/s of the fastest non-FP64 method that reaches the
residual gate; per-method figures, the roofline of the dominant kernel (the
trailing-update GEMM, timed with HIP events on the launch stream) and the
CPU baseline (LAPACK zgesv + R, reference solver type 1) are in the line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-scan]
"""
import argparse
import json
import os
import sys
import time

METRIC = "solves/s & TFLOP/s, N=25600 complex LU+solve: mixed vs FP64 at resid≤1e-12"
# Ceiling of the trailing update in algorithmic complex TFLOP/s (8 flops per
# complex multiply-add) on MI355X's dense matrix peaks: FP64 78.6, FP32 157.3;
# the split-BF16 kernels execute 3 (bf16x3) or 6 (bf16x6) BF16 products per
# real product, so their ceiling is the BF16 peak 2516.6 divided by that.
BF16_PEAK = 2516.6
PEAK = {'fp64': 78.6, 'fp32': 157.3, 'fp32_ir5': 157.3, 'fp32_gmres': 157.3, 'fp32_3m': 157.3,
        'fp32_3m_gmres': 157.3,
        'bf16x3': BF16_PEAK / 3, 'bf16x3_gmres': BF16_PEAK / 3,
        'bf16x6': BF16_PEAK / 6, 'bf16x6_gmres': BF16_PEAK / 6,
        'fp32_3m_gmres_op': 157.3, 'fp64_op': 78.6}
# name -> (factorization method, max refinement sweeps)
SPECS = {'fp64': ('fp64', 10, 'ir'), 'fp32': ('fp32', 30, 'ir'), 'fp32_ir5': ('fp32', 5, 'ir'),
         'fp32_gmres': ('fp32', 24, 'fgmres'), 'bf16x3': ('bf16x3', 30, 'ir'),
         'bf16x3_gmres': ('bf16x3', 32, 'fgmres'), 'fp32_3m': ('fp32_3m', 30, 'ir'),
         'fp32_3m_gmres': ('fp32_3m', 24, 'fgmres'),
         'bf16x6': ('bf16x6', 30, 'ir'), 'bf16x6_gmres': ('bf16x6', 24, 'fgmres'),
         'fp32_3m_gmres_op': ('fp32_3m', 24, 'fgmres'), 'fp64_op': ('fp64', 10, 'ir')}


def flops_solve(n, nrhs):
    return 8.0 / 3.0 * n ** 3 + 8.0 * n * n * nrhs


def traffic_record(prec):
    """Counter-measured HBM bytes of one trailing-update launch (rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised under profiles/), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                        'trailing_gemm_traffic.json')
    try:
        with open(path) as f:
            return json.load(f)[prec]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(nch=64, nr=200, reps=3):
    """Reference solver type 1 (ZGESV + R extraction) on host cores, bounded sample."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import lagrange, rmatrix as ormat, solve as osolve, synth as osynth
    threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    lval, qk, cpot, _ = osynth.system(nch, nr, 15.0, 8.0, 5.0, 0)
    mesh = lagrange.rmat_ini(nr, 1, 15.0)
    from oracle import channels
    _, _, _, xc1 = channels.channel_functions(lval, qk, 15.0)
    c = ormat.assemble(mesh, lval, qk, cpot, xc1)
    b = ormat.rhs(mesh['q2'][0], nch)
    n = nch * nr
    with threadpool_limits(limits=threads):
        osolve.zgesv(c[:512, :512], b[:512, :4])          # warm-up of the BLAS pool
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            x, info = osolve.zgesv(c, b)
            ormat.extract_r(x, mesh['q2'][0], nch)
            ts.append(time.perf_counter() - t0)
    ts.sort()
    t = ts[len(ts) // 2]
    tflops = flops_solve(n, nch) / t / 1e12
    return {'value': tflops * 1e12 / flops_solve(25600, 128), 'unit': 'solves/s',
            'cores': threads, 'kind': 'port',
            'tflops': tflops, 'seconds': t,
            'sample': ('LAPACK zgesv (scipy/OpenBLAS) + R extraction on N=%d (nch=%d x nr=%d, '
                       'nrhs=%d) of the same recipe, median of %d; value = measured TFLOP/s '
                       'divided by the N=25600/nrhs=128 flop count' % (n, nch, nr, nch, reps))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--nch', type=int, default=128)
    ap.add_argument('--nr', type=int, default=200)
    ap.add_argument('--methods', default='fp64,fp32_gmres,bf16x6_gmres,fp32_3m_gmres,fp32,'
                                         'fp32_ir5,bf16x3,bf16x3_gmres,fp64_op,fp32_3m_gmres_op')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--copy-probe', type=int, default=0,
                    help='copy this many bytes device-to-device once before the run (a '
                         'kernel of known traffic for calibrating HBM counters)')
    ap.add_argument('--no-scan', action='store_true',
                    help='skip the batched energy scan (64 waves x 256 energies)')
    ap.add_argument('--scan-waves', type=int, default=64)
    ap.add_argument('--scan-energies', type=int, default=256)
    ap.add_argument('--scan-lanes', type=int, default=2)
    ap.add_argument('--scan-refine', type=int, default=15)
    ap.add_argument('--scan-method', default='fp32_3m', choices=['fp32', 'fp32_3m'])
    ap.add_argument('--scan-graph', type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: this build has no multi-GPU path '
              '(--gpus %d requested); run with --gpus 1' % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from hprmat_amd import _lib, rmatrix, solver, synth
    lib = _lib.load()
    dev = torch.device('cuda:0')
    torch.cuda.set_device(dev)
    # all work on one non-default stream (the library's look-ahead stream must
    # not be implicitly ordered against the legacy default stream)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        if args.copy_probe > 0:
            src = torch.ones(args.copy_probe // 4, dtype=torch.float32, device=dev)
            dst = torch.empty_like(src)
            dst.copy_(src)
            torch.cuda.synchronize()
            del src, dst
        run(args, torch, lib, _lib, rmatrix, solver, synth)


def run(args, torch, lib, _lib, rmatrix, solver, synth):

    sysm = synth.System(args.nch, args.nr, 15.0, 8.0, 5.0, 0)
    n, nrhs = sysm.n, sysm.nch
    dmesh = rmatrix.DeviceMesh(sysm.mesh)
    open_idx, co, cop, xc1 = rmatrix.channel_data(sysm.lval, sysm.qk, sysm.a)
    a_cm = rmatrix.assemble(dmesh, sysm.lval, sysm.qk, xc1, sysm.cpot)
    b0 = rmatrix.rhs_colmajor(dmesh, nrhs)
    torch.cuda.synchronize()
    fl = flops_solve(n, nrhs)

    # the Lagrange-mesh parts of the matrix, for the *_op methods
    op_parts = (torch.as_tensor(sysm.lval, dtype=torch.int32, device='cuda'),
                torch.as_tensor(sysm.qk, dtype=torch.float64, device='cuda'),
                torch.as_tensor(xc1, dtype=torch.float64, device='cuda'),
                torch.as_tensor(sysm.cpot, dtype=torch.complex128, device='cuda').contiguous())
    methods = {}
    x_ref = s_ref = None
    for m in args.methods.split(','):
        lu = solver.LU()
        fm, max_refine, refinement = SPECS[m]
        if m.endswith('_op'):
            lu.set_operator(dmesh, *op_parts)
        b = b0.clone()

        def step():
            b.copy_(b0)
            lu.factor(a_cm, fm, equilibrate=True)
            lu.solve_(b, max_refine, 1e-12, 'rel_fro', True, stats=True, refinement=refinement)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        lib.hprmat_hip_prof_enable(1)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            step()
        e1.record()
        torch.cuda.synchronize()
        lib.hprmat_hip_prof_enable(0)
        import ctypes
        gms, gfl, gcnt = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        lib.hprmat_hip_prof_read(ctypes.byref(gms), ctypes.byref(gfl), ctypes.byref(gcnt))
        ms = e0.elapsed_time(e1) / args.steps
        # verification pass (outside the timed region): same inputs, with statistics
        b.copy_(b0)
        lu.factor(a_cm, fm, equilibrate=True)
        st = lu.solve_(b, max_refine, 1e-12, 'rel_fro', True, stats=True, refinement=refinement)
        _, s_m = rmatrix.r_to_s(dmesh, b, nrhs, sysm.qk, open_idx, co, cop)
        if m == 'fp64' or (m == 'fp64_op' and x_ref is None):
            x_ref, s_ref = b.clone(), s_m.clone()
        methods[m] = {
            'ms': ms, 'solves_per_s': 1000.0 / ms, 'tflops': fl / (ms * 1e-3) / 1e12,
            'res_rel': st['res_rel'], 'sweeps': st['sweeps'], 'converged': st['converged'],
            'info': st['info'],
            'gemm': {'ms_per_launch': gms.value / max(gcnt.value, 1),
                     'launches_per_step': gcnt.value // args.steps,
                     'tflops': gfl.value / (gms.value * 1e-3) / 1e12 if gms.value > 0 else 0.0,
                     'flops_per_launch': gfl.value / max(gcnt.value, 1)},
        }
        methods[m]['hist'] = st['hist']
        if x_ref is not None:
            methods[m]['err_x_vs_fp64'] = float((b - x_ref).abs().max() / x_ref.abs().max())
            methods[m]['dS_vs_fp64'] = float((s_m - s_ref).abs().max())
        print('# %s: %.1f ms/solve, %.2f solves/s, %.1f TFLOP/s, res_rel %.2e, sweeps %d, '
              'trailing GEMM %.1f TFLOP/s, hist %s' % (
                  m, ms, 1000.0 / ms, methods[m]['tflops'], st['res_rel'], st['sweeps'],
                  methods[m]['gemm']['tflops'], ' '.join('%.1e' % h for h in st['hist'])),
              file=sys.stderr, flush=True)
        del lu

    dense = {m: v for m, v in methods.items() if not m.endswith('_op')} or methods
    op_methods = {m: v for m, v in methods.items() if m.endswith('_op')}
    mixed = [m for m in dense if m != 'fp64' and methods[m]['res_rel'] <= 1e-12]
    # nopen actually drawn (the workload text quotes about 100)
    nopen = len(open_idx)
    head = min(mixed, key=lambda m: methods[m]['ms']) if mixed else min(dense, key=lambda m: methods[m]['ms'])
    hm = methods[head]
    g = hm['gemm']
    roof = {'bound': 'mfma', 'achieved': g['tflops'], 'peak': PEAK[head],
            'unit': 'TFLOP/s', 'frac': g['tflops'] / PEAK[head], 'traffic': None,
            'kernel': '%s (trailing update A22 -= L21 U12, %s)' % (
                'zgemm_bf16x3_kernel<64, 64, 32, 32>' if head.startswith('bf16x3') else
                'zgemm_bf16x6_kernel<128, 64, 64, 32>' if head.startswith('bf16x6') else
                'trailing_update_kernel<%s, 64, 64, 16, 32, 32, %s>' % (
                    'double' if head == 'fp64' else 'float',
                    'true' if head.startswith('fp32_3m') else 'false'), head),
            'flops_per_launch': g['flops_per_launch'], 'ms_per_launch': g['ms_per_launch']}
    tr = traffic_record('fp64_factor' if head == 'fp64' else
                        'fp32_3m_factor' if head.startswith('fp32_3m') else 'fp32')
    if head.startswith('bf16'):
        roof['peak_note'] = 'BF16 dense matrix peak %.1f TFLOP/s / %d BF16 products per real product' % (
            BF16_PEAK, 3 if head.startswith('bf16x3') else 6)
    if head.startswith('fp32_3m'):
        roof['peak_note'] = ('Gauss three-product complex GEMM: the matrix cores execute 6 of the 8 '
                             'algorithmic flops of a complex multiply-add; executed rate = 0.75 x '
                             'achieved = %.1f TFLOP/s (%.2f of the FP32 matrix peak)' % (
                                 0.75 * g['tflops'], 0.75 * g['tflops'] / PEAK[head]))
    if tr is not None and not head.startswith('bf16') and 'launches' in tr:
        # mean over the trailing updates of one factorization, like 'achieved'
        roof['traffic'] = tr['hbm_bytes']
        roof['traffic_launch'] = {
            'launches': tr['launches'], 'hbm_read_bytes': tr['hbm_read_bytes'],
            'hbm_write_bytes': tr['hbm_write_bytes'], 'alg_bytes': tr['alg_bytes'],
            'read_ratio': tr['read_ratio'], 'write_ratio': tr['write_ratio'],
            'source': 'profiles/trailing_gemm_counters_N25600.md (FETCH_SIZE, WRITE_SIZE of '
                      '%s over the %d trailing updates of one N = %d factorization, '
                      'copy-calibrated; mean per launch)' % (tr['kernel'], tr['launches'],
                                                             tr['N'])}
    elif tr is not None and not head.startswith('bf16'):
        roof['traffic'] = tr['hbm_bytes']
        roof['traffic_launch'] = {
            'M': tr['M'], 'N': tr['N'], 'K': tr['K'], 'hbm_read_bytes': tr['hbm_read_bytes'],
            'hbm_write_bytes': tr['hbm_write_bytes'], 'alg_bytes': tr['alg_bytes'],
            'source': 'profiles/trailing_gemm_counters_N25600.md (FETCH_SIZE, WRITE_SIZE, '
                      'copy-calibrated; one GEMM of the first update\'s shape)'}
    out = {
        'metric': METRIC, 'value': hm['solves_per_s'], 'unit': 'solves/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': hm['ms'],
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': ('c128' if head == 'fp64' else
                  'c64 LU (split-bf16 update) + c128 refinement' if head.startswith('bf16')
                  else 'c64 LU + c128 refinement'),
        'data': 'synthetic',
        'config': {'workload': 'single Lagrange-mesh solve N=%d (nch=%d x nr=%d), nrhs=%d, '
                               'gate ||r||_F/||b||_F<=1e-12' % (n, args.nch, args.nr, nrhs),
                   'N': n, 'nrhs': nrhs, 'nopen': nopen, 'method': head, 'flops_per_solve': fl},
        'tflops': hm['tflops'],
        'methods': methods,
        'speedup_mixed_vs_fp64': (methods['fp64']['ms'] / hm['ms']) if 'fp64' in methods else None,
        'roofline': roof,
    }
    if op_methods:
        # the same solves with the refinement's FP64 products formed from the
        # Lagrange-mesh parts of the matrix (not the headline)
        blk = {m: {'ms': v['ms'], 'solves_per_s': v['solves_per_s'], 'res_rel': v['res_rel'],
                   'sweeps': v['sweeps']} for m, v in op_methods.items()}
        if 'fp32_3m_gmres_op' in op_methods and 'fp64_op' in op_methods:
            blk['speedup_mixed_vs_fp64'] = (op_methods['fp64_op']['ms'] /
                                            op_methods['fp32_3m_gmres_op']['ms'])
        out['mesh_operator'] = blk
    if not args.no_scan:
        try:
            out['energy_scan'] = energy_scan(torch, args)
        except Exception as exc:     # the headline line is still printed
            out['energy_scan'] = {'error': repr(exc)}
            print('# scan failed: %r' % (exc,), file=sys.stderr, flush=True)
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline()
    print(json.dumps(out))


def energy_scan(torch, args):
    """Batched energy scan: 64 waves (nch in {2..64}, nr in {40..100}) x 256
    energies in [0.5, 10]; each wave's energies solved as uniform batches
    (complex-FP32 LU + FP64 refinement, assembly and R->S on the device),
    lanes of streams, captured in HIP graphs; systems left unconverged are
    re-solved in FP64 inside the timed region.  Against a loop of plain
    per-system FP64 solves timed on every 16th energy of every wave."""
    from hprmat_amd import scan
    t0 = time.perf_counter()
    waves = scan.make_waves(args.scan_waves, args.scan_energies, seed=0)
    t_prep = time.perf_counter() - t0
    sc = scan.BatchScan(waves, method=args.scan_method, lanes=args.scan_lanes,
                        max_refine=args.scan_refine,
                        log=sys.stderr)
    t0 = time.perf_counter()
    if args.scan_graph:
        sc.capture()
    else:
        sc.run()
        sc.finish()
    t_cap = time.perf_counter() - t0
    print('# scan: %d systems, %.3g flops, prep %.1f s, warm-up/capture %.1f s, lanes %d'
          % (sc.n_systems, sc.flops, t_prep, t_cap, len(sc.lanes)), file=sys.stderr, flush=True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    print('# scan: timed run', file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    sc.run()
    n_redo = sc.finish()
    ms = (time.perf_counter() - t0) * 1e3
    bad = sum(int((w.status != 1).sum()) for w in waves)
    worst = max(float(w.res.max()) for w in waves)
    s_scan = {(w.index, k): w.s_of(k).clone() for w in waves for k in range(0, len(w.energies), 16)}
    pairs = [(w, k) for w in waves for k in range(0, len(w.energies), 16)]
    torch.cuda.synchronize()
    print('# scan: serial FP64 sample', file=sys.stderr, flush=True)
    e0.record()
    scan.serial_fp64(waves, pairs)
    e1.record()
    torch.cuda.synchronize()
    ms_ser = e0.elapsed_time(e1)
    dss = [(float((w.s_of(k) - s_scan[(w.index, k)]).abs().max()) if w.nopen[k] else 0.0, w, k)
           for w, k in pairs]
    ds, ww, wk = max(dss, key=lambda t: t[0])
    worst_sys = {'wave': ww.index, 'energy_index': int(wk), 'energy': float(ww.energies[wk]),
                 'N': ww.n, 'nch': ww.nch, 'nr': ww.nr, 'nopen': int(ww.nopen[wk]),
                 'res_rel_scan': float(ww.res[wk]), 'dS': ds}
    unconv = [{'wave': w.index, 'energy_index': int(k), 'N': w.n, 'res_rel': float(w.res[k])}
              for w in waves for k in torch.nonzero(w.status != 1).flatten().tolist()]
    n_sys = sc.n_systems
    out = {'systems': n_sys, 'ms': ms, 'solves_per_s': n_sys / (ms * 1e-3),
           'tflops': sc.flops / (ms * 1e-3) / 1e12, 'flops': sc.flops, 'lanes': len(sc.lanes),
           'graphs': bool(args.scan_graph), 'max_refine': args.scan_refine,
           'method': args.scan_method,
           'fp64_fallbacks': n_redo, 'unconverged': bad, 'max_res_rel': worst,
           'serial_fp64': {'systems_timed': len(pairs), 'ms': ms_ser,
                           'solves_per_s': len(pairs) / (ms_ser * 1e-3),
                           'sample': 'every 16th energy of every wave, one stream, eager'},
           'max_dS_vs_fp64': ds, 'worst_dS_system': worst_sys, 'unconverged_systems': unconv[:16]}
    out['speedup_vs_serial_fp64'] = out['solves_per_s'] / out['serial_fp64']['solves_per_s']
    print('# scan: %.1f ms, %.0f solves/s, %.1f TFLOP/s; serial FP64 %.0f solves/s; x%.2f; '
          'max|dS| %.1e; fallbacks %d; max res %.1e' % (
              ms, out['solves_per_s'], out['tflops'], out['serial_fp64']['solves_per_s'],
              out['speedup_vs_serial_fp64'], ds, n_redo, worst), file=sys.stderr, flush=True)
    return out


if __name__ == '__main__':
    main()
