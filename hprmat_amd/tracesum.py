"""Summary of a rocprofv3 kernel-trace CSV (``--kernel-trace --output-format csv``).

Prints, as markdown, the kernels by total time (calls, total, average, min,
max, share) and, for the launches whose name contains ``--kernel`` and whose
grid has at least ``--min-wg`` workgroups, the same figures per HIP queue:
with the look-ahead factorization the bulk trailing updates are the large
launches on the second queue, so this isolates the launches that bench.py
times with HIP events for its roofline record.

With --counters the file is a rocprofv3 counter-collection CSV (``--pmc``)
instead: per kernel, the dispatch count and time and, from the counters it
holds, MFMA throughput (SQ_INSTS_VALU_MFMA_MOPS_* x 512 operations per
second) against the MI355X dense matrix peaks, the share of MFMA-busy cycles
(SQ_VALU_MFMA_BUSY_CYCLES / SQ_BUSY_CU_CYCLES), the clock seen by the kernel
(GRBM_GUI_ACTIVE / time / 8 XCDs), and HBM traffic (FETCH_SIZE, WRITE_SIZE)
scaled by what the same counter reads for the device-to-device copy in the
file (the largest ``__amd_rocclr_copyBuffer``, of --copy-bytes).  Several
counter files (one pass each) are merged per kernel.

With --traffic N the counter files (a FETCH_SIZE pass and a WRITE_SIZE pass)
hold the trailing-update launches of one blocked factorization of order N
(block --nb): the j-th launch of --kernel updates the nc x ncols block with
nc = N - (j + 1) nb and ncols = nc, or nc - nb on a look-ahead step (its grid
of --tile x --tile tiles tells which; 64 for the FP32/FP64 kernels, 128 for
the split-precision ones), and its algorithmic bytes are one read of L21
(nc x nb), U12 (nb x ncols) and C (nc x ncols) and one write of C, at
--elem-bytes per complex element (8: FP32 complex, or the four 16-bit planes
of the two-piece splits).  The
summary compares the copy-calibrated HBM bytes with those, per launch and in
total (--json writes the same figures).

Usage: python -m hprmat_amd.tracesum TRACE.csv [--kernel zgemm_kernel<float] [--min-wg 1000]
       python -m hprmat_amd.tracesum --counters COUNTERS.csv [MORE.csv ...] [--copy-bytes N]
       python -m hprmat_amd.tracesum --traffic 25600 FETCH.csv WRITE.csv --kernel K [--tile 64]
                                     [--json OUT]
"""
import argparse
import collections
import csv
import json
import sys


def load(path):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            s, e = int(r['Start_Timestamp']), int(r['End_Timestamp'])
            wg = max(1, int(r['Workgroup_Size_X']) * int(r['Workgroup_Size_Y'])
                     * int(r['Workgroup_Size_Z']))
            grid = int(r['Grid_Size_X']) * int(r['Grid_Size_Y']) * int(r['Grid_Size_Z'])
            name = r['Kernel_Name'].replace('(anonymous namespace)::', '')
            rows.append({'name': name.split('(')[0].replace('void ', ''),
                         'queue': r['Queue_Id'], 'us': (e - s) / 1e3, 'wgs': grid // wg})
    return rows


def stats(us):
    return len(us), sum(us), sum(us) / len(us), min(us), max(us)


PEAK_TFLOPS = {'F32': 157.3, 'F64': 78.6, 'BF16': 2516.6, 'F16': 2516.6}   # MI355X dense peaks
MOPS = ('F32', 'F64', 'BF16', 'F16')


def counters(path, copy_bytes):
    """Per-kernel sums of a counter-collection CSV (one row per dispatch and counter)."""
    disp = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            d = disp.setdefault(r['Dispatch_Id'], {
                'name': r['Kernel_Name'].replace('(anonymous namespace)::', '')
                        .split('(')[0].replace('void ', ''),
                'ns': int(r['End_Timestamp']) - int(r['Start_Timestamp'])})
            d[r['Counter_Name']] = float(r['Counter_Value'])
    copies = [d for d in disp.values() if 'copyBuffer' in d['name']]
    scale = {}
    for c in ('FETCH_SIZE', 'WRITE_SIZE'):
        vals = [d[c] for d in copies if c in d]
        if vals:
            scale[c] = copy_bytes / 1024.0 / max(vals)
    agg = collections.OrderedDict()
    for d in disp.values():
        a = agg.setdefault(d['name'], collections.defaultdict(float))
        a['n'] += 1
        for k, v in d.items():
            if k != 'name':
                a[k] += v
    return agg, scale


def merged_counters(paths, copy_bytes):
    """Per-kernel sums over several counter files (one counter pass each):
    counters are added, dispatch counts and times taken from the first file
    that has the kernel, HBM counters scaled by each file's own copy."""
    agg = collections.OrderedDict()
    scales = {}
    for path in paths:
        part, scale = counters(path, copy_bytes)
        scales.update(scale)
        for name, a in part.items():
            dst = agg.setdefault(name, collections.defaultdict(float))
            if 'n' not in dst:
                dst['n'], dst['ns'] = a['n'], a['ns']
            for k, v in a.items():
                if k in ('n', 'ns'):
                    continue
                dst[k] += v * scale.get(k, 1.0)
    return agg, scales


def print_counters(paths, copy_bytes):
    agg, scale = merged_counters(paths, copy_bytes)
    has = lambda key: any(key in a for a in agg.values())
    if has('SQ_VALU_MFMA_BUSY_CYCLES') or any(has('SQ_INSTS_VALU_MFMA_MOPS_' + m) for m in MOPS):
        print('| kernel | dispatches | ms | MFMA TFLOP/s (type) | of peak | MFMA busy | clock GHz |')
        print('|---|---:|---:|---:|---:|---:|---:|')
        for name, a in agg.items():
            if 'copyBuffer' in name:
                continue
            sec = a['ns'] * 1e-9
            cells = []
            for m in MOPS:
                ops = a.get('SQ_INSTS_VALU_MFMA_MOPS_' + m, 0.0) * 512
                if ops > 0:
                    tf = ops / sec / 1e12
                    cells.append((tf, m, tf / PEAK_TFLOPS[m]))
            tfs = ', '.join('%.1f (%s)' % (t, m) for t, m, _ in cells) or '-'
            frac = ', '.join('%.2f' % f for _, _, f in cells) or '-'
            # SQ_BUSY_CU_CYCLES counts quad-cycles
            busy = (a['SQ_VALU_MFMA_BUSY_CYCLES'] / (4.0 * a['SQ_BUSY_CU_CYCLES'])
                    if a.get('SQ_BUSY_CU_CYCLES') else float('nan'))
            ghz = a.get('GRBM_GUI_ACTIVE', 0.0) / 8 / sec / 1e9 if sec > 0 else float('nan')
            print('| `%s` | %d | %.2f | %s | %s | %.2f | %.2f |'
                  % (name[:80], a['n'], a['ns'] / 1e6, tfs, frac, busy, ghz))
    if has('FETCH_SIZE') or has('WRITE_SIZE'):
        print()
        print('HBM scale factors from the copy: %s' % ', '.join(
            '%s x%.3f' % kv for kv in sorted(scale.items())))
        print()
        print('| kernel | dispatches | ms | HBM read GB | HBM write GB | GB/s |')
        print('|---|---:|---:|---:|---:|---:|')
        for name, a in agg.items():
            if 'copyBuffer' in name:
                continue
            rd = a.get('FETCH_SIZE', 0.0) * 1024 / 1e9
            wr = a.get('WRITE_SIZE', 0.0) * 1024 / 1e9
            sec = a['ns'] * 1e-9
            print('| `%s` | %d | %.2f | %.3f | %.3f | %.0f |'
                  % (name[:80], a['n'], a['ns'] / 1e6, rd, wr, (rd + wr) / sec if sec else 0))


def dispatches(path):
    """Dispatches of a counter-collection CSV in submission order."""
    disp = collections.OrderedDict()
    with open(path) as f:
        for r in csv.DictReader(f):
            d = disp.setdefault(int(r['Dispatch_Id']), {
                'name': r['Kernel_Name'].replace('(anonymous namespace)::', '')
                        .split('(')[0].replace('void ', ''),
                'wgs': int(r['Grid_Size']) // max(1, int(r['Workgroup_Size'])),
                'ns': int(r['End_Timestamp']) - int(r['Start_Timestamp'])})
            d[r['Counter_Name']] = float(r['Counter_Value'])
    return [disp[k] for k in sorted(disp)]


def factor_traffic(paths, copy_bytes, kernel, n, nb, elem, tile=64):
    """Copy-calibrated HBM bytes against algorithmic bytes for the trailing
    updates of one factorization (one counter per file).

    One launch per step is the update of every column right of the panel
    (nc x nc, or nc x (nc - nb) under look-ahead).  A kernel that also
    updates the next panel's nb columns on the main stream shows a strip
    launch (nc x nb) before the bulk one of the same step: the strips are
    left out, so that the record covers the launches the bench times."""
    cdiv = lambda a, b: (a + b - 1) // b
    per = None
    keep = None
    scales = {}
    for path in paths:
        ds = dispatches(path)
        copies = [d for d in ds if 'copyBuffer' in d['name']]
        sel = [d for d in ds if d['name'].startswith(kernel)]
        if per is None:
            per, keep = [], []
            j, i = 0, 0
            while i < len(sel):
                d = sel[i]
                nc = n - (j + 1) * nb
                if nc <= 0:
                    raise ValueError('more launches of %s than steps of N=%d' % (kernel, n))
                grid = lambda c: cdiv(nc, tile) * cdiv(c, tile)
                if (nc > nb and d['wgs'] == grid(nb) and i + 1 < len(sel)
                        and sel[i + 1]['wgs'] == grid(nc - nb)):
                    i += 1          # strip of the next panel's columns
                    continue
                shapes = [c for c in (nc, nc - nb) if c > 0 and grid(c) == d['wgs']]
                if not shapes:
                    raise ValueError('launch %d: %d workgroups fit no update of %d rows'
                                     % (i, d['wgs'], nc))
                ncols = shapes[0]
                per.append({'M': nc, 'N': ncols, 'K': nb, 'ns': d['ns'],
                            'alg_read': elem * (nc * nb + nb * ncols + nc * ncols),
                            'alg_write': elem * nc * ncols})
                keep.append(i)
                i += 1
                j += 1
            nsel = len(sel)
        elif len(sel) != nsel:
            raise ValueError('%s: %d launches, %d in the first file' % (path, len(sel), nsel))
        sel = [sel[i] for i in keep]
        for c in ('FETCH_SIZE', 'WRITE_SIZE'):
            vals = [d[c] for d in copies if c in d]
            if not vals or not any(c in d for d in sel):
                continue
            scales[c] = copy_bytes / 1024.0 / max(vals)
            for p, d in zip(per, sel):
                p[c] = d[c] * scales[c] * 1024.0
    if not per:
        raise ValueError('no launches of %s' % kernel)
    for c in ('FETCH_SIZE', 'WRITE_SIZE'):
        if c not in scales:
            raise ValueError('no %s pass with a calibration copy' % c)
    tot = lambda key: sum(p[key] for p in per)
    cnt = len(per)
    rd, wr, ard, awr = tot('FETCH_SIZE'), tot('WRITE_SIZE'), tot('alg_read'), tot('alg_write')
    first = per[0]
    return {
        'kernel': kernel, 'N': n, 'nb': nb, 'launches': cnt,
        'fetch_scale': scales['FETCH_SIZE'], 'write_scale': scales['WRITE_SIZE'],
        'hbm_read_bytes': rd / cnt, 'hbm_write_bytes': wr / cnt, 'hbm_bytes': (rd + wr) / cnt,
        'alg_read_bytes': ard / cnt, 'alg_write_bytes': awr / cnt, 'alg_bytes': (ard + awr) / cnt,
        'read_ratio': rd / ard, 'write_ratio': wr / awr,
        'kernel_ms': tot('ns') / 1e6,
        'first_launch': {'M': first['M'], 'N': first['N'], 'K': nb,
                         'hbm_read_bytes': first['FETCH_SIZE'],
                         'hbm_write_bytes': first['WRITE_SIZE'],
                         'alg_read_bytes': first['alg_read'], 'alg_write_bytes': first['alg_write'],
                         'read_ratio': first['FETCH_SIZE'] / first['alg_read'],
                         'write_ratio': first['WRITE_SIZE'] / first['alg_write']},
    }


def print_traffic(rec):
    f = rec['first_launch']
    print('HBM scale factors from the copy: FETCH_SIZE x%.3f, WRITE_SIZE x%.3f'
          % (rec['fetch_scale'], rec['write_scale']))
    print()
    print('| `%s`, N = %d | launches | HBM read GB | HBM write GB | alg read GB | '
          'alg write GB | read / alg | write / alg |' % (rec['kernel'], rec['N']))
    print('|---|---:|---:|---:|---:|---:|---:|---:|')
    c = rec['launches']
    print('| all launches (sum) | %d | %.3f | %.3f | %.3f | %.3f | %.2f | %.2f |'
          % (c, c * rec['hbm_read_bytes'] / 1e9, c * rec['hbm_write_bytes'] / 1e9,
             c * rec['alg_read_bytes'] / 1e9, c * rec['alg_write_bytes'] / 1e9,
             rec['read_ratio'], rec['write_ratio']))
    print('| per launch (mean) | 1 | %.3f | %.3f | %.3f | %.3f | %.2f | %.2f |'
          % (rec['hbm_read_bytes'] / 1e9, rec['hbm_write_bytes'] / 1e9,
             rec['alg_read_bytes'] / 1e9, rec['alg_write_bytes'] / 1e9,
             rec['read_ratio'], rec['write_ratio']))
    print('| first launch %d x %d x %d | 1 | %.3f | %.3f | %.3f | %.3f | %.2f | %.2f |'
          % (f['M'], f['N'], f['K'], f['hbm_read_bytes'] / 1e9, f['hbm_write_bytes'] / 1e9,
             f['alg_read_bytes'] / 1e9, f['alg_write_bytes'] / 1e9, f['read_ratio'],
             f['write_ratio']))


def kernel_source_hash(kernel=''):
    """First 16 hex digits of the SHA-256 of the source of a trailing-update
    GEMM (csrc/split_gemm.hpp for the split-precision kernels, csrc/gemm.hpp
    otherwise): stamped into traffic records so that a reader can tell
    whether a record was measured on the kernel it is used for."""
    import hashlib
    import os
    src = 'split_gemm.hpp' if 'split' in kernel else 'gemm.hpp'
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'csrc', src)
    with open(path, 'rb') as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('trace', nargs='+')
    ap.add_argument('--kernel', default='zgemm_kernel<float')
    ap.add_argument('--min-wg', type=int, default=1000)
    ap.add_argument('--top', type=int, default=16)
    ap.add_argument('--counters', action='store_true')
    ap.add_argument('--copy-bytes', type=int, default=1 << 30)
    ap.add_argument('--traffic', type=int, default=0, metavar='N')
    ap.add_argument('--nb', type=int, default=256)
    ap.add_argument('--elem-bytes', type=int, default=8)
    ap.add_argument('--tile', type=int, default=64)
    ap.add_argument('--json', default='')
    args = ap.parse_args(argv)
    if args.traffic:
        rec = factor_traffic(args.trace, args.copy_bytes, args.kernel, args.traffic, args.nb,
                             args.elem_bytes, args.tile)
        rec['build'] = kernel_source_hash(args.kernel)
        print_traffic(rec)
        if args.json:
            with open(args.json, 'w') as f:
                json.dump(rec, f, indent=1)
        return 0
    if args.counters:
        print_counters(args.trace, args.copy_bytes)
        return 0
    rows = load(args.trace[0])
    total = sum(r['us'] for r in rows)
    by = collections.defaultdict(list)
    for r in rows:
        by[r['name']].append(r['us'])
    print('| kernel | calls | total ms | avg us | min us | max us | % |')
    print('|---|---:|---:|---:|---:|---:|---:|')
    for name, us in sorted(by.items(), key=lambda t: -sum(t[1]))[:args.top]:
        n, t, a, lo, hi = stats(us)
        print('| `%s` | %d | %.2f | %.2f | %.2f | %.2f | %.2f |'
              % (name[:90], n, t / 1e3, a, lo, hi, 100.0 * t / total))
    sel = [r for r in rows if args.kernel in r['name'] and r['wgs'] >= args.min_wg]
    print()
    print('Launches of `%s` with >= %d workgroups, per queue:' % (args.kernel, args.min_wg))
    print()
    print('| queue | calls | total ms | avg us | min us | max us |')
    print('|---|---:|---:|---:|---:|---:|')
    byq = collections.defaultdict(list)
    for r in sel:
        byq[r['queue']].append(r['us'])
    for q, us in sorted(byq.items()):
        n, t, a, lo, hi = stats(us)
        print('| %s | %d | %.2f | %.2f | %.2f | %.2f |' % (q, n, t / 1e3, a, lo, hi))
    return 0


if __name__ == '__main__':
    sys.exit(main())
