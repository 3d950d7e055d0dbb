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

Usage: python -m hprmat_amd.tracesum TRACE.csv [--kernel zgemm_kernel<float] [--min-wg 1000]
       python -m hprmat_amd.tracesum --counters COUNTERS.csv [MORE.csv ...] [--copy-bytes N]
"""
import argparse
import collections
import csv
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


PEAK_TFLOPS = {'F32': 157.3, 'F64': 78.6, 'BF16': 2516.6}   # MI355X dense matrix peaks
MOPS = ('F32', 'F64', 'BF16')


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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('trace', nargs='+')
    ap.add_argument('--kernel', default='zgemm_kernel<float')
    ap.add_argument('--min-wg', type=int, default=1000)
    ap.add_argument('--top', type=int, default=16)
    ap.add_argument('--counters', action='store_true')
    ap.add_argument('--copy-bytes', type=int, default=1 << 30)
    args = ap.parse_args(argv)
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
