"""Summary of a rocprofv3 kernel-trace CSV (``--kernel-trace --output-format csv``).

Prints, as markdown, the kernels by total time (calls, total, average, min,
max, share) and, for the launches whose name contains ``--kernel`` and whose
grid has at least ``--min-wg`` workgroups, the same figures per HIP queue:
with the look-ahead factorization the bulk trailing updates are the large
launches on the second queue, so this isolates the launches that bench.py
times with HIP events for its roofline record.

Usage: python -m hprmat_amd.tracesum TRACE.csv [--kernel zgemm_kernel<float] [--min-wg 1000]
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('trace')
    ap.add_argument('--kernel', default='zgemm_kernel<float')
    ap.add_argument('--min-wg', type=int, default=1000)
    ap.add_argument('--top', type=int, default=16)
    args = ap.parse_args(argv)
    rows = load(args.trace)
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
