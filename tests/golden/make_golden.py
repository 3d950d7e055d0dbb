"""Extract the golden phase shifts and wave functions of the reference's
example outputs.

Reads (as text) rmat_pierre/Ex2/output2.txt, rmat_pierre/Ex2/wf2.txt and
rmat_pierre/Ex5/output5.txt from the HPRMAT source tree and writes
tests/golden/rmat_pierre.json with the inputs of each run and the printed
results (5 significant digits, es12.4): the one-interval (ns = 1) runs, the
multi-interval Ex2 runs (ns = 2 and 3) and the internal wave functions of the
ns = 3 Ex2 run (the rmax < 0 entry of rmat_pierre/Ex2/data2).  These are Descouvemont's original outputs,
which HPRMAT reproduces (README.md:113-121).

Usage: python tests/golden/make_golden.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'rmat_pierre.json')
NUM = r'[-+]?\d+\.\d+E[-+]\d+'


def blocks(path):
    """Split an output file into runs: header dict + result lines."""
    runs, cur = [], None
    for line in open(path):
        if 'angular momentum=' in line:
            cur = {'j': int(line.split('=')[1]), 'lines': []}
            runs.append(cur)
        elif 'basis functions per interval=' in line:
            cur['nr'] = int(line.split('=')[1])
        elif 'number of intervals=' in line:
            cur['ns'] = int(line.split('=')[1])
        elif 'channel radius=' in line:
            cur['rmax'] = float(line.split('=')[1])
        elif line.startswith('E (MeV)='):
            cur['lines'].append(line.strip())
    return runs


def ex2(multi=False):
    out = []
    for run in blocks(os.path.join(REF, 'rmat_pierre/Ex2/output2.txt')):
        if (run['ns'] != 1) != multi:
            continue
        for line in run['lines']:
            e = float(re.search(r'E \(MeV\)=\s*([\d.]+)', line).group(1))
            d1, d2, eta12 = [float(x) for x in re.findall(NUM, line)]
            rec = dict(j=run['j'], nr=run['nr'], rmax=run['rmax'], e=e,
                       delta1_rad=d1, delta2_rad=d2, eta12=eta12)
            if multi:
                rec['ns'] = run['ns']
            out.append(rec)
    return out


def ex2_wf():
    """Internal wave functions (r, channel 1, channel 2) per energy of
    rmat_pierre/Ex2/wf2.txt: j = 1, nr = 25, ns = 3, rmax = 8, entrance
    channel 1."""
    out, cur, internal = [], None, False
    for line in open(os.path.join(REF, 'rmat_pierre/Ex2/wf2.txt')):
        if line.startswith('Wave function, E='):
            e = float(re.search(r'E=\s*([\d.]+)', line).group(1))
            cur = dict(j=1, nr=25, ns=3, rmax=8.0, e=e, entrance=1, r=[], wf=[])
            out.append(cur)
        elif 'Internal wave functions' in line:
            internal = True
        elif 'External wave functions' in line:
            internal = False
        elif internal and line.strip():
            v = [float(x) for x in line.split()]
            cur['r'].append(v[0])
            cur['wf'].append([v[1], v[2], v[3], v[4]])
    return out


def ex5():
    out = []
    for run in blocks(os.path.join(REF, 'rmat_pierre/Ex5/output5.txt')):
        if run['ns'] != 1:
            continue
        for line in run['lines']:
            e = float(re.search(r'E \(MeV\)=\s*([\d.]+)', line).group(1))
            delta, amp = [float(x) for x in re.findall(NUM, line)]
            out.append(dict(nr=run['nr'], rmax=run['rmax'], e=e, delta_deg=delta, modulus=amp))
    return out


if __name__ == '__main__':
    data = {
        'source': 'rmat_pierre/Ex2/output2.txt, rmat_pierre/Ex2/wf2.txt, '
                  'rmat_pierre/Ex5/output5.txt (es12.4)',
        'ex2_reid_nn': ex2(),
        'ex2_reid_nn_intervals': ex2(multi=True),
        'ex2_wave_functions': ex2_wf(),
        'ex5_yamaguchi': ex5(),
    }
    with open(OUT, 'w') as f:
        json.dump(data, f, indent=1)
    print('wrote', OUT, len(data['ex2_reid_nn']), len(data['ex5_yamaguchi']))
