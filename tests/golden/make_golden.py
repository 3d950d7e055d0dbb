"""Extract the golden phase shifts of the reference's example outputs.

Reads (as text) rmat_pierre/Ex2/output2.txt and rmat_pierre/Ex5/output5.txt
from the HPRMAT source tree and writes tests/golden/rmat_pierre.json with the
inputs of each one-interval (ns = 1) run and the printed results (5
significant digits, es12.4).  These are Descouvemont's original outputs,
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


def ex2():
    out = []
    for run in blocks(os.path.join(REF, 'rmat_pierre/Ex2/output2.txt')):
        if run['ns'] != 1:
            continue
        for line in run['lines']:
            e = float(re.search(r'E \(MeV\)=\s*([\d.]+)', line).group(1))
            d1, d2, eta12 = [float(x) for x in re.findall(NUM, line)]
            out.append(dict(j=run['j'], nr=run['nr'], rmax=run['rmax'], e=e,
                            delta1_rad=d1, delta2_rad=d2, eta12=eta12))
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
        'source': 'rmat_pierre/Ex2/output2.txt, rmat_pierre/Ex5/output5.txt (es12.4)',
        'ex2_reid_nn': ex2(),
        'ex5_yamaguchi': ex5(),
    }
    with open(OUT, 'w') as f:
        json.dump(data, f, indent=1)
    print('wrote', OUT, len(data['ex2_reid_nn']), len(data['ex5_yamaguchi']))
