"""Build libhprmat_hip.so (gfx950 only) in-tree with hipcc.

Usage: ``python -m hprmat_amd.build`` or ``hprmat_amd.build.build_library()``.
Every .hip file under csrc/ is compiled to an object in parallel, then linked
into ``hprmat_amd/libhprmat_hip.so``.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libhprmat_hip.so')
ARCH = 'gfx950'

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
CFLAGS = [
    '-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH,
    # packed FP32 VALU forms are disabled for all device code
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-I', os.path.join(ROOT, 'include'), '-I', CSRC,
    '-Wno-unused-result', '-Wno-unused-value',
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + '.o')
    cmd = [HIPCC] + CFLAGS + ['-c', src, '-o', obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = '\n'.join(l for l in (p.stdout + p.stderr).splitlines()
                    if 'packed-fp32-ops' not in l)
    if p.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (src, out))
    return obj, out


def build_library(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    for _, out in results:
        if verbose and out.strip():
            print(out)
    tmp = LIB + '.tmp%d' % os.getpid()
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + [o for o, _ in results]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError('link failed:\n' + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build_library(verbose='-v' in sys.argv))
