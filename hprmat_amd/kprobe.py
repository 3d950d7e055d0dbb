"""Single-kernel driver for hardware-counter profiles of the dominant kernel.

Runs, on cuda:0, one device-to-device copy of a known size (a kernel whose
HBM traffic is known: the calibration for FETCH_SIZE / WRITE_SIZE) and then
the first trailing update of the N = 25600 FP32 LU, A22 -= L21 U12 with
M = N = 25344, K = 256, through the library's complex GEMM (the same kernel
and tile as getrf's update).  Prints the algorithmic bytes of both.

With --solve METHODS (comma-separated: fp32, fp64, bf16x3, fp32_3m) it then
builds the N = 25600 Lagrange-mesh system on the device (nch = 128 x nr =
200, seed 0) and, per method, factors it once (equilibrated, look-ahead) and
runs one substitution of its 128 right-hand sides without refinement,
printing a line after each step: the dispatches of one factorization and one
solve, for counter passes restricted to a few kernels by
rocprofv3's --kernel-include-regex.

Usage: python -m hprmat_amd.kprobe [--m 25344] [--k 256] [--copy-bytes 1073741824]
       [--precision 32] [--solve fp32,fp64,bf16x3]
"""
import argparse
import ctypes
import json
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--m', type=int, default=25344)
    ap.add_argument('--k', type=int, default=256)
    ap.add_argument('--precision', type=int, default=32, choices=(32, 64))
    ap.add_argument('--copy-bytes', type=int, default=1 << 30)
    ap.add_argument('--solve', default='')
    args = ap.parse_args()
    import torch
    from . import _lib, solver
    lib = _lib.load()
    torch.cuda.set_device(0)
    src = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print('copy: %d bytes read, %d bytes written' % (args.copy_bytes, args.copy_bytes), flush=True)
    del src, dst
    ct = torch.complex64 if args.precision == 32 else torch.complex128
    m, k = args.m, args.k
    g = torch.Generator(device='cuda').manual_seed(1)
    a = torch.randn(m, k, generator=g, device='cuda', dtype=ct)
    b = torch.randn(k, m, generator=g, device='cuda', dtype=ct)
    c = torch.randn(m, m, generator=g, device='cuda', dtype=ct)
    torch.cuda.synchronize()
    rc = lib.hprmat_hip_zgemm(args.precision, m, m, k, ctypes.c_void_p(a.data_ptr()), k, 0,
                              ctypes.c_void_p(b.data_ptr()), m, ctypes.c_void_p(c.data_ptr()), m,
                              -1, 1, solver._stream_ptr())
    _lib.check(rc, 'hprmat_hip_zgemm')
    torch.cuda.synchronize()
    eb = 8 if args.precision == 32 else 16
    alg = {'M': m, 'N': m, 'K': k, 'read_bytes': (m * k + k * m + m * m) * eb,
           'write_bytes': m * m * eb, 'flops': 8.0 * m * m * k}
    print('gemm: ' + json.dumps(alg), flush=True)
    if args.solve:
        del a, b, c
        solve_pass(args.solve.split(','), torch)


def solve_pass(methods, torch):
    """One factorization + one unrefined substitution per method at N = 25600."""
    from . import rmatrix, solver, synth
    sysm = synth.System(128, 200, 15.0, 8.0, 5.0, 0)
    dmesh = rmatrix.DeviceMesh(sysm.mesh)
    _, _, _, xc1 = rmatrix.channel_data(sysm.lval, sysm.qk, sysm.a)
    a_cm = rmatrix.assemble(dmesh, sysm.lval, sysm.qk, xc1, sysm.cpot)
    b0 = rmatrix.rhs_colmajor(dmesh, sysm.nch)
    torch.cuda.synchronize()
    print('system: N=%d nrhs=%d' % (sysm.n, sysm.nch), flush=True)
    for m in methods:
        lu = solver.LU().factor(a_cm, m, equilibrate=True)
        torch.cuda.synchronize()
        print('factor %s done' % m, flush=True)
        b = b0.clone()
        lu.solve_(b, 0, 0.0, 'rel_fro', False, stats=False)
        torch.cuda.synchronize()
        print('solve %s done' % m, flush=True)
        del lu


if __name__ == '__main__':
    sys.exit(main())
