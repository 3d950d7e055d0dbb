"""Single-kernel driver for hardware-counter profiles of the dominant kernel.

Runs, on cuda:0, one device-to-device copy of a known size (a kernel whose
HBM traffic is known: the calibration for FETCH_SIZE / WRITE_SIZE) and then
the first trailing update of the N = 25600 FP32 LU, A22 -= L21 U12 with
M = N = 25344, K = 256, through the library's complex GEMM (the same kernel
and tile as getrf's update).  Prints the algorithmic bytes of both.

Usage: python -m hprmat_amd.kprobe [--m 25344] [--k 256] [--copy-bytes 1073741824]
       [--precision 32]
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


if __name__ == '__main__':
    sys.exit(main())
