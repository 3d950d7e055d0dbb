// Split-BF16 three-product complex GEMM for the FP32 LU trailing update.
//
// Each FP32 operand x is split once, before the GEMM, into two BF16 numbers:
//   hi = bf16(x)   (round to nearest even),   lo = bf16(x - hi)
// (x - hi is exact in FP32).  A real product is then formed from three
// BF16 products accumulated in FP32 on the matrix cores,
//   x y ~ hi_x hi_y + hi_x lo_y + lo_x hi_y,
// dropping lo_x lo_y (relative size 2^-16).  The complex product uses four
// real products (Ar Br - Ai Bi, Ar Bi + Ai Br), so one complex multiply-add
// costs twelve BF16 MFMA products (v_mfma_f32_16x16x32_bf16, K = 32 per
// instruction) instead of four FP32 ones (v_mfma_f32_16x16x4_f32).
//
// Operand planes: an R x K operand is stored as four BF16 planes
// (p = 0 re_hi, 1 re_lo, 2 im_hi, 3 im_lo), each R rows of Kp elements with
// K contiguous (Kp = K rounded up to 32, zero padded); A is split from its
// row-major storage, B (K x N row-major) is split transposed so that both
// operands are read along K.
#pragma once
#include "common.hpp"

typedef unsigned short bfraw;
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ inline bfraw bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
  return (bfraw)(u >> 16);
}
__device__ inline float bf16_val(bfraw h) { return __uint_as_float((unsigned)h << 16); }

// Split a rows x K block (TRANS = false: element (r, k) at S[r * ld + k];
// TRANS = true: at S[k * ld + r]) into 2 PIECES planes at D (plane stride
// `plane`, row length Kp): planes 0 .. PIECES-1 hold the successive BF16
// pieces of the real parts (piece q = bf16 of what pieces 0 .. q-1 left),
// planes PIECES .. 2 PIECES-1 those of the imaginary parts.  PIECES = 2 is
// the hi/lo split above; PIECES = 3 (hi, mid, lo: 24 significant bits)
// represents an FP32 value to within its own rounding.  32 x 32 tiles
// through LDS, coalesced on both sides.  grid = (Kp / 32, ceil(rows / 32)),
// block 256.
template <bool TRANS, int PIECES = 2>
__global__ void __launch_bounds__(256)
split_bf16_kernel(int rows, int K, const cx<float>* __restrict__ S, int64_t ld,
                  bfraw* __restrict__ D, size_t plane, int Kp) {
  __shared__ cx<float> tile[32][33];   // [r][k]
  const int k0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int q = ty; q < 32; q += 8) {
    // TRANS: read along r (contiguous in S); else along k
    const int r = TRANS ? r0 + tx : r0 + q, k = TRANS ? k0 + q : k0 + tx;
    cx<float> v = mkc<float>(0.f, 0.f);
    if (r < rows && k < K) v = TRANS ? S[(int64_t)k * ld + r] : S[(int64_t)r * ld + k];
    if (TRANS) tile[tx][q] = v; else tile[q][tx] = v;
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const int r = r0 + q, k = k0 + tx;
    if (r >= rows) continue;
    const cx<float> v = tile[q][tx];
    const size_t o = (size_t)r * Kp + k;
    float xr = v.x, xi = v.y;   // what the pieces so far leave (exact in FP32)
#pragma unroll
    for (int p = 0; p < PIECES; ++p) {
      const bfraw hr = bf16_rne(xr), hi = bf16_rne(xi);
      D[p * plane + o] = hr;
      D[(PIECES + p) * plane + o] = hi;
      xr -= bf16_val(hr);
      xi -= bf16_val(hi);
    }
  }
}

__device__ inline bf16x8_t as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8_t, v); }
__device__ inline uint4 neg_bf16x8(uint4 v) {
  return make_uint4(v.x ^ 0x80008000u, v.y ^ 0x80008000u, v.z ^ 0x80008000u, v.w ^ 0x80008000u);
}

template <int BM, int BN>
constexpr size_t zgemm_bf16x3_lds_bytes() {
  return (size_t)2 * 4 * (BM + BN) * (32 + 8) * sizeof(bfraw);
}

// C = beta C + alpha A B over split planes (A: M x Kp planes, B: N x Kp
// planes), C row-major FP32 complex.  One workgroup of 256 threads computes a
// BM x BN tile, each of the four waves a WM x WN sub-tile of 16 x 16 blocks.
// LDS double-buffered ([buf][plane][row][32 + 8 pad]), one barrier per
// K-slice of 32.  Tile order as zgemm_kernel (XCD-aware, grouped).
template <int BM, int BN, int WM, int WN>
__global__ void __launch_bounds__(256)
zgemm_bf16x3_kernel(int M, int N, int Kp, const bfraw* __restrict__ Ap, size_t pa,
                    const bfraw* __restrict__ Bp, size_t pb, cx<float>* __restrict__ C,
                    int64_t ldc, float alpha, int beta) {
  constexpr int BK = 32, NT = 256, RS = BK + 8;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int LA = BM * BK / 8 / NT, LB = BN * BK / 8 / NT;   // uint4 per thread per plane
  static_assert((BM / WM) * (BN / WN) == 4 && LA >= 1 && LB >= 1, "tile shape");
  typedef float v4f __attribute__((ext_vector_type(4)));
  extern __shared__ __align__(16) unsigned char bsm[];
  bfraw* As = reinterpret_cast<bfraw*>(bsm);      // [2][4][BM][RS]
  bfraw* Bs = As + 2 * 4 * BM * RS;                // [2][4][BN][RS]

  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const int total = tiles_m * tiles_n;
  int t = blockIdx.x;
  if ((total & 7) == 0) t = (t & 7) * (total >> 3) + (t >> 3);
  constexpr int GROUP = 8;
  const int group = t / (GROUP * tiles_n);
  const int first_m = group * GROUP;
  const int gsz = min(tiles_m - first_m, GROUP);
  const int tm = first_m + (t % (GROUP * tiles_n)) % gsz;
  const int tn = (t % (GROUP * tiles_n)) / gsz;
  const int row0 = tm * BM, col0 = tn * BN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave / (BN / WN), wc = wave % (BN / WN);
  const int li = lane & 15, lg = lane >> 4;

  v4f acc_re[FM][FN], acc_im[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      acc_re[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
      acc_im[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    }

  uint4 ra[4][LA], rb[4][LB];
  auto load = [&](int k0) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
#pragma unroll
      for (int e = 0; e < LA; ++e) {
        const int idx = tid + e * NT, r = idx >> 2, kq = (idx & 3) * 8, gr = row0 + r;
        ra[p][e] = gr < M ? *reinterpret_cast<const uint4*>(Ap + p * pa + (size_t)gr * Kp + k0 + kq)
                          : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for (int e = 0; e < LB; ++e) {
        const int idx = tid + e * NT, r = idx >> 2, kq = (idx & 3) * 8, gr = col0 + r;
        rb[p][e] = gr < N ? *reinterpret_cast<const uint4*>(Bp + p * pb + (size_t)gr * Kp + k0 + kq)
                          : make_uint4(0, 0, 0, 0);
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
#pragma unroll
      for (int e = 0; e < LA; ++e) {
        const int idx = tid + e * NT, r = idx >> 2, kq = (idx & 3) * 8;
        *reinterpret_cast<uint4*>(As + ((size_t)(buf * 4 + p) * BM + r) * RS + kq) = ra[p][e];
      }
#pragma unroll
      for (int e = 0; e < LB; ++e) {
        const int idx = tid + e * NT, r = idx >> 2, kq = (idx & 3) * 8;
        *reinterpret_cast<uint4*>(Bs + ((size_t)(buf * 4 + p) * BN + r) * RS + kq) = rb[p][e];
      }
    }
  };

  load(0);
  store(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < Kp; k0 += BK) {
    const bool more = k0 + BK < Kp;
    if (more) load(k0 + BK);
    uint4 fa[4][FM], fb[4][FN];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
#pragma unroll
      for (int a = 0; a < FM; ++a)
        fa[p][a] = *reinterpret_cast<const uint4*>(
            As + ((size_t)(buf * 4 + p) * BM + wr * WM + a * 16 + li) * RS + lg * 8);
#pragma unroll
      for (int b = 0; b < FN; ++b)
        fb[p][b] = *reinterpret_cast<const uint4*>(
            Bs + ((size_t)(buf * 4 + p) * BN + wc * WN + b * 16 + li) * RS + lg * 8);
    }
    uint4 nih[FM], nil[FM];
#pragma unroll
    for (int a = 0; a < FM; ++a) {
      nih[a] = neg_bf16x8(fa[2][a]);
      nil[a] = neg_bf16x8(fa[3][a]);
    }
    // six terms per accumulator, each pass over all (a, b) blocks so that
    // dependent MFMAs are separated by the independent ones
#define HPR_BF3(ACC, XA, XB)                                                         \
  _Pragma("unroll") for (int a = 0; a < FM; ++a)                                     \
  _Pragma("unroll") for (int b = 0; b < FN; ++b)                                     \
      ACC[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(XA), as_bf16x8(XB), \
                                                          ACC[a][b], 0, 0, 0);
    HPR_BF3(acc_re, fa[0][a], fb[0][b])   // re_hi re_hi
    HPR_BF3(acc_im, fa[0][a], fb[2][b])   // re_hi im_hi
    HPR_BF3(acc_re, fa[0][a], fb[1][b])   // re_hi re_lo
    HPR_BF3(acc_im, fa[0][a], fb[3][b])   // re_hi im_lo
    HPR_BF3(acc_re, fa[1][a], fb[0][b])   // re_lo re_hi
    HPR_BF3(acc_im, fa[1][a], fb[2][b])   // re_lo im_hi
    HPR_BF3(acc_re, nih[a], fb[2][b])     // -im_hi im_hi
    HPR_BF3(acc_im, fa[2][a], fb[0][b])   // im_hi re_hi
    HPR_BF3(acc_re, nih[a], fb[3][b])     // -im_hi im_lo
    HPR_BF3(acc_im, fa[2][a], fb[1][b])   // im_hi re_lo
    HPR_BF3(acc_re, nil[a], fb[2][b])     // -im_lo im_hi
    HPR_BF3(acc_im, fa[3][a], fb[0][b])   // im_lo re_hi
#undef HPR_BF3
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = row0 + wr * WM + a * 16 + mfma_row<float>(lg, r);
        const int gj = col0 + wc * WN + b * 16 + li;
        if (gi < M && gj < N) {
          cx<float>* pc = C + (int64_t)gi * ldc + gj;
          const cx<float> c0 = beta ? *pc : mkc<float>(0.f, 0.f);
          *pc = mkc<float>(c0.x + alpha * acc_re[a][b][r], c0.y + alpha * acc_im[a][b][r]);
        }
      }
}

// ---------------------------------------------------------------------------
// Three-piece split-BF16 complex GEMM ("bf16x6"): operands split into hi,
// mid and lo BF16 pieces (split_bf16_kernel<., 3>, six planes: re hi/mid/lo,
// im hi/mid/lo); a real product keeps the six piece products of weight
// >= 2^-16 relative,
//   x y ~ h h + h m + m h + h l + m m + l h,
// accumulated in FP32 on v_mfma_f32_16x16x32_bf16.  The dropped terms
// (m l, l m, l l) are below 2^-24 relative, so the product is as accurate as
// an FP32 one (componentwise |error| <= ~0.5 x the FP32 GEMM's on random
// operands), at 24 BF16 MFMA products per complex multiply-add.
//
// Tile: BM x BN = 128 x 64 per workgroup (four waves of 64 x 32), BK = 32 (one
// MFMA K-depth); the six planes of both operands are staged through a single
// LDS buffer of padded rows (row stride 40 BF16 = 80 bytes, so the 16 rows a
// fragment read touches fall on distinct banks) with the next K-slice
// prefetched into registers during the current one (192 MFMAs per wave per
// slice).  Per A piece, the MFMAs of all its B partners run over all 4 x 2
// (a, b) blocks before the next partner, so dependent MFMAs on one
// accumulator are eight instructions apart.  C += alpha A B (beta = 1).
// Measured at N = 25600: about the FP32 kernel's speed (one wave per SIMD,
// 280 registers and 92 KB of LDS per workgroup leave the load latency
// exposed; a double-buffered 147 KB variant was slower still).
// ---------------------------------------------------------------------------
template <int BM, int BN>
constexpr size_t zgemm_bf16x6_lds_bytes() {
  return (size_t)6 * (BM + BN) * (32 + 8) * sizeof(bfraw);
}

template <int BM, int BN, int WM, int WN>
__global__ void __launch_bounds__(256)
zgemm_bf16x6_kernel(int M, int N, int Kp, const bfraw* __restrict__ Ap, size_t pa,
                    const bfraw* __restrict__ Bp, size_t pb, cx<float>* __restrict__ C,
                    int64_t ldc, float alpha) {
  constexpr int BK = 32, NT = 256, RS = BK + 8, NP = 6;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int LA = BM * BK / 8 / NT, LB = BN * BK / 8 / NT;   // uint4 per thread per plane
  static_assert((BM / WM) * (BN / WN) == 4 && LA >= 1 && LB >= 1, "tile shape");
  static_assert(LA * NT * 8 == BM * BK && LB * NT * 8 == BN * BK, "load split");
  typedef float v4f __attribute__((ext_vector_type(4)));
  extern __shared__ __align__(16) unsigned char b6sm[];
  bfraw* As = reinterpret_cast<bfraw*>(b6sm);      // [NP][BM][RS]
  bfraw* Bs = As + NP * BM * RS;                    // [NP][BN][RS]

  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const int total = tiles_m * tiles_n;
  int t = blockIdx.x;
  if ((total & 7) == 0) t = (t & 7) * (total >> 3) + (t >> 3);
  constexpr int GROUP = 8;
  const int group = t / (GROUP * tiles_n);
  const int first_m = group * GROUP;
  const int gsz = min(tiles_m - first_m, GROUP);
  const int tm = first_m + (t % (GROUP * tiles_n)) % gsz;
  const int tn = (t % (GROUP * tiles_n)) / gsz;
  const int row0 = tm * BM, col0 = tn * BN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave / (BN / WN), wc = wave % (BN / WN);
  const int li = lane & 15, lg = lane >> 4;

  v4f acc_re[FM][FN], acc_im[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      acc_re[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
      acc_im[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    }

  uint4 ra[NP][LA], rb[NP][LB];
  auto load = [&](int k0) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
      for (int e = 0; e < LA; ++e) {
        const int idx = tid + e * NT, r = idx >> 2, kq = (idx & 3) * 8, gr = row0 + r;
        ra[p][e] = gr < M ? *reinterpret_cast<const uint4*>(Ap + p * pa + (size_t)gr * Kp + k0 + kq)
                          : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for (int e = 0; e < LB; ++e) {
        const int idx = tid + e * NT, r = idx >> 2, kq = (idx & 3) * 8, gr = col0 + r;
        rb[p][e] = gr < N ? *reinterpret_cast<const uint4*>(Bp + p * pb + (size_t)gr * Kp + k0 + kq)
                          : make_uint4(0, 0, 0, 0);
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
      for (int e = 0; e < LA; ++e) {
        const int idx = tid + e * NT, r = idx >> 2, kq = (idx & 3) * 8;
        *reinterpret_cast<uint4*>(As + ((size_t)p * BM + r) * RS + kq) = ra[p][e];
      }
#pragma unroll
      for (int e = 0; e < LB; ++e) {
        const int idx = tid + e * NT, r = idx >> 2, kq = (idx & 3) * 8;
        *reinterpret_cast<uint4*>(Bs + ((size_t)p * BN + r) * RS + kq) = rb[p][e];
      }
    }
  };

  load(0);
  for (int k0 = 0; k0 < Kp; k0 += BK) {
    __syncthreads();   // the previous slice's fragment reads are done
    store();
    __syncthreads();
    if (k0 + BK < Kp) load(k0 + BK);
    // every B fragment of the slice (6 planes x FN), then one A piece at a time
    uint4 fb[NP][FN];
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int b = 0; b < FN; ++b)
        fb[p][b] = *reinterpret_cast<const uint4*>(
            Bs + ((size_t)p * BN + wc * WN + b * 16 + li) * RS + lg * 8);
#define HPR_BF6(ACC, XA, XB)                                                          \
  _Pragma("unroll") for (int a = 0; a < FM; ++a)                                      \
  _Pragma("unroll") for (int b = 0; b < FN; ++b)                                      \
      ACC[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(XA), as_bf16x8(XB), \
                                                          ACC[a][b], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {   // piece of A: 0 hi, 1 mid, 2 lo
      uint4 fr[FM], fi[FM], nfi[FM];
#pragma unroll
      for (int a = 0; a < FM; ++a) {
        const int r = wr * WM + a * 16 + li;
        fr[a] = *reinterpret_cast<const uint4*>(As + ((size_t)i * BM + r) * RS + lg * 8);
        fi[a] = *reinterpret_cast<const uint4*>(As + ((size_t)(3 + i) * BM + r) * RS + lg * 8);
        nfi[a] = neg_bf16x8(fi[a]);
      }
#pragma unroll
      for (int j = 0; j + i < 3; ++j) {   // partner piece of B
        HPR_BF6(acc_re, fr[a], fb[j][b])        // Ar_i Br_j
        HPR_BF6(acc_im, fr[a], fb[3 + j][b])    // Ar_i Bi_j
        HPR_BF6(acc_re, nfi[a], fb[3 + j][b])   // -Ai_i Bi_j
        HPR_BF6(acc_im, fi[a], fb[j][b])        // Ai_i Br_j
      }
    }
#undef HPR_BF6
  }

#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = row0 + wr * WM + a * 16 + mfma_row<float>(lg, r);
        const int gj = col0 + wc * WN + b * 16 + li;
        if (gi < M && gj < N) {
          cx<float>* pc = C + (int64_t)gi * ldc + gj;
          const cx<float> c0 = *pc;
          *pc = mkc<float>(c0.x + alpha * acc_re[a][b][r], c0.y + alpha * acc_im[a][b][r]);
        }
      }
}

