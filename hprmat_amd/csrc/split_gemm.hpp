// Split-precision complex GEMM on 32x32x16 matrix cores for the FP32 LU
// trailing update C -= L21 U12.
//
// Each FP32 operand is split once per panel, before the GEMM, into P pieces
// of a 16-bit matrix-core type; a real product x y is then formed from the
// piece products x_i y_j with i + j < P, accumulated in FP32:
//
//   SPLIT_F16X2  P = 2 FP16 pieces (hi = fp16(s x), lo = fp16(s x - hi)):
//                22 significant bits; products hh + hl + lh (the dropped
//                lo lo term is below 2^-22 relative).  FP16 has a narrow
//                exponent range, so the operands are scaled by powers of
//                two (exact, v_ldexp) in two stages.  First the K index:
//                column k of L21 is divided by 2^e_k, the binade of its
//                largest |Re|, |Im|, and row k of U12 multiplied by it (the
//                product is unchanged; without this a diagonal row scaling
//                D A of the matrix, which puts the factor D_k into both L's
//                column k (inversely) and U's row k, would leave terms of
//                equal weight at magnitudes far apart inside one row of U12).
//                Then every row of L21 and every column of U12 by 2^k that
//                brings its largest |Re|, |Im| into [2^14, 2^15); the product
//                is scaled back by 2^-(k_row + k_col) in the epilogue.  The
//                exponents are kept as integers: they span more than an FP32
//                factor 2^k can hold.  Three products per real product.
//   SPLIT_BF16X3 P = 3 BF16 pieces (hi, mid, lo: 24 significant bits, no
//                scaling: BF16 has FP32's exponent range); products hh hm mh
//                hl mm lh (dropped terms below 2^-24 relative).  Six
//                products per real product.
//   SPLIT_BF16X2 P = 2 BF16 pieces (hi, lo: 16 significant bits); products
//                hh hl lh (accurate to about 2^-16).  Three products.
//
// A complex multiply-add uses four real products (Ar Br - Ai Bi,
// Ar Bi + Ai Br); the minus sign is applied by flipping the sign bits of the
// Ai pieces in registers.
//
// Packed operand layout (written by split_rows_kernel / split_cols_kernel,
// read by zgemm_split_kernel): an R x 256 operand (K = 256 = one LU panel,
// zero padded past the panel's width) is stored row by row; a row holds 8
// K-slices of 32, and a slice holds the 2P planes (real pieces 0 .. P-1,
// then imaginary pieces) of 32 consecutive K, so one row's slice is
// 2P x 64 contiguous bytes:
//   half index ((r * 8 + kb) * 2P + plane) * 32 + kk.
// Scales (SPLIT_F16X2 only): kexp[r] = k of row r (int32); the K scales
// come from colmax[k] = bits of max_i max(|Re L_ik|, |Im L_ik|)
// (split_colmax_kernel).
#pragma once
#include "common.hpp"

namespace hpr {

enum SplitKind { SPLIT_F16X2 = 0, SPLIT_BF16X3 = 1, SPLIT_BF16X2 = 2 };

template <int KIND> struct SplitTraits;
template <> struct SplitTraits<SPLIT_F16X2> {
  static constexpr int P = 2;          // pieces per real number
  static constexpr bool SCALED = true;
};
template <> struct SplitTraits<SPLIT_BF16X3> {
  static constexpr int P = 3;
  static constexpr bool SCALED = false;
};
template <> struct SplitTraits<SPLIT_BF16X2> {
  static constexpr int P = 2;
  static constexpr bool SCALED = false;
};

// halves per packed row (K = 256)
template <int KIND> constexpr int split_row_halves() { return 256 * 2 * SplitTraits<KIND>::P; }

typedef unsigned short half_bits;

__device__ inline half_bits f16_bits_rne(float x) {
  const _Float16 h = (_Float16)x;   // v_cvt_f16_f32: round to nearest even
  return __builtin_bit_cast(half_bits, h);
}
__device__ inline float f16_value(half_bits b) {
  return (float)__builtin_bit_cast(_Float16, b);
}
__device__ inline half_bits bf16_bits_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
  return (half_bits)(u >> 16);
}
__device__ inline float bf16_value(half_bits h) { return __uint_as_float((unsigned)h << 16); }

// Exponent k of the power-of-two scale 2^k bringing m = max |component| of
// a row into [2^14, 2^15) (FP32 subnormal m included); 0 for a zero (or
// non-finite) maximum.
__device__ inline int split_scale_exp(float m) {
  if (!(m > 0.f) || !(m < INFINITY)) return 0;
  int e;
  (void)frexpf(m, &e);   // m = f 2^e, f in [0.5, 1)
  return 15 - e;
}

// Pieces of one scaled real value into out[0 .. P-1].
template <int KIND>
__device__ inline void split_value(float x, half_bits* out) {
  if constexpr (KIND == SPLIT_F16X2) {
    const half_bits h = f16_bits_rne(x);
    out[0] = h;
    out[1] = f16_bits_rne(x - f16_value(h));   // x - hi is exact in FP32
  } else {
#pragma unroll
    for (int p = 0; p < SplitTraits<KIND>::P; ++p) {
      const half_bits h = bf16_bits_rne(x);
      out[p] = h;
      x -= bf16_value(h);
    }
  }
}

// Writes the pieces of four consecutive complex values v[0..3] of packed row
// `r` at K index k (k % 4 == 0), value e scaled by 2^ks[e].
template <int KIND>
__device__ inline void split_store4(half_bits* __restrict__ D, int r, int k, const cx<float>* v,
                                    const int* ks) {
  constexpr int P = SplitTraits<KIND>::P, NP = 2 * P;
  half_bits pr[4][P], pi[4][P];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    split_value<KIND>(ldexpf(v[e].x, ks[e]), pr[e]);
    split_value<KIND>(ldexpf(v[e].y, ks[e]), pi[e]);
  }
  half_bits* row = D + ((size_t)r * 8 + (k >> 5)) * NP * 32 + (k & 31);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    uint2 wr, wi;
    wr.x = (unsigned)pr[0][p] | ((unsigned)pr[1][p] << 16);
    wr.y = (unsigned)pr[2][p] | ((unsigned)pr[3][p] << 16);
    wi.x = (unsigned)pi[0][p] | ((unsigned)pi[1][p] << 16);
    wi.y = (unsigned)pi[2][p] | ((unsigned)pi[3][p] << 16);
    *reinterpret_cast<uint2*>(row + p * 32) = wr;
    *reinterpret_cast<uint2*>(row + (P + p) * 32) = wi;
  }
}

// Binade e_k of the K scale from colmax bits (0 for a zero column).
__device__ inline int split_kexp(const unsigned* colmax, int k) {
  const float m = __uint_as_float(colmax[k]);
  if (!(m > 0.f) || !(m < INFINITY)) return 0;
  int e;
  (void)frexpf(m, &e);
  return e;
}

// Column maxima of a row-major FP32 complex block (element (r, k) at
// S[r * ld + k], r < rows, k < K <= 256): colmax[k] = bits of the max over
// the rows of max(|Re S_rk|, |Im S_rk|) (non-negative floats order like
// their bit patterns), by atomic maxima into colmax, which the caller
// zeroes first.  Workgroup g (1024 threads: column k = t % 256, row phase
// q = t / 256) takes 64 rows, thread (k, q) rows q, q + 4, ... of them, 16
// loads in flight; the four phases combine in LDS, then one atomic per
// column.  The max is independent of the order of the atomics.
// grid = ceil(rows / 64), block 1024.
__global__ void __launch_bounds__(1024)
split_colmax_kernel(int rows, int K, const cx<float>* __restrict__ S, int64_t ld,
                    unsigned* __restrict__ colmax) {
  __shared__ float red[4][256];
  const int k = threadIdx.x & 255, q = threadIdx.x >> 8;
  const int r0 = blockIdx.x * 64;
  float m = 0.f;
  if (k < K) {
#pragma unroll
    for (int i = q; i < 64; i += 4) {
      if (r0 + i < rows) {
        const cx<float> v = S[(int64_t)(r0 + i) * ld + k];
        m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
      }
    }
  }
  red[q][k] = m;
  __syncthreads();
  if (q == 0 && k < K) {
    const float mk = fmaxf(fmaxf(red[0][k], red[1][k]), fmaxf(red[2][k], red[3][k]));
    atomicMax(colmax + k, __float_as_uint(mk));
  }
}

// Split `rows` rows of a row-major FP32 complex block S (element (r, k) at
// S[r * ld + k], k < K <= 256) into packed row r of D (+ kexp[r] when
// scaled, the K scales 2^-e_k applied first).  One wave per row, lane l
// owning K indices 4l .. 4l + 3 (zero past K).  grid = ceil(rows / 4),
// block 256.
template <int KIND>
__global__ void __launch_bounds__(256)
split_rows_kernel(int rows, int K, const cx<float>* __restrict__ S, int64_t ld,
                  half_bits* __restrict__ D, int* __restrict__ kexp,
                  const unsigned* __restrict__ colmax) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= rows) return;
  cx<float> v[4];
  int ks[4] = {0, 0, 0, 0};
  float m = 0.f;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int k = 4 * lane + e;
    v[e] = k < K ? S[(int64_t)r * ld + k] : mkc<float>(0.f, 0.f);
    if constexpr (SplitTraits<KIND>::SCALED) ks[e] = k < K ? -split_kexp(colmax, k) : 0;
    m = fmaxf(m, ldexpf(fmaxf(fabsf(v[e].x), fabsf(v[e].y)), ks[e]));
  }
  if constexpr (SplitTraits<KIND>::SCALED) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    const int kr = split_scale_exp(m);
    if (lane == 0) kexp[r] = kr;
#pragma unroll
    for (int e = 0; e < 4; ++e) ks[e] += kr;
  }
  split_store4<KIND>(D, r, 4 * lane, v, ks);
}

// Split the columns [0, cols) of a row-major FP32 complex block S (element
// (k, j) at S[k * ld + j], k < K <= 256) into packed rows j of D (the
// transposed operand, + kexp[j] when scaled, the K scales 2^e_k applied
// first).  A workgroup transposes 32 columns x 256 K through LDS (separate
// real / imaginary tiles, 33-float rows); wave w then packs columns w,
// w + 4, ...  grid = ceil(cols / 32), block 256.
template <int KIND>
__global__ void __launch_bounds__(256)
split_cols_kernel(int cols, int K, const cx<float>* __restrict__ S, int64_t ld,
                  half_bits* __restrict__ D, int* __restrict__ kexp,
                  const unsigned* __restrict__ colmax) {
  __shared__ float tre[256][33], tim[256][33];
  const int j0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 256; k += 8) {
    cx<float> v = mkc<float>(0.f, 0.f);
    if (k < K && j0 + tx < cols) v = S[(int64_t)k * ld + j0 + tx];
    tre[k][tx] = v.x;
    tim[k][tx] = v.y;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int kk[4] = {0, 0, 0, 0};
  if constexpr (SplitTraits<KIND>::SCALED) {
#pragma unroll
    for (int e = 0; e < 4; ++e) kk[e] = 4 * lane + e < K ? split_kexp(colmax, 4 * lane + e) : 0;
  }
  for (int jj = w; jj < 32; jj += 4) {
    const int j = j0 + jj;
    if (j >= cols) break;   // uniform across the wave
    cx<float> v[4];
    int ks[4] = {kk[0], kk[1], kk[2], kk[3]};
    float m = 0.f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * lane + e;
      v[e] = mkc<float>(tre[k][jj], tim[k][jj]);
      m = fmaxf(m, ldexpf(fmaxf(fabsf(v[e].x), fabsf(v[e].y)), ks[e]));
    }
    if constexpr (SplitTraits<KIND>::SCALED) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
      const int kc = split_scale_exp(m);
      if (lane == 0) kexp[j] = kc;
#pragma unroll
      for (int e = 0; e < 4; ++e) ks[e] += kc;
    }
    split_store4<KIND>(D, j, 4 * lane, v, ks);
  }
}

typedef unsigned u32x4_v __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8_v __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_v __attribute__((ext_vector_type(8)));
typedef float f32x16_v __attribute__((ext_vector_type(16)));

template <int KIND>
__device__ inline f32x16_v mfma_split(u32x4_v a, u32x4_v b, f32x16_v c) {
  if constexpr (KIND == SPLIT_F16X2)
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8_v, a),
                                                  __builtin_bit_cast(f16x8_v, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_v, a),
                                                   __builtin_bit_cast(bf16x8_v, b), c, 0, 0, 0);
}

__device__ inline u32x4_v neg_half8(u32x4_v v) { return v ^ 0x80008000u; }

// LDS bytes of zgemm_split_kernel: two buffers of NP planes, each plane
// (BM + BN) rows of BK halves plus 64 bytes of skew.
template <int KIND, int BM, int BN, int BK>
constexpr size_t zgemm_split_lds_bytes() {
  return (size_t)2 * 2 * SplitTraits<KIND>::P * ((BM + BN) * BK * 2 + 64);
}

// C += alpha * diag(2^-ka) (A B) diag(2^-kb) over packed operands: A is
// M x 256 (rows of the left factor), B is N x 256 (rows = columns of the
// right factor), C row-major FP32 complex with leading dimension ldc.
// ka / kb are the scale exponents (SPLIT_F16X2; ignored otherwise).
//
// Workgroup: NTH threads (NTH / 64 waves) computing BM x BN of C; wave (wr, wc)
// owns WM x WN, as (WM / 32) x (WN / 32) blocks of 32 x 32 with separate real
// and imaginary FP32 accumulators.  K runs in slices of BK (16 or 32): the
// slice NSET ahead is loaded from global memory into registers while the
// current one is multiplied out of LDS, and the next one is stored to the
// other LDS buffer at the end of the slice; one barrier per slice.  MINB
// workgroups per CU are requested from the register allocator (2: at most
// 256 registers per lane, so that one workgroup's epilogue overlaps
// another's products).  LDS row = one plane's BK halves of a row, its
// 16-byte chunks XOR-swizzled by row so that the 16 rows a fragment read
// covers per 16 lanes hit distinct banks; planes are skewed by 64 bytes.
// Fragment of a 32x32x16 product: lane l holds row l % 32, K 8 (l / 32) ..
// + 7 of the current 16-K step; result register r of lane l holds row
// 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
// Tile order: XCD-aware when the tile count is a multiple of 8 (consecutive
// workgroups go to the 8 XCDs in turn, so each XCD receives a contiguous
// range), then grouped by 8 tile rows for L2 reuse of the B rows.
template <int KIND, int BM, int BN, int WM, int WN, int BK, int NSET, int MINB, int NTH = 256>
__global__ void __launch_bounds__(NTH, MINB)
zgemm_split_kernel(int M, int N, const half_bits* __restrict__ A, const int* __restrict__ ka,
                   const half_bits* __restrict__ B, const int* __restrict__ kb_,
                   cx<float>* __restrict__ C, int64_t ldc, float alpha) {
  constexpr int P = SplitTraits<KIND>::P, NP = 2 * P;
  constexpr int RH = split_row_halves<KIND>();           // halves per packed row
  constexpr int NT = NTH, FM = WM / 32, FN = WN / 32;
  constexpr int ROWS = BM + BN;
  constexpr int CHK = BK / 8;                             // 16-byte chunks per LDS row
  constexpr int RSW = 16 / CHK;                           // rows per swizzle step
  constexpr int PLANE = ROWS * BK + 32;                   // halves per LDS plane
  constexpr int RPP = NT / CHK;                           // rows per load pass
  constexpr int LE = (ROWS + RPP - 1) / RPP;              // load passes per plane
  constexpr int KS = 256 / BK;                            // slices
  static_assert((BM / WM) * (BN / WN) == NT / 64, "one wave per WM x WN block");
  static_assert(WM % 32 == 0 && WN % 32 == 0, "32 x 32 blocks");
  static_assert(BK == 16 || BK == 32, "slice of one or two 16-K steps");
  static_assert(NSET == 1 || NSET == 2, "register sets");
  extern __shared__ __align__(16) unsigned char split_smem[];
  half_bits* lds = reinterpret_cast<half_bits*>(split_smem);   // [2][NP][PLANE]

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
  const int l32 = lane & 31, lh = lane >> 5;

  // load slots: per plane, thread t moves 16-byte chunk t % CHK of rows
  // t / CHK + RPP e < ROWS (A rows 0 .. BM-1, then B rows; a partial last
  // pass leaves whole waves idle).  Rows past M / N are clamped to the last
  // row: they only feed C entries that are not stored.
  const int lc = tid % CHK, lr = tid / CHK;
  const int lsw = (lc ^ ((lr / RSW) % CHK)) * 8;   // RPP is a multiple of RSW CHK
  const half_bits* src[LE];
#pragma unroll
  for (int e = 0; e < LE; ++e) {
    const int rr = lr + RPP * e;
    src[e] = (rr < BM ? A + (size_t)min(row0 + rr, M - 1) * RH
                      : B + (size_t)min(col0 + rr - BM, N - 1) * RH) + lc * 8;
  }

  f32x16_v acc_re[FM][FN], acc_im[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) { acc_re[a][b][r] = 0.f; acc_im[a][b][r] = 0.f; }

  // register sets: slice kb + NSET is loaded while slice kb is multiplied
  // out of LDS (and, with two sets, slice kb + 1 waits in the other set)
  u32x4_v stg0[NP][LE], stg1[NP][LE];
  half_bits* const lds_st = lds + lr * BK + lsw;
  // K offset of slice KB inside a packed row (32-K groups of NP planes)
#define HPR_SPLIT_KOFF(KB) ((((KB) * BK) >> 5) * NP * 32 + (((KB) * BK) & 31))
#define HPR_SPLIT_LOAD(ST, KB)                                                          \
  _Pragma("unroll") for (int e = 0; e < LE; ++e)                                        \
  _Pragma("unroll") for (int p = 0; p < NP; ++p)                                        \
      if (ROWS % RPP == 0 || e + 1 < LE || lr + RPP * e < ROWS)                        \
        ST[p][e] = *reinterpret_cast<const u32x4_v*>(src[e] + HPR_SPLIT_KOFF(KB) + p * 32);
#define HPR_SPLIT_STORE(ST, BUF)                                                        \
  _Pragma("unroll") for (int p = 0; p < NP; ++p)                                        \
  _Pragma("unroll") for (int e = 0; e < LE; ++e)                                        \
      if (ROWS % RPP == 0 || e + 1 < LE || lr + RPP * e < ROWS)                        \
        *reinterpret_cast<u32x4_v*>(lds_st + ((BUF) * NP + p) * PLANE + RPP * e * BK) = ST[p][e];

  // With two staging sets, C is read during the last two slices into the
  // registers the sets free once their last slice is in LDS (set 0 one slice
  // before set 1: the first CPRE column blocks then), so that the epilogue
  // does not wait for it; with one set, at the epilogue.  Every C element of
  // the wave is read before any store (stores to C would otherwise order
  // each read after the previous write).
  constexpr bool CEARLY = NSET == 2 && KS >= 3;
  constexpr int CPRE = FN / 2;
  cx<float> cv[FN][FM][16];
#define HPR_SPLIT_CLOAD(B0, B1)                                                          \
  _Pragma("unroll") for (int b = (B0); b < (B1); ++b) {                                  \
    const int gj = col0 + wc * WN + b * 32 + l32;                                        \
    _Pragma("unroll") for (int a = 0; a < FM; ++a)                                       \
    _Pragma("unroll") for (int r = 0; r < 16; ++r) {                                     \
      const int gi = row0 + wr * WM + a * 32 + 8 * (r / 4) + 4 * lh + (r % 4);           \
      cv[b][a][r] = (gi < M && gj < N) ? C[(int64_t)gi * ldc + gj] : mkc<float>(0.f, 0.f); \
    }                                                                                    \
  }

  HPR_SPLIT_LOAD(stg0, 0)
  if constexpr (NSET == 2) { HPR_SPLIT_LOAD(stg1, 1) }
  HPR_SPLIT_STORE(stg0, 0)
  __syncthreads();
#pragma unroll
  for (int kb = 0; kb < KS; ++kb) {
    if constexpr (NSET == 2) {
      // the set of slice kb (already in LDS) takes slice kb + 2
      if (kb + 2 < KS) {
        if (kb & 1) { HPR_SPLIT_LOAD(stg1, kb + 2) }
        else { HPR_SPLIT_LOAD(stg0, kb + 2) }
      }
    } else {
      if (kb + 1 < KS) { HPR_SPLIT_LOAD(stg0, kb + 1) }
    }
    if (CEARLY && kb == KS - 2) { HPR_SPLIT_CLOAD(0, CPRE) }
    if (CEARLY && kb == KS - 1) { HPR_SPLIT_CLOAD(CPRE, FN) }
    // keep the loads here, ahead of the products (the scheduler would
    // otherwise sink them to the LDS stores)
    __builtin_amdgcn_sched_barrier(0);
    const half_bits* base = lds + (kb & 1) * NP * PLANE;
#pragma unroll
    for (int s = 0; s < BK / 16; ++s) {   // 16-K steps of the slice
      const int c = 2 * s + lh;
      u32x4_v fa[FM][NP], fb[FN][NP];
#pragma unroll
      for (int a = 0; a < FM; ++a) {
        const int rr = wr * WM + a * 32 + l32;
#pragma unroll
        for (int p = 0; p < NP; ++p)
          fa[a][p] = *reinterpret_cast<const u32x4_v*>(base + p * PLANE + rr * BK +
                                                     ((c ^ ((rr / RSW) % CHK)) * 8));
      }
#pragma unroll
      for (int b = 0; b < FN; ++b) {
        const int rr = BM + wc * WN + b * 32 + l32;
#pragma unroll
        for (int p = 0; p < NP; ++p)
          fb[b][p] = *reinterpret_cast<const u32x4_v*>(base + p * PLANE + rr * BK +
                                                     ((c ^ ((rr / RSW) % CHK)) * 8));
      }
      // piece products i + j < P; each pass runs over every (a, b) block so
      // that dependent products on one accumulator are FM FN apart
#pragma unroll
      for (int i = 0; i < P; ++i) {
        u32x4_v nai[FM];
#pragma unroll
        for (int a = 0; a < FM; ++a) nai[a] = neg_half8(fa[a][P + i]);
#pragma unroll
        for (int j = 0; i + j < P; ++j) {
#pragma unroll
          for (int a = 0; a < FM; ++a)
#pragma unroll
            for (int b = 0; b < FN; ++b) {
              acc_re[a][b] = mfma_split<KIND>(fa[a][i], fb[b][j], acc_re[a][b]);
              acc_im[a][b] = mfma_split<KIND>(fa[a][i], fb[b][P + j], acc_im[a][b]);
            }
#pragma unroll
          for (int a = 0; a < FM; ++a)
#pragma unroll
            for (int b = 0; b < FN; ++b) {
              acc_re[a][b] = mfma_split<KIND>(nai[a], fb[b][P + j], acc_re[a][b]);
              acc_im[a][b] = mfma_split<KIND>(fa[a][P + i], fb[b][j], acc_im[a][b]);
            }
        }
      }
    }
    if (kb + 1 < KS) {
      if constexpr (NSET == 2) {
        if (kb & 1) { HPR_SPLIT_STORE(stg0, 0) }
        else { HPR_SPLIT_STORE(stg1, 1) }
      } else {
        HPR_SPLIT_STORE(stg0, (kb + 1) & 1)
      }
      __syncthreads();
    }
  }

  // epilogue: C += the scaled product
  if (!CEARLY) { HPR_SPLIT_CLOAD(0, FN) }
  int sa[FM][16], sb[FN];   // scale exponents of the rows / columns
#pragma unroll
  for (int b = 0; b < FN; ++b) {
    sb[b] = 0;
    if constexpr (SplitTraits<KIND>::SCALED) sb[b] = kb_[min(col0 + wc * WN + b * 32 + l32, N - 1)];
  }
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int gi = row0 + wr * WM + a * 32 + 8 * (r / 4) + 4 * lh + (r % 4);
      sa[a][r] = 0;
      if constexpr (SplitTraits<KIND>::SCALED) sa[a][r] = ka[min(gi, M - 1)];
    }
#pragma unroll
  for (int b = 0; b < FN; ++b) {
    const int gj = col0 + wc * WN + b * 32 + l32;
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int gi = row0 + wr * WM + a * 32 + 8 * (r / 4) + 4 * lh + (r % 4);
        if (gi < M && gj < N) {
          float pr = alpha * acc_re[a][b][r], pi = alpha * acc_im[a][b][r];
          if constexpr (SplitTraits<KIND>::SCALED) {
            const int k = -(sb[b] + sa[a][r]);
            pr = ldexpf(pr, k);
            pi = ldexpf(pi, k);
          }
          C[(int64_t)gi * ldc + gj] = mkc<float>(cv[b][a][r].x + pr, cv[b][a][r].y + pi);
        }
      }
  }
#undef HPR_SPLIT_CLOAD
#undef HPR_SPLIT_KOFF
#undef HPR_SPLIT_LOAD
#undef HPR_SPLIT_STORE
}

}  // namespace hpr
