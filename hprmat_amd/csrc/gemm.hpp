// Complex GEMM on MFMA: C = beta * C + alpha * A * B, alpha = +-1, beta in {0, 1}.
//
// A is M x K (row-major, or column-major when A_COL), B is K x N row-major,
// C is M x N row-major.  One workgroup of 256 threads (4 waves) computes a
// BM x BN tile; each wave owns a WM x WN sub-tile made of 16 x 16 MFMA blocks
// with separate real/imaginary accumulators.  The complex product is formed
// from real 16x16x4 MFMAs, either as four products (Ar Br - Ai Bi,
// Ar Bi + Ai Br) or, with GAUSS3M, as Gauss's three products
// P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi) (Higham, SIAM J. Matrix
// Anal. Appl. 13 (1992)), which saves a quarter of the matrix-core work.
//
// Tiles of A and B are staged through LDS split into real and imaginary
// planes, k-major, so every MFMA operand read is a conflict-free ds_read of
// 16 consecutive elements.  The next K-slice is prefetched into registers
// while the current one is multiplied.  Workgroup ids are remapped so that
// the workgroups of one XCD (ids congruent mod 8) sweep a compact group of
// tiles and share A row-blocks and B column-blocks in that XCD's L2.
#pragma once
#include "common.hpp"

template <class T, int BM, int BN, int BK, int WM, int WN, bool A_COL, bool GAUSS3M>
__device__ __forceinline__ void
zgemm_tile(int M, int N, int K, const cx<T>* __restrict__ A, int64_t lda,
           const cx<T>* __restrict__ B, int64_t ldb, cx<T>* __restrict__ C,
           int64_t ldc, T alpha, int beta, int t_in) {
  static_assert((BM / WM) * (BN / WN) == 4, "four waves per workgroup");
  static_assert(WM % 16 == 0 && WN % 16 == 0 && BK % 4 == 0, "tile shape");
  constexpr int NT = 256;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int PA = BM + 1, PB = BN + 1;  // padded LDS row lengths
  constexpr int EA = BM * BK / NT, EB = BK * BN / NT;
  static_assert(EA * NT == BM * BK && EB * NT == BK * BN, "load split");
  typedef typename acc_t<T>::type acc;

  __shared__ T As[2][BK][PA];  // [re/im][k][i]
  __shared__ T Bs[2][BK][PB];  // [re/im][k][j]

  // XCD-aware tile order: workgroups b with equal b % 8 run on one XCD.
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const int total = tiles_m * tiles_n;
  int t = t_in;
  if ((total & 7) == 0) t = (t & 7) * (total >> 3) + (t >> 3);
  constexpr int GROUP = 8;
  const int group = t / (GROUP * tiles_n);
  const int first_m = group * GROUP;
  const int gsz = min(tiles_m - first_m, GROUP);
  const int tm = first_m + (t % (GROUP * tiles_n)) % gsz;
  const int tn = (t % (GROUP * tiles_n)) / gsz;
  const int row0 = tm * BM, col0 = tn * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / (BN / WN), wc = wave % (BN / WN);
  const int li = lane & 15, lk = lane >> 4;

  acc acc_re[FM][FN], acc_im[FM][FN], acc_p3[GAUSS3M ? FM : 1][GAUSS3M ? FN : 1];
  for (int a = 0; a < FM; ++a)
    for (int b = 0; b < FN; ++b) {
      acc_re[a][b] = acc{};
      acc_im[a][b] = acc{};
      if (GAUSS3M) acc_p3[a][b] = acc{};
    }

  cx<T> ra[EA], rb[EB];
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      int idx = tid + e * NT, i, k;
      if (A_COL) { i = idx % BM; k = idx / BM; } else { i = idx / BK; k = idx % BK; }
      int gi = row0 + i, gk = k0 + k;
      ra[e] = (gi < M && gk < K)
                  ? (A_COL ? A[(int64_t)gk * lda + gi] : A[(int64_t)gi * lda + gk])
                  : mkc<T>(0, 0);
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      int idx = tid + e * NT, k = idx / BN, j = idx % BN;
      int gk = k0 + k, gj = col0 + j;
      rb[e] = (gk < K && gj < N) ? B[(int64_t)gk * ldb + gj] : mkc<T>(0, 0);
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      int idx = tid + e * NT, i, k;
      if (A_COL) { i = idx % BM; k = idx / BM; } else { i = idx / BK; k = idx % BK; }
      As[0][k][i] = ra[e].x;
      As[1][k][i] = ra[e].y;
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      int idx = tid + e * NT, k = idx / BN, j = idx % BN;
      Bs[0][k][j] = rb[e].x;
      Bs[1][k][j] = rb[e].y;
    }
  };

  load_tiles(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (k0 + BK < K) load_tiles(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T ar[FM], ai[FM], br[FN], bi[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) {
        int i = wr * WM + a * 16 + li;
        ar[a] = As[0][kk + lk][i];
        ai[a] = As[1][kk + lk][i];
      }
#pragma unroll
      for (int b = 0; b < FN; ++b) {
        int j = wc * WN + b * 16 + li;
        br[b] = Bs[0][kk + lk][j];
        bi[b] = Bs[1][kk + lk][j];
      }
      if (GAUSS3M) {
        T as[FM], bs[FN];
#pragma unroll
        for (int a = 0; a < FM; ++a) as[a] = ar[a] + ai[a];
#pragma unroll
        for (int b = 0; b < FN; ++b) bs[b] = br[b] + bi[b];
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
          for (int b = 0; b < FN; ++b) {
            acc_re[a][b] = mfma16x4(ar[a], br[b], acc_re[a][b]);  // P1
            acc_im[a][b] = mfma16x4(ai[a], bi[b], acc_im[a][b]);  // P2
            acc_p3[a][b] = mfma16x4(as[a], bs[b], acc_p3[a][b]);  // P3
          }
      } else {
        T nai[FM];
#pragma unroll
        for (int a = 0; a < FM; ++a) nai[a] = -ai[a];
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
          for (int b = 0; b < FN; ++b) {
            acc_re[a][b] = mfma16x4(ar[a], br[b], acc_re[a][b]);
            acc_re[a][b] = mfma16x4(nai[a], bi[b], acc_re[a][b]);
            acc_im[a][b] = mfma16x4(ar[a], bi[b], acc_im[a][b]);
            acc_im[a][b] = mfma16x4(ai[a], br[b], acc_im[a][b]);
          }
      }
    }
  }

#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int gi = row0 + wr * WM + a * 16 + mfma_row<T>(lk, r);
        int gj = col0 + wc * WN + b * 16 + li;
        if (gi < M && gj < N) {
          T pr, pi;
          if (GAUSS3M) {
            pr = acc_re[a][b][r] - acc_im[a][b][r];
            pi = acc_p3[a][b][r] - acc_re[a][b][r] - acc_im[a][b][r];
          } else {
            pr = acc_re[a][b][r];
            pi = acc_im[a][b][r];
          }
          cx<T>* pc = C + (int64_t)gi * ldc + gj;
          cx<T> c0 = beta ? *pc : mkc<T>(0, 0);
          *pc = mkc<T>(c0.x + alpha * pr, c0.y + alpha * pi);
        }
      }
}


template <class T, int BM, int BN, int BK, int WM, int WN, bool A_COL, bool GAUSS3M>
__global__ void __launch_bounds__(256)
zgemm_kernel(int M, int N, int K, const cx<T>* __restrict__ A, int64_t lda,
             const cx<T>* __restrict__ B, int64_t ldb, cx<T>* __restrict__ C,
             int64_t ldc, T alpha, int beta, const int* __restrict__ skip, int64_t sA = 0,
             int64_t sB = 0, int64_t sC = 0, int sSkip = 0, int prio = 0) {
  raise_wave_priority(prio);
  const int b = blockIdx.y;   // system of a batch
  if (skip && skip[(int64_t)b * sSkip]) return;
  zgemm_tile<T, BM, BN, BK, WM, WN, A_COL, GAUSS3M>(M, N, K, A + b * sA, lda, B + b * sB, ldb,
                                                   C + b * sC, ldc, alpha, beta, blockIdx.x);
}

// The LU's trailing update A22 -= L21 U12 (the launches timed for the
// roofline record): zgemm_kernel's tile under its own kernel name, so that
// kernel traces list these launches apart from the library's other GEMMs.
// FP64 four-product tiles are built for three workgroups per CU (150
// registers, accumulators out of AGPRs); the launch reserves LDS so that two
// run per CU beside a panel leaf (trail_lds_extra in lu.hip).
template <class T, int BM, int BN, int BK, int WM, int WN, bool GAUSS3M>
__global__ void __launch_bounds__(256, (sizeof(T) == 8 && !GAUSS3M) ? 3 : 1)
trailing_update_kernel(int M, int N, int K, const cx<T>* __restrict__ A, int64_t lda,
                       const cx<T>* __restrict__ B, int64_t ldb, cx<T>* __restrict__ C,
                       int64_t ldc, T alpha) {
  zgemm_tile<T, BM, BN, BK, WM, WN, false, GAUSS3M>(M, N, K, A, lda, B, ldb, C, ldc, alpha, 1,
                                                   blockIdx.x);
}

// Two independent GEMM problems (row-major A) in one launch: workgroups
// [0, tiles1) take problem 1, the rest problem 2.  Used where two products
// of one block step read the same operand and write disjoint outputs.
template <class T> struct GemmProb {
  int M, N, K;
  const cx<T>* A;
  int64_t lda;
  const cx<T>* B;
  int64_t ldb;
  cx<T>* C;
  int64_t ldc;
  T alpha;
  int beta;
  int64_t sA = 0, sB = 0, sC = 0;   // strides between the systems of a batch
};

template <class T, int BM, int BN, int BK, int WM, int WN>
__global__ void __launch_bounds__(256)
zgemm_dual_kernel(GemmProb<T> p1, GemmProb<T> p2, int tiles1, const int* __restrict__ skip,
                  int sSkip = 0) {
  const int b = blockIdx.y;   // system of a batch
  if (skip && skip[(int64_t)b * sSkip]) return;
  const int t = blockIdx.x;
  if (t < tiles1)
    zgemm_tile<T, BM, BN, BK, WM, WN, false, false>(p1.M, p1.N, p1.K, p1.A + b * p1.sA, p1.lda,
                                                   p1.B + b * p1.sB, p1.ldb, p1.C + b * p1.sC,
                                                   p1.ldc, p1.alpha, p1.beta, t);
  else
    zgemm_tile<T, BM, BN, BK, WM, WN, false, false>(p2.M, p2.N, p2.K, p2.A + b * p2.sA, p2.lda,
                                                   p2.B + b * p2.sB, p2.ldb, p2.C + b * p2.sC,
                                                   p2.ldc, p2.alpha, p2.beta, t - tiles1);
}

// Product with the inverse D of a triangular diagonal block (the LU's
// 256 x 256 block inverses, row-major, lower when LOWER, else upper) in
// problem p1, with p2 an independent product as in zgemm_dual_kernel.  p1 is
// tiled 32 x 32 (eight row tiles of a full block instead of four, four waves
// of 16 x 16 each) and each row tile reads only D's nonzero triangle: row
// tile i0 of D X needs k < i0 + 32 (lower) or k >= i0 (upper), 9/16 of the
// block's work on average.  The zero blocks skipped add exact zeros, so the
// result is the full product's.  Workgroups [0, tiles1) take p1, the rest p2
// (BM2 x BN2 tiles, waves of WM2 x WN2).
template <class T, bool LOWER, int BM2, int BN2, int WM2, int WN2>
__global__ void __launch_bounds__(256)
tri_dual_kernel(GemmProb<T> p1, GemmProb<T> p2, int tiles1, const int* __restrict__ skip,
                int sSkip = 0) {
  const int b = blockIdx.y;   // system of a batch
  if (skip && skip[(int64_t)b * sSkip]) return;
  p1.A += b * p1.sA;
  p1.B += b * p1.sB;
  p1.C += b * p1.sC;
  p2.A += b * p2.sA;
  p2.B += b * p2.sB;
  p2.C += b * p2.sC;
  const int t = blockIdx.x;
  if (t < tiles1) {
    constexpr int TB = 32;
    const int tn = (p1.N + TB - 1) / TB;
    const int i0 = (t / tn) * TB;
    const int k0 = LOWER ? 0 : i0;
    const int k1 = LOWER ? min(i0 + TB, p1.K) : p1.K;
    zgemm_tile<T, TB, TB, 16, 16, 16, false, false>(
        min(TB, p1.M - i0), p1.N, k1 - k0, p1.A + (int64_t)i0 * p1.lda + k0, p1.lda,
        p1.B + (int64_t)k0 * p1.ldb, p1.ldb, p1.C + (int64_t)i0 * p1.ldc, p1.ldc, p1.alpha,
        p1.beta, t % tn);
  } else {
    zgemm_tile<T, BM2, BN2, 16, WM2, WN2, false, false>(p2.M, p2.N, p2.K, p2.A, p2.lda, p2.B,
                                                       p2.ldb, p2.C, p2.ldc, p2.alpha, p2.beta,
                                                       t - tiles1);
  }
}

// ---------------------------------------------------------------------------
// Second-generation complex GEMM (same contract as zgemm_kernel).
//
// * LDS is double-buffered: the next K-slice is loaded into registers while
//   the current one is multiplied, then stored into the other buffer, so one
//   barrier per K-slice suffices.
// * Bank-conflict-free LDS: MFMA operand reads take 16 consecutive elements
//   from each of two k-rows per half-wavefront.  Planes filled along k from a
//   row-major A use an XOR swizzle of the column index (i ^ swz(k), swz(k) =
//   (2k mod 32) ^ 16 (k mod 2)) that separates the two k-rows of a read into
//   opposite 16-bank halves and spreads the k-strided stores over all banks;
//   planes filled along i or j are padded to a row length = 16 mod 32.
// * Dynamic LDS: 2 buffers x 2 planes x BK x (row) for A and B.
// ---------------------------------------------------------------------------
template <int BK, int BM, int BN, bool A_COL, class T>
constexpr size_t zgemm2_lds_bytes() {
  return (size_t)2 * 2 * BK * ((A_COL ? BM + 16 : BM) + (BN + 16)) * sizeof(T);
}

__device__ inline int zg2_swz(int k) { return ((2 * k) & 31) ^ ((k & 1) << 4); }

template <class T, int BM, int BN, int BK, int WM, int WN, bool A_COL, bool GAUSS3M>
__global__ void __launch_bounds__(256)
zgemm2_kernel(int M, int N, int K, const cx<T>* __restrict__ A, int64_t lda,
              const cx<T>* __restrict__ B, int64_t ldb, cx<T>* __restrict__ C,
              int64_t ldc, T alpha, int beta, const int* __restrict__ skip) {
  if (skip && *skip) return;
  static_assert((BM / WM) * (BN / WN) == 4, "four waves per workgroup");
  static_assert(WM % 16 == 0 && WN % 16 == 0 && BK % 4 == 0 && BM % 32 == 0, "tile shape");
  constexpr int NT = 256;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int LA = A_COL ? BM + 16 : BM, LB = BN + 16;
  constexpr int EA = BM * BK / NT, EB = BK * BN / NT;
  static_assert(EA * NT == BM * BK && EB * NT == BK * BN, "load split");
  typedef typename acc_t<T>::type acc;

  extern __shared__ __align__(16) unsigned char zg2_smem[];
  T* As = reinterpret_cast<T*>(zg2_smem);          // [buf][plane][BK][LA]
  T* Bs = As + 2 * 2 * BK * LA;                    // [buf][plane][BK][LB]

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

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / (BN / WN), wc = wave % (BN / WN);
  const int li = lane & 15, lk = lane >> 4;

  acc acc_re[FM][FN], acc_im[FM][FN], acc_p3[GAUSS3M ? FM : 1][GAUSS3M ? FN : 1];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      acc_re[a][b] = acc{};
      acc_im[a][b] = acc{};
      if (GAUSS3M) acc_p3[a][b] = acc{};
    }

  cx<T> ra[EA], rb[EB];
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const int idx = tid + e * NT;
      int i, k;
      if (A_COL) { i = idx % BM; k = idx / BM; } else { i = idx / BK; k = idx % BK; }
      const int gi = row0 + i, gk = k0 + k;
      ra[e] = (gi < M && gk < K)
                  ? (A_COL ? A[(int64_t)gk * lda + gi] : A[(int64_t)gi * lda + gk])
                  : mkc<T>(0, 0);
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const int idx = tid + e * NT, k = idx / BN, j = idx % BN;
      const int gk = k0 + k, gj = col0 + j;
      rb[e] = (gk < K && gj < N) ? B[(int64_t)gk * ldb + gj] : mkc<T>(0, 0);
    }
  };
  auto store_tiles = [&](int buf) {
    T* a0 = As + (size_t)(buf * 2) * BK * LA;
    T* a1 = a0 + BK * LA;
    T* b0 = Bs + (size_t)(buf * 2) * BK * LB;
    T* b1 = b0 + BK * LB;
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const int idx = tid + e * NT;
      int i, k;
      if (A_COL) { i = idx % BM; k = idx / BM; } else { i = idx / BK; k = idx % BK; i ^= zg2_swz(k); }
      a0[k * LA + i] = ra[e].x;
      a1[k * LA + i] = ra[e].y;
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const int idx = tid + e * NT, k = idx / BN, j = idx % BN;
      b0[k * LB + j] = rb[e].x;
      b1[k * LB + j] = rb[e].y;
    }
  };

  load_tiles(0);
  store_tiles(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += BK) {
    const bool more = k0 + BK < K;
    if (more) load_tiles(k0 + BK);
    const T* a0 = As + (size_t)(buf * 2) * BK * LA;
    const T* a1 = a0 + BK * LA;
    const T* b0 = Bs + (size_t)(buf * 2) * BK * LB;
    const T* b1 = b0 + BK * LB;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int k = kk + lk;
      T ar[FM], ai[FM], br[FN], bi[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) {
        int i = wr * WM + a * 16 + li;
        if (!A_COL) i ^= zg2_swz(k);
        ar[a] = a0[k * LA + i];
        ai[a] = a1[k * LA + i];
      }
#pragma unroll
      for (int b = 0; b < FN; ++b) {
        const int j = wc * WN + b * 16 + li;
        br[b] = b0[k * LB + j];
        bi[b] = b1[k * LB + j];
      }
      if (GAUSS3M) {
        T as[FM], bs[FN];
#pragma unroll
        for (int a = 0; a < FM; ++a) as[a] = ar[a] + ai[a];
#pragma unroll
        for (int b = 0; b < FN; ++b) bs[b] = br[b] + bi[b];
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
          for (int b = 0; b < FN; ++b) {
            acc_re[a][b] = mfma16x4(ar[a], br[b], acc_re[a][b]);
            acc_im[a][b] = mfma16x4(ai[a], bi[b], acc_im[a][b]);
            acc_p3[a][b] = mfma16x4(as[a], bs[b], acc_p3[a][b]);
          }
      } else {
        T nai[FM];
#pragma unroll
        for (int a = 0; a < FM; ++a) nai[a] = -ai[a];
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
          for (int b = 0; b < FN; ++b) {
            acc_re[a][b] = mfma16x4(ar[a], br[b], acc_re[a][b]);
            acc_im[a][b] = mfma16x4(ar[a], bi[b], acc_im[a][b]);
          }
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
          for (int b = 0; b < FN; ++b) {
            acc_re[a][b] = mfma16x4(nai[a], bi[b], acc_re[a][b]);
            acc_im[a][b] = mfma16x4(ai[a], br[b], acc_im[a][b]);
          }
      }
    }
    if (more) store_tiles(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = row0 + wr * WM + a * 16 + mfma_row<T>(lk, r);
        const int gj = col0 + wc * WN + b * 16 + li;
        if (gi < M && gj < N) {
          T pr, pi;
          if (GAUSS3M) {
            pr = acc_re[a][b][r] - acc_im[a][b][r];
            pi = acc_p3[a][b][r] - acc_re[a][b][r] - acc_im[a][b][r];
          } else {
            pr = acc_re[a][b][r];
            pi = acc_im[a][b][r];
          }
          cx<T>* pc = C + (int64_t)gi * ldc + gj;
          const cx<T> c0 = beta ? *pc : mkc<T>(0, 0);
          *pc = mkc<T>(c0.x + alpha * pr, c0.y + alpha * pi);
        }
      }
}

// ---------------------------------------------------------------------------
// FP32 complex GEMM on 32x32x2 MFMA (same contract as zgemm_kernel, T = float).
//
// v_mfma_f32_32x32x2f32 reaches the FP32 matrix peak where 16x16x4 stops at
// about 0.89 of it (measured in a register-only loop on gfx950).  Operand
// layout: lane l supplies A[i = l % 32][k = l / 32] and B[k = l / 32][j = l % 32];
// result register r of lane l holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
// Each wave owns WM x WN = (32 FM) x (32 FN) of the tile with separate real and
// imaginary accumulators (four real products).  LDS double-buffered as in
// zgemm2_kernel; a row-major A plane is stored k-major with the column index
// XOR-swizzled by (2k mod 32): a read of 32 consecutive i stays a permutation of
// the 32 banks, and the k-strided stores spread over all banks.
// ---------------------------------------------------------------------------
template <int BK, int BM, int BN>
constexpr size_t zgemm32_lds_bytes() {
  return (size_t)2 * 2 * BK * (BM + BN) * sizeof(float);
}

template <int BM, int BN, int BK, int WM, int WN, bool A_COL>
__global__ void __launch_bounds__(256)
zgemm32_kernel(int M, int N, int K, const cx<float>* __restrict__ A, int64_t lda,
               const cx<float>* __restrict__ B, int64_t ldb, cx<float>* __restrict__ C,
               int64_t ldc, float alpha, int beta, const int* __restrict__ skip) {
  if (skip && *skip) return;
  static_assert((BM / WM) * (BN / WN) == 4, "four waves per workgroup");
  static_assert(WM % 32 == 0 && WN % 32 == 0 && BK % 2 == 0 && BM % 32 == 0, "tile shape");
  constexpr int NT = 256, FM = WM / 32, FN = WN / 32;
  constexpr int EA = BM * BK / NT, EB = BK * BN / NT;
  static_assert(EA * NT == BM * BK && EB * NT == BK * BN, "load split");
  typedef float v16f __attribute__((ext_vector_type(16)));
  extern __shared__ __align__(16) unsigned char z32_smem[];
  float* As = reinterpret_cast<float*>(z32_smem);   // [buf][plane][BK][BM]
  float* Bs = As + 2 * 2 * BK * BM;                  // [buf][plane][BK][BN]

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

  v16f acc_re[FM][FN], acc_im[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) { acc_re[a][b][r] = 0.f; acc_im[a][b][r] = 0.f; }

  cx<float> ra[EA], rb[EB];
  auto load = [&](int k0) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const int idx = tid + e * NT;
      int i, k;
      if (A_COL) { i = idx % BM; k = idx / BM; } else { i = idx / BK; k = idx % BK; }
      const int gi = row0 + i, gk = k0 + k;
      ra[e] = (gi < M && gk < K)
                  ? (A_COL ? A[(int64_t)gk * lda + gi] : A[(int64_t)gi * lda + gk])
                  : mkc<float>(0.f, 0.f);
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const int idx = tid + e * NT, k = idx / BN, j = idx % BN;
      const int gk = k0 + k, gj = col0 + j;
      rb[e] = (gk < K && gj < N) ? B[(int64_t)gk * ldb + gj] : mkc<float>(0.f, 0.f);
    }
  };
  auto store = [&](int buf) {
    float* a0 = As + (size_t)(buf * 2) * BK * BM;
    float* a1 = a0 + BK * BM;
    float* b0 = Bs + (size_t)(buf * 2) * BK * BN;
    float* b1 = b0 + BK * BN;
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const int idx = tid + e * NT;
      int i, k;
      if (A_COL) { i = idx % BM; k = idx / BM; } else { i = idx / BK; k = idx % BK; i ^= (2 * k) & 31; }
      a0[k * BM + i] = ra[e].x;
      a1[k * BM + i] = ra[e].y;
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const int idx = tid + e * NT, k = idx / BN, j = idx % BN;
      b0[k * BN + j] = rb[e].x;
      b1[k * BN + j] = rb[e].y;
    }
  };

  load(0);
  store(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += BK) {
    const bool more = k0 + BK < K;
    if (more) load(k0 + BK);
    const float* a0 = As + (size_t)(buf * 2) * BK * BM;
    const float* a1 = a0 + BK * BM;
    const float* b0 = Bs + (size_t)(buf * 2) * BK * BN;
    const float* b1 = b0 + BK * BN;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 2) {
      const int k = kk + lh;
      float ar[FM], ai[FM], nai[FM], br[FN], bi[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) {
        int i = wr * WM + a * 32 + l32;
        if (!A_COL) i ^= (2 * k) & 31;
        ar[a] = a0[k * BM + i];
        ai[a] = a1[k * BM + i];
        nai[a] = -ai[a];
      }
#pragma unroll
      for (int b = 0; b < FN; ++b) {
        const int j = wc * WN + b * 32 + l32;
        br[b] = b0[k * BN + j];
        bi[b] = b1[k * BN + j];
      }
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          acc_re[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[a], br[b], acc_re[a][b], 0, 0, 0);
          acc_im[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[a], bi[b], acc_im[a][b], 0, 0, 0);
        }
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          acc_re[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(nai[a], bi[b], acc_re[a][b], 0, 0, 0);
          acc_im[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ai[a], br[b], acc_im[a][b], 0, 0, 0);
        }
    }
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int gi = row0 + wr * WM + a * 32 + 8 * (r / 4) + 4 * lh + (r % 4);
        const int gj = col0 + wc * WN + b * 32 + l32;
        if (gi < M && gj < N) {
          cx<float>* pc = C + (int64_t)gi * ldc + gj;
          const cx<float> c0 = beta ? *pc : mkc<float>(0.f, 0.f);
          *pc = mkc<float>(c0.x + alpha * acc_re[a][b][r], c0.y + alpha * acc_im[a][b][r]);
        }
      }
}
