// Kernels of the blocked right-looking LU with partial pivoting (GEPP) and of
// the triangular solves, for row-major complex matrices on gfx950.
//
// Algorithm: LAPACK getrf/getf2/laswp/getrs (Anderson et al., LAPACK Users'
// Guide 3rd ed. 1999; Golub & Van Loan 4th ed. 3.4).  The pivot of column j
// is the first row holding the largest |Re| + |Im| (the i?amax rule), so
// pivot sequences match LAPACK on matrices without near ties.
#pragma once
#include "common.hpp"

// ---------------------------------------------------------------------------
// Exchange area of the cooperative panel leaves (device memory), made of
// tagged 64-bit words: high 32 bits = column tag (unique per column within a
// factorization, the area is zeroed at its start), low 32 bits = payload.  A
// reader accepts a word only when its tag is the current column's, so data
// and "arrival" travel together: no flag, no fence, no store-completion wait.
// Words are written and read with relaxed agent-scope atomics, which compile
// to single coherent (sc1) global loads/stores; an agent-scope fence would
// instead emit buffer_wbl2/buffer_inv and stall on (and evict) the dirty
// tiles of the trailing-update GEMM that runs concurrently on the side stream.
// Slots are double-buffered by column parity: a workgroup can only write
// parity p for column j+2 after all workgroups published column j+1, which
// each does only after it finished reading column j.
// ---------------------------------------------------------------------------
typedef unsigned long long u64;
template <class T> struct words_per;                 // 32-bit payload words per real
template <> struct words_per<float> { static constexpr int n = 1; };
template <> struct words_per<double> { static constexpr int n = 2; };

struct LeafXchg {
  u64* val;    // [2][Gmax]            candidate |Re|+|Im| (payload words: words_per)
  u64* idx;    // [2][Gmax]            candidate panel-local row
  u64* row;    // [2][Gmax][W*2*wpr]   candidate row (all W leaf columns)
  u64* rowj;   // [2][W*2*wpr]         current row j (published by its owner)
  int gstride; // Gmax
};

__device__ inline void put_word(u64* p, unsigned tag, unsigned payload) {
  __hip_atomic_store(p, ((u64)tag << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline u64 get_word(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool tag_ok(u64 w, unsigned tag) { return (unsigned)(w >> 32) == tag; }

// Real value <-> tagged words.
__device__ inline void put_real(u64* p, unsigned tag, float v) { put_word(p, tag, __float_as_uint(v)); }
__device__ inline void put_real(u64* p, unsigned tag, double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  put_word(p, tag, (unsigned)b);
  put_word(p + 1, tag, (unsigned)(b >> 32));
}
// Spin (bounded, watching *err) until every word of a real carries the tag.
// A wait that gives up clears `ok`: the value it returns is then stale and
// must not be used as an index.
template <class T> __device__ inline T get_real(const u64* p, unsigned tag, int* err, bool& ok);
template <> __device__ inline float get_real<float>(const u64* p, unsigned tag, int* err, bool& ok) {
  unsigned spins = 0;
  u64 w = get_word(p);
  while (!tag_ok(w, tag)) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 255u) == 0) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { ok = false; break; }
      if (spins > (1u << 22)) { atomicExch(err, 1); ok = false; break; }
    }
    w = get_word(p);
  }
  return __uint_as_float((unsigned)w);
}
template <> __device__ inline double get_real<double>(const u64* p, unsigned tag, int* err, bool& ok) {
  unsigned spins = 0;
  u64 lo = get_word(p), hi = get_word(p + 1);
  while (!tag_ok(lo, tag) || !tag_ok(hi, tag)) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 255u) == 0) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { ok = false; break; }
      if (spins > (1u << 22)) { atomicExch(err, 1); ok = false; break; }
    }
    lo = get_word(p);
    hi = get_word(p + 1);
  }
  return __longlong_as_double((long long)((lo & 0xffffffffull) | ((hi & 0xffffffffull) << 32)));
}
__device__ inline int get_int(const u64* p, unsigned tag, int* err, bool& ok) {
  unsigned spins = 0;
  u64 w = get_word(p);
  while (!tag_ok(w, tag)) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 255u) == 0) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { ok = false; break; }
      if (spins > (1u << 22)) { atomicExch(err, 1); ok = false; break; }
    }
    w = get_word(p);
  }
  return (int)(unsigned)w;
}

// ---------------------------------------------------------------------------
// Wavefront reductions on DPP.  Four DPP steps (swap neighbours, swap pairs,
// mirror halves of 8, mirror the 16-lane row) give every lane the result over
// its row; the four row results are combined through v_readlane, so the value
// returned is uniform across the wavefront.  All 64 lanes must be active.
// ---------------------------------------------------------------------------
constexpr int DPP_QUAD_1032 = 0xB1, DPP_QUAD_2301 = 0x4E, DPP_ROW_HALF_MIRROR = 0x141,
              DPP_ROW_MIRROR = 0x140;

template <int CTRL> __device__ inline int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL> __device__ inline float dpp_r(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL> __device__ inline double dpp_r(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_i<CTRL>((int)b), hi = (unsigned)dpp_i<CTRL>((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline int lane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ inline float lane_r(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double lane_r(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <class T> __device__ inline T wave_max(T v) {
  v = fmax(v, dpp_r<DPP_QUAD_1032>(v));
  v = fmax(v, dpp_r<DPP_QUAD_2301>(v));
  v = fmax(v, dpp_r<DPP_ROW_HALF_MIRROR>(v));
  v = fmax(v, dpp_r<DPP_ROW_MIRROR>(v));
  return fmax(fmax(lane_r(v, 0), lane_r(v, 16)), fmax(lane_r(v, 32), lane_r(v, 48)));
}
__device__ inline int wave_min(int v) {
  v = min(v, dpp_i<DPP_QUAD_1032>(v));
  v = min(v, dpp_i<DPP_QUAD_2301>(v));
  v = min(v, dpp_i<DPP_ROW_HALF_MIRROR>(v));
  v = min(v, dpp_i<DPP_ROW_MIRROR>(v));
  return min(min(lane_i(v, 0), lane_i(v, 16)), min(lane_i(v, 32), lane_i(v, 48)));
}
// (largest value, lowest index among the lanes holding it), uniform.
template <class T> __device__ inline void wave_argmax(T& v, int& i) {
  const T mx = wave_max(v);
  i = wave_min(v == mx ? i : 0x7fffffff);
  v = mx;
}

// Spin (bounded, watching *err) until the word carries the tag; payload.
__device__ inline unsigned get_u32(const u64* p, unsigned tag, int* err) {
  unsigned spins = 0;
  u64 w = get_word(p);
  while (!tag_ok(w, tag)) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 255u) == 0) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      if (spins > (1u << 22)) { atomicExch(err, 1); break; }
    }
    w = get_word(p);
  }
  return (unsigned)w;
}

// ---------------------------------------------------------------------------
// Cooperative GEPP of an m x w leaf (w <= W) at P (row-major, ld).  G
// workgroups of 256 threads own 256 consecutive rows each, one row per thread;
// a thread keeps its row's W leaf columns in registers for the whole leaf.
// The footprint (about 100 VGPRs for FP32, 170 for FP64, under 30 KB of LDS)
// leaves room on each CU for the trailing-update GEMM of the previous panel,
// which runs concurrently on another stream.
//
// Rows live in registers in ROTATED order: before step j, register c holds
// column (c + j) mod W, so step j always works on register 0 and every
// register index is a compile-time constant; the update writes register c
// from register c + 1 (the rotation costs nothing) and appends the finished
// multiplier at W - 1.  Registers W - j .. W - 1 hold finished L entries, so
// the copy of the pivot row used by the update ("masked") is zero there.
//
// Per column j:
//  1. each wavefront reduces |Re| + |Im| over its rows >= j on DPP (largest
//     value, lowest row on ties, the i?amax rule); the winning lane stages
//     its row (full and masked) in LDS, the owner of row j stages row j; one
//     barrier; every wave combines the four wave results the same way.
//  2. G > 1: wavefront 0 publishes the workgroup's candidate row, row j (from
//     its owner), then the candidate value and row index as tagged words;
//     polls the G candidates and reduces them on DPP; fetches the winning row
//     and, in the workgroup that owns the pivot row, row j; second barrier.
//  3. interchange, multiplier and rank-1 update of the thread's row.
// LDS is double-buffered by column parity, so no barrier closes a column.
// ipiv receives global 0-based rows.
// ---------------------------------------------------------------------------
template <class T, int W>
__global__ void __launch_bounds__(256)
leaf_getf2_kernel(cx<T>* __restrict__ P, int64_t ld, int m, int w, int row_base,
                  int* __restrict__ ipiv, int* __restrict__ info, LeafXchg x, unsigned seq_base,
                  int* __restrict__ err, int64_t sP = 0, int sIv = 0, int64_t sX = 0,
                  int prio = 0) {
  raise_wave_priority(prio);
  // batch of independent systems: blockIdx.y selects the system (strides sP
  // elements of P, sIv entries of ipiv, sX exchange words; info per system)
  {
    const int b = blockIdx.y;
    P += b * sP;
    ipiv += (int64_t)b * sIv;
    info += b;
    x.val += b * sX;
    x.idx += b * sX;
    x.row += b * sX;
    x.rowj += b * sX;
  }
  constexpr int NT = 256, NW = NT / 64, STAGE = 32;
  constexpr int WPR = words_per<T>::n, RW = 2 * WPR;   // 32-bit words per real / complex
  // column buffers, aliased with the staging tile used before and after the loop
  constexpr int COLBUF = 2 * (2 * NW + 3) * W, STAGEBUF = STAGE * (W + 1);
  __shared__ cx<T> lds[COLBUF > STAGEBUF ? COLBUF : STAGEBUF];
  __shared__ T red_v[2][NW];
  __shared__ int red_i[2][NW];
  __shared__ T sel_v[2];
  __shared__ int sel_i[2];
  cx<T>(*cand)[NW][W] = reinterpret_cast<cx<T>(*)[NW][W]>(lds);                // [2][NW][W]
  cx<T>(*candm)[NW][W] = reinterpret_cast<cx<T>(*)[NW][W]>(lds + 2 * NW * W);  // [2][NW][W]
  cx<T>(*prow)[W] = reinterpret_cast<cx<T>(*)[W]>(lds + 4 * NW * W);           // [2][W]
  cx<T>(*pupd)[W] = reinterpret_cast<cx<T>(*)[W]>(lds + 4 * NW * W + 2 * W);   // [2][W]
  cx<T>(*oldj)[W] = reinterpret_cast<cx<T>(*)[W]>(lds + 4 * NW * W + 4 * W);   // [2][W]
  cx<T>(*stage)[W + 1] = reinterpret_cast<cx<T>(*)[W + 1]>(lds);               // [STAGE][W+1]

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int G = gridDim.x, g = blockIdx.x;
  const int lo = g * NT, hi = min(m, lo + NT);
  const int row = lo + tid;
  const bool mine = row < hi;
  const int gs = x.gstride;

  // coalesced load through LDS, STAGE rows at a time
  cx<T> a[W];
  for (int s0 = 0; s0 < NT; s0 += STAGE) {
    __syncthreads();
    for (int e = tid; e < STAGE * W; e += NT) {
      const int i = e / W, c = e % W, r = lo + s0 + i;
      stage[i][c] = (r < hi && c < w) ? P[(int64_t)r * ld + c] : mkc<T>(0, 0);
    }
    __syncthreads();
    if (tid >= s0 && tid < s0 + STAGE) {
#pragma unroll
      for (int c = 0; c < W; ++c) a[c] = stage[tid - s0][c];
    }
  }
  __syncthreads();   // the staging tile is reused as column buffers

  for (int j = 0; j < w; ++j) {
    const int par = j & 1;
    const unsigned tag = seq_base + (unsigned)j + 1u;
    const int live = W - j;   // registers 0 .. live-1 hold unfinished columns
    // -- 1. pivot search inside the workgroup
    T v = (mine && row >= j) ? cabs1<T>(a[0]) : T(-1);
    int vi = row;
    wave_argmax(v, vi);
    if (row == vi) {
#pragma unroll
      for (int c = 0; c < W; ++c) {
        cand[par][wave][c] = a[c];
        candm[par][wave][c] = (c < live) ? a[c] : mkc<T>(0, 0);
      }
    }
    if (row == j) {
#pragma unroll
      for (int c = 0; c < W; ++c) oldj[par][c] = a[c];
    }
    if (lane == 0) { red_v[par][wave] = v; red_i[par][wave] = vi; }
    __syncthreads();
    T bv = red_v[par][0];
    int bi = red_i[par][0], bw = 0;
#pragma unroll
    for (int q = 1; q < NW; ++q) {
      const T ov = red_v[par][q];
      if (ov > bv) { bv = ov; bi = red_i[par][q]; bw = q; }   // waves ascend in rows
    }
    T best;
    int p;
    const cx<T>* F;   // pivot row (full)
    const cx<T>* U;   // pivot row, finished columns zeroed
    if (G == 1) {
      best = bv;
      p = (bv >= T(0)) ? bi : j;
      F = cand[par][bw];
      U = candm[par][bw];
    } else {
      if (wave == 0) {
        const bool own_j = (j >= lo && j < hi);
        // publish: candidate row, row j, then value and index
        const unsigned* cw = reinterpret_cast<const unsigned*>(cand[par][bw]);
        u64* dst = x.row + ((size_t)par * gs + g) * (W * RW);
        for (int e = lane; e < W * RW; e += 64) put_word(dst + e, tag, cw[e]);
        if (own_j) {
          const unsigned* jw = reinterpret_cast<const unsigned*>(oldj[par]);
          u64* dj = x.rowj + (size_t)par * (W * RW);
          for (int e = lane; e < W * RW; e += 64) put_word(dj + e, tag, jw[e]);
        }
        if (lane == 0) {
          put_real(x.val + ((size_t)par * gs + g) * WPR, tag, bv);
          put_word(x.idx + (size_t)par * gs + g, tag, (unsigned)bi);
        }
        // gather the G candidates and reduce
        T qv = T(-1);
        int qi = 0x7fffffff;
        bool ok = true;
        for (int q = lane; q < G; q += 64) {
          const T ov = get_real<T>(x.val + ((size_t)par * gs + q) * WPR, tag, err, ok);
          const int oi = get_int(x.idx + (size_t)par * gs + q, tag, err, ok);
          if (ov > qv || (ov == qv && oi < qi)) { qv = ov; qi = oi; }
        }
        // a wait that gave up (the error flag is then set and reported to the
        // host) leaves a stale candidate: take no pivot and no update in this
        // column, so every index used below stays in range and the kernel ends
        if (__ballot(!ok) != 0ull) { qv = T(-1); qi = 0x7fffffff; }
        wave_argmax(qv, qi);
        const bool have = qv >= T(0);
        const int pp = have ? qi : j;
        if (lane == 0) { sel_v[par] = qv; sel_i[par] = pp; }
        // winning row (full and masked)
        const int win = have ? qi / NT : 0;
        const u64* src = x.row + ((size_t)par * gs + win) * (W * RW);
        unsigned* fw = reinterpret_cast<unsigned*>(prow[par]);
        unsigned* uw = reinterpret_cast<unsigned*>(pupd[par]);
        for (int e = lane; e < W * RW; e += 64) {
          const unsigned u = have ? get_u32(src + e, tag, err) : 0u;   // values only
          fw[e] = u;
          uw[e] = (e / RW < live) ? u : 0u;
        }
        // row j, needed only by the owner of the pivot row
        if (pp != j && pp >= lo && pp < hi && !own_j) {
          const u64* sj = x.rowj + (size_t)par * (W * RW);
          unsigned* ow = reinterpret_cast<unsigned*>(oldj[par]);
          for (int e = lane; e < W * RW; e += 64) ow[e] = get_u32(sj + e, tag, err);
        }
      }
      __syncthreads();
      best = sel_v[par];
      p = sel_i[par];
      F = prow[par];
      U = pupd[par];
    }
    if (g == 0 && tid == 0) {
      ipiv[j] = row_base + p;
      if (best == T(0) && *info == 0) *info = row_base + j + 1;
    }
    // -- 3. interchange, multiplier, rank-1 update (rotating by one register)
    if (p != j) {
      if (row == j) {
#pragma unroll
        for (int c = 0; c < W; ++c) a[c] = F[c];
      } else if (row == p) {
#pragma unroll
        for (int c = 0; c < W; ++c) a[c] = oldj[par][c];
      }
    }
    const bool upd = best > T(0) && mine && row > j;
    const cx<T> l = upd ? cmul<T>(a[0], crecip<T>(F[0])) : mkc<T>(0, 0);
    const cx<T> head = upd ? l : a[0];
#pragma unroll
    for (int c = 0; c < W - 1; ++c) a[c] = cfms<T>(a[c + 1], l, U[c + 1]);
    a[W - 1] = head;
  }

  // undo the rotation (register c now holds column (c + w) mod W) and store
  // coalesced through LDS
  for (int s0 = 0; s0 < NT; s0 += STAGE) {
    __syncthreads();
    if (tid >= s0 && tid < s0 + STAGE) {
#pragma unroll
      for (int c = 0; c < W; ++c) {
        int col = c + w;
        if (col >= W) col -= W;
        stage[tid - s0][col] = a[c];
      }
    }
    __syncthreads();
    for (int e = tid; e < STAGE * W; e += NT) {
      const int i = e / W, c = e % W, r = lo + s0 + i;
      if (r < hi && c < w) P[(int64_t)r * ld + c] = stage[i][c];
    }
  }
}

// ---------------------------------------------------------------------------
// Communication-avoiding leaf with tournament pivoting (CALU: Grigori, Demmel
// & Xiang, SIAM J. Matrix Anal. Appl. 32 (2011) 1317) of an m x w FP32 leaf
// at P (row-major, ld), w <= W.  G workgroups of 256 threads own 256
// consecutive rows each, one row per thread, as in leaf_getf2_kernel; the
// pivots are chosen with a handful of exchanges instead of one per column:
//  1. every workgroup runs GEPP on its own rows (registers and LDS only) and
//     publishes the rows it took as pivots, in order: its candidate set;
//  2. tournament: at level l, workgroup q < ceil(G_l / 8) gathers the
//     candidate sets 8q .. 8q+7 of level l-1 (up to 256 rows, read in their
//     original values from P), runs GEPP on them and publishes its pivots;
//     the single set of the last level is the leaf's pivot sequence, and the
//     GEPP rows of that workgroup (always workgroup 0) are the factors
//     L11 \ U11 of the pivoted w x w block, which it publishes;
//  3. every workgroup turns the sequence into LAPACK's interchanges (row i
//     <-> the current row of pivot i), then writes the rows it owns: rows
//     i < w get the published LU rows, every other row its multipliers
//     l = a U11^-1 from its original values a (original rows i < w that the
//     interchanges displace are republished by workgroup 0, their owner).
// Pivots differ from GEPP's (each is the largest entry among its candidates,
// not among all rows) and are for the FP32 factors only; the FP64 path keeps
// leaf_getf2_kernel and LAPACK's pivots.  Exchanges use the tagged words of
// leaf_getf2_kernel with tags seq_base + 1 + stage.  P is written only in
// step 3, after every read of original rows by the tournament (each read's
// value feeds a publication that step 3 waits for).
// ---------------------------------------------------------------------------
struct CaluXchg {
  u64* sets;   // [CALU_LEVELS][gstride][W]  candidate rows (panel-local), 0xffffffff = none
  u64* lu;     // [W][W][2]                  LU rows of the pivoted block
  u64* disp;   // [W][W][2]                  original rows 0 .. w-1 (republished)
  int gstride;
};
constexpr int CALU_LEVELS = 4;   // G <= 8^3 = 512 workgroups
constexpr unsigned CALU_TAG_LU = 8, CALU_TAG_DISP = 9;

template <int W> struct CaluShared {
  cx<float> cand[2][4][W];   // the waves' candidate pivot rows, by step parity
  cx<float> crcp[2][4];      // 1 / (their entry in the step's column)
  float red_v[2][4];
  int red_i[2][4];
  int sel[W];                // pivot thread of each step (-1: none)
  float selv[W];             // its |Re| + |Im|
  int rowid[256];            // panel-local row held by each thread (-1: none)
  cx<float> u[W][W];         // L11 \ U11 of the pivoted block
  cx<float> rdiag[W];        // 1 / U11[c][c] (0 for a zero pivot)
  int piv[W];                // pivot sequence (panel-local rows)
  int dpos[W], drow[W];      // displaced original rows: position dpos holds row drow
  int fpos[W];               // final position of original row r < W
  int ndisp;
};

// a - b c for FP32 complex values as two fused multiply-adds per component.
__device__ inline cx<float> cfms_fused(cx<float> a, cx<float> b, cx<float> c) {
  return mkc<float>(fmaf(-b.x, c.x, fmaf(b.y, c.y, a.x)), fmaf(-b.x, c.y, fmaf(-b.y, c.x, a.y)));
}

// GEPP over the rows of the 256 threads (one row each, `live` marks rows that
// take part), w <= W steps, in registers.  The step loop is unrolled, so that
// step j updates only the W - 1 - j columns right of j and every register
// index is a constant; the winning lane of each wave stages its row and the
// reciprocal of its pivot entry.  On return (after a barrier) sh.sel[j] is
// the thread holding the pivot of step j (-1 once no row is left) and
// sh.selv[j] its |Re| + |Im|.  The pivot thread of step j stops updating at
// step j, so its registers hold its LU row: multipliers in columns < j, U in
// columns >= j.  Returns the step at which this thread's row was taken (-1:
// never).  Ties go to the lowest thread, as the i?amax rule.
template <int W>
__device__ __forceinline__ int calu_gepp(cx<float> (&a)[W], bool live, int w,
                                         CaluShared<W>& sh) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int mystep = -1;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    if (j < w) {   // uniform
      const int par = j & 1;
      float v = live ? cabs1<float>(a[j]) : -1.f;
      int vi = tid;
      wave_argmax(v, vi);
      if (tid == vi && v >= 0.f) {
#pragma unroll
        for (int c = j; c < W; ++c) sh.cand[par][wave][c] = a[c];
        sh.crcp[par][wave] = v > 0.f ? crecip<float>(a[j]) : mkc<float>(0.f, 0.f);
      }
      if (lane == 0) { sh.red_v[par][wave] = v; sh.red_i[par][wave] = vi; }
      __syncthreads();
      float bv = sh.red_v[par][0];
      int bi = sh.red_i[par][0], bw = 0;
#pragma unroll
      for (int q = 1; q < 4; ++q) {
        const float ov = sh.red_v[par][q];
        if (ov > bv) { bv = ov; bi = sh.red_i[par][q]; bw = q; }   // waves ascend in threads
      }
      if (tid == 0) { sh.sel[j] = bv >= 0.f ? bi : -1; sh.selv[j] = bv; }
      if (bv >= 0.f) {   // uniform: a row is left
        if (tid == bi) {
          live = false;
          mystep = j;
        } else if (live) {
          const cx<float>* F = sh.cand[par][bw];
          const cx<float> l = cmul<float>(a[j], sh.crcp[par][bw]);   // 0 for a zero pivot
          a[j] = l;
#pragma unroll
          for (int c = j + 1; c < W; ++c) a[c] = cfms_fused(a[c], l, F[c]);
        }
      }
    }
  }
  __syncthreads();
  return mystep;
}

// Loads row `r` (panel-local, r < 0: zeros) of P, w <= W columns, into a,
// 16 bytes at a time when `vec` (even ld, 16-byte aligned P).
template <int W>
__device__ __forceinline__ void calu_load_row(cx<float> (&a)[W], const cx<float>* P, int64_t ld,
                                              int r, int w, bool vec) {
  if (vec) {
#pragma unroll
    for (int c = 0; c < W; c += 2) {
      if (r >= 0 && c + 1 < w) {
        const float4 q = *reinterpret_cast<const float4*>(P + (int64_t)r * ld + c);
        a[c] = mkc<float>(q.x, q.y);
        a[c + 1] = mkc<float>(q.z, q.w);
      } else {
        a[c] = (r >= 0 && c < w) ? P[(int64_t)r * ld + c] : mkc<float>(0.f, 0.f);
        a[c + 1] = mkc<float>(0.f, 0.f);
      }
    }
  } else {
#pragma unroll
    for (int c = 0; c < W; ++c)
      a[c] = (r >= 0 && c < w) ? P[(int64_t)r * ld + c] : mkc<float>(0.f, 0.f);
  }
}

template <int W>
__global__ void __launch_bounds__(256)
leaf_calu_kernel(cx<float>* __restrict__ P, int64_t ld, int m, int w, int row_base,
                 int* __restrict__ ipiv, int* __restrict__ info, CaluXchg x, unsigned seq_base,
                 int* __restrict__ err, int prio = 0) {
  raise_wave_priority(prio);
  constexpr int NT = 256;
  static_assert(W % 2 == 0 && 8 * W <= NT, "eight candidate sets fill a workgroup");
  __shared__ CaluShared<W> sh;
  const int tid = threadIdx.x;
  const int G = gridDim.x, g = blockIdx.x;
  const int lo = g * NT, hi = min(m, lo + NT);
  const int row = lo + tid;
  const bool mine = row < hi;
  const int gs = x.gstride;
  const bool vec = (ld % 2 == 0) && ((reinterpret_cast<uintptr_t>(P) & 15) == 0);

  // levels: G_0 = G sets, G_l = ceil(G_{l-1} / 8), until one set is left
  int nlev = 0;
  for (int gl = G; gl > 1; gl = (gl + 7) / 8) ++nlev;

  cx<float> a[W];
  int mystep = -1;
  bool final_here = false;   // this workgroup holds the last level's GEPP
  // level 0: local GEPP on the workgroup's own rows (with G = 1 already the
  // leaf's GEPP); level l >= 1: the tournament's reduction of 8 sets.  One
  // call site of calu_gepp, so that it is inlined and `a` stays in registers.
  for (int l = 0, gl = G; l <= nlev; ++l) {
    const int gn = l == 0 ? G : (gl + 7) / 8;
    if (g < gn) {
      int r = -1;
      if (l == 0) {
        r = mine ? row : -1;
      } else {
        const int set = 8 * g + tid / W, slot = tid % W;
        if (set < gl && slot < w) {
          const unsigned u = get_u32(x.sets + ((size_t)(l - 1) * gs + set) * W + slot,
                                     seq_base + (unsigned)l, err);
          r = (u < (unsigned)m) ? (int)u : -1;   // none, or a wait that gave up
        }
        __syncthreads();   // sh.rowid / sh.sel of the previous GEPP are read no more
      }
      sh.rowid[tid] = r;
      calu_load_row<W>(a, P, ld, r, w, vec);
      __syncthreads();
      mystep = calu_gepp<W>(a, r >= 0, w, sh);
      if (l < nlev && tid < w) {
        const int s = sh.sel[tid];
        put_word(x.sets + ((size_t)l * gs + g) * W + tid, seq_base + 1u + (unsigned)l,
                 s >= 0 ? (unsigned)sh.rowid[s] : 0xffffffffu);
      }
      final_here = l == nlev;
    }
    if (l > 0) gl = gn;
  }
  // -- the last level (workgroup 0): pivot sequence, LU rows, ipiv, info
  if (final_here) {
    if (tid < W) sh.piv[tid] = (tid < w && sh.sel[tid] >= 0) ? sh.rowid[sh.sel[tid]] : -1;
    for (int e = tid; e < W * W; e += NT) sh.u[e / W][e % W] = mkc<float>(0.f, 0.f);
    __syncthreads();
    if (mystep >= 0) {   // LU row `mystep`
#pragma unroll
      for (int c = 0; c < W; ++c) sh.u[mystep][c] = a[c];
    }
    __syncthreads();
    if (G > 1) {
      for (int e = tid; e < W * W; e += NT) {
        const int i = e / W, c = e % W;
        const cx<float> v = sh.u[i][c];
        put_word(x.lu + (size_t)e * 2, seq_base + 1u + CALU_TAG_LU, __float_as_uint(v.x));
        put_word(x.lu + (size_t)e * 2 + 1, seq_base + 1u + CALU_TAG_LU, __float_as_uint(v.y));
      }
      if (tid < W) {
        const int pv = tid < w ? sh.piv[tid] : -1;
        put_word(x.sets + ((size_t)nlev * gs) * W + tid, seq_base + 1u + (unsigned)nlev,
                 pv >= 0 ? (unsigned)pv : 0xffffffffu);
      }
    }
  } else {
    // pivot sequence and LU rows from the last level
    if (tid < W) {
      int pv = -1;
      if (tid < w) {
        const unsigned u = get_u32(x.sets + ((size_t)nlev * gs) * W + tid,
                                   seq_base + 1u + (unsigned)nlev, err);
        pv = u < (unsigned)m ? (int)u : -1;
      }
      sh.piv[tid] = pv;
    }
    for (int e = tid; e < W * W; e += NT) {
      const int i = e / W, c = e % W;
      const float re = __uint_as_float(get_u32(x.lu + (size_t)e * 2, seq_base + 1u + CALU_TAG_LU, err));
      const float im = __uint_as_float(get_u32(x.lu + (size_t)e * 2 + 1, seq_base + 1u + CALU_TAG_LU, err));
      sh.u[i][c] = mkc<float>(re, im);
    }
  }
  __syncthreads();
  // -- 3. interchanges: wave 0 replays them with lane r < W holding the
  // position of original row r and the original row at position r (rows
  // >= w are still in place when chosen: only rows < w move before their
  // turn).  The row a displaced row lands on is recorded (position dpos
  // holds original row drow; a later record for the same position wins).
  if (tid < 64) {
    const int lane = tid;
    const int pivl = lane < W ? sh.piv[lane] : -1;   // lane i: pivot i
    int cur = lane, at = lane, nd = 0;
    for (int i = 0; i < w; ++i) {
      int r = __builtin_amdgcn_readlane(pivl, i);
      const int at_i = __builtin_amdgcn_readlane(at, i);
      if (r < 0) r = at_i;   // a failed wait: keep the row in place
      const int p = r < w ? __builtin_amdgcn_readlane(cur, r) : r;
      const int xr = at_i;   // the small row at position i moves to p
      if (p < w) {
        if (lane == p) at = xr;
      } else {
        if (lane == 0) { sh.dpos[nd] = p; sh.drow[nd] = xr; }
        ++nd;
      }
      if (lane == xr) cur = p;
      if (lane == i) at = r;
      if (r < w && lane == r) cur = i;
      if (g == 0 && lane == 0) {
        ipiv[i] = row_base + p;
        if (sh.selv[i] == 0.f && *info == 0 && final_here) *info = row_base + i + 1;
      }
    }
    if (lane == 0) sh.ndisp = nd;
    if (lane < W) sh.fpos[lane] = cur;
  }
  if (tid < W) {
    const cx<float> d = tid < w ? sh.u[tid][tid] : mkc<float>(1.f, 0.f);
    sh.rdiag[tid] = (d.x == 0.f && d.y == 0.f) ? mkc<float>(0.f, 0.f) : crecip<float>(d);
  }
  __syncthreads();
  // workgroup 0 republishes the displaced original rows (it owns rows < w and
  // has not written them yet)
  if (g == 0 && G > 1) {
    for (int e = tid; e < W * W; e += NT) {
      const int i = e / W, c = e % W;
      const cx<float> v = (i < w && c < w && i < m) ? P[(int64_t)i * ld + c] : mkc<float>(0.f, 0.f);
      put_word(x.disp + (size_t)e * 2, seq_base + 1u + CALU_TAG_DISP, __float_as_uint(v.x));
      put_word(x.disp + (size_t)e * 2 + 1, seq_base + 1u + CALU_TAG_DISP, __float_as_uint(v.y));
    }
  }
  if (G == 1) {
    // one workgroup: its GEPP ran on all rows, so every register row already
    // is a row of the factors (pivot rows their LU row, the others their
    // multipliers); each goes to its final position
    if (mine) {
      const int dest = mystep >= 0 ? mystep : (row < w ? sh.fpos[row] : row);
      if (vec) {
#pragma unroll
        for (int c = 0; c < W; c += 2) {
          if (c + 1 < w)
            *reinterpret_cast<float4*>(P + (int64_t)dest * ld + c) =
                make_float4(a[c].x, a[c].y, a[c + 1].x, a[c + 1].y);
          else if (c < w)
            P[(int64_t)dest * ld + c] = a[c];
        }
      } else {
#pragma unroll
        for (int c = 0; c < W; ++c)
          if (c < w) P[(int64_t)dest * ld + c] = a[c];
      }
    }
    return;
  }
  // the original values of the row that ends at this thread's position
  int src = row;
  for (int k = 0; k < sh.ndisp; ++k)
    if (sh.dpos[k] == row) src = sh.drow[k];
  if (mine && row >= w) {
    if (src == row || g == 0 || G == 1) {
      // (workgroup 0, the owner of rows < w, reads them before writing)
      calu_load_row<W>(a, P, ld, src, w, vec);
    } else {
#pragma unroll
      for (int c = 0; c < W; ++c) {
        const size_t e = (size_t)src * W + c;
        const float re = __uint_as_float(get_u32(x.disp + e * 2, seq_base + 1u + CALU_TAG_DISP, err));
        const float im = __uint_as_float(get_u32(x.disp + e * 2 + 1, seq_base + 1u + CALU_TAG_DISP, err));
        a[c] = mkc<float>(re, im);
      }
    }
    // multipliers: l U11 = a, row by row of U11 (right-looking)
#pragma unroll
    for (int c = 0; c < W; ++c) {
      const cx<float> lc = cmul<float>(a[c], sh.rdiag[c]);
      a[c] = lc;
#pragma unroll
      for (int c2 = c + 1; c2 < W; ++c2) a[c2] = cfms_fused(a[c2], lc, sh.u[c][c2]);
    }
  } else if (mine) {
#pragma unroll
    for (int c = 0; c < W; ++c) a[c] = sh.u[row][c];
  }
  __syncthreads();   // workgroup 0: every read of rows < w is done
  if (mine) {
    if (vec) {
#pragma unroll
      for (int c = 0; c < W; c += 2) {
        if (c + 1 < w)
          *reinterpret_cast<float4*>(P + (int64_t)row * ld + c) =
              make_float4(a[c].x, a[c].y, a[c + 1].x, a[c + 1].y);
        else if (c < w)
          P[(int64_t)row * ld + c] = a[c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < W; ++c)
        if (c < w) P[(int64_t)row * ld + c] = a[c];
    }
  }
}

// ---------------------------------------------------------------------------
// Row interchanges k1 .. k1+nk-1 (global 0-based ipiv, ipiv[k] >= k) applied
// to the columns [c0, c1) (and [c2, c3)) of a row-major matrix.  Each workgroup takes a
// 128-byte column strip, stages every row the sequence touches in LDS (the nk
// rows k1.. and each distinct pivot row below them), replays the swaps in
// order there and writes the rows back: two coalesced passes instead of a
// chain of dependent global swaps.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256)
laswp_kernel(cx<T>* __restrict__ A, int64_t ld, int c0, int c1, int k1, int nk,
             const int* __restrict__ ipiv, const int* __restrict__ skip, int64_t sA = 0,
             int sIv = 0, int sSkip = 0, int prio = 0, int c2 = 0, int c3 = 0) {
  raise_wave_priority(prio);
  if (skip && skip[(int64_t)blockIdx.y * sSkip]) return;
  A += blockIdx.y * sA;
  ipiv += (int64_t)blockIdx.y * sIv;
  constexpr int CW = 128 / sizeof(cx<T>);
  extern __shared__ unsigned char smem_raw[];
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem_raw);  // [2 * nk][CW]
  int* slot = reinterpret_cast<int*>(buf + 2 * nk * CW);  // [nk]
  int* srcrow = slot + nk;                                 // [2 * nk]
  int* piv = srcrow + 2 * nk;                              // [nk]
  const int tid = threadIdx.x;
  // strips of [c0, c1), then of the optional second range [c2, c3)
  const int g1 = (c1 - c0 + CW - 1) / CW;
  const int cb = (int)blockIdx.x < g1 ? c0 + blockIdx.x * CW : c2 + (blockIdx.x - g1) * CW;
  const int ncol = min(CW, ((int)blockIdx.x < g1 ? c1 : c3) - cb);
  if (ncol <= 0) return;

  for (int t = tid; t < nk; t += 256) piv[t] = ipiv[k1 + t];
  __syncthreads();
  for (int t = tid; t < nk; t += 256) {
    int p = piv[t];
    int s;
    if (p < k1 + nk) {
      s = p - k1;
    } else {
      int first = t;
      for (int q = 0; q < t; ++q)
        if (piv[q] == p) { first = q; break; }
      s = nk + first;
    }
    slot[t] = s;
    srcrow[t] = k1 + t;
    srcrow[nk + t] = (s == nk + t) ? p : -1;
  }
  __syncthreads();
  for (int e = tid; e < 2 * nk * CW; e += 256) {
    int r = e / CW, c = e % CW;
    int gr = srcrow[r];
    if (gr >= 0 && c < ncol) buf[e] = A[(int64_t)gr * ld + cb + c];
  }
  __syncthreads();
  if (tid < ncol) {
    for (int t = 0; t < nk; ++t) {
      int s = slot[t];
      if (s != t) {
        cx<T> a = buf[t * CW + tid];
        buf[t * CW + tid] = buf[s * CW + tid];
        buf[s * CW + tid] = a;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < 2 * nk * CW; e += 256) {
    int r = e / CW, c = e % CW;
    int gr = srcrow[r];
    if (gr >= 0 && c < ncol) A[(int64_t)gr * ld + cb + c] = buf[e];
  }
}

// ---------------------------------------------------------------------------
// The net effect of the interchanges k1 .. k1+nk-1 (nk <= 256) as row moves:
// map[2 r] = destination row of staging position r, map[2 r + 1] = its source
// row (-1 / -1: nothing moves), r < 2 nk.  One workgroup: the staging slots
// of laswp_kernel, then one thread replays the swaps on slot indices instead
// of every workgroup of every interchange launch replaying them on data
// (256 dependent LDS swaps per column strip).  The moves then apply with
// laswp_apply_kernel to any column range.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
laswp_map_kernel(const int* __restrict__ ipiv, int k1, int nk, int* __restrict__ map) {
  __shared__ __align__(16) int piv[256];
  __shared__ __align__(16) int slot[256];
  __shared__ int srcrow[512], idx[512];
  const int tid = threadIdx.x;
  for (int t = tid; t < nk; t += 256) piv[t] = ipiv[k1 + t];
  for (int r = tid; r < 2 * nk; r += 256) idx[r] = r;
  __syncthreads();
  for (int t = tid; t < nk; t += 256) {
    const int p = piv[t];
    int sl;
    if (p < k1 + nk) {
      sl = p - k1;
    } else {
      // the first earlier pivot to the same row, four entries per LDS read
      // (entries at or beyond t are not compared)
      int first = t;
      for (int q0 = 0; q0 < t && first == t; q0 += 4) {
        const int4 v = *reinterpret_cast<const int4*>(piv + q0);
        if (v.x == p) first = q0;
        else if (v.y == p && q0 + 1 < t) first = q0 + 1;
        else if (v.z == p && q0 + 2 < t) first = q0 + 2;
        else if (v.w == p && q0 + 3 < t) first = q0 + 3;
      }
      sl = nk + first;
    }
    slot[t] = sl;
    srcrow[t] = k1 + t;
    srcrow[nk + t] = (sl == nk + t) ? p : -1;
  }
  __syncthreads();
  if (tid == 0)   // the slots of four swaps per LDS read, so each swap waits once
    for (int t0 = 0; t0 < nk; t0 += 4) {
      const int4 s4 = *reinterpret_cast<const int4*>(slot + t0);
      const int sv[4] = {s4.x, s4.y, s4.z, s4.w};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = t0 + u, sl = sv[u];
        if (t < nk && sl != t) {
          const int a = idx[t];
          idx[t] = idx[sl];
          idx[sl] = a;
        }
      }
    }
  __syncthreads();
  for (int r = tid; r < 2 * nk; r += 256) {
    const int dst = srcrow[r];
    const int src = dst >= 0 ? srcrow[idx[r]] : -1;
    const bool moves = dst >= 0 && src != dst;
    map[2 * r] = moves ? dst : -1;
    map[2 * r + 1] = moves ? src : -1;
  }
}

// Row moves of a laswp_map_kernel map (npos positions) over the columns
// [c0, c1): each workgroup takes 16 columns, sixteen lanes per row; every
// read lands in registers before the barrier and every write follows it (a
// row can be both a source and a destination).  No LDS staging of rows: many
// workgroups per CU, one memory round trip each way.
template <class T>
__global__ void __launch_bounds__(256)
laswp_apply_kernel(cx<T>* __restrict__ A, int64_t ld, int c0, int c1,
                   const int* __restrict__ map, int npos, int prio = 0) {
  raise_wave_priority(prio);
  constexpr int CW = 16, GR = 256 / CW, MAXP = 512 / GR;
  __shared__ int m[1024];
  const int tid = threadIdx.x;
  for (int e = tid; e < 2 * npos; e += 256) m[e] = map[e];
  __syncthreads();
  const int q = tid / CW, col = c0 + blockIdx.x * CW + tid % CW;
  const bool ok = col < c1;
  cx<T> v[MAXP];
#pragma unroll
  for (int i = 0; i < MAXP; ++i) {
    const int r = q + i * GR;
    if (ok && r < npos && m[2 * r] >= 0) v[i] = A[(int64_t)m[2 * r + 1] * ld + col];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MAXP; ++i) {
    const int r = q + i * GR;
    if (ok && r < npos && m[2 * r] >= 0) A[(int64_t)m[2 * r] * ld + col] = v[i];
  }
}

// ---------------------------------------------------------------------------
// In-place triangular solve with one diagonal block D (nb x nb, nb <= 64) of a
// row-major factor: lower unit (forward) or upper non-unit (backward), on the
// nb x ncols block row S of a row-major right-hand side.  Each workgroup
// stages D and a 16-column strip of S in LDS and substitutes there.
// ---------------------------------------------------------------------------
template <class T, bool UPPER>
__global__ void __launch_bounds__(256)
trsm_diag_kernel(const cx<T>* __restrict__ D, int64_t ldd, int nb, cx<T>* __restrict__ S,
                 int64_t lds, int ncols, const int* __restrict__ skip, int64_t sA = 0,
                 int prio = 0) {
  raise_wave_priority(prio);
  if (skip && *skip) return;
  D += blockIdx.y * sA;
  S += blockIdx.y * sA;
  constexpr int CW = 16, NBM = 64;
  __shared__ cx<T> d[NBM][NBM + 1];
  __shared__ cx<T> s[NBM][CW];
  __shared__ cx<T> rdiag[NBM];
  const int tid = threadIdx.x;
  const int cb = blockIdx.x * CW;
  const int ncol = min(CW, ncols - cb);
  for (int e = tid; e < nb * nb; e += 256) {
    int i = e / nb, k = e % nb;
    bool need = UPPER ? (k >= i) : (k < i);
    d[i][k] = need ? D[(int64_t)i * ldd + k] : mkc<T>(0, 0);
  }
  for (int e = tid; e < nb * CW; e += 256) {
    int i = e / CW, c = e % CW;
    s[i][c] = (c < ncol) ? S[(int64_t)i * lds + cb + c] : mkc<T>(0, 0);
  }
  __syncthreads();
  if (UPPER) {
    if (tid < nb) rdiag[tid] = crecip<T>(d[tid][tid]);
    __syncthreads();
    for (int i = nb - 1; i >= 0; --i) {
      if (tid < CW) s[i][tid] = cmul<T>(s[i][tid], rdiag[i]);
      __syncthreads();
      for (int e = tid; e < i * CW; e += 256) {
        int r = e / CW, c = e % CW;
        s[r][c] = cfms<T>(s[r][c], d[r][i], s[i][c]);
      }
      __syncthreads();
    }
  } else {
    for (int i = 0; i < nb - 1; ++i) {
      int rows = nb - 1 - i;
      for (int e = tid; e < rows * CW; e += 256) {
        int r = i + 1 + e / CW, c = e % CW;
        s[r][c] = cfms<T>(s[r][c], d[r][i], s[i][c]);
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < nb * CW; e += 256) {
    int i = e / CW, c = e % CW;
    if (c < ncol) S[(int64_t)i * lds + cb + c] = s[i][c];
  }
}

// ---------------------------------------------------------------------------
// Final row permutation of a factorization: perm[i] = source row of row i of
// P A, from the interchange sequence ipiv (LAPACK order).  One workgroup;
// the array lives in LDS when n fits, ipiv is staged in chunks so the single
// sequential swap loop never waits on global memory.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
perm_from_ipiv_kernel(const int* __restrict__ ipiv, int n, int* __restrict__ perm, int in_lds,
                      int sIv = 0) {
  ipiv += (int64_t)blockIdx.x * sIv;   // one workgroup per system
  perm += (int64_t)blockIdx.x * sIv;
  extern __shared__ int sh_perm[];
  __shared__ __align__(16) int chunk[1024];
  int* p = in_lds ? sh_perm : perm;
  for (int i = threadIdx.x; i < n; i += 256) p[i] = i;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += 1024) {
    const int cn = min(1024, n - c0);
    for (int i = threadIdx.x; i < cn; i += 256) chunk[i] = ipiv[c0 + i];
    __syncthreads();
    if (threadIdx.x == 0) {   // four interchange rows per LDS read: one wait per swap
      for (int i0 = 0; i0 < cn; i0 += 4) {
        const int4 k4 = *reinterpret_cast<const int4*>(chunk + i0);
        const int kv[4] = {k4.x, k4.y, k4.z, k4.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int i = i0 + u, k = kv[u], a = c0 + i;
          if (i < cn && k != a) { int t = p[a]; p[a] = p[k]; p[k] = t; }
        }
      }
    }
    __syncthreads();
  }
  if (in_lds)
    for (int i = threadIdx.x; i < n; i += 256) perm[i] = sh_perm[i];
}

// out[i][:] = (Tout) (s[perm[i]] in[perm[i]][:]) for row-major n x nrhs arrays
// (s = null: no scaling).
template <class Tin, class Tout>
__global__ void __launch_bounds__(256)
permute_cast_kernel(int n, int nrhs, const int* __restrict__ perm, const cx<Tin>* __restrict__ in,
                    cx<Tout>* __restrict__ out, const double* __restrict__ scale,
                    const int* __restrict__ skip, int sIv = 0, int64_t sV = 0, int sSkip = 0) {
  {
    const int b = blockIdx.y;   // system of a batch: sIv perm/scale entries, sV elements
    if (skip && skip[(int64_t)b * sSkip]) return;
    perm += (int64_t)b * sIv;
    if (scale) scale += (int64_t)b * sIv;
    in += b * sV;
    out += b * sV;
  }
  const size_t total = (size_t)n * nrhs;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const int i = (int)(e / nrhs), c = (int)(e % nrhs);
    const int src = perm[i];
    cx<Tin> v = in[(size_t)src * nrhs + c];
    const double f = scale ? scale[src] : 1.0;
    out[e] = mkc<Tout>((Tout)(v.x * f), (Tout)(v.y * f));
  }
}

// ---------------------------------------------------------------------------
// Inverses of the 64 x 64 diagonal blocks of a row-major LU factor, one block
// per 64-thread workgroup: blockIdx.y = 0 -> inverse of the unit lower
// triangle, 1 -> inverse of the upper triangle.  Output: dinv[(b*2+y)*4096 +
// i*64 + k], zero outside the triangle and beyond a partial block (unit
// diagonal for L).
//
// The inverse X is formed from X T = I (LAPACK ztrti2's equations, whose
// small left residual makes X T_12 behave like a triangular solve): row i of
// X depends only on row i and on T, so lane i forms row i on its own, with
// no barrier between the 64 steps.  Upper: X_ii = 1 / U_ii, then for j > i
// X_ij = -X_jj sum_{i <= k < j} X_ik U_kj.  Unit lower: for j < i, from the
// right, X_ij = -(L_ij + sum_{j < k < i} X_ik L_kj).  The sums run in
// ztrti2's order; the lanes run the same loops, the zeros of row i outside
// its triangle adding exact zeros.  The triangle is staged in LDS and read
// as broadcasts; the rows of X are kept transposed in LDS (xt[k][i] = X_ik,
// lane i's column), so every lane's reads are conflict-free.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(64)
diag_inv_kernel(const cx<T>* __restrict__ LU, int64_t ld, int n, int blk0,
                cx<T>* __restrict__ dinv, int64_t sA = 0, int64_t sD = 0, int prio = 0,
                int part0 = 0) {
  raise_wave_priority(prio);
  LU += blockIdx.z * sA;   // system of a batch
  dinv += blockIdx.z * sD;
  __shared__ cx<T> tr[64][65];   // the triangle, zero elsewhere
  __shared__ cx<T> xt[64][65];   // xt[k][i] = X_ik (padded: the output pass reads rows)
  // part0 + blockIdx.y: 0 the unit-lower factor, 1 the upper factor
  const int bi = blk0 + blockIdx.x, upper = part0 + blockIdx.y, i = threadIdx.x;
  const int r0 = bi * 64, bs = min(64, n - r0);
  // 16 rows in flight per pass (one wave: the loads would otherwise wait
  // for one another)
  for (int r1 = 0; r1 < 64; r1 += 16) {
    cx<T> v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int r = r1 + u;
      const bool in = r < bs && i < bs && (upper ? (i >= r) : (i < r));
      v[u] = in ? LU[(int64_t)(r0 + r) * ld + r0 + i] : mkc<T>(0, 0);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) tr[r1 + u][i] = v[u];
  }
  for (int k = 0; k < 64; ++k) xt[k][i] = mkc<T>(0, 0);
  __syncthreads();
  auto fma_c = [](cx<T> s, cx<T> a, cx<T> b) {   // s + a b
    return mkc<T>(s.x + a.x * b.x - a.y * b.y, s.y + a.x * b.y + a.y * b.x);
  };
  if (upper) {
    for (int j = 0; j < bs; ++j) {
      const cx<T> xjj = crecip<T>(tr[j][j]);
      if (i == j) xt[j][i] = xjj;
      cx<T> s = mkc<T>(0, 0);
#pragma unroll 8
      for (int k = 0; k < j; ++k) s = fma_c(s, xt[k][i], tr[k][j]);
      if (i < j) xt[j][i] = mkc<T>(-(xjj.x * s.x - xjj.y * s.y), -(xjj.x * s.y + xjj.y * s.x));
    }
  } else {
    // the diagonal of X stays zero until the end, so that the sums stop at k < i
    for (int j = bs - 2; j >= 0; --j) {
      cx<T> s = tr[i][j];
#pragma unroll 8
      for (int k = j + 1; k < bs; ++k) s = fma_c(s, xt[k][i], tr[k][j]);
      if (i > j && i < bs) xt[j][i] = mkc<T>(-s.x, -s.y);
    }
    if (i < bs) xt[i][i] = mkc<T>(1, 0);
  }
  __syncthreads();
  // out[r][k] = X_rk = xt[k][r]: lane i writes column i (coalesced rows)
  cx<T>* out = dinv + ((int64_t)bi * 2 + upper) * 4096;
  for (int r = 0; r < 64; ++r) out[r * 64 + i] = xt[i][r];
}

// C[16 w + i][j] += sum_k A[16 w + i][k] B[k][j] (j < 16) for one 64 x 64 x 16
// complex block product on MFMA (wave w: one 16 x 16 tile), operands read
// from global memory (small, L2-resident blocks); A is zero beyond arows x
// acols.  Accumulator element r of lane l is row mfma_row(l/16, r), column
// l % 16.
template <class T>
__device__ inline void prod64x16_g(const cx<T>* __restrict__ A, int64_t lda, int arows,
                                   int acols, const cx<T>* __restrict__ B, int64_t ldb,
                                   typename acc_t<T>::type& re, typename acc_t<T>::type& im) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  const int i = 16 * w + li;
  // all 32 loads in flight at once: the chain of products in
  // blk_inv_kernel is bound by their latency
  cx<T> av[16], bv[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int k = 4 * s + lk;
    av[s] = (i < arows && k < acols) ? A[(int64_t)i * lda + k] : mkc<T>(0, 0);
    bv[s] = B[(int64_t)k * ldb + li];
  }
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const cx<T> a = av[s], b = bv[s];
    re = mfma16x4(a.x, b.x, re);
    re = mfma16x4(-a.y, b.y, re);
    im = mfma16x4(a.x, b.y, im);
    im = mfma16x4(a.y, b.x, im);
  }
}

// ---------------------------------------------------------------------------
// Inverse of one 256 x 256 diagonal block of a row-major LU factor from the
// 64 x 64 inverses of its diagonal sub-blocks (block forward / backward
// substitution on MFMA), one 16-column slice of a 64-column block column per
// workgroup (grid: blocks x {L, U} x 16: z = 4 jb + q, slice q of block
// column jb; the columns of X are independent):
//   L: X_jj = inv(L_jj),  X_ij = -inv(L_ii) sum_{k=j}^{i-1} L_ik X_kj  (i > j)
//   U: X_jj = inv(U_jj),  X_ij = -inv(U_ii) sum_{k=i+1}^{j} U_ik X_kj  (i < j)
// Output dinv256[(b*2+u)*65536 + r*256 + c], zero outside the triangle and
// outside a trailing partial block.
//
// The kernel runs on the panel's critical path while the previous trailing
// update occupies the CUs, so it uses no LDS and few registers: it can start
// wherever one GEMM workgroup retires.  The intermediate sum is parked in
// the output block X_ij itself.  A workgroup only reads X entries of its own
// slice, written by itself: barriers (workgroup-scope release/acquire; the
// waves share the CU's L1) order them.  The longest chain (j = 0 of L) is
// nine dependent products of 64 x 64 x 16.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256)
blk_inv_kernel(const cx<T>* __restrict__ LU, int64_t ld, int n, int bb0,
               const cx<T>* __restrict__ dinv64, cx<T>* __restrict__ dinv256, int64_t sA = 0,
               int64_t sD64 = 0, int64_t sD256 = 0, int prio = 0, int part0 = 0) {
  raise_wave_priority(prio);
  typedef typename acc_t<T>::type acc;
  // part0 + blockIdx.y: 0 the unit-lower factor, 1 the upper factor
  const int bb = bb0 + blockIdx.x, upper = part0 + blockIdx.y;
  const int jb = (blockIdx.z >> 2) & 3, q = blockIdx.z & 3;
  {
    const int b = blockIdx.z >> 4;   // system of a batch (z = 16 b + 4 jb + q)
    LU += b * sA;
    dinv64 += b * sD64;
    dinv256 += b * sD256;
  }
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
  const int r0 = bb * 256, kb = min(256, n - r0), nsb = (kb + 63) / 64;
  const int c0 = jb * 64 + q * 16;   // first column of the slice
  cx<T>* X = dinv256 + ((int64_t)bb * 2 + upper) * 65536;
  for (int ib = 0; ib < 4; ++ib) {
    const bool zero = jb >= nsb || ib >= nsb || (upper ? ib > jb : ib < jb);
    if (!zero) continue;
    for (int e = tid; e < 1024; e += 256) X[(ib * 64 + e / 16) * 256 + c0 + e % 16] = mkc<T>(0, 0);
  }
  if (jb >= nsb) return;
  auto inv64 = [&](int sb) { return dinv64 + ((int64_t)(bb * 4 + sb) * 2 + upper) * 4096; };
  {
    const cx<T>* src = inv64(jb) + q * 16;
    for (int e = tid; e < 1024; e += 256)
      X[(jb * 64 + e / 16) * 256 + c0 + e % 16] = src[(e / 16) * 64 + e % 16];
  }
  __syncthreads();
  const int step = upper ? -1 : 1;
  for (int ib = jb + step; ib >= 0 && ib < nsb; ib += step) {
    acc re = acc{}, im = acc{};
    const int k_lo = upper ? ib + 1 : jb, k_hi = upper ? jb : ib - 1;
    for (int kk = k_lo; kk <= k_hi; ++kk)
      prod64x16_g<T>(LU + (int64_t)(r0 + 64 * ib) * ld + r0 + 64 * kk, ld,
                     min(64, kb - 64 * ib), min(64, kb - 64 * kk),
                     X + (int64_t)(64 * kk) * 256 + c0, 256, re, im);
    cx<T>* Xij = X + (int64_t)(64 * ib) * 256 + c0;
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * w + mfma_row<T>(lk, r);
      Xij[(int64_t)row * 256 + li] = mkc<T>(re[r], im[r]);
    }
    __syncthreads();
    re = acc{};
    im = acc{};
    prod64x16_g<T>(inv64(ib), 64, 64, 64, Xij, 256, re, im);
    __syncthreads();   // every wave has read the parked sum before it is overwritten
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * w + mfma_row<T>(lk, r);
      Xij[(int64_t)row * 256 + li] = mkc<T>(-re[r], -im[r]);
    }
    __syncthreads();
  }
}

// dst[r * ldd + c] = src[r * lds + c] for r < rows, c < cols.
template <class T>
__global__ void __launch_bounds__(256)
copy2d_kernel(int rows, int cols, const cx<T>* __restrict__ src, int64_t lds, cx<T>* __restrict__ dst,
              int64_t ldd, int64_t sSrc = 0, int64_t sDst = 0) {
  src += blockIdx.y * sSrc;
  dst += blockIdx.y * sDst;
  const size_t total = (size_t)rows * cols;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const int r = (int)(e / cols), c = (int)(e % cols);
    dst[(int64_t)r * ldd + c] = src[(int64_t)r * lds + c];
  }
}
