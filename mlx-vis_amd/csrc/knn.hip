// Exact brute-force kNN for gfx950 (MI355X).
//
// Replaces the brute path of compute_knn — mlx_vis/_knn.py:47-88. The
// reference forms the full (chunk x n) matrix max(|x_i|^2 + |x_j|^2 - 2 x_i.x_j, 0)
// (_knn.py:60-63), masks the diagonal with +1e30 (_knn.py:64-69), argsorts every
// row and keeps k columns (_knn.py:71-72), then takes sqrt (_knn.py:86-87).
// Here no row is ever sorted, and the n x n matrix is materialised only by the
// symmetric screen (self kNN up to 48 GiB of matrix), where it halves the
// multiplies:
//
//   1. prepass     column means (FP64 partial sums), then a centred, zero-padded
//                  copy of X (and of Q when Q != X) with FP32 row norms and the
//                  largest norm. Centring is exact in real arithmetic (distances
//                  are translation invariant) and shrinks the screen's error bound.
//   2. screen      MFMA tiles of -2 Q X^T (v_mfma_f32_16x16x4_f32, 128x128 per
//                  workgroup, LDS-staged 32-wide K chunks, double buffered), the
//                  distance epilogue, and a fused per-row running top-C list
//                  (C = rerank*k <= 64) kept in LDS and updated by wavefront
//                  ballot/shuffle insertion. Columns are split S ways; the
//                  blockIdx -> (row block, split) map keeps one split per XCD so
//                  the 32 CUs of an XCD stream the same X tiles through its L2.
//   2'. symmetric  self kNN (Q = X): MFMA tiles on and above the diagonal only
//                  (k_dist_sym, 8 wavefronts per 128x128 tile), each stored with
//                  its transpose into the n_pad x n_pad matrix together with
//                  per-row minima of every 128-column block; k_select then keeps
//                  each row's C smallest keys, reading only the blocks whose
//                  minimum can hold one. The screen values are the fused
//                  screen's (FP8: up to the order of the two row scales in
//                  the transposed half), and the re-rank makes the results
//                  identical.
//   3. rerank      one wavefront per query row: merge the S lists into the global
//                  top-C, FP32 direct-difference distances of the C candidates,
//                  FP64 (sequential, unfused) only where two FP32 values cannot be
//                  separated by their error bound, and a certificate that no row
//                  outside the candidates can enter the top k.
//   4. exact scan  rows without a certificate are recomputed exhaustively in
//                  FP64 (rare; guarantees the canonical result).
#include <math.h>

#include <mutex>

#include "common.hpp"
#include "lowd.hpp"

namespace mv {
namespace knn {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int BM = 128;               // query rows per workgroup
constexpr int BN = 128;               // X rows (columns of D) per tile
constexpr int KC = 32;                // K chunk (floats) staged per step
constexpr int LDSA = KC + 4;          // padded LDS row stride (floats): conflict-free b128 reads
constexpr int DTS = BN + 4;           // padded stride of the D tile in LDS
constexpr int CMAX = 64;              // max candidates per row of the fused screen (one per lane)
constexpr int CMAX_SYM = 128;         // max candidates per row of the symmetric screen (two per lane)
constexpr int NT = 256;               // threads per workgroup
constexpr int ROW_CHUNK = 256;        // rows per column-sum partial
constexpr int MAX_SPLIT = 8;
constexpr uint64_t KEY_NONE = ~0ull;

constexpr int STAGE_FLOATS = 2 * 2 * BM * LDSA;   // two buffers x (A, B)
static_assert(BM * DTS <= STAGE_FLOATS, "D tile must fit in the staging area");

// ---------------------------------------------------------------- prepass --

__global__ void k_zero_i32(int32_t* p, int count) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) p[t] = 0;
}


// part[c, t] = sum over the rows of chunk c (rows c*rows_per .. ) of X[r, t]
// in FP64. A block covers 64 columns x 4 row lanes; the row lanes are summed
// through LDS.
__global__ __launch_bounds__(256) void k_colsum_partial(const float* __restrict__ X, int64_t n, int64_t d,
                                                        int64_t rows_per, double* __restrict__ part) {
  __shared__ double s_p[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * 64 + tx;
  const int64_t c = blockIdx.y;
  const int64_t r0 = c * rows_per;
  const int64_t r1 = r0 + rows_per < n ? r0 + rows_per : n;
  double acc = 0.0;
  if (t < d)
    for (int64_t r = r0 + ty; r < r1; r += 4) acc += static_cast<double>(X[r * d + t]);
  s_p[ty][tx] = acc;
  __syncthreads();
  if (ty == 0 && t < d) part[c * d + t] = s_p[0][tx] + s_p[1][tx] + s_p[2][tx] + s_p[3][tx];
}

// mu[t] = (sum over chunks of part[., t]) / n, one wavefront per column;
// padding columns get 0.
__global__ __launch_bounds__(256) void k_colsum_final(const double* __restrict__ part, int64_t nchunks,
                                                      int64_t n, int64_t d, int64_t d_pad,
                                                      float* __restrict__ mu) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (t >= d_pad) return;
  if (t >= d) {
    if (l == 0) mu[t] = 0.f;
    return;
  }
  double acc = 0.0;
  for (int64_t c = l; c < nchunks; c += 64) acc += part[c * d + t];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m);
  if (l == 0) mu[t] = static_cast<float>(acc / static_cast<double>(n));
}

// Largest of a non-negative float over a block (bit patterns order like the
// values), then one atomicMax per block.
__device__ __forceinline__ void block_max_bits(uint32_t bits, uint32_t* dst) {
  __shared__ uint32_t s_m[4];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const uint32_t o = static_cast<uint32_t>(__shfl_xor(static_cast<int>(bits), m));
    bits = o > bits ? o : bits;
  }
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t v = s_m[0];
    for (int w = 1; w < 4; ++w) v = s_m[w] > v ? s_m[w] : v;
    if (dst) atomicMax(dst, v);
  }
}

// Centred, zero-padded FP32 copy: one wavefront per row, rows strided over a
// bounded grid. Rows in [n, n_pad) become zero rows. sqmax receives the bit
// pattern of the largest row norm (one atomic per block).
__global__ __launch_bounds__(256) void k_center_f32(const float* __restrict__ X, int64_t n, int64_t d,
                                                    int64_t d_pad, int64_t n_pad,
                                                    const float* __restrict__ mu, float* __restrict__ Xc,
                                                    float* __restrict__ sq, uint32_t* __restrict__ sqmax) {
  const int l = threadIdx.x & 63;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * 4;
  uint32_t mbits = 0;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6); r < n_pad; r += waves) {
    float acc = 0.f;
    float* out = Xc + r * d_pad;
    if (r < n) {
      const float* in = X + r * d;
      for (int64_t t = l; t < d_pad; t += 64) {
        float v = t < d ? in[t] - mu[t] : 0.f;
        out[t] = v;
        acc = fmaf(v, v, acc);
      }
    } else {
      for (int64_t t = l; t < d_pad; t += 64) out[t] = 0.f;
    }
    acc = wave_sum_f32(acc);
    if (l == 0) sq[r] = acc;
    if (r < n) mbits = __float_as_uint(acc) > mbits ? __float_as_uint(acc) : mbits;
  }
  block_max_bits(mbits, sqmax);
}

__global__ __launch_bounds__(256) void k_row_sqnorms(const float* __restrict__ X, int64_t n, int64_t d,
                                                     float* __restrict__ out) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (r >= n) return;
  const float* in = X + r * d;
  float acc = 0.f;
  for (int64_t t = l; t < d; t += 64) acc = fmaf(in[t], in[t], acc);
  acc = wave_sum_f32(acc);
  if (l == 0) out[r] = acc;
}

enum { DT_F32 = 0, DT_BF16 = 1, DT_FP8 = 2, DT_BF16X3 = 3 };
// Bytes per element of the screen operand (the split BF16 form stores two
// BF16 values per element).
__host__ __device__ constexpr int elem_bytes(int dt) {
  return dt == DT_F32 || dt == DT_BF16X3 ? 4 : dt == DT_BF16 ? 2 : 1;
}
// Scale operand of the FP8 MFMA that leaves products unscaled.
constexpr int FP8_UNIT_SCALE = 127;

__device__ __forceinline__ uint16_t bf16_rne(float v) {
  uint32_t u = __float_as_uint(v);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// Centred, quantised, zero-padded copy for the BF16 / FP8 screens, one
// wavefront per row. Per row: sq = |x_hat|^2 of the dequantised vector,
// err = an upper bound of |x_c - x_hat| (x_c the FP32 centred row), and for FP8
// inv = 1 / scale with scale = 448 / max|x_c| (per-row absmax). sqmax / errmax
// receive the bit patterns of the largest sq / err over rows < n.
template <int DT>
__global__ __launch_bounds__(256) void k_center_q(const float* __restrict__ X, int64_t n, int64_t d, int64_t d_pad,
                                                  int64_t n_pad, const float* __restrict__ mu, char* __restrict__ Xq,
                                                  float* __restrict__ sq, float* __restrict__ err,
                                                  float* __restrict__ inv, uint32_t* __restrict__ sqmax,
                                                  uint32_t* __restrict__ errmax) {
  const int l = threadIdx.x & 63;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * 4;
  uint32_t sbits = 0, ebits = 0;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6); r < n_pad; r += waves) {
    const bool live = r < n;
    const float* in = X + r * d;
    float scale = 1.f;
    if (DT == DT_FP8 && live) {
      float amax = 0.f;
      for (int64_t t = l; t < d; t += 64) amax = fmaxf(amax, fabsf(in[t] - mu[t]));
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) amax = fmaxf(amax, __shfl_xor(amax, m));
      scale = amax > 0.f ? 448.f / amax : 1.f;
    }
    const float iv = 1.f / scale;
    float acc = 0.f, e2 = 0.f, v2 = 0.f;
    for (int64_t t = l; t < d_pad; t += 64) {
      const float v = (live && t < d) ? in[t] - mu[t] : 0.f;
      float xh;
      if (DT == DT_BF16) {
        const uint16_t q = bf16_rne(v);
        reinterpret_cast<uint16_t*>(Xq + r * d_pad * 2)[t] = q;
        xh = bf16_to_f32(q);
      } else {
        const int pk = __builtin_amdgcn_cvt_pk_fp8_f32(v * scale, 0.f, 0, false);
        reinterpret_cast<uint8_t*>(Xq + r * d_pad)[t] = static_cast<uint8_t>(pk & 0xFF);
        xh = __builtin_amdgcn_cvt_f32_fp8(pk, 0) * iv;
      }
      acc = fmaf(xh, xh, acc);
      const float dv = v - xh;
      e2 = fmaf(dv, dv, e2);
      v2 = fmaf(v, v, v2);
    }
    acc = wave_sum_f32(acc);
    e2 = wave_sum_f32(e2);
    v2 = wave_sum_f32(v2);
    // sqrt of the FP32 sum with 0.1 % slack, plus the centring rounding bound
    const float eb = live ? sqrtf(e2) * 1.001f + 1.2e-7f * sqrtf(v2) + 1e-20f : 0.f;
    if (l == 0) {
      sq[r] = acc;
      err[r] = eb;
      if (inv) inv[r] = iv;
    }
    if (live) {
      sbits = __float_as_uint(acc) > sbits ? __float_as_uint(acc) : sbits;
      ebits = __float_as_uint(eb) > ebits ? __float_as_uint(eb) : ebits;
    }
  }
  block_max_bits(sbits, sqmax);
  __syncthreads();
  block_max_bits(ebits, errmax);
}

// Split BF16 copy of the centred FP32 rows: v = hi + lo + r with hi = bf16(v),
// lo = bf16(v - hi) (both subtractions exact in FP32), |r| <= 2^-16 |v|. Row
// layout per 32-element group g (128 bytes): hi[32g .. 32g+31] then
// lo[32g .. 32g+31], so one 128-byte chunk of the screen holds one K group.
// gate (device, nullable): nothing is written while *gate == 0 (the re-screen
// preparation of a call in which every row was certified).
__global__ __launch_bounds__(256) void k_split_bf16(const float* __restrict__ Xc, int64_t count, int64_t d_pad,
                                                    uint16_t* __restrict__ Xs, const int32_t* __restrict__ gate) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= count || (gate && *gate == 0)) return;
  const int64_t r = t / d_pad, c = t % d_pad;
  const float v = Xc[t];
  const uint16_t hi = bf16_rne(v);
  const uint16_t lo = bf16_rne(v - bf16_to_f32(hi));
  uint16_t* row = Xs + r * 2 * d_pad + (c >> 5) * 64 + (c & 31);
  row[0] = hi;
  row[32] = lo;
}

// Gathers the FP32 centred rows listed in rows[0:*count] (original query ids)
// for the re-screen, with their norms and self indices.
// A few rows that the quantised certificate could not settle (at most
// `small`) go straight to the exhaustive FP64 scan: a re-screen costs a full
// pass over X whatever its row count. Sets the re-screen's row count (0 then)
// and, for the direct route, the scan's row list.
__global__ void k_route_small(const int32_t* __restrict__ fb1_count, const int32_t* __restrict__ fb1_rows,
                              int32_t* __restrict__ fb2_count, int32_t* __restrict__ fb2_rows,
                              int32_t* __restrict__ rs_count, int32_t* __restrict__ stats, int small) {
  const int c = *fb1_count;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (c <= small) {
    if (t < c) fb2_rows[t] = fb1_rows[t];
    if (t == 0) {
      *fb2_count = c;
      *rs_count = 0;
      if (c > 0) atomicAdd(&stats[3], c);
    }
  } else if (t == 0) {
    *rs_count = c;
  }
}

__global__ __launch_bounds__(256) void k_gather_rows(const float* __restrict__ Qc, const float* __restrict__ sqq,
                                                     int64_t d_pad, const int32_t* __restrict__ rows,
                                                     const int32_t* __restrict__ count,
                                                     const int32_t* __restrict__ q_self, int64_t self_base,
                                                     float* __restrict__ out, float* __restrict__ sq_out,
                                                     int32_t* __restrict__ self_out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (i >= *count) return;
  const int64_t r = rows[i];
  for (int64_t t = l; t < d_pad; t += 64) out[i * d_pad + t] = Qc[r * d_pad + t];
  if (l == 0) {
    sq_out[i] = sqq[r];
    self_out[i] = q_self ? q_self[r] : (self_base >= 0 ? static_cast<int32_t>(self_base + r) : -1);
  }
}

// ----------------------------------------------------------------- screen --

struct ScreenArgs {
  const char* Qq;         // [nq_pad, d_pad] centred (quantised) queries
  const char* Xq;         // [n_pad, d_pad] centred (quantised) data
  const float* sqq;       // [nq_pad] norms of the (dequantised) queries
  const float* sqx;       // [n_pad]
  const float* invq;      // FP8: per-row 1/scale of the queries (else unused)
  const float* invx;      // FP8: per-row 1/scale of the data
  const int32_t* q_self;  // nullable
  int64_t self_base;      // used when q_self == NULL; < 0: no exclusion
  int64_t row_bytes;      // bytes per padded row (multiple of 128)
  const int32_t* nq_dev;  // nullable: device-side query count (re-screen)
  int nq, n, C, S;
  int tiles_per_split, n_col_tiles, n_row_blocks, nq_pad;
  uint64_t* cand;         // [S, nq_pad, C] keys
  uint64_t* pend;         // [S, nq_pad, PEND] survivors not yet merged
};

// Survivors of the screen wait in a per-row buffer of PEND keys (global
// memory next to the candidate lists) and are merged into the row's sorted
// list PEND at a time, with one wavefront sort, instead of one insertion each.
constexpr int PEND = 64;

// Appends the keys of a wavefront's two column halves (KEY_NONE = no key;
// m0, m1 = their ballots) at positions base.. of buf, lane order, half 0 first.
__device__ __forceinline__ void pend_append(uint64_t* buf, int base, uint64_t k0, uint64_t m0, uint64_t k1,
                                            uint64_t m1) {
  const int p0 = base + static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                            static_cast<uint32_t>(m0 >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m0), 0u)));
  const int p1 = base + __popcll(m0) +
                 static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m1 >> 32),
                                                            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m1), 0u)));
  if (k0 != KEY_NONE) buf[p0] = k0;
  if (k1 != KEY_NONE) buf[p1] = k1;
}

// Merges unsorted keys (one per lane, KEY_NONE padded) into the row's list of
// C keys and returns the list's new C-th key.
__device__ __forceinline__ uint64_t list_absorb(uint64_t& lk, uint64_t keys, int C, int l) {
  lk = list_merge64_u64(lk, keys);
  if (l >= C) lk = KEY_NONE;
  return readlane_u64(lk, C - 1);
}

// The same for a list of up to 128 keys held in two registers: lk0 the
// ranks 0..63, lk1 the ranks 64..127, each ascending in lane order. With the
// batch sorted descending and padded in front with 64 KEY_NONE, the position-
// wise minimum of the list and the batch holds the 128 smallest of both and
// is bitonic (Batcher); one compare-exchange between the registers and a
// half-cleaner sort of each register sort it.
__device__ __forceinline__ uint64_t list_absorb2(uint64_t& lk0, uint64_t& lk1, uint64_t keys, int C, int l) {
  const uint64_t b = wave_sort_u64<true>(keys);
  const uint64_t t1 = lk1 < b ? lk1 : b;
  const uint64_t lo = lk0 < t1 ? lk0 : t1, hi = lk0 < t1 ? t1 : lk0;
  lk0 = bitonic_clean_u64(lo);
  lk1 = bitonic_clean_u64(hi);
  if (l >= C) lk0 = KEY_NONE;
  if (64 + l >= C) lk1 = KEY_NONE;
  return C > 64 ? readlane_u64(lk1, C - 65) : readlane_u64(lk0, C - 1);
}

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef int v8i __attribute__((ext_vector_type(8)));

// One 128-byte K chunk of the staged A (rows) and B (columns) tiles into the
// 4x4 grid of 16x16 accumulators of this wavefront.
template <int DT>
__device__ __forceinline__ void mfma_chunk(const float* As, const float* Bs, int wr, int wc, int l,
                                           v4f (&acc)[4][4]) {
  if constexpr (DT == DT_F32) {
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      v4f af[4], bf[4];
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
        af[mi] = *reinterpret_cast<const v4f*>(As + (wr * 64 + mi * 16 + (l & 15)) * LDSA + g * 16 + (l >> 4) * 4);
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
        bf[ni] = *reinterpret_cast<const v4f*>(Bs + (wc * 64 + ni * 16 + (l & 15)) * LDSA + g * 16 + (l >> 4) * 4);
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
          for (int ni = 0; ni < 4; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[mi][j], bf[ni][j], acc[mi][ni], 0, 0, 0);
    }
  } else if constexpr (DT == DT_BF16X3) {
    // one K group of 32: hi in floats [0, 16) of the row, lo in [16, 32)
    v8bf ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      const float* p = As + (wr * 64 + mi * 16 + (l & 15)) * LDSA + (l >> 4) * 4;
      ah[mi] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4f*>(p));
      al[mi] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4f*>(p + 16));
    }
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      const float* p = Bs + (wc * 64 + ni * 16 + (l & 15)) * LDSA + (l >> 4) * 4;
      bh[ni] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4f*>(p));
      bl[ni] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4f*>(p + 16));
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) {
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[mi], bh[ni], acc[mi][ni], 0, 0, 0);
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mi], bl[ni], acc[mi][ni], 0, 0, 0);
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mi], bh[ni], acc[mi][ni], 0, 0, 0);
      }
  } else if constexpr (DT == DT_BF16) {
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      v8bf af[4], bf[4];
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
        af[mi] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4f*>(
                                              As + (wr * 64 + mi * 16 + (l & 15)) * LDSA + g * 16 + (l >> 4) * 4));
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
        bf[ni] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4f*>(
                                              Bs + (wc * 64 + ni * 16 + (l & 15)) * LDSA + g * 16 + (l >> 4) * 4));
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mi], bf[ni], acc[mi][ni], 0, 0, 0);
    }
  } else {
    v8i af[4], bf[4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      const float* p = As + (wr * 64 + mi * 16 + (l & 15)) * LDSA + (l >> 4) * 8;
      const v4f lo = *reinterpret_cast<const v4f*>(p), hi = *reinterpret_cast<const v4f*>(p + 4);
      af[mi] = __builtin_bit_cast(v8i, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
    }
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      const float* p = Bs + (wc * 64 + ni * 16 + (l & 15)) * LDSA + (l >> 4) * 8;
      const v4f lo = *reinterpret_cast<const v4f*>(p), hi = *reinterpret_cast<const v4f*>(p + 4);
      bf[ni] = __builtin_bit_cast(v8i, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
        acc[mi][ni] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(af[mi], bf[ni], acc[mi][ni], 0, 0, 0,
                                                                        FP8_UNIT_SCALE, 0, FP8_UNIT_SCALE);
  }
}

template <int DT>
__global__ __launch_bounds__(NT, 2) void k_screen(ScreenArgs a) {
  __shared__ __attribute__((aligned(16))) float s_stage[STAGE_FLOATS];
  __shared__ float4 s_row[BM];          // |q|^2, FP8 1/scale, C-th list distance (+inf while
                                        // the list is not full), self column (int bits)
  __shared__ uint64_t s_tauk[BM];       // C-th list key (KEY_NONE while the list is not full)
  __shared__ int s_cnt[BM];             // slots of the row's buffer reserved this tile
  __shared__ int s_valid[BM];           // keys waiting in the row's buffer
  __shared__ uint64_t s_hit[2][2];      // [column half][row half]: rows with survivors

  // Workgroups are dispatched round-robin over the 8 XCDs (bid mod 8); XCD x
  // works on column split x mod S, so the co-resident workgroups of an XCD
  // stream the same data tiles through its L2.
  const int bid = blockIdx.x;
  const int xcd = bid & 7;
  const int per = 8 / a.S;
  const int split = xcd % a.S;
  const int rb = (bid >> 3) * per + xcd / a.S;
  const int nq = a.nq_dev ? *a.nq_dev : a.nq;
  if (rb >= a.n_row_blocks || static_cast<int64_t>(rb) * BM >= nq) return;

  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int wr = w >> 1, wc = w & 1;
  const int C = a.C;
  const int64_t rbytes = a.row_bytes;

  for (int r = tid; r < BM; r += NT) {
    const int64_t grow = static_cast<int64_t>(rb) * BM + r;
    int64_t self = -1;
    if (grow < nq) {
      if (a.q_self) self = a.q_self[grow];
      else if (a.self_base >= 0) self = a.self_base + grow;
    }
    s_row[r] = make_float4(a.sqq[grow], DT == DT_FP8 ? a.invq[grow] : 1.f, INFINITY,
                           __int_as_float(static_cast<int>(self)));
    s_tauk[r] = KEY_NONE;
    s_cnt[r] = 0;
    s_valid[r] = 0;
  }
  // The candidate lists live in the output buffer (this workgroup's rows of
  // its split, contiguous, L2-resident); only their C-th keys stay in LDS.
  uint64_t* lists = a.cand + (static_cast<int64_t>(split) * a.nq_pad + static_cast<int64_t>(rb) * BM) * C;
  uint64_t* pend = a.pend + (static_cast<int64_t>(split) * a.nq_pad + static_cast<int64_t>(rb) * BM) * PEND;
  for (int e = tid; e < BM * C; e += NT) lists[e] = KEY_NONE;
  __syncthreads();

  const int t0 = split * a.tiles_per_split;
  const int t1 = min(t0 + a.tiles_per_split, a.n_col_tiles);
  const char* Ab = a.Qq + static_cast<int64_t>(rb) * BM * rbytes;
  const int lr = tid >> 3;            // staging row (plus 32*i)
  const int lkk = (tid & 7) * 4;      // staging offset within the chunk (floats)
  const int nk = static_cast<int>(rbytes / 128);

  // Register prefetch one chunk deep into a double-buffered LDS stage: while
  // chunk kc is multiplied out of one buffer, chunk kc + 1 is in flight from
  // memory and then stored to the other buffer (the co-resident workgroup of
  // the CU hides the remaining latency). The first chunk of the next tile is
  // loaded between the two epilogue phases and stays in registers through
  // phase 2. Chunk indices past the end are clamped (re-read, never used).
  float4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
#define MV_GLOAD(BPTR, KCC)                                                                \
  do {                                                                                   \
    const int kq_ = (KCC) < nk ? (KCC) : nk - 1;                                         \
    const int64_t ko_ = static_cast<int64_t>(kq_) * 128 + lkk * 4;                       \
    const char* B_ = (BPTR);                                                             \
    ra0 = *reinterpret_cast<const float4*>(Ab + static_cast<int64_t>(lr) * rbytes + ko_);        \
    ra1 = *reinterpret_cast<const float4*>(Ab + static_cast<int64_t>(lr + 32) * rbytes + ko_);   \
    ra2 = *reinterpret_cast<const float4*>(Ab + static_cast<int64_t>(lr + 64) * rbytes + ko_);   \
    ra3 = *reinterpret_cast<const float4*>(Ab + static_cast<int64_t>(lr + 96) * rbytes + ko_);   \
    rb0 = *reinterpret_cast<const float4*>(B_ + static_cast<int64_t>(lr) * rbytes + ko_);        \
    rb1 = *reinterpret_cast<const float4*>(B_ + static_cast<int64_t>(lr + 32) * rbytes + ko_);   \
    rb2 = *reinterpret_cast<const float4*>(B_ + static_cast<int64_t>(lr + 64) * rbytes + ko_);   \
    rb3 = *reinterpret_cast<const float4*>(B_ + static_cast<int64_t>(lr + 96) * rbytes + ko_);   \
  } while (0)
#define MV_LSTORE(BUF)                                                                     \
  do {                                                                                   \
    float* As2_ = s_stage + (BUF) * (2 * BM * LDSA);                                     \
    float* Bs2_ = As2_ + BM * LDSA;                                                      \
    *reinterpret_cast<float4*>(As2_ + (lr) * LDSA + lkk) = ra0;                          \
    *reinterpret_cast<float4*>(As2_ + (lr + 32) * LDSA + lkk) = ra1;                     \
    *reinterpret_cast<float4*>(As2_ + (lr + 64) * LDSA + lkk) = ra2;                     \
    *reinterpret_cast<float4*>(As2_ + (lr + 96) * LDSA + lkk) = ra3;                     \
    *reinterpret_cast<float4*>(Bs2_ + (lr) * LDSA + lkk) = rb0;                          \
    *reinterpret_cast<float4*>(Bs2_ + (lr + 32) * LDSA + lkk) = rb1;                     \
    *reinterpret_cast<float4*>(Bs2_ + (lr + 64) * LDSA + lkk) = rb2;                     \
    *reinterpret_cast<float4*>(Bs2_ + (lr + 96) * LDSA + lkk) = rb3;                     \
  } while (0)
  // (a split past the last tile runs no tile; its prefetch address is clamped)
  MV_GLOAD(a.Xq + static_cast<int64_t>(t0 < a.n_col_tiles ? t0 : a.n_col_tiles - 1) * BN * rbytes, 0);

  for (int t = t0; t < t1; ++t) {
    const int col0 = t * BN;
    const char* Bb = a.Xq + static_cast<int64_t>(col0) * rbytes;
    const char* Bnext = t + 1 < t1 ? Bb + static_cast<int64_t>(BN) * rbytes : Bb;

    v4f acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = v4f{0.f, 0.f, 0.f, 0.f};

    MV_LSTORE(0);
    __syncthreads();
    for (int kc = 0; kc < nk; ++kc) {
      const int cur = kc & 1;
      const bool last = kc + 1 == nk;
      if (!last) MV_GLOAD(Bb, kc + 1);
      // keep the loads ahead of the multiplies (the scheduler would otherwise
      // sink them past the MFMAs to reuse operand registers, exposing their latency)
      __builtin_amdgcn_sched_barrier(0);
      mfma_chunk<DT>(s_stage + cur * (2 * BM * LDSA), s_stage + cur * (2 * BM * LDSA) + BM * LDSA, wr, wc, l, acc);
      // and the next chunk's LDS stores and the barrier behind them (the
      // scheduler would otherwise hoist both above the MFMAs, so that every
      // wavefront waits for its loads with the multiplies not yet issued)
      __builtin_amdgcn_sched_barrier(0);
      if (!last) MV_LSTORE(cur ^ 1);
      __syncthreads();
    }

    // Epilogue, phase 1: screen values D = |q|^2 + |x|^2 - 2 q.x (clamped at 0;
    // excluded columns +inf) stay in registers. Keys (distance, index) below
    // the row's C-th key are appended to the row's buffer straight from here:
    // per (mi, r) the 16 lanes of group g hold 64 columns of one row, lane
    // (16g) reserves their slots with an LDS atomic, and each lane stores its
    // keys at the reserved base plus its prefix count. The two wavefronts
    // sharing a row reserve from the same counter, so all successful
    // reservations of a tile precede the first one that does not fit. A row
    // half that does not fit writes its values to the D tile and is merged in
    // phase 2.
    float* Dt = s_stage;
    float sqxv[4], invc[4];
    int colv[4];
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      colv[ni] = col0 + wc * 64 + ni * 16 + (l & 15);
      sqxv[ni] = a.sqx[colv[ni]];
      invc[ni] = DT == DT_FP8 ? a.invx[colv[ni]] : 1.f;
    }
    const int g = l >> 4;
    uint64_t hits = 0;   // bit mi*16 + g*4 + r: row wr*64 + mi*16 + g*4 + r did not fit its buffer
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wr * 64 + mi * 16 + g * 4 + r;
        const float4 rv = s_row[row];
        const uint64_t tk = s_tauk[row];
        const int self = __float_as_int(rv.w);
        int c = 0;
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
          const float dot = DT == DT_FP8 ? acc[mi][ni][r] * rv.y * invc[ni] : acc[mi][ni][r];
          float v = fmaxf(rv.x + sqxv[ni] - 2.f * dot, 0.f);
          if (colv[ni] >= a.n || colv[ni] == self) v = INFINITY;
          // values failing the filter become +inf: no later, tighter threshold passes them
          const bool pass = v <= rv.z && v < INFINITY && dkey(v, static_cast<uint32_t>(colv[ni])) < tk;
          acc[mi][ni][r] = pass ? v : INFINITY;
          c += pass ? 1 : 0;
        }
        if (ballot(c > 0) == 0) continue;
        // inclusive prefix count over the group's 16 lanes (one DPP row); lane
        // 15 of the row holds the group total and reserves the slots
        int incl = c;
        incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xF, 0xF, false);   // row_shr:1
        incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xF, 0xF, false);   // row_shr:2
        incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xF, 0xF, false);   // row_shr:4
        incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xF, 0xF, false);   // row_shr:8
        int base = -2;   // lane 15: the reserved base; -1: the keys do not fit; -2: no keys
        if ((l & 15) == 15 && incl > 0) {
          base = atomicAdd(&s_cnt[row], incl);
          if (base + incl > PEND) base = -1;
          else atomicAdd(&s_valid[row], incl);
        }
        const int b0 = __builtin_amdgcn_readlane(base, 15), b1 = __builtin_amdgcn_readlane(base, 31);
        const int b2 = __builtin_amdgcn_readlane(base, 47), b3 = __builtin_amdgcn_readlane(base, 63);
        const int gb = g == 0 ? b0 : (g == 1 ? b1 : (g == 2 ? b2 : b3));
        if (gb >= 0) {
          uint64_t* buf = pend + row * PEND + gb + incl - c;
          int j = 0;
#pragma unroll
          for (int ni = 0; ni < 4; ++ni)
            if (acc[mi][ni][r] < INFINITY) buf[j++] = dkey(acc[mi][ni][r], static_cast<uint32_t>(colv[ni]));
        }
        // a group with keys that did not fit: the row half goes to the D tile
        const bool spill = gb == -1;
        const uint64_t sb = ballot(spill);
        if (sb) {
          if (spill) {
#pragma unroll
            for (int ni = 0; ni < 4; ++ni) Dt[row * DTS + wc * 64 + ni * 16 + (l & 15)] = acc[mi][ni][r];
          }
#pragma unroll
          for (int gg = 0; gg < 4; ++gg)
            if ((sb >> (16 * gg)) & 0xFFFFull) hits |= 1ull << (mi * 16 + gg * 4 + r);
        }
      }
    }
    if (l == 0) s_hit[wc][wr] = hits;
    // the first chunk of the next tile is in flight through phase 2 (not
    // earlier: phase 1 needs the registers)
    MV_GLOAD(Bnext, 0);
    __syncthreads();

    // Phase 2: wavefront w owns rows 32w .. 32w + 31 and merges those whose
    // buffer overflowed: the buffered keys first (which tightens the C-th
    // key), then the overflowing halves' values from the D tile.
    {
      const int hr = w >> 1, sh = (w & 1) * 32;
      const uint32_t h0 = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(s_hit[0][hr] >> sh));
      const uint32_t h1 = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(s_hit[1][hr] >> sh));
      uint32_t full = h0 | h1;
      while (full) {
        const int rr = __builtin_ctz(full);
        full &= full - 1;
        const int row = w * 32 + rr;
        uint64_t* buf = pend + row * PEND;
        const int valid = s_valid[row];
        uint64_t tau = s_tauk[row];
        uint64_t lk = l < C ? lists[row * C + l] : KEY_NONE;
        const uint64_t pk = l < valid ? buf[l] : KEY_NONE;
        const float v0 = (h0 >> rr) & 1u ? Dt[row * DTS + l] : INFINITY;
        const float v1 = (h1 >> rr) & 1u ? Dt[row * DTS + 64 + l] : INFINITY;
        if (valid > 0) tau = list_absorb(lk, pk, C, l);
        uint64_t k0 = v0 == INFINITY ? KEY_NONE : dkey(v0, static_cast<uint32_t>(col0 + l));
        uint64_t k1 = v1 == INFINITY ? KEY_NONE : dkey(v1, static_cast<uint32_t>(col0 + 64 + l));
        k0 = k0 < tau ? k0 : KEY_NONE;
        k1 = k1 < tau ? k1 : KEY_NONE;
        const uint64_t m0 = ballot(k0 != KEY_NONE);
        const uint64_t m1 = ballot(k1 != KEY_NONE);
        int cnt = __popcll(m0) + __popcll(m1);
        if (cnt <= PEND) {
          pend_append(buf, 0, k0, m0, k1, m1);
        } else {
          tau = list_absorb(lk, k0, C, l);
          k1 = k1 < tau ? k1 : KEY_NONE;
          if (ballot(k1 != KEY_NONE)) tau = list_absorb(lk, k1, C, l);
          cnt = 0;
        }
        if (l < C) lists[row * C + l] = lk;
        if (l == 0) {
          s_tauk[row] = tau;
          if (tau != KEY_NONE) s_row[row].z = key_dist(tau);
          s_cnt[row] = cnt;
          s_valid[row] = cnt;
        }
      }
    }
    __syncthreads();
  }

  // Buffered keys of the last tiles go into the lists.
  for (int rr = 0; rr < 32; ++rr) {
    const int row = w * 32 + rr;
    const int cnt = s_valid[row];
    if (cnt == 0) continue;
    uint64_t lk = l < C ? lists[row * C + l] : KEY_NONE;
    list_absorb(lk, l < cnt ? pend[row * PEND + l] : KEY_NONE, C, l);
    if (l < C) lists[row * C + l] = lk;
  }
#undef MV_GLOAD
#undef MV_LSTORE

}

// ------------------------------------------------------- symmetric screen --
//
// Self kNN (Q = X) of up to SYM_MAX_BYTES of distance matrix: D is symmetric,
// so only the tiles (I, J) with J >= I are multiplied (half the MFMA work of
// the screen above), each stored twice (D[I, J] and its transpose D[J, I]) to
// an n_pad x n_pad FP32 matrix in HBM; k_select then reads each row once and
// keeps its C smallest keys. The screen values are the screen's own
// expression (|x_i|^2 + |x_j|^2 - 2 x_i.x_j, clamped at 0, same K order), so
// the candidate lists, and everything downstream of them, are those of the
// fused screen.
constexpr size_t SYM_MAX_BYTES = size_t(48) << 30;
constexpr int64_t SYM_MIN_N = 4096;

struct SymArgs {
  const char* Xq;      // [n_pad, d_pad] centred data (FP32 or split BF16)
  const float* sqx;    // [n_pad]
  const float* invx;   // FP8: per-row 1/scale (else unused)
  int64_t row_bytes;
  int nb;              // row blocks (n_pad / 128)
  int nsb;             // super-blocks of SYM_G x SYM_G tiles per dimension (ceil(nb / SYM_G))
  int64_t n_slots;     // nsb (nsb + 1) / 2 * SYM_G^2 tile slots (some empty)
  int64_t per_xcd;     // slots per XCD (ceil(n_slots / 8))
  int64_t ldd;         // n_pad
  float* D;            // [n_pad, n_pad], FP32 (BF16 bit patterns for the BF16 and FP8 screens)
  float* bmin;         // [n_pad, nb]: per row, the smallest value of each 128-column block
  int n;
  int diag_self;       // the diagonal is each row's excluded self column
};

// Minimum over each 32-lane half of the wavefront, left in lanes 31 and 63:
// row_shr:1, 2, 4, 8 (lanes without a source keep their value) give each
// 16-lane row's minimum in its lane 15, row_bcast:15 folds row 0 into row 1
// and row 2 into row 3.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float dpp_keep_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, ROWS, 0xF, false));
}
__device__ __forceinline__ float min32_to_lane31(float v) {
  v = fminf(v, dpp_keep_f32<0x111>(v));
  v = fminf(v, dpp_keep_f32<0x112>(v));
  v = fminf(v, dpp_keep_f32<0x114>(v));
  v = fminf(v, dpp_keep_f32<0x118>(v));
  v = fminf(v, dpp_keep_f32<0x142, 0xA>(v));
  return v;
}

constexpr int SYM_G = 8;   // tiles per super-block side

// Tile t of the upper triangle in row-major order: row block I holds tiles
// start(I) .. start(I) + nb - I - 1 with start(I) = I nb - I (I - 1) / 2.
__device__ __forceinline__ void sym_tile(int64_t t, int nb, int& I, int& J) {
  const double b = 2.0 * nb + 1.0;
  int i = static_cast<int>((b - sqrt(b * b - 8.0 * static_cast<double>(t))) * 0.5);
  if (i < 0) i = 0;
  if (i > nb - 1) i = nb - 1;
  auto start = [nb](int64_t r) { return r * nb - r * (r - 1) / 2; };
  while (i > 0 && start(i) > t) --i;
  while (i + 1 < nb && start(i + 1) <= t) ++i;
  I = i;
  J = i + static_cast<int>(t - start(i));
}

// Workgroup of W wavefronts: W = 4 gives 64 x 64 wavefront tiles (2 x 2),
// W = 8 gives 64 x 32 tiles (2 x 4) with half the registers, so that four
// wavefronts share each SIMD.
#ifndef MV_SYM_WAVES
#define MV_SYM_WAVES 8
#endif
constexpr int SYM_W = MV_SYM_WAVES;

// The reduced-precision screens keep their matrix in BF16 (round to nearest
// even to 8 significant bits: s' <= s (1 + 2^-8), so the screen value
// s >= s' (1 - 2^-8) for s >= 0), which halves the matrix traffic; the
// re-rank's bounds scale the stored values by that factor (STORE16_LO).
constexpr double STORE16_LO = 1.0 - 1.0 / 256.0;
__device__ __forceinline__ float screen_round16(float v) {
  return __uint_as_float(static_cast<uint32_t>(__builtin_bit_cast(uint16_t, static_cast<__bf16>(v))) << 16);
}
__device__ __forceinline__ uint2 pack16x4(v4f x) {
  const uint32_t b0 = __float_as_uint(x[0]) >> 16, b1 = __float_as_uint(x[1]) >> 16;
  const uint32_t b2 = __float_as_uint(x[2]) >> 16, b3 = __float_as_uint(x[3]) >> 16;
  return make_uint2(b0 | (b1 << 16), b2 | (b3 << 16));
}
__device__ __forceinline__ float unpack16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float unpack16_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

// Staging of k_dist_sym: each 128-byte K chunk of a 128-row operand block
// lands in LDS straight from global memory (global_load_lds_dwordx4: lane L
// of an instruction writes bytes [16 L, 16 L + 16) past the wavefront-uniform
// base), 128 bytes per row with no padding. Row r keeps its 16-byte piece c in
// slot c ^ f(r), f(r) = ((r >> 1) ^ ((r & 1) << 2)) & 7, so that the MFMA
// operand reads (consecutive rows at one piece index) meet no bank conflict
// whether the LDS serves 8 lanes of a 16-byte read per cycle over 32 banks
// (f distinct over any 8 consecutive rows) or 16 lanes over 64 banks (f
// distinct over the even and over the odd rows of 16); the swizzle is applied
// on the load side, each lane fetching the piece its slot holds.
constexpr int SYM_OPB = BM * 128;                    // bytes per staged operand block
constexpr int SYM_STAGE_BYTES = 2 * 2 * SYM_OPB;     // two buffers x (A, B)
constexpr int SYM_LDS_FLOATS = (SYM_STAGE_BYTES / 4 > BM * (BN + 1)) ? SYM_STAGE_BYTES / 4 : BM * (BN + 1);

__device__ __forceinline__ int sym_swz_f(int r) { return ((r >> 1) ^ ((r & 1) << 2)) & 7; }
__device__ __forceinline__ int sym_swz(int r, int c) { return r * 128 + ((c ^ sym_swz_f(r)) << 4); }
__device__ __forceinline__ v4f sym_ld(const char* base, int r, int c) {
  return *reinterpret_cast<const v4f*>(base + sym_swz(r, c));
}

// One staged K chunk (128 bytes per row) of A rows wr*64 .. +64 and B rows
// wc*16*NI .. +16*NI into the 4 x NI accumulators of this wavefront.
template <int DT, int NI>
__device__ __forceinline__ void mfma_chunk_sym(const char* As, const char* Bs, int wr, int wc, int l,
                                               v4f (&acc)[4][NI]) {
  const int ra = wr * 64 + (l & 15), rb = wc * (16 * NI) + (l & 15), q = l >> 4;
  if constexpr (DT == DT_F32 || DT == DT_BF16) {
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      v4f af[4], bf[NI];
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) af[mi] = sym_ld(As, ra + mi * 16, g * 4 + q);
#pragma unroll
      for (int ni = 0; ni < NI; ++ni) bf[ni] = sym_ld(Bs, rb + ni * 16, g * 4 + q);
      if constexpr (DT == DT_F32) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int mi = 0; mi < 4; ++mi)
#pragma unroll
            for (int ni = 0; ni < NI; ++ni)
              acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[mi][j], bf[ni][j], acc[mi][ni], 0, 0, 0);
      } else {
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
          for (int ni = 0; ni < NI; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, af[mi]),
                                                                  __builtin_bit_cast(v8bf, bf[ni]), acc[mi][ni], 0, 0, 0);
      }
    }
  } else if constexpr (DT == DT_BF16X3) {
    // one K group of 32: hi in pieces 0..3 of the row, lo in pieces 4..7
    v8bf ah[4], al[4], bh[NI], bl[NI];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      ah[mi] = __builtin_bit_cast(v8bf, sym_ld(As, ra + mi * 16, q));
      al[mi] = __builtin_bit_cast(v8bf, sym_ld(As, ra + mi * 16, q + 4));
    }
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      bh[ni] = __builtin_bit_cast(v8bf, sym_ld(Bs, rb + ni * 16, q));
      bl[ni] = __builtin_bit_cast(v8bf, sym_ld(Bs, rb + ni * 16, q + 4));
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < NI; ++ni) {
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[mi], bh[ni], acc[mi][ni], 0, 0, 0);
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mi], bl[ni], acc[mi][ni], 0, 0, 0);
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mi], bh[ni], acc[mi][ni], 0, 0, 0);
      }
  } else {
    v8i af[4], bf[NI];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      const v4f lo = sym_ld(As, ra + mi * 16, 2 * q), hi = sym_ld(As, ra + mi * 16, 2 * q + 1);
      af[mi] = __builtin_bit_cast(v8i, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
    }
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      const v4f lo = sym_ld(Bs, rb + ni * 16, 2 * q), hi = sym_ld(Bs, rb + ni * 16, 2 * q + 1);
      bf[ni] = __builtin_bit_cast(v8i, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < NI; ++ni)
        acc[mi][ni] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(af[mi], bf[ni], acc[mi][ni], 0, 0, 0,
                                                                        FP8_UNIT_SCALE, 0, FP8_UNIT_SCALE);
  }
}

template <int DT, int W>
__global__ __launch_bounds__(64 * W, 2) void k_dist_sym(SymArgs a) {
  constexpr bool STORE16 = DT == DT_BF16 || DT == DT_FP8;
  constexpr int NTS = 64 * W;
  constexpr int NI = W == 4 ? 4 : 2;          // 16-column accumulator blocks per wavefront
  constexpr int WCOLS = W == 4 ? 2 : 4;       // wavefront columns
  constexpr int NQ = BM / (8 * W);            // load instructions per operand and chunk (8 rows each)
  __shared__ __attribute__((aligned(16))) float s_stage[SYM_LDS_FLOATS];
  // Workgroups are dispatched round-robin over the 8 XCDs (bid mod 8); XCD x
  // takes a contiguous run of tile slots. Slots are ordered by super-block
  // (SYM_G x SYM_G tiles, upper triangle of super-blocks in row-major order),
  // so the ~64 workgroups resident on an XCD work on one super-block: its
  // SYM_G row blocks and SYM_G column blocks are each read by SYM_G
  // workgroups at about the same K step, and all but one read hit the XCD's
  // L2 (a row-major tile order reads every column block from beyond the L2).
  const int bid = blockIdx.x;
  const int64_t u = static_cast<int64_t>(bid & 7) * a.per_xcd + (bid >> 3);
  if (u >= a.n_slots) return;
  int SI, SJ;
  sym_tile(u / (SYM_G * SYM_G), a.nsb, SI, SJ);
  const int I = SI * SYM_G + static_cast<int>(u % (SYM_G * SYM_G)) / SYM_G;
  const int J = SJ * SYM_G + static_cast<int>(u % SYM_G);
  if (I >= a.nb || J >= a.nb || J < I) return;

  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int wr = w / WCOLS, wc = w % WCOLS;
  const int64_t rbytes = a.row_bytes;
  const char* Ab = a.Xq + static_cast<int64_t>(I) * BM * rbytes;
  const char* Bb = a.Xq + static_cast<int64_t>(J) * BN * rbytes;
  const int nk = static_cast<int>(rbytes / 128);
  char* stage = reinterpret_cast<char*>(s_stage);

  // load instruction i of this wavefront: rows (i W + w) 8 + L / 8, slot L % 8
  int64_t goff[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    const int r = (i * W + w) * 8 + (l >> 3);
    goff[i] = static_cast<int64_t>(r) * rbytes + (((l & 7) ^ sym_swz_f(r)) << 4);
  }
  auto issue = [&](int kc, int buf) {
    char* dA = stage + buf * (2 * SYM_OPB);
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      const int64_t o = goff[i] + static_cast<int64_t>(kc) * 128;
      __builtin_amdgcn_global_load_lds(Ab + o, dA + (i * W + w) * 1024, 16, 0, 0);
      __builtin_amdgcn_global_load_lds(Bb + o, dA + SYM_OPB + (i * W + w) * 1024, 16, 0, 0);
    }
  };

  v4f acc[4][NI];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = v4f{0.f, 0.f, 0.f, 0.f};

  // Chunk kc + 1 streams into the other buffer while chunk kc is multiplied;
  // that buffer was last read in step kc - 1, which every wavefront has left
  // (barrier at its end). Each wavefront waits for its own loads, then the
  // barrier makes every wavefront's loads visible.
  issue(0, 0);
  __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int cur = kc & 1;
    if (kc + 1 < nk) issue(kc + 1, cur ^ 1);
    __builtin_amdgcn_sched_barrier(0);
    const char* As = stage + cur * (2 * SYM_OPB);
    // raised wave priority while issuing the multiplies: the arbiter then
    // prefers a wavefront with MFMAs ready over one that would only issue
    // loads or wait at the barrier (2.5 % faster at 70,000 x 784)
    __builtin_amdgcn_s_setprio(1);
    mfma_chunk_sym<DT, NI>(As, As + SYM_OPB, wr, wc, l, acc);
    __builtin_amdgcn_s_setprio(0);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
  }

  // Epilogue: lane (g = l / 16, c = l % 16) of (mi, ni) holds rows
  // wr*64 + mi*16 + 4g + r (r = 0..3) of column wc*16*NI + ni*16 + c. The
  // screen values go through LDS (the staging area is free after the last
  // K step; row stride 129 floats spreads the column reads over the banks),
  // so that both the block and its transpose leave as rows of 512 bytes (32
  // lanes x 16 bytes) instead of 64-byte pieces.
  constexpr int TS = BN + 1;
  static_assert(BM * TS <= SYM_LDS_FLOATS, "the screen tile must fit in the staging area");
  float* Dt = s_stage;
  {
    const int g = l >> 4, c = l & 15;
    const int64_t rb0 = static_cast<int64_t>(I) * BM, cb0 = static_cast<int64_t>(J) * BN + wc * (16 * NI);
    float sqc[NI], sqr[4][4], ivc[NI], ivr[4][4];
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      sqc[ni] = a.sqx[cb0 + ni * 16 + c];
      ivc[ni] = DT == DT_FP8 ? a.invx[cb0 + ni * 16 + c] : 1.f;
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        sqr[mi][r] = a.sqx[rb0 + wr * 64 + mi * 16 + g * 4 + r];
        ivr[mi][r] = DT == DT_FP8 ? a.invx[rb0 + wr * 64 + mi * 16 + g * 4 + r] : 1.f;
      }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
#pragma unroll
      for (int ni = 0; ni < NI; ++ni) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // the screen's expression (FP8: the two per-row scales applied to the product)
          const float dot = DT == DT_FP8 ? acc[mi][ni][r] * ivr[mi][r] * ivc[ni] : acc[mi][ni][r];
          Dt[(wr * 64 + mi * 16 + g * 4 + r) * TS + wc * (16 * NI) + ni * 16 + c] =
              fmaxf(sqr[mi][r] + sqc[ni] - 2.f * dot, 0.f);
        }
      }
    }
  }
  __syncthreads();
  const int64_t r0 = static_cast<int64_t>(I) * BM, c0 = static_cast<int64_t>(J) * BN;
  // Each row of the block (and, off the diagonal, of its transpose) is
  // stored by 32 lanes, which also reduce its minimum for the row selection
  // (bmin[row][column block]); padding columns and each row's self column are
  // left out of the minimum.
#pragma unroll
  for (int q = 0; q < BM * BN / 4 / NTS; ++q) {
    const int e = tid + q * NTS;
    const int row = e >> 5, c4 = (e & 31) * 4;
    v4f x;
    float mn = INFINITY;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      x[k] = STORE16 ? screen_round16(Dt[row * TS + c4 + k]) : Dt[row * TS + c4 + k];
      const bool ok = c0 + c4 + k < a.n && !(a.diag_self && I == J && c4 + k == row);
      mn = ok ? fminf(mn, x[k]) : mn;
    }
    if constexpr (STORE16)
      *reinterpret_cast<uint2*>(reinterpret_cast<uint16_t*>(a.D) + (r0 + row) * a.ldd + c0 + c4) = pack16x4(x);
    else
      *reinterpret_cast<v4f*>(a.D + (r0 + row) * a.ldd + c0 + c4) = x;
    mn = min32_to_lane31(mn);
    if ((e & 31) == 31) a.bmin[(r0 + row) * a.nb + J] = mn;
  }
  if (I != J) {
#pragma unroll
    for (int q = 0; q < BM * BN / 4 / NTS; ++q) {
      const int e = tid + q * NTS;
      const int col = e >> 5, r4 = (e & 31) * 4;
      v4f x;
      float mn = INFINITY;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        x[k] = STORE16 ? screen_round16(Dt[(r4 + k) * TS + col]) : Dt[(r4 + k) * TS + col];
        mn = r0 + r4 + k < a.n ? fminf(mn, x[k]) : mn;
      }
      if constexpr (STORE16)
        *reinterpret_cast<uint2*>(reinterpret_cast<uint16_t*>(a.D) + (c0 + col) * a.ldd + r0 + r4) = pack16x4(x);
      else
        *reinterpret_cast<v4f*>(a.D + (c0 + col) * a.ldd + r0 + r4) = x;
      mn = min32_to_lane31(mn);
      if ((e & 31) == 31) a.bmin[(c0 + col) * a.nb + I] = mn;
    }
  }
}

struct SelectArgs {
  const float* D;         // [n_pad, ldd] (S16: BF16 bit patterns)
  int64_t ldd;
  int n, nq, C;
  const int32_t* q_self;  // nullable
  int64_t self_base;      // used when q_self == NULL; < 0: no exclusion
  const float* bmin;      // [n_pad, nb] block minima (nullptr: read every block)
  int nb;
  uint64_t* cand;         // [nq_pad, C] keys (split 0 of the screen's layout)
};

// Offers the key of each lane (KEY_NONE: none) to the row's list: keys below
// the list's C-th key wait in the LDS buffer, which is merged into the list
// when it would overflow. KPL = 2 keeps a list of up to 128 keys (lk, lk1).
template <int KPL>
__device__ __forceinline__ uint64_t select_absorb(uint64_t& lk, uint64_t& lk1, uint64_t keys, int C, int l) {
  if constexpr (KPL == 1) return list_absorb(lk, keys, C, l);
  else return list_absorb2(lk, lk1, keys, C, l);
}

template <int KPL>
__device__ __forceinline__ void select_offer(uint64_t key, uint64_t& lk, uint64_t& lk1, uint64_t& tk, float& tf,
                                             int& cnt, uint64_t* pend, int C, int l) {
  key = key < tk ? key : KEY_NONE;
  uint64_t m = ballot(key != KEY_NONE);
  if (m == 0) return;
  if (cnt + __popcll(m) > 64) {
    tk = select_absorb<KPL>(lk, lk1, l < cnt ? pend[l] : KEY_NONE, C, l);
    if (tk != KEY_NONE) tf = key_dist(tk);
    cnt = 0;
    key = key < tk ? key : KEY_NONE;
    m = ballot(key != KEY_NONE);
  }
  const int pos = cnt + static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                            static_cast<uint32_t>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u)));
  if (key != KEY_NONE) pend[pos] = key;
  cnt += __popcll(m);
}

// The C smallest keys (distance, index) of each row of D, one wavefront per
// row. With block minima, the C-th smallest of the row's block minima, tau,
// bounds its C-th smallest value (C blocks hold a value <= tau), so only the
// blocks whose minimum is <= tau are read: every value of another block is
// > tau. Within the blocks read, a float threshold (the list's C-th distance)
// rejects values; keys that pass wait in an LDS buffer of 64 and are merged
// into the sorted list 64 at a time.
template <int KPL, bool S16>
__global__ __launch_bounds__(256) void k_select(SelectArgs a) {
  __shared__ uint64_t s_pend[4][64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t row = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (row >= a.nq) return;
  int64_t self = -1;
  if (a.q_self) self = a.q_self[row];
  else if (a.self_base >= 0) self = a.self_base + row;
  const float* Dr = a.D + row * a.ldd;
  uint64_t* pend = s_pend[w];
  const int C = a.C;
  uint64_t lk = KEY_NONE, lk1 = KEY_NONE;
  uint64_t tk = KEY_NONE;
  float tf = INFINITY;
  int cnt = 0;
  float tau = INFINITY;
  if (a.bmin && a.nb > C) {
    // tau: the C-th smallest block minimum (its key list reuses the row list's machinery)
    const float* bm = a.bmin + row * a.nb;
    for (int b0 = 0; b0 < a.nb; b0 += 64) {
      const int b = b0 + l;
      const float v = b < a.nb ? bm[b] : INFINITY;
      select_offer<KPL>(v < INFINITY ? dkey(v, static_cast<uint32_t>(b)) : KEY_NONE, lk, lk1, tk, tf, cnt, pend, C, l);
    }
    if (cnt > 0) tk = select_absorb<KPL>(lk, lk1, l < cnt ? pend[l] : KEY_NONE, C, l);
    if (tk != KEY_NONE) tau = key_dist(tk);
    lk = KEY_NONE;
    lk1 = KEY_NONE;
    tk = KEY_NONE;
    tf = INFINITY;
    cnt = 0;
    // the blocks to read, 64 at a time; each block: 2 columns per lane
    for (int b0 = 0; b0 < a.nb; b0 += 64) {
      const int b = b0 + l;
      uint64_t need = ballot(b < a.nb && bm[b] <= tau);
      while (need) {
        const int jb = (b0 + __builtin_ctzll(need)) * BN;
        need &= need - 1;
        const int64_t j = jb + 2 * l;
        float2 v;
        if constexpr (S16) {
          const uint32_t w2 = reinterpret_cast<const uint32_t*>(reinterpret_cast<const uint16_t*>(a.D) + row * a.ldd)[j >> 1];
          v = make_float2(unpack16_lo(w2), unpack16_hi(w2));
        } else {
          v = *reinterpret_cast<const float2*>(Dr + j);
        }
        const bool ok0 = j < a.n && j != self && v.x <= tf;
        const bool ok1 = j + 1 < a.n && j + 1 != self && v.y <= tf;
        select_offer<KPL>(ok0 ? dkey(v.x, static_cast<uint32_t>(j)) : KEY_NONE, lk, lk1, tk, tf, cnt, pend, C, l);
        select_offer<KPL>(ok1 ? dkey(v.y, static_cast<uint32_t>(j + 1)) : KEY_NONE, lk, lk1, tk, tf, cnt, pend, C, l);
      }
    }
  } else {
    // the loop bound is uniform: every lane takes part in the merges
    for (int64_t base = 0; base < a.n; base += 256) {
      const int64_t j0 = base + l * 4;
      v4f v = v4f{INFINITY, INFINITY, INFINITY, INFINITY};
      if (j0 < a.ldd) {
        if constexpr (S16) {
          const uint2 w2 = *reinterpret_cast<const uint2*>(reinterpret_cast<const uint16_t*>(a.D) + row * a.ldd + j0);
          v = v4f{unpack16_lo(w2.x), unpack16_hi(w2.x), unpack16_lo(w2.y), unpack16_hi(w2.y)};
        } else {
          v = *reinterpret_cast<const v4f*>(Dr + j0);
        }
      }
      bool any = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) any |= v[q] <= tf;
      if (ballot(any) == 0) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t j = j0 + q;
        const bool ok = j < a.n && j != self && v[q] <= tf;
        select_offer<KPL>(ok ? dkey(v[q], static_cast<uint32_t>(j)) : KEY_NONE, lk, lk1, tk, tf, cnt, pend, C, l);
      }
    }
  }
  if (cnt > 0) select_absorb<KPL>(lk, lk1, l < cnt ? pend[l] : KEY_NONE, C, l);
  if (l < C) a.cand[row * C + l] = lk;
  if (KPL == 2 && 64 + l < C) a.cand[row * C + 64 + l] = lk1;
}

// ----------------------------------------------------------------- rerank --

struct RerankArgs {
  const float* Q;
  const float* X;
  int64_t d;
  int nq, n, k, C, S, nq_pad;
  const uint64_t* cand;
  const float* sqq;            // centred query norms (screen space)
  const uint32_t* sqmax_bits;  // largest centred data norm
  float err_fac;               // screen arithmetic bound: err_fac * (sqq[i] + sqmax)
  double store_lo;             // screen values as listed are at most 1 / store_lo times the
                               // screen's (1: FP32 lists; STORE16_LO: the BF16-stored matrix)
  const float* eq;             // quantisation error bound per query (nullptr: FP32 screen)
  const uint32_t* emax_bits;   // largest quantisation error bound of the data
  const int32_t* nq_dev;       // nullable: device-side row count (re-screen)
  const int32_t* row_map;      // nullable: original query id of row i (re-screen)
  int fail_stat;               // stats slot counting rows without a certificate
  int vec4;                    // Q and X rows are 16-byte aligned (d % 4 == 0)
  int euclid;
  int32_t* out_idx;
  float* out_dist;
  int32_t* fb_count;
  int32_t* fb_rows;
  int32_t* stats;
};

// Sequential, unfused FP64 squared distance: the canonical pair distance.
__device__ __forceinline__ double dist64_seq(const float* __restrict__ q, const float* __restrict__ x,
                                             int64_t d) {
  double acc = 0.0;
  int64_t t = 0;
  // sixteen loads in flight per step, then the terms in order
  for (; t + 16 <= d; t += 16) {
    float qv[16], xv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      qv[u] = q[t + u];
      xv[u] = x[t + u];
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const double df = static_cast<double>(qv[u]) - static_cast<double>(xv[u]);
      acc = __dadd_rn(acc, __dmul_rn(df, df));
    }
  }
  for (; t < d; ++t) {
    const double df = static_cast<double>(q[t]) - static_cast<double>(x[t]);
    acc = __dadd_rn(acc, __dmul_rn(df, df));
  }
  return acc;
}

// Orders (distance, index) pairs with FP64 distances; one pair per lane.
struct Pair64 {
  uint64_t dbits;  // bit pattern of a non-negative double (monotone)
  uint32_t idx;
};
__device__ __forceinline__ bool pair_less(uint64_t ad, uint32_t ai, uint64_t bd, uint32_t bi) {
  return ad < bd || (ad == bd && ai < bi);
}
__device__ __forceinline__ void wave_sort_pair64(uint64_t& dbits, uint32_t& idx) {
  const int l = lane_id();
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint64_t od = shfl_xor_u64(dbits, j);
      const uint32_t oi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(idx), j));
      const bool up = (l & k) == 0;
      const bool lower = (l & j) == 0;
      const bool other_less = pair_less(od, oi, dbits, idx);
      const bool take_other = (lower == up) ? other_less : !other_less;
      if (take_other) {
        dbits = od;
        idx = oi;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_rerank(RerankArgs a) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + w;
  const int nq = a.nq_dev ? *a.nq_dev : a.nq;
  if (i >= nq) return;
  const int orow = a.row_map ? a.row_map[i] : i;
  const int C = a.C, k = a.k;
  const int M = a.S * C;

  // 1. Global top-C (by screen key) of the S per-split lists.
  uint64_t e[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int f = l + 64 * q;
    e[q] = f < M ? a.cand[(static_cast<int64_t>(f / C) * a.nq_pad + i) * C + (f % C)] : KEY_NONE;
  }
  uint64_t mine = KEY_NONE;
  for (int j = 0; j < C; ++j) {
    uint64_t lm = e[0];
#pragma unroll
    for (int q = 1; q < 4; ++q) lm = e[q] < lm ? e[q] : lm;
    const uint64_t m = wave_min_u64(lm);
    if (m == KEY_NONE) break;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (e[q] == m) e[q] = KEY_NONE;
    if (l == j) mine = m;
  }
  const int nvalid = __popcll(ballot(mine != KEY_NONE));
  const uint64_t tau_key = readlane_u64(mine, C - 1);
  const int cidx = mine != KEY_NONE ? static_cast<int>(key_idx(mine)) : 0;

  // 2. FP32 direct-difference distances, 4 candidates at a time (16 lanes
  //    each), of the candidates that can still reach the top k: first the k
  //    with the smallest screen values, whose largest distance plus its bound
  //    (thr) bounds the k-th true distance from above; then only the others
  //    whose screen value, less the screen's error bound, is not above thr.
  //    The rest are farther than all of the first k and are left out of the
  //    sort (ncomp candidates are computed); they still count for the
  //    certificate through tau.
  const float* qrow = a.Q + static_cast<int64_t>(orow) * a.d;
  const int g = l >> 4, sub = l & 15;
  // Each lane sums ceil(d/16) fmaf terms in order, then 4 butterfly adds;
  // every term carries the rounding of q - x (relative 2u after squaring).
  const int64_t terms = a.vec4 ? 4 * ((a.d + 63) / 64) : (a.d + 15) / 16;
  const float bfac = static_cast<float>(terms + 8) * 5.9604645e-8f * 1.01f;
  float d32 = INFINITY;
  auto compute = [&](uint64_t m) {
    while (m) {
      int p[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        p[q] = m ? __builtin_ctzll(m) : -1;
        m = m ? m & (m - 1) : 0;
      }
      const int c = g == 0 ? p[0] : (g == 1 ? p[1] : (g == 2 ? p[2] : p[3]));
      const int j = __shfl(cidx, c >= 0 ? c : 0);
      float acc = 0.f;
      if (c >= 0) {
        const float* xr = a.X + static_cast<int64_t>(j) * a.d;
        if (a.vec4) {
          // 16 lanes x 16-byte loads cover 64 contiguous floats per step.
          const float4* q4 = reinterpret_cast<const float4*>(qrow);
          const float4* x4 = reinterpret_cast<const float4*>(xr);
          for (int64_t t = sub; t < a.d / 4; t += 16) {
            const float4 qa = q4[t], xb = x4[t];
            float df = qa.x - xb.x;
            acc = fmaf(df, df, acc);
            df = qa.y - xb.y;
            acc = fmaf(df, df, acc);
            df = qa.z - xb.z;
            acc = fmaf(df, df, acc);
            df = qa.w - xb.w;
            acc = fmaf(df, df, acc);
          }
        } else {
          for (int64_t t = sub; t < a.d; t += 16) {
            const float df = qrow[t] - xr[t];
            acc = fmaf(df, df, acc);
          }
        }
      }
      acc += __shfl_xor(acc, 8);
      acc += __shfl_xor(acc, 4);
      acc += __shfl_xor(acc, 2);
      acc += __shfl_xor(acc, 1);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float got = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(acc), 16 * q));
        if (l == p[q]) d32 = got;
      }
    }
  };
  const int kfirst = nvalid < k ? nvalid : k;
  const uint64_t first = kfirst >= 64 ? ~0ull : ((1ull << kfirst) - 1);
  compute(first);
  {
    float ub = l < kfirst ? d32 + (d32 * bfac + 1e-35f) : 0.f;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) ub = fmaxf(ub, __shfl_xor(ub, m));
    const double thr = static_cast<double>(ub);
    // lower bound of the true squared distance of candidate l from its screen value
    const double sv = static_cast<double>(key_dist(mine)) * a.store_lo;
    const float sqmax0 = __uint_as_float(*a.sqmax_bits);
    const double ar = static_cast<double>(a.err_fac * (a.sqq[i] + sqmax0) + 1e-30f);
    double lb = sv - ar;
    if (a.eq) {
      const double r = sqrt(fmax(sv - ar, 0.0)) - static_cast<double>(a.eq[i]) -
                       static_cast<double>(__uint_as_float(*a.emax_bits));
      lb = r > 0.0 ? r * r * (1.0 - 1e-6) : 0.0;
    }
    const bool need = l >= kfirst && l < nvalid && lb <= thr;
    compute(ballot(need));
  }
  const int ncomp = __popcll(ballot(d32 < INFINITY));

  // 3. Sort by (FP32 distance, index); flag pairs FP32 cannot separate.
  uint64_t k32 = d32 < INFINITY ? dkey(d32, static_cast<uint32_t>(cidx)) : KEY_NONE;
  k32 = wave_sort_u64(k32);
  const float ds = key_dist(k32);
  const float bd = ds * bfac + 1e-35f;
  const float dn = __shfl_down(ds, 1);
  const float bn = __shfl_down(bd, 1);
  const bool amb = (l < k) && (l + 1 < ncomp) && (dn - bn <= ds + bd);
  const bool any_amb = ballot(amb) != 0;

  uint32_t out_i = key_idx(k32);
  float out_d = ds;
  float kth_ub;
  if (any_amb) {
    // 4. FP64 tie check: canonical distances for every candidate, then sort.
    double d64 = INFINITY;
    if (l < ncomp) d64 = dist64_seq(qrow, a.X + static_cast<int64_t>(out_i) * a.d, a.d);
    uint64_t dbits = l < ncomp ? static_cast<uint64_t>(__double_as_longlong(d64)) : KEY_NONE;
    uint32_t di = l < ncomp ? out_i : 0xFFFFFFFFu;
    wave_sort_pair64(dbits, di);
    const double dv = __longlong_as_double(static_cast<long long>(dbits));
    out_i = di;
    out_d = a.euclid ? static_cast<float>(sqrt(dv)) : static_cast<float>(dv);
    const double kth = __longlong_as_double(static_cast<long long>(
        shfl_u64(dbits, k - 1)));
    kth_ub = static_cast<float>(kth * (1.0 + 1e-6));
    if (l == 0) atomicAdd(&a.stats[1], 1);
  } else {
    kth_ub = __shfl(ds + bd, k - 1);
    out_d = a.euclid ? sqrtf(ds) : ds;
  }

  // 5. Certificate: every non-candidate j has screen value >= tau. With an FP32
  //    screen its true squared distance is >= tau - arith_i; with a quantised
  //    screen |x_i - x_j| >= sqrt(tau - arith_i) - E_i - E_max (E: norms of the
  //    quantisation errors). The k-th candidate must lie strictly below.
  bool cert = nvalid < C;
  if (!cert) {
    const double tau = static_cast<double>(key_dist(tau_key)) * a.store_lo;
    const float sqmax = __uint_as_float(*a.sqmax_bits);
    const float arith = a.err_fac * (a.sqq[i] + sqmax) + 1e-30f;
    if (!a.eq) {
      cert = static_cast<double>(kth_ub) < tau - static_cast<double>(arith);
    } else {
      const double t = fmax(tau - static_cast<double>(arith), 0.0);
      const double r = sqrt(t) - static_cast<double>(a.eq[i]) - static_cast<double>(__uint_as_float(*a.emax_bits));
      cert = r > 0.0 && static_cast<double>(kth_ub) < r * r * (1.0 - 1e-6);
    }
  }
  if (cert) {
    if (l < k) {
      a.out_idx[static_cast<int64_t>(orow) * k + l] = static_cast<int32_t>(out_i);
      a.out_dist[static_cast<int64_t>(orow) * k + l] = out_d;
    }
  } else if (l == 0) {
    const int p = atomicAdd(a.fb_count, 1);
    a.fb_rows[p] = orow;
    atomicAdd(&a.stats[a.fail_stat], 1);
  }
}

// Re-rank of up to 128 candidates: the symmetric screen's single list per row
// (sorted by screen key), for the wider candidate sets a BF16 screen needs
// before its certificate holds. Same steps as k_rerank with two candidates
// per lane: FP32 direct differences of the first 64 candidates; of the next
// 64 only those whose screen value, less the screen's error bound, is not
// above the k-th FP32 distance so far (plus its bound): the others cannot
// enter the top k, they only raise the certificate's threshold to the C-th
// screen key. Then a sort by (FP32 distance, index), FP64 only where FP32
// cannot separate two of the first k + 1, and the certificate. When FP64 is
// needed it orders the first 64; a row whose 65th computed candidate is not
// separated from them goes to the exhaustive scan.
__device__ __forceinline__ float rerank_dist32(const RerankArgs& a, const float* qrow, int j, int sub) {
  float acc = 0.f;
  const float* xr = a.X + static_cast<int64_t>(j) * a.d;
  if (a.vec4) {
    const float4* q4 = reinterpret_cast<const float4*>(qrow);
    const float4* x4 = reinterpret_cast<const float4*>(xr);
    for (int64_t t = sub; t < a.d / 4; t += 16) {
      const float4 qa = q4[t], xb = x4[t];
      float df = qa.x - xb.x;
      acc = fmaf(df, df, acc);
      df = qa.y - xb.y;
      acc = fmaf(df, df, acc);
      df = qa.z - xb.z;
      acc = fmaf(df, df, acc);
      df = qa.w - xb.w;
      acc = fmaf(df, df, acc);
    }
  } else {
    for (int64_t t = sub; t < a.d; t += 16) {
      const float df = qrow[t] - xr[t];
      acc = fmaf(df, df, acc);
    }
  }
  return acc;
}

__global__ __launch_bounds__(256) void k_rerank2(RerankArgs a) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + w;
  if (i >= a.nq) return;
  const int orow = i;
  const int C = a.C, k = a.k;
  const uint64_t* cl = a.cand + static_cast<int64_t>(i) * C;
  const uint64_t c0 = l < C ? cl[l] : KEY_NONE;
  const uint64_t c1 = 64 + l < C ? cl[64 + l] : KEY_NONE;
  const int n0 = __popcll(ballot(c0 != KEY_NONE));
  const int nall = n0 + __popcll(ballot(c1 != KEY_NONE));
  const uint64_t tau_key = C > 64 ? readlane_u64(c1, C - 65) : readlane_u64(c0, C - 1);
  const int ci0 = c0 != KEY_NONE ? static_cast<int>(key_idx(c0)) : 0;
  const int ci1 = c1 != KEY_NONE ? static_cast<int>(key_idx(c1)) : 0;
  const float* qrow = a.Q + static_cast<int64_t>(orow) * a.d;
  const int g = l >> 4, sub = l & 15;
  const int64_t terms = a.vec4 ? 4 * ((a.d + 63) / 64) : (a.d + 15) / 16;
  const float bfac = static_cast<float>(terms + 8) * 5.9604645e-8f * 1.01f;
  const float sqmax = __uint_as_float(*a.sqmax_bits);
  const float arith = a.err_fac * (a.sqq[i] + sqmax) + 1e-30f;

  // FP32 direct differences, 4 candidates at a time, 16 lanes each: lane l
  // ends with candidate l (d0).
  float d0 = INFINITY, d1 = INFINITY;
  for (int r = 0; r * 4 < n0; ++r) {
    const int c = r * 4 + g;
    const int j = __shfl(ci0, c & 63);
    float acc = c < n0 ? rerank_dist32(a, qrow, j, sub) : 0.f;
    acc += __shfl_xor(acc, 8);
    acc += __shfl_xor(acc, 4);
    acc += __shfl_xor(acc, 2);
    acc += __shfl_xor(acc, 1);
    const float got = __shfl(acc, (l & 3) * 16);
    if ((l >> 2) == r) d0 = got;
  }
  uint64_t x0 = c0 != KEY_NONE ? dkey(d0, static_cast<uint32_t>(ci0)) : KEY_NONE;
  x0 = wave_sort_u64(x0);
  // upper bound of the true k-th distance: the k-th FP32 value plus its bound
  const float kd = key_dist(readlane_u64(x0, k - 1));
  const float kth32_ub = kd + kd * bfac + 1e-35f;
  bool need1 = false;
  if (c1 != KEY_NONE) {
    const double sv = static_cast<double>(key_dist(c1)) * a.store_lo;
    double lb;
    if (!a.eq) {
      lb = sv - static_cast<double>(arith);
    } else {
      const double t = fmax(sv - static_cast<double>(arith), 0.0);
      const double r = sqrt(t) - static_cast<double>(a.eq[i]) - static_cast<double>(__uint_as_float(*a.emax_bits));
      lb = r > 0.0 ? r * r * (1.0 - 1e-6) : 0.0;
    }
    need1 = lb <= static_cast<double>(kth32_ub);
  }
  // candidates are sorted by screen key, so the ones needed form a prefix
  const int n1 = __popcll(ballot(need1));
  for (int r = 0; r * 4 < n1; ++r) {
    const int c = r * 4 + g;
    const int j = __shfl(ci1, c & 63);
    float acc = c < n1 ? rerank_dist32(a, qrow, j, sub) : 0.f;
    acc += __shfl_xor(acc, 8);
    acc += __shfl_xor(acc, 4);
    acc += __shfl_xor(acc, 2);
    acc += __shfl_xor(acc, 1);
    const float got = __shfl(acc, (l & 3) * 16);
    if ((l >> 2) == r) d1 = got;
  }
  const int nvalid = n0 + n1;

  // Sort the computed keys: x0 is sorted, x1 sorted descending, then one
  // compare-exchange between them and a half-cleaner sort of each.
  uint64_t x1 = l < n1 ? dkey(d1, static_cast<uint32_t>(ci1)) : KEY_NONE;
  if (n1 > 0) {
    x1 = wave_sort_u64<true>(x1);
    const uint64_t lo = x0 < x1 ? x0 : x1, hi = x0 < x1 ? x1 : x0;
    x0 = bitonic_clean_u64(lo);
    x1 = bitonic_clean_u64(hi);
  }
  const float ds = key_dist(x0);
  const float bd = ds * bfac + 1e-35f;
  const float ds1 = key_dist(x1);
  const float bd1 = ds1 * bfac + 1e-35f;
  const float first1 = __shfl(ds1, 0), firstb1 = __shfl(bd1, 0);
  float dn = __shfl_down(ds, 1), bn = __shfl_down(bd, 1);
  if (l == 63) {
    dn = first1;
    bn = firstb1;
  }
  const bool amb = (l < k) && (l + 1 < nvalid) && (dn - bn <= ds + bd);
  const bool any_amb = ballot(amb) != 0;

  uint32_t out_i = key_idx(x0);
  float out_d = ds;
  float kth_ub;
  bool beyond_ok = true;   // no computed candidate past the 64th can enter the top k
  if (any_amb) {
    double d64 = INFINITY;
    const bool v = l < nvalid;
    if (v) d64 = dist64_seq(qrow, a.X + static_cast<int64_t>(out_i) * a.d, a.d);
    uint64_t dbits = v ? static_cast<uint64_t>(__double_as_longlong(d64)) : KEY_NONE;
    uint32_t di = v ? out_i : 0xFFFFFFFFu;
    wave_sort_pair64(dbits, di);
    const double dv = __longlong_as_double(static_cast<long long>(dbits));
    out_i = di;
    out_d = a.euclid ? static_cast<float>(sqrt(dv)) : static_cast<float>(dv);
    const double kth = __longlong_as_double(static_cast<long long>(shfl_u64(dbits, k - 1)));
    kth_ub = static_cast<float>(kth * (1.0 + 1e-6));
    if (nvalid > 64) beyond_ok = first1 - firstb1 > kth_ub;
    if (l == 0) atomicAdd(&a.stats[1], 1);
  } else {
    kth_ub = __shfl(ds + bd, k - 1);
    out_d = a.euclid ? sqrtf(ds) : ds;
  }

  bool cert = beyond_ok && nall < C;
  if (beyond_ok && !cert) {
    const double tau = static_cast<double>(key_dist(tau_key)) * a.store_lo;
    if (!a.eq) {
      cert = static_cast<double>(kth_ub) < tau - static_cast<double>(arith);
    } else {
      const double t = fmax(tau - static_cast<double>(arith), 0.0);
      const double r = sqrt(t) - static_cast<double>(a.eq[i]) - static_cast<double>(__uint_as_float(*a.emax_bits));
      cert = r > 0.0 && static_cast<double>(kth_ub) < r * r * (1.0 - 1e-6);
    }
  }
  if (cert) {
    if (l < k) {
      a.out_idx[static_cast<int64_t>(orow) * k + l] = static_cast<int32_t>(out_i);
      a.out_dist[static_cast<int64_t>(orow) * k + l] = out_d;
    }
  } else if (l == 0) {
    const int p = atomicAdd(a.fb_count, 1);
    a.fb_rows[p] = orow;
    atomicAdd(&a.stats[a.fail_stat], 1);
  }
}

// ------------------------------------------------------- k above 64 (self) --

constexpr int KMAX_LARGE = 1024;   // largest k: selected straight from the symmetric screen's matrix
// From this k on, a self kNN on the symmetric screen selects from the matrix
// (at 70,000 x 784: 38.7 ms at k = 56 through the candidate lists, 84 ms at
// k = 64, where most rows miss the certificate; 44.5 ms at k = 65 here).
constexpr int K_LARGE_FROM = 57;

// Large-k self kNN, one workgroup per row, from the symmetric screen's n x n
// matrix (FP32 or split-BF16 screen values s_j, |s_j - t_j| <= a with t_j the
// true squared distance and a = err_fac (|q|^2 + max |x|^2), the screen's
// arithmetic bound):
//   1. B = the k-th smallest block minimum (k distinct columns are <= B, so the
//      k-th smallest value tau <= B and only blocks with minimum <= B hold
//      values that can rank below it);
//   2. tau exactly, by a radix select (11, 11, 10 bits of the order-preserving
//      key) over the values <= B of those blocks;
//   3. every column with s_j <= T = tau + 2a is a candidate: the k columns of
//      smallest screen value have t <= tau + a, so the k-th smallest true
//      distance is <= tau + a, and any column at or below it has s <= tau + 2a;
//   4. the candidates' canonical FP64 distances, merged batch by batch into a
//      sorted list of (distance, index) in LDS; its first k are the result.
// The result is the canonical order itself, whatever the screen's precision
// within its bound; the cost is about k candidates' exact distances per row.
struct LargeArgs {
  const float* D;   // [n_pad, ldd]
  int64_t ldd;
  const float* bmin;   // [n_pad, nb] block minima over the selectable columns, or nullptr
  int nb;
  int n, nq, k;
  const int32_t* q_self;
  int64_t self_base;
  const float* sqq;             // centred squared norms of the rows
  const uint32_t* sqmax_bits;   // largest centred squared norm
  float err_fac;
  const float* Q;   // original rows: canonical distances
  const float* X;
  int64_t d;
  int euclid;
  int32_t* out_idx;
  float* out_dist;
};

__device__ __forceinline__ uint32_t fkey(float v) {
  const uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fkey_value(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Block-wide bitonic sort (ascending) of P = power of two <= 2048 entries in
// LDS, 256 threads; (distance, index) pairs or 32-bit keys.
__device__ void lds_sort_pairs(double* kd, int32_t* ki, int P) {
  for (int size = 2; size <= P; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (int t = threadIdx.x; t < P / 2; t += 256) {
        const int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
        const double a = kd[lo], b = kd[hi];
        const int32_t ia = ki[lo], ib = ki[hi];
        const bool gt = a > b || (a == b && ia > ib);
        if (gt == ((lo & size) == 0)) {
          kd[lo] = b;
          kd[hi] = a;
          ki[lo] = ib;
          ki[hi] = ia;
        }
      }
    }
  __syncthreads();
}
__device__ void lds_sort_u32(uint32_t* key, int P) {
  for (int size = 2; size <= P; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (int t = threadIdx.x; t < P / 2; t += 256) {
        const int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
        const uint32_t a = key[lo], b = key[hi];
        if ((a > b) == ((lo & size) == 0)) {
          key[lo] = b;
          key[hi] = a;
        }
      }
    }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_knn_large(LargeArgs g) {
  constexpr int LIST = 2048;   // running list (<= k) + one batch (< 512)
  __shared__ double s_md[LIST];
  __shared__ int32_t s_mi[LIST];
  __shared__ uint32_t s_u[2048];   // block-minimum keys, then the radix histogram
  __shared__ int32_t s_blk[1024];
  __shared__ int32_t s_bc[512];
  __shared__ int32_t s_cnt, s_nblk;
  __shared__ uint32_t s_digit, s_below;
  const int tid = threadIdx.x;
  const int64_t row = blockIdx.x;
  if (row >= g.nq) return;
  const int k = g.k;
  int64_t self = -1;
  if (g.q_self) self = g.q_self[row];
  else if (g.self_base >= 0) self = g.self_base + row;
  const float* Dr = g.D + row * g.ldd;
  const bool lists = g.bmin != nullptr && g.nb <= 1024;

  // 1. B: the k-th smallest block minimum (+inf: read every block)
  uint32_t bkey = 0xFFFFFFFFu;
  if (lists && g.nb > k) {
    for (int b = tid; b < 1024; b += 256) s_u[b] = b < g.nb ? fkey(g.bmin[row * g.nb + b]) : 0xFFFFFFFFu;
    lds_sort_u32(s_u, 1024);
    bkey = s_u[k - 1];
    __syncthreads();
  }
  // the blocks whose minimum is at most the bound key
  auto build_list = [&](uint32_t bound) {
    if (tid == 0) s_nblk = 0;
    __syncthreads();
    for (int b = tid; b < g.nb; b += 256)
      if (fkey(g.bmin[row * g.nb + b]) <= bound) s_blk[atomicAdd(&s_nblk, 1)] = b;
    __syncthreads();
  };
  int nblk = g.nb;
  if (lists) {
    build_list(bkey);
    nblk = s_nblk;
  }
  // visits the selectable values of the listed blocks, two blocks (256 columns)
  // per step; f(j, key, step_valid) for every thread (valid: a column to offer)
  auto visit = [&](int count, auto&& f) {
    for (int i = 0; i < count; i += 2) {
      const int li = i + (tid >> 7);
      bool ok = li < count;
      int64_t j = 0;
      float v = INFINITY;
      if (ok) {
        const int b = lists ? s_blk[li] : li;
        j = static_cast<int64_t>(b) * 128 + (tid & 127);
        ok = j < g.n && j != self;
        if (ok) v = Dr[j];
      }
      f(j, fkey(v), ok);
    }
  };

  // 2. tau: the k-th smallest key among those <= B, three radix digits
  uint32_t prefix = 0;
  uint32_t krem = static_cast<uint32_t>(k);
  const int shifts[3] = {21, 10, 0};
  const int widths[3] = {11, 11, 10};
  for (int p = 0; p < 3; ++p) {
    const int sh = shifts[p], wd = widths[p];
    const uint32_t hi_mask = sh + wd >= 32 ? 0u : ~((1u << (sh + wd)) - 1u);
    const uint32_t dmask = (1u << wd) - 1u;
    for (int b = tid; b < 2048; b += 256) s_u[b] = 0;
    __syncthreads();
    visit(nblk, [&](int64_t, uint32_t key, bool ok) {
      if (ok && key <= bkey && (key & hi_mask) == (prefix & hi_mask)) atomicAdd(&s_u[(key >> sh) & dmask], 1u);
    });
    __syncthreads();
    // the digit holding rank krem: sums of 8 bins per thread, 4 of those per
    // lane of wavefront 0, a wavefront scan, then the lane that crosses krem
    // walks its 32 bins
    {
      uint32_t part = 0;
#pragma unroll
      for (int e = 0; e < 8; ++e) part += s_u[tid * 8 + e];
      s_bc[tid] = static_cast<int32_t>(part);
    }
    __syncthreads();
    if (tid < 64) {
      uint32_t p4[4];
      uint32_t ls = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        p4[e] = static_cast<uint32_t>(s_bc[tid * 4 + e]);
        ls += p4[e];
      }
      uint32_t incl = ls;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if (tid >= off) incl += o;
      }
      const uint64_t hit = __ballot(incl >= krem);
      const int lane = hit ? __builtin_ctzll(hit) : 63;
      if (tid == lane) {
        uint32_t below = incl - ls;
        uint32_t dg = static_cast<uint32_t>(tid) * 32;
        const uint32_t end = dg + 32 <= dmask + 1 ? dg + 32 : dmask + 1;
        for (; dg + 1 < end; ++dg) {
          if (below + s_u[dg] >= krem) break;
          below += s_u[dg];
        }
        s_digit = dg;
        s_below = below;
      }
    }
    __syncthreads();
    prefix |= s_digit << sh;
    krem -= s_below;
    __syncthreads();
  }
  const float tau = fkey_value(prefix);

  // 3. candidates: s <= T = tau + 2a (rounded up)
  const double a = static_cast<double>(g.err_fac) *
                   (static_cast<double>(g.sqq[row]) + static_cast<double>(__uint_as_float(*g.sqmax_bits))) + 1e-30;
  const double Td = static_cast<double>(tau) + 2.0 * a;
  float T = static_cast<float>(Td);
  if (static_cast<double>(T) < Td) T = nextafterf(T, INFINITY);
  const uint32_t tkey = fkey(T);
  if (lists) {
    build_list(tkey);
    nblk = s_nblk;
  }

  // 4. canonical distances of the candidates, merged into the sorted list
  const float* q = g.Q + row * g.d;
  int L = 0;
  if (tid == 0) s_cnt = 0;
  __syncthreads();
  auto flush = [&]() {
    const int m = s_cnt;
    for (int e = tid; e < m; e += 256) {
      const int32_t j = s_bc[e];
      s_md[L + e] = dist64_seq(q, g.X + static_cast<int64_t>(j) * g.d, g.d);
      s_mi[L + e] = j;
    }
    int P = 1;
    while (P < L + m) P <<= 1;
    for (int e = L + m + tid; e < P; e += 256) {
      s_md[e] = INFINITY;
      s_mi[e] = 0x7FFFFFFF;
    }
    lds_sort_pairs(s_md, s_mi, P);
    L = L + m < k ? L + m : k;
    if (tid == 0) s_cnt = 0;
    __syncthreads();
  };
  for (int i = 0; i < nblk; i += 2) {
    const int li = i + (tid >> 7);
    if (li < nblk) {
      const int b = lists ? s_blk[li] : li;
      const int64_t j = static_cast<int64_t>(b) * 128 + (tid & 127);
      if (j < g.n && j != self && fkey(Dr[j]) <= tkey) s_bc[atomicAdd(&s_cnt, 1)] = static_cast<int32_t>(j);
    }
    __syncthreads();
    const int c = s_cnt;
    __syncthreads();   // every thread has read the count before the next step adds to it
    if (c >= 256) flush();   // a step adds at most 256 (buffer of 512)
  }
  __syncthreads();
  if (s_cnt > 0) flush();
  for (int t = tid; t < k; t += 256) {
    const double dv = s_md[t];
    g.out_idx[row * k + t] = s_mi[t];
    g.out_dist[row * k + t] = static_cast<float>(g.euclid ? sqrt(dv) : dv);
  }
}

// ------------------------------------------------------------- exact scan --

constexpr int SCAN_CHUNKS = 64;    // column chunks per failing row in the split scan
constexpr int CMAX_K = 64;         // largest k
constexpr int SCAN_ROWS = 256;     // failing rows handled by the split scan (the rest: one block per row)

// Split exhaustive scan, part 1: block (chunk c, slot y) handles failing rows
// f = y, y + SCAN_ROWS', ... < min(count, SCAN_ROWS) over the columns of chunk c.
// Every lane computes canonical FP64 distances (sequential, unfused) for its
// columns; each wavefront keeps a sorted top-k list; the block's four lists are
// merged and its k best (distance bits, index) pairs stored for row f.
struct ScanPartArgs {
  const float* Q;
  const float* X;
  int64_t d;
  int n, k;
  const int32_t* q_self;
  int64_t self_base;
  const int32_t* fb_count;
  const int32_t* fb_rows;
  uint64_t* part_d;   // [SCAN_ROWS, SCAN_CHUNKS, k]
  uint32_t* part_i;
};

__device__ __forceinline__ void wave_list_insert64(uint64_t& ld, uint32_t& li, uint64_t cd, uint32_t ci, int k,
                                                   int l) {
  const int pos = __popcll(ballot(ld < cd || (ld == cd && li < ci)));
  const uint64_t upd = shfl_up_u64(ld, 1);
  const uint32_t upi = static_cast<uint32_t>(__shfl_up(static_cast<int>(li), 1));
  if (l > pos) {
    ld = upd;
    li = upi;
  } else if (l == pos) {
    ld = cd;
    li = ci;
  }
  if (l >= k) {
    ld = KEY_NONE;
    li = 0xFFFFFFFFu;
  }
}

__global__ __launch_bounds__(256) void k_scan_part(ScanPartArgs a) {
  __shared__ uint64_t s_d[4 * 64];
  __shared__ uint32_t s_i[4 * 64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int count = min(*a.fb_count, SCAN_ROWS);
  const int k = a.k;
  const int64_t per = (a.n + SCAN_CHUNKS - 1) / SCAN_CHUNKS;
  const int c0 = static_cast<int>(blockIdx.x * per);
  const int c1 = static_cast<int>(min(static_cast<int64_t>(a.n), c0 + per));
  for (int f = blockIdx.y; f < count; f += gridDim.y) {
    const int i = a.fb_rows[f];
    int64_t self = -1;
    if (a.q_self) self = a.q_self[i];
    else if (a.self_base >= 0) self = a.self_base + i;
    const float* qrow = a.Q + static_cast<int64_t>(i) * a.d;
    uint64_t ld = KEY_NONE, td = KEY_NONE;
    uint32_t li = 0xFFFFFFFFu, ti = 0xFFFFFFFFu;
    for (int j0 = c0 + w * 64; j0 < c1; j0 += 256) {
      const int j = j0 + l;
      const bool valid = j < c1 && j != self;
      uint64_t vb = KEY_NONE;
      if (valid) vb = static_cast<uint64_t>(__double_as_longlong(dist64_seq(qrow, a.X + static_cast<int64_t>(j) * a.d, a.d)));
      uint64_t m = ballot(valid && pair_less(vb, static_cast<uint32_t>(j), td, ti));
      while (m) {
        const int src = __ffsll(static_cast<long long>(m)) - 1;
        m &= m - 1;
        const uint64_t cd = readlane_u64(vb, src);
        const uint32_t ci = static_cast<uint32_t>(j0 + src);
        if (!pair_less(cd, ci, td, ti)) continue;
        wave_list_insert64(ld, li, cd, ci, k, l);
        td = readlane_u64(ld, k - 1);
        ti = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(li), k - 1));
      }
    }
    s_d[w * 64 + l] = ld;
    s_i[w * 64 + l] = li;
    __syncthreads();
    if (w == 0) {
      // merge: insert the other three lists into wave 0's list
      for (int q = 1; q < 4; ++q)
        for (int e = 0; e < k; ++e) {
          const uint64_t cd = s_d[q * 64 + e];
          const uint32_t ci = s_i[q * 64 + e];
          if (cd == KEY_NONE) break;
          if (!pair_less(cd, ci, td, ti)) break;   // lists are sorted: the rest cannot enter
          wave_list_insert64(ld, li, cd, ci, k, l);
          td = readlane_u64(ld, k - 1);
          ti = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(li), k - 1));
        }
      if (l < k) {
        const int64_t o = (static_cast<int64_t>(f) * SCAN_CHUNKS + blockIdx.x) * k + l;
        a.part_d[o] = ld;
        a.part_i[o] = li;
      }
    }
    __syncthreads();
  }
}

// Split exhaustive scan, part 2: one wavefront per failing row f <
// min(count, SCAN_ROWS) merges the SCAN_CHUNKS partial lists into the top-k.
struct ScanMergeArgs {
  const int32_t* fb_count;
  const int32_t* fb_rows;
  const uint64_t* part_d;
  const uint32_t* part_i;
  int k, euclid;
  int32_t* out_idx;
  float* out_dist;
  int32_t* stats;
};

__global__ __launch_bounds__(256) void k_scan_merge(ScanMergeArgs a) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int f = blockIdx.x * 4 + w;
  const int count = min(*a.fb_count, SCAN_ROWS);
  if (f >= count) return;
  const int k = a.k;
  const int M = SCAN_CHUNKS * k;   // <= 64 * 64 entries
  uint64_t outd = KEY_NONE;
  uint32_t outi = 0xFFFFFFFFu;
  int taken = 0;   // per lane: entries of its column consumed (lists are sorted per chunk)
  // lane l < SCAN_CHUNKS walks chunk l's sorted list: a k-way merge by wave minima
  const int64_t base = static_cast<int64_t>(f) * SCAN_CHUNKS * k;
  for (int j = 0; j < k; ++j) {
    uint64_t hd = KEY_NONE;
    uint32_t hi = 0xFFFFFFFFu;
    if (l < SCAN_CHUNKS && taken < k) {
      hd = a.part_d[base + static_cast<int64_t>(l) * k + taken];
      hi = a.part_i[base + static_cast<int64_t>(l) * k + taken];
    }
    uint64_t bd = hd;
    uint32_t bi = hi;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      const uint64_t od = shfl_xor_u64(bd, m);
      const uint32_t oi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(bi), m));
      if (pair_less(od, oi, bd, bi)) {
        bd = od;
        bi = oi;
      }
    }
    if (hd == bd && hi == bi && hd != KEY_NONE) ++taken;
    if (l == j) {
      outd = bd;
      outi = bi;
    }
  }
  (void)M;
  const int i = a.fb_rows[f];
  if (l < k) {
    const double dv = __longlong_as_double(static_cast<long long>(outd));
    a.out_idx[static_cast<int64_t>(i) * k + l] = static_cast<int32_t>(outi);
    a.out_dist[static_cast<int64_t>(i) * k + l] = a.euclid ? static_cast<float>(sqrt(dv)) : static_cast<float>(dv);
  }
  if (l == 0) atomicAdd(&a.stats[0], 1);
}

struct ScanArgs {
  const float* Q;
  const float* X;
  int64_t d;
  int n, k;
  const int32_t* q_self;
  int64_t self_base;
  const int32_t* fb_count;
  const int32_t* fb_rows;
  int euclid;
  int32_t* out_idx;
  float* out_dist;
  int32_t* stats;
};

// Exhaustive canonical scan of one failing row (index >= SCAN_ROWS in the
// failure list) per workgroup iteration: every
// lane computes FP64 sequential distances for its columns; each wavefront
// keeps a sorted top-k list; the four lists are merged at the end.
__global__ __launch_bounds__(256) void k_exact_scan(ScanArgs a) {
  __shared__ uint64_t s_d[4 * 64];
  __shared__ uint32_t s_i[4 * 64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int count = *a.fb_count;
  const int k = a.k;
  // rows below SCAN_ROWS are handled by the split scan (k_scan_part / k_scan_merge)
  for (int f = SCAN_ROWS + blockIdx.x; f < count; f += gridDim.x) {
    const int i = a.fb_rows[f];
    int64_t self = -1;
    if (a.q_self) self = a.q_self[i];
    else if (a.self_base >= 0) self = a.self_base + i;
    const float* qrow = a.Q + static_cast<int64_t>(i) * a.d;

    uint64_t ld = KEY_NONE;   // this lane's list entry (distance bits)
    uint32_t li = 0xFFFFFFFFu;
    uint64_t td = KEY_NONE;   // list's k-th entry
    uint32_t ti = 0xFFFFFFFFu;
    for (int j0 = w * 64; j0 < a.n; j0 += 256) {
      const int j = j0 + l;
      const bool valid = j < a.n && j != self;
      uint64_t vb = KEY_NONE;
      if (valid) {
        const double v = dist64_seq(qrow, a.X + static_cast<int64_t>(j) * a.d, a.d);
        vb = static_cast<uint64_t>(__double_as_longlong(v));
      }
      uint64_t m = ballot(valid && pair_less(vb, static_cast<uint32_t>(j), td, ti));
      while (m) {
        const int src = __ffsll(static_cast<long long>(m)) - 1;
        m &= m - 1;
        const uint64_t cd = readlane_u64(vb, src);
        const uint32_t ci = static_cast<uint32_t>(j0 + src);
        if (!pair_less(cd, ci, td, ti)) continue;
        const int pos = __popcll(ballot(pair_less(ld, li, cd, ci)));
        const uint64_t upd = shfl_up_u64(ld, 1);
        const uint32_t upi = static_cast<uint32_t>(__shfl_up(static_cast<int>(li), 1));
        if (l > pos) {
          ld = upd;
          li = upi;
        } else if (l == pos) {
          ld = cd;
          li = ci;
        }
        if (l >= k) {
          ld = KEY_NONE;
          li = 0xFFFFFFFFu;
        }
        td = readlane_u64(ld, k - 1);
        ti = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(li), k - 1));
      }
    }
    s_d[w * 64 + l] = ld;
    s_i[w * 64 + l] = li;
    __syncthreads();
    if (w == 0) {
      uint64_t ed[4];
      uint32_t ei[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ed[q] = s_d[q * 64 + l];
        ei[q] = s_i[q * 64 + l];
      }
      uint64_t outd = KEY_NONE;
      uint32_t outi = 0xFFFFFFFFu;
      for (int j = 0; j < k; ++j) {
        uint64_t bd = ed[0];
        uint32_t bi = ei[0];
#pragma unroll
        for (int q = 1; q < 4; ++q)
          if (pair_less(ed[q], ei[q], bd, bi)) {
            bd = ed[q];
            bi = ei[q];
          }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
          const uint64_t od = shfl_xor_u64(bd, m);
          const uint32_t oi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(bi), m));
          if (pair_less(od, oi, bd, bi)) {
            bd = od;
            bi = oi;
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (ed[q] == bd && ei[q] == bi) {
            ed[q] = KEY_NONE;
            ei[q] = 0xFFFFFFFFu;
          }
        if (l == j) {
          outd = bd;
          outi = bi;
        }
      }
      if (l < k) {
        const double dv = __longlong_as_double(static_cast<long long>(outd));
        a.out_idx[static_cast<int64_t>(i) * k + l] = static_cast<int32_t>(outi);
        a.out_dist[static_cast<int64_t>(i) * k + l] =
            a.euclid ? static_cast<float>(sqrt(dv)) : static_cast<float>(dv);
      }
      if (l == 0) atomicAdd(&a.stats[0], 1);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------- host --

struct Layout {
  int64_t nq_pad, n_pad, d_pad, nchunks;   // d_pad: FP32 row (multiple of 32)
  int64_t rows_per;                        // rows per column-sum partial
  int64_t dq_pad;                          // quantised row elements (128-byte rows)
  int C, dtype;
  bool same;
  bool sym;                                // symmetric screen: full distance matrix in the workspace
  bool lowd;                               // d <= LOWD_MAX_D: exact grid search (lowd.hip), nothing else
  size_t bytes;
};

struct Buffers {
  double* part;
  float* mu;
  float *Xc, *sqx, *Qc, *sqq;                   // FP32 centred copies
  char *Xq, *Qq;                                // quantised copies (dtype != FP32)
  float *sqxq, *sqqq, *ex, *eq, *invx, *invq;   // quantised norms, error bounds, 1/scale
  uint64_t* cand;
  uint64_t* pend;   // screen survivor buffers
  int32_t* ctr;   // [0] fb1 count, [1] sqmax32, [2] sqmaxq, [3] emax, [4] fb2 count, [6] re-screen rows, [8..11] stats
  int32_t *fb1, *fb2;
  uint64_t* scan_d;                             // split exhaustive scan partial lists
  uint32_t* scan_i;
  float *Qc2, *sqq2;                            // FP32 re-screen rows
  char *Xs, *Qs2;                               // split BF16 copies for the re-screen
  int32_t* self2;
  float* Dfull;                                 // symmetric screen: [n_pad, n_pad] screen values
  float* bmin;                                  // symmetric screen: [n_pad, n_pad / 128] block minima
};

static Buffers carve(void* ws, int64_t nq, int64_t n, int64_t d, const Layout& L, size_t* total) {
  Carve c(ws);
  Buffers b{};
  b.part = c.take<double>(static_cast<size_t>(L.nchunks) * d);
  b.mu = c.take<float>(L.d_pad > L.dq_pad ? L.d_pad : L.dq_pad);
  b.Xc = c.take<float>(static_cast<size_t>(L.n_pad) * L.d_pad);
  b.sqx = c.take<float>(L.n_pad);
  if (L.same) {
    b.Qc = b.Xc;
    b.sqq = b.sqx;
  } else {
    b.Qc = c.take<float>(static_cast<size_t>(L.nq_pad) * L.d_pad);
    b.sqq = c.take<float>(L.nq_pad);
  }
  if (L.dtype == MV_BF16X3) {
    const size_t rb = static_cast<size_t>(L.dq_pad) * elem_bytes(L.dtype);
    b.Xq = c.take<char>(static_cast<size_t>(L.n_pad) * rb);
    b.Qq = L.same ? b.Xq : c.take<char>(static_cast<size_t>(L.nq_pad) * rb);
  } else if (L.dtype != MV_F32) {
    const size_t rb = static_cast<size_t>(L.dq_pad) * elem_bytes(L.dtype);
    b.Xq = c.take<char>(static_cast<size_t>(L.n_pad) * rb);
    b.sqxq = c.take<float>(L.n_pad);
    b.ex = c.take<float>(L.n_pad);
    b.invx = c.take<float>(L.n_pad);
    if (L.same) {
      b.Qq = b.Xq;
      b.sqqq = b.sqxq;
      b.eq = b.ex;
      b.invq = b.invx;
    } else {
      b.Qq = c.take<char>(static_cast<size_t>(L.nq_pad) * rb);
      b.sqqq = c.take<float>(L.nq_pad);
      b.eq = c.take<float>(L.nq_pad);
      b.invq = c.take<float>(L.nq_pad);
    }
    b.Qc2 = c.take<float>(static_cast<size_t>(L.nq_pad) * L.d_pad);
    b.sqq2 = c.take<float>(L.nq_pad);
    b.self2 = c.take<int32_t>(L.nq_pad);
    b.Xs = c.take<char>(static_cast<size_t>(L.n_pad) * L.d_pad * 4);
    b.Qs2 = c.take<char>(static_cast<size_t>(L.nq_pad) * L.d_pad * 4);
  }
  b.cand = c.take<uint64_t>(static_cast<size_t>(MAX_SPLIT) * L.nq_pad * L.C);
  b.pend = c.take<uint64_t>(static_cast<size_t>(MAX_SPLIT) * L.nq_pad * PEND);
  b.ctr = c.take<int32_t>(16);
  b.fb1 = c.take<int32_t>(nq > 0 ? nq : 1);
  b.fb2 = c.take<int32_t>(nq > 0 ? nq : 1);
  const int64_t srows = nq < SCAN_ROWS ? (nq > 0 ? nq : 1) : SCAN_ROWS;
  b.scan_d = c.take<uint64_t>(static_cast<size_t>(srows) * SCAN_CHUNKS * CMAX_K);
  b.scan_i = c.take<uint32_t>(static_cast<size_t>(srows) * SCAN_CHUNKS * CMAX_K);
  if (L.sym) {
    // BF16 and FP8 screens: BF16 values, half the bytes
    const bool s16 = L.dtype == MV_BF16 || L.dtype == MV_FP8;
    b.Dfull = c.take<float>(static_cast<size_t>(L.n_pad) * L.n_pad / (s16 ? 2 : 1));
    b.bmin = c.take<float>(static_cast<size_t>(L.n_pad) * (L.n_pad / BN));
  }
  *total = c.off + 256;
  return b;
}

// Symmetric screen for self kNN (mv_knn_set_symmetric; on by default).
static bool g_symmetric = true;

// Static scratch of mv_knn_brute_noalloc: the workspace of its calls, for
// callers that cannot allocate. Each GPU the library runs on holds one copy
// (128 MiB) until the process ends; every call on a GPU shares it, so calls
// that overlap in time on one GPU race. Host code never passes `scratch`
// itself to a kernel (in host code the name is only a stand-in): it uses the
// device address that hipGetSymbolAddress returns for the current device.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
static int noalloc_scratch(void** out) {
  static std::mutex mu;
  static void* addr[64] = {nullptr};
  int dev = 0;
  MV_TRY_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    mv::set_error("mv_knn_brute_noalloc: device ordinal %d outside [0, 64)", dev);
    return MV_EUNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    MV_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      mv::set_error("mv_knn_brute_noalloc: hipGetSymbolAddress returned no address for the scratch array");
      return MV_EHIP;
    }
    addr[dev] = p;
  }
  *out = addr[dev];
  return MV_OK;
}

static Layout make_layout(int64_t nq, int64_t n, int64_t d, int k, int rerank, bool same, int dtype,
                          bool sym_ok) {
  Layout L{};
  if (d <= LOWD_MAX_D && k <= CMAX_K) {
    L.lowd = true;
    L.bytes = lowd_knn_ws_bytes(n);
    return L;
  }
  L.nq_pad = static_cast<int64_t>(align_up(nq > 0 ? nq : 1, BM));
  L.n_pad = static_cast<int64_t>(align_up(n, BN));
  L.d_pad = static_cast<int64_t>(align_up(d, KC));
  L.dq_pad = dtype == MV_F32 ? L.d_pad : static_cast<int64_t>(align_up(d, 128 / elem_bytes(dtype)));
  L.rows_per = n > 512 * ROW_CHUNK ? ceil_div(n, 512) : ROW_CHUNK;   // at most 512 column-sum partials
  L.nchunks = ceil_div(n, L.rows_per);
  L.dtype = dtype;
  L.same = same;
  // k above 64 is selected from the symmetric screen's matrix only, at any n
  L.sym = same && sym_ok && (n >= SYM_MIN_N || k > CMAX_K) &&
          static_cast<size_t>(L.n_pad) * L.n_pad * 4 <= SYM_MAX_BYTES;
  // candidates per row: up to 128 from the symmetric screen's selection, 64
  // from the fused screen's per-lane lists
  int C = rerank * k;
  if (C < k) C = k;
  const int cmax = L.sym ? CMAX_SYM : CMAX;
  if (C > cmax) C = cmax;
  L.C = C;
  size_t total = 0;
  carve(nullptr, nq, n, d, L, &total);
  L.bytes = total;
  return L;
}

// Optional phase timing (mv_knn_set_timing): HIP events recorded on the launch
// stream around the prepass, screen, rerank and exact-scan launches, and
// between the distance kernel and the selection of the symmetric screen.
struct PhaseTimer {
  bool on = false;
  bool valid = false;
  bool last_sym = false;   // the last timed call used the symmetric screen (ev[5]: between its two kernels)
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int ensure() {
    for (auto& e : ev)
      if (!e) MV_TRY_HIP(hipEventCreate(&e));
    return MV_OK;
  }
  int mark(int i, hipStream_t s) {
    if (!on) return MV_OK;
    MV_TRY_HIP(hipEventRecord(ev[i], s));
    return MV_OK;
  }
};
static thread_local PhaseTimer g_timer;

// Blocks for the row-strided prepass kernels: one wavefront per row up to 4,096
// blocks, so that their per-block atomics stay few.
static int64_t prep_grid(int64_t rows) {
  const int64_t b = ceil_div(rows, 4);
  return b < 4096 ? b : 4096;
}

static int cu_count_cached() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cached[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      v = 256;
    cached[dev] = v;
  }
  return cached[dev];
}

// Column split count: balances whole waves of workgroups (one per CU) against
// the per-split merge work (S * C <= 256 keys per row in the re-rank).
static int choose_split(int64_t row_blocks, int64_t col_tiles, int C, int cus) {
  int best = 1;
  double best_score = -1.0;
  for (int S = 1; S <= MAX_SPLIT; S <<= 1) {
    if (S > col_tiles || S * C > 256) break;
    const double wgs = static_cast<double>(row_blocks * S);
    const double waves = ceil(wgs / cus);
    const double wave_eff = wgs / (waves * cus);
    const int64_t tps = ceil_div(col_tiles, S);
    const double tile_eff = static_cast<double>(col_tiles) / static_cast<double>(tps * S);
    const double score = wave_eff * tile_eff;
    if (score > best_score + 0.02) {
      best_score = score;
      best = S;
    }
  }
  return best;
}

template <int DT>
int launch_screen(const ScreenArgs& sa, int64_t grid, hipStream_t s) {
  hipLaunchKernelGGL((k_screen<DT>), dim3(static_cast<unsigned>(grid)), dim3(NT), 0, s, sa);
  MV_LAUNCHED("k_screen");
  return MV_OK;
}

// The whole brute-force kNN of mv_knn_brute on a given workspace. `sym_ok`
// permits the symmetric screen (self kNN), `timed` the phase events of
// mv_knn_set_timing; `who` names the entry point in error messages.
// Screen arithmetic bound, in units of (|q|^2 + max |x|^2). FP32: two
// roundings per product term plus the norms and the final combination.
// Split BF16: products are exact, one rounding per each of the 3 d_pad
// accumulated terms, plus the dropped lo*lo term and the split residuals
// (each <= 2^-16 relative: 3 * 256 units of 2^-24).
static float screen_err_fac(int dtype, const Layout& L, int64_t row_bytes) {
  const double units = dtype == MV_BF16X3 ? 3.0 * static_cast<double>(L.d_pad) + 3.0 * 256.0 + 40.0
                                          : 2.0 * static_cast<double>(row_bytes / elem_bytes(dtype)) + 10.0;
  return static_cast<float>(units * 5.9604644775390625e-8 * 1.01);
}

// The selection from the matrix needs screen values that carry only the
// arithmetic bound: there, BF16 and FP8 requests run the split-BF16 screen
// (the result does not depend on the screen).
static int large_k_dtype(int k, int dtype, bool sym) {
  return sym && k >= K_LARGE_FROM && (dtype == MV_BF16 || dtype == MV_FP8) ? MV_BF16X3 : dtype;
}

static int knn_run(const char* who, const float* Q, int64_t nq, const float* X, int64_t n, int64_t d, int k,
                   int dtype, int rerank, const int32_t* q_self, int64_t self_base, int euclid, int32_t* out_idx,
                   float* out_dist, int32_t* out_stats, void* ws, size_t ws_bytes, void* stream, bool sym_ok,
                   bool timed) {
  const bool same = (Q == X) && nq == n;
  Layout L = make_layout(nq, n, d, k, rerank, same, dtype, sym_ok);
  if (large_k_dtype(k, dtype, L.sym) != dtype) {
    dtype = large_k_dtype(k, dtype, L.sym);
    L = make_layout(nq, n, d, k, rerank, same, dtype, sym_ok);
  }
  if (k > CMAX_K && !L.sym) {
    mv::set_error("%s: k=%d above %d needs the symmetric self-kNN screen (Q == X, an n x n matrix of at most %zu "
                  "bytes, mv_knn_set_symmetric on)", who, k, CMAX_K, SYM_MAX_BYTES);
    return MV_EUNSUPPORTED;
  }
  if (ws_bytes < L.bytes || !ws) {
    mv::set_error("%s: workspace of %zu bytes < required %zu", who, ws_bytes, L.bytes);
    return MV_EWORKSPACE;
  }
  if (nq == 0) return MV_OK;
  if (L.lowd) {
    // one to three dimensions: the exact grid search (no screen, no certificate)
    if (out_stats) {
      hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), out_stats, 4);
      MV_LAUNCHED("k_zero_i32");
    }
    return lowd_knn(Q, nq, X, n, static_cast<int>(d), k, q_self, self_base, euclid, out_idx, out_dist, ws, ws_bytes,
                    static_cast<hipStream_t>(stream));
  }
  size_t total = 0;
  Buffers B = carve(ws, nq, n, d, L, &total);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int32_t* fb1_count = B.ctr;
  uint32_t* sqmax32 = reinterpret_cast<uint32_t*>(B.ctr + 1);
  uint32_t* sqmaxq = reinterpret_cast<uint32_t*>(B.ctr + 2);
  uint32_t* emax = reinterpret_cast<uint32_t*>(B.ctr + 3);
  int32_t* fb2_count = B.ctr + 4;
  int32_t* stats = out_stats ? out_stats : B.ctr + 8;
  const int64_t self_b = q_self ? -1 : self_base;

  PhaseTimer quiet;                              // records nothing
  PhaseTimer& tm = timed ? g_timer : quiet;
  if (tm.on) {
    if (int rc = tm.ensure()) return rc;
    tm.valid = false;
  }
  if (int rc = tm.mark(0, s)) return rc;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, B.ctr, 16);
  MV_LAUNCHED("k_zero_i32");
  if (out_stats) {
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, out_stats, 4);
    MV_LAUNCHED("k_zero_i32");
  }
  hipLaunchKernelGGL(k_colsum_partial, dim3(static_cast<unsigned>(ceil_div(d, 64)), static_cast<unsigned>(L.nchunks)),
                     dim3(256), 0, s, X, n, d, L.rows_per, B.part);
  MV_LAUNCHED("k_colsum_partial");
  const int64_t mu_len = L.d_pad > L.dq_pad ? L.d_pad : L.dq_pad;
  hipLaunchKernelGGL(k_colsum_final, dim3(static_cast<unsigned>(ceil_div(mu_len, 4))), dim3(256), 0, s, B.part,
                     L.nchunks, n, d, mu_len, B.mu);
  MV_LAUNCHED("k_colsum_final");
  hipLaunchKernelGGL(k_center_f32, dim3(static_cast<unsigned>(prep_grid(L.n_pad))), dim3(256), 0, s, X, n, d,
                     L.d_pad, L.n_pad, B.mu, B.Xc, B.sqx, sqmax32);
  MV_LAUNCHED("k_center_f32");
  if (!same) {
    hipLaunchKernelGGL(k_center_f32, dim3(static_cast<unsigned>(prep_grid(L.nq_pad))), dim3(256), 0, s, Q, nq, d,
                       L.d_pad, L.nq_pad, B.mu, B.Qc, B.sqq, static_cast<uint32_t*>(nullptr));
    MV_LAUNCHED("k_center_f32");
  }
  if (dtype == MV_BF16X3) {
    const int64_t cx = L.n_pad * L.d_pad;
    hipLaunchKernelGGL(k_split_bf16, dim3(static_cast<unsigned>(ceil_div(cx, 256))), dim3(256), 0, s, B.Xc, cx,
                       L.d_pad, reinterpret_cast<uint16_t*>(B.Xq), static_cast<const int32_t*>(nullptr));
    MV_LAUNCHED("k_split_bf16");
    if (!same) {
      const int64_t cq = L.nq_pad * L.d_pad;
      hipLaunchKernelGGL(k_split_bf16, dim3(static_cast<unsigned>(ceil_div(cq, 256))), dim3(256), 0, s, B.Qc, cq,
                         L.d_pad, reinterpret_cast<uint16_t*>(B.Qq), static_cast<const int32_t*>(nullptr));
      MV_LAUNCHED("k_split_bf16");
    }
  } else if (dtype != MV_F32) {
    auto kq = dtype == MV_BF16 ? k_center_q<DT_BF16> : k_center_q<DT_FP8>;
    hipLaunchKernelGGL(kq, dim3(static_cast<unsigned>(prep_grid(L.n_pad))), dim3(256), 0, s, X, n, d, L.dq_pad, L.n_pad,
                       B.mu, B.Xq, B.sqxq, B.ex, B.invx, sqmaxq, emax);
    MV_LAUNCHED("k_center_q");
    if (!same) {
      uint32_t* dummy = reinterpret_cast<uint32_t*>(B.ctr + 5);
      hipLaunchKernelGGL(kq, dim3(static_cast<unsigned>(prep_grid(L.nq_pad))), dim3(256), 0, s, Q, nq, d, L.dq_pad,
                         L.nq_pad, B.mu, B.Qq, B.sqqq, B.eq, B.invq, dummy, dummy);
      MV_LAUNCHED("k_center_q");
    }
  }

  const int64_t row_blocks = L.nq_pad / BM;
  const int64_t col_tiles = L.n_pad / BN;
  const int Cf = L.C < CMAX ? L.C : CMAX;   // candidates of a fused screen (also the re-screen)
  const int S = choose_split(row_blocks, col_tiles, Cf, 2 * cu_count_cached());   // two workgroups per CU
  const int64_t grid = ceil_div(row_blocks, 8 / S) * 8;
  ScreenArgs sa;
  sa.q_self = q_self;
  sa.self_base = self_b;
  sa.nq_dev = nullptr;
  sa.nq = static_cast<int>(nq);
  sa.n = static_cast<int>(n);
  sa.C = L.C;
  sa.S = S;
  sa.tiles_per_split = static_cast<int>(ceil_div(col_tiles, S));
  sa.n_col_tiles = static_cast<int>(col_tiles);
  sa.n_row_blocks = static_cast<int>(row_blocks);
  sa.nq_pad = static_cast<int>(L.nq_pad);
  sa.cand = B.cand;
  sa.pend = B.pend;
  if (dtype == MV_F32 || dtype == MV_BF16X3) {
    sa.Qq = dtype == MV_F32 ? reinterpret_cast<const char*>(B.Qc) : B.Qq;
    sa.Xq = dtype == MV_F32 ? reinterpret_cast<const char*>(B.Xc) : B.Xq;
    sa.sqq = B.sqq;
    sa.sqx = B.sqx;
    sa.invq = sa.invx = nullptr;
    sa.row_bytes = L.d_pad * 4;
  } else {
    sa.Qq = B.Qq;
    sa.Xq = B.Xq;
    sa.sqq = B.sqqq;
    sa.sqx = B.sqxq;
    sa.invq = B.invq;
    sa.invx = B.invx;
    sa.row_bytes = L.dq_pad * elem_bytes(dtype);
  }
  if (int rc = tm.mark(1, s)) return rc;
  if (L.sym) {
    // symmetric screen: upper-triangle tiles into the full matrix, then one
    // list per row (a single split)
    SymArgs ya;
    ya.Xq = sa.Xq;
    ya.sqx = sa.sqx;
    ya.invx = sa.invx;
    ya.row_bytes = sa.row_bytes;
    ya.nb = static_cast<int>(col_tiles);
    ya.nsb = static_cast<int>(ceil_div(col_tiles, SYM_G));
    ya.n_slots = static_cast<int64_t>(ya.nsb) * (ya.nsb + 1) / 2 * SYM_G * SYM_G;
    ya.per_xcd = ceil_div(ya.n_slots, 8);
    ya.ldd = L.n_pad;
    ya.D = B.Dfull;
    ya.bmin = B.bmin;
    ya.n = static_cast<int>(n);
    ya.diag_self = q_self == nullptr && self_b == 0;
    const unsigned sgrid = static_cast<unsigned>(ya.per_xcd * 8);
    auto ks = dtype == MV_F32 ? k_dist_sym<DT_F32, SYM_W>
              : dtype == MV_BF16 ? k_dist_sym<DT_BF16, SYM_W>
              : dtype == MV_BF16X3 ? k_dist_sym<DT_BF16X3, SYM_W> : k_dist_sym<DT_FP8, SYM_W>;
    hipLaunchKernelGGL(ks, dim3(sgrid), dim3(64 * SYM_W), 0, s, ya);
    MV_LAUNCHED("k_dist_sym");
    if (int rc = tm.mark(5, s)) return rc;
    if (k >= K_LARGE_FROM) {
      // exact selection straight from the matrix (FP32 or split-BF16 screen)
      LargeArgs ga;
      ga.D = B.Dfull;
      ga.ldd = L.n_pad;
      ga.bmin = (q_self == nullptr && self_b <= 0) ? B.bmin : nullptr;
      ga.nb = static_cast<int>(col_tiles);
      ga.n = static_cast<int>(n);
      ga.nq = static_cast<int>(nq);
      ga.k = k;
      ga.q_self = q_self;
      ga.self_base = self_b;
      ga.sqq = sa.sqq;
      ga.sqmax_bits = sqmax32;
      ga.err_fac = screen_err_fac(dtype, L, sa.row_bytes);
      ga.Q = Q;
      ga.X = X;
      ga.d = d;
      ga.euclid = euclid;
      ga.out_idx = out_idx;
      ga.out_dist = out_dist;
      hipLaunchKernelGGL(k_knn_large, dim3(static_cast<unsigned>(nq)), dim3(256), 0, s, ga);
      MV_LAUNCHED("k_knn_large");
      if (int rc = tm.mark(2, s)) return rc;
      if (int rc = tm.mark(3, s)) return rc;
      if (int rc = tm.mark(4, s)) return rc;
      tm.valid = tm.on;
      tm.last_sym = true;
      return MV_OK;
    }
    SelectArgs xa;
    xa.D = B.Dfull;
    xa.ldd = L.n_pad;
    xa.n = static_cast<int>(n);
    xa.nq = static_cast<int>(nq);
    xa.C = L.C;
    xa.q_self = q_self;
    xa.self_base = self_b;
    // block pruning needs the block minima to range over exactly the
    // selectable columns: the diagonal excluded (self_base 0, as the minima
    // leave it out) or no exclusion at all; an arbitrary q_self reads every block
    xa.bmin = (q_self == nullptr && self_b <= 0) ? B.bmin : nullptr;
    xa.nb = static_cast<int>(col_tiles);
    xa.cand = B.cand;
    const bool s16 = dtype == MV_BF16 || dtype == MV_FP8;
    auto ksel = L.C > 64 ? (s16 ? k_select<2, true> : k_select<2, false>)
                         : (s16 ? k_select<1, true> : k_select<1, false>);
    hipLaunchKernelGGL(ksel, dim3(static_cast<unsigned>(ceil_div(nq, 4))), dim3(256), 0, s, xa);
    MV_LAUNCHED("k_select");
  } else {
    int rc = dtype == MV_F32      ? launch_screen<DT_F32>(sa, grid, s)
             : dtype == MV_BF16   ? launch_screen<DT_BF16>(sa, grid, s)
             : dtype == MV_BF16X3 ? launch_screen<DT_BF16X3>(sa, grid, s)
                                  : launch_screen<DT_FP8>(sa, grid, s);
    if (rc) return rc;
  }
  if (int rc2 = tm.mark(2, s)) return rc2;

  const float fac = screen_err_fac(dtype, L, sa.row_bytes);
  RerankArgs ra;
  ra.Q = Q;
  ra.X = X;
  ra.d = d;
  ra.nq = static_cast<int>(nq);
  ra.n = static_cast<int>(n);
  ra.k = k;
  ra.C = L.C;
  ra.S = L.sym ? 1 : S;
  ra.nq_pad = static_cast<int>(L.nq_pad);
  ra.cand = B.cand;
  ra.sqq = sa.sqq;
  const bool exact_form = dtype == MV_F32 || dtype == MV_BF16X3;   // no quantisation-error term
  ra.sqmax_bits = exact_form ? sqmax32 : sqmaxq;
  ra.err_fac = fac;
  ra.store_lo = L.sym && (dtype == MV_BF16 || dtype == MV_FP8) ? STORE16_LO : 1.0;
  ra.eq = exact_form ? nullptr : B.eq;
  ra.emax_bits = emax;
  ra.nq_dev = nullptr;
  ra.row_map = nullptr;
  ra.fail_stat = 2;
  ra.vec4 = (d % 4 == 0) && (reinterpret_cast<uintptr_t>(Q) % 16 == 0) && (reinterpret_cast<uintptr_t>(X) % 16 == 0);
  ra.euclid = euclid;
  ra.out_idx = out_idx;
  ra.out_dist = out_dist;
  ra.fb_count = fb1_count;
  ra.fb_rows = B.fb1;
  ra.stats = stats;
  if (L.C > 64) hipLaunchKernelGGL(k_rerank2, dim3(static_cast<unsigned>(ceil_div(nq, 4))), dim3(256), 0, s, ra);
  else hipLaunchKernelGGL(k_rerank, dim3(static_cast<unsigned>(ceil_div(nq, 4))), dim3(256), 0, s, ra);
  MV_LAUNCHED("k_rerank");
  if (int rc2 = tm.mark(3, s)) return rc2;

  const int32_t* scan_count = fb1_count;
  const int32_t* scan_rows = B.fb1;
  if (!exact_form) {
    // Re-screen of the rows the quantised certificate could not settle, with
    // the split-BF16 screen (FP32-class error bound, no quantisation term),
    // when there are more than SCAN_ROWS of them; fewer go to the scan.
    int32_t* rs_count = B.ctr + 6;
    hipLaunchKernelGGL(k_route_small, dim3(static_cast<unsigned>(ceil_div(SCAN_ROWS, 256))), dim3(256), 0, s,
                       fb1_count, B.fb1, fb2_count, B.fb2, rs_count, stats, SCAN_ROWS);
    MV_LAUNCHED("k_route_small");
    hipLaunchKernelGGL(k_gather_rows, dim3(static_cast<unsigned>(ceil_div(nq, 4))), dim3(256), 0, s, B.Qc, B.sqq,
                       L.d_pad, B.fb1, rs_count, q_self, self_b, B.Qc2, B.sqq2, B.self2);
    MV_LAUNCHED("k_gather_rows");
    const int64_t cxs = L.n_pad * L.d_pad, cqs = L.nq_pad * L.d_pad;
    hipLaunchKernelGGL(k_split_bf16, dim3(static_cast<unsigned>(ceil_div(cxs, 256))), dim3(256), 0, s, B.Xc, cxs,
                       L.d_pad, reinterpret_cast<uint16_t*>(B.Xs), static_cast<const int32_t*>(rs_count));
    MV_LAUNCHED("k_split_bf16");
    hipLaunchKernelGGL(k_split_bf16, dim3(static_cast<unsigned>(ceil_div(cqs, 256))), dim3(256), 0, s, B.Qc2, cqs,
                       L.d_pad, reinterpret_cast<uint16_t*>(B.Qs2), static_cast<const int32_t*>(rs_count));
    MV_LAUNCHED("k_split_bf16");
    ScreenArgs sb = sa;
    sb.C = Cf;
    sb.Qq = B.Qs2;
    sb.Xq = B.Xs;
    sb.sqq = B.sqq2;
    sb.sqx = B.sqx;
    sb.invq = sb.invx = nullptr;
    sb.row_bytes = L.d_pad * 4;
    sb.q_self = B.self2;
    sb.self_base = -1;
    sb.nq_dev = rs_count;
    if (int rc2 = launch_screen<DT_BF16X3>(sb, grid, s)) return rc2;
    RerankArgs rb2 = ra;
    rb2.S = S;   // the fused screen's lists (also after a symmetric first pass)
    rb2.C = Cf;
    rb2.sqq = B.sqq2;
    rb2.sqmax_bits = sqmax32;
    rb2.err_fac = static_cast<float>((3.0 * static_cast<double>(L.d_pad) + 3.0 * 256.0 + 40.0) *
                                     5.9604644775390625e-8 * 1.01);
    rb2.store_lo = 1.0;   // the fused screen's FP32 lists
    rb2.eq = nullptr;
    rb2.nq_dev = rs_count;
    rb2.row_map = B.fb1;
    rb2.fail_stat = 3;
    rb2.fb_count = fb2_count;
    rb2.fb_rows = B.fb2;
    hipLaunchKernelGGL(k_rerank, dim3(static_cast<unsigned>(ceil_div(nq, 4))), dim3(256), 0, s, rb2);
    MV_LAUNCHED("k_rerank");
    scan_count = fb2_count;
    scan_rows = B.fb2;
  }

  const int64_t srows = nq < SCAN_ROWS ? nq : SCAN_ROWS;
  ScanPartArgs pa;
  pa.Q = Q;
  pa.X = X;
  pa.d = d;
  pa.n = static_cast<int>(n);
  pa.k = k;
  pa.q_self = q_self;
  pa.self_base = self_b;
  pa.fb_count = scan_count;
  pa.fb_rows = scan_rows;
  pa.part_d = B.scan_d;
  pa.part_i = B.scan_i;
  hipLaunchKernelGGL(k_scan_part, dim3(SCAN_CHUNKS, static_cast<unsigned>(srows)), dim3(256), 0, s, pa);
  MV_LAUNCHED("k_scan_part");
  ScanMergeArgs ma;
  ma.fb_count = scan_count;
  ma.fb_rows = scan_rows;
  ma.part_d = B.scan_d;
  ma.part_i = B.scan_i;
  ma.k = k;
  ma.euclid = euclid;
  ma.out_idx = out_idx;
  ma.out_dist = out_dist;
  ma.stats = stats;
  hipLaunchKernelGGL(k_scan_merge, dim3(static_cast<unsigned>(ceil_div(srows, 4))), dim3(256), 0, s, ma);
  MV_LAUNCHED("k_scan_merge");
  if (nq > SCAN_ROWS) {
    ScanArgs xa;
    xa.Q = Q;
    xa.X = X;
    xa.d = d;
    xa.n = static_cast<int>(n);
    xa.k = k;
    xa.q_self = q_self;
    xa.self_base = self_b;
    xa.fb_count = scan_count;
    xa.fb_rows = scan_rows;
    xa.euclid = euclid;
    xa.out_idx = out_idx;
    xa.out_dist = out_dist;
    xa.stats = stats;
    const int64_t rest = nq - SCAN_ROWS;
    const int64_t sgrid = rest < 2048 ? rest : 2048;
    hipLaunchKernelGGL(k_exact_scan, dim3(static_cast<unsigned>(sgrid)), dim3(256), 0, s, xa);
    MV_LAUNCHED("k_exact_scan");
  }
  if (int rc2 = tm.mark(4, s)) return rc2;
  tm.valid = tm.on;
  tm.last_sym = L.sym;
  return MV_OK;
}

}  // namespace knn
}  // namespace mv

using namespace mv;
using namespace mv::knn;

extern "C" {

int mv_knn_ws_bytes(int64_t nq, int64_t n, int64_t d, int k, int dtype, int rerank, int same_qx,
                    size_t* out_bytes) {
  MV_REQUIRE(out_bytes, "mv_knn_ws_bytes: out_bytes is NULL");
  MV_REQUIRE(nq >= 0 && n >= 1 && d >= 1, "mv_knn_ws_bytes: bad shape nq=%lld n=%lld d=%lld",
             (long long)nq, (long long)n, (long long)d);
  MV_REQUIRE(k >= 1 && k <= KMAX_LARGE, "mv_knn_ws_bytes: k=%d outside [1, %d]", k, KMAX_LARGE);
  MV_REQUIRE(rerank >= 1, "mv_knn_ws_bytes: rerank=%d < 1", rerank);
  if (dtype != MV_F32 && dtype != MV_BF16 && dtype != MV_FP8 && dtype != MV_BF16X3) {
    mv::set_error("mv_knn_ws_bytes: dtype %d not supported", dtype);
    return MV_EUNSUPPORTED;
  }
  Layout L = make_layout(nq, n, d, k, rerank, same_qx != 0, dtype, g_symmetric);
  if (large_k_dtype(k, dtype, L.sym) != dtype)
    L = make_layout(nq, n, d, k, rerank, same_qx != 0, large_k_dtype(k, dtype, L.sym), g_symmetric);
  if (k > CMAX_K && !L.sym) {
    mv::set_error("mv_knn_ws_bytes: k=%d above %d needs the symmetric self-kNN screen (same_qx, an n x n matrix of "
                  "at most %zu bytes, mv_knn_set_symmetric on)", k, CMAX_K, SYM_MAX_BYTES);
    return MV_EUNSUPPORTED;
  }
  *out_bytes = L.bytes;
  return MV_OK;
}

int mv_row_sqnorms(const float* X, int64_t n, int64_t d, float* out, void* stream) {
  MV_FAULT_POINT("launch");
  MV_REQUIRE(X && out, "mv_row_sqnorms: null pointer");
  MV_REQUIRE(n >= 0 && d >= 1, "mv_row_sqnorms: bad shape");
  if (n == 0) return MV_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_row_sqnorms, dim3(static_cast<unsigned>(ceil_div(n, 4))), dim3(256), 0, s, X, n, d,
                     out);
  MV_LAUNCHED("k_row_sqnorms");
  return MV_OK;
}

// Argument checks shared by mv_knn_brute and mv_knn_brute_noalloc.
static int knn_check(const char* who, const float* Q, int64_t nq, const float* X, int64_t n, int64_t d, int k,
                     int dtype, int rerank, const int32_t* q_self, int64_t self_base, int32_t* out_idx,
                     float* out_dist) {
  MV_REQUIRE(Q && X && out_idx && out_dist, "%s: null pointer argument", who);
  MV_REQUIRE(nq >= 0 && n >= 1 && d >= 1, "%s: bad shape nq=%lld n=%lld d=%lld", who, (long long)nq,
             (long long)n, (long long)d);
  MV_REQUIRE(n < (1ll << 31) - 256 && nq < (1ll << 31) - 256, "%s: n or nq too large", who);
  MV_REQUIRE(k >= 1 && k <= KMAX_LARGE, "%s: k=%d outside [1, %d]", who, k, KMAX_LARGE);
  MV_REQUIRE(rerank >= 1, "%s: rerank=%d < 1", who, rerank);
  const bool excl = q_self != nullptr || self_base >= 0;
  MV_REQUIRE(k <= n - (excl ? 1 : 0), "%s: k=%d exceeds the %lld selectable rows", who, k,
             (long long)(n - (excl ? 1 : 0)));
  if (dtype != MV_F32 && dtype != MV_BF16 && dtype != MV_FP8 && dtype != MV_BF16X3) {
    mv::set_error("%s: dtype %d not supported", who, dtype);
    return MV_EUNSUPPORTED;
  }
  return MV_OK;
}

int mv_knn_brute(const float* Q, int64_t nq, const float* X, int64_t n, int64_t d, int k, int dtype,
                 int rerank, const int32_t* q_self, int64_t self_base, int euclid, int32_t* out_idx,
                 float* out_dist, int32_t* out_stats, void* ws, size_t ws_bytes, void* stream) {
  if (int rc = knn_check("mv_knn_brute", Q, nq, X, n, d, k, dtype, rerank, q_self, self_base, out_idx, out_dist))
    return rc;
  MV_FAULT_POINT("launch");
  return knn_run("mv_knn_brute", Q, nq, X, n, d, k, dtype, rerank, q_self, self_base, euclid, out_idx, out_dist,
                 out_stats, ws, ws_bytes, stream, g_symmetric, true);
}

int mv_knn_brute_noalloc(const float* Q, int64_t nq, const float* X, int64_t n, int64_t d, int k, int dtype,
                         int rerank, const int32_t* q_self, int64_t self_base, int euclid, int32_t* out_idx,
                         float* out_dist, int32_t* out_stats, void* stream) {
  static const char* who = "mv_knn_brute_noalloc";
  if (int rc = knn_check(who, Q, nq, X, n, d, k, dtype, rerank, q_self, self_base, out_idx, out_dist)) return rc;
  if (k > CMAX_K) {
    mv::set_error("%s: k=%d above %d needs the symmetric screen's n x n matrix; use mv_knn_brute", who, k, CMAX_K);
    return MV_EUNSUPPORTED;
  }
  const Layout L = make_layout(nq, n, d, k, rerank, (Q == X) && nq == n, dtype, false);
  if (L.bytes > NOALLOC_SCRATCH_BYTES) {
    mv::set_error("%s: nq=%lld n=%lld d=%lld k=%d needs %zu bytes of scratch, more than the %zu-byte static "
                  "scratch of this entry point; use mv_knn_brute with a workspace",
                  who, (long long)nq, (long long)n, (long long)d, k, L.bytes, NOALLOC_SCRATCH_BYTES);
    return MV_EUNSUPPORTED;
  }
  MV_FAULT_POINT("launch");
  void* ws = nullptr;
  if (int rc = noalloc_scratch(&ws)) return rc;
  return knn_run(who, Q, nq, X, n, d, k, dtype, rerank, q_self, self_base, euclid, out_idx, out_dist, out_stats,
                 ws, NOALLOC_SCRATCH_BYTES, stream, false, false);
}

int mv_knn_set_symmetric(int on) {
  const int prev = g_symmetric ? 1 : 0;
  g_symmetric = on != 0;
  return prev;
}

int mv_knn_set_timing(int on) {
  g_timer.on = on != 0;
  g_timer.valid = false;
  return MV_OK;
}

int mv_knn_last_phase_ms(float* out4) {
  MV_REQUIRE(out4, "mv_knn_last_phase_ms: null output");
  if (!g_timer.valid) {
    mv::set_error("mv_knn_last_phase_ms: no timed mv_knn_brute call (enable with mv_knn_set_timing)");
    return MV_EINVAL;
  }
  MV_TRY_HIP(hipEventSynchronize(g_timer.ev[4]));
  for (int i = 0; i < 4; ++i) MV_TRY_HIP(hipEventElapsedTime(&out4[i], g_timer.ev[i], g_timer.ev[i + 1]));
  return MV_OK;
}

int mv_knn_last_screen_ms(float* out2, int32_t* out_sym) {
  MV_REQUIRE(out2 && out_sym, "mv_knn_last_screen_ms: null output");
  if (!g_timer.valid) {
    mv::set_error("mv_knn_last_screen_ms: no timed mv_knn_brute call (enable with mv_knn_set_timing)");
    return MV_EINVAL;
  }
  MV_TRY_HIP(hipEventSynchronize(g_timer.ev[4]));
  *out_sym = g_timer.last_sym ? 1 : 0;
  if (g_timer.last_sym) {
    MV_TRY_HIP(hipEventElapsedTime(&out2[0], g_timer.ev[1], g_timer.ev[5]));
    MV_TRY_HIP(hipEventElapsedTime(&out2[1], g_timer.ev[5], g_timer.ev[2]));
  } else {
    MV_TRY_HIP(hipEventElapsedTime(&out2[0], g_timer.ev[1], g_timer.ev[2]));
    out2[1] = 0.f;
  }
  return MV_OK;
}

}  // extern "C"
