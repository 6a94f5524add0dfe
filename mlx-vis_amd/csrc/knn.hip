// Exact brute-force kNN for gfx950 (MI355X).
//
// Replaces the brute path of compute_knn — mlx_vis/_knn.py:47-88. The
// reference forms the full (chunk x n) matrix max(|x_i|^2 + |x_j|^2 - 2 x_i.x_j, 0)
// (_knn.py:60-63), masks the diagonal with +1e30 (_knn.py:64-69), argsorts every
// row and keeps k columns (_knn.py:71-72), then takes sqrt (_knn.py:86-87).
// Here nothing of size n x n is ever materialised:
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
//   3. rerank      one wavefront per query row: merge the S lists into the global
//                  top-C, FP32 direct-difference distances of the C candidates,
//                  FP64 (sequential, unfused) only where two FP32 values cannot be
//                  separated by their error bound, and a certificate that no row
//                  outside the candidates can enter the top k.
//   4. exact scan  rows without a certificate are recomputed exhaustively in
//                  FP64 (rare; guarantees the canonical result).
#include <math.h>

#include "common.hpp"

namespace mv {
namespace knn {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int BM = 128;               // query rows per workgroup
constexpr int BN = 128;               // X rows (columns of D) per tile
constexpr int KC = 32;                // K chunk (floats) staged per step
constexpr int LDSA = KC + 4;          // padded LDS row stride (floats): conflict-free b128 reads
constexpr int DTS = BN + 4;           // padded stride of the D tile in LDS
constexpr int CMAX = 64;              // max candidates per row (one per lane)
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

// part[c, t] = sum over rows of chunk c of X[r, t] (FP64).
__global__ __launch_bounds__(256) void k_colsum_partial(const float* __restrict__ X, int64_t n,
                                                        int64_t d, double* __restrict__ part) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  const int64_t c = blockIdx.y;
  if (t >= d) return;
  const int64_t r0 = c * ROW_CHUNK;
  const int64_t r1 = r0 + ROW_CHUNK < n ? r0 + ROW_CHUNK : n;
  double s = 0.0;
  for (int64_t r = r0; r < r1; ++r) s += static_cast<double>(X[r * d + t]);
  part[c * d + t] = s;
}

__global__ __launch_bounds__(256) void k_colsum_final(const double* __restrict__ part, int64_t nchunks,
                                                      int64_t n, int64_t d, int64_t d_pad,
                                                      float* __restrict__ mu) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (t >= d_pad) return;
  if (t >= d) {
    mu[t] = 0.f;
    return;
  }
  double s = 0.0;
  for (int64_t c = 0; c < nchunks; ++c) s += part[c * d + t];
  mu[t] = static_cast<float>(s / static_cast<double>(n));
}

// Centred, zero-padded FP32 copy: one wavefront per row. Rows in [n, n_pad)
// become zero rows. sqmax holds the bit pattern of the largest row norm.
__global__ __launch_bounds__(256) void k_center_f32(const float* __restrict__ X, int64_t n, int64_t d,
                                                    int64_t d_pad, int64_t n_pad,
                                                    const float* __restrict__ mu, float* __restrict__ Xc,
                                                    float* __restrict__ sq, uint32_t* __restrict__ sqmax) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (r >= n_pad) return;
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
  if (l == 0) {
    sq[r] = acc;
    if (r < n && sqmax) atomicMax(sqmax, __float_as_uint(acc));
  }
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

// ----------------------------------------------------------------- screen --

struct ScreenArgs {
  const float* Qc;        // [nq_pad, d_pad] centred queries
  const float* Xc;        // [n_pad, d_pad] centred data
  const float* sqq;       // [nq_pad]
  const float* sqx;       // [n_pad]
  const int32_t* q_self;  // nullable
  int64_t self_base;      // used when q_self == NULL; < 0: no exclusion
  int64_t d_pad;
  int nq, n, C, S;
  int tiles_per_split, n_col_tiles, n_row_blocks, nq_pad;
  uint64_t* cand;         // [S, nq_pad, C] keys
};

// Inserts key kv into the sorted wave-held list (lane l holds entry l, lanes
// >= C hold KEY_NONE) and refreshes the list's C-th key `tau`.
__device__ __forceinline__ void list_insert(uint64_t& lk, uint64_t kv, uint64_t& tau, int C, int l) {
  if (kv >= tau) return;
  const int pos = __popcll(ballot(lk < kv));
  const uint64_t up = shfl_up_u64(lk, 1);
  if (l > pos) lk = up;
  else if (l == pos) lk = kv;
  if (l >= C) lk = KEY_NONE;
  tau = readlane_u64(lk, C - 1);
}

__global__ __launch_bounds__(NT, 1) void k_screen_f32(ScreenArgs a) {
  __shared__ __attribute__((aligned(16))) float s_stage[STAGE_FLOATS];
  __shared__ uint64_t s_list[BM * CMAX];
  __shared__ uint64_t s_tau[BM];
  __shared__ float s_sqq[BM];
  __shared__ int s_self[BM];

  const int bid = blockIdx.x;
  const int xcd = bid & 7;
  const int per = 8 / a.S;
  const int split = xcd % a.S;
  const int rb = (bid >> 3) * per + xcd / a.S;
  if (rb >= a.n_row_blocks) return;

  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int wr = w >> 1, wc = w & 1;
  const int C = a.C;
  const int64_t d_pad = a.d_pad;

  for (int r = tid; r < BM; r += NT) {
    const int64_t grow = static_cast<int64_t>(rb) * BM + r;
    s_sqq[r] = a.sqq[grow];
    int64_t self = -1;
    if (grow < a.nq) {
      if (a.q_self) self = a.q_self[grow];
      else if (a.self_base >= 0) self = a.self_base + grow;
    }
    s_self[r] = static_cast<int>(self);
    s_tau[r] = KEY_NONE;
  }
  for (int e = tid; e < BM * CMAX; e += NT) s_list[e] = KEY_NONE;
  __syncthreads();

  const int t0 = split * a.tiles_per_split;
  const int t1 = min(t0 + a.tiles_per_split, a.n_col_tiles);
  const float* Ab = a.Qc + static_cast<int64_t>(rb) * BM * d_pad;
  const int lr = tid >> 3;            // staging row (plus 32*i)
  const int lkk = (tid & 7) * 4;      // staging k offset within the chunk
  const int nk = static_cast<int>(d_pad / KC);

  for (int t = t0; t < t1; ++t) {
    const int col0 = t * BN;
    const float* Bb = a.Xc + static_cast<int64_t>(col0) * d_pad;

    v4f acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = v4f{0.f, 0.f, 0.f, 0.f};

    float4 ra[4], rv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ra[i] = *reinterpret_cast<const float4*>(Ab + static_cast<int64_t>(lr + 32 * i) * d_pad + lkk);
      rv[i] = *reinterpret_cast<const float4*>(Bb + static_cast<int64_t>(lr + 32 * i) * d_pad + lkk);
    }
    {
      float* As = s_stage;
      float* Bs = s_stage + BM * LDSA;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        *reinterpret_cast<float4*>(As + (lr + 32 * i) * LDSA + lkk) = ra[i];
        *reinterpret_cast<float4*>(Bs + (lr + 32 * i) * LDSA + lkk) = rv[i];
      }
    }
    __syncthreads();

    for (int kc = 0; kc < nk; ++kc) {
      const int cur = kc & 1;
      const bool more = kc + 1 < nk;
      if (more) {
        const int64_t ko = static_cast<int64_t>(kc + 1) * KC + lkk;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          ra[i] = *reinterpret_cast<const float4*>(Ab + static_cast<int64_t>(lr + 32 * i) * d_pad + ko);
          rv[i] = *reinterpret_cast<const float4*>(Bb + static_cast<int64_t>(lr + 32 * i) * d_pad + ko);
        }
      }
      const float* As = s_stage + cur * (2 * BM * LDSA);
      const float* Bs = As + BM * LDSA;
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        v4f af[4], bf[4];
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
          af[mi] = *reinterpret_cast<const v4f*>(As + (wr * 64 + mi * 16 + (l & 15)) * LDSA + g * 16 +
                                                 (l >> 4) * 4);
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
          bf[ni] = *reinterpret_cast<const v4f*>(Bs + (wc * 64 + ni * 16 + (l & 15)) * LDSA + g * 16 +
                                                 (l >> 4) * 4);
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int mi = 0; mi < 4; ++mi)
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
              acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[mi][j], bf[ni][j], acc[mi][ni], 0, 0, 0);
      }
      if (more) {
        float* As2 = s_stage + (cur ^ 1) * (2 * BM * LDSA);
        float* Bs2 = As2 + BM * LDSA;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          *reinterpret_cast<float4*>(As2 + (lr + 32 * i) * LDSA + lkk) = ra[i];
          *reinterpret_cast<float4*>(Bs2 + (lr + 32 * i) * LDSA + lkk) = rv[i];
        }
      }
      __syncthreads();
    }

    // Epilogue: D = |q|^2 + |x|^2 - 2 q.x, clamped at 0; excluded columns -> +inf.
    float* Dt = s_stage;
    float sqxv[4];
    int colv[4];
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      colv[ni] = col0 + wc * 64 + ni * 16 + (l & 15);
      sqxv[ni] = a.sqx[colv[ni]];
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wr * 64 + mi * 16 + (l >> 4) * 4 + r;
        const float sr = s_sqq[row];
        const int self = s_self[row];
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
          float v = fmaxf(sr + sqxv[ni] - 2.f * acc[mi][ni][r], 0.f);
          if (colv[ni] >= a.n || colv[ni] == self) v = INFINITY;
          Dt[row * DTS + wc * 64 + ni * 16 + (l & 15)] = v;
        }
      }
    }
    __syncthreads();

    for (int rr = 0; rr < BM / 4; ++rr) {
      const int row = w * (BM / 4) + rr;
      uint64_t tau = s_tau[row];
      const float v0 = Dt[row * DTS + l];
      const float v1 = Dt[row * DTS + 64 + l];
      const uint64_t k0 = v0 == INFINITY ? KEY_NONE : dkey(v0, static_cast<uint32_t>(col0 + l));
      const uint64_t k1 = v1 == INFINITY ? KEY_NONE : dkey(v1, static_cast<uint32_t>(col0 + 64 + l));
      uint64_t m0 = ballot(k0 < tau);
      uint64_t m1 = ballot(k1 < tau);
      if ((m0 | m1) == 0) continue;
      uint64_t lk = l < C ? s_list[row * CMAX + l] : KEY_NONE;
      while (m0) {
        const int src = __ffsll(static_cast<long long>(m0)) - 1;
        m0 &= m0 - 1;
        list_insert(lk, readlane_u64(k0, src), tau, C, l);
      }
      while (m1) {
        const int src = __ffsll(static_cast<long long>(m1)) - 1;
        m1 &= m1 - 1;
        list_insert(lk, readlane_u64(k1, src), tau, C, l);
      }
      if (l < C) s_list[row * CMAX + l] = lk;
      if (l == 0) s_tau[row] = tau;
    }
    __syncthreads();
  }

  for (int rr = 0; rr < BM / 4; ++rr) {
    const int row = w * (BM / 4) + rr;
    const int64_t grow = static_cast<int64_t>(rb) * BM + row;
    if (grow < a.nq && l < C)
      a.cand[(static_cast<int64_t>(split) * a.nq_pad + grow) * C + l] = s_list[row * CMAX + l];
  }
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
  float err_fac;               // screen error bound: err_fac * (sqq[i] + sqmax)
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
  for (int64_t t = 0; t < d; ++t) {
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
  if (i >= a.nq) return;
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

  // 2. FP32 direct-difference distances: 4 candidates at a time, 16 lanes each.
  const float* qrow = a.Q + static_cast<int64_t>(i) * a.d;
  const int g = l >> 4, sub = l & 15;
  float d32 = INFINITY;
  for (int r = 0; r * 4 < nvalid; ++r) {
    const int c = r * 4 + g;
    const int j = __shfl(cidx, c < 64 ? c : 0);
    float acc = 0.f;
    if (c < nvalid) {
      const float* xr = a.X + static_cast<int64_t>(j) * a.d;
      for (int64_t t = sub; t < a.d; t += 16) {
        const float df = qrow[t] - xr[t];
        acc = fmaf(df, df, acc);
      }
    }
    acc += __shfl_xor(acc, 8);
    acc += __shfl_xor(acc, 4);
    acc += __shfl_xor(acc, 2);
    acc += __shfl_xor(acc, 1);
    const float got = __shfl(acc, (l & 3) * 16);
    if ((l >> 2) == r) d32 = got;
  }

  // 3. Sort by (FP32 distance, index); flag pairs FP32 cannot separate.
  uint64_t k32 = l < nvalid ? dkey(d32, static_cast<uint32_t>(cidx)) : KEY_NONE;
  k32 = wave_sort_u64(k32);
  const float bfac = static_cast<float>(a.d + 8) * 5.9604645e-8f * 1.001f;
  const float ds = key_dist(k32);
  const float bd = ds * bfac + 1e-35f;
  const float dn = __shfl_down(ds, 1);
  const float bn = __shfl_down(bd, 1);
  const bool amb = (l < k) && (l + 1 < nvalid) && (dn - bn <= ds + bd);
  const bool any_amb = ballot(amb) != 0;

  uint32_t out_i = key_idx(k32);
  float out_d = ds;
  float kth_ub;
  if (any_amb) {
    // 4. FP64 tie check: canonical distances for every candidate, then sort.
    double d64 = INFINITY;
    if (l < nvalid) d64 = dist64_seq(qrow, a.X + static_cast<int64_t>(out_i) * a.d, a.d);
    uint64_t dbits = l < nvalid ? static_cast<uint64_t>(__double_as_longlong(d64)) : KEY_NONE;
    uint32_t di = l < nvalid ? out_i : 0xFFFFFFFFu;
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

  // 5. Certificate: every non-candidate j has screen value >= tau, hence true
  //    distance >= tau - err_i; the k-th candidate must lie strictly below.
  bool cert = nvalid < C;
  if (!cert) {
    const float tau = key_dist(tau_key);
    const float sqmax = __uint_as_float(*a.sqmax_bits);
    const float err_i = a.err_fac * (a.sqq[i] + sqmax) + 1e-30f;
    cert = kth_ub < tau - err_i;
  }
  if (cert) {
    if (l < k) {
      a.out_idx[static_cast<int64_t>(i) * k + l] = static_cast<int32_t>(out_i);
      a.out_dist[static_cast<int64_t>(i) * k + l] = out_d;
    }
  } else if (l == 0) {
    const int p = atomicAdd(a.fb_count, 1);
    a.fb_rows[p] = i;
    atomicAdd(&a.stats[2], 1);
  }
}

// ------------------------------------------------------------- exact scan --

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

// Exhaustive canonical scan of one query row per workgroup iteration: every
// lane computes FP64 sequential distances for its columns; each wavefront
// keeps a sorted top-k list; the four lists are merged at the end.
__global__ __launch_bounds__(256) void k_exact_scan(ScanArgs a) {
  __shared__ uint64_t s_d[4 * 64];
  __shared__ uint32_t s_i[4 * 64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int count = *a.fb_count;
  const int k = a.k;
  for (int f = blockIdx.x; f < count; f += gridDim.x) {
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
  int64_t nq_pad, n_pad, d_pad, nchunks;
  int C;
  bool same;
  size_t bytes;
  // offsets are produced by carving in a fixed order (see carve()).
};

struct Buffers {
  double* part;
  float* mu;
  float* Xc;
  float* sqx;
  float* Qc;
  float* sqq;
  uint64_t* cand;
  int32_t* ctr;     // [0] fallback count, [1] sqmax bits, [2..5] stats
  int32_t* fb_rows;
};

static Buffers carve(void* ws, int64_t nq, int64_t n, int64_t d, const Layout& L, size_t* total) {
  Carve c(ws);
  Buffers b;
  b.part = c.take<double>(static_cast<size_t>(L.nchunks) * d);
  b.mu = c.take<float>(L.d_pad);
  b.Xc = c.take<float>(static_cast<size_t>(L.n_pad) * L.d_pad);
  b.sqx = c.take<float>(L.n_pad);
  if (L.same) {
    b.Qc = b.Xc;
    b.sqq = b.sqx;
  } else {
    b.Qc = c.take<float>(static_cast<size_t>(L.nq_pad) * L.d_pad);
    b.sqq = c.take<float>(L.nq_pad);
  }
  b.cand = c.take<uint64_t>(static_cast<size_t>(MAX_SPLIT) * L.nq_pad * L.C);
  b.ctr = c.take<int32_t>(8);
  b.fb_rows = c.take<int32_t>(nq > 0 ? nq : 1);
  *total = c.off + 256;
  return b;
}

static Layout make_layout(int64_t nq, int64_t n, int64_t d, int k, int rerank, bool same) {
  Layout L;
  L.nq_pad = static_cast<int64_t>(align_up(nq > 0 ? nq : 1, BM));
  L.n_pad = static_cast<int64_t>(align_up(n, BN));
  L.d_pad = static_cast<int64_t>(align_up(d, KC));
  L.nchunks = ceil_div(n, ROW_CHUNK);
  int C = rerank * k;
  if (C < k) C = k;
  if (C > CMAX) C = CMAX;
  L.C = C;
  L.same = same;
  size_t total = 0;
  carve(nullptr, nq, n, d, L, &total);
  L.bytes = total;
  return L;
}

// Optional phase timing (mv_knn_set_timing): HIP events recorded on the launch
// stream around the prepass, screen, rerank and exact-scan launches.
struct PhaseTimer {
  bool on = false;
  bool valid = false;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
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
// the per-split merge work.
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
  MV_REQUIRE(k >= 1 && k <= 32, "mv_knn_ws_bytes: k=%d outside [1, 32]", k);
  MV_REQUIRE(rerank >= 1, "mv_knn_ws_bytes: rerank=%d < 1", rerank);
  if (dtype != MV_F32) {
    mv::set_error("mv_knn_ws_bytes: dtype %d not supported", dtype);
    return MV_EUNSUPPORTED;
  }
  Layout L = make_layout(nq, n, d, k, rerank, same_qx != 0);
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

int mv_knn_brute(const float* Q, int64_t nq, const float* X, int64_t n, int64_t d, int k, int dtype,
                 int rerank, const int32_t* q_self, int64_t self_base, int euclid, int32_t* out_idx,
                 float* out_dist, int32_t* out_stats, void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(Q && X && out_idx && out_dist, "mv_knn_brute: null pointer argument");
  MV_REQUIRE(nq >= 0 && n >= 1 && d >= 1, "mv_knn_brute: bad shape nq=%lld n=%lld d=%lld", (long long)nq,
             (long long)n, (long long)d);
  MV_REQUIRE(n < (1ll << 31) - 256 && nq < (1ll << 31) - 256, "mv_knn_brute: n or nq too large");
  MV_REQUIRE(k >= 1 && k <= 32, "mv_knn_brute: k=%d outside [1, 32]", k);
  MV_REQUIRE(rerank >= 1, "mv_knn_brute: rerank=%d < 1", rerank);
  const bool excl = q_self != nullptr || self_base >= 0;
  MV_REQUIRE(k <= n - (excl ? 1 : 0), "mv_knn_brute: k=%d exceeds the %lld selectable rows", k,
             (long long)(n - (excl ? 1 : 0)));
  if (dtype != MV_F32) {
    mv::set_error("mv_knn_brute: dtype %d not supported", dtype);
    return MV_EUNSUPPORTED;
  }
  MV_FAULT_POINT("launch");
  const bool same = (Q == X) && nq == n;
  Layout L = make_layout(nq, n, d, k, rerank, same);
  if (ws_bytes < L.bytes || !ws) {
    mv::set_error("mv_knn_brute: workspace of %zu bytes < required %zu", ws_bytes, L.bytes);
    return MV_EWORKSPACE;
  }
  if (nq == 0) return MV_OK;
  size_t total = 0;
  Buffers B = carve(ws, nq, n, d, L, &total);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int32_t* fb_count = B.ctr;
  uint32_t* sqmax = reinterpret_cast<uint32_t*>(B.ctr + 1);
  int32_t* stats = out_stats ? out_stats : B.ctr + 2;

  if (g_timer.on) {
    if (int rc = g_timer.ensure()) return rc;
    g_timer.valid = false;
  }
  if (int rc = g_timer.mark(0, s)) return rc;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, B.ctr, 8);
  MV_LAUNCHED("k_zero_i32");
  if (out_stats) {
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, out_stats, 4);
    MV_LAUNCHED("k_zero_i32");
  }
  hipLaunchKernelGGL(k_colsum_partial, dim3(static_cast<unsigned>(ceil_div(d, 256)),
                                            static_cast<unsigned>(L.nchunks)),
                     dim3(256), 0, s, X, n, d, B.part);
  MV_LAUNCHED("k_colsum_partial");
  hipLaunchKernelGGL(k_colsum_final, dim3(static_cast<unsigned>(ceil_div(L.d_pad, 256))), dim3(256), 0, s,
                     B.part, L.nchunks, n, d, L.d_pad, B.mu);
  MV_LAUNCHED("k_colsum_final");
  hipLaunchKernelGGL(k_center_f32, dim3(static_cast<unsigned>(L.n_pad / 4)), dim3(256), 0, s, X, n, d,
                     L.d_pad, L.n_pad, B.mu, B.Xc, B.sqx, sqmax);
  MV_LAUNCHED("k_center_f32");
  if (!same) {
    hipLaunchKernelGGL(k_center_f32, dim3(static_cast<unsigned>(L.nq_pad / 4)), dim3(256), 0, s, Q, nq, d,
                       L.d_pad, L.nq_pad, B.mu, B.Qc, B.sqq, static_cast<uint32_t*>(nullptr));
    MV_LAUNCHED("k_center_f32");
  }

  const int64_t row_blocks = L.nq_pad / BM;
  const int64_t col_tiles = L.n_pad / BN;
  const int S = choose_split(row_blocks, col_tiles, L.C, cu_count_cached());
  ScreenArgs sa;
  sa.Qc = B.Qc;
  sa.Xc = B.Xc;
  sa.sqq = B.sqq;
  sa.sqx = B.sqx;
  sa.q_self = q_self;
  sa.self_base = q_self ? -1 : self_base;
  sa.d_pad = L.d_pad;
  sa.nq = static_cast<int>(nq);
  sa.n = static_cast<int>(n);
  sa.C = L.C;
  sa.S = S;
  sa.tiles_per_split = static_cast<int>(ceil_div(col_tiles, S));
  sa.n_col_tiles = static_cast<int>(col_tiles);
  sa.n_row_blocks = static_cast<int>(row_blocks);
  sa.nq_pad = static_cast<int>(L.nq_pad);
  sa.cand = B.cand;
  const int per = 8 / S;
  const int64_t grid = ceil_div(row_blocks, per) * 8;
  if (int rc = g_timer.mark(1, s)) return rc;
  hipLaunchKernelGGL(k_screen_f32, dim3(static_cast<unsigned>(grid)), dim3(NT), 0, s, sa);
  MV_LAUNCHED("k_screen_f32");
  if (int rc = g_timer.mark(2, s)) return rc;

  RerankArgs ra;
  ra.Q = Q;
  ra.X = X;
  ra.d = d;
  ra.nq = static_cast<int>(nq);
  ra.n = static_cast<int>(n);
  ra.k = k;
  ra.C = L.C;
  ra.S = S;
  ra.nq_pad = static_cast<int>(L.nq_pad);
  ra.cand = B.cand;
  ra.sqq = B.sqq;
  ra.sqmax_bits = sqmax;
  ra.err_fac = static_cast<float>((2.0 * static_cast<double>(L.d_pad) + 10.0) * 5.9604644775390625e-8 * 1.01);
  ra.euclid = euclid;
  ra.out_idx = out_idx;
  ra.out_dist = out_dist;
  ra.fb_count = fb_count;
  ra.fb_rows = B.fb_rows;
  ra.stats = stats;
  hipLaunchKernelGGL(k_rerank, dim3(static_cast<unsigned>(ceil_div(nq, 4))), dim3(256), 0, s, ra);
  MV_LAUNCHED("k_rerank");
  if (int rc = g_timer.mark(3, s)) return rc;

  ScanArgs xa;
  xa.Q = Q;
  xa.X = X;
  xa.d = d;
  xa.n = static_cast<int>(n);
  xa.k = k;
  xa.q_self = q_self;
  xa.self_base = q_self ? -1 : self_base;
  xa.fb_count = fb_count;
  xa.fb_rows = B.fb_rows;
  xa.euclid = euclid;
  xa.out_idx = out_idx;
  xa.out_dist = out_dist;
  xa.stats = stats;
  const int64_t sgrid = nq < 2048 ? nq : 2048;
  hipLaunchKernelGGL(k_exact_scan, dim3(static_cast<unsigned>(sgrid)), dim3(256), 0, s, xa);
  MV_LAUNCHED("k_exact_scan");
  if (int rc = g_timer.mark(4, s)) return rc;
  g_timer.valid = g_timer.on;
  return MV_OK;
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

}  // extern "C"
