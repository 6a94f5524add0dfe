// NNDescent k-NN graph for gfx950.
//
// Replaces NNDescent.build — mlx_vis/_nndescent/nndescent.py:34-403 — and its
// helpers _compute_reverse_candidates (406-436), _dedup_topk(_block) (439-486),
// _gather_dists (573-612) and _gather_dists_half (615-641). The iteration
// schedule (nndescent.py:94-398) runs natively in mv_nnd_build; the host reads
// one integer per iteration (the number of changed slots), as the reference
// does at nndescent.py:384.
//
// Kernels:
//   k_nnd_update   one wavefront per row: builds the row's candidate slots
//                  (current k, nn-of-nn mc_new x j_nn, reverse k), computes
//                  the missing distances with a coalesced gather-dot (four
//                  candidates at a time, 16 lanes x 16-byte loads each, the
//                  source row in LDS), masks self, resolves duplicate ids the
//                  way the reference's sort-by-id pass does (the earliest slot
//                  keeps its distance, later copies get 1e30), and extracts
//                  the k smallest (distance, id) keys. Writes the new row,
//                  positional NEW flags and the changed count.
//   reverse        counting sort of the n*k edges by destination (atomic
//                  counts, exclusive scan, atomic fill) and a per-destination
//                  selection of the k smallest flat edge positions — the
//                  first k sources in the reference's argsort order; empty
//                  slots hold node 0 as in the reference (nndescent.py:431-434).
//   projections    FP32 tiled GEMM for X @ P (nndescent.py:75-76, 231).
//   half copy      FP16 copy for the FP16 iterations; the dot is accumulated
//                  in FP32 and rounded to FP16 like an FP16 matmul output
//                  (nndescent.py:628).
#include <hip/hip_fp16.h>
#include <math.h>
#include <stdio.h>

#include <algorithm>

#include "common.hpp"

namespace mv {
namespace nnd {

constexpr uint64_t KEY_NONE = ~0ull;
constexpr float BIG = 1e30f;
constexpr int MAX_LD = 4096;   // largest padded row (elements) kept in LDS

// ------------------------------------------------------------------ misc --

__global__ void k_zero_u32(uint32_t* p, int64_t count) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) p[t] = 0;
}

// Zero-padded FP32 copy (ld % 4 == 0) and FP32 norms, one wavefront per row.
__global__ __launch_bounds__(256) void k_pad_f32(const float* __restrict__ X, int64_t n, int64_t d, int64_t ld,
                                                 float* __restrict__ out, float* __restrict__ sq) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (r >= n) return;
  float acc = 0.f;
  for (int64_t t = l; t < ld; t += 64) {
    const float v = t < d ? X[r * d + t] : 0.f;
    out[r * ld + t] = v;
    acc = fmaf(v, v, acc);
  }
  acc = wave_sum_f32(acc);
  if (l == 0 && sq) sq[r] = acc;
}

// FP16 copy of a padded FP32 matrix (ldh % 8 == 0, zero padding).
__global__ __launch_bounds__(256) void k_to_half(const float* __restrict__ X, int64_t n, int64_t ld, int64_t ldh,
                                                 __half* __restrict__ out) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (r >= n) return;
  for (int64_t t = l; t < ldh; t += 64) out[r * ldh + t] = __float2half_rn(t < ld ? X[r * ld + t] : 0.f);
}

// C[n, ldc] = A[n, lda] @ B[d, p] (FP32, 64x64 tiles, 4x4 per thread); the
// columns [p, ldc) are written as zeros.
__global__ __launch_bounds__(256) void k_gemm_f32(const float* __restrict__ A, int64_t n, int64_t d, int64_t lda,
                                                  const float* __restrict__ B, int64_t p, float* __restrict__ C,
                                                  int64_t ldc) {
  __shared__ float As[16][64 + 4];
  __shared__ float Bs[16][64 + 4];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int64_t r0 = static_cast<int64_t>(blockIdx.y) * 64;
  const int64_t c0 = static_cast<int64_t>(blockIdx.x) * 64;
  float acc[4][4] = {};
  for (int64_t k0 = 0; k0 < d; k0 += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int kk = e & 15, rr = e >> 4;
      const int64_t gr = r0 + rr, gk = k0 + kk;
      As[kk][rr] = (gr < n && gk < d) ? A[gr * lda + gk] : 0.f;
      const int cc = e & 63, kb = e >> 6;
      const int64_t gc = c0 + cc, gk2 = k0 + kb;
      Bs[kb][cc] = (gc < p && gk2 < d) ? B[gk2 * p + gc] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t gr = r0 + ty * 4 + i;
    if (gr >= n) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t gc = c0 + tx * 4 + j;
      if (gc < ldc) C[gr * ldc + gc] = gc < p ? acc[i][j] : 0.f;
    }
  }
}

__global__ __launch_bounds__(256) void k_sqrt(const float* __restrict__ in, float* __restrict__ out, int64_t count) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) out[t] = sqrtf(fmaxf(in[t], 0.f));
}

// ----------------------------------------------------- distance source --

struct Src {
  const float* xf;   // FP32 rows (ld % 4 == 0), or nullptr
  const __half* xh;  // FP16 rows (ld % 8 == 0), or nullptr
  const float* sq;   // FP32 norms
  int64_t ld;        // row stride (elements)
};

// Stages row i of the source into this wavefront's LDS slice.
__device__ __forceinline__ void stage_row(const Src& s, int64_t i, float* lds_f, __half* lds_h, int l) {
  if (s.xh) {
    const uint4* src = reinterpret_cast<const uint4*>(s.xh + i * s.ld);
    uint4* dst = reinterpret_cast<uint4*>(lds_h);
    for (int64_t t = l; t < s.ld / 8; t += 64) dst[t] = src[t];
  } else {
    const float4* src = reinterpret_cast<const float4*>(s.xf + i * s.ld);
    float4* dst = reinterpret_cast<float4*>(lds_f);
    for (int64_t t = l; t < s.ld / 4; t += 64) dst[t] = src[t];
  }
}

// Dot of the staged row with row j by a 16-lane group (sub = lane % 16); the
// result is reduced over the group in a fixed order, so a pair's value does
// not depend on which slot or row computes it.
// Dot product of the staged row (LDS) with row j, over the 16 lanes of a
// group (lane `sub` takes the 16-byte pieces sub, sub + 16, ...). The pieces
// of row j are fetched GD_BATCH at a time with no branch around the loads
// (an index past the row reads its last piece and meets a zero query piece),
// so that a batch is in flight together instead of one memory round trip per
// piece; the FMAs run in the same order as piece by piece (adding 0 * b
// changes nothing), so the dot is the same to the bit.
constexpr int GD_BATCH = 4;
__device__ __forceinline__ float group_dot(const Src& s, int64_t j, const float* lds_f, const __half* lds_h,
                                           int sub, bool live) {
  float acc = 0.f;
  if (live) {
    if (s.xh) {
      const uint4* xr = reinterpret_cast<const uint4*>(s.xh + j * s.ld);
      const uint4* qr = reinterpret_cast<const uint4*>(lds_h);
      const int64_t nch = s.ld / 8;
      for (int64_t t0 = sub; t0 < nch; t0 += 16 * GD_BATCH) {
        uint4 bb[GD_BATCH];
#pragma unroll
        for (int u = 0; u < GD_BATCH; ++u) {
          const int64_t t = t0 + 16 * u;
          bb[u] = xr[t < nch ? t : nch - 1];
        }
#pragma unroll
        for (int u = 0; u < GD_BATCH; ++u) {
          const int64_t t = t0 + 16 * u;
          const uint4 a = t < nch ? qr[t] : make_uint4(0u, 0u, 0u, 0u);
          const uint32_t av[4] = {a.x, a.y, a.z, a.w};
          const uint32_t bv[4] = {bb[u].x, bb[u].y, bb[u].z, bb[u].w};
#pragma unroll
          for (int h = 0; h < 4; ++h) {
            const __half2 ha = *reinterpret_cast<const __half2*>(&av[h]);
            const __half2 hb = *reinterpret_cast<const __half2*>(&bv[h]);
            const float2 fa = __half22float2(ha), fb = __half22float2(hb);
            acc = fmaf(fa.x, fb.x, acc);
            acc = fmaf(fa.y, fb.y, acc);
          }
        }
      }
    } else {
      const float4* xr = reinterpret_cast<const float4*>(s.xf + j * s.ld);
      const float4* qr = reinterpret_cast<const float4*>(lds_f);
      const int64_t nch = s.ld / 4;
      for (int64_t t0 = sub; t0 < nch; t0 += 16 * GD_BATCH) {
        float4 bb[GD_BATCH];
#pragma unroll
        for (int u = 0; u < GD_BATCH; ++u) {
          const int64_t t = t0 + 16 * u;
          bb[u] = xr[t < nch ? t : nch - 1];
        }
#pragma unroll
        for (int u = 0; u < GD_BATCH; ++u) {
          const int64_t t = t0 + 16 * u;
          const float4 a = t < nch ? qr[t] : make_float4(0.f, 0.f, 0.f, 0.f);
          acc = fmaf(a.x, bb[u].x, acc);
          acc = fmaf(a.y, bb[u].y, acc);
          acc = fmaf(a.z, bb[u].z, acc);
          acc = fmaf(a.w, bb[u].w, acc);
        }
      }
    }
  }
  acc += __shfl_xor(acc, 8);
  acc += __shfl_xor(acc, 4);
  acc += __shfl_xor(acc, 2);
  acc += __shfl_xor(acc, 1);
  return acc;
}

__device__ __forceinline__ float pair_dist(const Src& s, float sqi, int64_t j, float dot) {
  if (s.xh) dot = __half2float(__float2half_rn(dot));   // FP16 matmul output
  return fmaxf(sqi + s.sq[j] - 2.0f * dot, 0.f);
}

// --------------------------------------------------------------- update --

constexpr int NND_NQ_MAX = 16;   // candidate slots per lane of the largest update variant

struct UpdateArgs {
  Src src;
  const int32_t* idx;       // current graph (n, k)
  const float* dist;        // current distances (n, k)
  const uint8_t* flags;     // current NEW flags (n, k)
  const int32_t* rev;       // reverse candidates (n, k) or nullptr
  int32_t* out_idx;
  float* out_dist;
  uint8_t* out_flags;
  uint32_t* changed;
  int64_t n;
  int k, mc_new, j_nn;
  int converging;           // 1: NEW-first sources, only rows with a NEW flag
  int recompute_cur;        // distances of the current k recomputed with src
};

// NQ candidate slots per lane: 64 NQ per row (slot s = 64 q + l).
template <int NQ>
__global__ __launch_bounds__(256) void k_nnd_update(UpdateArgs a) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (i >= a.n) return;
  const int k = a.k;
  const int64_t bytes_per_wave = a.src.xh ? a.src.ld * 2 : a.src.ld * 4;
  float* lds_f = reinterpret_cast<float*>(smem + w * bytes_per_wave);
  __half* lds_h = reinterpret_cast<__half*>(smem + w * bytes_per_wave);

  const int32_t cur_id = l < k ? a.idx[i * k + l] : -1;
  const float cur_d = l < k ? a.dist[i * k + l] : 0.f;
  const bool cur_new = l < k ? a.flags[i * k + l] != 0 : false;

  if (a.converging) {
    if (ballot(cur_new) == 0) {
      if (l < k) {
        a.out_idx[i * k + l] = cur_id;
        a.out_dist[i * k + l] = cur_d;
        a.out_flags[i * k + l] = 0;
      }
      return;
    }
  }
  // Source order of the nn-of-nn expansion: positional (full path) or NEW
  // entries first, each group in positional order (nndescent.py:105-106).
  int32_t first = cur_id;
  if (a.converging) {
    const uint64_t nm = ballot(cur_new && l < k);
    const uint64_t below = l == 0 ? 0ull : (~0ull >> (64 - l));
    const int n_new = __popcll(nm);
    int pos;
    if (l >= k) pos = l;
    else if (cur_new) pos = __popcll(nm & below);
    else pos = n_new + (l - __popcll(nm & below));
    first = __builtin_amdgcn_ds_permute(pos * 4, cur_id);
  }

  const int n_nn = a.mc_new * a.j_nn;
  const int n_rev = a.rev ? k : 0;
  const int total = k + n_nn + n_rev;   // <= 64 NQ (checked on the host)

  // Slot ids (slot s = 64 q + l). The source-row shuffle runs on every lane.
  int32_t sid[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int s = 64 * q + l;
    const int rel = s - k;
    const int m = (rel >= 0 && rel < n_nn) ? rel / a.j_nn : 0;
    const int t = (rel >= 0 && rel < n_nn) ? rel % a.j_nn : 0;
    const int32_t srow = __shfl(first, m);
    int32_t id = -1;
    if (s < k) id = cur_id;
    else if (s < k + n_nn) id = a.idx[static_cast<int64_t>(srow) * k + t];
    else if (s < total) id = a.rev[i * k + (s - k - n_nn)];
    sid[q] = id;
  }

  stage_row(a.src, i, lds_f, lds_h, l);
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const float sqi = a.src.sq[i];

  // Distances: four slots per round, one per 16-lane group; register q holds
  // the slots of rounds 16 q .. 16 q + 15 (every register index is static:
  // a dynamically indexed register array would live in scratch memory).
  float sd[NQ];
  const int g = l >> 4, sub = l & 15;
  const int rounds = (total + 3) / 4;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    sd[q] = 0.f;
    for (int rr = 0; rr < 16 && 16 * q + rr < rounds; ++rr) {
      const int r = 16 * q + rr;
      const int s = 4 * r + g;
      const int32_t id = __shfl(sid[q], s & 63);
      const bool stored = s < k && !a.recompute_cur;
      const bool live = s < total && id >= 0 && id != i && !stored;
      const float dot = group_dot(a.src, id, lds_f, lds_h, sub, live);
      float v = 0.f;
      if (live) v = pair_dist(a.src, sqi, id, dot);
      const float got = __shfl(v, (l & 3) * 16);
      if ((l >> 2) == rr) sd[q] = got;
    }
  }

  // Keys: (distance, id); self and padding get 1e30 / none; stored current
  // distances are used as is.
  uint64_t key[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int s = 64 * q + l;
    float dv = sd[q];
    if (s < k && !a.recompute_cur) dv = cur_d;
    if (sid[q] == static_cast<int32_t>(i)) dv = BIG;
    key[q] = (s < total && sid[q] >= 0) ? dkey(dv, static_cast<uint32_t>(sid[q])) : KEY_NONE;
  }

  // Extract the k smallest keys; a duplicated id keeps the distance of its
  // earliest slot, later copies become 1e30 (nndescent.py:441-449).
  uint64_t outk = KEY_NONE;
  int produced = 0;
  for (int guard = 0; guard < NQ * 64 && produced < k; ++guard) {
    uint64_t lm = key[0];
#pragma unroll
    for (int q = 1; q < NQ; ++q) lm = key[q] < lm ? key[q] : lm;
    const uint64_t m = wave_min_u64(lm);
    if (m == KEY_NONE) break;
    const int32_t mid = static_cast<int32_t>(key_idx(m));
    int first_slot = 1 << 20;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      if (key[q] != KEY_NONE && sid[q] == mid) first_slot = min(first_slot, 64 * q + l);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) first_slot = min(first_slot, __shfl_xor(first_slot, off));
    const int fl = first_slot & 63;
    uint64_t fkv = KEY_NONE;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      if (64 * q + l == first_slot) fkv = key[q];
    const uint64_t fk = readlane_u64(fkv, fl);
    // later copies of this id -> 1e30
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      if (key[q] != KEY_NONE && sid[q] == mid && (64 * q + l) != first_slot)
        key[q] = dkey(BIG, static_cast<uint32_t>(mid));
    if (fk == m) {
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (64 * q + l == first_slot) key[q] = KEY_NONE;
      if (l == produced) outk = m;
      ++produced;
    }
  }
  const bool have = l < k && outk != KEY_NONE;
  const int32_t nid = have ? static_cast<int32_t>(key_idx(outk)) : cur_id;
  const float nd = have ? key_dist(outk) : cur_d;
  const bool flag = l < k && nid != cur_id;
  if (l < k) {
    a.out_idx[i * k + l] = nid;
    a.out_dist[i * k + l] = nd;
    a.out_flags[i * k + l] = flag ? 1 : 0;
  }
  const uint64_t fm = ballot(flag);
  if (l == 0 && fm) atomicAdd(a.changed, static_cast<uint32_t>(__popcll(fm)));
}

// Rows with more candidate slots than one pass holds (k + mc k + k > 1024,
// from k = 32 at the default max_candidates): the slots are taken in chunks
// of 1024 (NQ = 16 per lane) and each chunk is merged into the row's running
// top-k list, smallest key first, one copy of each key. Copies of an id carry
// the same key, since a pair's distance does not depend on the slot that
// computes it, except a current neighbour kept with its stored distance
// (recompute_cur == 0): a new candidate equal to a current neighbour is then
// skipped, which is the reference's rule that the earliest slot of an id keeps
// its distance and later copies drop out (nndescent.py:441-449). The self
// slot drops out as well; a row with fewer than k distinct candidates (only
// possible for tiny n) keeps its remaining current entries.
__global__ __launch_bounds__(256) void k_nnd_update_chunked(UpdateArgs a) {
  constexpr int NQ = NND_NQ_MAX;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (i >= a.n) return;
  const int k = a.k;
  const int64_t bytes_per_wave = a.src.xh ? a.src.ld * 2 : a.src.ld * 4;
  float* lds_f = reinterpret_cast<float*>(smem + w * bytes_per_wave);
  __half* lds_h = reinterpret_cast<__half*>(smem + w * bytes_per_wave);

  const int32_t cur_id = l < k ? a.idx[i * k + l] : -1;
  const float cur_d = l < k ? a.dist[i * k + l] : 0.f;
  const bool cur_new = l < k ? a.flags[i * k + l] != 0 : false;
  if (a.converging && ballot(cur_new) == 0) {
    if (l < k) {
      a.out_idx[i * k + l] = cur_id;
      a.out_dist[i * k + l] = cur_d;
      a.out_flags[i * k + l] = 0;
    }
    return;
  }
  int32_t first = cur_id;
  if (a.converging) {
    const uint64_t nm = ballot(cur_new && l < k);
    const uint64_t below = l == 0 ? 0ull : (~0ull >> (64 - l));
    const int n_new = __popcll(nm);
    int pos;
    if (l >= k) pos = l;
    else if (cur_new) pos = __popcll(nm & below);
    else pos = n_new + (l - __popcll(nm & below));
    first = __builtin_amdgcn_ds_permute(pos * 4, cur_id);
  }
  const int n_nn = a.mc_new * a.j_nn;
  const int n_rev = a.rev ? k : 0;
  const int total = k + n_nn + n_rev;

  stage_row(a.src, i, lds_f, lds_h, l);
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const float sqi = a.src.sq[i];
  const int g = l >> 4, sub = l & 15;

  uint64_t lk = KEY_NONE;   // running list: lane l holds the l-th smallest key (l < k)
  for (int base = 0; base < total; base += 64 * NQ) {
    int32_t sid[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int s = base + 64 * q + l;
      const int rel = s - k;
      const int m = (rel >= 0 && rel < n_nn) ? rel / a.j_nn : 0;
      const int t = (rel >= 0 && rel < n_nn) ? rel % a.j_nn : 0;
      const int32_t srow = __shfl(first, m);
      int32_t id = -1;
      if (s < k) id = cur_id;
      else if (s < k + n_nn) id = a.idx[static_cast<int64_t>(srow) * k + t];
      else if (s < total) id = a.rev[i * k + (s - k - n_nn)];
      sid[q] = id;
    }
    // a new candidate equal to a current neighbour kept with its stored distance
    if (!a.recompute_cur) {
      for (int u = 0; u < k; ++u) {
        const int32_t cu = __shfl(cur_id, u);
#pragma unroll
        for (int q = 0; q < NQ; ++q)
          if (base + 64 * q + l >= k && sid[q] == cu) sid[q] = -1;
      }
    }
    float sd[NQ];
    const int chunk = total - base < 64 * NQ ? total - base : 64 * NQ;
    const int rounds = (chunk + 3) / 4;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      sd[q] = 0.f;
      for (int rr = 0; rr < 16 && 16 * q + rr < rounds; ++rr) {
        const int r = 16 * q + rr;
        const int sl = 4 * r + g;   // slot within the chunk
        const int32_t id = __shfl(sid[q], sl & 63);
        const bool stored = base + sl < k && !a.recompute_cur;
        const bool live = sl < chunk && id >= 0 && id != i && !stored;
        const float dot = group_dot(a.src, id, lds_f, lds_h, sub, live);
        float v = 0.f;
        if (live) v = pair_dist(a.src, sqi, id, dot);
        const float got = __shfl(v, (l & 3) * 16);
        if ((l >> 2) == rr) sd[q] = got;
      }
    }
    uint64_t key[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int s = base + 64 * q + l;
      float dv = sd[q];
      if (s < k && !a.recompute_cur) dv = cur_d;
      const bool ok = s < total && sid[q] >= 0 && sid[q] != static_cast<int32_t>(i);
      key[q] = ok ? dkey(dv, static_cast<uint32_t>(sid[q])) : KEY_NONE;
    }
    // merge: the k smallest distinct keys of the list and the chunk
    uint64_t outk = KEY_NONE;
    for (int produced = 0; produced < k; ++produced) {
      uint64_t lm = lk;
#pragma unroll
      for (int q = 0; q < NQ; ++q) lm = key[q] < lm ? key[q] : lm;
      const uint64_t m = wave_min_u64(lm);
      if (m == KEY_NONE) break;
      lk = lk == m ? KEY_NONE : lk;
#pragma unroll
      for (int q = 0; q < NQ; ++q) key[q] = key[q] == m ? KEY_NONE : key[q];
      if (l == produced) outk = m;
    }
    lk = outk;
  }
  const bool have = l < k && lk != KEY_NONE;
  const int32_t nid = have ? static_cast<int32_t>(key_idx(lk)) : cur_id;
  const float nd = have ? key_dist(lk) : cur_d;
  const bool flag = l < k && nid != cur_id;
  if (l < k) {
    a.out_idx[i * k + l] = nid;
    a.out_dist[i * k + l] = nd;
    a.out_flags[i * k + l] = flag ? 1 : 0;
  }
  const uint64_t fm = ballot(flag);
  if (l == 0 && fm) atomicAdd(a.changed, static_cast<uint32_t>(__popcll(fm)));
}

// k from 65 to NND_K_WIDE: lane l holds current entries l and 64 + l, and the
// running top-k list the same way (two keys per lane). Candidate slots are
// taken in chunks of 1024 and merged as in k_nnd_update_chunked (same slot
// order, same duplicate and self rules). The source rows of the nn-of-nn
// expansion (positional, or NEW first in the converging path) are put in
// order through LDS, since mc_new may exceed 64.
constexpr int NND_K_WIDE = 128;
__global__ __launch_bounds__(256) void k_nnd_update_wide(UpdateArgs a) {
  constexpr int NQ = NND_NQ_MAX;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  __shared__ int32_t s_first[4][NND_K_WIDE];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (i >= a.n) return;
  const int k = a.k;
  const int64_t bytes_per_wave = a.src.xh ? a.src.ld * 2 : a.src.ld * 4;
  float* lds_f = reinterpret_cast<float*>(smem + w * bytes_per_wave);
  __half* lds_h = reinterpret_cast<__half*>(smem + w * bytes_per_wave);
  int32_t* first = s_first[w];

  int32_t cur_id[2];
  float cur_d[2];
  bool cur_new[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int e = 64 * j + l;
    cur_id[j] = e < k ? a.idx[i * k + e] : -1;
    cur_d[j] = e < k ? a.dist[i * k + e] : 0.f;
    cur_new[j] = e < k ? a.flags[i * k + e] != 0 : false;
  }
  const uint64_t nm0 = ballot(cur_new[0]), nm1 = ballot(cur_new[1]);
  if (a.converging && (nm0 | nm1) == 0) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int e = 64 * j + l;
      if (e < k) {
        a.out_idx[i * k + e] = cur_id[j];
        a.out_dist[i * k + e] = cur_d[j];
        a.out_flags[i * k + e] = 0;
      }
    }
    return;
  }
  {
    const uint64_t below = l == 0 ? 0ull : (~0ull >> (64 - l));
    const int n_new = __popcll(nm0) + __popcll(nm1);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int e = 64 * j + l;
      if (e < k) {
        int pos = e;
        if (a.converging) {
          const int new_before = j == 0 ? __popcll(nm0 & below) : __popcll(nm0) + __popcll(nm1 & below);
          pos = cur_new[j] ? new_before : n_new + (e - new_before);
        }
        first[pos] = cur_id[j];
      }
    }
  }
  const int n_nn = a.mc_new * a.j_nn;
  const int n_rev = a.rev ? k : 0;
  const int total = k + n_nn + n_rev;

  stage_row(a.src, i, lds_f, lds_h, l);
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const float sqi = a.src.sq[i];
  const int g = l >> 4, sub = l & 15;

  uint64_t lk0 = KEY_NONE, lk1 = KEY_NONE;   // running list: lane l holds ranks l and 64 + l
  for (int base = 0; base < total; base += 64 * NQ) {
    int32_t sid[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int s = base + 64 * q + l;
      const int rel = s - k;
      const int m = (rel >= 0 && rel < n_nn) ? rel / a.j_nn : 0;
      const int t = (rel >= 0 && rel < n_nn) ? rel % a.j_nn : 0;
      int32_t id = -1;
      if (s < k) id = q == 0 ? cur_id[0] : cur_id[1];   // s < k <= 128: base 0, q 0 or 1
      else if (s < k + n_nn) id = a.idx[static_cast<int64_t>(first[m]) * k + t];
      else if (s < total) id = a.rev[i * k + (s - k - n_nn)];
      sid[q] = id;
    }
    if (!a.recompute_cur) {
      for (int u = 0; u < k; ++u) {
        const int32_t cu = u < 64 ? __shfl(cur_id[0], u) : __shfl(cur_id[1], u - 64);
#pragma unroll
        for (int q = 0; q < NQ; ++q)
          if (base + 64 * q + l >= k && sid[q] == cu) sid[q] = -1;
      }
    }
    float sd[NQ];
    const int chunk = total - base < 64 * NQ ? total - base : 64 * NQ;
    const int rounds = (chunk + 3) / 4;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      sd[q] = 0.f;
      for (int rr = 0; rr < 16 && 16 * q + rr < rounds; ++rr) {
        const int r = 16 * q + rr;
        const int sl = 4 * r + g;
        const int32_t id = __shfl(sid[q], sl & 63);
        const bool stored = base + sl < k && !a.recompute_cur;
        const bool live = sl < chunk && id >= 0 && id != i && !stored;
        const float dot = group_dot(a.src, id, lds_f, lds_h, sub, live);
        float v = 0.f;
        if (live) v = pair_dist(a.src, sqi, id, dot);
        const float got = __shfl(v, (l & 3) * 16);
        if ((l >> 2) == rr) sd[q] = got;
      }
    }
    uint64_t key[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int s = base + 64 * q + l;
      float dv = sd[q];
      if (s < k && !a.recompute_cur) dv = q == 0 ? cur_d[0] : cur_d[1];
      const bool ok = s < total && sid[q] >= 0 && sid[q] != static_cast<int32_t>(i);
      key[q] = ok ? dkey(dv, static_cast<uint32_t>(sid[q])) : KEY_NONE;
    }
    uint64_t o0 = KEY_NONE, o1 = KEY_NONE;
    for (int produced = 0; produced < k; ++produced) {
      uint64_t lm = lk0 < lk1 ? lk0 : lk1;
#pragma unroll
      for (int q = 0; q < NQ; ++q) lm = key[q] < lm ? key[q] : lm;
      const uint64_t m = wave_min_u64(lm);
      if (m == KEY_NONE) break;
      lk0 = lk0 == m ? KEY_NONE : lk0;
      lk1 = lk1 == m ? KEY_NONE : lk1;
#pragma unroll
      for (int q = 0; q < NQ; ++q) key[q] = key[q] == m ? KEY_NONE : key[q];
      if (l == produced) o0 = m;
      if (l + 64 == produced) o1 = m;
    }
    lk0 = o0;
    lk1 = o1;
  }
  int nflag = 0;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int e = 64 * j + l;
    const uint64_t lk = j == 0 ? lk0 : lk1;
    const bool have = e < k && lk != KEY_NONE;
    const int32_t nid = have ? static_cast<int32_t>(key_idx(lk)) : cur_id[j];
    const float nd = have ? key_dist(lk) : cur_d[j];
    const bool flag = e < k && nid != cur_id[j];
    if (e < k) {
      a.out_idx[i * k + e] = nid;
      a.out_dist[i * k + e] = nd;
      a.out_flags[i * k + e] = flag ? 1 : 0;
    }
    nflag += __popcll(ballot(flag));
  }
  if (l == 0 && nflag) atomicAdd(a.changed, static_cast<uint32_t>(nflag));
}

// Distances of (row, idx[row, c]) for all c with src; FP32 (or FP16) like
// _gather_dists(_half). Optionally sorts each row by (distance, position)
// (the init sort, nndescent.py:62-65).
__global__ __launch_bounds__(256) void k_nnd_gather(Src src, const int32_t* __restrict__ idx, int64_t n, int c,
                                                    int sort_rows, int32_t* __restrict__ out_idx,
                                                    float* __restrict__ out_dist) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (i >= n) return;
  const int64_t bpw = src.xh ? src.ld * 2 : src.ld * 4;
  float* lds_f = reinterpret_cast<float*>(smem + w * bpw);
  __half* lds_h = reinterpret_cast<__half*>(smem + w * bpw);
  const int32_t my = l < c ? idx[i * c + l] : -1;
  stage_row(src, i, lds_f, lds_h, l);
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const float sqi = src.sq[i];
  float dv = 0.f;
  const int g = l >> 4, sub = l & 15;
  for (int r = 0; 4 * r < c; ++r) {
    const int s = 4 * r + g;
    const int32_t id = __shfl(my, s & 63);
    const bool live = s < c && id >= 0;
    const float dot = group_dot(src, id, lds_f, lds_h, sub, live);
    float v = 0.f;
    if (live) v = pair_dist(src, sqi, id, dot);
    const float got = __shfl(v, (l & 3) * 16);
    if ((l >> 2) == r) dv = got;
  }
  if (sort_rows) {
    uint64_t key = l < c ? dkey(dv, static_cast<uint32_t>(l)) : KEY_NONE;
    key = wave_sort_u64(key);
    const int from = static_cast<int>(key_idx(key)) & 63;
    const int32_t sid = __shfl(my, from);
    if (l < c) {
      out_idx[i * c + l] = sid;
      out_dist[i * c + l] = key_dist(key);
    }
  } else if (l < c) {
    out_dist[i * c + l] = dv;
    if (out_idx) out_idx[i * c + l] = my;
  }
}

// k_nnd_gather for c from 65 to NND_K_WIDE: lane l holds positions l and
// 64 + l. The sort places each entry at its rank among the row's
// (distance, position) keys, which are distinct.
__global__ __launch_bounds__(256) void k_nnd_gather_wide(Src src, const int32_t* __restrict__ idx, int64_t n,
                                                         int c, int sort_rows, int32_t* __restrict__ out_idx,
                                                         float* __restrict__ out_dist) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (i >= n) return;
  const int64_t bpw = src.xh ? src.ld * 2 : src.ld * 4;
  float* lds_f = reinterpret_cast<float*>(smem + w * bpw);
  __half* lds_h = reinterpret_cast<__half*>(smem + w * bpw);
  const int32_t my0 = l < c ? idx[i * c + l] : -1;
  const int32_t my1 = 64 + l < c ? idx[i * c + 64 + l] : -1;
  stage_row(src, i, lds_f, lds_h, l);
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const float sqi = src.sq[i];
  float dv0 = 0.f, dv1 = 0.f;
  const int g = l >> 4, sub = l & 15;
  for (int r = 0; 4 * r < c; ++r) {
    const int s = 4 * r + g;   // rounds 0..15 hold positions 0..63, rounds 16..31 positions 64..127
    const int32_t id = r < 16 ? __shfl(my0, s & 63) : __shfl(my1, s & 63);
    const bool live = s < c && id >= 0;
    const float dot = group_dot(src, id, lds_f, lds_h, sub, live);
    float v = 0.f;
    if (live) v = pair_dist(src, sqi, id, dot);
    const float got = __shfl(v, (l & 3) * 16);
    if ((l >> 2) == (r & 15)) {
      if (r < 16) dv0 = got;
      else dv1 = got;
    }
  }
  if (sort_rows) {
    const uint64_t k0 = l < c ? dkey(dv0, static_cast<uint32_t>(l)) : KEY_NONE;
    const uint64_t k1 = 64 + l < c ? dkey(dv1, static_cast<uint32_t>(64 + l)) : KEY_NONE;
    int r0 = 0, r1 = 0;
    for (int f = 0; f < c; ++f) {
      const uint64_t kf = f < 64 ? readlane_u64(k0, f) : readlane_u64(k1, f - 64);
      r0 += kf < k0 ? 1 : 0;
      r1 += kf < k1 ? 1 : 0;
    }
    if (l < c) {
      out_idx[i * c + r0] = my0;
      out_dist[i * c + r0] = key_dist(k0);
    }
    if (64 + l < c) {
      out_idx[i * c + r1] = my1;
      out_dist[i * c + r1] = key_dist(k1);
    }
  } else {
    if (l < c) {
      out_dist[i * c + l] = dv0;
      if (out_idx) out_idx[i * c + l] = my0;
    }
    if (64 + l < c) {
      out_dist[i * c + 64 + l] = dv1;
      if (out_idx) out_idx[i * c + 64 + l] = my1;
    }
  }
}

// ------------------------------------------------------------- reverse --

__global__ __launch_bounds__(256) void k_rev_count(const int32_t* __restrict__ idx, int64_t m,
                                                   uint32_t* __restrict__ cnt) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f < m) atomicAdd(&cnt[idx[f]], 1u);
}

// Exclusive scan, level 1: 1024-element blocks (4 per thread).
__global__ __launch_bounds__(256) void k_scan_block(const uint32_t* __restrict__ in, int64_t count,
                                                    uint32_t* __restrict__ out, uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t s[256];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * 1024 + threadIdx.x * 4;
  uint32_t v[4];
  uint32_t tot = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    v[j] = base + j < count ? in[base + j] : 0u;
    tot += v[j];
  }
  s[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint32_t add = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - tot;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (base + j < count) out[base + j] = run;
    run += v[j];
  }
  if (threadIdx.x == 255) block_sums[blockIdx.x] = s[255];
}

// Level 2: exclusive scan of the block sums by one block (serial per thread chunk).
__global__ __launch_bounds__(256) void k_scan_sums(uint32_t* __restrict__ sums, int64_t nb) {
  __shared__ uint32_t s[256];
  const int64_t per = (nb + 255) / 256;
  const int64_t b0 = threadIdx.x * per;
  uint32_t tot = 0;
  for (int64_t j = 0; j < per; ++j)
    if (b0 + j < nb) tot += sums[b0 + j];
  s[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint32_t add = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - tot;
  for (int64_t j = 0; j < per; ++j)
    if (b0 + j < nb) {
      const uint32_t v = sums[b0 + j];
      sums[b0 + j] = run;
      run += v;
    }
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* __restrict__ out, int64_t count,
                                                  const uint32_t* __restrict__ sums) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) out[t] += sums[t / 1024];
}

__global__ __launch_bounds__(256) void k_rev_fill(const int32_t* __restrict__ idx, int64_t m,
                                                  const uint32_t* __restrict__ off, uint32_t* __restrict__ cursor,
                                                  uint32_t* __restrict__ bucket) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f < m) {
    const int32_t j = idx[f];
    const uint32_t p = off[j] + atomicAdd(&cursor[j], 1u);
    bucket[p] = static_cast<uint32_t>(f);
  }
}

constexpr uint32_t REV_THREAD_MAX = 64;   // in-degree handled by one thread in k_rev_select

// rev[j, t] = source row of the t-th smallest flat edge position pointing at
// j (t < min(k, count)); remaining slots hold 0.
__global__ __launch_bounds__(256) void k_rev_select(const uint32_t* __restrict__ bucket,
                                                    const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ cnt, int64_t n, int k,
                                                    int32_t* __restrict__ rev) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t c = cnt[j];
  if (c > REV_THREAD_MAX) return;   // hubs: k_rev_select_hub
  const uint32_t* b = bucket + off[j];
  int64_t prev = -1;
  for (int t = 0; t < k; ++t) {
    int64_t best = -1;
    if (static_cast<uint32_t>(t) < c) {
      uint32_t bv = 0xFFFFFFFFu;
      for (uint32_t e = 0; e < c; ++e) {
        const uint32_t v = b[e];
        if (static_cast<int64_t>(v) > prev && v < bv) bv = v;
      }
      best = bv;
      prev = best;
    }
    rev[j * k + t] = best >= 0 ? static_cast<int32_t>(best / k) : 0;
  }
}

// Destinations with more than REV_THREAD_MAX in-edges (hubs): one wavefront
// each, k rounds of a wave-wide minimum over the entries above the previous
// pick. Wavefronts stride over the destinations of a bounded grid.
__global__ __launch_bounds__(256) void k_rev_select_hub(const uint32_t* __restrict__ bucket,
                                                        const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ cnt, int64_t n, int k,
                                                        int32_t* __restrict__ rev) {
  const int l = threadIdx.x & 63;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * 4;
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6); j < n; j += waves) {
    const uint32_t c = cnt[j];
    if (c <= REV_THREAD_MAX) continue;
    const uint32_t* b = bucket + off[j];
    int64_t prev = -1;
    for (int t = 0; t < k; ++t) {
      uint32_t lm = 0xFFFFFFFFu;
      for (uint32_t e = l; e < c; e += 64) {
        const uint32_t v = b[e];
        if (static_cast<int64_t>(v) > prev && v < lm) lm = v;
      }
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
        const uint32_t o = static_cast<uint32_t>(__shfl_xor(static_cast<int>(lm), m));
        lm = o < lm ? o : lm;
      }
      prev = lm;
      // slots past the in-degree hold 0 (a hub may have fewer than k in-edges when k > 64)
      if (l == 0) rev[j * k + t] = static_cast<uint32_t>(t) < c ? static_cast<int32_t>(lm / static_cast<uint32_t>(k)) : 0;
    }
  }
}

// ----------------------------------------------------------------- host --

struct Plan {
  int64_t n, d, k;
  int64_t ld, ldh, ld_lo, ldh_lo, ld_mid, ldh_mid;
  int64_t p1, p2;
  size_t bytes;
  // buffers
  float *Xp, *sq, *Xlo, *sqlo, *Xmid, *sqmid;
  __half *H, *Hlo, *Hmid;
  int32_t *idx0, *idx1, *rev;
  float *dist0, *dist1;
  uint8_t *fl0, *fl1;
  uint32_t *cnt, *off, *cursor, *bucket, *sums, *changed;
};

static void carve_plan(Plan& P, void* ws) {
  Carve c(ws);
  const int64_t n = P.n, k = P.k;
  P.Xp = c.take<float>(n * P.ld);
  P.sq = c.take<float>(n);
  P.H = c.take<__half>(n * P.ldh);
  P.Xlo = P.p1 ? c.take<float>(n * P.ld_lo) : nullptr;
  P.sqlo = P.p1 ? c.take<float>(n) : nullptr;
  P.Hlo = nullptr;
  P.Xmid = P.p2 ? c.take<float>(n * P.ld_mid) : nullptr;
  P.sqmid = P.p2 ? c.take<float>(n) : nullptr;
  P.Hmid = P.p2 ? c.take<__half>(n * P.ldh_mid) : nullptr;
  P.idx0 = c.take<int32_t>(n * k);
  P.idx1 = c.take<int32_t>(n * k);
  P.rev = c.take<int32_t>(n * k);
  P.dist0 = c.take<float>(n * k);
  P.dist1 = c.take<float>(n * k);
  P.fl0 = c.take<uint8_t>(n * k);
  P.fl1 = c.take<uint8_t>(n * k);
  P.cnt = c.take<uint32_t>(n);
  P.off = c.take<uint32_t>(n);
  P.cursor = c.take<uint32_t>(n);
  P.bucket = c.take<uint32_t>(n * k);
  P.sums = c.take<uint32_t>(ceil_div(n, 1024) + 1);
  P.changed = c.take<uint32_t>(4);
  P.bytes = c.off + 256;
}

static Plan make_plan(int64_t n, int64_t d, int k, int64_t p1, int64_t p2) {
  Plan P{};
  P.n = n;
  P.d = d;
  P.k = k;
  P.ld = static_cast<int64_t>(align_up(d, 8));
  P.ldh = P.ld;
  P.p1 = p1;
  P.p2 = p2;
  P.ld_lo = static_cast<int64_t>(align_up(p1 > 0 ? p1 : 1, 8));
  P.ld_mid = static_cast<int64_t>(align_up(p2 > 0 ? p2 : 1, 8));
  P.ldh_mid = P.ld_mid;
  carve_plan(P, nullptr);
  return P;
}

static int launch_update(const Plan& P, const Src& src, const int32_t* idx, const float* dist, const uint8_t* fl,
                         const int32_t* rev, int32_t* oidx, float* odist, uint8_t* ofl, int mc_new, int j_nn,
                         int converging, int recompute_cur, hipStream_t s) {
  UpdateArgs a;
  a.src = src;
  a.idx = idx;
  a.dist = dist;
  a.flags = fl;
  a.rev = rev;
  a.out_idx = oidx;
  a.out_dist = odist;
  a.out_flags = ofl;
  a.changed = P.changed;
  a.n = P.n;
  a.k = static_cast<int>(P.k);
  a.mc_new = mc_new;
  a.j_nn = j_nn;
  a.converging = converging;
  a.recompute_cur = recompute_cur;
  const size_t lds = 4 * static_cast<size_t>(src.xh ? src.ld * 2 : src.ld * 4);
  const int total = a.k + mc_new * j_nn + (rev ? a.k : 0);
  const dim3 grid(static_cast<unsigned>(ceil_div(P.n, 4)));
  if (a.k > 64) hipLaunchKernelGGL(k_nnd_update_wide, grid, dim3(256), lds, s, a);
  else if (total <= 64 * 4) hipLaunchKernelGGL(k_nnd_update<4>, grid, dim3(256), lds, s, a);
  else if (total <= 64 * 8) hipLaunchKernelGGL(k_nnd_update<8>, grid, dim3(256), lds, s, a);
  else if (total <= 64 * NND_NQ_MAX) hipLaunchKernelGGL(k_nnd_update<NND_NQ_MAX>, grid, dim3(256), lds, s, a);
  else hipLaunchKernelGGL(k_nnd_update_chunked, grid, dim3(256), lds, s, a);
  MV_LAUNCHED("k_nnd_update");
  return MV_OK;
}

static int launch_gather(const Plan& P, const Src& src, const int32_t* idx, int sort_rows, int32_t* oidx,
                         float* odist, hipStream_t s) {
  const size_t lds = 4 * static_cast<size_t>(src.xh ? src.ld * 2 : src.ld * 4);
  if (P.k > 64)
    hipLaunchKernelGGL(k_nnd_gather_wide, dim3(static_cast<unsigned>(ceil_div(P.n, 4))), dim3(256), lds, s, src,
                       idx, P.n, static_cast<int>(P.k), sort_rows, oidx, odist);
  else
    hipLaunchKernelGGL(k_nnd_gather, dim3(static_cast<unsigned>(ceil_div(P.n, 4))), dim3(256), lds, s, src, idx,
                       P.n, static_cast<int>(P.k), sort_rows, oidx, odist);
  MV_LAUNCHED("k_nnd_gather");
  return MV_OK;
}

static int reverse_candidates(const Plan& P, const int32_t* idx, int32_t* rev, hipStream_t s) {
  const int64_t n = P.n, m = P.n * P.k;
  const unsigned bn = static_cast<unsigned>(ceil_div(n, 256));
  const unsigned bm = static_cast<unsigned>(ceil_div(m, 256));
  hipLaunchKernelGGL(k_zero_u32, dim3(bn), dim3(256), 0, s, P.cnt, n);
  MV_LAUNCHED("k_zero_u32");
  hipLaunchKernelGGL(k_zero_u32, dim3(bn), dim3(256), 0, s, P.cursor, n);
  MV_LAUNCHED("k_zero_u32");
  hipLaunchKernelGGL(k_rev_count, dim3(bm), dim3(256), 0, s, idx, m, P.cnt);
  MV_LAUNCHED("k_rev_count");
  const int64_t nb = ceil_div(n, 1024);
  hipLaunchKernelGGL(k_scan_block, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, P.cnt, n, P.off, P.sums);
  MV_LAUNCHED("k_scan_block");
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, P.sums, nb);
  MV_LAUNCHED("k_scan_sums");
  hipLaunchKernelGGL(k_scan_add, dim3(bn), dim3(256), 0, s, P.off, n, P.sums);
  MV_LAUNCHED("k_scan_add");
  hipLaunchKernelGGL(k_rev_fill, dim3(bm), dim3(256), 0, s, idx, m, P.off, P.cursor, P.bucket);
  MV_LAUNCHED("k_rev_fill");
  hipLaunchKernelGGL(k_rev_select, dim3(bn), dim3(256), 0, s, P.bucket, P.off, P.cnt, n, static_cast<int>(P.k),
                     rev);
  MV_LAUNCHED("k_rev_select");
  const int64_t hgrid = ceil_div(n, 4) < 2048 ? ceil_div(n, 4) : 2048;
  hipLaunchKernelGGL(k_rev_select_hub, dim3(static_cast<unsigned>(hgrid)), dim3(256), 0, s, P.bucket, P.off, P.cnt, n,
                     static_cast<int>(P.k), rev);
  MV_LAUNCHED("k_rev_select_hub");
  return MV_OK;
}

static int project(const Plan& P, const float* Pm, int64_t p, float* out, int64_t ldo, float* sq_out,
                   hipStream_t s) {
  dim3 grid(static_cast<unsigned>(ceil_div(ldo, 64)), static_cast<unsigned>(ceil_div(P.n, 64)));
  hipLaunchKernelGGL(k_gemm_f32, grid, dim3(256), 0, s, P.Xp, P.n, P.d, P.ld, Pm, p, out, ldo);
  MV_LAUNCHED("k_gemm_f32");
  hipLaunchKernelGGL(k_pad_f32, dim3(static_cast<unsigned>(ceil_div(P.n, 4))), dim3(256), 0, s, out, P.n, ldo, ldo,
                     out, sq_out);
  MV_LAUNCHED("k_pad_f32");
  return MV_OK;
}

}  // namespace nnd
}  // namespace mv

using namespace mv;
using namespace mv::nnd;

extern "C" {

int mv_nnd_ws_bytes(int64_t n, int64_t d, int k, int64_t p1, int64_t p2, size_t* out_bytes) {
  MV_REQUIRE(out_bytes, "mv_nnd_ws_bytes: out_bytes is NULL");
  MV_REQUIRE(n >= 2 && d >= 1 && k >= 1 && k <= NND_K_WIDE && k < n, "mv_nnd_ws_bytes: bad n=%lld d=%lld k=%d",
             (long long)n, (long long)d, k);
  MV_REQUIRE(p1 >= 0 && p2 >= 0, "mv_nnd_ws_bytes: bad projection sizes");
  Plan P = make_plan(n, d, k, p1, p2);
  *out_bytes = P.bytes;
  return MV_OK;
}

int mv_nnd_reverse(const int32_t* idx, int64_t n, int k, int32_t* rev, void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(idx && rev, "mv_nnd_reverse: null pointer");
  MV_REQUIRE(n >= 2 && k >= 1 && k <= NND_K_WIDE && k < n, "mv_nnd_reverse: bad n=%lld k=%d", (long long)n, k);
  MV_REQUIRE(n * k < (1ll << 31), "mv_nnd_reverse: n * k must stay below 2^31");
  Plan P = make_plan(n, 1, k, 0, 0);
  if (!ws || ws_bytes < P.bytes) {
    mv::set_error("mv_nnd_reverse: workspace of %zu bytes < required %zu", ws_bytes, P.bytes);
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  carve_plan(P, ws);
  return reverse_candidates(P, idx, rev, static_cast<hipStream_t>(stream));
}

int mv_nnd_build(const float* X, int64_t n, int64_t d, int k, int n_iters, int max_candidates, float delta,
                 const int32_t* init_idx, const float* P1, int64_t p1, const float* P2, int64_t p2, int verbose,
                 int32_t* out_idx, float* out_dist, int64_t* out_changed, int* out_iters, void* ws,
                 size_t ws_bytes, void* stream) {
  MV_REQUIRE(X && init_idx && out_idx && out_dist, "mv_nnd_build: null pointer");
  MV_REQUIRE(n >= 2 && d >= 1 && k >= 1 && k <= NND_K_WIDE && k < n, "mv_nnd_build: bad n=%lld d=%lld k=%d",
             (long long)n, (long long)d, k);
  MV_REQUIRE(n * k < (1ll << 31), "mv_nnd_build: n * k must stay below 2^31");
  MV_REQUIRE(n_iters >= 0, "mv_nnd_build: n_iters < 0");
  MV_REQUIRE((p1 == 0) == (P1 == nullptr) && (p2 == 0) == (P2 == nullptr), "mv_nnd_build: projection mismatch");
  MV_REQUIRE(p1 < d && p2 < d, "mv_nnd_build: projections must reduce the dimension");
  const int mc = max_candidates > 0 ? max_candidates : k;
  if (align_up(d, 8) > static_cast<size_t>(MAX_LD)) {
    mv::set_error("mv_nnd_build: d=%lld exceeds the %d-element LDS row", (long long)d, MAX_LD);
    return MV_EUNSUPPORTED;
  }
  if (k > 64 && align_up(d, 8) > static_cast<size_t>(MAX_LD - NND_K_WIDE)) {
    // k_nnd_update_wide keeps 4 x NND_K_WIDE source ids in LDS beside the four staged rows
    mv::set_error("mv_nnd_build: d=%lld exceeds the %d-element LDS row of k > 64", (long long)d,
                  MAX_LD - NND_K_WIDE);
    return MV_EUNSUPPORTED;
  }
  Plan P = make_plan(n, d, k, p1, p2);
  if (!ws || ws_bytes < P.bytes) {
    mv::set_error("mv_nnd_build: workspace of %zu bytes < required %zu", ws_bytes, P.bytes);
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  carve_plan(P, ws);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned rows4 = static_cast<unsigned>(ceil_div(n, 4));

  // Full-precision source (padded copy) and its norms (nndescent.py:51).
  hipLaunchKernelGGL(k_pad_f32, dim3(rows4), dim3(256), 0, s, X, n, d, P.ld, P.Xp, P.sq);
  MV_LAUNCHED("k_pad_f32");
  const Src full{P.Xp, nullptr, P.sq, P.ld};

  // Random init sorted by distance (nndescent.py:53-66).
  MV_TRY_HIP(hipMemcpyAsync(P.idx1, init_idx, sizeof(int32_t) * n * k, hipMemcpyDeviceToDevice, s));
  if (int rc = launch_gather(P, full, P.idx1, 1, P.idx0, P.dist0, s)) return rc;
  hipLaunchKernelGGL(k_zero_u32, dim3(static_cast<unsigned>(ceil_div(n * k, 256))), dim3(256), 0, s,
                     reinterpret_cast<uint32_t*>(P.fl0), ceil_div(n * k, 4));
  MV_LAUNCHED("k_zero_u32");

  // Projections (nndescent.py:72-82).
  if (p1) {
    if (int rc = project(P, P1, p1, P.Xlo, P.ld_lo, P.sqlo, s)) return rc;
  }
  const Src lowsrc{p1 ? P.Xlo : P.Xp, nullptr, p1 ? P.sqlo : P.sq, p1 ? P.ld_lo : P.ld};
  bool half_full_ready = false;

  int32_t *ci = P.idx0, *ni = P.idx1;
  float *cd = P.dist0, *nd = P.dist1;
  uint8_t *cf = P.fl0, *nf = P.fl1;
  double uf = 1.0;
  int it_done = 0;
  for (int it = 0; it < n_iters; ++it) {
    const int kk = k;
    const double sc = fmin(1.0, sqrt(uf) * 3.0);
    const int mc_nn = std::max(3, static_cast<int>(std::min(kk, mc) * sc));
    const int j_nn = std::max(kk / 2, std::min(kk, static_cast<int>(kk * sc)));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, P.changed, 1);
    MV_LAUNCHED("k_zero_u32");
    if (uf < 0.5) {
      // Converging path (nndescent.py:103-220): NEW-first sources, active rows
      // only, FP32 distances in the full space against stored current ones.
      const int mc_new = std::max(3, static_cast<int>(kk * uf));
      const bool use_rev = uf >= 0.10;
      if (use_rev)
        if (int rc = reverse_candidates(P, ci, P.rev, s)) return rc;
      const int jn = std::min(j_nn, kk);
      if (int rc = launch_update(P, full, ci, cd, cf, use_rev ? P.rev : nullptr, ni, nd, nf, std::min(mc_new, kk),
                                 jn, 1, 0, s))
        return rc;
    } else {
      // Full path (nndescent.py:221-378).
      const int mc_new = std::min(mc_nn, kk);
      Src src = full;
      const bool use_half = it >= 1 && it <= 4;
      if (it == 0) {
        src = lowsrc;
      } else if (it == 1 && p2) {
        if (int rc = project(P, P2, p2, P.Xmid, P.ld_mid, P.sqmid, s)) return rc;
        hipLaunchKernelGGL(k_to_half, dim3(rows4), dim3(256), 0, s, P.Xmid, n, P.ld_mid, P.ldh_mid, P.Hmid);
        MV_LAUNCHED("k_to_half");
        src = Src{nullptr, P.Hmid, P.sqmid, P.ldh_mid};
      } else if (use_half) {
        if (!half_full_ready) {
          hipLaunchKernelGGL(k_to_half, dim3(rows4), dim3(256), 0, s, P.Xp, n, P.ld, P.ldh, P.H);
          MV_LAUNCHED("k_to_half");
          half_full_ready = true;
        }
        src = Src{nullptr, P.H, P.sq, P.ldh};
      }
      const bool recompute = it == 0 || use_half;
      const bool use_rev = uf >= 0.10;
      if (use_rev)
        if (int rc = reverse_candidates(P, ci, P.rev, s)) return rc;
      if (int rc = launch_update(P, src, ci, cd, cf, use_rev ? P.rev : nullptr, ni, nd, nf, mc_new, j_nn, 0,
                                 recompute ? 1 : 0, s))
        return rc;
      if (it == 4) {
        // FP32 full-space distances after the last FP16 iteration (nndescent.py:376-378).
        if (int rc = launch_gather(P, full, ni, 0, nullptr, nd, s)) return rc;
      }
    }
    uint32_t changed = 0;
    MV_TRY_HIP(hipMemcpyAsync(&changed, P.changed, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MV_TRY_HIP(hipStreamSynchronize(s));
    uf = static_cast<double>(changed) / static_cast<double>(n * k);
    if (out_changed) out_changed[it] = changed;
    std::swap(ci, ni);
    std::swap(cd, nd);
    std::swap(cf, nf);
    it_done = it + 1;
    if (verbose) fprintf(stderr, "Iter %d/%d: %u updates (%.4f)\n", it + 1, n_iters, changed, uf);
    if (uf < delta) {
      if (verbose) fprintf(stderr, "Converged at iteration %d\n", it + 1);
      break;
    }
  }
  if (out_iters) *out_iters = it_done;
  MV_TRY_HIP(hipMemcpyAsync(out_idx, ci, sizeof(int32_t) * n * k, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_sqrt, dim3(static_cast<unsigned>(ceil_div(n * k, 256))), dim3(256), 0, s, cd, out_dist,
                     n * k);
  MV_LAUNCHED("k_sqrt");
  return MV_OK;
}

int mv_nnd_gather_dists(const float* X, int64_t n, int64_t d, const int32_t* idx, int k, int half, float* out_dist,
                        void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(X && idx && out_dist, "mv_nnd_gather_dists: null pointer");
  MV_REQUIRE(n >= 2 && d >= 1 && k >= 1 && k <= NND_K_WIDE && k < n, "mv_nnd_gather_dists: bad sizes");
  if (align_up(d, 8) > static_cast<size_t>(MAX_LD)) {
    mv::set_error("mv_nnd_gather_dists: d too large");
    return MV_EUNSUPPORTED;
  }
  Plan P = make_plan(n, d, k, 0, 0);
  if (!ws || ws_bytes < P.bytes) {
    mv::set_error("mv_nnd_gather_dists: workspace of %zu bytes < required %zu", ws_bytes, P.bytes);
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  carve_plan(P, ws);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned rows4 = static_cast<unsigned>(ceil_div(n, 4));
  hipLaunchKernelGGL(k_pad_f32, dim3(rows4), dim3(256), 0, s, X, n, d, P.ld, P.Xp, P.sq);
  MV_LAUNCHED("k_pad_f32");
  Src src{P.Xp, nullptr, P.sq, P.ld};
  if (half) {
    hipLaunchKernelGGL(k_to_half, dim3(rows4), dim3(256), 0, s, P.Xp, n, P.ld, P.ldh, P.H);
    MV_LAUNCHED("k_to_half");
    src = Src{nullptr, P.H, P.sq, P.ldh};
  }
  return launch_gather(P, src, idx, 0, nullptr, out_dist, s);
}

}  // extern "C"
