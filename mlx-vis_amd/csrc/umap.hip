// UMAP fuzzy graph and edge-wise layout SGD for gfx950.
//
// Replaces, in mlx_vis/_umap/umap.py:
//   _fuzzy_simplicial_set (199-270)  -> k_fuzzy_rows, k_fuzzy_sym, k_fuzzy_sched
//   _optimize schedule    (400-414)  -> the per-edge epochs_per_next test in k_epoch
//   _optimize epoch loop  (423-447)  -> k_epoch + k_tail_apply per epoch
//   _sgd_step             (359-387)  -> k_sgd_step (explicit negatives, for parity)
//
// Graph layout in HBM: the reference keeps the forward edges (i, knn[i, c]) in
// row-major order (umap.py:243-247) and compacts the survivors of the prune
// (umap.py:265-270). Here the n x k slot array is kept dense: slot e = i*k + c
// has head i = e / k, tail knn[e], weight w[e], epochs_per_sample eps[e] (-1 for
// a pruned slot) and epochs_per_next epn[e]. No compaction or sort is needed:
// the reverse weight of (i, j) is found by scanning row j's k entries.
//
// Semantics kept from the reference:
//   - sigma bisection per row in FP32, 64 iterations, tolerance 1e-5 around
//     log2(k), bounds 1e-20 / 1e3, doubling while hi >= 1e3 (umap.py:214-236);
//   - symmetrised weight w + w' - w*w' on forward edges only (umap.py:249-262);
//   - prune at max(w_sym)/n_epochs (umap.py:265); schedule in FP32 with
//     eps = n_epochs / (n_epochs * (w / w_max)) (umap.py:400-404);
//   - an epoch is synchronous: every gradient of epoch e reads the embedding as
//     it was at the start of e; updates are accumulated into a delta buffer and
//     applied at the end (umap.py:361-387 returns a new Y from the old one);
//   - per active edge: one attractive update of head and tail, and
//     negative_sample_rate repulsive updates of the head only, each against a
//     uniformly drawn node (umap.py:433-436, 384-386).
// Negatives come from Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed
// by the seed and counted by (epoch, edge slot), in-edge lists (up to 16,384
// entries per node) are in slot order, and every tail sum is formed in an
// order that does not depend on scheduling (in-wavefront scans; two terms
// onto zero; fixed-point totals for nodes spanning three or more chunks of
// the tail sums), so a run repeats bit for bit (unless a node has over 16,384
// in-edges).
#include <math.h>

#include <algorithm>

#include "common.hpp"

namespace mv {
namespace umap {

constexpr int KMAX = 64;          // neighbours per row of the per-thread fuzzy-row kernel
constexpr int KMAX_GRAPH = 256;   // largest k of the graph, the spectral init and the epochs
constexpr int DIM_MAX = 8;   // largest embedding dimension of the layout kernels

__global__ void k_zero_u32(uint32_t* p, int count) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) p[t] = 0;
}

// ------------------------------------------------------------ fuzzy rows --
// One thread per row: rho, sigma bisection, membership weights
// (umap.py:208-240).
// `target` is log2(k) rounded to FP32 on the host (umap.py:204 compares FP32
// sums with it).
__global__ __launch_bounds__(256) void k_fuzzy_rows(const float* __restrict__ dist, int64_t n, int k,
                                                    float target, float* __restrict__ w,
                                                    float* __restrict__ sigma_out, float* __restrict__ rho_out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* d = dist + i * k;
  float dv[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) dv[j] = j < k ? d[j] : 0.f;

  float rho = INFINITY;
#pragma unroll
  for (int j = 0; j < KMAX; ++j)
    if (j < k && dv[j] > 0.f) rho = fminf(rho, dv[j]);
  rho = fmaxf(rho, 1e-8f);

  float sh[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) sh[j] = fmaxf(dv[j] - rho, 0.f);

  float lo = 1e-20f, hi = 1e3f, sigma = 1.f;
  for (int it = 0; it < 64; ++it) {
    float s = 0.f;
#pragma unroll
    for (int j = 1; j < KMAX; ++j)
      if (j < k) s += expf(-__fdiv_rn(sh[j], sigma));
    const bool conv = fabsf(s - target) < 1e-5f;
    if (conv) break;
    if (s > target) {
      hi = sigma;
      sigma = (lo + sigma) / 2.f;
    } else if (s < target) {
      lo = sigma;
      sigma = hi >= 1e3f ? sigma * 2.f : (sigma + hi) / 2.f;
    } else {
      break;  // NaN sums never converge; the reference leaves sigma unchanged
    }
  }
  const float den = fmaxf(sigma, 1e-10f);
  float* wr = w + i * k;
#pragma unroll
  for (int j = 0; j < KMAX; ++j)
    if (j < k) wr[j] = expf(-__fdiv_rn(sh[j], den));
  if (sigma_out) sigma_out[i] = sigma;
  if (rho_out) rho_out[i] = rho;
}

// The same for 64 < k <= 256: one wavefront per row, lane l holding columns
// l, l + 64, ...; the bisection's sums are wavefront reductions (a different
// summation order than the one-thread kernel's sequential sum).
constexpr int FW_PER_LANE = KMAX_GRAPH / 64;
__global__ __launch_bounds__(256) void k_fuzzy_rows_wave(const float* __restrict__ dist, int64_t n, int k,
                                                         float target, float* __restrict__ w,
                                                         float* __restrict__ sigma_out, float* __restrict__ rho_out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (i >= n) return;
  const float* d = dist + i * k;
  float dv[FW_PER_LANE];
  float rho = INFINITY;
#pragma unroll
  for (int u = 0; u < FW_PER_LANE; ++u) {
    const int j = l + 64 * u;
    dv[u] = j < k ? d[j] : 0.f;
    if (j < k && dv[u] > 0.f) rho = fminf(rho, dv[u]);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) rho = fminf(rho, __shfl_xor(rho, m));
  rho = fmaxf(rho, 1e-8f);
  float sh[FW_PER_LANE];
#pragma unroll
  for (int u = 0; u < FW_PER_LANE; ++u) sh[u] = fmaxf(dv[u] - rho, 0.f);
  float lo = 1e-20f, hi = 1e3f, sigma = 1.f;
  for (int it = 0; it < 64; ++it) {
    float s = 0.f;
#pragma unroll
    for (int u = 0; u < FW_PER_LANE; ++u) {
      const int j = l + 64 * u;
      if (j >= 1 && j < k) s += expf(-__fdiv_rn(sh[u], sigma));
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
    const bool conv = fabsf(s - target) < 1e-5f;
    if (conv) break;
    if (s > target) {
      hi = sigma;
      sigma = (lo + sigma) / 2.f;
    } else if (s < target) {
      lo = sigma;
      sigma = hi >= 1e3f ? sigma * 2.f : (sigma + hi) / 2.f;
    } else {
      break;
    }
  }
  const float den = fmaxf(sigma, 1e-10f);
  float* wr = w + i * k;
#pragma unroll
  for (int u = 0; u < FW_PER_LANE; ++u) {
    const int j = l + 64 * u;
    if (j < k) wr[j] = expf(-__fdiv_rn(sh[u], den));
  }
  if (l == 0) {
    if (sigma_out) sigma_out[i] = sigma;
    if (rho_out) rho_out[i] = rho;
  }
}

// w_sym[e] = w + w' - w*w' where w' is the weight of the reverse edge if the
// tail's own list contains the head (first matching column), else 0.
__global__ __launch_bounds__(256) void k_fuzzy_sym(const int32_t* __restrict__ knn, const float* __restrict__ w,
                                                   int64_t n, int k, float* __restrict__ wsym,
                                                   uint32_t* __restrict__ wmax_bits) {
  // slots strided over a bounded grid; the block's maximum is added with one
  // atomic (weights are >= 0, so their IEEE bit patterns order like the values)
  __shared__ uint32_t s_m[4];
  uint32_t bits = 0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n * k; e += stride) {
    const int64_t i = e / k;
    const int64_t j = knn[e];
    const float a = w[e];
    float b = 0.f;
    if (j >= 0 && j < n) {
      const int32_t* row = knn + j * k;
      for (int c = 0; c < k; ++c)
        if (row[c] == i) {
          b = w[j * k + c];
          break;
        }
    }
    const float v = a + b - a * b;
    wsym[e] = v;
    bits = __float_as_uint(v) > bits ? __float_as_uint(v) : bits;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const uint32_t o = static_cast<uint32_t>(__shfl_xor(static_cast<int>(bits), m));
    bits = o > bits ? o : bits;
  }
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t v = s_m[0];
    for (int q = 1; q < 4; ++q) v = s_m[q] > v ? s_m[q] : v;
    atomicMax(wmax_bits, v);
  }
}

// Prune and schedule each slot: kept when w_sym >= w_max / n_epochs
// (umap.py:265), eps = n_epochs / (n_epochs * (w / w_max)) (umap.py:402-403),
// epn = eps; pruned slots get eps = -1, epn = +inf. n_edges counts kept slots
// (one atomic per block).
__global__ __launch_bounds__(256) void k_fuzzy_sched(const float* __restrict__ wsym, int64_t count,
                                                     int n_epochs, const uint32_t* __restrict__ wmax_bits,
                                                     float* __restrict__ eps, float* __restrict__ epn,
                                                     uint32_t* __restrict__ n_edges) {
  __shared__ uint32_t s_c[4];
  const float wmax = __uint_as_float(*wmax_bits);
  const float ne = static_cast<float>(n_epochs);
  const float thr = __fdiv_rn(wmax, ne);
  uint32_t kept = 0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < count; e += stride) {
    const float v = wsym[e];
    const bool keep = v >= thr;
    float sv = -1.f;
    if (keep) {
      const float ns = __fmul_rn(ne, __fdiv_rn(v, wmax));
      sv = ns > 0.f ? __fdiv_rn(ne, ns) : -1.f;
    }
    eps[e] = keep ? sv : -1.f;
    epn[e] = keep ? sv : INFINITY;
    kept += keep ? 1u : 0u;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) kept += static_cast<uint32_t>(__shfl_xor(static_cast<int>(kept), m));
  if ((threadIdx.x & 63) == 0) s_c[threadIdx.x >> 6] = kept;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t t = s_c[0] + s_c[1] + s_c[2] + s_c[3];
    if (t) atomicAdd(n_edges, t);
  }
}

// -------------------------------------------------------------- gradients --

// Gradient coefficients on the hardware v_log_f32 / v_exp_f32 / v_rcp_f32
// (about 1 ulp each; no range scaling is needed: d2 is clamped to >= 1e-6, so
// log2 d2 >= -19.94 and the exponents b log2 d2 stay far from the denormal
// range). log2 d2 is shared by d2^b and d2^(b-1). Relative error of a
// coefficient ~ (|b log2 d2| + 3) 2^-23, below 3e-6 over the clamped range.
__device__ __forceinline__ float hw_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

struct Law {
  float a, b, alpha;
  // Attractive coefficient of umap.py:367-370: -2ab d2^(b-1) / (1 + a d2^b).
  __device__ __forceinline__ float attract(float d2) const {
    d2 = fmaxf(d2, 1e-6f);
    const float l = hw_log2(d2);
    const float pw = hw_exp2(b * l);
    const float pw1 = hw_exp2((b - 1.0f) * l);
    return -2.0f * a * b * pw1 * hw_rcp(1.0f + a * pw);
  }
  // Repulsive coefficient of umap.py:377-380: 2b / ((0.001 + d2)(1 + a d2^b)).
  __device__ __forceinline__ float repulse(float d2) const {
    d2 = fmaxf(d2, 1e-6f);
    const float pw = hw_exp2(b * hw_log2(d2));
    return 2.0f * b * hw_rcp((0.001f + d2) * (1.0f + a * pw));
  }
};

__device__ __forceinline__ float clip4(float v) { return fminf(fmaxf(v, -4.0f), 4.0f); }

// One synchronous epoch, phase 1, over the dense slot array. Each head row
// is served by G lanes, each lane holding S slots (c, c + G, ...; G S >= k),
// so that the slot schedules and tail indices load with several gathers in
// flight per lane. Only the active slots of the epoch (on average a third at
// 70,000 x 784) carry the expensive part, the tail and negative gathers and
// the gradients: they are compacted per wavefront through LDS (slot, tail
// index, head coordinates, in slot order u-major, lane-minor) and processed
// one item per lane, usually in a single round instead of S partly idle ones.
// Per item, the tail's position and the negatives (Philox-4x32-10 counter
// (slot, epoch, group of 4)) are gathered together, so that an item waits for
// one memory round trip after its slot's schedule; the item's head update goes
// back through LDS to the slot's lane. The head's updates are reduced over the
// row's G lanes into the next positions Ynext[i] (Y is double buffered). The
// tail updates are phase 2's: no atomics here, and nothing is stored per slot.
constexpr int NEG_CHUNK = 8;

// NR >= 0: negatives per edge fixed at compile time, -1: neg_rate; MASKED:
// movable is read (frozen nodes keep their position)
template <int D, int G, int S, int NR, bool MASKED>
__global__ __launch_bounds__(256) void k_epoch(const float* __restrict__ Y, float* __restrict__ Ynext,
                                               float* __restrict__ wbase, const int32_t* __restrict__ span,
                                               const uint8_t* __restrict__ movable, const int32_t* __restrict__ tail,
                                               const float* __restrict__ eps, float* __restrict__ epn, int64_t n,
                                               int k, int epoch, int neg_rate, Law law, uint32_t seed_lo,
                                               uint32_t seed_hi) {
  constexpr int NI = 64 * S;   // active items of a wavefront, at most
  __shared__ int32_t s_e[4][NI];
  __shared__ float s_yi[4][NI][D];   // head position, then the item's head update
  __shared__ int32_t s_j[4][NI];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t i = gid / G;      // head row
  const int c = static_cast<int>(gid % G);
  // No load is under a branch (a lane outside the slot array reads slot 0 and
  // row 0, and its values are masked): a branch per load makes the compiler
  // wait for the loads before it at each join, so the slots' schedules, tail
  // indices, the head position and the row owner's span entry and movable
  // flag are all in flight together.
  int64_t e[S];
  float sch[S], nx[S];
  int32_t j[S];
  bool in_row[S], act[S];
  const int64_t ic = i < n ? i : 0;
#pragma unroll
  for (int u = 0; u < S; ++u) {
    in_row[u] = i < n && c + u * G < k;
    e[u] = in_row[u] ? i * k + c + u * G : 0;
    sch[u] = eps[e[u]];
    nx[u] = epn[e[u]];
    j[u] = tail[e[u]];
  }
  float yi[D];
#pragma unroll
  for (int q = 0; q < D; ++q) yi[q] = Y[ic * D + q];
  const bool owner = c == 0 && i < n;
  const int32_t sp_r = span[ic];
  bool frozen = false;
  if constexpr (MASKED) frozen = movable[ic] == 0;
  const int32_t sp = owner ? sp_r : -1;
  frozen = owner && frozen;
#pragma unroll
  for (int u = 0; u < S; ++u) {
    act[u] = in_row[u] && sch[u] > 0.f && nx[u] <= static_cast<float>(epoch);
    if (act[u]) epn[e[u]] = nx[u] + sch[u];
  }
  // compaction of the active slots: item rank[u] of this wavefront
  int rank[S];
  int A = 0;
#pragma unroll
  for (int u = 0; u < S; ++u) {
    const uint64_t m = __ballot(act[u]);
    rank[u] = A + static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                             __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u)));
    A += __popcll(m);
    if (act[u]) {
      s_e[w][rank[u]] = static_cast<int32_t>(e[u]);
      s_j[w][rank[u]] = j[u];
#pragma unroll
      for (int q = 0; q < D; ++q) s_yi[w][rank[u]][q] = yi[q];
    }
  }
  __syncthreads();
  for (int r = 0; r * 64 < A; ++r) {
    const int pi = r * 64 + l;
    if (pi < A) {
      const int32_t ei = s_e[w][pi];
      const int64_t ji = s_j[w][pi];
      float hi_[D], hj[D], h[D];
#pragma unroll
      for (int q = 0; q < D; ++q) {
        hj[q] = Y[ji * D + q];
        hi_[q] = s_yi[w][pi][q];
        h[q] = 0.f;
      }
      if constexpr (NR >= 0) {
        // a count known at compile time: the gathers are straight-line code,
        // all in flight together (a branch per negative makes the compiler
        // wait for earlier gathers at each join); same Philox words, same
        // order of the additions as the general loop below
        constexpr int NG = (NR + 3) / 4;
        uint32_t words[NG > 0 ? 4 * NG : 1];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          const uint4 r = Philox::gen(make_uint4(static_cast<uint32_t>(ei), 0u, static_cast<uint32_t>(epoch),
                                                 static_cast<uint32_t>(g)),
                                      seed_lo, seed_hi);
          words[4 * g] = r.x;
          words[4 * g + 1] = r.y;
          words[4 * g + 2] = r.z;
          words[4 * g + 3] = r.w;
        }
        float yt[NR > 0 ? NR : 1][D];
#pragma unroll
        for (int s2 = 0; s2 < NR; ++s2) {
          const int64_t t = uniform_below(words[s2], static_cast<uint32_t>(n));
#pragma unroll
          for (int q = 0; q < D; ++q) yt[s2][q] = Y[t * D + q];
        }
#pragma unroll
        for (int s2 = 0; s2 < NR; ++s2) {
          float nd[D];
          float n2 = 0.f;
#pragma unroll
          for (int q = 0; q < D; ++q) {
            nd[q] = hi_[q] - yt[s2][q];
            n2 += nd[q] * nd[q];
          }
          const float cr = law.repulse(n2);
#pragma unroll
          for (int q = 0; q < D; ++q) h[q] += clip4(cr * nd[q]) * law.alpha;
        }
      }
      for (int base = 0; NR < 0 && base < neg_rate; base += NEG_CHUNK) {
        float yt[NEG_CHUNK][D];
        const uint4 r0 = Philox::gen(make_uint4(static_cast<uint32_t>(ei), 0u, static_cast<uint32_t>(epoch),
                                                static_cast<uint32_t>(base >> 2)),
                                     seed_lo, seed_hi);
        const uint4 r1 = Philox::gen(make_uint4(static_cast<uint32_t>(ei), 0u, static_cast<uint32_t>(epoch),
                                                static_cast<uint32_t>((base >> 2) + 1)),
                                     seed_lo, seed_hi);
        const uint32_t words[NEG_CHUNK] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
#pragma unroll
        for (int s2 = 0; s2 < NEG_CHUNK; ++s2) {
          if (base + s2 < neg_rate) {
            const int64_t t = uniform_below(words[s2], static_cast<uint32_t>(n));
#pragma unroll
            for (int q = 0; q < D; ++q) yt[s2][q] = Y[t * D + q];
          }
        }
#pragma unroll
        for (int s2 = 0; s2 < NEG_CHUNK; ++s2) {
          if (base + s2 < neg_rate) {
            float nd[D];
            float n2 = 0.f;
#pragma unroll
            for (int q = 0; q < D; ++q) {
              nd[q] = hi_[q] - yt[s2][q];
              n2 += nd[q] * nd[q];
            }
            const float cr = law.repulse(n2);
#pragma unroll
            for (int q = 0; q < D; ++q) h[q] += clip4(cr * nd[q]) * law.alpha;
          }
        }
      }
      float diff[D];
      float d2 = 0.f;
#pragma unroll
      for (int q = 0; q < D; ++q) {
        diff[q] = hi_[q] - hj[q];
        d2 += diff[q] * diff[q];
      }
      const float ca = law.attract(d2);
#pragma unroll
      for (int q = 0; q < D; ++q) {
        // rounded product (not fused into the add): phase 2 subtracts the same value
        h[q] += __fmul_rn(clip4(ca * diff[q]), law.alpha);
        s_yi[w][pi][q] = h[q];
      }
    }
  }
  __syncthreads();
  float acc[D];
#pragma unroll
  for (int q = 0; q < D; ++q) acc[q] = 0.f;
#pragma unroll
  for (int u = 0; u < S; ++u) {
    if (act[u]) {
#pragma unroll
      for (int q = 0; q < D; ++q) acc[q] += s_yi[w][rank[u]][q];
    }
  }
  // Reduce the head's updates over its G lanes (contiguous in the wavefront).
#pragma unroll
  for (int q = 0; q < D; ++q) {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) acc[q] += __shfl_xor(acc[q], m);
  }
  // Row owner: the next positions start as Y + head update. A node whose
  // in-edge entries cross a chunk boundary of phase 2 (span[i] = index of the
  // chunk holding its first entry) starts at 0 instead, and Y + head update
  // goes to wbase[span[i]] for that chunk to add with its sum; a frozen node
  // keeps Y.
  if (owner) {
    if (frozen) {
#pragma unroll
      for (int q = 0; q < D; ++q) Ynext[i * D + q] = yi[q];
    } else if (sp >= 0) {
#pragma unroll
      for (int q = 0; q < D; ++q) {
        wbase[static_cast<int64_t>(sp) * D + q] = yi[q] + acc[q];
        Ynext[i * D + q] = 0.f;
      }
    } else {
#pragma unroll
      for (int q = 0; q < D; ++q) Ynext[i * D + q] = yi[q] + acc[q];
    }
  }
}

// Phase 2: tail sums, applied, over the in-edge entries (entries of a node
// contiguous, in slot order; in_node[t] the node of entry t, in_head[t] the
// head of its slot). The entries keep their own copy of the schedule (eps_in,
// epn_in in entry order), advanced by the same FP32 additions as the slots'
// epn in phase 1, so both phases agree on the active edges of every epoch. An
// active entry computes minus the attractive update that phase 1 applied to
// its head, from the same two positions with the same expression, so the
// value is the same to the bit. Each wavefront takes a chunk of TAIL_CHUNK
// (128) entries, two per lane (the whole grid is then resident at once); a
// segmented inclusive scan adds each node's entries (the segment running
// from the first half of the chunk into the second carries its sum across),
// and the last lane of a segment adds the sum to Ynext with one atomic, onto
// the value phase 1 left there: Y + head update for a node whose entries lie
// in one chunk (one term onto a fixed value), 0 for a node crossing a chunk
// boundary, whose first segment adds wbase[c] (its Y + head update, indexed
// by the chunk) with its sum. Over two chunks the result does not depend on
// the order of the two atomics (two terms onto zero commute). A node whose
// entries span three or more chunks (a hub: more than 128 in-edges) would
// depend on their order, so there each chunk rounds its segment sum (the
// first one with the base) to 64-bit fixed point (units of 2^-32; integer
// additions are exact and commute) and adds it plus a counting bias to the
// node's accumulator with one atomic; the value it returns tells the chunk
// that completed the count, which stores the total as the node's next
// position and clears the accumulator for the next epoch. With the in-edge
// lists in slot order every segment sum is the same on every run, so a run
// repeats bit for bit (a node spanning more than 1,022 chunks, over 130,000
// in-edges, falls back to float atomics in arrival order). Frozen nodes
// (movable[j] == 0) are skipped.
constexpr double FIX_ONE = 4294967296.0;   // fixed-point unit of the hub tail sums: 2^-32
constexpr unsigned long long FIX_BIAS = 1ull << 54;   // one chunk counted into a hub accumulator
constexpr int HUB_MAX_WAVES = 1023;                   // chunks a fixed-point hub sum can count
constexpr int TAIL_CHUNK = 128;                       // in-edge entries per wavefront of k_tail_apply (two per lane)

template <int D, bool MASKED>   // MASKED: movable is read (frozen nodes keep their position)
__global__ __launch_bounds__(256) void k_tail_apply(const int32_t* __restrict__ in_node,
                                                    const int32_t* __restrict__ in_head,
                                                    const float* __restrict__ eps_in, float* __restrict__ epn_in,
                                                    int64_t m, const float* __restrict__ Y, float* __restrict__ Ynext,
                                                    const float* __restrict__ wbase, const uint8_t* __restrict__ movable,
                                                    const int32_t* __restrict__ span_w,
                                                    unsigned long long* __restrict__ hub_acc, int epoch, Law law) {
  const int64_t wv = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;   // chunk of this wavefront
  const int l = threadIdx.x & 63;
  const int64_t c0 = wv * TAIL_CHUNK;   // first entry of the chunk
  if (c0 >= m) return;                  // a wavefront past the last chunk (the grid is whole workgroups)
  // No load is under a branch (lanes past the end read the last entry and
  // their values are masked): a branch per load makes the compiler wait for
  // the loads before it at each join. The entry loads come first, then those
  // that depend on the node, then the two halves' position gathers together.
  bool valid[2];
  int64_t tc[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t t = c0 + h * 64 + l;
    valid[h] = t < m;
    tc[h] = valid[h] ? t : m - 1;
  }
  int rn[2];
  float es[2], ne[2];
  int64_t hd[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    rn[h] = in_node[tc[h]];
    es[h] = eps_in[tc[h]];
    ne[h] = epn_in[tc[h]];
    hd[h] = in_head[tc[h]];
  }
  // the node before the chunk (lane 0) and after it (lane 63); the base of the
  // node crossing out of the chunk (same address in every lane)
  const int rb = in_node[c0 > 0 ? c0 - 1 : 0];
  const int ra = in_node[c0 + TAIL_CHUNK < m ? c0 + TAIL_CHUNK : m - 1];
  float bq[D];
#pragma unroll
  for (int q = 0; q < D; ++q) bq[q] = wbase[wv * D + q];
  int node[2], crossings[2];
  bool mov[2], act[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    crossings[h] = span_w[rn[h]];   // chunk boundaries inside the node's entries
    mov[h] = true;
    if constexpr (MASKED) mov[h] = movable[rn[h]] != 0;
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    node[h] = valid[h] ? rn[h] : -1 - h;   // distinct out-of-range keys end the segments
    act[h] = valid[h] && es[h] > 0.f && ne[h] <= static_cast<float>(epoch);
    mov[h] = mov[h] && valid[h];
    crossings[h] = valid[h] ? crossings[h] : 0;
  }
  const int before = (l == 0 && c0 > 0) ? rb : -1;
  const int after = (l == 63 && c0 + TAIL_CHUNK < m) ? ra : -1;
  float base[D];
#pragma unroll
  for (int q = 0; q < D; ++q) base[q] = l == 63 && valid[1] ? bq[q] : 0.f;
  // an inactive entry gathers its own node's position twice (a line the
  // wavefront reads anyway) and contributes 0
  float yh[2][D], yn[2][D];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t src = act[h] ? hd[h] : static_cast<int64_t>(rn[h]);
#pragma unroll
    for (int q = 0; q < D; ++q) {
      yh[h][q] = Y[src * D + q];
      yn[h][q] = Y[static_cast<int64_t>(rn[h]) * D + q];
    }
  }
  float v[2][D];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (act[h]) epn_in[tc[h]] = ne[h] + es[h];
    float diff[D];
    float d2 = 0.f;
#pragma unroll
    for (int q = 0; q < D; ++q) {
      diff[q] = yh[h][q] - yn[h][q];
      d2 += diff[q] * diff[q];
    }
    const float ca = law.attract(d2);
#pragma unroll
    for (int q = 0; q < D; ++q) v[h][q] = act[h] ? -__fmul_rn(clip4(ca * diff[q]), law.alpha) : 0.f;
  }
  // Segmented inclusive scans of the two halves; the segment running from
  // the end of the first half into the second carries its sum into lane 0 of
  // the second half before that half is scanned.
  const int first_half1 = __shfl(node[1], 0);
  const int last_half0 = __shfl(node[0], 63);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h == 1) {
#pragma unroll
      for (int q = 0; q < D; ++q) {
        const float carry = __shfl(v[0][q], 63);
        if (l == 0 && first_half1 == last_half0) v[1][q] += carry;
      }
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int on = __shfl_up(node[h], off);
#pragma unroll
      for (int q = 0; q < D; ++q) {
        const float ov = __shfl_up(v[h][q], off);
        if (l >= off && on == node[h]) v[h][q] += ov;
      }
    }
  }
  const int first_node = __shfl(node[0], 0);
  const int before_node = __shfl(before, 0);
  const int after_node = __shfl(after, 63);
  const int down0 = __shfl_down(node[0], 1), down1 = __shfl_down(node[1], 1);   // every lane takes part
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    // next entry's node: the lane above, across the halves, or (lane 63 of
    // the second half) the chunk's end, where a continuing node counts as
    // crossing into the next chunk
    const int nxt = l < 63 ? (h == 0 ? down0 : down1) : (h == 0 ? first_half1 : -3);
    const bool last = valid[h] && nxt != node[h];
    if (last && mov[h]) {
    const bool from_before = first_node == node[h] && before_node == node[h];
    const bool to_after = h == 1 && l == 63 && after_node == node[h];
    const bool first_seg = to_after && !from_before;   // first segment of a node crossing into the next chunk
    const bool hub = crossings[h] >= 2 && crossings[h] < HUB_MAX_WAVES;
    const int64_t nd = static_cast<int64_t>(node[h]) * D;
    if (!hub) {
      if (first_seg) {
#pragma unroll
        for (int q = 0; q < D; ++q) v[h][q] += base[q];
      }
#pragma unroll
      for (int q = 0; q < D; ++q) atomicAdd(&Ynext[nd + q], v[h][q]);
    } else {
      // hub: the fixed-point total of all its chunks, stored by the last
      // one. Each chunk adds its partial plus FIX_BIAS, so the bits from 2^54
      // up count the chunks in (the partials stay below 2^53 in magnitude);
      // the returned value tells each one whether it completed the count and,
      // if so, the total, so one round trip suffices and no other location
      // has to be read.
      long long fv[D];
#pragma unroll
      for (int q = 0; q < D; ++q) {
        fv[q] = __double2ll_rn(static_cast<double>(v[h][q]) * FIX_ONE);
        if (first_seg) fv[q] += __double2ll_rn(static_cast<double>(base[q]) * FIX_ONE);
      }
      unsigned long long* acc = hub_acc + nd;
      const unsigned long long W = static_cast<unsigned long long>(crossings[h] + 1);
      unsigned long long nw[D];
#pragma unroll
      for (int q = 0; q < D; ++q) {
        const unsigned long long add = static_cast<unsigned long long>(fv[q]) + FIX_BIAS;
        nw[q] = atomicAdd(&acc[q], add) + add;
      }
#pragma unroll
      for (int q = 0; q < D; ++q) {
        if ((((nw[q] + FIX_BIAS / 2) >> 54) & 1023ull) == W) {
          const long long tot = static_cast<long long>(nw[q] - W * FIX_BIAS);
          Ynext[nd + q] = static_cast<float>(static_cast<double>(tot) / FIX_ONE);
          atomicExch(&acc[q], 0ull);   // ready for the next epoch
        }
      }
    }
    }
  }
}

// span[j] = index of the chunk (TAIL_CHUNK in-edge entries) holding node j's
// first entry if its entries cross a chunk boundary, else -1 (span pre-filled
// with -1, i.e. 0xFFFFFFFF as unsigned, so the unsigned minimum keeps the
// first boundary's chunk); span_w[j] counts the boundaries crossed.
__global__ __launch_bounds__(256) void k_span(const int32_t* __restrict__ in_node, int64_t m,
                                              int32_t* __restrict__ span, int32_t* __restrict__ span_w) {
  const int64_t t = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x + 1) * TAIL_CHUNK;
  if (t < m && in_node[t - 1] == in_node[t]) {
    atomicMin(reinterpret_cast<uint32_t*>(&span[in_node[t]]), static_cast<uint32_t>(t / TAIL_CHUNK - 1));
    atomicAdd(&span_w[in_node[t]], 1);
  }
}


__global__ __launch_bounds__(256) void k_fill_i32(int32_t* __restrict__ p, int64_t count, int32_t v) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) p[t] = v;
}

// In-edge lists of the kept slots: counts per tail, exclusive scan, atomic
// fill, then k_in_sort (up to 128 entries) and k_in_sort_big (up to 16,384)
// put the entries of each node in slot order.
__global__ __launch_bounds__(256) void k_in_count(const int32_t* __restrict__ tail, const float* __restrict__ eps,
                                                  int64_t m, uint32_t* __restrict__ cnt) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < m && eps[e] > 0.f) atomicAdd(&cnt[tail[e]], 1u);
}

__global__ __launch_bounds__(256) void k_in_fill(const int32_t* __restrict__ tail, const float* __restrict__ eps,
                                                 int64_t m, const uint32_t* __restrict__ off,
                                                 uint32_t* __restrict__ cursor, int32_t* __restrict__ in_slots,
                                                 int32_t* __restrict__ in_node) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < m && eps[e] > 0.f) {
    const int32_t j = tail[e];
    const uint32_t p = off[j] + atomicAdd(&cursor[j], 1u);
    in_slots[p] = static_cast<int32_t>(e);
    in_node[p] = j;
  }
}

__global__ __launch_bounds__(256) void k_scan_block_u32(const uint32_t* __restrict__ in, int64_t count,
                                                        uint32_t* __restrict__ out, uint32_t* __restrict__ sums) {
  __shared__ uint32_t sh[256];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * 1024 + threadIdx.x * 4;
  uint32_t v[4];
  uint32_t tot = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    v[j] = base + j < count ? in[base + j] : 0u;
    tot += v[j];
  }
  sh[threadIdx.x] = tot;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const uint32_t add = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t run = sh[threadIdx.x] - tot;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (base + j < count) out[base + j] = run;
    run += v[j];
  }
  if (threadIdx.x == 255) sums[blockIdx.x] = sh[255];
}

__global__ __launch_bounds__(256) void k_scan_sums_u32(uint32_t* __restrict__ sums, int64_t nb) {
  __shared__ uint32_t sh[256];
  const int64_t per = (nb + 255) / 256;
  const int64_t b0 = threadIdx.x * per;
  uint32_t tot = 0;
  for (int64_t j = 0; j < per; ++j)
    if (b0 + j < nb) tot += sums[b0 + j];
  sh[threadIdx.x] = tot;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const uint32_t add = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t run = sh[threadIdx.x] - tot;
  for (int64_t j = 0; j < per; ++j)
    if (b0 + j < nb) {
      const uint32_t v = sums[b0 + j];
      sums[b0 + j] = run;
      run += v;
    }
}

__global__ __launch_bounds__(256) void k_scan_add_u32(uint32_t* __restrict__ out, int64_t count,
                                                      const uint32_t* __restrict__ sums) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) out[t] += sums[t / 1024];
}

// Orders each node's in-edge segment [off[j], off[j] + cnt[j]) by slot
// (k_in_fill places them in atomic order), so the summation order of a tail
// sum is the same on every run. One wavefront per node: lane t holds entries
// t and t + 64, the rank of an entry is the number of smaller slots (slots are
// distinct), and it is stored at its rank. Segments of up to 128 entries are
// ordered here, longer ones by k_in_sort_big.
__global__ __launch_bounds__(256) void k_in_sort(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off,
                                                 int64_t n, int32_t* __restrict__ in_slots) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (j >= n) return;
  const int c = static_cast<int>(cnt[j]);
  if (c < 2 || c > 128) return;
  int32_t* seg = in_slots + off[j];
  const int32_t v0 = l < c ? seg[l] : INT32_MAX;
  const int32_t v1 = l + 64 < c ? seg[l + 64] : INT32_MAX;
  int r0 = 0, r1 = 0;
  for (int u = 0; u < 64; ++u) {
    const int32_t a = __shfl(v0, u), b = __shfl(v1, u);
    r0 += (a < v0) + (b < v0);
    r1 += (a < v1) + (b < v1);
  }
  if (l < c) seg[r0] = v0;
  if (l + 64 < c) seg[r1] = v1;
}

// Segments of 129 to IN_SORT_BIG entries: a bitonic sort of the slots in LDS
// (padded with INT32_MAX to a power of two), written back in slot order, one
// node at a time per workgroup. Workgroups stride over the nodes 256 at a
// time (one count per thread, the long segments listed in LDS). Longer
// segments stay in fill order (their per-wavefront sums may then differ in
// rounding from run to run).
constexpr int IN_SORT_BIG = 16384;
__global__ __launch_bounds__(256) void k_in_sort_big(const uint32_t* __restrict__ cnt,
                                                     const uint32_t* __restrict__ off, int64_t n,
                                                     int32_t* __restrict__ in_slots) {
  __shared__ int32_t s_v[IN_SORT_BIG];
  __shared__ int32_t s_list[256];
  __shared__ int s_nlist;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * 256; base < n; base += static_cast<int64_t>(gridDim.x) * 256) {
    if (threadIdx.x == 0) s_nlist = 0;
    __syncthreads();
    const int64_t jt = base + threadIdx.x;
    if (jt < n) {
      const uint32_t ct = cnt[jt];
      if (ct > 128 && ct <= IN_SORT_BIG) s_list[atomicAdd(&s_nlist, 1)] = static_cast<int32_t>(threadIdx.x);
    }
    __syncthreads();
    const int nl = s_nlist;
    __syncthreads();   // every thread has read the count before the next round resets it
    for (int b = 0; b < nl; ++b) {
      const int64_t j = base + s_list[b];
      const int c = static_cast<int>(cnt[j]);
      int P = 256;
      while (P < c) P <<= 1;
      int32_t* seg = in_slots + off[j];
      for (int i = threadIdx.x; i < P; i += blockDim.x) s_v[i] = i < c ? seg[i] : INT32_MAX;
      __syncthreads();
      for (int kk = 2; kk <= P; kk <<= 1) {
        for (int jj = kk >> 1; jj > 0; jj >>= 1) {
          for (int i = threadIdx.x; i < P; i += blockDim.x) {
            const int ixj = i ^ jj;
            if (ixj > i) {
              const int32_t x = s_v[i], y = s_v[ixj];
              if ((x > y) == ((i & kk) == 0)) {
                s_v[i] = y;
                s_v[ixj] = x;
              }
            }
          }
          __syncthreads();
        }
      }
      for (int i = threadIdx.x; i < c; i += blockDim.x) seg[i] = s_v[i];
      __syncthreads();   // s_v is refilled for the next node
    }
  }
}

// Y += dY; dY = 0.
__global__ __launch_bounds__(256) void k_apply(float* __restrict__ Y, float* __restrict__ dY, int64_t count) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) {
    Y[t] += dY[t];
    dY[t] = 0.f;
  }
}

__global__ __launch_bounds__(256) void k_zero_f32(float* __restrict__ p, int64_t count) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) p[t] = 0.f;
}

__global__ __launch_bounds__(256) void k_copy_f32(float* __restrict__ dst, const float* __restrict__ src,
                                                  int64_t count) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) dst[t] = src[t];
}

// Exact _sgd_step (umap.py:361-387) with explicit edge and negative lists:
// Y_out must hold a copy of Y_in; every update reads Y_in.
__global__ __launch_bounds__(256) void k_sgd_step(const float* __restrict__ Yin, float* __restrict__ Yout,
                                                  int64_t n, int dim, const int32_t* __restrict__ ef,
                                                  const int32_t* __restrict__ et, int64_t A,
                                                  const int32_t* __restrict__ nf, const int32_t* __restrict__ nt,
                                                  int64_t n_neg, Law law) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < A) {
    const int64_t f = ef[t], g = et[t];
    float d2 = 0.f;
    for (int q = 0; q < dim; ++q) {
      const float df = Yin[f * dim + q] - Yin[g * dim + q];
      d2 += df * df;
    }
    const float ca = law.attract(d2);
    for (int q = 0; q < dim; ++q) {
      const float df = Yin[f * dim + q] - Yin[g * dim + q];
      const float gr = clip4(ca * df) * law.alpha;
      atomicAdd(&Yout[f * dim + q], gr);
      atomicAdd(&Yout[g * dim + q], -gr);
    }
  } else if (t < A + n_neg) {
    const int64_t u = t - A;
    const int64_t f = nf[u], g = nt[u];
    float d2 = 0.f;
    for (int q = 0; q < dim; ++q) {
      const float df = Yin[f * dim + q] - Yin[g * dim + q];
      d2 += df * df;
    }
    const float cr = law.repulse(d2);
    for (int q = 0; q < dim; ++q) {
      const float df = Yin[f * dim + q] - Yin[g * dim + q];
      atomicAdd(&Yout[f * dim + q], clip4(cr * df) * law.alpha);
    }
  }
}

// ---------------------------------------------------------------- spectral --
// Power iteration of D^-1/2 W D^-1/2 on the kept forward edges with
// re-orthonormalisation after every product (umap.py:298-337). The reference
// orthonormalises with modified Gram-Schmidt; here the Gram matrix of the
// product is accumulated during the product (FP64) and the Cholesky factor R
// (U = Q R, positive diagonal) gives the same Q = U R^-1, applied lazily by the
// next product. Degrees come from the heads only (umap.py:307-311).

__global__ __launch_bounds__(256) void k_spec_prep(const float* __restrict__ wsym, const float* __restrict__ eps,
                                                   int64_t n, int k, float* __restrict__ dis) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float deg = 0.f;
  for (int c = 0; c < k; ++c) {
    const int64_t e = i * k + c;
    if (eps[e] > 0.f) deg += wsym[e];
  }
  dis[i] = 1.0f / sqrtf(fmaxf(deg, 1e-10f));
}

__global__ __launch_bounds__(256) void k_spec_wnorm(const float* __restrict__ wsym, const float* __restrict__ eps,
                                                    const int32_t* __restrict__ tail, const float* __restrict__ dis,
                                                    int64_t n, int k, float* __restrict__ wn) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  const int64_t i = e / k;
  wn[e] = eps[e] > 0.f ? wsym[e] * dis[i] * dis[tail[e]] : 0.f;
}

constexpr int NC_MAX = DIM_MAX + 1;   // spectral vectors (the first is dropped)
constexpr int GRAM_STRIDE = 48;       // >= NC_MAX (NC_MAX + 1) / 2 packed entries
// Each Gram buffer holds GRAM_COPIES partial sums: block b adds into copy
// b mod GRAM_COPIES, so that few atomics meet on one address; the reader adds
// the copies in copy order.
constexpr int GRAM_COPIES = 32;

// Sum over the copies of the packed Gram entry p = threadIdx.x (threads
// p < NP; 0 elsewhere, and 0 when g is null), in copy order. The loads are
// not under a branch (every thread reads a valid entry of `any`, a Gram
// buffer, when g is null or p is out of range), so they stay in flight with
// the loads that follow.
template <int NC>
__device__ __forceinline__ double gram_sum(const double* __restrict__ g, const double* __restrict__ any) {
  constexpr int NP = NC * (NC + 1) / 2;
  const int p = threadIdx.x < NP ? static_cast<int>(threadIdx.x) : 0;
  const double* src = g ? g : any;
  double c8[GRAM_COPIES];
#pragma unroll
  for (int c = 0; c < GRAM_COPIES; ++c) c8[c] = src[c * GRAM_STRIDE + p];
  double v = 0.0;
#pragma unroll
  for (int c = 0; c < GRAM_COPIES; ++c) v += c8[c];
  return g && threadIdx.x < NP ? v : 0.0;
}

// s_g[p] = the summed Gram entry p (threads p < NP hold it in v); followed by
// a barrier.
template <int NC>
__device__ __forceinline__ void gram_store(double v, double* s_g) {
  constexpr int NP = NC * (NC + 1) / 2;
  if (threadIdx.x < NP) s_g[threadIdx.x] = v;
  __syncthreads();
}

// R^-1 of the Cholesky factor (G = R^T R, R upper, pivots sqrt(. + 1e-10) as
// the reference's norms, umap.py:335) of a packed Gram matrix; identity if g
// is null.
template <int NC>
__device__ void chol_inverse(const double* __restrict__ g, float* __restrict__ out) {
  double R[NC][NC], Ri[NC][NC];
  for (int p = 0; p < NC; ++p)
    for (int q = 0; q < NC; ++q) {
      R[p][q] = 0.0;
      Ri[p][q] = p == q ? 1.0 : 0.0;
    }
  if (g) {
    double G[NC][NC];
    int idx = 0;
    for (int p = 0; p < NC; ++p)
      for (int q = p; q < NC; ++q) {
        G[p][q] = g[idx];
        G[q][p] = g[idx];
        ++idx;
      }
    for (int j = 0; j < NC; ++j) {
      double s = G[j][j];
      for (int i = 0; i < j; ++i) s -= R[i][j] * R[i][j];
      const double rjj = sqrt(fmax(s, 0.0) + 1e-10);
      R[j][j] = rjj;
      for (int q = j + 1; q < NC; ++q) {
        double v = G[j][q];
        for (int i = 0; i < j; ++i) v -= R[i][j] * R[i][q];
        R[j][q] = v / rjj;
      }
    }
    for (int q = 0; q < NC; ++q) {
      Ri[q][q] = 1.0 / R[q][q];
      for (int p = q - 1; p >= 0; --p) {
        double v = 0.0;
        for (int m = p + 1; m <= q; ++m) v += R[p][m] * Ri[m][q];
        Ri[p][q] = -v / R[p][p];
      }
    }
  }
  for (int p = 0; p < NC; ++p)
    for (int q = 0; q < NC; ++q) out[p * NC + q] = static_cast<float>(Ri[p][q]);
}

// One power-iteration step: Unext = (W_norm U) R^-1(gram_prev), accumulating
// the Gram matrix of Unext into gram_cur (one atomic per block and entry) and
// clearing gram_zero for the next step. Row i is served by G lanes holding S
// adjacent slots each (G S >= k), so that the weights and tail indices of a
// wavefront's rows load as contiguous runs and all of a lane's gathers are in
// flight at once; the G partial sums are added by a butterfly over the row's
// lanes.
template <int NC, int G, int S>
__global__ __launch_bounds__(256) void k_spec_iter(const float* __restrict__ U, const double* __restrict__ gram_prev,
                                                   const float* __restrict__ wn, const int32_t* __restrict__ tail,
                                                   int64_t n, int k, float* __restrict__ Unext,
                                                   double* __restrict__ gram_cur, double* __restrict__ gram_zero) {
  constexpr int NP = NC * (NC + 1) / 2;
  __shared__ float s_R[NC * NC];
  __shared__ double s_part[4][NP];
  __shared__ double s_g[NP];
  if (blockIdx.x == 0)
    for (int z = threadIdx.x; z < GRAM_COPIES * GRAM_STRIDE; z += blockDim.x) gram_zero[z] = 0.0;
  // the previous step's Gram matrix, loaded before the gathers so that both
  // arrive together
  const double gsum = gram_sum<NC>(gram_prev, gram_cur);
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t i = gid / G;
  const int c = static_cast<int>(gid % G);
  float t[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) t[q] = 0.f;
  {
    // no load under a branch: slots past the row (or rows past n) read the
    // row's (row 0's) last slot again with weight 0
    const int64_t ic = i < n ? i : 0;
    float w[S];
    int64_t j[S];
#pragma unroll
    for (int v = 0; v < S; ++v) {
      const int cs = c * S + v;
      const int64_t e = ic * k + (cs < k ? cs : k - 1);
      w[v] = wn[e];
      j[v] = tail[e];
      if (!(i < n && cs < k)) w[v] = 0.f;
    }
    float u8[S][NC];
#pragma unroll
    for (int v = 0; v < S; ++v)
#pragma unroll
      for (int q = 0; q < NC; ++q) u8[v][q] = U[j[v] * NC + q];
#pragma unroll
    for (int v = 0; v < S; ++v)
#pragma unroll
      for (int q = 0; q < NC; ++q) t[q] += w[v] * u8[v][q];
  }
#pragma unroll
  for (int q = 0; q < NC; ++q)
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) t[q] += __shfl_xor(t[q], m);
  // R^-1 after the product: the Cholesky factorisation of the previous Gram
  // matrix (a dependent load and serial FP64 work) overlaps the gathers of
  // the other wavefronts instead of delaying them all.
  gram_store<NC>(gsum, s_g);
  if (threadIdx.x == 0) chol_inverse<NC>(gram_prev ? s_g : nullptr, s_R);
  __syncthreads();
  const bool owner = c == 0 && i < n;
  float u[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    float acc = 0.f;
#pragma unroll
    for (int p = 0; p <= q; ++p) acc += t[p] * s_R[p * NC + q];
    u[q] = owner ? acc : 0.f;
  }
  if (owner) {
#pragma unroll
    for (int q = 0; q < NC; ++q) Unext[i * NC + q] = u[q];
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int idx = 0;
#pragma unroll
  for (int p = 0; p < NC; ++p) {
#pragma unroll
    for (int q = p; q < NC; ++q) {
      double v = static_cast<double>(u[p]) * static_cast<double>(u[q]);
#pragma unroll
      for (int m = 32; m >= G; m >>= 1) v += __shfl_xor(v, m);   // only the row owners (lanes c = 0) hold terms
      if (l == 0) s_part[w][idx] = v;
      ++idx;
    }
  }
  __syncthreads();
  if (threadIdx.x < NP) {
    const double v = s_part[0][threadIdx.x] + s_part[1][threadIdx.x] + s_part[2][threadIdx.x] +
                     s_part[3][threadIdx.x];
    atomicAdd(&gram_cur[(blockIdx.x % GRAM_COPIES) * GRAM_STRIDE + threadIdx.x], v);
  }
}

// V = U R^-1(gram).
template <int NC>
__global__ __launch_bounds__(256) void k_spec_apply(const float* __restrict__ U, const double* __restrict__ gram,
                                                    int64_t n, float* __restrict__ V) {
  constexpr int NP = NC * (NC + 1) / 2;
  __shared__ float s_R[NC * NC];
  __shared__ double s_g[NP];
  gram_store<NC>(gram_sum<NC>(gram, gram), s_g);
  if (threadIdx.x == 0) chol_inverse<NC>(gram ? s_g : nullptr, s_R);
  __syncthreads();
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    float acc = 0.f;
#pragma unroll
    for (int p = 0; p <= q; ++p) acc += U[i * NC + p] * s_R[p * NC + q];
    V[i * NC + q] = acc;
  }
}


template <int NC>
static int spectral_run(const float* wsym, const float* eps, const int32_t* tail, int64_t n, int k, int iters,
                        float* V, char* ws, hipStream_t s) {
  Carve cv(ws);
  float* dis = cv.take<float>(n);
  float* wn = cv.take<float>(static_cast<size_t>(n) * k);
  float* U0 = cv.take<float>(static_cast<size_t>(n) * NC);
  float* U1 = cv.take<float>(static_cast<size_t>(n) * NC);
  double* gram = cv.take<double>(3 * GRAM_COPIES * GRAM_STRIDE);   // three rotating packed Gram buffers
  const unsigned rb = static_cast<unsigned>(ceil_div(n, 256));
  hipLaunchKernelGGL(k_spec_prep, dim3(rb), dim3(256), 0, s, wsym, eps, n, k, dis);
  MV_LAUNCHED("k_spec_prep");
  hipLaunchKernelGGL(k_spec_wnorm, dim3(static_cast<unsigned>(ceil_div(n * k, 256))), dim3(256), 0, s, wsym, eps,
                     tail, dis, n, k, wn);
  MV_LAUNCHED("k_spec_wnorm");
  constexpr int64_t gram_floats = 2 * 3 * GRAM_COPIES * GRAM_STRIDE;   // the three Gram buffers, as floats
  hipLaunchKernelGGL(k_zero_f32, dim3(static_cast<unsigned>(ceil_div(gram_floats, 256))), dim3(256), 0, s,
                     reinterpret_cast<float*>(gram), gram_floats);
  MV_LAUNCHED("k_zero_f32");
  hipLaunchKernelGGL(k_copy_f32, dim3(static_cast<unsigned>(ceil_div(n * NC, 256))), dim3(256), 0, s, U0, V,
                     n * NC);
  MV_LAUNCHED("k_copy_f32");
  if (iters == 0) return MV_OK;
  float* cur = U0;
  float* nxt = U1;
  for (int it = 0; it < iters; ++it) {
    constexpr int GB = GRAM_COPIES * GRAM_STRIDE;   // one Gram buffer
    const double* prev = it == 0 ? nullptr : gram + GB * ((it + 2) % 3);
    double* gcur = gram + GB * (it % 3);
    double* gzero = gram + GB * ((it + 1) % 3);
    if (k <= 16)
      hipLaunchKernelGGL((k_spec_iter<NC, 4, 4>), dim3(static_cast<unsigned>(ceil_div(n * 4, 256))), dim3(256), 0, s,
                         cur, prev, wn, tail, n, k, nxt, gcur, gzero);
    else if (k <= 64)
      hipLaunchKernelGGL((k_spec_iter<NC, 16, 4>), dim3(static_cast<unsigned>(ceil_div(n * 16, 256))), dim3(256), 0,
                         s, cur, prev, wn, tail, n, k, nxt, gcur, gzero);
    else
      hipLaunchKernelGGL((k_spec_iter<NC, 64, 4>), dim3(static_cast<unsigned>(ceil_div(n * 64, 256))), dim3(256), 0,
                         s, cur, prev, wn, tail, n, k, nxt, gcur, gzero);
    MV_LAUNCHED("k_spec_iter");
    float* t = cur;
    cur = nxt;
    nxt = t;
  }
  hipLaunchKernelGGL((k_spec_apply<NC>), dim3(rb), dim3(256), 0, s, cur,
                     gram + GRAM_COPIES * GRAM_STRIDE * ((iters - 1) % 3), n,
                     V);
  MV_LAUNCHED("k_spec_apply");
  return MV_OK;
}

static size_t spectral_ws(int64_t n, int k, int nc) {
  return align_up(n * 4, 256) + align_up(static_cast<size_t>(n) * k * 4, 256) +
         2 * align_up(static_cast<size_t>(n) * nc * 4, 256) + align_up(3 * GRAM_COPIES * GRAM_STRIDE * 8, 256) + 1024;
}

struct EpochBufs {
  float* wbase;      // ceil(n*k/TAIL_CHUNK)*dim: Y + head update of the node crossing out of each chunk
  float* ybuf;       // n*dim, the other half of the double-buffered positions
  int32_t* span;     // n: chunk of the node's first in-edge entry if its entries cross a boundary, else -1
  int32_t* span_w;   // n: chunk boundaries inside the node's entries
  unsigned long long* hub_acc;   // n*dim: biased fixed-point tail sums of nodes spanning three or more chunks
  int32_t* in_head;  // n*k (n_in used): head of each in-edge entry's slot
  float* eps_in;     // n*k: the entry's epochs_per_sample
  float* epn_in;     // n*k: the entry's epochs_per_next, advanced by phase 2
};

static size_t epochs_ws(int64_t n, int k, int dim) {
  return align_up(static_cast<size_t>(ceil_div(n * k, TAIL_CHUNK)) * dim * 4, 256) +
         align_up(static_cast<size_t>(n) * dim * 4, 256) + 2 * align_up(static_cast<size_t>(n) * 4, 256) +
         align_up(static_cast<size_t>(n) * dim * 8, 256) + 3 * align_up(static_cast<size_t>(n) * k * 4, 256) + 256;
}

static EpochBufs carve_epochs(void* ws, int64_t n, int k, int dim) {
  Carve c(ws);
  EpochBufs b;
  b.wbase = c.take<float>(static_cast<size_t>(ceil_div(n * k, TAIL_CHUNK)) * dim);
  b.ybuf = c.take<float>(static_cast<size_t>(n) * dim);
  b.span = c.take<int32_t>(static_cast<size_t>(n));
  b.span_w = c.take<int32_t>(static_cast<size_t>(n));
  b.hub_acc = c.take<unsigned long long>(static_cast<size_t>(n) * dim);
  b.in_head = c.take<int32_t>(static_cast<size_t>(n) * k);
  b.eps_in = c.take<float>(static_cast<size_t>(n) * k);
  b.epn_in = c.take<float>(static_cast<size_t>(n) * k);
  return b;
}

// Phase 2's copy of the schedule, in entry order: the head, eps and the
// current epn of each in-edge entry's slot (the caller's epn at the start of a
// call, which phase 1 keeps advancing in slot order).
__global__ __launch_bounds__(256) void k_in_prep(const int32_t* __restrict__ in_slots, int64_t n_in, int k,
                                                 const float* __restrict__ eps, const float* __restrict__ epn,
                                                 int32_t* __restrict__ in_head, float* __restrict__ eps_in,
                                                 float* __restrict__ epn_in) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n_in) return;
  const int64_t e = in_slots[t];
  in_head[t] = static_cast<int32_t>(e / k);
  eps_in[t] = eps[e];
  epn_in[t] = epn[e];
}

template <int D, int G, int S>
static int launch_epoch(const float* Y, float* Ynext, const EpochBufs& B, const uint8_t* movable,
                        const int32_t* tail, const float* eps, float* epn, int64_t n, int k, int epoch, int neg_rate,
                        Law law, uint64_t seed, hipStream_t s) {
  const int64_t threads = n * G;
  const unsigned blocks = static_cast<unsigned>(ceil_div(threads, 256));
  // the reference's default of 5 negatives per edge runs a specialised copy
  // (the masked epochs, used for refinement, run the general copy)
  auto kern = movable ? k_epoch<D, G, S, -1, true> : neg_rate == 5 ? k_epoch<D, G, S, 5, false> : k_epoch<D, G, S, -1, false>;
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, Y, Ynext, B.wbase, B.span, movable, tail, eps, epn, n, k,
                     epoch, neg_rate, law, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
  MV_LAUNCHED("k_epoch");
  return MV_OK;
}

// One epoch from positions Y into Ynext (two launches: head updates, then
// tail updates).
template <int D>
static int run_epoch(const float* Y, float* Ynext, const EpochBufs& B, const int32_t* tail, const float* eps,
                     float* epn, const int32_t* in_node, int64_t n_in, int64_t n, int k, int epoch, int neg_rate,
                     Law law, uint64_t seed, const uint8_t* movable, hipStream_t s) {
  int rc;
  // two slots per lane from k = 9 on (G lanes per row, G * S >= k)
  if (k <= 8) rc = launch_epoch<D, 8, 1>(Y, Ynext, B, movable, tail, eps, epn, n, k, epoch, neg_rate, law, seed, s);
  else if (k <= 16) rc = launch_epoch<D, 8, 2>(Y, Ynext, B, movable, tail, eps, epn, n, k, epoch, neg_rate, law, seed, s);
  else if (k <= 32) rc = launch_epoch<D, 16, 2>(Y, Ynext, B, movable, tail, eps, epn, n, k, epoch, neg_rate, law, seed, s);
  else if (k <= 64) rc = launch_epoch<D, 32, 2>(Y, Ynext, B, movable, tail, eps, epn, n, k, epoch, neg_rate, law, seed, s);
  else if (k <= 128) rc = launch_epoch<D, 32, 4>(Y, Ynext, B, movable, tail, eps, epn, n, k, epoch, neg_rate, law, seed, s);
  else rc = launch_epoch<D, 64, 4>(Y, Ynext, B, movable, tail, eps, epn, n, k, epoch, neg_rate, law, seed, s);
  if (rc) return rc;
  if (n_in > 0) {
    // one wavefront per chunk of TAIL_CHUNK entries, four per workgroup
    auto tail_kernel = movable ? k_tail_apply<D, true> : k_tail_apply<D, false>;
    hipLaunchKernelGGL(tail_kernel, dim3(static_cast<unsigned>(ceil_div(n_in, 4 * TAIL_CHUNK))), dim3(256), 0, s,
                       in_node,
                       B.in_head, B.eps_in, B.epn_in, n_in, Y, Ynext, B.wbase, movable, B.span_w, B.hub_acc,
                       epoch, law);
    MV_LAUNCHED("k_tail_apply");
  }
  return MV_OK;
}

// Standard normal draws for the spectral start vectors: Philox-4x32-10 on the
// counter (i, 0x5A11, 0, 0) under key (seed, seed >> 32), two Box-Muller pairs
// per call (uniforms on (0, 1] from the top 24 bits).
__global__ __launch_bounds__(256) void k_random_normal(float* __restrict__ out, int64_t count, uint64_t seed) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i * 4 >= count) return;
  const uint4 r = Philox::gen(make_uint4(static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), 0x5A11u, 0u),
                              static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
  const float u0 = (static_cast<float>(r.x >> 8) + 1.f) * (1.f / 16777216.f);
  const float u1 = static_cast<float>(r.y >> 8) * (1.f / 16777216.f);
  const float u2 = (static_cast<float>(r.z >> 8) + 1.f) * (1.f / 16777216.f);
  const float u3 = static_cast<float>(r.w >> 8) * (1.f / 16777216.f);
  const float m0 = sqrtf(-2.f * logf(u0)), m1 = sqrtf(-2.f * logf(u2));
  const float a0 = 6.2831853071795865f * u1, a1 = 6.2831853071795865f * u3;
  const float v[4] = {m0 * cosf(a0), m0 * sinf(a0), m1 * cosf(a1), m1 * sinf(a1)};
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (i * 4 + q < count) out[i * 4 + q] = v[q];
}

// Spectral embedding to [0, 10] (umap.py:344-348): e = V[:, 1:nc] * (10 / max|V[:, 1:nc]|)
// + noise, then per column 10 (e - min) / (max - min + 1e-10). `scal` holds the
// maximum |v| (float bits), the per-column minima and maxima (order-preserving
// integer keys) and a non-finite flag.
__device__ __forceinline__ uint32_t ordered_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ordered_value(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}
struct EmbedScal {
  uint32_t absmax;
  uint32_t lo[DIM_MAX];   // min keys
  uint32_t hi[DIM_MAX];   // max keys
  int32_t nonfinite;
};

__global__ void k_embed_init(EmbedScal* sc) {
  if (threadIdx.x == 0) {
    sc->absmax = 0u;
    for (int q = 0; q < DIM_MAX; ++q) {
      sc->lo[q] = 0xFFFFFFFFu;
      sc->hi[q] = 0u;
    }
    sc->nonfinite = 0;
  }
}

__device__ __forceinline__ uint32_t block_reduce_max_u32(uint32_t v, uint32_t* red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = max(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), off)));
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  uint32_t r = red[0];
  for (int i = 1; i < static_cast<int>(blockDim.x >> 6); ++i) r = max(r, red[i]);
  return r;
}

__global__ __launch_bounds__(256) void k_embed_absmax(const float* __restrict__ V, int64_t n, int nc,
                                                      EmbedScal* __restrict__ sc) {
  __shared__ uint32_t red[4];
  const int dim = nc - 1;
  uint32_t m = 0u;
  int bad = 0;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n * dim;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const float v = V[(t / dim) * nc + 1 + t % dim];
    if (!isfinite(v)) bad = 1;
    else m = max(m, __float_as_uint(fabsf(v)));
  }
  m = block_reduce_max_u32(m, red);
  const uint32_t anybad = block_reduce_max_u32(static_cast<uint32_t>(bad), red);
  if (threadIdx.x == 0) {
    atomicMax(&sc->absmax, m);
    if (anybad) atomicOr(&sc->nonfinite, 1);
  }
}

// Y holds the noise draws on entry (standard normal); Y = e * expansion + 1e-4 * noise.
__global__ __launch_bounds__(256) void k_embed_scale(const float* __restrict__ V, int64_t n, int nc,
                                                     float* __restrict__ Y, EmbedScal* __restrict__ sc) {
  __shared__ uint32_t red[4];
  const int dim = nc - 1;
  const float amax = __uint_as_float(sc->absmax);
  const float expansion = 10.0f / amax;
  uint32_t lo[DIM_MAX], hi[DIM_MAX];
#pragma unroll
  for (int c = 0; c < DIM_MAX; ++c) {
    lo[c] = 0xFFFFFFFFu;
    hi[c] = 0u;
  }
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n * dim;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int q = static_cast<int>(t % dim);
    const float e = V[(t / dim) * nc + 1 + q] * expansion + Y[t] * 0.0001f;
    Y[t] = e;
    const uint32_t key = ordered_key(e);
#pragma unroll
    for (int c = 0; c < DIM_MAX; ++c)
      if (c == q) {
        lo[c] = min(lo[c], key);
        hi[c] = max(hi[c], key);
      }
  }
#pragma unroll
  for (int c = 0; c < DIM_MAX; ++c) {
    if (c >= dim) break;
    const uint32_t l = ~block_reduce_max_u32(~lo[c], red);
    const uint32_t h = block_reduce_max_u32(hi[c], red);
    if (threadIdx.x == 0) {
      atomicMin(&sc->lo[c], l);
      atomicMax(&sc->hi[c], h);
    }
  }
}

__global__ __launch_bounds__(256) void k_embed_norm(float* __restrict__ Y, int64_t n, int dim,
                                                    const EmbedScal* __restrict__ sc) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n * dim) return;
  const int q = static_cast<int>(t % dim);
  const float lo = ordered_value(sc->lo[q]), hi = ordered_value(sc->hi[q]);
  Y[t] = 10.0f * (Y[t] - lo) / (hi - lo + 1e-10f);
}

}  // namespace umap
}  // namespace mv

using namespace mv;
using namespace mv::umap;

extern "C" {

int mv_umap_graph_ws_bytes(int64_t n, int k, size_t* out_bytes) {
  MV_REQUIRE(out_bytes, "mv_umap_graph_ws_bytes: out_bytes is NULL");
  MV_REQUIRE(n >= 1 && k >= 1 && k <= KMAX_GRAPH, "mv_umap_graph_ws_bytes: bad n=%lld k=%d", (long long)n, k);
  *out_bytes = align_up(static_cast<size_t>(n) * k * sizeof(float), 256) + 256;
  return MV_OK;
}

int mv_umap_fuzzy_graph(const int32_t* knn_idx, const float* knn_dist, int64_t n, int k, int n_epochs,
                        float* w_sym, float* eps, float* epn, uint32_t* counters, float* sigma_out,
                        float* rho_out, void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(knn_idx && knn_dist && w_sym && eps && epn && counters, "mv_umap_fuzzy_graph: null pointer");
  MV_REQUIRE(n >= 1 && k >= 1 && k <= KMAX_GRAPH, "mv_umap_fuzzy_graph: bad n=%lld k=%d (k <= %d)", (long long)n, k,
             KMAX_GRAPH);
  MV_REQUIRE(n_epochs >= 1, "mv_umap_fuzzy_graph: n_epochs=%d < 1", n_epochs);
  size_t need = 0;
  mv_umap_graph_ws_bytes(n, k, &need);
  if (!ws || ws_bytes < need) {
    mv::set_error("mv_umap_fuzzy_graph: workspace of %zu bytes < required %zu", ws_bytes, need);
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  float* w = static_cast<float*>(ws);
  const int64_t count = n * k;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, counters, 2);
  MV_LAUNCHED("k_zero_u32");
  if (k <= KMAX)
    hipLaunchKernelGGL(k_fuzzy_rows, dim3(static_cast<unsigned>(ceil_div(n, 256))), dim3(256), 0, s, knn_dist, n, k,
                       static_cast<float>(log2(static_cast<double>(k))), w, sigma_out, rho_out);
  else
    hipLaunchKernelGGL(k_fuzzy_rows_wave, dim3(static_cast<unsigned>(ceil_div(n, 4))), dim3(256), 0, s, knn_dist, n,
                       k, static_cast<float>(log2(static_cast<double>(k))), w, sigma_out, rho_out);
  MV_LAUNCHED("k_fuzzy_rows");
  const int64_t fgrid = ceil_div(count, 256) < 4096 ? ceil_div(count, 256) : 4096;
  hipLaunchKernelGGL(k_fuzzy_sym, dim3(static_cast<unsigned>(fgrid)), dim3(256), 0, s, knn_idx, w, n, k, w_sym,
                     counters + 1);
  MV_LAUNCHED("k_fuzzy_sym");
  hipLaunchKernelGGL(k_fuzzy_sched, dim3(static_cast<unsigned>(fgrid)), dim3(256), 0, s, w_sym, count, n_epochs,
                     counters + 1, eps, epn, counters);
  MV_LAUNCHED("k_fuzzy_sched");
  return MV_OK;
}

int mv_umap_in_edges(const int32_t* tail, const float* eps, int64_t n, int k, int32_t* in_slots,
                     int32_t* in_node, void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(tail && eps && in_slots && in_node, "mv_umap_in_edges: null pointer");
  MV_REQUIRE(n >= 1 && k >= 1 && k <= KMAX_GRAPH, "mv_umap_in_edges: bad n=%lld k=%d", (long long)n, k);
  const int64_t nb = ceil_div(n, 1024);
  const size_t need = 3 * align_up(static_cast<size_t>(n) * 4, 256) + align_up(static_cast<size_t>(nb) * 4, 256);
  if (!ws || ws_bytes < need) {
    mv::set_error("mv_umap_in_edges: workspace of %zu bytes < required %zu", ws_bytes, need);
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Carve c(ws);
  uint32_t* cnt = c.take<uint32_t>(n);
  uint32_t* off = c.take<uint32_t>(n);
  uint32_t* cur = c.take<uint32_t>(n);
  uint32_t* sums = c.take<uint32_t>(nb);
  const int64_t m = n * k;
  const unsigned bn = static_cast<unsigned>(ceil_div(n, 256));
  const unsigned bm = static_cast<unsigned>(ceil_div(m, 256));
  hipLaunchKernelGGL(k_zero_u32, dim3(bn), dim3(256), 0, s, cnt, static_cast<int>(n));
  MV_LAUNCHED("k_zero_u32");
  hipLaunchKernelGGL(k_zero_u32, dim3(bn), dim3(256), 0, s, cur, static_cast<int>(n));
  MV_LAUNCHED("k_zero_u32");
  hipLaunchKernelGGL(k_in_count, dim3(bm), dim3(256), 0, s, tail, eps, m, cnt);
  MV_LAUNCHED("k_in_count");
  hipLaunchKernelGGL(k_scan_block_u32, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, cnt, n, off, sums);
  MV_LAUNCHED("k_scan_block_u32");
  hipLaunchKernelGGL(k_scan_sums_u32, dim3(1), dim3(256), 0, s, sums, nb);
  MV_LAUNCHED("k_scan_sums_u32");
  hipLaunchKernelGGL(k_scan_add_u32, dim3(bn), dim3(256), 0, s, off, n, sums);
  MV_LAUNCHED("k_scan_add_u32");
  hipLaunchKernelGGL(k_in_fill, dim3(bm), dim3(256), 0, s, tail, eps, m, off, cur, in_slots, in_node);
  MV_LAUNCHED("k_in_fill");
  hipLaunchKernelGGL(k_in_sort, dim3(static_cast<unsigned>(ceil_div(n, 4))), dim3(256), 0, s, cnt, off, n, in_slots);
  MV_LAUNCHED("k_in_sort");
  hipLaunchKernelGGL(k_in_sort_big, dim3(static_cast<unsigned>(std::min<int64_t>(ceil_div(n, 256), 1024))), dim3(256),
                     0, s, cnt, off, n, in_slots);
  MV_LAUNCHED("k_in_sort_big");
  return MV_OK;
}

int mv_umap_epochs_ws_bytes(int64_t n, int k, int dim, size_t* out_bytes) {
  MV_REQUIRE(out_bytes, "mv_umap_epochs_ws_bytes: out_bytes is NULL");
  MV_REQUIRE(n >= 1 && k >= 1 && k <= KMAX_GRAPH && dim >= 1 && dim <= DIM_MAX, "mv_umap_epochs_ws_bytes: bad sizes");
  *out_bytes = epochs_ws(n, k, dim);
  return MV_OK;
}

int mv_umap_epochs_masked(float* Y, int64_t n, int dim, const int32_t* tail, int k, const float* eps, float* epn,
                          const int32_t* in_slots, const int32_t* in_node, int64_t n_in, int e0, int e1,
                          int n_epochs, int neg_rate, float lr, float a, float b, uint64_t seed,
                          const uint8_t* movable, void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(Y && tail && eps && epn, "mv_umap_epochs: null pointer");
  MV_REQUIRE(n_in == 0 || (in_slots && in_node), "mv_umap_epochs: null in-edge list");
  MV_REQUIRE(n >= 1 && k >= 1 && k <= KMAX_GRAPH, "mv_umap_epochs: bad n=%lld k=%d", (long long)n, k);
  MV_REQUIRE(n_in >= 0 && n_in <= n * k, "mv_umap_epochs: bad in-edge count %lld", (long long)n_in);
  MV_REQUIRE(dim >= 1 && dim <= DIM_MAX, "mv_umap_epochs: dim=%d outside [1, %d]", dim, DIM_MAX);
  MV_REQUIRE(n * k < (1ll << 31), "mv_umap_epochs: n * k = %lld slots, at most 2^31 - 1", (long long)(n * k));
  MV_REQUIRE(0 <= e0 && e0 <= e1 && e1 <= n_epochs, "mv_umap_epochs: bad epoch range [%d, %d) of %d", e0, e1,
             n_epochs);
  MV_REQUIRE(neg_rate >= 0, "mv_umap_epochs: negative_sample_rate < 0");
  if (!ws || ws_bytes < epochs_ws(n, k, dim)) {
    mv::set_error("mv_umap_epochs: workspace of %zu bytes < required %zu", ws_bytes, epochs_ws(n, k, dim));
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  EpochBufs B = carve_epochs(ws, n, k, dim);
  if (n_in > 0) {
    hipLaunchKernelGGL(k_in_prep, dim3(static_cast<unsigned>(ceil_div(n_in, 256))), dim3(256), 0, s, in_slots, n_in, k,
                       eps, epn, B.in_head, B.eps_in, B.epn_in);
    MV_LAUNCHED("k_in_prep");
  }
  hipLaunchKernelGGL(k_fill_i32, dim3(static_cast<unsigned>(ceil_div(n, 256))), dim3(256), 0, s, B.span, n, -1);
  MV_LAUNCHED("k_fill_i32");
  // crossings and fixed-point accumulators start at zero in every call
  hipLaunchKernelGGL(k_zero_u32, dim3(static_cast<unsigned>(ceil_div(n, 256))), dim3(256), 0, s,
                     reinterpret_cast<uint32_t*>(B.span_w), static_cast<int>(n));
  MV_LAUNCHED("k_zero_u32");
  hipLaunchKernelGGL(k_zero_u32, dim3(static_cast<unsigned>(ceil_div(2 * n * dim, 256))), dim3(256), 0, s,
                     reinterpret_cast<uint32_t*>(B.hub_acc), static_cast<int>(2 * n * dim));
  MV_LAUNCHED("k_zero_u32");
  if (n_in > TAIL_CHUNK) {
    hipLaunchKernelGGL(k_span, dim3(static_cast<unsigned>(ceil_div(ceil_div(n_in, TAIL_CHUNK), 256))), dim3(256), 0, s,
                       in_node,
                       n_in, B.span, B.span_w);
    MV_LAUNCHED("k_span");
  }
  // positions alternate between Y and ybuf; an odd number of epochs ends in ybuf
  float* cur = Y;
  float* nxt = B.ybuf;
  for (int ep = e0; ep < e1; ++ep) {
    Law law;
    law.a = a;
    law.b = b;
    law.alpha = static_cast<float>(static_cast<double>(lr) * (1.0 - static_cast<double>(ep) / n_epochs));
    int rc = MV_OK;
    switch (dim) {
      case 1: rc = run_epoch<1>(cur, nxt, B, tail, eps, epn, in_node, n_in, n, k, ep, neg_rate, law, seed, movable, s); break;
      case 2: rc = run_epoch<2>(cur, nxt, B, tail, eps, epn, in_node, n_in, n, k, ep, neg_rate, law, seed, movable, s); break;
      case 3: rc = run_epoch<3>(cur, nxt, B, tail, eps, epn, in_node, n_in, n, k, ep, neg_rate, law, seed, movable, s); break;
      case 4: rc = run_epoch<4>(cur, nxt, B, tail, eps, epn, in_node, n_in, n, k, ep, neg_rate, law, seed, movable, s); break;
      case 5: rc = run_epoch<5>(cur, nxt, B, tail, eps, epn, in_node, n_in, n, k, ep, neg_rate, law, seed, movable, s); break;
      case 6: rc = run_epoch<6>(cur, nxt, B, tail, eps, epn, in_node, n_in, n, k, ep, neg_rate, law, seed, movable, s); break;
      case 7: rc = run_epoch<7>(cur, nxt, B, tail, eps, epn, in_node, n_in, n, k, ep, neg_rate, law, seed, movable, s); break;
      default: rc = run_epoch<8>(cur, nxt, B, tail, eps, epn, in_node, n_in, n, k, ep, neg_rate, law, seed, movable, s); break;
    }
    if (rc != MV_OK) return rc;
    float* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (cur != Y) {
    hipLaunchKernelGGL(k_copy_f32, dim3(static_cast<unsigned>(ceil_div(n * dim, 256))), dim3(256), 0, s, Y, cur,
                       n * dim);
    MV_LAUNCHED("k_copy_f32");
  }
  return MV_OK;
}

int mv_umap_epochs(float* Y, int64_t n, int dim, const int32_t* tail, int k, const float* eps, float* epn,
                   const int32_t* in_slots, const int32_t* in_node, int64_t n_in, int e0, int e1, int n_epochs,
                   int neg_rate, float lr, float a, float b, uint64_t seed, void* ws, size_t ws_bytes,
                   void* stream) {
  return mv_umap_epochs_masked(Y, n, dim, tail, k, eps, epn, in_slots, in_node, n_in, e0, e1, n_epochs, neg_rate, lr,
                               a, b, seed, nullptr, ws, ws_bytes, stream);
}

int mv_umap_spectral_ws_bytes(int64_t n, int k, int nc, size_t* out_bytes) {
  MV_REQUIRE(out_bytes, "mv_umap_spectral_ws_bytes: out_bytes is NULL");
  MV_REQUIRE(n >= 1 && k >= 1 && nc >= 2 && nc <= NC_MAX, "mv_umap_spectral_ws_bytes: bad n=%lld k=%d nc=%d",
             (long long)n, k, nc);
  *out_bytes = spectral_ws(n, k, nc);
  return MV_OK;
}

int mv_umap_spectral(const float* w_sym, const float* eps, const int32_t* tail, int64_t n, int k, int nc,
                     int iters, float* V, void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(w_sym && eps && tail && V, "mv_umap_spectral: null pointer");
  MV_REQUIRE(n >= 1 && k >= 1 && k <= KMAX_GRAPH && nc >= 2 && nc <= NC_MAX && iters >= 0,
             "mv_umap_spectral: bad n=%lld k=%d nc=%d iters=%d", (long long)n, k, nc, iters);
  if (!ws || ws_bytes < spectral_ws(n, k, nc)) {
    mv::set_error("mv_umap_spectral: workspace of %zu bytes < required %zu", ws_bytes, spectral_ws(n, k, nc));
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* w = static_cast<char*>(ws);
  switch (nc) {
    case 2: return spectral_run<2>(w_sym, eps, tail, n, k, iters, V, w, s);
    case 3: return spectral_run<3>(w_sym, eps, tail, n, k, iters, V, w, s);
    case 4: return spectral_run<4>(w_sym, eps, tail, n, k, iters, V, w, s);
    case 5: return spectral_run<5>(w_sym, eps, tail, n, k, iters, V, w, s);
    case 6: return spectral_run<6>(w_sym, eps, tail, n, k, iters, V, w, s);
    case 7: return spectral_run<7>(w_sym, eps, tail, n, k, iters, V, w, s);
    case 8: return spectral_run<8>(w_sym, eps, tail, n, k, iters, V, w, s);
    default: return spectral_run<9>(w_sym, eps, tail, n, k, iters, V, w, s);
  }
}

int mv_umap_spectral_embed(const float* V, int64_t n, int nc, uint64_t seed, float* Y, int32_t* out_nonfinite,
                           void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(V && Y, "mv_umap_spectral_embed: null pointer");
  MV_REQUIRE(n >= 1 && nc >= 2 && nc <= NC_MAX, "mv_umap_spectral_embed: bad n=%lld nc=%d", (long long)n, nc);
  if (!ws || ws_bytes < sizeof(EmbedScal)) {
    mv::set_error("mv_umap_spectral_embed: workspace of %zu bytes < required %zu", ws_bytes, sizeof(EmbedScal));
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  EmbedScal* sc = static_cast<EmbedScal*>(ws);
  const int dim = nc - 1;
  const int64_t count = n * dim;
  const unsigned red_grid = static_cast<unsigned>(std::min<int64_t>(ceil_div(count, 256), 1024));
  hipLaunchKernelGGL(k_embed_init, dim3(1), dim3(64), 0, s, sc);
  MV_LAUNCHED("k_embed_init");
  hipLaunchKernelGGL(k_random_normal, dim3(static_cast<unsigned>(ceil_div(ceil_div(count, 4), 256))), dim3(256), 0, s,
                     Y, count, seed);
  MV_LAUNCHED("k_random_normal");
  hipLaunchKernelGGL(k_embed_absmax, dim3(red_grid), dim3(256), 0, s, V, n, nc, sc);
  MV_LAUNCHED("k_embed_absmax");
  hipLaunchKernelGGL(k_embed_scale, dim3(red_grid), dim3(256), 0, s, V, n, nc, Y, sc);
  MV_LAUNCHED("k_embed_scale");
  hipLaunchKernelGGL(k_embed_norm, dim3(static_cast<unsigned>(ceil_div(count, 256))), dim3(256), 0, s, Y, n, dim, sc);
  MV_LAUNCHED("k_embed_norm");
  if (out_nonfinite) MV_TRY_HIP(hipMemcpyAsync(out_nonfinite, &sc->nonfinite, sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  return MV_OK;
}

int mv_random_normal(float* out, int64_t count, uint64_t seed, void* stream) {
  MV_REQUIRE(out || count == 0, "mv_random_normal: null pointer");
  MV_REQUIRE(count >= 0, "mv_random_normal: count < 0");
  MV_FAULT_POINT("launch");
  if (count == 0) return MV_OK;
  hipLaunchKernelGGL(k_random_normal, dim3(static_cast<unsigned>(ceil_div(ceil_div(count, 4), 256))), dim3(256), 0,
                     static_cast<hipStream_t>(stream), out, count, seed);
  MV_LAUNCHED("k_random_normal");
  return MV_OK;
}

int mv_zero_f32(float* p, int64_t count, void* stream) {
  MV_REQUIRE(p || count == 0, "mv_zero_f32: null pointer");
  MV_FAULT_POINT("launch");
  if (count == 0) return MV_OK;
  hipLaunchKernelGGL(k_zero_f32, dim3(static_cast<unsigned>(ceil_div(count, 256))), dim3(256), 0,
                     static_cast<hipStream_t>(stream), p, count);
  MV_LAUNCHED("k_zero_f32");
  return MV_OK;
}

int mv_umap_sgd_step(float* Y_out, const float* Y_in, int64_t n, int dim, const int32_t* ef, const int32_t* et,
                     int64_t A, const int32_t* neg_from, const int32_t* neg_to, int64_t n_neg, float alpha,
                     float a, float b, void* stream) {
  MV_REQUIRE(Y_out && Y_in, "mv_umap_sgd_step: null embedding pointer");
  MV_REQUIRE(A == 0 || (ef && et), "mv_umap_sgd_step: null edge list");
  MV_REQUIRE(n_neg == 0 || (neg_from && neg_to), "mv_umap_sgd_step: null negative list");
  MV_REQUIRE(n >= 1 && dim >= 1 && A >= 0 && n_neg >= 0, "mv_umap_sgd_step: bad sizes");
  MV_REQUIRE(Y_out != Y_in, "mv_umap_sgd_step: Y_out must not alias Y_in");
  MV_FAULT_POINT("launch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t count = n * dim;
  hipLaunchKernelGGL(k_copy_f32, dim3(static_cast<unsigned>(ceil_div(count, 256))), dim3(256), 0, s, Y_out, Y_in,
                     count);
  MV_LAUNCHED("k_copy_f32");
  Law law;
  law.a = a;
  law.b = b;
  law.alpha = alpha;
  const int64_t total = A + n_neg;
  if (total > 0) {
    hipLaunchKernelGGL(k_sgd_step, dim3(static_cast<unsigned>(ceil_div(total, 256))), dim3(256), 0, s, Y_in, Y_out,
                       n, dim, ef, et, A, neg_from, neg_to, n_neg, law);
    MV_LAUNCHED("k_sgd_step");
  }
  return MV_OK;
}

}  // extern "C"
