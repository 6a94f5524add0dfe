// Trustworthiness ranks for gfx950 (Venna & Kaski 2001).
//
// T = 1 - 2 / (m k (2n - 3k - 1)) * sum_i sum_{j in lo(i)} max(r(i, j) - k, 0)
// where lo(i) are the embedding's k neighbours of point i and r(i, j) the rank
// of j among i's neighbours in the input space (1 = nearest). The reference
// defines no metric; this is the textbook statistic (sklearn.manifold.
// trustworthiness computes the same sum from a full argsort).
//
// A rank needs the input-space distance from i to every other point, so the
// work is the full m x n distance matrix, counted and never stored:
//   r(i, j) = 1 + #{ l != i, l != j : D(i, l) < D(i, j)
//                                     or (D(i, l) == D(i, j) and l < j) },
// i.e. the position of j in the order (distance, index) that a stable
// argsort gives. D(i, l) = sum_f (x_i[f] - x_l[f])^2 in FP32, one fmaf per
// feature in feature order; the target distances D(i, j) are computed first
// by the same sequence, so they equal what the scan computes for l = j.
//
// Kernels:
//   k_trust_targets  D(i, j) for the m x k targets (thread per target).
//   k_trust_count    64 query rows x 64 candidate rows per tile, 16-feature
//                    chunks through LDS, a 4 x 4 register tile per thread
//                    (VALU: the subtraction before the square keeps the
//                    distances free of cancellation); after each tile every
//                    target's count of closer candidates, reduced over the 16
//                    lanes sharing a query row with cross-lane swizzles, into
//                    LDS; one global atomic add per (row, target) per block.
//   k_trust_finish   per-row sum of max(r - k, 0).
#include <algorithm>
#include <math.h>

#include "common.hpp"

namespace mv {
namespace metrics {

constexpr int TQ = 64;    // query rows per tile
constexpr int TC = 64;    // candidate rows per tile
constexpr int KC = 16;    // features per LDS chunk
constexpr int KMAX = 64;  // targets per row
constexpr int NT = 256;

__global__ __launch_bounds__(256) void k_trust_targets(const float* __restrict__ X, int64_t n, int64_t d,
                                                       const int32_t* __restrict__ rows, int64_t m,
                                                       const int32_t* __restrict__ lo, int64_t ldl, int k,
                                                       float* __restrict__ T) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m * k) return;
  const int64_t q = e / k;
  const int t = static_cast<int>(e - q * k);
  const int64_t i = rows ? rows[q] : q;
  const int64_t j = lo[q * ldl + t];
  if (j < 0 || j >= n || i < 0 || i >= n) {
    T[e] = -1.f;   // nothing is closer than -1: the target counts nothing
    return;
  }
  const float* xi = X + i * d;
  const float* xj = X + j * d;
  float acc = 0.f;
  for (int64_t f = 0; f < d; ++f) {
    const float df = xi[f] - xj[f];
    acc = fmaf(df, df, acc);
  }
  T[e] = acc;
}

// Sum over the 16 lanes of a 16-lane group (a DPP row): xor swizzles 8, 4, 2, 1.
__device__ __forceinline__ int row16_sum(int v) {
  v += __shfl_xor(v, 8, 16);
  v += __shfl_xor(v, 4, 16);
  v += __shfl_xor(v, 2, 16);
  v += __shfl_xor(v, 1, 16);
  return v;
}

// grid (ceil(m / 64), S): block (bx, by) takes query tile bx against the
// candidate tiles by, by + S, ...; dynamic LDS: 3 * 64 * k words.
__global__ __launch_bounds__(NT) void k_trust_count(const float* __restrict__ X, int64_t n, int64_t d,
                                                    const int32_t* __restrict__ rows, int64_t m,
                                                    const int32_t* __restrict__ lo, int64_t ldl, int k,
                                                    const float* __restrict__ T, int32_t* __restrict__ cnt) {
  __shared__ __align__(16) float sA[KC][TQ];
  __shared__ __align__(16) float sB[KC][TC];
  __shared__ int64_t sI[TQ];
  extern __shared__ int dyn[];
  float* sT = reinterpret_cast<float*>(dyn);   // [TQ][k]
  int* sJ = dyn + TQ * k;                       // [TQ][k]
  int* sC = dyn + 2 * TQ * k;                   // [TQ][k]

  const int tid = threadIdx.x;
  const int ty = tid >> 4, tx = tid & 15;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * TQ;
  for (int r = tid; r < TQ; r += NT) {
    const int64_t q = q0 + r;
    const int64_t i = q < m ? (rows ? static_cast<int64_t>(rows[q]) : q) : -1;
    sI[r] = (i >= 0 && i < n) ? i : -1;   // a row outside [0, n) counts nothing
  }
  for (int e = tid; e < TQ * k; e += NT) {
    const int r = e / k, t = e - r * k;
    const int64_t q = q0 + r;
    if (q < m) {
      sT[e] = T[q * k + t];
      sJ[e] = lo[q * ldl + t];
    } else {
      sT[e] = -1.f;
      sJ[e] = -1;
    }
    sC[e] = 0;
  }
  __syncthreads();

  // loader role: row lr of the tile, features lf .. lf + 3 of the chunk
  const int lr = tid >> 2, lf = (tid & 3) * 4;
  const int64_t li = sI[lr];
  const float* qa = X + (li >= 0 ? li : 0) * d;
  const int64_t ctiles = (n + TC - 1) / TC;

  for (int64_t ct = blockIdx.y; ct < ctiles; ct += gridDim.y) {
    const int64_t c0 = ct * TC;
    const int64_t lc = c0 + lr;
    const float* cb = X + (lc < n ? lc : 0) * d;
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;

    for (int64_t f0 = 0; f0 < d; f0 += KC) {
      __syncthreads();
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t f = f0 + lf + e;
        sA[lf + e][lr] = (f < d && li >= 0) ? qa[f] : 0.f;
        sB[lf + e][lr] = (f < d && lc < n) ? cb[f] : 0.f;
      }
      __syncthreads();
#pragma unroll
      for (int f = 0; f < KC; ++f) {
        const float4 av = *reinterpret_cast<const float4*>(&sA[f][ty * 4]);
        const float4 bv = *reinterpret_cast<const float4*>(&sB[f][tx * 4]);
        const float ar[4] = {av.x, av.y, av.z, av.w};
        const float br[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const float df = ar[a] - br[b];
            acc[a][b] = fmaf(df, df, acc[a][b]);
          }
      }
    }

    // count, for each of this thread's 4 query rows and every target, the
    // candidates of its 4 columns that precede the target
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int r = ty * 4 + a;
      const int64_t i = sI[r];
      for (int t = 0; t < k; ++t) {
        const float tv = sT[r * k + t];
        const int64_t j = sJ[r * k + t];
        int p = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int64_t l = c0 + tx * 4 + b;
          const float dv = acc[a][b];
          const bool before = dv < tv || (dv == tv && l < j);
          p += (l < n && l != i && l != j && before) ? 1 : 0;
        }
        p = row16_sum(p);
        if (tx == 0) sC[r * k + t] += p;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < TQ * k; e += NT) {
    const int r = e / k, t = e - r * k;
    const int64_t q = q0 + r;
    if (q < m && sC[e] != 0) atomicAdd(&cnt[q * k + t], sC[e]);
  }
}

__global__ __launch_bounds__(256) void k_trust_finish(const int32_t* __restrict__ rows, int64_t m,
                                                      const int32_t* __restrict__ lo, int64_t ldl, int k,
                                                      int64_t n, const int32_t* __restrict__ cnt,
                                                      int64_t* __restrict__ out) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const int64_t i = rows ? rows[q] : q;
  int64_t s = 0;
  for (int t = 0; t < k; ++t) {
    const int64_t j = lo[q * ldl + t];
    if (j < 0 || j >= n) continue;
    // the point itself sorts last (its distance is excluded as +inf): rank n
    const int64_t r = (j == i) ? n : 1 + static_cast<int64_t>(cnt[q * k + t]);
    s += r > k ? r - k : 0;
  }
  out[q] = s;
}

__global__ void k_zero_i32(int32_t* __restrict__ p, int64_t count) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < count) p[e] = 0;
}

size_t trust_ws(int64_t m, int k) {
  const size_t a = static_cast<size_t>(m) * k;
  return ((a * sizeof(float) + 255) / 256) * 256 + ((a * sizeof(int32_t) + 255) / 256) * 256;
}

}  // namespace metrics
}  // namespace mv

using namespace mv;
using namespace mv::metrics;

extern "C" {

int mv_trust_ws_bytes(int64_t m, int k, size_t* out) {
  MV_REQUIRE(out, "mv_trust_ws_bytes: null output");
  MV_REQUIRE(m >= 0 && k >= 1 && k <= KMAX, "mv_trust_ws_bytes: need m >= 0 and 1 <= k <= %d", KMAX);
  *out = trust_ws(m, k);
  return MV_OK;
}

int mv_trust_excess(const float* X, int64_t n, int64_t d, const int32_t* rows, int64_t m, const int32_t* lo,
                    int64_t ldl, int k, int64_t* out_excess, void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(X && lo && out_excess, "mv_trust_excess: null pointer");
  MV_REQUIRE(n >= 1 && d >= 1 && m >= 0 && k >= 1 && k <= KMAX && ldl >= k,
             "mv_trust_excess: need n, d >= 1, m >= 0, 1 <= k <= %d, ldl >= k", KMAX);
  MV_REQUIRE(n <= INT32_MAX, "mv_trust_excess: n must fit int32 indices");
  MV_REQUIRE(rows != nullptr || m <= n, "mv_trust_excess: without rows, m must not exceed n");
  const size_t need = trust_ws(m, k);
  MV_REQUIRE(ws != nullptr || need == 0 || m == 0, "mv_trust_excess: null workspace");
  MV_REQUIRE(ws_bytes >= need, "mv_trust_excess: workspace %zu bytes < %zu", ws_bytes, need);
  MV_FAULT_POINT("launch");
  if (m == 0) return MV_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t mk = m * k;
  float* T = static_cast<float*>(ws);
  int32_t* cnt = reinterpret_cast<int32_t*>(static_cast<char*>(ws) + ((mk * sizeof(float) + 255) / 256) * 256);
  hipLaunchKernelGGL(k_zero_i32, dim3(static_cast<unsigned>(ceil_div(mk, 256))), dim3(256), 0, s, cnt, mk);
  MV_LAUNCHED("k_zero_i32");
  hipLaunchKernelGGL(k_trust_targets, dim3(static_cast<unsigned>(ceil_div(mk, 256))), dim3(256), 0, s, X, n, d, rows,
                     m, lo, ldl, k, T);
  MV_LAUNCHED("k_trust_targets");
  const int64_t qt = ceil_div(m, TQ), ctiles = ceil_div(n, TC);
  MV_REQUIRE(qt <= INT32_MAX, "mv_trust_excess: too many query rows");
  const int64_t S = std::max<int64_t>(1, std::min<int64_t>(ctiles, std::min<int64_t>(65535, ceil_div(4096, qt))));
  const size_t lds = static_cast<size_t>(3) * TQ * k * sizeof(int);
  hipLaunchKernelGGL(k_trust_count, dim3(static_cast<unsigned>(qt), static_cast<unsigned>(S)), dim3(NT), lds, s, X,
                     n, d, rows, m, lo, ldl, k, T, cnt);
  MV_LAUNCHED("k_trust_count");
  hipLaunchKernelGGL(k_trust_finish, dim3(static_cast<unsigned>(ceil_div(m, 256))), dim3(256), 0, s, rows, m, lo,
                     ldl, k, n, cnt, out_excess);
  MV_LAUNCHED("k_trust_finish");
  return MV_OK;
}

}  // extern "C"
