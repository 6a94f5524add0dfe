// PCA preprocessing of UMAP (mlx_vis/_umap/umap.py:92-113) on the device:
// the column means, the covariance (X - mean)^T (X - mean) / (n - 1) and the
// projection (X - mean) V onto the leading eigenvectors. The eigen-
// decomposition of the d x d covariance runs on the host, as in the
// reference (mx.linalg.eigh on the CPU stream, umap.py:100).
//
//   k_pca_colsum     FP64 partial column sums over row chunks
//   k_pca_mean       their FP64 sum / n
//   k_pca_cov        the covariance's tiles on and above the diagonal
//                    (64 x 64 per workgroup, 4 x 4 per thread, the centred
//                    operands staged through LDS 16 rows at a time), one FP32
//                    partial per row chunk (split K)
//   k_pca_cov_sum    the partials summed in FP64 in chunk order (the result
//                    does not depend on scheduling), / (n - 1), mirrored
//   k_pca_project    (X - mean) V, the same tiling, centring fused
#include "common.hpp"

namespace mv {
namespace pca {

constexpr int T = 64;           // covariance / projection tile
constexpr int KS = 16;          // rows (K) staged per step
constexpr int64_t ROWS_PER = 1024;   // rows per covariance partial (chunk)
constexpr int MAX_CHUNKS = 64;

__global__ __launch_bounds__(256) void k_pca_colsum(const float* __restrict__ X, int64_t n, int64_t d,
                                                    int64_t rows_per, double* __restrict__ part) {
  const int64_t col = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y;
  if (col >= d) return;
  const int64_t r0 = c * rows_per, r1 = r0 + rows_per < n ? r0 + rows_per : n;
  double s = 0.0;
  for (int64_t r = r0; r < r1; ++r) s += static_cast<double>(X[r * d + col]);
  part[c * d + col] = s;
}

__global__ __launch_bounds__(256) void k_pca_mean(const double* __restrict__ part, int64_t nchunks, int64_t n,
                                                  int64_t d, double* __restrict__ mean, float* __restrict__ mean32) {
  const int64_t col = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (col >= d) return;
  double s = 0.0;
  for (int64_t c = 0; c < nchunks; ++c) s += part[c * d + col];
  const double m = s / static_cast<double>(n);
  mean[col] = m;
  mean32[col] = static_cast<float>(m);
}

// Tile (I, J), I <= J, of rows [r0, r1): part[chunk][tile] (T x T FP32).
__global__ __launch_bounds__(256) void k_pca_cov(const float* __restrict__ X, int64_t n, int64_t d,
                                                 const float* __restrict__ mean32, int nt, int64_t rows_per,
                                                 float* __restrict__ part) {
  __shared__ float As[KS][T + 4];
  __shared__ float Bs[KS][T + 4];
  // tile index over the upper triangle, row-major
  int t = blockIdx.x, I = 0;
  while (t >= nt - I) {
    t -= nt - I;
    ++I;
  }
  const int J = I + t;
  const int64_t chunk = blockIdx.y;
  const int64_t r0 = chunk * rows_per;
  const int64_t r1 = r0 + rows_per < n ? r0 + rows_per : n;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int64_t ca = static_cast<int64_t>(I) * T, cb = static_cast<int64_t>(J) * T;
  float acc[4][4] = {};
  for (int64_t k0 = r0; k0 < r1; k0 += KS) {
    for (int e = threadIdx.x; e < KS * T; e += 256) {
      const int kk = e / T, cc = e % T;
      const int64_t r = k0 + kk;
      const int64_t ga = ca + cc, gb = cb + cc;
      As[kk][cc] = (r < r1 && ga < d) ? X[r * d + ga] - mean32[ga] : 0.f;
      Bs[kk][cc] = (r < r1 && gb < d) ? X[r * d + gb] - mean32[gb] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
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
  float* out = part + (chunk * gridDim.x + blockIdx.x) * static_cast<int64_t>(T * T);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) out[(ty * 4 + i) * T + tx * 4 + j] = acc[i][j];
}

__global__ __launch_bounds__(256) void k_pca_cov_sum(const float* __restrict__ part, int64_t nchunks, int ntiles,
                                                     int nt, int64_t n, int64_t d, double* __restrict__ cov) {
  const int tile = blockIdx.x;
  int t = tile, I = 0;
  while (t >= nt - I) {
    t -= nt - I;
    ++I;
  }
  const int J = I + t;
  const double inv = 1.0 / static_cast<double>(n - 1 > 0 ? n - 1 : 1);
  for (int e = threadIdx.x; e < T * T; e += blockDim.x) {
    const int64_t gi = static_cast<int64_t>(I) * T + e / T, gj = static_cast<int64_t>(J) * T + e % T;
    if (gi >= d || gj >= d) continue;
    double s = 0.0;
    for (int64_t c = 0; c < nchunks; ++c) s += static_cast<double>(part[(c * ntiles + tile) * T * T + e]);
    s *= inv;
    cov[gi * d + gj] = s;
    cov[gj * d + gi] = s;
  }
}

// out[n, p] = (X - mean) V, V: d x p FP32 row-major.
__global__ __launch_bounds__(256) void k_pca_project(const float* __restrict__ X, int64_t n, int64_t d,
                                                     const float* __restrict__ mean32, const float* __restrict__ V,
                                                     int64_t p, float* __restrict__ out) {
  __shared__ float As[KS][T + 4];
  __shared__ float Bs[KS][T + 4];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int64_t r0 = static_cast<int64_t>(blockIdx.y) * T;
  const int64_t c0 = static_cast<int64_t>(blockIdx.x) * T;
  float acc[4][4] = {};
  for (int64_t k0 = 0; k0 < d; k0 += KS) {
    for (int e = threadIdx.x; e < KS * T; e += 256) {
      const int kk = e & (KS - 1), rr = e / KS;
      const int64_t gr = r0 + rr, gk = k0 + kk;
      As[kk][rr] = (gr < n && gk < d) ? X[gr * d + gk] - mean32[gk] : 0.f;
      const int cc = e % T, kb = e / T;
      const int64_t gc = c0 + cc, gk2 = k0 + kb;
      Bs[kb][cc] = (gc < p && gk2 < d) ? V[gk2 * p + gc] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
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
      if (gc < p) out[gr * p + gc] = acc[i][j];
    }
  }
}

struct Layout {
  int64_t rows_per, nchunks;
  int nt, ntiles;
  size_t bytes;
  double* colpart;
  float* mean32;
  float* part;
};

static Layout carve(void* ws, int64_t n, int64_t d) {
  Layout L;
  L.rows_per = ROWS_PER;
  L.nchunks = ceil_div(n, ROWS_PER);
  if (L.nchunks > MAX_CHUNKS) {
    L.nchunks = MAX_CHUNKS;
    L.rows_per = ceil_div(n, MAX_CHUNKS);
  }
  L.nt = static_cast<int>(ceil_div(d, T));
  L.ntiles = L.nt * (L.nt + 1) / 2;
  Carve c(ws);
  L.colpart = c.take<double>(static_cast<size_t>(L.nchunks) * d);
  L.mean32 = c.take<float>(d);
  L.part = c.take<float>(static_cast<size_t>(L.nchunks) * L.ntiles * T * T);
  L.bytes = c.off + 256;
  return L;
}

}  // namespace pca
}  // namespace mv

using namespace mv;
using namespace mv::pca;

extern "C" {

int mv_pca_ws_bytes(int64_t n, int64_t d, size_t* out_bytes) {
  MV_REQUIRE(out_bytes, "mv_pca_ws_bytes: out_bytes is NULL");
  MV_REQUIRE(n >= 1 && d >= 1 && d <= 8192, "mv_pca_ws_bytes: bad shape n=%lld d=%lld", (long long)n, (long long)d);
  *out_bytes = carve(nullptr, n, d).bytes;
  return MV_OK;
}

int mv_pca_cov(const float* X, int64_t n, int64_t d, double* mean_out, double* cov_out, void* ws, size_t ws_bytes,
               void* stream) {
  MV_REQUIRE(X && mean_out && cov_out, "mv_pca_cov: null pointer");
  MV_REQUIRE(n >= 1 && d >= 1 && d <= 8192, "mv_pca_cov: bad shape n=%lld d=%lld", (long long)n, (long long)d);
  const Layout need = carve(nullptr, n, d);
  if (!ws || ws_bytes < need.bytes) {
    mv::set_error("mv_pca_cov: workspace of %zu bytes < required %zu", ws_bytes, need.bytes);
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  const Layout L = carve(ws, n, d);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned cb = static_cast<unsigned>(ceil_div(d, 256));
  hipLaunchKernelGGL(k_pca_colsum, dim3(cb, static_cast<unsigned>(L.nchunks)), dim3(256), 0, s, X, n, d, L.rows_per,
                     L.colpart);
  MV_LAUNCHED("k_pca_colsum");
  hipLaunchKernelGGL(k_pca_mean, dim3(cb), dim3(256), 0, s, L.colpart, L.nchunks, n, d, mean_out, L.mean32);
  MV_LAUNCHED("k_pca_mean");
  hipLaunchKernelGGL(k_pca_cov, dim3(static_cast<unsigned>(L.ntiles), static_cast<unsigned>(L.nchunks)), dim3(256), 0,
                     s, X, n, d, L.mean32, L.nt, L.rows_per, L.part);
  MV_LAUNCHED("k_pca_cov");
  hipLaunchKernelGGL(k_pca_cov_sum, dim3(static_cast<unsigned>(L.ntiles)), dim3(256), 0, s, L.part, L.nchunks,
                     L.ntiles, L.nt, n, d, cov_out);
  MV_LAUNCHED("k_pca_cov_sum");
  return MV_OK;
}

int mv_pca_project(const float* X, int64_t n, int64_t d, const float* mean, const float* V, int64_t p, float* out,
                   void* stream) {
  MV_REQUIRE(X && mean && V && out, "mv_pca_project: null pointer");
  MV_REQUIRE(n >= 0 && d >= 1 && p >= 1 && p <= d, "mv_pca_project: bad shape n=%lld d=%lld p=%lld", (long long)n,
             (long long)d, (long long)p);
  MV_FAULT_POINT("launch");
  if (n == 0) return MV_OK;
  const dim3 grid(static_cast<unsigned>(ceil_div(p, T)), static_cast<unsigned>(ceil_div(n, T)));
  hipLaunchKernelGGL(k_pca_project, grid, dim3(256), 0, static_cast<hipStream_t>(stream), X, n, d, mean, V, p, out);
  MV_LAUNCHED("k_pca_project");
  return MV_OK;
}

}  // extern "C"
