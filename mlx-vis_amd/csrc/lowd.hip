// Exact kNN in one to three dimensions on a uniform grid (gfx950).
//
// The brute path of compute_knn (mlx_vis/_knn.py:47-88) forms
// |q|^2 + |x|^2 - 2 q.x. In two or three dimensions (the embeddings whose
// neighbours the layout checks and the quality metrics need) that expansion
// cannot separate near neighbours: its rounding error grows with |q|^2 +
// |x|^2, the neighbour gaps with the point density. Here the canonical pair
// distance itself (the sequential, unfused FP64 sum of ((double)q_t -
// (double)x_t)^2, the order every other kNN path of the library produces) is
// computed for every point the search visits, so no certificate or fallback
// is needed:
//
//   1. bounds    per-dimension min / max of X (order-preserving integer atomics)
//   2. setup     cubic cells, about two points per cell (at most n / 2 cells)
//   3. bucket    cell of every point, counts, exclusive scan, fill: points
//                sorted by cell (coordinates as float4, original indices)
//   4. search    one thread per query: rings of cells around the query's cell,
//                r = 0, 1, 2, ...; every visited point enters a sorted
//                register list of (FP64 distance, index) pairs; a cell whose
//                box is farther than the current k-th distance is skipped; the
//                search stops once the k-th distance is below the distance
//                to every unvisited cell (strictly: a tie with a lower index
//                there could still enter).
//
// Box and gap bounds carry a margin of 1e-9 of the coordinate scale, which
// covers the rounding of the cell assignment floor((x - lo) / h), so every
// bound is a true lower bound and the result is exactly the canonical one.
#include <math.h>

#include <algorithm>

#include "common.hpp"
#include "lowd.hpp"

namespace mv {
namespace lowd {

struct Grid {
  double lo[3];
  double h, inv_h;
  double margin;
  int g[3];
  int ncell;
};

// Order-preserving map of a float onto uint32 (NaN not supported).
__device__ __forceinline__ uint32_t f2key(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__global__ void k_bounds_init(uint32_t* mm) {
  if (threadIdx.x < 3) {
    mm[threadIdx.x] = 0xFFFFFFFFu;   // min keys
    mm[3 + threadIdx.x] = 0u;        // max keys
  }
}

__global__ __launch_bounds__(256) void k_bounds(const float* __restrict__ X, int64_t n, int d,
                                                uint32_t* __restrict__ mm) {
  uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    for (int t = 0; t < d; ++t) {
      const uint32_t kk = f2key(X[i * d + t]);
      mn[t] = kk < mn[t] ? kk : mn[t];
      mx[t] = kk > mx[t] ? kk : mx[t];
    }
  }
  for (int t = 0; t < d; ++t) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const uint32_t a = __shfl_xor(mn[t], off), b = __shfl_xor(mx[t], off);
      mn[t] = a < mn[t] ? a : mn[t];
      mx[t] = b > mx[t] ? b : mx[t];
    }
  }
  if ((threadIdx.x & 63) == 0) {
    for (int t = 0; t < d; ++t) {
      atomicMin(&mm[t], mn[t]);
      atomicMax(&mm[3 + t], mx[t]);
    }
  }
}

// Cubic cells of edge h: the largest extent split into G = floor(target^(1/d))
// cells, every other dimension into ceil(extent / h) <= G, so that at most
// target cells exist. A dimension without extent has one cell.
__global__ void k_grid_setup(const uint32_t* __restrict__ mm, int d, int64_t target, Grid* __restrict__ gr) {
  if (threadIdx.x != 0) return;
  double ext[3] = {0.0, 0.0, 0.0}, emax = 0.0, scale = 0.0;
  for (int t = 0; t < 3; ++t) {
    gr->lo[t] = 0.0;
    gr->g[t] = 1;
  }
  for (int t = 0; t < d; ++t) {
    const double lo = key2f(mm[t]), hi = key2f(mm[3 + t]);
    gr->lo[t] = lo;
    ext[t] = hi - lo;
    emax = fmax(emax, ext[t]);
    scale = fmax(scale, fmax(fabs(lo), fabs(hi)));
  }
  int G = static_cast<int>(floor(pow(static_cast<double>(target), 1.0 / d) + 1e-9));
  if (G < 1) G = 1;
  double h = emax > 0.0 ? emax / G : 1.0;
  int ncell = 1;
  for (int t = 0; t < d; ++t) {
    int gt = ext[t] > 0.0 ? static_cast<int>(ceil(ext[t] / h)) : 1;
    gt = gt < 1 ? 1 : (gt > G ? G : gt);
    gr->g[t] = gt;
    ncell *= gt;
  }
  gr->h = h;
  gr->inv_h = 1.0 / h;
  gr->ncell = ncell;
  gr->margin = 1e-9 * (scale + emax + h);
}

template <int D>
__device__ __forceinline__ void cell_of(const Grid& gr, const float* x, int* c) {
#pragma unroll
  for (int t = 0; t < D; ++t) {
    const double f = floor((static_cast<double>(x[t]) - gr.lo[t]) * gr.inv_h);
    int ct = f < 0.0 ? 0 : (f >= gr.g[t] ? gr.g[t] - 1 : static_cast<int>(f));
    c[t] = ct;
  }
}

template <int D>
__device__ __forceinline__ int cell_linear(const Grid& gr, const int* c) {
  int lin = 0;
#pragma unroll
  for (int t = D - 1; t >= 0; --t) lin = lin * gr.g[t] + c[t];
  return lin;
}

template <int D>
__global__ __launch_bounds__(256) void k_cell_count(const float* __restrict__ X, int64_t n,
                                                    const Grid* __restrict__ grp, int32_t* __restrict__ cellid,
                                                    uint32_t* __restrict__ cnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Grid gr = *grp;
  int c[3];
  cell_of<D>(gr, X + i * D, c);
  const int lin = cell_linear<D>(gr, c);
  cellid[i] = lin;
  atomicAdd(&cnt[lin], 1u);
}

// Exclusive scan of cnt[0:count) into off[0:count] (off[count] = total):
// per-block scans of 1,024 entries, a scan of the block sums, the add.
__global__ __launch_bounds__(256) void k_scan1(const uint32_t* __restrict__ in, int64_t count,
                                               uint32_t* __restrict__ out, uint32_t* __restrict__ sums) {
  __shared__ uint32_t sh[256];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * 1024 + threadIdx.x * 4;
  uint32_t v[4], tot = 0;
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

__global__ __launch_bounds__(256) void k_scan2(uint32_t* __restrict__ sums, int64_t nb) {
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
  for (int64_t j = 0; j < per; ++j) {
    if (b0 + j < nb) {
      const uint32_t v = sums[b0 + j];
      sums[b0 + j] = run;
      run += v;
    }
  }
  if (threadIdx.x == 255) sums[nb] = sh[255];   // grand total
}

__global__ __launch_bounds__(256) void k_scan3(uint32_t* __restrict__ out, int64_t count,
                                               const uint32_t* __restrict__ sums, int64_t nb) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) out[t] += sums[t / 1024];
  if (t == count) out[count] = sums[nb];
}

template <int D>
__global__ __launch_bounds__(256) void k_cell_fill(const float* __restrict__ X, int64_t n,
                                                   const int32_t* __restrict__ cellid,
                                                   const uint32_t* __restrict__ off, uint32_t* __restrict__ cursor,
                                                   float4* __restrict__ sx, int32_t* __restrict__ sidx) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cellid[i];
  const uint32_t p = off[c] + atomicAdd(&cursor[c], 1u);
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  v.x = X[i * D];
  if (D > 1) v.y = X[i * D + 1];
  if (D > 2) v.z = X[i * D + 2];
  sx[p] = v;
  sidx[p] = static_cast<int32_t>(i);
}

struct SearchArgs {
  const Grid* grid;
  const float4* sx;        // points sorted by cell
  const int32_t* sidx;     // their original indices
  const uint32_t* off;     // cell offsets (ncell + 1)
  const float* Q;          // queries (nq x D); nullptr: the queries are the points themselves, in cell order
  int64_t nq;
  const int32_t* q_self;   // nullable
  int64_t self_base;       // used when q_self == NULL; < 0: no exclusion
  int k;
  int euclid;
  int32_t* out_idx;
  float* out_dist;
};

__device__ __forceinline__ double coord(const float4& v, int t) { return t == 0 ? v.x : (t == 1 ? v.y : v.z); }

// Canonical squared distance: the sequential, unfused FP64 sum over t.
template <int D>
__device__ __forceinline__ double dist64(const double* q, const float4& x) {
  double acc = 0.0;
#pragma unroll
  for (int t = 0; t < D; ++t) {
    const double df = q[t] - coord(x, t);
    acc = __dadd_rn(acc, __dmul_rn(df, df));
  }
  return acc;
}

// Lower bound of the squared distance from q to any point of cell c (its
// box widened by the margin; outer cells are unbounded outwards).
template <int D>
__device__ __forceinline__ double cell_min_d2(const Grid& gr, const double* q, const int* c) {
  double s = 0.0;
#pragma unroll
  for (int t = 0; t < D; ++t) {
    double gap = 0.0;
    if (c[t] > 0) {
      const double low = gr.lo[t] + c[t] * gr.h - gr.margin;
      if (q[t] < low) gap = low - q[t];
    }
    if (c[t] < gr.g[t] - 1) {
      const double high = gr.lo[t] + (c[t] + 1) * gr.h + gr.margin;
      if (q[t] > high) gap = q[t] - high;
    }
    s += gap * gap;
  }
  return s * (1.0 - 1e-12);
}

template <int D, int KB>
__global__ __launch_bounds__(256) void k_grid_knn(SearchArgs a) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= a.nq) return;
  const Grid gr = *a.grid;
  int64_t row;
  double q[3] = {0.0, 0.0, 0.0};
  float qf[3] = {0.f, 0.f, 0.f};
  if (a.Q) {
    row = t;
#pragma unroll
    for (int u = 0; u < D; ++u) qf[u] = a.Q[t * D + u];
  } else {
    row = a.sidx[t];
    const float4 v = a.sx[t];
    qf[0] = v.x;
    qf[1] = v.y;
    qf[2] = v.z;
  }
#pragma unroll
  for (int u = 0; u < D; ++u) q[u] = qf[u];
  const int64_t excl = a.q_self ? static_cast<int64_t>(a.q_self[row]) : (a.self_base >= 0 ? a.self_base + row : -1);
  int cq[3] = {0, 0, 0};
  cell_of<D>(gr, qf, cq);

  double ld[KB];
  int32_t li[KB];
#pragma unroll
  for (int m = 0; m < KB; ++m) {
    ld[m] = INFINITY;
    li[m] = INT32_MAX;
  }
  const int k = a.k;
  int64_t found = 0;
  int gmax = gr.g[0];
#pragma unroll
  for (int u = 1; u < D; ++u) gmax = gr.g[u] > gmax ? gr.g[u] : gmax;

  for (int r = 0; r <= gmax; ++r) {
    // the cells at Chebyshev distance r from cq (clipped to the grid)
    const int span = 2 * r + 1;
    const int64_t cells = D == 1 ? (r == 0 ? 1 : 2) : (D == 2 ? static_cast<int64_t>(span) * span
                                                                : static_cast<int64_t>(span) * span * span);
    // D == 2, 3: enumerate the cube and keep its surface; D == 3 steps the
    // last coordinate over {-r, r} only inside the surface's interior columns
    for (int64_t e = 0; e < cells; ++e) {
      int off3[3] = {0, 0, 0};
      if (D == 1) {
        off3[0] = r == 0 ? 0 : (e == 0 ? -r : r);
      } else if (D == 2) {
        off3[0] = static_cast<int>(e % span) - r;
        off3[1] = static_cast<int>(e / span) - r;
        if (r > 0 && abs(off3[0]) != r && abs(off3[1]) != r) {
          // interior of the square: jump to the right edge of this row
          e += span - 2 - (off3[0] + r);
          continue;
        }
      } else {
        off3[0] = static_cast<int>(e % span) - r;
        off3[1] = static_cast<int>((e / span) % span) - r;
        off3[2] = static_cast<int>(e / (static_cast<int64_t>(span) * span)) - r;
        if (r > 0 && abs(off3[0]) != r && abs(off3[1]) != r && abs(off3[2]) != r) {
          e += span - 2 - (off3[0] + r);
          continue;
        }
      }
      int c[3];
      bool inside = true;
#pragma unroll
      for (int u = 0; u < D; ++u) {
        c[u] = cq[u] + off3[u];
        inside = inside && c[u] >= 0 && c[u] < gr.g[u];
      }
      if (!inside) continue;
      if (found >= k && cell_min_d2<D>(gr, q, c) > ld[k - 1]) continue;
      const int lin = cell_linear<D>(gr, c);
      const uint32_t p0 = a.off[lin], p1 = a.off[lin + 1];
      for (uint32_t p = p0; p < p1; ++p) {
        const int32_t j = a.sidx[p];
        if (j == excl) continue;
        const double dd = dist64<D>(q, a.sx[p]);
        ++found;
        if (dd < ld[KB - 1] || (dd == ld[KB - 1] && j < li[KB - 1])) {
          double cd = dd;
          int32_t ci = j;
#pragma unroll
          for (int m = 0; m < KB; ++m) {
            const bool less = cd < ld[m] || (cd == ld[m] && ci < li[m]);
            const double td = less ? ld[m] : cd;
            const int32_t ti = less ? li[m] : ci;
            ld[m] = less ? cd : ld[m];
            li[m] = less ? ci : li[m];
            cd = td;
            ci = ti;
          }
        }
      }
    }
    // lower bound of the distance to every cell outside the visited cube
    double gap = INFINITY;
#pragma unroll
    for (int u = 0; u < D; ++u) {
      if (cq[u] - r > 0) gap = fmin(gap, q[u] - (gr.lo[u] + (cq[u] - r) * gr.h + gr.margin));
      if (cq[u] + r + 1 < gr.g[u]) gap = fmin(gap, gr.lo[u] + (cq[u] + r + 1) * gr.h - gr.margin - q[u]);
    }
    if (gap == INFINITY) break;   // every cell visited
    if (found >= k) {
      const double g2 = gap > 0.0 ? gap * gap * (1.0 - 1e-12) : 0.0;
      if (ld[k - 1] < g2) break;
    }
  }
  int32_t* oi = a.out_idx + row * k;
  float* od = a.out_dist + row * k;
#pragma unroll
  for (int m = 0; m < KB; ++m) {
    if (m < k) {
      oi[m] = li[m];
      od[m] = a.euclid ? static_cast<float>(sqrt(ld[m])) : static_cast<float>(ld[m]);
    }
  }
}

struct Buffers {
  uint32_t* mm;
  Grid* grid;
  int32_t* cellid;
  uint32_t *cnt, *off, *cursor, *sums;
  float4* sx;
  int32_t* sidx;
};

static int64_t cell_cap(int64_t n) { return n / 2 > 1 ? n / 2 : 1; }

static Buffers carve(void* ws, int64_t n, size_t* total) {
  Carve c(ws);
  Buffers b;
  const int64_t cap = cell_cap(n);
  b.mm = c.take<uint32_t>(8);
  b.grid = c.take<Grid>(1);
  b.cellid = c.take<int32_t>(n);
  b.cnt = c.take<uint32_t>(cap);
  b.off = c.take<uint32_t>(cap + 1);
  b.cursor = c.take<uint32_t>(cap);
  b.sums = c.take<uint32_t>(ceil_div(cap, 1024) + 1);
  b.sx = c.take<float4>(n);
  b.sidx = c.take<int32_t>(n);
  *total = c.off + 256;
  return b;
}

__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* __restrict__ p, int64_t count) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) p[t] = 0u;
}

template <int D>
static int search(const SearchArgs& sa, int k, hipStream_t s) {
  const dim3 grid(static_cast<unsigned>(ceil_div(sa.nq, 128))), block(128);
  if (k <= 16) hipLaunchKernelGGL((k_grid_knn<D, 16>), grid, block, 0, s, sa);
  else if (k <= 32) hipLaunchKernelGGL((k_grid_knn<D, 32>), grid, block, 0, s, sa);
  else hipLaunchKernelGGL((k_grid_knn<D, 64>), grid, block, 0, s, sa);
  MV_LAUNCHED("k_grid_knn");
  return MV_OK;
}

template <int D>
static int build(const float* X, int64_t n, const Buffers& B, hipStream_t s) {
  const int64_t cap = cell_cap(n);
  const unsigned bn = static_cast<unsigned>(ceil_div(n, 256));
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, s, B.mm);
  MV_LAUNCHED("k_bounds_init");
  hipLaunchKernelGGL(k_bounds, dim3(static_cast<unsigned>(std::min<int64_t>(ceil_div(n, 256), 1024))), dim3(256), 0, s,
                     X, n, D, B.mm);
  MV_LAUNCHED("k_bounds");
  hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(64), 0, s, B.mm, D, cap, B.grid);
  MV_LAUNCHED("k_grid_setup");
  const unsigned bc = static_cast<unsigned>(ceil_div(cap, 256));
  hipLaunchKernelGGL(k_zero_u32, dim3(bc), dim3(256), 0, s, B.cnt, cap);
  MV_LAUNCHED("k_zero_u32");
  hipLaunchKernelGGL(k_zero_u32, dim3(bc), dim3(256), 0, s, B.cursor, cap);
  MV_LAUNCHED("k_zero_u32");
  hipLaunchKernelGGL((k_cell_count<D>), dim3(bn), dim3(256), 0, s, X, n, B.grid, B.cellid, B.cnt);
  MV_LAUNCHED("k_cell_count");
  const int64_t nb = ceil_div(cap, 1024);
  hipLaunchKernelGGL(k_scan1, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, B.cnt, cap, B.off, B.sums);
  MV_LAUNCHED("k_scan1");
  hipLaunchKernelGGL(k_scan2, dim3(1), dim3(256), 0, s, B.sums, nb);
  MV_LAUNCHED("k_scan2");
  hipLaunchKernelGGL(k_scan3, dim3(static_cast<unsigned>(ceil_div(cap + 1, 256))), dim3(256), 0, s, B.off, cap, B.sums,
                     nb);
  MV_LAUNCHED("k_scan3");
  hipLaunchKernelGGL((k_cell_fill<D>), dim3(bn), dim3(256), 0, s, X, n, B.cellid, B.off, B.cursor, B.sx, B.sidx);
  MV_LAUNCHED("k_cell_fill");
  return MV_OK;
}

}  // namespace lowd

size_t lowd_knn_ws_bytes(int64_t n) {
  size_t total = 0;
  lowd::carve(nullptr, n, &total);
  return total;
}

int lowd_knn(const float* Q, int64_t nq, const float* X, int64_t n, int d, int k, const int32_t* q_self,
             int64_t self_base, int euclid, int32_t* out_idx, float* out_dist, void* ws, size_t ws_bytes,
             hipStream_t s) {
  using namespace lowd;
  if (d < 1 || d > LOWD_MAX_D || k < 1 || k > 64) {
    mv::set_error("lowd_knn: d=%d or k=%d outside the grid search", d, k);
    return MV_EUNSUPPORTED;
  }
  size_t need = 0;
  Buffers B = carve(ws, n, &need);
  if (!ws || ws_bytes < need) {
    mv::set_error("lowd_knn: workspace of %zu bytes < required %zu", ws_bytes, need);
    return MV_EWORKSPACE;
  }
  const bool same = Q == X && nq == n;
  int rc = d == 1 ? build<1>(X, n, B, s) : (d == 2 ? build<2>(X, n, B, s) : build<3>(X, n, B, s));
  if (rc) return rc;
  SearchArgs sa;
  sa.grid = B.grid;
  sa.sx = B.sx;
  sa.sidx = B.sidx;
  sa.off = B.off;
  sa.Q = same ? nullptr : Q;
  sa.nq = nq;
  sa.q_self = q_self;
  sa.self_base = q_self ? -1 : self_base;
  sa.k = k;
  sa.euclid = euclid;
  sa.out_idx = out_idx;
  sa.out_dist = out_dist;
  return d == 1 ? search<1>(sa, k, s) : (d == 2 ? search<2>(sa, k, s) : search<3>(sa, k, s));
}

}  // namespace mv
