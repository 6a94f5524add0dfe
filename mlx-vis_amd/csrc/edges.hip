// Edge-gradient kernels with pluggable force laws: the layout gradient family
// that the reference's CNE, t-SNE and PaCMAP/LocalMAP share with UMAP.
//
//   k_cne_grad<LOSS, D>  contrastive losses of CNE over edges (i, j) with m
//                        negatives each, Cauchy similarity s = 1 / (1 + d^2):
//                        InfoNCE, NCE, NEG (mlx_vis/_cne/cne.py:219-313). The
//                        head gets attraction + repulsion, the tail minus the
//                        attraction, every negative its repulsive term.
//   k_pair_grad<LAW, D>  one force law over a list of pairs (src, dst) with
//                        weight w: PaCMAP's neighbour, mid-near and further
//                        pairs and LocalMAP's neighbour law
//                        (mlx_vis/_pacmap/pacmap.py:275-354), and t-SNE's
//                        attractive term (mlx_vis/_tsne/tsne.py:511-517).
//   k_adam_y             the Adam update of the embedding used by CNE
//                        (cne.py:355-360) and PaCMAP (pacmap.py:300-303).
//
// Gradients accumulate into a caller-zeroed FP32 buffer with atomics (the
// reference's scatter-adds, `.at[].add`); edges of CNE's symmetrised list come
// sorted by head, so a wavefront first sums the head terms of equal heads.
// Negatives and mini-batches (uniform, with replacement, as
// mx.random.randint at cne.py:338, 346) come from Philox-4x32-10 keyed by the
// seed and counted by (edge, iteration, slot).
#include <math.h>

#include "common.hpp"

namespace mv {
namespace edges {

enum { CNE_INFONCE = 0, CNE_NCE = 1, CNE_NEG = 2 };
enum { LAW_PACMAP_NN = 0, LAW_PACMAP_MN = 1, LAW_PACMAP_FP = 2, LAW_LOCALMAP_NN = 3, LAW_TSNE_ATTR = 4 };
constexpr int MNEG_MAX = 32;

struct CneArgs {
  const float* Y;
  int64_t n;
  const int32_t* ei;
  const int32_t* ej;
  int64_t E;           // edges in the list
  int64_t B;           // edges processed (E, or the batch size)
  int batch;           // nonzero: edge b is a uniform draw from [0, E)
  const int32_t* neg;  // nullable: explicit negatives [B, m]; else Philox draws from [0, n)
  int m;
  float inv_m;         // NCE: 1 / m
  float scale;         // multiplies every term (E / B for mini-batches)
  uint32_t k0, k1;     // Philox key
  uint32_t iter;
  float* grad;         // [n, D], accumulated
};

__device__ __forceinline__ uint32_t draw(uint32_t b, uint32_t iter, uint32_t slot, uint32_t k0, uint32_t k1) {
  return Philox::gen(make_uint4(b, iter, slot, 0xC0E5u), k0, k1).x;
}

template <int LOSS, int D>
__global__ __launch_bounds__(256) void k_cne_grad(CneArgs a) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool live = b < a.B;
  int64_t e = b;
  if (live && a.batch) e = uniform_below(draw(static_cast<uint32_t>(b), a.iter, 0xFFFFu, a.k0, a.k1),
                                         static_cast<uint32_t>(a.E));
  int32_t i = 0, j = 0;
  float yi[D] = {}, dp[D] = {}, gh[D];
#pragma unroll
  for (int q = 0; q < D; ++q) gh[q] = 0.f;
  if (live) {
    i = a.ei[e];
    j = a.ej[e];
    float d2 = 0.f;
#pragma unroll
    for (int q = 0; q < D; ++q) {
      yi[q] = a.Y[static_cast<int64_t>(i) * D + q];
      dp[q] = yi[q] - a.Y[static_cast<int64_t>(j) * D + q];
      d2 += dp[q] * dp[q];
    }
    const float sp = 1.f / (1.f + d2);
    // Negative k of edge b (recomputed in the second pass instead of stored:
    // no per-thread arrays).
#define MV_NEG(K) (a.neg ? a.neg[b * a.m + (K)]                                                             \
                         : static_cast<int32_t>(uniform_below(draw(static_cast<uint32_t>(b), a.iter,             \
                                                                   static_cast<uint32_t>(K), a.k0, a.k1),        \
                                                              static_cast<uint32_t>(a.n))))
    float S = 0.f;   // sum of the negatives' similarities (InfoNCE's normaliser)
    if (LOSS == CNE_INFONCE) {
      for (int k = 0; k < a.m; ++k) {
        const int32_t nk = MV_NEG(k);
        float n2 = 0.f;
#pragma unroll
        for (int q = 0; q < D; ++q) {
          const float df = yi[q] - a.Y[static_cast<int64_t>(nk) * D + q];
          n2 += df * df;
        }
        S += 1.f / (1.f + n2);
      }
    }
    const float Z = sp + S;
    float wa;   // attraction coefficient on (y_i - y_j)
    if (LOSS == CNE_INFONCE) wa = (1.f - sp / Z) * 2.f * sp;
    else if (LOSS == CNE_NCE) wa = 2.f * a.inv_m * sp / (sp + a.inv_m);
    else wa = 2.f * sp;
    wa *= a.scale;
#pragma unroll
    for (int q = 0; q < D; ++q) {
      gh[q] = wa * dp[q];
      atomicAdd(&a.grad[static_cast<int64_t>(j) * D + q], -gh[q]);
    }
    for (int k = 0; k < a.m; ++k) {
      const int32_t nk = MV_NEG(k);
      float dn[D], n2 = 0.f;
#pragma unroll
      for (int q = 0; q < D; ++q) {
        dn[q] = yi[q] - a.Y[static_cast<int64_t>(nk) * D + q];
        n2 += dn[q] * dn[q];
      }
      const float sn = 1.f / (1.f + n2);
      float wr;   // repulsion coefficient on (y_i - y_k): the head gets -wr, the negative +wr
      if (LOSS == CNE_INFONCE) wr = (sn / Z) * 2.f * sn;
      else if (LOSS == CNE_NCE) wr = 2.f * sn * sn / (sn + a.inv_m);
      else wr = 2.f * sn * sn / fmaxf(n2 * sn, 1e-8f);   // 1 - s = d^2 s, without the cancellation
      wr *= a.scale;
#pragma unroll
      for (int q = 0; q < D; ++q) {
        gh[q] -= wr * dn[q];
        atomicAdd(&a.grad[static_cast<int64_t>(nk) * D + q], wr * dn[q]);
      }
    }
#undef MV_NEG
  }
  // head terms: each run of adjacent lanes with the same head (a head-sorted
  // list makes runs long; mini-batch draws leave mostly runs of one) is summed
  // by a segmented suffix scan and added with one atomic by the run's first
  // lane. A lane only takes a farther partial sum while no run ends between
  // the two lanes, so equal heads in separate runs are never merged.
  const int l = threadIdx.x & 63;
  const int32_t key = live ? i : -1 - l;   // dead lanes never match
  const int32_t prev = __shfl_up(key, 1);
  const int32_t next = __shfl_down(key, 1);
  const bool start = l == 0 || prev != key;
  const bool run_end = l == 63 || next != key;
#pragma unroll
  for (int q = 0; q < D; ++q) {
    float v = gh[q];
    bool closed = run_end;   // a run ends within [l, l + off)
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const float o = __shfl_down(v, off);
      const bool oc = __shfl_down(static_cast<int>(closed), off) != 0;
      if (l + off < 64 && !closed) v += o;
      if (l + off < 64) closed = closed || oc;
    }
    gh[q] = v;
  }
  if (live && start) {
#pragma unroll
    for (int q = 0; q < D; ++q) atomicAdd(&a.grad[static_cast<int64_t>(i) * D + q], gh[q]);
  }
}

struct PairArgs {
  const float* Y;
  const int32_t* src;
  const int32_t* dst;
  int64_t P;
  const float* w_pair;   // nullable: per-pair weight (t-SNE edge weights)
  float w;               // weight for all pairs (PaCMAP phase weight)
  float scale;           // LocalMAP neighbour scale
  float* grad;
};

// g = coef(d) (y_src - y_dst), d = |y_src - y_dst|^2 + 1 (PaCMAP) or the plain
// squared distance (t-SNE). Attractive laws add g to src and -g to dst; the
// further-pair law adds -g to src and g to dst; t-SNE adds g to src only.
template <int LAW, int D>
__global__ __launch_bounds__(256) void k_pair_grad(PairArgs a) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= a.P) return;
  const int64_t s = a.src[p], t = a.dst[p];
  float df[D], d2 = 0.f;
#pragma unroll
  for (int q = 0; q < D; ++q) {
    df[q] = a.Y[s * D + q] - a.Y[t * D + q];
    d2 += df[q] * df[q];
  }
  const float w = a.w_pair ? a.w_pair[p] : a.w;
  const float d = d2 + 1.f;
  float c;
  if (LAW == LAW_PACMAP_NN) c = w * 20.f / ((10.f + d) * (10.f + d));
  else if (LAW == LAW_PACMAP_MN) c = w * 20000.f / ((10000.f + d) * (10000.f + d));
  else if (LAW == LAW_PACMAP_FP) c = -(w * 2.f / ((1.f + d) * (1.f + d)));
  else if (LAW == LAW_LOCALMAP_NN) c = w * 20.f * a.scale / (((10.f + d) * (10.f + d)) * sqrtf(d));
  else c = 4.f * w / (1.f + d2);
#pragma unroll
  for (int q = 0; q < D; ++q) {
    atomicAdd(&a.grad[s * D + q], c * df[q]);
    if (LAW != LAW_TSNE_ATTR) atomicAdd(&a.grad[t * D + q], -c * df[q]);
  }
}

// Y -= step * m / (sqrt(v) + eps') after m = b1 m + (1 - b1) g,
// v = b2 v + (1 - b2) g^2; the host folds the bias corrections into step and
// eps'. Also zeroes g for the next iteration.
__global__ __launch_bounds__(256) void k_adam_y(float* __restrict__ Y, float* __restrict__ g, float* __restrict__ m,
                                                float* __restrict__ v, int64_t count, float b1, float b2, float step,
                                                float eps) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.f - b1) * gt;
  const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  Y[t] -= step * mt / (sqrtf(vt) + eps);
  g[t] = 0.f;
}

__global__ void k_zero(float* __restrict__ p, int64_t count) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) p[t] = 0.f;
}

template <int D>
static int launch_cne(int loss, const CneArgs& a, hipStream_t s) {
  const dim3 grid(static_cast<unsigned>(ceil_div(a.B, 256)));
  switch (loss) {
    case CNE_INFONCE: hipLaunchKernelGGL((k_cne_grad<CNE_INFONCE, D>), grid, dim3(256), 0, s, a); break;
    case CNE_NCE: hipLaunchKernelGGL((k_cne_grad<CNE_NCE, D>), grid, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((k_cne_grad<CNE_NEG, D>), grid, dim3(256), 0, s, a); break;
  }
  MV_LAUNCHED("k_cne_grad");
  return MV_OK;
}

static int cne_grad(int dim, int loss, const CneArgs& a, hipStream_t s) {
  switch (dim) {
    case 1: return launch_cne<1>(loss, a, s);
    case 2: return launch_cne<2>(loss, a, s);
    case 3: return launch_cne<3>(loss, a, s);
    default: return launch_cne<4>(loss, a, s);
  }
}

template <int D>
static int launch_pair(int law, const PairArgs& a, hipStream_t s) {
  const dim3 grid(static_cast<unsigned>(ceil_div(a.P, 256)));
  switch (law) {
    case LAW_PACMAP_NN: hipLaunchKernelGGL((k_pair_grad<LAW_PACMAP_NN, D>), grid, dim3(256), 0, s, a); break;
    case LAW_PACMAP_MN: hipLaunchKernelGGL((k_pair_grad<LAW_PACMAP_MN, D>), grid, dim3(256), 0, s, a); break;
    case LAW_PACMAP_FP: hipLaunchKernelGGL((k_pair_grad<LAW_PACMAP_FP, D>), grid, dim3(256), 0, s, a); break;
    case LAW_LOCALMAP_NN: hipLaunchKernelGGL((k_pair_grad<LAW_LOCALMAP_NN, D>), grid, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((k_pair_grad<LAW_TSNE_ATTR, D>), grid, dim3(256), 0, s, a); break;
  }
  MV_LAUNCHED("k_pair_grad");
  return MV_OK;
}

}  // namespace edges
}  // namespace mv

using namespace mv;
using namespace mv::edges;

extern "C" {

int mv_cne_grad(const float* Y, int64_t n, int dim, const int32_t* ei, const int32_t* ej, int64_t E,
                const int32_t* neg, int m, int loss, float scale, uint64_t seed, int iter, int64_t batch,
                float* grad, void* stream) {
  MV_REQUIRE(Y && ei && ej && grad, "mv_cne_grad: null pointer");
  MV_REQUIRE(n >= 1 && n < (1ll << 31) && dim >= 1 && dim <= 4 && E >= 0 && E < (1ll << 31),
             "mv_cne_grad: bad sizes n=%lld dim=%d E=%lld", (long long)n, dim, (long long)E);
  MV_REQUIRE(m >= 1 && m <= MNEG_MAX, "mv_cne_grad: negatives per edge %d outside [1, %d]", m, MNEG_MAX);
  MV_REQUIRE(loss >= CNE_INFONCE && loss <= CNE_NEG, "mv_cne_grad: unknown loss %d", loss);
  MV_REQUIRE(batch >= 0 && !(neg && batch), "mv_cne_grad: explicit negatives need the full edge list");
  MV_FAULT_POINT("launch");
  CneArgs a;
  a.Y = Y;
  a.n = n;
  a.ei = ei;
  a.ej = ej;
  a.E = E;
  a.batch = batch > 0 ? 1 : 0;
  a.B = batch > 0 ? batch : E;
  a.neg = neg;
  a.m = m;
  a.inv_m = 1.f / static_cast<float>(m);
  a.scale = scale;
  a.k0 = static_cast<uint32_t>(seed);
  a.k1 = static_cast<uint32_t>(seed >> 32);
  a.iter = static_cast<uint32_t>(iter);
  a.grad = grad;
  if (a.B == 0) return MV_OK;
  return cne_grad(dim, loss, a, static_cast<hipStream_t>(stream));
}

int mv_cne_optimize(float* Y, int64_t n, int dim, const int32_t* ei, const int32_t* ej, int64_t E, int m, int loss,
                    int it0, int it1, float lr, int64_t batch, uint64_t seed, float* adam_m, float* adam_v,
                    float* grad, void* stream) {
  MV_REQUIRE(Y && ei && ej && adam_m && adam_v && grad, "mv_cne_optimize: null pointer");
  MV_REQUIRE(n >= 1 && n < (1ll << 31) && dim >= 1 && dim <= 4 && E >= 1 && E < (1ll << 31),
             "mv_cne_optimize: bad sizes n=%lld dim=%d E=%lld", (long long)n, dim, (long long)E);
  MV_REQUIRE(m >= 1 && m <= MNEG_MAX, "mv_cne_optimize: negatives per edge %d outside [1, %d]", m, MNEG_MAX);
  MV_REQUIRE(loss >= CNE_INFONCE && loss <= CNE_NEG, "mv_cne_optimize: unknown loss %d", loss);
  MV_REQUIRE(1 <= it0 && it0 <= it1 && batch >= 0, "mv_cne_optimize: bad iteration range [%d, %d)", it0, it1);
  MV_FAULT_POINT("launch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t count = n * dim;
  const bool mini = batch > 0 && batch < E;
  CneArgs a;
  a.Y = Y;
  a.n = n;
  a.ei = ei;
  a.ej = ej;
  a.E = E;
  a.batch = mini ? 1 : 0;
  a.B = mini ? batch : E;
  a.neg = nullptr;
  a.m = m;
  a.inv_m = 1.f / static_cast<float>(m);
  a.scale = mini ? static_cast<float>(static_cast<double>(E) / static_cast<double>(batch)) : 1.f;
  a.k0 = static_cast<uint32_t>(seed);
  a.k1 = static_cast<uint32_t>(seed >> 32);
  a.grad = grad;
  const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
  for (int it = it0; it < it1; ++it) {
    a.iter = static_cast<uint32_t>(it);
    if (int rc = cne_grad(dim, loss, a, s)) return rc;
    // Y -= lr * (m / c1) / (sqrt(v / c2) + eps)  (cne.py:355-360)
    const double c1 = 1.0 - pow(b1, it), c2 = 1.0 - pow(b2, it);
    const float step = static_cast<float>(lr * sqrt(c2) / c1);
    const float epsp = static_cast<float>(eps * sqrt(c2));
    hipLaunchKernelGGL(k_adam_y, dim3(static_cast<unsigned>(ceil_div(count, 256))), dim3(256), 0, s, Y, grad, adam_m,
                       adam_v, count, static_cast<float>(b1), static_cast<float>(b2), step, epsp);
    MV_LAUNCHED("k_adam_y");
  }
  return MV_OK;
}

int mv_pair_grad(const float* Y, int64_t n, int dim, const int32_t* src, const int32_t* dst, int64_t P,
                 const float* w_pair, float w, float scale, int law, float* grad, void* stream) {
  MV_REQUIRE(Y && src && dst && grad, "mv_pair_grad: null pointer");
  MV_REQUIRE(n >= 1 && dim >= 1 && dim <= 4 && P >= 0, "mv_pair_grad: bad sizes");
  MV_REQUIRE(law >= LAW_PACMAP_NN && law <= LAW_TSNE_ATTR, "mv_pair_grad: unknown law %d", law);
  MV_FAULT_POINT("launch");
  if (P == 0) return MV_OK;
  PairArgs a;
  a.Y = Y;
  a.src = src;
  a.dst = dst;
  a.P = P;
  a.w_pair = w_pair;
  a.w = w;
  a.scale = scale;
  a.grad = grad;
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (dim) {
    case 1: return launch_pair<1>(law, a, s);
    case 2: return launch_pair<2>(law, a, s);
    case 3: return launch_pair<3>(law, a, s);
    default: return launch_pair<4>(law, a, s);
  }
}

int mv_adam_update(float* Y, float* grad, float* adam_m, float* adam_v, int64_t count, float b1, float b2, float step,
                   float eps, void* stream) {
  MV_REQUIRE(Y && grad && adam_m && adam_v, "mv_adam_update: null pointer");
  MV_REQUIRE(count >= 0, "mv_adam_update: count < 0");
  MV_FAULT_POINT("launch");
  if (count == 0) return MV_OK;
  hipLaunchKernelGGL(k_adam_y, dim3(static_cast<unsigned>(ceil_div(count, 256))), dim3(256), 0,
                     static_cast<hipStream_t>(stream), Y, grad, adam_m, adam_v, count, b1, b2, step, eps);
  MV_LAUNCHED("k_adam_y");
  return MV_OK;
}

}  // extern "C"
