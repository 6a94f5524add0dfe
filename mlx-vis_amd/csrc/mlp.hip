// Parametric embedding (learned form) for gfx950: an MLP din -> h1 -> h2 -> dout
// with GELU, BF16 weights and activations, FP32 master weights, FP32
// accumulation, and Adam. Trained on the UMAP edge loss over a kNN graph plus a
// regression to a plain-layout embedding; maps new points in one pass.
//
// The reference has no parametric UMAP; its only MLP is the MMAE autoencoder
// (mlx_vis/_mmae/mmae.py:22-60, trained with AdamW at mmae.py:258-281), the
// precedent for the API shape (fit on a reference embedding, encode new
// points). The loss head reuses the layout's gradient law: the attractive
// coefficient -2ab d^(2(b-1)) / (1 + a d^(2b)) and the repulsive coefficient
// 2b / ((0.001 + d^2)(1 + a d^(2b))) of umap.py:367-380, here without the
// layout's +-4 clip, so that it is the exact gradient of the edge
// cross-entropy.
//
// Kernels:
//   k_gemm_nt<EPI>   C = A B^T, A (M x K) and B (N x K) BF16 row-major (K
//                    contiguous), v_mfma_f32_16x16x32_bf16, 128x128 tiles,
//                    64-element K chunks double-buffered in LDS; epilogues:
//                    FP32 store, bias + GELU (stores pre-activation and
//                    activation in BF16), GELU-gradient product (BF16).
//   k_transpose_bf16, k_gather_x, k_to_bf16, k_head_fwd, k_loss_grad,
//   k_head_bwd, k_adam_blob, k_sample_edges.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <math.h>

#include "common.hpp"

namespace mv {
namespace mlp {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef uint16_t bf16_t;   // storage type: IEEE bfloat16 bit pattern

constexpr int BM = 128, BN = 128, LDSA = 36;   // LDS row stride in floats (128 B data + 16 B pad)
constexpr int NT = 256;

typedef __bf16 v2bf __attribute__((ext_vector_type(2)));

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ bf16_t f2bf(float v) { return __builtin_bit_cast(bf16_t, static_cast<__bf16>(v)); }
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// GELU with the standard normal CDF Phi(x) = 1 - erfc(x / sqrt2) / 2, where
// erfc(z) (z >= 0) follows Abramowitz & Stegun 7.1.26:
//   erfc(z) = t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-z^2), t = 1 / (1 + p z),
// absolute error at most 1.5e-7 (far below the BF16 resolution the network
// stores its activations in). Negative x use Phi(x) = erfc(|x| / sqrt2) / 2,
// which has no cancellation.
struct NormCdf {
  float cdf;   // Phi(x)
  float e;     // exp(-x^2 / 2)
};
__device__ __forceinline__ NormCdf norm_cdf(float x) {
  const float z = fabsf(x) * 0.70710678118654752f;
  const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, z, 1.f));
  const float poly = t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 1.061405429f, -1.453152027f), 1.421413741f), -0.284496736f),
                              0.254829592f);
  const float e = __expf(-z * z);
  // 1 - half_erfc is written as one fused operation, so that every inlined
  // copy rounds alike (the lookup table below is built from this function).
  const float hp = 0.5f * poly;
  return NormCdf{x >= 0.f ? fmaf(-hp, e, 1.f) : hp * e, e};
}
__device__ __forceinline__ float gelu(float x) { return x * norm_cdf(x).cdf; }

// GELU of a BF16 input rounded to BF16, by table for |x| in [2^-12, 16): the
// 2 x 16 x 128 BF16 values there (sign, exponent field 115..130, mantissa)
// index a 4,096-entry table that holds f2bf(gelu(x)) computed by the function
// above, so the lookup returns the same bits as the formula. For finite
// |x| >= 16 the formula's exp(-x^2 / 2) underflows to 0, so it returns x
// (x >= 16) or -0 (x <= -16) exactly, and so does the rule here. Other
// inputs (zero, |x| < 2^-12, Inf/NaN) take the formula.
constexpr int GELU_LUT = 4096;
constexpr uint32_t GELU_LUT_EXP0 = 115;
__device__ __forceinline__ uint32_t gelu_lut_bits(int i) {
  return (static_cast<uint32_t>(i >> 11) << 15) | (static_cast<uint32_t>(i & 2047) + (GELU_LUT_EXP0 << 7));
}
// off = (exponent field - 115) * 128 + mantissa: < 2,048 in the table,
// [2,048, 17,920) for 16 <= |x| < Inf, larger (or wrapped) otherwise
constexpr uint32_t GELU_OFF_BIG_END = (255u - GELU_LUT_EXP0) << 7;
__device__ __forceinline__ uint32_t gelu_lut_off(uint32_t pb) { return (pb & 0x7FFFu) - (GELU_LUT_EXP0 << 7); }
__device__ __forceinline__ uint32_t gelu_big(uint32_t pb) { return (pb & 0x8000u) ? 0x8000u : pb; }
__device__ __forceinline__ bf16_t gelu_bf16(const bf16_t* lut, bf16_t pb) {
  const uint32_t off = gelu_lut_off(pb);
  if (off < 2048u) return lut[((static_cast<uint32_t>(pb) >> 15) << 11) + off];
  if (off < GELU_OFF_BIG_END) return static_cast<bf16_t>(gelu_big(pb));
  return f2bf(gelu(bf2f(pb)));
}
// N inputs at once: when every lane's inputs lie in the table's range (a
// wave-uniform test) the N table reads are issued unconditionally, so that
// they are in flight together instead of one read per branch.
template <int N>
__device__ __forceinline__ void gelu_bf16_n(const bf16_t* lut, const uint32_t (&pb)[N], uint32_t (&out)[N]) {
  bool miss = false;
#pragma unroll
  for (int k = 0; k < N; ++k) miss |= gelu_lut_off(pb[k]) >= GELU_OFF_BIG_END;
  if (__builtin_amdgcn_ballot_w64(miss) == 0) {
    uint32_t t[N];
#pragma unroll
    for (int k = 0; k < N; ++k) t[k] = lut[((pb[k] >> 15) << 11) + min(gelu_lut_off(pb[k]), 2047u)];
    // The empty statements use every read before any select: without them
    // the compiler sinks each read under its select, one branch per value.
#pragma unroll
    for (int k = 0; k < N; ++k) asm volatile("" : "+v"(t[k]));
#pragma unroll
    for (int k = 0; k < N; ++k) out[k] = gelu_lut_off(pb[k]) < 2048u ? t[k] : gelu_big(pb[k]);
  } else {
#pragma unroll
    for (int k = 0; k < N; ++k) out[k] = gelu_bf16(lut, static_cast<bf16_t>(pb[k]));
  }
}
__device__ __forceinline__ float gelu_grad(float x) {
  const NormCdf n = norm_cdf(x);
  return n.cdf + x * 0.39894228040143268f * n.e;
}

enum { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_GELU_GRAD = 2, EPI_BF16 = 3, EPI_F32_ADD = 4, EPI_HEAD = 5 };

struct GemmArgs {
  const void* A;     // M x K, row stride lda (elements): BF16, or FP32 when the kernel's AF32 is set
  const bf16_t* B;   // N x K, row stride ldb
  int M, N, K;       // K multiple of 64
  int kchunks;       // K chunks (of 64) per split
  int m_tiles, n_tiles, splits;   // tile grid (set by the launcher)
  int kvalid;        // AF32: A elements per row that exist (a multiple of 8); later ones read as 0
  int64_t lda, ldb, ldc, ldx;
  void* C;           // F32 or BF16 output, row stride ldc
  const float* bias;       // EPI_BIAS_GELU / EPI_HEAD: bias[N]
  bf16_t* pre_out;         // EPI_BIAS_GELU (nullable): pre-activation output (row stride ldx)
  const bf16_t* pre_in;    // EPI_GELU_GRAD: pre-activation input (row stride ldx)
  float* colsum;           // EPI_GELU_GRAD (nullable): colsum[col] += sum over rows of the output
  const float* W3;         // EPI_HEAD: output layer (dout x N, FP32)
  int dout;                // EPI_HEAD: outputs (<= 4); C is Z (M x dout FP32), pre-filled with b3
  bf16_t* ct;              // BF16 epilogues (nullable): transposed copy, ct[col * ldct + row]; C may then be null
  int64_t ldct;
  int64_t split_stride;    // EPI_F32: K split s stores its partial product to C + s * split_stride
};

// Sum over the 16 lanes of a DPP row (row_ror 8, 4, 2, 1); every lane of the
// row receives the total.
__device__ __forceinline__ float row16_sum(float v) {
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x124, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x122, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x121, 0xF, 0xF, false));
  return v;
}

// Eight FP32 values rounded to BF16 and packed into 16 bytes.
__device__ __forceinline__ float4 pack_bf16x8(float4 a, float4 b) {
  const v2bf p0 = {static_cast<__bf16>(a.x), static_cast<__bf16>(a.y)};
  const v2bf p1 = {static_cast<__bf16>(a.z), static_cast<__bf16>(a.w)};
  const v2bf p2 = {static_cast<__bf16>(b.x), static_cast<__bf16>(b.y)};
  const v2bf p3 = {static_cast<__bf16>(b.z), static_cast<__bf16>(b.w)};
  return make_float4(__builtin_bit_cast(float, p0), __builtin_bit_cast(float, p1), __builtin_bit_cast(float, p2),
                     __builtin_bit_cast(float, p3));
}

// C = A B^T on v_mfma_f32_16x16x32_bf16: 128 x 128 output tiles, four
// wavefronts of 64 x 64, 64-element K chunks double-buffered in LDS with a one
// chunk deep register prefetch. With AF32 the A operand is read as FP32 and
// rounded to BF16 (ties to even) while it is staged. Grid z splits K
// (EPI_F32_ADD accumulates the partial products with atomics). Epilogues:
//   EPI_F32 / EPI_BF16   store C;
//   EPI_F32_ADD          C += partial (C zeroed beforehand);
//   EPI_BIAS_GELU        pre = bf16(acc + bias), C = bf16(gelu(pre));
//   EPI_GELU_GRAD        C = bf16(acc * gelu'(pre_in)), colsum += C;
//   EPI_HEAD             h = bf16(gelu(bf16(acc + bias))), Z[row, o] += sum_col h W3[o, col].
template <int EPI, bool AF32, int DO = 4>
__global__ __launch_bounds__(NT, 2) void k_gemm_nt(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) float s_stage[2 * 2 * BM * LDSA];
  constexpr bool LUT = EPI == EPI_BIAS_GELU || EPI == EPI_HEAD;
  __shared__ bf16_t s_gelu[LUT ? GELU_LUT : 1];
  const int tid = threadIdx.x;
  if constexpr (LUT) {
    for (int i = tid; i < GELU_LUT; i += NT) s_gelu[i] = f2bf(gelu(bf2f(static_cast<bf16_t>(gelu_lut_bits(i)))));
    __syncthreads();
  }
  const int w = tid >> 6, l = tid & 63;
  const int wr = w >> 1, wc = w & 1;
  const int lr = tid >> 3, lq = tid & 7;   // staging row, 16-byte slot in the 128-byte BF16 chunk row
  const int nk_all = g.K / 64;
  // Persistent grid (a multiple of 8 workgroups): workgroup b takes tiles b,
  // b + G, ... Tile t runs on XCD t mod 8 and covers a row block m with
  // m mod 8 = t mod 8, so that the column tiles and K splits of a row block
  // share one XCD's L2.
  const int ncols = g.n_tiles * g.splits;
  const int slots = 8 * ((g.m_tiles + 7) / 8) * ncols;
  for (int t = blockIdx.x; t < slots; t += gridDim.x) {
  const int mt = (t & 7) + 8 * ((t >> 3) / ncols);
  if (mt >= g.m_tiles) continue;
  const int cc = (t >> 3) % ncols;
  const int bm = mt * BM, bn = (cc % g.n_tiles) * BN;
  const int ks = cc / g.n_tiles;   // K split
  const int kc0 = ks * g.kchunks;
  const int kc1 = min(nk_all, kc0 + g.kchunks);
  if (kc0 >= kc1) continue;

  // Rows past M / N are clamped (read in bounds, never stored).
  const char* Arow[4];
  const char* Brow[4];
  constexpr int AEB = AF32 ? 4 : 2;   // bytes per A element
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int ra = min(bm + lr + 32 * i, g.M - 1);
    const int rbn = min(bn + lr + 32 * i, g.N - 1);
    Arow[i] = static_cast<const char*>(g.A) + static_cast<int64_t>(ra) * g.lda * AEB + lq * 8 * AEB;
    Brow[i] = reinterpret_cast<const char*>(g.B + static_cast<int64_t>(rbn) * g.ldb) + lq * 16;
  }
  v4f acc[4][4];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = v4f{0.f, 0.f, 0.f, 0.f};

  // Register ring of NPF chunks in flight: chunk c (relative to kc0) is
  // loaded into set c % NPF NPF chunks before it is staged to LDS, so a
  // chunk's global-memory latency is covered by the multiplies of the NPF
  // chunks before it (a BF16 chunk is only ~512 MFMA cycles per wavefront).
  constexpr int NPF = AF32 ? 1 : 2;   // the FP32-input variant needs its registers for the wider slots
  // Named registers, not arrays: identical steps on different arrays get
  // merged by the compiler into one path indexing memory.
  float4 ra0_0, ra0_1, ra0_2, ra0_3, ra1_0, ra1_1, ra1_2, ra1_3;   // A slots
  float4 rx0_0, rx0_1, rx0_2, rx0_3, rx1_0, rx1_1, rx1_2, rx1_3;   // FP32 A, second halves
  float4 rv0_0, rv0_1, rv0_2, rv0_3, rv1_0, rv1_1, rv1_2, rv1_3;   // B slots
  const int nch = kc1 - kc0;
  // FP32 A: an 8-element slot past the row's valid length is read at the row
  // start (in bounds) and zeroed when staged.
#define MV_GLOAD1(SET, I, KQ, AVAL)                                                                 \
  do {                                                                                            \
    const char* pa_ = (AVAL) ? Arow[I] + static_cast<int64_t>(KQ) * 64 * AEB : Arow[I] - lq * 8 * AEB; \
    ra##SET##_##I = *reinterpret_cast<const float4*>(pa_);                                           \
    if (AF32) rx##SET##_##I = *reinterpret_cast<const float4*>(pa_ + 16);                            \
    rv##SET##_##I = *reinterpret_cast<const float4*>(Brow[I] + static_cast<int64_t>(KQ) * 128);      \
  } while (0)
#define MV_GLOAD(SET, C)                                                                            \
  do {                                                                                            \
    const int kq_ = kc0 + ((C) < nch ? (C) : nch - 1);                                            \
    const bool aval_ = !AF32 || kq_ * 64 + lq * 8 < g.kvalid;                                     \
    MV_GLOAD1(SET, 0, kq_, aval_);                                                                \
    MV_GLOAD1(SET, 1, kq_, aval_);                                                                \
    MV_GLOAD1(SET, 2, kq_, aval_);                                                                \
    MV_GLOAD1(SET, 3, kq_, aval_);                                                                \
  } while (0)
#define MV_LSTORE1(SET, I, KQ, AS, BS)                                                              \
  do {                                                                                            \
    float4 av_ = ra##SET##_##I;                                                                      \
    if (AF32) {                                                                                   \
      /* masked, not branched: a branch here makes the compiler wait for every load in flight */  \
      const uint32_t m_ = (KQ) * 64 + lq * 8 < g.kvalid ? 0xFFFFFFFFu : 0u;                       \
      const uint4 p_ = __builtin_bit_cast(uint4, pack_bf16x8(ra##SET##_##I, rx##SET##_##I));         \
      av_ = __builtin_bit_cast(float4, make_uint4(p_.x & m_, p_.y & m_, p_.z & m_, p_.w & m_));   \
    }                                                                                             \
    *reinterpret_cast<float4*>((AS) + (lr + 32 * (I)) * LDSA + lq * 4) = av_;                     \
    *reinterpret_cast<float4*>((BS) + (lr + 32 * (I)) * LDSA + lq * 4) = rv##SET##_##I;              \
  } while (0)
#define MV_LSTORE(SET, C, BUF)                                                                      \
  do {                                                                                            \
    float* As2_ = s_stage + (BUF) * (2 * BM * LDSA);                                              \
    float* Bs2_ = As2_ + BM * LDSA;                                                               \
    const int kq_ = kc0 + (C);                                                                    \
    MV_LSTORE1(SET, 0, kq_, As2_, Bs2_);                                                          \
    MV_LSTORE1(SET, 1, kq_, As2_, Bs2_);                                                          \
    MV_LSTORE1(SET, 2, kq_, As2_, Bs2_);                                                          \
    MV_LSTORE1(SET, 3, kq_, As2_, Bs2_);                                                          \
  } while (0)
  // The scheduling barriers keep the loads of chunk 1 ahead of those of
  // chunk NPF, the order the loop's wait counts assume.
  MV_GLOAD(0, 0);
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (NPF > 1) MV_GLOAD(1, 1);
  __builtin_amdgcn_sched_barrier(0);
  MV_LSTORE(0, 0, 0);
  MV_GLOAD(0, NPF);
  __syncthreads();
  // One step per chunk t = base + U; NS = (U + 1) % NPF names the register
  // set holding chunk t + 1.
#define MV_STEP(U, NS)                                                                                \
  do {                                                                                            \
    const int t_ = base + (U);                                                                    \
    if (t_ < nch) {                                                                               \
      const float* As = s_stage + (t_ & 1) * (2 * BM * LDSA);                                     \
      const float* Bs = As + BM * LDSA;                                                           \
      _Pragma("unroll") for (int gk = 0; gk < 2; ++gk) {                                          \
        v8bf af[4], bf[4];                                                                        \
        _Pragma("unroll") for (int mi = 0; mi < 4; ++mi) af[mi] = __builtin_bit_cast(             \
            v8bf, *reinterpret_cast<const v4f*>(As + (wr * 64 + mi * 16 + (l & 15)) * LDSA + gk * 16 + (l >> 4) * 4)); \
        _Pragma("unroll") for (int ni = 0; ni < 4; ++ni) bf[ni] = __builtin_bit_cast(             \
            v8bf, *reinterpret_cast<const v4f*>(Bs + (wc * 64 + ni * 16 + (l & 15)) * LDSA + gk * 16 + (l >> 4) * 4)); \
        _Pragma("unroll") for (int mi = 0; mi < 4; ++mi)                                          \
          _Pragma("unroll") for (int ni = 0; ni < 4; ++ni)                                        \
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mi], bf[ni], acc[mi][ni], 0, 0, 0); \
      }                                                                                           \
    }                                                                                             \
    /* The staging of the next chunk is unconditional (past the end: the     */                   \
    /* clamped last chunk into the buffer no step reads): loads and stores   */                   \
    /* under a branch make the compiler wait for every load in flight.       */                   \
    MV_LSTORE(NS, t_ + 1, (t_ + 1) & 1);                                                          \
    MV_GLOAD(NS, t_ + 1 + NPF);                                                                   \
    __syncthreads();                                                                              \
  } while (0)
  for (int base = 0; base < nch; base += NPF) {
    if constexpr (NPF == 2) {
      MV_STEP(0, 1);
      MV_STEP(1, 0);
    } else {
      MV_STEP(0, 0);
    }
  }
#undef MV_STEP
#undef MV_GLOAD
#undef MV_GLOAD1
#undef MV_LSTORE
#undef MV_LSTORE1

  // Epilogue. Per-column operands (bias, W3) and per-element inputs (pre_in)
  // are read into registers before any output is written: the compiler
  // cannot prove that the output does not alias them and would otherwise
  // re-load them after every store, each load paying a full memory latency.
  float bv[4];
#pragma unroll
  for (int ni = 0; ni < 4; ++ni) {
    const int col = min(bn + wc * 64 + ni * 16 + (l & 15), g.N - 1);
    bv[ni] = (EPI == EPI_BIAS_GELU || EPI == EPI_HEAD) ? g.bias[col] : 0.f;
  }
  if constexpr (EPI == EPI_HEAD) {
    // DO outputs (= g.dout); columns past N carry zero weights, so that
    // every lane runs the same instructions
    float w3[DO][4];   // [o][ni]
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      const int col = bn + wc * 64 + ni * 16 + (l & 15);
#pragma unroll
      for (int o = 0; o < DO; ++o) w3[o][ni] = col < g.N ? g.W3[o * g.N + col] : 0.f;
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      uint32_t pb[16], hb[16];   // [r][ni]
#pragma unroll
      for (int k = 0; k < 16; ++k) pb[k] = f2bf(acc[mi][k & 3][k >> 2] + bv[k & 3]);
      gelu_bf16_n<16>(s_gelu, pb, hb);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = bm + wr * 64 + mi * 16 + (l >> 4) * 4 + r;
        float part[DO];
#pragma unroll
        for (int o = 0; o < DO; ++o) part[o] = 0.f;
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
          const float h = bf2f(static_cast<bf16_t>(hb[r * 4 + ni]));
#pragma unroll
          for (int o = 0; o < DO; ++o) part[o] = fmaf(h, w3[o][ni], part[o]);
        }
        // the 16 lanes of a row group (one DPP row) hold its 64 columns
        float* z = static_cast<float*>(g.C) + static_cast<int64_t>(row) * DO;
#pragma unroll
        for (int o = 0; o < DO; ++o) {
          const float v = row16_sum(part[o]);
          if ((l & 15) == 0 && row < g.M) atomicAdd(&z[o], v);
        }
      }
    }
    continue;
  }

  // BF16 outputs go through LDS (the staging area is free after the K loop)
  // and leave as 16-byte row segments instead of one 2-byte store per value.
  constexpr bool BF16_OUT = EPI == EPI_BF16 || EPI == EPI_BIAS_GELU || EPI == EPI_GELU_GRAD;
  constexpr int TS = BN + 8;   // LDS row stride of the output tile (BF16 elements)
  bf16_t* s_out = reinterpret_cast<bf16_t*>(s_stage);
  bf16_t* s_pre = s_out + BM * TS;
  const bool vec_out = BF16_OUT && g.N % 8 == 0 && g.ldc % 8 == 0 && (!g.pre_out || g.ldx % 8 == 0) &&
                       (!g.ct || (g.M % 8 == 0 && g.ldct % 8 == 0));
  float csum[4] = {0.f, 0.f, 0.f, 0.f};   // EPI_GELU_GRAD: column sums of this lane's outputs
  // VEC: BF16 outputs to the LDS tile (uniform choice, made once per tile)
#define MV_EPI_LOOP(VEC)                                                                            \
  _Pragma("unroll") for (int mi = 0; mi < 4; ++mi) {                                               \
    bf16_t pre_v[4][4];   /* EPI_GELU_GRAD: [r][ni], loaded before the row group's outputs */     \
    uint32_t gv_[16];     /* EPI_BIAS_GELU: [r][ni], the row group's activations */                \
    if constexpr (EPI == EPI_BIAS_GELU) {                                                         \
      uint32_t pbv_[16];                                                                          \
      _Pragma("unroll") for (int k = 0; k < 16; ++k) pbv_[k] = f2bf(acc[mi][k & 3][k >> 2] + bv[k & 3]); \
      gelu_bf16_n<16>(s_gelu, pbv_, gv_);                                                         \
    }                                                                                             \
    if constexpr (EPI == EPI_GELU_GRAD) {                                                         \
      _Pragma("unroll") for (int r = 0; r < 4; ++r) {                                             \
        const int row = min(bm + wr * 64 + mi * 16 + (l >> 4) * 4 + r, g.M - 1);                  \
        _Pragma("unroll") for (int ni = 0; ni < 4; ++ni) {                                        \
          const int col = min(bn + wc * 64 + ni * 16 + (l & 15), g.N - 1);                        \
          pre_v[r][ni] = g.pre_in[static_cast<int64_t>(row) * g.ldx + col];                       \
        }                                                                                         \
      }                                                                                           \
    }                                                                                             \
    _Pragma("unroll") for (int r = 0; r < 4; ++r) {                                               \
      const int lrow = wr * 64 + mi * 16 + (l >> 4) * 4 + r;                                      \
      const int row = bm + lrow;                                                                  \
      _Pragma("unroll") for (int ni = 0; ni < 4; ++ni) {                                          \
        const int lcol = wc * 64 + ni * 16 + (l & 15);                                            \
        const int col = bn + lcol;                                                                \
        const bool ok = row < g.M && col < g.N;                                                   \
        const float v = acc[mi][ni][r];                                                           \
        if constexpr (EPI == EPI_F32) {                                                           \
          if (ok) static_cast<float*>(g.C)[ks * g.split_stride + static_cast<int64_t>(row) * g.ldc + col] = v; \
        } else if constexpr (EPI == EPI_F32_ADD) {                                                \
          if (ok) atomicAdd(&static_cast<float*>(g.C)[static_cast<int64_t>(row) * g.ldc + col], v); \
        } else {                                                                                  \
          bf16_t o, pb = 0;                                                                       \
          if constexpr (EPI == EPI_BF16) {                                                        \
            o = f2bf(v);                                                                          \
          } else if constexpr (EPI == EPI_BIAS_GELU) {                                            \
            pb = f2bf(v + bv[ni]);                                                                \
            o = static_cast<bf16_t>(gv_[r * 4 + ni]);                                             \
          } else {                                                                                \
            o = f2bf(v * gelu_grad(bf2f(pre_v[r][ni])));                                          \
            if (ok) csum[ni] += bf2f(o);                                                          \
          }                                                                                       \
          if (VEC) {                                                                              \
            s_out[lrow * TS + lcol] = o;                                                          \
            if (EPI == EPI_BIAS_GELU && g.pre_out) s_pre[lrow * TS + lcol] = pb;                  \
          } else if (ok) {                                                                        \
            if (g.C) static_cast<bf16_t*>(g.C)[static_cast<int64_t>(row) * g.ldc + col] = o;      \
            if (g.ct) g.ct[static_cast<int64_t>(col) * g.ldct + row] = o;                         \
            if (EPI == EPI_BIAS_GELU && g.pre_out) g.pre_out[static_cast<int64_t>(row) * g.ldx + col] = pb; \
          }                                                                                       \
        }                                                                                         \
      }                                                                                           \
    }                                                                                             \
  }
  if (vec_out) {
    MV_EPI_LOOP(true)
  } else {
    MV_EPI_LOOP(false)
  }
#undef MV_EPI_LOOP
  if (vec_out) {
    __syncthreads();
    // 128 rows x 16 segments of 8 values; thread tid takes segments tid + 256 j
#pragma unroll
    for (int j = 0; j < (BM * BN / 8) / NT; ++j) {
      const int sgi = tid + NT * j;
      const int lrow = sgi >> 4, lcol = (sgi & 15) * 8;
      const int row = bm + lrow, col = bn + lcol;
      if (row < g.M && col < g.N) {
        if (g.C)
          *reinterpret_cast<float4*>(static_cast<bf16_t*>(g.C) + static_cast<int64_t>(row) * g.ldc + col) =
              *reinterpret_cast<const float4*>(s_out + lrow * TS + lcol);
        if (EPI == EPI_BIAS_GELU && g.pre_out)
          *reinterpret_cast<float4*>(g.pre_out + static_cast<int64_t>(row) * g.ldx + col) =
              *reinterpret_cast<const float4*>(s_pre + lrow * TS + lcol);
      }
    }
    if (g.ct) {
      // transposed copy: 64 column pairs x 16 segments of 8 rows; a lane
      // reads its pair of columns 4 bytes per row (consecutive lanes,
      // consecutive words: conflict-free) and splits them into the two
      // columns' 16-byte row segments (N is even here, so col + 1 < N)
#pragma unroll
      for (int j = 0; j < (BM * BN / 16) / NT; ++j) {
        const int sgi = tid + NT * j;
        const int lcol = (sgi & (BN / 2 - 1)) * 2, lrow = (sgi / (BN / 2)) * 8;
        const int row = bm + lrow, col = bn + lcol;
        if (row < g.M && col < g.N) {
          uint32_t x[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) x[e] = *reinterpret_cast<const uint32_t*>(s_out + (lrow + e) * TS + lcol);
          uint32_t w0[4], w1[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            w0[e] = (x[2 * e] & 0xFFFFu) | (x[2 * e + 1] << 16);
            w1[e] = (x[2 * e] >> 16) | (x[2 * e + 1] & 0xFFFF0000u);
          }
          *reinterpret_cast<uint4*>(g.ct + static_cast<int64_t>(col) * g.ldct + row) = make_uint4(w0[0], w0[1], w0[2], w0[3]);
          *reinterpret_cast<uint4*>(g.ct + static_cast<int64_t>(col + 1) * g.ldct + row) = make_uint4(w1[0], w1[1], w1[2], w1[3]);
        }
      }
    }
    __syncthreads();   // the next tile's staging overwrites the output tile
  }
  if (EPI == EPI_GELU_GRAD && g.colsum) {
    // lanes l, l^16, l^32, l^48 hold the same columns
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      float v = csum[ni];
      v += __shfl_xor(v, 16);
      v += __shfl_xor(v, 32);
      const int col = bn + wc * 64 + ni * 16 + (l & 15);
      if (l < 16 && col < g.N) atomicAdd(&g.colsum[col], v);
    }
  }
  }   // tiles
}

// out[c, r] = in[r, c] for a rows x cols BF16 matrix (64 x 64 tiles through LDS).
// 64 x 64 tiles through LDS. Where the rows are whole 16-byte groups (cols,
// ldi and ldo multiples of 8, the MLP's padded widths) each thread moves
// 8-element vectors both ways; otherwise 2-byte elements.
__global__ __launch_bounds__(256) void k_transpose_bf16(const bf16_t* __restrict__ in, int rows, int cols,
                                                        int64_t ldi, bf16_t* __restrict__ out, int64_t ldo) {
  __shared__ bf16_t tile[64][66];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const bool vec = (cols % 8) == 0 && (ldi % 8) == 0 && (ldo % 8) == 0 && (rows % 8) == 0;
  if (vec) {
    // load: thread t takes row t / 8 (+ 32), columns 8 (t % 8) .. + 7
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int y = (threadIdx.x >> 3) + 32 * h, x8 = (threadIdx.x & 7) * 8;
      const int r = r0 + y, c = c0 + x8;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (r < rows && c < cols) v = *reinterpret_cast<const uint4*>(in + static_cast<int64_t>(r) * ldi + c);
      const bf16_t* e = reinterpret_cast<const bf16_t*>(&v);
#pragma unroll
      for (int q = 0; q < 8; ++q) tile[y][x8 + q] = e[q];
    }
    __syncthreads();
    // store: thread t writes transposed row t / 8 (+ 32), entries 8 (t % 8) .. + 7
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int y = (threadIdx.x >> 3) + 32 * h, x8 = (threadIdx.x & 7) * 8;
      const int c = c0 + y, r = r0 + x8;
      uint4 v;
      bf16_t* e = reinterpret_cast<bf16_t*>(&v);
#pragma unroll
      for (int q = 0; q < 8; ++q) e[q] = tile[x8 + q][y];
      if (c < cols && r < rows) *reinterpret_cast<uint4*>(out + static_cast<int64_t>(c) * ldo + r) = v;
    }
    return;
  }
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int y = ty; y < 64; y += 4) {
    const int r = r0 + y, c = c0 + tx;
    tile[y][tx] = (r < rows && c < cols) ? in[static_cast<int64_t>(r) * ldi + c] : static_cast<bf16_t>(0);
  }
  __syncthreads();
  for (int y = ty; y < 64; y += 4) {
    const int c = c0 + y, r = r0 + tx;
    if (c < cols && r < rows) out[static_cast<int64_t>(c) * ldo + r] = tile[tx][y];
  }
}

// BF16, zero-padded copy of FP32 rows: out[r, 0:dp] from X[r, 0:d].
__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ X, int64_t n, int d, int dp,
                                                 bf16_t* __restrict__ out) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (r >= n) return;
  for (int t = l; t < dp; t += 64) out[r * dp + t] = t < d ? f2bf(X[r * d + t]) : static_cast<bf16_t>(0);
}

// out[p] = src[idx[p]] (rows of dp BF16), and yb[p] = Y[idx[p]] (dout FP32).
__global__ __launch_bounds__(256) void k_gather_x(const bf16_t* __restrict__ src, int dp, const int32_t* __restrict__ idx,
                                                  int P, bf16_t* __restrict__ out, const float* __restrict__ Y,
                                                  int dout, float* __restrict__ yb) {
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (p >= P) return;
  const int64_t s = idx[p];
  const uint4* in = reinterpret_cast<const uint4*>(src + s * dp);
  uint4* o = reinterpret_cast<uint4*>(out + static_cast<int64_t>(p) * dp);
  for (int t = l; t < dp / 8; t += 64) o[t] = in[t];
  if (Y && l < dout) yb[static_cast<int64_t>(p) * dout + l] = Y[s * dout + l];
}

// Z[p] = H2[p] W3^T + b3 (one wavefront per point; dout <= 4; W3 FP32).
// Output layer forward, Z[p, o] = sum_c H2[p, c] W3[o, c] + b3[o]: four
// points per wavefront, whose rows are all loaded before any sum (one memory
// latency per four points); lane l takes columns l + 64 j, and each (point,
// output) sum is reduced over the wavefront with DPP (lane 63 ends with it).
constexpr int HEADF_PPW = 4;
__device__ __forceinline__ float wave_sum_to_lane63(float v) {
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xF, 0xF, false));   // row_shr:1
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x112, 0xF, 0xF, false));   // row_shr:2
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x114, 0xF, 0xF, false));   // row_shr:4
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x118, 0xF, 0xF, false));   // row_shr:8
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0xA, 0xF, false));   // row_bcast:15
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x143, 0xC, 0xF, false));   // row_bcast:31
  return v;
}
template <int DOUT, int JM>   // JM: 64-column groups of the hidden layer (h2 <= 64 JM)
__global__ __launch_bounds__(256) void k_head_fwd(const bf16_t* __restrict__ H2, int P, int h2,
                                                  const float* __restrict__ W3, const float* __restrict__ b3,
                                                  float* __restrict__ Z) {
  constexpr int dout = DOUT;
  const int p0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * HEADF_PPW;
  const int l = threadIdx.x & 63;
  if (p0 >= P) return;
  float w3r[DOUT][JM];
#pragma unroll
  for (int j = 0; j < JM; ++j) {
    const int c = l + 64 * j;
#pragma unroll
    for (int o = 0; o < DOUT; ++o) w3r[o][j] = c < h2 ? W3[o * h2 + c] : 0.f;
  }
  float h[HEADF_PPW][JM];
#pragma unroll
  for (int i = 0; i < HEADF_PPW; ++i) {
    const int p = min(p0 + i, P - 1);
#pragma unroll
    for (int j = 0; j < JM; ++j) {
      const int c = l + 64 * j;
      h[i][j] = c < h2 ? bf2f(H2[static_cast<int64_t>(p) * h2 + c]) : 0.f;
    }
  }
#pragma unroll
  for (int i = 0; i < HEADF_PPW; ++i) {
    const int p = p0 + i;
#pragma unroll
    for (int o = 0; o < DOUT; ++o) {
      float acc = 0.f;
#pragma unroll
      for (int j = 0; j < JM; ++j) acc = fmaf(h[i][j], w3r[o][j], acc);
      acc = wave_sum_to_lane63(acc);
      if (l == 63 && p < P) Z[static_cast<int64_t>(p) * dout + o] = acc + b3[o];
    }
  }
}

// Loss gradient dZ for a batch of B edges: rows [0, B) are heads, [B, 2B)
// tails. The loss is the UMAP edge cross-entropy plus a regression term,
//   L = (1/B) sum_e [ log(1 + a d_e^(2b))
//                     - sum_s log(1 - 1/(1 + a d_es^(2b))) ]
//     + (lam/B) sum_rows ||z - y||^2,
// with d_e = ||z_head - z_tail|| (attraction) and d_es = ||z_head - z_t(s)||
// for the negatives t(s) = tails of edges (e + 1 + 97 s) mod B, s < neg (the
// in-batch tails, which are drawn in proportion to the edge weights). Its
// z-gradient is the layout's law (umap.py:367-380) without the +-4 clip:
// attraction -2ab d^(2(b-1)) / (1 + a d^(2b)) (z_i - z_j) on the head and the
// opposite on the tail; repulsion 2b / ((0.001 + d^2)(1 + a d^(2b))) (z_i - z_t)
// on the head only (negatives move the head only, as the layout). Squared
// distances are floored at 1e-6 as the layout does. loss_out (nullable)
// accumulates the two terms.
__global__ __launch_bounds__(256) void k_loss_grad(const float* __restrict__ Z, const float* __restrict__ Yb, int B,
                                                   int dout, int neg, float a, float b, float lam,
                                                   float* __restrict__ dZ, float* __restrict__ loss_out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  float ce = 0.f, reg = 0.f;
  if (e < B) {
    const float inv_b = 1.f / static_cast<float>(B);
    float zi[4] = {0, 0, 0, 0}, zj[4] = {0, 0, 0, 0}, gi[4] = {0, 0, 0, 0}, gj[4] = {0, 0, 0, 0};
    float d2 = 0.f;
    for (int q = 0; q < dout; ++q) {
      zi[q] = Z[static_cast<int64_t>(e) * dout + q];
      zj[q] = Z[static_cast<int64_t>(B + e) * dout + q];
      d2 += (zi[q] - zj[q]) * (zi[q] - zj[q]);
    }
    {
      const float dd = fmaxf(d2, 1e-6f);
      const float pw = exp2f(b * __log2f(dd));
      const float coef = -2.0f * a * b * exp2f((b - 1.0f) * __log2f(dd)) / (1.0f + a * pw);
      ce += logf(1.0f + a * pw);
      for (int q = 0; q < dout; ++q) {
        const float gq = coef * (zi[q] - zj[q]) * inv_b;   // minus the head's gradient
        gi[q] -= gq;
        gj[q] += gq;
      }
    }
    for (int s = 0; s < neg; ++s) {
      const int te = (e + 1 + s * 97) % B;
      if (te == e) continue;
      const int t = B + te;
      float n2 = 0.f, zt[4] = {0, 0, 0, 0};
      for (int q = 0; q < dout; ++q) {
        zt[q] = Z[static_cast<int64_t>(t) * dout + q];
        n2 += (zi[q] - zt[q]) * (zi[q] - zt[q]);
      }
      const float dd = fmaxf(n2, 1e-6f);
      const float pw = exp2f(b * __log2f(dd));
      const float coef = 2.0f * b / ((0.001f + dd) * (1.0f + a * pw));
      ce -= logf(fmaxf(1.0f - 1.0f / (1.0f + a * pw), 1e-30f));
      for (int q = 0; q < dout; ++q) gi[q] -= coef * (zi[q] - zt[q]) * inv_b;
    }
    for (int q = 0; q < dout; ++q) {
      const float ri = zi[q] - Yb[static_cast<int64_t>(e) * dout + q];
      const float rj = zj[q] - Yb[static_cast<int64_t>(B + e) * dout + q];
      reg += ri * ri + rj * rj;
      dZ[static_cast<int64_t>(e) * dout + q] = gi[q] + 2.f * lam * ri * inv_b;
      dZ[static_cast<int64_t>(B + e) * dout + q] = gj[q] + 2.f * lam * rj * inv_b;
    }
  }
  if (loss_out) {
    ce = wave_sum_f32(ce);
    reg = wave_sum_f32(reg);
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&loss_out[0], ce / static_cast<float>(B));
      atomicAdd(&loss_out[1], lam * reg / static_cast<float>(B));
    }
  }
}

constexpr int HMAX = 1024;        // widest hidden layer
constexpr int HEAD_PPW = 4;       // points per wavefront group in k_head_bwd
constexpr int HEAD_G = 2;         // groups per wavefront (fewer blocks: fewer gradient atomics)

// Output-layer backward for DOUT outputs. Per point p:
//   dA2[p, c] = (sum_o dZ[p, o] W3[o, c]) * gelu'(A2pre[p, c])   (stored BF16)
// and the block's partial sums of dW3 = dZ^T H2, db3 = sum dZ and
// db2 = sum dA2 (of the BF16-rounded values, as the GEMMs see them), reduced
// in LDS and added with one atomic per block and parameter.
template <int DOUT, int JM>   // JM: 64-column groups of the hidden layer (h2 <= 64 JM)
__global__ __launch_bounds__(256) void k_head_bwd(const float* __restrict__ dZ, const bf16_t* __restrict__ H2,
                                                  const bf16_t* __restrict__ A2pre, int P, int h2,
                                                  const float* __restrict__ W3, bf16_t* __restrict__ dA2,
                                                  float* __restrict__ gW3, float* __restrict__ gb3,
                                                  float* __restrict__ gb2) {
  __shared__ float s_w3[DOUT * HMAX];
  __shared__ float s_b2[HMAX];
  __shared__ float s_b3[DOUT];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  for (int t = tid; t < DOUT * h2; t += 256) s_w3[t] = 0.f;
  for (int t = tid; t < h2; t += 256) s_b2[t] = 0.f;
  if (tid < DOUT) s_b3[tid] = 0.f;
  __syncthreads();
  float pw3[DOUT][JM], pb2[JM], pb3[DOUT];
#pragma unroll
  for (int j = 0; j < JM; ++j) {
    pb2[j] = 0.f;
#pragma unroll
    for (int o = 0; o < DOUT; ++o) pw3[o][j] = 0.f;
  }
#pragma unroll
  for (int o = 0; o < DOUT; ++o) pb3[o] = 0.f;
  // W3 columns of this lane, read once
  float w3r[DOUT][JM];
#pragma unroll
  for (int j = 0; j < JM; ++j) {
    const int c = l + 64 * j;
#pragma unroll
    for (int o = 0; o < DOUT; ++o) w3r[o][j] = c < h2 ? W3[o * h2 + c] : 0.f;
  }
  for (int gi = 0; gi < HEAD_G; ++gi) {
    // the wavefront's HEAD_PPW points: every load first (one memory latency),
    // then the arithmetic
    const int p0 = ((blockIdx.x * HEAD_G + gi) * 4 + w) * HEAD_PPW;
    float dzs[HEAD_PPW][DOUT], a2v[HEAD_PPW][JM], hv[HEAD_PPW][JM];
  #pragma unroll
    for (int i = 0; i < HEAD_PPW; ++i) {
      const int p = min(p0 + i, P - 1);
  #pragma unroll
      for (int o = 0; o < DOUT; ++o) dzs[i][o] = dZ[static_cast<int64_t>(p) * DOUT + o];
  #pragma unroll
      for (int j = 0; j < JM; ++j) {
        const int c = l + 64 * j;
        const int64_t at = static_cast<int64_t>(p) * h2 + (c < h2 ? c : 0);
        a2v[i][j] = bf2f(A2pre[at]);
        hv[i][j] = bf2f(H2[at]);
      }
    }
  #pragma unroll
    for (int i = 0; i < HEAD_PPW; ++i) {
      const int p = p0 + i;
      if (p >= P) break;
  #pragma unroll
      for (int o = 0; o < DOUT; ++o) pb3[o] += dzs[i][o];
  #pragma unroll
      for (int j = 0; j < JM; ++j) {
        const int c = l + 64 * j;
        if (c < h2) {
          float s = 0.f;
  #pragma unroll
          for (int o = 0; o < DOUT; ++o) s = fmaf(dzs[i][o], w3r[o][j], s);
          const int64_t at = static_cast<int64_t>(p) * h2 + c;
          const bf16_t g = f2bf(s * gelu_grad(a2v[i][j]));
          dA2[at] = g;
          pb2[j] += bf2f(g);
  #pragma unroll
          for (int o = 0; o < DOUT; ++o) pw3[o][j] = fmaf(dzs[i][o], hv[i][j], pw3[o][j]);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < JM; ++j) {
    const int c = l + 64 * j;
    if (c < h2) {
      atomicAdd(&s_b2[c], pb2[j]);
#pragma unroll
      for (int o = 0; o < DOUT; ++o) atomicAdd(&s_w3[o * h2 + c], pw3[o][j]);
    }
  }
  if (l == 0) {
#pragma unroll
    for (int o = 0; o < DOUT; ++o) atomicAdd(&s_b3[o], pb3[o]);
  }
  __syncthreads();
  for (int t = tid; t < DOUT * h2; t += 256) atomicAdd(&gW3[t], s_w3[t]);
  for (int t = tid; t < h2; t += 256) atomicAdd(&gb2[t], s_b2[t]);
  if (tid < DOUT) atomicAdd(&gb3[tid], s_b3[tid]);
}

// Adam on the FP32 master weights (bias corrections c1 = 1 - b1^t,
// c2 = 1 - b2^t) over the whole parameter blob in one launch: W1 (with its
// BF16 copy), b1, W2 (BF16 copy and BF16 transpose), then b2, W3, b3.
struct AdamBlob {
  float* p;
  float* g;               // gradient blob; entries outside W1 / W2 are cleared after use
  const float* part1;     // W1 gradient: s1 partial slices of stride st1
  const float* part2;     // W2 gradient: s2 partial slices of stride st2
  int64_t st1, st2;
  int s1, s2;
  float* m;
  float* v;
  int64_t np, o_b1, o_W2, o_b2;
  int h1, h2;
  bf16_t* W1b;
  bf16_t* W2b;
  bf16_t* W2tb;
};
__global__ __launch_bounds__(256) void k_adam_blob(AdamBlob a, float lr, float b1, float b2, float eps, float c1,
                                                   float c2) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= a.np) return;
  float g = 0.f;
  if (t < a.o_b1) {
    for (int q = 0; q < a.s1; ++q) g += a.part1[q * a.st1 + t];
  } else if (t >= a.o_W2 && t < a.o_b2) {
    const int64_t e = t - a.o_W2;
    for (int q = 0; q < a.s2; ++q) g += a.part2[q * a.st2 + e];
  } else {
    g = a.g[t];
    a.g[t] = 0.f;   // ready for the next step's atomic sums
  }
  const float mt = b1 * a.m[t] + (1.f - b1) * g;
  const float vt = b2 * a.v[t] + (1.f - b2) * g * g;
  a.m[t] = mt;
  a.v[t] = vt;
  const float nv = a.p[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
  a.p[t] = nv;
  if (t < a.o_b1) {
    a.W1b[t] = f2bf(nv);
  } else if (t >= a.o_W2 && t < a.o_b2) {
    const int64_t e = t - a.o_W2;
    a.W2b[e] = f2bf(nv);
    const int r = static_cast<int>(e / a.h1), c = static_cast<int>(e % a.h1);
    a.W2tb[static_cast<int64_t>(c) * a.h2 + r] = f2bf(nv);
  }
}

// The W1 / W2 entries of the gradient blob from their partial slices (split
// order, as k_adam_blob sums them).
__global__ __launch_bounds__(256) void k_reduce_parts(AdamBlob a) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t nW2 = a.o_b2 - a.o_W2;
  if (t < a.o_b1) {
    float g = 0.f;
    for (int q = 0; q < a.s1; ++q) g += a.part1[q * a.st1 + t];
    a.g[t] = g;
  } else if (t - a.o_b1 < nW2) {
    const int64_t e = t - a.o_b1;
    float g = 0.f;
    for (int q = 0; q < a.s2; ++q) g += a.part2[q * a.st2 + e];
    a.g[a.o_W2 + e] = g;
  }
}

// Edge batches drawn from a fuzzy graph in its dense slot layout (slot e =
// i * k + c, head i, tail tail[e], eps[e] = epochs per sample = w_max / w,
// -1 for a pruned slot), in proportion to the weights, by rejection: a slot
// uniform over n * k is accepted with probability 1 / eps. Draw j of step t
// writes the head to out[t * 2B + j] and the tail to out[t * 2B + B + j].
// Randomness is Philox-4x32-10 on the counter (j, step0 + t, try, tag) under
// key (seed, ~seed), two candidates per call; after 64 rejected candidates the
// highest-weight kept candidate seen is taken (probability (1 - mean w)^64).
__global__ __launch_bounds__(256) void k_sample_edges(const float* __restrict__ eps, const int32_t* __restrict__ tail,
                                                      int64_t nk, int k, uint32_t seed, uint32_t step0, int steps,
                                                      int B, int32_t* __restrict__ out) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= static_cast<int64_t>(steps) * B) return;
  const int t = static_cast<int>(g / B), j = static_cast<int>(g % B);
  int64_t pick = -1, best = -1;
  float best_eps = INFINITY;
  for (int tr = 0; tr < 32 && pick < 0; ++tr) {
    const uint4 r = Philox::gen(make_uint4(static_cast<uint32_t>(j), step0 + static_cast<uint32_t>(t),
                                           static_cast<uint32_t>(tr), 0x4D4C5030u),
                                seed, ~seed);
    const uint32_t ra[2] = {r.x, r.z}, ru[2] = {r.y, r.w};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (pick >= 0) break;
      const int64_t e = uniform_below(ra[h], static_cast<uint32_t>(nk));
      const float ep = eps[e];
      if (ep >= 1.f) {
        const float u = static_cast<float>(ru[h] >> 8) * (1.f / 16777216.f);
        if (u * ep < 1.f) pick = e;
        if (ep < best_eps) {
          best_eps = ep;
          best = e;
        }
      }
    }
  }
  if (pick < 0) pick = best;
  if (pick < 0) {
    // none of the 64 draws hit a kept slot (a heavily pruned graph): probe
    // forward from a drawn slot to the next kept one. The maximum-weight slot
    // is always kept, so the probe ends within nk steps unless the graph has
    // no kept slot at all (then slot 0 is returned).
    const uint4 r = Philox::gen(make_uint4(static_cast<uint32_t>(j), step0 + static_cast<uint32_t>(t), 0xFFFFu,
                                           0x4D4C5030u),
                                seed, ~seed);
    const int64_t e0 = uniform_below(r.x, static_cast<uint32_t>(nk));
    for (int64_t q = 0; q < nk && pick < 0; ++q) {
      const int64_t e = e0 + q < nk ? e0 + q : e0 + q - nk;
      if (eps[e] >= 1.f) pick = e;
    }
    if (pick < 0) pick = 0;
  }
  int32_t* o = out + static_cast<int64_t>(t) * 2 * B;
  o[j] = static_cast<int32_t>(pick / k);
  o[B + j] = tail[pick];
}

// Per-row overlap of two k-NN lists (the 15-NN preservation count):
// count[i] = |{a[i, 0:ka]} & {b[i, 0:kb]}|, flag[i] = count[i] < min_count.
__global__ __launch_bounds__(256) void k_knn_overlap(const int32_t* __restrict__ A, int64_t lda, int ka,
                                                     const int32_t* __restrict__ Bm, int64_t ldb, int kb, int64_t n,
                                                     int min_count, int32_t* __restrict__ count,
                                                     uint8_t* __restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t* a = A + i * lda;
  const int32_t* b = Bm + i * ldb;
  int c = 0;
  for (int x = 0; x < ka; ++x) {
    const int32_t v = a[x];
    bool hit = false;
    for (int y = 0; y < kb; ++y) hit |= (b[y] == v);
    c += hit ? 1 : 0;
  }
  if (count) count[i] = c;
  if (flag) flag[i] = c < min_count ? 1 : 0;
}

__global__ void k_fill_f32(float* __restrict__ p, int64_t count, float v) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) p[t] = v;
}

__global__ void k_f32_to_bf16(const float* __restrict__ in, bf16_t* __restrict__ out, int64_t count) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < count) out[t] = f2bf(in[t]);
}

__global__ void k_f32_to_bf16_t(const float* __restrict__ in, int rows, int cols, bf16_t* __restrict__ out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < static_cast<int64_t>(rows) * cols) {
    const int r = static_cast<int>(t / cols), c = static_cast<int>(t % cols);
    out[static_cast<int64_t>(c) * rows + r] = f2bf(in[t]);
  }
}

// ------------------------------------------------------------------ host --

struct Dims {
  int din, dp, h1, h2, dout;
};

// Parameter blob: FP32 master (W1 h1 x dp, b1, W2 h2 x h1, b2, W3 dout x h2, b3),
// Adam m and v of the same shape, then BF16 copies W1b (h1 x dp), W2b (h2 x h1),
// W2tb (h1 x h2).
struct Params {
  float *W1, *b1, *W2, *b2, *W3, *b3;
  float *m, *v;
  bf16_t *W1b, *W2b, *W2tb;
  int64_t n_f32;   // FP32 parameter count (master region)
};

static int64_t n_params(const Dims& D) {
  return static_cast<int64_t>(D.h1) * D.dp + D.h1 + static_cast<int64_t>(D.h2) * D.h1 + D.h2 +
         static_cast<int64_t>(D.dout) * D.h2 + D.dout;
}

static Params carve_params(void* blob, const Dims& D, size_t* total) {
  Carve c(blob);
  Params P{};
  const int64_t np = n_params(D);
  P.n_f32 = np;
  float* base = c.take<float>(np);
  P.W1 = base;
  P.b1 = P.W1 + static_cast<int64_t>(D.h1) * D.dp;
  P.W2 = P.b1 + D.h1;
  P.b2 = P.W2 + static_cast<int64_t>(D.h2) * D.h1;
  P.W3 = P.b2 + D.h2;
  P.b3 = P.W3 + static_cast<int64_t>(D.dout) * D.h2;
  P.m = c.take<float>(np);
  P.v = c.take<float>(np);
  P.W1b = c.take<bf16_t>(static_cast<int64_t>(D.h1) * D.dp);
  P.W2b = c.take<bf16_t>(static_cast<int64_t>(D.h2) * D.h1);
  P.W2tb = c.take<bf16_t>(static_cast<int64_t>(D.h1) * D.h2);
  if (total) *total = c.off + 256;
  return P;
}

// K splits actually used for a request of `req` over K (every split non-empty).
static int split_count(int K, int req) {
  const int nk = K / 64;
  if (nk < 1) return 1;
  if (req > nk) req = nk;
  if (req < 1) req = 1;
  const int kch = static_cast<int>(ceil_div(nk, req));
  return static_cast<int>(ceil_div(nk, kch));
}

// K splits for a weight-gradient GEMM: about 256 workgroups in total.
static int grad_splits(int M, int N, int K) {
  const int64_t tiles = ceil_div(M, BM) * ceil_div(N, BN);
  int sp = static_cast<int>(256 / (tiles > 0 ? tiles : 1));
  return sp < 1 ? 1 : sp;
}

// Per-batch buffers for P points.
struct Batch {
  bf16_t *Xb, *A1, *H1, *A2, *H2, *dA2, *dA2t, *H1t, *dA1t, *Xbt;
  float *Z, *Yb, *dZ, *grad, *loss;
  float *part1, *part2;   // K-split partial slices of the W1 / W2 gradients
  int s1, s2;             // their split counts
  int32_t* idx;
};

static Batch carve_batch(void* ws, const Dims& D, int P, int64_t np, size_t* total) {
  Carve c(ws);
  Batch b{};
  b.Xb = c.take<bf16_t>(static_cast<int64_t>(P) * D.dp);
  b.A1 = c.take<bf16_t>(static_cast<int64_t>(P) * D.h1);
  b.H1 = c.take<bf16_t>(static_cast<int64_t>(P) * D.h1);
  b.A2 = c.take<bf16_t>(static_cast<int64_t>(P) * D.h2);
  b.H2 = c.take<bf16_t>(static_cast<int64_t>(P) * D.h2);
  b.dA2 = c.take<bf16_t>(static_cast<int64_t>(P) * D.h2);
  b.dA2t = c.take<bf16_t>(static_cast<int64_t>(P) * D.h2);
  b.H1t = c.take<bf16_t>(static_cast<int64_t>(P) * D.h1);
  b.dA1t = c.take<bf16_t>(static_cast<int64_t>(P) * D.h1);
  b.Xbt = c.take<bf16_t>(static_cast<int64_t>(P) * D.dp);
  b.Z = c.take<float>(static_cast<int64_t>(P) * D.dout);
  b.Yb = c.take<float>(static_cast<int64_t>(P) * D.dout);
  b.dZ = c.take<float>(static_cast<int64_t>(P) * D.dout);
  b.grad = c.take<float>(np);
  b.s1 = split_count(P, grad_splits(D.h1, D.dp, P));
  b.s2 = split_count(P, grad_splits(D.h2, D.h1, P));
  b.part1 = c.take<float>(static_cast<int64_t>(b.s1) * D.h1 * D.dp);
  b.part2 = c.take<float>(static_cast<int64_t>(b.s2) * D.h2 * D.h1);
  b.loss = c.take<float>(2);
  b.idx = c.take<int32_t>(P);
  if (total) *total = align_up(c.off, 256) + 256;   // a multiple of 256: data placed after it stays aligned
  return b;
}

static int cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cached[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cached[dev] = v;
  }
  return cached[dev];
}

// Launches k_gemm_nt over ceil(N/128) x ceil(M/128) tiles and `splits` K
// ranges (splits > 1 with EPI_F32_ADD, or EPI_F32 into split_stride-spaced
// partial slices) on a persistent grid.
template <int EPI, bool AF32 = false, int DO = 4>
static int gemm(GemmArgs g, hipStream_t s, int splits = 1) {
  const int nk = g.K / 64;
  splits = split_count(g.K, splits);
  g.kchunks = static_cast<int>(ceil_div(nk, splits));
  g.m_tiles = static_cast<int>(ceil_div(g.M, BM));
  g.n_tiles = static_cast<int>(ceil_div(g.N, BN));
  g.splits = splits;
  const int64_t slots = 8 * ceil_div(g.m_tiles, 8) * static_cast<int64_t>(g.n_tiles) * splits;
  const int64_t resident = 2 * static_cast<int64_t>(cu_count());   // two workgroups per CU (LDS)
  const int64_t grid = std::max<int64_t>(8, (std::min(slots, resident) + 7) / 8 * 8);
  hipLaunchKernelGGL((k_gemm_nt<EPI, AF32, DO>), dim3(static_cast<unsigned>(grid)), dim3(NT), 0, s, g);
  MV_LAUNCHED("k_gemm_nt");
  return MV_OK;
}

static int transpose(const bf16_t* in, int rows, int cols, int64_t ldi, bf16_t* out, int64_t ldo, hipStream_t s) {
  dim3 grid(static_cast<unsigned>(ceil_div(cols, 64)), static_cast<unsigned>(ceil_div(rows, 64)));
  hipLaunchKernelGGL(k_transpose_bf16, grid, dim3(256), 0, s, in, rows, cols, ldi, out, ldo);
  MV_LAUNCHED("k_transpose_bf16");
  return MV_OK;
}

// Forward of P rows already in Xb (BF16, stride dp): A1/H1, A2/H2, Z.
static int forward(const Params& W, const Dims& D, const bf16_t* Xb, int P, bf16_t* A1, bf16_t* H1, bf16_t* H1t,
                   bf16_t* A2, bf16_t* H2, float* Z, hipStream_t s) {
  GemmArgs g{};
  g.A = Xb;
  g.B = W.W1b;
  g.M = P;
  g.N = D.h1;
  g.K = D.dp;
  g.lda = D.dp;
  g.ldb = D.dp;
  g.C = H1;
  g.ldc = D.h1;
  g.bias = W.b1;
  g.pre_out = A1;
  g.ldx = D.h1;
  g.ct = H1t;   // H1^T for the weight gradient of layer 2 (rows of P)
  g.ldct = P;
  if (int rc = gemm<EPI_BIAS_GELU>(g, s)) return rc;
  g.ct = nullptr;
  g.A = H1;
  g.B = W.W2b;
  g.N = D.h2;
  g.K = D.h1;
  g.lda = D.h1;
  g.ldb = D.h1;
  g.C = H2;
  g.ldc = D.h2;
  g.bias = W.b2;
  g.pre_out = A2;
  g.ldx = D.h2;
  if (int rc = gemm<EPI_BIAS_GELU>(g, s)) return rc;
  {
    const dim3 hg(static_cast<unsigned>(ceil_div(P, 4 * HEADF_PPW)));
    const bool narrow = D.h2 <= 512;
#define MV_HEADF(DO)                                                                                 \
  if (narrow) hipLaunchKernelGGL((k_head_fwd<DO, 8>), hg, dim3(256), 0, s, H2, P, D.h2, W.W3, W.b3, Z); \
  else hipLaunchKernelGGL((k_head_fwd<DO, 16>), hg, dim3(256), 0, s, H2, P, D.h2, W.W3, W.b3, Z)
    switch (D.dout) {
      case 1: MV_HEADF(1); break;
      case 2: MV_HEADF(2); break;
      case 3: MV_HEADF(3); break;
      default: MV_HEADF(4); break;
    }
#undef MV_HEADF
  }
  MV_LAUNCHED("k_head_fwd");
  return MV_OK;
}

// Z[r, o] = b3[o] for the rows of an inference chunk (the fused head adds to it).
__global__ void k_init_rows(float* __restrict__ Z, int64_t rows, int dout, const float* __restrict__ b3) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < rows * dout) Z[t] = b3[t % dout];
}

// Inference of `rows` FP32 rows (row stride din): GEMM 1 reads the FP32 rows
// directly (rounded to BF16 while staged; din a multiple of 8), GEMM 2 ends in
// the output layer (no H2 in memory). Xb is used only when din % 8 != 0.
static int forward_infer(const Params& W, const Dims& D, const float* X, int rows, bf16_t* Xb, bf16_t* H1, float* Z,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_init_rows, dim3(static_cast<unsigned>(ceil_div(static_cast<int64_t>(rows) * D.dout, 256))),
                     dim3(256), 0, s, Z, static_cast<int64_t>(rows), D.dout, W.b3);
  MV_LAUNCHED("k_init_rows");
  GemmArgs g{};
  g.B = W.W1b;
  g.M = rows;
  g.N = D.h1;
  g.K = D.dp;
  g.ldb = D.dp;
  g.C = H1;
  g.ldc = D.h1;
  g.bias = W.b1;
  if (D.din % 8 == 0) {
    g.A = X;
    g.lda = D.din;
    g.kvalid = D.din;
    if (int rc = gemm<EPI_BIAS_GELU, true>(g, s)) return rc;
  } else {
    hipLaunchKernelGGL(k_to_bf16, dim3(static_cast<unsigned>(ceil_div(rows, 4))), dim3(256), 0, s, X,
                       static_cast<int64_t>(rows), D.din, D.dp, Xb);
    MV_LAUNCHED("k_to_bf16");
    g.A = Xb;
    g.lda = D.dp;
    if (int rc = gemm<EPI_BIAS_GELU>(g, s)) return rc;
  }
  g = GemmArgs{};
  g.A = H1;
  g.B = W.W2b;
  g.M = rows;
  g.N = D.h2;
  g.K = D.h1;
  g.lda = D.h1;
  g.ldb = D.h1;
  g.C = Z;
  g.bias = W.b2;
  g.W3 = W.W3;
  g.dout = D.dout;
  switch (D.dout) {
    case 1: return gemm<EPI_HEAD, false, 1>(g, s);
    case 2: return gemm<EPI_HEAD, false, 2>(g, s);
    case 3: return gemm<EPI_HEAD, false, 3>(g, s);
    default: return gemm<EPI_HEAD, false, 4>(g, s);
  }
}

// Backward from dZ into the gradient buffer (zeroed here).
static int backward(const Params& W, const Dims& D, const Batch& b, int P, int64_t np, hipStream_t s) {
  float* gW1 = b.grad;
  float* gb1 = gW1 + static_cast<int64_t>(D.h1) * D.dp;
  float* gW2 = gb1 + D.h1;
  float* gb2 = gW2 + static_cast<int64_t>(D.h2) * D.h1;
  float* gW3 = gb2 + D.h2;
  float* gb3 = gW3 + static_cast<int64_t>(D.dout) * D.h2;
  // b.grad's bias and output-layer entries receive atomic sums and must be
  // zero here (the caller clears them); the W1 / W2 gradients go to the
  // K-split partial slices b.part1 / b.part2 (summed by k_adam_blob or
  // k_reduce_parts in split order).
  {
    const dim3 grid(static_cast<unsigned>(ceil_div(P, 4 * HEAD_PPW * HEAD_G)));
    const bool narrow = D.h2 <= 512;
    switch (D.dout) {
      case 1: if (narrow) hipLaunchKernelGGL((k_head_bwd<1, 8>), grid, dim3(256), 0, s, b.dZ, b.H2, b.A2, P, D.h2, W.W3, b.dA2, gW3, gb3, gb2);
              else hipLaunchKernelGGL((k_head_bwd<1, 16>), grid, dim3(256), 0, s, b.dZ, b.H2, b.A2, P, D.h2, W.W3, b.dA2, gW3, gb3, gb2);
              break;
      case 2: if (narrow) hipLaunchKernelGGL((k_head_bwd<2, 8>), grid, dim3(256), 0, s, b.dZ, b.H2, b.A2, P, D.h2, W.W3, b.dA2, gW3, gb3, gb2);
              else hipLaunchKernelGGL((k_head_bwd<2, 16>), grid, dim3(256), 0, s, b.dZ, b.H2, b.A2, P, D.h2, W.W3, b.dA2, gW3, gb3, gb2);
              break;
      case 3: if (narrow) hipLaunchKernelGGL((k_head_bwd<3, 8>), grid, dim3(256), 0, s, b.dZ, b.H2, b.A2, P, D.h2, W.W3, b.dA2, gW3, gb3, gb2);
              else hipLaunchKernelGGL((k_head_bwd<3, 16>), grid, dim3(256), 0, s, b.dZ, b.H2, b.A2, P, D.h2, W.W3, b.dA2, gW3, gb3, gb2);
              break;
      default: if (narrow) hipLaunchKernelGGL((k_head_bwd<4, 8>), grid, dim3(256), 0, s, b.dZ, b.H2, b.A2, P, D.h2, W.W3, b.dA2, gW3, gb3, gb2);
               else hipLaunchKernelGGL((k_head_bwd<4, 16>), grid, dim3(256), 0, s, b.dZ, b.H2, b.A2, P, D.h2, W.W3, b.dA2, gW3, gb3, gb2);
               break;
    }
    MV_LAUNCHED("k_head_bwd");
  }
  if (int rc = transpose(b.dA2, P, D.h2, D.h2, b.dA2t, P, s)) return rc;
  GemmArgs g{};
  // gW2 (h2 x h1) = dA2^T H1
  g.A = b.dA2t;
  g.B = b.H1t;
  g.M = D.h2;
  g.N = D.h1;
  g.K = P;
  g.lda = P;
  g.ldb = P;
  g.C = b.part2;
  g.ldc = D.h1;
  g.split_stride = static_cast<int64_t>(D.h2) * D.h1;
  if (int rc = gemm<EPI_F32>(g, s, b.s2)) return rc;
  (void)gW2;
  // dA1 (P x h1) = (dA2 W2) * gelu'(A1) = dA2 (W2^T)^T
  g = GemmArgs{};
  g.A = b.dA2;
  g.B = W.W2tb;
  g.M = P;
  g.N = D.h1;
  g.K = D.h2;
  g.lda = D.h2;
  g.ldb = D.h2;
  g.C = nullptr;   // only dA1^T is needed (the weight gradient of layer 1)
  g.ldc = D.h1;
  g.ct = b.dA1t;
  g.ldct = P;
  g.pre_in = b.A1;
  g.ldx = D.h1;
  g.colsum = gb1;
  if (int rc = gemm<EPI_GELU_GRAD>(g, s)) return rc;
  if (int rc = transpose(b.Xb, P, D.dp, D.dp, b.Xbt, P, s)) return rc;
  // gW1 (h1 x dp) = dA1^T Xb
  g = GemmArgs{};
  g.A = b.dA1t;
  g.B = b.Xbt;
  g.M = D.h1;
  g.N = D.dp;
  g.K = P;
  g.lda = P;
  g.ldb = P;
  g.C = b.part1;
  g.ldc = D.dp;
  g.split_stride = static_cast<int64_t>(D.h1) * D.dp;
  return gemm<EPI_F32>(g, s, b.s1);
}

static AdamBlob adam_blob(const Params& W, const Dims& D, const Batch& b, int64_t np) {
  const int64_t nW1 = static_cast<int64_t>(D.h1) * D.dp;
  const int64_t nW2 = static_cast<int64_t>(D.h2) * D.h1;
  AdamBlob ab;
  ab.p = W.W1;
  ab.g = b.grad;
  ab.part1 = b.part1;
  ab.part2 = b.part2;
  ab.st1 = nW1;
  ab.st2 = nW2;
  ab.s1 = b.s1;
  ab.s2 = b.s2;
  ab.m = W.m;
  ab.v = W.v;
  ab.np = np;
  ab.o_b1 = nW1;
  ab.o_W2 = nW1 + D.h1;
  ab.o_b2 = ab.o_W2 + nW2;
  ab.h1 = D.h1;
  ab.h2 = D.h2;
  ab.W1b = W.W1b;
  ab.W2b = W.W2b;
  ab.W2tb = W.W2tb;
  return ab;
}

static int adam_all(const Params& W, const Dims& D, const Batch& bt, int64_t np, float lr, int step, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, static_cast<float>(step));
  const float c2 = 1.f - powf(b2, static_cast<float>(step));
  const AdamBlob ab = adam_blob(W, D, bt, np);
  hipLaunchKernelGGL(k_adam_blob, dim3(static_cast<unsigned>(ceil_div(np, 256))), dim3(256), 0, s, ab, lr, b1, b2,
                     eps, c1, c2);
  MV_LAUNCHED("k_adam_blob");
  return MV_OK;
}

static int check_dims(int din, int h1, int h2, int dout) {
  MV_REQUIRE(din >= 1 && din <= 65536 && h1 >= 64 && h2 >= 64 && h1 % 64 == 0 && h2 % 64 == 0 && h1 <= HMAX &&
                 h2 <= HMAX && dout >= 1 && dout <= 4,
             "mlp: unsupported dims din=%d h1=%d h2=%d dout=%d (hidden: multiples of 64 up to %d; dout <= 4)", din,
             h1, h2, dout, HMAX);
  return MV_OK;
}

static Dims make_dims(int din, int h1, int h2, int dout) {
  Dims D;
  D.din = din;
  D.dp = static_cast<int>(align_up(din, 64));
  D.h1 = h1;
  D.h2 = h2;
  D.dout = dout;
  return D;
}

}  // namespace mlp
}  // namespace mv

using namespace mv;
using namespace mv::mlp;

extern "C" {

int mv_mlp_param_bytes(int din, int h1, int h2, int dout, size_t* out_bytes, int64_t* out_n_params) {
  MV_REQUIRE(out_bytes, "mv_mlp_param_bytes: null output");
  if (int rc = check_dims(din, h1, h2, dout)) return rc;
  Dims D = make_dims(din, h1, h2, dout);
  size_t total = 0;
  carve_params(nullptr, D, &total);
  *out_bytes = total;
  if (out_n_params) *out_n_params = n_params(D);
  return MV_OK;
}

int mv_mlp_batch_bytes(int din, int h1, int h2, int dout, int points, size_t* out_bytes) {
  MV_REQUIRE(out_bytes, "mv_mlp_batch_bytes: null output");
  if (int rc = check_dims(din, h1, h2, dout)) return rc;
  MV_REQUIRE(points >= 1, "mv_mlp_batch_bytes: points < 1");
  Dims D = make_dims(din, h1, h2, dout);
  size_t total = 0;
  carve_batch(nullptr, D, points, n_params(D), &total);
  *out_bytes = total;
  return MV_OK;
}

int mv_mlp_init(void* blob, int din, int h1, int h2, int dout, const float* W1, const float* b1, const float* W2,
                const float* b2, const float* W3, const float* b3, void* stream) {
  MV_REQUIRE(blob && W1 && b1 && W2 && b2 && W3 && b3, "mv_mlp_init: null pointer");
  if (int rc = check_dims(din, h1, h2, dout)) return rc;
  MV_FAULT_POINT("launch");
  Dims D = make_dims(din, h1, h2, dout);
  Params W = carve_params(blob, D, nullptr);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t np = W.n_f32;
  for (float* region : {W.W1, W.m, W.v}) {
    hipLaunchKernelGGL(k_fill_f32, dim3(static_cast<unsigned>(ceil_div(np, 256))), dim3(256), 0, s, region, np, 0.f);
    MV_LAUNCHED("k_fill_f32");
  }
  // W1 rows are padded from din to dp columns (zeros).
  MV_TRY_HIP(hipMemcpy2DAsync(W.W1, sizeof(float) * D.dp, W1, sizeof(float) * din, sizeof(float) * din, h1,
                              hipMemcpyDeviceToDevice, s));
  MV_TRY_HIP(hipMemcpyAsync(W.b1, b1, sizeof(float) * h1, hipMemcpyDeviceToDevice, s));
  MV_TRY_HIP(hipMemcpyAsync(W.W2, W2, sizeof(float) * h2 * h1, hipMemcpyDeviceToDevice, s));
  MV_TRY_HIP(hipMemcpyAsync(W.b2, b2, sizeof(float) * h2, hipMemcpyDeviceToDevice, s));
  MV_TRY_HIP(hipMemcpyAsync(W.W3, W3, sizeof(float) * dout * h2, hipMemcpyDeviceToDevice, s));
  MV_TRY_HIP(hipMemcpyAsync(W.b3, b3, sizeof(float) * dout, hipMemcpyDeviceToDevice, s));
  const int64_t nW1 = static_cast<int64_t>(h1) * D.dp, nW2 = static_cast<int64_t>(h2) * h1;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(static_cast<unsigned>(ceil_div(nW1, 256))), dim3(256), 0, s, W.W1, W.W1b, nW1);
  MV_LAUNCHED("k_f32_to_bf16");
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(static_cast<unsigned>(ceil_div(nW2, 256))), dim3(256), 0, s, W.W2, W.W2b, nW2);
  MV_LAUNCHED("k_f32_to_bf16");
  hipLaunchKernelGGL(k_f32_to_bf16_t, dim3(static_cast<unsigned>(ceil_div(nW2, 256))), dim3(256), 0, s, W.W2, h2, h1,
                     W.W2tb);
  MV_LAUNCHED("k_f32_to_bf16_t");
  return MV_OK;
}

// One forward/backward on a batch of B edges (rows idx[0:2B] of X, heads
// then tails), writing the FP32 gradient (parameter layout of the blob) into
// grad_out and the two loss terms into loss_out (both device, nullable).
// Used to check gradients; mv_mlp_train runs the same step with Adam.
int mv_mlp_grad(const void* blob, int din, int h1, int h2, int dout, const float* X, int64_t n, const float* Yp,
                const int32_t* idx, int B, int neg, float a, float b, float lam, float* grad_out, float* loss_out,
                void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(blob && X && Yp && idx, "mv_mlp_grad: null pointer");
  if (int rc = check_dims(din, h1, h2, dout)) return rc;
  MV_REQUIRE(B >= 32 && B % 32 == 0 && neg >= 0 && n >= 1 && n <= INT32_MAX,
             "mv_mlp_grad: batch B=%d must be a positive multiple of 32 (n=%lld)", B, (long long)n);
  Dims D = make_dims(din, h1, h2, dout);
  const int P = 2 * B;
  const int64_t np = n_params(D);
  size_t need = 0;
  carve_batch(nullptr, D, P, np, &need);
  const size_t xbytes = align_up(static_cast<size_t>(n) * D.dp * 2, 256);
  if (!ws || ws_bytes < need + xbytes) {
    mv::set_error("mv_mlp_grad: workspace of %zu bytes < required %zu", ws_bytes, need + xbytes);
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Params W = carve_params(const_cast<void*>(blob), D, nullptr);
  Batch bt = carve_batch(ws, D, P, np, nullptr);
  bf16_t* Xbf = reinterpret_cast<bf16_t*>(static_cast<char*>(ws) + need);
  hipLaunchKernelGGL(k_to_bf16, dim3(static_cast<unsigned>(ceil_div(n, 4))), dim3(256), 0, s, X, n, din, D.dp, Xbf);
  MV_LAUNCHED("k_to_bf16");
  hipLaunchKernelGGL(k_gather_x, dim3(static_cast<unsigned>(ceil_div(P, 4))), dim3(256), 0, s, Xbf, D.dp, idx, P,
                     bt.Xb, Yp, dout, bt.Yb);
  MV_LAUNCHED("k_gather_x");
  if (int rc = forward(W, D, bt.Xb, P, bt.A1, bt.H1, bt.H1t, bt.A2, bt.H2, bt.Z, s)) return rc;
  hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, s, bt.loss, static_cast<int64_t>(2), 0.f);
  MV_LAUNCHED("k_fill_f32");
  hipLaunchKernelGGL(k_loss_grad, dim3(static_cast<unsigned>(ceil_div(B, 256))), dim3(256), 0, s, bt.Z, bt.Yb, B, dout,
                     neg, a, b, lam, bt.dZ, bt.loss);
  MV_LAUNCHED("k_loss_grad");
  hipLaunchKernelGGL(k_fill_f32, dim3(static_cast<unsigned>(ceil_div(np, 256))), dim3(256), 0, s, bt.grad, np, 0.f);
  MV_LAUNCHED("k_fill_f32");
  if (int rc = backward(W, D, bt, P, np, s)) return rc;
  {
    const AdamBlob ab = adam_blob(W, D, bt, np);
    const int64_t nw = ab.o_b1 + (ab.o_b2 - ab.o_W2);
    hipLaunchKernelGGL(k_reduce_parts, dim3(static_cast<unsigned>(ceil_div(nw, 256))), dim3(256), 0, s, ab);
    MV_LAUNCHED("k_reduce_parts");
  }
  if (grad_out) MV_TRY_HIP(hipMemcpyAsync(grad_out, bt.grad, sizeof(float) * np, hipMemcpyDeviceToDevice, s));
  if (loss_out) MV_TRY_HIP(hipMemcpyAsync(loss_out, bt.loss, sizeof(float) * 2, hipMemcpyDeviceToDevice, s));
  return MV_OK;
}

// Trains for `steps` Adam steps; batch t uses idx[t * 2B : (t + 1) * 2B]
// (heads then tails, rows of X / Yp). loss_trace (device, nullable, 2 per
// step) receives the loss terms of every step.
int mv_mlp_train(void* blob, int din, int h1, int h2, int dout, const float* X, int64_t n, const float* Yp,
                 const int32_t* idx, int steps, int B, int neg, float a, float b, float lam, float lr, int step0,
                 float* loss_trace, void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(blob && X && Yp && idx, "mv_mlp_train: null pointer");
  if (int rc = check_dims(din, h1, h2, dout)) return rc;
  MV_REQUIRE(B >= 32 && B % 32 == 0 && neg >= 0 && n >= 1 && n <= INT32_MAX && steps >= 0 && step0 >= 0,
             "mv_mlp_train: batch B=%d must be a positive multiple of 32 (n=%lld, steps=%d)", B, (long long)n, steps);
  Dims D = make_dims(din, h1, h2, dout);
  const int P = 2 * B;
  const int64_t np = n_params(D);
  size_t need = 0;
  carve_batch(nullptr, D, P, np, &need);
  const size_t xbytes = align_up(static_cast<size_t>(n) * D.dp * 2, 256);
  if (!ws || ws_bytes < need + xbytes) {
    mv::set_error("mv_mlp_train: workspace of %zu bytes < required %zu", ws_bytes, need + xbytes);
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Params W = carve_params(blob, D, nullptr);
  Batch bt = carve_batch(ws, D, P, np, nullptr);
  bf16_t* Xbf = reinterpret_cast<bf16_t*>(static_cast<char*>(ws) + need);
  hipLaunchKernelGGL(k_to_bf16, dim3(static_cast<unsigned>(ceil_div(n, 4))), dim3(256), 0, s, X, n, din, D.dp, Xbf);
  MV_LAUNCHED("k_to_bf16");
  // the gradient blob starts at zero; k_adam_blob clears what it consumed
  hipLaunchKernelGGL(k_fill_f32, dim3(static_cast<unsigned>(ceil_div(np, 256))), dim3(256), 0, s, bt.grad, np, 0.f);
  MV_LAUNCHED("k_fill_f32");
  if (loss_trace && steps > 0) {
    const int64_t nl = 2 * static_cast<int64_t>(steps);
    hipLaunchKernelGGL(k_fill_f32, dim3(static_cast<unsigned>(ceil_div(nl, 256))), dim3(256), 0, s, loss_trace, nl, 0.f);
    MV_LAUNCHED("k_fill_f32");
  }
  for (int t = 0; t < steps; ++t) {
    hipLaunchKernelGGL(k_gather_x, dim3(static_cast<unsigned>(ceil_div(P, 4))), dim3(256), 0, s, Xbf, D.dp,
                       idx + static_cast<int64_t>(t) * P, P, bt.Xb, Yp, dout, bt.Yb);
    MV_LAUNCHED("k_gather_x");
    if (int rc = forward(W, D, bt.Xb, P, bt.A1, bt.H1, bt.H1t, bt.A2, bt.H2, bt.Z, s)) return rc;
    float* lo = loss_trace ? loss_trace + 2 * static_cast<int64_t>(t) : nullptr;
    hipLaunchKernelGGL(k_loss_grad, dim3(static_cast<unsigned>(ceil_div(B, 256))), dim3(256), 0, s, bt.Z, bt.Yb, B,
                       dout, neg, a, b, lam, bt.dZ, lo);
    MV_LAUNCHED("k_loss_grad");
    if (int rc = backward(W, D, bt, P, np, s)) return rc;
    if (int rc = adam_all(W, D, bt, np, lr, step0 + t + 1, s)) return rc;
  }
  return MV_OK;
}

// Inference: Z (m x dout) for rows of X (m x din FP32), in chunks of `chunk`
// rows (ws sized by mv_mlp_batch_bytes(..., chunk) plus chunk * dp * 2 bytes).
int mv_mlp_forward(const void* blob, int din, int h1, int h2, int dout, const float* X, int64_t m, float* Z,
                   int chunk, void* ws, size_t ws_bytes, void* stream) {
  MV_REQUIRE(blob && X && Z, "mv_mlp_forward: null pointer");
  if (int rc = check_dims(din, h1, h2, dout)) return rc;
  MV_REQUIRE(m >= 0 && chunk >= 1, "mv_mlp_forward: bad sizes");
  Dims D = make_dims(din, h1, h2, dout);
  const int64_t np = n_params(D);
  size_t need = 0;
  carve_batch(nullptr, D, chunk, np, &need);
  if (!ws || ws_bytes < need) {
    mv::set_error("mv_mlp_forward: workspace of %zu bytes < required %zu", ws_bytes, need);
    return MV_EWORKSPACE;
  }
  MV_FAULT_POINT("launch");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Params W = carve_params(const_cast<void*>(blob), D, nullptr);
  Batch bt = carve_batch(ws, D, chunk, np, nullptr);
  for (int64_t r0 = 0; r0 < m; r0 += chunk) {
    const int rows = static_cast<int>(m - r0 < chunk ? m - r0 : chunk);
    if (int rc = forward_infer(W, D, X + r0 * din, rows, bt.Xb, bt.H1, Z + r0 * dout, s)) return rc;
  }
  return MV_OK;
}

int mv_mlp_sample_edges(const float* eps, const int32_t* tail, int64_t n, int k, uint32_t seed, int step0, int steps,
                        int B, int32_t* out_idx, void* stream) {
  MV_REQUIRE(eps && tail && out_idx, "mv_mlp_sample_edges: null pointer");
  MV_REQUIRE(n >= 1 && k >= 1 && n * k < (int64_t(1) << 32) && steps >= 0 && B >= 1 && step0 >= 0,
             "mv_mlp_sample_edges: bad sizes n=%lld k=%d steps=%d B=%d", (long long)n, k, steps, B);
  MV_FAULT_POINT("launch");
  const int64_t count = static_cast<int64_t>(steps) * B;
  if (count == 0) return MV_OK;
  hipLaunchKernelGGL(k_sample_edges, dim3(static_cast<unsigned>(ceil_div(count, 256))), dim3(256), 0,
                     static_cast<hipStream_t>(stream), eps, tail, n * k, k, seed, static_cast<uint32_t>(step0), steps,
                     B, out_idx);
  MV_LAUNCHED("k_sample_edges");
  return MV_OK;
}

int mv_knn_overlap(const int32_t* a, int64_t lda, int ka, const int32_t* b, int64_t ldb, int kb, int64_t n,
                   int min_count, int32_t* out_count, uint8_t* out_flag, void* stream) {
  MV_REQUIRE(a && b && (out_count || out_flag), "mv_knn_overlap: null pointer");
  MV_REQUIRE(n >= 0 && ka >= 0 && kb >= 0 && lda >= ka && ldb >= kb, "mv_knn_overlap: bad sizes");
  MV_FAULT_POINT("launch");
  if (n == 0) return MV_OK;
  hipLaunchKernelGGL(k_knn_overlap, dim3(static_cast<unsigned>(ceil_div(n, 256))), dim3(256), 0,
                     static_cast<hipStream_t>(stream), a, lda, ka, b, ldb, kb, n, min_count, out_count, out_flag);
  MV_LAUNCHED("k_knn_overlap");
  return MV_OK;
}

}  // extern "C"
