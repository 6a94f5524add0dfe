// Shared helpers for the libmivis HIP sources: status/error plumbing, launch
// checks, fault injection, and wavefront primitives for gfx950 (64 lanes).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/mivis.h"

namespace mv {

// Records a printf-style message as the calling thread's last error.
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
// Returns nonzero when MIVIS_FAULT names `site` ("launch", "alloc", "all").
int fault_injected(const char* site);

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Bump allocator over a caller-provided workspace.
struct Carve {
  char* base;
  size_t off = 0;
  explicit Carve(void* p) : base(static_cast<char*>(p)) {}
  template <class T>
  T* take(size_t count) {
    off = align_up(off, 256);
    T* p = reinterpret_cast<T*>(base ? base + off : nullptr);
    off += count * sizeof(T);
    return p;
  }
};

}  // namespace mv

#define MV_TRY_HIP(expr)                                                              \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) {                                                           \
      mv::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                    __LINE__);                                                        \
      return MV_EHIP;                                                                 \
    }                                                                                 \
  } while (0)

#define MV_LAUNCHED(name)                                                               \
  do {                                                                                  \
    hipError_t _e = hipGetLastError();                                                  \
    if (_e != hipSuccess) {                                                             \
      mv::set_error("launch of %s failed: %s", name, hipGetErrorString(_e));           \
      return MV_EHIP;                                                                   \
    }                                                                                   \
  } while (0)

#define MV_REQUIRE(cond, ...)   \
  do {                          \
    if (!(cond)) {              \
      mv::set_error(__VA_ARGS__); \
      return MV_EINVAL;         \
    }                           \
  } while (0)

#define MV_FAULT_POINT(site)                                              \
  do {                                                                    \
    if (mv::fault_injected(site)) {                                       \
      mv::set_error("fault injected at '%s' (MIVIS_FAULT)", site);        \
      return MV_EINJECTED;                                                \
    }                                                                     \
  } while (0)

// ----------------------------------------------------------------------------
// Device-side wavefront helpers (wave64).
namespace mv {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
  int lo = __shfl(static_cast<int>(static_cast<uint32_t>(v)), src);
  int hi = __shfl(static_cast<int>(static_cast<uint32_t>(v >> 32)), src);
  return (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo);
}
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  int lo = __shfl_xor(static_cast<int>(static_cast<uint32_t>(v)), m);
  int hi = __shfl_xor(static_cast<int>(static_cast<uint32_t>(v >> 32)), m);
  return (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo);
}
__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t v, int delta) {
  int lo = __shfl_up(static_cast<int>(static_cast<uint32_t>(v)), delta);
  int hi = __shfl_up(static_cast<int>(static_cast<uint32_t>(v >> 32)), delta);
  return (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo);
}
// Lane i receives lane i - 1's value (lane 0 keeps its own): the DPP
// wave_shr:1 move, a single VALU operation instead of a permute through LDS.
__device__ __forceinline__ uint64_t wave_shr1_u64(uint64_t v) {
  const int lo = static_cast<int>(static_cast<uint32_t>(v));
  const int hi = static_cast<int>(static_cast<uint32_t>(v >> 32));
  const int slo = __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xF, 0xF, false);
  const int shi = __builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xF, 0xF, false);
  return (static_cast<uint64_t>(static_cast<uint32_t>(shi)) << 32) | static_cast<uint32_t>(slo);
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int src) {
  uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), src);
  uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), src);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// ---- bitonic networks over one 64-bit key per lane ----------------------
// Partners within a 16-lane row come from DPP moves (quad_perm for distances
// 1 and 2, row_half_mirror then a quad reversal for 4, row_ror:8 for 8);
// across rows, v_permlane16_swap / v_permlane32_swap deliver both elements of
// every pair (the lower lane's in result [0], the upper lane's in [1]) to both
// lanes. No stage goes through LDS.
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  const int lo = static_cast<int>(static_cast<uint32_t>(v));
  const int hi = static_cast<int>(static_cast<uint32_t>(v >> 32));
  const int rlo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  const int rhi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return (static_cast<uint64_t>(static_cast<uint32_t>(rhi)) << 32) | static_cast<uint32_t>(rlo);
}

// Compare-exchange at distance J (a power of two <= 32): the lane keeps the
// smaller of its key and its partner's (lane ^ J) when take_min, else the larger.
template <int J>
__device__ __forceinline__ uint64_t cmpx_u64(uint64_t x, bool take_min) {
  if constexpr (J >= 16) {
    const uint32_t lo = static_cast<uint32_t>(x), hi = static_cast<uint32_t>(x >> 32);
    uint32_t alo, blo, ahi, bhi;
    if constexpr (J == 16) {
      const auto r0 = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
      const auto r1 = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
      alo = r0[0]; blo = r0[1]; ahi = r1[0]; bhi = r1[1];
    } else {
      const auto r0 = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
      const auto r1 = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
      alo = r0[0]; blo = r0[1]; ahi = r1[0]; bhi = r1[1];
    }
    const uint64_t a = (static_cast<uint64_t>(ahi) << 32) | alo;   // the pair's lower-lane key
    const uint64_t b = (static_cast<uint64_t>(bhi) << 32) | blo;   // the pair's upper-lane key
    return ((a < b) == take_min) ? a : b;
  } else {
    uint64_t p;
    if constexpr (J == 1) p = dpp_u64<0xB1>(x);                    // quad_perm [1,0,3,2]
    else if constexpr (J == 2) p = dpp_u64<0x4E>(x);               // quad_perm [2,3,0,1]
    else if constexpr (J == 4) p = dpp_u64<0x1B>(dpp_u64<0x141>(x));   // half mirror, then [3,2,1,0]
    else p = dpp_u64<0x128>(x);                                    // row_ror:8
    return ((p < x) == take_min) ? p : x;
  }
}

// Stage (K, J) of the bitonic sorting network, ascending (descending if DESC).
template <int K, int J, bool DESC>
__device__ __forceinline__ uint64_t bitonic_stage_u64(uint64_t x, int l) {
  const bool up = ((l & K) == 0) != DESC;
  return cmpx_u64<J>(x, ((l & J) == 0) == up);
}

// Bitonic sort of one 64-bit key per lane across the wave (ascending, or
// descending if DESC).
template <bool DESC = false>
__device__ __forceinline__ uint64_t wave_sort_u64(uint64_t x) {
  const int l = lane_id();
  x = bitonic_stage_u64<2, 1, DESC>(x, l);
  x = bitonic_stage_u64<4, 2, DESC>(x, l);
  x = bitonic_stage_u64<4, 1, DESC>(x, l);
  x = bitonic_stage_u64<8, 4, DESC>(x, l);
  x = bitonic_stage_u64<8, 2, DESC>(x, l);
  x = bitonic_stage_u64<8, 1, DESC>(x, l);
  x = bitonic_stage_u64<16, 8, DESC>(x, l);
  x = bitonic_stage_u64<16, 4, DESC>(x, l);
  x = bitonic_stage_u64<16, 2, DESC>(x, l);
  x = bitonic_stage_u64<16, 1, DESC>(x, l);
  x = bitonic_stage_u64<32, 16, DESC>(x, l);
  x = bitonic_stage_u64<32, 8, DESC>(x, l);
  x = bitonic_stage_u64<32, 4, DESC>(x, l);
  x = bitonic_stage_u64<32, 2, DESC>(x, l);
  x = bitonic_stage_u64<32, 1, DESC>(x, l);
  x = bitonic_stage_u64<64, 32, DESC>(x, l);
  x = bitonic_stage_u64<64, 16, DESC>(x, l);
  x = bitonic_stage_u64<64, 8, DESC>(x, l);
  x = bitonic_stage_u64<64, 4, DESC>(x, l);
  x = bitonic_stage_u64<64, 2, DESC>(x, l);
  x = bitonic_stage_u64<64, 1, DESC>(x, l);
  return x;
}

// Ascending sort of a bitonic sequence of one key per lane (half-cleaners).
__device__ __forceinline__ uint64_t bitonic_clean_u64(uint64_t x) {
  const int l = lane_id();
  x = bitonic_stage_u64<64, 32, false>(x, l);
  x = bitonic_stage_u64<64, 16, false>(x, l);
  x = bitonic_stage_u64<64, 8, false>(x, l);
  x = bitonic_stage_u64<64, 4, false>(x, l);
  x = bitonic_stage_u64<64, 2, false>(x, l);
  x = bitonic_stage_u64<64, 1, false>(x, l);
  return x;
}

// The 64 smallest of the ascending list `lk` and the unsorted keys `a` (one
// per lane), ascending: `a` is sorted descending, so that the lane-wise
// minimum of the two is a bitonic sequence holding the 64 smallest.
__device__ __forceinline__ uint64_t list_merge64_u64(uint64_t lk, uint64_t a) {
  a = wave_sort_u64<true>(a);
  return bitonic_clean_u64(lk < a ? lk : a);
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    uint64_t o = shfl_xor_u64(v, m);
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ float wave_sum_f32(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

// Key of a (non-negative distance, index) pair: orders by distance, then index.
__device__ __forceinline__ uint64_t dkey(float dist, uint32_t idx) {
  return (static_cast<uint64_t>(__float_as_uint(dist)) << 32) | idx;
}
__device__ __forceinline__ float key_dist(uint64_t k) {
  return __uint_as_float(static_cast<uint32_t>(k >> 32));
}
__device__ __forceinline__ uint32_t key_idx(uint64_t k) { return static_cast<uint32_t>(k); }

// Philox-4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11):
// ten rounds of the 4x32 multiply-xor round with the published multipliers
// and Weyl key increments.
struct Philox {
  __device__ static __forceinline__ void round(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                                               uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    // one 32 x 32 -> 64 multiply per product (v_mad_u64_u32) gives both halves
    const uint64_t p0 = static_cast<uint64_t>(M0) * c0, p1 = static_cast<uint64_t>(M1) * c2;
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  __device__ static __forceinline__ uint4 gen(uint4 ctr, uint32_t k0, uint32_t k1) {
    uint32_t c0 = ctr.x, c1 = ctr.y, c2 = ctr.z, c3 = ctr.w;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      round(c0, c1, c2, c3, k0, k1);
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    return make_uint4(c0, c1, c2, c3);
  }
};

__device__ __forceinline__ uint32_t uniform_below(uint32_t r, uint32_t n) {
  return static_cast<uint32_t>((static_cast<uint64_t>(r) * n) >> 32);
}

}  // namespace mv
