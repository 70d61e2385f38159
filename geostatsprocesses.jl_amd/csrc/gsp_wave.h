// gsp_wave.h — wavefront helpers for gfx950 (64 lanes).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace gsp {

// value of `x` in lane `lane` (uniform lane index)
__device__ inline double rl(double x, int lane) {
  uint64_t u = __builtin_bit_cast(uint64_t, x);
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ inline float rlf(float x, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), lane));
}

__device__ inline int rli(int x, int lane) { return __builtin_amdgcn_readlane(x, lane); }

// number of set bits of `mask` below this lane
__device__ inline int mbcnt(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// orders LDS traffic of one wavefront (no cross-wave barrier)
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP row_newbcast: the value of lane M of this lane's 16-lane row
template <int M>
__device__ inline double rbd(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x150 + M, 0xF,
                                                            0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32),
                                                            0x150 + M, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

template <int M>
__device__ inline float rbf(float x) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x150 + M, 0xF, 0xF, false));
}

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// visits N-1, N-2, ..., 0
template <int N, class F>
__device__ __forceinline__ void static_for_down(F&& f) {
  if constexpr (N > 0) {
    f(std::integral_constant<int, N - 1>{});
    static_for_down<N - 1>(f);
  }
}

// (j, m) with j >= m of the e-th entry of a packed lower triangle
__device__ inline void tri_unpack(int e, int& j, int& m) {
  int jj = (int)((__builtin_sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
  while (jj * (jj + 1) / 2 > e) --jj;
  while ((jj + 1) * (jj + 2) / 2 <= e) ++jj;
  j = jj;
  m = e - jj * (jj + 1) / 2;
}

__device__ inline bool key_less(double da, int ia, double db, int ib) {
  return da < db || (da == db && ia < ib);
}

}  // namespace gsp
