// wave.hpp -- wave64 / workgroup helpers shared by the tgtp kernels.
// All reductions use a fixed shuffle tree and a fixed cross-wave order, so
// results are the same on every run.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tgtp_dev {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }
__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x >> 6; }
__device__ __forceinline__ uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, static_cast<uint32_t>(__shfl_xor(v, off)));
  return v;
}

// Workgroup reductions for blocks of kWaves waves; `scratch` holds kWaves
// entries in LDS.  Every thread returns the result.
template <int kWaves, typename T, typename Op>
__device__ __forceinline__ T block_reduce(T v, T* scratch, Op op) {
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off));
  if (lane_id() == 0) scratch[wave_id()] = v;
  __syncthreads();
  T r = scratch[0];
#pragma unroll
  for (int w = 1; w < kWaves; ++w) r = op(r, scratch[w]);
  __syncthreads();
  return r;
}

struct OpSum {
  template <typename T>
  __device__ T operator()(T a, T b) const { return a + b; }
};
struct OpMaxF {
  __device__ float operator()(float a, float b) const { return fmaxf(a, b); }
};
struct OpMinF {
  __device__ float operator()(float a, float b) const { return fminf(a, b); }
};
struct OpMaxU {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

// Exclusive prefix sum over a workgroup of kWaves waves (one value per
// thread); returns the thread's offset and writes the total to *total.
template <int kWaves>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* scratch,
                                                         uint32_t* total) {
  const uint32_t lane = lane_id();
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= static_cast<uint32_t>(off)) x += y;
  }
  if (lane == 63) scratch[wave_id()] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const uint32_t s = scratch[w];
    if (static_cast<uint32_t>(w) < wave_id()) before += s;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

}  // namespace tgtp_dev
