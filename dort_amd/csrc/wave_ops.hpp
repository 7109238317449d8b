// wave_ops.hpp -- wave and workgroup primitives shared by the wavefront kernels
// (kernels.hip, bdpt.hip).
#pragma once
#include "dgpu_internal.hpp"

namespace dgpu {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Appends `pred` lanes of a whole workgroup of SHADE_BLOCK threads with one atomic: the
// waves' counts go through LDS, thread 0 takes their exclusive prefix and adds the
// total.  Every thread of the workgroup must make the same sequence of calls.  Queue
// counters are single addresses that every wave of the chip appends to, and atomics on
// one address serialise in L2: one atomic per wave capped the shade kernel at ~90M
// appends/s per counter (measured on MI355X); per workgroup the cap moves out by the
// number of waves in the workgroup.
__device__ __forceinline__ uint32_t block_append(uint32_t* counter, bool pred, uint32_t* sh) {
  constexpr uint32_t WAVES = SHADE_BLOCK / 64;
  unsigned long long mask = __ballot(pred);
  uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = (uint32_t)__popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < WAVES; ++w) { uint32_t c = sh[w]; sh[w] = tot; tot += c; }
    sh[WAVES] = tot ? atomicAdd(counter, tot) : 0u;
  }
  __syncthreads();
  unsigned long long below = lane == 0 ? 0ull : (mask & ((1ull << lane) - 1ull));
  return sh[WAVES] + sh[wave] + (uint32_t)__popcll(below);
}

// Appends `cnt` entries per thread for a whole workgroup of SHADE_BLOCK threads with one
// atomic: returns the thread's first slot (its entries are consecutive).  A wave scans its
// counts with shuffles; thread 0 takes the exclusive prefix of the wave totals.  Every
// thread of the workgroup must call it.
__device__ __forceinline__ uint32_t block_append_n(uint32_t* counter, uint32_t cnt, uint32_t* sh) {
  constexpr uint32_t WAVES = SHADE_BLOCK / 64;
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  uint32_t x = cnt;
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63u) sh[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < WAVES; ++w) { uint32_t c = sh[w]; sh[w] = tot; tot += c; }
    sh[WAVES] = tot ? atomicAdd(counter, tot) : 0u;
  }
  __syncthreads();
  return sh[WAVES] + sh[wave] + x - cnt;
}

}  // namespace dgpu
