// Wavefront-level helpers for gfx950 (wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mh {

constexpr int kWave = 64;

// Orders LDS traffic between the lanes of one wavefront (the kernels that use
// it run one wavefront per workgroup). LDS instructions of a wavefront execute
// in issue order, so only compiler reordering has to be prevented.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value from lane (lane - 1); lane 0 receives its own value.
__device__ inline double lane_from_left(double x) {
    const int src = (int)((threadIdx.x & 63u) == 0 ? 0 : (threadIdx.x & 63u) - 1) << 2;
    const int2 v = __builtin_bit_cast(int2, x);
    int2 r;
    r.x = __builtin_amdgcn_ds_bpermute(src, v.x);
    r.y = __builtin_amdgcn_ds_bpermute(src, v.y);
    return __builtin_bit_cast(double, r);
}
__device__ inline float lane_from_left(float x) {
    const int src = (int)((threadIdx.x & 63u) == 0 ? 0 : (threadIdx.x & 63u) - 1) << 2;
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src, __builtin_bit_cast(int, x)));
}
// Value from lane (lane + 1); lane 63 receives its own value.
__device__ inline double lane_from_right(double x) {
    const int src = (int)((threadIdx.x & 63u) == 63u ? 63u : (threadIdx.x & 63u) + 1) << 2;
    const int2 v = __builtin_bit_cast(int2, x);
    int2 r;
    r.x = __builtin_amdgcn_ds_bpermute(src, v.x);
    r.y = __builtin_amdgcn_ds_bpermute(src, v.y);
    return __builtin_bit_cast(double, r);
}
__device__ inline float lane_from_right(float x) {
    const int src = (int)((threadIdx.x & 63u) == 63u ? 63u : (threadIdx.x & 63u) + 1) << 2;
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src, __builtin_bit_cast(int, x)));
}

// Wavefront max of a non-negative 64-bit pattern (non-negative doubles order
// like their bit patterns).
__device__ inline uint64_t wave_max_u64(uint64_t x) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int src = (int)(((threadIdx.x & 63u) ^ (unsigned)off) << 2);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)x);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)(x >> 32));
        const uint64_t y = ((uint64_t)hi << 32) | lo;
        x = y > x ? y : x;
    }
    return x;
}

__device__ inline unsigned int wave_sum_u32(unsigned int x) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int src = (int)(((threadIdx.x & 63u) ^ (unsigned)off) << 2);
        x += (unsigned int)__builtin_amdgcn_ds_bpermute(src, (int)x);
    }
    return x;
}

// Workgroup ids are dealt round-robin to the 8 XCDs of an MI355X. Remap so
// that each XCD receives a contiguous run of logical ids: neighbouring tiles
// (which share halo cells) then share an L2.
__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t n) {
    constexpr uint32_t kXcd = 8;
    const uint32_t per = n / kXcd, rem = n % kXcd;
    const uint32_t xcd = b % kXcd, pos = b / kXcd;
    return xcd < rem ? xcd * (per + 1) + pos : rem * (per + 1) + (xcd - rem) * per + pos;
}

}  // namespace mh
