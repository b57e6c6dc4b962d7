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

// Value of x in lane `src` (clamped to the wave).
__device__ inline double lane_value(double x, int src) {
    const int a = min(max(src, 0), kWave - 1) << 2;
    const int2 v = __builtin_bit_cast(int2, x);
    int2 r;
    r.x = __builtin_amdgcn_ds_bpermute(a, v.x);
    r.y = __builtin_amdgcn_ds_bpermute(a, v.y);
    return __builtin_bit_cast(double, r);
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

// Raw buffer resources (gfx9 V#: stride 0, num_records 0xFFFFFFFF, 32-bit
// data format word). Callers keep scalar offset + vector offset below 2^31
// bytes; the resource's range check is not relied on.
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, -1, 0x00020000);
}
// Direct-to-LDS load: lane t's dword at (voff + soff) lands in lds[t]
// without passing through registers. The compiler does not track it: the
// reader issues wait_vm_loads() first, and the LDS words it overwrites must
// have been read (wait_lds_ops()) before it is issued.
__device__ inline void buf_load_lds(__amdgpu_buffer_rsrc_t r, uint32_t* lds, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 4, (int)voff,
                                             (int)soff, 0, 0);
}
// s_waitcnt vmcnt(0): every vector-memory operation of this wave has completed.
__device__ inline void wait_vm_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }
// s_waitcnt vmcnt(N): at most the N most recently issued vector-memory
// operations of this wave are still outstanding (they complete in order).
template <int N>
__device__ inline void wait_vm_at_most() {
    static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}
// s_waitcnt lgkmcnt(0): every LDS operation of this wave has completed.
__device__ inline void wait_lds_ops() { __builtin_amdgcn_s_waitcnt(0xC07F); }

__device__ inline float buf_load_f32(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ inline void buf_store_f32(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 0);
}
__device__ inline double buf_load_f64(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
__device__ inline void buf_store_f64(double v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    using V2 = __attribute__((ext_vector_type(2))) uint32_t;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, (int)voff, (int)soff, 0);
}

}  // namespace mh
