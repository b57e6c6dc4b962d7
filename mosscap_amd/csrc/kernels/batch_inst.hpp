// Kernels and launchers of the batched 1-D ensemble, instantiated per form
// (batch_f64.hip, batch_f32c.hip).
#pragma once

#include "../device/batch.hpp"
#include "../launch_batch.hpp"
#include "stage_inst.hpp"

namespace mh {

// Occupancy the register allocator must allow for the persistent stage
// (waves per SIMD); the autotuner picks how many of them to launch.
#ifndef MH_BATCH_WAVES_PER_EU
#define MH_BATCH_WAVES_PER_EU 1
#endif

template <class Form, Reconstruction R, SlopeLimiter SL, RiemannSolver RS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MH_BATCH_WAVES_PER_EU))) void batch_stage_kernel(
    BatchGrid g, typename Form::Ptrs p, BatchArrays b) {
    batch_stage<Form, R, SL, RS>(g, p, b);
}

template <class Form>
__global__ __launch_bounds__(256) void batch_ghost_kernel(GhostGrid g, int nprob, typename Form::Planes in,
                                                          typename Form::Planes out, BatchArrays b) {
    batch_ghost<Form>(g, nprob, in, out, b);
}

// CFL max over every cell (ghosts included) of each problem: one block per problem.
template <class Form>
__global__ __launch_bounds__(256) void batch_cfl_kernel(typename Form::Planes q, double gamma, BatchArrays b) {
    using Raw = typename Form::raw;
    const int pr = blockIdx.x;
    const int xc = b.xc[pr], base = b.base[pr];
    double amax = 0.0;
    for (int i = threadIdx.x; i < xc; i += blockDim.x) {
        Raw v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = Form::get(q, base + k * xc + i);
        const double a = Form::template signal_speed_nd<1>(gamma, v);
        amax = a > amax ? a : amax;
    }
    const uint64_t m = wave_max_u64(__builtin_bit_cast(uint64_t, amax));
    if ((threadIdx.x & 63) == 0 && m != 0) atomicMax(&b.amax[pr], (unsigned long long)m);
}

template <class Form>
struct BatchPick {
    using Fn = void (*)(BatchGrid, typename Form::Ptrs, BatchArrays);
    template <Reconstruction R, SlopeLimiter SL>
    static Fn pick(RiemannSolver rs) {
        switch (rs) {
            case RiemannSolver::Rusanov: return batch_stage_kernel<Form, R, SL, RiemannSolver::Rusanov>;
            case RiemannSolver::Hll: return batch_stage_kernel<Form, R, SL, RiemannSolver::Hll>;
            default: return batch_stage_kernel<Form, R, SL, RiemannSolver::Hllc>;
        }
    }
};

template <class Form>
hipError_t launch_batch_stage_impl(const Schemes& s, const BatchGrid& g, int grid, const typename Form::Ptrs& p,
                                   const BatchArrays& b, hipStream_t st) {
    if (g.nitems < 1 || grid < 1 || g.ng < min_ghost_cells(s.recon)) return hipErrorInvalidValue;
    auto pick = [&]<Reconstruction R, SlopeLimiter SL>() { return BatchPick<Form>::template pick<R, SL>(s.solver); };
    auto fn = pick_scheme<Form>(s, pick);
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(kWave), 0, st, g, p, b);
    return hipGetLastError();
}

template <class Form>
hipError_t launch_batch_ghost_impl(const GhostGrid& g, int nprob, typename Form::Planes in,
                                   typename Form::Planes out, const BatchArrays& b, hipStream_t st) {
    if (nprob < 1 || g.ng < 1) return hipErrorInvalidValue;
    const int n = nprob * 2 * g.ng;
    hipLaunchKernelGGL((batch_ghost_kernel<Form>), dim3((n + 255) / 256), dim3(256), 0, st, g, nprob, in, out, b);
    return hipGetLastError();
}

template <class Form>
hipError_t launch_batch_cfl_impl(int nprob, typename Form::Planes q, double gamma, const BatchArrays& b,
                                 hipStream_t st) {
    if (nprob < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL((batch_cfl_kernel<Form>), dim3((unsigned)nprob), dim3(256), 0, st, q, gamma, b);
    return hipGetLastError();
}

}  // namespace mh
