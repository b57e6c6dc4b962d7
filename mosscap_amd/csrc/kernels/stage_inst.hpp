// Kernel entry points and scheme dispatch for the fused stages, instantiated
// once per state form (stage_f64.hip, stage_f32c.hip).
#pragma once

#include "../launch.hpp"
#include "../device/tracer.hpp"

// Occupancy target of the fused stage (waves per SIMD the register allocator
// must allow); each form's translation unit sets its own value.
#ifndef MH_STAGE_WAVES_PER_EU
#define MH_STAGE_WAVES_PER_EU 1
#endif
// ... of the instance with the gravity source and tracer stores
#ifndef MH_STAGE_WAVES_EXTRAS
#define MH_STAGE_WAVES_EXTRAS MH_STAGE_WAVES_PER_EU
#endif
// ... and of the plain instance with the positivity floors
#ifndef MH_STAGE_WAVES_FLOORS
#define MH_STAGE_WAVES_FLOORS MH_STAGE_WAVES_PER_EU
#endif

namespace mh {

template <class Form, Reconstruction R, SlopeLimiter SL, RiemannSolver RS, int Inst>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(Inst == kStageExtras   ? MH_STAGE_WAVES_EXTRAS
                                   : Inst == kStageFloors ? MH_STAGE_WAVES_FLOORS
                                                          : MH_STAGE_WAVES_PER_EU)))
void stage2d_kernel(StageGrid g, typename Form::Ptrs p) {
    stage2d_march<Form, R, SL, RS, 2, Inst>(g, p);
}

struct Stage1dGrid {
    int32_t xc, ng, nx, kind, reduce_dt;
};

template <class Form, Reconstruction R, SlopeLimiter SL, RiemannSolver RS>
__global__ __launch_bounds__(64) void stage1d_kernel(Stage1dGrid g, typename Form::Ptrs p) {
    using T = typename Form::real;
    constexpr int ORD = stencil_order(R);
    __shared__ T lds[3 * (kWave + 2 * ORD)];
    const int i0 = g.ng + (int)blockIdx.x * kStripOut;
    const T dt = (T)p.dt[0];
    const T dtdx = dt / (T)p.dx;
    double amax = 0.0;
    unsigned int hits = 0;
    stage1d_segment<Form, R, SL, RS>(p, g.kind, g.reduce_dt, 0, g.xc, g.ng, g.nx, i0, dtdx, dt, lds, amax, hits);
    if (g.reduce_dt) {
        const uint64_t m = wave_max_u64(__builtin_bit_cast(uint64_t, amax));
        if (threadIdx.x == 0) atomicMax(p.amax_bits, (unsigned long long)m);
    }
    if (p.floor_hits != nullptr) {
        const unsigned int total = wave_sum_u32(hits);
        if (threadIdx.x == 0 && total != 0u) atomicAdd(p.floor_hits, total);
    }
}

template <class Form>
using Stage2dFn = void (*)(StageGrid, typename Form::Ptrs);
template <class Form>
using Stage1dFn = void (*)(Stage1dGrid, typename Form::Ptrs);

template <class Form, Reconstruction R, SlopeLimiter SL>
struct SolverPick {
    template <int Inst>
    static Stage2dFn<Form> two(RiemannSolver rs) {
        switch (rs) {
            case RiemannSolver::Rusanov: return stage2d_kernel<Form, R, SL, RiemannSolver::Rusanov, Inst>;
            case RiemannSolver::Hll: return stage2d_kernel<Form, R, SL, RiemannSolver::Hll, Inst>;
            default: return stage2d_kernel<Form, R, SL, RiemannSolver::Hllc, Inst>;
        }
    }
    static Stage1dFn<Form> one(RiemannSolver rs) {
        switch (rs) {
            case RiemannSolver::Rusanov: return stage1d_kernel<Form, R, SL, RiemannSolver::Rusanov>;
            case RiemannSolver::Hll: return stage1d_kernel<Form, R, SL, RiemannSolver::Hll>;
            default: return stage1d_kernel<Form, R, SL, RiemannSolver::Hllc>;
        }
    }
};

// (reconstruction, limiter) -> instance; FOG and WENO5-Z ignore the limiter
// (is_limiter_agnostic, Reconstruct.hpp:34-45).
template <class Form, class Pick>
auto pick_scheme(const Schemes& s, Pick pick) {
    using SL = SlopeLimiter;
    using RC = Reconstruction;
    switch (s.recon) {
        case RC::Fog: return pick.template operator()<RC::Fog, SL::MonotonizedCentral>();
        case RC::Weno5Z: return pick.template operator()<RC::Weno5Z, SL::MonotonizedCentral>();
        case RC::Muscl:
            if (s.limiter == SL::VanLeer) return pick.template operator()<RC::Muscl, SL::VanLeer>();
            if (s.limiter == SL::Minmod) return pick.template operator()<RC::Muscl, SL::Minmod>();
            return pick.template operator()<RC::Muscl, SL::MonotonizedCentral>();
        default:
            if (s.limiter == SL::VanLeer) return pick.template operator()<RC::Ppm, SL::VanLeer>();
            if (s.limiter == SL::Minmod) return pick.template operator()<RC::Ppm, SL::Minmod>();
            return pick.template operator()<RC::Ppm, SL::MonotonizedCentral>();
    }
}

template <class Form>
hipError_t launch_stage2d_impl(const Schemes& s, const StageGrid& g, const typename Form::Ptrs& p, hipStream_t st) {
    if (g.nx < 1 || g.ny < 1 || g.ng < min_ghost_cells(s.recon) || g.xc != g.nx + 2 * g.ng ||
        g.yc != g.ny + 2 * g.ng || g.nspans < 1 || g.nspans > g.ny || g.strips != (g.nx + kStripOut - 1) / kStripOut) {
        return hipErrorInvalidValue;
    }
    // the plain instance (buffer-resource addressing, no source, tracer
    // stores or floors), with the floors compiled in when they are on (forms
    // that have them), and the general instance for a gravity source or
    // tracer flux stores or when the state's byte offsets do not fit the
    // buffer offsets' 31 bits
    const bool big = (int64_t)4 * g.xc * g.yc * Form::kElem >= ((int64_t)1 << 31);
    const bool extras = p.gravity != 0 || p.frx != nullptr || p.fry != nullptr || big;
    const bool floors = Form::floors_on(p);
    auto pick = [&]<Reconstruction R, SlopeLimiter SL>() -> Stage2dFn<Form> {
        if (extras) return SolverPick<Form, R, SL>::template two<kStageExtras>(s.solver);
        if constexpr (Form::kHasFloors) {
            if (floors) return SolverPick<Form, R, SL>::template two<kStageFloors>(s.solver);
        }
        return SolverPick<Form, R, SL>::template two<kStagePlain>(s.solver);
    };
    Stage2dFn<Form> fn = pick_scheme<Form>(s, pick);
    hipLaunchKernelGGL(fn, dim3((unsigned)(g.strips * g.nspans)), dim3(kWave), 0, st, g, p);
    return hipGetLastError();
}

template <class Form>
hipError_t launch_stage1d_impl(const Schemes& s, int kind, int reduce_dt, int xc, int ng,
                               const typename Form::Ptrs& p, hipStream_t st) {
    const int nx = xc - 2 * ng;
    if (nx < 1 || ng < min_ghost_cells(s.recon)) return hipErrorInvalidValue;
    auto pick = [&]<Reconstruction R, SlopeLimiter SL>() { return SolverPick<Form, R, SL>::one(s.solver); };
    Stage1dFn<Form> fn = pick_scheme<Form>(s, pick);
    const Stage1dGrid g{xc, ng, nx, kind, reduce_dt};
    const int segs = (nx + kStripOut - 1) / kStripOut;
    hipLaunchKernelGGL(fn, dim3((unsigned)segs), dim3(kWave), 0, st, g, p);
    return hipGetLastError();
}

#ifdef MH_STAGE3D_INSTANCES
template <class Form, Reconstruction R, SlopeLimiter SL, RiemannSolver RS, bool Extras>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(Extras ? MH_STAGE_WAVES_EXTRAS : MH_STAGE_WAVES_PER_EU)))
void stage3d_kernel(StageGrid g, typename Form::Ptrs p) {
    stage2d_march<Form, R, SL, RS, 3, Extras ? kStageExtras : kStagePlain>(g, p);
}

template <class Form, Reconstruction R, SlopeLimiter SL, RiemannSolver RS>
__global__ __launch_bounds__(128) void zsweep_kernel(ZSweepGrid g, typename Form::Ptrs p, typename Form::real* dz) {
    zsweep_column<Form, R, SL, RS>(g, p, dz);
}

template <class Form>
hipError_t launch_stage3d_impl(const Schemes& s, const StageGrid& g, const typename Form::Ptrs& p,
                               typename Form::real* dz, hipStream_t st) {
    if (g.nx < 1 || g.ny < 1 || g.nz < 1 || g.ng < min_ghost_cells(s.recon) || g.xc != g.nx + 2 * g.ng ||
        g.yc != g.ny + 2 * g.ng || g.zc != g.nz + 2 * g.ng || g.nspans < 1 || g.nspans > g.ny ||
        g.strips != (g.nx + kStripOut - 1) / kStripOut || !dz) {
        return hipErrorInvalidValue;
    }
    using ZFn = void (*)(ZSweepGrid, typename Form::Ptrs, typename Form::real*);
    auto zpick = [&]<Reconstruction R, SlopeLimiter SL>() -> ZFn {
        switch (s.solver) {
            case RiemannSolver::Rusanov: return zsweep_kernel<Form, R, SL, RiemannSolver::Rusanov>;
            case RiemannSolver::Hll: return zsweep_kernel<Form, R, SL, RiemannSolver::Hll>;
            default: return zsweep_kernel<Form, R, SL, RiemannSolver::Hllc>;
        }
    };
    // plain instance (buffer addressing, no source or floors) unless gravity
    // or the floors are on or a byte offset could pass 2^31
    const bool big = (int64_t)5 * g.xc * g.yc * g.zc * Form::kElem >= ((int64_t)1 << 31);
    const bool extras = p.gravity != 0 || Form::floors_on(p) || big;
    auto mpick = [&]<Reconstruction R, SlopeLimiter SL>() -> Stage2dFn<Form> {
        switch (s.solver) {
            case RiemannSolver::Rusanov:
                return extras ? stage3d_kernel<Form, R, SL, RiemannSolver::Rusanov, true>
                              : stage3d_kernel<Form, R, SL, RiemannSolver::Rusanov, false>;
            case RiemannSolver::Hll:
                return extras ? stage3d_kernel<Form, R, SL, RiemannSolver::Hll, true>
                              : stage3d_kernel<Form, R, SL, RiemannSolver::Hll, false>;
            default:
                return extras ? stage3d_kernel<Form, R, SL, RiemannSolver::Hllc, true>
                              : stage3d_kernel<Form, R, SL, RiemannSolver::Hllc, false>;
        }
    };
    // z sweep: chunks of at most 32 planes per thread, columns of the interior across the grid's x
    const int chunks = (g.nz + 31) / 32;
    ZSweepGrid zg{g.xc, g.yc, g.zc, g.ng, g.nx, g.ny, g.nz, (g.nz + chunks - 1) / chunks};
    zg.virt = g.virt;
    zg.zper = g.zper;
    const int cols = g.nx * g.ny;
    hipLaunchKernelGGL(pick_scheme<Form>(s, zpick), dim3((unsigned)((cols + 127) / 128), (unsigned)chunks), dim3(128),
                       0, st, zg, p, dz);
    typename Form::Ptrs pm = p;
    pm.dz3 = dz;
    hipLaunchKernelGGL(pick_scheme<Form>(s, mpick), dim3((unsigned)(g.strips * g.nspans * g.nz)), dim3(kWave), 0, st,
                       g, pm);
    return hipGetLastError();
}
#endif

// Tracer planes of one 2-D stage (after the hydro stage stored the density
// fluxes): W_t of the stage input, then the interior update.
template <class Form>
hipError_t launch_tracer_impl(const Schemes& s, const TracerGrid& g, const typename Form::Ptrs& p,
                              typename Form::real* wtr, hipStream_t st) {
    if (g.ntr < 1 || g.nx < 1 || g.ny < 1 || g.ng < min_ghost_cells(s.recon) || g.xc != g.nx + 2 * g.ng ||
        g.yc != g.ny + 2 * g.ng || !wtr || !p.frx || !p.fry)
        return hipErrorInvalidValue;
    const int cells = g.xc * g.yc, inner = g.nx * g.ny;
    hipLaunchKernelGGL(tracer_prim_kernel<Form>, dim3((cells + 255) / 256), dim3(256), 0, st, g, p.qs, wtr);
    using Fn = void (*)(TracerGrid, typename Form::Ptrs, const typename Form::real*, const typename Form::real*,
                        const typename Form::real*);
    auto pick = [&]<Reconstruction R, SlopeLimiter SL>() -> Fn { return tracer_stage_kernel<Form, R, SL>; };
    Fn fn = pick_scheme<Form>(s, pick);
    hipLaunchKernelGGL(fn, dim3((inner + 255) / 256), dim3(256), 0, st, g, p, wtr, p.frx, p.fry);
    return hipGetLastError();
}

}  // namespace mh
