// Batched ensemble of independent 1-D problems of uneven length.
//
// Layout: problem p owns a contiguous block [3][xc_p] (xc_p = n_p + 2 ng) at
// element offset base_p; the blocks are packed back to back. A host-built work
// list holds one item per (problem, 62-cell segment). One persistent launch
// per stage walks the list (wave w takes items w, w + grid, ...), so every
// wavefront does equal work whatever the length mix; each problem keeps its
// own dx, dt and time, and the CFL max is reduced per problem with atomics.
#pragma once

#include "ghost_dt.hpp"
#include "stage_core.hpp"

namespace mh {

struct BatchArrays {
    const int4* items;          // (problem, first output cell i0, block offset, cells incl. ghosts)
    const int* base;            // element offset of each problem's block
    const int* xc;              // cells incl. ghosts
    const double* dx;
    double* dt;                 // per-problem dt of the current step
    unsigned long long* amax;   // per-problem CFL max (bit pattern)
    double* time;               // per-problem clock
};

struct BatchGrid {
    int32_t ng, nitems, nprob, kind, reduce_dt;
    int32_t virt, per;  // ghosts read from their sources (periodic `per`, else zerograd); no fill per stage
    // The step's per-problem clock work folded into its stages (graph-replayed
    // persistent steps of three or more stages): 1 = this first stage takes
    // each problem's dt from its CFL max and clock itself and the wave with
    // the problem's first segment stores it; 2 = that wave of this second
    // stage advances the problem's time and clears its CFL max. 0 = none.
    int32_t fin = 0;
    double max_cfl = 0.0, max_time = 0.0;
};

// A problem's next dt from its CFL max and clock (compute_dt_impl's
// dx / max(a) times max_cfl, clamped to end exactly at max_time with the
// reference's nextafter loop). Reads only.
__device__ inline double batch_next_dt(const BatchArrays& b, int pr, double max_cfl, double max_time) {
    const double amax = __builtin_bit_cast(double, b.amax[pr]);
    double dt = (b.dx[pr] / amax) * max_cfl;
    const double t = b.time[pr];
    if (t + dt >= max_time) {
        dt = max_time - t;
        while (t + dt < max_time) dt = nextafter(dt, 1.7976931348623157e308);
    }
    return dt;
}

// Each wave walks its items with the next item's loads issued before the
// current item's arithmetic (the item carries the block offset and length,
// so those loads wait for nothing but the item itself).
template <class Form, Reconstruction R, SlopeLimiter SL, RiemannSolver RS>
__device__ inline void batch_stage(const BatchGrid& g, const typename Form::Ptrs& p, const BatchArrays& b) {
    using T = typename Form::real;
    constexpr int ORD = stencil_order(R);
    __shared__ T lds[3 * (kWave + 2 * ORD)];
    const int lane = threadIdx.x & 63;
    const int virt = Form::kVirtualGhosts ? g.virt : 0;
    const bool per = g.per != 0;
    unsigned int hits = 0;
    int it = blockIdx.x;
    int4 w = it < g.nitems ? b.items[it] : make_int4(0, 0, 0, 0);
    SegRaw<Form> cur;
    if (it < g.nitems) seg_load<Form, R>(p, g.kind, w.z, w.w, g.ng, w.w - 2 * g.ng, w.y, virt, per, cur);
    for (; it < g.nitems; it += gridDim.x) {
        const int nit = it + (int)gridDim.x;
        int4 wn = w;
        SegRaw<Form> nxt = cur;
        if (nit < g.nitems) {
            wn = b.items[nit];
            seg_load<Form, R>(p, g.kind, wn.z, wn.w, g.ng, wn.w - 2 * g.ng, wn.y, virt, per, nxt);
        }
        const int pr = w.x, i0 = w.y, base = w.z, xc = w.w;
        double dt_step;
        if (g.fin == 1) {
            dt_step = batch_next_dt(b, pr, g.max_cfl, g.max_time);
            if (i0 == g.ng && lane == 0) b.dt[pr] = dt_step;
        } else {
            dt_step = b.dt[pr];
            if (g.fin == 2 && i0 == g.ng && lane == 0) {
                b.time[pr] = b.time[pr] + dt_step;
                b.amax[pr] = 0ull;
            }
        }
        const T dt = (T)dt_step;
        const T dtdx = dt / (T)b.dx[pr];
        double amax = 0.0;
        seg_compute<Form, R, SL, RS>(p, g.kind, g.reduce_dt, base, xc, g.ng, xc - 2 * g.ng, i0, dtdx, dt, lds, amax,
                                     hits, cur);
        if (g.reduce_dt) {
            const uint64_t m = wave_max_u64(__builtin_bit_cast(uint64_t, amax));
            if (lane == 0) atomicMax(&b.amax[pr], (unsigned long long)m);
        }
        w = wn;
        cur = nxt;
    }
    if (p.floor_hits != nullptr) {
        const unsigned int total = wave_sum_u32(hits);
        if (lane == 0 && total != 0u) atomicAdd(p.floor_hits, total);
    }
}

// Ghost fill of every problem (Boundaries.hpp:11-119 with Axis = 0, NumDim =
// 1): one thread per (problem, ghost cell); sources are interior cells.
template <class Form>
__device__ inline void batch_ghost(const GhostGrid& g, int nprob, typename Form::Planes in, typename Form::Planes out,
                                   const BatchArrays& b) {
    using Raw = typename Form::raw;
    constexpr int NV = 3;
    const int ng = g.ng;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0;
    int pr = -1;
    if (tid < nprob * 2 * ng) {
        pr = tid / (2 * ng);
        const int pidx = tid % (2 * ng);
        const int xc = b.xc[pr], base = b.base[pr];
        const int i = pidx < ng ? pidx : (xc - 1) - (pidx - ng);
        const bool start = i < ng;
        const BoundaryType bt = (BoundaryType)(start ? g.bc[0] : g.bc[1]);
        Raw q[NV];
        if (bt == BoundaryType::UserFn) {
#pragma unroll
            for (int k = 0; k < NV; ++k) q[k] = Form::get(in, base + k * xc + i);
        } else {
            if (bt != BoundaryType::Constant) {
                const int is = ghost_source(i, xc, ng, bt);
#pragma unroll
                for (int k = 0; k < NV; ++k) q[k] = Form::get(out, base + k * xc + is);
            }
            ghost_modify<Form, 1, NV, 0>(q, bt, start, start ? g.bconst[0] : g.bconst[1]);
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) Form::put(out, base + k * xc + i, q[k]);
        if (g.reduce_dt) a = Form::template signal_speed_nd<1>(g.gamma, q);
    }
    if (g.reduce_dt && pr >= 0 && a > 0.0) atomicMax(&b.amax[pr], (unsigned long long)__builtin_bit_cast(uint64_t, a));
}

}  // namespace mh
