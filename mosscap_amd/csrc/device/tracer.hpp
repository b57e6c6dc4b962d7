// Passive tracer fields (simulation.n_extra_fields) of a 2-D stage.
//
// The reference carries each tracer as a conserved density Q_t; its primitive
// is the mass fraction W_t = Q_t * (1/rho) (Hydro.cpp:27-33), reconstructed
// like the hydro variables (Hydro.cpp:55-102), and its face flux is the
// density flux times the upwind face value: F_t = F_rho * RR(left) when
// F_rho >= 0, else F_rho * RL(right) (Hydro.cpp:155-162). The divergence and
// the stage combination are those of the hydro variables with a zero source
// (TimeStepping.cpp stage lambdas, `var >= n_hydro`).
//
// Here the fused hydro stage stores F_rho of every x and y face it computes;
// the tracer pass then reads the stage input's tracer and density planes,
// forms W_t once per cell, and updates each interior cell from its four
// faces. Tracers never feed back into the hydro variables or the time step.
#pragma once

#include "functors.hpp"
#include "ghost_dt.hpp"

namespace mh {

struct TracerGrid {
    int32_t xc, yc, ng, nx, ny;
    int32_t ntr;   // tracer count; planes 4 .. 3+ntr of the state
    int32_t kind;  // StageKind
};

// W_t = Q_t * (1/rho) on every cell (ghosts included), FP64 reciprocal as in
// global_cons_to_prim_impl.
template <class Form>
__global__ __launch_bounds__(256) void tracer_prim_kernel(TracerGrid g, typename Form::Planes qs,
                                                          typename Form::real* wtr) {
    using T = typename Form::real;
    const int plane = g.xc * g.yc;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= plane) return;
    const double inv_rho = 1.0 / Form::value(Form::get(qs, cell));
    for (int k = 0; k < g.ntr; ++k) {
        const double q = Form::value(Form::get(qs, (4 + k) * plane + cell));
        wtr[k * plane + cell] = (T)(q * inv_rho);
    }
}

// One thread per interior cell: reconstruction of the cell and its two
// neighbours along each axis, upwinded face fluxes, divergence, combination.
template <class Form, Reconstruction R, SlopeLimiter SL>
__global__ __launch_bounds__(256) void tracer_stage_kernel(TracerGrid g, typename Form::Ptrs p,
                                                           const typename Form::real* __restrict__ wtr,
                                                           const typename Form::real* __restrict__ frx,
                                                           const typename Form::real* __restrict__ fry) {
    using T = typename Form::real;
    using Raw = typename Form::raw;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nx * g.ny) return;
    const int i = g.ng + t % g.nx, j = g.ng + t / g.nx;
    const int xc = g.xc, plane = g.xc * g.yc;
    const int cell = j * xc + i;
    const T dt = (T)p.dt[0];
    const T dtdx = dt / (T)p.dx;
    const T src = T(0) * dt;
    const T fx0 = frx[cell], fx1 = frx[cell + 1];
    const T fy0 = fry[cell], fy1 = fry[cell + xc];
    for (int k = 0; k < g.ntr; ++k) {
        const T* w = wtr + (size_t)k * plane + cell;
        T l[3], r[3];
        // x: cells i-1, i, i+1
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            auto s = [&](int o) -> T { return w[(c - 1) + o]; };
            reconstruct_stencil<R, SL, T>(s, l[c], r[c]);
        }
        const T Fx0 = fx0 >= T(0) ? fx0 * r[0] : fx0 * l[1];
        const T Fx1 = fx1 >= T(0) ? fx1 * r[1] : fx1 * l[2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            auto s = [&](int o) -> T { return w[((c - 1) + o) * xc]; };
            reconstruct_stencil<R, SL, T>(s, l[c], r[c]);
        }
        const T Gy0 = fy0 >= T(0) ? fy0 * r[0] : fy0 * l[1];
        const T Gy1 = fy1 >= T(0) ? fy1 * r[1] : fy1 * l[2];
        T q_update = T(0);
        q_update += Fx0 - Fx1;
        q_update += Gy0 - Gy1;
        q_update *= dtdx;
        const int idx = (4 + k) * plane + cell;
        const Raw qs = Form::get(p.qs, idx);
        const Raw qn = Form::needs_qn(g.kind) ? Form::get(p.qn, idx) : qs;
        Form::put(p.out, idx, Form::combine(g.kind, qs, qn, q_update, src));
    }
}

// Ghost fill of the tracer planes: the hydro fill's index maps (x pass, then
// y pass reading the x-filled rows) with no modification, since wall,
// symmetric and diode faces only change the normal momentum and the energy.
template <class Form>
__global__ __launch_bounds__(256) void tracer_ghost_kernel(GhostGrid g, int ntr, typename Form::Planes in,
                                                           typename Form::Planes out) {
    const int xc = g.xc, yc = g.yc, ng = g.ng;
    const int plane = xc * yc;
    const int nxband = 2 * ng * yc, nyband = 2 * ng * xc;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    // (source buffer is `in`, source cell); x pass value at (j, i), i an x-ghost column
    auto xsource = [&](int j, int i, bool& from_in) -> int {
        const bool start = i < ng;
        const BoundaryType b = (BoundaryType)(start ? g.bc[0] : g.bc[1]);
        if (b == BoundaryType::UserFn) {
            from_in = true;
            return j * xc + i;
        }
        const bool yghost = j < ng || j >= yc - ng;
        from_in = yghost;
        return j * xc + ghost_source(i, xc, ng, b);
    };
    int dst = -1, srcc = 0;
    bool from_in = false;
    if (tid < nxband) {
        const int pp = tid % (2 * ng);
        const int j = tid / (2 * ng);
        const int i = pp < ng ? pp : (xc - 1) - (pp - ng);
        bool skip = false;
        if (j < ng || j >= yc - ng) skip = (BoundaryType)(j < ng ? g.bc[2] : g.bc[3]) != BoundaryType::UserFn;
        if (!skip) {
            dst = j * xc + i;
            srcc = xsource(j, i, from_in);
        }
    } else if (tid < nxband + nyband) {
        const int tt = tid - nxband;
        const int pp = tt / xc;
        const int i = tt % xc;
        const int j = pp < ng ? pp : (yc - 1) - (pp - ng);
        const BoundaryType b = (BoundaryType)(j < ng ? g.bc[2] : g.bc[3]);
        dst = j * xc + i;
        if (b == BoundaryType::UserFn) {
            from_in = true;
            srcc = dst;
        } else {
            const int js = ghost_source(j, yc, ng, b);
            if (i < ng || i >= xc - ng) {
                srcc = xsource(js, i, from_in);
            } else {
                from_in = false;
                srcc = js * xc + i;
            }
        }
    }
    if (dst < 0) return;
    for (int k = 0; k < ntr; ++k) {
        const int off = (4 + k) * plane;
        Form::put(out, off + dst, Form::get(from_in ? in : out, off + srcc));
    }
}

}  // namespace mh
