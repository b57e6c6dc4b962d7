// 3-D stage pieces: the z sweep and the per-axis ghost passes.
//
// A 3-D stage is two launches. The z sweep below walks each interior (i, j)
// column up the k axis with a (2R+1)-cell register window (the y window of
// the 2-D march turned to z), reconstructs every cell once, solves the
// Riemann problem on each z face and stores the cell's flux difference
// Fz_k - Fz_{k+1} (compute_recon_impl / compute_flux_impl with Axis = 2,
// Hydro.cpp:55-166). The x-y march then runs once per interior z-plane and
// adds that difference after the y one, the order of integrate_flux
// (TimeStepping.cpp:9-35). HBM traffic per cell and stage: the sweep reads
// Q^(s) and writes NV values, the march reads Q^(s), Q^n and the NV values and
// writes Q^(s+1).
//
// Ghost cells in 3-D are filled by three passes, x then y then z, each over
// the full extent of the other two axes, exactly as fill_bcs_impl does
// (Boundaries.hpp:11-131). A pass reads only interior cells of its own axis,
// so its threads never read what another thread of the same pass writes.
#pragma once

#include "functors.hpp"
#include "ghost_dt.hpp"

namespace mh {

struct ZSweepGrid {
    int32_t xc, yc, zc, ng;
    int32_t nx, ny, nz;
    int32_t kchunk;  // interior planes per thread (blockIdx.y selects the chunk)
    int32_t virt = 0, zper = 0;  // ghost planes read from their sources (StageGrid::virt)
};

template <class Form, Reconstruction R, SlopeLimiter SL, RiemannSolver RS>
__device__ inline void zsweep_column(const ZSweepGrid& g, const typename Form::Ptrs& p, typename Form::real* dz) {
    using T = typename Form::real;
    using Raw = typename Form::raw;
    constexpr int ND = 3;
    constexpr int NV = 5;
    constexpr int ORD = stencil_order(R);
    constexpr int WIN = 2 * ORD + 1;

    const int col = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (col >= g.nx * g.ny) return;
    const int i = g.ng + col % g.nx;
    const int j = g.ng + col / g.nx;
    const int k0 = g.ng + (int)blockIdx.y * g.kchunk;
    if (k0 >= g.ng + g.nz) return;
    const int k1 = min(k0 + g.kchunk, g.ng + g.nz);  // faces k0 .. k1, cells k0 .. k1-1 written
    const int pl = g.xc * g.yc;
    const int vs = pl * g.zc;
    const int base = j * g.xc + i;
    const EosT<T> eos = make_eos<T>((double)p.gamma);

    T win[WIN][NV];
    auto load = [&](int k, T (&w)[NV]) {
        const int ks = (Form::kVirtualGhosts && g.virt) ? copy_ghost_source(k, g.nz, g.ng, g.zper) : k;
        Raw q[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = Form::get(p.qs, v * vs + ks * pl + base);
        Form::template to_prim_nd<ND>(p, q, w);
    };
    auto recon = [&](T (&wl)[NV], T (&wr)[NV]) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            auto s = [&](int o) -> T { return win[ORD + o][v]; };
            reconstruct_stencil<R, SL, T>(s, wl[v], wr[v]);
        }
        Form::template floor_face<ND>(p, wl);
        Form::template floor_face<ND>(p, wr);
    };

#pragma unroll
    for (int m = 0; m < WIN; ++m) load(k0 - 1 - ORD + m, win[m]);
    T wl[NV], wr[NV], rprev[NV], fprev[NV], f[NV];
    recon(wl, wr);  // cell k0 - 1
#pragma unroll
    for (int v = 0; v < NV; ++v) rprev[v] = wr[v];

    for (int k = k0; k <= k1; ++k) {
#pragma unroll
        for (int m = 0; m + 1 < WIN; ++m)
#pragma unroll
            for (int v = 0; v < NV; ++v) win[m][v] = win[m + 1][v];
        load(k + ORD, win[WIN - 1]);
        recon(wl, wr);  // cell k
        riemann_flux<RS, 2, ND>(eos, rprev, wl, f);  // face k: RR of cell k-1, RL of cell k
        if (k > k0) {
#pragma unroll
            for (int v = 0; v < NV; ++v) dz[v * vs + (k - 1) * pl + base] = fprev[v] - f[v];
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            fprev[v] = f[v];
            rprev[v] = wr[v];
        }
    }
}

struct Ghost3Grid {
    int32_t xc, yc, zc, ng;
    int32_t bc[6];          // xs, xe, ys, ye, zs, ze (BoundaryType)
    double bconst[6][5];    // Constant-boundary conserved values per face
    double gamma;
    int32_t reduce_dt;      // fold the cells this pass leaves final into the CFL max
};

// One pass of fill_one_bc_impl<Axis, 3> (Boundaries.hpp:11-115) over buffer q.
// With reduce_dt the pass folds c_s + |v| of the ghost cells no later pass
// rewrites (x: interior j and k; y: interior k; z: all), so with the stage's
// interior fold every cell enters the CFL max exactly once.
template <class Form, int Axis>
__global__ __launch_bounds__(256) void ghost_pass3d_kernel(Ghost3Grid g, typename Form::Planes q,
                                                           unsigned long long* amax_bits) {
    using Raw = typename Form::raw;
    constexpr int ND = 3;
    constexpr int NV = 5;
    const int dims[3] = {g.xc, g.yc, g.zc};
    int ld[3] = {g.xc, g.yc, g.zc};
    ld[Axis] = 2 * g.ng;
    const int n = ld[0] * ld[1] * ld[2];
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    double a = 0.0;
    if (t < n) {
        int c[3] = {t % ld[0], (t / ld[0]) % ld[1], t / (ld[0] * ld[1])};
        const int pencil = c[Axis];
        const bool start = pencil < g.ng;
        c[Axis] = start ? pencil : (dims[Axis] - 1) - (pencil - g.ng);
        const int face = 2 * Axis + (start ? 0 : 1);
        const BoundaryType b = (BoundaryType)g.bc[face];
        if (b != BoundaryType::UserFn) {  // userfn: left to the caller (3-D contexts refuse it)
            const int vs = g.xc * g.yc * g.zc;
            Raw v[NV];
            if (b != BoundaryType::Constant) {
                int s[3] = {c[0], c[1], c[2]};
                s[Axis] = ghost_source(c[Axis], dims[Axis], g.ng, b);
                const int src = (s[2] * g.yc + s[1]) * g.xc + s[0];
#pragma unroll
                for (int k = 0; k < NV; ++k) v[k] = Form::get(q, k * vs + src);
            }
            ghost_modify<Form, ND, NV, Axis>(v, b, start, g.bconst[face]);
            const int cell = (c[2] * g.yc + c[1]) * g.xc + c[0];
#pragma unroll
            for (int k = 0; k < NV; ++k) Form::put(q, k * vs + cell, v[k]);
            const bool kin = c[2] >= g.ng && c[2] < g.zc - g.ng;
            const bool jin = c[1] >= g.ng && c[1] < g.yc - g.ng;
            const bool final_value = Axis == 2 || (Axis == 1 && kin) || (Axis == 0 && kin && jin);
            if (g.reduce_dt && final_value) a = Form::template signal_speed_nd<ND>(g.gamma, v);
        }
    }
    if (g.reduce_dt) {
        const uint64_t m = wave_max_u64(__builtin_bit_cast(uint64_t, a));
        if ((threadIdx.x & 63) == 0 && m != 0) fold_amax(amax_bits, blockIdx.x, m);
    }
}

}  // namespace mh
