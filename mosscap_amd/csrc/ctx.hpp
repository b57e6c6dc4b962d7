// Internal: the context behind the opaque mh_ctx handle (shared by the C ABI
// translation units).
#pragma once

#include <cstring>
#include <string>
#include <vector>

#include "../../include/mosscap_hip.h"
#include "launch.hpp"

namespace mh {

inline int stage_kind(TimeStepScheme ts, int stage) {
    switch (ts) {
        case TimeStepScheme::Rk2: return stage == 0 ? kAddFull : kRk2Final;
        case TimeStepScheme::SspRk3: return stage == 0 ? kAddFull : stage == 1 ? kSsp3Second : kSsp3Third;
        default: return stage == 2 ? kSsp4Third : kAddHalf;
    }
}

inline bool valid_bc(int32_t b) { return b >= 0 && b <= 6; }

// Spans per plane of the 2-D march (each wavefront marches one span of a
// strip; the ny rows are split into spans whose heights differ by at most one
// row). rows > 0: spans of about that many rows. rows <= 0, the default
// layout: the most spans whose wavefronts all fit 1024 SIMDs at 4 waves each
// (every SIMD then holds 4 wavefronts and none waits for a second round) with
// at least 8 rows each (the y halo costs 2R+2 rows per span), and spans of at
// most 32 rows when even that needs several rounds.
inline int march_spans(int rows, int strips, int ny, int nz) {
    if (rows > 0) return (ny + rows - 1) / rows;
    const int64_t fill = 4096 / ((int64_t)strips * nz);
    const int64_t n = fill < ny / 8 ? fill : ny / 8;
    const int64_t floor32 = (ny + 31) / 32;
    return (int)(n > floor32 ? n : (floor32 > 0 ? floor32 : 1));
}

// Message returned by mh_last_error(NULL) (functions without a context).
void set_last_error(const std::string& msg);

}  // namespace mh

using namespace mh;

struct mh_ctx {
    int nd = 0, nvar = 0;
    int ntr = 0;                     // tracer planes after the hydro variables
    int last_in = -1;                // input buffer of the most recent stage (update_eos sees it)
    double* eos_y = nullptr;         // analytic LTE EOS fields (mh_eos_fields)
    double* eos_T = nullptr;
    int* eos_flag = nullptr;
    int32_t tab_on = 0;              // tabulated LTE EOS (mh_set_tabulated_eos): update_eos before each stage
    TabParams tab{};
    double* tab_y = nullptr;         // its device tables, [n_rho][n_eint_rho]
    double* tab_lT = nullptr;
    double* tab_yspace = nullptr;    // the y and T fields update_eos keeps (ghosts included)
    double* tab_Tspace = nullptr;
    int* tab_neg = nullptr;          // set when a cell's internal energy is negative
    int32_t gravity = 0;             // registered gravity source (mh_set_gravity)
    double grav[3] = {0.0, 0.0, 0.0};
    double* frx64 = nullptr;         // tracers: density fluxes of the stage (FP64 form)
    double* fry64 = nullptr;
    double* wtr64 = nullptr;         // tracers: mass fractions of the stage input
    float* frx32 = nullptr;          // ... compensated form
    float* fry32 = nullptr;
    float* wtr32 = nullptr;
    double* dz64 = nullptr;          // 3-D: z flux differences of the stage (FP64 form)
    float* dz32 = nullptr;           // ... compensated form
    mh_grid_size sz{};
    double dx = 0.0;
    mh_grid_loc loc{};
    mh_boundaries bdry{};
    Schemes sch{};
    double gamma = 1.4;
    mh_options opt{};
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t ncells = 0, nvals = 0;
    double* q64[3] = {nullptr, nullptr, nullptr};
    float* qhi[3] = {nullptr, nullptr, nullptr};
    float* qlo[3] = {nullptr, nullptr, nullptr};
    double* staging = nullptr;  // FP64 staging for the compensated form
    Clock* clock = nullptr;
    unsigned int* floor_hits = nullptr;
    int cur = 0;         // buffer holding the current state
    int qn = 0;          // buffer holding Q^n during a step
    int next_stage = 0;  // mh_stage bookkeeping
    // The clock's CFL slots hold the largest signal speed of buffer `cur`
    // (set by mh_run, whose last stage folds it; cleared by every call that
    // changes the state or the slots otherwise), so the next mh_run needs no
    // separate CFL pass.
    bool amax_valid = false;
    int nspans = 1;   // spans per plane of the 2-D march (march_spans)
    int strips = 0;
    std::string err;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    // event log for mh_run_timed: (kind, start, stop); kind 0 stage, 1 ghost, 2 clock
    struct Ev {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Ev>* evlog = nullptr;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_next = 0;

    hipError_t ev_get(hipEvent_t& e) {
        if (ev_next < ev_pool.size()) {
            e = ev_pool[ev_next++];
            return hipSuccess;
        }
        hipError_t r = hipEventCreate(&e);
        if (r == hipSuccess) {
            ev_pool.push_back(e);
            ev_next = ev_pool.size();
        }
        return r;
    }
    // Runs `launch` between two recorded events when timing is on.
    template <class F>
    hipError_t timed(int kind, F&& launch) {
        if (!evlog) return launch();
        Ev ev{kind, nullptr, nullptr};
        hipError_t e = ev_get(ev.a);
        if (e == hipSuccess) e = ev_get(ev.b);
        if (e == hipSuccess) e = hipEventRecord(ev.a, stream);
        if (e == hipSuccess) e = launch();
        if (e == hipSuccess) e = hipEventRecord(ev.b, stream);
        if (e == hipSuccess) evlog->push_back(ev);
        return e;
    }
    // user_bc hook (mh_set_user_bc)
    mh_user_bc_fn user_fn = nullptr;
    void* user_data = nullptr;
    int view_buf = -1;  // buffer the running hook acts on
    int hook_rc = 0;    // non-zero return of the last hook call
    int64_t steps = 0;
    double last_dt = 0.0;
    double time = 0.0, prev_output_time = 0.0;

    bool fp32c() const { return opt.precision_form == MH_FORM_FP32C; }
    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        if (hook_rc != 0) {
            err = std::string(what) + ": the user_bc hook returned " + std::to_string(hook_rc);
            hook_rc = 0;
            return MH_ERR_INVALID;
        }
        err = std::string(what) + ": " + hipGetErrorString(e);
        return MH_ERR_HIP;
    }
    FormF64::Planes p64(int k) const { return FormF64::Planes{q64[k]}; }
    FormF32c::Planes p32(int k) const { return FormF32c::Planes{qhi[k], qlo[k]}; }

    GhostGrid ghost_grid(int reduce) const {
        GhostGrid g{};
        g.xc = sz.xc;
        g.yc = sz.yc;
        g.ng = sz.ng;
        g.bc[0] = bdry.xs;
        g.bc[1] = bdry.xe;
        g.bc[2] = bdry.ys;
        g.bc[3] = bdry.ye;
        std::memcpy(g.bconst[0], bdry.xs_const, sizeof(double) * 5);
        std::memcpy(g.bconst[1], bdry.xe_const, sizeof(double) * 5);
        std::memcpy(g.bconst[2], bdry.ys_const, sizeof(double) * 5);
        std::memcpy(g.bconst[3], bdry.ye_const, sizeof(double) * 5);
        g.gamma = gamma;
        g.reduce_dt = reduce;
        return g;
    }

    Ghost3Grid ghost3_grid(int reduce) const {
        Ghost3Grid g{};
        g.gamma = gamma;
        g.reduce_dt = reduce;
        g.xc = sz.xc;
        g.yc = sz.yc;
        g.zc = sz.zc;
        g.ng = sz.ng;
        const int32_t b[6] = {bdry.xs, bdry.xe, bdry.ys, bdry.ye, bdry.zs, bdry.ze};
        const double* cv[6] = {bdry.xs_const, bdry.xe_const, bdry.ys_const, bdry.ye_const, bdry.zs_const,
                               bdry.ze_const};
        for (int f = 0; f < 6; ++f) {
            g.bc[f] = b[f];
            std::memcpy(g.bconst[f], cv[f], sizeof(double) * 5);
        }
        return g;
    }

    // Ghost fill of `out` after a stage from `in`. In 3-D the three axis
    // passes run on `out` alone.
    hipError_t ghost(int in, int out, int reduce) {
        unsigned long long* amax = amax_slots(clock);
        hipError_t e = timed(1, [&] {
            if (nd == 3)
                return fp32c() ? launch_ghost3d_f32c(ghost3_grid(reduce), p32(out), amax, stream)
                               : launch_ghost3d_f64(ghost3_grid(reduce), p64(out), amax, stream);
            hipError_t r = fp32c() ? launch_ghost_f32c(nd, ghost_grid(reduce), p32(in), p32(out), amax, stream)
                                   : launch_ghost_f64(nd, ghost_grid(reduce), p64(in), p64(out), amax, stream);
            if (r != hipSuccess || ntr == 0) return r;
            return fp32c() ? launch_tracer_ghost_f32c(ghost_grid(0), ntr, p32(in), p32(out), stream)
                           : launch_tracer_ghost_f64(ghost_grid(0), ntr, p64(in), p64(out), stream);
        });
        if (e != hipSuccess || !user_fn) return e;
        view_buf = out;
        hook_rc = user_fn(this, user_data);
        view_buf = -1;
        return hook_rc != 0 ? hipErrorUnknown : hipGetLastError();
    }

    hipError_t finalize(int advance) {
        return timed(2, [&] { return launch_clock_finalize(clock, advance, stream); });
    }

    hipError_t cfl_all(int k) {
        unsigned long long* amax = amax_slots(clock);
        if (fp32c()) return launch_cfl_f32c(nd, p32(k), (int)ncells, gamma, amax, stream);
        return launch_cfl_f64(nd, p64(k), (int)ncells, gamma, amax, stream);
    }

    // Tabulated EOS update of buffer `in` into `dst` (update_eos at the start of
    // compute_hydro_fluxes, Hydro.cpp:325-328).
    hipError_t tab_update(int in, int dst) {
        return fp32c() ? launch_tab_eos_f32c(nd, p32(in), p32(dst), (int)ncells, nvar, tab, tab_y, tab_lT, tab_yspace,
                                             tab_Tspace, tab_neg, stream)
                       : launch_tab_eos_f64(nd, p64(in), p64(dst), (int)ncells, nvar, tab, tab_y, tab_lT, tab_yspace,
                                            tab_Tspace, tab_neg, stream);
    }

    // One stage from buffer `in` into `out` with Q^n in `qn_k`, then ghost fill.
    // With the tabulated EOS the stage reads its input after update_eos. The
    // reference copies Q_old before the first stage's update_eos
    // (TimeStepping.cpp:52, 129, 226), so for stage 0 (in == qn_k) the updated
    // input goes to the third buffer, which no stage of the step has written
    // yet, and Q^n keeps the values it had.
    // With `fill` false and virtual_ghosts(), the output's ghosts are left
    // stale: the next stage reads ghost rows and columns from their sources.
    // fin: StageGrid::fin of a 2-D stage (the step's clock work inside it)
    hipError_t stage(int in, int qn_k, int out, int kind, int reduce, bool fill = true, int fin = 0) {
        const int src = !tab_on ? in : (in == qn_k ? 3 - in - out : in);
        hipError_t e = timed(0, [&] {
            hipError_t r = tab_on ? tab_update(in, src) : hipSuccess;
            if (r == hipSuccess) r = stage_kernel(src, qn_k, out, kind, reduce, fin);
            return r;
        });
        if (e != hipSuccess || (!fill && virtual_ghosts())) return e;
        return ghost(src, out, reduce);
    }

    // 2-D and 3-D grids whose faces are all periodic or zerograd, with no
    // user_bc hook, no tracer fields and no tabulated EOS (its update runs
    // over the ghosts of each stage's input): every ghost is a copy of an
    // interior cell, so the stage kernels read ghost rows, columns and planes
    // from their sources (StageGrid::virt) and a step needs one ghost fill,
    // at its end, instead of one per stage (each a graph node and a kernel
    // boundary). A stage reads the values the reference's fill would have
    // written (Boundaries.hpp), so the results are the same bits; leaving the
    // ghosts out of the CFL max changes nothing (they copy interior values).
    // mh_stage still fills after every stage. The adjoint's recomputation
    // skips the fills too; its reverse sweep reads ghosts from their sources.
    bool virtual_ghosts() const {
        const auto copy_bc = [](int32_t b) {
            return b == (int32_t)BoundaryType::Periodic || b == (int32_t)BoundaryType::ZeroGrad;
        };
        return nd >= 2 && !user_fn && ntr == 0 && !tab_on && copy_bc(bdry.xs) && copy_bc(bdry.xe) &&
               copy_bc(bdry.ys) && copy_bc(bdry.ye) && (nd < 3 || (copy_bc(bdry.zs) && copy_bc(bdry.ze)));
    }
    void set_virt(StageGrid& g) const {
        g.virt = virtual_ghosts() ? 1 : 0;
        g.xper = bdry.xs == (int32_t)BoundaryType::Periodic;
        g.yper = bdry.ys == (int32_t)BoundaryType::Periodic;
        g.zper = bdry.zs == (int32_t)BoundaryType::Periodic;
    }

    // 3-D march grid (the interior fold of the CFL max with `reduce`)
    StageGrid grid3(int kind, int reduce) const {
        StageGrid g{sz.xc, sz.yc, sz.ng, sz.xc - 2 * sz.ng, sz.yc - 2 * sz.ng, strips, nspans, kind, reduce};
        g.zc = sz.zc;
        g.nz = sz.zc - 2 * sz.ng;
        set_virt(g);
        return g;
    }

    hipError_t stage_kernel(int in, int qn_k, int out, int kind, int reduce, int fin = 0) {
        hipError_t e;
        const TracerGrid tg{sz.xc, sz.yc, sz.ng, sz.xc - 2 * sz.ng, sz.yc - 2 * sz.ng, ntr, kind};
        if (fp32c()) {
            FormF32c::Ptrs p{p32(in), p32(qn_k), p32(out), &clock->dt, dx, gamma, gamma - 1.0,
                             amax_slots(clock), floor_hits, opt.rho_floor, opt.p_floor, frx32, fry32,
                             gravity, grav[0], grav[1]};
            if (nd == 2) {
                StageGrid g{sz.xc, sz.yc, sz.ng, sz.xc - 2 * sz.ng, sz.yc - 2 * sz.ng, strips, nspans, kind, reduce};
                set_virt(g);
                g.fin = fin;
                g.clock = clock;
                e = launch_stage2d_f32c(sch, g, p, stream);
                if (e == hipSuccess && ntr > 0) e = launch_tracer_f32c(sch, tg, p, wtr32, stream);
            } else if (nd == 3) {
                p.gz = grav[2];
                e = launch_stage3d_f32c(sch, grid3(kind, reduce), p, dz32, stream);
            } else {
                e = launch_stage1d_f32c(sch, kind, reduce, sz.xc, sz.ng, p, stream);
            }
        } else {
            FormF64::Ptrs p{p64(in), p64(qn_k), p64(out), &clock->dt, dx, gamma, amax_slots(clock), nullptr,
                            frx64, fry64, gravity, grav[0], grav[1]};
            if (nd == 2) {
                StageGrid g{sz.xc, sz.yc, sz.ng, sz.xc - 2 * sz.ng, sz.yc - 2 * sz.ng, strips, nspans, kind, reduce};
                set_virt(g);
                g.fin = fin;
                g.clock = clock;
                e = launch_stage2d_f64(sch, g, p, stream);
                if (e == hipSuccess && ntr > 0) e = launch_tracer_f64(sch, tg, p, wtr64, stream);
            } else if (nd == 3) {
                p.gz = grav[2];
                e = launch_stage3d_f64(sch, grid3(kind, reduce), p, dz64, stream);
            } else {
                e = launch_stage1d_f64(sch, kind, reduce, sz.xc, sz.ng, p, stream);
            }
        }
        return e;
    }

    // FP64 stage between arbitrary device buffers (adjoint recomputation).
    // dt_dev: the step's dt in device memory. With virtual_ghosts() the
    // input's ghosts are read from their sources and the output's are left
    // unfilled (the reverse sweep reads them the same way).
    hipError_t stage_ptr(const double* in, const double* qn_p, double* out, int kind, const double* dt_dev) {
        FormF64::Ptrs p{FormF64::Planes{(double*)in}, FormF64::Planes{(double*)qn_p}, FormF64::Planes{out}, dt_dev,
                        dx, gamma, amax_slots(clock), nullptr, nullptr, nullptr, gravity, grav[0], grav[1]};
        StageGrid g{sz.xc, sz.yc, sz.ng, sz.xc - 2 * sz.ng, sz.yc - 2 * sz.ng, strips, nspans, kind, 0};
        set_virt(g);
        hipError_t e = nd == 2 ? launch_stage2d_f64(sch, g, p, stream)
                               : launch_stage1d_f64(sch, kind, 0, sz.xc, sz.ng, p, stream);
        if (e != hipSuccess || virtual_ghosts()) return e;
        return launch_ghost_f64(nd, ghost_grid(0), FormF64::Planes{(double*)in}, FormF64::Planes{out}, nullptr, stream);
    }

    // adjoint records (mh_adjoint_forward / mh_adjoint_backward)
    std::vector<double*> ckpt;
    std::vector<double> adj_dts;
    double* adj_dts_dev = nullptr;  // the same step sizes on the device (the recomputation reads them)
    int adj_every = 0;
    double adj_bwd_ms = 0.0;  // device time of the last reverse sweep
    // reverse-sweep scratch kept between mh_adjoint_backward calls (nvals
    // doubles each) and pinned host staging for its transfers
    std::vector<double*> adj_scratch;
    double* adj_host = nullptr;
    hipError_t adjoint_scratch(size_t nbuf) {
        hipError_t e = hipSuccess;
        if (adj_scratch.size() != nbuf) {
            free_adjoint_scratch();
            adj_scratch.assign(nbuf, nullptr);
            for (size_t k = 0; k < nbuf && e == hipSuccess; ++k)
                e = hipMalloc(&adj_scratch[k], (size_t)nvals * sizeof(double));
        }
        if (e == hipSuccess && !adj_host) e = hipHostMalloc(&adj_host, (size_t)nvals * sizeof(double));
        if (e != hipSuccess) free_adjoint_scratch();
        return e;
    }
    void free_adjoint_scratch() {
        for (double* p : adj_scratch)
            if (p) hipFree(p);
        adj_scratch.clear();
        if (adj_host) hipHostFree(adj_host);
        adj_host = nullptr;
    }
    void free_adjoint() {
        for (double* p : ckpt)
            if (p) hipFree(p);
        ckpt.clear();
        adj_dts.clear();
        if (adj_dts_dev) hipFree(adj_dts_dev);
        adj_dts_dev = nullptr;
    }

    // All stages of one step starting from `cur`; returns the new cur.
    // With a user_bc hook the reduction is a separate pass after the hook.
    // With virtual_ghosts() no stage fills ghosts: the caller fills those of
    // the last step's result (fill_final).
    // With fuse_clock (clock_in_stages()) the step's clock finalize runs
    // inside its first two stages (StageGrid::fin) instead of a kernel of its
    // own before them.
    int enqueue_step(int k, bool reduce_last, hipError_t& e, bool fuse_clock = false) {
        const int ns = num_stages(sch.stepper);
        const bool fused = reduce_last && !user_fn;
        int in = k;
        for (int s = 0; s < ns; ++s) {
            last_in = in;
            const int out = (in == k) ? (k + 1) % 3 : 3 - in - k;
            const int fin = fuse_clock ? (s == 0 ? 1 : (s == 1 ? 2 : 0)) : 0;
            e = stage(in, k, out, stage_kind(sch.stepper, s), (fused && s == ns - 1) ? 1 : 0, false, fin);
            if (e != hipSuccess) return k;
            in = out;
        }
        if (reduce_last && !fused) e = cfl_all(in);
        return in;
    }
    // 2-D steps of three or more stages (the second stage clears the CFL
    // slots that only the last one folds into), no user_bc hook
    bool clock_in_stages() const { return nd == 2 && num_stages(sch.stepper) >= 3 && !user_fn; }

    // ghosts of buffer k after steps run with virtual ghosts
    hipError_t fill_final(int k) { return virtual_ghosts() ? ghost(k, k, 0) : hipSuccess; }

    // kernel arguments changed: graphs captured before must be rebuilt
    void drop_graphs() {
        for (int k = 0; k < 3; ++k) {
            if (graph[k]) hipGraphExecDestroy(graph[k]);
            graph[k] = nullptr;
        }
    }

    void free_all() {
        free_adjoint();
        free_adjoint_scratch();
        for (int k = 0; k < 3; ++k) {
            if (graph[k]) hipGraphExecDestroy(graph[k]);
            graph[k] = nullptr;
            if (q64[k]) hipFree(q64[k]);
            if (qhi[k]) hipFree(qhi[k]);
            if (qlo[k]) hipFree(qlo[k]);
            q64[k] = nullptr;
            qhi[k] = qlo[k] = nullptr;
        }
        for (hipEvent_t ev : ev_pool) hipEventDestroy(ev);
        ev_pool.clear();
        for (void* t : {(void*)frx64, (void*)fry64, (void*)wtr64, (void*)frx32, (void*)fry32, (void*)wtr32,
                        (void*)eos_y, (void*)eos_T, (void*)eos_flag, (void*)dz64, (void*)dz32, (void*)tab_y,
                        (void*)tab_lT, (void*)tab_yspace, (void*)tab_Tspace, (void*)tab_neg})
            if (t) hipFree(t);
        tab_y = tab_lT = tab_yspace = tab_Tspace = nullptr;
        tab_neg = nullptr;
        dz64 = nullptr;
        dz32 = nullptr;
        frx64 = fry64 = wtr64 = nullptr;
        frx32 = fry32 = wtr32 = nullptr;
        eos_y = eos_T = nullptr;
        eos_flag = nullptr;
        if (staging) hipFree(staging);
        if (clock) hipFree(clock);
        if (floor_hits) hipFree(floor_hits);
        staging = nullptr;
        clock = nullptr;
        floor_hits = nullptr;
        if (stream) hipStreamDestroy(stream);
        stream = nullptr;
    }
};

