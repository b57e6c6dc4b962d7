// C ABI of the batched 1-D ensemble (include/mosscap_hip.h, mh_batch_*).
#include "../../include/mosscap_hip.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "launch.hpp"
#include "launch_batch.hpp"

using namespace mh;

namespace {

thread_local std::string g_batch_create_error;

int stage_kind_of(TimeStepScheme ts, int stage) {
    switch (ts) {
        case TimeStepScheme::Rk2: return stage == 0 ? kAddFull : kRk2Final;
        case TimeStepScheme::SspRk3: return stage == 0 ? kAddFull : stage == 1 ? kSsp3Second : kSsp3Third;
        default: return stage == 2 ? kSsp4Third : kAddHalf;
    }
}

}  // namespace

struct mh_batch {
    int nprob = 0, ng = 0;
    Schemes sch{};
    mh_boundaries bdry{};
    double gamma = 1.4;
    mh_options opt{};
    int device = 0, cus = 256;
    hipStream_t stream = nullptr;
    std::vector<int> len, base, xc;
    std::vector<double> dx;
    int64_t nvals = 0;
    double* q64[3] = {nullptr, nullptr, nullptr};
    float* qhi[3] = {nullptr, nullptr, nullptr};
    float* qlo[3] = {nullptr, nullptr, nullptr};
    double* staging = nullptr;
    int4* items = nullptr;
    int* d_base = nullptr;
    int* d_xc = nullptr;
    double* d_dx = nullptr;
    double* d_dt = nullptr;
    unsigned long long* d_amax = nullptr;
    double* d_time = nullptr;
    unsigned int* floor_hits = nullptr;
    int nitems = 0, cur = 0, waves = 0;
    hipGraphExec_t graph[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    int64_t steps = 0;
    std::string err;

    bool fp32c() const { return opt.precision_form == MH_FORM_FP32C; }
    int hip_fail(hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return MH_ERR_HIP;
    }
    BatchArrays arrays() const { return BatchArrays{items, d_base, d_xc, d_dx, d_dt, d_amax, d_time}; }
    FormF64::Planes p64(int k) const { return FormF64::Planes{q64[k]}; }
    FormF32c::Planes p32(int k) const { return FormF32c::Planes{qhi[k], qlo[k]}; }
    GhostGrid ghost_grid(int reduce) const {
        GhostGrid g{};
        g.xc = 0;
        g.yc = 1;
        g.ng = ng;
        g.bc[0] = bdry.xs;
        g.bc[1] = bdry.xe;
        g.bc[2] = g.bc[3] = (int32_t)BoundaryType::Periodic;
        std::memcpy(g.bconst[0], bdry.xs_const, sizeof(double) * 5);
        std::memcpy(g.bconst[1], bdry.xe_const, sizeof(double) * 5);
        g.gamma = gamma;
        g.reduce_dt = reduce;
        return g;
    }
    void drop_graphs() {
        for (auto& row : graph)
            for (auto& g : row) {
                if (g) hipGraphExecDestroy(g);
                g = nullptr;
            }
    }

    // Periodic or zerograd faces: the persistent stage reads ghosts from
    // their source cells, so the stages of a step fill no ghosts and
    // mh_batch_run fills the final state's once (the ghosts would add nothing
    // to the per-problem CFL max: they copy interior values).
    bool virtual_ghosts() const {
        const auto copy_bc = [](int32_t x) {
            return x == (int32_t)BoundaryType::Periodic || x == (int32_t)BoundaryType::ZeroGrad;
        };
        return copy_bc(bdry.xs) && copy_bc(bdry.xe);
    }

    // one stage for every problem: persistent (mode 0) or one launch per problem (mode 1)
    // fin, fin_cfl, fin_max_time: BatchGrid::fin and its arguments (persistent mode)
    hipError_t stage(int mode, int in, int qn_k, int out, int kind, int reduce, int fin = 0, double fin_cfl = 0.0,
                     double fin_max_time = 0.0) {
        hipError_t e = hipSuccess;
        const BatchArrays b = arrays();
        if (mode == 0) {
            BatchGrid g{ng, nitems, nprob, kind, reduce, 0, 0};
            g.fin = fin;
            g.max_cfl = fin_cfl;
            g.max_time = fin_max_time;
            g.virt = virtual_ghosts() ? 1 : 0;
            g.per = bdry.xs == (int32_t)BoundaryType::Periodic;
            const int grid = std::min(waves > 0 ? waves : cus * 12, nitems);
            if (fp32c()) {
                FormF32c::Ptrs p{p32(in), p32(qn_k), p32(out), d_dt, 0.0, gamma, gamma - 1.0, d_amax,
                                 floor_hits, opt.rho_floor, opt.p_floor};
                e = launch_batch_stage_f32c(sch, g, grid, p, b, stream);
                if (e == hipSuccess && !g.virt)
                    e = launch_batch_ghost_f32c(ghost_grid(reduce), nprob, p32(in), p32(out), b, stream);
            } else {
                FormF64::Ptrs p{p64(in), p64(qn_k), p64(out), d_dt, 0.0, gamma, d_amax, nullptr};
                e = launch_batch_stage_f64(sch, g, grid, p, b, stream);
                if (e == hipSuccess && !g.virt)
                    e = launch_batch_ghost_f64(ghost_grid(reduce), nprob, p64(in), p64(out), b, stream);
            }
            return e;
        }
        for (int pr = 0; pr < nprob && e == hipSuccess; ++pr) {
            GhostGrid gg = ghost_grid(reduce);
            gg.xc = xc[pr];
            const size_t o = (size_t)base[pr];
            if (fp32c()) {
                auto P = [&](int k) { return FormF32c::Planes{qhi[k] + o, qlo[k] + o}; };
                FormF32c::Ptrs p{P(in), P(qn_k), P(out), d_dt + pr, dx[pr], gamma, gamma - 1.0, d_amax + pr,
                                 floor_hits, opt.rho_floor, opt.p_floor};
                e = launch_stage1d_f32c(sch, kind, reduce, xc[pr], ng, p, stream);
                if (e == hipSuccess) e = launch_ghost_f32c(1, gg, P(in), P(out), d_amax + pr, stream);
            } else {
                auto P = [&](int k) { return FormF64::Planes{q64[k] + o}; };
                FormF64::Ptrs p{P(in), P(qn_k), P(out), d_dt + pr, dx[pr], gamma, d_amax + pr, nullptr};
                e = launch_stage1d_f64(sch, kind, reduce, xc[pr], ng, p, stream);
                if (e == hipSuccess) e = launch_ghost_f64(1, gg, P(in), P(out), d_amax + pr, stream);
            }
        }
        return e;
    }

    // With fuse_clock (clock_in_stages(mode)) the per-problem finalize runs
    // inside the step's first two stages (BatchGrid::fin), not as a kernel.
    int enqueue_step(int mode, int k, double max_cfl, double max_time, hipError_t& e, bool fuse_clock = false) {
        e = fuse_clock ? hipSuccess : launch_batch_finalize(nprob, max_cfl, max_time, 1, arrays(), d_dt, stream);
        const int ns = num_stages(sch.stepper);
        int in = k;
        for (int s = 0; s < ns && e == hipSuccess; ++s) {
            const int out = (in == k) ? (k + 1) % 3 : 3 - in - k;
            const int fin = fuse_clock ? (s == 0 ? 1 : (s == 1 ? 2 : 0)) : 0;
            e = stage(mode, in, k, out, stage_kind_of(sch.stepper, s), s == ns - 1 ? 1 : 0, fin, max_cfl, max_time);
            in = out;
        }
        return in;
    }
    // persistent steps of three or more stages (the second clears the CFL
    // maxima that only the last folds into)
    bool clock_in_stages(int mode) const { return mode == 0 && num_stages(sch.stepper) >= 3; }
    double graph_cfl = -1.0, graph_max_time = -1.0;  // the values the graphs were captured with

    void free_all() {
        drop_graphs();
        for (int k = 0; k < 3; ++k) {
            if (q64[k]) hipFree(q64[k]);
            if (qhi[k]) hipFree(qhi[k]);
            if (qlo[k]) hipFree(qlo[k]);
            q64[k] = nullptr;
            qhi[k] = qlo[k] = nullptr;
        }
        void* ps[] = {staging, items, d_base, d_xc, d_dx, d_dt, d_amax, d_time, floor_hits};
        for (void* p : ps)
            if (p) hipFree(p);
        staging = nullptr;
        items = nullptr;
        d_base = d_xc = nullptr;
        d_dx = d_dt = d_time = nullptr;
        d_amax = nullptr;
        floor_hits = nullptr;
        if (stream) hipStreamDestroy(stream);
        stream = nullptr;
    }
};

extern "C" {

const char* mh_batch_last_error(const mh_batch* b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int mh_batch_create(int32_t nprob, const int32_t* lengths, int32_t ng, double x_dim, const mh_boundaries* bdry,
                    const mh_schemes* schemes, double gamma, const mh_options* opts, mh_batch** out) {
    g_batch_create_error.clear();
    auto bad = [](const std::string& m) {
        g_batch_create_error = m;
        return MH_ERR_INVALID;
    };
    if (!out || !lengths || !bdry || !schemes || nprob < 1) return bad("mh_batch_create: bad argument");
    *out = nullptr;
    if (schemes->reconstruction < 0 || schemes->reconstruction > 3 || schemes->slope_limit < 0 ||
        schemes->slope_limit > 2 || schemes->riemann_solver < 0 || schemes->riemann_solver > 2 ||
        schemes->time_stepper < 0 || schemes->time_stepper > 2)
        return bad("unknown scheme enum value");
    if (ng < min_ghost_cells((Reconstruction)schemes->reconstruction))
        return bad("Insufficient ghost cells for the reconstruction scheme");
    if (bdry->xs < 0 || bdry->xs > 6 || bdry->xe < 0 || bdry->xe > 6 || ((bdry->xs == 1) != (bdry->xe == 1)))
        return bad("bad or unpaired boundary types");
    if (!(x_dim > 0.0) || !(gamma > 1.0)) return bad("x_dim must be > 0 and gamma > 1");
    mh_batch* b = new mh_batch();
    b->nprob = nprob;
    b->ng = ng;
    b->sch = Schemes{(Reconstruction)schemes->reconstruction, (SlopeLimiter)schemes->slope_limit,
                     (RiemannSolver)schemes->riemann_solver, (TimeStepScheme)schemes->time_stepper};
    b->bdry = *bdry;
    b->gamma = gamma;
    if (opts) b->opt = *opts;
    std::vector<int4> items;  // (problem, first output cell, block offset, cells incl. ghosts)
    int64_t off = 0;
    for (int p = 0; p < nprob; ++p) {
        const int n = lengths[p];
        if (n < ng) {
            delete b;
            return bad("every problem needs at least num_ghost interior cells");
        }
        b->len.push_back(n);
        b->xc.push_back(n + 2 * ng);
        b->base.push_back((int)off);
        b->dx.push_back(x_dim / n);  // setup_grid, SetupModules.hpp:29-31
        off += 3LL * (n + 2 * ng);
        for (int i0 = ng; i0 < ng + n; i0 += kStripOut)
            items.push_back(make_int4(p, i0, b->base.back(), n + 2 * ng));
    }
    if (off >= (int64_t)1 << 31) {
        delete b;
        return bad("ensemble larger than 2^31 values");
    }
    b->nvals = off;
    b->nitems = (int)items.size();
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) {
        if (b->fp32c()) {
            e = hipMalloc(&b->qhi[k], off * sizeof(float));
            if (e == hipSuccess) e = hipMalloc(&b->qlo[k], off * sizeof(float));
            if (e == hipSuccess) e = hipMemsetAsync(b->qhi[k], 0, off * sizeof(float), b->stream);
            if (e == hipSuccess) e = hipMemsetAsync(b->qlo[k], 0, off * sizeof(float), b->stream);
        } else {
            e = hipMalloc(&b->q64[k], off * sizeof(double));
            if (e == hipSuccess) e = hipMemsetAsync(b->q64[k], 0, off * sizeof(double), b->stream);
        }
    }
    if (e == hipSuccess && b->fp32c()) e = hipMalloc(&b->staging, off * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&b->items, items.size() * sizeof(int4));
    if (e == hipSuccess) e = hipMalloc(&b->d_base, nprob * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&b->d_xc, nprob * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&b->d_dx, nprob * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&b->d_dt, nprob * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&b->d_amax, nprob * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&b->d_time, nprob * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&b->floor_hits, sizeof(unsigned int));
    if (e == hipSuccess) e = hipMemcpy(b->items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_base, b->base.data(), nprob * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_xc, b->xc.data(), nprob * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_dx, b->dx.data(), nprob * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_dt, 0, nprob * sizeof(double), b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_amax, 0, nprob * sizeof(unsigned long long), b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_time, 0, nprob * sizeof(double), b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->floor_hits, 0, sizeof(unsigned int), b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        g_batch_create_error = std::string("mh_batch_create: ") + hipGetErrorString(e);
        b->free_all();
        delete b;
        return MH_ERR_HIP;
    }
    *out = b;
    return MH_OK;
}

void mh_batch_destroy(mh_batch* b) {
    if (!b) return;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    b->free_all();
    delete b;
}

int64_t mh_batch_num_values(const mh_batch* b) { return b ? b->nvals : 0; }

int mh_batch_offsets(const mh_batch* b, int64_t* offsets) {
    if (!b || !offsets) return MH_ERR_INVALID;
    for (int p = 0; p < b->nprob; ++p) offsets[p] = b->base[p];
    return MH_OK;
}

void* mh_batch_stream(mh_batch* b) { return b ? (void*)b->stream : nullptr; }

int mh_batch_upload(mh_batch* b, const double* q) {
    if (!b || !q) return MH_ERR_INVALID;
    hipError_t e = hipSetDevice(b->device);
    const size_t bytes = (size_t)b->nvals * sizeof(double);
    if (b->fp32c()) {
        if (e == hipSuccess) e = hipMemcpyAsync(b->staging, q, bytes, hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess) e = launch_split_f64_to_f32c(b->staging, b->p32(b->cur), (int)b->nvals, b->stream);
    } else if (e == hipSuccess) {
        e = hipMemcpyAsync(b->q64[b->cur], q, bytes, hipMemcpyHostToDevice, b->stream);
    }
    if (e == hipSuccess) e = hipMemsetAsync(b->d_time, 0, b->nprob * sizeof(double), b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? MH_OK : b->hip_fail(e, "mh_batch_upload");
}

int mh_batch_download(mh_batch* b, double* q) {
    if (!b || !q) return MH_ERR_INVALID;
    hipError_t e = hipSetDevice(b->device);
    const size_t bytes = (size_t)b->nvals * sizeof(double);
    if (b->fp32c()) {
        if (e == hipSuccess) e = launch_join_f32c_to_f64(b->p32(b->cur), b->staging, (int)b->nvals, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(q, b->staging, bytes, hipMemcpyDeviceToHost, b->stream);
    } else if (e == hipSuccess) {
        e = hipMemcpyAsync(q, b->q64[b->cur], bytes, hipMemcpyDeviceToHost, b->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? MH_OK : b->hip_fail(e, "mh_batch_download");
}

int mh_batch_fill_bcs(mh_batch* b) {
    if (!b) return MH_ERR_INVALID;
    hipError_t e = hipSetDevice(b->device);
    const BatchArrays arr = b->arrays();
    if (e == hipSuccess) {
        e = b->fp32c() ? launch_batch_ghost_f32c(b->ghost_grid(0), b->nprob, b->p32(b->cur), b->p32(b->cur), arr, b->stream)
                       : launch_batch_ghost_f64(b->ghost_grid(0), b->nprob, b->p64(b->cur), b->p64(b->cur), arr, b->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? MH_OK : b->hip_fail(e, "mh_batch_fill_bcs");
}

int mh_batch_run(mh_batch* b, int64_t nsteps, double max_cfl, double max_time, int32_t mode, int32_t use_graph) {
    if (!b || nsteps < 0 || (mode != 0 && mode != 1)) return MH_ERR_INVALID;
    hipError_t e = hipSetDevice(b->device);
    const BatchArrays arr = b->arrays();
    // CFL max of the starting state over every cell of every problem
    if (e == hipSuccess) e = hipMemsetAsync(b->d_amax, 0, b->nprob * sizeof(unsigned long long), b->stream);
    if (e == hipSuccess) {
        e = b->fp32c() ? launch_batch_cfl_f32c(b->nprob, b->p32(b->cur), b->gamma, arr, b->stream)
                       : launch_batch_cfl_f64(b->nprob, b->p64(b->cur), b->gamma, arr, b->stream);
    }
    // the captured steps carry max_cfl and max_time as kernel arguments
    if (use_graph && (max_cfl != b->graph_cfl || max_time != b->graph_max_time)) {
        b->drop_graphs();
        b->graph_cfl = max_cfl;
        b->graph_max_time = max_time;
    }
    const int advance = num_stages(b->sch.stepper) == 3 ? 1 : 2;  // buffer rotation per step
    for (int64_t s = 0; s < nsteps && e == hipSuccess; ++s) {
        const int k = b->cur;
        if (use_graph) {
            hipGraphExec_t& ge = b->graph[mode][k];
            if (!ge) {
                hipGraph_t g = nullptr;
                e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) break;
                hipError_t ce = hipSuccess;
                b->enqueue_step(mode, k, max_cfl, max_time, ce, b->clock_in_stages(mode));
                e = hipStreamEndCapture(b->stream, &g);
                if (e == hipSuccess) e = ce;
                if (e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                if (g) hipGraphDestroy(g);
                if (e != hipSuccess) break;
            }
            e = hipGraphLaunch(ge, b->stream);
            b->cur = (k + advance) % 3;
        } else {
            b->cur = b->enqueue_step(mode, k, max_cfl, max_time, e);
        }
        b->steps += 1;
    }
    if (e == hipSuccess && mode == 0 && nsteps > 0 && b->virtual_ghosts())  // the final state's ghosts
        e = b->fp32c() ? launch_batch_ghost_f32c(b->ghost_grid(0), b->nprob, b->p32(b->cur), b->p32(b->cur), arr,
                                                 b->stream)
                       : launch_batch_ghost_f64(b->ghost_grid(0), b->nprob, b->p64(b->cur), b->p64(b->cur), arr,
                                                b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? MH_OK : b->hip_fail(e, "mh_batch_run");
}

int mh_batch_times(mh_batch* b, double* times, double* last_dt) {
    if (!b) return MH_ERR_INVALID;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess && times) e = hipMemcpy(times, b->d_time, b->nprob * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && last_dt) e = hipMemcpy(last_dt, b->d_dt, b->nprob * sizeof(double), hipMemcpyDeviceToHost);
    return e == hipSuccess ? MH_OK : b->hip_fail(e, "mh_batch_times");
}

int mh_batch_set_waves(mh_batch* b, int32_t waves) {
    if (!b || waves < 0) return MH_ERR_INVALID;
    b->waves = waves;
    b->drop_graphs();
    return MH_OK;
}

int mh_batch_autotune(mh_batch* b, int32_t steps, const int32_t* waves_per_cu, int32_t* best_waves, double* ms_out) {
    if (!b || !best_waves || steps < 1) return MH_ERR_INVALID;
    static const int32_t kDefault[] = {4, 8, 12, 16, 20, 24, 28, 32, 0};
    const int32_t* cand = waves_per_cu ? waves_per_cu : kDefault;
    hipError_t e = hipSetDevice(b->device);
    const size_t bytes = (size_t)b->nvals * (b->fp32c() ? sizeof(float) : sizeof(double));
    void* keep_a = nullptr;
    void* keep_b = nullptr;
    double* keep_t = nullptr;  // per-problem time, then last dt
    unsigned int keep_hits = 0;
    const size_t tbytes = (size_t)b->nprob * sizeof(double);
    if (e == hipSuccess) e = hipMalloc(&keep_a, bytes);
    if (e == hipSuccess && b->fp32c()) e = hipMalloc(&keep_b, bytes);
    if (e == hipSuccess) e = hipMalloc(&keep_t, 2 * tbytes);
    const int start = b->cur;
    // what the trial runs change: the state, the per-problem clocks and the floor-hit count
    auto stash = [&](bool restore) -> hipError_t {
        const auto copy = [&](void* live, void* keep, size_t n) {
            return restore ? hipMemcpyAsync(live, keep, n, hipMemcpyDeviceToDevice, b->stream)
                           : hipMemcpyAsync(keep, live, n, hipMemcpyDeviceToDevice, b->stream);
        };
        void* d0 = b->fp32c() ? (void*)b->qhi[start] : (void*)b->q64[start];
        hipError_t r = copy(d0, keep_a, bytes);
        if (r == hipSuccess && b->fp32c()) r = copy(b->qlo[start], keep_b, bytes);
        if (r == hipSuccess) r = copy(b->d_time, keep_t, tbytes);
        if (r == hipSuccess) r = copy(b->d_dt, keep_t + b->nprob, tbytes);
        if (r == hipSuccess)
            r = restore ? hipMemcpyAsync(b->floor_hits, &keep_hits, sizeof(keep_hits), hipMemcpyHostToDevice, b->stream)
                        : hipMemcpyAsync(&keep_hits, b->floor_hits, sizeof(keep_hits), hipMemcpyDeviceToHost, b->stream);
        if (r == hipSuccess) r = hipStreamSynchronize(b->stream);  // keep_hits is a host word
        return r;
    };
    if (e == hipSuccess) e = stash(false);
    const int old = b->waves;
    const int64_t old_steps = b->steps;
    double best = 1e300;
    int best_w = old;
    hipEvent_t ea = nullptr, eb = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&ea);
    if (e == hipSuccess) e = hipEventCreate(&eb);
    for (int k = 0; cand[k] > 0 && e == hipSuccess; ++k) {
        b->waves = cand[k] * b->cus;
        b->drop_graphs();
        int rc = mh_batch_run(b, 1, 0.4, 1e30, 0, 1);
        if (rc == MH_OK) e = hipEventRecord(ea, b->stream);
        if (rc == MH_OK && e == hipSuccess) rc = mh_batch_run(b, steps, 0.4, 1e30, 0, 1);
        if (rc != MH_OK) {
            e = hipErrorUnknown;
            break;
        }
        e = hipEventRecord(eb, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(eb);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ea, eb);
        if (ms_out) ms_out[k] = ms / steps;
        if (ms / steps < best) {
            best = ms / steps;
            best_w = b->waves;
        }
        b->cur = start;
        if (e == hipSuccess) e = stash(true);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (ea) hipEventDestroy(ea);
    if (eb) hipEventDestroy(eb);
    if (keep_a) hipFree(keep_a);
    if (keep_b) hipFree(keep_b);
    if (keep_t) hipFree(keep_t);
    b->drop_graphs();
    b->cur = start;
    b->steps = old_steps;
    if (e != hipSuccess) {
        b->waves = old;
        return b->hip_fail(e, "mh_batch_autotune");
    }
    b->waves = best_w;
    *best_waves = best_w;
    return MH_OK;
}

int mh_batch_stats(mh_batch* b, mh_stats* out) {
    if (!b || !out) return MH_ERR_INVALID;
    unsigned int hits = 0;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipMemcpy(&hits, b->floor_hits, sizeof(hits), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return b->hip_fail(e, "mh_batch_stats");
    *out = mh_stats{};
    out->floor_hits = hits;
    out->steps = b->steps;
    return MH_OK;
}

}  // extern "C"
