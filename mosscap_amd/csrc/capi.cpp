// C ABI of the hydro library (include/mosscap_hip.h): context, buffers,
// stage sequencing, device step loop and HIP-graph replay.
//
// Buffer plan: three state buffers rotate through a step, so a stage never
// writes the buffer it reads (a fused stage reads its neighbours' halo
// cells) and the reference's Q_old copy (TimeStepping.cpp:52, 129, 226)
// disappears: Q^n stays in its buffer for the whole step.
#include "ctx.hpp"

namespace {

thread_local std::string g_create_error;

}  // namespace

void mh::set_last_error(const std::string& msg) { g_create_error = msg; }

extern "C" {

const char* mh_last_error(const mh_ctx* ctx) {
    if (ctx == nullptr) return g_create_error.c_str();
    return ctx->err.c_str();
}

int mh_create(const mh_grid_size* sz, double dx, const mh_grid_loc* loc, const mh_boundaries* bdry,
              int32_t num_dim, int32_t num_tracers, const mh_schemes* schemes, double gamma,
              const mh_options* opts, mh_ctx** out) {
    g_create_error.clear();
    auto bad = [](int code, const std::string& m) {
        g_create_error = m;
        return code;
    };
    if (!sz || !bdry || !schemes || !out) return bad(MH_ERR_INVALID, "mh_create: null argument");
    *out = nullptr;
    if (num_dim < 1 || num_dim > 3) {
        return bad(MH_ERR_INVALID, "simulation.num_dim = " + std::to_string(num_dim) + ", must be in range [1, 3]");
    }
    if (num_tracers < 0 || num_tracers > 16) return bad(MH_ERR_INVALID, "num_tracers must be in [0, 16]");
    if (num_tracers > 0 && num_dim != 2)
        return bad(MH_ERR_UNSUPPORTED, "tracer fields (simulation.n_extra_fields) are built for 2-D grids only");
    if (schemes->reconstruction < 0 || schemes->reconstruction > 3 || schemes->slope_limit < 0 ||
        schemes->slope_limit > 2 || schemes->riemann_solver < 0 || schemes->riemann_solver > 2 ||
        schemes->time_stepper < 0 || schemes->time_stepper > 2) {
        return bad(MH_ERR_INVALID, "unknown scheme enum value");
    }
    const Reconstruction recon = (Reconstruction)schemes->reconstruction;
    const int min_ghost = min_ghost_cells(recon);
    if (sz->ng < min_ghost) {
        static const char* names[] = {"fog", "muscl", "ppm", "weno5z"};
        return bad(MH_ERR_INVALID, std::string("Insufficient ghost cells for reconstruction scheme ") +
                                       names[schemes->reconstruction] + ", need minimum of " +
                                       std::to_string(min_ghost) + ".");
    }
    const int nx = sz->xc - 2 * sz->ng;
    const int ny = num_dim > 1 ? sz->yc - 2 * sz->ng : 1;
    const int nz = num_dim > 2 ? sz->zc - 2 * sz->ng : 1;
    if (nx < sz->ng || ny < (num_dim > 1 ? sz->ng : 1) || nz < (num_dim > 2 ? sz->ng : 1) ||
        (num_dim < 3 && sz->zc != 1) || (num_dim == 1 && sz->yc != 1)) {
        return bad(MH_ERR_INVALID, "grid size inconsistent with num_dim / num_ghost (need n >= num_ghost)");
    }
    const int32_t bcs[6] = {bdry->xs, bdry->xe, bdry->ys, bdry->ye, bdry->zs, bdry->ze};
    for (int k = 0; k < 2 * num_dim; ++k) {
        if (!valid_bc(bcs[k])) return bad(MH_ERR_INVALID, "unknown boundary type");
        if (num_dim == 3 && bcs[k] == (int32_t)BoundaryType::UserFn)
            return bad(MH_ERR_UNSUPPORTED, "userfn boundaries on 3-D grids");
    }
    if (bdry->xs == 1 ? bdry->xe != 1 : bdry->xe == 1) {
        return bad(MH_ERR_INVALID, "periodic boundaries must be paired (xs and xe)");
    }
    if (num_dim > 1 && (bdry->ys == 1 ? bdry->ye != 1 : bdry->ye == 1)) {
        return bad(MH_ERR_INVALID, "periodic boundaries must be paired (ys and ye)");
    }
    if (num_dim > 2 && (bdry->zs == 1 ? bdry->ze != 1 : bdry->ze == 1)) {
        return bad(MH_ERR_INVALID, "periodic boundaries must be paired (zs and ze)");
    }
    if (!(dx > 0.0) || !(gamma > 1.0)) return bad(MH_ERR_INVALID, "dx must be > 0 and gamma > 1");
    if (num_tracers > 0) {
        for (int k = 0; k < 4; ++k)
            if (bcs[k] == (int32_t)BoundaryType::Constant)
                // the reference's constant vector holds the 5 hydro values only (State.hpp:93-98)
                return bad(MH_ERR_UNSUPPORTED, "constant boundaries with tracer fields");
    }
    const int64_t ncells = (int64_t)sz->xc * sz->yc * sz->zc;
    const int nvar = 2 + num_dim + num_tracers;
    if (ncells * nvar >= (int64_t)1 << 31) return bad(MH_ERR_INVALID, "state larger than 2^31 values");

    mh_ctx* c = new mh_ctx();
    c->nd = num_dim;
    c->nvar = nvar;
    c->ntr = num_tracers;
    c->sz = *sz;
    c->dx = dx;
    if (loc) c->loc = *loc;
    c->bdry = *bdry;
    if (num_dim == 1) c->bdry.ys = c->bdry.ye = (int32_t)BoundaryType::Periodic;  // unused
    if (num_dim < 3) c->bdry.zs = c->bdry.ze = (int32_t)BoundaryType::Periodic;  // unused
    c->sch = Schemes{recon, (SlopeLimiter)schemes->slope_limit, (RiemannSolver)schemes->riemann_solver,
                     (TimeStepScheme)schemes->time_stepper};
    c->gamma = gamma;
    if (opts) c->opt = *opts;
    if (c->opt.precision_form != MH_FORM_FP64 && c->opt.precision_form != MH_FORM_FP32C) {
        delete c;
        return bad(MH_ERR_INVALID, "unknown precision_form");
    }
    c->ncells = ncells;
    c->nvals = ncells * nvar;
    c->strips = (nx + kStripOut - 1) / kStripOut;
    c->nspans = march_spans(c->opt.span_rows, c->strips, ny, nz);

    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    const size_t bytes64 = (size_t)c->nvals * sizeof(double);
    const size_t bytes32 = (size_t)c->nvals * sizeof(float);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) {
        if (c->fp32c()) {
            e = hipMalloc(&c->qhi[k], bytes32);
            if (e == hipSuccess) e = hipMalloc(&c->qlo[k], bytes32);
        } else {
            e = hipMalloc(&c->q64[k], bytes64);
        }
    }
    if (e == hipSuccess && c->fp32c()) e = hipMalloc(&c->staging, bytes64);
    if (e == hipSuccess && num_tracers > 0) {
        const size_t n = (size_t)ncells;
        if (c->fp32c()) {
            e = hipMalloc(&c->frx32, n * sizeof(float));
            if (e == hipSuccess) e = hipMalloc(&c->fry32, n * sizeof(float));
            if (e == hipSuccess) e = hipMalloc(&c->wtr32, n * num_tracers * sizeof(float));
        } else {
            e = hipMalloc(&c->frx64, n * sizeof(double));
            if (e == hipSuccess) e = hipMalloc(&c->fry64, n * sizeof(double));
            if (e == hipSuccess) e = hipMalloc(&c->wtr64, n * num_tracers * sizeof(double));
        }
    }
    if (e == hipSuccess && num_dim == 3) {
        const size_t n = (size_t)ncells * 5;
        if (c->fp32c()) e = hipMalloc(&c->dz32, n * sizeof(float));
        else e = hipMalloc(&c->dz64, n * sizeof(double));
    }
    if (e == hipSuccess) e = hipMalloc(&c->clock, kClockBytes);
    if (e == hipSuccess) e = hipMalloc(&c->floor_hits, sizeof(unsigned int));
    if (e == hipSuccess) e = hipMemsetAsync(c->clock, 0, kClockBytes, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->floor_hits, 0, sizeof(unsigned int), c->stream);
    // defined contents everywhere (ghosts of the rotating buffers included)
    for (int k = 0; k < 3 && e == hipSuccess; ++k) {
        if (c->fp32c()) {
            e = hipMemsetAsync(c->qhi[k], 0, bytes32, c->stream);
            if (e == hipSuccess) e = hipMemsetAsync(c->qlo[k], 0, bytes32, c->stream);
        } else {
            e = hipMemsetAsync(c->q64[k], 0, bytes64, c->stream);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        g_create_error = std::string("mh_create: ") + hipGetErrorString(e);
        c->free_all();
        delete c;
        return MH_ERR_HIP;
    }
    *out = c;
    return MH_OK;
}

void mh_destroy(mh_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    ctx->free_all();
    delete ctx;
}

int64_t mh_num_values(const mh_ctx* ctx) { return ctx ? ctx->nvals : 0; }

void* mh_stream(mh_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int mh_set_user_bc(mh_ctx* ctx, mh_user_bc_fn fn, void* user) {
    if (!ctx) return MH_ERR_INVALID;
    if (ctx->next_stage != 0) return ctx->fail(MH_ERR_INVALID, "a step is in progress (mh_stage)");
    ctx->user_fn = fn;
    ctx->user_data = fn ? user : nullptr;
    ctx->amax_valid = false;
    return MH_OK;
}

int mh_set_gravity(mh_ctx* ctx, int32_t enable, double gx, double gy, double gz) {
    if (!ctx) return MH_ERR_INVALID;
    if (ctx->next_stage != 0) return ctx->fail(MH_ERR_INVALID, "a step is in progress (mh_stage)");
    if (enable && ((ctx->nd < 3 && gz != 0.0) || (ctx->nd < 2 && gy != 0.0)))
        return ctx->fail(MH_ERR_INVALID, "gravity components beyond the grid's dimensions must be 0");
    ctx->gravity = enable ? 1 : 0;
    ctx->grav[0] = enable ? gx : 0.0;
    ctx->grav[1] = enable ? gy : 0.0;
    ctx->grav[2] = enable ? gz : 0.0;
    ctx->drop_graphs();
    return MH_OK;
}

int mh_eos_fields(mh_ctx* ctx, int32_t which, double avg_mass, int32_t ionisation_e, double* y, double* T) {
    if (!ctx || !y || !T || (which != 0 && which != 1) || !(avg_mass > 0.0)) return MH_ERR_INVALID;
    hipError_t e = hipSetDevice(ctx->device);
    const size_t n = (size_t)ctx->ncells;
    if (e == hipSuccess && !ctx->eos_y) e = hipMalloc(&ctx->eos_y, n * sizeof(double));
    if (e == hipSuccess && !ctx->eos_T) e = hipMalloc(&ctx->eos_T, n * sizeof(double));
    if (e == hipSuccess && !ctx->eos_flag) e = hipMalloc(&ctx->eos_flag, sizeof(int));
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_eos_fields");
    const int k = (which == 1 && ctx->last_in >= 0) ? ctx->last_in : ctx->cur;
    const LteParams prm{ctx->gamma, 1.0 / avg_mass, ionisation_e ? 1 : 0};
    int flag = 0;
    // Steps with ghosts read from their sources leave the ghosts of the
    // intermediate stage inputs unfilled; the fields cover every cell, so
    // fill them here (plain copies of interior cells, the values those
    // stages read).
    e = (k != ctx->cur && ctx->virtual_ghosts()) ? ctx->ghost(k, k, 0) : hipSuccess;
    if (e == hipSuccess) e = hipMemsetAsync(ctx->eos_flag, 0, sizeof(int), ctx->stream);
    if (e == hipSuccess) {
        e = ctx->fp32c() ? launch_lte_eos_f32c(ctx->nd, ctx->p32(k), (int)ctx->ncells, prm, ctx->eos_y, ctx->eos_T,
                                               ctx->eos_flag, ctx->stream)
                         : launch_lte_eos_f64(ctx->nd, ctx->p64(k), (int)ctx->ncells, prm, ctx->eos_y, ctx->eos_T,
                                              ctx->eos_flag, ctx->stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, ctx->eos_y, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(T, ctx->eos_T, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, ctx->eos_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_eos_fields");
    if (flag) return ctx->fail(MH_ERR_INVALID, "Temperature bounds flipped!");
    return MH_OK;
}

int mh_device_state(mh_ctx* ctx, mh_device_view* out) {
    if (!ctx || !out) return MH_ERR_INVALID;
    const int k = ctx->view_buf >= 0 ? ctx->view_buf : ctx->cur;
    *out = mh_device_view{};
    out->precision_form = ctx->opt.precision_form;
    out->num_values = ctx->nvals;
    if (ctx->fp32c()) {
        out->q_hi = ctx->qhi[k];
        out->q_lo = ctx->qlo[k];
    } else {
        out->q = ctx->q64[k];
    }
    return MH_OK;
}

static int upload_impl(mh_ctx* c, const double* q, hipMemcpyKind kind) {
    if (!c || !q) return MH_ERR_INVALID;
    hipError_t e = hipSetDevice(c->device);
    const size_t bytes = (size_t)c->nvals * sizeof(double);
    if (c->fp32c()) {
        if (e == hipSuccess) e = hipMemcpyAsync(c->staging, q, bytes, kind, c->stream);
        if (e == hipSuccess) e = launch_split_f64_to_f32c(c->staging, c->p32(c->cur), (int)c->nvals, c->stream);
    } else {
        if (e == hipSuccess) e = hipMemcpyAsync(c->q64[c->cur], q, bytes, kind, c->stream);
    }
    c->amax_valid = false;
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "mh_upload");
    c->next_stage = 0;
    c->last_in = -1;
    return MH_OK;
}

static int download_impl(mh_ctx* c, double* q, hipMemcpyKind kind) {
    if (!c || !q) return MH_ERR_INVALID;
    hipError_t e = hipSetDevice(c->device);
    const size_t bytes = (size_t)c->nvals * sizeof(double);
    if (c->fp32c()) {
        if (e == hipSuccess) e = launch_join_f32c_to_f64(c->p32(c->cur), c->staging, (int)c->nvals, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(q, c->staging, bytes, kind, c->stream);
    } else {
        if (e == hipSuccess) e = hipMemcpyAsync(q, c->q64[c->cur], bytes, kind, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "mh_download");
    return MH_OK;
}

int mh_upload(mh_ctx* ctx, const double* q_host) { return upload_impl(ctx, q_host, hipMemcpyHostToDevice); }
int mh_download(mh_ctx* ctx, double* q_host) { return download_impl(ctx, q_host, hipMemcpyDeviceToHost); }
int mh_upload_device(mh_ctx* ctx, const double* q_dev) { return upload_impl(ctx, q_dev, hipMemcpyDeviceToDevice); }
int mh_download_device(mh_ctx* ctx, double* q_dev) { return download_impl(ctx, q_dev, hipMemcpyDeviceToDevice); }

int mh_fill_bcs(mh_ctx* ctx) {
    if (!ctx) return MH_ERR_INVALID;
    hipError_t e = hipSetDevice(ctx->device);
    ctx->amax_valid = false;  // a hook may change ghost values; state written through mh_device_state
    if (e == hipSuccess) e = ctx->ghost(ctx->cur, ctx->cur, 0);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? MH_OK : ctx->hip_fail(e, "mh_fill_bcs");
}

// Writes the clock's fields; clears the CFL slots unless `keep_max`.
static hipError_t write_clock(mh_ctx* c, double max_cfl, double time, double max_time, double prev, double delta,
                              bool keep_max = false) {
    Clock h{};
    h.time = time;
    h.dt = 0.0;
    h.max_time = max_time;
    h.prev_output_time = prev;
    h.delta = delta;
    h.max_cfl = max_cfl;
    h.dx = c->dx;
    h.step = c->steps;
    const hipError_t e = hipMemcpyAsync(c->clock, &h, sizeof(Clock), hipMemcpyHostToDevice, c->stream);
    return (e == hipSuccess && !keep_max) ? launch_clock_reset_max(c->clock, c->stream) : e;
}

// The tabulated EOS's abort condition (TabulatedLteH.hpp:140-142) after a
// synchronised step: a cell whose internal energy went negative.
static int tab_check(mh_ctx* ctx) {
    if (!ctx->tab_on) return MH_OK;
    int neg = 0;
    const hipError_t e = hipMemcpy(&neg, ctx->tab_neg, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return ctx->hip_fail(e, "tabulated EOS");
    return neg ? ctx->fail(MH_ERR_INVALID, "eint negative") : MH_OK;
}

int mh_set_tabulated_eos(mh_ctx* ctx, const mh_lte_table* t) {
    if (!ctx) return MH_ERR_INVALID;
    if (ctx->next_stage != 0) return ctx->fail(MH_ERR_INVALID, "a step is in progress (mh_stage)");
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    for (double** p : {&ctx->tab_y, &ctx->tab_lT, &ctx->tab_yspace, &ctx->tab_Tspace}) {
        if (*p) hipFree(*p);
        *p = nullptr;
    }
    if (ctx->tab_neg) hipFree(ctx->tab_neg);
    ctx->tab_neg = nullptr;
    ctx->tab_on = 0;
    ctx->amax_valid = false;
    ctx->drop_graphs();
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_set_tabulated_eos");
    if (!t) return MH_OK;
    if (t->n_eint_rho < 1 || t->n_rho < 1 || !t->y_table || !t->log_T_table || !(t->log_eint_rho_step > 0.0) ||
        !(t->log_rho_step > 0.0) || (int64_t)t->n_eint_rho * t->n_rho > (int64_t)1 << 28)
        return ctx->fail(MH_ERR_INVALID, "mh_set_tabulated_eos: bad table axes or null tables");
    const size_t m = (size_t)t->n_eint_rho * t->n_rho, n = (size_t)ctx->ncells;
    e = hipMalloc(&ctx->tab_y, m * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->tab_lT, m * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->tab_yspace, n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->tab_Tspace, n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->tab_neg, sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(ctx->tab_y, t->y_table, m * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ctx->tab_lT, t->log_T_table, m * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {  // y and T start at -1 (Eos.cpp:80)
        const std::vector<double> minus_one(n, -1.0);
        e = hipMemcpy(ctx->tab_yspace, minus_one.data(), n * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(ctx->tab_Tspace, minus_one.data(), n * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemset(ctx->tab_neg, 0, sizeof(int));
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_set_tabulated_eos");
    ctx->tab = TabParams{t->log_eint_rho_start, t->log_eint_rho_step, t->n_eint_rho, t->log_rho_start, t->log_rho_step,
                         t->n_rho};
    ctx->tab_on = 1;
    return MH_OK;
}

int mh_tabulated_eos_update(mh_ctx* ctx) {
    if (!ctx) return MH_ERR_INVALID;
    if (!ctx->tab_on) return ctx->fail(MH_ERR_INVALID, "no tabulated EOS is installed");
    if (ctx->next_stage != 0) return ctx->fail(MH_ERR_INVALID, "a step is in progress (mh_stage)");
    hipError_t e = hipSetDevice(ctx->device);
    ctx->amax_valid = false;
    if (e == hipSuccess) e = ctx->tab_update(ctx->cur, ctx->cur);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_tabulated_eos_update");
    return tab_check(ctx);
}

int mh_tabulated_eos_fields(mh_ctx* ctx, double* y, double* T) {
    if (!ctx || !y || !T) return MH_ERR_INVALID;
    if (!ctx->tab_on) return ctx->fail(MH_ERR_INVALID, "no tabulated EOS is installed");
    const size_t n = (size_t)ctx->ncells;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(y, ctx->tab_yspace, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(T, ctx->tab_Tspace, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_tabulated_eos_fields");
    return tab_check(ctx);
}

int mh_compute_dt(mh_ctx* ctx, double max_cfl, double time, double max_time, double prev_output_time,
                  double delta, double* dt_out) {
    if (!ctx || !dt_out) return MH_ERR_INVALID;
    hipError_t e = hipSetDevice(ctx->device);
    ctx->amax_valid = false;
    if (e == hipSuccess) e = write_clock(ctx, max_cfl, time, max_time, prev_output_time, delta);
    if (e == hipSuccess) e = ctx->cfl_all(ctx->cur);
    if (e == hipSuccess) e = launch_clock_finalize(ctx->clock, 0, ctx->stream);
    Clock h{};
    if (e == hipSuccess) e = hipMemcpyAsync(&h, ctx->clock, sizeof(Clock), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_compute_dt");
    *dt_out = h.dt;
    return MH_OK;
}

int mh_stage(mh_ctx* ctx, int32_t stage, double dt) {
    if (!ctx) return MH_ERR_INVALID;
    const int ns = num_stages(ctx->sch.stepper);
    if (stage < 0 || stage >= ns) return ctx->fail(MH_ERR_INVALID, "stage out of range for the time stepper");
    if (stage != ctx->next_stage) return ctx->fail(MH_ERR_INVALID, "stages must run in order starting at 0");
    hipError_t e = hipSetDevice(ctx->device);
    ctx->amax_valid = false;
    if (stage == 0) ctx->qn = ctx->cur;
    const int in = ctx->cur;
    const int out = (in == ctx->qn) ? (ctx->qn + 1) % 3 : 3 - in - ctx->qn;
    if (e == hipSuccess) e = launch_clock_set_dt(ctx->clock, dt, ctx->stream);
    if (e == hipSuccess) e = ctx->stage(in, ctx->qn, out, stage_kind(ctx->sch.stepper, stage), 0);
    ctx->last_in = in;
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_stage");
    if (const int rc = tab_check(ctx); rc != MH_OK) return rc;
    ctx->cur = out;
    ctx->next_stage = (stage + 1) % ns;
    if (ctx->next_stage == 0) {
        ctx->steps += 1;
        ctx->last_dt = dt;
    }
    return MH_OK;
}

int mh_time_step(mh_ctx* ctx, double dt) {
    if (!ctx) return MH_ERR_INVALID;
    if (ctx->next_stage != 0) return ctx->fail(MH_ERR_INVALID, "a step is already in progress (mh_stage)");
    hipError_t e = hipSetDevice(ctx->device);
    ctx->amax_valid = false;
    if (e == hipSuccess) e = launch_clock_set_dt(ctx->clock, dt, ctx->stream);
    int nc = ctx->cur;
    if (e == hipSuccess) nc = ctx->enqueue_step(ctx->cur, false, e);
    if (e == hipSuccess) e = ctx->fill_final(nc);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_time_step");
    ctx->cur = nc;
    if (const int rc = tab_check(ctx); rc != MH_OK) return rc;
    ctx->steps += 1;
    ctx->last_dt = dt;
    return MH_OK;
}

int mh_run(mh_ctx* ctx, int64_t nsteps, double max_cfl, double* time, double max_time, double* prev_output_time,
           double delta, int32_t use_graph) {
    if (!ctx || !time || !prev_output_time || nsteps < 0) return MH_ERR_INVALID;
    if (ctx->next_stage != 0) return ctx->fail(MH_ERR_INVALID, "a step is already in progress (mh_stage)");
    hipError_t e = hipSetDevice(ctx->device);
    // Steady state: the previous mh_run's last stage left the max signal
    // speed of the current state in the CFL slots, so only the clock's
    // fields are written; otherwise a CFL pass over the state fills them.
    const bool keep_max = ctx->amax_valid;
    ctx->amax_valid = false;
    if (e == hipSuccess) e = write_clock(ctx, max_cfl, *time, max_time, *prev_output_time, delta, keep_max);
    if (e == hipSuccess && !keep_max) e = ctx->cfl_all(ctx->cur);
    for (int64_t s = 0; s < nsteps && e == hipSuccess; ++s) {
        const int k = ctx->cur;
        if (use_graph && !ctx->user_fn) {
            if (!ctx->graph[k]) {
                hipGraph_t g = nullptr;
                e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) break;
                hipError_t ce = hipSuccess;
                if (ctx->clock_in_stages()) {
                    ctx->enqueue_step(k, true, ce, true);
                } else {
                    ce = launch_clock_finalize(ctx->clock, 1, ctx->stream);
                    if (ce == hipSuccess) ctx->enqueue_step(k, true, ce);
                }
                e = hipStreamEndCapture(ctx->stream, &g);
                if (e == hipSuccess) e = ce;
                if (e == hipSuccess) e = hipGraphInstantiate(&ctx->graph[k], g, nullptr, nullptr, 0);
                if (g) hipGraphDestroy(g);
                if (e != hipSuccess) break;
            }
            e = hipGraphLaunch(ctx->graph[k], ctx->stream);
            {  // the buffers the captured step uses, as enqueue_step assigns them
                int in = k;
                for (int st = 0; st < num_stages(ctx->sch.stepper); ++st) {
                    ctx->last_in = in;
                    in = (in == k) ? (k + 1) % 3 : 3 - in - k;
                }
                ctx->cur = in;
            }
        } else {
            e = ctx->finalize(1);
            if (e == hipSuccess) ctx->cur = ctx->enqueue_step(k, true, e);
        }
        ctx->steps += 1;
    }
    if (e == hipSuccess && nsteps > 0) e = ctx->fill_final(ctx->cur);
    Clock h{};
    if (e == hipSuccess) e = hipMemcpyAsync(&h, ctx->clock, sizeof(Clock), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_run");
    if (const int rc = tab_check(ctx); rc != MH_OK) return rc;
    ctx->amax_valid = true;
    *time = h.time;
    *prev_output_time = h.prev_output_time;
    ctx->time = h.time;
    ctx->prev_output_time = h.prev_output_time;
    ctx->last_dt = h.dt;
    return MH_OK;
}

int mh_run_timed(mh_ctx* ctx, int64_t nsteps, double max_cfl, double* time, double max_time,
                 double* prev_output_time, double delta, mh_timing* out) {
    if (!ctx || !out) return MH_ERR_INVALID;
    std::vector<mh_ctx::Ev> log;
    log.reserve((size_t)nsteps * 10 + 4);
    ctx->evlog = &log;
    ctx->ev_next = 0;
    const int rc = mh_run(ctx, nsteps, max_cfl, time, max_time, prev_output_time, delta, 0);
    ctx->evlog = nullptr;
    if (rc != MH_OK) return rc;
    *out = mh_timing{};
    hipError_t e = hipSuccess;
    for (const auto& ev : log) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, ev.a, ev.b);
        if (e != hipSuccess) break;
        if (ev.kind == 0) {
            out->stage_ms += ms;
            out->stage_launches += 1;
        } else if (ev.kind == 1) {
            out->ghost_ms += ms;
            out->ghost_launches += 1;
        } else {
            out->clock_ms += ms;
            out->clock_launches += 1;
        }
    }
    if (e == hipSuccess && !log.empty()) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, log.front().a, log.back().b);
        out->step_ms = ms;
    }
    return e == hipSuccess ? MH_OK : ctx->hip_fail(e, "mh_run_timed");
}

int mh_autotune(mh_ctx* ctx, int32_t steps, const int32_t* candidates, int32_t* best_span, double* ms_out) {
    if (!ctx || !best_span || steps < 1) return MH_ERR_INVALID;
    if (ctx->next_stage != 0) return ctx->fail(MH_ERR_INVALID, "a step is in progress (mh_stage)");
    // rows per span; -1: the default layout (march_spans with rows <= 0)
    static const int32_t kDefault[] = {4, 8, 12, 16, 18, 20, 24, 28, 32, 48, 64, -1, 0};
    const int32_t* cand = candidates ? candidates : kDefault;
    hipError_t e = hipSetDevice(ctx->device);
    // keep the state in a scratch copy and restore it after each trial
    void* keep_a = nullptr;
    void* keep_b = nullptr;
    const size_t bytes = ctx->fp32c() ? ctx->nvals * sizeof(float) : ctx->nvals * sizeof(double);
    if (e == hipSuccess) e = hipMalloc(&keep_a, bytes);
    if (e == hipSuccess && ctx->fp32c()) e = hipMalloc(&keep_b, bytes);
    const int start = ctx->cur;
    auto stash = [&](bool restore) -> hipError_t {
        void* d0 = ctx->fp32c() ? (void*)ctx->qhi[ctx->cur] : (void*)ctx->q64[ctx->cur];
        if (restore) {
            hipError_t r = hipMemcpyAsync(d0, keep_a, bytes, hipMemcpyDeviceToDevice, ctx->stream);
            if (r == hipSuccess && ctx->fp32c())
                r = hipMemcpyAsync(ctx->qlo[ctx->cur], keep_b, bytes, hipMemcpyDeviceToDevice, ctx->stream);
            return r;
        }
        hipError_t r = hipMemcpyAsync(keep_a, d0, bytes, hipMemcpyDeviceToDevice, ctx->stream);
        if (r == hipSuccess && ctx->fp32c())
            r = hipMemcpyAsync(keep_b, ctx->qlo[ctx->cur], bytes, hipMemcpyDeviceToDevice, ctx->stream);
        return r;
    };
    if (e == hipSuccess) e = stash(false);
    // what the trial steps change besides the state: the tabulated EOS's fields,
    // the floor-hit count, the stage-input record
    std::vector<double> keep_y, keep_T;
    unsigned int keep_hits = 0;
    if (e == hipSuccess && ctx->tab_on) {
        keep_y.resize((size_t)ctx->ncells);
        keep_T.resize((size_t)ctx->ncells);
        e = hipMemcpyAsync(keep_y.data(), ctx->tab_yspace, keep_y.size() * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(keep_T.data(), ctx->tab_Tspace, keep_T.size() * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(&keep_hits, ctx->floor_hits, sizeof(keep_hits), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    const int old_spans = ctx->nspans;
    const int64_t old_steps = ctx->steps;
    double best = 1e300;
    int best_h = 0, best_spans = old_spans;
    hipEvent_t a = nullptr, b = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    const int ny = ctx->nd > 1 ? ctx->sz.yc - 2 * ctx->sz.ng : 1;
    const int nz = ctx->nd > 2 ? ctx->sz.zc - 2 * ctx->sz.ng : 1;
    for (int k = 0; cand[k] != 0 && e == hipSuccess; ++k) {
        if (cand[k] > ny && k > 0) continue;
        ctx->nspans = march_spans(cand[k], ctx->strips, ny, nz);
        for (int g = 0; g < 3; ++g) {  // graphs bake the launch shape in
            if (ctx->graph[g]) hipGraphExecDestroy(ctx->graph[g]);
            ctx->graph[g] = nullptr;
        }
        double t = 0.0, p = 0.0;
        int rc = mh_run(ctx, 1, 0.4, &t, 1e30, &p, 1e30, 1);  // warm-up and graph capture
        if (rc == MH_OK) e = hipEventRecord(a, ctx->stream);
        if (rc == MH_OK && e == hipSuccess) rc = mh_run(ctx, steps, 0.4, &t, 1e30, &p, 1e30, 1);
        if (rc != MH_OK) {
            e = hipErrorUnknown;
            break;
        }
        e = hipEventRecord(b, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        const double per = ms / steps;
        if (ms_out) ms_out[k] = per;
        if (per < best) {
            best = per;
            best_h = cand[k];
            best_spans = ctx->nspans;
        }
        ctx->cur = start;
        if (e == hipSuccess) e = stash(true);
        if (e == hipSuccess && ctx->tab_on) {
            e = hipMemcpyAsync(ctx->tab_yspace, keep_y.data(), keep_y.size() * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(ctx->tab_Tspace, keep_T.data(), keep_T.size() * sizeof(double),
                                   hipMemcpyHostToDevice, ctx->stream);
            if (e == hipSuccess) e = hipMemsetAsync(ctx->tab_neg, 0, sizeof(int), ctx->stream);
        }
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->floor_hits, &keep_hits, sizeof(keep_hits), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    // the trial steps overwrote every buffer but `start`: the stage-input
    // record is void, and mh_eos_fields(which = 1) now reads the current state
    ctx->last_in = -1;
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    if (keep_a) hipFree(keep_a);
    if (keep_b) hipFree(keep_b);
    for (int g = 0; g < 3; ++g) {
        if (ctx->graph[g]) hipGraphExecDestroy(ctx->graph[g]);
        ctx->graph[g] = nullptr;
    }
    ctx->steps = old_steps;
    ctx->cur = start;
    ctx->amax_valid = false;  // the slots hold a trial state's maximum
    if (e != hipSuccess) {
        ctx->nspans = old_spans;
        return ctx->hip_fail(e, "mh_autotune");
    }
    ctx->nspans = best_spans;
    *best_span = best_h;
    return MH_OK;
}

int mh_get_stats(mh_ctx* ctx, mh_stats* out) {
    if (!ctx || !out) return MH_ERR_INVALID;
    unsigned int hits = 0;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipMemcpyAsync(&hits, ctx->floor_hits, sizeof(hits), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_get_stats");
    out->floor_hits = hits;
    out->steps = ctx->steps;
    out->last_dt = ctx->last_dt;
    out->time = ctx->time;
    out->prev_output_time = ctx->prev_output_time;
    return MH_OK;
}

// ---- probes -------------------------------------------------------------------
namespace {

struct DevBuf {
    double* p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc(&p, n * sizeof(double)); }
    ~DevBuf() {
        if (p) hipFree(p);
    }
};

thread_local std::string g_probe_error;

int probe_fail(hipError_t e) {
    g_create_error = std::string("probe: ") + hipGetErrorString(e);
    return MH_ERR_HIP;
}

}  // namespace

int mh_probe_reconstruct(int32_t recon, int32_t limiter, int32_t prec, int64_t n, const double* stencils,
                         double* wl, double* wr) {
    if (n < 1 || n > (1 << 28) || !stencils || !wl || !wr || recon < 0 || recon > 3 || limiter < 0 || limiter > 2)
        return MH_ERR_INVALID;
    DevBuf s, l, r;
    hipError_t e = s.alloc(5 * n);
    if (e == hipSuccess) e = l.alloc(n);
    if (e == hipSuccess) e = r.alloc(n);
    if (e == hipSuccess) e = hipMemcpy(s.p, stencils, 5 * n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe_reconstruct(recon, limiter, prec, (int)n, s.p, l.p, r.p, nullptr);
    if (e == hipSuccess) e = hipMemcpy(wl, l.p, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(wr, r.p, n * sizeof(double), hipMemcpyDeviceToHost);
    return e == hipSuccess ? MH_OK : probe_fail(e);
}

int mh_probe_riemann(int32_t solver, int32_t axis, int32_t nd, int32_t prec, double gamma, int64_t n,
                     const double* wl, const double* wr, double* flux) {
    if (n < 1 || n > (1 << 26) || !wl || !wr || !flux || nd < 1 || nd > 3 || axis < 0 || axis >= nd) return MH_ERR_INVALID;
    const size_t m = (size_t)n * (2 + nd);
    DevBuf a, b, f;
    hipError_t e = a.alloc(m);
    if (e == hipSuccess) e = b.alloc(m);
    if (e == hipSuccess) e = f.alloc(m);
    if (e == hipSuccess) e = hipMemcpy(a.p, wl, m * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.p, wr, m * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe_riemann(solver, axis, nd, prec, gamma, (int)n, a.p, b.p, f.p, nullptr);
    if (e == hipSuccess) e = hipMemcpy(flux, f.p, m * sizeof(double), hipMemcpyDeviceToHost);
    return e == hipSuccess ? MH_OK : probe_fail(e);
}

int mh_probe_math(int32_t op, int64_t n, const double* a, const double* b, double* out) {
    if (n < 1 || n > (1 << 28) || !a || !out || op < 0 || op > 3 || ((op == 0 || op == 2) && !b)) return MH_ERR_INVALID;
    DevBuf da, db, dout;
    hipError_t e = da.alloc(n);
    if (e == hipSuccess) e = db.alloc(n);
    if (e == hipSuccess) e = dout.alloc(n);
    if (e == hipSuccess) e = hipMemcpy(da.p, a, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db.p, b ? b : a, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe_math(op, (int)n, da.p, db.p, dout.p, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dout.p, n * sizeof(double), hipMemcpyDeviceToHost);
    return e == hipSuccess ? MH_OK : probe_fail(e);
}

}  // extern "C"
