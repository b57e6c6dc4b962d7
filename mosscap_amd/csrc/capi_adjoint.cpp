// C ABI of the discrete adjoint (mh_adjoint_*): checkpointed forward with
// recorded dt, then the reverse sweep segment by segment (recompute the
// segment's stage states from its checkpoint, then run the stages backwards).
#include <vector>

#include "ctx.hpp"
#include "launch_adjoint.hpp"

namespace {

// Stage k maps (Q_s, Q^n) to a_qn*Q^n + a_qs*Q_s + a_L*L(Q_s) (TimeStepping.cpp stage lambdas).
struct StageCoef {
    double a_qn, a_qs, a_l;
};

StageCoef coef_of(int kind) {
    switch (kind) {
        case kAddFull: return {0.0, 1.0, 1.0};
        case kRk2Final: return {0.5, 0.5, 0.5};
        case kSsp3Second: return {0.75, 0.25, 0.25};
        case kSsp3Third: return {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0};
        case kAddHalf: return {0.0, 1.0, 0.5};
        default: return {2.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    }
}

// a device buffer owned elsewhere
struct DevPtr {
    double* p = nullptr;
};

struct DevVec {
    double* p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc(&p, n * sizeof(double)); }
    ~DevVec() {
        if (p) hipFree(p);
    }
};

}  // namespace

extern "C" {

int mh_adjoint_forward(mh_ctx* ctx, int64_t nsteps, int32_t checkpoint_every, double max_cfl, const double* dts_in,
                       double* dts_out) {
    if (!ctx || nsteps < 1 || checkpoint_every < 1) return MH_ERR_INVALID;
    if (ctx->fp32c() || ctx->nd != 2) return ctx->fail(MH_ERR_UNSUPPORTED, "the adjoint runs on FP64 2-D contexts");
    if (ctx->tab_on) return ctx->fail(MH_ERR_UNSUPPORTED, "no adjoint for the tabulated EOS");
    const int32_t bcs[4] = {ctx->bdry.xs, ctx->bdry.xe, ctx->bdry.ys, ctx->bdry.ye};
    for (int32_t b : bcs) {
        if (b == (int32_t)BoundaryType::SymmetricOutflowDiode)
            return ctx->fail(MH_ERR_UNSUPPORTED, "no adjoint for symmetricoutflowdiode boundaries");
        if (b == (int32_t)BoundaryType::UserFn)
            return ctx->fail(MH_ERR_UNSUPPORTED, "no adjoint for userfn boundaries");
    }
    if (ctx->user_fn) return ctx->fail(MH_ERR_UNSUPPORTED, "no adjoint while a user_bc hook is set");
    if (ctx->ntr > 0) return ctx->fail(MH_ERR_UNSUPPORTED, "no adjoint of tracer fields");
    if (ctx->sz.ng > 8) return ctx->fail(MH_ERR_UNSUPPORTED, "the adjoint supports at most 8 ghost cells");
    if (ctx->next_stage != 0) return ctx->fail(MH_ERR_INVALID, "a step is in progress (mh_stage)");
    hipError_t e = hipSetDevice(ctx->device);
    ctx->amax_valid = false;  // the clock's CFL slots are reused here
    ctx->free_adjoint();
    const int nck = (int)((nsteps + checkpoint_every - 1) / checkpoint_every);
    const size_t bytes = (size_t)ctx->nvals * sizeof(double);
    ctx->ckpt.assign(nck, nullptr);
    for (int k = 0; k < nck && e == hipSuccess; ++k) e = hipMalloc(&ctx->ckpt[k], bytes);
    if (e != hipSuccess) {
        ctx->free_adjoint();
        return ctx->hip_fail(e, "mh_adjoint_forward: checkpoints");
    }
    ctx->adj_every = checkpoint_every;
    ctx->adj_dts.resize(nsteps);
    // a failed forward leaves no record, so a reverse sweep cannot start from it
    const auto failed = [&](int rc) {
        ctx->free_adjoint();
        return rc;
    };
    for (int64_t n = 0; n < nsteps; ++n) {
        if (n % checkpoint_every == 0) {
            e = hipMemcpyAsync(ctx->ckpt[n / checkpoint_every], ctx->q64[ctx->cur], bytes, hipMemcpyDeviceToDevice,
                               ctx->stream);
            if (e != hipSuccess) return failed(ctx->hip_fail(e, "mh_adjoint_forward"));
        }
        double dt = 0.0;
        if (dts_in) {
            dt = dts_in[n];
        } else {
            const int rc = mh_compute_dt(ctx, max_cfl, 0.0, 1e300, 0.0, 1e300, &dt);
            if (rc != MH_OK) return failed(rc);
        }
        ctx->adj_dts[n] = dt;
        const int rc = mh_time_step(ctx, dt);
        if (rc != MH_OK) return failed(rc);
    }
    // the recorded step sizes on the device, for the reverse sweep's recomputation
    e = hipMalloc(&ctx->adj_dts_dev, (size_t)nsteps * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(ctx->adj_dts_dev, ctx->adj_dts.data(), (size_t)nsteps * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return failed(ctx->hip_fail(e, "mh_adjoint_forward"));
    if (dts_out)
        for (int64_t n = 0; n < nsteps; ++n) dts_out[n] = ctx->adj_dts[n];
    return MH_OK;
}

int mh_adjoint_backward(mh_ctx* ctx, const double* lambda_final, double* lambda_initial) {
    if (!ctx || !lambda_final || !lambda_initial) return MH_ERR_INVALID;
    const int64_t nsteps = (int64_t)ctx->adj_dts.size();
    if (nsteps == 0 || ctx->ckpt.empty() || !ctx->adj_dts_dev)
        return ctx->fail(MH_ERR_INVALID, "run mh_adjoint_forward first");
    hipError_t e = hipSetDevice(ctx->device);
    ctx->amax_valid = false;  // the clock's CFL slots are reused here
    const int ns = num_stages(ctx->sch.stepper);
    const int every = ctx->adj_every;
    const size_t nv = (size_t)ctx->nvals;
    AdjGrid g{};
    g.xc = ctx->sz.xc;
    g.yc = ctx->sz.yc;
    g.ng = ctx->sz.ng;
    g.nx = g.xc - 2 * g.ng;
    g.ny = g.yc - 2 * g.ng;
    g.ncells = (int)ctx->ncells;
    g.bc[0] = ctx->bdry.xs;
    g.bc[1] = ctx->bdry.xe;
    g.bc[2] = ctx->bdry.ys;
    g.bc[3] = ctx->bdry.ye;
    g.gamma = ctx->gamma;
    g.gravity = ctx->gravity;
    g.gx = ctx->grav[0];
    g.gy = ctx->grav[1];
    g.virt = ctx->virtual_ghosts() ? 1 : 0;  // the recomputed stage inputs keep unfilled ghosts
    g.xper = ctx->bdry.xs == (int32_t)BoundaryType::Periodic;
    g.yper = ctx->bdry.ys == (int32_t)BoundaryType::Periodic;
    // scratch, allocated by the first reverse sweep and kept by the context:
    // the x sweep's primitive cotangents, three cotangent states, the stage
    // states of one segment ((every) steps x ns stages) and pinned host
    // staging for the cotangents' transfers
    if (e == hipSuccess) e = ctx->adjoint_scratch(4 + (size_t)every * ns);
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_adjoint_backward: scratch");
    DevPtr lamx{ctx->adj_scratch[0]}, acc{ctx->adj_scratch[1]}, cur{ctx->adj_scratch[2]}, nxt{ctx->adj_scratch[3]};
    std::vector<DevPtr> seg((size_t)every * ns);
    for (size_t k = 0; k < seg.size(); ++k) seg[k].p = ctx->adj_scratch[4 + k];
    std::memcpy(ctx->adj_host, lambda_final, nv * sizeof(double));
    e = hipMemcpyAsync(cur.p, ctx->adj_host, nv * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_adjoint_axpby(g, 1.0, cur.p, 0.0, cur.p, ctx->stream);  // clear ghost entries
    // ... of the other cotangent buffer too: the line sweeps of a periodic
    // axis never write its ghost entries, which must read as zero
    if (e == hipSuccess) e = launch_adjoint_axpby(g, 1.0, nxt.p, 0.0, nxt.p, ctx->stream);
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&ev0);
    if (e == hipSuccess) e = hipEventCreate(&ev1);
    if (e == hipSuccess) e = hipEventRecord(ev0, ctx->stream);
    const int nseg = (int)ctx->ckpt.size();
    for (int sgi = nseg - 1; sgi >= 0 && e == hipSuccess; --sgi) {
        const int64_t n0 = (int64_t)sgi * every;
        const int64_t n1 = std::min<int64_t>(n0 + every, nsteps);
        // recompute the segment: seg[(n-n0)*ns + s] holds the input of stage s of step n
        e = hipMemcpyAsync(seg[0].p, ctx->ckpt[sgi], nv * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream);
        for (int64_t n = n0; n < n1 && e == hipSuccess; ++n) {
            const double* qn = seg[(n - n0) * ns].p;
            for (int s = 0; s < ns && e == hipSuccess; ++s) {
                const double* in = seg[(n - n0) * ns + s].p;
                const bool last = (s == ns - 1);
                if (last && n + 1 == n1) break;  // the segment's end state is not needed
                double* out = last ? seg[(n + 1 - n0) * ns].p : seg[(n - n0) * ns + s + 1].p;
                e = ctx->stage_ptr(in, qn, out, stage_kind(ctx->sch.stepper, s), ctx->adj_dts_dev + n);
            }
        }
        // reverse the segment's steps
        // cur holds the cotangent of step n's output; after the stages it holds
        // dJ/dQ^n (stage 0, reversed last, adds the Q^n terms gathered in acc)
        for (int64_t n = n1 - 1; n >= n0 && e == hipSuccess; --n) {
            g.dtdx = ctx->adj_dts[n] / ctx->dx;
            g.dt = ctx->adj_dts[n];
            for (int s = ns - 1; s >= 0 && e == hipSuccess; --s) {
                const StageCoef c = coef_of(stage_kind(ctx->sch.stepper, s));
                const int32_t mode = s == ns - 1 ? 1 : (s == 0 ? 2 : 0);
                const AdjStage st{seg[(n - n0) * ns + s].p, cur.p, lamx.p, nxt.p, acc.p, c.a_qn, c.a_qs, c.a_l,
                                  mode};
                e = launch_adjoint_stage(ctx->sch, g, st, ctx->stream);
                std::swap(cur.p, nxt.p);
            }
        }
    }
    if (e == hipSuccess) e = hipEventRecord(ev1, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->adj_host, cur.p, nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) std::memcpy(lambda_initial, ctx->adj_host, nv * sizeof(double));
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
    ctx->adj_bwd_ms = e == hipSuccess ? (double)ms : -1.0;
    if (ev0) hipEventDestroy(ev0);
    if (ev1) hipEventDestroy(ev1);
    return e == hipSuccess ? MH_OK : ctx->hip_fail(e, "mh_adjoint_backward");
}

int mh_adjoint_tangent(mh_ctx* ctx, const double* dq0, double* dq_final, double* q_final) {
    if (!ctx || !dq0 || !dq_final) return MH_ERR_INVALID;
    const int64_t nsteps = (int64_t)ctx->adj_dts.size();
    if (nsteps == 0 || ctx->ckpt.empty()) return ctx->fail(MH_ERR_INVALID, "run mh_adjoint_forward first");
    const int32_t bcs[4] = {ctx->bdry.xs, ctx->bdry.xe, ctx->bdry.ys, ctx->bdry.ye};
    for (int32_t b : bcs)
        if (b == (int32_t)BoundaryType::Constant)  // its ghosts are fixed values, their derivative zero
            return ctx->fail(MH_ERR_UNSUPPORTED, "no tangent for constant boundaries");
    hipError_t e = hipSetDevice(ctx->device);
    const size_t nv = (size_t)ctx->nvals;
    // values and derivatives of three rotating states
    DevVec val[3], der[3];
    for (int k = 0; k < 3 && e == hipSuccess; ++k) {
        e = val[k].alloc(nv);
        if (e == hipSuccess) e = der[k].alloc(nv);
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "mh_adjoint_tangent: scratch");
    // the remaining boundary types fill ghosts by copies and sign changes of
    // interior cells, so the same fill maps derivatives to ghost derivatives
    const auto fill = [&](double* q) {
        return launch_ghost_f64(2, ctx->ghost_grid(0), FormF64::Planes{q}, FormF64::Planes{q}, nullptr, ctx->stream);
    };
    e = hipMemcpyAsync(val[0].p, ctx->ckpt[0], nv * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(der[0].p, dq0, nv * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = fill(der[0].p);
    TanStage a{};
    a.xc = ctx->sz.xc;
    a.yc = ctx->sz.yc;
    a.ng = ctx->sz.ng;
    a.nx = a.xc - 2 * a.ng;
    a.ny = a.yc - 2 * a.ng;
    a.gamma = ctx->gamma;
    a.dx = ctx->dx;
    a.gravity = ctx->gravity;
    a.gx = ctx->grav[0];
    a.gy = ctx->grav[1];
    const int ns = num_stages(ctx->sch.stepper);
    int k = 0;  // buffer of Q^n
    for (int64_t n = 0; n < nsteps && e == hipSuccess; ++n) {
        a.dt = ctx->adj_dts[n];
        int in = k;
        for (int s = 0; s < ns && e == hipSuccess; ++s) {
            const int out = (in == k) ? (k + 1) % 3 : 3 - in - k;
            a.qv = val[in].p;
            a.qd = der[in].p;
            a.nv = val[k].p;
            a.nd = der[k].p;
            a.ov = val[out].p;
            a.od = der[out].p;
            a.kind = stage_kind(ctx->sch.stepper, s);
            e = launch_tangent_stage(ctx->sch, a, ctx->stream);
            if (e == hipSuccess) e = fill(val[out].p);
            if (e == hipSuccess) e = fill(der[out].p);
            in = out;
        }
        k = in;
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(dq_final, der[k].p, nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && q_final)
        e = hipMemcpyAsync(q_final, val[k].p, nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? MH_OK : ctx->hip_fail(e, "mh_adjoint_tangent");
}

int mh_adjoint_backward_ms(const mh_ctx* ctx, double* ms) {
    if (!ctx || !ms) return MH_ERR_INVALID;
    *ms = ctx->adj_bwd_ms;
    return MH_OK;
}

int mh_probe_vjp(int32_t what, int64_t n, const double* in, const double* cot, double* out, double gamma) {
    if (n < 1 || n > (1 << 24) || !in || !cot || !out || what < 0 || what > 9) return MH_ERR_INVALID;
    const int nin = what < 2 ? 5 : 8, ncot = what < 2 ? 2 : 4;
    DevVec a, b, c;
    hipError_t e = a.alloc(n * nin);
    if (e == hipSuccess) e = b.alloc(n * ncot);
    if (e == hipSuccess) e = c.alloc(n * nin);
    if (e == hipSuccess) e = hipMemcpy(a.p, in, n * nin * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.p, cot, n * ncot * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe_vjp(what, (int)n, a.p, b.p, c.p, gamma, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, c.p, n * nin * sizeof(double), hipMemcpyDeviceToHost);
    return e == hipSuccess ? MH_OK : MH_ERR_HIP;
}

}  // extern "C"
