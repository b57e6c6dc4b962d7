// mh_step_noalloc: one step (CFL dt + every stage with its ghost fill) of a
// caller-owned FP64 state, with no context and no device allocation. The
// step clock and the two intermediate stage states live in `scratch`, a
// file-scope device array of fixed size; each device holds its own copy,
// whose address is looked up once per device (hipGetSymbolAddress) and kept.
// Kernels get that address, never the host-side name of the array.
#include <mutex>

#include "ctx.hpp"

namespace {

constexpr size_t kScratchBytes = (size_t)MH_NOALLOC_SCRATCH_BYTES;
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (looked up on first use).
hipError_t scratch_on(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    return hipSuccess;
}

constexpr size_t round_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

}  // namespace

extern "C" {

int mh_step_noalloc(const mh_grid_size* sz, double dx, const mh_boundaries* bdry, int32_t num_dim,
                    const mh_schemes* schemes, double gamma, double max_cfl, double dt, const double* q_in,
                    double* q_out, double* dt_used, void* stream) {
    const auto bad = [](int code, const std::string& m) {
        mh::set_last_error(m);
        return code;
    };
    if (!sz || !bdry || !schemes || !q_in || !q_out) return bad(MH_ERR_INVALID, "mh_step_noalloc: null argument");
    if (q_in == q_out) return bad(MH_ERR_INVALID, "mh_step_noalloc: q_in and q_out must be distinct buffers");
    if (num_dim < 1 || num_dim > 2) return bad(MH_ERR_UNSUPPORTED, "mh_step_noalloc: 1-D and 2-D grids only");
    if (schemes->reconstruction < 0 || schemes->reconstruction > 3 || schemes->slope_limit < 0 ||
        schemes->slope_limit > 2 || schemes->riemann_solver < 0 || schemes->riemann_solver > 2 ||
        schemes->time_stepper < 0 || schemes->time_stepper > 2)
        return bad(MH_ERR_INVALID, "unknown scheme enum value");
    const Schemes sch{(Reconstruction)schemes->reconstruction, (SlopeLimiter)schemes->slope_limit,
                      (RiemannSolver)schemes->riemann_solver, (TimeStepScheme)schemes->time_stepper};
    if (sz->ng < min_ghost_cells(sch.recon)) return bad(MH_ERR_INVALID, "insufficient ghost cells for the scheme");
    const int nx = sz->xc - 2 * sz->ng;
    const int ny = num_dim > 1 ? sz->yc - 2 * sz->ng : 1;
    if (nx < sz->ng || (num_dim > 1 && ny < sz->ng) || sz->zc != 1 || (num_dim == 1 && sz->yc != 1))
        return bad(MH_ERR_INVALID, "grid size inconsistent with num_dim / num_ghost");
    const int32_t bcs[4] = {bdry->xs, bdry->xe, bdry->ys, bdry->ye};
    for (int k = 0; k < 2 * num_dim; ++k) {
        if (!valid_bc(bcs[k])) return bad(MH_ERR_INVALID, "unknown boundary type");
        if (bcs[k] == (int32_t)BoundaryType::UserFn)
            return bad(MH_ERR_UNSUPPORTED, "mh_step_noalloc: userfn boundaries need a context (mh_set_user_bc)");
    }
    if ((bdry->xs == 1) != (bdry->xe == 1) || (num_dim > 1 && (bdry->ys == 1) != (bdry->ye == 1)))
        return bad(MH_ERR_INVALID, "periodic boundaries must be paired");
    if (!(dx > 0.0) || !(gamma > 1.0)) return bad(MH_ERR_INVALID, "dx must be > 0 and gamma > 1");
    if (!(dt > 0.0) && !(max_cfl > 0.0)) return bad(MH_ERR_INVALID, "mh_step_noalloc: need dt > 0 or max_cfl > 0");

    const int64_t ncells = (int64_t)sz->xc * sz->yc;
    const int64_t nvals = ncells * (2 + num_dim);
    const size_t state_bytes = round_up((size_t)nvals * sizeof(double));
    const size_t need = round_up(kClockBytes) + 2 * state_bytes;
    if (need > kScratchBytes)
        return bad(MH_ERR_INVALID, "mh_step_noalloc: grid too large for the scratch array (" + std::to_string(need) +
                                       " bytes needed, " + std::to_string(kScratchBytes) +
                                       " available: at most 2 stage states of " +
                                       std::to_string((kScratchBytes - round_up(kClockBytes)) / 2 / sizeof(double)) +
                                       " values)");

    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    unsigned char* base = nullptr;
    if (e == hipSuccess) e = scratch_on(dev, &base);
    if (e != hipSuccess) return bad(MH_ERR_HIP, std::string("mh_step_noalloc: ") + hipGetErrorString(e));
    Clock* clock = reinterpret_cast<Clock*>(base);
    double* stage_buf[2] = {reinterpret_cast<double*>(base + round_up(kClockBytes)),
                            reinterpret_cast<double*>(base + round_up(kClockBytes) + state_bytes)};
    const hipStream_t st = static_cast<hipStream_t>(stream);

    mh_boundaries b = *bdry;
    if (num_dim == 1) b.ys = b.ye = (int32_t)BoundaryType::Periodic;  // unused
    GhostGrid gg{};
    gg.xc = sz->xc;
    gg.yc = sz->yc;
    gg.ng = sz->ng;
    gg.bc[0] = b.xs;
    gg.bc[1] = b.xe;
    gg.bc[2] = b.ys;
    gg.bc[3] = b.ye;
    std::memcpy(gg.bconst[0], b.xs_const, sizeof(double) * 5);
    std::memcpy(gg.bconst[1], b.xe_const, sizeof(double) * 5);
    std::memcpy(gg.bconst[2], b.ys_const, sizeof(double) * 5);
    std::memcpy(gg.bconst[3], b.ye_const, sizeof(double) * 5);
    gg.gamma = gamma;
    gg.reduce_dt = 0;

    // the clock: fields written, CFL slots cleared by a kernel (no memset)
    Clock h{};
    h.max_time = 1e300;
    h.delta = 1e300;
    h.max_cfl = max_cfl;
    h.dx = dx;
    e = hipMemcpyAsync(clock, &h, sizeof(Clock), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_clock_reset_max(clock, st);
    if (e == hipSuccess) {
        if (dt > 0.0) {
            e = launch_clock_set_dt(clock, dt, st);
        } else {  // dt_reducer over Q^n, ghosts included (Hydro.cpp:208-229)
            e = launch_cfl_f64(num_dim, FormF64::Planes{const_cast<double*>(q_in)}, (int)ncells, gamma,
                               amax_slots(clock), st);
            if (e == hipSuccess) e = launch_clock_finalize(clock, 0, st);
        }
    }

    // the stages: Q^n stays in q_in; intermediate states rotate through the
    // two scratch states; the last stage writes q_out
    const int strips = (nx + kStripOut - 1) / kStripOut;
    const int nspans = march_spans(0, strips, ny, 1);
    const int ns = num_stages(sch.stepper);
    const double* in = q_in;
    for (int s = 0; s < ns && e == hipSuccess; ++s) {
        double* out = (s == ns - 1) ? q_out : stage_buf[s & 1];
        const int kind = stage_kind(sch.stepper, s);
        FormF64::Ptrs p{FormF64::Planes{const_cast<double*>(in)}, FormF64::Planes{const_cast<double*>(q_in)},
                        FormF64::Planes{out}, &clock->dt, dx, gamma, amax_slots(clock), nullptr};
        if (num_dim == 2) {
            StageGrid g{sz->xc, sz->yc, sz->ng, nx, ny, strips, nspans, kind, 0};
            e = launch_stage2d_f64(sch, g, p, st);
        } else {
            e = launch_stage1d_f64(sch, kind, 0, sz->xc, sz->ng, p, st);
        }
        if (e == hipSuccess)
            e = launch_ghost_f64(num_dim, gg, FormF64::Planes{const_cast<double*>(in)}, FormF64::Planes{out}, nullptr,
                                 st);
        in = out;
    }
    double used = 0.0;
    if (e == hipSuccess) e = hipMemcpyAsync(&used, &clock->dt, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return bad(MH_ERR_HIP, std::string("mh_step_noalloc: ") + hipGetErrorString(e));
    if (dt_used) *dt_used = used;
    return MH_OK;
}

}  // extern "C"
