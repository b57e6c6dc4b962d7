// Ghost fill, CFL reduction, step clock, FP64 <-> compensated conversions and
// the functor probes used by the unit tests. Built with -ffp-contract=off.
#include "../launch.hpp"

namespace mh {

namespace {

constexpr int kBlock = 256;

int blocks_for(long long n) { return (int)((n + kBlock - 1) / kBlock); }

template <class Form>
hipError_t launch_ghost_impl(int nd, const GhostGrid& g, typename Form::Planes in, typename Form::Planes out,
                             unsigned long long* amax, hipStream_t st) {
    if (g.ng < 1 || g.xc < 3 * g.ng || (nd > 1 && g.yc < 3 * g.ng) || (nd == 1 && g.yc != 1)) {
        return hipErrorInvalidValue;
    }
    const long long n = 2LL * g.ng * g.yc + (nd > 1 ? 2LL * g.ng * g.xc : 0LL);
    if (nd == 1) {
        hipLaunchKernelGGL((ghost_fill_kernel<Form, 1>), dim3(blocks_for(n)), dim3(kBlock), 0, st, g, in, out, amax);
    } else if (nd == 2) {
        hipLaunchKernelGGL((ghost_fill_kernel<Form, 2>), dim3(blocks_for(n)), dim3(kBlock), 0, st, g, in, out, amax);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// 3-D ghost fill: the x, y and z passes in order on one buffer.
template <class Form>
hipError_t launch_ghost3d_impl(const Ghost3Grid& g, typename Form::Planes q, unsigned long long* amax,
                               hipStream_t st) {
    if (g.ng < 1 || g.xc < 3 * g.ng || g.yc < 3 * g.ng || g.zc < 3 * g.ng || (g.reduce_dt && !amax))
        return hipErrorInvalidValue;
    const long long nx = 2LL * g.ng * g.yc * g.zc, ny = 2LL * g.ng * g.xc * g.zc, nz = 2LL * g.ng * g.xc * g.yc;
    hipLaunchKernelGGL((ghost_pass3d_kernel<Form, 0>), dim3(blocks_for(nx)), dim3(kBlock), 0, st, g, q, amax);
    hipLaunchKernelGGL((ghost_pass3d_kernel<Form, 1>), dim3(blocks_for(ny)), dim3(kBlock), 0, st, g, q, amax);
    hipLaunchKernelGGL((ghost_pass3d_kernel<Form, 2>), dim3(blocks_for(nz)), dim3(kBlock), 0, st, g, q, amax);
    return hipGetLastError();
}

template <class Form>
hipError_t launch_cfl_impl(int nd, typename Form::Planes q, int ncells, double gamma, unsigned long long* amax,
                           hipStream_t st) {
    if (ncells < 1) return hipErrorInvalidValue;
    const int blocks = blocks_for(ncells) < 2048 ? blocks_for(ncells) : 2048;
    if (nd == 1) {
        hipLaunchKernelGGL((cfl_reduce_kernel<Form, 1>), dim3(blocks), dim3(kBlock), 0, st, q, ncells, gamma, amax);
    } else if (nd == 2) {
        hipLaunchKernelGGL((cfl_reduce_kernel<Form, 2>), dim3(blocks), dim3(kBlock), 0, st, q, ncells, gamma, amax);
    } else if (nd == 3) {
        hipLaunchKernelGGL((cfl_reduce_kernel<Form, 3>), dim3(blocks), dim3(kBlock), 0, st, q, ncells, gamma, amax);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

__global__ void split_kernel(const double* __restrict__ src, float* hi, float* lo, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float2 v = df::from_double(src[i]);
        hi[i] = v.x;
        lo[i] = v.y;
    }
}

__global__ void join_kernel(const float* __restrict__ hi, const float* __restrict__ lo, double* dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double)hi[i] + (double)lo[i];
}

// ---- functor probes ----------------------------------------------------------
template <typename T, Reconstruction R, SlopeLimiter SL>
__global__ void probe_reconstruct_kernel(int n, const double* st5, double* wl, double* wr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T s[5];
    for (int k = 0; k < 5; ++k) s[k] = (T)st5[5 * i + k];
    auto acc = [&](int o) -> T { return s[2 + o]; };
    T l, r;
    reconstruct_stencil<R, SL, T>(acc, l, r);
    wl[i] = (double)l;
    wr[i] = (double)r;
}

template <typename T, RiemannSolver RS, int Axis, int ND>
__global__ void probe_riemann_kernel(double gamma, int n, const double* wl, const double* wr, double* flux) {
    constexpr int NH = 2 + ND;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T a[NH], b[NH], f[NH];
    for (int k = 0; k < NH; ++k) {
        a[k] = (T)wl[NH * i + k];
        b[k] = (T)wr[NH * i + k];
    }
    riemann_flux<RS, Axis, ND>(make_eos<T>(gamma), a, b, f);
    for (int k = 0; k < NH; ++k) flux[NH * i + k] = (double)f[k];
}

__global__ void probe_math_kernel(int op, int n, const double* a, const double* b, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (op == 0) out[i] = Arith<double>::div(a[i], b[i]);
    else if (op == 1) out[i] = Arith<double>::sqrt(a[i]);
    else if (op == 2) out[i] = (double)Arith<float>::div((float)a[i], (float)b[i]);
    else out[i] = (double)Arith<float>::sqrt((float)a[i]);
}

template <typename T>
hipError_t probe_reconstruct_t(int recon, int limiter, int n, const double* s, double* wl, double* wr, hipStream_t st) {
    using RC = Reconstruction;
    using SL = SlopeLimiter;
    const dim3 grid(blocks_for(n)), block(kBlock);
    const SL sl = (SL)limiter;
    switch ((RC)recon) {
        case RC::Fog:
            hipLaunchKernelGGL((probe_reconstruct_kernel<T, RC::Fog, SL::MonotonizedCentral>), grid, block, 0, st, n, s, wl, wr);
            break;
        case RC::Weno5Z:
            hipLaunchKernelGGL((probe_reconstruct_kernel<T, RC::Weno5Z, SL::MonotonizedCentral>), grid, block, 0, st, n, s, wl, wr);
            break;
        case RC::Muscl:
            if (sl == SL::VanLeer) hipLaunchKernelGGL((probe_reconstruct_kernel<T, RC::Muscl, SL::VanLeer>), grid, block, 0, st, n, s, wl, wr);
            else if (sl == SL::Minmod) hipLaunchKernelGGL((probe_reconstruct_kernel<T, RC::Muscl, SL::Minmod>), grid, block, 0, st, n, s, wl, wr);
            else hipLaunchKernelGGL((probe_reconstruct_kernel<T, RC::Muscl, SL::MonotonizedCentral>), grid, block, 0, st, n, s, wl, wr);
            break;
        case RC::Ppm:
            if (sl == SL::VanLeer) hipLaunchKernelGGL((probe_reconstruct_kernel<T, RC::Ppm, SL::VanLeer>), grid, block, 0, st, n, s, wl, wr);
            else if (sl == SL::Minmod) hipLaunchKernelGGL((probe_reconstruct_kernel<T, RC::Ppm, SL::Minmod>), grid, block, 0, st, n, s, wl, wr);
            else hipLaunchKernelGGL((probe_reconstruct_kernel<T, RC::Ppm, SL::MonotonizedCentral>), grid, block, 0, st, n, s, wl, wr);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T, RiemannSolver RS>
hipError_t probe_riemann_rs(int axis, int nd, double gamma, int n, const double* wl, const double* wr, double* f,
                            hipStream_t st) {
    const dim3 grid(blocks_for(n)), block(kBlock);
    if (nd == 1 && axis == 0) hipLaunchKernelGGL((probe_riemann_kernel<T, RS, 0, 1>), grid, block, 0, st, gamma, n, wl, wr, f);
    else if (nd == 2 && axis == 0) hipLaunchKernelGGL((probe_riemann_kernel<T, RS, 0, 2>), grid, block, 0, st, gamma, n, wl, wr, f);
    else if (nd == 2 && axis == 1) hipLaunchKernelGGL((probe_riemann_kernel<T, RS, 1, 2>), grid, block, 0, st, gamma, n, wl, wr, f);
    else if (nd == 3 && axis == 0) hipLaunchKernelGGL((probe_riemann_kernel<T, RS, 0, 3>), grid, block, 0, st, gamma, n, wl, wr, f);
    else if (nd == 3 && axis == 1) hipLaunchKernelGGL((probe_riemann_kernel<T, RS, 1, 3>), grid, block, 0, st, gamma, n, wl, wr, f);
    else if (nd == 3 && axis == 2) hipLaunchKernelGGL((probe_riemann_kernel<T, RS, 2, 3>), grid, block, 0, st, gamma, n, wl, wr, f);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename T>
hipError_t probe_riemann_t(int solver, int axis, int nd, double gamma, int n, const double* wl, const double* wr,
                           double* f, hipStream_t st) {
    switch ((RiemannSolver)solver) {
        case RiemannSolver::Rusanov: return probe_riemann_rs<T, RiemannSolver::Rusanov>(axis, nd, gamma, n, wl, wr, f, st);
        case RiemannSolver::Hll: return probe_riemann_rs<T, RiemannSolver::Hll>(axis, nd, gamma, n, wl, wr, f, st);
        case RiemannSolver::Hllc: return probe_riemann_rs<T, RiemannSolver::Hllc>(axis, nd, gamma, n, wl, wr, f, st);
        default: return hipErrorInvalidValue;
    }
}

// dt from the reduced signal speed, then the output-time clamp
// (Hydro.cpp:230-241). With `advance`, also runs main.cpp:54-71's bookkeeping:
// time += dt and, once an output time is reached, prev_output_time = time.
// Resets the running max for the next step.
// One wave: lane l reduces and clears slot l.
__global__ void clock_finalize_kernel(Clock* c, int advance) {
    if (blockIdx.x != 0 || threadIdx.x >= kWave) return;
    const double dt = clock_next_dt(c);
    amax_slots(c)[kAmaxStride * threadIdx.x] = 0ull;  // after every lane has read its slot (wave order)
    if (threadIdx.x != 0) return;
    c->dt = dt;
    if (advance) clock_advance(c);
}

__global__ void clock_set_dt_kernel(Clock* c, double dt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c->dt = dt;
}

__global__ void clock_reset_max_kernel(Clock* c) {
    if (blockIdx.x == 0 && threadIdx.x < kAmaxSlots) amax_slots(c)[kAmaxStride * threadIdx.x] = 0ull;
}

}  // namespace

hipError_t launch_ghost_f64(int nd, const GhostGrid& g, FormF64::Planes in, FormF64::Planes out,
                            unsigned long long* amax, hipStream_t st) {
    return launch_ghost_impl<FormF64>(nd, g, in, out, amax, st);
}
hipError_t launch_ghost_f32c(int nd, const GhostGrid& g, FormF32c::Planes in, FormF32c::Planes out,
                             unsigned long long* amax, hipStream_t st) {
    return launch_ghost_impl<FormF32c>(nd, g, in, out, amax, st);
}
template <class Form>
hipError_t launch_tab_eos_impl(int nd, typename Form::Planes in, typename Form::Planes out, int ncells, int nvar,
                               const TabParams& t, const double* ytab, const double* ltab, double* y_space,
                               double* T_space, int* neg, hipStream_t st) {
    if (ncells < 1 || nvar < 2 + nd || t.ex_num < 1 || t.r_num < 1 || !ytab || !ltab || !y_space || !T_space || !neg)
        return hipErrorInvalidValue;
    const dim3 grid(blocks_for(ncells)), block(kBlock);
    if (nd == 1) hipLaunchKernelGGL((tab_eos_kernel<Form, 1>), grid, block, 0, st, in, out, ncells, nvar, t, ytab, ltab, y_space, T_space, neg);
    else if (nd == 2) hipLaunchKernelGGL((tab_eos_kernel<Form, 2>), grid, block, 0, st, in, out, ncells, nvar, t, ytab, ltab, y_space, T_space, neg);
    else if (nd == 3) hipLaunchKernelGGL((tab_eos_kernel<Form, 3>), grid, block, 0, st, in, out, ncells, nvar, t, ytab, ltab, y_space, T_space, neg);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_tab_eos_f64(int nd, FormF64::Planes in, FormF64::Planes out, int ncells, int nvar,
                              const TabParams& t, const double* ytab, const double* ltab, double* y_space,
                              double* T_space, int* neg, hipStream_t st) {
    return launch_tab_eos_impl<FormF64>(nd, in, out, ncells, nvar, t, ytab, ltab, y_space, T_space, neg, st);
}
hipError_t launch_tab_eos_f32c(int nd, FormF32c::Planes in, FormF32c::Planes out, int ncells, int nvar,
                               const TabParams& t, const double* ytab, const double* ltab, double* y_space,
                               double* T_space, int* neg, hipStream_t st) {
    return launch_tab_eos_impl<FormF32c>(nd, in, out, ncells, nvar, t, ytab, ltab, y_space, T_space, neg, st);
}

hipError_t launch_ghost3d_f64(const Ghost3Grid& g, FormF64::Planes q, unsigned long long* amax, hipStream_t st) {
    return launch_ghost3d_impl<FormF64>(g, q, amax, st);
}
hipError_t launch_ghost3d_f32c(const Ghost3Grid& g, FormF32c::Planes q, unsigned long long* amax, hipStream_t st) {
    return launch_ghost3d_impl<FormF32c>(g, q, amax, st);
}
template <class Form>
hipError_t launch_lte_impl(int nd, typename Form::Planes q, int ncells, const LteParams& p, double* y, double* T,
                           int* flipped, hipStream_t st) {
    if (ncells < 1 || !y || !T || !flipped) return hipErrorInvalidValue;
    if (nd == 1) {
        hipLaunchKernelGGL((lte_eos_kernel<Form, 1>), dim3(blocks_for(ncells)), dim3(kBlock), 0, st, q, ncells, p, y, T,
                           flipped);
    } else if (nd == 2) {
        hipLaunchKernelGGL((lte_eos_kernel<Form, 2>), dim3(blocks_for(ncells)), dim3(kBlock), 0, st, q, ncells, p, y, T,
                           flipped);
    } else if (nd == 3) {
        hipLaunchKernelGGL((lte_eos_kernel<Form, 3>), dim3(blocks_for(ncells)), dim3(kBlock), 0, st, q, ncells, p, y, T,
                           flipped);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_lte_eos_f64(int nd, FormF64::Planes q, int ncells, const LteParams& p, double* y, double* T,
                              int* flipped, hipStream_t st) {
    return launch_lte_impl<FormF64>(nd, q, ncells, p, y, T, flipped, st);
}
hipError_t launch_lte_eos_f32c(int nd, FormF32c::Planes q, int ncells, const LteParams& p, double* y, double* T,
                               int* flipped, hipStream_t st) {
    return launch_lte_impl<FormF32c>(nd, q, ncells, p, y, T, flipped, st);
}

hipError_t launch_tracer_ghost_f64(const GhostGrid& g, int ntr, FormF64::Planes in, FormF64::Planes out,
                                   hipStream_t st) {
    if (ntr < 1 || g.ng < 1 || g.xc < 3 * g.ng || g.yc < 3 * g.ng) return hipErrorInvalidValue;
    const long long n = 2LL * g.ng * g.yc + 2LL * g.ng * g.xc;
    hipLaunchKernelGGL(tracer_ghost_kernel<FormF64>, dim3(blocks_for(n)), dim3(kBlock), 0, st, g, ntr, in, out);
    return hipGetLastError();
}
hipError_t launch_tracer_ghost_f32c(const GhostGrid& g, int ntr, FormF32c::Planes in, FormF32c::Planes out,
                                    hipStream_t st) {
    if (ntr < 1 || g.ng < 1 || g.xc < 3 * g.ng || g.yc < 3 * g.ng) return hipErrorInvalidValue;
    const long long n = 2LL * g.ng * g.yc + 2LL * g.ng * g.xc;
    hipLaunchKernelGGL(tracer_ghost_kernel<FormF32c>, dim3(blocks_for(n)), dim3(kBlock), 0, st, g, ntr, in, out);
    return hipGetLastError();
}
hipError_t launch_cfl_f64(int nd, FormF64::Planes q, int ncells, double gamma, unsigned long long* amax,
                          hipStream_t st) {
    return launch_cfl_impl<FormF64>(nd, q, ncells, gamma, amax, st);
}
hipError_t launch_cfl_f32c(int nd, FormF32c::Planes q, int ncells, double gamma, unsigned long long* amax,
                           hipStream_t st) {
    return launch_cfl_impl<FormF32c>(nd, q, ncells, gamma, amax, st);
}

hipError_t launch_clock_finalize(Clock* c, int advance, hipStream_t st) {
    hipLaunchKernelGGL(clock_finalize_kernel, dim3(1), dim3(64), 0, st, c, advance);
    return hipGetLastError();
}
hipError_t launch_clock_set_dt(Clock* c, double dt, hipStream_t st) {
    hipLaunchKernelGGL(clock_set_dt_kernel, dim3(1), dim3(64), 0, st, c, dt);
    return hipGetLastError();
}
hipError_t launch_clock_reset_max(Clock* c, hipStream_t st) {
    hipLaunchKernelGGL(clock_reset_max_kernel, dim3(1), dim3(64), 0, st, c);
    return hipGetLastError();
}

hipError_t launch_split_f64_to_f32c(const double* src, FormF32c::Planes dst, int n, hipStream_t st) {
    if (n < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(split_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, st, src, dst.hi, dst.lo, n);
    return hipGetLastError();
}
hipError_t launch_join_f32c_to_f64(FormF32c::Planes src, double* dst, int n, hipStream_t st) {
    if (n < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(join_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, st, src.hi, src.lo, dst, n);
    return hipGetLastError();
}

hipError_t launch_probe_reconstruct(int recon, int limiter, int prec, int n, const double* stencils, double* wl,
                                    double* wr, hipStream_t st) {
    if (n < 1) return hipErrorInvalidValue;
    return prec == 0 ? probe_reconstruct_t<double>(recon, limiter, n, stencils, wl, wr, st)
                     : probe_reconstruct_t<float>(recon, limiter, n, stencils, wl, wr, st);
}
hipError_t launch_probe_riemann(int solver, int axis, int nd, int prec, double gamma, int n, const double* wl,
                                const double* wr, double* flux, hipStream_t st) {
    if (n < 1) return hipErrorInvalidValue;
    return prec == 0 ? probe_riemann_t<double>(solver, axis, nd, gamma, n, wl, wr, flux, st)
                     : probe_riemann_t<float>(solver, axis, nd, gamma, n, wl, wr, flux, st);
}
hipError_t launch_probe_math(int op, int n, const double* a, const double* b, double* out, hipStream_t st) {
    if (n < 1 || op < 0 || op > 3) return hipErrorInvalidValue;
    hipLaunchKernelGGL(probe_math_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, st, op, n, a, b, out);
    return hipGetLastError();
}

}  // namespace mh
