// Tangent-linear (forward-mode) model of the FP64 2-D stage, dt frozen.
//
// One thread per interior cell evaluates the stage on dual numbers
// (value, derivative along one direction): cons->prim of the 2R+3 cells of
// its row and of its column, the reconstructions of the cell and its two
// neighbours along each axis, the Riemann fluxes on the cell's four faces,
// the flux differences x before y, and the stage combination with Q^(s)
// and Q^n, all through the same functors and expressions as the fused stage
// (functors.hpp, FormF64::combine / gravity_src). The value parts therefore
// repeat the drop-in's arithmetic bit for bit, and the derivative parts are
// the exact directional derivative of that arithmetic along the branches
// the values take. It is the reference against which the discrete adjoint
// is checked at full size (<J'(x)^T g, d> = <g, J'(x) d>); it is not tuned
// for speed (a thread reloads its stencil from HBM and L2).
#include "../device/adjoint.hpp"
#include "../device/form_f64.hpp"
#include "../device/stage_core.hpp"
#include "../launch_adjoint.hpp"

namespace mh {

namespace {

using D = Dual<1>;

__device__ inline D dual(const double* v, const double* d, int idx) {
    D r(v[idx]);
    r.d[0] = d[idx];
    return r;
}

// cons->prim of the cell at offset o from (i, j) along Axis
template <int Axis>
__device__ inline void prim_at(const TanStage& a, int i, int j, int o, D (&w)[4]) {
    const int idx = (Axis == 0) ? j * a.xc + i + o : (j + o) * a.xc + i;
    const int plane = a.xc * a.yc;
    D q[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) q[v] = dual(a.qv, a.qd, v * plane + idx);
    cons_to_prim<2>(D(a.gamma), q, w);
}

// Fm - Fp on the cell's two faces along Axis (Fm on the face towards lower
// index): the x difference Fx_c - Fx_{c+1} and the y difference
// Gy_j - Gy_{j+1} of the fused stage.
template <Reconstruction R, SlopeLimiter SL, RiemannSolver RS, int Axis>
__device__ inline void face_difference(const TanStage& a, int i, int j, D (&df)[4]) {
    constexpr int ORD = stencil_order(R);
    constexpr int NC = 2 * ORD + 3;  // cells -1-ORD .. 1+ORD
    D w[NC][4];
#pragma unroll
    for (int k = 0; k < NC; ++k) prim_at<Axis>(a, i, j, k - ORD - 1, w[k]);
    D wl[3][4], wr[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            auto s = [&](int o) -> D { return w[c + ORD + o][v]; };
            reconstruct_stencil<R, SL, D>(s, wl[c][v], wr[c][v]);
        }
    }
    const EosT<D> eos{D(a.gamma)};
    D fm[4], fp[4];
    riemann_flux<RS, Axis, 2>(eos, wr[0], wl[1], fm);
    riemann_flux<RS, Axis, 2>(eos, wr[1], wl[2], fp);
#pragma unroll
    for (int v = 0; v < 4; ++v) df[v] = fm[v] - fp[v];
}

template <Reconstruction R, SlopeLimiter SL, RiemannSolver RS>
__global__ void __launch_bounds__(256) tangent_stage_kernel(TanStage a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nx * a.ny) return;
    const int i = a.ng + t % a.nx;
    const int j = a.ng + t / a.nx;
    D dx[4], dy[4];
    face_difference<R, SL, RS, 0>(a, i, j, dx);
    face_difference<R, SL, RS, 1>(a, i, j, dy);
    const int plane = a.xc * a.yc;
    const int idx = j * a.xc + i;
    D qs[4], qn[4], src[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        qs[v] = dual(a.qv, a.qd, v * plane + idx);
        qn[v] = FormF64::needs_qn(a.kind) ? dual(a.nv, a.nd, v * plane + idx) : qs[v];
    }
    const D dt(a.dt);
    if (a.gravity) {
        FormF64::Ptrs p{};
        p.gx = a.gx;
        p.gy = a.gy;
        FormF64::gravity_src<2>(p, qs, dt, src);
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) src[v] = D(0.0) * dt;
    }
    const D dtdx(a.dt / a.dx);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        D q_update(0.0);
        q_update += dx[v];
        q_update += dy[v];
        q_update *= dtdx;
        const D qo = FormF64::combine(a.kind, qs[v], qn[v], q_update, src[v]);
        a.ov[v * plane + idx] = qo.v;
        a.od[v * plane + idx] = qo.d[0];
    }
}

template <Reconstruction R, SlopeLimiter SL>
hipError_t launch_r(RiemannSolver rs, const TanStage& a, hipStream_t st) {
    const dim3 grid((a.nx * a.ny + 255) / 256);
    switch (rs) {
        case RiemannSolver::Rusanov:
            hipLaunchKernelGGL((tangent_stage_kernel<R, SL, RiemannSolver::Rusanov>), grid, dim3(256), 0, st, a);
            break;
        case RiemannSolver::Hll:
            hipLaunchKernelGGL((tangent_stage_kernel<R, SL, RiemannSolver::Hll>), grid, dim3(256), 0, st, a);
            break;
        default:
            hipLaunchKernelGGL((tangent_stage_kernel<R, SL, RiemannSolver::Hllc>), grid, dim3(256), 0, st, a);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_tangent_stage(const Schemes& sch, const TanStage& a, hipStream_t st) {
    if (a.nx < a.ng || a.ny < a.ng || a.ng < min_ghost_cells(sch.recon) || a.xc != a.nx + 2 * a.ng ||
        a.yc != a.ny + 2 * a.ng || (int64_t)a.xc * a.yc * 4 >= ((int64_t)1 << 31) || !a.qv || !a.qd || !a.nv ||
        !a.nd || !a.ov || !a.od || a.kind < 0 || a.kind > 5)
        return hipErrorInvalidValue;
    using RC = Reconstruction;
    using SLm = SlopeLimiter;
    switch (sch.recon) {
        case RC::Fog: return launch_r<RC::Fog, SLm::MonotonizedCentral>(sch.solver, a, st);
        case RC::Weno5Z: return launch_r<RC::Weno5Z, SLm::MonotonizedCentral>(sch.solver, a, st);
        case RC::Muscl:
            if (sch.limiter == SLm::VanLeer) return launch_r<RC::Muscl, SLm::VanLeer>(sch.solver, a, st);
            if (sch.limiter == SLm::Minmod) return launch_r<RC::Muscl, SLm::Minmod>(sch.solver, a, st);
            return launch_r<RC::Muscl, SLm::MonotonizedCentral>(sch.solver, a, st);
        default:
            if (sch.limiter == SLm::VanLeer) return launch_r<RC::Ppm, SLm::VanLeer>(sch.solver, a, st);
            if (sch.limiter == SLm::Minmod) return launch_r<RC::Ppm, SLm::Minmod>(sch.solver, a, st);
            return launch_r<RC::Ppm, SLm::MonotonizedCentral>(sch.solver, a, st);
    }
}

}  // namespace mh
