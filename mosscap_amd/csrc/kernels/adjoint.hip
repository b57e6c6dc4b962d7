// Discrete adjoint of the 2-D stage operator L(Q) = (dt/dx) * (flux
// divergence of ghost-filled Q), FP64 (-ffp-contract=off). Given a cotangent
// mu on the interior, computes J_L(Q)^T mu on the interior in five passes:
//
//   c2p      W = cons_to_prim(Q) on every cell
//   faces    per x (y) face: lambda_F = (dt/dx)(mu_right - mu_left), the
//            Riemann VJP at the reconstructed face states -> cotangents of
//            RR(left cell) and RL(right cell)
//   recon    per reconstructed cell and variable: reconstruction VJP -> five
//            stencil contributions, stored by offset (no atomics)
//   gather   per cell: sum of the contributions aimed at it, then the
//            cons_to_prim VJP -> lambda_Q (ghosts included)
//   fold     transpose of the ghost fill: y ghosts into their sources, then
//            x ghosts (the fill runs x then y), ghosts cleared.
#include "../device/adjoint.hpp"
#include "../device/ghost_dt.hpp"
#include "../launch_adjoint.hpp"

namespace mh {

namespace {

constexpr int kB = 256;

__global__ void adj_c2p_kernel(const double* __restrict__ q, double* __restrict__ w, int ncells, double gamma) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncells) return;
    double qq[4], ww[4];
    for (int v = 0; v < 4; ++v) qq[v] = q[v * ncells + i];
    cons_to_prim<2>(gamma, qq, ww);
    for (int v = 0; v < 4; ++v) w[v * ncells + i] = ww[v];
}

template <Reconstruction R, SlopeLimiter SL>
__device__ inline void recon_cell(const double* w, int ncells, int cell, int stride, double (&l)[4], double (&r)[4]) {
    for (int v = 0; v < 4; ++v) {
        const double* p = w + v * ncells + cell;
        auto s = [&](int o) -> double { return p[o * stride]; };
        reconstruct_stencil<R, SL, double>(s, l[v], r[v]);
    }
}

// one thread per face of the given axis: faces f in [ng, ng + n] along the
// axis, interior along the other axis; face f lies between cells f-1 and f.
template <Reconstruction R, SlopeLimiter SL, RiemannSolver RS, int Axis>
__global__ void adj_face_kernel(AdjGrid g, const double* __restrict__ w, const double* __restrict__ mu,
                                double* __restrict__ lam_rr, double* __restrict__ lam_rl) {
    const int nface = (Axis == 0 ? g.nx + 1 : g.ny + 1);
    const int nline = (Axis == 0 ? g.ny : g.nx);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nface * nline) return;
    const int f = g.ng + t % nface;
    const int l = g.ng + t / nface;
    const int i = Axis == 0 ? f : l;
    const int j = Axis == 0 ? l : f;
    const int stride = Axis == 0 ? 1 : g.xc;
    const int cr = j * g.xc + i;
    const int cl = cr - stride;
    double wl_l[4], wr_l[4], wl_r[4], wr_r[4];
    recon_cell<R, SL>(w, g.ncells, cl, stride, wl_l, wr_l);
    recon_cell<R, SL>(w, g.ncells, cr, stride, wl_r, wr_r);
    const bool right_in = f < g.ng + (Axis == 0 ? g.nx : g.ny);
    const bool left_in = f > g.ng;
    double lf[4];
    for (int v = 0; v < 4; ++v) {
        const double mr = right_in ? mu[v * g.ncells + cr] : 0.0;
        const double ml = left_in ? mu[v * g.ncells + cl] : 0.0;
        lf[v] = g.dtdx * (mr - ml);
    }
    double ll[4] = {0.0, 0.0, 0.0, 0.0}, lr[4] = {0.0, 0.0, 0.0, 0.0};
    if constexpr (RS == RiemannSolver::Rusanov) {
        rusanov_vjp2<Axis>(g.gamma, wr_l, wl_r, lf, ll, lr);
    } else {
        riemann_vjp2_dual<RS, Axis>(g.gamma, wr_l, wl_r, lf, ll, lr);
    }
    for (int v = 0; v < 4; ++v) {
        lam_rr[v * g.ncells + cl] = ll[v];
        lam_rl[v * g.ncells + cr] = lr[v];
    }
}

// one thread per reconstructed cell of the axis ([ng-1, ng+n] along it,
// interior across): five stencil contributions per variable into gsc[o][v][cell].
template <Reconstruction R, SlopeLimiter SL, int Axis>
__global__ void adj_recon_kernel(AdjGrid g, const double* __restrict__ w, const double* __restrict__ lam_rr,
                                 const double* __restrict__ lam_rl, double* __restrict__ gsc) {
    const int nrec = (Axis == 0 ? g.nx + 2 : g.ny + 2);
    const int nline = (Axis == 0 ? g.ny : g.nx);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrec * nline) return;
    const int c = g.ng - 1 + t % nrec;
    const int l = g.ng + t / nrec;
    const int i = Axis == 0 ? c : l;
    const int j = Axis == 0 ? l : c;
    const int stride = Axis == 0 ? 1 : g.xc;
    const int cell = j * g.xc + i;
    const int nend = g.ng + (Axis == 0 ? g.nx : g.ny);
    for (int v = 0; v < 4; ++v) {
        // RR of cell c feeds face c+1 (exists for c <= ng+n-1); RL of c feeds face c (c >= ng)
        const double lR = (c <= nend - 1) ? lam_rr[v * g.ncells + cell] : 0.0;
        const double lL = (c >= g.ng) ? lam_rl[v * g.ncells + cell] : 0.0;
        double s[5], ls[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int o = 0; o < 5; ++o) s[o] = (o >= 2 - stencil_order(R) && o <= 2 + stencil_order(R))
                                                ? w[v * g.ncells + cell + (o - 2) * stride] : 0.0;
        if constexpr (R == Reconstruction::Weno5Z) {
            weno5z_vjp(s, lL, lR, ls);
        } else {
            reconstruct_vjp_dual<R, SL>(s, lL, lR, ls);
        }
        for (int o = 0; o < 5; ++o) gsc[(o * 4 + v) * g.ncells + cell] = ls[o];
    }
}

// per cell: gather the stencil contributions aimed at it from both axes, then
// the cons_to_prim VJP; lam_q gets J^T mu on every cell (ghosts not yet folded).
__global__ void adj_gather_kernel(AdjGrid g, const double* __restrict__ q, const double* __restrict__ gx,
                                  const double* __restrict__ gy, double* __restrict__ lam_q, int order) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= g.ncells) return;
    const int i = cell % g.xc, j = cell / g.xc;
    double lw[4] = {0.0, 0.0, 0.0, 0.0};
    const bool row_in = j >= g.ng && j < g.ng + g.ny;
    const bool col_in = i >= g.ng && i < g.ng + g.nx;
    for (int o = 2 - order; o <= 2 + order; ++o) {
        const int d = o - 2;  // this cell is at offset d of the stencil centred on cell - d
        const int cx = i - d;
        if (row_in && cx >= g.ng - 1 && cx <= g.ng + g.nx) {
            for (int v = 0; v < 4; ++v) lw[v] += gx[(o * 4 + v) * g.ncells + cell - d];
        }
        const int cy = j - d;
        if (col_in && cy >= g.ng - 1 && cy <= g.ng + g.ny) {
            for (int v = 0; v < 4; ++v) lw[v] += gy[(o * 4 + v) * g.ncells + cell - d * g.xc];
        }
    }
    double qq[4], lq[4];
    for (int v = 0; v < 4; ++v) qq[v] = q[v * g.ncells + cell];
    cons_to_prim_vjp2(g.gamma, qq, lw, lq);
    for (int v = 0; v < 4; ++v) lam_q[v * g.ncells + cell] = lq[v];
}

// Transpose of one ghost pass: thread per line across the axis, both ends
// handled sequentially (zerograd sends ng ghosts to one source).
template <int Axis>
__global__ void adj_fold_kernel(AdjGrid g, double* __restrict__ lam) {
    const int nline = Axis == 0 ? g.yc : g.xc;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nline) return;
    const int dims = Axis == 0 ? g.xc : g.yc;
    const int stride = Axis == 0 ? 1 : g.xc;
    const int base = Axis == 0 ? t * g.xc : t;
    constexpr int IM = 1 + Axis;
    for (int k = 0; k < 2 * g.ng; ++k) {
        const int gcoord = k < g.ng ? k : dims - 2 * g.ng + k;
        const bool start = gcoord < g.ng;
        const BoundaryType b = (BoundaryType)g.bc[2 * Axis + (start ? 0 : 1)];
        const int gc = base + gcoord * stride;
        if (b == BoundaryType::Wall || b == BoundaryType::Periodic || b == BoundaryType::Symmetric ||
            b == BoundaryType::ZeroGrad) {
            const int sc = base + ghost_source(gcoord, dims, g.ng, b) * stride;
            for (int v = 0; v < 4; ++v) {
                const double x = lam[v * g.ncells + gc];
                lam[v * g.ncells + sc] += (b == BoundaryType::Wall && v == IM) ? -x : x;
            }
        }
        for (int v = 0; v < 4; ++v) lam[v * g.ncells + gc] = 0.0;
    }
}

// y = a*x + b*y on the interior, y = 0 on ghosts
__global__ void adj_axpby_kernel(AdjGrid g, double a, const double* __restrict__ x, double b, double* y) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= g.ncells) return;
    const int i = cell % g.xc, j = cell / g.xc;
    const bool in = i >= g.ng && i < g.ng + g.nx && j >= g.ng && j < g.ng + g.ny;
    for (int v = 0; v < 4; ++v) {
        const int k = v * g.ncells + cell;
        y[k] = in ? (a * x[k] + (b == 0.0 ? 0.0 : b * y[k])) : 0.0;
    }
}

template <RiemannSolver RS, Reconstruction R, SlopeLimiter SL>
hipError_t jt_impl(const AdjGrid& g, const double* q, const double* mu, double* out, const AdjScratch& s,
                   hipStream_t st) {
    const int nc = g.ncells;
    hipLaunchKernelGGL(adj_c2p_kernel, dim3((nc + kB - 1) / kB), dim3(kB), 0, st, q, s.w, nc, g.gamma);
    const int fx = (g.nx + 1) * g.ny, fy = (g.ny + 1) * g.nx;
    hipLaunchKernelGGL((adj_face_kernel<R, SL, RS, 0>), dim3((fx + kB - 1) / kB), dim3(kB), 0, st, g, s.w, mu,
                       s.lam_rr, s.lam_rl);
    const int rx = (g.nx + 2) * g.ny, ry = (g.ny + 2) * g.nx;
    hipLaunchKernelGGL((adj_recon_kernel<R, SL, 0>), dim3((rx + kB - 1) / kB), dim3(kB), 0, st, g, s.w, s.lam_rr,
                       s.lam_rl, s.gx);
    hipLaunchKernelGGL((adj_face_kernel<R, SL, RS, 1>), dim3((fy + kB - 1) / kB), dim3(kB), 0, st, g, s.w, mu,
                       s.lam_rr, s.lam_rl);
    hipLaunchKernelGGL((adj_recon_kernel<R, SL, 1>), dim3((ry + kB - 1) / kB), dim3(kB), 0, st, g, s.w, s.lam_rr,
                       s.lam_rl, s.gy);
    hipLaunchKernelGGL(adj_gather_kernel, dim3((nc + kB - 1) / kB), dim3(kB), 0, st, g, q, s.gx, s.gy, out,
                       stencil_order(R));
    hipLaunchKernelGGL(adj_fold_kernel<1>, dim3((g.xc + kB - 1) / kB), dim3(kB), 0, st, g, out);
    hipLaunchKernelGGL(adj_fold_kernel<0>, dim3((g.yc + kB - 1) / kB), dim3(kB), 0, st, g, out);
    return hipGetLastError();
}

template <RiemannSolver RS>
hipError_t jt_rs(Reconstruction r, SlopeLimiter sl, const AdjGrid& g, const double* q, const double* mu, double* out,
                 const AdjScratch& s, hipStream_t st) {
    using RC = Reconstruction;
    using SLm = SlopeLimiter;
    switch (r) {
        case RC::Fog: return jt_impl<RS, RC::Fog, SLm::MonotonizedCentral>(g, q, mu, out, s, st);
        case RC::Weno5Z: return jt_impl<RS, RC::Weno5Z, SLm::MonotonizedCentral>(g, q, mu, out, s, st);
        case RC::Muscl:
            if (sl == SLm::VanLeer) return jt_impl<RS, RC::Muscl, SLm::VanLeer>(g, q, mu, out, s, st);
            if (sl == SLm::Minmod) return jt_impl<RS, RC::Muscl, SLm::Minmod>(g, q, mu, out, s, st);
            return jt_impl<RS, RC::Muscl, SLm::MonotonizedCentral>(g, q, mu, out, s, st);
        default:
            if (sl == SLm::VanLeer) return jt_impl<RS, RC::Ppm, SLm::VanLeer>(g, q, mu, out, s, st);
            if (sl == SLm::Minmod) return jt_impl<RS, RC::Ppm, SLm::Minmod>(g, q, mu, out, s, st);
            return jt_impl<RS, RC::Ppm, SLm::MonotonizedCentral>(g, q, mu, out, s, st);
    }
}

// probes: the hand VJPs against the dual-number Jacobians
__global__ void probe_vjp_kernel(int what, int n, const double* in, const double* cot, double* out, double gamma) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (what == 0 || what == 1) {  // WENO5-Z: in[5], cot[2] -> out[5]
        double s[5], ls[5] = {0, 0, 0, 0, 0};
        for (int k = 0; k < 5; ++k) s[k] = in[5 * t + k];
        if (what == 0) weno5z_vjp(s, cot[2 * t], cot[2 * t + 1], ls);
        else reconstruct_vjp_dual<Reconstruction::Weno5Z, SlopeLimiter::MonotonizedCentral>(s, cot[2 * t], cot[2 * t + 1], ls);
        for (int k = 0; k < 5; ++k) out[5 * t + k] = ls[k];
    } else {  // Rusanov axis 0 (2: hand, 3: dual) / axis 1 (4: hand, 5: dual): in[8], cot[4] -> out[8]
        double a[4], b[4], lf[4], la[4] = {0, 0, 0, 0}, lb[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) {
            a[k] = in[8 * t + k];
            b[k] = in[8 * t + 4 + k];
            lf[k] = cot[4 * t + k];
        }
        if (what == 2) rusanov_vjp2<0>(gamma, a, b, lf, la, lb);
        else if (what == 3) riemann_vjp2_dual<RiemannSolver::Rusanov, 0>(gamma, a, b, lf, la, lb);
        else if (what == 4) rusanov_vjp2<1>(gamma, a, b, lf, la, lb);
        else riemann_vjp2_dual<RiemannSolver::Rusanov, 1>(gamma, a, b, lf, la, lb);
        for (int k = 0; k < 4; ++k) {
            out[8 * t + k] = la[k];
            out[8 * t + 4 + k] = lb[k];
        }
    }
}

}  // namespace

hipError_t launch_adjoint_jt(const Schemes& sch, const AdjGrid& g, const double* q, const double* mu, double* out,
                             const AdjScratch& s, hipStream_t st) {
    if (g.nx < g.ng || g.ny < g.ng || g.ng < min_ghost_cells(sch.recon) || g.xc != g.nx + 2 * g.ng ||
        g.yc != g.ny + 2 * g.ng || g.ncells != g.xc * g.yc)
        return hipErrorInvalidValue;
    switch (sch.solver) {
        case RiemannSolver::Rusanov: return jt_rs<RiemannSolver::Rusanov>(sch.recon, sch.limiter, g, q, mu, out, s, st);
        case RiemannSolver::Hll: return jt_rs<RiemannSolver::Hll>(sch.recon, sch.limiter, g, q, mu, out, s, st);
        default: return jt_rs<RiemannSolver::Hllc>(sch.recon, sch.limiter, g, q, mu, out, s, st);
    }
}

hipError_t launch_adjoint_axpby(const AdjGrid& g, double a, const double* x, double b, double* y, hipStream_t st) {
    hipLaunchKernelGGL(adj_axpby_kernel, dim3((g.ncells + kB - 1) / kB), dim3(kB), 0, st, g, a, x, b, y);
    return hipGetLastError();
}

hipError_t launch_probe_vjp(int what, int n, const double* in, const double* cot, double* out, double gamma,
                            hipStream_t st) {
    if (n < 1 || what < 0 || what > 5) return hipErrorInvalidValue;
    hipLaunchKernelGGL(probe_vjp_kernel, dim3((n + kB - 1) / kB), dim3(kB), 0, st, what, n, in, cot, out, gamma);
    return hipGetLastError();
}

}  // namespace mh
