// Discrete adjoint of one 2-D stage, FP64 (-ffp-contract=off). A stage maps
// (Q_s, Q^n) to a_qn*Q^n + a_qs*Q_s + a_l*L(Q_s) with L(Q) = (dt/dx) * (flux
// divergence of ghost-filled Q). Given the cotangent mu of its output, the
// stage's reverse gives a_qn*mu to Q^n and a_qs*mu + a_l*J_L(Q_s)^T mu to Q_s.
//
// J_L^T mu runs as two line sweeps, one per axis, each a single pass over
// lines of the grid with one wavefront per 64-cell segment of a line (lanes
// along the line, every intermediate in LDS):
//
//   cons->prim of the segment; reconstruction of every cell; per face the
//   Riemann VJP of (dt/dx)(mu_right - mu_left) at the reconstructed states
//   (cotangents of RR of the left cell and RL of the right cell); per cell
//   the reconstruction VJP into its 2R+1 stencil cells; per cell the sum of
//   the contributions aimed at it.
//
// A segment's outer 2R+1 lanes on each side are halo (their sums are
// incomplete), so a wave emits 62 - 4R cells. The x sweep writes the
// primitive cotangents of the interior rows; the y sweep adds its own,
// applies the cons->prim VJP and the stage combination, and writes every cell
// of the output (ghost cells get a_l * their cotangent and are folded into
// their sources by the transposed ghost fill afterwards: y ghosts, then x
// ghosts, the reverse of the fill). An axis periodic on both faces runs its
// lines as circles instead: a wave's halo positions past either end read the
// periodic images, so the faces across the seam are differentiated by the
// waves that own the cells on either side and no cotangent reaches a ghost.
#include "../device/adjoint.hpp"
#include "../device/ghost_dt.hpp"
#include "../device/wave.hpp"
#include "../launch_adjoint.hpp"

namespace mh {

namespace {

constexpr int kB = 256;

struct LineArgs {
    AdjGrid g;
    const double* q;
    const double* mu;
    double* lamx;
    double* nxt;
    double* acc;
    double a_qn, a_qs, a_l;
    int32_t acc_mode;  // AdjStage::acc_mode
    int32_t nseg;
    // the axis is periodic on both faces: its lines are circular (positions
    // run over the interior, halo positions wrap), so no contribution lands
    // in a ghost cell and that axis needs no transposed ghost fill
    int32_t circ_x, circ_y;
};

template <RiemannSolver RS, int Axis>
__device__ inline void face_vjp(double gamma, const double (&wl)[4], const double (&wr)[4], const double (&lf)[4],
                                double (&ll)[4], double (&lr)[4]) {
    if constexpr (RS == RiemannSolver::Rusanov) {
        rusanov_vjp2<Axis>(gamma, wl, wr, lf, ll, lr);
    } else if constexpr (RS == RiemannSolver::Hllc) {
        hllc_vjp2<Axis>(gamma, wl, wr, lf, ll, lr);
    } else {
        riemann_vjp2_dual<RS, Axis>(gamma, wl, wr, lf, ll, lr);
    }
}

// One wavefront per (line, segment). Axis 0: lines are the interior rows,
// positions run over all columns (over the interior columns, wrapping, when x
// is circular). Axis 1: lines are all columns (ghost columns only carry the x
// sweep's cotangents to the epilogue; the interior columns alone when x is
// circular), positions run over all rows (the interior rows, wrapping, when
// y is circular).
template <Reconstruction R, SlopeLimiter SL, RiemannSolver RS, int Axis>
__global__ __launch_bounds__(64) void adj_line_kernel(LineArgs a) {
    constexpr int ORD = stencil_order(R);
    constexpr int HALO = 2 * ORD + 1;
    constexpr int OUT = kWave - 2 * HALO;
    // 8 KB of LDS per wave, so that 4 waves fit each SIMD
    __shared__ double sW[4][kWave], sWR[4][kWave], sMu[4][kWave], sLR[4][kWave];

    const AdjGrid& g = a.g;
    const int t = threadIdx.x;
    // Logical block ids run contiguously within each XCD (xcd_remap), so
    // neighbouring lines share an L2. The x sweep keeps a row's segments
    // adjacent (they share halo cells); the y sweep takes columns fastest, so
    // the 16 neighbouring columns that share each cache line of a row run
    // together.
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t nlines = gridDim.x / (uint32_t)a.nseg;
    const int seg = Axis == 0 ? (int)(lb % (uint32_t)a.nseg) : (int)(lb / nlines);
    const int line = Axis == 0 ? (int)(lb / (uint32_t)a.nseg) : (int)(lb % nlines);
    const int len = Axis == 0 ? g.xc : g.yc;
    const int nax = Axis == 0 ? g.nx : g.ny;
    const bool circ = (Axis == 0 ? a.circ_x : a.circ_y) != 0;
    // position along the line (with ghosts); circular lines number the
    // interior from 0 and read a position outside it at its periodic image
    const int p = circ ? g.ng + seg * OUT - HALO + t : seg * OUT - HALO + t;
    const int pc = circ ? g.ng + ((p - g.ng) % nax + nax) % nax : min(max(p, 0), len - 1);
    const int li = Axis == 0 ? g.ng + line : (a.circ_x ? g.ng + line : line);  // row (axis 0) or column (axis 1)
    const bool line_in = Axis == 0 ? true : (li >= g.ng && li < g.ng + g.nx);
    const int cell = Axis == 0 ? li * g.xc + pc : pc * g.xc + li;
    const int i = Axis == 0 ? pc : li, j = Axis == 0 ? li : pc;
    const bool cell_in = i >= g.ng && i < g.ng + g.nx && j >= g.ng && j < g.ng + g.ny;
    const int nc = g.ncells;

    const int qcell = g.virt ? copy_ghost_source(j, g.ny, g.ng, g.yper) * g.xc + copy_ghost_source(i, g.nx, g.ng, g.xper)
                             : cell;
    double qq[4], mu[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        qq[v] = a.q[v * nc + qcell];
        mu[v] = cell_in ? a.mu[v * nc + cell] : 0.0;
    }
    {
        double w[4];
        cons_to_prim<2>(g.gamma, qq, w);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            sW[v][t] = w[v];
            sMu[v][t] = mu[v];
            sLR[v][t] = 0.0;
        }
    }
    wave_lds_sync();

    // reconstruction of the cell at p (cells ng-1 .. ng+n along the axis)
    const bool rec_ok = line_in && t >= ORD && t <= kWave - 1 - ORD && (circ || (p >= g.ng - 1 && p <= g.ng + nax));
    double wl[4] = {0.0, 0.0, 0.0, 0.0}, wr[4] = {0.0, 0.0, 0.0, 0.0};
    if (rec_ok) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            auto s = [&](int o) -> double { return sW[v][t + o]; };
            reconstruct_stencil<R, SL, double>(s, wl[v], wr[v]);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) sWR[v][t] = wr[v];
    wave_lds_sync();

    // face at the left edge of p (faces ng .. ng+n): RR of p-1 against RL of p
    double lL[4] = {0.0, 0.0, 0.0, 0.0};
    if (line_in && t >= ORD + 1 && t <= kWave - 1 - ORD && (circ || (p >= g.ng && p <= g.ng + nax))) {
        double wrl[4], lf[4], lrl[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            wrl[v] = sWR[v][t - 1];
            lf[v] = g.dtdx * (sMu[v][t] - sMu[v][t - 1]);
        }
        face_vjp<RS, Axis>(g.gamma, wrl, wl, lf, lrl, lL);
#pragma unroll
        for (int v = 0; v < 4; ++v) sLR[v][t - 1] = lrl[v];
    }
    wave_lds_sync();

    // reconstruction VJP of the cell at p into its stencil, and per cell the
    // sum of the contributions aimed at it: stencil entry o of the cell in
    // lane t - (o - ORD), read across lanes, summed in order of o
    const bool vjp_ok = rec_ok && t >= ORD + 1 && t <= kWave - 2 - ORD;
    double lw[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double ls[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (vjp_ok) {
            double st[5];
#pragma unroll
            for (int o = 0; o < 5; ++o) st[o] = (o >= 2 - ORD && o <= 2 + ORD) ? sW[v][t + o - 2] : 0.0;
            if constexpr (R == Reconstruction::Weno5Z) {
                weno5z_vjp(st, lL[v], sLR[v][t], ls);
            } else {
                reconstruct_vjp_dual<R, SL>(st, lL[v], sLR[v][t], ls);
            }
        }
        double sum = 0.0;
#pragma unroll
        for (int o = 0; o < 2 * ORD + 1; ++o)
            sum += o == ORD ? ls[2] : lane_value(ls[2 - ORD + o], t - (o - ORD));
        lw[v] = sum;
    }

    if (t < HALO || t >= HALO + OUT || p >= (circ ? g.ng + nax : len)) return;
    if constexpr (Axis == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) a.lamx[v * nc + cell] = lw[v];
    } else {
        const bool row_in = j >= g.ng && j < g.ng + g.ny;
        if (row_in) {
#pragma unroll
            for (int v = 0; v < 4; ++v) lw[v] += a.lamx[v * nc + cell];
        }
        double lq[4];
        cons_to_prim_vjp2(g.gamma, qq, lw, lq);
        if (g.gravity && cell_in) {  // S*dt = dt*(0, rho gx, rho gy, mx gx + my gy) of this cell
            lq[0] += g.dt * (g.gx * mu[1] + g.gy * mu[2]);
            lq[1] += g.dt * (g.gx * mu[3]);
            lq[2] += g.dt * (g.gy * mu[3]);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int k = v * nc + cell;
            if (cell_in) {
                double r = a.a_qs * mu[v] + a.a_l * lq[v];
                if (a.acc_mode == 2) r += a.acc[k];
                a.nxt[k] = r;
                if (a.acc_mode == 1) a.acc[k] = a.a_qn * mu[v];
                else if (a.acc_mode == 0 && a.a_qn != 0.0) a.acc[k] += a.a_qn * mu[v];
            } else {
                a.nxt[k] = a.a_l * lq[v];
                if (a.acc_mode == 1) a.acc[k] = 0.0;
            }
        }
    }
}

// Transpose of one ghost pass (the fill copies a source into each ghost, with
// the normal momentum negated at walls): each ghost's cotangent is added to its
// source and the ghost cleared. One thread per (line across the axis,
// variable); a side's ghosts are read together and zerograd's ng ghosts, which
// share one source, are summed first. The two sides run in order, so they may
// share sources on a narrow grid.
constexpr int kMaxFoldGhosts = 8;

template <int Axis>
__global__ void adj_fold_kernel(AdjGrid g, double* __restrict__ lam) {
    const int nline = Axis == 0 ? g.yc : g.xc;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 4 * nline) return;
    const int v = t / nline, line = t % nline;
    const int dims = Axis == 0 ? g.xc : g.yc;
    const int stride = Axis == 0 ? 1 : g.xc;
    double* L = lam + (size_t)v * g.ncells + (Axis == 0 ? line * g.xc : line);
    constexpr int IM = 1 + Axis;
    const int ng = g.ng;
    for (int side = 0; side < 2; ++side) {
        const BoundaryType b = (BoundaryType)g.bc[2 * Axis + side];
        const bool maps = b == BoundaryType::Wall || b == BoundaryType::Periodic || b == BoundaryType::Symmetric ||
                          b == BoundaryType::ZeroGrad;
        double gv[kMaxFoldGhosts];
#pragma unroll
        for (int k = 0; k < kMaxFoldGhosts; ++k) {
            if (k >= ng) break;
            const int gc = side == 0 ? k : dims - 1 - k;
            gv[k] = L[gc * stride];
            L[gc * stride] = 0.0;
        }
        if (!maps) continue;
        const double sgn = (b == BoundaryType::Wall && v == IM) ? -1.0 : 1.0;
        if (b == BoundaryType::ZeroGrad) {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < kMaxFoldGhosts; ++k) {
                if (k >= ng) break;
                sum += gv[k];
            }
            const int sc = ghost_source(side == 0 ? 0 : dims - 1, dims, ng, b);
            L[sc * stride] += sum;
        } else {
            double sv[kMaxFoldGhosts];
            int si[kMaxFoldGhosts];
#pragma unroll
            for (int k = 0; k < kMaxFoldGhosts; ++k) {
                if (k >= ng) break;
                si[k] = ghost_source(side == 0 ? k : dims - 1 - k, dims, ng, b);
                sv[k] = L[si[k] * stride];
            }
#pragma unroll
            for (int k = 0; k < kMaxFoldGhosts; ++k) {
                if (k >= ng) break;
                L[si[k] * stride] = sv[k] + sgn * gv[k];
            }
        }
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
hipError_t stage_impl(const AdjGrid& g, const AdjStage& st, hipStream_t stream) {
    constexpr int OUT = kWave - 2 * (2 * stencil_order(R) + 1);
    LineArgs a{g, st.q, st.mu, st.lamx, st.nxt, st.acc, st.a_qn, st.a_qs, st.a_l, st.acc_mode, 0, 0, 0};
    const auto per = (int32_t)BoundaryType::Periodic;
    a.circ_x = g.bc[0] == per && g.bc[1] == per;
    a.circ_y = g.bc[2] == per && g.bc[3] == per;
    a.nseg = ((a.circ_x ? g.nx : g.xc) + OUT - 1) / OUT;
    hipLaunchKernelGGL((adj_line_kernel<R, SL, RS, 0>), dim3(a.nseg * g.ny), dim3(kWave), 0, stream, a);
    a.nseg = ((a.circ_y ? g.ny : g.yc) + OUT - 1) / OUT;
    hipLaunchKernelGGL((adj_line_kernel<R, SL, RS, 1>), dim3(a.nseg * (a.circ_x ? g.nx : g.xc)), dim3(kWave), 0, stream,
                       a);
    // ghost cotangents of the non-circular axes into their sources (the
    // ghost entries of a circular axis are never written and stay zero)
    if (!a.circ_y)
        hipLaunchKernelGGL(adj_fold_kernel<1>, dim3((4 * g.xc + kB - 1) / kB), dim3(kB), 0, stream, g, st.nxt);
    if (!a.circ_x)
        hipLaunchKernelGGL(adj_fold_kernel<0>, dim3((4 * g.yc + kB - 1) / kB), dim3(kB), 0, stream, g, st.nxt);
    return hipGetLastError();
}

template <RiemannSolver RS>
hipError_t stage_rs(Reconstruction r, SlopeLimiter sl, const AdjGrid& g, const AdjStage& st, hipStream_t stream) {
    using RC = Reconstruction;
    using SLm = SlopeLimiter;
    switch (r) {
        case RC::Fog: return stage_impl<RS, RC::Fog, SLm::MonotonizedCentral>(g, st, stream);
        case RC::Weno5Z: return stage_impl<RS, RC::Weno5Z, SLm::MonotonizedCentral>(g, st, stream);
        case RC::Muscl:
            if (sl == SLm::VanLeer) return stage_impl<RS, RC::Muscl, SLm::VanLeer>(g, st, stream);
            if (sl == SLm::Minmod) return stage_impl<RS, RC::Muscl, SLm::Minmod>(g, st, stream);
            return stage_impl<RS, RC::Muscl, SLm::MonotonizedCentral>(g, st, stream);
        default:
            if (sl == SLm::VanLeer) return stage_impl<RS, RC::Ppm, SLm::VanLeer>(g, st, stream);
            if (sl == SLm::Minmod) return stage_impl<RS, RC::Ppm, SLm::Minmod>(g, st, stream);
            return stage_impl<RS, RC::Ppm, SLm::MonotonizedCentral>(g, st, stream);
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
    } else {  // Rusanov x (2 hand, 3 dual), Rusanov y (4, 5), HLLC x (6, 7), HLLC y (8, 9): in[8], cot[4] -> out[8]
        double a[4], b[4], lf[4], la[4] = {0, 0, 0, 0}, lb[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) {
            a[k] = in[8 * t + k];
            b[k] = in[8 * t + 4 + k];
            lf[k] = cot[4 * t + k];
        }
        if (what == 2) rusanov_vjp2<0>(gamma, a, b, lf, la, lb);
        else if (what == 3) riemann_vjp2_dual<RiemannSolver::Rusanov, 0>(gamma, a, b, lf, la, lb);
        else if (what == 4) rusanov_vjp2<1>(gamma, a, b, lf, la, lb);
        else if (what == 5) riemann_vjp2_dual<RiemannSolver::Rusanov, 1>(gamma, a, b, lf, la, lb);
        else if (what == 6) hllc_vjp2<0>(gamma, a, b, lf, la, lb);
        else if (what == 7) riemann_vjp2_dual<RiemannSolver::Hllc, 0>(gamma, a, b, lf, la, lb);
        else if (what == 8) hllc_vjp2<1>(gamma, a, b, lf, la, lb);
        else riemann_vjp2_dual<RiemannSolver::Hllc, 1>(gamma, a, b, lf, la, lb);
        for (int k = 0; k < 4; ++k) {
            out[8 * t + k] = la[k];
            out[8 * t + 4 + k] = lb[k];
        }
    }
}

}  // namespace

hipError_t launch_adjoint_stage(const Schemes& sch, const AdjGrid& g, const AdjStage& st, hipStream_t stream) {
    if (g.nx < g.ng || g.ny < g.ng || g.ng < min_ghost_cells(sch.recon) || g.xc != g.nx + 2 * g.ng ||
        g.yc != g.ny + 2 * g.ng || g.ncells != g.xc * g.yc || g.ng > kMaxFoldGhosts || !st.q || !st.mu ||
        !st.lamx || !st.nxt || !st.acc || st.acc_mode < 0 || st.acc_mode > 2)
        return hipErrorInvalidValue;
    switch (sch.solver) {
        case RiemannSolver::Rusanov: return stage_rs<RiemannSolver::Rusanov>(sch.recon, sch.limiter, g, st, stream);
        case RiemannSolver::Hll: return stage_rs<RiemannSolver::Hll>(sch.recon, sch.limiter, g, st, stream);
        default: return stage_rs<RiemannSolver::Hllc>(sch.recon, sch.limiter, g, st, stream);
    }
}

hipError_t launch_adjoint_axpby(const AdjGrid& g, double a, const double* x, double b, double* y, hipStream_t st) {
    hipLaunchKernelGGL(adj_axpby_kernel, dim3((g.ncells + kB - 1) / kB), dim3(kB), 0, st, g, a, x, b, y);
    return hipGetLastError();
}

hipError_t launch_probe_vjp(int what, int n, const double* in, const double* cot, double* out, double gamma,
                            hipStream_t st) {
    if (n < 1 || what < 0 || what > 9) return hipErrorInvalidValue;
    hipLaunchKernelGGL(probe_vjp_kernel, dim3((n + kB - 1) / kB), dim3(kB), 0, st, what, n, in, cot, out, gamma);
    return hipGetLastError();
}

}  // namespace mh
