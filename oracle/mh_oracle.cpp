// TEST INFRASTRUCTURE — NOT PART OF THE PRODUCT.
//
// CPU oracle for the HIP hydro library: a plain restatement of Mosscap's
// unfused finite-volume hot path, op for op, in the reference's evaluation
// order (left-to-right as grouped in the source, no FMA: build with
// -ffp-contract=off). Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may load it. It shares no code with the product.
//
// Layout follows State.hpp:114-121 / Types.hpp:17: arrays are
// [var][k][j][i] with i (x) fastest and ghosts included.
//
// Parity pins (see tests/test_oracle_*.py): closed forms and invariants for
// every functor; the reference ships no golden vectors. Functions whose
// transcription only the invariants check (and which the tests therefore pin
// with hand-worked examples) are marked "pinned by worked example".

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

// Per-kernel wall-time accumulators named like the reference's kernels (the
// dex_parallel_for names YAKL's auto-profiler reports), for the CPU
// baseline's time breakdown.
namespace prof {
enum : int { QtoW, ReconX, FluxX, ReconY, FluxY, ReconZ, FluxZ, Integrate, FillBcs, Dt, Sources, QOld, UpdateEos, N };
const char* const kNames[N] = {"Q -> W", "Reconstruct x", "Flux x", "Reconstruct y", "Flux y", "Reconstruct z",
                               "Flux z", "integrate_flux (stage update)", "Fill BCs", "dt_reducer", "Source terms",
                               "Q_old copy", "update_eos"};
double ms[N] = {};
struct Scope {
    int k;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit Scope(int kk) : k(kk) {}
    ~Scope() { ms[k] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};
}  // namespace prof

namespace {

enum Recon { FOG = 0, MUSCL = 1, PPM = 2, WENO5Z = 3 };              // Reconstruct.hpp:10-15
enum Limiter { VANLEER = 0, MONOTONIZED_CENTRAL = 1, MINMOD = 2 };  // Reconstruct.hpp:22-26
enum Solver { RUSANOV = 0, HLL = 1, HLLC = 2 };                      // Riemann.hpp:10-14
enum Stepper { RK2 = 0, SSPRK3 = 1, SSPRK4 = 2 };                    // TimeStepping.hpp:10-14
enum Bc { WALL = 0, PERIODIC, SYMMETRIC, DIODE, ZEROGRAD, CONSTANT, USERFN };  // State.hpp:62-70

inline double sq(double x) { return x * x; }  // Types.hpp:106-109

// Prim/Cons index maps, State.hpp:10-60: Rho=0, V/Mom(axis)=1+axis, Pres/Ene=1+nd.
inline int i_pres(int nd) { return 1 + nd; }

// sound_speed, Eos.hpp:54-59
inline double sound_speed(double gamma, const double* w, int nd) {
    return std::sqrt(gamma * w[i_pres(nd)] / w[0]);
}

// cons_to_prim, Eos.hpp:61-79
inline void cons_to_prim(double gamma, const double* q, double* w, int nd) {
    w[0] = q[0];
    w[1] = q[1] / q[0];
    double v2 = sq(w[1]);
    if (nd > 1) {
        w[2] = q[2] / q[0];
        v2 += sq(w[2]);
    }
    if (nd > 2) {
        w[3] = q[3] / q[0];
        v2 += sq(w[3]);
    }
    const double e_kin = 0.5 * q[0] * v2;
    w[i_pres(nd)] = (gamma - 1.0) * ((q[i_pres(nd)] - e_kin));
}

// prim_to_cons, Eos.hpp:81-99
inline void prim_to_cons(double gamma, const double* w, double* q, int nd) {
    q[0] = w[0];
    q[1] = w[0] * w[1];
    double v2 = sq(w[1]);
    if (nd > 1) {
        q[2] = w[0] * w[2];
        v2 += sq(w[2]);
    }
    if (nd > 2) {
        q[3] = w[0] * w[3];
        v2 += sq(w[3]);
    }
    const double e_kin = 0.5 * w[0] * v2;
    const double e_int = w[i_pres(nd)] / (gamma - 1.0);
    q[i_pres(nd)] = e_int + e_kin;
}

// prim_to_flux, Eos.hpp:101-131
inline void prim_to_flux(double gamma, const double* w, double* f, int axis, int nd) {
    const int iv = 1 + axis;
    const double mass_flux = w[0] * w[iv];
    double e_kin = 0.0;
    f[0] = mass_flux;
    f[1] = mass_flux * w[1];
    e_kin += sq(w[1]);
    if (nd > 1) {
        f[2] = mass_flux * w[2];
        e_kin += sq(w[2]);
    }
    if (nd > 2) {
        f[3] = mass_flux * w[3];
        e_kin += sq(w[3]);
    }
    e_kin *= 0.5 * w[0];
    f[iv] += w[i_pres(nd)];
    const double e_tot = w[i_pres(nd)] / (gamma - 1.0) + e_kin;
    f[i_pres(nd)] = (e_tot + w[i_pres(nd)]) * w[iv];
}

// slope_limiter, Reconstruct.hpp:130-148
inline double slope_limiter(int sl, double um, double up) {
    using std::copysign;
    if (sl == VANLEER) {
        return (um * up > 0.0) ? 2.0 * um * up / (um + up) : 0.0;
    } else if (sl == MONOTONIZED_CENTRAL) {
        return (copysign(1.0, um) + copysign(1.0, up)) *
               std::min(std::abs(um), std::min(0.25 * std::abs(um + up), std::abs(up)));
    } else if (sl == MINMOD) {
        return 0.5 * (copysign(1.0, um) + copysign(1.0, up)) * std::min(std::abs(um), std::abs(up));
    }
    return 0.0;
}

// reconstruct<...>, Reconstruct.hpp:150-274. `s` points at the centre of a
// 5-point stencil (s[-2..2]) gathered along the axis (Stencil::fill 95-123).
inline void reconstruct(int recon, int sl, const double* s, double& wL, double& wR) {
    if (recon == FOG) {  // 150-154
        wL = s[0];
        wR = s[0];
    } else if (recon == MUSCL) {  // 156-167
        const double dwL = s[0] - s[-1];
        const double dwR = s[1] - s[0];
        const double delta = slope_limiter(sl, dwL, dwR);
        wL = s[0] - 0.5 * delta;
        wR = s[0] + 0.5 * delta;
    } else if (recon == PPM) {  // 169-215 (do_clamp is false)
        auto limited_slope = [&](int c) { return slope_limiter(sl, s[c] - s[c - 1], s[c + 1] - s[c]); };
        const double dw_m = limited_slope(-1);
        const double dw_0 = limited_slope(0);
        const double dw_p = limited_slope(1);
        wL = 0.5 * (s[-1] + s[0]) - (1.0 / 6.0) * (dw_0 - dw_m);
        wR = 0.5 * (s[0] + s[1]) - (1.0 / 6.0) * (dw_p - dw_0);
        // three sequential tests, not else-if (204-214)
        if ((wR - s[0]) * (s[0] - wL) <= 0.0) {
            wL = s[0];
            wR = s[0];
        }
        if (std::abs(wR - s[0]) >= 2.0 * std::abs(wL - s[0])) {
            wR = 3.0 * s[0] - 2.0 * wL;
        }
        if (std::abs(wL - s[0]) >= 2.0 * std::abs(wR - s[0])) {
            wL = 3.0 * s[0] - 2.0 * wR;
        }
    } else {  // WENO5-Z, 217-274 (pinned by worked example)
        const double c0 = 13.0 / 12.0, c1 = 0.25;
        const double beta0 = (c0 * sq(s[-2] - 2.0 * s[-1] + s[0]) + c1 * sq(s[-2] - 4.0 * s[-1] + 3.0 * s[0]));
        const double beta1 = (c0 * sq(s[-1] - 2.0 * s[0] + s[1]) + c1 * sq(s[-1] - s[1]));
        const double beta2 = (c0 * sq(s[0] - 2.0 * s[1] + s[2]) + c1 * sq(3.0 * s[0] - 4.0 * s[1] + s[2]));
        const double eps = 1e-36;
        const double tau5 = std::abs(beta0 - beta2);
        const double ind0 = tau5 / (beta0 + eps);
        const double ind1 = tau5 / (beta1 + eps);
        const double ind2 = tau5 / (beta2 + eps);

        double eno0 = (2.0 * s[0] + 5.0 * s[-1] - s[-2]);
        double eno1 = (-s[1] + 5.0 * s[0] + 2.0 * s[-1]);
        double eno2 = (2.0 * s[2] - 7.0 * s[1] + 11.0 * s[0]);
        double a0 = 3.0 * (1.0 + sq(ind0));
        double a1 = 6.0 * (1.0 + sq(ind1));
        double a2 = 1.0 * (1.0 + sq(ind2));
        double denom = 6.0 * (a0 + a1 + a2);
        wL = (eno0 * a0 + eno1 * a1 + eno2 * a2) / denom;

        eno0 = (2.0 * s[0] + 5.0 * s[1] - s[2]);
        eno1 = (-s[-1] + 5.0 * s[0] + 2.0 * s[1]);
        eno2 = (2.0 * s[-2] - 7.0 * s[-1] + 11.0 * s[0]);
        a0 = 3.0 * (1.0 + sq(ind2));
        a1 = 6.0 * (1.0 + sq(ind1));
        a2 = 1.0 * (1.0 + sq(ind0));
        denom = 6.0 * (a0 + a1 + a2);
        wR = (eno0 * a0 + eno1 * a1 + eno2 * a2) / denom;
    }
}

// riemann_flux<...>, Riemann.hpp:22-176. wL = RR of the left cell, wR = RL of
// the right cell (Hydro.cpp:138-144). Only the n_hydro = 2+nd entries are used.
inline void riemann_flux(int solver, double gamma, const double* wL, const double* wR, double* flux,
                         int axis, int nd) {
    const int n_hydro = 2 + nd;
    const int iv = 1 + axis;
    const int im = 1 + axis;
    const int ie = i_pres(nd);
    double qL[5], qR[5], fL[5], fR[5];
    if (solver == RUSANOV) {  // 22-46
        const double vL = wL[iv];
        const double vR = wR[iv];
        prim_to_cons(gamma, wL, qL, nd);
        prim_to_cons(gamma, wR, qR, nd);
        prim_to_flux(gamma, wL, fL, axis, nd);
        prim_to_flux(gamma, wR, fR, axis, nd);
        const double csL = sound_speed(gamma, wL, nd);
        const double csR = sound_speed(gamma, wR, nd);
        const double max_c = 0.5 * (csL + std::abs(vL) + csR + std::abs(vR));
        for (int i = 0; i < n_hydro; ++i) {
            flux[i] = 0.5 * (fL[i] + fR[i] - max_c * (qR[i] - qL[i]));
        }
    } else if (solver == HLL) {  // 48-73, including the sR wave-speed quirk of line 60
        const double csL = sound_speed(gamma, wL, nd);
        const double csR = sound_speed(gamma, wR, nd);
        const double tiny = 1e-20;
        const double sL = std::min(-tiny, std::min(wL[iv] - csL, wR[iv] - csR));
        const double sR = std::max(-tiny, std::max(wL[iv] + csL, wL[iv] + csR));
        const double sM = 1.0 / (sR - sL);
        prim_to_cons(gamma, wL, qL, nd);
        prim_to_cons(gamma, wR, qR, nd);
        prim_to_flux(gamma, wL, fL, axis, nd);
        prim_to_flux(gamma, wR, fR, axis, nd);
        for (int i = 0; i < n_hydro; ++i) {
            flux[i] = sM * (sR * fL[i] - sL * fR[i] + sR * sL * (qR[i] - qL[i]));
        }
    } else {  // HLLC 75-176 (pinned by worked example)
        const double csL = sound_speed(gamma, wL, nd);
        const double csR = sound_speed(gamma, wR, nd);
        const double tiny = 1e-20;
        prim_to_cons(gamma, wL, qL, nd);
        prim_to_cons(gamma, wR, qR, nd);
        const double pL = wL[ie];
        const double pR = wR[ie];
        const double rho_avg = 0.5 * (wL[0] + wR[0]);
        const double cs_avg = 0.5 * (csL + csR);
        const double p_mid = 0.5 * (pL + pR + (wL[iv] - wR[iv]) * rho_avg * cs_avg);
        const double cLstar = (p_mid <= wL[ie]) ? 1.0
                              : std::sqrt(1.0 + (gamma + 1.0) / (2.0 * gamma) * (p_mid / wL[ie] - 1.0));
        const double cRstar = (p_mid <= wR[ie]) ? 1.0
                              : std::sqrt(1.0 + (gamma + 1.0) / (2.0 * gamma) * (p_mid / wR[ie] - 1.0));
        const double aL = wL[iv] - csL * cLstar;
        const double aR = wR[iv] + csR * cRstar;
        const double bM = std::min(-tiny, aL);
        const double bP = std::max(tiny, aR);
        double vxL = wL[iv] - aL;
        double vxR = wR[iv] - aR;
        const double tL = wL[ie] + vxL * qL[im];
        const double tR = wR[ie] + vxR * qR[im];
        const double mL = wL[0] * vxL;
        const double mR = -(wR[0] * vxR);
        const double aM = (tL - tR) / (mL + mR);
        const double cp = std::max((mL * tR + mR * tL) / (mL + mR), 0.0);
        vxL = wL[iv] - bM;
        vxR = wR[iv] - bP;
        const double mfL = wL[0] * vxL;
        const double mfR = wR[0] * vxR;
        fL[0] = mfL;
        fR[0] = mfR;
        fL[1] = mfL * wL[1];
        fR[1] = mfR * wR[1];
        if (nd > 1) {
            fL[2] = mfL * wL[2];
            fR[2] = mfR * wR[2];
        }
        if (nd > 2) {
            fL[3] = mfL * wL[3];
            fR[3] = mfR * wR[3];
        }
        fL[im] += wL[ie];
        fR[im] += wR[ie];
        fL[ie] = qL[ie] * vxL + wL[ie] * wL[iv];
        fR[ie] = qR[ie] * vxR + wR[ie] * wR[iv];
        double sL, sR, sM;
        if (aM >= 0.0) {
            sL = aM / (aM - bM);
            sR = 0.0;
            sM = -bM / (aM - bM);
        } else {
            sL = 0.0;
            sR = -aM / (bP - aM);
            sM = bP / (bP - aM);
        }
        for (int v = 0; v < n_hydro; ++v) {
            flux[v] = sL * fL[v] + sR * fR[v];
        }
        flux[im] += sM * cp;
        flux[ie] += sM * cp * aM;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Simulation state: the subset of Simulation (Simulation.hpp:38-62) on the path.
// ---------------------------------------------------------------------------
extern "C" {

typedef struct {
    int32_t num_dim;
    int32_t num_tracers;
    int32_t xc, yc, zc, ng;  // GridSize, State.hpp:101-106 (counts include ghosts)
    double dx;
    double loc[3];           // GridLoc, State.hpp:108-112
    int32_t bc[6];           // xs xe ys ye zs ze, State.hpp:82-89
    double bc_const[6][5];   // State.hpp:92-97
    int32_t recon, limiter, solver, stepper;  // NumericalSchemes, Simulation.hpp:17-22
    double gamma;
    int32_t gravity;                          // a registered gravity source (Gravity.cpp:43-70)
    double grav[3];                           // sources.gravity.{x,y,z}
} orc_config;

}  // extern "C"

// TabulatedLteH (TabulatedLteH.hpp:10-28): two log-spaced axes and the y and
// log10 T tables, [n_rho][n_eint_rho].
struct orc_tab {
    double ex_start, ex_step;  // log10(eint / rho)
    int32_t ex_num;
    double r_start, r_step;    // log10(rho)
    int32_t r_num;
    std::vector<double> y, log_T;
};

struct orc_sim {
    orc_config c;
    int nvar;
    size_t cells;
    std::vector<double> Q, W, RL, RR, Fx, Fy, Fz, S, Q_old;
    int64_t current_step = 0;
    bool tab_on = false;               // eos.type = tabulatedlteh: update_eos before each stage
    orc_tab tab;
    std::vector<double> y_space, T_space;
    int eint_negative = 0;             // the reference aborts ("eint negative")

    double* at(std::vector<double>& a, int v, int k, int j, int i) {
        return &a[(((size_t)v * c.zc + k) * c.yc + j) * c.xc + i];
    }
};

namespace {

inline size_t lin(const orc_sim& s, int v, int k, int j, int i) {
    return (((size_t)v * s.c.zc + k) * s.c.yc + j) * s.c.xc + i;
}

// global_cons_to_prim_impl, Hydro.cpp:8-37 (over all cells incl. ghosts)
void global_cons_to_prim(orc_sim& s) {
    prof::Scope timer(prof::QtoW);
    const int nd = s.c.num_dim, n_hydro = 2 + nd, nv = s.nvar;
    const int64_t n = (int64_t)s.c.zc * s.c.yc * s.c.xc;
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < n; ++idx) {
        double q[16], w[16];
        for (int v = 0; v < nv; ++v) q[v] = s.Q[(size_t)v * n + idx];
        cons_to_prim(s.c.gamma, q, w, nd);
        const double inv_rho = 1.0 / q[0];
        for (int v = n_hydro; v < nv; ++v) w[v] = q[v] * inv_rho;
        for (int v = 0; v < n_hydro; ++v) s.W[(size_t)v * n + idx] = w[v];
        for (int v = n_hydro; v < nv; ++v) s.W[(size_t)v * n + idx] = w[v];
    }
}

// compute_recon_impl, Hydro.cpp:55-102
void compute_recon(orc_sim& s, int axis) {
    prof::Scope timer(prof::ReconX + 2 * axis);
    const auto& c = s.c;
    int nx = c.xc - 2 * c.ng;
    int ny = std::max(c.yc - 2 * c.ng, 1);
    int nz = std::max(c.zc - 2 * c.ng, 1);
    int xs = c.ng;
    int ys = ny == 1 ? 0 : c.ng;
    int zs = nz == 1 ? 0 : c.ng;
    if (axis == 0) { nx += 2; xs -= 1; }
    else if (axis == 1) { ny += 2; ys -= 1; }
    else { nz += 2; zs -= 1; }
    const int64_t n = (int64_t)nz * ny * nx;
    const int64_t stride = axis == 0 ? 1 : axis == 1 ? c.xc : (int64_t)c.xc * c.yc;
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < n; ++idx) {
        const int ii = idx % nx, ji = (idx / nx) % ny, ki = idx / ((int64_t)nx * ny);
        const int i = ii + xs, j = ji + ys, k = ki + zs;
        for (int v = 0; v < s.nvar; ++v) {
            const size_t base = lin(s, v, k, j, i);
            double st[5];
            // Stencil<Order>::fill gathers only 2*Order+1 points; gather the
            // in-range ones for the scheme in use.
            const int order = (c.recon == FOG) ? 0 : (c.recon == MUSCL) ? 1 : 2;
            for (int o = -2; o <= 2; ++o) {
                st[o + 2] = (o >= -order && o <= order) ? s.W[base + o * stride] : 0.0;
            }
            reconstruct(c.recon, c.limiter, st + 2, s.RL[base], s.RR[base]);
        }
    }
}

// compute_flux_impl, Hydro.cpp:104-166
void compute_flux(orc_sim& s, int axis) {
    prof::Scope timer(prof::FluxX + 2 * axis);
    const auto& c = s.c;
    const int nd = c.num_dim, n_hydro = 2 + nd;
    int nx = c.xc - 2 * c.ng;
    int ny = std::max(c.yc - 2 * c.ng, 1);
    int nz = std::max(c.zc - 2 * c.ng, 1);
    int xs = c.ng;
    int ys = ny == 1 ? 0 : c.ng;
    int zs = nz == 1 ? 0 : c.ng;
    std::vector<double>* F;
    if (axis == 0) { nx += 1; F = &s.Fx; }
    else if (axis == 1) { ny += 1; F = &s.Fy; }
    else { nz += 1; F = &s.Fz; }
    const int64_t n = (int64_t)nz * ny * nx;
    const int64_t stride = axis == 0 ? 1 : axis == 1 ? c.xc : (int64_t)c.xc * c.yc;
    const size_t vstride = (size_t)c.zc * c.yc * c.xc;
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < n; ++idx) {
        const int ii = idx % nx, ji = (idx / nx) % ny, ki = idx / ((int64_t)nx * ny);
        const int i = ii + xs, j = ji + ys, k = ki + zs;
        const size_t cell = lin(s, 0, k, j, i);
        const size_t cellm = cell - stride;
        double wl[16], wr[16], f[16];
        for (int v = 0; v < s.nvar; ++v) {
            wl[v] = s.RR[cellm + v * vstride];
            wr[v] = s.RL[cell + v * vstride];
        }
        riemann_flux(c.solver, c.gamma, wl, wr, f, axis, nd);
        for (int v = 0; v < n_hydro; ++v) (*F)[cell + v * vstride] = f[v];
        for (int v = n_hydro; v < s.nvar; ++v) {
            // upwind tracer flux, Hydro.cpp:155-162
            if (f[0] >= 0.0) {
                (*F)[cell + v * vstride] = f[0] * wl[v];
            } else {
                (*F)[cell + v * vstride] = f[0] * wr[v];
            }
        }
    }
}

// TabulatedLteH::sample (TabulatedLteH.hpp:57-118): bilinear interpolation
// in (log10 eint/rho, log10 rho), clamped to the table's edge outside it. The
// floor-based index is formed only inside the table, so a NaN coordinate (a
// state that already went NaN, where the reference's int conversion is
// undefined) takes the first edge instead of indexing out of the table; every
// other input gives the reference's indices.
void tab_sample(const orc_tab& t, double log_eint_rho, double log_rho, double& y, double& log_T) {
    const double frac_x = (log_eint_rho - t.ex_start) / t.ex_step;
    int ix;
    int ixp;
    double tx, txp;
    const int max_x = t.ex_num - 1;
    const int max_y = t.r_num - 1;
    if (!(frac_x >= 0.0) || frac_x >= max_x) {  // outside the table (or NaN): its edge
        ix = frac_x >= max_x ? max_x : 0;
        ixp = ix;
        tx = 1.0;
        txp = 0.0;
    } else {
        ix = int(std::floor(frac_x));
        ixp = ix + 1;
        txp = frac_x - ix;
        tx = 1.0 - txp;
    }
    const double frac_y = (log_rho - t.r_start) / t.r_step;
    int iy;
    int iyp;
    double ty, typ;
    if (!(frac_y >= 0.0) || frac_y >= max_y) {
        iy = frac_y >= max_y ? max_y : 0;
        iyp = iy;
        ty = 1.0;
        typ = 0.0;
    } else {
        iy = int(std::floor(frac_y));
        iyp = iy + 1;
        typ = frac_y - iy;
        ty = 1.0 - typ;
    }
    auto at = [&](const std::vector<double>& a, int r, int x) { return a[(size_t)r * t.ex_num + x]; };
    y = ty * (tx * at(t.y, iy, ix) + txp * at(t.y, iy, ixp)) + typ * (tx * at(t.y, iyp, ix) + txp * at(t.y, iyp, ixp));
    log_T = ty * (tx * at(t.log_T, iy, ix) + txp * at(t.log_T, iy, ixp)) +
            typ * (tx * at(t.log_T, iyp, ix) + txp * at(t.log_T, iyp, ixp));
}

// TabulatedLteH::update_eos (TabulatedLteH.hpp:120-161): per cell (ghosts
// included) the table's y and T for the cell's density and internal energy,
// and E += rho chi_H / m_H (y - y_prev) once y_prev has been set (the fields
// start at -1, Eos.cpp:80).
void update_eos_tab(orc_sim& s) {
    prof::Scope timer(prof::UpdateEos);
    constexpr double h_mass = 1.6737830080950003e-27, chi_H = 2.178710282685096e-18;
    const int nd = s.c.num_dim;
    const size_t n = s.cells;
    double* Q = s.Q.data();
    int neg = 0;
#pragma omp parallel for schedule(static) reduction(| : neg)
    for (size_t c = 0; c < n; ++c) {
        double mom2_sum = sq(Q[n + c]);
        if (nd > 1) mom2_sum += sq(Q[2 * n + c]);
        if (nd > 2) mom2_sum += sq(Q[3 * n + c]);
        const double rho = Q[c];
        const double e_kin = 0.5 * mom2_sum / rho;
        const double eint = Q[(1 + nd) * n + c] - e_kin;
        if (eint < 0.0) {
            neg = 1;
            continue;
        }
        const double log_rho = std::log10(rho);
        const double log_eint_rho = std::log10(eint) - log_rho;
        double y, log_T;
        tab_sample(s.tab, log_eint_rho, log_rho, y, log_T);
        const double prev_y = s.y_space[c];
        s.y_space[c] = y;
        const double delta_eint = (prev_y < 0.0) ? 0.0 : rho * (chi_H / h_mass) * (y - prev_y);
        s.T_space[c] = std::pow(10.0, log_T);
        Q[(1 + nd) * n + c] += delta_eint;
    }
    s.eint_negative |= neg;
}

// compute_hydro_fluxes, Hydro.cpp:325-343 (update_eos only for the tabulated EOS here)
void compute_hydro_fluxes(orc_sim& s) {
    if (s.tab_on) update_eos_tab(s);
    global_cons_to_prim(s);
    compute_recon(s, 0);
    compute_flux(s, 0);
    if (s.c.num_dim > 1) {
        compute_recon(s, 1);
        compute_flux(s, 1);
    }
    if (s.c.num_dim > 2) {
        compute_recon(s, 2);
        compute_flux(s, 2);
    }
}

// fill_one_bc_impl<Axis>, Boundaries.hpp:11-119 (one launch per axis). The
// per-axis pass is sequential over its flat range here; sources are interior
// cells only, so the order inside a pass does not matter.
void fill_one_bc(orc_sim& s, int axis) {
    const auto& c = s.c;
    const int nd = c.num_dim;
    const int ng = c.ng;
    const int dims[3] = {c.xc, c.yc, c.zc};
    int ld[3] = {c.xc, c.yc, c.zc};
    ld[axis] = 2 * ng;
    const int im = 1 + axis;
    const int ie = 1 + nd;
    const int64_t n = (int64_t)ld[0] * ld[1] * ld[2];
    const size_t vstride = (size_t)c.zc * c.yc * c.xc;
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < n; ++idx) {
        const int ii = idx % ld[0], ji = (idx / ld[0]) % ld[1], ki = idx / ((int64_t)ld[0] * ld[1]);
        int coord[3] = {ii, ji, ki};
        const int pencil_idx = coord[axis];
        int cflip = (2 * ng - 1) - coord[axis];
        int cedge = ng;
        if (pencil_idx >= ng) {
            coord[axis] = (dims[axis] - 1) - (pencil_idx - ng);
            cflip = (dims[axis] - 1) - (2 * ng - 1) + (pencil_idx - ng);
            cedge = (dims[axis] - 1) - ng;
        }
        int c_flip[3] = {coord[0], coord[1], coord[2]};
        c_flip[axis] = cflip;
        int c_per[3] = {coord[0], coord[1], coord[2]};
        c_per[axis] += (pencil_idx < ng ? 1 : -1) * (dims[axis] - 2 * ng);
        int c_edge[3] = {coord[0], coord[1], coord[2]};
        c_edge[axis] = cedge;
        const size_t dst = lin(s, 0, coord[2], coord[1], coord[0]);
        const size_t src_flip = lin(s, 0, c_flip[2], c_flip[1], c_flip[0]);
        const size_t src_per = lin(s, 0, c_per[2], c_per[1], c_per[0]);
        const size_t src_edge = lin(s, 0, c_edge[2], c_edge[1], c_edge[0]);
        const bool start = (coord[axis] < ng);
        const int bound = start ? c.bc[2 * axis] : c.bc[2 * axis + 1];
        const double* const_vals = start ? c.bc_const[2 * axis] : c.bc_const[2 * axis + 1];
        double* Q = s.Q.data();
        if (bound == WALL) {
            for (int v = 0; v < s.nvar; ++v) Q[dst + v * vstride] = Q[src_flip + v * vstride];
            Q[dst + im * vstride] = -Q[dst + im * vstride];
        } else if (bound == PERIODIC) {
            for (int v = 0; v < s.nvar; ++v) Q[dst + v * vstride] = Q[src_per + v * vstride];
        } else if (bound == SYMMETRIC) {
            for (int v = 0; v < s.nvar; ++v) Q[dst + v * vstride] = Q[src_flip + v * vstride];
        } else if (bound == DIODE) {
            for (int v = 0; v < s.nvar; ++v) Q[dst + v * vstride] = Q[src_flip + v * vstride];
            const double mom_pre = Q[dst + im * vstride];
            if (start) {
                Q[dst + im * vstride] = std::min(mom_pre, 0.0);
            } else {
                Q[dst + im * vstride] = std::max(mom_pre, 0.0);
            }
            Q[dst + ie * vstride] -= sq(Q[dst + im * vstride] - mom_pre) / Q[dst];
        } else if (bound == ZEROGRAD) {
            for (int v = 0; v < s.nvar; ++v) Q[dst + v * vstride] = Q[src_edge + v * vstride];
        } else if (bound == CONSTANT) {
            for (int v = 0; v < s.nvar; ++v) Q[dst + v * vstride] = const_vals[v];
        }
    }
}

// fill_bcs_impl / fill_bcs, Boundaries.hpp:121-149 (no user_bc)
void fill_bcs(orc_sim& s) {
    prof::Scope timer(prof::FillBcs);
    fill_one_bc(s, 0);
    if (s.c.num_dim > 1) fill_one_bc(s, 1);
    if (s.c.num_dim > 2) fill_one_bc(s, 2);
}

// gravity_kernel<NumDim>, Gravity.cpp:14-40: S accumulates the gravity source
// of the current Q on every cell (S was zeroed before the stage,
// SourceTerms.hpp:8-10).
void compute_source_terms(orc_sim& s) {
    prof::Scope timer(prof::Sources);
    const auto& c = s.c;
    if (!c.gravity) return;
    const int nd = c.num_dim;
    const int64_t n = (int64_t)s.cells;
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < n; ++idx) {
        const double rho = s.Q[idx];
        s.S[1 * n + idx] += rho * c.grav[0];
        double energy_update = s.Q[1 * n + idx] * c.grav[0];
        if (nd > 1) {
            s.S[2 * n + idx] += rho * c.grav[1];
            energy_update += s.Q[2 * n + idx] * c.grav[1];
        }
        if (nd > 2) {
            s.S[3 * n + idx] += rho * c.grav[2];
            energy_update += s.Q[3 * n + idx] * c.grav[2];
        }
        s.S[(1 + nd) * n + idx] += energy_update;
    }
}

// integrate_flux<NumDim>, TimeStepping.cpp:9-35, with the stage update
// selected by (stepper, stage); the source term S (zero unless gravity is
// registered) enters as S*dt for the hydro variables.
void integrate_flux(orc_sim& s, int stepper, int stage, double dt) {
    prof::Scope timer(prof::Integrate);
    const auto& c = s.c;
    const int nd = c.num_dim, n_hydro = 2 + nd;
    const int nx = c.xc - 2 * c.ng;
    const int ny = std::max(c.yc - 2 * c.ng, 1);
    const int nz = std::max(c.zc - 2 * c.ng, 1);
    const int64_t n = (int64_t)nz * ny * nx;
    const int64_t sy = c.xc, sz = (int64_t)c.xc * c.yc;
    const size_t vstride = (size_t)c.zc * c.yc * c.xc;
    const double dx = c.dx;
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < n; ++idx) {
        const int ii = idx % nx, ji = (idx / nx) % ny, ki = idx / ((int64_t)nx * ny);
        const int k = nz == 1 ? ki : ki + c.ng;
        const int j = ny == 1 ? ji : ji + c.ng;
        const int i = ii + c.ng;
        for (int var = 0; var < s.nvar; ++var) {
            const size_t cell = lin(s, var, k, j, i);
            double q_update = 0.0;
            q_update += s.Fx[cell] - s.Fx[cell + 1];
            if (nd > 1) q_update += s.Fy[cell] - s.Fy[cell + sy];
            if (nd > 2) q_update += s.Fz[cell] - s.Fz[cell + sz];
            // the updater lambdas
            q_update *= dt / dx;
            double source = 0.0;
            if (var < n_hydro) source = s.S[cell] * dt;
            double& Q = s.Q[cell];
            const double Qo = s.Q_old[cell];
            if (stepper == RK2) {  // TimeStepping.cpp:60-67, 80-87
                if (stage == 0) Q += q_update + source;
                else Q = 0.5 * (Qo + Q + q_update + source);
            } else if (stepper == SSPRK3) {  // 137-144, 157-164, 177-184
                if (stage == 0) Q += q_update + source;
                else if (stage == 1) Q = 0.75 * Qo + 0.25 * (Q + q_update + source);
                else Q = (1.0 / 3.0) * Qo + (2.0 / 3.0) * (Q + q_update + source);
            } else {  // SSPRK(4,3), 234-241, 254-261, 274-281, 294-301
                if (stage == 0 || stage == 1 || stage == 3) Q += 0.5 * (q_update + source);
                else Q = (2.0 / 3.0) * Qo + (1.0 / 3.0) * Q + (1.0 / 6.0) * (q_update + source);
            }
        }
    }
}

int num_stages(int stepper) { return stepper == RK2 ? 2 : stepper == SSPRK3 ? 3 : 4; }

}  // namespace

extern "C" {

orc_sim* orc_create(const orc_config* cfg) {
    orc_sim* s = new orc_sim();
    s->c = *cfg;
    s->nvar = 2 + cfg->num_dim + cfg->num_tracers;
    s->cells = (size_t)cfg->zc * cfg->yc * cfg->xc;
    const size_t n = s->cells * s->nvar;
    s->Q.assign(n, 0.0);
    s->W.assign(n, 0.0);
    s->RL.assign(n, 0.0);
    s->RR.assign(n, 0.0);
    s->Fx.assign(n, 0.0);
    if (cfg->num_dim > 1) s->Fy.assign(n, 0.0);
    if (cfg->num_dim > 2) s->Fz.assign(n, 0.0);
    s->S.assign(s->cells * (2 + cfg->num_dim), 0.0);
    s->Q_old.assign(n, 0.0);
    return s;
}

void orc_destroy(orc_sim* s) { delete s; }

int64_t orc_num_values(const orc_sim* s) { return (int64_t)(s->cells * s->nvar); }

void orc_set_q(orc_sim* s, const double* Q) { std::memcpy(s->Q.data(), Q, s->Q.size() * sizeof(double)); }
void orc_get_q(const orc_sim* s, double* Q) { std::memcpy(Q, s->Q.data(), s->Q.size() * sizeof(double)); }
void orc_get_w(const orc_sim* s, double* W) { std::memcpy(W, s->W.data(), s->W.size() * sizeof(double)); }

void orc_fill_bcs(orc_sim* s) { fill_bcs(*s); }
void orc_cons_to_prim_all(orc_sim* s) { global_cons_to_prim(*s); }

// compute_dt_impl, Hydro.cpp:192-242 with dt_reducer. Kokkos::Min writes the
// reduction result into dt_max (the 1e5 seed only survives an empty range).
double orc_compute_dt(const orc_sim* s, double max_cfl, double time, double max_time,
                      double prev_output_time, double delta) {
    prof::Scope timer(prof::Dt);
    const auto& c = s->c;
    const int nd = c.num_dim;
    const int64_t n = (int64_t)c.zc * c.yc * c.xc;
    double dt_max = std::numeric_limits<double>::max();
    if (n == 0) dt_max = 1e5;
#pragma omp parallel for reduction(min : dt_max) schedule(static)
    for (int64_t idx = 0; idx < n; ++idx) {
        double q[5], w[5];
        for (int v = 0; v < 2 + nd; ++v) q[v] = s->Q[(size_t)v * n + idx];
        cons_to_prim(c.gamma, q, w, nd);
        const double cs = sound_speed(c.gamma, w, nd);
        double vel2 = sq(w[1]);
        if (nd > 1) vel2 += sq(w[2]);
        if (nd > 2) vel2 += sq(w[3]);
        const double vel = std::sqrt(vel2);
        const double dt_local = c.dx / (cs + vel);
        dt_max = std::min(dt_local, dt_max);
    }
    double dt = dt_max * max_cfl;
    const double next_write_time = std::min(max_time, prev_output_time + delta);
    if (time + dt >= next_write_time) {
        dt = next_write_time - time;
        while (time + dt < next_write_time) {
            dt = std::nextafter(dt, std::numeric_limits<double>::max());
        }
    }
    return dt;
}

// TimeStepper<...>::time_step<NumDim>, TimeStepping.cpp:39-308: stage s runs
// fill_bcs (s>0), compute_hydro_fluxes, integrate_flux; the step ends with
// fill_bcs. Q_old is a copy of Q at the start of the step.
void orc_time_step(orc_sim* s, double dt) {
    const int ns = num_stages(s->c.stepper);
    {
        prof::Scope timer(prof::QOld);
        s->Q_old = s->Q;
    }
    for (int st = 0; st < ns; ++st) {
        {
            prof::Scope timer(prof::Sources);
            std::fill(s->S.begin(), s->S.end(), 0.0);  // zero_source_terms, every stage
        }
        if (st > 0) fill_bcs(*s);
        compute_hydro_fluxes(*s);
        compute_source_terms(*s);
        integrate_flux(*s, s->c.stepper, st, dt);
    }
    fill_bcs(*s);
    s->current_step += 1;
}

// One stage only (for stage-level parity tests): Q_old must have been set by
// orc_begin_step. Does not fill ghosts before the stage.
void orc_begin_step(orc_sim* s) { s->Q_old = s->Q; }
void orc_stage(orc_sim* s, int stage, double dt) {
    std::fill(s->S.begin(), s->S.end(), 0.0);
    compute_hydro_fluxes(*s);
    compute_source_terms(*s);
    integrate_flux(*s, s->c.stepper, stage, dt);
}

// main.cpp:54-71 step loop: while time < max_time { dt; step; output bookkeeping }.
// prev_output_time starts at the time of the initial write (setup_sim writes
// at t=0, SetupModules.hpp:259 + Output.cpp:274). Stops after max_steps steps
// when max_steps >= 0. Returns the number of steps taken; *time is updated.
int64_t orc_run(orc_sim* s, double max_cfl, double* time, double max_time, double* prev_output_time,
                double delta, int64_t max_steps, double* dts_out) {
    int64_t steps = 0;
    while (*time < max_time && (max_steps < 0 || steps < max_steps)) {
        const double dt = orc_compute_dt(s, max_cfl, *time, max_time, *prev_output_time, delta);
        orc_time_step(s, dt);
        *time += dt;
        if (dts_out) dts_out[steps] = dt;
        ++steps;
        if (*time >= *prev_output_time + delta) {
            *prev_output_time = *time;
        }
    }
    return steps;
}

// ---- analytic LTE hydrogen EOS (AnalyticLteH.hpp:9-125) ---------------------
// The update_eos pass of eos.type = analyticlteh: per cell, the temperature T
// and ionisation fraction y of a pure-hydrogen gas in LTE for the cell's
// density and internal energy, by bisection on T (Saha equation for y). The
// hydro step does not read them; they are output fields. Q is [var][cells]
// with nd + 2 hydro variables. Returns 1 if a cell's temperature bracket is
// inverted (the reference aborts there), else 0.
namespace {
constexpr double kHMass = 1.6737830080950003e-27;  // [kg]
constexpr double kKB = 1.380649e-23;               // [J/K]
constexpr double kChiH = 2.178710282685096e-18;    // [J]

double saha_rhs_h(double temp) {  // AnalyticLteH.hpp:9-13
    return 2.4146830395719654e+21 * std::pow(temp, 1.5) * std::exp(-157763.42386247337 / temp);
}
double y_from_nhtot(double nhtot, double temp) {  // 15-19
    const double X = saha_rhs_h(temp);
    return 0.5 * (-X + std::sqrt(sq(X) + 4 * nhtot * X)) / nhtot;
}
}  // namespace

int orc_lte_eos(int nd, int64_t ncells, const double* Q, double gamma, double avg_mass, int ionisation_e, double* y_out,
                double* T_out) {
    const double inv_avg_mass = 1.0 / avg_mass;
    int flipped = 0;
#pragma omp parallel for schedule(static) reduction(| : flipped)
    for (int64_t c = 0; c < ncells; ++c) {
        const double rho = Q[c];
        double mom2_sum = sq(Q[ncells + c]);
        if (nd > 1) mom2_sum += sq(Q[2 * ncells + c]);
        if (nd > 2) mom2_sum += sq(Q[3 * ncells + c]);
        const double e_kin = 0.5 * mom2_sum / rho;
        const double eint = Q[(1 + nd) * ncells + c] - e_kin;
        const double e_to_T = (gamma - 1.0) / (rho * (kKB / kHMass) * inv_avg_mass);
        auto temp_from_y = [&](double y) {
            return e_to_T / (1.0 + y) * (eint - (double)ionisation_e * y * rho * (kChiH / kHMass) * inv_avg_mass);
        };
        const double rho_to_nhtot = (1.0 / kHMass) * inv_avg_mass;
        const double min_temperature = 100.0;
        const double tb0 = std::max(temp_from_y(1.0), min_temperature), tb1 = temp_from_y(0.0);
        if (tb0 > tb1) {
            flipped = 1;
            y_out[c] = T_out[c] = std::nan("");
            continue;
        }
        double temp_step = tb1 - tb0, temp = tb0, y = 0.0;
        for (int iter = 0; iter < 100; ++iter) {
            temp_step *= 0.5;
            const double test_temp = temp + temp_step;
            y = y_from_nhtot(rho * rho_to_nhtot, test_temp);
            const double temp_err = test_temp - temp_from_y(y);
            if (temp_err <= 0.0) temp = test_temp;
            if (std::abs(temp_step) < 1e-1 || std::abs(temp_err) < 1e-1) break;
        }
        y_out[c] = y;
        T_out[c] = temp;
    }
    return flipped;
}

// ---- tabulated LTE hydrogen EOS (TabulatedLteH.hpp, Eos.cpp:66-88) ----------
// Installs the table (y and log10 T, [n_rho][n_eint_rho]) and sets the y and
// T fields to -1; from then on every stage starts with update_eos.
void orc_set_tabulated_eos(orc_sim* s, double ex_start, double ex_step, int32_t ex_num, double r_start,
                           double r_step, int32_t r_num, const double* y_table, const double* log_T_table) {
    s->tab_on = true;
    s->tab.ex_start = ex_start;
    s->tab.ex_step = ex_step;
    s->tab.ex_num = ex_num;
    s->tab.r_start = r_start;
    s->tab.r_step = r_step;
    s->tab.r_num = r_num;
    const size_t m = (size_t)ex_num * r_num;
    s->tab.y.assign(y_table, y_table + m);
    s->tab.log_T.assign(log_T_table, log_T_table + m);
    s->y_space.assign(s->cells, -1.0);
    s->T_space.assign(s->cells, -1.0);
    s->eint_negative = 0;
}
// The fields of the last update_eos; returns 1 if a cell had eint < 0.
int orc_tabulated_fields(const orc_sim* s, double* y, double* T) {
    std::memcpy(y, s->y_space.data(), s->cells * sizeof(double));
    std::memcpy(T, s->T_space.data(), s->cells * sizeof(double));
    return s->eint_negative;
}
void orc_set_y_space(orc_sim* s, const double* y) { std::memcpy(s->y_space.data(), y, s->cells * sizeof(double)); }
void orc_update_eos(orc_sim* s) { update_eos_tab(*s); }
void orc_tab_sample(double ex_start, double ex_step, int32_t ex_num, double r_start, double r_step, int32_t r_num,
                    const double* y_table, const double* log_T_table, int64_t n, const double* log_eint_rho,
                    const double* log_rho, double* y, double* log_T) {
    orc_tab t{ex_start, ex_step, ex_num, r_start, r_step, r_num, {}, {}};
    const size_t m = (size_t)ex_num * r_num;
    t.y.assign(y_table, y_table + m);
    t.log_T.assign(log_T_table, log_T_table + m);
    for (int64_t i = 0; i < n; ++i) tab_sample(t, log_eint_rho[i], log_rho[i], y[i], log_T[i]);
}

// ---- per-kernel time breakdown ----------------------------------------------
int orc_profile_count() { return prof::N; }
const char* orc_profile_name(int k) { return (k >= 0 && k < prof::N) ? prof::kNames[k] : ""; }
double orc_profile_ms(int k) { return (k >= 0 && k < prof::N) ? prof::ms[k] : 0.0; }
void orc_profile_reset() {
    for (double& v : prof::ms) v = 0.0;
}

// ---- single-functor entry points for the unit tests ------------------------
void orc_cons_to_prim(int nd, double gamma, const double* q, double* w) { cons_to_prim(gamma, q, w, nd); }
void orc_prim_to_cons(int nd, double gamma, const double* w, double* q) { prim_to_cons(gamma, w, q, nd); }
void orc_prim_to_flux(int axis, int nd, double gamma, const double* w, double* f) {
    prim_to_flux(gamma, w, f, axis, nd);
}
double orc_sound_speed(int nd, double gamma, const double* w) { return sound_speed(gamma, w, nd); }
double orc_slope_limiter(int sl, double um, double up) { return slope_limiter(sl, um, up); }
// n stencils of 5 values each (s[-2..2]) -> n (wL, wR)
void orc_reconstruct(int recon, int sl, int64_t n, const double* stencils, double* wL, double* wR) {
    for (int64_t i = 0; i < n; ++i) reconstruct(recon, sl, stencils + 5 * i + 2, wL[i], wR[i]);
}
// n interfaces; wL, wR, flux are [n][2+nd]
void orc_riemann(int solver, int axis, int nd, double gamma, int64_t n, const double* wL, const double* wR,
                 double* flux) {
    const int nh = 2 + nd;
    for (int64_t i = 0; i < n; ++i) riemann_flux(solver, gamma, wL + nh * i, wR + nh * i, flux + nh * i, axis, nd);
}

}  // extern "C"
