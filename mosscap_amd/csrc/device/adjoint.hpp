// Discrete adjoint building blocks (FP64): vector-Jacobian products of the
// hydro functors, with dt frozen.
//
// Conventions at non-differentiable points follow the forward code's branches:
// derivatives flow through the branch the forward values select (ternaries,
// min/max keep their first operand on ties), and |x| has derivative sign(x)
// with sign(0) = 0.
//
// Two tools:
//   * Dual<N>: forward-mode dual numbers. Instantiating a functor of
//     functors.hpp on Dual<N> gives its exact Jacobian (same code path, same
//     branches); used for HLLC and to pin the hand-written VJPs.
//   * hand-written reverse-mode VJPs for cons_to_prim, WENO5-Z and Rusanov.
#pragma once

#include "functors.hpp"

namespace mh {

template <int N>
struct Dual {
    double v;
    double d[N];
    __host__ __device__ constexpr Dual() : v(0.0), d{} {}
    __host__ __device__ constexpr Dual(double x) : v(x), d{} {}  // NOLINT: constants promote
    __host__ __device__ constexpr Dual(int x) : v((double)x), d{} {}
    __host__ __device__ static Dual var(double x, int k) {
        Dual r(x);
        r.d[k] = 1.0;
        return r;
    }
};

template <int N>
__host__ __device__ constexpr Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
__host__ __device__ constexpr Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
__host__ __device__ constexpr Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v);
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N>
__host__ __device__ constexpr Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
__host__ __device__ constexpr Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    const double q = a.v / b.v;
    Dual<N> r(q);
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - q * b.d[k]) / b.v;
    return r;
}
template <int N>
__host__ __device__ constexpr Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { return a = a + b; }
template <int N>
__host__ __device__ constexpr Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) { return a = a - b; }
template <int N>
__host__ __device__ constexpr Dual<N>& operator*=(Dual<N>& a, const Dual<N>& b) { return a = a * b; }
template <int N>
__host__ __device__ constexpr bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N>
__host__ __device__ constexpr bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v; }
template <int N>
__host__ __device__ constexpr bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N>
__host__ __device__ constexpr bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }

template <int N>
__host__ __device__ inline Dual<N> sabs(Dual<N> a) {
    const double s = a.v > 0.0 ? 1.0 : (a.v < 0.0 ? -1.0 : 0.0);
    Dual<N> r(__builtin_fabs(a.v));
    for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k];
    return r;
}
template <int N>
__host__ __device__ inline Dual<N> sign1(Dual<N> a) {  // copysign(1, a): piecewise constant
    return Dual<N>(__builtin_copysign(1.0, a.v));
}

template <int N>
struct Arith<Dual<N>> {
    __host__ __device__ static inline Dual<N> div(const Dual<N>& a, const Dual<N>& b) { return a / b; }
    __host__ __device__ static inline Dual<N> sqrt(const Dual<N>& a) {
        const double s = __builtin_sqrt(a.v);
        Dual<N> r(s);
        for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / (2.0 * s);
        return r;
    }
    static constexpr double weno_eps = 1e-36;
};

// ---- hand-written reverse mode ---------------------------------------------------

// cons_to_prim (Eos.hpp:61-79), 2-D: w = (rho, mx/rho, my/rho, (g-1)(E - (rho/2)|v|^2)).
__host__ __device__ inline void cons_to_prim_vjp2(double gamma, const double (&q)[4], const double (&lw)[4],
                                                  double (&lq)[4]) {
    const double rho = q[0];
    const double ir = 1.0 / rho;
    const double vx = q[1] * ir, vy = q[2] * ir;
    const double gm1 = gamma - 1.0;
    lq[0] = lw[0] - (vx * lw[1] + vy * lw[2]) * ir + gm1 * 0.5 * (vx * vx + vy * vy) * lw[3];
    lq[1] = lw[1] * ir - gm1 * vx * lw[3];
    lq[2] = lw[2] * ir - gm1 * vy * lw[3];
    lq[3] = gm1 * lw[3];
}

// WENO5-Z (Reconstruct.hpp:217-274): given the cotangents lL, lR of the
// left/right face values of the cell, accumulate d(lL*wL + lR*wR)/ds into ls.
__host__ __device__ inline void weno5z_vjp(const double (&s)[5], double lL, double lR, double (&ls)[5]) {
    const double sm2 = s[0], sm1 = s[1], s0 = s[2], sp1 = s[3], sp2 = s[4];
    const double c0 = 13.0 / 12.0, c1 = 0.25, eps = 1e-36;
    const double a0 = sm2 - 2.0 * sm1 + s0, b0 = sm2 - 4.0 * sm1 + 3.0 * s0;
    const double a1 = sm1 - 2.0 * s0 + sp1, b1 = sm1 - sp1;
    const double a2 = s0 - 2.0 * sp1 + sp2, b2 = 3.0 * s0 - 4.0 * sp1 + sp2;
    const double be[3] = {c0 * a0 * a0 + c1 * b0 * b0, c0 * a1 * a1 + c1 * b1 * b1, c0 * a2 * a2 + c1 * b2 * b2};
    const double dtau = be[0] - be[2];
    const double tau = __builtin_fabs(dtau);
    double ind[3], q[3];
    for (int k = 0; k < 3; ++k) {
        ind[k] = tau / (be[k] + eps);
        q[k] = 1.0 + ind[k] * ind[k];
    }
    const double eL[3] = {2.0 * s0 + 5.0 * sm1 - sm2, -sp1 + 5.0 * s0 + 2.0 * sm1, 2.0 * sp2 - 7.0 * sp1 + 11.0 * s0};
    const double eR[3] = {2.0 * s0 + 5.0 * sp1 - sp2, -sm1 + 5.0 * s0 + 2.0 * sp1, 2.0 * sm2 - 7.0 * sm1 + 11.0 * s0};
    const double aL[3] = {3.0 * q[0], 6.0 * q[1], q[2]};
    const double aR[3] = {3.0 * q[2], 6.0 * q[1], q[0]};
    const double NL = eL[0] * aL[0] + eL[1] * aL[1] + eL[2] * aL[2], DL = 6.0 * (aL[0] + aL[1] + aL[2]);
    const double NR = eR[0] * aR[0] + eR[1] * aR[1] + eR[2] * aR[2], DR = 6.0 * (aR[0] + aR[1] + aR[2]);
    // wL = NL/DL, wR = NR/DR
    const double lNL = lL / DL, lDL = -lL * (NL / DL) / DL;
    const double lNR = lR / DR, lDR = -lR * (NR / DR) / DR;
    double laL[3], laR[3], leL[3], leR[3];
    for (int k = 0; k < 3; ++k) {
        laL[k] = lNL * eL[k] + 6.0 * lDL;
        laR[k] = lNR * eR[k] + 6.0 * lDR;
        leL[k] = lNL * aL[k];
        leR[k] = lNR * aR[k];
    }
    double lq[3];
    lq[0] = 3.0 * laL[0] + laR[2];
    lq[1] = 6.0 * (laL[1] + laR[1]);
    lq[2] = laL[2] + 3.0 * laR[0];
    double ltau = 0.0, lbe[3];
    for (int k = 0; k < 3; ++k) {
        const double lind = 2.0 * ind[k] * lq[k];
        const double r = 1.0 / (be[k] + eps);
        ltau += lind * r;
        lbe[k] = -lind * ind[k] * r;
    }
    const double sg = dtau > 0.0 ? 1.0 : (dtau < 0.0 ? -1.0 : 0.0);
    lbe[0] += sg * ltau;
    lbe[2] -= sg * ltau;
    const double la0 = 2.0 * c0 * a0 * lbe[0], lb0 = 2.0 * c1 * b0 * lbe[0];
    const double la1 = 2.0 * c0 * a1 * lbe[1], lb1 = 2.0 * c1 * b1 * lbe[1];
    const double la2 = 2.0 * c0 * a2 * lbe[2], lb2 = 2.0 * c1 * b2 * lbe[2];
    // stencil coefficients of a, b and the ENO candidates over (s-2, s-1, s0, s1, s2)
    ls[0] += la0 + lb0 - leL[0] + 2.0 * leR[2];
    ls[1] += -2.0 * la0 - 4.0 * lb0 + la1 + lb1 + 5.0 * leL[0] + 2.0 * leL[1] - leR[1] - 7.0 * leR[2];
    ls[2] += la0 + 3.0 * lb0 - 2.0 * la1 + la2 + 3.0 * lb2 + 2.0 * leL[0] + 5.0 * leL[1] + 11.0 * leL[2] +
             2.0 * leR[0] + 5.0 * leR[1] + 11.0 * leR[2];
    ls[3] += la1 - lb1 - 2.0 * la2 - 4.0 * lb2 - leL[1] - 7.0 * leL[2] + 5.0 * leR[0] + 2.0 * leR[1];
    ls[4] += la2 + lb2 + 2.0 * leL[2] - leR[0];
}

// Rusanov (Riemann.hpp:22-46), 2-D along Axis: accumulate J^T lf into lwL, lwR.
template <int Axis>
__host__ __device__ inline void rusanov_vjp2(double gamma, const double (&wL)[4], const double (&wR)[4],
                                             const double (&lf)[4], double (&lwL)[4], double (&lwR)[4]) {
    constexpr int IV = 1 + Axis;
    const double gm1 = gamma - 1.0;
    auto side = [&](const double (&w)[4], double sgn_q, double (&lw)[4], double& cs_out, double (&q)[4]) {
        const double rho = w[0], vx = w[1], vy = w[2], p = w[3];
        const double vn = w[IV];
        const double v2 = vx * vx + vy * vy;
        const double E = p / gm1 + 0.5 * rho * v2;
        q[0] = rho;
        q[1] = rho * vx;
        q[2] = rho * vy;
        q[3] = E;
        cs_out = __builtin_sqrt(gamma * p / rho);
        const double H = E + p;
        // 0.5 * J_f^T lf
        double g[4];
        g[0] = 0.5 * (vn * lf[0] + vn * vx * lf[1] + vn * vy * lf[2] + 0.5 * v2 * vn * lf[3]);
        g[1] = 0.5 * ((Axis == 0 ? rho : 0.0) * lf[0] + (rho * vn + (Axis == 0 ? rho * vx : 0.0)) * lf[1] +
                      (Axis == 0 ? rho * vy : 0.0) * lf[2] + (rho * vx * vn + (Axis == 0 ? H : 0.0)) * lf[3]);
        g[2] = 0.5 * ((Axis == 1 ? rho : 0.0) * lf[0] + (Axis == 1 ? rho * vx : 0.0) * lf[1] +
                      (rho * vn + (Axis == 1 ? rho * vy : 0.0)) * lf[2] + (rho * vy * vn + (Axis == 1 ? H : 0.0)) * lf[3]);
        g[3] = 0.5 * ((Axis == 0 ? 1.0 : 0.0) * lf[1] + (Axis == 1 ? 1.0 : 0.0) * lf[2] + (1.0 / gm1 + 1.0) * vn * lf[3]);
        // -/+ 0.5 * max_c * J_q^T lf (sgn_q = +1 on the left, -1 on the right)
        // (stored for the caller to scale by max_c)
        lw[0] = g[0];
        lw[1] = g[1];
        lw[2] = g[2];
        lw[3] = g[3];
        (void)sgn_q;
    };
    double qL[4], qR[4], csL, csR, gl[4], gr[4];
    side(wL, 1.0, gl, csL, qL);
    side(wR, -1.0, gr, csR, qR);
    const double max_c = 0.5 * (csL + __builtin_fabs(wL[IV]) + csR + __builtin_fabs(wR[IV]));
    // J_q^T lf for prim_to_cons
    auto jq = [&](const double (&w)[4], double (&o)[4]) {
        const double rho = w[0], vx = w[1], vy = w[2];
        o[0] = lf[0] + vx * lf[1] + vy * lf[2] + 0.5 * (vx * vx + vy * vy) * lf[3];
        o[1] = rho * lf[1] + rho * vx * lf[3];
        o[2] = rho * lf[2] + rho * vy * lf[3];
        o[3] = lf[3] / gm1;
    };
    double jqL[4], jqR[4];
    jq(wL, jqL);
    jq(wR, jqR);
    // d(max_c): 0.5 * (dcs + sign(v) dv) on each side
    double dq = 0.0;
    for (int k = 0; k < 4; ++k) dq += lf[k] * (qR[k] - qL[k]);
    const double lc = -0.5 * dq;  // cotangent of max_c
    auto dcs = [&](const double (&w)[4], double cs, double (&o)[4]) {
        // cs = sqrt(g p / rho): dcs/drho = -cs/(2 rho), dcs/dp = g / (2 rho cs)
        const double lcs = 0.5 * lc;
        o[0] += lcs * (-cs / (2.0 * w[0]));
        o[3] += lcs * (gamma / (2.0 * w[0] * cs));
        const double vn = w[IV];
        const double sg = vn > 0.0 ? 1.0 : (vn < 0.0 ? -1.0 : 0.0);
        o[IV] += lcs * sg;
    };
    for (int k = 0; k < 4; ++k) {
        lwL[k] += gl[k] + 0.5 * max_c * jqL[k];
        lwR[k] += gr[k] - 0.5 * max_c * jqR[k];
    }
    dcs(wL, csL, lwL);
    dcs(wR, csR, lwR);
}

// HLLC (Riemann.hpp:75-176), 2-D along Axis: the forward sequence of
// riemann_flux<Hllc> replayed, then its reverse sweep; J^T lf accumulated into
// lwL, lwR. Branches (wave-speed switches, the floors on bM/bP and cp) pass the
// derivative of the branch the forward values took.
template <int Axis>
__host__ __device__ inline void hllc_vjp2(double gamma, const double (&wL)[4], const double (&wR)[4],
                                          const double (&lf)[4], double (&lwL)[4], double (&lwR)[4]) {
    constexpr int IV = 1 + Axis;  // normal velocity (prim) = normal momentum (cons) index
    constexpr double tiny = 1e-20;
    const double gm1 = gamma - 1.0;
    const double rL = wL[0], rR = wR[0], pL = wL[3], pR = wR[3];
    const double uL = wL[IV], uR = wR[IV];
    // ---- forward (same operations as the functor) ----
    const double csL = __builtin_sqrt(gamma * pL / rL);
    const double csR = __builtin_sqrt(gamma * pR / rR);
    const double v2L = wL[1] * wL[1] + wL[2] * wL[2];
    const double v2R = wR[1] * wR[1] + wR[2] * wR[2];
    const double qLm = rL * uL, qRm = rR * uR;
    const double qLe = pL / gm1 + 0.5 * rL * v2L;
    const double qRe = pR / gm1 + 0.5 * rR * v2R;
    const double rho_avg = 0.5 * (rL + rR);
    const double cs_avg = 0.5 * (csL + csR);
    const double K = uL - uR;
    const double p_mid = 0.5 * (pL + pR + K * rho_avg * cs_avg);
    const double g_fac = (gamma + 1.0) / (2.0 * gamma);
    const bool lo_L = p_mid <= pL, lo_R = p_mid <= pR;
    const double cLs = lo_L ? 1.0 : __builtin_sqrt(1.0 + g_fac * (p_mid / pL - 1.0));
    const double cRs = lo_R ? 1.0 : __builtin_sqrt(1.0 + g_fac * (p_mid / pR - 1.0));
    const double aL = uL - csL * cLs;
    const double aR = uR + csR * cRs;
    const bool bM_is_aL = aL < -tiny, bP_is_aR = tiny < aR;
    const double bM = bM_is_aL ? aL : -tiny;
    const double bP = bP_is_aR ? aR : tiny;
    const double vL1 = uL - aL, vR1 = uR - aR;
    const double tL = pL + vL1 * qLm, tR = pR + vR1 * qRm;
    const double mL = rL * vL1, mR = -(rR * vR1);
    const double inv_m = mL + mR;
    const double aM = (tL - tR) / inv_m;
    const double cp0 = (mL * tR + mR * tL) / inv_m;
    const bool cp_pass = !(cp0 < 0.0);
    const double cp = cp_pass ? cp0 : 0.0;
    const double vL2 = uL - bM, vR2 = uR - bP;
    const double mfL = rL * vL2, mfR = rR * vR2;
    double fL[4] = {mfL, mfL * wL[1], mfL * wL[2], qLe * vL2 + pL * uL};
    double fR[4] = {mfR, mfR * wR[1], mfR * wR[2], qRe * vR2 + pR * uR};
    fL[IV] += pL;
    fR[IV] += pR;
    const bool pos = aM >= 0.0;
    const double d = pos ? aM - bM : bP - aM;
    const double sL = pos ? aM / d : 0.0;
    const double sR = pos ? 0.0 : -aM / d;
    const double sM = pos ? -bM / d : bP / d;
    // ---- reverse ----
    double gL[4] = {0.0, 0.0, 0.0, 0.0}, gR[4] = {0.0, 0.0, 0.0, 0.0};  // cotangents of wL, wR
    const double smcp = sM * cp;
    double l_sL = 0.0, l_sR = 0.0;
    for (int v = 0; v < 4; ++v) {
        l_sL += lf[v] * fL[v];
        l_sR += lf[v] * fR[v];
    }
    const double l_smcp = lf[IV] + lf[3] * aM;  // flux[IM] += sM*cp; flux[E] += (sM*cp)*aM
    const double l_sM = l_smcp * cp;
    const double l_cp = l_smcp * sM;
    double l_aM = lf[3] * smcp;
    double l_bM = 0.0, l_bP = 0.0;
    if (pos) {
        const double l_d = -(l_sL * sL + l_sM * sM) / d;
        l_aM += l_sL / d + l_d;
        l_bM += -l_sM / d - l_d;
    } else {
        const double l_d = -(l_sR * sR + l_sM * sM) / d;
        l_aM += -l_sR / d - l_d;
        l_bP += l_sM / d + l_d;
    }
    // fL, fR
    auto side_f = [&](const double (&w)[4], double s, double mf, double v2, double qe, double p, double u,
                      double (&g)[4], double& l_b, double& l_qe_out) {
        const double l0 = s * lf[0], l1 = s * lf[1], l2 = s * lf[2], l3 = s * lf[3];
        const double l_mf = l0 + l1 * w[1] + l2 * w[2];
        g[1] += l1 * mf;
        g[2] += l2 * mf;
        g[3] += s * lf[IV];  // f[IM] += p
        // f[E] = qe*v2 + p*u (v2 here: the flux velocity u - b)
        l_qe_out = l3 * v2;
        double l_v2 = l3 * qe;
        g[3] += l3 * u;
        g[IV] += l3 * p;
        // mf = rho * v2
        g[0] += l_mf * v2;
        l_v2 += l_mf * w[0];
        // v2 = u - b
        g[IV] += l_v2;
        l_b -= l_v2;
    };
    double l_qLe = 0.0, l_qRe = 0.0;
    if (sL != 0.0) side_f(wL, sL, mfL, vL2, qLe, pL, uL, gL, l_bM, l_qLe);
    if (sR != 0.0) side_f(wR, sR, mfR, vR2, qRe, pR, uR, gR, l_bP, l_qRe);
    // cp = max(cp0, 0); cp0 = (mL*tR + mR*tL) / inv_m; aM = (tL - tR) / inv_m
    double l_mL = 0.0, l_mR = 0.0, l_tL = 0.0, l_tR = 0.0, l_inv = 0.0;
    if (cp_pass) {
        const double l_n = l_cp / inv_m;
        l_inv -= l_cp * cp0 / inv_m;
        l_mL += l_n * tR;
        l_tR += l_n * mL;
        l_mR += l_n * tL;
        l_tL += l_n * mR;
    }
    l_tL += l_aM / inv_m;
    l_tR -= l_aM / inv_m;
    l_inv -= l_aM * aM / inv_m;
    l_mL += l_inv;
    l_mR += l_inv;
    // mR = -(rR*vR1), mL = rL*vL1, tX = pX + vX1*qXm
    double l_vR1 = -l_mR * rR + l_tR * qRm;
    double l_vL1 = l_mL * rL + l_tL * qLm;
    gR[0] -= l_mR * vR1;
    gL[0] += l_mL * vL1;
    gR[3] += l_tR;
    gL[3] += l_tL;
    const double l_qRm = l_tR * vR1, l_qLm = l_tL * vL1;
    // vX1 = uX - aX
    gR[IV] += l_vR1;
    gL[IV] += l_vL1;
    double l_aR = -l_vR1, l_aL = -l_vL1;
    // bP = max(tiny, aR), bM = min(-tiny, aL)
    if (bP_is_aR) l_aR += l_bP;
    if (bM_is_aL) l_aL += l_bM;
    // aR = uR + csR*cRs, aL = uL - csL*cLs
    gR[IV] += l_aR;
    gL[IV] += l_aL;
    double l_csR = l_aR * cRs, l_csL = -l_aL * cLs;
    const double l_cRs = l_aR * csR, l_cLs = -l_aL * csL;
    double l_pmid = 0.0;
    if (!lo_R) {
        const double l_x = l_cRs / (2.0 * cRs) * g_fac;  // d/d(p_mid/pR)
        l_pmid += l_x / pR;
        gR[3] -= l_x * (p_mid / pR) / pR;
    }
    if (!lo_L) {
        const double l_x = l_cLs / (2.0 * cLs) * g_fac;
        l_pmid += l_x / pL;
        gL[3] -= l_x * (p_mid / pL) / pL;
    }
    // p_mid = 0.5*(pL + pR + (K*rho_avg)*cs_avg)
    const double l_h = 0.5 * l_pmid;
    gL[3] += l_h;
    gR[3] += l_h;
    const double l_K = l_h * rho_avg * cs_avg;
    const double l_ravg = l_h * K * cs_avg;
    const double l_cavg = l_h * K * rho_avg;
    gL[IV] += l_K;
    gR[IV] -= l_K;
    gL[0] += 0.5 * l_ravg;
    gR[0] += 0.5 * l_ravg;
    l_csL += 0.5 * l_cavg;
    l_csR += 0.5 * l_cavg;
    // q: qXm = rX*uX, qXe = pX/gm1 + 0.5*rX*|v|^2
    gL[0] += l_qLm * uL + l_qLe * 0.5 * v2L;
    gL[IV] += l_qLm * rL;
    gL[3] += l_qLe / gm1;
    gL[1] += l_qLe * rL * wL[1];
    gL[2] += l_qLe * rL * wL[2];
    gR[0] += l_qRm * uR + l_qRe * 0.5 * v2R;
    gR[IV] += l_qRm * rR;
    gR[3] += l_qRe / gm1;
    gR[1] += l_qRe * rR * wR[1];
    gR[2] += l_qRe * rR * wR[2];
    // cs = sqrt(gamma*p/rho)
    gL[3] += l_csL * gamma / (2.0 * rL * csL);
    gL[0] -= l_csL * csL / (2.0 * rL);
    gR[3] += l_csR * gamma / (2.0 * rR * csR);
    gR[0] -= l_csR * csR / (2.0 * rR);
    for (int k = 0; k < 4; ++k) {
        lwL[k] += gL[k];
        lwR[k] += gR[k];
    }
}

// Any Riemann solver through Dual<8>: J^T lf accumulated into lwL, lwR.
template <RiemannSolver RS, int Axis>
__host__ __device__ inline void riemann_vjp2_dual(double gamma, const double (&wL)[4], const double (&wR)[4],
                                                  const double (&lf)[4], double (&lwL)[4], double (&lwR)[4]) {
    using D = Dual<8>;
    D a[4], b[4], f[4];
    for (int k = 0; k < 4; ++k) {
        a[k] = D::var(wL[k], k);
        b[k] = D::var(wR[k], 4 + k);
    }
    riemann_flux<RS, Axis, 2>(EosT<D>{D(gamma)}, a, b, f);
    for (int m = 0; m < 4; ++m) {
        double sl = 0.0, sr = 0.0;
        for (int k = 0; k < 4; ++k) {
            sl += lf[k] * f[k].d[m];
            sr += lf[k] * f[k].d[4 + m];
        }
        lwL[m] += sl;
        lwR[m] += sr;
    }
}

// Any reconstruction through Dual<5>.
template <Reconstruction R, SlopeLimiter SL>
__host__ __device__ inline void reconstruct_vjp_dual(const double (&s)[5], double lL, double lR, double (&ls)[5]) {
    using D = Dual<5>;
    D st[5];
    for (int k = 0; k < 5; ++k) st[k] = D::var(s[k], k);
    auto acc = [&](int o) -> D { return st[2 + o]; };
    D wl, wr;
    reconstruct_stencil<R, SL, D>(acc, wl, wr);
    for (int k = 0; k < 5; ++k) ls[k] += lL * wl.d[k] + lR * wr.d[k];
}

}  // namespace mh
