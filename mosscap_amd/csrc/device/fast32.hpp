// FP32 forms of the two heaviest functors for the compensated precision form.
//
// Same mathematics as reconstruct<Weno5Z> (Reconstruct.hpp:217-274) and
// riemann_flux<Hllc> (Riemann.hpp:75-176), rearranged to spend fewer VALU
// and transcendental instructions; the FP64 drop-in never uses them.
//
// WENO5-Z in difference form. With d1..d4 the four consecutive differences of
// the stencil, the smoothness indicators are (13/12)(a_k^2 + (3/13) b_k^2) with
// a = (d2-d1, d3-d2, d4-d3), b = (3d2-d1, d2+d3, d4-3d3); the common factor
// 13/12 cancels in tau/(beta+eps) once eps is scaled by 12/13. The ENO
// candidates, relative to s0 and times 6, are g = (d1-4d2, -2d2-d3, 2d4-5d3)
// = (-(b0+d2), -(d2+b1), 2b2+d3) on the left face and h = (4d3-d4, d2+2d3,
// 5d2-2d1) = (d3-b2, b1+d3, 2b0-d2) on the right. With the reference's
// weights (3q0, 6q1, q2) on the left and (3q2, 6q1, q0) on the right,
// w = s0 + sum(alpha g) / (6 sum(alpha)); the 6 is divided into the
// candidates, so wL = s0 + (q0 g0/2 + q1 g1 + q2 g2/6) / (3q0 + 6q1 + q2).
//
// HLLC with (i) c_s = gamma p * rsq(gamma p rho), (ii) the star-state factor
// sqrt(1 + g (p*/p - 1)) = X rsq(X p) with X = p + g (p* - p), and (iii) only
// the upwind side's flux formed, since one of the reference's two weights is
// always zero: F = s_K F_K + s_M (0, cp, ..., cp aM) with K the side selected
// by the sign of the contact speed and s_K = aM / (aM - b_K),
// s_M = -b_K / (aM - b_K) (b_K = bM or bP) for both signs.
#pragma once

#include "layout.hpp"

namespace mh {

__device__ inline float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float rsq32(float x) { return __builtin_amdgcn_rsqf(x); }

template <typename S>
__device__ inline void weno5z_f32(const S& s, float& wL, float& wR) {
    const float sm2 = s(-2), sm1 = s(-1), s0 = s(0), sp1 = s(1), sp2 = s(2);
    const float d1 = sm1 - sm2, d2 = s0 - sm1, d3 = sp1 - s0, d4 = sp2 - sp1;
    constexpr float k = 3.0f / 13.0f;
    const float a0 = d2 - d1, b0 = __builtin_fmaf(3.0f, d2, -d1);
    const float a1 = d3 - d2, b1 = d2 + d3;
    const float a2 = d4 - d3, b2 = __builtin_fmaf(-3.0f, d3, d4);
    // (12/13) (beta_k + eps): eps, scaled alike, folded into the first product
    constexpr float eps = 1e-15f * (12.0f / 13.0f);
    const float be0 = __builtin_fmaf(k * b0, b0, __builtin_fmaf(a0, a0, eps));
    const float be1 = __builtin_fmaf(k * b1, b1, __builtin_fmaf(a1, a1, eps));
    const float be2 = __builtin_fmaf(k * b2, b2, __builtin_fmaf(a2, a2, eps));
    const float tau = __builtin_fabsf(be0 - be2);
    // tau/(beta+eps) reaches ~1e19 next to a flat stencil; clamped so that
    // its square (and the weighted sums) stay finite in FP32. A weight ratio
    // of 1e30 keeps the ENO choice exact to FP32 precision.
    constexpr float tmax = 1e15f;
    const float t0 = __builtin_fminf(tau * rcp32(be0), tmax);
    const float t1 = __builtin_fminf(tau * rcp32(be1), tmax);
    const float t2 = __builtin_fminf(tau * rcp32(be2), tmax);
    const float q0 = __builtin_fmaf(t0, t0, 1.0f);
    const float q1 = __builtin_fmaf(t1, t1, 1.0f);
    const float q2 = __builtin_fmaf(t2, t2, 1.0f);
    // candidates relative to s0 with the weights' common factor 1/6 taken
    // into them: g0/2, g1, g2/6 (left face) and h0/2, h1, h2/6 (right face)
    const float g0 = -0.5f * (b0 + d2);
    const float g1 = -(d2 + b1);
    const float g2 = __builtin_fmaf(2.0f, b2, d3) * (1.0f / 6.0f);
    const float h0 = 0.5f * (d3 - b2);
    const float h1 = b1 + d3;
    const float h2 = __builtin_fmaf(2.0f, b0, -d2) * (1.0f / 6.0f);
    const float nL = __builtin_fmaf(q0, g0, __builtin_fmaf(q1, g1, q2 * g2));
    const float nR = __builtin_fmaf(q2, h0, __builtin_fmaf(q1, h1, q0 * h2));
    const float sL = __builtin_fmaf(3.0f, q0, __builtin_fmaf(6.0f, q1, q2));
    const float sR = __builtin_fmaf(3.0f, q2, __builtin_fmaf(6.0f, q1, q0));
    wL = __builtin_fmaf(nL, rcp32(sL), s0);
    wR = __builtin_fmaf(nR, rcp32(sR), s0);
}

// EOS constants of the FP32 functors, rounded once from FP64 on the host so
// that E = p/(gamma-1) + ... in the fluxes inverts the FP64 pressure recovery.
struct Eos32 {
    float gamma;  // float(gamma)
    float igm1;   // float(1 / (gamma - 1))
    float gfac;   // float((gamma + 1) / (2 gamma))
};

template <int NumDim, typename W>
__device__ inline float energy_f32(const Eos32& e, const W& w) {
    float v2 = w[1] * w[1];
#pragma unroll
    for (int a = 2; a <= NumDim; ++a) v2 = __builtin_fmaf(w[a], w[a], v2);
    return __builtin_fmaf(w[1 + NumDim], e.igm1, 0.5f * w[0] * v2);
}

// Rusanov (Riemann.hpp:22-46) and HLL (48-73, with its sR quirk) in FP32.
template <bool HLL, int Axis, int NumDim, typename WL, typename WR, typename FT>
__device__ inline void rusanov_hll_f32(const Eos32& e, const WL& wL, const WR& wR, FT& flux) {
    constexpr int IV = 1 + Axis;
    constexpr int IE = 1 + NumDim;
    const float gpL = e.gamma * wL[IE], gpR = e.gamma * wR[IE];
    const float csL = gpL * rsq32(gpL * wL[0]);
    const float csR = gpR * rsq32(gpR * wR[0]);
    const float vL = wL[IV], vR = wR[IV];
    const float EL = energy_f32<NumDim>(e, wL), ER = energy_f32<NumDim>(e, wR);
    float qL[2 + NumDim], qR[2 + NumDim], fL[2 + NumDim], fR[2 + NumDim];
    qL[0] = wL[0];
    qR[0] = wR[0];
    fL[0] = wL[0] * vL;
    fR[0] = wR[0] * vR;
#pragma unroll
    for (int a = 1; a <= NumDim; ++a) {
        qL[a] = wL[0] * wL[a];
        qR[a] = wR[0] * wR[a];
        fL[a] = fL[0] * wL[a];
        fR[a] = fR[0] * wR[a];
    }
    fL[IV] += wL[IE];
    fR[IV] += wR[IE];
    qL[IE] = EL;
    qR[IE] = ER;
    fL[IE] = (EL + wL[IE]) * vL;
    fR[IE] = (ER + wR[IE]) * vR;
    if constexpr (HLL) {
        constexpr float tiny = 1e-20f;
        const float a = __builtin_fminf(vL - csL, vR - csR);
        const float sL = (a < -tiny) ? a : -tiny;
        const float b = __builtin_fmaxf(vL + csL, vL + csR);  // the reference's vL twice
        const float sR = (-tiny < b) ? b : -tiny;
        const float sM = rcp32(sR - sL);
        const float srl = sR * sL;
#pragma unroll
        for (int i = 0; i < 2 + NumDim; ++i)
            flux[i] = sM * __builtin_fmaf(srl, qR[i] - qL[i], __builtin_fmaf(sR, fL[i], -sL * fR[i]));
    } else {
        const float max_c = 0.5f * ((csL + __builtin_fabsf(vL)) + (csR + __builtin_fabsf(vR)));
#pragma unroll
        for (int i = 0; i < 2 + NumDim; ++i) flux[i] = 0.5f * __builtin_fmaf(-max_c, qR[i] - qL[i], fL[i] + fR[i]);
    }
}

template <int Axis, int NumDim, typename WL, typename WR, typename FT>
__device__ inline void hllc_f32(const Eos32& e, const WL& wL, const WR& wR, FT& flux) {
    constexpr int IV = 1 + Axis;
    constexpr int IE = 1 + NumDim;
    constexpr float tiny = 1e-20f;
    const float gamma = e.gamma, gfac = e.gfac;
    const float rL = wL[0], rR = wR[0], pL = wL[IE], pR = wR[IE], vL = wL[IV], vR = wR[IV];
    const float gpL = gamma * pL, gpR = gamma * pR;
    const float csL = gpL * rsq32(gpL * rL);
    const float csR = gpR * rsq32(gpR * rR);
    const float rho_avg = 0.5f * (rL + rR);
    const float cs_avg = 0.5f * (csL + csR);
    const float p_mid = 0.5f * __builtin_fmaf((vL - vR) * rho_avg, cs_avg, pL + pR);
    // star-state factor: 1 where p_mid <= p (X <= p, so max(X, p) = p and
    // p rsq(p p) = 1 to rounding), X rsq(X p) where p_mid > p
    const float XL = __builtin_fmaxf(__builtin_fmaf(gfac, p_mid - pL, pL), pL);
    const float XR = __builtin_fmaxf(__builtin_fmaf(gfac, p_mid - pR, pR), pR);
    const float cLs = XL * rsq32(XL * pL);
    const float cRs = XR * rsq32(XR * pR);
    const float aL = __builtin_fmaf(-csL, cLs, vL);
    const float aR = __builtin_fmaf(csR, cRs, vR);
    const float bM = (aL < -tiny) ? aL : -tiny;
    const float bP = (tiny < aR) ? aR : tiny;
    const float uL = vL - aL, uR = vR - aR;
    const float tL = __builtin_fmaf(uL, rL * vL, pL);
    const float tR = __builtin_fmaf(uR, rR * vR, pR);
    const float mL = rL * uL;
    const float mR = -(rR * uR);
    const float iD = rcp32(mL + mR);
    const float aM = (tL - tR) * iD;
    const float cp = __builtin_fmaxf(__builtin_fmaf(mL, tR, mR * tL) * iD, 0.0f);
    // upwind side of the contact: its primitives, then its total energy
    const bool left = aM >= 0.0f;
    const float rK = left ? rL : rR, pK = left ? pL : pR;
    const float bK = left ? bM : bP;
    float vKa[NumDim];
#pragma unroll
    for (int a = 0; a < NumDim; ++a) vKa[a] = left ? wL[1 + a] : wR[1 + a];
    const float vK = vKa[Axis];
    float v2 = vKa[0] * vKa[0];
#pragma unroll
    for (int a = 1; a < NumDim; ++a) v2 = __builtin_fmaf(vKa[a], vKa[a], v2);
    const float EK = __builtin_fmaf(pK, e.igm1, 0.5f * rK * v2);
    const float uK = vK - bK;
    const float mfK = rK * uK;
    const float r = rcp32(aM - bK);
    const float sK = aM * r;
    const float sMcp = -bK * r * cp;
    flux[0] = sK * mfK;
#pragma unroll
    for (int a = 1; a <= NumDim; ++a) flux[a] = sK * (mfK * vKa[a - 1]);
    flux[IV] = __builtin_fmaf(sK, pK, flux[IV]) + sMcp;
    flux[IE] = __builtin_fmaf(sK, __builtin_fmaf(EK, uK, pK * vK), sMcp * aM);
}

}  // namespace mh
