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
// on the left face and h = (4d3-d4, d2+2d3, 5d2-2d1) on the right, so
// w = s0 + sum(alpha g) / (6 sum(alpha)) with the reference's weights.
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
    const float be0 = __builtin_fmaf(k * b0, b0, a0 * a0);
    const float be1 = __builtin_fmaf(k * b1, b1, a1 * a1);
    const float be2 = __builtin_fmaf(k * b2, b2, a2 * a2);
    constexpr float eps = 1e-15f * (12.0f / 13.0f);
    const float tau = __builtin_fabsf(be0 - be2);
    // tau/(beta+eps) reaches ~1e19 next to a flat stencil; clamped so that
    // its square (and the weighted sums) stay finite in FP32. A weight ratio
    // of 1e30 keeps the ENO choice exact to FP32 precision.
    constexpr float tmax = 1e15f;
    const float t0 = __builtin_fminf(tau * rcp32(be0 + eps), tmax);
    const float t1 = __builtin_fminf(tau * rcp32(be1 + eps), tmax);
    const float t2 = __builtin_fminf(tau * rcp32(be2 + eps), tmax);
    const float q0 = __builtin_fmaf(t0, t0, 1.0f);
    const float q1 = __builtin_fmaf(t1, t1, 1.0f);
    const float q2 = __builtin_fmaf(t2, t2, 1.0f);
    const float A0 = 3.0f * q0, A1 = 6.0f * q1, A2 = 3.0f * q2;
    const float g0 = __builtin_fmaf(-4.0f, d2, d1);
    const float g1 = __builtin_fmaf(-2.0f, d2, -d3);
    const float g2 = __builtin_fmaf(2.0f, d4, -5.0f * d3);
    const float h0 = __builtin_fmaf(4.0f, d3, -d4);
    const float h1 = __builtin_fmaf(2.0f, d3, d2);
    const float h2 = __builtin_fmaf(5.0f, d2, -2.0f * d1);
    const float nL = __builtin_fmaf(A0, g0, __builtin_fmaf(A1, g1, q2 * g2));
    const float nR = __builtin_fmaf(A2, h0, __builtin_fmaf(A1, h1, q0 * h2));
    wL = __builtin_fmaf(nL, rcp32(6.0f * (A0 + A1 + q2)), s0);
    wR = __builtin_fmaf(nR, rcp32(6.0f * (A2 + A1 + q0)), s0);
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
    const float EL = energy_f32<NumDim>(e, wL);
    const float ER = energy_f32<NumDim>(e, wR);
    const float rho_avg = 0.5f * (rL + rR);
    const float cs_avg = 0.5f * (csL + csR);
    const float p_mid = 0.5f * __builtin_fmaf((vL - vR) * rho_avg, cs_avg, pL + pR);
    const float XL = __builtin_fmaf(gfac, p_mid - pL, pL);
    const float XR = __builtin_fmaf(gfac, p_mid - pR, pR);
    const float cLs = (p_mid <= pL) ? 1.0f : XL * rsq32(XL * pL);
    const float cRs = (p_mid <= pR) ? 1.0f : XR * rsq32(XR * pR);
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
    const float cp0 = __builtin_fmaf(mL, tR, mR * tL) * iD;
    const float cp = (cp0 < 0.0f) ? 0.0f : cp0;
    // upwind side of the contact
    const bool left = aM >= 0.0f;
    const float rK = left ? rL : rR, pK = left ? pL : pR, vK = left ? vL : vR, EK = left ? EL : ER;
    const float bK = left ? bM : bP;
    const float uK = vK - bK;
    const float mfK = rK * uK;
    const float r = rcp32(aM - bK);
    const float sK = aM * r;
    const float sMcp = -bK * r * cp;
    flux[0] = sK * mfK;
#pragma unroll
    for (int a = 1; a <= NumDim; ++a) {
        const float vKa = left ? wL[a] : wR[a];
        flux[a] = sK * (mfK * vKa);
    }
    flux[IV] = __builtin_fmaf(sK, pK, flux[IV]) + sMcp;
    flux[IE] = __builtin_fmaf(sK, __builtin_fmaf(EK, uK, pK * vK), sMcp * aM);
}

}  // namespace mh
