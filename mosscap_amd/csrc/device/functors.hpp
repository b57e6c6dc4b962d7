// Hydro functors of the finite-volume Euler path, written for gfx950.
//
// Same operator surface as Mosscap's Eos.hpp, Reconstruct.hpp and Riemann.hpp
// (names, template parameters, argument meaning), with two changes made for a
// fused LDS/register kernel:
//   * the real type is a template parameter (double for the drop-in path,
//     float for the compensated precision form), carried by `Arith<T>`,
//     which also chooses the divide/sqrt flavour;
//   * field access goes through accessor types (anything with operator[]),
//     so one functor reads registers, LDS rows or host arrays.
//
// In the double instantiation every expression keeps the reference's
// association order and the file is compiled with -ffp-contract=off, so
// results are bitwise those of the reference CPU build (x86-64, no FMA).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "layout.hpp"
#include "fast32.hpp"
#include <type_traits>

namespace mh {

// ---- arithmetic policy ------------------------------------------------------
template <typename T>
struct Arith;

template <>
struct Arith<double> {
    // IEEE-correct divide and sqrt (the gfx950 f64 expansions are correctly
    // rounded), matching glibc on the host.
    __host__ __device__ static inline double div(double a, double b) { return a / b; }
    __host__ __device__ static inline double sqrt(double a) { return __builtin_sqrt(a); }
    static constexpr double weno_eps = 1e-36;  // Reconstruct.hpp:237-239
};

template <>
struct Arith<float> {
    // Precision form: hardware reciprocal / sqrt (about 1 ulp).
    __device__ static inline float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static inline float sqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
    static constexpr float weno_eps = 1e-15f;  // single-precision value, Reconstruct.hpp:235-236
};

template <typename T>
__host__ __device__ inline T square(T t) { return t * t; }  // Types.hpp:106-109

// std::min / std::max semantics: min(a, b) = (b < a) ? b : a; max(a, b) = (a < b) ? b : a.
template <typename T>
__host__ __device__ inline T smin(T a, T b) { return (b < a) ? b : a; }
template <typename T>
__host__ __device__ inline T smax(T a, T b) { return (a < b) ? b : a; }
template <typename T>
__host__ __device__ inline T sabs(T a) { return __builtin_fabs(a); }
__host__ __device__ inline float sabs(float a) { return __builtin_fabsf(a); }
template <typename T>
__host__ __device__ inline T sign1(T a) { return __builtin_copysign(T(1), a); }
__host__ __device__ inline float sign1(float a) { return __builtin_copysignf(1.0f, a); }

// ---- EOS, Eos.hpp:54-131 ----------------------------------------------------
template <typename T>
struct EosT {
    T gamma;
};

// The FP32 functors take the EOS constants rounded once from FP64 (fast32.hpp).
template <>
struct EosT<float> : Eos32 {};

template <typename T>
__host__ __device__ inline EosT<T> make_eos(double gamma);
template <>
__host__ __device__ inline EosT<double> make_eos<double>(double gamma) {
    return EosT<double>{gamma};
}
template <>
__host__ __device__ inline EosT<float> make_eos<float>(double gamma) {
    EosT<float> e;
    e.gamma = (float)gamma;
    e.igm1 = (float)(1.0 / (gamma - 1.0));
    e.gfac = (float)((gamma + 1.0) / (2.0 * gamma));
    return e;
}

// sound_speed, Eos.hpp:54-59
template <int NumDim, typename T, typename WType>
__host__ __device__ inline T sound_speed(const T gamma, const WType& w) {
    using P = Prim<NumDim>;
    return Arith<T>::sqrt(Arith<T>::div(gamma * w[P::Pres], w[P::Rho]));
}

// cons_to_prim, Eos.hpp:61-79
template <int NumDim, typename T, typename QType, typename WType>
__host__ __device__ inline void cons_to_prim(const T gamma, const QType& q, WType& w) {
    using P = Prim<NumDim>;
    using C = Cons<NumDim>;
    w[P::Rho] = q[C::Rho];
    w[P::Vx] = Arith<T>::div(q[C::MomX], q[C::Rho]);
    T v2_sum = square(w[P::Vx]);
    if constexpr (NumDim > 1) {
        w[P::Vy] = Arith<T>::div(q[C::MomY], q[C::Rho]);
        v2_sum += square(w[P::Vy]);
    }
    if constexpr (NumDim > 2) {
        w[P::Vz] = Arith<T>::div(q[C::MomZ], q[C::Rho]);
        v2_sum += square(w[P::Vz]);
    }
    const T e_kin = T(0.5) * q[C::Rho] * v2_sum;
    w[P::Pres] = (gamma - T(1)) * ((q[C::Ene] - e_kin));
}

// prim_to_cons, Eos.hpp:81-99
template <int NumDim, typename T, typename WType, typename QType>
__host__ __device__ inline void prim_to_cons(const T gamma, const WType& w, QType& q) {
    using P = Prim<NumDim>;
    using C = Cons<NumDim>;
    q[C::Rho] = w[P::Rho];
    q[C::MomX] = w[P::Rho] * w[P::Vx];
    T v2_sum = square(w[P::Vx]);
    if constexpr (NumDim > 1) {
        q[C::MomY] = w[P::Rho] * w[P::Vy];
        v2_sum += square(w[P::Vy]);
    }
    if constexpr (NumDim > 2) {
        q[C::MomZ] = w[P::Rho] * w[P::Vz];
        v2_sum += square(w[P::Vz]);
    }
    const T e_kin = T(0.5) * w[P::Rho] * v2_sum;
    const T e_int = Arith<T>::div(w[P::Pres], gamma - T(1));
    q[C::Ene] = e_int + e_kin;
}

// prim_to_flux, Eos.hpp:101-131
template <int Axis, int NumDim, typename T, typename WType, typename FType>
__host__ __device__ inline void prim_to_flux(const T gamma, const WType& w, FType& f) {
    using P = Prim<NumDim>;
    using C = Cons<NumDim>;
    constexpr int IV1 = Velocity<Axis, NumDim>();
    constexpr int IM1 = Momentum<Axis, NumDim>();
    const T mass_flux = w[P::Rho] * w[IV1];
    T e_kin = T(0);
    f[C::Rho] = mass_flux;
    f[C::MomX] = mass_flux * w[P::Vx];
    e_kin += square(w[P::Vx]);
    if constexpr (NumDim > 1) {
        f[C::MomY] = mass_flux * w[P::Vy];
        e_kin += square(w[P::Vy]);
    }
    if constexpr (NumDim > 2) {
        f[C::MomZ] = mass_flux * w[P::Vz];
        e_kin += square(w[P::Vz]);
    }
    e_kin *= T(0.5) * w[P::Rho];
    f[IM1] += w[P::Pres];
    const T e_tot = Arith<T>::div(w[P::Pres], gamma - T(1)) + e_kin;
    f[C::Ene] = (e_tot + w[P::Pres]) * w[IV1];
}

// ---- reconstruction, Reconstruct.hpp:95-274 ---------------------------------

// slope_limiter<SlopeLimiter>, Reconstruct.hpp:130-148
template <SlopeLimiter scheme, typename T>
__host__ __device__ inline T slope_limiter(const T um, const T up) {
    if constexpr (scheme == SlopeLimiter::VanLeer) {
        return (um * up > T(0)) ? Arith<T>::div(T(2) * um * up, um + up) : T(0);
    } else if constexpr (scheme == SlopeLimiter::MonotonizedCentral) {
        return (sign1(um) + sign1(up)) * smin(sabs(um), smin(T(0.25) * sabs(um + up), sabs(up)));
    } else {
        return T(0.5) * (sign1(um) + sign1(up)) * smin(sabs(um), sabs(up));
    }
}

// Stencil radius each scheme reads (Stencil<Order>, Reconstruct.hpp:95-123).
__host__ __device__ constexpr int stencil_order(Reconstruction r) {
    return r == Reconstruction::Fog ? 0 : r == Reconstruction::Muscl ? 1 : 2;
}

// Core of reconstruct<...>: `s` is any accessor with s(o) = value at offset o
// in [-order, order] along the sweep axis. Produces the cell's left-face (wL)
// and right-face (wR) values.
template <Reconstruction recon, SlopeLimiter sl, typename T, typename S>
__host__ __device__ inline void reconstruct_stencil(const S& s, T& wL, T& wR) {
    if constexpr (recon == Reconstruction::Fog) {  // 150-154
        wL = s(0);
        wR = s(0);
    } else if constexpr (recon == Reconstruction::Muscl) {  // 156-167
        const T dwL = s(0) - s(-1);
        const T dwR = s(1) - s(0);
        const T delta = slope_limiter<sl>(dwL, dwR);
        wL = s(0) - T(0.5) * delta;
        wR = s(0) + T(0.5) * delta;
    } else if constexpr (recon == Reconstruction::Ppm) {  // 169-215, do_clamp = false
        const T dw_m = slope_limiter<sl>(s(-1) - s(-2), s(0) - s(-1));
        const T dw_0 = slope_limiter<sl>(s(0) - s(-1), s(1) - s(0));
        const T dw_p = slope_limiter<sl>(s(1) - s(0), s(2) - s(1));
        constexpr T sixth = T(1) / T(6);
        T l = T(0.5) * (s(-1) + s(0)) - sixth * (dw_0 - dw_m);
        T r = T(0.5) * (s(0) + s(1)) - sixth * (dw_p - dw_0);
        const T c = s(0);
        // sequential monotonicity fixes (not else-if), 204-214
        if ((r - c) * (c - l) <= T(0)) {
            l = c;
            r = c;
        }
        if (sabs(r - c) >= T(2) * sabs(l - c)) {
            r = T(3) * c - T(2) * l;
        }
        if (sabs(l - c) >= T(2) * sabs(r - c)) {
            l = T(3) * c - T(2) * r;
        }
        wL = l;
        wR = r;
    } else if constexpr (std::is_same_v<T, float>) {  // precision form
        weno5z_f32(s, wL, wR);
    } else {  // WENO5-Z (Borges et al. 2008), 217-274
        constexpr T b0c = T(13) / T(12);
        constexpr T b1c = T(0.25);
        const T sm2 = s(-2), sm1 = s(-1), s0 = s(0), sp1 = s(1), sp2 = s(2);
        const T beta0 = (b0c * square(sm2 - T(2) * sm1 + s0) + b1c * square(sm2 - T(4) * sm1 + T(3) * s0));
        const T beta1 = (b0c * square(sm1 - T(2) * s0 + sp1) + b1c * square(sm1 - sp1));
        const T beta2 = (b0c * square(s0 - T(2) * sp1 + sp2) + b1c * square(T(3) * s0 - T(4) * sp1 + sp2));
        const T eps = Arith<T>::weno_eps;
        const T tau5 = sabs(beta0 - beta2);
        const T ind0 = Arith<T>::div(tau5, beta0 + eps);
        const T ind1 = Arith<T>::div(tau5, beta1 + eps);
        const T ind2 = Arith<T>::div(tau5, beta2 + eps);
        const T q0 = T(1) + square(ind0);
        const T q1 = T(1) + square(ind1);
        const T q2 = T(1) + square(ind2);

        T eno0 = (T(2) * s0 + T(5) * sm1 - sm2);
        T eno1 = (-sp1 + T(5) * s0 + T(2) * sm1);
        T eno2 = (T(2) * sp2 - T(7) * sp1 + T(11) * s0);
        T a0 = T(3) * q0;
        T a1 = T(6) * q1;
        T a2 = q2;  // 1 * (1 + ind2^2) is exact
        T denom = T(6) * (a0 + a1 + a2);
        wL = Arith<T>::div(eno0 * a0 + eno1 * a1 + eno2 * a2, denom);

        eno0 = (T(2) * s0 + T(5) * sp1 - sp2);
        eno1 = (-sm1 + T(5) * s0 + T(2) * sp1);
        eno2 = (T(2) * sm2 - T(7) * sm1 + T(11) * s0);
        a0 = T(3) * q2;
        a1 = T(6) * q1;
        a2 = q0;
        denom = T(6) * (a0 + a1 + a2);
        wR = Arith<T>::div(eno0 * a0 + eno1 * a1 + eno2 * a2, denom);
    }
}

// reconstruct<Reconstruction, SlopeLimiter, Axis>(W, var, idx, wL, wR),
// Reconstruct.hpp:150-274, with W an accessor W(var, k, j, i).
template <Reconstruction recon, SlopeLimiter sl, int Axis, typename T, typename WAccess>
__host__ __device__ inline void reconstruct(const WAccess& W, const int var, const CellIndex& idx, T& wL, T& wR) {
    auto s = [&](int o) -> T {
        if constexpr (Axis == 0) return W(var, idx.k, idx.j, idx.i + o);
        else if constexpr (Axis == 1) return W(var, idx.k, idx.j + o, idx.i);
        else return W(var, idx.k + o, idx.j, idx.i);
    };
    reconstruct_stencil<recon, sl, T>(s, wL, wR);
}

// ---- Riemann solvers, Riemann.hpp:22-176 -------------------------------------
// wL: state on the left of the interface (RR of cell i-1), wR: on the right
// (RL of cell i), Hydro.cpp:138-144. Writes the N_HYDRO_VARS<NumDim> fluxes.
template <RiemannSolver rs, int Axis, int NumDim, typename T, typename WL, typename WR, typename FT>
__host__ __device__ inline void riemann_flux(const EosT<T>& eos, const WL& wL, const WR& wR, FT& flux) {
    using P = Prim<NumDim>;
    using C = Cons<NumDim>;
    constexpr int n_hydro = N_HYDRO_VARS<NumDim>;
    constexpr int IV = Velocity<Axis, NumDim>();
    constexpr int IM = Momentum<Axis, NumDim>();
    const T gamma = eos.gamma;
    T qL[n_hydro], qR[n_hydro], fL[n_hydro], fR[n_hydro];

    if constexpr (std::is_same_v<T, float> && rs != RiemannSolver::Hllc) {  // precision form
        rusanov_hll_f32<rs == RiemannSolver::Hll, Axis, NumDim>(eos, wL, wR, flux);
    } else if constexpr (rs == RiemannSolver::Rusanov) {  // 22-46
        const T vL = wL[IV];
        const T vR = wR[IV];
        prim_to_cons<NumDim>(gamma, wL, qL);
        prim_to_cons<NumDim>(gamma, wR, qR);
        prim_to_flux<Axis, NumDim>(gamma, wL, fL);
        prim_to_flux<Axis, NumDim>(gamma, wR, fR);
        const T csL = sound_speed<NumDim>(gamma, wL);
        const T csR = sound_speed<NumDim>(gamma, wR);
        const T max_c = T(0.5) * (csL + sabs(vL) + csR + sabs(vR));
#pragma unroll
        for (int i = 0; i < n_hydro; ++i) {
            flux[i] = T(0.5) * (fL[i] + fR[i] - max_c * (qR[i] - qL[i]));
        }
    } else if constexpr (rs == RiemannSolver::Hll) {  // 48-73
        const T csL = sound_speed<NumDim>(gamma, wL);
        const T csR = sound_speed<NumDim>(gamma, wR);
        constexpr T tiny = T(1e-20);
        const T sL = smin(-tiny, smin(wL[IV] - csL, wR[IV] - csR));
        // The right speed uses wL's velocity twice and floors at -tiny,
        // exactly as Riemann.hpp:60 does.
        const T sR = smax(-tiny, smax(wL[IV] + csL, wL[IV] + csR));
        const T sM = Arith<T>::div(T(1), sR - sL);
        prim_to_cons<NumDim>(gamma, wL, qL);
        prim_to_cons<NumDim>(gamma, wR, qR);
        prim_to_flux<Axis, NumDim>(gamma, wL, fL);
        prim_to_flux<Axis, NumDim>(gamma, wR, fR);
#pragma unroll
        for (int i = 0; i < n_hydro; ++i) {
            flux[i] = sM * (sR * fL[i] - sL * fR[i] + sR * sL * (qR[i] - qL[i]));
        }
    } else if constexpr (std::is_same_v<T, float>) {  // precision form
        hllc_f32<Axis, NumDim>(eos, wL, wR, flux);
    } else {  // HLLC, 75-176 (Athena-style pressure-based wave speeds)
        const T csL = sound_speed<NumDim>(gamma, wL);
        const T csR = sound_speed<NumDim>(gamma, wR);
        constexpr T tiny = T(1e-20);
        prim_to_cons<NumDim>(gamma, wL, qL);
        prim_to_cons<NumDim>(gamma, wR, qR);

        const T pL = wL[P::Pres];
        const T pR = wR[P::Pres];
        const T rho_avg = T(0.5) * (wL[P::Rho] + wR[P::Rho]);
        const T cs_avg = T(0.5) * (csL + csR);
        const T p_mid = T(0.5) * (pL + pR + (wL[IV] - wR[IV]) * rho_avg * cs_avg);
        const T g_fac = Arith<T>::div(gamma + T(1), T(2) * gamma);
        const T cLstar = (p_mid <= pL) ? T(1) : Arith<T>::sqrt(T(1) + g_fac * (Arith<T>::div(p_mid, pL) - T(1)));
        const T cRstar = (p_mid <= pR) ? T(1) : Arith<T>::sqrt(T(1) + g_fac * (Arith<T>::div(p_mid, pR) - T(1)));

        const T aL = wL[IV] - csL * cLstar;
        const T aR = wR[IV] + csR * cRstar;
        const T bM = smin(-tiny, aL);
        const T bP = smax(tiny, aR);

        T vxL = wL[IV] - aL;
        T vxR = wR[IV] - aR;
        const T tL = pL + vxL * qL[IM];
        const T tR = pR + vxR * qR[IM];
        const T mL = wL[P::Rho] * vxL;
        const T mR = -(wR[P::Rho] * vxR);
        const T inv_m = mL + mR;
        const T aM = Arith<T>::div(tL - tR, inv_m);
        const T cp = smax(Arith<T>::div(mL * tR + mR * tL, inv_m), T(0));

        vxL = wL[IV] - bM;
        vxR = wR[IV] - bP;
        const T mfL = wL[P::Rho] * vxL;
        const T mfR = wR[P::Rho] * vxR;
        fL[C::Rho] = mfL;
        fR[C::Rho] = mfR;
        fL[C::MomX] = mfL * wL[P::Vx];
        fR[C::MomX] = mfR * wR[P::Vx];
        if constexpr (NumDim > 1) {
            fL[C::MomY] = mfL * wL[P::Vy];
            fR[C::MomY] = mfR * wR[P::Vy];
        }
        if constexpr (NumDim > 2) {
            fL[C::MomZ] = mfL * wL[P::Vz];
            fR[C::MomZ] = mfR * wR[P::Vz];
        }
        fL[IM] += pL;
        fR[IM] += pR;
        fL[C::Ene] = qL[C::Ene] * vxL + pL * wL[IV];
        fR[C::Ene] = qR[C::Ene] * vxR + pR * wR[IV];

        T sL, sR, sM;
        if (aM >= T(0)) {
            const T d = aM - bM;
            sL = Arith<T>::div(aM, d);
            sR = T(0);
            sM = Arith<T>::div(-bM, d);
        } else {
            const T d = bP - aM;
            sL = T(0);
            sR = Arith<T>::div(-aM, d);
            sM = Arith<T>::div(bP, d);
        }
#pragma unroll
        for (int v = 0; v < n_hydro; ++v) {
            flux[v] = sL * fL[v] + sR * fR[v];
        }
        flux[IM] += sM * cp;
        flux[C::Ene] += sM * cp * aM;
    }
}

// dt_reducer's per-cell signal speed a = c_s + |v| (Hydro.cpp:192-206); the
// CFL reduction takes dx / max(a), which equals min(dx / a) bitwise because a
// correctly rounded divide is monotone in its denominator.
template <int NumDim, typename T, typename WType>
__host__ __device__ inline T signal_speed(const T gamma, const WType& w) {
    using P = Prim<NumDim>;
    const T cs = sound_speed<NumDim>(gamma, w);
    T vel2 = square(w[P::Vx]);
    if constexpr (NumDim > 1) vel2 += square(w[P::Vy]);
    if constexpr (NumDim > 2) vel2 += square(w[P::Vz]);
    const T vel = Arith<T>::sqrt(vel2);
    return cs + vel;
}

}  // namespace mh
