// Compensated FP32 state form (the precision form).
//
// * State: each conserved value is an unevaluated float pair hi + lo
//   (|lo| <= ulp(hi)/2), stored as two float planes: 8 bytes per value, the
//   same HBM footprint as FP64.
// * Reconstruction and Riemann solves run in FP32 with hardware reciprocal
//   and sqrt.
// * cons->prim recovers the pressure p = (gamma-1)(E - |m|^2/(2 rho)) in FP64
//   from the full hi+lo values: at high Mach number E and the kinetic energy
//   nearly cancel, which FP32 cannot resolve. A cell whose kinetic energy is
//   at most half of E (mild cancellation), and at least twice above the
//   floors when they are on, takes it in FP32.
// * The stage combine adds each stage's increment over Q^n, formed in FP32,
//   to the hi+lo Q^n with TwoSum, so the state accumulates the small
//   per-stage increments without FP32 rounding drift.
// * The CFL signal speed is evaluated in FP64 from hi+lo, or in FP32 for a
//   cell of moderate kinetic energy.
// * Optional positivity floors on rho and p of the reconstruction input and
//   of each stage's new state; each interior cell and stage that needed one
//   is counted once.
#pragma once

#include "functors.hpp"
#include "wave.hpp"

// 1: cons->prim and the CFL signal speed take cells of moderate kinetic
// energy in FP32 (see to_prim_nd, signal_speed_nd); 0: every cell through the
// FP64 recovery.
#ifndef MH_F32C_FAST_PRIM
#define MH_F32C_FAST_PRIM 1
#endif
// |m|^2/rho (twice the kinetic energy) at most MH_F32C_FAST_KIN * E for the
// FP32 path; with floors, MH_F32C_FAST_FLOORS = 1 lets a cell at least twice
// above both floors take it too (0: every cell of a floored run takes FP64).
#ifndef MH_F32C_FAST_KIN
#define MH_F32C_FAST_KIN 1.0f
#endif
#ifndef MH_F32C_FAST_FLOORS
#define MH_F32C_FAST_FLOORS 1
#endif

namespace mh {

// The error-free transforms below must not be contracted into FMAs (the
// translation unit otherwise allows contraction), hence the pragmas.
namespace df {  // double-float helpers (Dekker / Knuth error-free transforms)

__device__ inline float2 two_sum(float a, float b) {
#pragma clang fp contract(off)
    const float s = a + b;
    const float bb = s - a;
    const float e = (a - (s - bb)) + (b - bb);
    return make_float2(s, e);
}
__device__ inline float2 fast_two_sum(float a, float b) {  // requires |a| >= |b| or a == 0
#pragma clang fp contract(off)
    const float s = a + b;
    const float e = b - (s - a);
    return make_float2(s, e);
}
__device__ inline float2 add_f(float2 a, float b) {
#pragma clang fp contract(off)
    const float2 s = two_sum(a.x, b);
    const float e = s.y + a.y;
    return fast_two_sum(s.x, e);
}
__device__ inline double to_double(float2 a) { return (double)a.x + (double)a.y; }
__device__ inline float2 from_double(double d) {
    const float hi = (float)d;
    const float lo = (float)(d - (double)hi);
    return make_float2(hi, lo);
}

}  // namespace df

struct FormF32c {
    using real = float;
    using raw = float2;
    struct Planes {
        float* hi;
        float* lo;
    };
    struct Ptrs {
        Planes qs, qn, out;
        const double* dt;
        double dx;
        double gamma;
        double gm1;  // gamma - 1 in FP64, for the pressure recovery
        unsigned long long* amax_bits;
        unsigned int* floor_hits;
        float rho_floor;  // <= 0 disables the floors
        float p_floor;
        float* frx = nullptr;  // with tracers: density flux of x faces [j][i] (face i of row j)
        float* fry = nullptr;  // ... of y faces [j][i] (face j of column i)
        int32_t gravity = 0;   // a gravity source term is registered
        double gx = 0.0, gy = 0.0;
        const float* dz3 = nullptr;  // 3-D: z flux difference per interior cell [var][k][j][i]
        double gz = 0.0;
    };

    __device__ static inline float2 get(const Planes& a, int idx) { return make_float2(a.hi[idx], a.lo[idx]); }
    // Buffer-resource access for the plain march (see FormF64::Buf).
    struct Buf {
        __amdgpu_buffer_rsrc_t hi, lo;
    };
    // 2-D grids with copy-type faces read ghost rows and columns from their
    // sources instead of filling them after every stage (StageGrid::virt)
    static constexpr bool kVirtualGhosts = true;
    static constexpr bool kHasFloors = true;  // positivity floors (floors_on)
    // the instance without buffer resources loads Q^(s) and Q^n for the
    // combine at the top of the row (their latency hides behind the sweeps)
    static constexpr bool kEarlyOperands = true;
    static constexpr uint32_t kElem = 4;
    __device__ static inline Buf buf(const Planes& a) { return Buf{make_rsrc(a.hi), make_rsrc(a.lo)}; }
    __device__ static inline float2 bget(const Buf& b, uint32_t voff, uint32_t s_elem) {
        return make_float2(buf_load_f32(b.hi, voff, s_elem * kElem), buf_load_f32(b.lo, voff, s_elem * kElem));
    }
    __device__ static inline void bput(const Buf& b, uint32_t voff, uint32_t s_elem, float2 v) {
        buf_store_f32(v.x, b.hi, voff, s_elem * kElem);
        buf_store_f32(v.y, b.lo, voff, s_elem * kElem);
    }
    // direct-to-LDS load of an element's hi and lo words
    __device__ static inline void dma(const Buf& b, uint32_t* hi, uint32_t* lo, uint32_t voff, uint32_t s_elem) {
        buf_load_lds(b.hi, hi, voff, s_elem * kElem);
        buf_load_lds(b.lo, lo, voff, s_elem * kElem);
    }
    __device__ static inline float2 from_words(uint32_t hi, uint32_t lo) {
        return make_float2(__builtin_bit_cast(float, hi), __builtin_bit_cast(float, lo));
    }
    __device__ static inline void put(const Planes& a, int idx, float2 v) {
        a.hi[idx] = v.x;
        a.lo[idx] = v.y;
    }
    __device__ static inline bool needs_qn(int kind) { return kind == 1 || kind == 2 || kind == 3 || kind == 5; }
    __device__ static inline float2 negate(float2 v) { return make_float2(-v.x, -v.y); }
    __device__ static inline float2 from_double(double v) { return df::from_double(v); }
    __device__ static inline double value(float2 v) { return df::to_double(v); }
    __device__ static inline const void* get_ptr(const Planes& a) { return a.hi; }
    __device__ static inline float2 add_double(float2 v, double d) { return df::from_double(df::to_double(v) + d); }

    template <int ND, int NV, int IM>
    __device__ static inline void diode(float2 (&q)[NV], bool start) {
        const double mom_pre = df::to_double(q[IM]);
        const double mom = start ? smin(mom_pre, 0.0) : smax(mom_pre, 0.0);
        q[IM] = df::from_double(mom);
        const double e = df::to_double(q[Cons<ND>::Ene]) - square(mom - mom_pre) / df::to_double(q[Cons<ND>::Rho]);
        q[Cons<ND>::Ene] = df::from_double(e);
    }

    // cons->prim with FP64 pressure recovery. Returns true when a floor fired.
    // The density floor applies before 1/rho is formed, so a state whose
    // density went non-positive still yields finite velocities.
    template <int ND, bool Floors = true, int NV>
    __device__ static inline bool to_prim_nd(const Ptrs& p, const float2 (&q)[NV], float (&w)[NV]) {
#if MH_F32C_FAST_PRIM
        {
            // A cell whose kinetic energy is at most half of E takes its
            // primitives in FP32 from the hi words (and E's lo word):
            // E - |m|^2/(2 rho) then loses no more than ~4 ulp, a few times
            // the FP32 rounding its primitives carry anyway. With the floors
            // on, only a cell at least twice above both (where the FP64 path
            // would not floor it either). Other cells (high Mach, where the
            // two nearly cancel, or near the floors) take the FP64 recovery
            // below.
            const float rho = q[0].x;
            const float r = __builtin_amdgcn_rcpf(rho);
            float v[ND], kin2 = 0.0f;
#pragma unroll
            for (int a = 0; a < ND; ++a) {
                v[a] = q[1 + a].x * r;
                kin2 = __builtin_fmaf(q[1 + a].x, v[a], kin2);
            }
            const float E = q[1 + ND].x;
            const float pres = (float)p.gm1 * ((E - 0.5f * kin2) + q[1 + ND].y);
            const bool floors = Floors && p.rho_floor > 0.0f;
            const bool clear = MH_F32C_FAST_FLOORS ? (rho > 2.0f * p.rho_floor && pres > 2.0f * p.p_floor) : false;
            if (kin2 <= MH_F32C_FAST_KIN * E && (!floors || clear)) {
                w[0] = rho;
#pragma unroll
                for (int a = 0; a < ND; ++a) w[1 + a] = v[a];
                w[1 + ND] = pres;
                return false;
            }
        }
#endif
        double rho_d = df::to_double(q[0]);
        bool hit = false;
        const bool floors = Floors && p.rho_floor > 0.0f;
        if (floors && rho_d < (double)p.rho_floor) {
            rho_d = (double)p.rho_floor;
            hit = true;
        }
        // 1/rho: hardware FP32 reciprocal refined by one FP64 Newton step
        // (relative error ~1e-14, ample for the kinetic-energy subtraction).
        const double r0 = (double)__builtin_amdgcn_rcpf((float)rho_d);
        const double r = r0 * (2.0 - rho_d * r0);
        double m2 = 0.0;
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            const double m = df::to_double(q[1 + a]);
            m2 += m * m;
            w[1 + a] = (float)(m * r);
        }
        double pres = p.gm1 * (df::to_double(q[1 + ND]) - 0.5 * m2 * r);
        if (floors && pres < (double)p.p_floor) {
            pres = (double)p.p_floor;
            hit = true;
        }
        w[0] = (float)rho_d;
        w[1 + ND] = (float)pres;
        return hit;
    }

    // Gravity source S*dt (Gravity.cpp:22-38) from the stage input, formed in
    // FP64 from hi+lo and rounded to FP32 like the flux increment it joins.
    template <int ND, int NV>
    __device__ static inline void gravity_src(const Ptrs& p, const float2 (&q)[NV], float dt, float (&src)[NV]) {
        const double rho = df::to_double(q[0]), d = (double)dt;
        src[0] = 0.0f;
        src[1] = (float)(rho * p.gx * d);
        double energy_update = df::to_double(q[1]) * p.gx;
        if constexpr (ND > 1) {
            src[2] = (float)(rho * p.gy * d);
            energy_update += df::to_double(q[2]) * p.gy;
        }
        if constexpr (ND > 2) {
            src[3] = (float)(rho * p.gz * d);
            energy_update += df::to_double(q[3]) * p.gz;
        }
        src[1 + ND] = (float)(energy_update * d);
    }

    // The stage lambdas of TimeStepping.cpp, a_qn Q^n + a_qs Q + a_l dq,
    // rearranged around Q^n where they read it (a_qn + a_qs = 1):
    //   Q^n + a_qs ((Q - Q^n) + (a_l / a_qs) dq).
    // The increment over Q^n is formed in FP32 (Q - Q^n is exact in the hi
    // words while Q stays within a factor 2 of Q^n, and FP32's relative error
    // on the increment is what the FP32 fluxes carry anyway) and added to
    // Q^n with TwoSum, so the hi+lo state accumulates the small per-stage
    // increments without FP32 rounding drift. The source joins the flux
    // increment (dq + S*dt) in FP32.
    __device__ static inline float2 combine(int kind, float2 Q, float2 Qn, float dq_flux, float src) {
#pragma clang fp contract(off)
        return combine_nosrc(kind, Q, Qn, dq_flux + src);
    }
    // The same with the increment dq of the flux term alone (no source).
    __device__ static inline float2 combine_nosrc(int kind, float2 Q, float2 Qn, float dq) {
#pragma clang fp contract(off)
        const auto around_qn = [&](float a_qs, float b) {
            const float d = (Q.x - Qn.x) + (Q.y - Qn.y);
            return df::add_f(Qn, a_qs * (d + b * dq));
        };
        switch (kind) {
            case 0: return df::add_f(Q, dq);                    // Q + dq
            case 1: return around_qn(0.5f, 1.0f);               // (Q^n + Q + dq) / 2
            case 2: return around_qn(0.25f, 1.0f);              // 3/4 Q^n + 1/4 (Q + dq)
            case 3: return around_qn(2.0f / 3.0f, 1.0f);        // 1/3 Q^n + 2/3 (Q + dq)
            case 4: return df::add_f(Q, 0.5f * dq);             // Q + dq / 2
            default: return around_qn(1.0f / 3.0f, 0.5f);       // 2/3 Q^n + 1/3 Q + 1/6 dq
        }
    }
    // adds of zero (the flux sum's leading 0, a zero source) are left out
    static constexpr bool kExactZeros = false;

    // Positivity floors of a stage's new state (with the floors on): a
    // density below rho_floor (or NaN) puts the cell at rest at rho_floor and
    // p_floor; otherwise a pressure below p_floor (or NaN) raises the energy
    // to the kinetic energy plus p_floor/(gamma-1). Floors on the
    // reconstruction input alone cannot stop a near-vacuum cell from being
    // emptied below zero by its outflow, after which 1/rho of the floored
    // density turns its momentum into an unbounded velocity. Returns true
    // when it changed the cell.
    template <int ND, bool Floors = true, int NV>
    __device__ static inline bool floor_state(const Ptrs& p, float2 (&q)[NV]) {
        if (!Floors || !(p.rho_floor > 0.0f)) return false;
        const double rho = df::to_double(q[0]);
        const double pf = (double)p.p_floor;
        if (!(rho >= (double)p.rho_floor)) {
            q[0] = df::from_double((double)p.rho_floor);
#pragma unroll
            for (int a = 0; a < ND; ++a) q[1 + a] = make_float2(0.0f, 0.0f);
            q[1 + ND] = df::from_double(pf / p.gm1);
            return true;
        }
        const double r0 = (double)__builtin_amdgcn_rcpf((float)rho);
        const double r = (rho < 1e30) ? r0 * (2.0 - rho * r0) : 1.0 / rho;
        double m2 = 0.0;
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            const double m = df::to_double(q[1 + a]);
            m2 += m * m;
        }
        const double ke = 0.5 * m2 * r;
        const double pres = p.gm1 * (df::to_double(q[1 + ND]) - ke);
        if (!(pres >= pf)) {
            q[1 + ND] = df::from_double(ke + pf / p.gm1);
            return true;
        }
        return false;
    }

    // The floors are on: the stage launchers then pick the instance that
    // compiles them in (Floors = true); the plain one leaves them out.
    static inline bool floors_on(const Ptrs& p) { return p.rho_floor > 0.0f; }

    // Reconstructed face states get the same floors as the cell states
    // (WENO and PPM may undershoot between positive cell values).
    template <int ND, bool Floors = true, int NV>
    __device__ static inline void floor_face(const Ptrs& p, float (&w)[NV]) {
        if (Floors && p.rho_floor > 0.0f) {
            w[0] = w[0] < p.rho_floor ? p.rho_floor : w[0];
            w[1 + ND] = w[1 + ND] < p.p_floor ? p.p_floor : w[1 + ND];
        }
    }

    // sqrt(x) to ~1e-14 relative: the FP32 reciprocal square root refined
    // by one FP64 Newton step, s + (x - s^2) y / 2 with s = x y. Zero gives
    // zero; values outside [1e-30, 1e30] (FP32's range with margin), negative
    // or NaN ones take the correctly rounded FP64 square root.
    __device__ static inline double sqrt_fast(double x) {
#pragma clang fp contract(off)
        const bool fast = x > 1e-30 && x < 1e30;
        if (!fast && x != 0.0) return __builtin_sqrt(x);
        const double xs = fast ? x : 1.0;
        const double y = (double)__builtin_amdgcn_rsqf((float)xs);
        const double s = xs * y;
        const double r = __builtin_fma(0.5 * y, __builtin_fma(-s, s, xs), s);
        return fast ? r : 0.0;
    }

    // CFL signal speed in FP64; a state whose pressure went negative
    // contributes its sound speed at p = 0 (the floors keep the scheme going).
    template <int ND, int NV>
    __device__ static inline double signal_speed_nd(double gamma, const float2 (&q)[NV]) {
        // Not contracted, so the stage kernels (built with contraction) and
        // the ghost and CFL kernels (built without) give a cell the same speed.
#pragma clang fp contract(off)
#if MH_F32C_FAST_PRIM
        {
            // A cell of moderate kinetic energy (at most half of E, as in
            // to_prim_nd) with a normal density and positive pressure: the
            // speed in FP32 from the hi words (relative error ~1e-7, which
            // moves dt by as much; the FP32 fluxes carry that anyway).
            const float rho = q[0].x;
            const float r = __builtin_amdgcn_rcpf(rho);
            float m2 = 0.0f;
#pragma unroll
            for (int a = 0; a < ND; ++a) m2 = m2 + q[1 + a].x * q[1 + a].x;
            const float kin2 = m2 * r, E = q[1 + ND].x;
            const float pres = (float)(gamma - 1.0) * ((E - 0.5f * kin2) + q[1 + ND].y);
            if (rho > 1e-30f && rho < 1e30f && kin2 <= MH_F32C_FAST_KIN * E && pres > 0.0f)
                return (double)(__builtin_amdgcn_sqrtf((float)gamma * pres * r) + __builtin_amdgcn_sqrtf(m2) * r);
        }
#endif
        // FP64 throughout, with 1/rho from the FP32 reciprocal and one Newton
        // step (relative error ~1e-14) in place of three IEEE divisions, and
        // the square roots from sqrt_fast: |v| = sqrt(|m|^2) / rho.
        double rho = df::to_double(q[0]);
        rho = rho > 1e-300 ? rho : 1e-300;
        const double r0 = (double)__builtin_amdgcn_rcpf((float)rho);
        const double r = (rho > 1e-30 && rho < 1e30) ? r0 * (2.0 - rho * r0) : 1.0 / rho;
        double m2 = 0.0;
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            const double m = df::to_double(q[1 + a]);
            m2 += m * m;
        }
        double pres = (gamma - 1.0) * (df::to_double(q[1 + ND]) - 0.5 * m2 * r);
        pres = pres > 0.0 ? pres : 0.0;
        return sqrt_fast(gamma * pres * r) + sqrt_fast(m2) * r;
    }
};

}  // namespace mh
