// FP64 state form: the drop-in. State and arithmetic are double, every
// expression keeps the reference's grouping (compile with -ffp-contract=off).
#pragma once

#include "functors.hpp"
#include "wave.hpp"

namespace mh {

struct FormF64 {
    using real = double;
    using raw = double;
    struct Planes {
        double* p;  // [var][k][j][i]
    };
    struct Ptrs {
        Planes qs;         // Q^(s): input of the stage, ghosts filled
        Planes qn;         // Q^n: state at the start of the step
        Planes out;        // Q^(s+1)
        const double* dt;  // device copy of the step's dt
        double dx;
        double gamma;
        unsigned long long* amax_bits;  // CFL max of c_s + |v| (bit pattern)
        unsigned int* floor_hits;       // unused in this form
        double* frx = nullptr;          // with tracers: density flux of x faces [j][i] (face i of row j)
        double* fry = nullptr;          // ... of y faces [j][i] (face j of column i)
        int32_t gravity = 0;            // a gravity source term is registered
        double gx = 0.0, gy = 0.0;      // its acceleration
        const double* dz3 = nullptr;    // 3-D: z flux difference Fz_k - Fz_{k+1} per interior cell [var][k][j][i]
        double gz = 0.0;
    };

    __device__ static inline double get(const Planes& a, int idx) { return a.p[idx]; }
    // Buffer-resource access for the plain march: the wave-uniform part of
    // the element index goes in a scalar offset, the lane's part in a vector
    // offset that stays fixed over the march (no per-load 64-bit address math).
    struct Buf {
        __amdgpu_buffer_rsrc_t p;
    };
    // grids with copy-type faces read ghost rows and columns from their
    // sources instead of filling them after every stage (StageGrid::virt)
    static constexpr bool kVirtualGhosts = true;
    static constexpr bool kHasFloors = false;  // positivity floors (floors_on)
    // the instance without buffer resources loads Q^(s) and Q^n for the
    // combine just before it (fewer registers live across the sweeps)
    static constexpr bool kEarlyOperands = false;
    static constexpr uint32_t kElem = 8;
    __device__ static inline Buf buf(const Planes& a) { return Buf{make_rsrc(a.p)}; }
    __device__ static inline double bget(const Buf& b, uint32_t voff, uint32_t s_elem) {
        return buf_load_f64(b.p, voff, s_elem * kElem);
    }
    __device__ static inline void bput(const Buf& b, uint32_t voff, uint32_t s_elem, double v) {
        buf_store_f64(v, b.p, voff, s_elem * kElem);
    }
    // direct-to-LDS load of an element as its low and high words
    __device__ static inline void dma(const Buf& b, uint32_t* lo, uint32_t* hi, uint32_t voff, uint32_t s_elem) {
        buf_load_lds(b.p, lo, voff, s_elem * kElem);
        buf_load_lds(b.p, hi, voff + 4u, s_elem * kElem);
    }
    __device__ static inline double from_words(uint32_t lo, uint32_t hi) {
        return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
    }
    __device__ static inline void put(const Planes& a, int idx, double v) { a.p[idx] = v; }
    __device__ static inline bool needs_qn(int kind) { return kind == 1 || kind == 2 || kind == 3 || kind == 5; }
    __device__ static inline double negate(double v) { return -v; }
    __device__ static inline double from_double(double v) { return v; }
    __device__ static inline double value(double v) { return v; }
    __device__ static inline const void* get_ptr(const Planes& a) { return a.p; }
    __device__ static inline double add_double(double v, double d) { return v + d; }

    // SymmetricOutflowDiode momentum clamp and energy correction,
    // Boundaries.hpp:96-106.
    template <int ND, int NV, int IM>
    __device__ static inline void diode(double (&q)[NV], bool start) {
        const double mom_pre = q[IM];
        q[IM] = start ? smin(mom_pre, 0.0) : smax(mom_pre, 0.0);
        q[Cons<ND>::Ene] -= square(q[IM] - mom_pre) / q[Cons<ND>::Rho];
    }

    template <int ND, bool Floors = true, int NV>
    __device__ static inline bool to_prim_nd(const Ptrs& p, const double (&q)[NV], double (&w)[NV]) {
        cons_to_prim<ND>(p.gamma, q, w);
        return false;  // no floors in the drop-in
    }

    // no floors in the drop-in (the reference has none)
    template <int ND, bool Floors = true, int NV>
    __device__ static inline void floor_face(const Ptrs&, double (&)[NV]) {}
    template <int ND, bool Floors = true, int NV>
    __device__ static inline bool floor_state(const Ptrs&, double (&)[NV]) { return false; }
    static inline bool floors_on(const Ptrs&) { return false; }

    // TimeStepping.cpp stage lambdas (RK2 60-67/80-87, SSPRK3 137-144/157-164/
    // 177-184, SSPRK(4,3) 234-301), grouped exactly as written there. Generic
    // over the real type so that the tangent-linear stage (dual numbers)
    // evaluates the same expressions; T = double is the drop-in's.
    template <class T>
    __device__ static inline T combine(int kind, T Q, T Qn, T dq, T src) {
        switch (kind) {
            case 0: return Q + (dq + src);
            case 1: return T(0.5) * (Qn + Q + dq + src);
            case 2: return T(0.75) * Qn + T(0.25) * (Q + dq + src);
            case 3: return T(1.0 / 3.0) * Qn + T(2.0 / 3.0) * (Q + dq + src);
            case 4: return Q + T(0.5) * (dq + src);
            default: return T(2.0 / 3.0) * Qn + T(1.0 / 3.0) * Q + T(1.0 / 6.0) * (dq + src);
        }
    }

    // ... without a registered source: S*dt = 0*dt = +0, as the reference adds it
    template <class T>
    __device__ static inline T combine_nosrc(int kind, T Q, T Qn, T dq) {
        return combine(kind, Q, Qn, dq, T(0.0));
    }
    // the drop-in keeps every operation of the reference, adds of zeros included
    static constexpr bool kExactZeros = true;

    // Source S*dt of the stage lambdas for a registered gravity term:
    // gravity_kernel (Gravity.cpp:22-38) accumulated into a zeroed S from the
    // stage input Q of the cell, then multiplied by dt (generic like combine).
    template <int ND, int NV, class T>
    __device__ static inline void gravity_src(const Ptrs& p, const T (&q)[NV], T dt, T (&src)[NV]) {
        src[0] = T(0.0) * dt;
        src[1] = (T(0.0) + q[0] * T(p.gx)) * dt;
        T energy_update = q[1] * T(p.gx);
        if constexpr (ND > 1) {
            src[2] = (T(0.0) + q[0] * T(p.gy)) * dt;
            energy_update += q[2] * T(p.gy);
        }
        if constexpr (ND > 2) {
            src[3] = (T(0.0) + q[0] * T(p.gz)) * dt;
            energy_update += q[3] * T(p.gz);
        }
        src[1 + ND] = (T(0.0) + energy_update) * dt;
    }

    // dt_reducer, Hydro.cpp:192-206: cons_to_prim then c_s + |v|.
    template <int ND, int NV>
    __device__ static inline double signal_speed_nd(double gamma, const double (&q)[NV]) {
        double w[NV];
        cons_to_prim<ND>(gamma, q, w);
        return mh::signal_speed<ND>(gamma, w);
    }
};

}  // namespace mh
