// Ghost fill and CFL time step, generic over the state form.
//
// Ghost fill reproduces Boundaries.hpp:11-149 in one launch. The reference
// fills x ghosts over every row, then y ghosts over every column (reading the
// x ghosts it just wrote), so a corner ends up as ymap(xmap(interior)). Here
// each thread owns one ghost cell and composes the two maps itself; every
// value read is an interior cell of `out` or a cell of `in` that no thread of
// the launch writes, so the single launch has no read-after-write hazard and
// gives the reference's bits.
//
// `in` is the stage's input state and `out` the state being completed (they
// are the same buffer for a stand-alone fill). UserFn faces are left to the
// caller's hook, as in the reference: their ghosts keep the values of `in`.
#pragma once

#include "functors.hpp"
#include "wave.hpp"

namespace mh {

struct GhostGrid {
    int32_t xc, yc, ng;
    int32_t bc[4];          // xs, xe, ys, ye (BoundaryType)
    double bconst[4][5];    // Constant-boundary conserved values per face
    double gamma;
    int32_t reduce_dt;      // fold the ghost cells into the CFL max
};

// Device-resident step clock: dt, time and output bookkeeping of main.cpp's
// loop, so a whole step can be replayed from a graph without host syncs.
struct Clock {
    double time;
    double dt;
    double max_time;
    double prev_output_time;
    double delta;
    double max_cfl;
    double dx;
    long long step;
};

// The running CFL max (bit pattern of the largest c_s + |v|; non-negative
// doubles order like their bit patterns) lives in kAmaxSlots words that follow
// the Clock in its allocation, each on a 256-byte line of its own. A wave
// folds its maximum into the slot its block index selects and the step's
// finalize reduces the slots. Device-scope atomics on one address are carried
// out one after another (about 12 ns apart on MI355X), so the ~4000 waves of a
// 2048^2 stage folding into a single word would queue for tens of
// microseconds while the launch drains.
constexpr int kAmaxSlots = 64;         // one per lane of the finalize wave
constexpr int kAmaxStride = 32;        // 64-bit words per slot (256 B)
constexpr size_t kAmaxOffset = 256;    // byte offset of slot 0 from the Clock
constexpr size_t kClockBytes = kAmaxOffset + sizeof(unsigned long long) * kAmaxSlots * kAmaxStride;
static_assert(sizeof(Clock) <= kAmaxOffset, "the Clock overlaps its CFL slots");

__host__ __device__ inline unsigned long long* amax_slots(Clock* c) {
    return reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(c) + kAmaxOffset);
}

// Fold a wave's maximum into the slot of `block` (lane 0 of the wave calls).
__device__ inline void fold_amax(unsigned long long* slots, uint32_t block, uint64_t m) {
    atomicMax(slots + kAmaxStride * (block & (uint32_t)(kAmaxSlots - 1)), (unsigned long long)m);
}

// The next step's dt from the CFL slots and the clock (compute_dt_impl's
// dt_max = dx / max(a), times max_cfl, with the output-time clamp and its
// nextafter loop, Hydro.cpp:208-240). Called by every lane of one wave (lane
// l reads slot l); every lane returns the same value. Reads only.
__device__ inline double clock_next_dt(const Clock* c) {
    static_assert(kAmaxSlots == kWave, "one slot per lane");
    const unsigned long long* slot =
        amax_slots(const_cast<Clock*>(c)) + kAmaxStride * (threadIdx.x & (unsigned)(kWave - 1));
    const double amax = __builtin_bit_cast(double, wave_max_u64(*slot));
    const double dt_min = c->dx / amax;  // == min over cells of dx / a
    double dt = dt_min * c->max_cfl;
    const double t = c->time;
    const double next_write_time = smin(c->max_time, c->prev_output_time + c->delta);
    if (t + dt >= next_write_time) {
        dt = next_write_time - t;
        while (t + dt < next_write_time) {
            dt = nextafter(dt, 1.7976931348623157e308);
        }
    }
    return dt;
}

// time += dt and, once an output time is reached, prev_output_time = time
// (main.cpp:62-71). One thread.
__device__ inline void clock_advance(Clock* c) {
    c->time = c->time + c->dt;
    c->step += 1;
    if (c->time >= c->prev_output_time + c->delta) c->prev_output_time = c->time;
}

// Source index along an axis of extent `dims` for ghost coordinate `g`
// (fill_one_bc_impl, Boundaries.hpp:27-57).
__host__ __device__ inline int ghost_source(int g, int dims, int ng, BoundaryType b) {
    const bool start = g < ng;
    switch (b) {
        case BoundaryType::Periodic:
            return start ? g + (dims - 2 * ng) : g - (dims - 2 * ng);
        case BoundaryType::ZeroGrad:
            return start ? ng : (dims - 1) - ng;
        default:  // Wall, Symmetric, SymmetricOutflowDiode: mirror
            return start ? (2 * ng - 1) - g : 2 * dims - 2 * ng - 1 - g;
    }
}

// Source index of index i of an axis with n interior cells and ng ghosts on
// each side when its ghosts are copies: ghost_source() for Periodic and
// ZeroGrad; interior indices map to themselves. Readers of a state whose
// ghosts were not filled use it in place of the fill.
__host__ __device__ inline int copy_ghost_source(int i, int n, int ng, bool periodic) {
    if (i < ng) return periodic ? i + n : ng;
    if (i >= ng + n) return periodic ? i - n : ng + n - 1;
    return i;
}

// Modification after the copy (Boundaries.hpp:83-115).
template <class Form, int ND, int NV, int Axis>
__device__ inline void ghost_modify(typename Form::raw (&q)[NV], BoundaryType b, bool start, const double* cvals) {
    constexpr int IM = Momentum<Axis, ND>();
    if (b == BoundaryType::Wall) {
        q[IM] = Form::negate(q[IM]);
    } else if (b == BoundaryType::SymmetricOutflowDiode) {
        Form::template diode<ND, NV, IM>(q, start);
    } else if (b == BoundaryType::Constant) {
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = Form::from_double(cvals[v]);
    }
}

template <class Form, int ND>
__global__ __launch_bounds__(256) void ghost_fill_kernel(GhostGrid g, typename Form::Planes in,
                                                         typename Form::Planes out, unsigned long long* amax_bits) {
    using Raw = typename Form::raw;
    constexpr int NV = 2 + ND;
    const int xc = g.xc, yc = g.yc, ng = g.ng;
    const int plane = xc * yc;
    const int nxband = 2 * ng * yc;
    const int nyband = ND > 1 ? 2 * ng * xc : 0;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    bool owner = false;
    Raw q[NV];

    // value the x pass leaves at (j, i), i an x-ghost column
    auto xvalue = [&](int j, int i, Raw (&v)[NV]) {
        const bool start = i < ng;
        const BoundaryType b = (BoundaryType)(start ? g.bc[0] : g.bc[1]);
        if (b == BoundaryType::UserFn) {
#pragma unroll
            for (int k = 0; k < NV; ++k) v[k] = Form::get(in, k * plane + j * xc + i);
            return;
        }
        if (b != BoundaryType::Constant) {
            const int is = ghost_source(i, xc, ng, b);
            const bool yghost = ND > 1 && (j < ng || j >= yc - ng);
            const auto& src = yghost ? in : out;
#pragma unroll
            for (int k = 0; k < NV; ++k) v[k] = Form::get(src, k * plane + j * xc + is);
        }
        ghost_modify<Form, ND, NV, 0>(v, b, start, start ? g.bconst[0] : g.bconst[1]);
    };

    if (tid < nxband) {
        const int p = tid % (2 * ng);
        const int j = tid / (2 * ng);
        const int i = p < ng ? p : (xc - 1) - (p - ng);
        bool skip = false;
        if (ND > 1 && (j < ng || j >= yc - ng)) {
            // y-ghost rows are completed by the y pass unless that face is UserFn
            const BoundaryType by = (BoundaryType)(j < ng ? g.bc[2] : g.bc[3]);
            skip = by != BoundaryType::UserFn;
        }
        if (!skip) {
            xvalue(j, i, q);
#pragma unroll
            for (int k = 0; k < NV; ++k) Form::put(out, k * plane + j * xc + i, q[k]);
            owner = true;
        }
    } else if (tid < nxband + nyband) {
        const int t = tid - nxband;
        const int p = t / xc;
        const int i = t % xc;
        const int j = p < ng ? p : (yc - 1) - (p - ng);
        const bool start = j < ng;
        const BoundaryType b = (BoundaryType)(start ? g.bc[2] : g.bc[3]);
        if (b == BoundaryType::UserFn) {
#pragma unroll
            for (int k = 0; k < NV; ++k) q[k] = Form::get(in, k * plane + j * xc + i);
        } else {
            if (b != BoundaryType::Constant) {
                const int js = ghost_source(j, yc, ng, b);
                if (i < ng || i >= xc - ng) {
                    xvalue(js, i, q);  // the x pass wrote this cell first
                } else {
#pragma unroll
                    for (int k = 0; k < NV; ++k) q[k] = Form::get(out, k * plane + js * xc + i);
                }
            }
            if constexpr (ND > 1) {
                ghost_modify<Form, ND, NV, 1>(q, b, start, start ? g.bconst[2] : g.bconst[3]);
            }
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) Form::put(out, k * plane + j * xc + i, q[k]);
        owner = true;
    }
    if (g.reduce_dt) {
        double a = 0.0;
        if (owner) a = Form::template signal_speed_nd<ND>(g.gamma, q);
        const uint64_t m = wave_max_u64(__builtin_bit_cast(uint64_t, a));
        if ((threadIdx.x & 63) == 0 && m != 0) {
            // 1-D fills also serve single-word maxima (one per batched problem)
            if constexpr (ND > 1) fold_amax(amax_bits, blockIdx.x, m);
            else atomicMax(amax_bits, (unsigned long long)m);
        }
    }
}

// CFL reduction over every cell, ghosts included (compute_dt_impl,
// Hydro.cpp:208-229): max of c_s + |v|, folded into the clock's CFL slots.
template <class Form, int ND>
__global__ __launch_bounds__(256) void cfl_reduce_kernel(typename Form::Planes q_in, int ncells, double gamma,
                                                         unsigned long long* amax_bits) {
    using Raw = typename Form::raw;
    constexpr int NV = 2 + ND;
    double amax = 0.0;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < ncells; idx += gridDim.x * blockDim.x) {
        Raw q[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) q[k] = Form::get(q_in, k * ncells + idx);
        const double a = Form::template signal_speed_nd<ND>(gamma, q);
        amax = a > amax ? a : amax;
    }
    const uint64_t m = wave_max_u64(__builtin_bit_cast(uint64_t, amax));
    if ((threadIdx.x & 63) == 0 && m != 0) fold_amax(amax_bits, blockIdx.x, m);
}

}  // namespace mh
