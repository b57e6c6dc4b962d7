// Fused SSPRK stage for 2-D unsplit finite volumes: the "marching strip".
//
// One wavefront owns a strip of 62 output columns and marches down a span of
// rows. Lane t holds column c = i0 - 1 + t (the 62 outputs are lanes 1..62,
// lanes 0 and 63 supply the outer reconstructions and faces). Per output row j:
//
//   1. load row j+1+R of Q^(s) (coalesced, x fastest), cons->prim, and append
//      the lane's own column value to a (2R+2)-row register window (the y
//      stencil of cell j+1 and row j);
//   2. y sweep: reconstruct cell j+1 from the register window, Riemann flux on
//      face j+1 from (RR of cell j, RL of cell j+1), dy = Gy_j - Gy_{j+1};
//   3. x sweep: row j's primitives (and the halo columns) go to a one-row LDS
//      exchange; reconstruct row j from it, take RR of the left neighbour
//      lane, Riemann flux on the face, dx = Fx_c - Fx_{c+1} (flux of the
//      right neighbour lane);
//   4. combine in registers: dq = (0 + dx) + dy, dq *= dt/dx, stage update with
//      Q^(s) and Q^n, store Q^(s+1); optionally fold |v|+c_s into a running
//      max for the next CFL time step.
//
// HBM traffic per cell and stage is one read of Q^(s) (plus a 6-column /
// 6-row halo, mostly served from L2), one read of Q^n for stages >= 1 and one
// write of Q^(s+1). R is the stencil order of the reconstruction (FOG 0,
// MUSCL 1, PPM and WENO5-Z 2).
//
// The arithmetic follows Mosscap's unfused sequence (Hydro.cpp:8-166,
// TimeStepping.cpp:9-35 and the stage lambdas): both sweeps read the same
// primitive field (unsplit), the x difference is added before the y
// difference, and the stage formulas keep their grouping. The state form
// (`Form`) supplies loads, cons->prim, the stage combine and stores, so the
// same march serves the FP64 drop-in and the compensated FP32 form.
#pragma once

#include "functors.hpp"
#include "ghost_dt.hpp"
#include "wave.hpp"
#include <utility>

namespace mh {

constexpr int kStripOut = 62;  // output columns per wavefront

// Stencil rows of Q^(s) in flight per wavefront of the plain march (1 or 2);
// each translation unit of stage kernels may set its own.
#ifndef MH_STAGE_LEAD_ROWS
#define MH_STAGE_LEAD_ROWS 1
#endif
// The plain march keeps the raw Q^(s) of the last ORD+1 stencil rows in
// registers for the stage combine instead of loading row j again for it:
// 0 never, 1 always (but not in the floors instance of WENO5-Z, which has no
// registers to spare), 2 for WENO5-Z only (the FP64 form, whose other
// schemes would lose a wave per SIMD to the registers).
#ifndef MH_STAGE_QS_RING
#define MH_STAGE_QS_RING 0
#endif
// Plain instance with the Q^(s) ring: 1 = the march runs its rows in groups
// of the ring's length (the ring rotates by renaming); 0 = one row at a time
// with the ring's registers moved.
#ifndef MH_STAGE_RING_UNROLL
#define MH_STAGE_RING_UNROLL 0
#endif
// 3-D plain march: 1 = the z flux differences of the next row go straight
// into LDS with the combine operands; 0 = they are loaded into registers at
// the top of the row they are added in.
#ifndef MH_STAGE_DZ_LDS
#define MH_STAGE_DZ_LDS 0
#endif

// Stage update kinds: the lambdas of TimeStepping.cpp, src = S*dt with S = 0.
enum StageKind : int32_t {
    kAddFull = 0,      // Q + (dq + src)                       RK2/SSPRK3 stage 0
    kRk2Final = 1,     // 0.5*(((Qn + Q) + dq) + src)          RK2 stage 1
    kSsp3Second = 2,   // 0.75*Qn + 0.25*((Q + dq) + src)      SSPRK3 stage 1
    kSsp3Third = 3,    // (1/3)*Qn + (2/3)*((Q + dq) + src)    SSPRK3 stage 2
    kAddHalf = 4,      // Q + 0.5*(dq + src)                   SSPRK(4,3) stages 0, 1, 3
    kSsp4Third = 5,    // ((2/3)*Qn + (1/3)*Q) + (1/6)*(dq + src)   SSPRK(4,3) stage 2
};

// Instances of the fused stage. Plain: buffer-resource addressing, operands
// staged through LDS, no floors, gravity source or tracer stores. Floors: the
// plain instance with the compensated form's positivity floors. Extras:
// gravity source, tracers' density-flux stores, floors and 64-bit addressing.
enum StageInst : int { kStagePlain = 0, kStageFloors = 1, kStageExtras = 2 };

struct StageGrid {
    int32_t xc, yc, ng, nx, ny;
    int32_t strips, nspans;  // 62-column strips, spans of rows per plane
    int32_t kind;        // StageKind
    int32_t reduce_dt;   // fold the new state into the CFL max
    int32_t zc = 1, nz = 1;  // 3-D: planes (ghosts included) and interior planes
    // Every face periodic or zerograd: ghost rows, columns (and, for the z
    // sweep, planes) are read from their source cells (their fill is a plain
    // copy), so the stage needs no ghost fill of its input. xper / yper /
    // zper: periodic along that axis, else zerograd.
    int32_t virt = 0, xper = 0, yper = 0, zper = 0;
    // The step's clock work folded into its stages (graph-replayed steps of
    // three or more stages): 1 = this first stage takes the step's dt from
    // the CFL slots and the clock itself (every wave, read-only) and block 0
    // stores it in the clock; 2 = block 0 of this second stage advances the
    // clock's time and clears the slots for the last stage's fold. 0 = the
    // stage reads dt from the clock.
    int32_t fin = 0;
    Clock* clock = nullptr;
};


// The x-y march of one plane. ND = 2: the whole 2-D stage. ND = 3: one
// interior z-plane per (strip, span, plane) workgroup; the z flux difference
// of each cell comes from the z sweep (p.dz3) and is added after the y one,
// as integrate_flux adds Fz after Fy (TimeStepping.cpp:22-30).
// Inst selects what is compiled in (their addresses, operands and branches
// cost registers the plain hydro stage needs); the launcher picks the
// instance from the pointers and options it is given.
template <class Form, Reconstruction R, SlopeLimiter SL, RiemannSolver RS, int ND = 2, int Inst = kStageExtras>
__device__ inline void stage2d_march(const StageGrid& g, const typename Form::Ptrs& p) {
    constexpr bool Extras = Inst == kStageExtras;
    using T = typename Form::real;
    using Raw = typename Form::raw;
    constexpr int NV = 2 + ND;  // rho, m_x, m_y[, m_z], E
    constexpr int ORD = stencil_order(R);
    // register window of the lane's own column: rows j+1-ORD .. j+1+ORD (the
    // y stencil of cell j+1), which holds row j (whose x sweep runs this
    // iteration) except for FOG, whose window reaches one row further back
    constexpr int WIN = 2 * ORD + 1 + (ORD == 0 ? 1 : 0);
    constexpr int WY = WIN - 1 - ORD;  // window index of cell j+1
    constexpr int WX = WY - 1;         // ... and of row j
    constexpr int ROW = kWave + 2 * ORD;

    // row j's primitives with the halo columns, exchanged for the x stencils
    __shared__ T xrow[NV][ROW];

    const int lane = threadIdx.x;
    const uint32_t bb = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t per_plane = (uint32_t)g.strips * (uint32_t)g.nspans;
    const uint32_t b = bb % per_plane;
    const int strip = (int)(b % (uint32_t)g.strips);
    const int span = (int)(b / (uint32_t)g.strips);
    const int i0 = g.ng + strip * kStripOut;
    // rows of this span: the ny rows split into nspans spans of equal height (+-1)
    const int j0 = g.ng + (int)(((int64_t)span * g.ny) / g.nspans);
    const int j1 = g.ng + (int)(((int64_t)(span + 1) * g.ny) / g.nspans);
    const int c = i0 - 1 + lane;
    // ghosts read from their sources (StageGrid::virt; compiled for the forms that use it)
    constexpr bool kVirt = Form::kVirtualGhosts;
    const int c_ld = (kVirt && g.virt) ? copy_ghost_source(min(c, g.xc - 1), g.nx, g.ng, g.xper) : min(c, g.xc - 1);
    // source row of row r (wave-uniform)
    auto src_row = [&](int r) { return (kVirt && g.virt) ? copy_ghost_source(r, g.ny, g.ng, g.yper) : r; };
    const int xc = g.xc;
    // variables are [var][k][j][i]: var stride of the whole grid, then this plane's offset
    const int plane = g.xc * g.yc * (ND > 2 ? g.zc : 1);
    const int pk = ND > 2 ? (g.ng + (int)(bb / per_plane)) * g.xc * g.yc : 0;

    // Halo columns: exchange-row entries [0, ORD) (columns left of the strip)
    // and [64+ORD, 64+2*ORD) (right of it). They are converted to primitives
    // in batches of HB rows, one lane per (row, column) pair, into `hbuf`
    // (two batches deep: a batch is converted when its first row is
    // committed, up to ORD+1 rows before that row's x sweep), and lanes
    // [0, 2*ORD) copy a row's entries into the exchange row: one cons->prim
    // pass per HB rows instead of a divergent pass per row.
    constexpr int NH = 2 * ORD;
    // 4 rows: the batch reads at most 3 rows ahead of the neighbouring strips'
    // waves, whose interior these columns are, so the lines are still in L2
    // (16-row batches measured 11 % more HBM reads and a 1.5 % slower FP32C
    // stage); lanes [0, 4*2*ORD) take part.
    constexpr int HB = NH > 0 ? 4 : 1;
    static_assert(HB * NH <= kWave, "one lane per (row, column) of a batch");
    static_assert((HB & (HB - 1)) == 0, "batch rows must be a power of two");
    constexpr int HSLOTS = 2 * HB;
    // the batch holding row j + HSLOTS is converted at iteration j + HSLOTS -
    // (HB-1) - 1 - ORD or later, not before row j's x sweep (iteration j)
    static_assert(ORD <= HSLOTS - HB, "halo batches overwritten before their rows are swept");
    __shared__ T hbuf[HSLOTS][NV][NH > 0 ? NH : 1];
    const bool halo_lane = lane < NH;
    const int h_entry = lane < ORD ? lane : kWave + lane;
    const int r_first = j0 - 1 - ORD;  // first and last rows the march commits
    const int r_last = j1 + ORD;

    // The plain instance addresses Q through buffer resources: scalar offsets
    // for the wave-uniform row/plane/variable part, fixed per-lane vector
    // offsets (the launcher keeps every byte offset below 2^31).
    constexpr bool kBuf = !Extras;
    // positivity floors (the compensated form's option)
    constexpr bool kFloors = Inst != kStagePlain;
    const auto bqs = Form::buf(p.qs);
    const auto bqn = Form::buf(p.qn);
    const auto bout = Form::buf(p.out);
    constexpr uint32_t ES = Form::kElem;
    auto load_qs = [&](int col, int uniform) -> Raw {
        if constexpr (kBuf) return Form::bget(bqs, (uint32_t)col * ES, (uint32_t)uniform);
        else return Form::get(p.qs, uniform + col);
    };
    const EosT<T> eos = make_eos<T>((double)p.gamma);
    double dt_step;
    if (g.fin == 1) {
        dt_step = clock_next_dt(g.clock);
        if (blockIdx.x == 0 && lane == 0) g.clock->dt = dt_step;
    } else {
        dt_step = p.dt[0];
        if (g.fin == 2 && blockIdx.x == 0) {
            if (lane == 0) clock_advance(g.clock);
            amax_slots(g.clock)[kAmaxStride * lane] = 0ull;
        }
    }
    const T dt = (T)dt_step;
    const T dtdx = dt / (T)p.dx;
    const T src = T(0) * dt;

    T win[WIN][NV] = {};
    const bool out_lane = lane >= 1 && lane <= kStripOut && c < g.ng + g.nx;

    // Plain instance: the stage-combine operands of the next row (Q^(s) and
    // Q^n, two words per value) are loaded straight into LDS while this row
    // is swept: no registers held across the sweeps (the compensated form
    // held 16) and no load latency at the combine (the FP64 form's).
    constexpr bool kDmaOps = kBuf;
    // With the Q^(s) ring (kQsRing) only Q^n is loaded for the combine.
    constexpr bool kQsRing =
        kDmaOps && ((MH_STAGE_QS_RING == 1 && (Inst == kStagePlain || R != Reconstruction::Weno5Z)) ||
                    (MH_STAGE_QS_RING == 2 && R == Reconstruction::Weno5Z));
    constexpr int kQnSlot = kQsRing ? 0 : 1;
    __shared__ uint32_t opnd[kDmaOps && !kQsRing ? 2 : 1][kDmaOps ? NV : 1][2][kWave];  // [Q^(s), Q^n][var][word][lane]
    // 3-D: the z flux differences of the combined row (the z sweep's output,
    // one value per variable and interior cell; lanes outside read a clamped
    // column that they do not use)
    constexpr bool kDzLds = ND > 2 && kDmaOps && MH_STAGE_DZ_LDS != 0;
    constexpr int kDzWords = sizeof(T) / 4;
    __shared__ uint32_t dzo[kDzLds ? NV : 1][kDzLds ? kDzWords : 1][kWave];
    const int c_dz = min(c, g.xc - 1);
    auto issue_operands = [&](int jr) {
        if constexpr (kDmaOps) {
            const int row = pk + jr * xc;
            if constexpr (kDzLds) {
                const auto bdz = make_rsrc(p.dz3);
#pragma unroll
                for (int v = 0; v < NV; ++v)
#pragma unroll
                    for (int w = 0; w < kDzWords; ++w)
                        buf_load_lds(bdz, &dzo[v][w][0], (uint32_t)c_dz * (uint32_t)sizeof(T) + 4u * w,
                                     (uint32_t)(v * plane + row) * (uint32_t)sizeof(T));
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if (!kQsRing)
                    Form::dma(bqs, &opnd[0][v][0][0], &opnd[0][v][1][0], (uint32_t)c_ld * ES,
                              (uint32_t)(v * plane + row));
                if (Form::needs_qn(g.kind))
                    Form::dma(bqn, &opnd[kQnSlot][v][0][0], &opnd[kQnSlot][v][1][0], (uint32_t)c_ld * ES,
                              (uint32_t)(v * plane + row));
            }
        }
    };
    unsigned int floor_hits = 0;
    // bit k: a floor fired in the cons->prim of the row committed k rows ago
    // (the row combined next was committed ORD + 1 rows before it)
    uint32_t in_floor_bits = 0;

    auto commit_row = [&](int r, const Raw (&q)[NV]) {
        if constexpr (NH > 0) {
            if (((r - r_first) & (HB - 1)) == 0) {  // wave-uniform: convert the halo of rows r .. r+HB-1
                wave_lds_sync();  // the previous batch's reads are done
                const int hb_idx = lane % NH;  // this lane's halo column
                const int hb_row = lane / NH;  // and row offset in the batch
                int hb_col = min(i0 - 1 - ORD + (hb_idx < ORD ? hb_idx : kWave + hb_idx), g.xc - 1);
                const int rr = r + hb_row;
                if (hb_row < HB && rr <= r_last) {
                    int rs = rr;
                    if (kVirt && g.virt) {
                        hb_col = copy_ghost_source(hb_col, g.nx, g.ng, g.xper);
                        rs = copy_ghost_source(rr, g.ny, g.ng, g.yper);
                    }
                    Raw qh[NV];
                    // the lane's row goes in the vector offset (the scalar one is wave-uniform)
#pragma unroll
                    for (int v = 0; v < NV; ++v) qh[v] = load_qs(rs * xc + hb_col, v * plane + pk);
                    T wh[NV];
                    Form::template to_prim_nd<ND, kFloors>(p, qh, wh);
#pragma unroll
                    for (int v = 0; v < NV; ++v) hbuf[(rr - r_first) & (HSLOTS - 1)][v][hb_idx] = wh[v];
                }
                wave_lds_sync();
            }
        }
        T w[NV];
        const bool hit = Form::template to_prim_nd<ND, kFloors>(p, q, w);
        in_floor_bits = (in_floor_bits << 1) | (hit ? 1u : 0u);
#pragma unroll
        for (int k = 0; k + 1 < WIN; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) win[k][v] = win[k + 1][v];
#pragma unroll
        for (int v = 0; v < NV; ++v) win[WIN - 1][v] = w[v];
    };
    auto fetch_row = [&](int r, Raw (&q)[NV]) {
        const int base = pk + src_row(r) * xc;
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = load_qs(c_ld, v * plane + base);
    };
    // Plain instance: the stencil rows of Q^(s) go straight into LDS too, so
    // no registers hold them across a row. With kLeadRows = 1 a row is issued
    // at the top of the iteration before the one that reads it and is
    // complete by that iteration's combine, which waits for every load. With
    // kLeadRows = 2 (MH_STAGE_LEAD_ROWS) it is issued two iterations ahead
    // into one of two slots (row r in slot r & 1), and the combine waits for
    // every load but the lead row issued at its own top (loads complete in
    // order), so a lead row has two rows of sweeps to arrive. The 3-D FP64
    // march takes 2 (fewer spilled registers, 21 % faster); the others 1 (no
    // gain, and the 3-D FP32C march loses a wave per SIMD to the LDS).
    constexpr int kLeadRows = MH_STAGE_LEAD_ROWS;
    static_assert(kLeadRows == 1 || kLeadRows == 2, "lead rows in flight");
    // two slots with two rows in flight, or with the Q^(s) ring (a row's raw
    // words are read again at the end of the iteration that committed it)
    constexpr int kLeadSlots = (kLeadRows == 2 || kQsRing) ? 2 : 1;
    __shared__ uint32_t lead[kDmaOps ? kLeadSlots : 1][kDmaOps ? NV : 1][2][kWave];  // [slot][var][word][lane]
    constexpr int kLeadOps = 2 * NV;  // vector-memory operations of one issue_lead
    auto lead_slot = [&](int r) { return kLeadSlots == 2 ? (r & 1) : 0; };
    // raw Q^(s) of rows j .. j+ORD at the combine of row j (kQsRing)
    Raw qsr[kQsRing ? ORD + 1 : 1][NV];
    auto issue_lead = [&](int r) {
        if constexpr (kDmaOps) {
            const int base = pk + src_row(r) * xc;
            const int sl = lead_slot(r);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                Form::dma(bqs, &lead[sl][v][0][0], &lead[sl][v][1][0], (uint32_t)c_ld * ES,
                          (uint32_t)(v * plane + base));
        }
    };
    auto recon_y = [&](T (&wl)[NV], T (&wr)[NV]) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            auto s = [&](int o) -> T { return win[WY + o][v]; };
            reconstruct_stencil<R, SL, T>(s, wl[v], wr[v]);
        }
        Form::template floor_face<ND, kFloors>(p, wl);
        Form::template floor_face<ND, kFloors>(p, wr);
    };

    // ---- prologue: rows j0-1-ORD .. j0+ORD; y faces up to j0 ------------------
    Raw qa[NV];
    T yRp[NV], Gyp[NV];
    {
        Raw qb[NV];
#pragma unroll
        for (int k = 0; k < 2 * ORD + 1; ++k) {  // rows j0-1-ORD .. j0-1+ORD
            const int r = j0 - 1 - ORD + k;
            fetch_row(r, qb);
            commit_row(r, qb);
            if constexpr (kQsRing) {
                if (k >= ORD + 1) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) qsr[k - ORD - 1][v] = qb[v];
                }
            }
        }
        T yl[NV], yr[NV];
        recon_y(yl, yr);  // cell j0-1
#pragma unroll
        for (int v = 0; v < NV; ++v) yRp[v] = yr[v];
        fetch_row(j0 + ORD, qb);
        commit_row(j0 + ORD, qb);
        if constexpr (kQsRing) {
#pragma unroll
            for (int v = 0; v < NV; ++v) qsr[ORD][v] = qb[v];
        }
        recon_y(yl, yr);  // cell j0
        riemann_flux<RS, 1, ND>(eos, yRp, yl, Gyp);
        if (Extras && ND == 2 && p.fry != nullptr && out_lane && j0 < j1) p.fry[j0 * xc + c] = Gyp[0];
#pragma unroll
        for (int v = 0; v < NV; ++v) yRp[v] = yr[v];
        if constexpr (kDmaOps) {
            if (j0 < j1) {
                issue_lead(j0 + 1 + ORD);
                if (kLeadRows == 2 && j0 + 1 < j1) issue_lead(j0 + 2 + ORD);
                issue_operands(j0);
                wait_vm_loads();  // the first row reads both without a combine's wait before it
            }
        } else {
            if (j0 < j1) fetch_row(j0 + 1 + ORD, qa);
        }
    }

    double amax = 0.0;

    // one row of the march; rot: the Q^(s) ring's rotation (slot (k + rot) %
    // (ORD + 1) holds row j + k), or -1 to keep slot k = row j + k by moving
    // the ring's registers at the end of the row
    auto march_row = [&](const int j, auto rot_c) {
        constexpr int rot = decltype(rot_c)::value;
        // Stage-combine operands of this row: from LDS in the plain instance
        // (loaded during the previous row); otherwise loaded here or just
        // before the combine (Form::kEarlyOperands).
        const int row = pk + j * xc;
        Raw qs_c[NV], qn_c[NV];
        auto load_operands = [&] {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                qs_c[v] = load_qs(c_ld, v * plane + row);
                if constexpr (kBuf) {
                    qn_c[v] = Form::needs_qn(g.kind)
                                  ? Form::bget(bqn, (uint32_t)c_ld * ES, (uint32_t)(v * plane + row))
                                  : qs_c[v];
                } else {
                    qn_c[v] = Form::needs_qn(g.kind) ? Form::get(p.qn, v * plane + row + c_ld) : qs_c[v];
                }
            }
        };
        if constexpr (Form::kEarlyOperands && !kDmaOps) load_operands();
        if constexpr (kDmaOps) {
            wave_lds_sync();
            Raw ql[NV];
            const int sl = lead_slot(j + 1 + ORD);
#pragma unroll
            for (int v = 0; v < NV; ++v) ql[v] = Form::from_words(lead[sl][v][0][lane], lead[sl][v][1][lane]);
            commit_row(j + 1 + ORD, ql);
            if (j + kLeadRows < j1) {
                wait_lds_ops();  // the slot's words are read before the next row for it overwrites them
                wave_lds_sync();
                issue_lead(j + 1 + kLeadRows + ORD);
            }
        } else {
            commit_row(j + 1 + ORD, qa);
            if (j + 1 < j1) fetch_row(j + 2 + ORD, qa);  // prefetch the next row
        }

        // 3-D: this row's z flux differences into registers now, so the
        // load has the sweeps to arrive (without kDzLds)
        T dzr[ND > 2 && !kDzLds ? NV : 1];
        if constexpr (ND > 2 && !kDzLds) {
#pragma unroll
            for (int v = 0; v < NV; ++v) dzr[v] = p.dz3[v * plane + pk + j * xc + c_dz];
        }

        // y sweep: cell j+1, face j+1
        T yl[NV], yr[NV], Gy[NV], dy[NV];
        recon_y(yl, yr);
        riemann_flux<RS, 1, ND>(eos, yRp, yl, Gy);
        if (Extras && ND == 2 && p.fry != nullptr && out_lane) p.fry[(j + 1) * xc + c] = Gy[0];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            dy[v] = Gyp[v] - Gy[v];
            Gyp[v] = Gy[v];
            yRp[v] = yr[v];
        }

        // x sweep on row j: its primitives (window row WX) and halo entries
        // (their batch was converted ORD+1 or more rows ago) go to the
        // exchange row, which the previous row's x sweep has read (one
        // wavefront: its LDS operations complete in program order)
        wave_lds_sync();
#pragma unroll
        for (int v = 0; v < NV; ++v) xrow[v][lane + ORD] = win[WX][v];
        if constexpr (NH > 0) {
            if (halo_lane) {
#pragma unroll
                for (int v = 0; v < NV; ++v) xrow[v][h_entry] = hbuf[(j - r_first) & (HSLOTS - 1)][v][lane];
            }
        }
        wave_lds_sync();  // row j visible to every lane
        T xl[NV], xr[NV], xrm[NV], Fx[NV], Fxp[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            auto s = [&](int o) -> T { return xrow[v][lane + ORD + o]; };
            reconstruct_stencil<R, SL, T>(s, xl[v], xr[v]);
        }
        Form::template floor_face<ND, kFloors>(p, xl);
        Form::template floor_face<ND, kFloors>(p, xr);
#pragma unroll
        for (int v = 0; v < NV; ++v) xrm[v] = lane_from_left(xr[v]);
        riemann_flux<RS, 0, ND>(eos, xrm, xl, Fx);
        if (Extras && ND == 2 && p.frx != nullptr && lane >= 1 && c >= g.ng && c <= g.ng + g.nx) p.frx[j * xc + c] = Fx[0];
#pragma unroll
        for (int v = 0; v < NV; ++v) Fxp[v] = lane_from_right(Fx[v]);

        if constexpr (kDmaOps) {  // this row's operands, loaded into LDS during the previous row
            // every load (with two lead rows in flight: but the one issued at
            // this row's top, if any)
            if (kLeadRows == 2 && j + 2 < j1) wait_vm_at_most<kLeadOps>();
            else wait_vm_loads();
            wave_lds_sync();
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if constexpr (kQsRing) qs_c[v] = qsr[rot < 0 ? 0 : rot][v];
                else qs_c[v] = Form::from_words(opnd[0][v][0][lane], opnd[0][v][1][lane]);
                qn_c[v] = Form::needs_qn(g.kind)
                              ? Form::from_words(opnd[kQnSlot][v][0][lane], opnd[kQnSlot][v][1][lane])
                              : qs_c[v];
            }
        }
        if (out_lane) {
            if constexpr (!Form::kEarlyOperands && !kDmaOps) load_operands();
            Raw qo[NV];
            T srcv[NV];
            if (Extras && p.gravity) {
                Form::template gravity_src<ND>(p, qs_c, dt, srcv);
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) srcv[v] = src;
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                // integrate_flux's 0 + dx + dy (+ dz) (the compensated form
                // drops the leading 0, which only turns -0 into +0)
                T q_update = Form::kExactZeros ? T(0) + (Fx[v] - Fxp[v]) : Fx[v] - Fxp[v];
                q_update += dy[v];
                if constexpr (ND > 2) {
                    if constexpr (!kDzLds) q_update += dzr[v];
                    else if constexpr (kDzWords == 2) q_update += Form::from_words(dzo[v][0][lane], dzo[v][1][lane]);
                    else q_update += __builtin_bit_cast(T, dzo[v][0][lane]);
                }
                q_update *= dtdx;
                qo[v] = (Extras && p.gravity) ? Form::combine(g.kind, qs_c[v], qn_c[v], q_update, srcv[v])
                                              : Form::combine_nosrc(g.kind, qs_c[v], qn_c[v], q_update);
            }
            // each interior cell is combined by exactly one (strip, span) owner
            const bool st_hit = Form::template floor_state<ND, kFloors>(p, qo);
            floor_hits += (st_hit || ((in_floor_bits >> (ORD + 1)) & 1u)) ? 1u : 0u;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if constexpr (kBuf) Form::bput(bout, (uint32_t)c * ES, (uint32_t)(v * plane + row), qo[v]);
                else Form::put(p.out, v * plane + row + c, qo[v]);
            }
            if (g.reduce_dt) {
                const double a = Form::template signal_speed_nd<ND>((double)p.gamma, qo);
                amax = a > amax ? a : amax;
            }
        }
        if constexpr (kQsRing) {  // drop row j, take row j+1+ORD (committed at this row's top)
            if (j + 1 < j1) {
                if constexpr (rot < 0) {
#pragma unroll
                    for (int k = 0; k < ORD; ++k)
#pragma unroll
                        for (int v = 0; v < NV; ++v) qsr[k][v] = qsr[k + 1][v];
                }
                const int sl = lead_slot(j + 1 + ORD);
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    qsr[rot < 0 ? ORD : rot][v] = Form::from_words(lead[sl][v][0][lane], lead[sl][v][1][lane]);
            }
        }
        if constexpr (kDmaOps) {
            if (j + 1 < j1) {
                wait_lds_ops();  // this row's operand words are read before the next row's overwrite them
                wave_lds_sync();
                issue_operands(j + 1);
            }
        }
    };
    if constexpr (kQsRing && Inst == kStagePlain && MH_STAGE_RING_UNROLL) {
        // rows in groups of the ring's length: each copy of the row addresses
        // fixed slots, so the ring rotates without moving registers
        constexpr int kRing = ORD + 1;
        int j = j0;
        for (; j + kRing <= j1; j += kRing)
            [&]<int... I>(std::integer_sequence<int, I...>) {
                (march_row(j + I, std::integral_constant<int, I>{}), ...);
            }(std::make_integer_sequence<int, kRing>{});
        for (; j < j1; ++j) march_row(j, std::integral_constant<int, -1>{});
    } else {
        for (int j = j0; j < j1; ++j) march_row(j, std::integral_constant<int, -1>{});
    }

    if (g.reduce_dt) {
        const uint64_t m = wave_max_u64(__builtin_bit_cast(uint64_t, amax));
        if (lane == 0) fold_amax((unsigned long long*)p.amax_bits, blockIdx.x, m);
    }
    if (p.floor_hits != nullptr) {
        const unsigned long long any = __ballot(floor_hits != 0u);
        if (any != 0ull) {
            const unsigned int total = wave_sum_u32(floor_hits);
            if (lane == 0) atomicAdd(p.floor_hits, total);
        }
    }
}

// 1-D stage: one wavefront covers 62 output cells of a (segment of a) row;
// lanes are cells c = i0 - 1 + t, the halo entries come from lanes [0, 2R).
// seg_load issues the segment's loads (raw Q^(s), Q^n and the halo columns),
// seg_compute does the rest, so a caller can overlap one segment's loads with
// the previous segment's arithmetic.
template <class Form>
struct SegRaw {
    typename Form::raw q[3], qn[3], qh[3];
};

template <class Form, Reconstruction R>
__device__ inline void seg_load(const typename Form::Ptrs& p, int kind, int base, int xc, int ng, int nx, int i0,
                                int virt, bool per, SegRaw<Form>& s) {
    constexpr int NV = 3;  // rho, mx, E
    constexpr int ORD = stencil_order(R);
    const int lane = threadIdx.x & 63;
    const int c = i0 - 1 + lane;
    // virt: ghosts are read from their sources (periodic `per`, else zerograd)
    const int c_ld = virt ? copy_ghost_source(min(c, xc - 1), nx, ng, per) : min(c, xc - 1);
    const bool halo_lane = lane < 2 * ORD;
    const int h_entry = lane < ORD ? lane : kWave + lane;
    const int h_col = virt ? copy_ghost_source(min(i0 - 1 - ORD + h_entry, xc - 1), nx, ng, per)
                           : min(i0 - 1 - ORD + h_entry, xc - 1);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        s.q[v] = Form::get(p.qs, base + v * xc + c_ld);
        s.qn[v] = Form::needs_qn(kind) ? Form::get(p.qn, base + v * xc + c_ld) : s.q[v];
        if (ORD > 0 && halo_lane) s.qh[v] = Form::get(p.qs, base + v * xc + h_col);
    }
}

template <class Form, Reconstruction R, SlopeLimiter SL, RiemannSolver RS>
__device__ inline void seg_compute(const typename Form::Ptrs& p, int kind, int reduce_dt, int base, int xc, int ng,
                                   int nx, int i0, const typename Form::real dtdx, const typename Form::real dt,
                                   typename Form::real* lds_row /*[3][64+2R]*/, double& amax,
                                   unsigned int& floor_hits, const SegRaw<Form>& s) {
    using T = typename Form::real;
    using Raw = typename Form::raw;
    constexpr int NV = 3;  // rho, mx, E
    constexpr int ORD = stencil_order(R);
    constexpr int ROW = kWave + 2 * ORD;
    const int lane = threadIdx.x & 63;
    const int c = i0 - 1 + lane;
    const bool halo_lane = lane < 2 * ORD;
    const int h_entry = lane < ORD ? lane : kWave + lane;
    const EosT<T> eos = make_eos<T>((double)p.gamma);

    T w[NV];
    const bool out_lane = lane >= 1 && lane <= kStripOut && c < ng + nx;
    const bool in_hit = Form::template to_prim_nd<1>(p, s.q, w);
    wave_lds_sync();
#pragma unroll
    for (int v = 0; v < NV; ++v) lds_row[v * ROW + lane + ORD] = w[v];
    if (ORD > 0 && halo_lane) {
        T wh[NV];
        Form::template to_prim_nd<1>(p, s.qh, wh);
#pragma unroll
        for (int v = 0; v < NV; ++v) lds_row[v * ROW + h_entry] = wh[v];
    }
    wave_lds_sync();
    T xl[NV], xr[NV], xrm[NV], Fx[NV], Fxp[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        auto st = [&](int o) -> T { return lds_row[v * ROW + lane + ORD + o]; };
        reconstruct_stencil<R, SL, T>(st, xl[v], xr[v]);
    }
    Form::template floor_face<1>(p, xl);
    Form::template floor_face<1>(p, xr);
#pragma unroll
    for (int v = 0; v < NV; ++v) xrm[v] = lane_from_left(xr[v]);
    riemann_flux<RS, 0, 1>(eos, xrm, xl, Fx);
#pragma unroll
    for (int v = 0; v < NV; ++v) Fxp[v] = lane_from_right(Fx[v]);
    if (out_lane) {
        Raw qo[NV];
        T srcv[NV];
        if (p.gravity) {
            Form::template gravity_src<1>(p, s.q, dt, srcv);
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) srcv[v] = T(0) * dt;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            T q_update = Form::kExactZeros ? T(0) + (Fx[v] - Fxp[v]) : Fx[v] - Fxp[v];
            q_update *= dtdx;
            qo[v] = p.gravity ? Form::combine(kind, s.q[v], s.qn[v], q_update, srcv[v])
                              : Form::combine_nosrc(kind, s.q[v], s.qn[v], q_update);
        }
        const bool st_hit = Form::template floor_state<1>(p, qo);
        floor_hits += (in_hit || st_hit) ? 1u : 0u;
#pragma unroll
        for (int v = 0; v < NV; ++v) Form::put(p.out, base + v * xc + c, qo[v]);
        if (reduce_dt) {
            const double a = Form::template signal_speed_nd<1>((double)p.gamma, qo);
            amax = a > amax ? a : amax;
        }
    }
}

template <class Form, Reconstruction R, SlopeLimiter SL, RiemannSolver RS>
__device__ inline void stage1d_segment(const typename Form::Ptrs& p, int kind, int reduce_dt, int base, int xc,
                                       int ng, int nx, int i0, const typename Form::real dtdx,
                                       const typename Form::real dt, typename Form::real* lds_row /*[3][64+2R]*/,
                                       double& amax, unsigned int& floor_hits, int virt = 0, bool per = false) {
    SegRaw<Form> s;
    seg_load<Form, R>(p, kind, base, xc, ng, nx, i0, virt, per, s);
    seg_compute<Form, R, SL, RS>(p, kind, reduce_dt, base, xc, ng, nx, i0, dtdx, dt, lds_row, amax, floor_hits, s);
}

}  // namespace mh
