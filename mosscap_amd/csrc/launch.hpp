// Internal launch interface between the C ABI (capi.cpp) and the kernel
// translation units. Every launcher checks its shapes against what the
// kernel's grid assumes before launching and returns the HIP status.
#pragma once

#include <hip/hip_runtime.h>

#include "device/layout.hpp"
#include "device/form_f64.hpp"
#include "device/form_f32c.hpp"
#include "device/ghost_dt.hpp"
#include "device/stage_core.hpp"
#include "device/stage3d.hpp"
#include "device/tracer.hpp"
#include "device/lte_eos.hpp"
#include "device/lte_table.hpp"

namespace mh {

struct Schemes {
    Reconstruction recon;
    SlopeLimiter limiter;
    RiemannSolver solver;
    TimeStepScheme stepper;
};

// 2-D fused stage over the interior, one wavefront per (strip, span).
hipError_t launch_stage2d_f64(const Schemes& s, const StageGrid& g, const FormF64::Ptrs& p, hipStream_t st);
hipError_t launch_stage2d_f32c(const Schemes& s, const StageGrid& g, const FormF32c::Ptrs& p, hipStream_t st);
// 3-D stage: the z sweep into `dz` (NV values per cell, [var][k][j][i]),
// then the x-y march of every interior plane (g.zc, g.nz set).
hipError_t launch_stage3d_f64(const Schemes& s, const StageGrid& g, const FormF64::Ptrs& p, double* dz,
                              hipStream_t st);
hipError_t launch_stage3d_f32c(const Schemes& s, const StageGrid& g, const FormF32c::Ptrs& p, float* dz,
                               hipStream_t st);
// 1-D stage over one row of xc cells (ng ghosts each side).
hipError_t launch_stage1d_f64(const Schemes& s, int kind, int reduce_dt, int xc, int ng, const FormF64::Ptrs& p,
                              hipStream_t st);
hipError_t launch_stage1d_f32c(const Schemes& s, int kind, int reduce_dt, int xc, int ng, const FormF32c::Ptrs& p,
                               hipStream_t st);

// Tracer planes of a 2-D stage (p.frx / p.fry hold the stage's density fluxes).
hipError_t launch_tracer_f64(const Schemes& s, const TracerGrid& g, const FormF64::Ptrs& p, double* wtr,
                             hipStream_t st);
hipError_t launch_tracer_f32c(const Schemes& s, const TracerGrid& g, const FormF32c::Ptrs& p, float* wtr,
                              hipStream_t st);
hipError_t launch_tracer_ghost_f64(const GhostGrid& g, int ntr, FormF64::Planes in, FormF64::Planes out,
                                   hipStream_t st);
hipError_t launch_tracer_ghost_f32c(const GhostGrid& g, int ntr, FormF32c::Planes in, FormF32c::Planes out,
                                    hipStream_t st);

hipError_t launch_ghost_f64(int nd, const GhostGrid& g, FormF64::Planes in, FormF64::Planes out,
                            unsigned long long* amax, hipStream_t st);
hipError_t launch_ghost_f32c(int nd, const GhostGrid& g, FormF32c::Planes in, FormF32c::Planes out,
                             unsigned long long* amax, hipStream_t st);
// 3-D ghost fill of one buffer (x, y, z passes), folding the ghost cells
// into the CFL max when g.reduce_dt is set.
hipError_t launch_ghost3d_f64(const Ghost3Grid& g, FormF64::Planes q, unsigned long long* amax, hipStream_t st);
hipError_t launch_ghost3d_f32c(const Ghost3Grid& g, FormF32c::Planes q, unsigned long long* amax, hipStream_t st);
hipError_t launch_cfl_f64(int nd, FormF64::Planes q, int ncells, double gamma, unsigned long long* amax,
                          hipStream_t st);
hipError_t launch_cfl_f32c(int nd, FormF32c::Planes q, int ncells, double gamma, unsigned long long* amax,
                           hipStream_t st);

// Analytic LTE EOS fields (ionisation fraction, temperature) of every cell.
hipError_t launch_lte_eos_f64(int nd, FormF64::Planes q, int ncells, const LteParams& p, double* y, double* T,
                              int* flipped, hipStream_t st);
hipError_t launch_lte_eos_f32c(int nd, FormF32c::Planes q, int ncells, const LteParams& p, double* y, double* T,
                               int* flipped, hipStream_t st);

// Tabulated LTE EOS update_eos over every cell of `in` into `out` (same
// buffer, or a copy of all nvar planes with E updated).
hipError_t launch_tab_eos_f64(int nd, FormF64::Planes in, FormF64::Planes out, int ncells, int nvar,
                              const TabParams& t, const double* ytab, const double* ltab, double* y_space,
                              double* T_space, int* neg, hipStream_t st);
hipError_t launch_tab_eos_f32c(int nd, FormF32c::Planes in, FormF32c::Planes out, int ncells, int nvar,
                               const TabParams& t, const double* ytab, const double* ltab, double* y_space,
                               double* T_space, int* neg, hipStream_t st);

hipError_t launch_clock_finalize(Clock* c, int advance, hipStream_t st);
hipError_t launch_clock_set_dt(Clock* c, double dt, hipStream_t st);
hipError_t launch_clock_reset_max(Clock* c, hipStream_t st);

// Conversions between the FP64 layout and the compensated hi/lo planes.
hipError_t launch_split_f64_to_f32c(const double* src, FormF32c::Planes dst, int n, hipStream_t st);
hipError_t launch_join_f32c_to_f64(FormF32c::Planes src, double* dst, int n, hipStream_t st);

// Functor probes for the unit tests: run the device functors on n inputs.
hipError_t launch_probe_reconstruct(int recon, int limiter, int prec, int n, const double* stencils, double* wl,
                                    double* wr, hipStream_t st);
hipError_t launch_probe_riemann(int solver, int axis, int nd, int prec, double gamma, int n, const double* wl,
                                const double* wr, double* flux, hipStream_t st);
hipError_t launch_probe_math(int op, int n, const double* a, const double* b, double* out, hipStream_t st);

}  // namespace mh
