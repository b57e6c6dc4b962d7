// Launch interface of the batched 1-D ensemble (capi_batch.cpp <-> kernels).
#pragma once

#include "device/batch.hpp"
#include "device/form_f32c.hpp"
#include "device/form_f64.hpp"

namespace mh {

struct Schemes;

hipError_t launch_batch_stage_f64(const Schemes& s, const BatchGrid& g, int grid, const FormF64::Ptrs& p,
                                  const BatchArrays& b, hipStream_t st);
hipError_t launch_batch_stage_f32c(const Schemes& s, const BatchGrid& g, int grid, const FormF32c::Ptrs& p,
                                   const BatchArrays& b, hipStream_t st);
hipError_t launch_batch_ghost_f64(const GhostGrid& g, int nprob, FormF64::Planes in, FormF64::Planes out,
                                  const BatchArrays& b, hipStream_t st);
hipError_t launch_batch_ghost_f32c(const GhostGrid& g, int nprob, FormF32c::Planes in, FormF32c::Planes out,
                                   const BatchArrays& b, hipStream_t st);
hipError_t launch_batch_cfl_f64(int nprob, FormF64::Planes q, double gamma, const BatchArrays& b, hipStream_t st);
hipError_t launch_batch_cfl_f32c(int nprob, FormF32c::Planes q, double gamma, const BatchArrays& b, hipStream_t st);
// Per-problem dt = max_cfl * dx / max(c_s + |v|), clamped to max_time
// (Hydro.cpp:230-241 with no output times inside the run); resets the max;
// with `advance`, time += dt.
hipError_t launch_batch_finalize(int nprob, double max_cfl, double max_time, int advance, const BatchArrays& b,
                                 double* dt_out, hipStream_t st);

}  // namespace mh
