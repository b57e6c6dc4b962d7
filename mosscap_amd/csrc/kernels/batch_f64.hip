// Batched 1-D ensemble, FP64 form (-ffp-contract=off: bitwise the single-problem path).
#define MH_STAGE_WAVES_PER_EU 1
#include "batch_inst.hpp"

namespace mh {

namespace {

__global__ void batch_finalize_kernel(int nprob, double max_cfl, double max_time, int advance, BatchArrays b,
                                      double* dt_out) {
    const int pr = blockIdx.x * blockDim.x + threadIdx.x;
    if (pr >= nprob) return;
    const double dt = batch_next_dt(b, pr, max_cfl, max_time);
    const double t = b.time[pr];
    dt_out[pr] = dt;
    b.amax[pr] = 0ull;
    if (advance) b.time[pr] = t + dt;
}

}  // namespace

hipError_t launch_batch_stage_f64(const Schemes& s, const BatchGrid& g, int grid, const FormF64::Ptrs& p,
                                  const BatchArrays& b, hipStream_t st) {
    return launch_batch_stage_impl<FormF64>(s, g, grid, p, b, st);
}
hipError_t launch_batch_ghost_f64(const GhostGrid& g, int nprob, FormF64::Planes in, FormF64::Planes out,
                                  const BatchArrays& b, hipStream_t st) {
    return launch_batch_ghost_impl<FormF64>(g, nprob, in, out, b, st);
}
hipError_t launch_batch_cfl_f64(int nprob, FormF64::Planes q, double gamma, const BatchArrays& b, hipStream_t st) {
    return launch_batch_cfl_impl<FormF64>(nprob, q, gamma, b, st);
}
hipError_t launch_batch_finalize(int nprob, double max_cfl, double max_time, int advance, const BatchArrays& b,
                                 double* dt_out, hipStream_t st) {
    if (nprob < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_finalize_kernel, dim3((nprob + 255) / 256), dim3(256), 0, st, nprob, max_cfl, max_time,
                       advance, b, dt_out);
    return hipGetLastError();
}

}  // namespace mh
