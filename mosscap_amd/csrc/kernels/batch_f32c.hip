// Batched 1-D ensemble, compensated FP32 form (-ffp-contract=fast-honor-pragmas; the
// error-free transforms pin contraction off themselves).
#define MH_STAGE_WAVES_PER_EU 1
// 5 waves per SIMD fit without spills with the next item's loads in flight
// (96 VGPRs); the 4096-problem ensemble is fastest at 20 waves per CU.
#define MH_BATCH_WAVES_PER_EU 5
#include "batch_inst.hpp"

namespace mh {

hipError_t launch_batch_stage_f32c(const Schemes& s, const BatchGrid& g, int grid, const FormF32c::Ptrs& p,
                                   const BatchArrays& b, hipStream_t st) {
    return launch_batch_stage_impl<FormF32c>(s, g, grid, p, b, st);
}
hipError_t launch_batch_ghost_f32c(const GhostGrid& g, int nprob, FormF32c::Planes in, FormF32c::Planes out,
                                   const BatchArrays& b, hipStream_t st) {
    return launch_batch_ghost_impl<FormF32c>(g, nprob, in, out, b, st);
}
hipError_t launch_batch_cfl_f32c(int nprob, FormF32c::Planes q, double gamma, const BatchArrays& b, hipStream_t st) {
    return launch_batch_cfl_impl<FormF32c>(nprob, q, gamma, b, st);
}

}  // namespace mh
