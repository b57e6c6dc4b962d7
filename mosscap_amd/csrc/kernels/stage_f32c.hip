// Fused stage kernels of the compensated FP32 precision form. Built with
// -ffp-contract=fast-honor-pragmas: multiply-adds may fuse, except where
// form_f32c.hpp pins contraction off (the stage combine's error-free
// transforms, the CFL signal speed). The plain instance fits 128 VGPRs (4
// waves per SIMD); the instance with the gravity source and tracer stores
// needs more (3 waves).
#ifndef MH_STAGE_WAVES_F32C
#define MH_STAGE_WAVES_F32C 4
#endif
#define MH_STAGE_WAVES_PER_EU MH_STAGE_WAVES_F32C
#define MH_STAGE_WAVES_EXTRAS 3
#define MH_STAGE_WAVES_FLOORS 4
// The raw Q^(s) of the stencil rows stays in registers for the combine
// (no second load of row j).
#ifndef MH_STAGE_QS_RING_F32C
#define MH_STAGE_QS_RING_F32C 1
#endif
#define MH_STAGE_QS_RING MH_STAGE_QS_RING_F32C
// The plain instance marches its rows in groups of three, so the ring rotates
// by renaming (code 20 -> 47 KB, 7 spilled dwords; stage -1.3 %).
#ifndef MH_STAGE_RING_UNROLL
#define MH_STAGE_RING_UNROLL 1
#endif
#include "stage_inst.hpp"

namespace mh {

hipError_t launch_stage2d_f32c(const Schemes& s, const StageGrid& g, const FormF32c::Ptrs& p, hipStream_t st) {
    return launch_stage2d_impl<FormF32c>(s, g, p, st);
}

hipError_t launch_stage1d_f32c(const Schemes& s, int kind, int reduce_dt, int xc, int ng, const FormF32c::Ptrs& p,
                               hipStream_t st) {
    return launch_stage1d_impl<FormF32c>(s, kind, reduce_dt, xc, ng, p, st);
}

hipError_t launch_tracer_f32c(const Schemes& s, const TracerGrid& g, const FormF32c::Ptrs& p,
                              FormF32c::real* wtr, hipStream_t st) {
    return launch_tracer_impl<FormF32c>(s, g, p, wtr, st);
}

}  // namespace mh
