// Fused stage kernels of the FP64 drop-in path. Built with -ffp-contract=off:
// no multiply-add contraction, so results match the reference CPU build bitwise.
#ifndef MH_STAGE_WAVES_F64
#define MH_STAGE_WAVES_F64 1
#endif
#define MH_STAGE_WAVES_PER_EU MH_STAGE_WAVES_F64
// The raw Q^(s) of the stencil rows stays in registers for the combine (no
// second load of row j) with WENO5-Z; the other schemes keep 3 waves per
// SIMD without it (PPM + HLL: 154 VGPRs, 180 with it).
#ifndef MH_STAGE_QS_RING_F64
#define MH_STAGE_QS_RING_F64 2
#endif
#define MH_STAGE_QS_RING MH_STAGE_QS_RING_F64
#include "stage_inst.hpp"

namespace mh {

hipError_t launch_stage2d_f64(const Schemes& s, const StageGrid& g, const FormF64::Ptrs& p, hipStream_t st) {
    return launch_stage2d_impl<FormF64>(s, g, p, st);
}

hipError_t launch_stage1d_f64(const Schemes& s, int kind, int reduce_dt, int xc, int ng, const FormF64::Ptrs& p,
                              hipStream_t st) {
    return launch_stage1d_impl<FormF64>(s, kind, reduce_dt, xc, ng, p, st);
}

hipError_t launch_tracer_f64(const Schemes& s, const TracerGrid& g, const FormF64::Ptrs& p,
                              FormF64::real* wtr, hipStream_t st) {
    return launch_tracer_impl<FormF64>(s, g, p, wtr, st);
}

}  // namespace mh
