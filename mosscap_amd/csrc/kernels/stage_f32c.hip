// Fused stage kernels of the compensated FP32 precision form. Built with
// -ffp-contract=fast: multiply-adds may fuse (the error-free transforms of
// the combine contain no multiply-add that contraction could change).
#include "stage_inst.hpp"

namespace mh {

hipError_t launch_stage2d_f32c(const Schemes& s, const StageGrid& g, const FormF32c::Ptrs& p, hipStream_t st) {
    return launch_stage2d_impl<FormF32c>(s, g, p, st);
}

hipError_t launch_stage1d_f32c(const Schemes& s, int kind, int reduce_dt, int xc, int ng, const FormF32c::Ptrs& p,
                               hipStream_t st) {
    return launch_stage1d_impl<FormF32c>(s, kind, reduce_dt, xc, ng, p, st);
}

}  // namespace mh
