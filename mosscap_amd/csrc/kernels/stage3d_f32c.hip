// 3-D stage kernels (z sweep + per-plane x-y march) of the compensated FP32
// form, built with -ffp-contract=fast-honor-pragmas like the 2-D unit. Five variables per
// cell need more registers than the 2-D march: the plain instance is held to
// at least three waves per SIMD (it takes 112 VGPRs, four waves), the one with
// the gravity source to two.
#define MH_STAGE_WAVES_PER_EU 3
#define MH_STAGE_WAVES_EXTRAS 2
#ifndef MH_STAGE_QS_RING_F32C3
#define MH_STAGE_QS_RING_F32C3 0
#endif
#define MH_STAGE_QS_RING MH_STAGE_QS_RING_F32C3
#define MH_STAGE3D_INSTANCES 1
#include "stage_inst.hpp"

namespace mh {

hipError_t launch_stage3d_f32c(const Schemes& s, const StageGrid& g, const FormF32c::Ptrs& p, float* dz,
                               hipStream_t st) {
    return launch_stage3d_impl<FormF32c>(s, g, p, dz, st);
}

}  // namespace mh
