// 3-D stage kernels (z sweep + per-plane x-y march) of the FP64 drop-in path,
// built with -ffp-contract=off like the 2-D units.
#ifndef MH_STAGE_WAVES_F64
#define MH_STAGE_WAVES_F64 1
#endif
#define MH_STAGE_WAVES_PER_EU MH_STAGE_WAVES_F64
#define MH_STAGE3D_INSTANCES 1
#include "stage_inst.hpp"

namespace mh {

hipError_t launch_stage3d_f64(const Schemes& s, const StageGrid& g, const FormF64::Ptrs& p, double* dz,
                              hipStream_t st) {
    return launch_stage3d_impl<FormF64>(s, g, p, dz, st);
}

}  // namespace mh
