// 3-D stage kernels (z sweep + per-plane x-y march) of the FP64 drop-in path,
// built with -ffp-contract=off like the 2-D units.
// The 3-D march holds five variables per window row: it needs about 280
// VGPRs unconstrained (one wavefront per SIMD); held to 256 (two), it spills
// a few dwords. With two stencil rows in flight (loaded straight into LDS two
// rows ahead) it spills 4 dwords instead of 25 and its stage runs 21 %
// faster on the 256^3 blast than with one.
#ifndef MH_STAGE3D_WAVES_F64
#define MH_STAGE3D_WAVES_F64 2
#endif
#define MH_STAGE_LEAD_ROWS 2
// the z flux differences of a row reach LDS with the combine operands
#ifndef MH_STAGE_DZ_LDS
#define MH_STAGE_DZ_LDS 1
#endif
#define MH_STAGE_WAVES_PER_EU MH_STAGE3D_WAVES_F64
#define MH_STAGE3D_INSTANCES 1
#include "stage_inst.hpp"

namespace mh {

hipError_t launch_stage3d_f64(const Schemes& s, const StageGrid& g, const FormF64::Ptrs& p, double* dz,
                              hipStream_t st) {
    return launch_stage3d_impl<FormF64>(s, g, p, dz, st);
}

}  // namespace mh
