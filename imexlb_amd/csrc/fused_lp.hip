// FP32 and FP16 storage of f - w_i S, FP32 collision with an FP64 density sum.
#include "fused.h"

namespace imx {
hipError_t launch_fused_f32(int bc, int pattern, const FusedParams &P, int block, hipStream_t st)
{
    return launch_fused_t<float>(bc, pattern, P, block, st);
}
hipError_t launch_fused_f16(int bc, int pattern, const FusedParams &P, int block, hipStream_t st)
{
    return launch_fused_t<_Float16>(bc, pattern, P, block, st);
}
}  // namespace imx
