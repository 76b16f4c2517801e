// FP64 storage, FP64 collision: the drop-in path.  Built with
// -ffp-contract=off so every operation rounds exactly as the oracle's does.
#include "fused.h"

namespace imx {
hipError_t launch_fused_f64(int bc, int pattern, const FusedParams &P, int block, hipStream_t st)
{
    return launch_fused_t<double>(bc, pattern, P, block, st);
}
}  // namespace imx
