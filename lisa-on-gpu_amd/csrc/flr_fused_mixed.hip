// flr_fused_mixed.hip -- fused projection + TDI kernels with FP32 storage (complex64 h / scale
// in, float out, FP64 arithmetic inside), fast stencils, for 256-, 512- and 1024-sample tiles
// (one, two or four output samples per lane), the choice the autotuner makes per workload.
#include "flr_fused.hpp"

namespace flr {

hipError_t launch_fused_mixed(const ProjConfig &pc, const TdiConfig &tc,
                              const Batch<FusedSource> &b, const OrbitView &orb,
                              const LagrangeView &lg, int tile_passes, int grid_per_cu,
                              const StockTdi &st, hipStream_t stream)
{
    switch (tile_passes) {
    case 1:
        return launch_fused_variant<false, float2, float, 1>(pc, tc, b, orb, lg, st, grid_per_cu,
                                                             stream);
    case 2:
        return launch_fused_variant<false, float2, float, 2>(pc, tc, b, orb, lg, st, grid_per_cu,
                                                             stream);
    case 4:
        return launch_fused_variant<false, float2, float, 4>(pc, tc, b, orb, lg, st, grid_per_cu,
                                                             stream);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace flr
