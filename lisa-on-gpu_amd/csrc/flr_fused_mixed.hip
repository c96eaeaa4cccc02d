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

bool lti_filters_needed(const StockTdi &st, const LagrangeView &lg)
{
    return use_equal_arm(st, lg);
}

hipError_t launch_lti_filters(const TdiConfig &tc, const Batch<FusedSource> &b,
                              const OrbitView &orb, const LagrangeView &lg, const StockTdi &st,
                              hipStream_t stream)
{
    if (!lti_filters_needed(st, lg)) return hipSuccess;
    if (!tc.lti_fir) return hipErrorInvalidValue;
    return st.generation == 1
               ? launch(lti_filter_kernel<8>, b.num_sources, 0, stream, tc, b, orb, lg, st.rows)
               : launch(lti_filter_kernel<16>, b.num_sources, 0, stream, tc, b, orb, lg, st.rows);
}

}  // namespace flr
