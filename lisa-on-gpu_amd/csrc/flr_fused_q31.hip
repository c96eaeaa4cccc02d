// flr_fused_q31.hip -- fused projection + TDI kernels whose input is 32-bit fixed point: each
// component of h stored as round(h / s * (2^31 - 1)) in an int32 (8 bytes per complex sample,
// like complex64) with one scale per source, float out, FP64 arithmetic inside. For signals of
// bounded amplitude the quantisation step is 2^-31 of full scale instead of complex64's 2^-24
// of the value, so the storage error that TDI amplifies is ~2^7 times smaller at the same
// bandwidth. Same tile variants as the complex64 form.
#include "flr_fused.hpp"

namespace flr {

hipError_t launch_fused_q31(const ProjConfig &pc, const TdiConfig &tc,
                            const Batch<FusedSource> &b, const OrbitView &orb,
                            const LagrangeView &lg, int tile_passes, int grid_per_cu,
                            const StockTdi &st, hipStream_t stream)
{
    switch (tile_passes) {
    case 1:
        return launch_fused_variant<false, int2, float, 1>(pc, tc, b, orb, lg, st, grid_per_cu,
                                                           stream);
    case 2:
        return launch_fused_variant<false, int2, float, 2>(pc, tc, b, orb, lg, st, grid_per_cu,
                                                           stream);
    case 4:
        return launch_fused_variant<false, int2, float, 4>(pc, tc, b, orb, lg, st, grid_per_cu,
                                                           stream);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace flr
