// flr_fused_c128.hip -- the mixed fused kernels with a complex128 input (h / scale in FP64,
// 16 bytes per sample) and float32 outputs. The fused launch reads a few percent of HBM
// bandwidth, so the wider input costs little time, while it removes the storage error of the
// 8-byte formats -- white noise that TDI amplifies most in the T channel, where the signal
// cancels. The arithmetic is the other mixed forms' (FP64, fast stencils, composite filters
// for equal arms); the outputs are rounded to float32 once. Same tile variants.
#include "flr_fused.hpp"

namespace flr {

hipError_t launch_fused_c128(const ProjConfig &pc, const TdiConfig &tc,
                             const Batch<FusedSource> &b, const OrbitView &orb,
                             const LagrangeView &lg, int tile_passes, int grid_per_cu,
                             const StockTdi &st, hipStream_t stream)
{
    switch (tile_passes) {
    case 1:
        return launch_fused_variant<false, double2, float, 1>(pc, tc, b, orb, lg, st, grid_per_cu,
                                                              stream);
    case 2:
        return launch_fused_variant<false, double2, float, 2>(pc, tc, b, orb, lg, st, grid_per_cu,
                                                              stream);
    case 4:
        return launch_fused_variant<false, double2, float, 4>(pc, tc, b, orb, lg, st, grid_per_cu,
                                                              stream);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace flr
