// flr_fused_fp64.hip -- fused projection + TDI kernels with FP64 storage (complex128 in,
// double out): exact stencils (bitwise the unfused path) or the fast form, 512-sample tiles.
#include "flr_fused.hpp"

namespace flr {

hipError_t launch_fused_fp64(const ProjConfig &pc, const TdiConfig &tc,
                             const Batch<FusedSource> &b, const OrbitView &orb,
                             const LagrangeView &lg, bool exact, int grid_per_cu,
                             const StockTdi &st, hipStream_t stream)
{
    if (exact)
        return launch_fused_variant<true, double2, double, 2>(pc, tc, b, orb, lg, st, grid_per_cu,
                                                              stream);
    return launch_fused_variant<false, double2, double, 2>(pc, tc, b, orb, lg, st, grid_per_cu,
                                                           stream);
}

}  // namespace flr
