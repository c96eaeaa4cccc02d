// flr_kernels.hip -- gfx950 kernels of the LISA time-domain response.
//
// Work unit: a tile of kBlock consecutive output samples of one source, one sample per lane.
// Each tile stages its sample window (tile + delay halo) in LDS once and every lane reads all
// its stencil taps from there.
//
// projection_tile: all six links per lane. The complex h window is shared by the 3 receiver
//   and 6 emitter stencils of every lane (the reference restages it per link,
//   LISAResponse.cu:628-655). A receiver stencil depends on the receiving spacecraft only, so
//   each is computed once for its two links: 9 complex stencils per sample instead of 12,
//   with bit-identical results (6 with equal arms, where the emitter stencils are shared too).
// tdi_tile: any TDI table. The 6 y rows of the window are staged once; each lane walks every
//   unit of every channel for its sample, summing units of a channel in index order in
//   registers (no atomics), then stores X, Y, Z or A, E, T.
// tdi_stock_tile: the stock 1st/2nd-generation tables with compile-time chains
//   (flr_tdi_stock.hpp): shared delays per chain length when all light travel times are
//   equal, per-unit delays from registers otherwise.
//
// Each tile function is launched two ways: one workgroup per tile of a single source, or a
// persistent batch kernel whose workgroups walk contiguous runs of a flattened (source, tile)
// list -- many sources of uneven length in one launch.
#include "flr_paths.hpp"

namespace flr {

// ------------------------------------------------------------------ projection

template <bool EXACT, int H, bool EQ>
__device__ __forceinline__ void projection_tile(const ProjConfig &cfg, const ProjSource &src,
                                                int tile, const OrbitView &orb,
                                                const LagrangeView &lg, double2 *hwin)
{
    const int tid = threadIdx.x;
    const TileRange tr = tile_range(tile, src.num_delays, src.start, cfg.zero_edges);
    const int i = tr.i0 + tid;
    const bool valid = (i >= tr.lo_valid) && (i < tr.hi_valid);
    const bool in_range = (i >= 0) && (i < src.num_delays);

    if (!valid && in_range && cfg.zero_edges) {
#pragma unroll
        for (int l = 0; l < kNLinks; ++l) src.y[(size_t)l * src.num_delays + i] = 0.0;
    }
    if (tr.first > tr.last) return;  // whole tile outside the computed range

    // The window is sized from the orbit extent, so nothing has to be known about the delays
    // before it is staged.
    const int win_lo = tr.first - cfg.halo_lo;
    const int win_len = tr.last + cfg.halo_hi - win_lo + 1;
    if (win_len > cfg.win_cap || win_lo < 0 || win_lo + win_len > src.num_inputs) {
        if (valid) {
            atomicOr(cfg.err, kErrWindow);
            for (int l = 0; l < kNLinks; ++l) src.y[(size_t)l * src.num_delays + i] = __builtin_nan("");
        }
        return;
    }
    for (int jj = tid; jj < win_len; jj += kBlock) hwin[jj] = src.h[win_lo + jj];

    const int ic = valid ? i : tr.first;
    const double t = src.t_data ? src.t_data[ic] : src.t0 + (double)ic * cfg.dt;
    const double t0_offset = src.t_data ? src.t_data[0] : src.t0 + (double)0 * cfg.dt;
    __syncthreads();
    if (!valid) return;
    const double y_deps = EXACT ? rcp_nr2(lg.deps) : 0.0;
    double *y = src.y + i;
    const size_t stride = (size_t)src.num_delays;
    if (project_links<EXACT, H, EQ>(cfg, src.k, src.u, src.v, t, t0_offset, orb, lg, hwin, win_lo,
                                win_len, y_deps, nullptr, OrbitStage{},
                                [&](int row, double val) { y[row * stride] = val; }))
        atomicOr(cfg.err, kErrWindow);
}

// ------------------------------------------------------------------ TDI (any table)

__device__ __forceinline__ void tdi_store(const TdiConfig &cfg, const TdiSource &src, int i,
                                          const double (&acc)[3])
{
    if (cfg.aet) {
        const double X = acc[0], Y = acc[1], Z = acc[2];
        src.out[i] = (Z - X) / cfg.s2;
        src.out[(size_t)src.num_delays + i] = (X - 2.0 * Y + Z) / cfg.s6;
        src.out[(size_t)2 * src.num_delays + i] = (X + Y + Z) / cfg.s3;
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) src.out[(size_t)c * src.num_delays + i] = acc[c];
    }
}

// Zero the tile's edge samples if asked, check and stage the y window. Returns false when the
// tile has nothing (more) to do; every lane returns the same value.
__device__ __forceinline__ bool tdi_stage(const TdiConfig &cfg, const TdiSource &src,
                                          const TileRange &tr, int i, bool valid, double *ywin,
                                          int &win_lo, int &win_len)
{
    const bool in_range = (i >= 0) && (i < src.num_delays);
    if (!valid && in_range && cfg.zero_edges) {
#pragma unroll
        for (int c = 0; c < 3; ++c) src.out[(size_t)c * src.num_delays + i] = 0.0;
    }
    if (tr.first > tr.last) return false;
    win_lo = tr.first - cfg.halo_lo;
    win_len = tr.last + cfg.halo_hi - win_lo + 1;
    if (win_len > cfg.win_cap || win_lo < 0 || win_lo + win_len > src.num_inputs) {
        if (valid) {
            atomicOr(cfg.err, kErrWindow);
            for (int c = 0; c < 3; ++c) src.out[(size_t)c * src.num_delays + i] = __builtin_nan("");
        }
        return false;
    }
    for (int r = 0; r < kNLinks; ++r) {
        const double *from = src.y + (size_t)r * src.num_inputs + win_lo;
        double *to = ywin + r * cfg.win_cap;
        for (int jj = threadIdx.x; jj < win_len; jj += kBlock) to[jj] = from[jj];
    }
    return true;
}

template <bool EXACT, int H>
__device__ __forceinline__ void tdi_tile(const TdiConfig &cfg, const TdiSource &src, int tile,
                                         const OrbitView &orb, const LagrangeView &lg,
                                         double *smem)
{
    double *ywin = smem;                        // [6][win_cap]
    double *Ls = smem + kNLinks * cfg.win_cap;  // [6][kBlock]
    const int tid = threadIdx.x;
    const TileRange tr = tile_range(tile, src.num_delays, src.start, cfg.zero_edges);
    const int i = tr.i0 + tid;
    const bool valid = (i >= tr.lo_valid) && (i < tr.hi_valid);
    int win_lo = 0, win_len = 0;
    if (!tdi_stage(cfg, src, tr, i, valid, ywin, win_lo, win_len)) return;

    const int ic = valid ? i : tr.first;
    const double t = src.t_arr ? src.t_arr[ic] : src.t0 + (double)ic * cfg.dt;
    const double t0_offset = src.t_arr ? src.t_arr[0] : src.t0 + (double)0 * cfg.dt;
    {
        const OrbitNode nd = orbit_node<EXACT>(orb, t);
#pragma unroll
        for (int r = 0; r < kNLinks; ++r) Ls[r * kBlock + tid] = orbit_val(nd, r);
    }
    __syncthreads();
    if (!valid) return;
    const double y_deps = EXACT ? rcp_nr2(lg.deps) : 0.0;

    double acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        acc[c] = cfg.overwrite ? 0.0 : src.out[(size_t)c * src.num_delays + i];

    if (tdi_units<EXACT, H>(cfg, t, t0_offset, Ls + tid, kBlock, ywin, cfg.win_cap, win_lo, win_len,
                            lg, y_deps, acc))
        atomicOr(cfg.err, kErrWindow);
    tdi_store(cfg, src, i, acc);
}

// ------------------------------------------------------------------ equal-arm stock TDI
// Stock 1st/2nd-generation tables at order 25 when all six light travel times of the orbit
// table are bitwise equal (equal-arm orbits): the distinct delays t - m*L (m = chain length
// 0..3 or 0..7) are split once each and their stencil reciprocals (exact form) or weights
// (fast form) are shared by the units of that length -- 8 groups instead of 48 units for 2nd
// generation, with bit-identical results.

template <int NX, bool EXACT, bool EQ>
__device__ __forceinline__ void tdi_stock_tile(const TdiConfig &cfg, const TdiSource &src,
                                               int tile, const OrbitView &orb,
                                               const LagrangeView &lg, const RowMap &rm,
                                               double *ywin)
{
    const TileRange tr = tile_range(tile, src.num_delays, src.start, cfg.zero_edges);
    const int i = tr.i0 + (int)threadIdx.x;
    const bool valid = (i >= tr.lo_valid) && (i < tr.hi_valid);
    int win_lo = 0, win_len = 0;
    if (!tdi_stage(cfg, src, tr, i, valid, ywin, win_lo, win_len)) return;

    const int ic = valid ? i : tr.first;
    const double t = src.t_arr ? src.t_arr[ic] : src.t0 + (double)ic * cfg.dt;
    const double t0_offset = src.t_arr ? src.t_arr[0] : src.t0 + (double)0 * cfg.dt;
    const OrbitNode nd = orbit_node<EXACT>(orb, t);
    double L[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) L[k] = orbit_val(nd, rm.row[k]);
    int rowoff[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) rowoff[k] = rm.row[k] * cfg.win_cap - win_lo;
    __syncthreads();
    if (!valid) return;
    const double y_deps = EXACT ? rcp_nr2(lg.deps) : 0.0;

    double acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        acc[c] = cfg.overwrite ? 0.0 : src.out[(size_t)c * src.num_delays + i];
    bool ok;
    if constexpr (EQ)
        ok = tdi_equal_arm<NX, EXACT>(t, L[0], t0_offset, cfg.fs, ywin, rowoff, win_lo, win_len, lg,
                                      y_deps, acc);
    else
        ok = tdi_stock_unequal<NX, EXACT>(t, L, t0_offset, cfg.fs, ywin, rowoff, win_lo, win_len,
                                          lg, y_deps, acc);
    if (!ok) atomicOr(cfg.err, kErrWindow);
    tdi_store(cfg, src, i, acc);
}

// ------------------------------------------------------------------ kernels

template <bool EXACT, int H, bool EQ>
__global__ __launch_bounds__(kBlock) void projection_kernel(ProjConfig cfg, ProjSource src,
                                                            OrbitView orb, LagrangeView lg)
{
    extern __shared__ double2 hwin[];
    projection_tile<EXACT, H, EQ>(cfg, src, blockIdx.x, orb, lg, hwin);
}

template <bool EXACT, int H, bool EQ>
__global__ __launch_bounds__(kBlock) void projection_batch_kernel(ProjConfig cfg,
                                                                  Batch<ProjSource> batch,
                                                                  OrbitView orb, LagrangeView lg)
{
    extern __shared__ double2 hwin[];
    FLR_BATCH_LOOP(batch, (projection_tile<EXACT, H, EQ>(cfg, src, tile, orb, lg, hwin)));
}

template <bool EXACT, int H>
__global__ __launch_bounds__(kBlock) void tdi_kernel(TdiConfig cfg, TdiSource src, OrbitView orb,
                                                     LagrangeView lg)
{
    extern __shared__ double smem[];
    tdi_tile<EXACT, H>(cfg, src, blockIdx.x, orb, lg, smem);
}

template <bool EXACT, int H>
__global__ __launch_bounds__(kBlock) void tdi_batch_kernel(TdiConfig cfg, Batch<TdiSource> batch,
                                                           OrbitView orb, LagrangeView lg)
{
    extern __shared__ double smem[];
    FLR_BATCH_LOOP(batch, (tdi_tile<EXACT, H>(cfg, src, tile, orb, lg, smem)));
}

template <int NX, bool EXACT, bool EQ>
__global__ __launch_bounds__(kBlock) void tdi_stock_kernel(TdiConfig cfg, TdiSource src,
                                                               OrbitView orb, LagrangeView lg,
                                                               RowMap rm)
{
    extern __shared__ double smem[];
    tdi_stock_tile<NX, EXACT, EQ>(cfg, src, blockIdx.x, orb, lg, rm, smem);
}

template <int NX, bool EXACT, bool EQ>
__global__ __launch_bounds__(kBlock) void tdi_stock_batch_kernel(TdiConfig cfg,
                                                                     Batch<TdiSource> batch,
                                                                     OrbitView orb,
                                                                     LagrangeView lg, RowMap rm)
{
    extern __shared__ double smem[];
    FLR_BATCH_LOOP(batch, (tdi_stock_tile<NX, EXACT, EQ>(cfg, src, tile, orb, lg, rm, smem)));
}

// ------------------------------------------------------------------ small kernels

__global__ void orbit_eval_kernel(OrbitView orb, const double *t, int n, int row, int sc,
                                  double *ltt_out, double *pos_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const OrbitNode nd = orbit_node(orb, t[i]);
    if (ltt_out) ltt_out[i] = orbit_val(nd, row);
    if (pos_out) {
        for (int c = 0; c < 3; ++c) pos_out[3 * i + c] = orbit_val(nd, 6 + 3 * sc + c);
    }
}

__global__ void clear_kernel(int *word)
{
    if (threadIdx.x == 0) word[0] = 0;
}

// ------------------------------------------------------------------ launchers

int batch_grid(int total_tiles, int grid_per_cu)
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
    }
    const int grid = cus * (grid_per_cu > 0 ? grid_per_cu : 16);
    return total_tiles < grid ? total_tiles : grid;
}

static size_t proj_lds(const ProjConfig &cfg) { return (size_t)cfg.win_cap * sizeof(double2); }

// equal_arm: all six light travel time columns of the orbit table are bitwise equal (the
// emitter stencils are then shared per emitting spacecraft; order 25 only).
hipError_t launch_projection(const ProjConfig &cfg, const ProjSource &src, const OrbitView &orb,
                             const LagrangeView &lg, bool exact, bool equal_arm,
                             hipStream_t stream)
{
    const int nb = tiles_for(src.num_delays, src.start, cfg.zero_edges);
    const size_t shm = proj_lds(cfg);
    if (lg.h == 13 && equal_arm)
        return exact ? launch(projection_kernel<true, 13, true>, nb, shm, stream, cfg, src, orb, lg)
                     : launch(projection_kernel<false, 13, true>, nb, shm, stream, cfg, src, orb, lg);
    if (lg.h == 13)
        return exact ? launch(projection_kernel<true, 13, false>, nb, shm, stream, cfg, src, orb, lg)
                     : launch(projection_kernel<false, 13, false>, nb, shm, stream, cfg, src, orb, lg);
    return exact ? launch(projection_kernel<true, 0, false>, nb, shm, stream, cfg, src, orb, lg)
                 : launch(projection_kernel<false, 0, false>, nb, shm, stream, cfg, src, orb, lg);
}

hipError_t launch_projection_batch(const ProjConfig &cfg, const Batch<ProjSource> &b,
                                   const OrbitView &orb, const LagrangeView &lg, bool exact,
                                   bool equal_arm, int grid_per_cu, hipStream_t stream)
{
    const int nb = batch_grid(b.total_tiles, grid_per_cu);
    const size_t shm = proj_lds(cfg);
    if (lg.h == 13 && equal_arm)
        return exact ? launch(projection_batch_kernel<true, 13, true>, nb, shm, stream, cfg, b, orb, lg)
                     : launch(projection_batch_kernel<false, 13, true>, nb, shm, stream, cfg, b, orb, lg);
    if (lg.h == 13)
        return exact ? launch(projection_batch_kernel<true, 13, false>, nb, shm, stream, cfg, b, orb, lg)
                     : launch(projection_batch_kernel<false, 13, false>, nb, shm, stream, cfg, b, orb, lg);
    return exact ? launch(projection_batch_kernel<true, 0, false>, nb, shm, stream, cfg, b, orb, lg)
                 : launch(projection_batch_kernel<false, 0, false>, nb, shm, stream, cfg, b, orb, lg);
}

// The stock kernels keep the light travel times in registers; the generic one in LDS.
static size_t tdi_lds(const TdiConfig &cfg, bool stock)
{
    return ((size_t)kNLinks * cfg.win_cap + (stock ? 0 : (size_t)kNLinks * kBlock)) * sizeof(double);
}

hipError_t launch_tdi(const TdiConfig &cfg, const TdiSource &src, const OrbitView &orb,
                      const LagrangeView &lg, bool exact, const StockTdi &st, hipStream_t stream)
{
    const int nb = tiles_for(src.num_delays, src.start, cfg.zero_edges);
    if (use_stock(st, lg)) {
        const size_t shm = tdi_lds(cfg, true);
        const RowMap &rm = st.rows;
#define FLR_STOCK(NX, EQ)                                                                   \
    (exact ? launch(tdi_stock_kernel<NX, true, EQ>, nb, shm, stream, cfg, src, orb, lg, rm) \
           : launch(tdi_stock_kernel<NX, false, EQ>, nb, shm, stream, cfg, src, orb, lg, rm))
        if (st.equal_arm) return st.generation == 1 ? FLR_STOCK(8, true) : FLR_STOCK(16, true);
        return st.generation == 1 ? FLR_STOCK(8, false) : FLR_STOCK(16, false);
#undef FLR_STOCK
    }
    const size_t shm = tdi_lds(cfg, false);
    if (lg.h == 13)
        return exact ? launch(tdi_kernel<true, 13>, nb, shm, stream, cfg, src, orb, lg)
                     : launch(tdi_kernel<false, 13>, nb, shm, stream, cfg, src, orb, lg);
    return exact ? launch(tdi_kernel<true, 0>, nb, shm, stream, cfg, src, orb, lg)
                 : launch(tdi_kernel<false, 0>, nb, shm, stream, cfg, src, orb, lg);
}

hipError_t launch_tdi_batch(const TdiConfig &cfg, const Batch<TdiSource> &b, const OrbitView &orb,
                            const LagrangeView &lg, bool exact, const StockTdi &st,
                            int grid_per_cu, hipStream_t stream)
{
    const int nb = batch_grid(b.total_tiles, grid_per_cu);
    if (use_stock(st, lg)) {
        const size_t shm = tdi_lds(cfg, true);
        const RowMap &rm = st.rows;
#define FLR_STOCK(NX, EQ)                                                                        \
    (exact ? launch(tdi_stock_batch_kernel<NX, true, EQ>, nb, shm, stream, cfg, b, orb, lg, rm)  \
           : launch(tdi_stock_batch_kernel<NX, false, EQ>, nb, shm, stream, cfg, b, orb, lg, rm))
        if (st.equal_arm) return st.generation == 1 ? FLR_STOCK(8, true) : FLR_STOCK(16, true);
        return st.generation == 1 ? FLR_STOCK(8, false) : FLR_STOCK(16, false);
#undef FLR_STOCK
    }
    const size_t shm = tdi_lds(cfg, false);
    if (lg.h == 13)
        return exact ? launch(tdi_batch_kernel<true, 13>, nb, shm, stream, cfg, b, orb, lg)
                     : launch(tdi_batch_kernel<false, 13>, nb, shm, stream, cfg, b, orb, lg);
    return exact ? launch(tdi_batch_kernel<true, 0>, nb, shm, stream, cfg, b, orb, lg)
                 : launch(tdi_batch_kernel<false, 0>, nb, shm, stream, cfg, b, orb, lg);
}

hipError_t launch_orbit_eval(const OrbitView &orb, const double *t, int n, int row, int sc,
                             double *ltt_out, double *pos_out, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(orbit_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, orb, t, n,
                       row, sc, ltt_out, pos_out);
    return hipGetLastError();
}

hipError_t launch_clear(int *word, hipStream_t stream)
{
    hipLaunchKernelGGL(clear_kernel, dim3(1), dim3(64), 0, stream, word);
    return hipGetLastError();
}

size_t fused_lds_bytes(const ProjConfig &pc, const TdiConfig &tc, bool generic, bool lti,
                       int passes, bool utv)
{
    if (utv)  // h window or the filters, then the padded rows (fused_tile_utv)
        return ((size_t)std::max(2 * pc.win_cap, 2 * kUtvPairs * kFpad) +
                (size_t)kNLinks * blk_pitch(tc.win_cap)) *
               sizeof(double);
    const size_t doubles =
        lti ? (size_t)kNLinks * blk_pitch(tc.win_cap) + 2 * kFpad
            : (size_t)kNLinks * tc.win_cap + (generic ? (size_t)kNLinks * kBlock : 0);
    return (size_t)pc.win_cap * sizeof(double2) + doubles * sizeof(double) +
           (lti ? (size_t)passes * kBlock : 0);
}

// One workgroup per source: each thread adds the tile records g0 + t, g0 + t + kBlock, ... of
// its source, then a fixed tree over the threads -- the same order whatever ran before.
__global__ __launch_bounds__(kBlock) void inner_reduce_kernel(const double *part,
                                                              const int *tile_start,
                                                              int num_sources, double *sums)
{
    constexpr int K = 3 * kInnerSums;
    __shared__ double red[K][kBlock];
    const int s = blockIdx.x, t = threadIdx.x;
    if (s >= num_sources) return;  // whole workgroup
    const int g0 = tile_start[s], g1 = tile_start[s + 1];
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (int g = g0 + t; g < g1; g += kBlock)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += part[(size_t)g * K + k];
#pragma unroll
    for (int k = 0; k < K; ++k) red[k][t] = acc[k];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int k = 0; k < K; ++k) red[k][t] += red[k][t + w];
        __syncthreads();
    }
    if (t < K) sums[(size_t)s * K + t] = red[t][0];
}

hipError_t launch_inner_reduce(const double *part, const int *tile_start, int num_sources,
                               double *sums, hipStream_t stream)
{
    return launch(inner_reduce_kernel, num_sources, 0, stream, part, tile_start, num_sources, sums);
}

hipError_t launch_fused_batch(const ProjConfig &pc, const TdiConfig &tc,
                              const Batch<FusedSource> &b, const OrbitView &orb,
                              const LagrangeView &lg, bool exact, bool mixed, int input,
                              int tile_passes, int grid_per_cu, const StockTdi &st,
                              hipStream_t stream)
{
    if (lg.h != 13) return hipErrorInvalidValue;  // fused kernels are built for order 25
    if (mixed) {
        switch (input) {
        case kInC64:
            return launch_fused_mixed(pc, tc, b, orb, lg, tile_passes, grid_per_cu, st, stream);
        case kInQ31:
            return launch_fused_q31(pc, tc, b, orb, lg, tile_passes, grid_per_cu, st, stream);
        case kInC128:
            return launch_fused_c128(pc, tc, b, orb, lg, tile_passes, grid_per_cu, st, stream);
        default:
            return hipErrorInvalidValue;
        }
    }
    if (tile_passes != 2) return hipErrorInvalidValue;
    return launch_fused_fp64(pc, tc, b, orb, lg, exact, grid_per_cu, st, stream);
}

}  // namespace flr
