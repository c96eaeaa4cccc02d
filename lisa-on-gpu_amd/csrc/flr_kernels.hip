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
//   with bit-identical results.
// tdi_tile: any TDI table. The 6 y rows of the window are staged once; each lane walks every
//   unit of every channel for its sample, summing units of a channel in index order in
//   registers (no atomics), then stores X, Y, Z or A, E, T.
// tdi_equal_arm_tile: stock tables when all light travel times are equal (flr_tdi_stock.hpp).
//
// Each tile function is launched two ways: one workgroup per tile of a single source, or a
// persistent batch kernel whose workgroups walk contiguous runs of a flattened (source, tile)
// list -- many sources of uneven length in one launch.
#include "flr_device.hpp"
#include "flr_tdi_stock.hpp"

namespace flr {

// Sample range of tile `tile` of a source: lanes i in [i0, i0 + kBlock), computed samples
// [lo_valid, hi_valid), and the tile's first/last computed sample.
struct TileRange {
    int i0, lo_valid, hi_valid, first, last;
};

__device__ __forceinline__ TileRange tile_range(int tile, int num_delays, int start, int zero_edges)
{
    TileRange r;
    r.i0 = (zero_edges ? 0 : start) + tile * kBlock;
    r.lo_valid = start;
    r.hi_valid = num_delays - start;
    r.first = max(r.i0, r.lo_valid);
    r.last = min(r.i0 + kBlock, r.hi_valid) - 1;
    return r;
}

// ------------------------------------------------------------------ projection

template <bool EXACT, int H>
__device__ __forceinline__ void projection_tile(const ProjConfig &cfg, const ProjSource &src,
                                                int tile, const OrbitView &orb,
                                                const LagrangeView &lg, double2 *hwin)
{
    const int h = H > 0 ? H : lg.h;
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
    const OrbitNode nr = orbit_node(orb, t);
    __syncthreads();
    if (!valid) return;
    const double y_deps = EXACT ? rcp_nr2(lg.deps) : 0.0;

    // Per receiving spacecraft: its stencil (shared by its two links), then each link's
    // emitter geometry and stencil; the window is sized from the orbit extent, so nothing has
    // to be known about the delays before it is staged. Operation order per link follows
    // LISAResponse.cu:559-668.
    bool bad = false;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        double xr[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) xr[c] = orbit_val(nr, 6 + 3 * s + c);
        int d_rec;
        double e_rec;
        split_delay(t - dot3(src.k, xr) * kCInv, t0_offset, cfg.fs, d_rec, e_rec);
        const int lr = d_rec - win_lo;
        const bool br = (lr - (h - 1) < 0) || (lr + h >= win_len);
        double2 hr = make_double2(0.0, 0.0);
        if (!br) hr = stencil_c<EXACT, H>(hwin + lr, e_rec, lg, y_deps);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int row = cfg.rec_row[s][q];
            const int se = cfg.em_sc[s][q];
            const double L = orbit_val(nr, row);
            const double t_em = t - L;
            const OrbitNode ne = orbit_node(orb, t_em);
            double xe[3], n[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) xe[c] = orbit_val(ne, 6 + 3 * se + c);
            double norm = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double nt = xr[c] - xe[c];
                n[c] = nt;
                norm += nt * nt;
            }
            norm = sqrt(norm);
#pragma unroll
            for (int c = 0; c < 3; ++c) n[c] = n[c] / norm;
            const double un = dot3(src.u, n);
            const double vn = dot3(src.v, n);
            const double xi_p = 0.5 * ((un * un) - (vn * vn));
            const double xi_c = un * vn;
            const double kn = dot3(src.k, n);
            int d_em;
            double e_em;
            split_delay(t_em - dot3(src.k, xe) * kCInv, t0_offset, cfg.fs, d_em, e_em);
            const int le = d_em - win_lo;
            const bool bq = br || (le - (h - 1) < 0) || (le + h >= win_len);
            double out = __builtin_nan("");
            if (!bq) {
                const double2 he = stencil_c<EXACT, H>(hwin + le, e_em, lg, y_deps);
                const double pre = 1. / (1. - kn);
                const double large = (he.x - hr.x) * xi_p + (he.y - hr.y) * xi_c;
                out = pre * large;
            }
            bad |= bq;
            src.y[(size_t)row * src.num_delays + i] = out;
        }
    }
    if (bad) atomicOr(cfg.err, kErrWindow);
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
    const int h = H > 0 ? H : lg.h;
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
        const OrbitNode nd = orbit_node(orb, t);
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

    bool bad = false;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        for (int u = cfg.ch_begin[c]; u < cfg.ch_begin[c + 1]; ++u) {
            const TdiUnit un = cfg.units[u];
            // Same-t delay chain: every light travel time is taken at the undelayed t and
            // subtracted in chain order (LISAResponse.cu:314-333).
            double delay = t;
            for (int s = 0; s < un.chain_len; ++s) {
                const int row = cfg.chain_rows[un.chain_start + s];
                if (row >= 0) delay -= Ls[row * kBlock + tid];
            }
            int d;
            double e;
            split_delay(delay, t0_offset, cfg.fs, d, e);
            const int ld = d - win_lo;
            double val;
            if ((ld - (h - 1) >= 0) && (ld + h < win_len)) {
                val = stencil<EXACT, H>(ywin + un.base_row * cfg.win_cap + ld, e, lg, y_deps);
            } else {
                bad = true;
                val = __builtin_nan("");
            }
            val *= un.sign;
            acc[c] += val;
        }
    }
    if (bad) atomicOr(cfg.err, kErrWindow);
    tdi_store(cfg, src, i, acc);
}

// ------------------------------------------------------------------ equal-arm stock TDI
// Stock 1st/2nd-generation tables at order 25 when all six light travel times of the orbit
// table are bitwise equal (equal-arm orbits): the distinct delays t - m*L (m = chain length
// 0..3 or 0..7) are split once each and their stencil reciprocals (exact form) or weights
// (fast form) are shared by the units of that length -- 8 groups instead of 48 units for 2nd
// generation, with bit-identical results.

template <int NX, bool EXACT>
__device__ __forceinline__ void tdi_equal_arm_tile(const TdiConfig &cfg, const TdiSource &src,
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
    const double L = orbit_val(orbit_node(orb, t), rm.row[0]);
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
    if (!tdi_equal_arm<NX, EXACT>(t, L, t0_offset, cfg.fs, ywin, rowoff, win_lo, win_len, lg,
                                  y_deps, acc))
        atomicOr(cfg.err, kErrWindow);
    tdi_store(cfg, src, i, acc);
}

// ------------------------------------------------------------------ kernels

// Source of flattened tile g: tile_start is a prefix sum over sources (uniform search).
template <typename Source>
__device__ __forceinline__ int batch_source(const Batch<Source> &b, int g)
{
    int lo = 0, hi = b.num_sources;  // tile_start[lo] <= g < tile_start[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (b.tile_start[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Workgroup b of G walks the contiguous tiles [b*T/G, (b+1)*T/G): neighbouring tiles share
// halo samples, which then stay in the same XCD's L2.
#define FLR_BATCH_LOOP(batch, body)                                                        \
    do {                                                                                   \
        const long long T_ = (batch).total_tiles;                                          \
        const int g0_ = (int)((T_ * blockIdx.x) / gridDim.x);                              \
        const int g1_ = (int)((T_ * (blockIdx.x + 1)) / gridDim.x);                        \
        for (int g = g0_; g < g1_; ++g) {                                                  \
            __syncthreads(); /* the previous tile's LDS readers are done */                \
            const int s_ = batch_source(batch, g);                                         \
            const auto src = (batch).sources[s_];                                          \
            const int tile = g - (batch).tile_start[s_];                                   \
            body;                                                                          \
        }                                                                                  \
    } while (0)

template <bool EXACT, int H>
__global__ __launch_bounds__(kBlock) void projection_kernel(ProjConfig cfg, ProjSource src,
                                                            OrbitView orb, LagrangeView lg)
{
    extern __shared__ double2 hwin[];
    projection_tile<EXACT, H>(cfg, src, blockIdx.x, orb, lg, hwin);
}

template <bool EXACT, int H>
__global__ __launch_bounds__(kBlock) void projection_batch_kernel(ProjConfig cfg,
                                                                  Batch<ProjSource> batch,
                                                                  OrbitView orb, LagrangeView lg)
{
    extern __shared__ double2 hwin[];
    FLR_BATCH_LOOP(batch, (projection_tile<EXACT, H>(cfg, src, tile, orb, lg, hwin)));
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

template <int NX, bool EXACT>
__global__ __launch_bounds__(kBlock) void tdi_equal_arm_kernel(TdiConfig cfg, TdiSource src,
                                                               OrbitView orb, LagrangeView lg,
                                                               RowMap rm)
{
    extern __shared__ double smem[];
    tdi_equal_arm_tile<NX, EXACT>(cfg, src, blockIdx.x, orb, lg, rm, smem);
}

template <int NX, bool EXACT>
__global__ __launch_bounds__(kBlock) void tdi_equal_arm_batch_kernel(TdiConfig cfg,
                                                                     Batch<TdiSource> batch,
                                                                     OrbitView orb,
                                                                     LagrangeView lg, RowMap rm)
{
    extern __shared__ double smem[];
    FLR_BATCH_LOOP(batch, (tdi_equal_arm_tile<NX, EXACT>(cfg, src, tile, orb, lg, rm, smem)));
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

template <typename Kernel>
static hipError_t set_lds_limit(Kernel k, size_t bytes)
{
    if (bytes <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// Persistent batch grid: enough workgroups to fill every CU a few times over, never more than
// there are tiles.
static int batch_grid(int total_tiles)
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
    }
    const int grid = cus * 8;
    return total_tiles < grid ? total_tiles : grid;
}

template <typename Kernel, typename... Args>
static hipError_t launch(Kernel k, int nblocks, size_t shmem, hipStream_t stream, Args... args)
{
    if (nblocks <= 0) return hipSuccess;
    hipError_t err = set_lds_limit(k, shmem);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(kBlock), shmem, stream, args...);
    return hipGetLastError();
}

static size_t proj_lds(const ProjConfig &cfg) { return (size_t)cfg.win_cap * sizeof(double2); }

hipError_t launch_projection(const ProjConfig &cfg, const ProjSource &src, const OrbitView &orb,
                             const LagrangeView &lg, bool exact, hipStream_t stream)
{
    const int nb = tiles_for(src.num_delays, src.start, cfg.zero_edges);
    const size_t shm = proj_lds(cfg);
    if (lg.h == 13)
        return exact ? launch(projection_kernel<true, 13>, nb, shm, stream, cfg, src, orb, lg)
                     : launch(projection_kernel<false, 13>, nb, shm, stream, cfg, src, orb, lg);
    return exact ? launch(projection_kernel<true, 0>, nb, shm, stream, cfg, src, orb, lg)
                 : launch(projection_kernel<false, 0>, nb, shm, stream, cfg, src, orb, lg);
}

hipError_t launch_projection_batch(const ProjConfig &cfg, const Batch<ProjSource> &b,
                                   const OrbitView &orb, const LagrangeView &lg, bool exact,
                                   hipStream_t stream)
{
    const int nb = batch_grid(b.total_tiles);
    const size_t shm = proj_lds(cfg);
    if (lg.h == 13)
        return exact ? launch(projection_batch_kernel<true, 13>, nb, shm, stream, cfg, b, orb, lg)
                     : launch(projection_batch_kernel<false, 13>, nb, shm, stream, cfg, b, orb, lg);
    return exact ? launch(projection_batch_kernel<true, 0>, nb, shm, stream, cfg, b, orb, lg)
                 : launch(projection_batch_kernel<false, 0>, nb, shm, stream, cfg, b, orb, lg);
}

static bool use_equal_arm(const StockTdi &st, const LagrangeView &lg)
{
    return st.generation != 0 && st.equal_arm && lg.h == 13;
}

static size_t tdi_lds(const TdiConfig &cfg, bool equal_arm)
{
    return ((size_t)kNLinks * cfg.win_cap + (equal_arm ? 0 : (size_t)kNLinks * kBlock)) * sizeof(double);
}

hipError_t launch_tdi(const TdiConfig &cfg, const TdiSource &src, const OrbitView &orb,
                      const LagrangeView &lg, bool exact, const StockTdi &st, hipStream_t stream)
{
    const int nb = tiles_for(src.num_delays, src.start, cfg.zero_edges);
    if (use_equal_arm(st, lg)) {
        const size_t shm = tdi_lds(cfg, true);
        const RowMap &rm = st.rows;
        if (st.generation == 1)
            return exact ? launch(tdi_equal_arm_kernel<8, true>, nb, shm, stream, cfg, src, orb, lg, rm)
                         : launch(tdi_equal_arm_kernel<8, false>, nb, shm, stream, cfg, src, orb, lg, rm);
        return exact ? launch(tdi_equal_arm_kernel<16, true>, nb, shm, stream, cfg, src, orb, lg, rm)
                     : launch(tdi_equal_arm_kernel<16, false>, nb, shm, stream, cfg, src, orb, lg, rm);
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
                            hipStream_t stream)
{
    const int nb = batch_grid(b.total_tiles);
    if (use_equal_arm(st, lg)) {
        const size_t shm = tdi_lds(cfg, true);
        const RowMap &rm = st.rows;
        if (st.generation == 1)
            return exact ? launch(tdi_equal_arm_batch_kernel<8, true>, nb, shm, stream, cfg, b, orb, lg, rm)
                         : launch(tdi_equal_arm_batch_kernel<8, false>, nb, shm, stream, cfg, b, orb, lg, rm);
        return exact ? launch(tdi_equal_arm_batch_kernel<16, true>, nb, shm, stream, cfg, b, orb, lg, rm)
                     : launch(tdi_equal_arm_batch_kernel<16, false>, nb, shm, stream, cfg, b, orb, lg, rm);
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

}  // namespace flr
