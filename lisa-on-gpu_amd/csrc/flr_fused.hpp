// flr_fused.hpp -- the fused projection + TDI tile and its persistent kernel, instantiated by
// flr_fused_fp64.hip and flr_fused_mixed.hip.
#pragma once

#include "flr_paths.hpp"

namespace flr {

// ------------------------------------------------------------------ fused projection + TDI
// One tile = tc.tile output samples, sized on the host so that tile + TDI halos is exactly
// PASSES * kBlock window samples: the projection phase then runs PASSES full rounds of the
// workgroup and the TDI phase at most PASSES rounds (the last one partly idle). Stage h over the tile's TDI window plus
// the projection halo (converted to FP64), project the six links of every window sample into
// LDS (zero outside [proj_start, n - proj_start), as the reference's zeroed y_gw), then run TDI
// from LDS. y never touches HBM: per sample the launch reads h and writes the three outputs.
// The projected window overlaps the neighbouring tiles' by the TDI halo (~8% recompute at
// 512-sample tiles and 2nd-generation equal arms).
//
// TDI modes: kModeGeneric -- any table, unit by unit (tdi_units); kModeStock -- stock table,
// unequal arms, compile-time chains (tdi_stock_unequal); kModeGroups -- stock table
// with equal arms, shared delays per chain length (tdi_equal_arm, exact or fast); kModeLinear --
// the same, fast, with each group's units merged into one stencil per output (tdi_linear; the
// window columns are replaced by the combined series right after projection).
constexpr int kModeGeneric = 0;
constexpr int kModeGroups = 1;
constexpr int kModeLinear = 2;
constexpr int kModeStock = 3;  // stock table, unequal arms: tdi_stock_unequal

template <typename HT>
__device__ __forceinline__ double2 load_h(const void *h, int j)
{
    const HT v = reinterpret_cast<const HT *>(h)[j];
    return make_double2((double)v.x, (double)v.y);
}

template <typename OT>
__device__ __forceinline__ void store_out(void *out, size_t idx, double v)
{
    reinterpret_cast<OT *>(out)[idx] = (OT)v;
}

template <typename OT>
__device__ __forceinline__ void fused_store(const TdiConfig &tc, const FusedSource &src, int i,
                                            const double (&acc)[3], bool xyz_to_aet)
{
    const size_t n = (size_t)src.num_pts;
    double o[3];
    if (xyz_to_aet) {
        const double X = acc[0], Y = acc[1], Z = acc[2];
        o[0] = (Z - X) / tc.s2;
        o[1] = (X - 2.0 * Y + Z) / tc.s6;
        o[2] = (X + Y + Z) / tc.s3;
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = acc[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) store_out<OT>(src.out, c * n + i, o[c] * src.scale);
}

template <bool EXACT, int MODE, int NX, typename HT, typename OT, int PASSES>
__device__ __forceinline__ void fused_tile(const ProjConfig &pc, const TdiConfig &tc,
                                           const FusedSource &src, int tile, const OrbitView &orb,
                                           const LagrangeView &lg, const RowMap &rm, double *smem)
{
    constexpr int H = 13;
    constexpr int kPasses = PASSES;
    const int tile_len = tc.tile;  // <= PASSES * kBlock
    const int tid = threadIdx.x;
    const int n = src.num_pts;
    const int T = src.tdi_start;
    const int i0 = (tc.zero_edges ? 0 : T) + tile * tile_len;
    const int first = max(i0, T);
    const int last = min(i0 + tile_len, n - T) - 1;
    if (tc.zero_edges) {
#pragma unroll
        for (int r = 0; r < kPasses; ++r) {
            const int i = i0 + r * kBlock + tid;
            if (i < n && (i < T || i >= n - T))
                for (int c = 0; c < 3; ++c) store_out<OT>(src.out, (size_t)c * n + i, 0.0);
        }
    }
    if (first > last) return;  // whole tile outside the computed range

    const int cap = tc.win_cap;
    const int yw_lo = first - tc.halo_lo;
    const int yw_len = last + tc.halo_hi - yw_lo + 1;
    if (yw_len > cap || yw_lo < 0 || yw_lo + yw_len > n) {
        for (int r = 0; r < kPasses; ++r) {
            const int i = i0 + r * kBlock + tid;
            if (i < first || i > last) continue;
            atomicOr(tc.err, kErrWindow);
            for (int c = 0; c < 3; ++c) store_out<OT>(src.out, (size_t)c * n + i, __builtin_nan(""));
        }
        return;
    }
    const int hw_lo = max(0, yw_lo - pc.halo_lo);
    const int hw_len = min(n, yw_lo + yw_len + pc.halo_hi) - hw_lo;  // <= pc.win_cap (host-sized)
    double2 *hwin = reinterpret_cast<double2 *>(smem);
    double *ywin = smem + 2 * pc.win_cap;  // [6][cap]
    for (int jj = tid; jj < hw_len; jj += kBlock) hwin[jj] = load_h<HT>(src.h, hw_lo + jj);
    __syncthreads();

    const double t0_offset = src.t0 + (double)0 * pc.dt;
    const double y_deps = EXACT ? rcp_nr2(lg.deps) : 0.0;
    bool bad = false;
    for (int w = tid; w < yw_len; w += kBlock) {
        const int j = yw_lo + w;
        double *yc = ywin + w;
        if (j < src.proj_start || j >= n - src.proj_start) {
#pragma unroll
            for (int r = 0; r < kNLinks; ++r) yc[r * cap] = 0.0;
        } else {
            const double t = src.t0 + (double)j * pc.dt;
            bad |= project_links<EXACT, H, MODE == kModeGroups || MODE == kModeLinear>(pc, src.k, src.u, src.v, t, t0_offset, orb, lg, hwin,
                                           hw_lo, hw_len, y_deps,
                                           [&](int row, double val) { yc[row * cap] = val; });
        }
        if constexpr (MODE == kModeLinear) linear_series(yc, cap, rm, tc.aet, tc.s2, tc.s6, tc.s3);
    }
    __syncthreads();

    // Mixed precision, equal arms: composite time-invariant filters, referenced at the
    // source's first TDI sample so that every tile of a source (whatever the tile length or
    // the batch) applies the same filter.
    constexpr bool kLti = MODE == kModeLinear && std::is_same<OT, float>::value;
    double *fir = ywin + kNLinks * cap;  // [2][kFir]
    LtiRef<NX> lref;
    bool lti = false;
    if constexpr (kLti) {
        const double t_ref = src.t0 + (double)T * tc.dt;
        const double L_ref = orbit_val(orbit_node<EXACT>(orb, t_ref), rm.row[0]);
        lti_reference<NX>(t_ref, T, L_ref, t0_offset, tc.fs, lref);
        lti = lref.ok;
        if (lti && tid < 2 * kFir) fir[tid] = lti_filter_entry<NX>(lref, lg, tid);
        __syncthreads();
    }

#pragma unroll 1
    for (int r = 0; r < kPasses; ++r) {
        const int i = i0 + r * kBlock + tid;
        if (i < first || i > last) continue;
        const double t = src.t0 + (double)i * tc.dt;
        double acc[3] = {0.0, 0.0, 0.0};
        if constexpr (kLti) {
            const double L = orbit_val(orbit_node<EXACT>(orb, t), rm.row[0]);
            if (lti && lti_matches<NX>(t, i, L, t0_offset, tc.fs, lref, yw_lo, yw_len)) {
                lti_apply(ywin, cap, i + lref.kmin - yw_lo, fir, acc);
                fused_store<OT>(tc, src, i, acc, false);
                continue;
            }
        }
        if constexpr (MODE == kModeGeneric) {
            double *Ls = ywin + kNLinks * cap;  // [6][kBlock], one column per lane
            const OrbitNode nd = orbit_node<EXACT>(orb, t);
#pragma unroll
            for (int q = 0; q < kNLinks; ++q) Ls[q * kBlock + tid] = orbit_val(nd, q);
            bad |= tdi_units<EXACT, H>(tc, t, t0_offset, Ls + tid, kBlock, ywin, cap, yw_lo,
                                       yw_len, lg, y_deps, acc);
        } else if constexpr (MODE == kModeStock) {
            const OrbitNode nd = orbit_node<EXACT>(orb, t);
            double L[6];
#pragma unroll
            for (int q = 0; q < kNLinks; ++q) L[q] = orbit_val(nd, rm.row[q]);
            int rowoff[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) rowoff[k] = rm.row[k] * cap - yw_lo;
            bad |= !tdi_stock_unequal<NX, EXACT>(t, L, t0_offset, tc.fs, ywin, rowoff, yw_lo,
                                                 yw_len, lg, y_deps, acc);
        } else {
            const double L = orbit_val(orbit_node<EXACT>(orb, t), rm.row[0]);
            if constexpr (MODE == kModeGroups) {
                int rowoff[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) rowoff[k] = rm.row[k] * cap - yw_lo;
                bad |= !tdi_equal_arm<NX, EXACT>(t, L, t0_offset, tc.fs, ywin, rowoff, yw_lo,
                                                 yw_len, lg, y_deps, acc);
            } else {
                bad |= !tdi_linear<NX>(t, L, t0_offset, tc.fs, ywin, cap, yw_lo, yw_len, lg, acc);
            }
        }
        fused_store<OT>(tc, src, i, acc, MODE != kModeLinear && tc.aet);
    }
    if (bad) atomicOr(tc.err, kErrWindow);
}

template <bool EXACT, int MODE, int NX, typename HT, typename OT, int PASSES>
__global__ __launch_bounds__(kBlock) void fused_batch_kernel(ProjConfig pc, TdiConfig tc,
                                                             Batch<FusedSource> batch,
                                                             OrbitView orb, LagrangeView lg,
                                                             RowMap rm)
{
    extern __shared__ double smem[];
    FLR_BATCH_LOOP(batch, (fused_tile<EXACT, MODE, NX, HT, OT, PASSES>(pc, tc, src, tile, orb, lg, rm, smem)));
}

// Launch one fused variant over a batch: the TDI mode follows the table (stock + equal arms:
// groups for exact stencils, linear merging for fast ones; otherwise unit by unit).
template <bool EXACT, typename HT, typename OT, int PASSES>
hipError_t launch_fused_variant(const ProjConfig &pc, const TdiConfig &tc, const Batch<FusedSource> &b,
                                const OrbitView &orb, const LagrangeView &lg, const StockTdi &st,
                                int grid_per_cu, hipStream_t stream)
{
    const int nb = batch_grid(b.total_tiles, grid_per_cu);
    const size_t shm = fused_lds_bytes(pc, tc, !use_stock(st, lg));
    const RowMap &rm = st.rows;
    constexpr int kEq = EXACT ? kModeGroups : kModeLinear;
#define FLR_FUSED(M, NX) \
    launch(fused_batch_kernel<EXACT, M, NX, HT, OT, PASSES>, nb, shm, stream, pc, tc, b, orb, lg, rm)
    if (use_equal_arm(st, lg)) return st.generation == 1 ? FLR_FUSED(kEq, 8) : FLR_FUSED(kEq, 16);
    if (use_stock(st, lg))
        return st.generation == 1 ? FLR_FUSED(kModeStock, 8) : FLR_FUSED(kModeStock, 16);
    return FLR_FUSED(kModeGeneric, 0);
#undef FLR_FUSED
}

}  // namespace flr
