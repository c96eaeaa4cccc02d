// flr_fused.hpp -- the fused projection + TDI tile and its persistent kernel, instantiated by
// flr_fused_fp64.hip and flr_fused_mixed.hip.
#pragma once

#include "flr_paths.hpp"

namespace flr {

// ------------------------------------------------------------------ fused projection + TDI
// One tile = tc.tile output samples, sized on the host so that tile + TDI halos is exactly
// PASSES * TPB window samples (TPB threads per workgroup): the projection phase then runs PASSES full rounds of the
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

// Copy the E table into LDS (read back by the fast stencils after the next barrier), followed
// by the folded stencil's pairs {1/E[j-1], j(j+1)/E[j-1]} (zero where E is).
constexpr int kELdsWords = 3 * kMaxHalf;
__device__ __forceinline__ void load_E_lds(const LagrangeView &lg, double *s_E)
{
    const int x = threadIdx.x;
    if (x < kMaxHalf) {
        const double E = lg.E[x];
        const double inv = E != 0.0 ? 1.0 / E : 0.0;
        s_E[x] = E;
        s_E[kMaxHalf + 2 * x] = inv;
        s_E[kMaxHalf + 2 * x + 1] = (double)((x + 1) * (x + 2)) * inv;
    }
}

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

// ------------------------------------------------------------------ inner-product epilogue
// In an inner-product launch (tc.inner_part set) a tile's outputs are not stored: each output
// h (FP64, before the float32 rounding, times the source's scale) meets the data series d at
// the same place, and per channel the tile sums d h, h h, h and (-1)^(i - T) h -- per thread
// (InnerAcc), then per wave (a fixed butterfly) into per-wave LDS slots (InnerShared), then
// per tile in wave order into inner_part [tile][3][kInnerSums]; launch_inner_reduce adds the
// tiles of each source in a fixed order. The result does not depend on timing.
struct InnerAcc {
    double a[3][kInnerSums];
};
struct InnerShared {
    double v[3][kInnerSums][16];  // [channel][sum][wave], up to 16 waves per workgroup
};

__device__ __forceinline__ void inner_zero(InnerAcc &ia)
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < kInnerSums; ++k) ia.a[c][k] = 0.0;
}

// Channel c's output v (before the source's scale) at sample i, into a.
template <typename OT>
__device__ __forceinline__ void inner_add(const TdiConfig &tc, const FusedSource &src, int c,
                                          int i, double v, double (&a)[kInnerSums])
{
    const OT *d = reinterpret_cast<const OT *>(
        tc.inner_data + (reinterpret_cast<const char *>(src.out) - tc.inner_out));
    const double w = v * src.scale;
    const double x = (double)d[(size_t)c * src.num_pts + i];
    a[0] = __builtin_fma(x, w, a[0]);
    a[1] = __builtin_fma(w, w, a[1]);
    a[2] += w;
    a[3] += ((i - src.tdi_start) & 1) ? -w : w;
}

// The wave's sums of a (every lane of the wave takes part) added to its slots of channel c.
__device__ __forceinline__ void inner_commit(InnerShared &sh, int c, const double (&a)[kInnerSums])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kInnerSums; ++k) {
        double v = a[k];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) sh.v[c][k][wave] += v;
    }
}

__device__ __forceinline__ void inner_commit_all(InnerShared &sh, const InnerAcc &ia)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) inner_commit(sh, c, ia.a[c]);
}

__device__ __forceinline__ void inner_clear(InnerShared &sh)
{
    double *v = &sh.v[0][0][0];
    for (int x = threadIdx.x; x < 3 * kInnerSums * 16; x += blockDim.x) v[x] = 0.0;
}

// End of tile g: its record from the wave slots (in wave order), which are cleared for the
// next tile (ordered before its commits by the persistent loop's barrier).
__device__ __forceinline__ void inner_flush(const TdiConfig &tc, InnerShared &sh, int g)
{
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 3 * kInnerSums) {
        const int c = t / kInnerSums, k = t - c * kInnerSums;
        double sum = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            sum += sh.v[c][k][w];
            sh.v[c][k][w] = 0.0;
        }
        tc.inner_part[(size_t)g * (3 * kInnerSums) + t] = sum;
    }
}

// Store the three outputs of sample i (or, with ia in an inner-product launch, add them to
// its sums).
template <typename OT>
__device__ __forceinline__ void fused_store(const TdiConfig &tc, const FusedSource &src, int i,
                                            const double (&acc)[3], bool xyz_to_aet,
                                            InnerAcc *ia = nullptr)
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
    if (ia) {
#pragma unroll
        for (int c = 0; c < 3; ++c) inner_add<OT>(tc, src, c, i, o[c], ia->a[c]);
        return;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) store_out<OT>(src.out, c * n + i, o[c] * src.scale);
}

// The composite filters' reference for source src: its first TDI sample
// (t0 + tdi_start * dt) and the arm length there.
template <int NX, bool EXACT>
__device__ __forceinline__ double lti_source_reference(const TdiConfig &tc, const FusedSource &src,
                                                       const OrbitView &orb, const RowMap &rm,
                                                       LtiRef<NX> &lref)
{
    const int T = src.tdi_start;
    const double t_ref = src.t0 + (double)T * tc.dt;
    const double L_ref = orbit_val(orbit_node<EXACT>(orb, t_ref), rm.row[0]);
    lti_reference<NX>(t_ref, T, L_ref, src.t0, tc.fs, lref);
    return L_ref;
}

// Unit in the last place of x > 0.
__device__ __forceinline__ double ulp_of(double x) { return ldexp(1.0, ilogb(x) - 52); }

// True when lti_matches holds for every sample of [first, last], so that the per-sample test
// can be skipped (the outputs are then the same bits); kmin and L_ref are the source's
// (lti_source_reference): (1) the filter's taps fit the window at
// the tile's two ends, hence everywhere; (2) the arm length is L_ref, with zero slope, at every
// orbit node the tile's times fall in, so each sample's chain delays t - m L differ from the
// reference's only by rounding; (3) that rounding cannot move a position by kLtiTol: per
// sample, t = t0 + i dt (1 ulp), m subtractions of L, the offset and the product with fs
// each round by half an ulp of a magnitude below tmax (times fs), and dt fs may differ from 1
// by r, which moves sample i by (i - T) r against the reference.
template <int NX>
__device__ __forceinline__ bool lti_tile_matches(const TdiConfig &tc, const FusedSource &src,
                                                 const OrbitView &orb, const RowMap &rm, int kmin,
                                                 double L_ref, int first, int last)
{
    if (kmin + tc.halo_lo < 0 || kmin + kFir > tc.halo_hi + 1) return false;
    const int row = rm.row[0];
    const double *pa = orbit_node<false>(orb, src.t0 + (double)first * tc.dt).p;
    const double *pb = orbit_node<false>(orb, src.t0 + (double)last * tc.dt).p;
    for (const double *p = pa; p <= pb; p += kOrbitStride)
        if (p[kOrbitVals + row] != 0.0 || p[row] != L_ref) return false;
    constexpr int M = LtiRef<NX>::NG - 1;  // longest chain
    const double tmax = 1.001 * (2.0 * fabs(src.t0) + (double)last * tc.dt + M * fabs(L_ref));
    const double fs = tc.fs;
    const double per_sample = (0.5 * M + 1.5) * ulp_of(tmax) * fs + ulp_of(tmax * fs);
    const double drift = 1.01 * fabs(fma(tc.dt, fs, -1.0)) * (double)(last - src.tdi_start);
    return 2.0 * per_sample + drift + 1e-13 <= kLtiTol;
}

// One workgroup per source: the two composite filters of the equal-arm stock TDI and the
// reference they are built from (a kLtiRecord record per source in tc.lti_fir; the filters are
// zero when the reference does not fit).
template <int NX>
__global__ __launch_bounds__(kBlock) void lti_filter_kernel(TdiConfig tc,
                                                            Batch<FusedSource> batch,
                                                            OrbitView orb, LagrangeView lg,
                                                            RowMap rm)
{
    const int s = blockIdx.x;
    const int x = threadIdx.x;
    if (s >= batch.num_sources || x >= kLtiRecord) return;
    const FusedSource src = batch.sources[s];
    LtiRef<NX> lref;
    const double L_ref = lti_source_reference<NX, false>(tc, src, orb, rm, lref);
    double v = 0.0;
    if (x < 2 * kFir)
        v = lref.ok ? lti_filter_entry<NX>(lref, lg, x) : 0.0;
    else if (x == kLtiLref)
        v = L_ref;
    else if (x == kLtiKmin)
        v = (double)lref.kmin;
    else if (x == kLtiOk)
        v = lref.ok ? 1.0 : 0.0;
    tc.lti_fir[(size_t)s * kLtiRecord + x] = v;
}

// Composite-filter TDI of one tile on the matrix cores. Outputs are taken in blocks of 16
// aligned to the source's first TDI sample (block q = samples T + 16q .. T + 16q + 15, so a
// sample's arithmetic is the same whatever tile holds it); per output o and block,
//   out[i_b + r] = sum_kind sum_{m=0}^{55} F_kind[m - r] S_{o,kind}[i_b + kmin + m],
// which is D = A B with A[r][m] = F[m - r] (Toeplitz, the padded filters) and
// B[m][block] = the series: one v_mfma_f64_16x16x4f64 per 4 taps of 16 blocks x 16 outputs
// (28 per group of 16 blocks of one output). Operand layout: lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; the result's element v is D[4v + l / 16][l % 16]. The rows are
// block-interleaved (blk_col), so each B operand is one load at a per-lane base plus a
// compile-time offset; edge blocks read the zeroed guard and tail slots wherever F is zero,
// their rows outside the tile, and samples left to the per-sample path (lti_ok = 0 unless
// all_ok), are not stored.
template <typename OT, int TPB>
__device__ __forceinline__ void lti_mfma(const TdiConfig &tc, const FusedSource &src, int first,
                                         int last, int i0, int yw_lo, int yw_len, int kmin,
                                         const double *ywin, int pitch, int R, const double *fpad,
                                         const unsigned char *lti_ok, bool all_ok,
                                         InnerShared *ish)
{
    const bool inner = ish != nullptr;  // (inner-product kernels only)
    typedef double d4 __attribute__((ext_vector_type(4)));
    const int T = src.tdi_start;
    const int q0 = (first - T) >> 4, q1 = (last - T) >> 4;  // first >= T
    const int ngc = (q1 - q0 + 16) >> 4;                     // groups of 16 blocks per output
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nb = lane & 15, kq = lane >> 4;
    const size_t n = (size_t)src.num_pts;
    for (int g = wave; g < 3 * ngc; g += TPB / 64) {
        const int o = g / ngc;
        const int qb = q0 + 16 * (g - o * ngc) + nb;
        // window column of this lane's tap step 0 (tap m = kq of block min(qb, q1)); a block
        // with a sample that uses the filter starts at column -15 or later (lti_matches), one
        // whose samples all take the per-sample path is held to the guard
        const int w0 = max(T + 16 * min(qb, q1) + kmin - yw_lo + kq, -16);
        int b[4];  // tap step j = 4a + i at b[i] + a
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = blk_col(w0 + 4 * i, R);
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kind = 0; kind < 2; ++kind) {
            const double *row = ywin + (2 * o + kind) * pitch;
            const double *f = fpad + kind * kFpad + kFpadLo + kq - nb;  // f[4j] = F[4j + kq - nb]
#pragma unroll
            for (int j = 0; j < 14; ++j) {
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(f[4 * j], row[b[j & 3] + (j >> 2)], acc,
                                                          0, 0, 0);
                __asm__ volatile("" ::: "memory");  // one operand pair in flight: no register growth
            }
        }
        double a[kInnerSums] = {0.0, 0.0, 0.0, 0.0};
        if (qb <= q1) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = T + 16 * qb + 4 * v + kq;
                if (i >= first && i <= last && (all_ok || lti_ok[i - i0])) {
                    if (inner)
                        inner_add<OT>(tc, src, o, i, acc[v], a);
                    else
                        store_out<OT>(src.out, o * n + i, acc[v] * src.scale);
                }
            }
        }
        if (inner) inner_commit(*ish, o, a);  // the whole wave, after the per-lane branch
    }
}

// The slowly-varying-arm TDI of one tile on the matrix cores (utv_unit): per group of 16
// blocks of 16 samples (blocks aligned to the source's first TDI sample, as lti_mfma), the
// three channels' F0 and F1 sums over their four base links -- 24 accumulations of 14
// v_mfma_f64_16x16x4f64 each, the link series' operand shared by the pair's two filters --
// then X_c = F0 part + (i - m) F1 part, A, E, T formed from X, Y, Z in FP64, and stored.
// filt: [2][kUtvPairs][kFpad] (F0 then F1, padded as fpad); link rows canonical at ywin +
// k * pitch.
template <typename OT, int TPB>
__device__ __forceinline__ void utv_mfma(const TdiConfig &tc, const FusedSource &src, int first,
                                         int last, int m, int yw_lo, int yw_len, int kmin,
                                         const double *ywin, int pitch, int R, const double *filt,
                                         InnerShared *ish)
{
    const bool inner = ish != nullptr;  // (inner-product kernels only)
    typedef double d4 __attribute__((ext_vector_type(4)));
    const int T = src.tdi_start;
    const int q0 = (first - T) >> 4, q1 = (last - T) >> 4;
    const int ngc = (q1 - q0 + 16) >> 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nb = lane & 15, kq = lane >> 4;
    const size_t n = (size_t)src.num_pts;
    for (int g = wave; g < ngc; g += TPB / 64) {
        const int qb = q0 + 16 * g + nb;
        // window column of this lane's tap step 0 (>= -15: the tile's taps fit its window)
        const int w0 = T + 16 * min(qb, q1) + kmin - yw_lo + kq;
        int b[4];  // tap step j = 4a + i at b[i] + a
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = blk_col(w0 + 4 * i, R);
        d4 a0[3], a1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) a0[c] = a1[c] = d4{0.0, 0.0, 0.0, 0.0};
        const double *fl = filt + kFpadLo + kq - nb;  // fl[p * kFpad + 4j] = F0_p[4j + kq - nb]
#pragma unroll
        for (int j = 0; j < 14; ++j) {
            const int col = b[j & 3] + (j >> 2);
            static_for<6>([&](auto l) {  // canonical link l: one operand for its two pairs
                const double yv = ywin[l * pitch + col];
                static_for<3>([&](auto c) {
                    static_for<kUtvLinks>([&](auto q) {
                        if constexpr (stock::canon(stock::cyc(kUtvLink[q], c)) == l) {
                            constexpr int p = c * kUtvLinks + q;
                            a0[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(fl[p * kFpad + 4 * j], yv,
                                                                         a0[c], 0, 0, 0);
                            a1[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                fl[(kUtvPairs + p) * kFpad + 4 * j], yv, a1[c], 0, 0, 0);
                        }
                    });
                });
                __asm__ volatile("" ::: "memory");  // one link's operands in flight at a time
            });
        }
        InnerAcc ia;
        inner_zero(ia);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = T + 16 * qb + 4 * v + kq;
            if (qb > q1 || i < first || i > last) continue;
            const double di = (double)(i - m);
            const double X = __builtin_fma(di, a1[0][v], a0[0][v]);
            const double Y = __builtin_fma(di, a1[1][v], a0[1][v]);
            const double Z = __builtin_fma(di, a1[2][v], a0[2][v]);
            double o0 = X, o1 = Y, o2 = Z;
            if (tc.aet) {
                o0 = (Z - X) * kInvSqrt2;
                o1 = (X - 2.0 * Y + Z) * kInvSqrt6;
                o2 = (X + Y + Z) * kInvSqrt3;
            }
            if (inner) {
                inner_add<OT>(tc, src, 0, i, o0, ia.a[0]);
                inner_add<OT>(tc, src, 1, i, o1, ia.a[1]);
                inner_add<OT>(tc, src, 2, i, o2, ia.a[2]);
            } else {
                store_out<OT>(src.out, i, o0 * src.scale);
                store_out<OT>(src.out, n + i, o1 * src.scale);
                store_out<OT>(src.out, 2 * n + i, o2 * src.scale);
            }
        }
        if (inner) inner_commit_all(*ish, ia);
    }
}

// SAMPLE_PASS (composite-filter path): the per-sample test and fallback stencils for tiles that
// cannot be cleared as a whole are compiled in. Without them (when the host has proved that
// every tile clears, tc.lti_uniform) the kernel is smaller and its projection loop keeps its
// operands in registers; a tile that fails anyway is reported (kErrLti) and written as NaN.
template <bool EXACT, int MODE, int NX, typename HT, typename OT, int PASSES, int TPB = kBlock,
          bool SAMPLE_PASS = true, bool STAGED = false>
__device__ __forceinline__ void fused_tile(const ProjConfig &pc, const TdiConfig &tc,
                                           const FusedSource &src, int s, int tile,
                                           const OrbitView &orb, const LagrangeView &lg,
                                           const RowMap &rm, double *smem, const double *s_E,
                                           InnerShared *ish = nullptr)
{
    constexpr int H = 13;
    constexpr int kPasses = PASSES;
    const int tile_len = tc.tile;  // <= PASSES * TPB
    const int tid = threadIdx.x;
    const int n = src.num_pts;
    const int T = src.tdi_start;
    const int i0 = (tc.zero_edges ? 0 : T) + tile * tile_len;
    const int first = max(i0, T);
    const int last = min(i0 + tile_len, n - T) - 1;
    // an inner-product kernel (composite-filter form only): sums instead of stores
    const bool inner = ish != nullptr;
    InnerAcc ia;
    inner_zero(ia);
    if (tc.zero_edges && !inner) {
#pragma unroll
        for (int r = 0; r < kPasses; ++r) {
            const int i = i0 + r * TPB + tid;
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
            const int i = i0 + r * TPB + tid;
            if (i < first || i > last) continue;
            atomicOr(tc.err, kErrWindow);
            if (inner) ia.a[0][0] = __builtin_nan("");
            else
                for (int c = 0; c < 3; ++c) store_out<OT>(src.out, (size_t)c * n + i, __builtin_nan(""));
        }
        if (inner) inner_commit_all(*ish, ia);
        return;
    }
    // Mixed precision, equal arms: the stock TDI is two composite time-invariant filters per
    // output (built once per source by lti_filter_kernel, referenced at the source's first TDI
    // sample so that every tile applies the same filter) applied on the matrix cores; the
    // window rows are then block-interleaved (blk_col) for the matrix operands' loads.
    constexpr bool kLti = MODE == kModeLinear && std::is_same<OT, float>::value;
    const int R = blk_R(cap);
    const int pitch = kLti ? blk_pitch(cap) : cap;
    const int hw_lo = max(0, yw_lo - pc.halo_lo);
    const int hw_len = min(n, yw_lo + yw_len + pc.halo_hi) - hw_lo;  // <= pc.win_cap (host-sized)
    double2 *hwin = reinterpret_cast<double2 *>(smem);
    double *ywin = smem + 2 * pc.win_cap;  // [6][pitch]
    double *fpad = ywin + kNLinks * pitch;  // kLti: [2][kFpad]
    unsigned char *lti_ok = reinterpret_cast<unsigned char *>(fpad + 2 * kFpad);  // [PASSES * TPB]
    for (int jj = tid; jj < hw_len; jj += TPB) hwin[jj] = load_h<HT>(src.h, hw_lo + jj);
    if constexpr (kLti) blk_zero(ywin, pitch, R, yw_len, tid, TPB);  // before the window's stores
    OrbitStage ost{};
    if constexpr (STAGED) {  // the orbit nodes of the tile's projection
        __shared__ double s_orb[kOrbitStage * kOrbitStride];
        ost = stage_orbit_nodes(orb, src.t0 + (double)yw_lo * pc.dt, s_orb, tid, TPB);
    }
    LtiRef<NX> lref;
    bool lti = false, tile_ok = false;
    if constexpr (kLti) {
        // the source's reference, from its filter record (lti_filter_kernel), and its filters
        // copied beside the staging (read by lti_mfma after the projection's barrier); the
        // chain positions themselves are only needed when the per-sample test runs
        const double *fir = tc.lti_fir + (size_t)s * kLtiRecord;
        const double L_ref = fir[kLtiLref];
        lref.kmin = (int)fir[kLtiKmin];
        lti = fir[kLtiOk] != 0.0;
        tile_ok = lti && lti_tile_matches<NX>(tc, src, orb, rm, lref.kmin, L_ref, first, last);
        if (lti) {
            for (int x = tid; x < 2 * kFpad; x += TPB) {
                const int kind = x / kFpad, k = x - kind * kFpad - kFpadLo;
                fpad[x] = (k >= 0 && k < kFir) ? fir[kind * kFir + k] : 0.0;
            }
        }
    }
    __syncthreads();

    const double t0_offset = src.t0 + (double)0 * pc.dt;
    const double y_deps = EXACT ? rcp_nr2(lg.deps) : 0.0;
    bool bad = false;
    for (int w = tid; w < yw_len; w += TPB) {
        const int j = yw_lo + w;
        int col = kLti ? blk_col(w, R) : w;
        pin(col);  // an opaque column: the padded index is not re-derived inside the projection
        double *yc = ywin + col;
        if (j < src.proj_start || j >= n - src.proj_start) {
#pragma unroll
            for (int r = 0; r < kNLinks; ++r) yc[r * pitch] = 0.0;
        } else {
            const double t = src.t0 + (double)j * pc.dt;
            // linear mode: the link rows are only read back by linear_series just below, so they
            // are kept in canonical link order (compile-time rows)
            // fast forms: E read from the workgroup's LDS copy (s_E), not kept in registers;
            // float32 outputs: the folded stencil (stencil_fold_c_lo)
            bad |= project_links<EXACT, H, MODE == kModeGroups || MODE == kModeLinear,
                                 MODE == kModeLinear,
                                 EXACT ? kEArg : std::is_same<OT, float>::value ? kELdsFold : kELds,
                                 STAGED>(
                pc, src.k, src.u, src.v, t, t0_offset, orb, lg, hwin, hw_lo, hw_len, y_deps, s_E,
                ost, [&](int row, double val) { yc[row * pitch] = val; });
        }
        if constexpr (MODE == kModeLinear)
            linear_series<true>(yc, pitch, rm, tc.aet);
    }
    __syncthreads();

    if constexpr (kLti && SAMPLE_PASS) {
        if (lti && !tile_ok) (void)lti_source_reference<NX, EXACT>(tc, src, orb, rm, lref);
    }

    if constexpr (kLti && !SAMPLE_PASS) {
        if (!tile_ok) {  // cannot happen when the host's proof holds: report it loudly
            for (int r = 0; r < kPasses; ++r) {
                const int i = i0 + r * TPB + tid;
                if (i < first || i > last) continue;
                if (inner) ia.a[0][0] = __builtin_nan("");
                else
                    for (int c = 0; c < 3; ++c)
                        store_out<OT>(src.out, (size_t)c * n + i, __builtin_nan(""));
            }
            if (tid == 0) atomicOr(tc.err, kErrLti);
            lti = false;  // no matrix-core pass either
        }
    }
#pragma unroll 1
    for (int r = 0; r < ((tile_ok || (kLti && !SAMPLE_PASS)) ? 0 : kPasses); ++r) {
        // tile_ok: every sample goes to the matrix-core pass below
        const int i = i0 + r * TPB + tid;
        if (i < first || i > last) continue;
        const double t = src.t0 + (double)i * tc.dt;
        double acc[3] = {0.0, 0.0, 0.0};
        if constexpr (kLti) {
            if constexpr (SAMPLE_PASS) {
                // samples whose delays match the reference are left to the matrix-core pass
                // below; the others (if any) take the per-sample stencils here
                const double L = orbit_val(orbit_node<EXACT>(orb, t), rm.row[0]);
                const bool ok =
                    lti && lti_matches<NX>(t, i, L, t0_offset, tc.fs, lref, yw_lo, yw_len);
                lti_ok[i - i0] = ok;
                if (ok) continue;
                bad |= !tdi_linear_padded<NX, true>(t, L, t0_offset, tc.fs, ywin, pitch, R,
                                                    yw_lo, yw_len, lg, acc);
                fused_store<OT>(tc, src, i, acc, false, inner ? &ia : nullptr);
            }
            continue;
        }
        if constexpr (MODE == kModeGeneric) {
            static_assert(TPB == kBlock, "the generic TDI scratch has kBlock columns");
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
                bad |= !tdi_linear<NX, true>(t, L, t0_offset, tc.fs, ywin, cap, yw_lo, yw_len, lg,
                                             acc);
            }
        }
        fused_store<OT>(tc, src, i, acc, MODE != kModeLinear && tc.aet);
    }
    if constexpr (kLti) {
        if (inner) inner_commit_all(*ish, ia);  // the per-sample samples' sums
        if (lti) {
            // lti_ok and the per-sample samples' sums written (the rows and filters were, by
            // the projection's barrier)
            if (SAMPLE_PASS || inner) __syncthreads();
            lti_mfma<OT, TPB>(tc, src, first, last, i0, yw_lo, yw_len, lref.kmin, ywin, pitch, R,
                              fpad, lti_ok, tile_ok, ish);
        }
    }
    if (bad) atomicOr(tc.err, kErrWindow);
}

template <bool EXACT, int MODE, int NX, typename HT, typename OT, int PASSES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
void fused_batch_kernel(ProjConfig pc, TdiConfig tc,
                                                             Batch<FusedSource> batch,
                                                             OrbitView orb, LagrangeView lg,
                                                             RowMap rm)
{
    extern __shared__ double smem[];
    __shared__ __align__(16) double s_E[kELdsWords];  // the E table, read by the fast stencils
    load_E_lds(lg, s_E);  // visible after the loop's first barrier
    FLR_BATCH_LOOP(batch, (fused_tile<EXACT, MODE, NX, HT, OT, PASSES>(pc, tc, src, s_, tile, orb, lg, rm, smem, s_E)));
}

// The mixed equal-arm form (composite filters on the matrix cores), held to 4 waves per SIMD:
// its VGPRs and the matrix accumulator share one 512-entry register file per lane.
// TPB threads per workgroup: 256, or 512 for the 1024-sample window (two rounds of 512 lanes;
// its 72 KB of LDS then still leave 16 waves per CU, with half the halo recompute of 512).
// INNER: the inner-product kernel (tc.inner_part set) -- sums against the data series instead
// of stores.
template <int NX, typename HT, int PASSES, int TPB, bool SAMPLE_PASS, bool INNER, bool STAGED>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4)))
void fused_lti_kernel(ProjConfig pc, TdiConfig tc, Batch<FusedSource> batch, OrbitView orb,
                      LagrangeView lg, RowMap rm)
{
    extern __shared__ double smem[];
    __shared__ __align__(16) double s_E[kELdsWords];
    InnerShared *ish = nullptr;
    if constexpr (INNER) {
        __shared__ InnerShared ish_slots;
        ish = &ish_slots;
        inner_clear(*ish);  // visible after the loop's first barrier
    }
    load_E_lds(lg, s_E);
    FLR_BATCH_LOOP_UNIFORM(batch, {
        fused_tile<false, kModeLinear, NX, HT, float, PASSES, TPB, SAMPLE_PASS, STAGED>(
            pc, tc, src, s_, tile, orb, lg, rm, smem, s_E, ish);
        if constexpr (INNER) inner_flush(tc, *ish, g);
    });
}

// The mixed form with unequal stock arms (order 25): fused_tile's staging and projection (link
// rows canonical and padded), the tile's unit model (utv_unit) computed beside the projection,
// then either the slowly-varying-arm filters on the matrix cores (utv_mfma) or, for a tile
// that fails the model's tests, the per-sample stencils. LDS: the h window, whose space holds
// the filters once the projection is done ([2][kUtvPairs][kFpad] doubles at most), then the
// six padded rows.
template <int NX, typename HT, int PASSES, int TPB, bool STAGED>
__device__ __forceinline__ void fused_tile_utv(const ProjConfig &pc, const TdiConfig &tc,
                                               const FusedSource &src, int tile,
                                               const OrbitView &orb, const LagrangeView &lg,
                                               const RowMap &rm, double *smem, const double *s_E,
                                               const int *s_chain, UtvShared &sh,
                                               InnerShared *ish)
{
    constexpr int H = 13;
    const int tile_len = tc.tile;
    const int tid = threadIdx.x;
    const int n = src.num_pts;
    const int T = src.tdi_start;
    const int i0 = (tc.zero_edges ? 0 : T) + tile * tile_len;
    const int first = max(i0, T);
    const int last = min(i0 + tile_len, n - T) - 1;
    const bool inner = ish != nullptr;  // (inner-product kernels only)
    InnerAcc ia;
    inner_zero(ia);
    if (tc.zero_edges && !inner) {
#pragma unroll
        for (int r = 0; r < PASSES; ++r) {
            const int i = i0 + r * TPB + tid;
            if (i < n && (i < T || i >= n - T))
                for (int c = 0; c < 3; ++c) store_out<float>(src.out, (size_t)c * n + i, 0.0);
        }
    }
    if (first > last) return;
    const int cap = tc.win_cap;
    const int yw_lo = first - tc.halo_lo;
    const int yw_len = last + tc.halo_hi - yw_lo + 1;
    if (yw_len > cap || yw_lo < 0 || yw_lo + yw_len > n) {
        for (int r = 0; r < PASSES; ++r) {
            const int i = i0 + r * TPB + tid;
            if (i < first || i > last) continue;
            atomicOr(tc.err, kErrWindow);
            if (inner) ia.a[0][0] = __builtin_nan("");
            else
                for (int c = 0; c < 3; ++c) store_out<float>(src.out, (size_t)c * n + i, __builtin_nan(""));
        }
        if (inner) inner_commit_all(*ish, ia);
        return;
    }
    const int R = blk_R(cap);
    const int pitch = blk_pitch(cap);
    const int hw_lo = max(0, yw_lo - pc.halo_lo);
    const int hw_len = min(n, yw_lo + yw_len + pc.halo_hi) - hw_lo;
    double2 *hwin = reinterpret_cast<double2 *>(smem);
    double *filt = smem;  // after the projection
    double *ywin = smem + max(2 * pc.win_cap, 2 * kUtvPairs * kFpad);  // [6][pitch]
    const double t0_offset = src.t0 + (double)0 * pc.dt;
    const int m = (first + last) >> 1;
    // the times and light travel times at the tile's first, middle and last sample
    if (tid < 3)
        utv_reference_times(sh, tid, tid == 0 ? first : tid == 1 ? m : last, src.t0, tc.dt, orb, rm);
    if (tid == 0) {
        sh.kmin = 1 << 30;
        sh.kmax = -(1 << 30);
        sh.ok = 1;
    }
    for (int jj = tid; jj < hw_len; jj += TPB) hwin[jj] = load_h<HT>(src.h, hw_lo + jj);
    blk_zero(ywin, pitch, R, yw_len, tid, TPB);  // before the window's stores
    OrbitStage ost{};
    if constexpr (STAGED) {  // the orbit nodes of the tile's projection
        __shared__ double s_orb[kOrbitStage * kOrbitStride];
        ost = stage_orbit_nodes(orb, src.t0 + (double)yw_lo * pc.dt, s_orb, tid, TPB);
    }
    __syncthreads();
    if (tid < 3 * NX) utv_unit<NX>(sh, tid, s_chain[tid], first, m, last, t0_offset, tc.fs, lg);

    bool bad = false;
    for (int w = tid; w < yw_len; w += TPB) {
        const int j = yw_lo + w;
        int col = blk_col(w, R);
        pin(col);
        double *yc = ywin + col;
        if (j < src.proj_start || j >= n - src.proj_start) {
#pragma unroll
            for (int r = 0; r < kNLinks; ++r) yc[r * pitch] = 0.0;
        } else {
            const double t = src.t0 + (double)j * pc.dt;
            bad |= project_links<false, H, false, true, kELdsFold, STAGED>(
                pc, src.k, src.u, src.v, t, t0_offset, orb, lg, hwin, hw_lo, hw_len, 0.0, s_E,
                ost, [&](int row, double val) { yc[row * pitch] = val; });
        }
    }
    __syncthreads();  // rows and the unit model written; the h window is free

    const int kmin = sh.kmin;
    const bool tile_ok = sh.ok && sh.kmax - kmin + 1 <= kFir && kmin >= -tc.halo_lo &&
                         sh.kmax <= tc.halo_hi;
    if (tile_ok) {
        const int nt = sh.kmax - kmin + 1;
        for (int x = tid; x < kUtvPairs * kFpad; x += TPB) {
            const int p = x / kFpad, k = x - p * kFpad - kFpadLo;
            double f0 = 0.0, f1 = 0.0;
            if (k >= 0 && k < nt) utv_filter_entry<NX>(sh, lg, s_E, p, k, f0, f1);
            filt[x] = f0;
            filt[kUtvPairs * kFpad + x] = f1;
        }
        __syncthreads();
        utv_mfma<float, TPB>(tc, src, first, last, m, yw_lo, yw_len, kmin, ywin, pitch, R, filt,
                             ish);
    } else {
#pragma unroll 1
        for (int r = 0; r < PASSES; ++r) {
            const int i = i0 + r * TPB + tid;
            if (i < first || i > last) continue;
            const double t = src.t0 + (double)i * tc.dt;
            const OrbitNode nd = orbit_node<false>(orb, t);
            double L[6];
#pragma unroll
            for (int q = 0; q < kNLinks; ++q) L[q] = orbit_val(nd, rm.row[q]);
            double acc[3] = {0.0, 0.0, 0.0};
            bad |= !tdi_stock_unequal_padded<NX>(t, L, t0_offset, tc.fs, ywin, pitch, R, yw_lo,
                                                 yw_len, lg, acc);
            fused_store<float>(tc, src, i, acc, tc.aet, inner ? &ia : nullptr);
        }
        if (inner) inner_commit_all(*ish, ia);
    }
    if (bad) atomicOr(tc.err, kErrWindow);
}

template <int NX, typename HT, int PASSES, int TPB, bool INNER, bool STAGED>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4)))
void fused_utv_kernel(ProjConfig pc, TdiConfig tc, Batch<FusedSource> batch, OrbitView orb,
                      LagrangeView lg, RowMap rm)
{
    extern __shared__ double smem[];
    __shared__ __align__(16) double s_E[kELdsWords];
    __shared__ UtvShared sh;
    InnerShared *ish = nullptr;
    if constexpr (INNER) {
        __shared__ InnerShared ish_slots;
        ish = &ish_slots;
        inner_clear(*ish);
    }
    __shared__ int s_chain[3 * NX];  // utv_chain_word of every unit
    load_E_lds(lg, s_E);
    if (threadIdx.x < 3 * NX) s_chain[threadIdx.x] = utv_chain_word(threadIdx.x, NX);
    FLR_BATCH_LOOP(batch, {  // (the loop's first barrier publishes s_E and s_chain)
        fused_tile_utv<NX, HT, PASSES, TPB, STAGED>(pc, tc, src, tile, orb, lg, rm, smem, s_E,
                                                    s_chain, sh, ish);
        if constexpr (INNER) inner_flush(tc, *ish, g);
    });
}

// Launch one fused variant over a batch: the TDI mode follows the table (stock + equal arms:
// groups for exact stencils, linear merging for fast ones; otherwise unit by unit).
template <bool EXACT, typename HT, typename OT, int PASSES>
hipError_t launch_fused_variant(const ProjConfig &pc, const TdiConfig &tc, const Batch<FusedSource> &b,
                                const OrbitView &orb, const LagrangeView &lg, const StockTdi &st,
                                int grid_per_cu, hipStream_t stream)
{
    const int nb = batch_grid(b.total_tiles, grid_per_cu);
    const bool lti = std::is_same<OT, float>::value && !EXACT && use_equal_arm(st, lg);
    const size_t shm = fused_lds_bytes(pc, tc, !use_stock(st, lg), lti, PASSES);
    const RowMap &rm = st.rows;
    constexpr int kEq = EXACT ? kModeGroups : kModeLinear;
#define FLR_FUSED(M, NX) \
    launch(fused_batch_kernel<EXACT, M, NX, HT, OT, PASSES>, nb, shm, stream, pc, tc, b, orb, lg, rm)
    if (use_equal_arm(st, lg)) {
        if constexpr (kEq == kModeLinear && std::is_same<OT, float>::value) {
            // tc.lti_fir holds the sources' composite filters (launch_lti_filters)
            if (!tc.lti_fir) return hipErrorInvalidValue;
            // the 1024-sample window runs as two rounds of 512 lanes
            constexpr int TPB = PASSES == 4 ? 2 * kBlock : kBlock;
            constexpr int P = PASSES == 4 ? 2 : PASSES;
            // without the per-sample pass when the host proved every tile clears as a whole
            // with the orbit nodes staged in LDS where every tile's fit (tc.orbit_stage); the
            // unstaged kernel keeps the per-sample pass (orbit tables finer than the tiles)
#define FLR_LTI(NXV, SP, IN, ST) \
    launch_tpb(fused_lti_kernel<NXV, HT, P, TPB, SP, IN, ST>, nb, TPB, shm, stream, pc, tc, b, orb, \
               lg, rm)
#define FLR_LTI_PICK(NXV, IN)                                                                      \
    (!tc.orbit_stage     ? FLR_LTI(NXV, true, IN, false)                                         \
     : tc.lti_uniform    ? FLR_LTI(NXV, false, IN, true)                                         \
                         : FLR_LTI(NXV, true, IN, true))
            if (tc.inner_part)
                return st.generation == 1 ? FLR_LTI_PICK(8, true) : FLR_LTI_PICK(16, true);
            return st.generation == 1 ? FLR_LTI_PICK(8, false) : FLR_LTI_PICK(16, false);
#undef FLR_LTI_PICK
#undef FLR_LTI
        } else {
            return st.generation == 1 ? FLR_FUSED(kEq, 8) : FLR_FUSED(kEq, 16);
        }
    }
    if (use_stock(st, lg)) {
        if constexpr (!EXACT && std::is_same<OT, float>::value) {
          if (tc.utv) {
            // mixed, unequal arms: the slowly-varying-arm filters (fused_tile_utv)
            constexpr int TPB = PASSES == 4 ? 2 * kBlock : kBlock;
            constexpr int P = PASSES == 4 ? 2 : PASSES;
            const size_t shu = fused_lds_bytes(pc, tc, false, false, PASSES, true);
#define FLR_UTV(NXV, IN, ST) \
    launch_tpb(fused_utv_kernel<NXV, HT, P, TPB, IN, ST>, nb, TPB, shu, stream, pc, tc, b, orb, lg, \
               rm)
#define FLR_UTV_PICK(NXV, IN) (tc.orbit_stage ? FLR_UTV(NXV, IN, true) : FLR_UTV(NXV, IN, false))
            if (tc.inner_part)
                return st.generation == 1 ? FLR_UTV_PICK(8, true) : FLR_UTV_PICK(16, true);
            return st.generation == 1 ? FLR_UTV_PICK(8, false) : FLR_UTV_PICK(16, false);
#undef FLR_UTV_PICK
#undef FLR_UTV
          }
          if (tc.inner_part) return hipErrorInvalidValue;  // inner products: filter form only
        }
        return st.generation == 1 ? FLR_FUSED(kModeStock, 8) : FLR_FUSED(kModeStock, 16);
    }
    return FLR_FUSED(kModeGeneric, 0);
#undef FLR_FUSED
}

}  // namespace flr
