// flr_paths.hpp -- per-sample paths and launch plumbing shared by the kernel translation units:
// the tile range of a source, the six-link projection of one sample, the unit-by-unit TDI of
// one sample, the persistent (source, tile) loop and the launch helpers.
#pragma once

#include "flr_device.hpp"
#include "flr_tdi_stock.hpp"

#include <type_traits>

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

// Projections of all six links at time t from the staged h window (hwin[0] = sample win_lo):
// per receiving spacecraft its stencil (shared by its two links), then each link's emitter
// geometry and stencil; store(row, value) receives y of table row `row` (CANON: of the link's
// canonical index stock::canon, a compile-time row). Operation order per link follows
// LISAResponse.cu:559-668. Returns true if a stencil fell outside the window (its links get
// NaN). ESRC: where the fast stencils read E (stencil_E; E_lds for kELds); kELdsFold: the
// folded stencil and one-Newton-step reciprocals of the float32-output forms. STAGED (fast
// forms): orbit nodes from the tile's LDS stage `ost` (orbit_node_staged; a lookup outside it
// makes the sample's links NaN and returns true).
template <bool EXACT, int H, bool EQ, bool CANON = false, int ESRC = kEArg, bool STAGED = false,
          typename Store>
__device__ __forceinline__ bool project_links(const ProjConfig &cfg, const double (&k)[3],
                                              const double (&u)[3], const double (&v)[3],
                                              double t, double t0_offset, const OrbitView &orb,
                                              const LagrangeView &lg, const double2 *hwin,
                                              int win_lo, int win_len, double y_deps,
                                              const double *E_lds, const OrbitStage &ost,
                                              Store &&store)
{
    static_assert(!(STAGED && EXACT), "staged orbit nodes: fast forms only");
    const int h = H > 0 ? H : lg.h;
    bool miss = false;
    auto node = [&](double tt) -> OrbitNode {
        if constexpr (STAGED)
            return orbit_node_staged(orb, ost, tt, miss);
        else
            return orbit_node<EXACT>(orb, tt);
    };
    const OrbitNode nr = node(t);
    bool bad = false;
    // one complex stencil at window offset l (sample d = win_lo + l) and fraction e; the fast
    // form with a compile-time order reads every tap at a constant offset from the lowest one
    auto cstencil = [&](int l, double e) -> double2 {
        if constexpr (!EXACT && H > 0) {
            int lo = l - (H - 1);
            pin(lo);  // an opaque base: the taps are not re-addressed from another register
            if constexpr (ESRC == kELdsFold)
                return stencil_fold_c_lo<H>(hwin + lo, e, lg, E_lds);
            else
                return stencil_fast_c_lo<H, ESRC>(hwin + lo, e, lg, E_lds);
        } else {
            return stencil_c<EXACT, H>(hwin + l, e, lg, y_deps);
        }
    };
    // EQ (all six light travel times bitwise equal): the emission time t - L, the emitter's
    // position and so the emitter stencil depend on the emitting spacecraft only -- 3 emitter
    // stencils instead of 6, each bit-identical to the per-link one.
    double2 hem[3];
    bool bem[3];
    // float32-output forms: the emitters' positions kept for the links' geometry (the same
    // values; the other forms reload them from the node, with fewer registers live)
    constexpr bool kKeepXe = EQ && ESRC == kELdsFold;
    double xem[kKeepXe ? 3 : 1][3];
    OrbitNode ne_eq{};
    if constexpr (EQ) {
        const double t_em_eq = t - orbit_val(nr, cfg.link_row[0][1]);
        ne_eq = node(t_em_eq);
        static_for<3>([&](auto se) {
            double xe[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) xe[c] = orbit_val(ne_eq, 6 + 3 * se + c);
            if constexpr (kKeepXe)
#pragma unroll
                for (int c = 0; c < 3; ++c) xem[se][c] = xe[c];
            int d_em;
            double e_em;
            split_delay(t_em_eq - dot3(k, xe) * kCInv, t0_offset, cfg.fs, d_em, e_em);
            const int le = d_em - win_lo;
            bem[se] = miss || (le - (h - 1) < 0) || (le + h >= win_len);
            hem[se] = make_double2(0.0, 0.0);
            if (!bem[se]) hem[se] = cstencil(le, e_em);
        });
    }
    static_for<3>([&](auto s) {
        double xr[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) xr[c] = orbit_val(nr, 6 + 3 * s + c);
        int d_rec;
        double e_rec;
        split_delay(t - dot3(k, xr) * kCInv, t0_offset, cfg.fs, d_rec, e_rec);
        const int lr = d_rec - win_lo;
        const bool br = (lr - (h - 1) < 0) || (lr + h >= win_len);
        double2 hr = make_double2(0.0, 0.0);
        if (!br) hr = cstencil(lr, e_rec);
        static_for<2>([&](auto q) {
            constexpr int se = (s + 1 + q) % 3;  // the emitters of receiver s
            const int row = CANON ? stock::canon(10 * (s + 1) + se + 1) : cfg.link_row[s][se];
            double xe[3], n[3];
            double2 he = make_double2(0.0, 0.0);
            bool bq;
            if constexpr (EQ) {
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    xe[c] = kKeepXe ? xem[kKeepXe ? (int)se : 0][c] : orbit_val(ne_eq, 6 + 3 * se + c);
                bq = br || bem[se];
                he = hem[se];
            } else {
                const double t_em = t - orbit_val(nr, row);
                const OrbitNode ne = node(t_em);
#pragma unroll
                for (int c = 0; c < 3; ++c) xe[c] = orbit_val(ne, 6 + 3 * se + c);
                int d_em;
                double e_em;
                split_delay(t_em - dot3(k, xe) * kCInv, t0_offset, cfg.fs, d_em, e_em);
                const int le = d_em - win_lo;
                bq = br || miss || (le - (h - 1) < 0) || (le + h >= win_len);
                if (!bq) he = cstencil(le, e_em);
            }
            double norm = 0.0;
            if constexpr (EXACT) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double nt = xr[c] - xe[c];
                    n[c] = nt;
                    norm += nt * nt;
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) n[c] = xr[c] - xe[c];
                norm = dot3_form<false>(n, n);
            }
            if constexpr (EXACT) {
                norm = sqrt(norm);
#pragma unroll
                for (int c = 0; c < 3; ++c) n[c] = n[c] / norm;
            } else {  // one reciprocal square root instead of a root and three divisions
                const double inv = ESRC == kELdsFold ? rsqrt_pos1(norm) : rsqrt_pos(norm);
#pragma unroll
                for (int c = 0; c < 3; ++c) n[c] = n[c] * inv;
            }
            const double un = dot3_form<EXACT>(u, n);
            const double vn = dot3_form<EXACT>(v, n);
            const double xi_p = EXACT ? 0.5 * ((un * un) - (vn * vn))
                                      : 0.5 * __builtin_fma(un, un, -(vn * vn));
            const double xi_c = un * vn;
            const double kn = dot3_form<EXACT>(k, n);
            double out = __builtin_nan("");
            if (!bq) {
                const double pre = EXACT ? 1. / (1. - kn)
                                   : ESRC == kELdsFold ? rcp_nr1(1. - kn) : rcp_nr2(1. - kn);
                const double large =
                    EXACT ? (he.x - hr.x) * xi_p + (he.y - hr.y) * xi_c
                          : __builtin_fma(he.x - hr.x, xi_p, (he.y - hr.y) * xi_c);
                out = pre * large;
            }
            bad |= bq;
            store(row, out);
        });
    });
    return bad;
}

// ------------------------------------------------------------------ TDI

// Every unit of every channel at time t, summed per channel in index order into acc. Ls[r *
// ls_stride] is the light travel time of table row r at t; row r of the y window starts at
// ywin + r * cap. Returns true if a stencil fell outside the window (its channel gets NaN).
template <bool EXACT, int H>
__device__ __forceinline__ bool tdi_units(const TdiConfig &cfg, double t, double t0_offset,
                                          const double *Ls, int ls_stride, const double *ywin,
                                          int cap, int win_lo, int win_len, const LagrangeView &lg,
                                          double y_deps, double (&acc)[3])
{
    const int h = H > 0 ? H : lg.h;
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
                if (row >= 0) delay -= Ls[row * ls_stride];
            }
            int d;
            double e;
            split_delay(delay, t0_offset, cfg.fs, d, e);
            const int ld = d - win_lo;
            double val;
            if ((ld - (h - 1) >= 0) && (ld + h < win_len)) {
                val = stencil<EXACT, H>(ywin + un.base_row * cap + ld, e, lg, y_deps);
            } else {
                bad = true;
                val = __builtin_nan("");
            }
            val *= un.sign;
            acc[c] += val;
        }
    }
    return bad;
}

// ------------------------------------------------------------------ persistent loop

// Source of flattened tile g: tile_start is a prefix sum over sources (uniform search).
// A record every lane reads at the same address, made wave-uniform word by word: it then sits
// in scalar registers instead of ~30 vector registers (the compiler cannot prove the kernel's
// stores leave the record alone, so it loads it with vector loads).
template <typename R>
__device__ __forceinline__ R uniform_record(const R &r)
{
    static_assert(sizeof(R) % 4 == 0, "word-sized record");
    R out;
    const int *in = reinterpret_cast<const int *>(&r);
    int *o = reinterpret_cast<int *>(&out);
#pragma unroll
    for (int w = 0; w < (int)(sizeof(R) / 4); ++w) o[w] = __builtin_amdgcn_readfirstlane(in[w]);
    return out;
}

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
// halo samples, which then stay in the same XCD's L2. The body sees the tile's global index g,
// its source s_ and record src, and its index `tile` within the source.
#define FLR_BATCH_LOOP_AS(LOAD, batch, ...)                                                \
    do {                                                                                   \
        const long long T_ = (batch).total_tiles;                                          \
        const int g0_ = (int)((T_ * blockIdx.x) / gridDim.x);                              \
        const int g1_ = (int)((T_ * (blockIdx.x + 1)) / gridDim.x);                        \
        for (int g = g0_; g < g1_; ++g) {                                                  \
            __syncthreads(); /* the previous tile's LDS readers are done */                \
            const int s_ = batch_source(batch, g);                                         \
            const auto src = LOAD((batch).sources[s_]);                                    \
            const int tile = g - (batch).tile_start[s_];                                   \
            __VA_ARGS__;                                                                   \
        }                                                                                  \
    } while (0)
#define FLR_BATCH_LOOP(batch, ...) FLR_BATCH_LOOP_AS(, batch, __VA_ARGS__)
// The same with the source record in scalar registers (uniform_record).
#define FLR_BATCH_LOOP_UNIFORM(batch, ...) FLR_BATCH_LOOP_AS(uniform_record, batch, __VA_ARGS__)

// ------------------------------------------------------------------ launch helpers

template <typename Kernel>
inline hipError_t set_lds_limit(Kernel k, size_t bytes)
{
    if (bytes <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// Persistent batch grid: grid_per_cu workgroups per CU (enough to fill every CU a few times
// over), never more than there are tiles (flr_kernels.hip).
int batch_grid(int total_tiles, int grid_per_cu = 32);

template <typename Kernel, typename... Args>
inline hipError_t launch_tpb(Kernel k, int nblocks, int threads, size_t shmem, hipStream_t stream,
                             Args... args)
{
    if (nblocks <= 0) return hipSuccess;
    hipError_t err = set_lds_limit(k, shmem);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(threads), shmem, stream, args...);
    return hipGetLastError();
}

template <typename Kernel, typename... Args>
inline hipError_t launch(Kernel k, int nblocks, size_t shmem, hipStream_t stream, Args... args)
{
    if (nblocks <= 0) return hipSuccess;
    hipError_t err = set_lds_limit(k, shmem);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(kBlock), shmem, stream, args...);
    return hipGetLastError();
}

// The stock-table kernels (compile-time chains, order 25); equal arms select their shared-delay
// form.
inline bool use_stock(const StockTdi &st, const LagrangeView &lg)
{
    return st.generation != 0 && lg.h == 13;
}

inline bool use_equal_arm(const StockTdi &st, const LagrangeView &lg)
{
    return use_stock(st, lg) && st.equal_arm;
}

}  // namespace flr
