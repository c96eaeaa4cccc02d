// flr_internal.hpp -- parameter blocks shared by the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/flr.h"

namespace flr {

constexpr double kCInv = 3.3356409519815204e-09;  // 1/c in s/m (LISAResponse.hh:7)
constexpr int kNLinks = 6;
constexpr int kMaxHalf = 20;      // half point count h = (order+1)/2, order <= 39 (reference MAX_ORDER 40)
constexpr int kMaxUnits = 256;    // TDI units (signed, delayed link terms)
constexpr int kMaxChain = 4096;   // total delay-chain entries over all units
constexpr int kBlock = 256;       // threads per workgroup (4 wave64) = samples per tile

// Orbit table, packed per node: 15 values [ltt row 0..5 | x sc1 xyz | sc2 xyz | sc3 xyz]
// followed by their 15 slopes (v[k+1] - v[k]) / dt, so that one node's data sits in 240 B.
constexpr int kOrbitVals = 15;
constexpr int kOrbitStride = 2 * kOrbitVals;

// Device error word bits.
constexpr int kErrWindow = 1;     // stencil outside the staged window (or the input array)
constexpr int kErrLti = 2;        // a composite-filter tile failed a check the host had proved

struct OrbitView {
    const double *table;  // [N][kOrbitStride]
    double dt;
    int N;
    double inv_dt;        // 1 / dt (fast-form node lookup)
    double max_ltt;       // largest light travel time in the table (s)
};

struct LagrangeView {
    const double2 *Apair;   // [num_A-1]: (A[k], A[k+1]) -- exact form
    const double2 *Aslope;  // [num_A-1]: (A[k], A[k+1]-A[k]) -- fast form
    double deps;
    double inv_deps;
    int num_A;
    int h;                  // half point count
    double E[kMaxHalf];     // E[j-1], j = 1..h-1
};

// ------------------------------------------------------------------ projection
// Shared by every source of a launch. link_row[s][e] is the table row (and y row) of the link
// received by spacecraft s (0..2) from spacecraft e (-1 for s == e).
struct ProjConfig {
    double dt, fs;
    int link_row[3][3];
    int halo_lo, halo_hi;   // staged window = [first - halo_lo, last + halo_hi]
    int win_cap;            // LDS window capacity (complex samples)
    int zero_edges;
    int *err;
};

// One source: t_i = t0 + i*dt unless t_data is given.
struct ProjSource {
    double *y;              // [6][num_delays]
    const double2 *h;       // [num_inputs] h+ + i hx
    const double *t_data;   // [num_delays] or nullptr
    int num_delays;
    int num_inputs;
    int start;              // projections_start_ind
    int pad;
    double t0;
    double k[3], u[3], v[3];
};

// ------------------------------------------------------------------ TDI
struct TdiUnit {
    double sign;
    int base_row;           // y row of the delayed link
    int chain_start;        // into chain_rows
    int chain_len;
    int pad;
};

struct TdiConfig {
    double dt, fs;
    const TdiUnit *units;   // sorted by channel, index order kept within a channel
    const int *chain_rows;  // orbit-table rows, -1 for the undelayed sentinel
    int ch_begin[4];        // units [ch_begin[c], ch_begin[c+1]) feed channel c
    int halo_lo, halo_hi;
    int win_cap;            // LDS window capacity per link row (doubles)
    int overwrite;
    int aet;
    int zero_edges;
    double s2, s6, s3;      // sqrt(2), sqrt(6), sqrt(3)
    int *err;
    int tile;               // fused launches: output samples per tile
    double *lti_fir = nullptr;  // mixed fused launches: per source, the two composite TDI filters
                                // ([2][kFir] doubles, written by lti_filter_kernel before the launch)
    int lti_uniform = 0;        // mixed fused launches: the host proved that every tile of every
                                // source clears the composite-filter test as a whole, so the
                                // kernel without the per-sample pass runs
    // Inner-product launches of mixed fused batches (inner_part set): instead of storing the
    // outputs, every tile writes its per-channel sums against the data series (same layout as
    // the outputs, at inner_data + (src.out - inner_out)) to inner_part [tile][3][kInnerSums].
    double *inner_part = nullptr;
    const char *inner_out = nullptr;
    const char *inner_data = nullptr;
    int utv = 1;  // mixed stock tables, unequal arms: 1 = the slowly-varying-arm filter form,
                  // 0 = the per-sample kernel (utv_suitable)
    int orbit_stage = 0;  // mixed matrix-core kernels: 1 = every tile's orbit nodes fit the LDS
                          // stage (orbit_stage_fits), so the staged variants run
};

// Sums of one channel over a tile in an inner-product launch: d h, h h, h and (-1)^(i - T) h
// (h the output before float32 rounding, d the data series).
constexpr int kInnerSums = 4;

struct TdiSource {
    double *out;            // [3][num_delays]
    const double *y;        // [6][num_inputs]
    const double *t_arr;    // [num_delays] or nullptr
    int num_inputs;
    int num_delays;
    int start;              // tdi_start_ind
    int pad;
    double t0;
};

// Device row (orbit-table column / y row) of each link in the canonical order
// 12, 23, 31, 13, 32, 21.
struct RowMap {
    int row[6];
};

// Stock TDI dispatch: generation 1 or 2 and whether all six light travel times of the orbit
// table are bitwise equal (then units with equal chain lengths share one delay).
struct StockTdi {
    int generation;  // 0: not a stock table
    bool equal_arm;
    RowMap rows;
};

// A batch: sources[num_sources] and tile_start[num_sources + 1] (prefix sums of the tiles of
// each source) in device memory.
template <typename Source>
struct Batch {
    const Source *sources;
    const int *tile_start;
    int num_sources;
    int total_tiles;
};

// Tiles a source needs: one per `tile` samples of [start, n - start), or of [0, n) when the
// edges are zeroed too.
inline int tiles_for(int num_delays, int start, int zero_edges, int tile = kBlock)
{
    const int span = zero_edges ? num_delays : num_delays - 2 * start;
    return span > 0 ? (span + tile - 1) / tile : 0;
}

// ------------------------------------------------------------------ fused response
// Projection and TDI in one launch: a tile of output samples projects the six links over its
// TDI window into LDS and runs TDI from there, so y never goes to HBM.
constexpr int kFusedPasses = 2;  // default tile: kFusedPasses * kBlock output samples

struct FusedSource {
    const void *h;          // [num_pts] complex: double2 (FP64), float2 or int2 holding h / scale
    void *out;              // [3][num_pts] double (FP64) or float (mixed)
    int num_pts;
    int proj_start;         // projections_start_ind: y is zero outside [proj_start, n - proj_start)
    int tdi_start;          // tdi_start_ind
    int pad;
    double t0;
    double scale;           // outputs are scale * (response of the stored h)
    double k[3], u[3], v[3];
};

// Launchers (flr_kernels.hip).
hipError_t launch_projection(const ProjConfig &cfg, const ProjSource &src, const OrbitView &orb,
                             const LagrangeView &lg, bool exact, bool equal_arm,
                             hipStream_t stream);
hipError_t launch_projection_batch(const ProjConfig &cfg, const Batch<ProjSource> &batch,
                                   const OrbitView &orb, const LagrangeView &lg, bool exact,
                                   bool equal_arm, int grid_per_cu, hipStream_t stream);
hipError_t launch_tdi(const TdiConfig &cfg, const TdiSource &src, const OrbitView &orb,
                      const LagrangeView &lg, bool exact, const StockTdi &stock,
                      hipStream_t stream);
hipError_t launch_tdi_batch(const TdiConfig &cfg, const Batch<TdiSource> &batch,
                            const OrbitView &orb, const LagrangeView &lg, bool exact,
                            const StockTdi &stock, int grid_per_cu, hipStream_t stream);
// Input formats of the mixed fused kernels (h / scale per source): complex64, 32-bit fixed-point
// pairs, complex128.
constexpr int kInC64 = 0;
constexpr int kInQ31 = 1;
constexpr int kInC128 = 2;
// Fused batch: exact (FP64 only) or fast stencils; mixed = h in one of the kIn* formats, float
// out. Tiles of
// tile_passes * kBlock output samples (1, 2 or 4 for mixed, 2 for FP64); grid_per_cu
// persistent workgroups per CU.
hipError_t launch_fused_batch(const ProjConfig &pc, const TdiConfig &tc,
                              const Batch<FusedSource> &batch, const OrbitView &orb,
                              const LagrangeView &lg, bool exact, bool mixed, int input,
                              int tile_passes, int grid_per_cu, const StockTdi &stock,
                              hipStream_t stream);
hipError_t launch_fused_fp64(const ProjConfig &pc, const TdiConfig &tc,
                             const Batch<FusedSource> &batch, const OrbitView &orb,
                             const LagrangeView &lg, bool exact, int grid_per_cu,
                             const StockTdi &stock, hipStream_t stream);
hipError_t launch_fused_mixed(const ProjConfig &pc, const TdiConfig &tc,
                              const Batch<FusedSource> &batch, const OrbitView &orb,
                              const LagrangeView &lg, int tile_passes, int grid_per_cu,
                              const StockTdi &stock, hipStream_t stream);
hipError_t launch_fused_q31(const ProjConfig &pc, const TdiConfig &tc,
                            const Batch<FusedSource> &batch, const OrbitView &orb,
                            const LagrangeView &lg, int tile_passes, int grid_per_cu,
                            const StockTdi &stock, hipStream_t stream);
// The composite TDI filters of a mixed fused batch's sources (equal-arm stock tables only),
// into tc.lti_fir; they depend on the sources, orbits, Lagrange and TDI tables, not on h.
bool lti_filters_needed(const StockTdi &stock, const LagrangeView &lg);
hipError_t launch_lti_filters(const TdiConfig &tc, const Batch<FusedSource> &batch,
                              const OrbitView &orb, const LagrangeView &lg,
                              const StockTdi &stock, hipStream_t stream);
hipError_t launch_fused_c128(const ProjConfig &pc, const TdiConfig &tc,
                             const Batch<FusedSource> &batch, const OrbitView &orb,
                             const LagrangeView &lg, int tile_passes, int grid_per_cu,
                             const StockTdi &stock, hipStream_t stream);
// Dynamic LDS of a fused launch: h window, six window rows (padded, plus the matrix-core
// filters and per-sample flags, for the mixed equal-arm form), generic TDI scratch; utv: the
// mixed unequal-arm form (h window or filters, padded rows).
size_t fused_lds_bytes(const ProjConfig &pc, const TdiConfig &tc, bool generic, bool lti,
                       int passes, bool utv = false);
// Per-source totals of an inner-product launch's tile sums, in a fixed order:
// sums[s][3][kInnerSums] (flr_kernels.hip).
hipError_t launch_inner_reduce(const double *part, const int *tile_start, int num_sources,
                               double *sums, hipStream_t stream);
// Galactic-binary generator (flr_waveform.hip).
hipError_t launch_gb(const flr_gb_source *sources, int num_sources, int64_t total, double dt,
                     int format, void *out, hipStream_t stream);
hipError_t launch_orbit_eval(const OrbitView &orb, const double *t, int n, int row, int sc,
                             double *ltt_out, double *pos_out, hipStream_t stream);
hipError_t launch_clear(int *word, hipStream_t stream);

}  // namespace flr
