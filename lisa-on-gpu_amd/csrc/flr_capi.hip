// flr_capi.hip -- the extern "C" boundary of libflr_hip.so (declared in include/flr.h).
//
// Handles own their device tables: orbit tables are packed once at creation, Lagrange and TDI
// tables are uploaded once and re-uploaded only when a drop-in call passes different contents.
// No entry point allocates per call; the drop-in calls synchronise their stream at the end, as
// the reference launchers do (LISAResponse.cu:420,704), and report latched device errors.
#include "flr_internal.hpp"
#include "flr_tdi_stock.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace flr;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(FLR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define FLR_HIP(call)                                  \
    do {                                               \
        hipError_t _e = (call);                        \
        if (_e != hipSuccess) return hip_fail(_e, #call); \
    } while (0)

// Copy n elements from a host or device pointer into a host vector.
template <typename T>
int fetch(const T *src, size_t n, std::vector<T> &dst, const char *name)
{
    if (n > 0 && src == nullptr) return fail(FLR_ERR_INVALID_ARGUMENT, std::string(name) + " is NULL");
    dst.resize(n);
    if (n == 0) return FLR_OK;
    hipError_t e = hipMemcpy(dst.data(), src, n * sizeof(T), hipMemcpyDefault);
    if (e != hipSuccess) return hip_fail(e, name);
    return FLR_OK;
}

// Keep the caller's current device unchanged across a call.
struct DeviceGuard {
    int prev = -1;
    bool ok = true;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

const int kStdLinks[kNLinks] = {12, 23, 31, 13, 32, 21};

// Whether `stream` is capturing into a graph (events recorded there are graph-internal).
int stream_capturing(hipStream_t stream, bool &capturing)
{
    hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
    FLR_HIP(hipStreamIsCapturing(stream, &status));
    capturing = status != hipStreamCaptureStatusNone;
    return FLR_OK;
}

}  // namespace

struct flr_orbits {
    double dt = 0.0;
    int N = 0;
    double armlength = 0.0;
    int links[kNLinks] = {};
    int sc_r[kNLinks] = {};
    int sc_e[kNLinks] = {};
    double *d_table = nullptr;
    double max_radius = 0.0;
    double max_ltt = 0.0;
    bool equal_arm = false;  // all six light-travel-time columns bitwise equal at every node
    bool constant_arm = false;  // ... and equal at every node (zero slopes): ltt0 everywhere
    double ltt0 = 0.0;
    double max_rate = 0.0;  // largest |slope| of a light travel time over the table (s/s)
    double max_kink = 0.0;  // largest |change of that slope| at a node (s/s)
    int device = 0;

    int row_of(int link) const
    {
        for (int r = 0; r < kNLinks; ++r)
            if (links[r] == link) return r;
        return -1;
    }
    OrbitView view() const { return OrbitView{d_table, dt, N, 1.0 / dt, max_ltt}; }
};

struct flr_response {
    flr_orbits *orbits = nullptr;
    uint64_t version = 0;  // bumped when orbits, Lagrange or TDI tables change
    int device = 0;
    int *d_err = nullptr;
    bool exact_default = true;

    // Lagrange tables
    bool lag_set = false;
    int order = 0;
    int num_A = 0;
    double deps = 0.0;
    std::vector<double> A_host, E_host;
    double2 *d_Apair = nullptr;
    double2 *d_Aslope = nullptr;
    int A_cap = 0;

    // TDI tables
    bool tdi_set = false;
    std::vector<int> sig_starts, sig_lengths, sig_base, sig_comb, sig_chan;
    std::vector<double> sig_signs;
    TdiUnit *d_units = nullptr;
    int *d_chain = nullptr;
    int ch_begin[4] = {0, 0, 0, 0};
    int num_units = 0;
    int max_chain = 0;  // most delayed links in one unit
    int stock_gen = 0;  // 1 or 2 when the table is the stock 1st/2nd-generation one

    LagrangeView lagrange() const
    {
        LagrangeView lg{};
        lg.Apair = d_Apair;
        lg.Aslope = d_Aslope;
        lg.deps = deps;
        lg.inv_deps = 1.0 / deps;
        lg.num_A = num_A;
        lg.h = (order + 1) / 2;
        for (int j = 0; j < kMaxHalf; ++j) lg.E[j] = j < (int)E_host.size() ? E_host[j] : 0.0;
        return lg;
    }
};

struct flr_batch {
    flr_response *r = nullptr;
    int flags = 0;
    double dt = 0.0, fs = 0.0;
    int num_sources = 0;
    int64_t samples = 0;
    ProjSource *d_proj = nullptr;
    TdiSource *d_tdi = nullptr;
    int *d_proj_tiles = nullptr;  // [num_sources + 1] prefix sums
    int *d_tdi_tiles = nullptr;
    int proj_tiles = 0, tdi_tiles = 0;
    bool fused = false;           // one fused launch over d_fused / d_fused_tiles
    FusedSource *d_fused = nullptr;
    int *d_fused_tiles = nullptr;
    int fused_tiles = 0;
    int capacity = 0;             // fused: sources the tables hold
    int tile_passes = kFusedPasses;  // fused: the tile's window is tile_passes * kBlock samples
    int tile_len = 0;                // fused: output samples per tile (make_fused_config)
    int grid_per_cu = 32;            // persistent workgroups per CU (the autotuner's usual pick)
    FusedSource *h_fused = nullptr;  // pinned staging of the source table and tile prefix
    int *h_fused_tiles = nullptr;
    hipEvent_t uploaded = nullptr;   // the last bind's upload
    hipEvent_t ran = nullptr;        // the last run's launch: a bind on any stream waits for it
    bool ran_recorded = false;       // `ran` has been recorded at least once
    double *d_lti = nullptr;         // mixed fused: [capacity][kLtiRecord] composite TDI filters
    uint64_t lti_version = ~0ull;    // response version the filters were built for (none yet)
    bool lti_uniform = false;        // every bound source clears the composite-filter tile test
    uint64_t uniform_version = ~0ull;  // response version and tile length lti_uniform holds for
    int uniform_tile = 0;
    const void *out_base = nullptr;  // fused: the bound out_all (inner-product launches)
    double *d_inner = nullptr;       // fused inner products: per tile [3][kInnerSums] sums
    int64_t inner_cap = 0;           // tiles d_inner holds
};

namespace {

int set_lagrange_impl(flr_response *r, int order, const double *A_in, double deps, int num_A,
                      const double *E_in)
{
    const int h = (order + 1) / 2;
    if (order < 1 || h > kMaxHalf)
        return fail(FLR_ERR_UNSUPPORTED, "order must be in [1, " + std::to_string(2 * kMaxHalf - 1) + "]");
    if (num_A < 2) return fail(FLR_ERR_INVALID_ARGUMENT, "num_A must be >= 2");
    if (!(deps > 0.0)) return fail(FLR_ERR_INVALID_ARGUMENT, "deps must be > 0");
    std::vector<double> A, E;
    int st = fetch(A_in, (size_t)num_A, A, "A_in");
    if (st) return st;
    st = fetch(E_in, (size_t)(h > 1 ? h - 1 : 0), E, "E_in");
    if (st) return st;
    if (r->lag_set && r->order == order && r->num_A == num_A && r->deps == deps && r->A_host == A &&
        r->E_host == E)
        return FLR_OK;
    if (num_A - 1 > r->A_cap) {
        if (r->d_Apair) (void)hipFree(r->d_Apair);
        if (r->d_Aslope) (void)hipFree(r->d_Aslope);
        r->d_Apair = r->d_Aslope = nullptr;
        r->A_cap = 0;
        FLR_HIP(hipMalloc(&r->d_Apair, sizeof(double2) * (num_A - 1)));
        FLR_HIP(hipMalloc(&r->d_Aslope, sizeof(double2) * (num_A - 1)));
        r->A_cap = num_A - 1;
    }
    std::vector<double2> pair(num_A - 1), slope(num_A - 1);
    for (int k = 0; k + 1 < num_A; ++k) {
        pair[k] = make_double2(A[k], A[k + 1]);
        slope[k] = make_double2(A[k], A[k + 1] - A[k]);
    }
    FLR_HIP(hipMemcpy(r->d_Apair, pair.data(), sizeof(double2) * pair.size(), hipMemcpyHostToDevice));
    FLR_HIP(hipMemcpy(r->d_Aslope, slope.data(), sizeof(double2) * slope.size(), hipMemcpyHostToDevice));
    r->order = order;
    r->num_A = num_A;
    r->deps = deps;
    r->A_host = std::move(A);
    r->E_host = std::move(E);
    r->lag_set = true;
    ++r->version;
    return FLR_OK;
}

// True when the units, in channel-then-index order, are exactly the stock table of generation
// `gen` (same links, signs, chains and channels), which the stock kernels hard-code.
bool is_stock_table(int gen, const std::vector<int> &order_idx, const std::vector<int> &starts,
                    const std::vector<int> &lengths, const std::vector<int> &base,
                    const std::vector<int> &comb, const std::vector<double> &signs,
                    const std::vector<int> &chan)
{
    const int nx = gen == 1 ? 8 : 16;
    if ((int)order_idx.size() != 3 * nx) return false;
    int k = 0;
    for (int c = 0; c < 3; ++c)
        for (int x = 0; x < nx; ++x, ++k) {
            const int u = order_idx[k];
            const stock::Term &tm = stock::kX[x];
            if (chan[u] != c || base[u] != stock::cyc(tm.link, c) || signs[u] != (double)tm.sign)
                return false;
            if (tm.len == 0) {
                if (lengths[u] != 1 || comb[starts[u]] != -11) return false;
            } else {
                if (lengths[u] != tm.len) return false;
                for (int s = 0; s < tm.len; ++s)
                    if (comb[starts[u] + s] != stock::cyc(tm.chain[s], c)) return false;
            }
        }
    return true;
}

int set_tdi_impl(flr_response *r, const int32_t *unit_starts, const int32_t *unit_lengths,
                 const int32_t *tdi_base_link, const int32_t *tdi_link_combinations,
                 const double *tdi_signs_in, const int32_t *channels, int num_units, int num_channels)
{
    if (num_units < 1 || num_units > kMaxUnits)
        return fail(FLR_ERR_UNSUPPORTED, "num_units must be in [1, " + std::to_string(kMaxUnits) + "]");
    if (num_channels < 1 || num_channels > 3)
        return fail(FLR_ERR_INVALID_ARGUMENT, "num_channels must be in [1, 3]");
    std::vector<int> starts, lengths, base, chan, comb;
    std::vector<double> signs;
    int st;
    if ((st = fetch(unit_starts, num_units, starts, "unit_starts"))) return st;
    if ((st = fetch(unit_lengths, num_units, lengths, "unit_lengths"))) return st;
    if ((st = fetch(tdi_base_link, num_units, base, "tdi_base_link"))) return st;
    if ((st = fetch(channels, num_units, chan, "channels"))) return st;
    if ((st = fetch(tdi_signs_in, num_units, signs, "tdi_signs_in"))) return st;
    int ncomb = 0;
    for (int u = 0; u < num_units; ++u) {
        if (starts[u] < 0 || lengths[u] < 0)
            return fail(FLR_ERR_INVALID_ARGUMENT, "unit_starts/unit_lengths must be >= 0");
        ncomb = std::max(ncomb, starts[u] + lengths[u]);
    }
    if (ncomb > kMaxChain)
        return fail(FLR_ERR_UNSUPPORTED, "more than " + std::to_string(kMaxChain) + " delay-chain entries");
    if ((st = fetch(tdi_link_combinations, ncomb, comb, "tdi_link_combinations"))) return st;

    if (r->tdi_set && starts == r->sig_starts && lengths == r->sig_lengths && base == r->sig_base &&
        chan == r->sig_chan && comb == r->sig_comb && signs == r->sig_signs)
        return FLR_OK;

    // Units of a channel keep their index order (the order of the reference's `+=`); channels
    // are laid out one after the other.
    std::vector<int> order_idx(num_units);
    for (int u = 0; u < num_units; ++u) order_idx[u] = u;
    for (int u = 0; u < num_units; ++u)
        if (chan[u] < 0 || chan[u] >= num_channels || chan[u] > 2)
            return fail(FLR_ERR_INVALID_ARGUMENT, "channel index out of range");
    std::stable_sort(order_idx.begin(), order_idx.end(), [&](int a, int b) { return chan[a] < chan[b]; });

    std::vector<TdiUnit> units;
    std::vector<int> rows;
    int max_chain = 0;
    for (int u : order_idx) {
        TdiUnit un{};
        un.sign = signs[u];
        un.base_row = r->orbits->row_of(base[u]);
        if (un.base_row < 0)
            return fail(FLR_ERR_INVALID_ARGUMENT, "unknown TDI base link " + std::to_string(base[u]));
        un.chain_start = (int)rows.size();
        un.chain_len = lengths[u];
        int delayed = 0;
        for (int s = 0; s < lengths[u]; ++s) {
            const int l = comb[starts[u] + s];
            if (l == -11) {
                rows.push_back(-1);
                continue;
            }
            const int row = r->orbits->row_of(l);
            if (row < 0) return fail(FLR_ERR_INVALID_ARGUMENT, "unknown TDI delay link " + std::to_string(l));
            rows.push_back(row);
            ++delayed;
        }
        max_chain = std::max(max_chain, delayed);
        units.push_back(un);
    }
    if (rows.size() > (size_t)kMaxChain)
        return fail(FLR_ERR_UNSUPPORTED, "more than " + std::to_string(kMaxChain) + " delay-chain entries");
    int ch_begin[4] = {0, 0, 0, 0};
    for (int c = 0; c < 3; ++c) {
        int cnt = 0;
        for (int u = 0; u < num_units; ++u) cnt += chan[u] == c;
        ch_begin[c + 1] = ch_begin[c] + cnt;
    }
    if (!r->d_units) FLR_HIP(hipMalloc(&r->d_units, sizeof(TdiUnit) * kMaxUnits));
    if (!r->d_chain) FLR_HIP(hipMalloc(&r->d_chain, sizeof(int) * kMaxChain));
    FLR_HIP(hipMemcpy(r->d_units, units.data(), sizeof(TdiUnit) * units.size(), hipMemcpyHostToDevice));
    if (!rows.empty())
        FLR_HIP(hipMemcpy(r->d_chain, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice));
    std::memcpy(r->ch_begin, ch_begin, sizeof(ch_begin));
    r->num_units = num_units;
    r->max_chain = max_chain;
    r->stock_gen = 0;
    for (int gen = 1; gen <= 2 && r->stock_gen == 0; ++gen)
        if (is_stock_table(gen, order_idx, starts, lengths, base, comb, signs, chan)) r->stock_gen = gen;
    r->sig_starts = std::move(starts);
    r->sig_lengths = std::move(lengths);
    r->sig_base = std::move(base);
    r->sig_chan = std::move(chan);
    r->sig_comb = std::move(comb);
    r->sig_signs = std::move(signs);
    r->tdi_set = true;
    ++r->version;
    return FLR_OK;
}

constexpr size_t kMaxLds = 160 * 1024;

// Launch configuration shared by all sources of a projection launch: receiver/emitter rows,
// and halos sized from the orbit extent (max |x|/c and max light travel time).
int make_proj_config(const flr_response *r, double dt, double fs, int flags, ProjConfig &cfg)
{
    const flr_orbits *o = r->orbits;
    const int h = (r->order + 1) / 2;
    cfg = ProjConfig{};
    cfg.dt = dt;
    cfg.fs = fs;
    cfg.zero_edges = (flags & FLR_FLAG_ZERO_EDGES) ? 1 : 0;
    for (int s = 0; s < 3; ++s)
        for (int e = 0; e < 3; ++e) cfg.link_row[s][e] = -1;
    for (int row = 0; row < kNLinks; ++row) {
        const int s = o->sc_r[row] - 1, e = o->sc_e[row] - 1;
        if (s < 0 || s > 2 || e < 0 || e > 2 || s == e || cfg.link_row[s][e] >= 0)
            return fail(FLR_ERR_INVALID_ARGUMENT, "links must be the six ordered spacecraft pairs");
        cfg.link_row[s][e] = row;
    }
    const double rc = o->max_radius * kCInv;
    cfg.halo_lo = (int)std::ceil((o->max_ltt + rc) * fs) + h + 4;
    cfg.halo_hi = (int)std::ceil(rc * fs) + h + 4;
    cfg.win_cap = kBlock + cfg.halo_lo + cfg.halo_hi;
    if ((size_t)cfg.win_cap * sizeof(double2) > kMaxLds)
        return fail(FLR_ERR_UNSUPPORTED, "projection delay window exceeds LDS (sampling too fine)");
    cfg.err = r->d_err;
    return FLR_OK;
}

// Launch configuration of a TDI launch: the resident unit table, halos from the longest delay
// chain, output mode, and which kernel family applies.
int make_tdi_config(const flr_response *r, double dt, double fs, int flags, TdiConfig &cfg,
                    StockTdi &stock)
{
    const flr_orbits *o = r->orbits;
    const int h = (r->order + 1) / 2;
    cfg = TdiConfig{};
    cfg.dt = dt;
    cfg.fs = fs;
    cfg.units = r->d_units;
    cfg.chain_rows = r->d_chain;
    for (int c = 0; c < 4; ++c) cfg.ch_begin[c] = r->ch_begin[c];
    cfg.halo_lo = (int)std::ceil(r->max_chain * o->max_ltt * fs) + h + 4;
    cfg.halo_hi = h + 4;
    cfg.win_cap = kBlock + cfg.halo_lo + cfg.halo_hi;
    if (((size_t)kNLinks * cfg.win_cap + (size_t)kNLinks * kBlock) * sizeof(double) > kMaxLds)
        return fail(FLR_ERR_UNSUPPORTED, "TDI delay window exceeds LDS (sampling too fine)");
    cfg.zero_edges = (flags & FLR_FLAG_ZERO_EDGES) ? 1 : 0;
    cfg.aet = (flags & FLR_FLAG_AET) ? 1 : 0;
    cfg.overwrite = ((flags & FLR_FLAG_OVERWRITE) || cfg.aet) ? 1 : 0;
    cfg.s2 = std::sqrt(2.0);
    cfg.s6 = std::sqrt(6.0);
    cfg.s3 = std::sqrt(3.0);
    cfg.err = r->d_err;
    stock = StockTdi{};
    stock.generation = r->stock_gen;
    stock.equal_arm = o->equal_arm;
    const int canon_links[6] = {12, 23, 31, 13, 32, 21};
    for (int k = 0; k < 6; ++k) stock.rows.row[k] = o->row_of(canon_links[k]);
    return FLR_OK;
}

// The fused fast kernels take the order-25 E weights as compile-time constants (e25): true when
// the response's uploaded table holds exactly those values.
bool stock_e25(const flr_response *r)
{
    if (r->order != 25 || r->E_host.size() < 12) return false;
    for (int j = 0; j < 12; ++j)
        if (r->E_host[j] != e25(j)) return false;
    return true;
}

// Host replica of the composite-filter path's decisions for one source (flr_fused.hpp:
// lti_source_reference, lti_reference, lti_tile_matches), in the same FP64 operations: true when
// the source's filters fit and every one of its tiles clears the tile test as a whole. The test
// of a tile is monotone in its last sample and the orbit nodes it covers, so the whole TDI range
// [T, n - T - 1] decides for every tile. Needs constant, equal arms (the only case where the
// composite filters are used throughout).
// Whether the orbit nodes of every tile of a mixed matrix-core launch fit the kernels' LDS stage
// (stage_orbit_nodes): a tile projects at most win_cap samples from t_first on, and reads nodes
// of their times and of their emission times down to t_first - max_ltt; the stage starts one node
// below that, and float rounding of t/dt can move either end by one node, so the tile needs at
// most floor(span / dt_orbit) + 5 of them.
bool orbit_stage_fits(const flr_orbits *o, int win_cap, double dt)
{
    const double span = (double)(win_cap - 1) * dt + o->max_ltt;
    return std::floor(span / o->dt) + 5.0 <= (double)kOrbitStage;
}

// Whether the slowly-varying-arm form (fused_utv_kernel) suits these orbits at tiles of `tile`
// samples: from the table's largest arm rate and largest slope change at a node, with every
// unit taken at the worst case, the expected share of tiles that fail utv_unit's tests -- the
// chord test at a node (3 x the mid-tile deviation of a kink of the chain's summed slope
// changes), the drift bound, and a crossing of one of the A table's nodes (probability per unit
// ~ the tile's range of e over the table spacing) -- is below one half. Otherwise the per-sample
// kernel runs, which is faster than the filter form's per-sample path on its block-interleaved
// rows. Only the speed depends on this choice.
bool utv_suitable(const flr_orbits *o, int nx, int tile, double dt, double fs, double deps)
{
    const int M = nx == 8 ? 3 : 7;  // longest chain
    const double per_sample = dt * fs;
    const double sigma = M * o->max_rate * per_sample;   // position drift per sample
    const double half = 0.5 * tile;
    const double drift = sigma * half;
    if (drift > kUtvDrift) return false;
    const double kink = M * o->max_kink * per_sample;    // change of that drift at a node
    if (3.0 * kink * tile / 4.0 > kLtiTol) return false;
    const double p = std::min(1.0, 2.0 * drift / deps);  // per unit
    const double fail = 1.0 - std::pow(1.0 - p, 3.0 * nx);
    return fail < 0.5;
}

bool lti_source_uniform(const flr_orbits *o, const FusedSource &f, double dt, double fs, int nx,
                        int halo_lo, int halo_hi)
{
    if (!o->constant_arm) return false;
    const int n = f.num_pts, T = f.tdi_start;
    if (n - 2 * T <= 0) return true;  // no sample reaches the TDI stage
    const int M = nx == 8 ? 3 : 7;    // longest chain
    const int H = 13;
    const double L_ref = o->ltt0;
    // lti_reference at the source's first TDI sample
    const double t_ref = f.t0 + (double)T * dt;
    double Dm = t_ref;
    int lo = 1 << 30, hi = -(1 << 30);
    for (int m = 0; m <= M; ++m) {
        if (m > 0) Dm -= L_ref;
        const double clipped = Dm - f.t0;
        const double cf = clipped * fs;
        const int d = (int)std::ceil(cf) - 1;
        lo = std::min(lo, d - T);
        hi = std::max(hi, d - T);
    }
    const int kmin = lo - (H - 1);
    if (!((hi + H) - kmin + 1 <= kFir)) return false;
    // lti_tile_matches over the whole range
    if (kmin + halo_lo < 0 || kmin + kFir > halo_hi + 1) return false;
    const int last = n - T - 1;
    auto ulp = [](double x) { return std::ldexp(1.0, std::ilogb(x) - 52); };
    const double tmax = 1.001 * (2.0 * std::fabs(f.t0) + (double)last * dt + M * std::fabs(L_ref));
    const double per_sample = (0.5 * M + 1.5) * ulp(tmax) * fs + ulp(tmax * fs);
    const double drift = 1.01 * std::fabs(std::fma(dt, fs, -1.0)) * (double)(last - T);
    return 2.0 * per_sample + drift + 1e-13 <= kLtiTol;
}

// Fused launch: the TDI window of a tile of passes * kBlock samples, and the h window that
// projects it.
int make_fused_config(const flr_response *r, double dt, double fs, int flags, int passes,
                      ProjConfig &pc, TdiConfig &tc, StockTdi &stock)
{
    if (!(flags & FLR_FLAG_EXACT) && !stock_e25(r))
        return fail(FLR_ERR_UNSUPPORTED,
                    "the fast fused kernels are built for the order-25 Lagrange E table of "
                    "pyResponseTDI._fill_A_E and the uploaded table differs (use FLR_FLAG_EXACT "
                    "or the unfused entry points)");
    int st = make_proj_config(r, dt, fs, flags, pc);
    if (st) return st;
    if ((st = make_tdi_config(r, dt, fs, flags | FLR_FLAG_OVERWRITE, tc, stock))) return st;
    // mixed precision: room for the composite filter's kFir taps after the last sample
    if (flags & FLR_FLAG_MIXED) tc.halo_hi = std::max(tc.halo_hi, kFir - (r->order + 1) / 2 + 1);
    // tile + halos = passes * kBlock: every projection round uses the whole workgroup
    const int halos = tc.halo_lo + tc.halo_hi;
    tc.tile = passes * kBlock - halos >= (passes * kBlock) / 2 ? passes * kBlock - halos
                                                               : passes * kBlock;
    tc.win_cap = tc.tile + halos;
    pc.win_cap = tc.win_cap + pc.halo_lo + pc.halo_hi;
    const bool generic = stock.generation == 0 || r->order != 25;
    const bool lti = (flags & FLR_FLAG_MIXED) && !generic && stock.equal_arm;
    const bool utv = (flags & FLR_FLAG_MIXED) && !generic && !stock.equal_arm;
    if (fused_lds_bytes(pc, tc, generic, lti, passes, utv) > kMaxLds)
        return fail(FLR_ERR_UNSUPPORTED, "fused delay windows exceed LDS (sampling too fine)");
    return FLR_OK;
}

int projection_impl(flr_response *r, double *y_gw, const double *input_in, int num_delays,
                    int num_inputs, const double *t_data, double t0, double dt, double fs,
                    const double *k, const double *u, const double *v, int start, int flags,
                    hipStream_t stream)
{
    if (!r->lag_set) return fail(FLR_ERR_INVALID_ARGUMENT, "Lagrange tables not set");
    if (!y_gw || !input_in) return fail(FLR_ERR_INVALID_ARGUMENT, "y_gw and input_in must not be NULL");
    if (!(flags & FLR_FLAG_T_UNIFORM) && !t_data)
        return fail(FLR_ERR_INVALID_ARGUMENT, "t_data is NULL without FLR_FLAG_T_UNIFORM");
    if (num_delays < 1 || num_inputs < 1) return fail(FLR_ERR_INVALID_ARGUMENT, "empty input");
    if (start < 0) return fail(FLR_ERR_INVALID_ARGUMENT, "projections_start_ind must be >= 0");
    if (!(fs > 0.0)) return fail(FLR_ERR_INVALID_ARGUMENT, "sampling_frequency must be > 0");
    ProjConfig cfg;
    int st = make_proj_config(r, dt, fs, flags, cfg);
    if (st) return st;
    ProjSource src{};
    src.y = y_gw;
    src.h = reinterpret_cast<const double2 *>(input_in);
    src.t_data = (flags & FLR_FLAG_T_UNIFORM) ? nullptr : t_data;
    src.num_delays = num_delays;
    src.num_inputs = num_inputs;
    src.start = start;
    src.t0 = t0;
    for (int c = 0; c < 3; ++c) {
        src.k[c] = k[c];
        src.u[c] = u[c];
        src.v[c] = v[c];
    }
    hipError_t e = launch_projection(cfg, src, r->orbits->view(), r->lagrange(),
                                     (flags & FLR_FLAG_EXACT) != 0, r->orbits->equal_arm, stream);
    if (e != hipSuccess) return hip_fail(e, "projection launch");
    return FLR_OK;
}

int tdi_impl(flr_response *r, double *out, const double *y_gw, int num_inputs, int num_delays,
             const double *t_arr, double t0, double dt, double fs, int start, int flags,
             hipStream_t stream)
{
    if (!r->lag_set) return fail(FLR_ERR_INVALID_ARGUMENT, "Lagrange tables not set");
    if (!r->tdi_set) return fail(FLR_ERR_INVALID_ARGUMENT, "TDI tables not set");
    if (!out || !y_gw) return fail(FLR_ERR_INVALID_ARGUMENT, "output and y_gw must not be NULL");
    if (!(flags & FLR_FLAG_T_UNIFORM) && !t_arr)
        return fail(FLR_ERR_INVALID_ARGUMENT, "t_arr is NULL without FLR_FLAG_T_UNIFORM");
    if (num_delays < 1 || num_inputs < 1) return fail(FLR_ERR_INVALID_ARGUMENT, "empty input");
    if (start < 0) return fail(FLR_ERR_INVALID_ARGUMENT, "tdi_start_ind must be >= 0");
    if (!(fs > 0.0)) return fail(FLR_ERR_INVALID_ARGUMENT, "sampling_frequency must be > 0");
    TdiConfig cfg;
    StockTdi stock;
    int st = make_tdi_config(r, dt, fs, flags, cfg, stock);
    if (st) return st;
    TdiSource src{};
    src.out = out;
    src.y = y_gw;
    src.t_arr = (flags & FLR_FLAG_T_UNIFORM) ? nullptr : t_arr;
    src.num_inputs = num_inputs;
    src.num_delays = num_delays;
    src.start = start;
    src.t0 = t0;
    hipError_t e = launch_tdi(cfg, src, r->orbits->view(), r->lagrange(),
                              (flags & FLR_FLAG_EXACT) != 0, stock, stream);
    if (e != hipSuccess) return hip_fail(e, "tdi launch");
    return FLR_OK;
}

int check_errors_impl(flr_response *r, hipStream_t stream)
{
    FLR_HIP(hipStreamSynchronize(stream));
    int word = 0;
    FLR_HIP(hipMemcpy(&word, r->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (word != 0) {
        FLR_HIP(launch_clear(r->d_err, stream));
        FLR_HIP(hipStreamSynchronize(stream));
    }
    if (word & kErrWindow)
        return fail(FLR_ERR_WINDOW,
                    "a stencil reached outside the staged sample window: the start index is too "
                    "small for the delays, or the time grid is not uniform");
    if (word & kErrLti)
        return fail(FLR_ERR_HIP, "internal: a composite-filter tile failed the test the host had "
                                 "proved it passes (its outputs are NaN)");
    return FLR_OK;
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

const char *flr_last_error(void) { return g_last_error.c_str(); }

const char *flr_version(void) { return "lisa-on-gpu_amd 0.1.0 (gfx950)"; }

int flr_device_count(int *count)
{
    if (!count) return fail(FLR_ERR_INVALID_ARGUMENT, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return FLR_OK;
}

int flr_orbits_create(double dt, int N, const double *n_arr, const double *ltt_arr,
                      const double *x_arr, const int32_t *links, const int32_t *sc_r,
                      const int32_t *sc_e, double armlength, flr_orbits **out)
{
    if (!out) return fail(FLR_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!(dt > 0.0)) return fail(FLR_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (N < 2) return fail(FLR_ERR_INVALID_ARGUMENT, "N must be >= 2");
    std::vector<double> n_h, ltt_h, x_h;
    std::vector<int> links_h, scr_h, sce_h;
    int st;
    if ((st = fetch(n_arr, (size_t)N * 18, n_h, "n_arr"))) return st;
    if ((st = fetch(ltt_arr, (size_t)N * 6, ltt_h, "ltt_arr"))) return st;
    if ((st = fetch(x_arr, (size_t)N * 9, x_h, "x_arr"))) return st;
    if ((st = fetch(links, 6, links_h, "links"))) return st;
    if ((st = fetch(sc_r, 6, scr_h, "sc_r"))) return st;
    if ((st = fetch(sc_e, 6, sce_h, "sc_e"))) return st;

    auto o = new flr_orbits();
    o->dt = dt;
    o->N = N;
    o->armlength = armlength;
    bool seen[kNLinks] = {};
    for (int r = 0; r < kNLinks; ++r) {
        o->links[r] = links_h[r];
        o->sc_r[r] = scr_h[r];
        o->sc_e[r] = sce_h[r];
        int which = -1;
        for (int q = 0; q < kNLinks; ++q)
            if (kStdLinks[q] == links_h[r]) which = q;
        if (which < 0 || seen[which] || scr_h[r] != links_h[r] / 10 || sce_h[r] != links_h[r] % 10) {
            delete o;
            return fail(FLR_ERR_INVALID_ARGUMENT,
                        "links must be a permutation of 12,23,31,13,32,21 with sc_r = link/10 and "
                        "sc_e = link%10");
        }
        seen[which] = true;
    }
    std::vector<double> table((size_t)N * kOrbitStride, 0.0);
    double rmax = 0.0, lmax = 0.0;
    for (int w = 0; w < N; ++w) {
        double *rec = &table[(size_t)w * kOrbitStride];
        for (int r = 0; r < kNLinks; ++r) {
            rec[r] = ltt_h[(size_t)w * 6 + r];
            lmax = std::max(lmax, std::fabs(rec[r]));
        }
        for (int s = 0; s < 3; ++s) {
            double r2 = 0.0;
            for (int c = 0; c < 3; ++c) {
                const double xv = x_h[((size_t)w * 3 + s) * 3 + c];
                rec[6 + 3 * s + c] = xv;
                r2 += xv * xv;
            }
            rmax = std::max(rmax, std::sqrt(r2));
        }
    }
    for (int w = 0; w + 1 < N; ++w) {
        double *rec = &table[(size_t)w * kOrbitStride];
        const double *nxt = &table[(size_t)(w + 1) * kOrbitStride];
        for (int v = 0; v < kOrbitVals; ++v) rec[kOrbitVals + v] = (nxt[v] - rec[v]) / dt;
    }
    if (!std::isfinite(rmax) || !std::isfinite(lmax)) {
        delete o;
        return fail(FLR_ERR_INVALID_ARGUMENT, "orbit tables hold non-finite values");
    }
    o->max_radius = rmax;
    o->max_ltt = lmax;
    o->equal_arm = true;
    for (int w = 0; w < N && o->equal_arm; ++w)
        for (int r = 1; r < kNLinks; ++r)
            if (table[(size_t)w * kOrbitStride + r] != table[(size_t)w * kOrbitStride]) o->equal_arm = false;
    for (int w = 0; w + 1 < N; ++w)
        for (int r = 0; r < kNLinks; ++r) {
            const double sl = table[(size_t)w * kOrbitStride + kOrbitVals + r];
            o->max_rate = std::max(o->max_rate, std::fabs(sl));
            if (w > 0)
                o->max_kink = std::max(
                    o->max_kink, std::fabs(sl - table[(size_t)(w - 1) * kOrbitStride + kOrbitVals + r]));
        }
    o->ltt0 = table[0];
    o->constant_arm = o->equal_arm;
    for (int w = 1; w < N && o->constant_arm; ++w)
        if (table[(size_t)w * kOrbitStride] != o->ltt0) o->constant_arm = false;
    hipError_t e = hipGetDevice(&o->device);
    if (e == hipSuccess) e = hipMalloc(&o->d_table, table.size() * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(o->d_table, table.data(), table.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (o->d_table) (void)hipFree(o->d_table);
        delete o;
        return hip_fail(e, "orbit table upload");
    }
    *out = o;
    return FLR_OK;
}

void flr_orbits_destroy(flr_orbits *orbits)
{
    if (!orbits) return;
    DeviceGuard g(orbits->device);
    if (orbits->d_table) (void)hipFree(orbits->d_table);
    delete orbits;
}

int flr_orbits_extent(const flr_orbits *orbits, double *max_radius, double *max_ltt)
{
    if (!orbits) return fail(FLR_ERR_NO_ORBITS, "orbits is NULL");
    if (max_radius) *max_radius = orbits->max_radius;
    if (max_ltt) *max_ltt = orbits->max_ltt;
    return FLR_OK;
}

int flr_orbits_eval(const flr_orbits *orbits, const double *t, int n, int link, int sc,
                    double *ltt_out, double *pos_out, void *stream)
{
    if (!orbits) return fail(FLR_ERR_NO_ORBITS, "orbits is NULL");
    if (n < 0 || (n > 0 && !t)) return fail(FLR_ERR_INVALID_ARGUMENT, "bad t array");
    const int row = orbits->row_of(link);
    if (ltt_out && row < 0) return fail(FLR_ERR_INVALID_ARGUMENT, "unknown link " + std::to_string(link));
    if (pos_out && (sc < 1 || sc > 3)) return fail(FLR_ERR_INVALID_ARGUMENT, "sc must be 1, 2 or 3");
    DeviceGuard g(orbits->device);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = launch_orbit_eval(orbits->view(), t, n, row < 0 ? 0 : row, sc - 1, ltt_out, pos_out, s);
    if (e != hipSuccess) return hip_fail(e, "orbit eval launch");
    return FLR_OK;
}

int flr_response_create(flr_orbits *orbits, flr_response **out)
{
    if (!out) return fail(FLR_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!orbits) return fail(FLR_ERR_NO_ORBITS, "Must add orbits with add_orbit_information method.");
    DeviceGuard g(orbits->device);
    auto r = new flr_response();
    r->orbits = orbits;
    r->device = orbits->device;
    ++r->version;
    hipError_t e = hipMalloc(&r->d_err, sizeof(int));
    if (e == hipSuccess) {
        const int zero = 0;
        e = hipMemcpy(r->d_err, &zero, sizeof(int), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        if (r->d_err) (void)hipFree(r->d_err);
        delete r;
        return hip_fail(e, "response create");
    }
    *out = r;
    return FLR_OK;
}

void flr_response_destroy(flr_response *r)
{
    if (!r) return;
    DeviceGuard g(r->device);
    if (r->d_err) (void)hipFree(r->d_err);
    if (r->d_Apair) (void)hipFree(r->d_Apair);
    if (r->d_Aslope) (void)hipFree(r->d_Aslope);
    if (r->d_units) (void)hipFree(r->d_units);
    if (r->d_chain) (void)hipFree(r->d_chain);
    delete r;
}

int flr_set_lagrange(flr_response *r, int order, const double *A_in, double deps, int num_A,
                     const double *E_in)
{
    if (!r) return fail(FLR_ERR_INVALID_ARGUMENT, "response is NULL");
    DeviceGuard g(r->device);
    return set_lagrange_impl(r, order, A_in, deps, num_A, E_in);
}

int flr_set_tdi(flr_response *r, const int32_t *unit_starts, const int32_t *unit_lengths,
                const int32_t *tdi_base_link, const int32_t *tdi_link_combinations,
                const double *tdi_signs_in, const int32_t *channels, int num_units, int num_channels)
{
    if (!r) return fail(FLR_ERR_INVALID_ARGUMENT, "response is NULL");
    if (!r->orbits) return fail(FLR_ERR_NO_ORBITS, "Must add orbits with add_orbit_information method.");
    DeviceGuard g(r->device);
    return set_tdi_impl(r, unit_starts, unit_lengths, tdi_base_link, tdi_link_combinations,
                        tdi_signs_in, channels, num_units, num_channels);
}

int flr_projection(flr_response *r, double *y_gw, const double *input_in, int num_pts,
                   const double *t_data, double t0, double dt, double sampling_frequency,
                   const double *k, const double *u, const double *v, int projections_start_ind,
                   int flags, void *stream)
{
    if (!r) return fail(FLR_ERR_INVALID_ARGUMENT, "response is NULL");
    if (!r->orbits) return fail(FLR_ERR_NO_ORBITS, "Must add orbits with add_orbit_information method.");
    if (!k || !u || !v) return fail(FLR_ERR_INVALID_ARGUMENT, "k, u, v must not be NULL");
    DeviceGuard g(r->device);
    return projection_impl(r, y_gw, input_in, num_pts, num_pts, t_data, t0, dt, sampling_frequency,
                           k, u, v, projections_start_ind, flags, (hipStream_t)stream);
}

int flr_tdi(flr_response *r, double *out, const double *y_gw, int num_pts, const double *t_arr,
            double t0, double dt, double sampling_frequency, int tdi_start_ind, int flags,
            void *stream)
{
    if (!r) return fail(FLR_ERR_INVALID_ARGUMENT, "response is NULL");
    if (!r->orbits) return fail(FLR_ERR_NO_ORBITS, "Must add orbits with add_orbit_information method.");
    DeviceGuard g(r->device);
    return tdi_impl(r, out, y_gw, num_pts, num_pts, t_arr, t0, dt, sampling_frequency, tdi_start_ind,
                    flags, (hipStream_t)stream);
}

int flr_check_errors(flr_response *r, void *stream)
{
    if (!r) return fail(FLR_ERR_INVALID_ARGUMENT, "response is NULL");
    DeviceGuard g(r->device);
    return check_errors_impl(r, (hipStream_t)stream);
}

int flr_get_response(flr_response *r, double *y_gw, const double *t_data, const double *k_in,
                     const double *u_in, const double *v_in, double dt, int num_delays,
                     const double *input_in, int num_inputs, int order, double sampling_frequency,
                     int buffer_integer, const double *A_in, double deps, int num_A,
                     const double *E_in, int projections_start_ind, void *stream)
{
    (void)buffer_integer;
    if (!r || !r->orbits) return fail(FLR_ERR_NO_ORBITS, "Must add orbits with add_orbit_information method.");
    DeviceGuard g(r->device);
    hipStream_t s = (hipStream_t)stream;
    FLR_HIP(hipStreamSynchronize(s));  // the small tables may be produced on this stream
    std::vector<double> k, u, v;
    int st;
    if ((st = fetch(k_in, 3, k, "k_in"))) return st;
    if ((st = fetch(u_in, 3, u, "u_in"))) return st;
    if ((st = fetch(v_in, 3, v, "v_in"))) return st;
    if ((st = set_lagrange_impl(r, order, A_in, deps, num_A, E_in))) return st;
    if (num_inputs < num_delays)
        return fail(FLR_ERR_INVALID_ARGUMENT, "num_inputs must be >= num_delays");
    FLR_HIP(launch_clear(r->d_err, s));
    const int flags = r->exact_default ? FLR_FLAG_EXACT : 0;
    st = projection_impl(r, y_gw, input_in, num_delays, num_inputs, t_data, 0.0, dt,
                         sampling_frequency, k.data(), u.data(), v.data(), projections_start_ind,
                         flags, s);
    if (st) return st;
    return check_errors_impl(r, s);
}

int flr_get_tdi_delays(flr_response *r, double *delayed_links, const double *input_links,
                       int num_inputs, int num_delays, const double *t_arr,
                       const int32_t *unit_starts, const int32_t *unit_lengths,
                       const int32_t *tdi_base_link, const int32_t *tdi_link_combinations,
                       const double *tdi_signs_in, const int32_t *channels, int num_units,
                       int num_channels, int order, double sampling_frequency, int buffer_integer,
                       const double *A_in, double deps, int num_A, const double *E_in,
                       int tdi_start_ind, void *stream)
{
    (void)buffer_integer;
    if (!r || !r->orbits) return fail(FLR_ERR_NO_ORBITS, "Must add orbits with add_orbit_information method.");
    DeviceGuard g(r->device);
    hipStream_t s = (hipStream_t)stream;
    FLR_HIP(hipStreamSynchronize(s));
    int st;
    if ((st = set_lagrange_impl(r, order, A_in, deps, num_A, E_in))) return st;
    if ((st = set_tdi_impl(r, unit_starts, unit_lengths, tdi_base_link, tdi_link_combinations,
                           tdi_signs_in, channels, num_units, num_channels)))
        return st;
    FLR_HIP(launch_clear(r->d_err, s));
    const int flags = r->exact_default ? FLR_FLAG_EXACT : 0;
    st = tdi_impl(r, delayed_links, input_links, num_inputs, num_delays, t_arr, 0.0,
                  1.0 / sampling_frequency, sampling_frequency, tdi_start_ind, flags, s);
    if (st) return st;
    return check_errors_impl(r, s);
}

void flr_batch_destroy(flr_batch *b)
{
    if (!b) return;
    DeviceGuard g(b->r ? b->r->device : 0);
    if (b->d_proj) (void)hipFree(b->d_proj);
    if (b->d_tdi) (void)hipFree(b->d_tdi);
    if (b->d_proj_tiles) (void)hipFree(b->d_proj_tiles);
    if (b->d_tdi_tiles) (void)hipFree(b->d_tdi_tiles);
    if (b->uploaded) {
        (void)hipEventSynchronize(b->uploaded);
        (void)hipEventDestroy(b->uploaded);
    }
    if (b->ran) {
        if (b->ran_recorded) (void)hipEventSynchronize(b->ran);
        (void)hipEventDestroy(b->ran);
    }
    if (b->d_fused) (void)hipFree(b->d_fused);
    if (b->d_fused_tiles) (void)hipFree(b->d_fused_tiles);
    if (b->d_lti) (void)hipFree(b->d_lti);
    if (b->d_inner) (void)hipFree(b->d_inner);
    if (b->h_fused) (void)hipHostFree(b->h_fused);
    if (b->h_fused_tiles) (void)hipHostFree(b->h_fused_tiles);
    delete b;
}

int flr_batch_create(flr_response *r, const flr_source *sources, int num_sources,
                     const double *h_all, double *y_all, double *out_all, double dt,
                     double sampling_frequency, int flags, flr_batch **out)
{
    if (!out) return fail(FLR_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!r || !r->orbits) return fail(FLR_ERR_NO_ORBITS, "Must add orbits with add_orbit_information method.");
    if (!r->lag_set || !r->tdi_set) return fail(FLR_ERR_INVALID_ARGUMENT, "Lagrange and TDI tables must be set");
    if (num_sources < 1 || !sources) return fail(FLR_ERR_INVALID_ARGUMENT, "no sources");
    if (!h_all || !y_all || !out_all) return fail(FLR_ERR_INVALID_ARGUMENT, "NULL packed array");
    if (!(sampling_frequency > 0.0) || !(dt > 0.0)) return fail(FLR_ERR_INVALID_ARGUMENT, "dt and sampling_frequency must be > 0");
    DeviceGuard g(r->device);
    const int zero_edges = (flags & FLR_FLAG_ZERO_EDGES) ? 1 : 0;
    std::vector<ProjSource> ps(num_sources);
    std::vector<TdiSource> ts(num_sources);
    std::vector<int> pt(num_sources + 1, 0), tt(num_sources + 1, 0);
    int64_t samples = 0;
    const double2 *h2 = reinterpret_cast<const double2 *>(h_all);
    for (int s = 0; s < num_sources; ++s) {
        const flr_source &q = sources[s];
        if (q.num_pts < 1 || q.offset < 0 || q.projections_start_ind < 0 || q.tdi_start_ind < 0)
            return fail(FLR_ERR_INVALID_ARGUMENT, "source " + std::to_string(s) + ": bad size or offset");
        ProjSource &p = ps[s];
        p.y = y_all + 6 * q.offset;
        p.h = h2 + q.offset;
        p.t_data = nullptr;
        p.num_delays = p.num_inputs = q.num_pts;
        p.start = q.projections_start_ind;
        p.t0 = q.t0;
        for (int c = 0; c < 3; ++c) {
            p.k[c] = q.k[c];
            p.u[c] = q.u[c];
            p.v[c] = q.v[c];
        }
        TdiSource &t = ts[s];
        t.out = out_all + 3 * q.offset;
        t.y = p.y;
        t.t_arr = nullptr;
        t.num_inputs = t.num_delays = q.num_pts;
        t.start = q.tdi_start_ind;
        t.t0 = q.t0;
        const int64_t np = (int64_t)pt[s] + tiles_for(q.num_pts, p.start, zero_edges);
        const int64_t nt = (int64_t)tt[s] + tiles_for(q.num_pts, t.start, zero_edges);
        if (np > INT32_MAX || nt > INT32_MAX) return fail(FLR_ERR_UNSUPPORTED, "batch has too many tiles");
        pt[s + 1] = (int)np;
        tt[s + 1] = (int)nt;
        samples += q.num_pts;
    }
    auto b = new flr_batch();
    b->r = r;
    b->flags = flags;
    b->dt = dt;
    b->fs = sampling_frequency;
    b->num_sources = num_sources;
    b->samples = samples;
    b->proj_tiles = pt[num_sources];
    b->tdi_tiles = tt[num_sources];
    hipError_t e = hipMalloc(&b->d_proj, sizeof(ProjSource) * num_sources);
    if (e == hipSuccess) e = hipMalloc(&b->d_tdi, sizeof(TdiSource) * num_sources);
    if (e == hipSuccess) e = hipMalloc(&b->d_proj_tiles, sizeof(int) * (num_sources + 1));
    if (e == hipSuccess) e = hipMalloc(&b->d_tdi_tiles, sizeof(int) * (num_sources + 1));
    if (e == hipSuccess) e = hipMemcpy(b->d_proj, ps.data(), sizeof(ProjSource) * num_sources, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_tdi, ts.data(), sizeof(TdiSource) * num_sources, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_proj_tiles, pt.data(), sizeof(int) * (num_sources + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_tdi_tiles, tt.data(), sizeof(int) * (num_sources + 1), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        flr_batch_destroy(b);
        return hip_fail(e, "batch upload");
    }
    *out = b;
    return FLR_OK;
}

int flr_fused_reserve(flr_response *r, int capacity, double dt, double sampling_frequency,
                      int flags, flr_batch **out)
{
    if (!out) return fail(FLR_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!r || !r->orbits) return fail(FLR_ERR_NO_ORBITS, "Must add orbits with add_orbit_information method.");
    if (!r->lag_set || !r->tdi_set) return fail(FLR_ERR_INVALID_ARGUMENT, "Lagrange and TDI tables must be set");
    if (r->order != 25) return fail(FLR_ERR_UNSUPPORTED, "fused batches are built for order 25");
    if ((flags & FLR_FLAG_MIXED) && (flags & FLR_FLAG_EXACT))
        return fail(FLR_ERR_UNSUPPORTED, "FLR_FLAG_EXACT needs FP64 storage (no FLR_FLAG_MIXED)");
    if ((flags & (FLR_FLAG_Q31 | FLR_FLAG_C128)) && !(flags & FLR_FLAG_MIXED))
        return fail(FLR_ERR_INVALID_ARGUMENT, "FLR_FLAG_Q31 and FLR_FLAG_C128 select a mixed-precision input format");
    if ((flags & FLR_FLAG_Q31) && (flags & FLR_FLAG_C128))
        return fail(FLR_ERR_INVALID_ARGUMENT, "FLR_FLAG_Q31 and FLR_FLAG_C128 exclude each other");
    if (capacity < 1) return fail(FLR_ERR_INVALID_ARGUMENT, "capacity must be >= 1");
    if (!(sampling_frequency > 0.0) || !(dt > 0.0)) return fail(FLR_ERR_INVALID_ARGUMENT, "dt and sampling_frequency must be > 0");
    DeviceGuard g(r->device);
    ProjConfig pc;
    TdiConfig tc;
    StockTdi stock;
    int st = make_fused_config(r, dt, sampling_frequency, flags, kFusedPasses, pc, tc, stock);
    if (st) return st;
    auto b = new flr_batch();
    b->r = r;
    b->flags = flags;
    b->dt = dt;
    b->fs = sampling_frequency;
    b->fused = true;
    b->capacity = capacity;
    b->tile_len = tc.tile;
    hipError_t e = hipMalloc(&b->d_fused, sizeof(FusedSource) * capacity);
    if (e == hipSuccess) e = hipMalloc(&b->d_fused_tiles, sizeof(int) * (capacity + 1));
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_fused, sizeof(FusedSource) * capacity);
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_fused_tiles, sizeof(int) * (capacity + 1));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ran, hipEventDisableTiming);
    if (e == hipSuccess && (flags & FLR_FLAG_MIXED))
        e = hipMalloc(&b->d_lti, sizeof(double) * kLtiRecord * capacity);
    if (e != hipSuccess) {
        flr_batch_destroy(b);
        return hip_fail(e, "fused batch reserve");
    }
    *out = b;
    return FLR_OK;
}

namespace {

// The device record of source `s` of a fused batch (validated): pointers into the packed input
// (complex128, or 8-byte samples for the complex64 and q31 formats) and the packed float32 or
// double outputs.
int fused_source_record(const flr_source &q, int s, int flags, const void *h_all, void *out_all,
                        FusedSource &f)
{
    if (q.num_pts < 1 || q.offset < 0 || q.projections_start_ind < 0 || q.tdi_start_ind < 0)
        return fail(FLR_ERR_INVALID_ARGUMENT, "source " + std::to_string(s) + ": bad size or offset");
    if (!std::isfinite(q.scale))
        return fail(FLR_ERR_INVALID_ARGUMENT, "source " + std::to_string(s) + ": scale is not finite");
    const size_t hsz = ((flags & FLR_FLAG_MIXED) && !(flags & FLR_FLAG_C128)) ? 2 * sizeof(float)
                                                                               : 2 * sizeof(double);
    const size_t osz = (flags & FLR_FLAG_MIXED) ? sizeof(float) : sizeof(double);
    f = FusedSource{};
    f.h = static_cast<const char *>(h_all) + q.offset * hsz;
    f.out = static_cast<char *>(out_all) + 3 * q.offset * osz;
    f.num_pts = q.num_pts;
    f.proj_start = q.projections_start_ind;
    f.tdi_start = q.tdi_start_ind;
    f.t0 = q.t0;
    f.scale = q.scale == 0.0 ? 1.0 : q.scale;
    for (int c = 0; c < 3; ++c) {
        f.k[c] = q.k[c];
        f.u[c] = q.u[c];
        f.v[c] = q.v[c];
    }
    return FLR_OK;
}

// Tile prefix sums of a fused batch's sources for tiles of tile_len output samples: prefix[s]
// is the first tile of source s, prefix[num_sources] the total.
int fused_tile_prefix(const FusedSource *src, int num_sources, int zero_edges, int tile_len,
                      int *prefix)
{
    prefix[0] = 0;
    for (int s = 0; s < num_sources; ++s) {
        const int64_t nt = (int64_t)prefix[s] + tiles_for(src[s].num_pts, src[s].tdi_start,
                                                          zero_edges, tile_len);
        if (nt > INT32_MAX) return fail(FLR_ERR_UNSUPPORTED, "batch has too many tiles");
        prefix[s + 1] = (int)nt;
    }
    return FLR_OK;
}

// A fused batch's tile length follows its launch configuration, which the response's tables
// set too (a TDI table with longer delay chains has wider halos and so shorter tiles). When the
// configuration's tile differs from the one the bound sources were tiled for, re-tile them from
// the pinned staging copy and upload the prefix on `stream` (after any run still reading it).
int fused_retile(flr_batch *b, int tile_len, hipStream_t stream)
{
    if (tile_len == b->tile_len) return FLR_OK;
    bool cap = false;
    int cst = stream_capturing(stream, cap);
    if (cst) return cst;
    if (cap)
        return fail(FLR_ERR_UNSUPPORTED, "the response's tables changed the batch's tiling since "
                                         "the bind: run it once outside stream capture first");
    if (b->num_sources > 0) {
        FLR_HIP(hipEventSynchronize(b->uploaded));  // the staging copy is no longer being read
        const int zero_edges = (b->flags & FLR_FLAG_ZERO_EDGES) ? 1 : 0;
        int st = fused_tile_prefix(b->h_fused, b->num_sources, zero_edges, tile_len,
                                   b->h_fused_tiles);
        if (st) return st;
        if (b->ran_recorded) FLR_HIP(hipStreamWaitEvent(stream, b->ran, 0));
        FLR_HIP(hipMemcpyAsync(b->d_fused_tiles, b->h_fused_tiles,
                               sizeof(int) * (b->num_sources + 1), hipMemcpyHostToDevice, stream));
        FLR_HIP(hipEventRecord(b->uploaded, stream));
        b->fused_tiles = b->h_fused_tiles[b->num_sources];
    }
    b->tile_len = tile_len;
    return FLR_OK;
}

// Build the bound sources' composite TDI filters (mixed fused batches with equal-arm stock
// tables) unless they are current: they are set on bind and rebuilt when the response's
// orbits, Lagrange or TDI tables have changed since. Enqueued on `stream`.
int lti_refresh(flr_batch *b, hipStream_t stream)
{
    flr_response *r = b->r;
    if (!b->fused || !(b->flags & FLR_FLAG_MIXED) || b->num_sources < 1 ||
        b->lti_version == r->version)
        return FLR_OK;
    ProjConfig pc;
    TdiConfig tc;
    StockTdi stock;
    int st = make_fused_config(r, b->dt, b->fs, b->flags, b->tile_passes, pc, tc, stock);
    if (st) return st;
    tc.lti_fir = b->d_lti;
    Batch<FusedSource> bf{b->d_fused, b->d_fused_tiles, b->num_sources, b->fused_tiles};
    hipError_t e = launch_lti_filters(tc, bf, r->orbits->view(), r->lagrange(), stock, stream);
    if (e != hipSuccess) return hip_fail(e, "composite filter launch");
    b->lti_version = r->version;
    return FLR_OK;
}

}  // namespace

int flr_fused_bind(flr_batch *b, const flr_source *sources, int num_sources, const void *h_all,
                   void *out_all, void *stream)
{
    if (!b || !b->r || !b->fused) return fail(FLR_ERR_INVALID_ARGUMENT, "not a fused batch");
    if (num_sources < 1 || !sources) return fail(FLR_ERR_INVALID_ARGUMENT, "no sources");
    if (num_sources > b->capacity)
        return fail(FLR_ERR_INVALID_ARGUMENT, "more sources than the batch was reserved for");
    if (!h_all || !out_all) return fail(FLR_ERR_INVALID_ARGUMENT, "NULL packed array");
    DeviceGuard g(b->r->device);
    const int flags = b->flags;
    const int zero_edges = (flags & FLR_FLAG_ZERO_EDGES) ? 1 : 0;
    // the tile the response's current tables give (they may have changed since the reserve)
    ProjConfig pc;
    TdiConfig tc;
    StockTdi stock;
    int cst = make_fused_config(b->r, b->dt, b->fs, flags, b->tile_passes, pc, tc, stock);
    if (cst) return cst;
    std::vector<FusedSource> fs(num_sources);
    std::vector<int> ft(num_sources + 1, 0);
    int64_t samples = 0;
    for (int s = 0; s < num_sources; ++s) {
        if ((cst = fused_source_record(sources[s], s, flags, h_all, out_all, fs[s]))) return cst;
        samples += sources[s].num_pts;
    }
    if ((cst = fused_tile_prefix(fs.data(), num_sources, zero_edges, tc.tile, ft.data()))) return cst;
    hipStream_t st = (hipStream_t)stream;
    FLR_HIP(hipEventSynchronize(b->uploaded));  // the staging buffers are free again
    // a run still in flight (possibly on another stream) reads the tables about to be replaced
    if (b->ran_recorded) FLR_HIP(hipStreamWaitEvent(st, b->ran, 0));
    b->tile_len = tc.tile;
    // A source's composite filters depend on its first time and first TDI sample (and on the
    // response's tables, tracked by lti_version), not on its sky position or input: when these
    // match the previous bind's (still in the staging copy), the filters built then -- already
    // complete, their launch precedes the `uploaded` event waited on above -- are kept.
    bool same_refs = b->num_sources == num_sources;
    for (int s = 0; same_refs && s < num_sources; ++s)
        same_refs = b->h_fused[s].t0 == fs[s].t0 && b->h_fused[s].tdi_start == fs[s].tdi_start;
    std::memcpy(b->h_fused, fs.data(), sizeof(FusedSource) * num_sources);
    std::memcpy(b->h_fused_tiles, ft.data(), sizeof(int) * (num_sources + 1));
    FLR_HIP(hipMemcpyAsync(b->d_fused, b->h_fused, sizeof(FusedSource) * num_sources,
                           hipMemcpyHostToDevice, st));
    FLR_HIP(hipMemcpyAsync(b->d_fused_tiles, b->h_fused_tiles, sizeof(int) * (num_sources + 1),
                           hipMemcpyHostToDevice, st));
    b->num_sources = num_sources;
    b->samples = samples;
    b->fused_tiles = ft[num_sources];
    b->out_base = out_all;
    if (!same_refs) b->lti_version = ~0ull;  // new sources: their filters are built now
    b->uniform_version = ~0ull;               // and the tile-test proof is redone at the run
    const int fst = lti_refresh(b, st);
    // recorded after the filters too: flr_fused_create waits on it before returning
    FLR_HIP(hipEventRecord(b->uploaded, st));
    return fst;
}

int flr_fused_create(flr_response *r, const flr_source *sources, int num_sources,
                     const void *h_all, void *out_all, double dt, double sampling_frequency,
                     int flags, flr_batch **out)
{
    if (!out) return fail(FLR_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!r || !r->orbits) return fail(FLR_ERR_NO_ORBITS, "Must add orbits with add_orbit_information method.");
    if (num_sources < 1 || !sources) return fail(FLR_ERR_INVALID_ARGUMENT, "no sources");
    flr_batch *b = nullptr;
    int st = flr_fused_reserve(r, num_sources, dt, sampling_frequency, flags, &b);
    if (st) return st;
    DeviceGuard g(r->device);
    st = flr_fused_bind(b, sources, num_sources, h_all, out_all, nullptr);
    if (st == FLR_OK && hipEventSynchronize(b->uploaded) != hipSuccess)
        st = fail(FLR_ERR_HIP, "fused batch upload");
    if (st) {
        flr_batch_destroy(b);
        return st;
    }
    *out = b;
    return FLR_OK;
}

namespace {

// The fused launch of a bound batch on stream s: outputs, or (data set) the inner products of
// the outputs with data (laid out as out_all) into sums [num_sources][3][kInnerSums].
int fused_launch(flr_batch *b, hipStream_t s, const void *data, double *sums)
{
    flr_response *r = b->r;
    const bool exact = (b->flags & FLR_FLAG_EXACT) != 0;
    ProjConfig pc;
    TdiConfig tc;
    StockTdi stock;
    int st;
    if (b->num_sources < 1) return fail(FLR_ERR_INVALID_ARGUMENT, "no sources bound to the batch");
    if ((st = make_fused_config(r, b->dt, b->fs, b->flags, b->tile_passes, pc, tc, stock)))
        return st;
    bool cap = false;
    if ((st = stream_capturing(s, cap))) return st;
    if (data) {
        if (!(b->flags & FLR_FLAG_MIXED) || stock.generation == 0 || r->order != 25)
            return fail(FLR_ERR_UNSUPPORTED, "fused inner products run in the mixed-precision "
                                             "kernels with a stock TDI table at order 25");
    }
    if ((st = fused_retile(b, tc.tile, s))) return st;
    if (data && b->fused_tiles > b->inner_cap) {
        if (cap)
            return fail(FLR_ERR_UNSUPPORTED, "the inner-product buffer must grow: call "
                                             "flr_fused_inner once outside stream capture first");
        if (b->d_inner) {
            FLR_HIP(hipStreamSynchronize(s));
            if (b->ran_recorded) FLR_HIP(hipEventSynchronize(b->ran));
            (void)hipFree(b->d_inner);
            b->d_inner = nullptr;
            b->inner_cap = 0;
        }
        FLR_HIP(hipMalloc(&b->d_inner, sizeof(double) * 3 * kInnerSums * (size_t)b->fused_tiles));
        b->inner_cap = b->fused_tiles;
    }
    if ((st = lti_refresh(b, s))) return st;
    Batch<FusedSource> bf{b->d_fused, b->d_fused_tiles, b->num_sources, b->fused_tiles};
    tc.lti_fir = b->d_lti;
    if ((b->flags & FLR_FLAG_MIXED) && stock.generation != 0 && stock.equal_arm &&
        r->order == 25) {
        if (b->uniform_version != r->version || b->uniform_tile != tc.tile) {
            // the staging copy holds the bound sources (its upload may still be in flight:
            // reading it on the host is fine)
            bool all = true;
            for (int q = 0; q < b->num_sources && all; ++q)
                all = lti_source_uniform(r->orbits, b->h_fused[q], b->dt, b->fs,
                                         stock.generation == 1 ? 8 : 16, tc.halo_lo, tc.halo_hi);
            b->lti_uniform = all;
            b->uniform_version = r->version;
            b->uniform_tile = tc.tile;
        }
        tc.lti_uniform = b->lti_uniform ? 1 : 0;
    }
    if ((b->flags & FLR_FLAG_MIXED) && stock.generation != 0 && !stock.equal_arm &&
        r->order == 25)  // inner products run in the filter form only
        tc.utv = data ? 1 : utv_suitable(r->orbits, stock.generation == 1 ? 8 : 16, tc.tile, b->dt,
                                         b->fs, r->deps) ? 1 : 0;
    if (b->flags & FLR_FLAG_MIXED) tc.orbit_stage = orbit_stage_fits(r->orbits, tc.win_cap, b->dt);
    if (data) {
        tc.inner_part = b->d_inner;
        tc.inner_out = static_cast<const char *>(b->out_base);
        tc.inner_data = static_cast<const char *>(data);
    }
    hipError_t e = launch_fused_batch(pc, tc, bf, r->orbits->view(), r->lagrange(), exact,
                                      (b->flags & FLR_FLAG_MIXED) != 0,
                                      (b->flags & FLR_FLAG_Q31) ? kInQ31
                                      : (b->flags & FLR_FLAG_C128) ? kInC128 : kInC64,
                                      b->tile_passes, b->grid_per_cu, stock, s);
    if (e != hipSuccess) return hip_fail(e, "fused batch launch");
    if (data) {
        e = launch_inner_reduce(b->d_inner, b->d_fused_tiles, b->num_sources, sums, s);
        if (e != hipSuccess) return hip_fail(e, "inner-product reduction launch");
    }
    // an event recorded inside a capture cannot be waited on outside it: replays of a
    // captured run are ordered against later binds by the caller
    if (!cap) {
        FLR_HIP(hipEventRecord(b->ran, s));
        b->ran_recorded = true;
    }
    return FLR_OK;
}

}  // namespace

int flr_fused_inner(flr_batch *b, const void *data_all, double *sums, void *stream)
{
    if (!b || !b->r || !b->fused) return fail(FLR_ERR_INVALID_ARGUMENT, "not a fused batch");
    if (!data_all || !sums) return fail(FLR_ERR_INVALID_ARGUMENT, "NULL data or sums");
    DeviceGuard g(b->r->device);
    return fused_launch(b, (hipStream_t)stream, data_all, sums);
}

int flr_batch_run(flr_batch *b, void *stream)
{
    if (!b || !b->r) return fail(FLR_ERR_INVALID_ARGUMENT, "batch is NULL");
    flr_response *r = b->r;
    DeviceGuard g(r->device);
    hipStream_t s = (hipStream_t)stream;
    const bool exact = (b->flags & FLR_FLAG_EXACT) != 0;
    ProjConfig pc;
    TdiConfig tc;
    StockTdi stock;
    int st;
    if (b->fused) return fused_launch(b, s, nullptr, nullptr);
    st = make_proj_config(r, b->dt, b->fs, b->flags, pc);
    if (st) return st;
    st = make_tdi_config(r, b->dt, b->fs, b->flags, tc, stock);
    if (st) return st;
    Batch<ProjSource> bp{b->d_proj, b->d_proj_tiles, b->num_sources, b->proj_tiles};
    Batch<TdiSource> bt{b->d_tdi, b->d_tdi_tiles, b->num_sources, b->tdi_tiles};
    hipError_t e = launch_projection_batch(pc, bp, r->orbits->view(), r->lagrange(), exact,
                                           r->orbits->equal_arm, b->grid_per_cu, s);
    if (e != hipSuccess) return hip_fail(e, "batch projection launch");
    e = launch_tdi_batch(tc, bt, r->orbits->view(), r->lagrange(), exact, stock, b->grid_per_cu, s);
    if (e != hipSuccess) return hip_fail(e, "batch tdi launch");
    return FLR_OK;
}

int flr_batch_set_launch(flr_batch *b, int tile_samples, int grid_per_cu)
{
    if (!b || !b->r) return fail(FLR_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (grid_per_cu < 0 || grid_per_cu > 64)
        return fail(FLR_ERR_INVALID_ARGUMENT, "grid_per_cu must be in [1, 64] (0 keeps it)");
    int passes = b->tile_passes;
    if (tile_samples != 0) {
        if (!b->fused) {
            if (tile_samples != kBlock)
                return fail(FLR_ERR_UNSUPPORTED, "unfused batches use 256-sample tiles");
        } else {
            passes = tile_samples / kBlock;
            const bool mixed = (b->flags & FLR_FLAG_MIXED) != 0;
            const bool ok = tile_samples % kBlock == 0 &&
                            (mixed ? (passes == 1 || passes == 2 || passes == 4) : passes == kFusedPasses);
            if (!ok)
                return fail(FLR_ERR_UNSUPPORTED, mixed ? "mixed fused tiles are 256, 512 or 1024 samples"
                                                       : "FP64 fused tiles are 512 samples");
        }
    }
    DeviceGuard g(b->r->device);
    if (b->fused && passes != b->tile_passes) {
        ProjConfig pc;
        TdiConfig tc;
        StockTdi stock;
        int st = make_fused_config(b->r, b->dt, b->fs, b->flags, passes, pc, tc, stock);
        if (st) return st;
        // re-tile the bound sources (host copy in the staging buffer), then wait for the upload
        if ((st = fused_retile(b, tc.tile, nullptr))) return st;
        FLR_HIP(hipEventSynchronize(b->uploaded));
        b->tile_passes = passes;
    }
    if (grid_per_cu != 0) b->grid_per_cu = grid_per_cu;
    return FLR_OK;
}

int64_t flr_batch_samples(const flr_batch *b) { return b ? b->samples : 0; }

int flr_gb_generate(const flr_gb_source *sources, int num_sources, int64_t total_samples,
                    double dt, int format, void *out, void *stream)
{
    if (!sources || num_sources < 1) return fail(FLR_ERR_INVALID_ARGUMENT, "no sources");
    if (!out || total_samples < 1) return fail(FLR_ERR_INVALID_ARGUMENT, "empty output");
    if (!(dt > 0.0)) return fail(FLR_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (format != FLR_GB_COMPLEX128 && format != FLR_GB_COMPLEX64 && format != FLR_GB_Q31)
        return fail(FLR_ERR_INVALID_ARGUMENT, "unknown waveform format");
    hipError_t e = launch_gb(sources, num_sources, total_samples, dt, format, out, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "waveform launch");
    return FLR_OK;
}

int flr_response_set_exact(flr_response *r, int exact)
{
    if (!r) return fail(FLR_ERR_INVALID_ARGUMENT, "response is NULL");
    r->exact_default = exact != 0;
    return FLR_OK;
}

}  // extern "C"

// ====================================================================== allocation-free run
// flr_fused_run_static: the fused batch of flr_fused_create + flr_batch_run without any device
// allocation. Its device tables (source records, tile prefix, composite TDI filters) live in
// one file-scope device array per GPU, of fixed size: kStaticSources sources at most.

namespace {

constexpr int kStaticSources = 512;
constexpr size_t kStaticTileOff = (sizeof(FusedSource) * kStaticSources + 255) & ~size_t(255);
constexpr size_t kStaticLtiOff =
    (kStaticTileOff + sizeof(int) * (kStaticSources + 1) + 255) & ~size_t(255);
constexpr size_t kStaticBytes = kStaticLtiOff + sizeof(double) * kLtiRecord * kStaticSources;
static_assert(kStaticBytes <= 512 * 1024, "the static scratch is meant to stay under 512 KiB");

}  // namespace

// The scratch array. In host code its name is a host-side stand-in: kernels only ever receive
// the device address returned by hipGetSymbolAddress (static_scratch_address).
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on `device`, which must be current: looked up on the first call
// on that device and kept (each GPU holds its own copy of the array at its own address).
int static_scratch_address(int device, unsigned char **ptr)
{
    if (device < 0 || device >= kMaxScratchDevices)
        return fail(FLR_ERR_UNSUPPORTED, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(FLR_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[device] = p;
    }
    *ptr = static_cast<unsigned char *>(g_scratch_ptr[device]);
    return FLR_OK;
}

// Source records travel to the device as kernel arguments, kStageChunk per launch (no host
// staging buffer and no copy from host memory, so the sequence can be captured into a graph).
constexpr int kStageChunk = 16;

struct StageChunk {
    FusedSource src[kStageChunk];
    int tile_end[kStageChunk];  // tile prefix after source first + k
    int first;                  // table index of src[0]
    int count;                  // records used, <= kStageChunk
};
static_assert(sizeof(StageChunk) <= 2048, "kernel argument block stays small");

__global__ void stage_sources_kernel(StageChunk c, FusedSource *table, int *prefix)
{
    static_for<kStageChunk>([&](auto k) {
        if ((int)threadIdx.x == k && k < c.count) {
            table[c.first + k] = c.src[k];
            prefix[c.first + k + 1] = c.tile_end[k];
        }
    });
    if (threadIdx.x == 0 && c.first == 0) prefix[0] = 0;
}

}  // namespace

extern "C" {

int flr_fused_run_static(flr_response *r, const flr_source *sources, int num_sources,
                         const void *h_all, void *out_all, double dt, double sampling_frequency,
                         int flags, int tile_samples, int grid_per_cu, void *stream)
{
    if (!r || !r->orbits) return fail(FLR_ERR_NO_ORBITS, "Must add orbits with add_orbit_information method.");
    if (!r->lag_set || !r->tdi_set) return fail(FLR_ERR_INVALID_ARGUMENT, "Lagrange and TDI tables must be set");
    if (r->order != 25) return fail(FLR_ERR_UNSUPPORTED, "fused batches are built for order 25");
    if (!sources || num_sources < 1) return fail(FLR_ERR_INVALID_ARGUMENT, "no sources");
    if (num_sources > kStaticSources)
        return fail(FLR_ERR_UNSUPPORTED, "flr_fused_run_static takes at most " +
                                             std::to_string(kStaticSources) +
                                             " sources per call (its device tables have a fixed size)");
    if (!h_all || !out_all) return fail(FLR_ERR_INVALID_ARGUMENT, "NULL packed array");
    if (!(sampling_frequency > 0.0) || !(dt > 0.0)) return fail(FLR_ERR_INVALID_ARGUMENT, "dt and sampling_frequency must be > 0");
    if ((flags & FLR_FLAG_MIXED) && (flags & FLR_FLAG_EXACT))
        return fail(FLR_ERR_UNSUPPORTED, "FLR_FLAG_EXACT needs FP64 storage (no FLR_FLAG_MIXED)");
    if ((flags & (FLR_FLAG_Q31 | FLR_FLAG_C128)) && !(flags & FLR_FLAG_MIXED))
        return fail(FLR_ERR_INVALID_ARGUMENT, "FLR_FLAG_Q31 and FLR_FLAG_C128 select a mixed-precision input format");
    if ((flags & FLR_FLAG_Q31) && (flags & FLR_FLAG_C128))
        return fail(FLR_ERR_INVALID_ARGUMENT, "FLR_FLAG_Q31 and FLR_FLAG_C128 exclude each other");
    if (grid_per_cu < 0 || grid_per_cu > 64)
        return fail(FLR_ERR_INVALID_ARGUMENT, "grid_per_cu must be in [1, 64] (0: the default 32)");
    const bool mixed = (flags & FLR_FLAG_MIXED) != 0;
    int passes = kFusedPasses;
    if (tile_samples != 0) {
        passes = tile_samples / kBlock;
        const bool ok = tile_samples % kBlock == 0 &&
                        (mixed ? (passes == 1 || passes == 2 || passes == 4) : passes == kFusedPasses);
        if (!ok)
            return fail(FLR_ERR_UNSUPPORTED, mixed ? "mixed fused tiles are 256, 512 or 1024 samples"
                                                   : "FP64 fused tiles are 512 samples");
    }
    DeviceGuard g(r->device);
    hipStream_t s = (hipStream_t)stream;
    ProjConfig pc;
    TdiConfig tc;
    StockTdi stock;
    int st = make_fused_config(r, dt, sampling_frequency, flags, passes, pc, tc, stock);
    if (st) return st;
    // validate every source before anything is enqueued
    const int zero_edges = (flags & FLR_FLAG_ZERO_EDGES) ? 1 : 0;
    int64_t tiles = 0;
    for (int q = 0; q < num_sources; ++q) {
        FusedSource f;
        if ((st = fused_source_record(sources[q], q, flags, h_all, out_all, f))) return st;
        tiles += tiles_for(f.num_pts, f.tdi_start, zero_edges, tc.tile);
        if (tiles > INT32_MAX) return fail(FLR_ERR_UNSUPPORTED, "batch has too many tiles");
    }
    unsigned char *base = nullptr;
    if ((st = static_scratch_address(r->device, &base))) return st;
    FusedSource *d_src = reinterpret_cast<FusedSource *>(base);
    int *d_prefix = reinterpret_cast<int *>(base + kStaticTileOff);
    double *d_lti = reinterpret_cast<double *>(base + kStaticLtiOff);
    StageChunk chunk;
    int prefix = 0;
    for (int first = 0; first < num_sources; first += kStageChunk) {
        std::memset(&chunk, 0, sizeof(chunk));
        chunk.first = first;
        chunk.count = std::min(kStageChunk, num_sources - first);
        for (int k = 0; k < chunk.count; ++k) {
            (void)fused_source_record(sources[first + k], first + k, flags, h_all, out_all,
                                      chunk.src[k]);
            prefix += tiles_for(chunk.src[k].num_pts, chunk.src[k].tdi_start, zero_edges, tc.tile);
            chunk.tile_end[k] = prefix;
        }
        hipLaunchKernelGGL(stage_sources_kernel, dim3(1), dim3(64), 0, s, chunk, d_src, d_prefix);
        FLR_HIP(hipGetLastError());
    }
    Batch<FusedSource> bf{d_src, d_prefix, num_sources, prefix};
    const LagrangeView lg = r->lagrange();
    const OrbitView orb = r->orbits->view();
    if (mixed) {
        tc.lti_fir = d_lti;
        if (stock.generation != 0 && stock.equal_arm && r->order == 25) {
            bool all = true;
            for (int q = 0; q < num_sources && all; ++q) {
                FusedSource f;
                (void)fused_source_record(sources[q], q, flags, h_all, out_all, f);
                all = lti_source_uniform(r->orbits, f, dt, sampling_frequency,
                                         stock.generation == 1 ? 8 : 16, tc.halo_lo, tc.halo_hi);
            }
            tc.lti_uniform = all ? 1 : 0;
        }
        if (stock.generation != 0 && !stock.equal_arm && r->order == 25)
            tc.utv = utv_suitable(r->orbits, stock.generation == 1 ? 8 : 16, tc.tile, dt,
                                  sampling_frequency, r->deps) ? 1 : 0;
        tc.orbit_stage = orbit_stage_fits(r->orbits, tc.win_cap, dt);
        hipError_t e = launch_lti_filters(tc, bf, orb, lg, stock, s);
        if (e != hipSuccess) return hip_fail(e, "composite filter launch");
    }
    hipError_t e = launch_fused_batch(pc, tc, bf, orb, lg, (flags & FLR_FLAG_EXACT) != 0, mixed,
                                      (flags & FLR_FLAG_Q31) ? kInQ31
                                      : (flags & FLR_FLAG_C128) ? kInC128 : kInC64,
                                      passes, grid_per_cu != 0 ? grid_per_cu : 32, stock, s);
    if (e != hipSuccess) return hip_fail(e, "fused batch launch");
    return FLR_OK;
}

}  // extern "C"
