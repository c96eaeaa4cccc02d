// flr_kernels.hip -- gfx950 kernels of the LISA time-domain response.
//
// projection_kernel: one workgroup per tile of kBlock consecutive samples, one sample per lane,
//   all six links per lane. The complex h window of the tile (tile + delay halo) is staged once
//   in LDS and shared by the 3 receiver and 6 emitter stencils of every lane (the reference
//   restages it per link, LISAResponse.cu:628-655). Receiver stencils depend on the receiving
//   spacecraft only, so each is computed once for its two links: 9 complex stencils per sample
//   instead of 12, with bit-identical results.
// tdi_kernel: same tiling; the 6 y_ij rows of the tile window are staged once in LDS; each lane
//   walks every TDI unit of every channel for its sample, summing units of a channel in index
//   order in registers (no atomics), then stores X, Y, Z or A, E, T.
#include "flr_device.hpp"
#include "flr_tdi_stock.hpp"

namespace flr {

// ------------------------------------------------------------------ projection

template <bool EXACT, int H>
__global__ __launch_bounds__(kBlock) void projection_kernel(ProjParams p, OrbitView orb,
                                                            LagrangeView lg)
{
    extern __shared__ double2 hwin[];
    const int h = H > 0 ? H : lg.h;
    const int tid = threadIdx.x;
    const int i0 = p.grid_begin + (int)blockIdx.x * kBlock;
    const int i = i0 + tid;
    const int lo_valid = p.start;
    const int hi_valid = p.num_delays - p.start;
    const bool valid = (i >= lo_valid) && (i < hi_valid);
    const bool in_range = (i >= 0) && (i < p.num_delays);
    const int first = max(i0, lo_valid);
    const int last = min(i0 + kBlock, hi_valid) - 1;

    if (!valid && in_range && p.zero_edges) {
#pragma unroll
        for (int l = 0; l < kNLinks; ++l) p.y[(size_t)l * p.num_delays + i] = 0.0;
    }
    if (first > last) return;  // whole tile outside the computed range

    const int win_lo = first - p.halo_lo;
    const int win_len = last + p.halo_hi - win_lo + 1;
    if (win_len > p.win_cap || win_lo < 0 || win_lo + win_len > p.num_inputs) {
        if (valid) {
            atomicOr(p.err, kErrWindow);
            for (int l = 0; l < kNLinks; ++l) p.y[(size_t)l * p.num_delays + i] = __builtin_nan("");
        }
        return;
    }
    for (int jj = tid; jj < win_len; jj += kBlock) hwin[jj] = p.h[win_lo + jj];

    const int ic = valid ? i : first;
    const double t = p.t_data ? p.t_data[ic] : p.t0 + (double)ic * p.dt;
    const double t0_offset = p.t_data ? p.t_data[0] : p.t0 + (double)0 * p.dt;
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
        split_delay(t - dot3(p.k, xr) * kCInv, t0_offset, p.fs, d_rec, e_rec);
        const int lr = d_rec - win_lo;
        const bool br = (lr - (h - 1) < 0) || (lr + h >= win_len);
        double2 hr = make_double2(0.0, 0.0);
        if (!br) hr = stencil_c<EXACT, H>(hwin + lr, e_rec, lg, y_deps);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int row = p.rec_row[s][q];
            const int se = p.em_sc[s][q];
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
            const double un = dot3(p.u, n);
            const double vn = dot3(p.v, n);
            const double xi_p = 0.5 * ((un * un) - (vn * vn));
            const double xi_c = un * vn;
            const double kn = dot3(p.k, n);
            int d_em;
            double e_em;
            split_delay(t_em - dot3(p.k, xe) * kCInv, t0_offset, p.fs, d_em, e_em);
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
            p.y[(size_t)row * p.num_delays + i] = out;
        }
    }
    if (bad) atomicOr(p.err, kErrWindow);
}

// ------------------------------------------------------------------ TDI

template <bool EXACT, int H>
__global__ __launch_bounds__(kBlock) void tdi_kernel(TdiParams p, OrbitView orb, LagrangeView lg)
{
    extern __shared__ double smem[];
    const int h = H > 0 ? H : lg.h;
    double *ywin = smem;                    // [6][win_cap]
    double *Ls = smem + kNLinks * p.win_cap;  // [6][kBlock]
    const int tid = threadIdx.x;
    const int i0 = p.grid_begin + (int)blockIdx.x * kBlock;
    const int i = i0 + tid;
    const int lo_valid = p.start;
    const int hi_valid = p.num_delays - p.start;
    const bool valid = (i >= lo_valid) && (i < hi_valid);
    const bool in_range = (i >= 0) && (i < p.num_delays);
    const int first = max(i0, lo_valid);
    const int last = min(i0 + kBlock, hi_valid) - 1;

    if (!valid && in_range && p.zero_edges) {
#pragma unroll
        for (int c = 0; c < 3; ++c) p.out[(size_t)c * p.num_delays + i] = 0.0;
    }
    if (first > last) return;

    const int win_lo = first - p.halo_lo;
    const int win_len = last + p.halo_hi - win_lo + 1;
    if (win_len > p.win_cap || win_lo < 0 || win_lo + win_len > p.num_inputs) {
        if (valid) {
            atomicOr(p.err, kErrWindow);
            for (int c = 0; c < 3; ++c) p.out[(size_t)c * p.num_delays + i] = __builtin_nan("");
        }
        return;
    }
    for (int r = 0; r < kNLinks; ++r) {
        const double *src = p.y + (size_t)r * p.num_inputs + win_lo;
        double *dst = ywin + r * p.win_cap;
        for (int jj = tid; jj < win_len; jj += kBlock) dst[jj] = src[jj];
    }

    const int ic = valid ? i : first;
    const double t = p.t_arr ? p.t_arr[ic] : p.t0 + (double)ic * p.dt;
    const double t0_offset = p.t_arr ? p.t_arr[0] : p.t0 + (double)0 * p.dt;
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
        acc[c] = (p.overwrite || p.aet) ? 0.0 : p.out[(size_t)c * p.num_delays + i];

    bool bad = false;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        for (int u = p.ch_begin[c]; u < p.ch_begin[c + 1]; ++u) {
            const TdiUnit un = p.units[u];
            // Same-t delay chain: every light travel time is taken at the undelayed t and
            // subtracted in chain order (LISAResponse.cu:314-333).
            double delay = t;
            for (int s = 0; s < un.chain_len; ++s) {
                const int row = p.chain_rows[un.chain_start + s];
                if (row >= 0) delay -= Ls[row * kBlock + tid];
            }
            int d;
            double e;
            split_delay(delay, t0_offset, p.fs, d, e);
            const int ld = d - win_lo;
            double val;
            if ((ld - (h - 1) >= 0) && (ld + h < win_len)) {
                val = stencil<EXACT, H>(ywin + un.base_row * p.win_cap + ld, e, lg, y_deps);
            } else {
                bad = true;
                val = __builtin_nan("");
            }
            val *= un.sign;
            acc[c] += val;
        }
    }
    if (bad) atomicOr(p.err, kErrWindow);

    if (p.aet) {
        const double X = acc[0], Y = acc[1], Z = acc[2];
        p.out[i] = (Z - X) / p.s2;
        p.out[(size_t)p.num_delays + i] = (X - 2.0 * Y + Z) / p.s6;
        p.out[(size_t)2 * p.num_delays + i] = (X + Y + Z) / p.s3;
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) p.out[(size_t)c * p.num_delays + i] = acc[c];
    }
}

// ------------------------------------------------------------------ equal-arm stock TDI
// Same tiling and outputs as tdi_kernel, for the 1st/2nd-generation tables at order 25 when
// all six light travel times of the orbit table are bitwise equal (equal-arm orbits): the
// distinct delays t - m*L (m = chain length 0..3 or 0..7) are split once each and their
// stencil reciprocals (exact form) or weights (fast form) are shared by the units of that
// length -- 8 groups instead of 48 units for 2nd generation, with bit-identical results.

template <int NX, bool EXACT>
__global__ __launch_bounds__(kBlock) void tdi_equal_arm_kernel(TdiParams p, OrbitView orb,
                                                           LagrangeView lg, RowMap rm)
{
    extern __shared__ double smem[];
    double *ywin = smem;  // [6][win_cap]
    const int tid = threadIdx.x;
    const int i0 = p.grid_begin + (int)blockIdx.x * kBlock;
    const int i = i0 + tid;
    const int lo_valid = p.start;
    const int hi_valid = p.num_delays - p.start;
    const bool valid = (i >= lo_valid) && (i < hi_valid);
    const bool in_range = (i >= 0) && (i < p.num_delays);
    const int first = max(i0, lo_valid);
    const int last = min(i0 + kBlock, hi_valid) - 1;

    if (!valid && in_range && p.zero_edges) {
#pragma unroll
        for (int c = 0; c < 3; ++c) p.out[(size_t)c * p.num_delays + i] = 0.0;
    }
    if (first > last) return;

    const int win_lo = first - p.halo_lo;
    const int win_len = last + p.halo_hi - win_lo + 1;
    if (win_len > p.win_cap || win_lo < 0 || win_lo + win_len > p.num_inputs) {
        if (valid) {
            atomicOr(p.err, kErrWindow);
            for (int c = 0; c < 3; ++c) p.out[(size_t)c * p.num_delays + i] = __builtin_nan("");
        }
        return;
    }
    for (int r = 0; r < kNLinks; ++r) {
        const double *src = p.y + (size_t)r * p.num_inputs + win_lo;
        double *dst = ywin + r * p.win_cap;
        for (int jj = tid; jj < win_len; jj += kBlock) dst[jj] = src[jj];
    }

    const int ic = valid ? i : first;
    const double t = p.t_arr ? p.t_arr[ic] : p.t0 + (double)ic * p.dt;
    const double t0_offset = p.t_arr ? p.t_arr[0] : p.t0 + (double)0 * p.dt;
    double Lc[6];
    {
        const OrbitNode nd = orbit_node(orb, t);
#pragma unroll
        for (int k = 0; k < 6; ++k) Lc[k] = orbit_val(nd, rm.row[k]);
    }
    int rowoff[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) rowoff[k] = rm.row[k] * p.win_cap - win_lo;
    __syncthreads();
    if (!valid) return;
    const double y_deps = EXACT ? rcp_nr2(lg.deps) : 0.0;

    double acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        acc[c] = (p.overwrite || p.aet) ? 0.0 : p.out[(size_t)c * p.num_delays + i];
    if (!tdi_equal_arm<NX, EXACT>(t, Lc[0], t0_offset, p.fs, ywin, rowoff, win_lo, win_len, lg,
                                  y_deps, acc))
        atomicOr(p.err, kErrWindow);

    if (p.aet) {
        const double X = acc[0], Y = acc[1], Z = acc[2];
        p.out[i] = (Z - X) / p.s2;
        p.out[(size_t)p.num_delays + i] = (X - 2.0 * Y + Z) / p.s6;
        p.out[(size_t)2 * p.num_delays + i] = (X + Y + Z) / p.s3;
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) p.out[(size_t)c * p.num_delays + i] = acc[c];
    }
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

template <bool EXACT, int H>
static hipError_t launch_projection_t(const ProjParams &p, const OrbitView &orb,
                                      const LagrangeView &lg, int nblocks, size_t shmem,
                                      hipStream_t stream)
{
    auto k = projection_kernel<EXACT, H>;
    hipError_t err = set_lds_limit(k, shmem);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(kBlock), shmem, stream, p, orb, lg);
    return hipGetLastError();
}

hipError_t launch_projection(const ProjParams &p, const OrbitView &orb, const LagrangeView &lg,
                             bool exact, hipStream_t stream)
{
    const int span = p.zero_edges ? p.num_delays : p.num_delays - 2 * p.start;
    if (span <= 0) return hipSuccess;
    const int nblocks = (span + kBlock - 1) / kBlock;
    const size_t shmem = (size_t)p.win_cap * sizeof(double2);
    if (lg.h == 13)
        return exact ? launch_projection_t<true, 13>(p, orb, lg, nblocks, shmem, stream)
                     : launch_projection_t<false, 13>(p, orb, lg, nblocks, shmem, stream);
    return exact ? launch_projection_t<true, 0>(p, orb, lg, nblocks, shmem, stream)
                 : launch_projection_t<false, 0>(p, orb, lg, nblocks, shmem, stream);
}

template <bool EXACT, int H>
static hipError_t launch_tdi_t(const TdiParams &p, const OrbitView &orb, const LagrangeView &lg,
                               int nblocks, size_t shmem, hipStream_t stream)
{
    auto k = tdi_kernel<EXACT, H>;
    hipError_t err = set_lds_limit(k, shmem);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(kBlock), shmem, stream, p, orb, lg);
    return hipGetLastError();
}

template <int NX, bool EXACT>
static hipError_t launch_tdi_equal_t(const TdiParams &p, const OrbitView &orb,
                                     const LagrangeView &lg, const RowMap &rm, int nblocks,
                                     size_t shmem, hipStream_t stream)
{
    auto k = tdi_equal_arm_kernel<NX, EXACT>;
    hipError_t err = set_lds_limit(k, shmem);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(kBlock), shmem, stream, p, orb, lg, rm);
    return hipGetLastError();
}

hipError_t launch_tdi(const TdiParams &p, const OrbitView &orb, const LagrangeView &lg, bool exact,
                      const StockTdi &stock_tdi, hipStream_t stream)
{
    const int span = p.zero_edges ? p.num_delays : p.num_delays - 2 * p.start;
    if (span <= 0) return hipSuccess;
    const int nblocks = (span + kBlock - 1) / kBlock;
    if (stock_tdi.generation != 0 && stock_tdi.equal_arm && lg.h == 13) {
        const size_t shm = (size_t)kNLinks * p.win_cap * sizeof(double);
        const RowMap &rm = stock_tdi.rows;
        if (stock_tdi.generation == 1)
            return exact ? launch_tdi_equal_t<8, true>(p, orb, lg, rm, nblocks, shm, stream)
                         : launch_tdi_equal_t<8, false>(p, orb, lg, rm, nblocks, shm, stream);
        return exact ? launch_tdi_equal_t<16, true>(p, orb, lg, rm, nblocks, shm, stream)
                     : launch_tdi_equal_t<16, false>(p, orb, lg, rm, nblocks, shm, stream);
    }
    const size_t shmem = ((size_t)kNLinks * p.win_cap + (size_t)kNLinks * kBlock) * sizeof(double);
    if (lg.h == 13)
        return exact ? launch_tdi_t<true, 13>(p, orb, lg, nblocks, shmem, stream)
                     : launch_tdi_t<false, 13>(p, orb, lg, nblocks, shmem, stream);
    return exact ? launch_tdi_t<true, 0>(p, orb, lg, nblocks, shmem, stream)
                 : launch_tdi_t<false, 0>(p, orb, lg, nblocks, shmem, stream);
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
