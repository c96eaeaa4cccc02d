// flr_tdi_stock.hpp -- compile-time form of the stock 1st/2nd-generation TDI combinations.
//
// The Michelson X terms (link, sign, delay chain) are the ones pyResponseTDI builds for
// tdi="1st generation" / "2nd generation" (response.py:318-356); Y and Z are X with the
// spacecraft indices shifted cyclically by 1 and 2 (response.py:382-433). With the table
// known at compile time every link index is a constant, the six light travel times stay in
// registers, common chain prefixes are folded, and units whose delays are bitwise identical
// share one fraction's stencil weights ("groups").
#pragma once

#include "flr_device.hpp"

#include <utility>

namespace flr {
namespace stock {

// Canonical link order: index k of the register arrays and of RowMap.
constexpr int canon(int link)
{
    return link == 12 ? 0 : link == 23 ? 1 : link == 31 ? 2 : link == 13 ? 3 : link == 32 ? 4 : link == 21 ? 5 : -1;
}

// Both spacecraft indices advanced by s (1 -> 2 -> 3 -> 1).
constexpr int cyc(int link, int s)
{
    return 10 * ((link / 10 - 1 + s) % 3 + 1) + ((link % 10 - 1 + s) % 3 + 1);
}

struct Term {
    int link;
    int sign;
    int len;
    int chain[7];
};

// X of 2nd-generation TDI; the first 8 terms are 1st-generation X.
constexpr Term kX[16] = {
    {13, +1, 0, {}},
    {31, +1, 1, {13}},
    {12, +1, 2, {13, 31}},
    {21, +1, 3, {13, 31, 12}},
    {12, -1, 0, {}},
    {21, -1, 1, {12}},
    {13, -1, 2, {12, 21}},
    {31, -1, 3, {12, 21, 13}},
    {12, +1, 4, {13, 31, 12, 21}},
    {21, +1, 5, {13, 31, 12, 21, 12}},
    {13, +1, 6, {13, 31, 12, 21, 12, 21}},
    {31, +1, 7, {13, 31, 12, 21, 12, 21, 13}},
    {13, -1, 4, {12, 21, 13, 31}},
    {31, -1, 5, {12, 21, 13, 31, 13}},
    {12, -1, 6, {12, 21, 13, 31, 13, 31}},
    {21, -1, 7, {12, 21, 13, 31, 13, 31, 12}},
};

constexpr int base_of(int c, int x) { return canon(cyc(kX[x].link, c)); }
constexpr int chain_of(int c, int x, int s) { return canon(cyc(kX[x].chain[s], c)); }
constexpr int max_len(int nx) { return nx == 8 ? 3 : 7; }

// The two X terms of chain length m combine to +-(y_la - y_lb): +1 or -1, or 0 if they are not
// that pair.
constexpr int pair_sign(int m, int la, int lb)
{
    int sa = 0, sb = 0, n = 0;
    for (int x = 0; x < 16; ++x) {
        if (kX[x].len != m) continue;
        ++n;
        if (kX[x].link == la)
            sa += kX[x].sign;
        else if (kX[x].link == lb)
            sb += kX[x].sign;
        else
            return 0;
    }
    if (n != 2) return 0;
    return (sa == 1 && sb == -1) ? 1 : (sa == -1 && sb == 1) ? -1 : 0;
}

// With equal arms all units of chain length m share one delay, and (by linearity of the
// stencil) their sum is one stencil of a combined series: kind 0 = y13 - y12, kind 1 =
// y31 - y21 (cyclically shifted for Y and Z), times group_sign.
constexpr int group_kind(int m) { return pair_sign(m, 13, 12) != 0 ? 0 : 1; }
constexpr int group_sign(int m)
{
    return pair_sign(m, 13, 12) != 0 ? pair_sign(m, 13, 12) : pair_sign(m, 31, 21);
}
static_assert(group_sign(0) != 0 && group_sign(1) != 0 && group_sign(2) != 0 &&
                  group_sign(3) != 0 && group_sign(4) != 0 && group_sign(5) != 0 &&
                  group_sign(6) != 0 && group_sign(7) != 0,
              "every chain length pairs two links of one kind");


}  // namespace stock

// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1.
template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Is...>)
{
    (f(std::integral_constant<int, Is>{}), ...);
}

template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Make x's value materialise here: an empty asm that reads and writes it stops the compiler
// from sinking the arithmetic producing it past later loops (which would keep every group's
// intermediates alive at once).
__device__ __forceinline__ void pin(double &x) { __asm__ volatile("" : "+v"(x)); }
__device__ __forceinline__ void pin(int &x) { __asm__ volatile("" : "+v"(x)); }

// Number of leading X terms x = 0, 1, ... whose chain length is at most m: once the groups of
// chain lengths 0..m are evaluated, these units can be added to their channel in index order.
constexpr int done_prefix(int nx, int m)
{
    int n = 0;
    while (n < nx && stock::kX[n].len <= m) ++n;
    return n;
}

// Equal-arm TDI for one sample: all six light travel times are bitwise equal, so every chain
// of m links yields the same delay t - L - ... - L; each distinct delay is split once and the
// units of that chain length (2 per channel) share one fraction, i.e. one set of per-tap
// reciprocals (exact form) or Lagrange weights (fast form). Per unit the tap sum runs over j
// in the reference's order, and units are added to their channel sum in index order.
// Returns false if a stencil fell outside the window (the affected outputs are NaN).
template <int NX, bool EXACT>
__device__ __forceinline__ bool tdi_equal_arm(double t, double L, double t0_offset, double fs,
                                              const double *ywin, const int (&rowoff)[6],
                                              int win_lo, int win_len, const LagrangeView &lg,
                                              double y_deps, double (&acc)[3])
{
    constexpr int H = 13;
    constexpr int NG = stock::max_len(NX) + 1;
    bool ok = true;
    double Dm = t;
    double v[3][NX];
    static_for<NG>([&](auto m) {
        if constexpr (m > 0) Dm -= L;
        int d;
        double e;
        split_delay(Dm, t0_offset, fs, d, e);
        const bool gbad = (d - win_lo - (H - 1) < 0) || (d - win_lo + H >= win_len);
        ok = ok && !gbad;
        if (gbad) d = win_lo + H - 1;  // in-window stand-in; the result is discarded

        // the units of this group: (c, x) with len(x) == m, two per channel
        const double *xr[3][2];
        static_for<3>([&](auto c) {
            int k = 0;
            static_for<NX>([&](auto x) {
                // based at the lowest tap x[d - (H-1)], so every tap is a non-negative immediate
                // offset; the pinned index keeps the compiler from choosing another base
                if constexpr (stock::kX[x].len == m) {
                    int off = rowoff[stock::base_of(c, x)] + d - (H - 1);
                    pin(off);
                    xr[c][k++] = ywin + off;
                }
            });
        });
        double S[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
        double P = 1.0;
        // j fully unrolled: every tap address is the unit's base plus an immediate offset (the
        // up/down pair of a unit loads with one ds_read2_b64); the memory barrier and the pins
        // at the end of each j keep its loads and arithmetic together.
#pragma unroll
        for (int j = 1; j < H; ++j) {
            const double E = lg.E[j - 1];
            if constexpr (EXACT) {
                const double F = j + e;
                const double G = j + (1 - e);
                const double yF = rcp_nr2(F);
                const double yG = rcp_nr2(G);
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int k = 0; k < 2; ++k)
                        S[c][k] += E * (div_rcp(xr[c][k][H + j], F, yF) +
                                        div_rcp(xr[c][k][H - 1 - j], G, yG));
            } else {
                const double qu = (double)j + e;
                const double qd = (double)(j + 1) - e;
                const double Q = (double)(j * (j + 1)) + e * (1.0 - e);
                const double EP = E * P;
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const double num = __builtin_fma(xr[c][k][H + j], qd, xr[c][k][H - 1 - j] * qu);
                        S[c][k] = __builtin_fma(S[c][k], Q, num * EP);
                    }
                P = P * Q;
                pin(P);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                pin(S[c][0]);
                pin(S[c][1]);
            }
            __asm__ volatile("" ::: "memory");
        }
        double A, B, C, D, invP = 0.0;
        if constexpr (EXACT) {
            A = lagrange_A_exact(lg, e, y_deps);
            B = 1.0 - e;
            C = e;
            D = e * (1.0 - e);
        } else {
            A = lagrange_A_fast(lg, e);
            B = 1.0 - e;
            C = e;
            D = e * B;
            invP = 1.0 / P;
        }
        static_for<3>([&](auto c) {
            int k = 0;
            static_for<NX>([&](auto x) {
                if constexpr (stock::kX[x].len == m) {
                    const double *p = xr[c][k] + (H - 1);  // p[0] = x[d], p[1] = x[d+1]
                    double val;
                    if constexpr (EXACT)
                        val = A * (B * p[1] + C * p[0] + D * S[c][k]);
                    else
                        val = A * __builtin_fma(D, S[c][k] * invP, __builtin_fma(B, p[1], C * p[0]));
                    val *= (double)stock::kX[x].sign;
                    v[c][x] = gbad ? __builtin_nan("") : val;
                    pin(v[c][x]);
                    ++k;
                }
            });
        });
        // add every unit whose predecessors are all known, in index order
        constexpr int lo = m > 0 ? done_prefix(NX, m - 1) : 0;
        constexpr int hi = done_prefix(NX, m);
        static_for<3>([&](auto c) {
            static_for<NX>([&](auto x) {
                if constexpr (x >= lo && x < hi) acc[c] += v[c][x];
            });
            pin(acc[c]);
        });
    });
    return ok;
}

// Stock TDI for one sample with arbitrary (unequal) light travel times L[k] (canonical link
// order): the chains are compile-time, so each unit's delay is formed from registers in its
// chain order -- the generic tdi_units' operation sequence unit by unit, hence bitwise equal to
// it -- and every unit's stencil is addressed from its lowest tap. Units are added to their
// channel in index order. Returns false if a stencil fell outside the window.
template <int NX, bool EXACT>
__device__ __forceinline__ bool tdi_stock_unequal(double t, const double (&L)[6], double t0_offset,
                                                  double fs, const double *ywin,
                                                  const int (&rowoff)[6], int win_lo, int win_len,
                                                  const LagrangeView &lg, double y_deps,
                                                  double (&acc)[3])
{
    constexpr int H = 13;
    bool ok = true;
    static_for<3>([&](auto c) {
        static_for<NX>([&](auto x) {
            // a fresh copy of t per unit: its delay arithmetic cannot be hoisted above the
            // previous unit (which would keep every unit's split live at once)
            double delay = t;
            pin(delay);
            static_for<stock::kX[x].len>([&](auto s) { delay -= L[stock::chain_of(c, x, s)]; });
            int d;
            double e;
            split_delay(delay, t0_offset, fs, d, e);
            const bool bad = (d - win_lo - (H - 1) < 0) || (d - win_lo + H >= win_len);
            ok = ok && !bad;
            int off = rowoff[stock::base_of(c, x)] + (bad ? win_lo + H - 1 : d) - (H - 1);
            pin(off);
            double val = stencil<EXACT, H>(ywin + off + (H - 1), e, lg, y_deps);
            val = bad ? __builtin_nan("") : val;
            val *= (double)stock::kX[x].sign;
            acc[c] += val;
            pin(acc[c]);
            __asm__ volatile("" ::: "memory");
        });
    });
    return ok;
}

// 1/sqrt(2), 1/sqrt(6), 1/sqrt(3) rounded to double: the combined series form A, E, T with
// one multiplication each (the fast forms; within an ulp of the division by the rounded root).
constexpr double kInvSqrt2 = 0.70710678118654752440084436210484903928;
constexpr double kInvSqrt6 = 0.40824829046386301636621401245098189866;
constexpr double kInvSqrt3 = 0.57735026918962576450914878050195745565;

// Column of the combined series (in place of the six y rows, same stride `cap`): row
// 2*o + kind holds output o's combination of kind `kind` -- X, Y, Z themselves, or A, E, T
// formed from them when aet is set. yc points at the column; rm maps canonical links to rows
// (CANON: the rows are already in canonical order).
template <bool CANON = false>
__device__ __forceinline__ void linear_series(double *yc, int cap, const RowMap &rm, bool aet)
{
    double y[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) y[k] = yc[(CANON ? k : rm.row[k]) * cap];
    double kd[2][3];
    static_for<3>([&](auto c) {
        kd[0][c] = y[stock::canon(stock::cyc(13, c))] - y[stock::canon(stock::cyc(12, c))];
        kd[1][c] = y[stock::canon(stock::cyc(31, c))] - y[stock::canon(stock::cyc(21, c))];
    });
#pragma unroll
    for (int kind = 0; kind < 2; ++kind) {
        const double X = kd[kind][0], Y = kd[kind][1], Z = kd[kind][2];
        if (aet) {
            yc[(0 + kind) * cap] = (Z - X) * kInvSqrt2;
            yc[(2 + kind) * cap] = (X - 2.0 * Y + Z) * kInvSqrt6;
            yc[(4 + kind) * cap] = (X + Y + Z) * kInvSqrt3;
        } else {
            yc[(0 + kind) * cap] = X;
            yc[(2 + kind) * cap] = Y;
            yc[(4 + kind) * cap] = Z;
        }
    }
}

// Equal-arm stock TDI from the combined series (fast stencils): per chain length m one delay
// t - m*L, one set of stencil weights and one stencil per output -- 3 stencils per group
// instead of 6 units. Equal to the unit-by-unit sum up to rounding. Returns false if a stencil
// fell outside the window (the outputs are NaN). The taps are read through
// cols(o, kind, c0) -> a pointer p with p[k] = window column c0 + k of output o's series of
// that kind (k = 0 .. 2H-1).
template <int NX, bool CE = false, typename Cols>
__device__ __forceinline__ bool tdi_linear_cols(double t, double L, double t0_offset, double fs,
                                                int win_lo, int win_len, const LagrangeView &lg,
                                                Cols cols, double (&acc)[3])
{
    constexpr int H = 13;
    constexpr int NG = stock::max_len(NX) + 1;
    bool ok = true;
    double Dm = t;
    static_for<NG>([&](auto m) {
        if constexpr (m > 0) Dm -= L;
        int d;
        double e;
        split_delay(Dm, t0_offset, fs, d, e);
        const bool gbad = (d - win_lo - (H - 1) < 0) || (d - win_lo + H >= win_len);
        ok = ok && !gbad;
        if (gbad) d = win_lo + H - 1;  // in-window stand-in; the result is discarded
        constexpr int kind = stock::group_kind(m);
        const int c0 = d - win_lo - (H - 1);  // lowest tap
        using Taps = decltype(cols(0, kind, c0));
        const Taps xr[3] = {cols(0, kind, c0), cols(1, kind, c0), cols(2, kind, c0)};
        double S[3] = {0.0, 0.0, 0.0};
        double P = 1.0;
        const double D = e * (1.0 - e);
#pragma unroll
        for (int j = 1; j < H; ++j) {
            const double Q = (double)(j * (j + 1)) + D;
            const double EP = lagrange_E<CE>(lg, j - 1) * P;
            const double eEP = e * EP;
            // the pair's weights times E_{j-1} P_{j-1}, shared by the three outputs
            const double wu = __builtin_fma((double)(j + 1), EP, -eEP);
            const double wd = __builtin_fma((double)j, EP, eEP);
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                const Taps &x = xr[o];
                const double num = __builtin_fma(x[H + j], wu, x[H - 1 - j] * wd);
                S[o] = j == 1 ? num : __builtin_fma(S[o], Q, num);  // S_1 = n_1
                pin(S[o]);
            }
            P = P * Q;
            pin(P);
            __asm__ volatile("" ::: "memory");
        }
        const double A = lagrange_A_fast(lg, e);
        const double B = 1.0 - e;
        const double invP = 1.0 / P;
        constexpr double sgn = (double)stock::group_sign(m);
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            const Taps &x = xr[o];  // x[H - 1] = sample d, x[H] = d + 1
            const double val = A * __builtin_fma(D, S[o] * invP, __builtin_fma(B, x[H], e * x[H - 1]));
            acc[o] += gbad ? __builtin_nan("") : sgn * val;
            pin(acc[o]);
        }
    });
    return ok;
}

// Contiguous rows of pitch cap: every tap at a non-negative immediate offset from one pinned
// base address per output.
template <int NX, bool CE = false>
__device__ __forceinline__ bool tdi_linear(double t, double L, double t0_offset, double fs,
                                           const double *ywin, int cap, int win_lo, int win_len,
                                           const LagrangeView &lg, double (&acc)[3])
{
    return tdi_linear_cols<NX, CE>(t, L, t0_offset, fs, win_lo, win_len, lg,
                               [&](int o, int kind, int c0) {
                                   int off = (2 * o + kind) * cap + c0;
                                   pin(off);
                                   return ywin + off;
                               },
                               acc);
}

// Rows of the matrix-core forms, block-interleaved: window column w is stored at
//   blk_col(w, R) = (w & 15) * R + (w >> 4) + 1
// in a row of 16 R doubles (16 phases of R slots; slot 0 of every phase is a zero guard for
// columns -16 .. -1, and slots past the window are zero too). Columns 16 apart are adjacent
// doubles, so a matrix operand's 16 block columns are 16 consecutive doubles (no bank
// conflicts), and its tap steps j = 4a + i sit at blk_col(w0 + 4i) + a: four per-lane bases
// and compile-time offsets. R (odd, so that a wave storing consecutive columns spreads over
// the banks) leaves room for the operands' reach of up to 56 columns past a block start.
__host__ __device__ constexpr int blk_R(int cap) { return ((cap + 79) >> 4) | 1; }
__host__ __device__ constexpr int blk_pitch(int cap) { return 16 * blk_R(cap); }
__host__ __device__ constexpr int blk_col(int w, int R) { return (w & 15) * R + (w >> 4) + 1; }

// Zero the guard slots and the slots past a window of len columns in the six rows at ywin
// (before the projection fills the window, which overwrites the last partial slot).
__device__ __forceinline__ void blk_zero(double *ywin, int pitch, int R, int len, int tid,
                                         int nthreads)
{
    const int s0 = 1 + (len >> 4);  // first slot that may hold no window column
    const int per = R - s0 + 1;     // slot 0 and slots s0 .. R-1, per phase
    for (int x = tid; x < kNLinks * 16 * per; x += nthreads) {
        const int r = x / (16 * per), rem = x - r * 16 * per;
        const int ph = rem / per, k = rem - ph * per;
        ywin[r * pitch + ph * R + (k == 0 ? 0 : s0 + k - 1)] = 0.0;
    }
}

// Block-interleaved window column c0 + k of a row, read through operator[].
struct BlockTaps {
    const double *row;
    int R;
    int c0;
    __device__ __forceinline__ double operator[](int k) const { return row[blk_col(c0 + k, R)]; }
};

// tdi_linear on block-interleaved rows (the per-sample fallback of the matrix-core path).
template <int NX, bool CE = false>
__device__ __forceinline__ bool tdi_linear_padded(double t, double L, double t0_offset, double fs,
                                                  const double *ywin, int pitch, int R,
                                                  int win_lo, int win_len, const LagrangeView &lg,
                                                  double (&acc)[3])
{
    return tdi_linear_cols<NX, CE>(t, L, t0_offset, fs, win_lo, win_len, lg,
                               [&](int o, int kind, int c0) {
                                   return BlockTaps{ywin + (2 * o + kind) * pitch, R, c0};
                               },
                               acc);
}

// ------------------------------------------------------------------ time-invariant form
// With equal, constant arms on a uniform time grid the stock TDI is a linear time-invariant
// filter: the position of chain length m relative to sample i, delta_m = d_m - i + 1 - e_m
// (the stencil interpolates at sample d_m + 1 - e_m), is the same for every i up to FP64
// rounding of the times (~1e-9 samples). The groups of one series kind then add up to one
// composite filter of at most kFir taps, built once per source from a reference sample;
// each sample checks that its own positions agree within kLtiTol samples (and falls back to
// tdi_linear otherwise), so the result differs from per-sample weights by at most
// kLtiTol * |dx/dsample| per stencil -- used by the mixed-precision kernels only.
constexpr int kFir = 40;
constexpr double kLtiTol = 1e-8;
// Per-source record of a mixed batch's composite filters (tc.lti_fir): the two filters
// ([2][kFir]), then the reference arm length, the first tap kmin and the fit flag (1 or 0) of
// lti_source_reference, padded to a multiple of 4 doubles.
constexpr int kLtiLref = 2 * kFir;
constexpr int kLtiKmin = 2 * kFir + 1;
constexpr int kLtiOk = 2 * kFir + 2;
constexpr int kLtiRecord = 2 * kFir + 4;
// One kind's composite filter as the matrix-core path reads it: taps -15 .. 55 (zero outside
// 0 .. kFir-1), so that a 16-output block's Toeplitz operand F[q - r] (q in 0..55, r in 0..15)
// is one indexed load.
constexpr int kFpadLo = 15;
constexpr int kFpad = 72;

template <int NX>
struct LtiRef {
    static constexpr int NG = stock::max_len(NX) + 1;
    int off[NG];       // d_m - i at the reference sample
    double e[NG];
    double delta[NG];  // d_m - i + 1 - e_m
    int kmin;          // first tap of the composite filter, relative to i
    bool ok;           // the filter fits in kFir taps
};

template <int NX>
__device__ __forceinline__ void lti_reference(double t, int i, double L, double t0_offset,
                                              double fs, LtiRef<NX> &ref)
{
    constexpr int H = 13;
    double Dm = t;
    int lo = 1 << 30, hi = -(1 << 30);
    static_for<LtiRef<NX>::NG>([&](auto m) {
        if constexpr (m > 0) Dm -= L;
        int d;
        double e;
        split_delay(Dm, t0_offset, fs, d, e);
        ref.off[m] = d - i;
        ref.e[m] = e;
        ref.delta[m] = (double)(d - i) + 1.0 - e;
        lo = min(lo, d - i);
        hi = max(hi, d - i);
    });
    ref.kmin = lo - (H - 1);
    ref.ok = (hi + H) - ref.kmin + 1 <= kFir;
}

// Entry idx of the two composite filters: kind idx / kFir, tap kmin + idx % kFir; the groups
// are added in chain-length order.
template <int NX>
__device__ __forceinline__ double lti_filter_entry(const LtiRef<NX> &ref, const LagrangeView &lg,
                                                   int idx)
{
    constexpr int H = 13;
    const int kind = idx / kFir;
    const int k = ref.kmin + idx % kFir;
    double f = 0.0;
    static_for<LtiRef<NX>::NG>([&](auto m) {
        if (stock::group_kind(m) == kind) {
            const int r = k - ref.off[m];
            if (r >= -(H - 1) && r <= H)
                f += (double)stock::group_sign(m) * lagrange_weight(lg, ref.e[m], r);
        }
    });
    return f;
}

// Sample i at time t may use the composite filter: every chain length's position agrees with
// the reference within kLtiTol and the filter's taps lie in the window.
template <int NX>
__device__ __forceinline__ bool lti_matches(double t, int i, double L, double t0_offset,
                                            double fs, const LtiRef<NX> &ref, int win_lo,
                                            int win_len)
{
    const int first_tap = i + ref.kmin - win_lo;
    bool ok = first_tap >= 0 && first_tap + kFir <= win_len;
    double Dm = t;
    static_for<LtiRef<NX>::NG>([&](auto m) {
        if constexpr (m > 0) Dm -= L;
        int d;
        double e;
        split_delay(Dm, t0_offset, fs, d, e);
        const double delta = (double)(d - i) + 1.0 - e;
        ok = ok && fabs(delta - ref.delta[m]) <= kLtiTol;
    });
    return ok;
}

// ------------------------------------------------------------------ slowly varying arms
// With unequal, time-varying arms every unit u of the stock TDI has a position of its own,
// delta_u(i) = d_u - i + 1 - e_u. Within a tile of some hundred samples the light travel times
// are linear in t (the orbit tables are interpolated linearly, and where a tile crosses a node
// the slope changes by ~1e-13), so delta_u(i) = delta_u(m) + sigma_u (i - m) about the tile's
// middle sample m, with |sigma_u| ~ 1e-10 for LISA-like arm rates. Expanding each unit's stencil
// weights w_r(e) to first order in e_u(i) - e_u(m) = -sigma_u (i - m), channel c of every output
// sample of the tile is, over the channel's four base links q,
//   X_c(i) = sum_q sum_k (F0_cq[k] + (i - m) F1_cq[k]) y_q[i + kmin + k],
//   F0_cq[k] = sum_{u in (c,q)} sign_u w_{k+kmin-off_u}(e_u),
//   F1_cq[k] = -sum_{u in (c,q)} sign_u sigma_u w'_{k+kmin-off_u}(e_u),
// with off_u = d_u - m and e_u taken at m: 24 composite filters per tile, applied on the matrix
// cores. The reference positions are computed in double-double (utv_unit), so unlike the
// per-sample delays they carry no FP64 rounding of the times. The weights are the fast stencil's
// (lagrange_weight_d), so the result differs from stencils at exact delays by the model's
// position error (at most kLtiTol, tested per unit on the tile's two ends and middle), the
// second-order term (|sigma| (b - a) / 2 <= kUtvDrift keeps it
// below 1e-9 of the interpolated value) and a change of the A table's slope inside the tile's
// range of e (tolerated up to kUtvKnot past a node of the table). A tile that fails any test
// takes the per-sample stencils instead.
constexpr double kUtvDrift = 1e-4;
constexpr double kUtvKnot = 3e-8;
constexpr int kUtvLinks = 4;                  // base links per channel (q)
constexpr int kUtvPairs = 3 * kUtvLinks;      // (channel, base link) filters per kind
constexpr int kUtvLink[kUtvLinks] = {13, 31, 12, 21};  // X's base links in order of first use

// Base-link index q of X term x.
__host__ __device__ constexpr int utv_q(int x)
{
    return stock::kX[x].link == 13 ? 0 : stock::kX[x].link == 31 ? 1 : stock::kX[x].link == 12 ? 2 : 3;
}

// Double-double arithmetic (value hi + lo, |lo| <= ulp(hi)/2): the model's reference positions
// are computed to ~1e-20 relative, so the tile's model carries none of the FP64 rounding of a
// time ~1e7 s (~1e-9 s) that every per-sample delay carries.
struct DD {
    double hi, lo;
};
__device__ __forceinline__ DD dd_two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DD dd_norm(double hi, double lo)
{
    const double s = hi + lo;
    return {s, lo - (s - hi)};
}
__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    const DD s = dd_two_sum(a.hi, b.hi);
    return dd_norm(s.hi, s.lo + (a.lo + b.lo));
}
__device__ __forceinline__ DD dd_neg(DD a) { return {-a.hi, -a.lo}; }
__device__ __forceinline__ DD dd_mul(DD a, double b)
{
    const double p = a.hi * b;
    return dd_norm(p, __builtin_fma(a.hi, b, -p) + a.lo * b);
}
__device__ __forceinline__ DD dd_prod(double a, double b)
{
    const double p = a * b;
    return {p, __builtin_fma(a, b, -p)};
}

// Per-tile model of the units, in LDS: the tile's first, middle and last sample time and the six
// light travel times there (canonical order), in double-double; per unit (channel * NX + term)
// off_u, e_u, sigma_u; the filters' tap range over all units and whether every unit passed.
struct UtvShared {
    double th[3], tl[3];
    double Lh[3][6], Ll[3][6];
    double off[48], e[48], sigma[48];
    double A[48], dA[48];  // the interpolated A table at e_u and its slope in e
    int kmin, kmax, ok;
};

// Reference sample s (0: first, 1: middle, 2: last) of the tile at sample i: t0 + i dt and the
// orbit tables' linear interpolant L_k(t) = v_k + slope_k (t - dt w) of every link, in
// double-double (window w from t as orbit_node picks it; the interpolant is continuous at nodes).
__device__ __forceinline__ void utv_reference_times(UtvShared &sh, int s, int i, double t0,
                                                    double dt, const OrbitView &orb,
                                                    const RowMap &rm)
{
    const DD t = dd_add({t0, 0.0}, dd_prod((double)i, dt));
    const double q = t.hi * orb.inv_dt;
    const int w = !(q >= 0.0) ? 0 : (q >= (double)(orb.N - 1) ? orb.N - 2 : (int)q);
    const DD tau = dd_add(t, dd_neg(dd_prod(orb.dt, (double)w)));
    const double *p = orb.table + (size_t)w * kOrbitStride;
    sh.th[s] = t.hi;
    sh.tl[s] = t.lo;
#pragma unroll
    for (int k = 0; k < kNLinks; ++k) {
        const DD L = dd_add(dd_mul(tau, p[kOrbitVals + rm.row[k]]), {p[rm.row[k]], 0.0});
        sh.Lh[s][k] = L.hi;
        sh.Ll[s][k] = L.lo;
    }
}

// Unit u = c * NX + x of the tile [a, b] (middle m): its positions at a, m and b from the
// double-double times (the chain subtracted in the reference's order, then the offset and the
// scale by fs, LISAResponse.cu:314-341), and its model tests. Writes the unit's record and folds
// its tap range and verdict into sh.
// Unit u (channel u / NX, term u % NX) of the stock tables: the canonical rows of its chain, 3
// bits per link in chain order, and the chain's length in bits 24 and up (utv_unit's argument).
__device__ __forceinline__ int utv_chain_word(int u, int nx)
{
    const int c = u / nx, x = u - c * nx;
    const stock::Term &tm = stock::kX[x];
    int w = tm.len << 24;
    for (int k = 0; k < tm.len; ++k) w |= stock::canon(stock::cyc(tm.chain[k], c)) << (3 * k);
    return w;
}

template <int NX>
__device__ __forceinline__ void utv_unit(UtvShared &sh, int u, int chain, int a, int m, int b,
                                         double t0_offset, double fs, const LagrangeView &lg)
{
    constexpr int H = 13;
    double pos[3];
    int off_m = 0;
    double e_m = 0.0;
    // the chain's links in order, summed for the three reference samples side by side (each
    // sum in the chain's order)
    const int len = chain >> 24;
    DD delays[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) delays[s] = {sh.th[s], sh.tl[s]};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        if (k < len) {
            const int l = (chain >> (3 * k)) & 7;
#pragma unroll
            for (int s = 0; s < 3; ++s) delays[s] = dd_add(delays[s], {-sh.Lh[s][l], -sh.Ll[s][l]});
        }
    }
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int i = s == 0 ? a : s == 1 ? m : b;
        const DD cf = dd_mul(dd_add(delays[s], {-t0_offset, 0.0}), fs);
        // d = ceil(cf) - 1 and e = 1 + d - cf, the split of split_delay on the exact value
        double cu = ceil(cf.hi);
        if (cf.hi == cu && cf.lo > 0.0) cu += 1.0;
        const int d = (int)cu - 1;
        const double e = ((1.0 + (double)d) - cf.hi) - cf.lo;
        pos[s] = (double)(d - i) + 1.0 - e;
        if (s == 1) {
            off_m = d - m;
            e_m = e;
        }
    }
    const double sigma = b > a ? (pos[2] - pos[0]) / (double)(b - a) : 0.0;
    // (1) the line through the middle position with the chord's slope is off the positions at
    // a and b by dev, and off any position in between by at most 3 dev (the positions are
    // piecewise linear in i between the orbit nodes)
    const double dev = fabs(pos[0] + sigma * (double)(m - a) - pos[1]);
    bool ok = 3.0 * dev <= kLtiTol;
    // (2) drift across the tile
    const double drift = fabs(sigma) * (double)max(m - a, b - m);
    ok = ok && drift <= kUtvDrift;
    // (3) e over the tile, [e_m - drift, e_m + drift], within the A-table segment of e_m
    const double s_m = e_m * lg.inv_deps;
    const int ind = min(max((int)s_m, 0), lg.num_A - 2);
    const double over = fmax(fmax((double)ind - (s_m - drift * lg.inv_deps),
                                  (s_m + drift * lg.inv_deps) - (double)(ind + 1)),
                             0.0) / lg.inv_deps;
    ok = ok && over <= kUtvKnot;
    sh.off[u] = (double)off_m;
    sh.e[u] = e_m;
    sh.sigma[u] = sigma;
    const double2 as = lg.Aslope[ind];  // A as lagrange_A_fast interpolates it, and its slope
    sh.A[u] = __builtin_fma(as.y, s_m - (double)ind, as.x);
    sh.dA[u] = as.y * lg.inv_deps;
    atomicMin(&sh.kmin, off_m - (H - 1));
    atomicMax(&sh.kmax, off_m + H);
    if (!ok) sh.ok = 0;
}

// Weight of tap d + r (r = -(H-1) .. H) of the fast stencil at fraction e, and its derivative
// in e: w_1 = A (1 - e), w_0 = A e, w_{j+1} = A D E_{j-1} / (j + e), w_{-j} = A D E_{j-1} /
// (j + 1 - e), D = e (1 - e); A is the linearly interpolated table at e and dA its slope
// within e's segment (utv_unit).
__device__ __forceinline__ void lagrange_weight_d(double A, double dA, const double *E_lds,
                                                  double e, int r, double &w, double &dw)
{
    const double B = 1.0 - e;
    const double D = e * B;
    const double dD = 1.0 - 2.0 * e;
    if (r == 1) {
        w = A * B;
        dw = __builtin_fma(dA, B, -A);
    } else if (r == 0) {
        w = A * e;
        dw = __builtin_fma(dA, e, A);
    } else {
        const bool up = r > 1;
        const int j = up ? r - 1 : -r;
        const double inv_q = rcp_nr2(up ? (double)j + e : (double)(j + 1) - e);
        const double Ej = E_lds[j - 1];
        const double AD = A * D;
        w = AD * Ej * inv_q;
        // d/de [A D / q] = (A' D + A D') / q -+ A D / q^2, q = j + e (up) or j + 1 - e
        const double lin = __builtin_fma(dA, D, A * dD);
        dw = Ej * inv_q * (up ? __builtin_fma(-AD, inv_q, lin) : __builtin_fma(AD, inv_q, lin));
    }
}

// Entry k (0 <= k < kFir) of both filters of pair p = c * kUtvLinks + q of the tile's model.
template <int NX>
__device__ __forceinline__ void utv_filter_entry(const UtvShared &sh, const LagrangeView &lg,
                                                 const double *E_lds, int p, int k, double &f0,
                                                 double &f1)
{
    constexpr int H = 13;
    const int c = p / kUtvLinks, q = p - c * kUtvLinks;
    f0 = 0.0;
    f1 = 0.0;
    for (int x = 0; x < NX; ++x) {  // the pair's units in index order
        if (utv_q(x) != q) continue;
        const int u = c * NX + x;
        const int r = k + sh.kmin - (int)sh.off[u];
        if (r < -(H - 1) || r > H) continue;
        double w, dw;
        lagrange_weight_d(sh.A[u], sh.dA[u], E_lds, sh.e[u], r, w, dw);
        const double sg = (double)stock::kX[x].sign;
        f0 = __builtin_fma(sg, w, f0);
        f1 = __builtin_fma(-sg * sh.sigma[u], dw, f1);
    }
}

// tdi_stock_unequal (fast stencils) on block-interleaved rows in canonical link order (row k at
// ywin + k * pitch): the per-sample path of the slowly-varying-arm form.
template <int NX>
__device__ __forceinline__ bool tdi_stock_unequal_padded(double t, const double (&L)[6],
                                                         double t0_offset, double fs,
                                                         const double *ywin, int pitch, int R,
                                                         int win_lo, int win_len,
                                                         const LagrangeView &lg, double (&acc)[3])
{
    constexpr int H = 13;
    bool ok = true;
    static_for<3>([&](auto c) {
        static_for<NX>([&](auto x) {
            double delay = t;
            pin(delay);
            static_for<stock::kX[x].len>([&](auto s) { delay -= L[stock::chain_of(c, x, s)]; });
            int d;
            double e;
            split_delay(delay, t0_offset, fs, d, e);
            const bool bad = (d - win_lo - (H - 1) < 0) || (d - win_lo + H >= win_len);
            ok = ok && !bad;
            // at[k] = window column of sample d + k
            const BlockTaps at{ywin + stock::base_of(c, x) * pitch, R, bad ? H - 1 : d - win_lo};
            double val = stencil_fast<H>(at, e, lg);
            val = bad ? __builtin_nan("") : val;
            val *= (double)stock::kX[x].sign;
            acc[c] += val;
            pin(acc[c]);
        });
    });
    return ok;
}

}  // namespace flr
