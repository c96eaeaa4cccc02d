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
                if constexpr (stock::kX[x].len == m) xr[c][k++] = ywin + rowoff[stock::base_of(c, x)] + d;
            });
        });
        double S[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
        double P = 1.0;
#pragma unroll 1
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
                        S[c][k] += E * (div_rcp(xr[c][k][1 + j], F, yF) + div_rcp(xr[c][k][-j], G, yG));
            } else {
                const double qu = (double)j + e;
                const double qd = (double)(j + 1) - e;
                const double Q = (double)(j * (j + 1)) + e * (1.0 - e);
                const double EP = E * P;
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const double num = __builtin_fma(xr[c][k][1 + j], qd, xr[c][k][-j] * qu);
                        S[c][k] = __builtin_fma(S[c][k], Q, num * EP);
                    }
                P = P * Q;
            }
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
                    const double *p = xr[c][k];
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

}  // namespace flr
