// flr_device.hpp -- device building blocks of the response kernels.
//
// Everything that feeds a time, a delay or the integer/fraction index split repeats the
// reference CPU operation sequence exactly (this file is compiled with -ffp-contract=off, so
// no multiply-add is fused behind our back); the fast forms only pick the orbit-table window by
// a multiplication (the interpolant is continuous across nodes). The stencils and the link
// geometry come in two forms:
//   exact: the reference's operation sequence with correctly rounded divisions -- bitwise
//          equal to the FP64 CPU path;
//   fast:  the same interpolant without per-tap divisions, and explicit fused multiply-adds in
//          the link geometry, equal to ~1e-15 relative.
#pragma once

#include "flr_internal.hpp"

namespace flr {

// ------------------------------------------------------------------ division
// gfx950 evaluates a / b as: v_rcp_f64 seed, two Newton steps y = y + y(1 - b y), quotient
// q = a y, remainder r = a - b q (exact), corrected quotient q + r y, wrapped in
// v_div_scale / v_div_fmas / v_div_fixup, which only act when an operand needs exponent
// rescaling (|a| below 2^-969, denormal or huge operands) or is special. Splitting the
// sequence at the reciprocal lets one y serve several numerators with the same denominator
// while producing the same bits as `a / b` -- the correctly rounded quotient -- for operands
// in the normal range, which signals and projections always are (|h| ~ 1e-21).

__device__ __forceinline__ double rcp_nr2(double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    return __builtin_fma(y, e, y);
}

// One Newton step from the hardware reciprocal: relative error below ~1e-13 (the square of the
// seed's), for the forms whose outputs are rounded to float32 (the other forms take rcp_nr2).
__device__ __forceinline__ double rcp_nr1(double b)
{
    const double y = __builtin_amdgcn_rcp(b);
    return __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
}

__device__ __forceinline__ double div_rcp(double a, double b, double y)
{
    double q = a * y;
    double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, y, q);
}

// ------------------------------------------------------------------ orbit interpolation
// Uniform-grid linear interpolation of the orbit tables (the lisatools Orbits methods called
// at LISAResponse.cu:331,565,570,576): window k = int(t/dt), clamped to [0, N-2], value
// slope_k * (t - dt*k) + v_k with slope_k = (v_{k+1} - v_k)/dt precomputed on the host.

struct OrbitNode {
    const double *p;  // node record: 15 values then 15 slopes
    double tau;       // t - dt*k
};

// EXACT: the window index from t / dt as the CPU path; otherwise from t * (1/dt), which can
// pick the neighbouring window at a node -- the interpolant is continuous there.
template <bool EXACT = true>
__device__ __forceinline__ int orbit_window(const OrbitView &o, double t)
{
    double q = EXACT ? t / o.dt : t * o.inv_dt;
    if (!(q >= 0.0)) return 0;
    if (q >= (double)(o.N - 1)) return o.N - 2;
    return (int)q;
}

template <bool EXACT = true>
__device__ __forceinline__ OrbitNode orbit_node(const OrbitView &o, double t)
{
    const int w = orbit_window<EXACT>(o, t);
    OrbitNode n;
    n.p = o.table + (size_t)w * kOrbitStride;
    n.tau = t - (o.dt * w);
    return n;
}

// A tile's orbit node records staged in LDS (the fused matrix-core kernels): every sample of a
// tile reads the same few nodes, which LDS serves at a fraction of the latency of the vector
// memory loads of orbit_node. Nodes [k_lo, k_hi], record of k_lo first.
constexpr int kOrbitStage = 8;

struct OrbitStage {
    const double *rec;
    int k_lo, k_hi;
};

// Stage from one below the node of (t_first - max_ltt): the nodes of a tile's sample times
// t >= t_first and of their emission times t - L, L <= max_ltt, up to kOrbitStage of them. The
// host launches a staged kernel only where the tile's whole span fits (orbit_stage_fits);
// orbit_node_staged checks every lookup anyway. Called by all nthreads threads before a barrier.
__device__ __forceinline__ OrbitStage stage_orbit_nodes(const OrbitView &o, double t_first,
                                                        double *s_orb, int tid, int nthreads)
{
    const double q = (t_first - o.max_ltt) * o.inv_dt;
    const int k_lo = !(q >= 1.0) ? 0 : q >= (double)(o.N - 1) ? o.N - 2 : (int)q - 1;
    const int k_hi = min(k_lo + kOrbitStage - 1, o.N - 2);
    const int words = (k_hi - k_lo + 1) * kOrbitStride;
    const double *src = o.table + (size_t)k_lo * kOrbitStride;
    for (int x = tid; x < words; x += nthreads) s_orb[x] = src[x];
    return OrbitStage{s_orb, k_lo, k_hi};
}

// orbit_node<false> read from the staged records: the same window and tau. A window outside
// the staged nodes sets `miss` (its caller writes NaN and latches an error) and reads the
// nearest staged node instead, so no access leaves the stage.
__device__ __forceinline__ OrbitNode orbit_node_staged(const OrbitView &o, const OrbitStage &st,
                                                       double t, bool &miss)
{
    const int w = orbit_window<false>(o, t);
    miss |= (w < st.k_lo) || (w > st.k_hi);
    const int ws = min(max(w, st.k_lo), st.k_hi);
    OrbitNode n;
    n.p = st.rec + (ws - st.k_lo) * kOrbitStride;
    n.tau = t - (o.dt * w);
    return n;
}

// Column: 0..5 light travel time of table row r; 6 + 3*s + c position of spacecraft s (0..2).
// Multiply then add, in every form: positions feed the arrival-time delays, which follow the
// reference's operation sequence.
__device__ __forceinline__ double orbit_val(const OrbitNode &n, int col)
{
    return n.p[kOrbitVals + col] * n.tau + n.p[col];
}

__device__ __forceinline__ double dot3(const double *a, const double *b)
{
    double out = 0.0;
    out += a[0] * b[0];
    out += a[1] * b[1];
    out += a[2] * b[2];
    return out;
}

// The fast forms' link geometry (unit vector, projections): the same dot product with fused
// multiply-adds (equal to the exact form up to rounding). Arrival times t - (k.x)/c keep the
// exact form in both: at t ~ 1e7 s one rounding step of t moves the signal by ~1e-10 of itself.
template <bool EXACT>
__device__ __forceinline__ double dot3_form(const double *a, const double *b)
{
    if constexpr (EXACT)
        return dot3(a, b);
    else
        return __builtin_fma(a[2], b[2], __builtin_fma(a[1], b[1], a[0] * b[0]));
}

// Integer/fraction split of a delay (LISAResponse.cu:339-341, :616-622): the stencil sits at
// fractional sample d + 1 - e, with e in [0, 1).
__device__ __forceinline__ void split_delay(double delay, double t0_offset, double fs, int &d,
                                            double &e)
{
    double clipped = delay - t0_offset;
    double cf = clipped * fs;
    int k = (int)ceil(cf) - 1;
    d = k;
    e = (1.0 + (double)k) - cf;
}

// Reciprocal square root of a positive normal x: the hardware seed y (relative error t' below
// 2^-22) refined by the binomial series (1 - t)^(-1/2) = 1 + t/2 + 3t^2/8 + ... with
// t = 1 - x y^2, to third order in t -- full double precision, without the special-value
// fix-ups of the library routine (the callers pass squared link lengths).
__device__ __forceinline__ double rsqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double t = __builtin_fma(-x * y, y, 1.0);
    return __builtin_fma(y * t, __builtin_fma(t, 0.375, 0.5), y);
}

// The same to first order in t (relative error 3t^2/8 < 1e-13), for the float32-output forms.
__device__ __forceinline__ double rsqrt_pos1(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double t = __builtin_fma(-x * y, y, 1.0);
    return __builtin_fma(y, 0.5 * t, y);
}

// ------------------------------------------------------------------ Lagrange prefactor A(e)
// Linear interpolation in the 1001-point table of A(e) (response.py:279-298), indexed like
// LISAResponse.cu:114-117; y_deps = rcp_nr2(deps). The index is clamped so that no e can read
// outside the table (the clamp never triggers for e in [0, 1)).
__device__ __forceinline__ double lagrange_A_exact(const LagrangeView &lg, double e, double y_deps)
{
    int ind = (int)div_rcp(e, lg.deps, y_deps);
    double frac = div_rcp(e - ind * lg.deps, lg.deps, y_deps);
    ind = ind < 0 ? 0 : (ind > lg.num_A - 2 ? lg.num_A - 2 : ind);
    double2 a = lg.Apair[ind];
    return a.x * (1. - frac) + a.y * frac;
}

// Same table, multiply instead of divide and one fused step: differs from the exact form by
// rounding only (the table is continuous across segment ends).
__device__ __forceinline__ double lagrange_A_fast(const LagrangeView &lg, double e)
{
    double s = e * lg.inv_deps;
    int ind = (int)s;
    ind = ind < 0 ? 0 : (ind > lg.num_A - 2 ? lg.num_A - 2 : ind);
    double frac = s - (double)ind;
    double2 a = lg.Aslope[ind];
    return __builtin_fma(a.y, frac, a.x);
}

// The order-25 (h = 13) weights E[j-1] = (-1)^j (h-1)!/(h-1-j)! h!/(h+j)!, j = 1..12, with
// the factorials rounded to double first, as pyResponseTDI._fill_A_E builds them
// (response.py:300-310): the exact bits of the table the host uploads. Kernels compiled for
// order 25 take them as constants (no kernel-argument registers); the host launches them only
// when the uploaded table holds these values (stock_e25).
__host__ __device__ constexpr double e25(int jm1)
{
    return jm1 == 0 ? -0x1.b6db6db6db6dbp-1 : jm1 == 1 ? 0x1.41d41d41d41d5p-1
         : jm1 == 2 ? -0x1.924924924924ap-2 : jm1 == 3 ? 0x1.a9f317a9f317ap-3
         : jm1 == 4 ? -0x1.7a9f317a9f318p-4 : jm1 == 5 ? 0x1.16fc0982c6248p-5
         : jm1 == 6 ? -0x1.4ec80b69ba923p-7 : jm1 == 7 ? 0x1.3ed6e64c5029bp-9
         : jm1 == 8 ? -0x1.cfc437b4d1b0fp-12 : jm1 == 9 ? 0x1.e3ee23de1275ep-15
         : jm1 == 10 ? -0x1.429ec2940c4eap-18 : 0x1.9cf42c429f1cep-23;
}

// E[j-1]: the compile-time order-25 table (CE) or the uploaded one.
template <bool CE>
__device__ __forceinline__ double lagrange_E(const LagrangeView &lg, int jm1)
{
    if constexpr (CE)
        return e25(jm1);
    else
        return lg.E[jm1];
}

// Where a stencil reads E[j-1]: the uploaded table in the kernel arguments (registers), the
// compile-time order-25 table, or a copy of the uploaded table in LDS (read per use, keeping
// the table out of the scalar registers of long loops).
constexpr int kEArg = 0;
constexpr int kEConst = 1;
constexpr int kELds = 2;
// The LDS copy plus, at E_lds + kMaxHalf, pairs {1/E[j-1], j(j+1)/E[j-1]} (load_E_lds) for the
// folded stencil of the float32-output forms (stencil_fold_c_lo).
constexpr int kELdsFold = 3;

template <int ESRC>
__device__ __forceinline__ double stencil_E(const LagrangeView &lg, const double *E_lds, int jm1)
{
    if constexpr (ESRC == kEConst)
        return e25(jm1);
    else if constexpr (ESRC == kELds)
        return E_lds[jm1];
    else
        return lg.E[jm1];
}

// Explicit weight of tap x[d + r], r in [-(h-1), h], in the fast stencil at fraction e (the
// weights the fast form applies implicitly): A B for x[d+1], A e for x[d], A D E_{j-1}/(j+e) for
// x[d+1+j] and A D E_{j-1}/(j+1-e) for x[d-j].
__device__ __forceinline__ double lagrange_weight(const LagrangeView &lg, double e, int r)
{
    const double A = lagrange_A_fast(lg, e);
    if (r == 1) return A * (1.0 - e);
    if (r == 0) return A * e;
    const double D = e * (1.0 - e);
    if (r > 1) return A * D * lg.E[r - 2] / ((double)(r - 1) + e);
    return A * D * lg.E[-r - 1] / ((double)(1 - r) - e);
}

// ------------------------------------------------------------------ stencils
// x points at sample d of a window; taps x[d-(h-1)] .. x[d+h] are read.
//
// Exact form: the reference's sequence, out = A(B x[d+1] + C x[d] + D S),
//   S = sum_{j=1}^{h-1} E_{j-1} (x[d+1+j]/(j+e) + x[d-j]/(j+1-e)),  B = 1-e, C = e, D = e(1-e)
// (LISAResponse.cu:110-147 real, :149-207 complex).
//
// Fast form: same S without per-tap divisions. Pairing tap j up and down,
//   x_u/(j+e) + x_d/(j+1-e) = (x_u (j+1-e) + x_d (j+e)) / Q_j,  Q_j = (j+e)(j+1-e) = j(j+1) + D,
// and S = T/P with the running product P = prod Q_j and T_j = T_{j-1} Q_j + E_{j-1} n_j P_{j-1},
// so one reciprocal per stencil replaces 2(h-1) divisions.

template <int H>
__device__ __forceinline__ double stencil_exact(const double *x, double e, const LagrangeView &lg,
                                                double y_deps)
{
    const int h = H > 0 ? H : lg.h;
    constexpr int kUnroll = H > 0 ? H : 1;
    double A = lagrange_A_exact(lg, e, y_deps);
    double B = 1.0 - e;
    double C = e;
    double D = e * (1.0 - e);
    double sum = 0.0;
#pragma unroll kUnroll
    for (int j = 1; j < h; j += 1) {
        double E = lg.E[j - 1];
        double F = j + e;
        double G = j + (1 - e);
        double up = x[1 + j];
        double down = x[-j];
        sum += E * (div_rcp(up, F, rcp_nr2(F)) + div_rcp(down, G, rcp_nr2(G)));
    }
    return A * (B * x[1] + C * x[0] + D * sum);
}

// x: a pointer at sample d, or any accessor with x[k] = sample d + k.
template <int H, typename X>
__device__ __forceinline__ double stencil_fast(const X &x, double e, const LagrangeView &lg)
{
    const int h = H > 0 ? H : lg.h;
    constexpr int kUnroll = H > 0 ? H : 1;
    double A = lagrange_A_fast(lg, e);
    double B = 1.0 - e;
    double D = e * B;
    double T = 0.0;
    double P = 1.0;
#pragma unroll kUnroll
    for (int j = 1; j < h; j += 1) {
        double up = x[1 + j];
        double down = x[-j];
        double qu = (double)j + e;
        double qd = (double)(j + 1) - e;
        double Q = (double)(j * (j + 1)) + D;
        double num = __builtin_fma(up, qd, down * qu);
        const double term = num * (lg.E[j - 1] * P);
        T = j == 1 ? term : __builtin_fma(T, Q, term);  // T_1 = E_0 n_1 (T_0 = 0, P_0 = 1)
        P = P * Q;
    }
    return A * __builtin_fma(D, T * rcp_nr2(P), __builtin_fma(B, x[1], e * x[0]));
}

// Complex: the h+ and hx sums share each denominator's reciprocal.
template <int H>
__device__ __forceinline__ double2 stencil_exact_c(const double2 *x, double e, const LagrangeView &lg,
                                                   double y_deps)
{
    const int h = H > 0 ? H : lg.h;
    constexpr int kUnroll = H > 0 ? H : 1;
    double A = lagrange_A_exact(lg, e, y_deps);
    double B = 1.0 - e;
    double C = e;
    double D = e * (1.0 - e);
    double sum_re = 0.0, sum_im = 0.0;
#pragma unroll kUnroll
    for (int j = 1; j < h; j += 1) {
        double E = lg.E[j - 1];
        double F = j + e;
        double G = j + (1 - e);
        double yF = rcp_nr2(F);
        double yG = rcp_nr2(G);
        double2 up = x[1 + j];
        double2 down = x[-j];
        sum_re += E * (div_rcp(up.x, F, yF) + div_rcp(down.x, G, yG));
        sum_im += E * (div_rcp(up.y, F, yF) + div_rcp(down.y, G, yG));
    }
    double2 up = x[1];
    double2 down = x[0];
    double2 out;
    out.x = A * (B * up.x + C * down.x + D * sum_re);
    out.y = A * (B * up.y + C * down.y + D * sum_im);
    return out;
}

template <int H>
__device__ __forceinline__ double2 stencil_fast_c(const double2 *x, double e, const LagrangeView &lg)
{
    const int h = H > 0 ? H : lg.h;
    constexpr int kUnroll = H > 0 ? H : 1;
    double A = lagrange_A_fast(lg, e);
    double B = 1.0 - e;
    double D = e * B;
    double Tr = 0.0, Ti = 0.0;
    double P = 1.0;
#pragma unroll kUnroll
    for (int j = 1; j < h; j += 1) {
        double2 up = x[1 + j];
        double2 down = x[-j];
        double Q = (double)(j * (j + 1)) + D;
        double EP = lg.E[j - 1] * P;
        double eEP = e * EP;
        // the pair's weights times E_{j-1} P_{j-1}, formed once for both components:
        // (j+1-e) EP for the upper tap, (j+e) EP for the lower
        double wu = __builtin_fma((double)(j + 1), EP, -eEP);
        double wd = __builtin_fma((double)j, EP, eEP);
        const double nr = __builtin_fma(up.x, wu, down.x * wd);
        const double ni = __builtin_fma(up.y, wu, down.y * wd);
        if (j == 1) {  // T_1 = n_1 (T_0 = 0, P_0 = 1)
            Tr = nr;
            Ti = ni;
        } else {
            Tr = __builtin_fma(Tr, Q, nr);
            Ti = __builtin_fma(Ti, Q, ni);
        }
        P = P * Q;
        // one tap pair in registers at a time (the loads of the next j are not hoisted): keeps
        // the fused kernels at 4 waves/SIMD
        __asm__ volatile("" ::: "memory");
    }
    double invP = rcp_nr2(P);
    double2 x1 = x[1], x0 = x[0];
    double2 out;
    out.x = A * __builtin_fma(D, Tr * invP, __builtin_fma(B, x1.x, e * x0.x));
    out.y = A * __builtin_fma(D, Ti * invP, __builtin_fma(B, x1.y, e * x0.y));
    return out;
}

// The fast complex stencil addressed from its lowest tap, lo = x - (H-1): every tap sits at a
// non-negative constant offset from one base address (x[d+1+j] = lo[H+j], x[d-j] = lo[H-1-j]),
// so no per-tap address arithmetic is needed. Same operation sequence as stencil_fast_c; E from
// where ESRC says (stencil_E).
template <int H, int ESRC>
__device__ __forceinline__ double2 stencil_fast_c_lo(const double2 *lo, double e,
                                                     const LagrangeView &lg, const double *E_lds)
{
    static_assert(H > 0, "compile-time half point count");
    double A = lagrange_A_fast(lg, e);
    double B = 1.0 - e;
    double D = e * B;
    double Tr = 0.0, Ti = 0.0;
    double P = 1.0;
#pragma unroll
    for (int j = 1; j < H; j += 1) {
        double2 up = lo[H + j];
        double2 down = lo[H - 1 - j];
        double Q = (double)(j * (j + 1)) + D;
        double EP = stencil_E<ESRC>(lg, E_lds, j - 1) * P;
        double eEP = e * EP;
        double wu = __builtin_fma((double)(j + 1), EP, -eEP);
        double wd = __builtin_fma((double)j, EP, eEP);
        const double nr = __builtin_fma(up.x, wu, down.x * wd);
        const double ni = __builtin_fma(up.y, wu, down.y * wd);
        if (j == 1) {
            Tr = nr;
            Ti = ni;
        } else {
            Tr = __builtin_fma(Tr, Q, nr);
            Ti = __builtin_fma(Ti, Q, ni);
        }
        P = P * Q;
        __asm__ volatile("" ::: "memory");  // one tap pair in registers at a time
    }
    double invP = rcp_nr2(P);
    double2 x1 = lo[H], x0 = lo[H - 1];
    double2 out;
    out.x = A * __builtin_fma(D, Tr * invP, __builtin_fma(B, x1.x, e * x0.x));
    out.y = A * __builtin_fma(D, Ti * invP, __builtin_fma(B, x1.y, e * x0.y));
    return out;
}

// The fast complex stencil with each E[j-1] folded into the running product, for the forms
// whose outputs are rounded to float32. With Q'_j = Q_j / E[j-1] = D/E[j-1] + j(j+1)/E[j-1]
// (one FMA from the LDS pair table), S = sum_j n_j / Q'_j = T/P with P = prod Q'_j and
// T_j = T_{j-1} Q'_j + n_j P_{j-1}, and the pair's weights are (j+1-e) P and (j+e) P -- one FP64
// operation per tap pair fewer than stencil_fast_c_lo. Each term's only extra rounding is that of
// 1/E[j-1], the size of E's own; |P| stays below ~1e47 at order 25. The final reciprocal takes
// one Newton step (rcp_nr1).
template <int H>
__device__ __forceinline__ double2 stencil_fold_c_lo(const double2 *lo, double e,
                                                     const LagrangeView &lg, const double *E_lds)
{
    static_assert(H > 0, "compile-time half point count");
    const double2 *F = reinterpret_cast<const double2 *>(E_lds + kMaxHalf);
    double A = lagrange_A_fast(lg, e);
    double B = 1.0 - e;
    double D = e * B;
    double Tr = 0.0, Ti = 0.0;
    double P = 1.0;
#pragma unroll
    for (int j = 1; j < H; j += 1) {
        double2 up = lo[H + j];
        double2 down = lo[H - 1 - j];
        const double2 f = F[j - 1];
        const double Q = __builtin_fma(D, f.x, f.y);
        const double eP = e * P;
        const double wu = __builtin_fma((double)(j + 1), P, -eP);
        const double wd = __builtin_fma((double)j, P, eP);
        const double nr = __builtin_fma(up.x, wu, down.x * wd);
        const double ni = __builtin_fma(up.y, wu, down.y * wd);
        if (j == 1) {
            Tr = nr;
            Ti = ni;
        } else {
            Tr = __builtin_fma(Tr, Q, nr);
            Ti = __builtin_fma(Ti, Q, ni);
        }
        P = P * Q;
        __asm__ volatile("" ::: "memory");  // one tap pair in registers at a time
    }
    double invP = rcp_nr1(P);
    double2 x1 = lo[H], x0 = lo[H - 1];
    double2 out;
    out.x = A * __builtin_fma(D, Tr * invP, __builtin_fma(B, x1.x, e * x0.x));
    out.y = A * __builtin_fma(D, Ti * invP, __builtin_fma(B, x1.y, e * x0.y));
    return out;
}

template <bool EXACT, int H>
__device__ __forceinline__ double stencil(const double *x, double e, const LagrangeView &lg,
                                          double y_deps)
{
    if constexpr (EXACT)
        return stencil_exact<H>(x, e, lg, y_deps);
    else
        return stencil_fast<H>(x, e, lg);
}

template <bool EXACT, int H>
__device__ __forceinline__ double2 stencil_c(const double2 *x, double e, const LagrangeView &lg,
                                             double y_deps)
{
    if constexpr (EXACT)
        return stencil_exact_c<H>(x, e, lg, y_deps);
    else
        return stencil_fast_c<H>(x, e, lg);
}

}  // namespace flr
