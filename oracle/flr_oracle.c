/*
 * flr_oracle.c -- FP64 CPU restatement of the fastlisaresponse time-domain response path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP kernels: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it. The product
 * (the HIP library and its Python host layer) never calls, links or imports it.
 *
 * It follows the CPU branch of the reference, operation by operation, so that the GPU
 * kernels' time/index arithmetic can be compared bit for bit:
 *   interp_single   <- src/fastlisaresponse/cutils/LISAResponse.cu:110-147
 *   interp (cplx)   <- src/fastlisaresponse/cutils/LISAResponse.cu:149-207
 *   TDI loop        <- src/fastlisaresponse/cutils/LISAResponse.cu:218-397 (CPU branch, :300-393)
 *   projection loop <- src/fastlisaresponse/cutils/LISAResponse.cu:431-672 (CPU branch, :519-668)
 *   XYZ -> AET      <- lisatools.utils.utility.AET, called at src/fastlisaresponse/response.py:665
 *
 * Orbit interpolation restates the lisatools `Orbits` device class (third-party package
 * lisaanalysistools >= 1.2.8, not present here; call sites LISAResponse.cu:311,326,331,565,
 * 570,576): uniform-grid linear interpolation of the tabulated light-travel times [N][6]
 * and spacecraft positions [N][3][3]. Its exact lisatools formula cannot be checked offline:
 * parity unpinned against lisatools (the tests pin it against numpy.interp instead). The XYZ ->
 * AET convention (A = (Z - X)/sqrt 2, E = (X - 2Y + Z)/sqrt 6, T = (X + Y + Z)/sqrt 3) is
 * likewise parity unpinned against lisatools; the tests pin its orthonormality and the images
 * of unit inputs.
 *
 * Build with -ffp-contract=off: the reference's operation sequence has no fused multiply-adds
 * and the time arithmetic must round exactly as written.
 *
 * Wide build (-DORACLE_WIDE, liboracle_flr_wide.so): the value arithmetic -- stencil sums and
 * prefactors, link geometry, projections, TDI sums, AET -- runs in x87 extended precision
 * (long double, 64-bit significand), while the time, delay and index arithmetic stays the FP64
 * sequence above, so every stencil sits at the same sample position with the same fraction as
 * in the FP64 build. Its results approximate the FP64 path's exact values with ~2^11 less
 * rounding: the reference against which the FP64 path's own rounding floor is measured.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_WIDE
typedef long double real_t; /* value arithmetic, extended precision */
#else
typedef double real_t;      /* value arithmetic, FP64 as the reference */
#endif

#define ORACLE_NLINKS 6
#define ORACLE_C_INV 3.3356409519815204e-09 /* LISAResponse.hh:7 */

/* Mutation hooks for the tests that show the oracle's checks have teeth: a build with
 * -DORACLE_MUTANT=k plants one plausible mistake (never set in the normal build):
 *   1 receiver delay with the wrong sign (t + k.x/c)
 *   2 the last link of every TDI delay chain dropped
 *   3 cross polarisation projection with the wrong sign
 *   4 receiver and emitter spacecraft swapped in the projection geometry
 *   5 A channel with the wrong sign
 *   6 stencil taps shifted by one sample */
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

typedef struct {
    double dt;          /* orbit grid spacing (s) */
    int N;              /* number of orbit nodes */
    const double *ltt;  /* [N][6] light travel times, rows in links[] order */
    const double *x;    /* [N][3][3] spacecraft positions (m), spacecraft 1..3 */
    int links[ORACLE_NLINKS];
    int sc_r[ORACLE_NLINKS];
    int sc_e[ORACLE_NLINKS];
    double armlength;
} oracle_orbits;

/* ---------------------------------------------------------------- orbits */

/* Row of a link number (12, 23, ...) in the orbit tables; -1 if unknown. */
int oracle_link_row(const oracle_orbits *o, int link)
{
    for (int r = 0; r < ORACLE_NLINKS; ++r)
        if (o->links[r] == link) return r;
    return -1;
}

/* Interval of the uniform orbit grid holding t, clamped to [0, N-2] (linear
 * extrapolation outside the table). */
static int orbit_window(const oracle_orbits *o, double t)
{
    double q = t / o->dt;
    if (!(q >= 0.0)) return 0;
    if (q >= (double)(o->N - 1)) return o->N - 2;
    return (int)q;
}

static double orbit_lerp(const oracle_orbits *o, double t, const double *arr, int w, int stride, int off)
{
    double up = arr[(w + 1) * stride + off];
    double down = arr[w * stride + off];
    return ((up - down) / o->dt) * (t - (o->dt * w)) + down;
}

double oracle_ltt(const oracle_orbits *o, double t, int link)
{
    int row = oracle_link_row(o, link);
    if (row < 0) return 0.0;
    int w = orbit_window(o, t);
    return orbit_lerp(o, t, o->ltt, w, ORACLE_NLINKS, row);
}

void oracle_pos(const oracle_orbits *o, double t, int sc, double out[3])
{
    int w = orbit_window(o, t);
    int s = sc - 1;
    for (int c = 0; c < 3; ++c) out[c] = orbit_lerp(o, t, o->x, w, 9, 3 * s + c);
}

/* ---------------------------------------------------------------- stencils */

/* Real order-(2h-1) Lagrange fractional-delay stencil, LISAResponse.cu:110-147. */
static real_t interp_single(const real_t *input, int h, int d, double e, const double *A_arr,
                            double deps, const double *E_arr)
{
#if ORACLE_MUTANT == 6
    d += 1;
#endif
    int ind = (int)(e / deps);
    real_t frac = ((real_t)e - ind * (real_t)deps) / (real_t)deps;
    real_t A = A_arr[ind] * (1. - frac) + A_arr[ind + 1] * frac;
    real_t B = 1.0 - (real_t)e;
    real_t C = e;
    real_t D = (real_t)e * (1.0 - (real_t)e);
    real_t sum = 0.0;
    for (int j = 1; j < h; j += 1) {
        real_t E = E_arr[j - 1];
        real_t F = j + (real_t)e;
        real_t G = j + (1 - (real_t)e);
        real_t up = input[d + 1 + j];
        real_t down = input[d - j];
        sum += E * (up / F + down / G);
    }
    real_t up = input[d + 1];
    real_t down = input[d];
    return A * (B * up + C * down + D * sum);
}

/* Complex stencil on interleaved (re, im) doubles, LISAResponse.cu:149-207. */
static void interp_cplx(real_t *out_re, real_t *out_im, const double *input, int h, int d, double e,
                        const double *A_arr, double deps, const double *E_arr)
{
    int ind = (int)(e / deps);
    real_t frac = ((real_t)e - ind * (real_t)deps) / (real_t)deps;
    real_t A = A_arr[ind] * (1. - frac) + A_arr[ind + 1] * frac;
    real_t B = 1.0 - (real_t)e;
    real_t C = e;
    real_t D = (real_t)e * (1.0 - (real_t)e);
    real_t sum_re = 0.0, sum_im = 0.0;
    for (int j = 1; j < h; j += 1) {
        real_t E = E_arr[j - 1];
        real_t F = j + (real_t)e;
        real_t G = j + (1 - (real_t)e);
        const double *up = &input[2 * (d + 1 + j)];
        const double *down = &input[2 * (d - j)];
        sum_re += E * ((real_t)up[0] / F + (real_t)down[0] / G);
        sum_im += E * ((real_t)up[1] / F + (real_t)down[1] / G);
    }
    const double *up = &input[2 * (d + 1)];
    const double *down = &input[2 * d];
    *out_re = A * (B * (real_t)up[0] + C * (real_t)down[0] + D * sum_re);
    *out_im = A * (B * (real_t)up[1] + C * (real_t)down[1] + D * sum_im);
}

/* The standalone stencils, exposed for the known-answer tests. */
real_t oracle_interp_single(const real_t *input, int h, int d, double e, const double *A_arr,
                            double deps, const double *E_arr)
{
    return interp_single(input, h, d, e, A_arr, deps, E_arr);
}

void oracle_interp_cplx(real_t *out2, const double *input, int h, int d, double e, const double *A_arr,
                        double deps, const double *E_arr)
{
    interp_cplx(&out2[0], &out2[1], input, h, d, e, A_arr, deps, E_arr);
}

/* Integer/fraction split of a delay, LISAResponse.cu:339-341 and :616-622. */
static void split_delay(double delay, double t0_offset, double fs, int *integer_delay, double *fraction)
{
    double clipped = delay - t0_offset;
    int k = (int)ceil(clipped * fs) - 1;
    *integer_delay = k;
    *fraction = 1.0 + k - clipped * fs;
}

void oracle_split_delay(double delay, double t0_offset, double fs, int *integer_delay, double *fraction)
{
    split_delay(delay, t0_offset, fs, integer_delay, fraction);
}

static double dot3(const double *a, const double *b)
{
    double out = 0.0;
    for (int i = 0; i < 3; i++) out += a[i] * b[i];
    return out;
}

/* The same dot product in the value arithmetic's type (the link geometry). */
static real_t dot3_r(const real_t *a, const double *b)
{
    real_t out = 0.0;
    for (int i = 0; i < 3; i++) out += a[i] * (real_t)b[i];
    return out;
}

/* ---------------------------------------------------------------- projection */

/* LISAResponse::get_response CPU path, LISAResponse.cu:431-672 and :674-718.
 * input_in: complex128 as interleaved doubles [num_inputs][2]. y_gw: [6][num_delays], only
 * samples in [projections_start_ind, num_delays - projections_start_ind) are written.
 * nthreads > 1 splits the sample loop; every output element is still produced by exactly
 * the same operation sequence, so the result does not depend on nthreads. */
int oracle_get_response(real_t *y_gw, const double *t_data, const double *k, const double *u,
                        const double *v, double dt, int num_delays, const double *input_in,
                        int num_inputs, int order, double sampling_frequency, int buffer_integer,
                        const double *A_in, double deps, int num_A, const double *E_in,
                        int projections_start_ind, const oracle_orbits *orbits, int nthreads)
{
    (void)dt; (void)num_inputs; (void)buffer_integer; (void)num_A;
    if (orbits == NULL) return -1;
    const int point_count = order + 1;
    const int h = point_count / 2;
    const double t0_offset = t_data[0];
    const int P = projections_start_ind;
    if (nthreads < 1) nthreads = 1;

    for (int link_i = 0; link_i < ORACLE_NLINKS; link_i++) {
        const int sc_r = orbits->sc_r[link_i];
        const int sc_e = orbits->sc_e[link_i];
        const int link = orbits->links[link_i];
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
        for (int i = P; i < num_delays - P; i++) {
            double x_rec[3], x_em[3];
            real_t n[3];
            double t = t_data[i];
            double L = oracle_ltt(orbits, t, link);
            double t_rec = t;
            double t_em = t - L;
#if ORACLE_MUTANT == 4
            oracle_pos(orbits, t_rec, sc_e, x_rec);
            oracle_pos(orbits, t_em, sc_r, x_em);
#else
            oracle_pos(orbits, t_rec, sc_r, x_rec);
            oracle_pos(orbits, t_em, sc_e, x_em);
#endif
            real_t norm = 0.0;
            for (int c = 0; c < 3; c++) {
                real_t nt = (real_t)x_rec[c] - (real_t)x_em[c];
                n[c] = nt;
                norm += nt * nt;
            }
#ifdef ORACLE_WIDE
            norm = sqrtl(norm);
#else
            norm = sqrt(norm);
#endif
            for (int c = 0; c < 3; c++) n[c] = n[c] / norm;

            real_t u_dot_n = dot3_r(n, u);
            real_t v_dot_n = dot3_r(n, v);
            real_t xi_p = 0.5 * ((u_dot_n * u_dot_n) - (v_dot_n * v_dot_n));
            real_t xi_c = u_dot_n * v_dot_n;
#if ORACLE_MUTANT == 3
            xi_c = -xi_c;
#endif

            real_t k_dot_n = dot3_r(n, k);
            double k_dot_x_rec = dot3(k, x_rec);
            double k_dot_x_em = dot3(k, x_em);

#if ORACLE_MUTANT == 1
            double delay_rec = t_rec + k_dot_x_rec * ORACLE_C_INV;
#else
            double delay_rec = t_rec - k_dot_x_rec * ORACLE_C_INV;
#endif
            double delay_em = t_em - k_dot_x_em * ORACLE_C_INV;

            int d_rec, d_em;
            double e_rec, e_em;
            split_delay(delay_rec, t0_offset, sampling_frequency, &d_rec, &e_rec);
            split_delay(delay_em, t0_offset, sampling_frequency, &d_em, &e_em);

            real_t hp_rec, hc_rec, hp_em, hc_em;
            interp_cplx(&hp_rec, &hc_rec, input_in, h, d_rec, e_rec, A_in, deps, E_in);
            interp_cplx(&hp_em, &hc_em, input_in, h, d_em, e_em, A_in, deps, E_in);

            real_t pre_factor = 1. / (1. - k_dot_n);
            real_t large_factor = (hp_em - hp_rec) * xi_p + (hc_em - hc_rec) * xi_c;
            y_gw[link_i * num_delays + i] = pre_factor * large_factor;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------- TDI */

/* LISAResponse::get_tdi_delays CPU path, LISAResponse.cu:218-397 and :399-429.
 * delayed_links [num_channels][num_delays] is accumulated into (the caller zeroes it),
 * unit by unit in index order, exactly as the reference's `+=` does. */
int oracle_get_tdi_delays(real_t *delayed_links, const real_t *input_links, int num_inputs,
                          int num_delays, const double *t_arr, const int *unit_starts,
                          const int *unit_lengths, const int *tdi_base_link,
                          const int *tdi_link_combinations, const double *tdi_signs_in,
                          const int *channels, int num_units, int num_channels, int order,
                          double sampling_frequency, int buffer_integer, const double *A_in,
                          double deps, int num_A, const double *E_in, int tdi_start_ind,
                          const oracle_orbits *orbits, int nthreads)
{
    (void)num_channels; (void)buffer_integer; (void)num_A;
    if (orbits == NULL) return -1;
    const int point_count = order + 1;
    const int h = point_count / 2;
    const double t0_offset = t_arr[0];
    if (nthreads < 1) nthreads = 1;

    for (int unit_i = 0; unit_i < num_units; unit_i++) {
        if (oracle_link_row(orbits, tdi_base_link[unit_i]) < 0) return -2;
        for (int s = 0; s < unit_lengths[unit_i]; s++) {
            int l = tdi_link_combinations[unit_starts[unit_i] + s];
            if (l != -11 && oracle_link_row(orbits, l) < 0) return -2;
        }
    }

#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int i = tdi_start_ind; i < num_delays - tdi_start_ind; i++) {
        double t = t_arr[i];
        for (int unit_i = 0; unit_i < num_units; unit_i++) {
            int unit_start = unit_starts[unit_i];
            int unit_length = unit_lengths[unit_i];
            int base_row = oracle_link_row(orbits, tdi_base_link[unit_i]);
            int channel = channels[unit_i];
            double sign = tdi_signs_in[unit_i];
            double delay = t;
#if ORACLE_MUTANT == 2
            if (unit_length > 1) unit_length -= 1;
#endif
            for (int sub_i = 0; sub_i < unit_length; sub_i++) {
                int combination_link = tdi_link_combinations[unit_start + sub_i];
                if (combination_link != -11) delay -= oracle_ltt(orbits, t, combination_link);
            }
            int d;
            double e;
            split_delay(delay, t0_offset, sampling_frequency, &d, &e);
            real_t val = interp_single(&input_links[(int64_t)base_row * num_inputs], h, d, e, A_in,
                                       deps, E_in);
            val *= sign;
            delayed_links[(int64_t)channel * num_delays + i] += val;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------- AET */

/* XYZ -> AET as lisatools.utils.utility.AET computes it (call site response.py:665):
 * A = (Z - X)/sqrt(2), E = (X - 2Y + Z)/sqrt(6), T = (X + Y + Z)/sqrt(3). */
void oracle_aet(const real_t *X, const real_t *Y, const real_t *Z, real_t *A, real_t *E, real_t *T,
                int64_t n)
{
#ifdef ORACLE_WIDE
    const real_t s2 = sqrtl(2.0L), s6 = sqrtl(6.0L), s3 = sqrtl(3.0L);
#else
    const double s2 = sqrt(2.0), s6 = sqrt(6.0), s3 = sqrt(3.0);
#endif
    for (int64_t i = 0; i < n; i++) {
        real_t x = X[i], y = Y[i], z = Z[i];
#if ORACLE_MUTANT == 5
        A[i] = (x - z) / s2;
#else
        A[i] = (z - x) / s2;
#endif
        E[i] = (x - 2.0 * y + z) / s6;
        T[i] = (x + y + z) / s3;
    }
}

/* 1 in the wide build (liboracle_flr_wide.so), 0 in the FP64 one. */
int oracle_is_wide(void)
{
#ifdef ORACLE_WIDE
    return 1;
#else
    return 0;
#endif
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
