/*
 * flr.h -- C ABI of the MI355X (gfx950) LISA time-domain response library, libflr_hip.so.
 *
 * The library computes the fastlisaresponse hot path on one GPU:
 *   (a) orbit interpolation onto the data grid,
 *   (b) the 6-link projections y_ij from order-(2h-1) Lagrange fractional-delay stencils of h+/hx,
 *   (c) 1st/2nd-generation (or custom) TDI delay chains -> X, Y, Z, optionally fused into A, E, T.
 *
 * All array arguments are plain pointers with explicit sizes; no framework types cross the
 * boundary. Arrays read by kernels (signals, projections, outputs, t arrays) must be device
 * memory. Small configuration arrays (orbit tables, sky vectors, Lagrange tables, TDI tables)
 * may be host or device memory: the library copies them.
 *
 * Every entry point returns FLR_OK (0) or a nonzero status; flr_last_error() then holds a
 * message for the calling thread. Stream arguments are hipStream_t passed as void* (NULL is
 * the null stream). Handles are not safe for concurrent use from several host threads.
 */
#ifndef FLR_H
#define FLR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define FLR_OK 0
#define FLR_ERR_INVALID_ARGUMENT 1 /* bad size, null pointer, unknown link number, ... */
#define FLR_ERR_HIP 2              /* a HIP runtime call failed */
#define FLR_ERR_WINDOW 3           /* a stencil reached outside the staged sample window */
#define FLR_ERR_NO_ORBITS 4        /* response handle without orbit information */
#define FLR_ERR_UNSUPPORTED 5      /* configuration outside the compiled kernels' limits */

/* Flags for the stream-ordered entry points flr_projection() and flr_tdi(). */
#define FLR_FLAG_EXACT 0x1      /* reference stencil operation order: bitwise equal to the FP64 CPU path */
#define FLR_FLAG_T_UNIFORM 0x2  /* synthesize t_i = t0 + i*dt in-kernel instead of reading t_data */
#define FLR_FLAG_ZERO_EDGES 0x4 /* also write 0 to samples outside the computed range */
#define FLR_FLAG_OVERWRITE 0x8  /* TDI: store the sums instead of adding them to the output */
#define FLR_FLAG_AET 0x10       /* TDI: store A, E, T (implies OVERWRITE) instead of X, Y, Z */
#define FLR_FLAG_MIXED 0x20     /* fused batches: complex64 input (h / scale), float32 output */
#define FLR_FLAG_Q31 0x40       /* with FLR_FLAG_MIXED: input is int32 pairs, h / scale in 32-bit fixed point */
#define FLR_FLAG_C128 0x80      /* with FLR_FLAG_MIXED: input is complex128 h / scale (no storage error) */

typedef struct flr_orbits flr_orbits;
typedef struct flr_response flr_response;

/* Message describing the last failing call on this host thread ("" if none). */
const char *flr_last_error(void);

/* Library version string. */
const char *flr_version(void);

/* Number of visible HIP devices (0 when no GPU is present); does not initialise a context. */
int flr_device_count(int *count);

/* ------------------------------------------------------------------ orbits
 * Replaces OrbitsWrap{CPU,GPU}_responselisa(dt, N, n_arr, ltt_arr, x_arr, links, sc_r, sc_e,
 * armlength): src/fastlisaresponse/cutils/binding_flr.hpp:35-53 (bound at binding_flr.cxx:127-128),
 * which wraps the lisatools Orbits device class.
 *   dt       orbit grid spacing (s); node k sits at t = k*dt
 *   N        number of nodes (>= 2)
 *   n_arr    [N][6][3] link unit vectors (length-checked and stored; the response never reads it,
 *            as in the reference, which recomputes n)
 *   ltt_arr  [N][6] light travel times (s), column r belongs to links[r]
 *   x_arr    [N][3][3] spacecraft positions (m), spacecraft 1, 2, 3
 *   links    [6] link numbers ij (receiver i, emitter j), a permutation of 12,23,31,13,32,21
 *   sc_r     [6] receiver spacecraft of each link (1..3); sc_e [6] emitter spacecraft
 * The tables are copied to device memory, packed per node with precomputed slopes. */
int flr_orbits_create(double dt, int N, const double *n_arr, const double *ltt_arr,
                      const double *x_arr, const int32_t *links, const int32_t *sc_r,
                      const int32_t *sc_e, double armlength, flr_orbits **out);
void flr_orbits_destroy(flr_orbits *orbits);

/* Max |x| over all nodes and spacecraft (m) and max light travel time (s): sizes the halos. */
int flr_orbits_extent(const flr_orbits *orbits, double *max_radius, double *max_ltt);

/* Light travel time of `link` and position of spacecraft `sc` (1..3) at the `n` device times
 * t[] (device outputs ltt_out[n], pos_out[n][3]); exposes the device orbit interpolation. */
int flr_orbits_eval(const flr_orbits *orbits, const double *t, int n, int link, int sc,
                    double *ltt_out, double *pos_out, void *stream);

/* ------------------------------------------------------------------ response
 * Replaces LISAResponseWrap{CPU,GPU}(orbits): binding_flr.hpp:81-85 / binding_flr.cxx:99-100.
 * The response handle borrows `orbits`, which must outlive it. */
int flr_response_create(flr_orbits *orbits, flr_response **out);
void flr_response_destroy(flr_response *response);

/* Drop-in projection. Replaces LISAResponseWrap::get_response_wrap (binding_flr.cxx:38-58) ->
 * LISAResponse::get_response (src/fastlisaresponse/cutils/LISAResponse.cu:674-718), same
 * argument order and meaning:
 *   y_gw [6][num_delays] device output; samples [projections_start_ind, num_delays -
 *        projections_start_ind) of each link row are written, the rest is left as it is
 *   t_data [num_delays] device sample times; t_data[0] is the time origin of input_in
 *   k_in, u_in, v_in [3] sky propagation and polarisation basis
 *   input_in [num_inputs] complex128 (interleaved re, im doubles) h+ + i hx, device
 *   A_in [num_A], E_in [(order+1)/2] Lagrange tables (pyResponseTDI._fill_A_E)
 * dt, num_inputs and buffer_integer are accepted for interface parity; buffer_integer is unused
 * because the staged window is sized from the orbit extent. Synchronous, like the reference:
 * returns after the kernel finished and reports FLR_ERR_WINDOW if a stencil left the window. */
int flr_get_response(flr_response *response, double *y_gw, const double *t_data, const double *k_in,
                     const double *u_in, const double *v_in, double dt, int num_delays,
                     const double *input_in, int num_inputs, int order,
                     double sampling_frequency, int buffer_integer, const double *A_in, double deps,
                     int num_A, const double *E_in, int projections_start_ind, void *stream);

/* Drop-in TDI. Replaces LISAResponseWrap::get_tdi_delays_wrap (binding_flr.cxx:16-36) ->
 * LISAResponse::get_tdi_delays (LISAResponse.cu:399-429), same argument order and meaning:
 *   delayed_links [num_channels][num_delays] device output, ACCUMULATED into like the reference
 *   input_links [6][num_inputs] device projections (rows in the orbits' link order)
 *   t_arr [num_delays] device times; unit_starts/unit_lengths [num_units]; tdi_base_link,
 *   tdi_signs_in, channels [num_units]; tdi_link_combinations [sum(unit_lengths)] with -11
 *   marking an undelayed unit (pyResponseTDI.tdi_combinations setter, response.py:372-419)
 * Units of one channel are summed in index order; no atomics, so results are deterministic. */
int flr_get_tdi_delays(flr_response *response, double *delayed_links, const double *input_links,
                       int num_inputs, int num_delays, const double *t_arr,
                       const int32_t *unit_starts, const int32_t *unit_lengths,
                       const int32_t *tdi_base_link, const int32_t *tdi_link_combinations,
                       const double *tdi_signs_in, const int32_t *channels, int num_units,
                       int num_channels, int order, double sampling_frequency, int buffer_integer,
                       const double *A_in, double deps, int num_A, const double *E_in,
                       int tdi_start_ind, void *stream);

/* ------------------------------------------------------------------ stream-ordered entry points
 * No counterpart in the reference (whose launchers allocate, copy and synchronise per call):
 * these keep the tables resident in the handle and only enqueue kernels, for callers that run
 * many responses back to back or capture them into a HIP graph. */

/* Upload the Lagrange tables once (host or device pointers). */
int flr_set_lagrange(flr_response *response, int order, const double *A_in, double deps, int num_A,
                     const double *E_in);

/* Upload a TDI combination table once (same arrays as flr_get_tdi_delays). */
int flr_set_tdi(flr_response *response, const int32_t *unit_starts, const int32_t *unit_lengths,
                const int32_t *tdi_base_link, const int32_t *tdi_link_combinations,
                const double *tdi_signs_in, const int32_t *channels, int num_units,
                int num_channels);

/* Projection over num_pts samples: y_gw [6][num_pts] from input_in [num_pts] complex128.
 * With FLR_FLAG_T_UNIFORM, t_data may be NULL and t_i = t0 + i*dt (bitwise numpy's
 * t0 + arange(n)*dt); otherwise t_data [num_pts] is read and t0 is ignored. k, u, v are host
 * arrays of 3. Errors found on the device are latched; read them with flr_check_errors(). */
int flr_projection(flr_response *response, double *y_gw, const double *input_in, int num_pts,
                   const double *t_data, double t0, double dt, double sampling_frequency,
                   const double *k, const double *u, const double *v, int projections_start_ind,
                   int flags, void *stream);

/* TDI over num_pts samples from y_gw [6][num_pts] into out [3][num_pts] (X, Y, Z, or A, E, T
 * with FLR_FLAG_AET). Same time convention as flr_projection. */
int flr_tdi(flr_response *response, double *out, const double *y_gw, int num_pts,
            const double *t_arr, double t0, double dt, double sampling_frequency, int tdi_start_ind,
            int flags, void *stream);

/* Synchronise `stream`, then read and clear the handle's latched device error word. */
int flr_check_errors(flr_response *response, void *stream);

/* ------------------------------------------------------------------ batches
 * No counterpart in the reference (one source per call there): many sources of uneven length
 * run as one persistent projection launch and one persistent TDI launch over a flattened
 * (source, tile) list. Sources are packed back to back: source s occupies samples
 * [offset, offset + num_pts) of h_all (complex128), rows [6*offset, 6*(offset + num_pts)) of
 * y_all laid out [6][num_pts], and [3*offset, 3*(offset + num_pts)) of out_all ([3][num_pts]).
 * Sample times are t0 + i*dt per source. */
typedef struct {
    int64_t offset;                /* first sample of the source in the packed arrays */
    int32_t num_pts;               /* samples of the source */
    int32_t projections_start_ind; /* as pyResponseTDI.get_projections computes it */
    int32_t tdi_start_ind;
    int32_t reserved;
    double t0;                     /* time of the source's first sample (s) */
    double k[3], u[3], v[3];       /* sky propagation and polarisation basis */
    double scale;                  /* fused batches: outputs are scale * response of the stored
                                      input (0 means 1); ignored by flr_batch_create */
} flr_source;

typedef struct flr_batch flr_batch;

/* Validate the sources, upload their tables and bind the packed device arrays. `flags` is fixed
 * for the batch (FLR_FLAG_EXACT, FLR_FLAG_ZERO_EDGES, FLR_FLAG_AET, FLR_FLAG_OVERWRITE).
 * Allocates device memory once here; flr_batch_run never allocates (graph-capturable). */
int flr_batch_create(flr_response *response, const flr_source *sources, int num_sources,
                     const double *h_all, double *y_all, double *out_all, double dt,
                     double sampling_frequency, int flags, flr_batch **out);
/* Fused batch: no counterpart in the reference. One launch computes projection and TDI tile
 * by tile with the six projections kept on chip (y is never written to memory), for order 25.
 * FLR_FLAG_EXACT selects the exact stencils (FP64 only; then the outputs are bitwise those of
 * flr_batch_create + flr_batch_run), otherwise the fast form. With FLR_FLAG_MIXED, h_all holds
 * complex64 samples h / scale (scale per source) and out_all float32 outputs; adding
 * FLR_FLAG_Q31, h_all holds int32 pairs (re, im) = h / scale in fixed point (the same 8 bytes,
 * 2^-31 full-scale steps); adding FLR_FLAG_C128 instead, h_all holds complex128 h / scale
 * (16 bytes, no storage error; outputs still float32); without FLR_FLAG_MIXED h_all is
 * complex128 and out_all double. Outputs are overwritten (X, Y, Z, or A, E, T with
 * FLR_FLAG_AET); FLR_FLAG_ZERO_EDGES also zeroes samples outside [tdi_start_ind,
 * num_pts - tdi_start_ind). Run and destroy it with flr_batch_run / flr_batch_destroy. */
int flr_fused_create(flr_response *response, const flr_source *sources, int num_sources,
                     const void *h_all, void *out_all, double dt, double sampling_frequency,
                     int flags, flr_batch **out);
/* Fused batch with room for `capacity` sources and none bound yet (device and pinned host
 * tables are allocated here, once). Same flags as flr_fused_create. For serving loops, which
 * must not allocate or synchronise per request. */
int flr_fused_reserve(flr_response *response, int capacity, double dt, double sampling_frequency,
                      int flags, flr_batch **out);
/* Bind num_sources <= capacity sources (packed as for flr_fused_create) to a reserved fused
 * batch: validates them on the host, then uploads the source table on `stream` without
 * blocking (from pinned staging; a bind first waits on the host for the previous bind's upload,
 * and `stream` waits on the device for the batch's last flr_batch_run, whatever stream it ran
 * on). Runs replayed from a captured graph are not tracked: order those before a bind. The
 * tiling follows the response's current tables; flr_batch_run re-tiles the bound sources if
 * the tables change it later (a TDI table with longer delay chains shortens the tiles). */
int flr_fused_bind(flr_batch *batch, const flr_source *sources, int num_sources,
                   const void *h_all, void *out_all, void *stream);
/* Enqueue the batch's launches on `stream` (projection then TDI, or the one fused launch). */
int flr_batch_run(flr_batch *batch, void *stream);
/* Fused inner products against a data series: no counterpart in the reference, where the
 * response is written out and utils/utility.py:16-82 (get_overlap) or a likelihood reads it
 * back. Runs the batch's fused launch without storing its outputs: every output h (FP64, before
 * the float32 rounding) meets data_all at the same place -- data_all is laid out as the bound
 * out_all and has its element type (float32 for the mixed forms) -- and per source s and output
 * channel c, over [tdi_start_ind, num_pts - tdi_start_ind), sums[(s*3 + c)*4 + k] (device
 * memory, double) receives k = 0: sum d h, 1: sum h^2, 2: sum h, 3: sum (-1)^(i - tdi_start_ind)
 * h. From these the white-noise inner products and log-likelihood follow, and the real part of
 * get_overlap's rfft overlap by Parseval (the DC and Nyquist bins from sums 2 and 3). Mixed
 * fused batches (FLR_FLAG_MIXED) with a stock TDI table at order 25 (FLR_ERR_UNSUPPORTED
 * otherwise); deterministic (fixed reduction order). The per-tile sums live in a device buffer
 * of the batch, allocated at the first call and when the batch's tile count grows (so the
 * first call must not be under stream capture); later calls allocate nothing. */
int flr_fused_inner(flr_batch *batch, const void *data_all, double *sums, void *stream);
/* Launch shape of a batch (no counterpart in the reference): tile_samples window samples per
 * workgroup tile (fused mixed: 256, 512 or 1024; fused FP64: 512; unfused: 256; 0 keeps the
 * current; a fused tile outputs tile_samples minus the TDI delay halos, so that its projection
 * window is whole rounds of the workgroup) and grid_per_cu persistent workgroups per CU
 * (1..64; 0 keeps the current, default 32). Results do not depend on either. Call it when no
 * launch of the batch is in flight; the autotuner (fastlisaresponse_amd.autotune) times the
 * candidates and keeps the fastest. */
int flr_batch_set_launch(flr_batch *batch, int tile_samples, int grid_per_cu);
/* Total samples of the batch. */
int64_t flr_batch_samples(const flr_batch *batch);
void flr_batch_destroy(flr_batch *batch);

/* Allocation-free fused run. No counterpart in the reference: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Computes what flr_fused_create(response, sources, num_sources, h_all, out_all, dt,
 * sampling_frequency, flags) followed by flr_batch_run would (same kernels, same bits), with the
 * launch shape of flr_batch_set_launch(tile_samples, grid_per_cu) (0, 0: the defaults), but
 * allocates nothing: the source table, its tile prefix and the composite TDI filters go to one
 * fixed device array per GPU (under 512 KiB, held until the process ends), and the source
 * records reach it as kernel arguments. At most 512 sources per call (FLR_ERR_UNSUPPORTED
 * beyond). Everything is enqueued on `stream`; device errors are latched as for the other
 * stream-ordered calls (flr_check_errors). Calls on one GPU share that array, so calls that
 * overlap on one GPU (concurrent streams or host threads) race: order them. */
int flr_fused_run_static(flr_response *response, const flr_source *sources, int num_sources,
                         const void *h_all, void *out_all, double dt, double sampling_frequency,
                         int flags, int tile_samples, int grid_per_cu, void *stream);

/* ------------------------------------------------------------------ waveforms
 * No counterpart in the reference library (its tests generate galactic binaries with numpy,
 * tests/test_fastlisaresponse.py:41-63): the same generator on the device, writing straight into
 * a packed response input. One record per source, in DEVICE memory, sorted by offset. */
typedef struct {
    int64_t offset;       /* first sample of the source in the packed output */
    int32_t num_pts;
    int32_t reserved;
    double amplitude;     /* A */
    double f0, fdot, fddot; /* Hz, Hz/s, Hz/s^2 */
    double cos_iota;      /* cos(inclination) */
    double cos2psi, sin2psi; /* polarisation rotation */
    double phi0;          /* phase at t = 0 */
    double t0;            /* time of the first sample (s) */
    double scale;         /* complex64 / q31 formats store h / scale (q31: times 2^31 - 1) */
} flr_gb_source;

#define FLR_GB_COMPLEX128 0  /* h */
#define FLR_GB_COMPLEX64 1   /* h / scale as complex64 */
#define FLR_GB_Q31 2         /* h / scale as int32 pairs, full scale 2^31 - 1 */

/* Generate num_sources galactic binaries (t = t0 + k*dt, k < num_pts) into `out` (packed, of
 * total_samples entries in `format`) on `stream`. Allocates nothing; samples not covered by a
 * source are left untouched. */
int flr_gb_generate(const flr_gb_source *sources, int num_sources, int64_t total_samples,
                    double dt, int format, void *out, void *stream);

/* Stencil form used by the drop-in entry points flr_get_response / flr_get_tdi_delays:
 * nonzero (default) the reference operation order, bitwise equal to the FP64 CPU path;
 * 0 the faster reciprocal-free form (equal to ~1e-15 relative per stencil). */
int flr_response_set_exact(flr_response *response, int exact);

#ifdef __cplusplus
}
#endif

#endif /* FLR_H */
