/*
 * atg_bls.h -- C ABI of the MI355X Box Least Squares engine
 * (libatg_bls.so, built for gfx950).
 *
 * One call runs the BLS periodogram over a ragged batch of light curves
 * that already sit in device memory:
 *   per light curve:  t (FP64), y and dy (FP32 or FP64), its own period
 *                     grid and duration set (CSR offsets);
 *   per period:       power, depth, depth_err, duration, transit_time,
 *                     depth_snr, log_likelihood (FP64);
 *   per light curve:  index of the best period and a status code.
 *
 * The results follow astropy's BoxLeastSquares.power(method="fast")
 * (the reference's oracle, src/astrotransit_gpu/search/cpu_reference_bls.py:
 * 11-16): t_ref = min(t), ivar = 1/dy^2, y - median(y), bins of
 * min(durations)/oversample, the likelihood or snr objective.
 *
 * Conventions: every call returns an int status (ATG_OK = 0, negative on
 * failure; atg_last_error() gives the message).  Nothing throws across the
 * ABI.  All device pointers belong to the caller; the library never frees
 * them.  Work is enqueued on the context's stream and is asynchronous
 * unless the function name ends in _sync.  A context is used by one host
 * thread at a time.
 */
#ifndef ATG_BLS_H
#define ATG_BLS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATG_ABI_VERSION 1

/* return codes */
#define ATG_OK 0
#define ATG_ERR_ARG (-1)      /* bad argument (null pointer, size, enum)   */
#define ATG_ERR_BUDGET (-2)   /* workspace would exceed the HBM budget     */
#define ATG_ERR_HIP (-3)      /* HIP runtime or launch failure             */
#define ATG_ERR_LDS (-4)      /* a period needs more bins than LDS holds   */
#define ATG_ERR_DEVICE (-5)   /* device is not gfx950 / not available      */

/* per-light-curve status written by the device */
#define ATG_LC_OK 0
#define ATG_LC_BAD_PERIOD 1   /* a period < DBL_EPSILON (astropy code 1)   */
#define ATG_LC_BAD_DURATION 2 /* max duration > min period, or a duration
                                 < DBL_EPSILON (astropy code 2)            */
#define ATG_LC_NAN 3          /* NaN in t or y                             */
#define ATG_LC_BAD_DY 4       /* dy <= 0 or NaN                            */
#define ATG_LC_TOO_MANY_BINS 5 /* a period's bins exceed the LDS budget
                                  (or 2^30 bins)                          */
#define ATG_LC_EMPTY 6        /* no points, periods or durations           */
#define ATG_LC_TOO_MANY_DURATIONS 7 /* more than 64 durations             */
#define ATG_LC_TOO_MANY_POINTS 8 /* more than 2^28 points in one curve     */

/* element types */
#define ATG_F32 1
#define ATG_F64 2

/* objective: astropy power(objective=...) */
#define ATG_OBJ_LIKELIHOOD 0
#define ATG_OBJ_SNR 1

/* mode: reference power(method=...) */
/* PARITY: the oracle's exact bin rule, int64 fixed-point bins, near-tied
 *   periods recomputed in the oracle's order: bit-reproducible, the best
 *   period and its fields equal to the oracle's.
 * FAST: fused-multiply-add fold without the bin-edge test, no
 *   recomputation, int64 bins (packed for uniform weights): reproducible,
 *   with a stated drift for points within ~1 ulp of a bin edge. */
#define ATG_MODE_PARITY 0
#define ATG_MODE_FAST 1

typedef struct atg_ctx atg_ctx;

/* A ragged batch of n_lc light curves; every pointer is a device pointer.
 * Light curve i owns points [lc_offsets[i], lc_offsets[i+1]), periods
 * [period_offsets[i], period_offsets[i+1]) and durations
 * [dur_offsets[i], dur_offsets[i+1]). */
typedef struct atg_bls_batch {
    int64_t n_lc;
    int64_t n_points;        /* total points   = lc_offsets[n_lc]          */
    int64_t n_periods;       /* total periods  = period_offsets[n_lc]      */
    int64_t n_durations;     /* total durations = dur_offsets[n_lc]        */
    const double *t;         /* [n_points] FP64 times                      */
    const void *y;           /* [n_points] flux, y_dtype                   */
    const void *dy;          /* [n_points] flux error, dy_dtype; NULL = 1  */
    int32_t y_dtype;         /* ATG_F32 / ATG_F64                          */
    int32_t dy_dtype;        /* ATG_F32 / ATG_F64 (ignored if dy == NULL)  */
    const int64_t *lc_offsets;     /* [n_lc + 1]                           */
    const double *periods;         /* [n_periods]                          */
    const int64_t *period_offsets; /* [n_lc + 1]                           */
    const double *durations;       /* [n_durations]                        */
    const int64_t *dur_offsets;    /* [n_lc + 1]                           */
    int32_t oversample;      /* bins per minimum duration (astropy: 10)    */
    int32_t objective;       /* ATG_OBJ_*                                  */
    int32_t mode;            /* ATG_MODE_*                                 */
    int32_t reserved;
} atg_bls_batch;

/* Caller-allocated device outputs aligned with the periods of the batch. */
typedef struct atg_bls_out {
    double *power;           /* [n_periods]                                */
    double *depth;
    double *depth_err;
    double *duration;
    double *transit_time;    /* absolute: phase + t_ref                    */
    double *depth_snr;
    double *log_likelihood;
    int64_t *best_index;     /* [n_lc] argmax of power within the curve's
                                grid (first maximum); may be NULL          */
    int32_t *status;         /* [n_lc] ATG_LC_*; may be NULL               */
} atg_bls_out;

/* Kernel tiling (chosen by the autotuner; 0 fields keep the current value).
 */
typedef struct atg_tuning {
    int32_t periods_per_item; /* periods per work item (1..64; a context
                                 starts at 0: min(24, max(4, periods /
                                 (8 * workgroups))) per batch)            */
    int32_t lds_bins;         /* LDS histogram: lds_bins * 16 bytes        */
    int32_t wg_per_cu;        /* persistent workgroups per CU              */
    int32_t reserved;
} atg_tuning;

/* ABI version (ATG_ABI_VERSION). */
int atg_abi_version(void);

/* Creates a context on `device` that enqueues on `stream` (a hipStream_t;
 * NULL = the null stream) and caps the library's own device workspace at
 * `hbm_budget_bytes` (0 = no cap).  Replaces the per-call CuPy set-up of the
 * reference wrappers (search/gpu_bls.py:18-40, search/v42_parity.py:7-25:
 * NVRTC compile and module cache): kernels here are compiled ahead of time.
 */
int atg_ctx_create(int device, void *stream, uint64_t hbm_budget_bytes,
                   atg_ctx **out);
void atg_ctx_destroy(atg_ctx *ctx);

/* Replaces the stream the context enqueues on. */
int atg_ctx_set_stream(atg_ctx *ctx, void *stream);

/* Pins (pin != 0) or releases one pin of (pin == 0) the context's
 * workspace; returns the number of pins held.  No counterpart in the
 * reference.  A HIP graph captured from this context's launches holds the
 * workspace's address: while pinned, a call that would need a larger
 * workspace fails with ATG_ERR_BUDGET instead of moving it.  Replays and
 * eager calls of one context share that workspace, so they must not
 * overlap (one stream, or a separate context for eager work). */
int atg_ctx_pin_workspace(atg_ctx *ctx, int pin);

/* Last error message of the context (never NULL). */
const char *atg_last_error(const atg_ctx *ctx);

/* Tiling of the main kernel. */
int atg_ctx_set_tuning(atg_ctx *ctx, const atg_tuning *tuning);
int atg_ctx_get_tuning(const atg_ctx *ctx, atg_tuning *tuning);

/* Device workspace a batch of this shape needs, in bytes. */
int atg_bls_workspace_size(const atg_ctx *ctx, const atg_bls_batch *batch,
                           uint64_t *bytes);

/* Runs the periodogram of the batch (asynchronous on the context stream).
 * Replaces the reference's kernel launches behind
 * BoxLeastSquaresGPU.power: run_vbls_exact_parity's
 * vbls_v42_parity_kernel launch (search/v42_parity.py:27-107, kernel
 * search/kernels/vbls_v42_parity.cu:13-31) for mode PARITY and
 * run_gpu_bls's bls_kernel launch (search/gpu_bls.py:42-144, kernel
 * search/kernels/bls.cu:12-27) for mode FAST, including their host-side
 * preparation (weights, totals, t_start) and argmax. */
int atg_bls_run(atg_ctx *ctx, const atg_bls_batch *batch,
                const atg_bls_out *out);

/* atg_bls_run followed by a synchronisation of the context stream.
 * Per-curve problems stay in out->status. */
int atg_bls_run_sync(atg_ctx *ctx, const atg_bls_batch *batch,
                     const atg_bls_out *out);

/* Size of the static device workspace of atg_bls_run_noalloc (per GPU,
 * held from the library's first context on that GPU until the process
 * ends). */
#define ATG_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* atg_bls_run without any device allocation: the same launch sequence and
 * results bit for bit, with the library's workspace in a static device
 * array of ATG_NOALLOC_SCRATCH_BYTES instead of memory the context
 * allocates.  No counterpart in the reference; it is for callers that
 * cannot allocate: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The caller still passes the batch and
 * the outputs.  A batch whose workspace (atg_bls_workspace_size) exceeds
 * the scratch returns ATG_ERR_BUDGET before anything is enqueued.  Every
 * call on one GPU shares that GPU's scratch (any context): calls that
 * overlap on one GPU race, so serialise them (one stream, or events).
 * No timing events are recorded for these calls. */
int atg_bls_run_noalloc(atg_ctx *ctx, const atg_bls_batch *batch,
                        const atg_bls_out *out);

/* Caller-allocated device outputs of a best-only run: one entry per light
 * curve, the fields of its best period (the period at the index
 * atg_bls_out.best_index would hold; NaN fields for a curve without
 * periods). */
typedef struct atg_bls_best {
    double *power;           /* [n_lc]                                     */
    double *depth;
    double *depth_err;
    double *duration;
    double *transit_time;    /* absolute: phase + t_ref                    */
    double *depth_snr;
    double *log_likelihood;
    double *period;          /* [n_lc] the best period itself              */
    int64_t *best_index;     /* [n_lc] within the curve's grid; may be NULL */
    int32_t *status;         /* [n_lc] ATG_LC_*; may be NULL               */
} atg_bls_best;

/* Best-only screening (asynchronous on the context stream).  Replaces
 * run_vbls_massive (search/vbls.py:62-188, kernels loaded at vbls.py:12-55
 * from search/kernels/vbls_blackwell_v39.cu or vbls.cu), which keeps
 * only each target's winner: the same launch sequence as atg_bls_run, with
 * the per-period arrays in the library's workspace (56 more bytes per
 * period) and the winners' fields gathered on the device, so that only
 * n_lc records leave it.  Every field equals the one atg_bls_run writes at
 * out->best_index. */
int atg_bls_run_best(atg_ctx *ctx, const atg_bls_batch *batch,
                     const atg_bls_best *best);

/* Device workspace atg_bls_run_best needs for a batch of this shape. */
int atg_bls_workspace_size_best(const atg_ctx *ctx,
                                const atg_bls_batch *batch, uint64_t *bytes);

/* Top-k peaks of every curve's periodogram (asynchronous on the context
 * stream).  Replaces get_top_k_candidates (search/gpu_bls.py:146-174) as
 * BoxLeastSquaresGPU.power uses it (search/api.py:113-123): the local
 * maxima of out->power (plateaus give their midpoint) thinned so that no
 * two kept peaks are closer than `distance` periods, highest power first
 * (scipy.signal.find_peaks(power, distance=distance)), then the k kept
 * peaks of highest power, in decreasing power.  A curve without a peak
 * gives its k highest powers instead.  Equal powers rank the later period
 * first (the order of the reference's reversed stable sorts; in the
 * no-peak case the final order puts the earlier period first).
 *   index_out[c*k + r]: period index of the r-th candidate of curve c,
 *                       relative to the curve's first period; -1 past
 *                       count_out[c] candidates.
 *   count_out[c]:       candidates found (0 for a curve whose status in
 *                       out->status, if given, is not ATG_LC_OK).
 * batch gives n_lc, n_periods and period_offsets; out gives power (and
 * optionally status), typically those of a preceding atg_bls_run.
 * 1 <= k <= 64, distance >= 1. */
int atg_bls_top_k(atg_ctx *ctx, const atg_bls_batch *batch,
                  const atg_bls_out *out, int k, int distance,
                  int64_t *index_out, int32_t *count_out);

/* Vetting statistics (see atg_bls_stats). */
#define ATG_STATS_FIELDS 16
#define ATG_STATS_OK 0
#define ATG_STATS_NO_TRANSIT 1        /* no point inside the box          */
#define ATG_STATS_TOO_MANY_TRANSITS 2 /* per-transit arrays not written   */

/* Vetting statistics of candidates (asynchronous on the context stream).
 * Replaces astropy BoxLeastSquares.compute_stats (astropy timeseries/
 * periodograms/bls/core.py:427-570), which the reference's vetting calls
 * on its BLS candidates, for n_cand candidates at once: candidate i
 * belongs to curve cand_lc[i] of `batch` (t, y, dy and lc_offsets are
 * used; raw t and y, as astropy) and has period[i], duration[i] and the
 * mid-transit time transit_time[i] (same time origin as t).
 * stats[i*ATG_STATS_FIELDS + f], f =
 *   0 depth, 1 its error, 2 depth_phased, 3 error, 4 depth_half,
 *   5 error, 6 depth_odd, 7 error, 8 depth_even, 9 error,
 *   10 harmonic_amplitude, 11 harmonic_delta_log_likelihood,
 *   12 weighted mean inside the box, 13 weighted mean outside (the two
 *   levels of astropy's model), 14 number of the first transit
 *   (round((t - T0)/P) of the earliest in-transit point), 15 number of
 *   transits n_tr (transit_times = T0 + P*(first + 0..n_tr-1)).
 * per_transit_count / per_transit_ll[i*max_transits + j], j < n_tr: points
 * and log likelihood of transit j.  status[i]: ATG_STATS_*.
 * 1 <= max_transits <= 4096.  All pointers are device pointers. */
int atg_bls_stats(atg_ctx *ctx, const atg_bls_batch *batch, int64_t n_cand,
                  const int32_t *cand_lc, const double *period,
                  const double *duration, const double *transit_time,
                  int max_transits, double *stats,
                  int32_t *per_transit_count, double *per_transit_ll,
                  int32_t *status);

/* Transit mask and box model at n times t (device pointers; either output
 * may be NULL): mask[i] = 1 inside the box of (period, duration,
 * transit_time), y_model[i] = y_in inside, y_out outside.  Replaces astropy
 * BoxLeastSquares.transit_mask (core.py:572-603) and, with y_in / y_out
 * from atg_bls_stats fields 12 / 13, BoxLeastSquares.model
 * (core.py:375-425). */
int atg_bls_model(atg_ctx *ctx, int64_t n, const double *t, double period,
                  double duration, double transit_time, double y_in,
                  double y_out, uint8_t *mask, double *y_model);

/* Average duration (ms) of the periodogram stage (the main kernel, the
 * phase-tile kernel and the result-field pass that follow it) over the runs
 * since the last reset,
 * timed with HIP events on the context stream; resets when `reset` != 0.
 * Returns the number of timed runs (0 = none). */
int atg_timing_main_kernel(atg_ctx *ctx, int enable, int reset,
                           double *avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* ATG_BLS_H */
