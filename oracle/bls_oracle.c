/*
 * bls_oracle.c -- CPU FP64 restatement of astropy's BoxLeastSquares "fast"
 * periodogram.  TEST INFRASTRUCTURE ONLY: the checker for the HIP path and
 * the CPU baseline timed by bench.py.  Nothing in astro_transit_gpu_amd/
 * links, loads or calls this file.
 *
 * What it restates.  The reference's oracle is astropy
 * BoxLeastSquares.power(method="fast")
 * (src/astrotransit_gpu/search/cpu_reference_bls.py:11-16), whose arithmetic
 * is astropy's compiled bls_impl (astropy 4.3.1 in this image; the reference
 * pins no version, pyproject.toml:13, and benchmarked 7.2.0).  Its C source
 * is not shipped, so the steps below are restated from astropy's documented
 * algorithm (bls/methods.py:98-143 gives the brute-force definition of every
 * field) and from the reference's transcription of the same binning
 * mechanics (src/astrotransit_gpu/search/kernels/vbls_v42_parity.cu:39-129,
 * docs/V42_PARITY_KERNEL_SPEC.md:32-52).  The details that only the compiled
 * code fixes (padding assignment, sign test, epsilon, loop order, rounding
 * of the likelihood expression) were pinned by comparing with outputs of
 * astropy's own bls_impl (fixtures under tests/golden, checked bit for bit
 * by tests/test_oracle.py).
 *
 * Per period P (bd = min(durations)/oversample):
 *   n_bins = ceil(P/bd) + oversample
 *   bin(t') = (int)((t' - P*floor(t'/P)) / bd) + 1       t' = t - min(t)
 *   bins[n_bins-os+j-1] = bins[j], j = 1..os  (assignment, after binning)
 *   inclusive cumulative sums over bins 0..n_bins
 *   for each duration k (m = round(d_k/bd)), each start n = 0..n_bins-m:
 *     W_in = C_w[n+m]-C_w[n], Y_in = C_y[n+m]-C_y[n]
 *     skip if W_in or W_out < DBL_EPSILON
 *     depth = Y_out/W_out - Y_in/W_in   accept if depth >= 0 and
 *     objective strictly greater than the best so far (first wins ties)
 *   transit phase = fmod(n*bd + 0.5*m*bd + min(t'), P)
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double wrap_into(double x, double period)
{
    return x - period * floor(x / period);
}

/* Per-period kernel of the restatement.  mean_y / mean_ivar are scratch of
 * at least n_bins+1 doubles. */
static void one_period(int64_t n, const double *t, const double *y,
                       const double *ivar, double min_t, double sum_y,
                       double sum_ivar, double period, double bin_duration,
                       int64_t n_durations, const double *durations,
                       int oversample, int use_likelihood, double *mean_y,
                       double *mean_ivar, double *o_power, double *o_depth,
                       double *o_depth_err, double *o_duration,
                       double *o_phase, double *o_depth_snr,
                       double *o_log_like)
{
    int n_bins = (int)ceil(period / bin_duration) + oversample;
    for (int b = 0; b <= n_bins; ++b) {
        mean_y[b] = 0.0;
        mean_ivar[b] = 0.0;
    }
    for (int64_t i = 0; i < n; ++i) {
        int ind = (int)(wrap_into(t[i] - min_t, period) / bin_duration) + 1;
        if (ind < 0) ind = 0;          /* unreachable for finite input */
        if (ind > n_bins) ind = n_bins;
        mean_y[ind] += y[i] * ivar[i];
        mean_ivar[ind] += ivar[i];
    }
    for (int j = 1, ind = n_bins - oversample; j <= oversample; ++j, ++ind) {
        mean_y[ind] = mean_y[j];
        mean_ivar[ind] = mean_ivar[j];
    }
    for (int b = 1; b <= n_bins; ++b) {
        mean_y[b] += mean_y[b - 1];
        mean_ivar[b] += mean_ivar[b - 1];
    }

    double best = -INFINITY;
    double b_depth = 0.0, b_err = 0.0, b_dur = 0.0, b_phase = 0.0;
    double b_snr = 0.0, b_ll = 0.0;
    for (int64_t k = 0; k < n_durations; ++k) {
        int dur = (int)round(durations[k] / bin_duration);
        int n_max = n_bins - dur;
        for (int s = 0; s <= n_max; ++s) {
            double y_in = mean_y[s + dur] - mean_y[s];
            double ivar_in = mean_ivar[s + dur] - mean_ivar[s];
            double y_out = sum_y - y_in;
            double ivar_out = sum_ivar - ivar_in;
            if (ivar_in < DBL_EPSILON || ivar_out < DBL_EPSILON) continue;
            y_in /= ivar_in;
            y_out /= ivar_out;
            double depth = y_out - y_in;
            double depth_err = sqrt(1.0 / ivar_in + 1.0 / ivar_out);
            double snr = depth / depth_err;
            double ll = 0.5 * ivar_in * depth * depth;
            double obj = use_likelihood ? ll : snr;
            if (depth >= 0.0 && obj > best) {
                best = obj;
                b_depth = depth;
                b_err = depth_err;
                b_snr = snr;
                b_ll = ll;
                b_dur = dur * bin_duration;
                b_phase = fmod(s * bin_duration + 0.5 * b_dur + min_t,
                               period);
            }
        }
    }
    *o_power = best;
    *o_depth = b_depth;
    *o_depth_err = b_err;
    *o_duration = b_dur;
    *o_phase = b_phase;
    *o_depth_snr = b_snr;
    *o_log_like = b_ll;
}

/*
 * atg_oracle_bls -- the restated bls_impl(t, y, ivar, period, duration,
 * oversample, use_likelihood) (astropy bls/methods.py:87-95).  ``t`` is
 * already relative (t - t_ref) and ``y`` already median-subtracted, as
 * astropy's power() passes them (bls/core.py:326-328).  Outputs the
 * transit *phase*; the caller adds t_ref.
 *
 * Returns 0 on success, 1 for a non-positive period, 2 for a duration
 * longer than the shortest period or non-positive, 3 for empty grids.
 * n_threads <= 0 uses the OpenMP default.  Periods are independent, so the
 * result does not depend on the thread count.
 */
int atg_oracle_bls(int64_t n, const double *t, const double *y,
                   const double *ivar, int64_t n_periods,
                   const double *periods, int64_t n_durations,
                   const double *durations, int oversample,
                   int use_likelihood, int n_threads, double *power,
                   double *depth, double *depth_err, double *duration,
                   double *transit_phase, double *depth_snr,
                   double *log_like)
{
    if (n_periods <= 0 || n_durations <= 0 || oversample < 1) return 3;
    double min_period = periods[0], max_period = periods[0];
    for (int64_t p = 1; p < n_periods; ++p) {
        if (periods[p] < min_period) min_period = periods[p];
        if (periods[p] > max_period) max_period = periods[p];
    }
    if (min_period < DBL_EPSILON) return 1;
    double min_dur = durations[0], max_dur = durations[0];
    for (int64_t k = 1; k < n_durations; ++k) {
        if (durations[k] < min_dur) min_dur = durations[k];
        if (durations[k] > max_dur) max_dur = durations[k];
    }
    if (max_dur > min_period || min_dur < DBL_EPSILON) return 2;

    double bin_duration = min_dur / (double)oversample;
    int max_n_bins = (int)ceil(max_period / bin_duration) + oversample;

    double min_t = INFINITY, sum_y = 0.0, sum_ivar = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        min_t = fmin(min_t, t[i]);
        sum_y += y[i] * ivar[i];
        sum_ivar += ivar[i];
    }

#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
    n_threads = 1;
#endif
    int failed = 0;
#pragma omp parallel num_threads(n_threads)
    {
        double *my = (double *)malloc(sizeof(double) * (size_t)(max_n_bins + 1));
        double *mi = (double *)malloc(sizeof(double) * (size_t)(max_n_bins + 1));
        if (!my || !mi) {
#pragma omp atomic write
            failed = 1;
        }
#pragma omp for schedule(dynamic, 8)
        for (int64_t p = 0; p < n_periods; ++p) {
            if (!my || !mi) continue;
            one_period(n, t, y, ivar, min_t, sum_y, sum_ivar, periods[p],
                       bin_duration, n_durations, durations, oversample,
                       use_likelihood, my, mi, power + p, depth + p,
                       depth_err + p, duration + p, transit_phase + p,
                       depth_snr + p, log_like + p);
        }
        free(my);
        free(mi);
    }
    return failed ? 4 : 0;
}

int atg_oracle_abi_version(void) { return 1; }


/*
 * Precision-study variants (what the GPU design avoids; no astropy
 * counterpart).  They restate the reference's single-precision V42 kernel
 * (SCALAR_T = float, src/astrotransit_gpu/search/kernels/vbls_v42_parity.cu:
 * 36-46 grid constants, 68-74 fold and binning, 78-84 padding and prefix
 * sums, 105-125 box scan; totals summed in float, v42_parity.py:66-67):
 *   fold32: t' = t - t_ref, the period, 1/P, bd and 1/bd in float and
 *           ph = t' - P*floor(t'*(1/P)), bin = (int)(ph*(1/bd)) + 1 in float;
 *   acc32:  the per-point products, the bins, the padding, the cumulative
 *           sums, the totals and the box scan in float.
 * With both flags 0 the result equals atg_oracle_bls (checked by the tests).
 */
static void one_period_variant(int64_t n, const double *t, const double *y,
                               const double *ivar, double sum_y,
                               double sum_ivar, float sy32, float si32,
                               double period, double bin_duration,
                               int64_t n_durations, const double *durations,
                               int oversample, int use_likelihood,
                               int fold32, int acc32, double *mean_y,
                               double *mean_ivar, float *fy, float *fi,
                               double *o_power, double *o_depth,
                               double *o_duration, double *o_phase)
{
    int n_bins = (int)ceil(period / bin_duration) + oversample;
    const float pf = (float)period, ipf = 1.0f / pf;
    const float bdf = (float)bin_duration, ibdf = 1.0f / bdf;
    if (fold32) n_bins = (int)ceilf(pf * ibdf) + oversample;
    for (int b = 0; b <= n_bins; ++b) {
        mean_y[b] = mean_ivar[b] = 0.0;
        fy[b] = fi[b] = 0.0f;
    }
    for (int64_t i = 0; i < n; ++i) {
        int ind;
        if (fold32) {
            const float tf = (float)t[i];
            const float ph = tf - pf * floorf(tf * ipf);
            ind = (int)(ph * ibdf) + 1;
        } else {
            ind = (int)(wrap_into(t[i], period) / bin_duration) + 1;
        }
        if (ind < 0) ind = 0;
        if (ind > n_bins) ind = n_bins;
        if (acc32) {
            const float wf = (float)ivar[i];
            fi[ind] += wf;
            fy[ind] += (float)y[i] * wf;
        } else {
            mean_y[ind] += y[i] * ivar[i];
            mean_ivar[ind] += ivar[i];
        }
    }
    for (int j = 1, ind = n_bins - oversample; j <= oversample; ++j, ++ind) {
        mean_y[ind] = mean_y[j];
        mean_ivar[ind] = mean_ivar[j];
        fy[ind] = fy[j];
        fi[ind] = fi[j];
    }
    for (int b = 1; b <= n_bins; ++b) {
        mean_y[b] += mean_y[b - 1];
        mean_ivar[b] += mean_ivar[b - 1];
        fy[b] += fy[b - 1];
        fi[b] += fi[b - 1];
    }
    const double bdv = fold32 ? (double)bdf : bin_duration;
    double best = -INFINITY, b_depth = 0.0, b_dur = 0.0, b_phase = 0.0;
    for (int64_t k = 0; k < n_durations; ++k) {
        const int dur = fold32 ? (int)roundf((float)durations[k] * ibdf)
                               : (int)round(durations[k] / bin_duration);
        for (int s = 0; s <= n_bins - dur; ++s) {
            double depth, obj;
            if (acc32) {
                float y_in = fy[s + dur] - fy[s];
                float ivar_in = fi[s + dur] - fi[s];
                float y_out = sy32 - y_in;
                float ivar_out = si32 - ivar_in;
                if (ivar_in < (float)DBL_EPSILON ||
                    ivar_out < (float)DBL_EPSILON)
                    continue;
                y_in /= ivar_in;
                y_out /= ivar_out;
                const float d = y_out - y_in;
                const float snr = d / sqrtf(1.0f / ivar_in + 1.0f / ivar_out);
                depth = d;
                obj = use_likelihood ? 0.5 * (double)(ivar_in * d * d)
                                     : (double)snr;
            } else {
                double y_in = mean_y[s + dur] - mean_y[s];
                double ivar_in = mean_ivar[s + dur] - mean_ivar[s];
                double y_out = sum_y - y_in;
                double ivar_out = sum_ivar - ivar_in;
                if (ivar_in < DBL_EPSILON || ivar_out < DBL_EPSILON) continue;
                y_in /= ivar_in;
                y_out /= ivar_out;
                depth = y_out - y_in;
                obj = use_likelihood
                          ? 0.5 * ivar_in * depth * depth
                          : depth / sqrt(1.0 / ivar_in + 1.0 / ivar_out);
            }
            if (depth >= 0.0 && obj > best) {
                best = obj;
                b_depth = depth;
                b_dur = dur * bdv;
                b_phase = fmod(s * bdv + 0.5 * b_dur, period);
            }
        }
    }
    *o_power = best;
    *o_depth = b_depth;
    *o_duration = b_dur;
    *o_phase = b_phase;
}

int atg_oracle_bls_variant(int64_t n, const double *t, const double *y,
                           const double *ivar, int64_t n_periods,
                           const double *periods, int64_t n_durations,
                           const double *durations, int oversample,
                           int use_likelihood, int fold32, int acc32,
                           int n_threads, double *power, double *depth,
                           double *duration, double *transit_phase)
{
    if (n_periods <= 0 || n_durations <= 0 || oversample < 1) return 3;
    double min_dur = durations[0], max_period = periods[0];
    for (int64_t k = 1; k < n_durations; ++k)
        if (durations[k] < min_dur) min_dur = durations[k];
    for (int64_t p = 1; p < n_periods; ++p)
        if (periods[p] > max_period) max_period = periods[p];
    double bin_duration = min_dur / (double)oversample;
    if (fold32) bin_duration = (double)((float)min_dur / (float)oversample);
    /* +2: the float bin count may exceed the double one by one */
    int max_n_bins = (int)ceil(max_period / bin_duration) + oversample + 2;
    double sum_y = 0.0, sum_ivar = 0.0;
    float sy32 = 0.0f, si32 = 0.0f;
    for (int64_t i = 0; i < n; ++i) {
        sum_y += y[i] * ivar[i];
        sum_ivar += ivar[i];
        sy32 += (float)y[i] * (float)ivar[i];
        si32 += (float)ivar[i];
    }
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
    n_threads = 1;
#endif
#pragma omp parallel num_threads(n_threads)
    {
        size_t nb1 = (size_t)(max_n_bins + 1);
        double *my = (double *)malloc(sizeof(double) * nb1);
        double *mi = (double *)malloc(sizeof(double) * nb1);
        float *fy = (float *)malloc(sizeof(float) * nb1);
        float *fi = (float *)malloc(sizeof(float) * nb1);
#pragma omp for schedule(dynamic, 8)
        for (int64_t p = 0; p < n_periods; ++p) {
            if (!my || !mi || !fy || !fi) continue;
            one_period_variant(n, t, y, ivar, sum_y, sum_ivar, sy32, si32,
                               periods[p], bin_duration, n_durations,
                               durations, oversample, use_likelihood, fold32,
                               acc32, my, mi, fy, fi, power + p, depth + p,
                               duration + p, transit_phase + p);
        }
        free(my);
        free(mi);
        free(fy);
        free(fi);
    }
    return 0;
}
