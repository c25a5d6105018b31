// atg_bls.hip -- Box Least Squares periodogram for MI355X (gfx950, wave64).
//
// Pipeline of atg_bls_run (all on the context's HIP stream):
//   k_prep     one workgroup per light curve: t_ref = min(t), NaN / dy /
//              grid checks, time order, exact median of y (radix select on
//              the order-preserving key of the FP64 value), fixed-point
//              scales, the lane-major point workspace (t' = t - t_ref in
//              FP64, quantized weight and weight*flux as int64, or one packed
//              int64 for uniform weights), point-order copies for the phase
//              tiles, and in parity mode the oracle's FP64 totals.
//   k_items    per-queue work-item offsets (curve c in queue c mod 8, one
//              queue per XCD) and the queue resets.
//   k_main     persistent workgroups pull (curve, period tile) items from
//              their XCD's queue, then steal.  Up to 8 periods share a pass
//              over the points: each lane folds its contiguous run of points
//              in FP64 with the oracle's exact bin rule and adds each point
//              to the LDS histogram with 64-bit integer atomics
//              (order-independent, so bit-reproducible).  Then per period:
//              padding copy, int64 prefix scan, conversion to FP64, sliding
//              box scan over every duration, argmax with the oracle's
//              first-wins order, the 7 output fields.  Periods whose bins
//              exceed the histogram take phase tiles (big_period).
//   k_cands    parity mode: the periods that can hold the oracle's maximum
//              (near-ties within the int64 path's error bound) into a queue.
//   k_exact    recomputes the queued periods in the oracle's operation order.
//   k_final    per-curve argmax of power and status (and, for a best-only
//              run, the winner's fields).
// atg_bls_run_best runs the same sequence with the per-period arrays in the
// workspace and returns only the winners.
// Further entry points: atg_bls_top_k (k_topk, peaks of each periodogram),
// atg_bls_stats (k_stats, vetting statistics per candidate), atg_bls_model
// (k_model, transit mask and box model).
//
// Semantics follow astropy BoxLeastSquares.power(method="fast") (the
// reference's oracle, src/astrotransit_gpu/search/cpu_reference_bls.py:
// 11-16); the binning mechanics are those the reference transcribes in
// src/astrotransit_gpu/search/kernels/vbls_v42_parity.cu:39-129.
//
// Compiled with -ffp-contract=off: every multiply and add that mirrors the
// oracle is rounded separately; fused multiply-adds appear only where fma()
// is written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "atg_bls.h"

namespace {

constexpr int NT = 256;      // threads per workgroup = 4 wave64
constexpr int NW = NT / 64;  // waves per workgroup
constexpr int SUBMAX = 8;    // periods folded together in one pass
constexpr int DMAX = 64;     // durations per light curve
// k_exact: EXW waves per workgroup, whose LDS holds EXACT_BINS FP64 bins
// (at most; the tuning's lds_bins caps it), EXW regions of equal size
constexpr int EXW = 4;
constexpr int EXACT_BINS = 1920;
constexpr int exact_total(int lds_bins) {
    return (lds_bins < EXACT_BINS ? lds_bins : EXACT_BINS) & ~(EXW - 1);
}
constexpr int NQ = 8;        // work queues = XCDs of an MI355X
constexpr int EHIST = 1075 + 1024;  // k_prep's exponent histogram
// entries past a period's prefix arrays that k_main's box scan reads (the
// starts of a duration's last stride beyond its final window)
constexpr int PADN = NT - 1;
// bins of k_main's histogram in counted form (12 B each: an int64 flux sum
// and an int32 point count) in the LDS of lds_bins 16-byte bins
constexpr int counted_cap(int lds_bins) { return lds_bins * 4 / 3; }

struct LcParams {
    double t_ref;     // min(t)
    double y_med;     // median(y)
    double bd;        // bin duration = min(durations) / oversample
    double w_unit;    // 2^-sw: value of one weight quantum
    double y_unit;    // 2^-sy: value of one weight*flux quantum
    double W_tot;     // total weight (real units, from the quantized sum)
    double Y_tot;     // total weight*flux
    long long Wq_tot; // quantized totals
    long long Yq_tot;
    long long ws_off; // first slot of the curve in the point workspace
    int L;            // rows of the lane-major layout = ceil(n / NT)
    int n;            // points
    int status;       // ATG_LC_*
    int n_dur;
    double tol_x;     // near-integer tolerance of phase/bd (uniform bound)
    int sorted;       // t non-decreasing (enables the ranged phase tiles)
    int lc_self;      // index of the curve
    int natural;      // natural-order copies of the points were written
    int packed;       // uniform weights: one packed int64 per bin
    int ybits;        // packed bin = count << ybits + weight*flux quantum
    long long wq_one; // the common quantized weight of a packed curve
    int uniform;      // every point has the same weight (wq_one)
    double t_pad;     // t' of the zero-weight padding slots
};

struct Batch {
    long long n_lc;
    const double *t;
    const void *y;
    const void *dy;
    int y_dt, dy_dt;
    const long long *lc_off;
    const double *periods;
    const long long *per_off;
    const double *durs;
    const long long *dur_off;
    int oversample, objective, mode;
};

struct Out {
    double *power, *depth, *depth_err, *duration, *transit_time, *depth_snr,
        *log_like;
    long long *best_index;
    int *status;
};

struct Work {
    LcParams *prm;
    double2 *seq;         // [n_lc] parity mode: the oracle's totals, summed
                          // in FP64 in point order (sum of y*ivar, of ivar)
    int *lc_flags;        // per-curve flags raised by k_main (LDS overflow)
    long long *item_off;  // [n_lc] first item of the curve in its queue
    long long *qtot;      // [NQ] items per queue
    unsigned int *qnext;  // [NQ] queue heads
    unsigned int *counter;
    double *ws_t;         // lane-major t'
    long long *ws_w;      // lane-major quantized ivar (packed curves: the
                          // bin increment (1 << ybits) + ws_y, 0 for padding)
    long long *ws_y;      // lane-major quantized ivar * (y - median)
    // the same values in point order, written only for time-sorted curves
    // with periods whose bins exceed the LDS histogram: t' and the
    // quantized ivar * (y - median) side by side in 16-byte records (slot
    // i: nat_ty[2i] = t', nat_ty[2i + 1] = its bits), so that one 16-byte
    // load brings a point; the quantized ivar apart (read only for
    // varying weights)
    double *nat_ty;
    long long *nat_w;
    unsigned int *cand_count;  // periods queued for refinement
    unsigned int *cand_next;   // refinement queue head
    // periods left to k_big (bins exceed LDS), in NQ lists, one per
    // queue: list q holds at most the periods of queue q's curves, from
    // big_base[q] on
    long long *big_base;       // [NQ]
    unsigned int *big_count;   // [NQ] entries of each list
    unsigned int *big_next;    // [NQ] k_big's heads
    long long *big_p;          // [n_periods] those periods ...
    int *big_lc;               // ... and their curves
    int *cand_lc;         // [n_periods] curve of each queued period
    long long *cand_p;    // [n_periods] its global period index
};

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ double load_val(const void *p, int dt,
                                           long long i) {
    return dt == ATG_F32 ? (double)((const float *)p)[i]
                         : ((const double *)p)[i];
}

// The raw inputs with their types fixed at compile time: y FP32 or FP64
// (YT), dy absent (DT 0), FP32 or FP64.  The kernels that stream a curve's
// raw points (k_prep, k_exact) are instantiated per type pair, so that their
// loads carry no branch: a branch around each load makes the compiler wait
// for every load before the next, which serialises the memory latency.
template <int YT, int DT>
struct Raw {
    static constexpr int Y_T = YT;
    static constexpr bool HAS_DY = DT != 0;
    static __device__ __forceinline__ double y(const Batch &b, long long i) {
        if (YT == ATG_F32) return (double)((const float *)b.y)[i];
        return ((const double *)b.y)[i];
    }
    static __device__ __forceinline__ double dy(const Batch &b, long long i) {
        if (DT == ATG_F32) return (double)((const float *)b.dy)[i];
        if (DT == ATG_F64) return ((const double *)b.dy)[i];
        return 1.0;
    }
};

// Calls f(Raw<y type, dy type>()) for the batch's input types (host side).
template <class F>
void with_raw_types(const Batch &b, F f) {
    const int dt = b.dy ? b.dy_dt : 0;
    if (b.y_dt == ATG_F32) {
        if (dt == 0) f(Raw<ATG_F32, 0>());
        else if (dt == ATG_F32) f(Raw<ATG_F32, ATG_F32>());
        else f(Raw<ATG_F32, ATG_F64>());
    } else {
        if (dt == 0) f(Raw<ATG_F64, 0>());
        else if (dt == ATG_F32) f(Raw<ATG_F64, ATG_F32>());
        else f(Raw<ATG_F64, ATG_F64>());
    }
}

// Value of lane j (uniform j) of an FP64 register, as a wave-uniform value.
__device__ __forceinline__ double readlane_d(double v, int j) {
    const long long x = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)x, j);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(x >> 32), j);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) |
                                            lo));
}

// (double)v, exactly, for |v| < 2^51: the bits of 1.5 * 2^52 plus v are
// the FP64 value 1.5 * 2^52 + v (ulp 1 there), and subtracting 1.5 * 2^52
// leaves v (one integer add and one FP64 subtraction instead of the
// two-word conversion).
__device__ __forceinline__ double small_i64_to_f64(long long v) {
    return __longlong_as_double(v + 0x4338000000000000ll) - 0x1.8p52;
}

// Order-preserving unsigned key of an FP64 value (non-NaN).
__device__ __forceinline__ unsigned long long okey(double v) {
    unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double ikey(unsigned long long k) {
    unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}

template <typename T, typename Op>
__device__ T block_reduce(T v, Op op, T *scratch) {
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) scratch[w] = v;
    __syncthreads();
    T r = scratch[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) r = op(r, scratch[i]);
    return r;
}

// Exclusive scan over the workgroup in thread order; *total gets the sum.
template <typename T>
__device__ T block_excl_scan(T v, T *scratch, T *total) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    T inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T u = __shfl_up(inc, o);
        if (l >= o) inc += u;
    }
    __syncthreads();
    if (l == 63) scratch[w] = inc;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        T s = scratch[i];
        if (i < w) off += s;
        tot += s;
    }
    *total = tot;
    return off + inc - v;
}

// One step of the int64 DPP scan below: v + (the value CTRL names), 0 for
// lanes outside ROW_MASK or reading past their row.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ long long dpp_add_i64(long long v) {
    const int lo = (int)(unsigned long long)v;
    const int hi = (int)((unsigned long long)v >> 32);
    const unsigned int l2 =
        (unsigned int)__builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xF, true);
    const unsigned int h2 =
        (unsigned int)__builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xF, true);
    return v + (long long)(((unsigned long long)h2 << 32) | l2);
}

// Inclusive int64 scan over the wave in lane order with DPP lane moves (row
// shifts 1, 2, 4, 8, then row broadcasts 15 and 31), no LDS traffic.
__device__ __forceinline__ long long wave_incl_scan_dpp(long long v) {
    long long inc = v;
    inc = dpp_add_i64<0x111, 0xF>(inc);  // row_shr:1
    inc = dpp_add_i64<0x112, 0xF>(inc);  // row_shr:2
    inc = dpp_add_i64<0x114, 0xF>(inc);  // row_shr:4
    inc = dpp_add_i64<0x118, 0xF>(inc);  // row_shr:8
    inc = dpp_add_i64<0x142, 0xA>(inc);  // row_bcast:15
    inc = dpp_add_i64<0x143, 0xC>(inc);  // row_bcast:31
    return inc;
}

// Orders one wave's LDS accesses before its later ones, across lanes (the
// wave's LDS operations complete in order; this keeps the compiler from
// moving them across).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Exclusive int64 scan over the workgroup in thread order, the wave part
// wave_incl_scan_dpp; *total gets the sum.  LEAD
// false omits the barrier that keeps the scratch's previous readers ahead
// of its new writers, for callers with a barrier since its last use.
template <bool LEAD = true>
__device__ __forceinline__ long long block_excl_scan_dpp(long long v,
                                                         long long *scratch,
                                                         long long *total) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const long long inc = wave_incl_scan_dpp(v);
    if (LEAD) __syncthreads();
    if (l == 63) scratch[w] = inc;
    __syncthreads();
    long long off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const long long t = scratch[i];
        if (i < w) off += t;
        tot += t;
    }
    *total = tot;
    return off + inc - v;
}

struct MinOp {
    __device__ double operator()(double a, double b) const { return fmin(a, b); }
};
struct MaxOp {
    __device__ double operator()(double a, double b) const { return fmax(a, b); }
};
struct OrOp {
    __device__ int operator()(int a, int b) const { return a | b; }
};
template <typename T>
struct AddOp {
    __device__ T operator()(T a, T b) const { return a + b; }
};

// Bin index of relative time t (>= 0) for one period, equal to the oracle's
// (int)((t - P*floor(t/P)) / bd) + 1.  The two divisions are replaced by
// products with the rounded reciprocals; a product can differ from the
// correctly rounded quotient by a few ulp, which changes floor() only when
// the quotient is within 2^-50 (relative) of an integer, and exactly then
// the true division is taken.
__device__ __forceinline__ int fold_bin(double t, double P, double invP,
                                        double bd, double inv_bd) {
    double r = t * invP;
    double q = floor(r);
    double fr = r - q;
    double tol = r * 0x1p-50;
    if (fr <= tol || fr >= 1.0 - tol) q = floor(t / P);
    double ph = t - P * q;
    double x = ph * inv_bd;
    double xi = trunc(x);
    double fx = x - xi;
    double tolx = fabs(x) * 0x1p-50;
    if (fx <= tolx || fx >= 1.0 - tolx) xi = trunc(ph / bd);
    return (int)xi + 1;
}

// The oracle's bin rule with true divisions.
__device__ __forceinline__ int fold_exact(double t, double P, double bd) {
    const double ph = t - P * floor(t / P);
    return (int)trunc(ph / bd) + 1;
}

// Cycle index floor(t/P) of the oracle's fold, with the same near-integer
// fallback as fold_bin, as a double.
__device__ __forceinline__ double fold_cycle(double t, double P,
                                            double invP, bool fast = false) {
    if (fast) return floor(t * invP);  // fast mode's fold, no fallback
    const double r = t * invP;
    double q = floor(r);
    const double fr = r - q;
    const double tol = r * 0x1p-50;
    if (fr <= tol || fr >= 1.0 - tol) q = floor(t / P);
    return q;
}

// Bin of t within cycle q (q = fold_cycle(t)), clamped to [0, nb]:
// equal to min(max(fold_bin(t), 0), nb).
__device__ __forceinline__ int fold_bin_in(double t, double P, double q,
                                           double bd, double inv_bd, int nb,
                                           bool fast = false) {
    if (fast)  // fast mode's fold: FMA phase, reciprocal only
        return min(max((int)(fma(-P, q, t) * inv_bd) + 1, 0), nb);
    const double ph = t - P * q;
    const double x = ph * inv_bd;
    double xi = trunc(x);
    const double fx = x - xi;
    const double tolx = fabs(x) * 0x1p-50;
    if (fx <= tolx || fx >= 1.0 - tolx) xi = trunc(ph / bd);
    return min(max((int)xi + 1, 0), nb);
}

// Point-order copies are stored in blocks of NT points: within a block,
// point 4*l + v (l = 0..63, v = 0..NW-1) sits in slot 64*v + l, so that the
// thread tid = 64*v + l of a workgroup walking a block reads slot tid (one
// contiguous 512-byte line set per wave) while the points of a wave's lanes
// are NW apart in time and mostly fall in distinct bins.  Each curve's copy
// starts at lc_off[c] + NT*c and is padded to whole blocks.
static_assert(NT == 256 && NW == 4, "natural-copy blocks assume 4 wave64");
__device__ __forceinline__ int nat_slot(int i) {
    return (i & ~(NT - 1)) | ((i & (NW - 1)) << 6) | ((i >> 2) & 63);
}
__device__ __forceinline__ long long nat_base(const Batch &b, int lc) {
    return b.lc_off[lc] + (long long)NT * lc;
}

// For time-sorted points the key (cycle, bin) of the oracle's fold is
// non-decreasing in the point index: floor(t/P) is monotone in t and, for a
// fixed cycle, so are t - P*q, the division by bd and the truncation.  The
// first index whose key is >= (c, s) is therefore a binary search.  The
// same holds for fast mode's fold (products and an FMA, each monotone).
__device__ int lower_key(const double *tn, int n, double P, double invP,
                         double bd, double inv_bd, int nb, double c, int s,
                         bool fast = false) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const double t = tn[2 * nat_slot(mid)];
        const double q = fold_cycle(t, P, invP, fast);
        const bool less =
            q < c ||
            (q == c && fold_bin_in(t, P, q, bd, inv_bd, nb, fast) < s);
        if (less) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// A workgroup-uniform global pointer moved to scalar registers, and loads
// from it at a 32-bit byte offset (the base + offset addressing of global
// loads, with no 64-bit address per lane).
typedef __attribute__((address_space(1))) const char gchar;
__device__ __forceinline__ gchar *scalar_base(const void *p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (gchar *)(((unsigned long long)hi << 32) | lo);
}
// a 16-byte record of two int64, one 16-byte load (byte_off a multiple of
// 16 from a 16-byte aligned base)
typedef long long I64x2 __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ T load_at(gchar *base, unsigned byte_off) {
    return *(const __attribute__((address_space(1))) T *)(base + byte_off);
}

// Bins every point of the lane for C periods at once.  C is a compile-time
// constant so that the C fold chains are independent straight-line code the
// scheduler can interleave; the next point's loads are issued before the
// current point is folded.  The near-integer test uses the curve's uniform
// tolerance tol_x (a bound of the per-point ones of fold_bin, widened for the
// shared quotient below): an unflagged point's reciprocal bin is the
// oracle's and lies in [1, nb - os]; a flagged point (a superset of the
// uncertain ones) is first added to its reciprocal bin, within [1, nb], and
// then moved to the bin the oracle's divisions give.
// KIND: 0 unpacked bins (t', weight and weight*flux read per point; int64
// weight sums in binW, int64 flux sums in binY), 1 packed bins (t' and the
// packed increment, into binW), 2 counted bins of a curve whose points
// share one weight (t' and weight*flux read; int32 point counts in binW,
// read as int, int64 flux sums in binY: 12 bytes a bin).  KIND 2 counts
// the npad padding slots (zero weight, at time t_pad) too, and takes them
// back from t_pad's bin at the end.
template <int C, bool EXACT, int KIND>
__device__ __forceinline__ void bin_points(
    const double *tp, const long long *wp, const long long *yp, int L,
    const double *per, const double *inv, const int *nb, const int *base,
    double bd, double inv_bd, double tol_x, long long *binW,
    long long *binY, const double *pg, int npad, double t_pad) {
    constexpr bool PACKED = KIND == 1, UW = KIND == 2;
    // tp, wp, yp: the curve's (uniform) slot bases, in scalar registers;
    // lane tid loads slot tid + i * NT at a 32-bit byte offset from them
    // (base + offset addressing: no per-lane 64-bit address arithmetic; a
    // curve of at most 2^28 points spans less than 2^32 bytes)
    gchar *bt = scalar_base(tp), *bw = scalar_base(wp), *by = scalar_base(yp);
    unsigned int lo = threadIdx.x * 8u;
    double tv_n = load_at<double>(bt, lo);
    long long wv_n = UW ? 1 : load_at<long long>(bw, lo);
    long long yv_n = PACKED ? 0 : load_at<long long>(by, lo);
    // Exact fold: the shifted quotient is computed as XB + x (XB = 2^20,
    // above every bin of the histogram), so that its exponent is fixed and
    // its fraction is exactly the low 32-bit word of the mantissa (units of
    // 2^-32): a fraction of at most 2*tolk is a word of at most tq, one
    // unsigned compare.  tolk adds 2^-29 to the curve's tolerance for the
    // roundings the shift adds (at most 4*XB*2^-53 = 2^-31).
    constexpr double XB = 1048576.0;
    constexpr int XBI = 1 << 20;
    const double tolk = tol_x + 0x1p-29, tbc = XB + tolk;
    const unsigned tq = (unsigned)(2.0 * tolk * 0x1p32) + 1u;
    // per-period origins at bin 1 - XB, so that bin at - XB + 1's LDS
    // address is one shift-and-add of at (32-bit LDS addresses: the
    // origin wraps, the sum does not)
    long long *ow[C], *oy[C];
    int *oc[C];
#pragma unroll
    for (int s = 0; s < C; ++s) {
        const int o = base[s] + 1 - (EXACT ? XBI : 0);
        ow[s] = binW + o;
        oy[s] = binY + o;
        oc[s] = (int *)binW + o;
    }
    // two points per iteration: the prefetched registers alternate instead
    // of being copied
#pragma unroll 2
    for (int i = 0; i < L; ++i) {
        const double tv = tv_n;
        const long long wv = wv_n, yv = yv_n;
        if (i + 1 < L) {
            lo += NT * 8;
            tv_n = load_at<double>(bt, lo);
            if (!UW) wv_n = load_at<long long>(bw, lo);
            if (!PACKED) yv_n = load_at<long long>(by, lo);
        }
        int at[C];
        // the least low word of the C shifted quotients: some period's
        // quotient is near an integer if it is <= tq; the flags themselves
        // are recomputed on the rare path below rather than held through
        // the fold (C lane masks of scalar registers, which then spill)
        unsigned mlo = 0xffffffffu;
        // XB + t'/bd + tolk, shared by the C periods of the pass
        const double tb = EXACT ? fma(tv, inv_bd, tbc) : 0.0;
#pragma unroll
        for (int s = 0; s < C; ++s) {
            if (EXACT) {
                // bin from the rounded reciprocals; a quotient within the
                // uniform tolerance of an integer is flagged for the
                // oracle's divisions.  Unflagged bins lie in [1, nb - os].
                // per[s] holds P/bd here: x = XB + t'/bd - q*P/bd + tolk
                // is within u*(7*t'/bd + 6*P/bd) + 4*XB*u of the oracle's
                // XB + (t' - P*q)/bd + tolk (u = 2^-53: the roundings of
                // the products, the shift and the FMA against the oracle's
                // three, and the shift's own), less than tolk.  So an x
                // whose fraction exceeds 2*tolk lies in the oracle's
                // integer interval, and the conversion's truncation is the
                // oracle's floor; a fraction up to 2*tolk is flagged.
                // inv[s] is 1/P scaled up by 1 + 2^-48, so t' * inv[s]
                // exceeds the oracle's rounded t'/P and the cycle q is the
                // oracle's or one too high; one too high (t'/P within ~36
                // ulp below an integer) puts the unshifted quotient within
                // u*(40*t'/bd + 39*P/bd) below 0, so x within
                // [XB, XB + 2*tolk]: flagged.
                const double q = floor(tv * inv[s]);
                const double x = fma(q, -per[s], tb);
                mlo = min(mlo, (unsigned)__double2loint(x));
                at[s] = (int)x;  // bin at[s] - XB + 1
            } else {
                // fast mode: FMA phase and rounded reciprocals only; a point
                // within ~1e-15 of a bin edge may land in the neighbour bin.
                // No clamp is needed: t' >= 0, and t' * inv is within
                // t'/P * 2^-52 of t'/P, so the cycle is off by at most one
                // and then the phase lies within t' * 2^-52 (far below bd)
                // outside [0, P): the truncated quotient is in
                // [0, ceil(P/bd)], at most nb - 1 (oversample >= 1), and
                // the bin (at + 1) is the clamped one of fold_bin_in
                const double ph = fma(-per[s], floor(tv * inv[s]), tv);
                at[s] = (int)(ph * inv_bd);
            }
        }
#pragma unroll
        for (int s = 0; s < C; ++s) {
            if (UW)
                __hip_atomic_fetch_add(oc[s] + at[s], 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            else
                __hip_atomic_fetch_add(ow[s] + at[s], wv, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            if (!PACKED)
                __hip_atomic_fetch_add(oy[s] + at[s], yv, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        if (EXACT && mlo <= tq) {
            // Every point went to its reciprocal bin above, so the atomics
            // share the fold's basic block.  A flagged point whose oracle
            // bin (both divisions) differs moves there: integer adds
            // commute, so the histogram is the one the oracle's bins give.
            // tx hides the folds above from common-subexpression
            // elimination: keeping their values live into this rare path
            // would cost registers in the main loop.
            double tx = tv;
            asm volatile("" : "+v"(tx));
            // the periods themselves, read here rather than kept in
            // registers through the loop (pg is opaque to hoisting)
            const double *pgx = pg;
            asm volatile("" : "+s"(pgx));
            // the oracle's bin of a point in its first cycle, the same for
            // every period of the pass: t' * inv[s] < 0.9375 gives
            // floor(t'/P) = 0 and the phase t' itself.  On a regular
            // cadence about half of the flags come from such points (t'/bd
            // an exact integer).
            const int e0 = (int)trunc(tx / bd) + 1;
#pragma unroll
            for (int s = 0; s < C; ++s) {
                // the main path's fold again, and its flag
                const double xs =
                    fma(floor(tx * inv[s]), -per[s], fma(tx, inv_bd, tbc));
                if ((unsigned)__double2loint(xs) <= tq) {
                    const int a = base[s] + 1 + ((int)xs - XBI);
                    int e0s = e0;
                    if (!(tx * inv[s] < 0.9375)) {
                        const double Ps = pgx[s];
                        const double ph = tx - Ps * floor(tx / Ps);
                        e0s = (int)trunc(ph / bd) + 1;
                    }
                    const int e = base[s] + min(max(e0s, 0), nb[s]);
                    if (e != a) {
                        if (UW) {
                            __hip_atomic_fetch_add(&((int *)binW)[a], -1,
                                                   __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                            __hip_atomic_fetch_add(&((int *)binW)[e], 1,
                                                   __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                        } else {
                            __hip_atomic_fetch_add(&binW[a], -wv,
                                                   __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                            __hip_atomic_fetch_add(&binW[e], wv,
                                                   __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                        if (!PACKED) {
                            __hip_atomic_fetch_add(
                                &binY[a], -yv, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_WORKGROUP);
                            __hip_atomic_fetch_add(
                                &binY[e], yv, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
                }
            }
        }
    }
    if (UW && npad > 0 && threadIdx.x == 0) {
        // the padding slots, counted above in t_pad's bin (the bin the fold
        // of this mode gives it), taken back (integer adds commute)
#pragma unroll
        for (int s = 0; s < C; ++s) {
            int r;
            if (EXACT) {
                r = min(max(fold_exact(t_pad, pg[s], bd), 0), nb[s]);
            } else {
                const double ph =
                    fma(-per[s], floor(t_pad * inv[s]), t_pad);
                r = min(max((int)(ph * inv_bd) + 1, 0), nb[s]);
            }
            __hip_atomic_fetch_add(&((int *)binW)[base[s] + r], -npad,
                                   __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
}

// ------------------------------------------------------------------ k_prep

// Two order statistics k0 <= k1 of the curve's y values (as FP64) by a
// radix select over the order-preserving 64-bit key, 8 bits a pass.  FP32
// values widened to FP64 have their 29 lowest mantissa bits zero, so their
// keys' lowest 24 bits are all zeros (positive values) or all ones
// (negative: the key inverts every bit); five passes fix the other 40 bits
// and the sign of the prefix fixes the rest.
template <class R>
__device__ void select2(const Batch &b, long long off, int n, int k0, int k1,
                        double *v0, double *v1, unsigned int *hist, int *sel,
                        unsigned int *uscratch) {
    constexpr int PASSES = R::Y_T == ATG_F32 ? 5 : 8;
    unsigned long long pre0 = 0, pre1 = 0, mask = 0;
    int kk0 = k0, kk1 = k1;
    for (int pass = 0; pass < PASSES; ++pass) {
        const int shift = 56 - 8 * pass;
        for (int i = threadIdx.x; i < 512; i += NT) hist[i] = 0;
        __syncthreads();
        unsigned c0 = 0, c1 = 0, d0 = 0, d1 = 0;
        // eight values per step, loaded before the first is counted (the
        // thread's order of points is unchanged)
        for (int i0 = threadIdx.x; i0 < n; i0 += 8 * NT) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                v[u] = R::y(b, off + min(i0 + u * NT, n - 1));
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (i0 + u * NT >= n) break;
                unsigned long long key = okey(v[u]);
                unsigned d = (unsigned)(key >> shift) & 255u;
                if ((key & mask) == pre0) {
                    if (c0 && d != d0) { atomicAdd(&hist[d0], c0); c0 = 0; }
                    d0 = d;
                    ++c0;
                }
                if ((key & mask) == pre1) {
                    if (c1 && d != d1) {
                        atomicAdd(&hist[256 + d1], c1);
                        c1 = 0;
                    }
                    d1 = d;
                    ++c1;
                }
            }
        }
        if (c0) atomicAdd(&hist[d0], c0);
        if (c1) atomicAdd(&hist[256 + d1], c1);
        __syncthreads();
        unsigned h0 = hist[threadIdx.x], h1 = hist[256 + threadIdx.x], tot;
        unsigned e0 = block_excl_scan(h0, uscratch, &tot);
        unsigned e1 = block_excl_scan(h1, uscratch, &tot);
        if (e0 <= (unsigned)kk0 && (unsigned)kk0 < e0 + h0) {
            sel[0] = threadIdx.x;
            sel[1] = kk0 - (int)e0;
        }
        if (e1 <= (unsigned)kk1 && (unsigned)kk1 < e1 + h1) {
            sel[2] = threadIdx.x;
            sel[3] = kk1 - (int)e1;
        }
        __syncthreads();
        pre0 |= (unsigned long long)sel[0] << shift;
        kk0 = sel[1];
        pre1 |= (unsigned long long)sel[2] << shift;
        kk1 = sel[3];
        mask |= 0xffull << shift;
        __syncthreads();
    }
    if (PASSES == 5) {
        if (!(pre0 >> 63)) pre0 |= 0xffffffull;
        if (!(pre1 >> 63)) pre1 |= 0xffffffull;
    }
    *v0 = ikey(pre0);
    *v1 = ikey(pre1);
}

// Raw y and dy of point i (dy = 1 when the curve has none).
template <class R>
__device__ __forceinline__ void load_point(const Batch &b, long long i,
                                           double &y, double &d) {
    y = R::y(b, i);
    d = R::dy(b, i);
}

// Fixed-point weight and weight*flux of a point from its raw y and dy.
template <class R>
__device__ __forceinline__ void quantize_raw(double y, double d, double med,
                                             double wscale, double yscale,
                                             long long &wq, long long &yq) {
    double iv = 1.0;
    if (R::HAS_DY) iv = 1.0 / (d * d);
    const double yc = y - med;
    wq = llrint(iv * wscale);
    yq = llrint((yc * iv) * yscale);
}

template <class R>
__global__ __launch_bounds__(NT) void k_prep(Batch b, Work w, int lds_bins) {
    __shared__ unsigned int hist[512];
    // the sequential totals' chunks: two buffers of SEQ_CH (term, weight)
    // pairs, each padded by 8 for the summing wave's read-ahead
    constexpr int SEQ_CH = 3 * 64 * 4;
    __shared__ double2 seq_v[2 * (SEQ_CH + 8)];
    __shared__ int sel[4];
    __shared__ unsigned int uscr[NW];
    __shared__ double dscr[NW];
    __shared__ int iscr[NW];
    __shared__ long long lscr[NW];
    // exponent histogram of |(y - median) * ivar|: [0] zeros, [e + 1075]
    // values with ilogb e (every finite double: -1074 <= e <= 1023)
    __shared__ unsigned int ehist[EHIST];
    __shared__ int s_e50;

    const int lc = blockIdx.x;
    const long long off = b.lc_off[lc];
    const long long n64 = b.lc_off[lc + 1] - off;
    const long long po = b.per_off[lc], np = b.per_off[lc + 1] - po;
    const long long dof = b.dur_off[lc], nd = b.dur_off[lc + 1] - dof;
    // point indices within a curve are 32-bit, and k_main addresses a
    // curve's lane-major slots (8 B each) by 32-bit byte offsets: at most
    // 2^28 points per curve
    const bool too_long = n64 > ((long long)1 << 28);
    const int n = too_long ? 0 : (int)n64;
    const int tid = threadIdx.x;
    if (tid == 0) w.lc_flags[lc] = 0;  // workspace bytes are reused across shapes
    LcParams P;
    memset(&P, 0, sizeof(P));
    P.n = n;
    P.n_dur = (int)nd;
    P.lc_self = lc;
    P.ws_off = off + (long long)lc * (2 * NT);
    P.L = (n + NT - 1) / NT;
    int status = ATG_LC_OK;
    if (too_long) status = ATG_LC_TOO_MANY_POINTS;
    else if (n <= 0 || np <= 0 || nd <= 0) status = ATG_LC_EMPTY;
    else if (nd > DMAX) status = ATG_LC_TOO_MANY_DURATIONS;

    // grid checks in the oracle's order: periods, then durations
    double pmin = INFINITY, pmax = -INFINITY, dmin = INFINITY,
           dmax = -INFINITY;
    int bad_grid = 0;
    for (long long i = tid; i < np; i += NT) {
        double p = b.periods[po + i];
        if (!(p >= DBL_EPSILON)) bad_grid |= 1;
        pmin = fmin(pmin, p);
        pmax = fmax(pmax, p);
    }
    for (long long i = tid; i < nd; i += NT) {
        double d = b.durs[dof + i];
        if (!(d >= DBL_EPSILON) || isnan(d)) bad_grid |= 2;
        dmin = fmin(dmin, d);
        dmax = fmax(dmax, d);
    }
    pmin = block_reduce(pmin, MinOp(), dscr);
    pmax = block_reduce(pmax, MaxOp(), dscr);
    dmin = block_reduce(dmin, MinOp(), dscr);
    dmax = block_reduce(dmax, MaxOp(), dscr);
    bad_grid = block_reduce(bad_grid, OrOp(), iscr);
    if (status == ATG_LC_OK) {
        if (bad_grid & 1) status = ATG_LC_BAD_PERIOD;
        else if ((bad_grid & 2) || dmax > pmin) status = ATG_LC_BAD_DURATION;
        // bin indices are 32-bit: n_bins = ceil(P / bd) + oversample with
        // bd = min(durations) / oversample must stay within 2^30
        else if (!(ceil(pmax / (dmin / (double)b.oversample)) + b.oversample <=
                   (double)(1 << 30)))
            status = ATG_LC_TOO_MANY_BINS;
    }

    // pass A: min(t), NaN and dy checks, time order
    double tmin = INFINITY, tmax = -INFINITY;
    int bad = 0;
    // four points per step, loaded first (indices past the end re-read
    // the last point, which changes no minimum, maximum or flag)
    for (int i0 = tid; i0 < n; i0 += 4 * NT) {
        double tv[4], tp[4], yv[4], dv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = min(i0 + u * NT, n - 1);
            tv[u] = b.t[off + i];
            tp[u] = b.t[off + max(i - 1, 0)];
            load_point<R>(b, off + i, yv[u], dv[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = min(i0 + u * NT, n - 1);
            tmax = fmax(tmax, tv[u]);
            tmin = fmin(tmin, tv[u]);
            if (isnan(tv[u]) || isnan(yv[u])) bad |= 1;
            if (i > 0 && !(tv[u] >= tp[u])) bad |= 4;
            if (R::HAS_DY && !(dv[u] > 0.0)) bad |= 2;
        }
    }
    tmin = block_reduce(tmin, MinOp(), dscr);
    tmax = block_reduce(tmax, MaxOp(), dscr);
    bad = block_reduce(bad, OrOp(), iscr);
    if (status == ATG_LC_OK) {
        if (bad & 1) status = ATG_LC_NAN;
        else if (bad & 2) status = ATG_LC_BAD_DY;
    }
    P.t_ref = tmin;
    P.status = status;
    P.sorted = (bad & 4) ? 0 : 1;
    if (status != ATG_LC_OK) {
        if (tid == 0) w.prm[lc] = P;
        return;
    }

    // exact median (numpy: mean of the two middle values for even n)
    double v0, v1;
    select2<R>(b, off, n, (n - 1) / 2, n / 2, &v0, &v1, hist, sel, uscr);
    const double med = (n & 1) ? v0 : (v0 + v1) * 0.5;
    P.y_med = med;

    // parity mode: the oracle's FP64 totals in its order (the refinement
    // kernel uses them).  Each total is one chain of dependent FP64 adds
    // (~20 cycles each), so one wave adds while the others stage: waves 1-3
    // put the next chunk of SEQ_CH (term, weight) pairs in LDS while wave 0
    // adds up the current one in point order from broadcast reads, eight
    // pairs read ahead of their adds.
    if (b.mode == ATG_MODE_PARITY) {
        constexpr int SPAN = 3 * 64;  // staging threads
        const int nch = (n + SEQ_CH - 1) / SEQ_CH;
        auto stage = [&](int c) {
            double2 *q = seq_v + (c & 1) * (SEQ_CH + 8);
            const int j = tid - 64, c0 = c * SEQ_CH;
            double yr[4], dr[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                load_point<R>(b, off + min(c0 + u * SPAN + j, n - 1), yr[u],
                              dr[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double iv = R::HAS_DY ? 1.0 / (dr[u] * dr[u]) : 1.0;
                q[u * SPAN + j] = make_double2((yr[u] - med) * iv, iv);
            }
        };
        double sy = 0.0, si = 0.0;
        if (tid >= 64) stage(0);
        __syncthreads();
        for (int c = 0; c < nch; ++c) {
            if (tid >= 64) {
                if (c + 1 < nch) stage(c + 1);
            } else {
                const double2 *q = seq_v + (c & 1) * (SEQ_CH + 8);
                const int cnt = min(SEQ_CH, n - c * SEQ_CH);
                const int full = cnt & ~7;
                // reads reach index full + 7 at most: within the padding;
                // the scheduling barriers keep each group's reads ahead of
                // the other group's adds
                double2 A[8], B[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) A[u] = q[u];
                int k = 0;
                for (; k + 16 <= full; k += 16) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) B[u] = q[k + 8 + u];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        sy = sy + A[u].x;
                        si = si + A[u].y;
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 8; ++u) A[u] = q[k + 16 + u];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        sy = sy + B[u].x;
                        si = si + B[u].y;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (k + 8 <= full) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        sy = sy + A[u].x;
                        si = si + A[u].y;
                    }
                    k += 8;
                }
                for (; k < cnt; ++k) {
                    sy = sy + q[k].x;
                    si = si + q[k].y;
                }
            }
            __syncthreads();
        }
        if (tid == 0) w.seq[lc] = make_double2(sy, si);
    }

    // pass B: magnitudes for the fixed-point scales, weight uniformity, and
    // the exponent histogram of the flux magnitudes
    for (int i = tid; i < EHIST; i += NT) ehist[i] = 0u;
    __syncthreads();
    double sw_sum = 0.0, sy_sum = 0.0, iv_min = INFINITY, iv_max = -INFINITY;
    for (int i0 = tid; i0 < n; i0 += 4 * NT) {
        double yr[4], dr[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            load_point<R>(b, off + min(i0 + u * NT, n - 1), yr[u], dr[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (i0 + u * NT >= n) break;
            double iv = 1.0;
            if (R::HAS_DY) iv = 1.0 / (dr[u] * dr[u]);
            double yc = yr[u] - med;
            const double a = fabs(yc * iv);
            sw_sum += iv;
            sy_sum += a;
            iv_min = fmin(iv_min, iv);
            iv_max = fmax(iv_max, iv);
            atomicAdd(&ehist[a > 0.0 && a <= DBL_MAX ? ilogb(a) + 1075 : 0],
                      1u);
        }
    }
    sw_sum = block_reduce(sw_sum, AddOp<double>(), dscr);
    sy_sum = block_reduce(sy_sum, AddOp<double>(), dscr);
    iv_min = block_reduce(iv_min, MinOp(), dscr);
    iv_max = block_reduce(iv_max, MaxOp(), dscr);
    // e50: exponent bin of the median magnitude (the first bin whose
    // cumulative count reaches half the points)
    {
        constexpr int PER = (EHIST + NT - 1) / NT;
        const int h0 = min(tid * PER, EHIST), h1 = min(h0 + PER, EHIST);
        unsigned int c = 0;
        for (int i = h0; i < h1; ++i) c += ehist[i];
        unsigned int tot;
        unsigned int before = block_excl_scan(c, uscr, &tot);
        const unsigned int half = (unsigned int)(n + 1) / 2;
        if (tid == 0) s_e50 = 0;
        __syncthreads();
        for (int i = h0; i < h1; ++i) {
            const unsigned int after = before + ehist[i];
            if (before < half && after >= half) s_e50 = i;
            before = after;
        }
        __syncthreads();
    }
    // Uniform weights: a bin is (point count, weight*flux sum) packed into
    // one int64 as count << ybits + sum.  Prefix sums (with the padding
    // copies) reach at most 2n points, which takes cbits bits; the flux
    // field keeps 63 - cbits bits and its scale leaves 2 bits of headroom.
    const int cbits = ilogb(2.0 * n + 1.0) + 1;
    // The flux field keeps 63 - cbits bits (48 at 15k points, 44 at 200k),
    // which moves power by up to ~1e-9 (15k) or ~1e-7 (200k) relative; the
    // refinement widens its near-tie window by that bound (k_cands).  In
    // parity mode packing is limited to n <= 65536 points, so that the
    // error stays near 1e-9 and the window small.
    // The packed flux quantum is proportional to the sum of magnitudes, so
    // one extreme value (an outlier far above the noise) coarsens every
    // window sum.  Packing is kept only while that sum stays within 64 times
    // n points of the median magnitude (R = sum / (n * 2^e50) is ~1.2-2.4
    // for Gaussian noise, ~100+ with a single 1e6-sigma point at 15k
    // points); beyond it the 60-bit unpacked bins keep the error at the
    // noise-only level.
    const bool spread_ok =
        s_e50 > 0 ? sy_sum <= ldexp(64.0 * (double)n, s_e50 - 1075)
                  : !(sy_sum > 0.0);
    P.uniform = iv_min == iv_max ? 1 : 0;
    P.packed = (iv_min == iv_max && 63 - cbits >= 40 && spread_ok &&
                (b.mode == ATG_MODE_FAST || n <= 65536)) ? 1 : 0;
    P.ybits = 63 - cbits;
    // scale so that the sum of |quantized| values stays below 2^60 (packed:
    // 2^(ybits-3), and 2^49 at most, so that every prefix, which the
    // padding copies at most double, stays below 2^51 for the exact
    // conversion of the scan, small_i64_to_f64).
    const int ytop = P.packed ? min(P.ybits - 4, 48) : 59;
    int sw = 59 - ilogb(sw_sum);
    int sy = (sy_sum > 0.0 && isfinite(sy_sum)) ? ytop - ilogb(sy_sum) : 0;
    sw = max(-1000, min(1000, sw));
    sy = max(-1000, min(1000, sy));
    const double wscale = ldexp(1.0, sw), yscale = ldexp(1.0, sy);
    P.w_unit = ldexp(1.0, -sw);
    P.y_unit = ldexp(1.0, -sy);
    P.bd = dmin / (double)b.oversample;
    // k_main's x = t'/bd - q*P/bd is within u*(5*t'/bd + 6*P/bd) of the
    // oracle's quotient (u = 2^-53), and a cycle one too high puts x within
    // u*(40*t'/bd + 39*P/bd) below 0; 2^-47 * (t'/bd + P/bd + ...) covers
    // both
    P.tol_x = ldexp((tmax - tmin) / P.bd + ceil(pmax / P.bd) +
                        b.oversample + 2.0,
                    -47);

    // pass C: lane-major workspace (slot = step * NT + lane).  Points whose
    // first-cycle bin quotient t'/bd lies at a bin edge (exactly gridded
    // times put every few dozenth point there, and t' = 0 too) are flagged
    // by the near-integer test of every period whose first cycle holds
    // them, which sends their whole wave through the divisions.  They go to
    // the last E_steps steps, lanes side by side, so that they share a few
    // wave steps.  The other points fill the steps in time order, 256 per
    // step, lane l of wave v taking point 4*l + v of the step: the lanes of
    // a wave hold points 4 apart in time, which for every period fall in
    // distinct, nearby bins (few LDS bank conflicts among the atomics; the
    // phase wraps at most once within a step).  The atomics are
    // order-independent, so the placement changes nothing but the time.
    const double inv_bd0 = 1.0 / P.bd, edge_tol = 4.0 * P.tol_x;
    auto is_edge = [&](double tp) {
        const double f = __builtin_amdgcn_fract(tp * inv_bd0);
        return f <= edge_tol || f >= 1.0 - edge_tol;
    };
    // Each wave walks a contiguous quarter of the points, 64 at a time, so
    // that the rank of a point among the edge points (or among the rest) is
    // a ballot count plus the wave's running carry, and the 64 points of a
    // step land in four runs of 16 slots.
    const int wv = tid >> 6, ln = tid & 63;
    const int qlen = (n + NW - 1) / NW;
    const int q0 = min(wv * qlen, n), q1 = min(q0 + qlen, n);
    const unsigned long long below = (1ull << ln) - 1ull;
    int my_edges = 0;  // edge points of this wave's quarter (every lane)
    int close = 0;     // points less than a bin after the point 4 before
    for (int i0 = q0; i0 < q1; i0 += 4 * 64) {
        double tr[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            tr[u] = b.t[off + min(i0 + u * 64 + ln, n - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool in = i0 + u * 64 + ln < q1;
            my_edges += __popcll(__ballot(in && is_edge(tr[u] - tmin)));
            const double t4 = __shfl_up(tr[u], 4);
            close += (in && ln >= 4 && tr[u] - t4 < P.bd) ? 1 : 0;
        }
    }
    __syncthreads();
    if (ln == 0) iscr[wv] = my_edges;
    __syncthreads();
    int edge_before = 0, n_edge = 0;
    for (int k = 0; k < NW; ++k) {
        if (k < wv) edge_before += iscr[k];
        n_edge += iscr[k];
    }
    __syncthreads();
    // Interleaved lanes (below) put points 4 apart side by side; where that
    // is often less than a bin (dense irregular sampling), lanes of one
    // atomic would hit the same bin, so such curves keep contiguous runs
    // per lane (lane j owns points [j*L0, (j+1)*L0) of the rest), whose
    // lanes are far apart in time.
    const bool ilv = 50 * block_reduce(close, AddOp<int>(), iscr) < n;
    const int n0 = n - n_edge;
    const int L0 = (n0 + NT - 1) / NT, L1 = (n_edge + NT - 1) / NT;
    P.L = L0 + L1;
    long long wq_sum = 0, yq_sum = 0;
    const double t_pad = tmax - tmin;
    P.t_pad = t_pad;
    const long long yq_pad = 0;
    // padding slots of this lane: zero weight, an in-range time
    for (int i = 0; i < L0; ++i)
        if ((ilv ? i * NT + (tid & 63) * 4 + (tid >> 6) : tid * L0 + i) >=
            n0) {
            const long long slot = P.ws_off + (long long)i * NT + tid;
            w.ws_t[slot] = t_pad;
            w.ws_w[slot] = 0;
            w.ws_y[slot] = yq_pad;
        }
    for (int i = 0; i < L1; ++i)
        if (i * NT + tid >= n_edge) {
            const long long slot = P.ws_off + (long long)(L0 + i) * NT + tid;
            w.ws_t[slot] = t_pad;
            w.ws_w[slot] = 0;
            w.ws_y[slot] = yq_pad;
        }
    int carry = edge_before;  // edge points before the wave's next step
    for (int i0 = q0; i0 < q1; i0 += 4 * 64) {
        double tr[4], yr[4], dr[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int src = min(i0 + u * 64 + ln, n - 1);
            tr[u] = b.t[off + src];
            load_point<R>(b, off + src, yr[u], dr[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int src = i0 + u * 64 + ln;
            const bool valid = src < q1;
            const double tp = tr[u] - tmin;
            const bool edge = valid && is_edge(tp);
            const unsigned long long m = __ballot(edge);
            const int e = carry + __popcll(m & below);  // edges before src
            carry += __popcll(m);
            if (!valid) continue;
            long long wq, yq;
            quantize_raw<R>(yr[u], dr[u], med, wscale, yscale, wq, yq);
            long long slot;
            if (edge) {
                slot = P.ws_off + (long long)(L0 + e / NT) * NT + e % NT;
            } else {
                const int d = src - e;  // rank among the rest
                const int q = d & (NT - 1);
                slot = P.ws_off +
                       (ilv ? (long long)(d >> 8) * NT + ((q & 3) << 6) +
                                  (q >> 2)
                            : (long long)(d % L0) * NT + d / L0);
            }
            w.ws_t[slot] = tp;
            // packed curves keep the packed bin increment in ws_w
            w.ws_w[slot] = P.packed ? (1ll << P.ybits) + yq : wq;
            w.ws_y[slot] = yq;
            if (src == 0) P.wq_one = wq;
            wq_sum += wq;
            yq_sum += yq;
        }
    }
    wq_sum = block_reduce(wq_sum, AddOp<long long>(), lscr);
    yq_sum = block_reduce(yq_sum, AddOp<long long>(), lscr);
    // pass D: point-order copies for the ranged phase tiles (time-sorted
    // curves only): of k_big, for periods whose bins exceed k_main's
    // histogram (packed bins take 12 bytes each in its scan, 16 otherwise,
    // out of lds_bins * 16), and of k_exact in parity mode, whose histogram
    // holds at most exact_total(lds_bins) bins
    P.natural = 0;
    const double nbmax = ceil(pmax / P.bd) + b.oversample;
    // as k_main's capacity test for the longest period alone (with the
    // overhang of its box scan)
    const double ov =
        fmax(0.0, PADN + fmin(round(dmax / P.bd), 32767.0) - nbmax);
    const bool big =
        P.packed ? nbmax + 1 + fmax(floor((nbmax + 1 + ov + 1) / 2), ov) >
                       2.0 * lds_bins
                 : nbmax + 1 + ov > (double)(P.uniform ? counted_cap(lds_bins)
                                                       : lds_bins);
    const bool big_exact = b.mode == ATG_MODE_PARITY &&
                           nbmax + 1 > (double)exact_total(lds_bins);
    if (P.sorted && (big || big_exact)) {
        P.natural = 1;
        const long long nbase = nat_base(b, lc);
        const int nfull = (n + NT - 1) & ~(NT - 1);
        for (int i0 = tid; i0 < nfull; i0 += 4 * NT) {
            double tr[4], yr[4], dr[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = min(i0 + u * NT, n - 1);
                tr[u] = b.t[off + i];
                load_point<R>(b, off + i, yr[u], dr[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = i0 + u * NT;
                if (i >= nfull) break;
                long long wq = 0, yq = 0;
                double tv = 0.0;
                if (i < n) {
                    quantize_raw<R>(yr[u], dr[u], med, wscale, yscale, wq,
                                 yq);
                    tv = tr[u] - tmin;
                }
                const long long sl = nbase + nat_slot(i);
                ((longlong2 *)w.nat_ty)[sl] =
                    make_longlong2(__double_as_longlong(tv), yq);
                w.nat_w[sl] = wq;
            }
        }
    }
    P.Wq_tot = wq_sum;
    P.Yq_tot = yq_sum;
    P.W_tot = (double)wq_sum * P.w_unit;
    P.Y_tot = (double)yq_sum * P.y_unit;
    if (tid == 0) w.prm[lc] = P;
}

// ----------------------------------------------------------------- k_items

// Work items of curve c go to queue c mod NQ.  Persistent workgroups are
// dispatched round-robin over the XCDs (workgroup i on XCD i mod NQ) and
// drain the queue of their own XCD first, so the points of a curve are
// fetched into one XCD's L2 instead of all eight.
__global__ __launch_bounds__(NT) void k_items(long long n_lc,
                                              const long long *per_off,
                                              int per_item, Work w) {
    __shared__ long long lscr[NW];
    long long pbase = 0;  // periods of the queues before q
    for (int q = 0; q < NQ; ++q) {
        const long long nq = (n_lc - q + NQ - 1) / NQ;
        long long carry = 0, pq = 0;
        for (long long base = 0; base < nq; base += NT) {
            const long long j = base + threadIdx.x;
            const long long lc = q + NQ * j;
            long long c = 0, np = 0;
            if (j < nq) {
                np = per_off[lc + 1] - per_off[lc];
                c = (np + per_item - 1) / per_item;
            }
            long long tot;
            const long long e = block_excl_scan(c, lscr, &tot);
            if (j < nq) w.item_off[lc] = carry + e;
            carry += tot;
            pq += block_reduce(np, AddOp<long long>(), lscr);
        }
        if (threadIdx.x == 0) {
            w.qtot[q] = carry;
            w.qnext[q] = 0u;
            w.big_base[q] = pbase;
            w.big_count[q] = 0u;
            w.big_next[q] = 0u;
        }
        pbase += pq;
    }
    if (threadIdx.x == 0) {
        *w.counter = 0u;
        *w.cand_count = 0u;
        *w.cand_next = 0u;
    }
}

// ------------------------------------------------------------------ k_main

struct Best {
    double s;
    int idx;
};

__device__ __forceinline__ bool better(double s, int i, double bs, int bi) {
    return s > bs || (s == bs && i < bi);
}

// Running best window of the box scan.  A window's score is q/den with
// q = num^2 (num = depth * W_in * W_out / W_tot^2) and den = w_in * w_out^2
// (likelihood) or w_in * w_out (snr), a monotone transform of the oracle's
// objective; a lane compares its windows by cross-multiplication (no
// division), equal scores keep the lowest (duration, start) index as the
// oracle's first-wins loop does; the lanes' winners are compared by the
// quotient (block_argmax).
struct Score {
    double q = -1.0, den = 1.0;
    int idx = INT_MAX;
    __device__ __forceinline__ void offer(double q2, double d2, int i2) {
        const double lhs = q2 * den, rhs = q * d2;
        if (lhs > rhs || (lhs == rhs && i2 < idx)) {
            q = q2;
            den = d2;
            idx = i2;
        }
    }
    // offer() of a window that counts only if ok; returns whether it was
    // taken
    __device__ __forceinline__ bool offer_if(bool ok, double q2, double d2,
                                             int i2) {
        const double lhs = q2 * den, rhs = q * d2;
        const bool take = ok & ((lhs > rhs) | ((lhs == rhs) & (i2 < idx)));
        if (take) {
            q = q2;
            den = d2;
            idx = i2;
        }
        return take;
    }
    // for a candidate whose index exceeds every one offered before; ok
    // false leaves the score as it is
    __device__ __forceinline__ void offer_later(bool ok, double q2, double d2,
                                                int i2) {
        if (ok & (q2 * den > q * d2)) {
            q = q2;
            den = d2;
            idx = i2;
        }
    }
};

__device__ __forceinline__ void write_nan(const Out &o, long long p) {
    const double nan = __longlong_as_double(0x7ff8000000000000ll);
    o.power[p] = nan;
    o.depth[p] = nan;
    o.depth_err[p] = nan;
    o.duration[p] = nan;
    o.transit_time[p] = nan;
    o.depth_snr[p] = nan;
    o.log_like[p] = nan;
}

// fmod(x, P) for x >= 0: x itself below P and x - P below 2P (exact by
// Sterbenz's lemma, so equal to fmod), the library call beyond.
__device__ __forceinline__ double fold_mod(double x, double P) {
    return x < P ? x : (x < 2.0 * P ? x - P : fmod(x, P));
}

// Sliding-box scan of one period's prefix arrays (Cw or the point counts
// Cn, and Cy; index 0..nb).  Duration-major: a thread meets its windows in
// increasing index (k << 16) + s (k < 64 durations, s < 2^15 starts: the
// order of (k, s), decoded with a shift), so a later equal score never
// replaces the held one.  Every lane runs the same number of iterations per
// duration (the last one predicated off past the final start, its loads
// clamped to it), so that a wave never runs a lane's remainder iterations on
// their own.
// One window of the scan: start s, m bins, (duration, start) index idx;
// counts only if in.  ntot > 0 (counted weights with one point's weight at
// least epsn, as box_scan_pad's CFORM): the weight tests are 0 < n < ntot
// on the window's point count n, exactly.
template <bool PACKED, bool LIKE>
__device__ __forceinline__ void box_window(Score &best, const double *Cw,
                                           const int *Cn, const double *Cy,
                                           double c1, double ytn, double epsn,
                                           int ntot, int s, int m, int idx,
                                           bool in) {
    const int ni = PACKED ? Cn[s + m] - Cn[s] : 0;
    const double win = PACKED ? (double)ni * c1 : Cw[s + m] - Cw[s];
    const double yin = Cy[s + m] - Cy[s];
    const double wout = 1.0 - win;
    // depth * W_in * W_out / W_tot^2 = y_tot*w_in - y_in
    const double num = fma(ytn, win, -yin);
    const bool wok = ntot > 0 ? (ni > 0) & (ni < ntot)
                              : !(win < epsn) & !(wout < epsn);
    const bool ok = in & wok & !(num < 0.0);
    best.offer_later(ok, num * num, LIKE ? win * wout * wout : win * wout,
                     idx);
}

template <bool PACKED, bool LIKE>
__device__ __forceinline__ void box_scan(Score &best, const double *Cw,
                                         const int *Cn, const double *Cy,
                                         double c1, const short *s_m, int nd,
                                         int nb, double ytn, double epsn,
                                         int ntot) {
    for (int k = 0; k < nd; ++k) {
        const int m = s_m[k];
        const int send = min(nb - 1, nb - m);
        // iterations in which every lane has a window, then one in which
        // the lanes past the final start are predicated off (their loads
        // clamped to it)
        const int full = (send + 1) / NT;
        const int kbase = k << 16;
        const int s0 = (int)threadIdx.x;
#pragma unroll 4
        for (int i = 0; i < full; ++i) {
            const int s = s0 + i * NT;
            box_window<PACKED, LIKE>(best, Cw, Cn, Cy, c1, ytn, epsn, ntot, s,
                                     m, kbase + s, true);
        }
        if (full * NT <= send) {
            const int s = s0 + full * NT;
            box_window<PACKED, LIKE>(best, Cw, Cn, Cy, c1, ytn, epsn, ntot,
                                     min(s, send), m, kbase + s, s <= send);
        }
    }
}

// The box scan of k_main.  For duration m the starts 0..send (send =
// nb - m) are split among the four waves in shares of q starts, q the
// multiple of 64 with 4q >= send + 1; wave w takes the q starts from
// min(w*q, send + 1 - q), so the shares cover every start, none runs past
// send, and where they overlap a window is evaluated twice (the argmax
// keeps one: same score, same index).  Every lane of a wave runs the same
// q/64 strides, with no predicate or tail.  Only a duration with fewer than
// 64 starts reads past its final window: PADN sentinel entries after the
// period's prefix values make those windows fail the weight test (ending in
// the pad they hold more than the whole weight, starting in it none).  The
// score is num * |num| / den, so a window of negative depth never beats the
// initial best (q = 0): only windows of positive depth are taken here, and
// a period without one is rescanned with box_scan (the depth-0 windows the
// oracle accepts).
template <bool PACKED, bool LIKE, bool CFORM>
__device__ __forceinline__ void box_scan_pad(Score &best, const double *Cw,
                                             const int *Cn, const double *Cy,
                                             double c1, const short *s_m,
                                             int nd, int nb, double ytn,
                                             double epsn, int ntot) {
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int lane = (int)threadIdx.x & 63;
    best.q = 0.0;
    // CFORM (counted weights, packed or counted bins, with one point's
    // weight c1 at least epsn): a window of n of the curve's ntot points
    // weighs n*c1 and leaves (ntot - n)*c1 outside, so the weight tests
    // fail exactly for n = 0 (an empty window, whose flux sum is 0 and
    // whose score, 0, never beats the initial best) and for n >= ntot (the
    // window holds every point, and padding copies), one integer compare.
    // c1 then moves into the constants: wout = 1 - c1*n and
    // num = ytn*c1*n - yin, and the score q/den is the window's divided by
    // c1 (one factor for the whole period, so its windows keep their
    // order).
    const double ytc = ytn * c1;
    for (int k = 0; k < nd; ++k) {
        const int m = s_m[k];
        const int nst = nb - m + 1;  // starts 0..nb-m
        if (nst <= 0) continue;
        const int q = (nst + 4 * 64 - 1) / (4 * 64) * 64;
        const int a = max(0, min(wv * q, nst - q));
        const int s0 = a + lane;
        const int *pn = Cn + s0, *pnm = Cn + s0 + m;
        const double *pw = Cw + s0, *pwm = Cw + s0 + m;
        const double *py = Cy + s0, *pym = Cy + s0 + m;
        const int i0 = (k << 16) + s0;
        const int trips = q / 64;
        if (CFORM) {
#pragma unroll 4
            for (int i = 0; i < trips; ++i) {
                const int o = i * 64;
                const int ni = pnm[o] - pn[o];
                const double n = (double)ni;
                const double yin = pym[o] - py[o];
                const double wout = fma(-c1, n, 1.0);
                const double num = fma(ytc, n, -yin);
                best.offer_later(ni < ntot, num * fabs(num),
                                 LIKE ? n * wout * wout : n * wout, i0 + o);
            }
            continue;
        }
#pragma unroll 4
        for (int i = 0; i < trips; ++i) {
            const int o = i * 64;
            const double win = PACKED ? (double)(pnm[o] - pn[o]) * c1
                                      : pwm[o] - pw[o];
            const double yin = pym[o] - py[o];
            const double wout = 1.0 - win;
            // depth * W_in * W_out / W_tot^2 = y_tot*w_in - y_in
            const double num = fma(ytn, win, -yin);
            best.offer_later(!(fmin(win, wout) < epsn), num * fabs(num),
                             LIKE ? win * wout * wout : win * wout, i0 + o);
        }
    }
}

// Value of lane 63 of an FP64 register after a DPP max-reduction.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(
        __double2loint(v), __double2loint(v), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(
        __double2hiint(v), __double2hiint(v), CTRL, ROW_MASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Workgroup argmax of the lanes' best windows: the highest score, then the
// lowest index.  Each lane's (q, den) becomes one FP64 score q / den (-1
// for a lane without a window); the wave's maximum goes to lane 63 with DPP
// lane moves (quad swaps, half-row and row mirrors, then row broadcasts 15
// and 31: every lane combines with lanes whose partial results cover the
// rest of the wave), and then the lowest index among the lanes holding it
// the same way.  Both operations are commutative and associative, so the
// result does not depend on the order of combination.  wave_argmax
// returns the wave's (score, index) in every lane; block_argmax returns the
// workgroup's index (INT_MAX: no window) in every thread (LEAD as in
// block_excl_scan_dpp).
__device__ __forceinline__ int wave_argmax(const Score &best, double &mx) {
    const double sc = best.idx == INT_MAX ? -1.0 : best.q / best.den;
    double v = sc;
    v = fmax(v, dpp_d<0xB1, 0xF>(v));   // quad_perm [1,0,3,2]
    v = fmax(v, dpp_d<0x4E, 0xF>(v));   // quad_perm [2,3,0,1]
    v = fmax(v, dpp_d<0x141, 0xF>(v));  // row_half_mirror
    v = fmax(v, dpp_d<0x140, 0xF>(v));  // row_mirror
    v = fmax(v, dpp_d<0x142, 0xA>(v));  // row_bcast:15
    v = fmax(v, dpp_d<0x143, 0xC>(v));  // row_bcast:31
    mx = readlane_d(v, 63);
    int c = sc == mx ? best.idx : INT_MAX;
    c = min(c, __builtin_amdgcn_update_dpp(c, c, 0xB1, 0xF, 0xF, false));
    c = min(c, __builtin_amdgcn_update_dpp(c, c, 0x4E, 0xF, 0xF, false));
    c = min(c, __builtin_amdgcn_update_dpp(c, c, 0x141, 0xF, 0xF, false));
    c = min(c, __builtin_amdgcn_update_dpp(c, c, 0x140, 0xF, 0xF, false));
    c = min(c, __builtin_amdgcn_update_dpp(c, c, 0x142, 0xA, 0xF, false));
    c = min(c, __builtin_amdgcn_update_dpp(c, c, 0x143, 0xC, 0xF, false));
    return __builtin_amdgcn_readlane(c, 63);
}
template <bool LEAD = true>
__device__ __forceinline__ int block_argmax(const Score &best, double *dscr,
                                            int *iscr) {
    double mx;
    const int ci = wave_argmax(best, mx);
    const int wv = threadIdx.x >> 6;
    if (LEAD) __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        dscr[wv] = mx;
        iscr[wv] = ci;
    }
    __syncthreads();
    double bm = dscr[0];
    int bi = iscr[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) {
        const double d = dscr[i];
        const int j = iscr[i];
        if (d > bm || (d == bm && j < bi)) {
            bm = d;
            bi = j;
        }
    }
    return bi;
}

// Padding, prefix scan and box scan of one period whose int64 histogram
// occupies bins [base, base + nb] of binW / binY; writes period p.  A
// packed histogram (P.packed) lives in binW alone (binY is unused): its
// scan writes the FP64 flux prefix in place and the int32 point counts into
// the (nb + 2) / 2 slots above it.  Counted bins (cnt non-null: a
// uniform-weight curve with unpacked bins) keep the int64 flux sums in binW
// and the int32 point counts in cnt; both prefixes are written in place.
// Entries past a period's last prefix value that its box scan reads: the
// lanes past the final window of a duration with fewer than 64 starts reach
// index 63 + m (m the duration's bins), so max(0, PADN + m_max - nb).
__device__ __forceinline__ int scan_overhang(int nb, int mmax) {
    return max(0, PADN + mmax - nb);
}


template <bool LIKE>
__device__ __forceinline__ void finish_period(const Batch &b, const Out &o, const LcParams &P,
                              long long p, int nb, int base,
                              long long *binW, long long *binY, int *cnt,
                              const short *s_m, int nd, int mmax,
                              long long *lscr, double *dscr, int *iscr) {
    const int tid = threadIdx.x;
    const int os = b.oversample;
    const bool packed = P.packed;
    const bool counted = cnt != nullptr;
    // padding: bins[nb-os+j-1] = bins[j], j = 1..os (an assignment)
    if (nb - os > os) {
        for (int j = 1 + tid; j <= os; j += NT) {
            binW[base + nb - os + j - 1] = binW[base + j];
            if (counted) cnt[base + nb - os + j - 1] = cnt[base + j];
            else if (!packed) binY[base + nb - os + j - 1] = binY[base + j];
        }
    } else if (tid == 0) {  // overlapping ranges: the oracle's order
        for (int j = 1, d = nb - os; j <= os; ++j, ++d) {
            binW[base + d] = binW[base + j];
            if (counted) cnt[base + d] = cnt[base + j];
            else if (!packed) binY[base + d] = binY[base + j];
        }
    }
    __syncthreads();

    // inclusive prefix over bins 0..nb in int64 (exact), then FP64
    const int nval = nb + 1;
    // an odd span per thread: the threads' first bins then fall in different
    // LDS banks (an even span put 2-8 threads on one bank; stage -0.6%)
    const int seg = ((nval + NT - 1) / NT) | 1;
    const int b0 = min(tid * seg, nval), b1 = min(b0 + seg, nval);
    long long sw = 0, sy = 0;
    for (int i = b0; i < b1; ++i) {
        if (counted) {
            sw += cnt[base + i];
            sy += binW[base + i];
        } else {
            sw += binW[base + i];
            if (!packed) sy += binY[base + i];
        }
    }
    long long totw, toty;
    // the barrier after the padding keeps every reader of the scratch (the
    // previous period's) ahead; the second scan has its own scratch
    long long ow = block_excl_scan_dpp<false>(sw, lscr, &totw);
    long long oy = packed ? 0 : block_excl_scan_dpp<false>(sy, lscr + NW,
                                                           &toty);
    // cumulative sums in units of the total weight (FP64).  Packed bins:
    // the flux prefix replaces the histogram in place and the weight prefix
    // is kept as the int32 point count Cn just above it (12 bytes a bin);
    // a window's weight is then its exact count times c1, the weight of one
    // point.  Unpacked bins: weight and flux prefixes in place.
    const double fw = P.w_unit / P.W_tot, fy = P.y_unit / P.W_tot;
    const double c1 = (double)P.wq_one * fw;
    double *Cw = (double *)binW + base;
    double *Cy = (packed || counted) ? (double *)binW + base
                                     : (double *)binY + base;
    int *Cn = counted ? cnt + base : (int *)(binW + base + nval);
    // packed prefix = (points << ybits) + flux sum, exactly; ybits >= 40,
    // so the low word is the flux field's and the high word holds the
    // count above its top hb = ybits - 32 (sign-extended) flux bits
    const int hb = P.ybits - 32;
    for (int i = b0; i < b1; ++i) {
        if (counted) {
            ow += cnt[base + i];
            oy += binW[base + i];
            Cn[i] = (int)ow;
            Cy[i] = (double)oy * fy;
            continue;
        }
        ow += binW[base + i];
        if (packed) {
            const int hi = (int)((unsigned long long)ow >> 32);
            const int yhi = (int)((unsigned)hi << (32 - hb)) >> (32 - hb);
            Cn[i] = (int)((unsigned)(hi - yhi) >> hb);
            const long long ys =
                (long long)(((unsigned long long)(unsigned)yhi << 32) |
                            (unsigned)ow);
            Cy[i] = small_i64_to_f64(ys) * fy;
        } else {
            oy += binY[base + i];
            Cw[i] = (double)ow * fw;
            Cy[i] = (double)oy * fy;
        }
    }
    // the box scan's sentinels after the last prefix value: the total
    // weight plus far more than all of it (the flux entries there are read
    // but never decide a window)
    const int over = scan_overhang(nb, mmax);
    for (int t = tid; t < over; t += NT) {
        if (counted) {
            Cn[nval + t] = (int)totw + (1 << 30);
        } else if (packed) {
            const int hi = (int)((unsigned long long)totw >> 32);
            const int yhi = (int)((unsigned)hi << (32 - hb)) >> (32 - hb);
            Cn[nval + t] = (int)((unsigned)(hi - yhi) >> hb) + (1 << 30);
        } else {
            Cw[nval + t] = (double)totw * fw + 4.0;
        }
    }
    __syncthreads();

    // box scan: start n, duration k (m bins), window (n, n+m]
    const double Wt = P.W_tot, Yt = P.Y_tot;
    const double ytn = Yt / Wt, epsn = DBL_EPSILON / Wt;
    Score best;
    if ((packed || counted) && c1 >= epsn)
        box_scan_pad<true, LIKE, true>(best, Cw, Cn, Cy, c1, s_m, nd, nb, ytn,
                                       epsn, P.n);
    else if (packed || counted)
        box_scan_pad<true, LIKE, false>(best, Cw, Cn, Cy, c1, s_m, nd, nb, ytn,
                                        epsn, 0);
    else
        box_scan_pad<false, LIKE, false>(best, Cw, Cn, Cy, c1, s_m, nd, nb,
                                         ytn, epsn, 0);
    // workgroup argmax (highest score, then lowest (k, n))
    // the scratch's last readers (the previous period's argmax) are behind
    // this period's barriers
    int bi = block_argmax<false>(best, dscr, iscr);
    if (bi == INT_MAX) {  // no window of positive depth: the oracle's choice
                          // among those of depth 0, if any
        Score zero;
        if (packed || counted)
            box_scan<true, LIKE>(zero, Cw, Cn, Cy, c1, s_m, nd, nb, ytn, epsn,
                                 c1 >= epsn ? P.n : 0);
        else
            box_scan<false, LIKE>(zero, Cw, Cn, Cy, c1, s_m, nd, nb, ytn,
                                  epsn, 0);
        bi = block_argmax(zero, dscr, iscr);
    }
    if (tid == 0) {
        if (bi == INT_MAX) {  // no window qualifies: power -inf as astropy
            o.power[p] = -INFINITY;
            o.depth[p] = 0.0;
            o.depth_err[p] = 0.0;
            o.duration[p] = 0.0;
            o.transit_time[p] = P.t_ref;
            o.depth_snr[p] = 0.0;
            o.log_like[p] = 0.0;
        } else {
            // the winner's window sums; k_fields derives the seven fields
            // from them in parallel over periods (the duration slot carries
            // -(index + 1) until then)
            const int k = bi >> 16, s = bi & 0xffff, m = s_m[k];
            o.depth[p] = ((packed || counted) ? (double)(Cn[s + m] - Cn[s]) * c1
                                              : Cw[s + m] - Cw[s]) *
                         Wt;
            o.depth_err[p] = (Cy[s + m] - Cy[s]) * Wt;
            o.duration[p] = -(double)bi - 1.0;
        }
    }
    // no barrier here: the next period's writes follow its padding barrier
    // (which thread 0 reaches after these reads), and k_main puts one
    // before the next pass clears the histogram
}

// Box scans of one period by one wave (finish_period_wave): the starts of
// each duration 64 at a time, one per lane, the lanes past the last start
// of the final stride repeating it (same score and index).  The window
// arithmetic, scores and tests are box_scan_pad's (positive depth) and
// box_window's (the depth-0 rescan).
template <bool PACKED, bool LIKE, bool CFORM>
__device__ __forceinline__ void wave_box_scan(Score &best, const double *Cw,
                                              const int *Cn, const double *Cy,
                                              double c1, const short *s_m,
                                              int nd, int nb, double ytn,
                                              double epsn, int ntot) {
    const int lane = (int)threadIdx.x & 63;
    best.q = 0.0;
    const double ytc = ytn * c1;
    auto window = [&](int st, int m, int idx) {
        if (CFORM) {
            const int ni = Cn[st + m] - Cn[st];
            const double n = (double)ni;
            const double yin = Cy[st + m] - Cy[st];
            const double wout = fma(-c1, n, 1.0);
            const double num = fma(ytc, n, -yin);
            best.offer_later(ni < ntot, num * fabs(num),
                             LIKE ? n * wout * wout : n * wout, idx);
        } else {
            const double win = PACKED ? (double)(Cn[st + m] - Cn[st]) * c1
                                      : Cw[st + m] - Cw[st];
            const double yin = Cy[st + m] - Cy[st];
            const double wout = 1.0 - win;
            const double num = fma(ytn, win, -yin);
            best.offer_later(!(fmin(win, wout) < epsn), num * fabs(num),
                             LIKE ? win * wout * wout : win * wout, idx);
        }
    };
    for (int k = 0; k < nd; ++k) {
        const int m = s_m[k];
        const int nst = nb - m + 1;  // starts 0..nb-m
        if (nst <= 0) continue;
        const int full = nst >> 6, kb = k << 16;
        for (int i = 0; i < full; ++i) {
            const int st = (i << 6) + lane;
            window(st, m, kb + st);
        }
        if (nst & 63) {
            const int st = min((full << 6) + lane, nst - 1);
            window(st, m, kb + st);
        }
    }
}

template <bool PACKED, bool LIKE>
__device__ __forceinline__ void wave_box_rescan(Score &best, const double *Cw,
                                                const int *Cn, const double *Cy,
                                                double c1, const short *s_m,
                                                int nd, int nb, double ytn,
                                                double epsn, int ntot) {
    const int lane = (int)threadIdx.x & 63;
    for (int k = 0; k < nd; ++k) {
        const int m = s_m[k];
        const int send = min(nb - 1, nb - m);
        for (int st = lane; st - lane <= send; st += 64)
            box_window<PACKED, LIKE>(best, Cw, Cn, Cy, c1, ytn, epsn, ntot,
                                     min(st, send), m, (k << 16) + st,
                                     st <= send);
    }
}

// finish_period by one wave, with no workgroup barrier: k_main's passes of
// short periods (four or more to a pass) give every period the space for
// its own point counts, so that the four waves finish different periods
// side by side.  The padding, the exact int64 prefix (a wave scan), the
// FP64 conversion, the box scan and the argmax are finish_period's, in the
// same order within the period.

template <bool LIKE>
__device__ __forceinline__ void finish_period_wave(
    const Batch &b, const Out &o, const LcParams &P, long long p, int nb,
    int base, long long *binW, long long *binY, int *cnt, const short *s_m,
    int nd) {
    const int lane = (int)threadIdx.x & 63;
    const int os = b.oversample;
    const bool packed = P.packed;
    const bool counted = cnt != nullptr;
    // padding: bins[nb-os+j-1] = bins[j], j = 1..os (an assignment)
    if (nb - os > os) {
        for (int j = 1 + lane; j <= os; j += 64) {
            binW[base + nb - os + j - 1] = binW[base + j];
            if (counted) cnt[base + nb - os + j - 1] = cnt[base + j];
            else if (!packed) binY[base + nb - os + j - 1] = binY[base + j];
        }
    } else if (lane == 0) {  // overlapping ranges: the oracle's order
        for (int j = 1, d = nb - os; j <= os; ++j, ++d) {
            binW[base + d] = binW[base + j];
            if (counted) cnt[base + d] = cnt[base + j];
            else if (!packed) binY[base + d] = binY[base + j];
        }
    }
    wave_lds_sync();

    const int nval = nb + 1;
    const int seg = (nval + 63) >> 6;
    const int b0 = min(lane * seg, nval), b1 = min(b0 + seg, nval);
    long long sw = 0, sy = 0;
    for (int i = b0; i < b1; ++i) {
        if (counted) {
            sw += cnt[base + i];
            sy += binW[base + i];
        } else {
            sw += binW[base + i];
            if (!packed) sy += binY[base + i];
        }
    }
    long long ow = wave_incl_scan_dpp(sw) - sw;
    long long oy = packed ? 0 : wave_incl_scan_dpp(sy) - sy;
    const double fw = P.w_unit / P.W_tot, fy = P.y_unit / P.W_tot;
    const double c1 = (double)P.wq_one * fw;
    double *Cw = (double *)binW + base;
    double *Cy = (packed || counted) ? (double *)binW + base
                                     : (double *)binY + base;
    int *Cn = counted ? cnt + base : (int *)(binW + base + nval);
    const int hb = P.ybits - 32;
    for (int i = b0; i < b1; ++i) {
        if (counted) {
            ow += cnt[base + i];
            oy += binW[base + i];
            Cn[i] = (int)ow;
            Cy[i] = (double)oy * fy;
            continue;
        }
        ow += binW[base + i];
        if (packed) {
            const int hi = (int)((unsigned long long)ow >> 32);
            const int yhi = (int)((unsigned)hi << (32 - hb)) >> (32 - hb);
            Cn[i] = (int)((unsigned)(hi - yhi) >> hb);
            const long long ys =
                (long long)(((unsigned long long)(unsigned)yhi << 32) |
                            (unsigned)ow);
            Cy[i] = small_i64_to_f64(ys) * fy;
        } else {
            oy += binY[base + i];
            Cw[i] = (double)ow * fw;
            Cy[i] = (double)oy * fy;
        }
    }
    wave_lds_sync();

    const double Wt = P.W_tot, Yt = P.Y_tot;
    const double ytn = Yt / Wt, epsn = DBL_EPSILON / Wt;
    Score best;
    if ((packed || counted) && c1 >= epsn)
        wave_box_scan<true, LIKE, true>(best, Cw, Cn, Cy, c1, s_m, nd, nb, ytn,
                                        epsn, P.n);
    else if (packed || counted)
        wave_box_scan<true, LIKE, false>(best, Cw, Cn, Cy, c1, s_m, nd, nb,
                                         ytn, epsn, 0);
    else
        wave_box_scan<false, LIKE, false>(best, Cw, Cn, Cy, c1, s_m, nd, nb,
                                          ytn, epsn, 0);
    double mx;
    int bi = wave_argmax(best, mx);
    if (bi == INT_MAX) {  // no window of positive depth: the depth-0 rule
        Score zero;
        if (packed || counted)
            wave_box_rescan<true, LIKE>(zero, Cw, Cn, Cy, c1, s_m, nd, nb, ytn,
                                        epsn, c1 >= epsn ? P.n : 0);
        else
            wave_box_rescan<false, LIKE>(zero, Cw, Cn, Cy, c1, s_m, nd, nb,
                                         ytn, epsn, 0);
        bi = wave_argmax(zero, mx);
    }
    if (lane == 0) {
        if (bi == INT_MAX) {  // no window qualifies: power -inf as astropy
            o.power[p] = -INFINITY;
            o.depth[p] = 0.0;
            o.depth_err[p] = 0.0;
            o.duration[p] = 0.0;
            o.transit_time[p] = P.t_ref;
            o.depth_snr[p] = 0.0;
            o.log_like[p] = 0.0;
        } else {
            const int k = bi >> 16, st = bi & 0xffff, m = s_m[k];
            o.depth[p] = ((packed || counted)
                              ? (double)(Cn[st + m] - Cn[st]) * c1
                              : Cw[st + m] - Cw[st]) *
                         Wt;
            o.depth_err[p] = (Cy[st + m] - Cy[st]) * Wt;
            o.duration[p] = -(double)bi - 1.0;
        }
    }
}

// Destinations of a point in raw bin r once the padding assignment
// bins[nb-os+j-1] = bins[j] (j = 1..os) has been applied: its own bin unless
// that bin is overwritten, plus the copy of bins 1..os.  Valid when the copy
// range does not overlap its source (nb - os > os).
__device__ __forceinline__ void pad_targets(int r, int nb, int os, int &j1,
                                            int &j2) {
    j1 = (r < nb - os || r == nb) ? r : -1;
    j2 = (r >= 1 && r <= os) ? r + nb - os - 1 : -1;
}

// One point into bin l of a tile: weight (or count) at bw[l], weight*flux
// at by[l].  Two arrays of 8-byte bins rather than interleaved pairs: the
// lanes of a wave hold points ~1.4 bins apart, which an 8-byte stride
// spreads over 32 of the 64 LDS bank pairs, a 16-byte one over 16.
__device__ __forceinline__ void bin_add(long long *bw, long long *by, int l,
                                        long long wv, long long yv) {
    __hip_atomic_fetch_add(bw + l, wv, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(by + l, yv, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void seg_add(int j, int s0, int hi, long long wv,
                                        long long yv, long long &cw,
                                        long long &cy, long long *bw,
                                        long long *by) {
    if (j < 0) return;
    if (j < s0) {
        cw += wv;
        cy += yv;
    } else if (j <= hi) {
        bin_add(bw, by, j - s0, wv, yv);
    }
}

// Ranged fold of one cycle's point range [a, e) of a time-sorted curve
// into the LDS segment of bins [s0, hi] (bin s0 + l at bw[l] and by[l]),
// one block of the point-order copy per step:
// thread tid takes slot tid, which holds point nat_slot^-1(tid) of the
// block.  The range comes from lower_key with the fold of this mode, so
// every point in it has its bin r (the oracle's in parity mode) in
// [s0, hi]: the blocks wholly inside [a, e) are folded with no per-point
// test, four blocks' loads issued before the first point is folded; the
// partial blocks at either end test the point index.  PAD sends a point of
// bin r (1..os) to its padding copy r + nb - os - 1 instead, which may lie
// outside the segment (then it is skipped).  A direct point whose own bin
// is one the padding assignment overwrites (nb - os .. nb - 1) is added
// like the others; big_period clears those bins before the copies go in.
// (the curve's padded copy must span less than 2^32 bytes)
// UW: every point of the curve has the same weight; it adds wone (1: the
// weight channel counts points) without a load (16 instead of 24 bytes
// read per point and period).
template <bool FAST, bool PAD, bool UW>
__device__ __forceinline__ void big_fold_range(
    const double *tn, const long long *wn, int a, int e,
    double period, double qc, double bd, double inv_bd, double tol_x, int nb,
    int os, int s0, int hi, long long *bw, long long *by, long long wone) {
    if (a >= e) return;
    const unsigned tid = threadIdx.x;
    const int pt = (tid & 63) * NW + (tid >> 6);  // point of slot tid
    // tn: the interleaved (t', y) records, one 16-byte load a point, at a
    // 32-bit byte offset (the ranged form takes curves of fewer than 2^27
    // points); wn: the weights (8 bytes a point)
    gchar *tb = scalar_base(tn), *wb = scalar_base(wn);
    const unsigned span = (unsigned)(hi - s0), cnt = (unsigned)(e - a);
    const double pq = period * qc;
    // Parity mode: x = t * inv_bd + (tol_x - pq * inv_bd), one FMA, is within
    // u*(4*t/bd + 2*nb + 2) (u = 2^-53: the reciprocal, the product pq *
    // inv_bd, the shift and the FMA) of the oracle's (t - pq)/bd + tol_x;
    // tol_x (2^-47 times the curve's span plus its longest period, in bins)
    // exceeds that 16 times.  So an x whose fraction exceeds 2*tol_x lies in
    // the same open unit interval as the oracle's quotient and truncates to
    // its bin; the others take the oracle's division.
    const double off = tol_x - pq * inv_bd, tol2 = 2.0 * tol_x;
    // destination of bin r, relative to s0
    const int dsh = PAD ? nb - os - 1 - s0 : -s0;
    auto rough = [&](double t, bool &flag) -> int {
        if (FAST) {  // lower_key's fold, clamps included
            flag = false;
            return fold_bin_in(t, period, qc, bd, inv_bd, nb, true);
        }
        const double x = fma(t, inv_bd, off);
        flag = __builtin_amdgcn_fract(x) <= tol2;
        return (int)x + 1;
    };
    auto exact = [&](double t) -> int {
        return min(max((int)trunc((t - pq) / bd) + 1, 0), nb);
    };
    // masked: test the point index (partial blocks); PAD also the segment
    auto fold_block = [&](int B, bool masked) {
        const unsigned o = ((unsigned)B + tid) * 8u;
        const I64x2 ty = load_at<I64x2>(tb, 2u * o);
        const double t = __longlong_as_double(ty.x);
        const long long wv = UW ? wone : load_at<long long>(wb, o);
        const long long yv = ty.y;
        bool fl;
        int r = rough(t, fl);
        if (fl) r = exact(t);
        const unsigned l = (unsigned)(r + dsh);
        const bool in = (!masked || (unsigned)(B + pt - a) < cnt) &&
                        (!PAD || l <= span);
        if (in) bin_add(bw, by, min(l, span), wv, yv);
    };
    const int Bf = (a + NT - 1) & ~(NT - 1);  // first whole block
    const int Be = e & ~(NT - 1);             // end of the whole blocks
    if (Bf > Be) {  // [a, e) within one block
        fold_block(a & ~(NT - 1), true);
        return;
    }
    if (a < Bf) fold_block(Bf - NT, true);
    int B = Bf;
    for (; B + 3 * NT < Be; B += 4 * NT) {
        const unsigned o = ((unsigned)B + tid) * 8u;
        double t[4];
        long long wv[4], yv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const I64x2 ty = load_at<I64x2>(tb, 2u * o + u * NT * 16);
            t[u] = __longlong_as_double(ty.x);
            wv[u] = UW ? wone : load_at<long long>(wb, o + u * NT * 8);
            yv[u] = ty.y;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            bool fl;
            int r = rough(t[u], fl);
            if (fl) r = exact(t[u]);
            const unsigned l = (unsigned)(r + dsh);
            if (!PAD || l <= span)
                bin_add(bw, by, min(l, span), wv[u], yv[u]);
        }
    }
    for (; B < Be; B += NT) fold_block(B, false);
    if (Be < e) fold_block(Be, true);
}

// Box scan of one segment of a phase-tiled period: starts s0..nend-1 of
// every duration over the segment's prefix values Cw[l] (weight; uniform
// weights: the point count, times c1) and Cy[l] (flux) of bin s0 + l, the
// thread taking starts s0 + tid, s0 + tid + NT, ...  sb receives the
// thread's best window (the first of equal scores: its indices increase).
template <bool PACKED, bool LIKE, bool CFORM>
__device__ __forceinline__ void big_box_scan(Score &sb, const double *Cw,
                                             const double *Cy,
                                             int s0, int nend, int nb,
                                             const int *s_m, int nd,
                                             double c1, double ytn,
                                             double epsn, int ntot) {
    const int nval = nb + 1;
    // CFORM (uniform weights, one point's weight c1 at least epsn): the
    // weight tests are 0 < n < ntot on the window's point count n, and c1
    // moves into the constants, as in k_main's box_scan_pad (the same
    // scores, so both kernels rank a period's windows alike)
    const double ytc = ytn * c1, ntop = (double)ntot - 0.5;
    for (int k = 0; k < nd; ++k) {
        const int m = s_m[k];
        const int last = min(nend - 1, nb - m);
#pragma unroll 4
        for (int n = s0 + (int)threadIdx.x; n <= last; n += NT) {
            const int l = n - s0;
            const double yin = Cy[l + m] - Cy[l];
            if (CFORM) {
                const double cnt = Cw[l + m] - Cw[l];
                const double wout = fma(-c1, cnt, 1.0);
                const double num = fma(ytc, cnt, -yin);
                sb.offer_later((cnt > 0.5) & (cnt < ntop) & !(num < 0.0),
                               num * num,
                               LIKE ? cnt * wout * wout : cnt * wout,
                               k * nval + n);
                continue;
            }
            const double win = PACKED ? (Cw[l + m] - Cw[l]) * c1
                                      : Cw[l + m] - Cw[l];
            const double wout = 1.0 - win;
            const double num = fma(ytn, win, -yin);
            sb.offer_later(!(fmin(win, wout) < epsn) & !(num < 0.0),
                           num * num, LIKE ? win * wout * wout : win * wout,
                           k * nval + n);
        }
    }
}

// One period whose n_bins + 1 exceeds the LDS histogram: the start
// positions are processed in segments of S start positions; segment k keeps
// bins [s0, s0 + S + mmax] in LDS (weight and weight*flux sums in two
// arrays, 16 B per bin) on top of the exact int64 sum of all bins below s0
// (the carried prefix).
//
// Ranged form (time-sorted curves): the points of cycle c whose bin lies in
// [s0, hi] are the index range [lower_key(c, s0), lower_key(c, hi + 1)), so
// a segment reads only its own points, all 256 threads stride each range in
// point order, and the carried prefix is the previous segment's scan value
// at its last start position.  The range table (Q+1 cycles) takes the tail
// of the histogram.  Otherwise every segment folds every point again and
// sums the ones below s0.
__device__ void big_period(const Batch &b, const Out &o, const Work &w,
                           const LcParams &P, long long p, int nb,
                           int lds_bins, long long *bins, const int *s_m,
                           int nd, int mmax, long long *lscr, double *dscr,
                           int *iscr, double *dscr2, double *dscr3,
                           long long *s_carry) {
    const int tid = threadIdx.x, os = b.oversample;
    const double period = b.periods[p], invP = 1.0 / period;
    const double bd = P.bd, inv_bd = 1.0 / bd;
    const int nval = nb + 1;
    // the bins of the mode's fold, so that results do not depend on which
    // periods the tiling sends here
    const bool fastm = b.mode == ATG_MODE_FAST;
    const bool like = b.objective == ATG_OBJ_LIKELIHOOD;
    const double Wt = P.W_tot, Yt = P.Y_tot;
    const double ytn = Yt / Wt, epsn = DBL_EPSILON / Wt;
    const double fw = P.w_unit / Wt, fy = P.y_unit / Wt;
    const double c1 = (double)P.wq_one * (P.w_unit / P.W_tot);
    // ranged form: cycles 0..Q of the sorted points
    const long long nbase = nat_base(b, P.lc_self);
    const double *tn = w.nat_ty + 2 * nbase;  // interleaved (t', y) records
    int ncyc = 0;
    if (P.natural)
        ncyc = (int)fold_cycle(tn[2 * nat_slot(P.n - 1)], period, invP,
                               fastm) + 1;
    const bool ranged = P.natural && ncyc <= (1 << 20) &&
                        P.n < (1 << 27) &&
                        lds_bins - ncyc - mmax - 1 >= 64;
    int S = ranged ? lds_bins - ncyc - mmax - 1 : lds_bins - mmax - 1;
    // The range table takes R 16-byte bins at the tail of the tile.  Where
    // it costs at most an eighth of the tile, the ranges of every segment
    // are searched at once, before the first segment (one round of binary
    // searches per period instead of one per segment): [k][c][2] for
    // segment k and cycle c, then the padding ranges [c][2]; otherwise
    // each segment searches its own ranges into [c][2].
    int R = ncyc, nseg = 0;
    bool batched = false;
    if (ranged) {
        const int nseg0 = (nb + S - 1) / S;
        const int Rb = (2 * ncyc * (nseg0 + 2) + 3) / 4;
        const int Sb = lds_bins - Rb - mmax - 1;
        if (Sb >= 64 && Rb <= lds_bins / 8 && (nb + Sb - 1) / Sb <= nseg0 + 1) {
            batched = true;
            R = Rb;
            S = Sb;
            nseg = (nb + Sb - 1) / Sb;
        }
    }
    int *tab = (int *)(bins + 2 * (lds_bins - R));
    // the tile: weights (uniform weights: point counts) in bw, weight*flux
    // sums in by, TB bins each, then the range table
    const int TB = lds_bins - R;
    long long *bw = bins, *by = bins + TB;
    int *erng = tab + 2 * ncyc * (batched ? nseg : 1);
    const long long *wn = w.nat_w + nbase;
    Score best;
    double bwin = 0.0, byin = 0.0;
    const double *tp = w.ws_t + P.ws_off + tid;
    const long long *wp = w.ws_w + P.ws_off + tid;
    const long long *yp = w.ws_y + P.ws_off + tid;
    // the prefix values, in place: C[2i] the weight (uniform weights: the
    // point count), C[2i + 1] the weight*flux sum
    double *Cw = (double *)bw, *Cy = (double *)by;
    if (tid == 0) {
        s_carry[0] = 0;
        s_carry[1] = 0;
    }
    if (batched) {
        const int ntask = 2 * ncyc * (nseg + 1);
        for (int e = tid; e < ntask; e += NT) {
            const int k = e / (2 * ncyc), c = (e - k * 2 * ncyc) >> 1;
            const int sb = k < nseg ? ((e & 1) ? min(k * S + S + mmax, nb) + 1
                                               : k * S)
                                    : ((e & 1) ? os + 1 : 1);
            tab[e] = lower_key(tn, P.n, period, invP, bd, inv_bd, nb,
                               (double)c, sb, fastm);
        }
    }
    __syncthreads();
    for (int s0 = 0, seg_k = 0; s0 < nb; s0 += S, ++seg_k) {
        const int hi = min(s0 + S + mmax, nb), len = hi - s0 + 1;
        int *rng = tab + (batched ? 2 * ncyc * seg_k : 0);
        long long cw = 0, cy = 0;
        if (ranged) {
            cw = s_carry[0];
            cy = s_carry[1];
        }
        for (int i = tid; i < len; i += NT) {
            bw[i] = 0;
            by[i] = 0;
        }
        const bool extra = hi >= nb - os;  // padding copies land here
        if (ranged && !batched) {
            const int ntask = (extra ? 4 : 2) * ncyc;
            for (int e = tid; e < ntask; e += NT) {
                const int c = (e >> 1) % ncyc;
                const bool ex = e >= 2 * ncyc;
                const int sb = ex ? ((e & 1) ? os + 1 : 1)
                                  : ((e & 1) ? hi + 1 : s0);
                const int v = lower_key(tn, P.n, period, invP, bd, inv_bd,
                                        nb, (double)c, sb, fastm);
                (ex ? erng : rng)[2 * c + (e & 1)] = v;
            }
        }
        __syncthreads();
        if (ranged) {
            // direct destinations: bins [s0, hi] of every cycle; then the
            // padding copies, bins 1..os of every cycle to nb-os..nb-1
#define ATG_RANGE(F, PD, U, J0, E)                                           \
    big_fold_range<F, PD, U>(tn, wn, J0, E, period, (double)c, bd,           \
                             inv_bd, P.tol_x, nb, os, s0, hi, bw, by, 1)
#define ATG_RANGES(PD, TAB)                                                  \
    for (int c = 0; c < ncyc; ++c) {                                         \
        const int j0 = TAB[2 * c], e = TAB[2 * c + 1];                       \
        if (fastm) {                                                         \
            if (P.uniform) ATG_RANGE(true, PD, true, j0, e);                 \
            else ATG_RANGE(true, PD, false, j0, e);                          \
        } else {                                                             \
            if (P.uniform) ATG_RANGE(false, PD, true, j0, e);                \
            else ATG_RANGE(false, PD, false, j0, e);                         \
        }                                                                    \
    }
            ATG_RANGES(false, rng)
            if (extra) {
                // the padding assignment: bins nb-os..nb-1 hold only the
                // copies of bins 1..os
                __syncthreads();
                const int z0 = max(nb - os, s0), z1 = min(nb - 1, hi);
                for (int i = z0 + tid; i <= z1; i += NT) {
                    bw[i - s0] = 0;
                    by[i - s0] = 0;
                }
                __syncthreads();
                ATG_RANGES(true, erng)
            }
#undef ATG_RANGES
#undef ATG_RANGE
            __syncthreads();
        } else {
            for (int i = 0; i < P.L; ++i) {
                long long wv = wp[(long long)i * NT];
                const long long yv = yp[(long long)i * NT];
                if (P.uniform) wv = wv != 0 ? 1 : 0;  // point counts
                if ((wv | yv) == 0) continue;
                const double tv = tp[(long long)i * NT];
                const int r =
                    fastm ? fold_bin_in(tv, period, floor(tv * invP), bd,
                                        inv_bd, nb, true)
                          : min(max(fold_bin(tv, period, invP, bd, inv_bd), 0),
                                nb);
                int j1, j2;
                pad_targets(r, nb, os, j1, j2);
                seg_add(j1, s0, hi, wv, yv, cw, cy, bw, by);
                seg_add(j2, s0, hi, wv, yv, cw, cy, bw, by);
            }
            __syncthreads();
            cw = block_reduce(cw, AddOp<long long>(), lscr);
            cy = block_reduce(cy, AddOp<long long>(), lscr);
        }
        // the segment's prefix and box scan above the normal priority, as
        // k_main's finish (long-baseline stage -1%)
        __builtin_amdgcn_s_setprio(1);
        const int seg = (len + NT - 1) / NT;
        const int b0 = min(tid * seg, len), b1 = min(b0 + seg, len);
        long long sw = 0, sy = 0;
        for (int i = b0; i < b1; ++i) {
            sw += bw[i];
            sy += by[i];
        }
        long long tw, ty;
        long long ow = block_excl_scan_dpp(sw, lscr, &tw) + cw;
        long long oy = block_excl_scan_dpp(sy, lscr, &ty) + cy;
        for (int i = b0; i < b1; ++i) {
            ow += bw[i];
            oy += by[i];
            if (i == S - 1) {  // prefix of the next segment's first bin
                s_carry[0] = ow;
                s_carry[1] = oy;
            }
            // uniform weights: the point count, as in finish_period (the
            // weight channel counts points)
            Cw[i] = P.uniform ? (double)ow : (double)ow * fw;
            Cy[i] = (double)oy * fy;
        }
        __syncthreads();
        const int nend = min(s0 + S, nb);
        // duration-major, as in finish_period: a thread meets the segment's
        // windows in increasing index, so the segment's best needs no tie
        // rule; it joins the period's best with the full one (indices are
        // not monotone across segments)
        Score sb;
#define ATG_BBS(U, L, CF)                                                    \
    big_box_scan<U, L, CF>(sb, Cw, Cy, s0, nend, nb, s_m, nd, c1, ytn, epsn,  \
                           P.n)
        if (P.uniform && c1 >= epsn) {
            if (like) ATG_BBS(true, true, true);
            else ATG_BBS(true, false, true);
        } else if (P.uniform) {
            if (like) ATG_BBS(true, true, false);
            else ATG_BBS(true, false, false);
        } else {
            if (like) ATG_BBS(false, true, false);
            else ATG_BBS(false, false, false);
        }
#undef ATG_BBS
        if (sb.idx != INT_MAX) {
            const int before = best.idx;
            best.offer(sb.q, sb.den, sb.idx);
            if (best.idx != before) {  // the window's sums, read back
                const int k = sb.idx / nval, n = sb.idx - k * nval;
                const int l = n - s0, m = s_m[k];
                bwin = P.uniform ? (Cw[l + m] - Cw[l]) * c1 : Cw[l + m] - Cw[l];
                byin = Cy[l + m] - Cy[l];
            }
        }
        __syncthreads();
        __builtin_amdgcn_s_setprio(0);
    }
    for (int ofs = 32; ofs > 0; ofs >>= 1) {
        const double oq = __shfl_xor(best.q, ofs), od = __shfl_xor(best.den, ofs);
        const int oi = __shfl_xor(best.idx, ofs);
        const double ow_ = __shfl_xor(bwin, ofs), oy_ = __shfl_xor(byin, ofs);
        const int before = best.idx;
        best.offer(oq, od, oi);
        if (best.idx != before) {
            bwin = ow_;
            byin = oy_;
        }
    }
    const int wv = tid >> 6, l = tid & 63;
    __syncthreads();
    if (l == 0) {
        dscr[wv] = best.q;
        dscr[NW + wv] = best.den;
        iscr[wv] = best.idx;
        dscr2[wv] = bwin;
        dscr3[wv] = byin;
    }
    __syncthreads();
    if (tid == 0) {
        for (int i = 1; i < NW; ++i) {
            const int before = best.idx;
            best.offer(dscr[i], dscr[NW + i], iscr[i]);
            if (best.idx != before) {
                bwin = dscr2[i];
                byin = dscr3[i];
            }
        }
        const int bi = best.idx;
        if (bi == INT_MAX) {
            o.power[p] = -INFINITY;
            o.depth[p] = 0.0;
            o.depth_err[p] = 0.0;
            o.duration[p] = 0.0;
            o.transit_time[p] = P.t_ref;
            o.depth_snr[p] = 0.0;
            o.log_like[p] = 0.0;
        } else {
            const int k = bi / nval, n = bi - k * nval, m = s_m[k];
            const double win = bwin * Wt;
            const double yin_s = byin * Wt;
            const double wout = Wt - win;
            const double y_in = yin_s / win;
            const double y_out = (Yt - yin_s) / wout;
            const double depth = y_out - y_in;
            const double derr = sqrt(1.0 / win + 1.0 / wout);
            const double snr = depth / derr;
            const double ll = 0.5 * win * depth * depth;
            const double dur = m * P.bd;
            o.power[p] = like ? ll : snr;
            o.depth[p] = depth;
            o.depth_err[p] = derr;
            o.duration[p] = dur;
            o.transit_time[p] = fold_mod(n * P.bd + 0.5 * dur + 0.0, period) +
                                P.t_ref;
            o.depth_snr[p] = snr;
            o.log_like[p] = ll;
        }
    }
    __syncthreads();
}

// Four waves per SIMD (at most 128 VGPRs): the fold, the LDS atomics and
// the box scan's loads are latency-bound at three; four workgroups of
// lds_bins <= 2528 share a CU's 160 KiB of LDS.
// EXACT: parity mode's fold (fast mode's otherwise); LIKE: the likelihood
// objective (the snr otherwise).
// Zeroes n bytes of LDS at a (a multiple of 4), 16 bytes per store where a
// is 16-byte aligned; the workgroup's threads share the stores.
__device__ __forceinline__ void lds_zero(void *a, int n) {
    unsigned char *c = (unsigned char *)a;
    int done = 0;
    if (((unsigned)(size_t)c & 15u) == 0) {
        const int n16 = n >> 4;
        for (int i = threadIdx.x; i < n16; i += NT)
            ((int4 *)c)[i] = make_int4(0, 0, 0, 0);
        done = n16 << 4;
    }
    for (int i = done + 4 * (int)threadIdx.x; i < n; i += 4 * NT)
        *(int *)(c + i) = 0;
}

template <bool EXACT, bool LIKE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4, 4)))
void k_main(Batch b, Out o, Work w, int per_item, int lds_bins) {
    extern __shared__ __align__(16) unsigned char smem[];
    long long *binW = (long long *)smem;
    long long *binY = binW + lds_bins;
    // duration bins; int16 is exact here: this kernel takes only periods of
    // at most 2 * lds_bins < 32767 bins, and a window of m >= nb bins has
    // no start, so larger m are clamped without effect
    __shared__ short s_m[DMAX];
    __shared__ unsigned int s_item;
    __shared__ int s_queue;
    __shared__ long long lscr[2 * NW];  // two scans' scratch
    __shared__ double dscr[2 * NW];
    __shared__ int iscr[NW];

    const int tid = threadIdx.x;
    const long long n_lc = b.n_lc;
    const int home = blockIdx.x % NQ;
    int queue = home;
    for (;;) {
        __syncthreads();
        if (tid == 0) {
            // own XCD's queue first, then steal from the others in turn
            int qq = queue, left = NQ;
            unsigned int it = 0xffffffffu;
            while (left > 0) {
                // the current queue is tried at once; a queue to steal from
                // is passed over with a load, not an atomic, once it is
                // exhausted (at the end every workgroup visits all of them)
                if (left == NQ ||
                    (long long)__hip_atomic_load(&w.qnext[qq], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) <
                        w.qtot[qq]) {
                    it = atomicAdd(&w.qnext[qq], 1u);
                    if ((long long)it < w.qtot[qq]) break;
                }
                it = 0xffffffffu;
                qq = (qq + 1) % NQ;
                --left;
            }
            s_item = it;
            s_queue = qq;
        }
        __syncthreads();
        const unsigned int item = s_item;
        if (item == 0xffffffffu) break;
        queue = s_queue;
        // curve of the item: last curve of the queue whose first item <= item
        const long long nq = (n_lc - queue + NQ - 1) / NQ;
        long long lo = 0, hi = nq;
        while (hi - lo > 1) {
            long long mid = (lo + hi) >> 1;
            if ((unsigned long long)w.item_off[queue + NQ * mid] <= item)
                lo = mid;
            else
                hi = mid;
        }
        const int lc = (int)(queue + NQ * lo);
        const LcParams P = w.prm[lc];
        const long long pbeg = b.per_off[lc] +
                               (long long)(item - w.item_off[lc]) * per_item;
        const long long pend = min(pbeg + per_item, b.per_off[lc + 1]);
        if (P.status != ATG_LC_OK) {
            for (long long p = pbeg + tid; p < pend; p += NT) write_nan(o, p);
            continue;
        }
        const int nd = P.n_dur;
        const long long dof = b.dur_off[lc];
        if (tid < nd)
            s_m[tid] = (short)fmin(round(b.durs[dof + tid] / P.bd), 32767.0);
        __syncthreads();
        int mmax = 0;
        for (int k = 0; k < nd; ++k) mmax = max(mmax, (int)s_m[k]);

        const double bd = P.bd;
        const double inv_bd = 1.0 / bd;
        const int os = b.oversample;
        // lane j of each wave prepares the item's period pbeg + j (an item
        // holds at most 64): its bins, and the exact fold's P/bd and 1/P
        // scaled up by 1 + 2^-48 (see bin_points); the passes read them
        // back as uniform values
        const int lj = tid & 63;
        double per_l = 1.0, inv_l = 1.0;
        int nb_l = 0;
        if (pbeg + lj < pend) {
            const double ps = b.periods[pbeg + lj];
            nb_l = (int)ceil(ps / bd) + os;
            per_l = EXACT ? ps * inv_bd : ps;
            inv_l = EXACT ? (1.0 / ps) * (1.0 + 0x1p-48) : 1.0 / ps;
        }

        long long p = pbeg;
        while (p < pend) {
            // greedy sub-batch of periods that fit in the LDS histogram.  A
            // packed histogram takes 8 B per bin, so both halves of the LDS
            // hold bins; each packed period also needs (nb + 2) / 2 free
            // slots above it for its point counts (finished last to first).
            // Counted bins (uniform weights, unpacked) take 12 B: int64
            // flux sums from the start of the LDS, int32 counts after
            // lds_bins * 4 / 3 of them.
            const bool packed = P.packed;
            const bool counted = !packed && P.uniform;
            const int cap2 = counted_cap(lds_bins);
            const int cap = packed ? 2 * lds_bins : (counted ? cap2 : lds_bins);
            int *cntb = (int *)(binW + cap2);
            // the pass's slot s is the item's period l0 + s (lane l0 + s)
            const int l0 = (int)(p - pbeg);
            // a pass whose first period leaves room for four of its size,
            // each with the room for its own point counts (packed), takes
            // periods so, and its waves finish them side by side
            // (finish_period_wave); longer periods take the workgroup's
            // finish, one at a time
            const int nb0 = __builtin_amdgcn_readlane(nb_l, l0);
            const bool wmode =
                NW * (nb0 + 1 + (packed ? (nb0 + 2) / 2 : 0)) <= cap;
            double per[SUBMAX], inv[SUBMAX];
            int nb[SUBMAX], base[SUBMAX];
            int cnt = 0, used = 0;
#pragma unroll
            for (int s = 0; s < SUBMAX; ++s) {
                per[s] = 1.0;
                inv[s] = 1.0;
                nb[s] = 0;
                base[s] = 0;
            }
#pragma unroll
            for (int s = 0; s < SUBMAX; ++s) {
                if (cnt == s && p + s < pend) {
                    const int nbs = __builtin_amdgcn_readlane(nb_l, l0 + s);
                    // the period on top also needs room for its box-scan
                    // sentinels: the overhang after its point counts
                    // (packed) or weights (unpacked)
                    const int ov = scan_overhang(nbs, mmax);
                    const int foot =
                        wmode ? nbs + 1 + (packed ? (nbs + 2) / 2 : 0)
                              : nbs + 1;
                    const int need =
                        wmode ? foot
                              : nbs + 1 + (packed ? max((nbs + 1 + ov + 1) / 2,
                                                        ov)
                                                  : ov);
                    if (used + need <= cap) {
                        per[s] = readlane_d(per_l, l0 + s);
                        inv[s] = readlane_d(inv_l, l0 + s);
                        nb[s] = nbs;
                        base[s] = used;
                        used += foot;
                        ++cnt;
                    }
                }
            }
            if (cnt == 0) {  // this period alone exceeds the LDS histogram:
                             // k_big takes it (its phase tiles need more
                             // registers than this kernel's four waves
                             // per SIMD leave)
                if (tid == 0) {
                    const int q = lc % NQ;
                    const long long j =
                        w.big_base[q] + atomicAdd(&w.big_count[q], 1u);
                    w.big_p[j] = p;
                    w.big_lc[j] = lc;
                }
                ++p;
                continue;
            }
            // the pass's binning at the normal priority, its finish above it
            // (latency-bound chains of LDS reads: ready finish instructions
            // issue ahead of the other workgroups' binning on the SIMD;
            // stage 48.1 -> 47.4 ms at the headline shape)
            __builtin_amdgcn_s_setprio(0);
            __syncthreads();  // the previous pass's last reads
            lds_zero(binW, used * 8);
            if (counted) lds_zero(cntb, used * 4);
            else if (!packed) lds_zero(binY, used * 8);
            __syncthreads();

            // fold + binning over the lane's points: the lane-major layout
            // puts the 64 lanes of a wave ~N/256 points apart in time, so
            // their bins rarely collide and every point goes straight to
            // the LDS histogram (padding slots carry zero weights)
            {
                // uniform bases; each lane adds its 32-bit slot offset
                const double *tp = w.ws_t + P.ws_off;
                const long long *wp = w.ws_w + P.ws_off;
                const long long *yp = w.ws_y + P.ws_off;
                const int L = P.L;
                // counted bins: the counts take the weight channel
                long long *hW = counted ? (long long *)cntb : binW;
                long long *hY = counted ? binW : binY;
#define ATG_BIN(C, E, K)                                                    \
    case C:                                                                 \
        bin_points<C, E, K>(tp, wp, yp, L, per, inv, nb, base, bd, inv_bd,  \
                            P.tol_x, hW, hY, b.periods + p, L * NT - P.n,   \
                            P.t_pad);                                       \
        break;
#define ATG_BIN8(E, K)                                                      \
    switch (cnt) {                                                          \
        ATG_BIN(1, E, K) ATG_BIN(2, E, K) ATG_BIN(3, E, K) ATG_BIN(4, E, K) \
        ATG_BIN(5, E, K) ATG_BIN(6, E, K) ATG_BIN(7, E, K) ATG_BIN(8, E, K) \
        default: break;                                                     \
    }
                if (packed) ATG_BIN8(EXACT, 1)
                else if (P.uniform) ATG_BIN8(EXACT, 2)
                else ATG_BIN8(EXACT, 0)
#undef ATG_BIN8
#undef ATG_BIN
            }
            __syncthreads();
            __builtin_amdgcn_s_setprio(1);

            if (wmode) {
                // wave w finishes the pass's periods w, w + 4, ...
                const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
                int bse = 0;
                for (int s = 0; s < cnt; ++s) {
                    const int nbs = __builtin_amdgcn_readlane(nb_l, l0 + s);
                    if ((s & (NW - 1)) == wv) {
                        if (packed)
                            finish_period_wave<LIKE>(b, o, P, p + s, nbs, bse,
                                                     binW, binW, nullptr, s_m,
                                                     nd);
                        else
                            finish_period_wave<LIKE>(
                                b, o, P, p + s, nbs, bse, binW, binY,
                                counted ? cntb : nullptr, s_m, nd);
                    }
                    bse += nbs + 1 + (packed ? (nbs + 2) / 2 : 0);
                }
            } else if (packed) {
                int end = used;
                for (int s = cnt - 1; s >= 0; --s) {
                    const int nbs = __builtin_amdgcn_readlane(nb_l, l0 + s);
                    const int bse = end - (nbs + 1);
                    finish_period<LIKE>(b, o, P, p + s, nbs, bse, binW, binW,
                                        nullptr, s_m, nd, mmax, lscr, dscr,
                                        iscr);
                    end = bse;
                }
            } else {
                // last to first too: a period's sentinels overwrite only
                // periods already finished
                int end = used;
                for (int s = cnt - 1; s >= 0; --s) {
                    const int nbs = __builtin_amdgcn_readlane(nb_l, l0 + s);
                    const int bse = end - (nbs + 1);
                    finish_period<LIKE>(b, o, P, p + s, nbs, bse, binW, binY,
                                        counted ? cntb : nullptr, s_m, nd, mmax,
                                        lscr, dscr, iscr);
                    end = bse;
                }
            }
            p += cnt;
        }
    }
}

}  // namespace

// ------------------------------------------------------------------- k_big

// Periods whose bins exceed the LDS histogram, queued by k_main, phase tile
// by phase tile (big_period).  Persistent workgroups drain the list of
// their own XCD's queue first (the curve's points are in that XCD's L2),
// then the others, so that a workload of many long periods stays balanced.
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4, 4)))
void k_big(Batch b, Out o, Work w, int lds_bins) {
    extern __shared__ __align__(16) unsigned char smem[];
    long long *bins = (long long *)smem;  // 2 * lds_bins int64 (the tile)
    __shared__ int s_m[DMAX];
    __shared__ unsigned int s_j;
    __shared__ long long lscr[NW];
    __shared__ double dscr[2 * NW], dscr2[NW], dscr3[NW];
    __shared__ int iscr[NW];
    __shared__ long long s_carry[2];
    const int tid = threadIdx.x;
    int q = blockIdx.x % NQ, empty = 0;
    while (empty < NQ) {
        __syncthreads();
        if (tid == 0) {
            // an exhausted list (all of them, mostly) costs a load, not an
            // atomic on one of eight counters shared by every workgroup
            const unsigned int cnt = w.big_count[q];
            s_j = __hip_atomic_load(&w.big_next[q], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) < cnt
                      ? atomicAdd(&w.big_next[q], 1u)
                      : cnt;
        }
        __syncthreads();
        const unsigned int j = s_j;
        if (j >= w.big_count[q]) {  // list q is done (k_main has ended)
            q = (q + 1) % NQ;
            ++empty;
            continue;
        }
        empty = 0;
        const long long e = w.big_base[q] + j;
        const long long p = w.big_p[e];
        const int lc = w.big_lc[e];
        const LcParams P = w.prm[lc];
        const int nd = P.n_dur;
        const long long dof = b.dur_off[lc];
        if (tid < nd) s_m[tid] = (int)round(b.durs[dof + tid] / P.bd);
        __syncthreads();
        int mmax = 0;
        for (int k = 0; k < nd; ++k) mmax = max(mmax, s_m[k]);
        const int os = b.oversample;
        const int nbp = (int)ceil(b.periods[p] / P.bd) + os;
        if (nbp - os > os && lds_bins - mmax - 1 >= 1) {
            big_period(b, o, w, P, p, nbp, lds_bins, bins, s_m, nd,
                       mmax, lscr, dscr, iscr, dscr2, dscr3, s_carry);
        } else if (tid == 0) {
            write_nan(o, p);
            atomicOr(&w.lc_flags[lc], 1);
        }
    }
}

// ---------------------------------------------------------------- k_fields
//
// The seven result fields of the periods whose winner k_main found: k_main
// leaves the winner's window weight W_in and flux sum in depth and
// depth_err and -(index + 1) in duration, and this pass evaluates the
// oracle's formulas for them in parallel over periods (astropy
// timeseries/periodograms/bls/methods.py:110-139, core.py:719).  Periods
// finished elsewhere (the phase tiles of k_big, no qualifying window,
// rejected curves) hold a duration >= 0 or NaN and are left alone.
__global__ __launch_bounds__(NT) void k_fields(Batch b, Out o, Work w) {
    const int lc = blockIdx.x;
    const LcParams P = w.prm[lc];
    if (P.status != ATG_LC_OK) return;
    const long long po = b.per_off[lc], np = b.per_off[lc + 1] - po;
    const long long dof = b.dur_off[lc];
    const bool like = b.objective == ATG_OBJ_LIKELIHOOD;
    const double Wt = P.W_tot, Yt = P.Y_tot;
    for (long long i = threadIdx.x; i < np; i += NT) {
        const long long p = po + i;
        const double mark = o.duration[p];
        if (!(mark < 0.0)) continue;
        const int bi = (int)(-mark - 1.0);
        const double period = b.periods[p];
        const int k = bi >> 16, s = bi & 0xffff;  // k_main's window index
        const int m =
            (short)fmin(round(b.durs[dof + k] / P.bd), 32767.0);  // as k_main
        const double win = o.depth[p], yin_s = o.depth_err[p];
        const double wout = Wt - win;
        const double y_in = yin_s / win;
        const double y_out = (Yt - yin_s) / wout;
        const double depth = y_out - y_in;
        const double derr = sqrt(1.0 / win + 1.0 / wout);
        const double snr = depth / derr;
        const double ll = 0.5 * win * depth * depth;
        const double dur = m * P.bd;
        const double phase = fold_mod(s * P.bd + 0.5 * dur + 0.0, period);
        o.power[p] = like ? ll : snr;
        o.depth[p] = depth;
        o.depth_err[p] = derr;
        o.duration[p] = dur;
        o.transit_time[p] = phase + P.t_ref;
        o.depth_snr[p] = snr;
        o.log_like[p] = ll;
    }
}

// ------------------------------------------------------- k_cands / k_exact
//
// The int64 bins make every window sum exact, so two periods whose best
// windows hold the same points get exactly equal power.  The oracle sums in
// FP64 in point order, and its rounding decides such ties (and near-ties
// below ~1e-13).  In parity mode the periods of each curve whose power is
// within TIE_REL of the curve's maximum are therefore recomputed with the
// oracle's own operation order: per-bin FP64 sums in point order, the
// padding assignment, the sequential cumulative sum and the box scan with
// the oracle's formulas.  Their outputs then equal the oracle's bit for bit
// and the argmax follows the oracle's.

namespace {

// Periods whose best windows hold the same points have equal int64 power,
// so the near-tie set can be a long plateau of neighbouring periods (59 at
// the 5,000-period benchmark grid, every period of a sparse curve whose best
// window holds one deep point); all of it is recomputed, or an unrefined
// member (exact integer value) could outrank the refined ones.  The queue
// holds up to every period of the batch, so nothing is dropped.
constexpr double TIE_REL = 1e-9;   // >> relative error of the int64 path
// k_exact's workgroups: the sums are chains in point order, so the kernel is
// latency-bound; each queued period gets EXW waves (a quarter of its bins,
// or a segment, each), three workgroups to a CU (its ~170 VGPRs allow three
// waves per SIMD; capped at 128 it spills and runs slower).
constexpr int EXACT_WG_PER_CU = 3;
constexpr int EXACT_TAB = 1024;  // the ranged form's table of range bounds

__global__ __launch_bounds__(NT) void k_cands(Batch b, Out o, Work w) {
    __shared__ double dscr[NW];
    __shared__ int iscr[NW];
    __shared__ unsigned int s_base;
    const int lc = blockIdx.x;
    const long long po = b.per_off[lc], np = b.per_off[lc + 1] - po;
    if (b.mode != ATG_MODE_PARITY || w.prm[lc].status != ATG_LC_OK ||
        w.lc_flags[lc])
        return;
    const LcParams P = w.prm[lc];
    // Relative error bound of a period's int64 power.  Unpacked bins: far
    // below TIE_REL (0 here).  Packed bins (uniform weight w): every yq is
    // within half a quantum q = y_unit of y*w, so the window statistic
    // num = W_in*Y_tot/W_tot - Y_in is off by at most n_in*q with
    // n_in = W_in/w, while num = depth*W_in*W_out/W_tot; the power (square
    // of num for the likelihood, proportional to it for the snr) is then off
    // by at most 2.5*q*W_tot/(w*|depth|*W_out) relative.  W_in follows from
    // the period's own log likelihood 0.5*W_in*depth^2.
    const double w1 = (double)P.wq_one * P.w_unit;
    auto bound = [&](long long i) -> double {
        if (!P.packed) return 0.0;
        const double dep = o.depth[i];
        if (!(fabs(dep) > 0.0)) return 0.0;
        const double w_in = 2.0 * o.log_like[i] / (dep * dep);
        const double w_out = fmax(P.W_tot - w_in, 1e-300);
        return 2.5 * P.y_unit * P.W_tot / (w1 * fabs(dep) * w_out);
    };
    // lower bound of the exact maximum; a period is a candidate when its
    // upper bound reaches it (TIE_REL covers the oracle's own FP64 rounding)
    double m = -INFINITY;
    for (long long i = threadIdx.x; i < np; i += NT) {
        const double v = o.power[po + i];
        if (v > -INFINITY) {  // NaN and -inf never count
            const double lo = v - bound(po + i) * fabs(v);
            if (lo > m) m = lo;
        }
    }
    m = block_reduce(m, MaxOp(), dscr);
    if (!(m > -INFINITY)) return;
    const double thr = m - TIE_REL * fabs(m);
    auto cand = [&](long long i) -> bool {
        if (i >= np) return false;
        const double v = o.power[po + i];
        return v > -INFINITY && v + bound(po + i) * fabs(v) >= thr;
    };
    // count, reserve the curve's slots, then write them in period order
    int count = 0;
    for (long long i = threadIdx.x; i < np; i += NT) count += cand(i) ? 1 : 0;
    count = block_reduce(count, AddOp<int>(), iscr);
    if (threadIdx.x == 0) s_base = atomicAdd(w.cand_count, (unsigned)count);
    __syncthreads();
    const unsigned int base0 = s_base;
    int done = 0;
    for (long long base = 0; base < np; base += NT) {
        const long long i = base + threadIdx.x;
        const int f = cand(i) ? 1 : 0;
        int tot;
        const int e = block_excl_scan(f, iscr, &tot);
        if (f) {
            w.cand_lc[base0 + done + e] = lc;
            w.cand_p[base0 + done + e] = po + i;
        }
        done += tot;
    }
}

// A compile-time flag passed as a value (to a generic lambda).
template <bool B>
struct BoolC {
    static constexpr bool value = B;
};

// A raw point of k_exact's walk (range index k for the ranged form), and how
// many chunks of 64 points ahead of the current one the walk loads.
struct ExactPt {
    double t = 0.0, y = 0.0, d = 1.0;
    int k = 0;
};
constexpr int EXACT_AHEAD = 4;

// One wave's chunk of points added into the LDS bins my/mi so that every
// bin receives its points in chunk order (the oracle's point order): lane j
// (j < cnt) adds (yv, iv) to bin dst (-1: to none).  (cur, ay, ai) is the
// bin whose running sums the item-by-item walk keeps in registers (cur -1:
// none).  The chunk is cut before every lane whose destination is below its
// predecessor's (a cycle wrap of time-sorted points, or a point that goes
// nowhere after others): within each piece the destinations never decrease,
// so each bin's points are one run of lanes, preceded only by the points
// that go nowhere.  The chunk's values are staged in LDS and the first lane
// of every run adds its run's values to the bin's sums one after another,
// all runs of a piece at once; the values are read eight at a time ahead of
// their adds (a dense curve puts dozens of points in a bin per cycle, and
// one LDS round trip per add would dominate).  The pieces follow one
// another.  A chunk of more than four pieces is walked point by point.  The
// sums are the same either way, add for add.
__device__ __forceinline__ void exact_add_chunk(int dst, double yv, double iv,
                                                int cnt, double *my,
                                                double *mi, double *st_y,
                                                double *st_i, int &cur,
                                                double &ay, double &ai) {
    const int lane = threadIdx.x & 63;
    const bool act = lane < cnt;
    const int prev = __shfl_up(dst, 1);
    const bool down = act && lane > 0 && dst < prev;
    unsigned long long cuts = __ballot(down);
    if (__popcll(cuts) <= 3) {
        if (cur >= 0) {
            if (lane == 0) {
                my[cur] = ay;
                mi[cur] = ai;
            }
            cur = -1;
        }
        st_y[lane] = yv;
        st_i[lane] = iv;
        __builtin_amdgcn_wave_barrier();
        int lo = 0;
        for (;;) {
            const int hi = cuts ? __builtin_ctzll(cuts) : cnt;
            const bool in = lane >= lo && lane < hi;
            const bool start = in && (lane == lo || dst != prev);
            const unsigned long long starts = __ballot(start);
            if (start && dst >= 0) {
                const unsigned long long later =
                    lane < 63 ? starts >> (lane + 1) : 0ull;
                const int end = later ? lane + 1 + __builtin_ctzll(later) : hi;
                double a = my[dst], c = mi[dst];
                for (int k = lane; k < end; k += 8) {
                    double vy[8], vi[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int j = min(k + u, 63);
                        vy[u] = st_y[j];
                        vi[u] = st_i[j];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (k + u < end) {
                            a += vy[u];
                            c += vi[u];
                        }
                    }
                }
                my[dst] = a;
                mi[dst] = c;
            }
            // the piece's sums are stored before the next piece reads them
            // (one wave: its LDS accesses complete in order)
            __builtin_amdgcn_wave_barrier();
            if (!cuts) break;
            lo = hi;
            cuts &= cuts - 1;
        }
        return;
    }
    for (int j = 0; j < cnt; ++j) {
        const int tj = __builtin_amdgcn_readlane(dst, j);
        if (tj < 0) continue;
        const double yvj = readlane_d(yv, j);
        const double ivj = readlane_d(iv, j);
        if (tj != cur) {
            if (cur >= 0 && lane == 0) {
                my[cur] = ay;
                mi[cur] = ai;
            }
            cur = tj;
            ay = my[tj];
            ai = mi[tj];
        }
        ay += yvj;
        ai += ivj;
    }
}

// Sequential inclusive prefix sums of y[0, len) and v[0, len) in place (one
// thread), continuing from the previous segment's totals (cy, cv) unless
// first.  Bins are read eight at a time, one group ahead of the other
// group's adds, and a group is stored after its adds (a store waits for its
// add, and the wave issues in order); the reads run up to 7 past len, into
// the next region or the arrays' padding, and are unused there.
__device__ __forceinline__ void exact_prefix(double *y, double *v, int len,
                                             bool first, double cy,
                                             double cv) {
    cy = first ? y[0] : cy + y[0];
    cv = first ? v[0] : cv + v[0];
    y[0] = cy;
    v[0] = cv;
    int l = 1;
    double ay[8], av[8], by[8], bv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        ay[u] = y[l + u];
        av[u] = v[l + u];
    }
    for (; l + 16 <= len; l += 16) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            by[u] = y[l + 8 + u];
            bv[u] = v[l + 8 + u];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            cy = cy + ay[u];
            cv = cv + av[u];
            ay[u] = cy;
            av[u] = cv;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            y[l + u] = ay[u];
            v[l + u] = av[u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            ay[u] = y[l + 16 + u];
            av[u] = v[l + 16 + u];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            cy = cy + by[u];
            cv = cv + bv[u];
            by[u] = cy;
            bv[u] = cv;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            y[l + 8 + u] = by[u];
            v[l + 8 + u] = bv[u];
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if (l + 8 <= len) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            cy = cy + ay[u];
            cv = cv + av[u];
            y[l + u] = cy;
            v[l + u] = cv;
        }
        l += 8;
    }
    for (; l < len; ++l) {
        cy = cy + y[l];
        cv = cv + v[l];
        y[l] = cy;
        v[l] = cv;
    }
}

// One queued period recomputed by a workgroup of EXW waves.  In LDS: my and
// mi (sum of y*ivar and of ivar per bin), EXW regions of rb bins each, plus 8
// bins of padding for the prefix's read-ahead.  Three forms:
// - in LDS (nb + 1 <= EXW * rb): every wave walks every point in point order
//   and adds those whose bin lies in its quarter of the bins, so each bin
//   still receives its points in point order; then the padding assignment,
//   the sequential prefix (one thread) and the box scan over all threads;
// - ranged (time-sorted curves with point-order copies, longer periods):
//   segments of S start positions (plus the longest duration's halo), one
//   per wave and region at a time; a wave walks only its segment's points
//   (the ranges of every cycle, then those of bins 1..os whose padding copies
//   land in the segment); after each round of EXW segments one thread runs
//   the prefix through them in order, carrying the total from segment to
//   segment, and each wave scans its segment.  The ranges of all segments
//   are searched in one round before the first when their table fits;
// - otherwise (unsorted curves, longer periods): segments over the whole
//   LDS, the first wave walking every point with its padded destinations.
// Outputs are the oracle's, as before: each bin's FP64 sum in point order,
// the oracle's padding, prefix order and formulas; the best window is the
// maximum over all of them by (objective, then lowest index).
template <class R>
__global__ __launch_bounds__(64 * EXW) void k_exact(Batch b, Out o, Work w,
                                                    int rb) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lds_bins = EXW * rb;
    double *my = (double *)smem;         // sum of y*ivar per bin
    double *mi = my + lds_bins + 8;      // sum of ivar per bin
    __shared__ int s_m[DMAX];
    __shared__ double s_tot[2];
    __shared__ double st_y[EXW][64], st_i[EXW][64];  // chunk values, staged
    __shared__ int s_tab[EXACT_TAB];
    __shared__ double r_bo[EXW], r_y[EXW], r_i[EXW];
    __shared__ int r_bi[EXW];
    __shared__ unsigned int s_q;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    for (;;) {
    __syncthreads();
    if (tid == 0) {
        // the drained queue costs a load, not an atomic on its one counter
        const unsigned int cnt = *w.cand_count;
        s_q = __hip_atomic_load(w.cand_next, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT) < cnt
                  ? atomicAdd(w.cand_next, 1u)
                  : cnt;
    }
    __syncthreads();
    const unsigned int q = s_q;
    if (q >= *w.cand_count) break;
    const int lc = w.cand_lc[q];
    const LcParams P = w.prm[lc];
    const long long p = w.cand_p[q];
    const double period = b.periods[p];
    const double bd = P.bd, inv_bd = 1.0 / bd, invP = 1.0 / period;
    const int os = b.oversample;
    const int nb = (int)ceil(period / bd) + os;
    const long long off = b.lc_off[lc];
    const int n = P.n, nd = P.n_dur;
    if (tid < nd) s_m[tid] = (int)round(b.durs[b.dur_off[lc] + tid] / bd);
    if (tid == 0) {
        s_tot[0] = w.seq[lc].x;
        s_tot[1] = w.seq[lc].y;
    }
    __syncthreads();
    int mmax = 0;
    for (int k = 0; k < nd; ++k) mmax = max(mmax, s_m[k]);
    const bool fits = nb + 1 <= lds_bins;
    const double *tn = w.nat_ty + 2 * nat_base(b, lc);
    int ncyc = 0;
    if (!fits && P.natural)
        ncyc = (int)fold_cycle(tn[2 * nat_slot(n - 1)], period, invP) + 1;
    const int T = 2 * ncyc + 2;  // a region's tail (doubles): its ranges
    const bool ranged = !fits && P.natural && nb - os > os &&
                        ncyc <= (1 << 20) && rb - T - mmax - 1 >= 64;
    if (!fits && !ranged && !(nb - os > os && lds_bins - mmax - 1 >= 1))
        continue;
    const int S = fits ? nb + 1
                       : (ranged ? rb - T - mmax - 1 : lds_bins - mmax - 1);
    const bool like = b.objective == ATG_OBJ_LIKELIHOOD;
    const int nval = nb + 1;
    const double ty = s_tot[0], ti = s_tot[1];
    double bo = -INFINITY, byin = 0.0, bivin = 0.0;
    int bi = INT_MAX;
    // box scan of starts st0, st0 + stride, ... below nend over the prefix
    // values of the segment at s0 held in (y, v): the oracle's formulas; a
    // window of negative depth never wins, so its objective is not formed,
    // and each objective forms only its own terms (the depth error and snr
    // of the best window follow)
    auto scan = [&](auto lk, const double *y, const double *v, int s0,
                    int st0, int nend, int stride) {
        constexpr bool LK = decltype(lk)::value;
        for (int st = st0; st < nend; st += stride) {
            const int l = st - s0;
            for (int k = 0; k < nd; ++k) {
                const int m = s_m[k];
                if (st + m > nb) continue;
                const double ys = y[l + m] - y[l];
                const double is = v[l + m] - v[l];
                double y_in = ys, iv_in = is;
                double y_out = ty - y_in;
                double iv_out = ti - iv_in;
                if (iv_in < DBL_EPSILON || iv_out < DBL_EPSILON) continue;
                y_in /= iv_in;
                y_out /= iv_out;
                const double depth = y_out - y_in;
                if (!(depth >= 0.0)) continue;
                double obj;
                if (LK) {
                    obj = 0.5 * iv_in * depth * depth;
                } else {
                    const double derr = sqrt(1.0 / iv_in + 1.0 / iv_out);
                    obj = depth / derr;
                }
                const int idx = k * nval + st;
                if (better(obj, idx, bo, bi)) {
                    bo = obj;
                    bi = idx;
                    byin = ys;
                    bivin = is;
                }
            }
        }
    };
    auto scan_seg = [&](const double *y, const double *v, int s0, int st0,
                        int nend, int stride) {
        if (like) scan(BoolC<true>(), y, v, s0, st0, nend, stride);
        else scan(BoolC<false>(), y, v, s0, st0, nend, stride);
    };
    if (fits) {
        for (int i = tid; i <= nb; i += 64 * EXW) {
            my[i] = 0.0;
            mi[i] = 0.0;
        }
        __syncthreads();
        // this wave's quarter of the bins
        const int qn = (nb + EXW) / EXW;
        const int qlo = wv * qn, qhi = min(qlo + qn, nb + 1) - 1;
        auto fetch = [&](int i, ExactPt &r) {
            if (i < n) {
                r.t = b.t[off + i];
                r.y = R::y(b, off + i);
                r.d = R::dy(b, off + i);
            }
        };
        double ay = 0.0, ai = 0.0;
        int cur = -1;
        ExactPt qp[EXACT_AHEAD];
#pragma unroll
        for (int u = 0; u < EXACT_AHEAD; ++u) fetch(64 * u + lane, qp[u]);
        for (int base0 = 0; base0 < n; base0 += 64 * EXACT_AHEAD) {
#pragma unroll
            for (int u = 0; u < EXACT_AHEAD; ++u) {
                const int base = base0 + 64 * u;
                if (base >= n) break;
                const ExactPt r = qp[u];
                fetch(base + 64 * EXACT_AHEAD + lane, qp[u]);
                const double tp = r.t - P.t_ref;
                const double iv = R::HAS_DY ? 1.0 / (r.d * r.d) : 1.0;
                const double yv = (r.y - P.y_med) * iv;
                const int ind =
                    min(max(fold_bin(tp, period, invP, bd, inv_bd), 0), nb);
                const int loc = (ind >= qlo && ind <= qhi) ? ind : -1;
                exact_add_chunk(loc, yv, iv, min(64, n - base), my, mi,
                                st_y[wv], st_i[wv], cur, ay, ai);
            }
        }
        if (cur >= 0 && lane == 0) {
            my[cur] = ay;
            mi[cur] = ai;
        }
        __syncthreads();
        if (tid == 0) {
            for (int j = 1, d = nb - os; j <= os; ++j, ++d) {
                my[d] = my[j];
                mi[d] = mi[j];
            }
            exact_prefix(my, mi, nb + 1, true, 0.0, 0.0);
        }
        __syncthreads();
        scan_seg(my, mi, 0, tid, nb, 64 * EXW);
    } else if (ranged) {
        const int nseg = (nb + S - 1) / S;
        // the range searches of all segments in one round (all threads)
        // when their table fits
        const int main_e = 2 * ncyc * nseg;
        const bool tab_ok = main_e + 2 * ncyc <= EXACT_TAB;
        if (tab_ok) {
            for (int e = tid; e < main_e + 2 * ncyc; e += 64 * EXW) {
                int v = 0;
                if (e < main_e) {
                    const int j = e / (2 * ncyc), r = e - j * (2 * ncyc);
                    const int s0j = j * S, hij = min(s0j + S + mmax, nb);
                    v = lower_key(tn, n, period, invP, bd, inv_bd, nb,
                                  (double)(r >> 1), (r & 1) ? hij + 1 : s0j);
                } else {
                    const int r = e - main_e;
                    v = lower_key(tn, n, period, invP, bd, inv_bd, nb,
                                  (double)(r >> 1), (r & 1) ? os + 1 : 1);
                }
                s_tab[e] = v;
            }
        }
        __syncthreads();
        double cpy = 0.0, cpi = 0.0;  // thread 0: the carried prefix
        for (int r0 = 0; r0 < nseg; r0 += EXW) {
            const int seg = r0 + wv, s0 = seg * S;
            const bool act = seg < nseg;
            double *myw = my + wv * rb, *miw = mi + wv * rb;
            if (act) {
                const int hi = min(s0 + S + mmax, nb);
                const int len = hi - s0 + 1;
                int *lo_t = (int *)(myw + (rb - T));  // [2*ncyc] range starts
                int *pre_t = lo_t + 2 * ncyc;         // [2*ncyc+1] prefix
                for (int i = lane; i < len; i += 64) {
                    myw[i] = 0.0;
                    miw[i] = 0.0;
                }
                const bool extra = hi >= nb - os;
                for (int e = lane; e < 4 * ncyc; e += 64) {
                    const int k = e >> 1, c = k % ncyc;
                    int v = 0;
                    if (k < ncyc)
                        v = tab_ok ? s_tab[2 * ncyc * seg + 2 * c + (e & 1)]
                                   : lower_key(tn, n, period, invP, bd, inv_bd,
                                               nb, (double)c,
                                               (e & 1) ? hi + 1 : s0);
                    else if (extra)
                        v = tab_ok ? s_tab[main_e + 2 * c + (e & 1)]
                                   : lower_key(tn, n, period, invP, bd, inv_bd,
                                               nb, (double)c,
                                               (e & 1) ? os + 1 : 1);
                    if (e & 1) pre_t[k] = v;  // range end, until the prefix
                    else lo_t[k] = v;
                }
                wave_lds_sync();
                if (lane == 0) {
                    int acc = 0;
                    for (int k = 0; k < 2 * ncyc; ++k) {
                        const int l = pre_t[k] - lo_t[k];
                        pre_t[k] = acc;
                        acc += l;
                    }
                    pre_t[2 * ncyc] = acc;
                }
                wave_lds_sync();
                const int total = pre_t[2 * ncyc];
                // item -> (range, point): a lane's items rise by 64 from one
                // chunk to the next, so its range cursor only moves forward;
                // the raw points of the next EXACT_AHEAD chunks are in flight
                // while the wave walks the current one
                int kc = 0;
                auto fetch = [&](int item, ExactPt &r) {
                    if (item < total) {
                        while (pre_t[kc + 1] <= item) ++kc;
                        const int j = lo_t[kc] + (item - pre_t[kc]);
                        r.k = kc;
                        r.t = b.t[off + j];
                        r.y = R::y(b, off + j);
                        r.d = R::dy(b, off + j);
                    }
                };
                double ay = 0.0, ai = 0.0;
                int cur = -1;
                ExactPt qp[EXACT_AHEAD];
#pragma unroll
                for (int u = 0; u < EXACT_AHEAD; ++u)
                    fetch(64 * u + lane, qp[u]);
                for (int cb0 = 0; cb0 < total; cb0 += 64 * EXACT_AHEAD) {
#pragma unroll
                    for (int u = 0; u < EXACT_AHEAD; ++u) {
                        const int cb = cb0 + 64 * u;
                        if (cb >= total) break;
                        const ExactPt r = qp[u];
                        fetch(cb + 64 * EXACT_AHEAD + lane, qp[u]);
                        const int k = r.k;
                        const bool valid = cb + lane < total;
                        const double tp = r.t - P.t_ref;
                        const double iv =
                            R::HAS_DY ? 1.0 / (r.d * r.d) : 1.0;
                        const double yv = (r.y - P.y_med) * iv;
                        int loc = -1;
                        if (valid) {
                            const int c = k < ncyc ? k : k - ncyc;
                            const int rbn = fold_bin_in(tp, period, (double)c,
                                                        bd, inv_bd, nb);
                            int dst;
                            if (k < ncyc)
                                dst = (rbn < nb - os || rbn == nb) ? rbn : -1;
                            else
                                dst = (rbn >= 1 && rbn <= os)
                                          ? rbn + nb - os - 1
                                          : -1;
                            if (dst >= s0 && dst <= hi) loc = dst - s0;
                        }
                        exact_add_chunk(loc, yv, iv, min(64, total - cb), myw,
                                        miw, st_y[wv], st_i[wv], cur, ay, ai);
                    }
                }
                if (cur >= 0 && lane == 0) {
                    myw[cur] = ay;
                    miw[cur] = ai;
                }
            }
            __syncthreads();
            if (tid == 0) {
                for (int ww = 0; ww < EXW && r0 + ww < nseg; ++ww) {
                    const int sg = r0 + ww, s0j = sg * S;
                    const int hij = min(s0j + S + mmax, nb);
                    double *yw = my + ww * rb, *vw = mi + ww * rb;
                    exact_prefix(yw, vw, hij - s0j + 1, sg == 0, cpy, cpi);
                    if (s0j + S < nb) {
                        cpy = yw[S - 1];
                        cpi = vw[S - 1];
                    }
                }
            }
            __syncthreads();
            if (act) scan_seg(myw, miw, s0, s0 + lane, min(s0 + S, nb), 64);
            __syncthreads();
        }
    } else {
        double cpy = 0.0, cpi = 0.0;  // thread 0: the carried prefix
        for (int s0 = 0; s0 < nb; s0 += S) {
            const int hi = min(s0 + S + mmax, nb);
            const int len = hi - s0 + 1;
            for (int i = tid; i < len; i += 64 * EXW) {
                my[i] = 0.0;
                mi[i] = 0.0;
            }
            __syncthreads();
            if (wv == 0) {
                auto fetch = [&](int i, ExactPt &r) {
                    if (i < n) {
                        r.t = b.t[off + i];
                        r.y = R::y(b, off + i);
                        r.d = R::dy(b, off + i);
                    }
                };
                ExactPt qp[EXACT_AHEAD];
#pragma unroll
                for (int u = 0; u < EXACT_AHEAD; ++u)
                    fetch(64 * u + lane, qp[u]);
                for (int base0 = 0; base0 < n; base0 += 64 * EXACT_AHEAD) {
#pragma unroll
                    for (int u = 0; u < EXACT_AHEAD; ++u) {
                        const int base = base0 + 64 * u;
                        if (base >= n) break;
                        const ExactPt r = qp[u];
                        fetch(base + 64 * EXACT_AHEAD + lane, qp[u]);
                        const double tp = r.t - P.t_ref;
                        const double iv =
                            R::HAS_DY ? 1.0 / (r.d * r.d) : 1.0;
                        const double yv = (r.y - P.y_med) * iv;
                        const int ind = min(
                            max(fold_bin(tp, period, invP, bd, inv_bd), 0), nb);
                        const int cnt = min(64, n - base);
                        // the wave walks the chunk in point order; every lane
                        // holds the same running values, lane 0 stores them
                        for (int j = 0; j < cnt; ++j) {
                            const double yvj = readlane_d(yv, j),
                                         ivj = readlane_d(iv, j);
                            const int indj = __builtin_amdgcn_readlane(ind, j);
                            int j1, j2;
                            pad_targets(indj, nb, os, j1, j2);
                            if (j1 >= s0 && j1 <= hi) {
                                const double a = my[j1 - s0] + yvj;
                                const double bb = mi[j1 - s0] + ivj;
                                if (lane == 0) {
                                    my[j1 - s0] = a;
                                    mi[j1 - s0] = bb;
                                }
                            }
                            if (j2 >= s0 && j2 <= hi) {
                                const double a = my[j2 - s0] + yvj;
                                const double bb = mi[j2 - s0] + ivj;
                                if (lane == 0) {
                                    my[j2 - s0] = a;
                                    mi[j2 - s0] = bb;
                                }
                            }
                        }
                    }
                }
            }
            __syncthreads();
            if (tid == 0) {
                exact_prefix(my, mi, len, s0 == 0, cpy, cpi);
                if (s0 + S < nb) {
                    cpy = my[S - 1];
                    cpi = mi[S - 1];
                }
            }
            __syncthreads();
            scan_seg(my, mi, s0, s0 + tid, min(s0 + S, nb), 64 * EXW);
            __syncthreads();
        }
    }
    // the best window over the workgroup
    for (int ofs = 32; ofs > 0; ofs >>= 1) {
        const double ob = __shfl_xor(bo, ofs);
        const int oi = __shfl_xor(bi, ofs);
        const double oy = __shfl_xor(byin, ofs), ov = __shfl_xor(bivin, ofs);
        if (better(ob, oi, bo, bi)) {
            bo = ob;
            bi = oi;
            byin = oy;
            bivin = ov;
        }
    }
    if (lane == 0) {
        r_bo[wv] = bo;
        r_bi[wv] = bi;
        r_y[wv] = byin;
        r_i[wv] = bivin;
    }
    __syncthreads();
    if (tid == 0) {
        for (int k = 1; k < EXW; ++k) {
            if (better(r_bo[k], r_bi[k], bo, bi)) {
                bo = r_bo[k];
                bi = r_bi[k];
                byin = r_y[k];
                bivin = r_i[k];
            }
        }
    }
    if (tid == 0 && bi != INT_MAX) {
        const int k = bi / nval, st = bi - k * nval, m = s_m[k];
        double y_in = byin, iv_in = bivin;
        double y_out = ty - y_in;
        double iv_out = ti - iv_in;
        y_in /= iv_in;
        y_out /= iv_out;
        const double depth = y_out - y_in;
        const double derr = sqrt(1.0 / iv_in + 1.0 / iv_out);
        const double snr = depth / derr;
        const double ll = 0.5 * iv_in * depth * depth;
        const double dur = m * bd;
        o.power[p] = like ? ll : snr;
        o.depth[p] = depth;
        o.depth_err[p] = derr;
        o.duration[p] = dur;
        o.transit_time[p] = fold_mod(st * bd + 0.5 * dur + 0.0, period) + P.t_ref;
        o.depth_snr[p] = snr;
        o.log_like[p] = ll;
    }
    }  // queue loop
}

// ------------------------------------------------------------------ k_topk
//
// Peaks of one curve's periodogram with the rule of scipy's find_peaks
// (local maxima; a plateau gives its midpoint; first and last samples never
// count) thinned by the minimum distance, then the k highest kept peaks.
// scipy thins greedily from the highest peak down; here the same result
// comes from rounds: an undecided peak is kept when no kept peak lies within
// the distance and it ranks above every undecided peak there, and an
// undecided peak within the distance of a kept one is dropped.  A peak is
// only ever dropped by a kept peak of higher rank, and the highest
// undecided peak is decided in every round, so the rounds end with exactly
// the greedy selection.  Rank: power, then period index (higher first).

struct Cand {
    double v;
    int i;
};

// higher rank: larger value, then larger index; i < 0 ranks lowest
__device__ __forceinline__ bool cand_above(double v, int i, double bv,
                                           int bi) {
    if (i < 0) return false;
    if (bi < 0) return true;
    return v > bv || (v == bv && i > bi);
}

__device__ Cand block_best(Cand c, double *dscr, int *iscr) {
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(c.v, o);
        const int oi = __shfl_xor(c.i, o);
        if (cand_above(ov, oi, c.v, c.i)) {
            c.v = ov;
            c.i = oi;
        }
    }
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        dscr[wv] = c.v;
        iscr[wv] = c.i;
    }
    __syncthreads();
    Cand r{dscr[0], iscr[0]};
    for (int k = 1; k < NW; ++k)
        if (cand_above(dscr[k], iscr[k], r.v, r.i)) {
            r.v = dscr[k];
            r.i = iscr[k];
        }
    return r;
}

__global__ __launch_bounds__(NT) void k_topk(const long long *per_off,
                                             const double *power,
                                             const int *status, int *tk_pos,
                                             unsigned char *tk_st, int k,
                                             int dist, long long *idx_out,
                                             int *cnt_out) {
    __shared__ int iscr[NW];
    __shared__ double dscr[NW];
    __shared__ int s_pick[64];
    const int lc = blockIdx.x, tid = threadIdx.x;
    const long long po = per_off[lc];
    const int np = (int)(per_off[lc + 1] - po);
    const double *x = power + po;
    int *pos = tk_pos + po;
    unsigned char *st = tk_st + po;
    if (tid < k) idx_out[(long long)lc * k + tid] = -1;
    if (status && status[lc] != ATG_LC_OK) {
        if (tid == 0) cnt_out[lc] = 0;
        return;
    }
    // local maxima in order of their left edges (= order of midpoints)
    int npk = 0;
    for (int base = 1; base < np - 1; base += NT) {
        const int i = base + tid;
        int mid = -1;
        if (i < np - 1 && x[i - 1] < x[i]) {
            int a = i + 1;
            while (a < np - 1 && x[a] == x[i]) ++a;
            if (x[a] < x[i]) mid = (i + a - 1) / 2;
        }
        int tot;
        const int e = block_excl_scan(mid >= 0 ? 1 : 0, iscr, &tot);
        if (mid >= 0) {
            pos[npk + e] = mid;
            st[npk + e] = 0;
        }
        npk += tot;
    }
    __syncthreads();
    int count = 0;
    if (npk > 0) {
        // 0 undecided, 1 kept, 2 dropped, 3 already reported
        for (;;) {
            for (int j = tid; j < npk; j += NT) {
                if (st[j] != 0) continue;
                const double vj = x[pos[j]];
                bool ok = true;
                for (int q = j - 1; ok && q >= 0 && pos[j] - pos[q] < dist;
                     --q)
                    ok = !(st[q] == 1 ||
                           (st[q] == 0 && cand_above(x[pos[q]], pos[q], vj,
                                                     pos[j])));
                for (int q = j + 1; ok && q < npk && pos[q] - pos[j] < dist;
                     ++q)
                    ok = !(st[q] == 1 ||
                           (st[q] == 0 && cand_above(x[pos[q]], pos[q], vj,
                                                     pos[j])));
                if (ok) st[j] = 1;
            }
            __syncthreads();
            int und = 0;
            for (int j = tid; j < npk; j += NT) {
                if (st[j] != 0) continue;
                bool drop = false;
                for (int q = j - 1; !drop && q >= 0 && pos[j] - pos[q] < dist;
                     --q)
                    drop = st[q] == 1;
                for (int q = j + 1; !drop && q < npk && pos[q] - pos[j] < dist;
                     ++q)
                    drop = st[q] == 1;
                if (drop) st[j] = 2;
                else ++und;
            }
            und = block_reduce(und, AddOp<int>(), iscr);
            if (und == 0) break;
        }
        // the k kept peaks of highest rank
        for (int r = 0; r < k; ++r) {
            Cand c{-INFINITY, -1};
            for (int j = tid; j < npk; j += NT)
                if (st[j] == 1 && cand_above(x[pos[j]], pos[j], c.v, c.i)) {
                    c.v = x[pos[j]];
                    c.i = pos[j];
                }
            const Cand bst = block_best(c, dscr, iscr);
            if (bst.i < 0) break;
            for (int j = tid; j < npk; j += NT)
                if (pos[j] == bst.i) st[j] = 3;
            if (tid == 0) idx_out[(long long)lc * k + r] = bst.i;
            ++count;
            __syncthreads();
        }
    } else {
        // no peak: the k highest values, listed in decreasing value with the
        // earlier index first among equal values
        for (int r = 0; r < k && r < np; ++r) {
            Cand c{-INFINITY, -1};
            for (int i = tid; i < np; i += NT) {
                bool taken = false;
                for (int q = 0; q < r; ++q) taken |= s_pick[q] == i;
                if (!taken && cand_above(x[i], i, c.v, c.i)) {
                    c.v = x[i];
                    c.i = i;
                }
            }
            const Cand bst = block_best(c, dscr, iscr);
            if (tid == 0) s_pick[r] = bst.i;
            __syncthreads();
            ++count;
        }
        if (tid == 0) {
            for (int a = 1; a < count; ++a) {  // insertion sort, k <= 64
                const int v = s_pick[a];
                int z = a - 1;
                while (z >= 0 && (x[s_pick[z]] < x[v] ||
                                  (x[s_pick[z]] == x[v] && s_pick[z] > v))) {
                    s_pick[z + 1] = s_pick[z];
                    --z;
                }
                s_pick[z + 1] = v;
            }
            for (int r = 0; r < count; ++r)
                idx_out[(long long)lc * k + r] = s_pick[r];
        }
    }
    if (tid == 0) cnt_out[lc] = count;
}

// ------------------------------------------------------ k_stats / k_model
//
// Vetting statistics of one (curve, period, duration, transit time)
// candidate, astropy BoxLeastSquares.compute_stats (astropy timeseries/
// periodograms/bls/core.py:427-570) on the raw t and y of the curve:
// weighted levels inside and outside the box and of the odd, even,
// half-period and half-phase boxes, the per-transit point counts and log
// likelihoods, and a least-squares sinusoid at the period.  Sums are FP64
// block reductions in a fixed order (deterministic); the per-transit sums
// are taken in point order by one wave.

constexpr int MAXTR = 4096;  // per-transit accumulators in LDS
constexpr int NSUM = 25;     // FP64 sums of the first pass
constexpr int NCNT = 8;      // point counts of the eight masks

// numpy's float remainder (floored modulo via fmod)
__device__ __forceinline__ double np_mod(double a, double b) {
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0.0) != (m < 0.0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

__device__ __forceinline__ bool in_box(double t, double T0, double P,
                                       double d) {
    const double hp = 0.5 * P;
    return fabs(np_mod(t - T0 + hp, P) - hp) < 0.5 * d;
}

template <typename T>
__device__ T block_sum_bcast(T v, T *scr) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) scr[threadIdx.x >> 6] = v;
    __syncthreads();
    T r = scr[0];
    for (int k = 1; k < NW; ++k) r += scr[k];
    return r;
}

// (level, variance) of a mask, or the depth below `ref` with its error;
// (0, inf) when the mask is empty or the reference undefined
__device__ void level_of(int cnt, double W, double Y, const double *ref,
                         double &a, double &b) {
    if (cnt > 0 && (ref == nullptr || isfinite(ref[1]))) {
        const double var = 1.0 / W;
        const double mean = Y * var;
        if (!ref) {
            a = mean;
            b = var;
        } else {
            a = ref[0] - mean;
            b = sqrt(var + ref[1]);
        }
        return;
    }
    a = 0.0;
    b = INFINITY;
}

__global__ __launch_bounds__(NT) void k_stats(
    Batch b, long long n_cand, const int *cand_lc, const double *cper,
    const double *cdur, const double *ct0, int cap, double *stats,
    int *pt_count, double *pt_ll, int *status) {
    __shared__ double dsc[NW];
    __shared__ long long lsc[NW];
    __shared__ int isc[NW];
    __shared__ double s_par[6];
    __shared__ double ll_l[MAXTR];
    __shared__ int cnt_l[MAXTR];
    const long long c = blockIdx.x;
    const int tid = threadIdx.x;
    const int lc = cand_lc[c];
    const long long off = b.lc_off[lc];
    const int n = (int)(b.lc_off[lc + 1] - off);
    const double P = cper[c], d = cdur[c], T0 = ct0[c];
    const double hp = 0.5 * P, twoP = 2.0 * P, qP = 0.25 * P, hP = 0.5 * P;
    const double twopi = 6.283185307179586;

    // pass 1: weights and weighted sums of every mask, the sinusoid's
    // normal equations, the range of transit numbers
    double a[NSUM];
    int cnt[NCNT];
#pragma unroll
    for (int k = 0; k < NSUM; ++k) a[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NCNT; ++k) cnt[k] = 0;
    long long idmin = LLONG_MAX, idmax = LLONG_MIN;
    for (int i = tid; i < n; i += NT) {
        const double t = b.t[off + i];
        const double y = load_val(b.y, b.y_dt, off + i);
        double iv = 1.0;
        if (b.dy) {
            const double e = load_val(b.dy, b.dy_dt, off + i);
            iv = 1.0 / (e * e);
        }
        const double yw = y * iv;
        const bool m_in = fabs(np_mod(t - T0 + hp, P) - hp) < 0.5 * d;
        const bool m_odd = fabs(np_mod(t - T0, twoP) - P) < 0.5 * d;
        const bool m_even = fabs(np_mod(t - T0 + P, twoP) - P) < 0.5 * d;
        const bool m_ph = fabs(np_mod(t - T0, P) - hp) < 0.5 * d;
        const bool m_half = fabs(np_mod(t - T0 + qP, hP) - qP) < 0.5 * d;
        const bool msk[NCNT] = {m_in,  !m_in, m_odd,  m_even,
                                m_ph,  !m_ph && !m_in, m_half, !m_half};
#pragma unroll
        for (int k = 0; k < NCNT; ++k)
            if (msk[k]) {
                a[2 * k] += iv;
                a[2 * k + 1] += yw;
                cnt[k] += 1;
            }
        const double arg = twopi * t / P;
        const double sn = sin(arg), cs = cos(arg);
        a[16] += sn * sn * iv;
        a[17] += sn * cs * iv;
        a[18] += sn * iv;
        a[19] += cs * cs * iv;
        a[20] += cs * iv;
        a[21] += iv;
        a[22] += sn * yw;
        a[23] += cs * yw;
        a[24] += yw;
        if (m_in) {
            const long long id = llrint((t - T0) / P);
            idmin = min(idmin, id);
            idmax = max(idmax, id);
        }
    }
#pragma unroll
    for (int k = 0; k < NSUM; ++k) a[k] = block_sum_bcast(a[k], dsc);
#pragma unroll
    for (int k = 0; k < NCNT; ++k) cnt[k] = block_sum_bcast(cnt[k], isc);
    idmin = block_reduce(idmin, MinOp(), lsc);
    __syncthreads();
    idmax = block_reduce(idmax, MaxOp(), lsc);
    double *so = stats + c * ATG_STATS_FIELDS;
    if (cnt[0] == 0) {
        if (tid < ATG_STATS_FIELDS) so[tid] = 0.0;
        if (tid == 0) status[c] = ATG_STATS_NO_TRANSIT;
        return;
    }
    if (tid == 0) {
        double out[2], dep[2], odd[2], even[2], phref[2], ph[2], nh[2],
            half[2];
        level_of(cnt[1], a[2], a[3], nullptr, out[0], out[1]);
        level_of(cnt[0], a[0], a[1], out, dep[0], dep[1]);
        level_of(cnt[2], a[4], a[5], out, odd[0], odd[1]);
        level_of(cnt[3], a[6], a[7], out, even[0], even[1]);
        level_of(cnt[5], a[10], a[11], nullptr, phref[0], phref[1]);
        level_of(cnt[4], a[8], a[9], phref, ph[0], ph[1]);
        level_of(cnt[7], a[14], a[15], nullptr, nh[0], nh[1]);
        level_of(cnt[6], a[12], a[13], nh, half[0], half[1]);
        // sinusoid: 3x3 normal equations, Gaussian elimination with
        // partial pivoting
        double M[3][4] = {{a[16], a[17], a[18], a[22]},
                          {a[17], a[19], a[20], a[23]},
                          {a[18], a[20], a[21], a[24]}};
        for (int col = 0; col < 3; ++col) {
            int piv = col;
            for (int r = col + 1; r < 3; ++r)
                if (fabs(M[r][col]) > fabs(M[piv][col])) piv = r;
            if (piv != col)
                for (int k = 0; k < 4; ++k) {
                    const double tmp = M[col][k];
                    M[col][k] = M[piv][k];
                    M[piv][k] = tmp;
                }
            for (int r = col + 1; r < 3; ++r) {
                const double f = M[r][col] / M[col][col];
                for (int k = col; k < 4; ++k) M[r][k] -= f * M[col][k];
            }
        }
        double wv[3];
        for (int r = 2; r >= 0; --r) {
            double acc = M[r][3];
            for (int k = r + 1; k < 3; ++k) acc -= M[r][k] * wv[k];
            wv[r] = acc / M[r][r];
        }
        so[0] = dep[0];
        so[1] = dep[1];
        so[2] = ph[0];
        so[3] = ph[1];
        so[4] = half[0];
        so[5] = half[1];
        so[6] = odd[0];
        so[7] = odd[1];
        so[8] = even[0];
        so[9] = even[1];
        so[10] = sqrt(wv[0] * wv[0] + wv[1] * wv[1]);
        so[12] = a[1] / a[0];  // model level inside (direct quotient)
        so[13] = cnt[1] > 0 ? a[3] / a[2] : 0.0;
        so[14] = (double)idmin;
        so[15] = (double)(idmax - idmin + 1);
        s_par[0] = out[0] - dep[0];  // y_in
        s_par[1] = out[0];           // y_out
        s_par[2] = wv[0];
        s_par[3] = wv[1];
        s_par[4] = wv[2];
    }
    __syncthreads();
    const double y_in = s_par[0], y_out = s_par[1];
    const double w0 = s_par[2], w1 = s_par[3], w2 = s_par[4];

    // pass 2: the box and sinusoid log likelihoods
    double sin_s = 0.0, in_s = 0.0, out_s = 0.0;
    for (int i = tid; i < n; i += NT) {
        const double t = b.t[off + i];
        const double y = load_val(b.y, b.y_dt, off + i);
        double iv = 1.0;
        if (b.dy) {
            const double e = load_val(b.dy, b.dy_dt, off + i);
            iv = 1.0 / (e * e);
        }
        const double arg = twopi * t / P;
        const double mod = sin(arg) * w0 + cos(arg) * w1 + w2;
        const double r = y - mod;
        sin_s += r * r * iv;
        if (in_box(t, T0, P, d)) {
            const double q = y - y_in;
            in_s += iv * (q * q);
        } else {
            const double q = y - y_out;
            out_s += iv * (q * q);
        }
    }
    sin_s = block_sum_bcast(sin_s, dsc);
    in_s = block_sum_bcast(in_s, dsc);
    out_s = block_sum_bcast(out_s, dsc);
    const long long n_tr = idmax - idmin + 1;
    const bool fits = n_tr <= cap && n_tr <= MAXTR;
    if (tid == 0) {
        double full_ll = -0.5 * in_s;
        full_ll -= 0.5 * out_s;
        so[11] = -0.5 * sin_s - full_ll;
        status[c] = fits ? ATG_STATS_OK : ATG_STATS_TOO_MANY_TRANSITS;
    }
    if (!fits) return;

    // per-transit counts and log likelihoods, in point order (wave 0)
    for (int k = tid; k < n_tr; k += NT) {
        ll_l[k] = 0.0;
        cnt_l[k] = 0;
    }
    __syncthreads();
    if (tid < 64) {
        int cur = -1, ccur = 0;
        double acc = 0.0;
        for (int base = 0; base < n; base += 64) {
            const int i = base + tid;
            int id = -1;
            double ll = 0.0;
            if (i < n) {
                const double t = b.t[off + i];
                if (in_box(t, T0, P, d)) {
                    const double y = load_val(b.y, b.y_dt, off + i);
                    double iv = 1.0;
                    if (b.dy) {
                        const double e = load_val(b.dy, b.dy_dt, off + i);
                        iv = 1.0 / (e * e);
                    }
                    const double qi = y - y_in, qo = y - y_out;
                    ll = -0.5 * iv * (qi * qi - qo * qo);
                    id = (int)(llrint((t - T0) / P) - idmin);
                }
            }
            const int cntb = min(64, n - base);
            for (int j = 0; j < cntb; ++j) {
                const int idj = __builtin_amdgcn_readlane(id, j);
                if (idj < 0) continue;
                const double llj = readlane_d(ll, j);
                if (idj != cur) {
                    if (cur >= 0 && tid == 0) {
                        ll_l[cur] = acc;
                        cnt_l[cur] = ccur;
                    }
                    cur = idj;
                    acc = ll_l[idj];
                    ccur = cnt_l[idj];
                }
                acc += llj;
                ccur += 1;
            }
        }
        if (cur >= 0 && tid == 0) {
            ll_l[cur] = acc;
            cnt_l[cur] = ccur;
        }
    }
    __syncthreads();
    for (int k = tid; k < n_tr; k += NT) {
        pt_count[c * cap + k] = cnt_l[k];
        pt_ll[c * cap + k] = ll_l[k];
    }
}

// astropy's transit_mask (core.py:572-603) and model (core.py:375-425)
// evaluated at arbitrary times, given the model's two levels.
__global__ __launch_bounds__(NT) void k_model(long long n, const double *t,
                                              double P, double d, double T0,
                                              double y_in, double y_out,
                                              unsigned char *mask,
                                              double *y_model) {
    const long long i = (long long)blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    const bool m = in_box(t[i], T0, P, d);
    if (mask) mask[i] = m ? 1 : 0;
    if (y_model) y_model[i] = m ? y_in : y_out;
}

// ----------------------------------------------------------------- k_final

// Winner records of a best-only run (per curve); power == nullptr in a
// full run.
struct Winners {
    double *power, *depth, *depth_err, *duration, *transit_time, *depth_snr,
        *log_like, *period;
    long long *best_index;
    int *status;
};

__global__ __launch_bounds__(NT) void k_final(Batch b, Out o, Work w,
                                              Winners bo) {
    __shared__ double dscr[NW];
    __shared__ long long lscr[NW];
    __shared__ int iscr[NW];
    const int lc = blockIdx.x;
    const long long po = b.per_off[lc], np = b.per_off[lc + 1] - po;
    // numpy argmax order: a NaN wins (first NaN), else the first maximum
    double bv = -INFINITY;
    long long bi = LLONG_MAX;
    int bn = 0;
    for (long long i = threadIdx.x; i < np; i += NT) {
        const double v = o.power[po + i];
        const int vn = isnan(v) ? 1 : 0;
        bool take = vn > bn || (vn == bn && (vn ? i < bi
                                               : (v > bv || (v == bv && i < bi))));
        if (take) {
            bv = v;
            bi = i;
            bn = vn;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        double ov = __shfl_xor(bv, off);
        long long oi = __shfl_xor(bi, off);
        int on = __shfl_xor(bn, off);
        bool take = on > bn || (on == bn && (on ? oi < bi
                                               : (ov > bv || (ov == bv && oi < bi))));
        if (take) {
            bv = ov;
            bi = oi;
            bn = on;
        }
    }
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        dscr[wv] = bv;
        lscr[wv] = bi;
        iscr[wv] = bn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < NW; ++i) {
            bool take = iscr[i] > bn ||
                        (iscr[i] == bn &&
                         (bn ? lscr[i] < bi
                             : (dscr[i] > bv || (dscr[i] == bv && lscr[i] < bi))));
            if (take) {
                bv = dscr[i];
                bi = lscr[i];
                bn = iscr[i];
            }
        }
        // bi stays LLONG_MAX only for a curve without periods
        const bool have = bi != LLONG_MAX;
        int st = w.prm[lc].status;
        if (st == ATG_LC_OK && w.lc_flags[lc]) st = ATG_LC_TOO_MANY_BINS;
        if (o.best_index) o.best_index[lc] = have ? bi : 0;
        if (o.status) o.status[lc] = st;
        if (bo.power) {  // best-only epilogue: gather the winner's fields
            const double nan = __longlong_as_double(0x7ff8000000000000ll);
            const long long q = po + (have ? bi : 0);
            bo.power[lc] = have ? o.power[q] : nan;
            bo.depth[lc] = have ? o.depth[q] : nan;
            bo.depth_err[lc] = have ? o.depth_err[q] : nan;
            bo.duration[lc] = have ? o.duration[q] : nan;
            bo.transit_time[lc] = have ? o.transit_time[q] : nan;
            bo.depth_snr[lc] = have ? o.depth_snr[q] : nan;
            bo.log_like[lc] = have ? o.log_like[q] : nan;
            bo.period[lc] = have ? b.periods[q] : nan;
            if (bo.best_index) bo.best_index[lc] = have ? bi : 0;
            if (bo.status) bo.status[lc] = st;
        }
        w.lc_flags[lc] = 0;
    }
}

}  // namespace

// ===================================================================== ABI

// The workspace of atg_bls_run_noalloc: a static device array, so that the
// call allocates nothing.  Every GPU the library runs on holds its own copy
// at its own address until the process ends.  Host code never passes the
// array's name to a kernel (on the host it names a stand-in, not device
// memory): each device's address comes from hipGetSymbolAddress, resolved
// once per device at context creation.
static __device__ __align__(256) unsigned char scratch[ATG_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on `device` (resolved with that device
// current, the caller's current device restored); nullptr on failure.
void *scratch_on(int device, hipError_t *err) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    *err = hipSuccess;
    if (device < 0 || device >= MAX_DEVICES) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    if (g_scratch_ptr[device]) return g_scratch_ptr[device];
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess) e = hipSetDevice(device);
    if (e != hipSuccess) {
        *err = e;
        return nullptr;
    }
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    hipSetDevice(prev);
    if (e != hipSuccess || !ptr) {
        *err = e != hipSuccess ? e : hipErrorInvalidSymbol;
        return nullptr;
    }
    g_scratch_ptr[device] = ptr;
    return ptr;
}
}  // namespace

struct atg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    std::string err;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    int pins = 0;  // captured graphs that hold ws: it must not move
    atg_tuning tune{0, 2540, 4, 0};  // periods per item: per batch
    int num_cu = 0;
    size_t lds_max = 65536;
    size_t lds_attr_set = 65536;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> spare;
    double t_sum = 0.0;
    int t_cnt = 0;
};

namespace {

int fail(atg_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(atg_ctx *c, hipError_t e, const char *what) {
    return fail(c, ATG_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Layout {
    size_t prm, seq, flags, items, queues, counter, cand_lc, cand_p, t, w, y,
        nt, nw, tk_pos, tk_st, big_q, big_p, big_lc, total;
};

Layout layout_of(const atg_bls_batch *b) {
    Layout L;
    const size_t nlc = (size_t)b->n_lc;
    const size_t slots = (size_t)b->n_points + nlc * (2 * NT);
    size_t o = 0;
    L.prm = o;     o = align256(o + nlc * sizeof(LcParams));
    L.seq = o;     o = align256(o + nlc * 2 * sizeof(double));
    L.flags = o;   o = align256(o + nlc * sizeof(int));
    L.items = o;   o = align256(o + (nlc + 1) * sizeof(long long));
    L.queues = o;  o = align256(o + NQ * (sizeof(long long) + sizeof(int)));
    L.counter = o; o = align256(o + 3 * sizeof(unsigned int));
    L.big_q = o;   o = align256(o + NQ * (sizeof(long long) + 2 * sizeof(int)));
    const size_t nper = (size_t)b->n_periods;
    L.cand_lc = o; o = align256(o + nper * sizeof(int));
    L.cand_p = o;  o = align256(o + nper * sizeof(long long));
    L.t = o;       o = align256(o + slots * sizeof(double));
    L.w = o;       o = align256(o + slots * sizeof(long long));
    L.y = o;       o = align256(o + slots * sizeof(long long));
    const size_t npt = (size_t)b->n_points + nlc * NT;  // whole blocks
    L.nt = o;      o = align256(o + npt * 2 * sizeof(double));
    L.nw = o;      o = align256(o + npt * sizeof(long long));
    L.tk_pos = o;  o = align256(o + nper * sizeof(int));
    L.tk_st = o;   o = align256(o + nper);
    L.big_p = o;   o = align256(o + nper * sizeof(long long));
    L.big_lc = o;  o = align256(o + nper * sizeof(int));
    L.total = o;
    return L;
}

int check_batch(atg_ctx *c, const atg_bls_batch *b) {
    if (!b) return fail(c, ATG_ERR_ARG, "null batch");
    if (b->n_lc < 1 || b->n_lc > INT_MAX)
        return fail(c, ATG_ERR_ARG, "n_lc must be in [1, 2^31)");
    if (b->n_points < 0 || b->n_periods < 1 || b->n_durations < 1)
        return fail(c, ATG_ERR_ARG, "empty periods/durations or negative size");
    if (b->n_points > ((long long)1 << 40))
        return fail(c, ATG_ERR_ARG, "too many points");
    if (!b->lc_offsets || !b->periods || !b->period_offsets ||
        !b->durations || !b->dur_offsets)
        return fail(c, ATG_ERR_ARG, "null offsets or grid pointer");
    if (b->n_points > 0 && (!b->t || !b->y))
        return fail(c, ATG_ERR_ARG, "null t or y");
    if (b->y_dtype != ATG_F32 && b->y_dtype != ATG_F64)
        return fail(c, ATG_ERR_ARG, "y_dtype must be ATG_F32 or ATG_F64");
    if (b->dy && b->dy_dtype != ATG_F32 && b->dy_dtype != ATG_F64)
        return fail(c, ATG_ERR_ARG, "dy_dtype must be ATG_F32 or ATG_F64");
    if (b->oversample < 1) return fail(c, ATG_ERR_ARG, "oversample must be >= 1");
    if (b->objective != ATG_OBJ_LIKELIHOOD && b->objective != ATG_OBJ_SNR)
        return fail(c, ATG_ERR_ARG, "objective must be likelihood or snr");
    if (b->mode != ATG_MODE_PARITY && b->mode != ATG_MODE_FAST)
        return fail(c, ATG_ERR_ARG, "mode must be parity or fast");
    return ATG_OK;
}

// The 7 per-period arrays of a best-only run, in the workspace after the
// layout's own regions.
size_t best_bytes(const atg_bls_batch *b) {
    return align256(7 * (size_t)b->n_periods * sizeof(double));
}

int ensure_workspace(atg_ctx *c, size_t need) {
    if (need <= c->ws_bytes) return ATG_OK;
    if (c->pins > 0)  // a captured graph holds the current address
        return fail(c, ATG_ERR_BUDGET,
                    "workspace of " + std::to_string(need) +
                        " bytes needed, but the context's " +
                        std::to_string(c->ws_bytes) +
                        "-byte workspace is pinned by " +
                        std::to_string(c->pins) +
                        " captured graph(s); release them or use another "
                        "context");
    size_t want = need + need / 4;
    if (c->budget && want > c->budget) want = need;
    if (c->budget && need > c->budget)
        return fail(c, ATG_ERR_BUDGET,
                    "workspace of " + std::to_string(need) +
                        " bytes exceeds the HBM budget of " +
                        std::to_string(c->budget));
    if (c->ws) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "hipStreamSynchronize");
        hipFree(c->ws);
        c->ws = nullptr;
        c->ws_bytes = 0;
    }
    hipError_t e = hipMalloc(&c->ws, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c->ws = nullptr;
        return fail(c, ATG_ERR_BUDGET,
                    "hipMalloc of " + std::to_string(want) + " bytes failed");
    }
    c->ws_bytes = want;
    return ATG_OK;
}

Batch to_batch(const atg_bls_batch *b) {
    Batch k;
    k.n_lc = b->n_lc;
    k.t = b->t;
    k.y = b->y;
    k.dy = b->dy;
    k.y_dt = b->y_dtype;
    k.dy_dt = b->dy_dtype;
    k.lc_off = (const long long *)b->lc_offsets;
    k.periods = b->periods;
    k.per_off = (const long long *)b->period_offsets;
    k.durs = b->durations;
    k.dur_off = (const long long *)b->dur_offsets;
    k.oversample = b->oversample;
    k.objective = b->objective;
    k.mode = b->mode;
    return k;
}

// The launch sequence of atg_bls_run (bst == nullptr: o holds the caller's
// per-period arrays) and atg_bls_run_best (o == nullptr: the per-period
// arrays live in the workspace, k_final gathers the winners into bst).
// ws == nullptr: the context's workspace (grown on demand); otherwise a
// workspace of at least the layout's size that the caller provides (the
// static scratch of atg_bls_run_noalloc: then nothing is allocated, and no
// timing events are recorded).
int run_impl(atg_ctx *c, const atg_bls_batch *b, const atg_bls_out *o,
             const atg_bls_best *bst, char *ws = nullptr) {
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    const Layout L = layout_of(b);
    const bool own_ws = ws == nullptr;
    if (own_ws) {
        int rc = ensure_workspace(c, L.total + (bst ? best_bytes(b) : 0));
        if (rc != ATG_OK) return rc;
    }
    char *base = own_ws ? (char *)c->ws : ws;
    const bool timing = c->timing && own_ws;
    Work w;
    w.prm = (LcParams *)(base + L.prm);
    w.seq = (double2 *)(base + L.seq);
    w.lc_flags = (int *)(base + L.flags);
    w.item_off = (long long *)(base + L.items);
    w.counter = (unsigned int *)(base + L.counter);
    w.qtot = (long long *)(base + L.queues);
    w.qnext = (unsigned int *)(base + L.queues + NQ * sizeof(long long));
    w.ws_t = (double *)(base + L.t);
    w.ws_w = (long long *)(base + L.w);
    w.ws_y = (long long *)(base + L.y);
    w.nat_ty = (double *)(base + L.nt);
    w.nat_w = (long long *)(base + L.nw);
    w.cand_count = w.counter + 1;
    w.cand_next = w.counter + 2;
    w.big_base = (long long *)(base + L.big_q);
    w.big_count = (unsigned int *)(base + L.big_q + NQ * sizeof(long long));
    w.big_next = w.big_count + NQ;
    w.big_p = (long long *)(base + L.big_p);
    w.big_lc = (int *)(base + L.big_lc);
    w.cand_lc = (int *)(base + L.cand_lc);
    w.cand_p = (long long *)(base + L.cand_p);
    Batch kb = to_batch(b);
    Out ko;
    Winners kbest{};
    if (bst) {
        double *f = (double *)(base + L.total);
        const size_t np = (size_t)b->n_periods;
        ko.power = f;
        ko.depth = f + np;
        ko.depth_err = f + 2 * np;
        ko.duration = f + 3 * np;
        ko.transit_time = f + 4 * np;
        ko.depth_snr = f + 5 * np;
        ko.log_like = f + 6 * np;
        ko.best_index = nullptr;
        ko.status = nullptr;
        kbest.power = bst->power;
        kbest.depth = bst->depth;
        kbest.depth_err = bst->depth_err;
        kbest.duration = bst->duration;
        kbest.transit_time = bst->transit_time;
        kbest.depth_snr = bst->depth_snr;
        kbest.log_like = bst->log_likelihood;
        kbest.period = bst->period;
        kbest.best_index = (long long *)bst->best_index;
        kbest.status = bst->status;
    } else {
        ko.power = o->power;
        ko.depth = o->depth;
        ko.depth_err = o->depth_err;
        ko.duration = o->duration;
        ko.transit_time = o->transit_time;
        ko.depth_snr = o->depth_snr;
        ko.log_like = o->log_likelihood;
        ko.best_index = (long long *)o->best_index;
        ko.status = o->status;
    }

    const int nlc = (int)b->n_lc;
    const size_t lds = (size_t)c->tune.lds_bins * 16;
    const int grid = c->num_cu * c->tune.wg_per_cu;
    // periods per work item, unless the tiling fixes it: long items (fewer
    // passes over each curve's points) when the batch has periods to spare
    // for eight items per workgroup, short ones for a small batch, so that
    // its few curves still spread over the whole GPU
    // at most 64 (set_tuning checks it): k_main holds an item's periods one
    // per lane of a wave
    const int per_item =
        c->tune.periods_per_item > 0
            ? c->tune.periods_per_item
            : (int)std::min<long long>(
                  24, std::max<long long>(4, b->n_periods / (8LL * grid)));
    if (lds > c->lds_attr_set) {
        for (const void *f :
             {(const void *)k_main<true, true>, (const void *)k_main<true, false>,
              (const void *)k_main<false, true>,
              (const void *)k_main<false, false>, (const void *)k_big}) {
            e = hipFuncSetAttribute(
                f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return hip_fail(c, e, "hipFuncSetAttribute");
        }
        c->lds_attr_set = lds;
    }
    with_raw_types(kb, [&](auto r) {
        hipLaunchKernelGGL(k_prep<decltype(r)>, dim3(nlc), dim3(NT), 0,
                           c->stream, kb, w, c->tune.lds_bins);
    });
    hipLaunchKernelGGL(k_items, dim3(1), dim3(NT), 0, c->stream,
                       (long long)b->n_lc, kb.per_off, per_item, w);
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (timing) {
        for (int i = 0; i < 2; ++i) {
            hipEvent_t ev;
            if (!c->spare.empty()) {
                ev = c->spare.back();
                c->spare.pop_back();
            } else if (hipEventCreate(&ev) != hipSuccess) {
                return fail(c, ATG_ERR_HIP, "hipEventCreate");
            }
            (i ? ev1 : ev0) = ev;
        }
        hipEventRecord(ev0, c->stream);
    }
    {
        const bool ex = b->mode == ATG_MODE_PARITY;
        const bool lk = b->objective == ATG_OBJ_LIKELIHOOD;
        auto km = ex ? (lk ? k_main<true, true> : k_main<true, false>)
                     : (lk ? k_main<false, true> : k_main<false, false>);
        hipLaunchKernelGGL(km, dim3(grid), dim3(NT), lds, c->stream, kb, ko,
                           w, per_item, c->tune.lds_bins);
    }
    // k_big's histogram is 32 bins (512 B) smaller where that is what its
    // static LDS needs for as many workgroups per CU as k_main (its results
    // do not depend on the tile size)
    const int big_bins =
        ((size_t)c->tune.lds_bins * 16 + 512) * c->tune.wg_per_cu >
                160 * 1024
            ? c->tune.lds_bins - 32
            : c->tune.lds_bins;
    hipLaunchKernelGGL(k_big, dim3(grid), dim3(NT), (size_t)big_bins * 16,
                       c->stream, kb, ko, w, big_bins);
    hipLaunchKernelGGL(k_fields, dim3(nlc), dim3(NT), 0, c->stream, kb, ko, w);
    if (timing) {  // the periodogram stage: k_main, k_big and k_fields
        hipEventRecord(ev1, c->stream);
        c->pending.emplace_back(ev0, ev1);
    }
    hipLaunchKernelGGL(k_cands, dim3(nlc), dim3(NT), 0, c->stream, kb, ko, w);
    const int ex_total = exact_total(c->tune.lds_bins);
    with_raw_types(kb, [&](auto r) {
        hipLaunchKernelGGL(k_exact<decltype(r)>,
                           dim3(c->num_cu * EXACT_WG_PER_CU), dim3(64 * EXW),
                           (size_t)16 * (ex_total + 8), c->stream, kb, ko, w,
                           ex_total / EXW);
    });
    hipLaunchKernelGGL(k_final, dim3(nlc), dim3(NT), 0, c->stream, kb, ko, w,
                       kbest);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(c, e, "kernel launch");
    return ATG_OK;
}

}  // namespace

extern "C" {

int atg_abi_version(void) { return ATG_ABI_VERSION; }

int atg_ctx_create(int device, void *stream, uint64_t hbm_budget_bytes,
                   atg_ctx **out) {
    if (!out) return ATG_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
        (void)hipGetLastError();
        return ATG_ERR_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return ATG_ERR_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return ATG_ERR_DEVICE;
    hipError_t se;
    if (!scratch_on(device, &se)) {
        (void)hipGetLastError();
        return ATG_ERR_HIP;
    }
    atg_ctx *c = new atg_ctx();
    c->device = device;
    c->stream = (hipStream_t)stream;
    c->budget = hbm_budget_bytes;
    c->num_cu = prop.multiProcessorCount;
    c->lds_max = prop.sharedMemPerBlock;
    *out = c;
    return ATG_OK;
}

void atg_ctx_destroy(atg_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    else hipDeviceSynchronize();
    for (auto &pr : c->pending) {
        hipEventDestroy(pr.first);
        hipEventDestroy(pr.second);
    }
    for (auto e : c->spare) hipEventDestroy(e);
    if (c->ws) hipFree(c->ws);
    delete c;
}

int atg_ctx_pin_workspace(atg_ctx *c, int pin) {
    if (!c) return ATG_ERR_ARG;
    if (pin) ++c->pins;
    else if (c->pins > 0) --c->pins;
    return c->pins;
}

int atg_ctx_set_stream(atg_ctx *c, void *stream) {
    if (!c) return ATG_ERR_ARG;
    c->stream = (hipStream_t)stream;
    return ATG_OK;
}

const char *atg_last_error(const atg_ctx *c) {
    return c ? c->err.c_str() : "null context";
}

int atg_ctx_set_tuning(atg_ctx *c, const atg_tuning *t) {
    if (!c || !t) return ATG_ERR_ARG;
    atg_tuning n = c->tune;
    if (t->periods_per_item) n.periods_per_item = t->periods_per_item;
    if (t->lds_bins) n.lds_bins = t->lds_bins;
    if (t->wg_per_cu) n.wg_per_cu = t->wg_per_cu;
    if (n.periods_per_item < 0 || n.periods_per_item > 64)
        return fail(c, ATG_ERR_ARG, "periods_per_item must be in [1, 64]");
    if (n.lds_bins < 64 || (size_t)n.lds_bins * 16 > c->lds_max)
        return fail(c, ATG_ERR_ARG,
                    "lds_bins must be >= 64 and fit " +
                        std::to_string(c->lds_max) + " bytes of LDS");
    if (n.wg_per_cu < 1 || n.wg_per_cu > 32)
        return fail(c, ATG_ERR_ARG, "wg_per_cu must be in [1, 32]");
    c->tune = n;
    return ATG_OK;
}

int atg_ctx_get_tuning(const atg_ctx *c, atg_tuning *t) {
    if (!c || !t) return ATG_ERR_ARG;
    *t = c->tune;
    return ATG_OK;
}

int atg_bls_workspace_size(const atg_ctx *c, const atg_bls_batch *b,
                           uint64_t *bytes) {
    (void)c;
    if (!b || !bytes || b->n_lc < 1 || b->n_points < 0) return ATG_ERR_ARG;
    *bytes = layout_of(b).total;
    return ATG_OK;
}

int atg_bls_run(atg_ctx *c, const atg_bls_batch *b, const atg_bls_out *o) {
    if (!c) return ATG_ERR_ARG;
    int rc = check_batch(c, b);
    if (rc != ATG_OK) return rc;
    if (!o || !o->power || !o->depth || !o->depth_err || !o->duration ||
        !o->transit_time || !o->depth_snr || !o->log_likelihood)
        return fail(c, ATG_ERR_ARG, "null output array");
    return run_impl(c, b, o, nullptr);
}

int atg_bls_run_best(atg_ctx *c, const atg_bls_batch *b,
                     const atg_bls_best *best) {
    if (!c) return ATG_ERR_ARG;
    int rc = check_batch(c, b);
    if (rc != ATG_OK) return rc;
    if (!best || !best->power || !best->depth || !best->depth_err ||
        !best->duration || !best->transit_time || !best->depth_snr ||
        !best->log_likelihood || !best->period)
        return fail(c, ATG_ERR_ARG, "null output array");
    return run_impl(c, b, nullptr, best);
}

int atg_bls_workspace_size_best(const atg_ctx *c, const atg_bls_batch *b,
                                uint64_t *bytes) {
    (void)c;
    if (!b || !bytes || b->n_lc < 1 || b->n_points < 0 || b->n_periods < 0)
        return ATG_ERR_ARG;
    *bytes = layout_of(b).total + best_bytes(b);
    return ATG_OK;
}

int atg_bls_run_sync(atg_ctx *c, const atg_bls_batch *b,
                     const atg_bls_out *o) {
    int rc = atg_bls_run(c, b, o);
    if (rc != ATG_OK) return rc;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamSynchronize");
    return ATG_OK;
}

int atg_bls_run_noalloc(atg_ctx *c, const atg_bls_batch *b,
                        const atg_bls_out *o) {
    if (!c) return ATG_ERR_ARG;
    int rc = check_batch(c, b);
    if (rc != ATG_OK) return rc;
    if (!o || !o->power || !o->depth || !o->depth_err || !o->duration ||
        !o->transit_time || !o->depth_snr || !o->log_likelihood)
        return fail(c, ATG_ERR_ARG, "null output array");
    const size_t need = layout_of(b).total;
    if (need > (size_t)ATG_NOALLOC_SCRATCH_BYTES)
        return fail(c, ATG_ERR_BUDGET,
                    "atg_bls_run_noalloc: the batch needs a workspace of " +
                        std::to_string(need) +
                        " bytes, more than the static scratch of " +
                        std::to_string((size_t)ATG_NOALLOC_SCRATCH_BYTES) +
                        " bytes; use atg_bls_run");
    hipError_t e;
    void *ws = scratch_on(c->device, &e);
    if (!ws) return hip_fail(c, e, "hipGetSymbolAddress(scratch)");
    return run_impl(c, b, o, nullptr, (char *)ws);
}

int atg_bls_top_k(atg_ctx *c, const atg_bls_batch *b, const atg_bls_out *o,
                  int k, int distance, int64_t *index_out,
                  int32_t *count_out) {
    if (!c) return ATG_ERR_ARG;
    if (!b || !o || !o->power || !index_out || !count_out ||
        !b->period_offsets)
        return fail(c, ATG_ERR_ARG, "null batch, power or output pointer");
    if (b->n_lc < 1 || b->n_lc > INT_MAX || b->n_periods < 1)
        return fail(c, ATG_ERR_ARG, "n_lc and n_periods must be >= 1");
    if (b->n_periods > INT_MAX)
        return fail(c, ATG_ERR_ARG, "n_periods must be < 2^31");
    if (k < 1 || k > 64) return fail(c, ATG_ERR_ARG, "k must be in [1, 64]");
    if (distance < 1) return fail(c, ATG_ERR_ARG, "distance must be >= 1");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    const Layout L = layout_of(b);
    int rc = ensure_workspace(c, L.total);
    if (rc != ATG_OK) return rc;
    char *base = (char *)c->ws;
    hipLaunchKernelGGL(k_topk, dim3((unsigned)b->n_lc), dim3(NT), 0,
                       c->stream, (const long long *)b->period_offsets,
                       (const double *)o->power, (const int *)o->status,
                       (int *)(base + L.tk_pos),
                       (unsigned char *)(base + L.tk_st), k, distance,
                       (long long *)index_out, (int *)count_out);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(c, e, "kernel launch");
    return ATG_OK;
}

int atg_bls_stats(atg_ctx *c, const atg_bls_batch *b, int64_t n_cand,
                  const int32_t *cand_lc, const double *period,
                  const double *duration, const double *transit_time,
                  int max_transits, double *stats,
                  int32_t *per_transit_count, double *per_transit_ll,
                  int32_t *status) {
    if (!c) return ATG_ERR_ARG;
    if (!b || !b->t || !b->y || !b->lc_offsets)
        return fail(c, ATG_ERR_ARG, "null batch, t, y or lc_offsets");
    if (b->n_lc < 1 || b->n_lc > INT_MAX)
        return fail(c, ATG_ERR_ARG, "n_lc must be in [1, 2^31)");
    if (b->y_dtype != ATG_F32 && b->y_dtype != ATG_F64)
        return fail(c, ATG_ERR_ARG, "y_dtype must be ATG_F32 or ATG_F64");
    if (b->dy && b->dy_dtype != ATG_F32 && b->dy_dtype != ATG_F64)
        return fail(c, ATG_ERR_ARG, "dy_dtype must be ATG_F32 or ATG_F64");
    if (n_cand < 1 || n_cand > INT_MAX)
        return fail(c, ATG_ERR_ARG, "n_cand must be in [1, 2^31)");
    if (!cand_lc || !period || !duration || !transit_time || !stats ||
        !per_transit_count || !per_transit_ll || !status)
        return fail(c, ATG_ERR_ARG, "null candidate or output pointer");
    if (max_transits < 1 || max_transits > MAXTR)
        return fail(c, ATG_ERR_ARG, "max_transits must be in [1, " +
                                        std::to_string(MAXTR) + "]");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    hipLaunchKernelGGL(k_stats, dim3((unsigned)n_cand), dim3(NT), 0,
                       c->stream, to_batch(b), (long long)n_cand,
                       (const int *)cand_lc, period, duration, transit_time,
                       max_transits, stats, (int *)per_transit_count,
                       per_transit_ll, (int *)status);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(c, e, "kernel launch");
    return ATG_OK;
}

int atg_bls_model(atg_ctx *c, int64_t n, const double *t, double period,
                  double duration, double transit_time, double y_in,
                  double y_out, uint8_t *mask, double *y_model) {
    if (!c) return ATG_ERR_ARG;
    if (n < 0 || (n > 0 && !t) || (!mask && !y_model))
        return fail(c, ATG_ERR_ARG, "bad size or null pointer");
    if (n == 0) return ATG_OK;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    const long long blocks = (n + NT - 1) / NT;
    if (blocks > INT_MAX) return fail(c, ATG_ERR_ARG, "too many points");
    hipLaunchKernelGGL(k_model, dim3((unsigned)blocks), dim3(NT), 0,
                       c->stream, (long long)n, t, period, duration,
                       transit_time, y_in, y_out, (unsigned char *)mask,
                       y_model);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(c, e, "kernel launch");
    return ATG_OK;
}

int atg_timing_main_kernel(atg_ctx *c, int enable, int reset,
                           double *avg_ms) {
    if (!c) return ATG_ERR_ARG;
    for (auto &pr : c->pending) {
        hipEventSynchronize(pr.second);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
            c->t_sum += ms;
            c->t_cnt += 1;
        }
        c->spare.push_back(pr.first);
        c->spare.push_back(pr.second);
    }
    c->pending.clear();
    int n = c->t_cnt;
    if (avg_ms) *avg_ms = n ? c->t_sum / n : 0.0;
    if (reset) {
        c->t_sum = 0.0;
        c->t_cnt = 0;
    }
    c->timing = enable != 0;
    return n;
}

}  // extern "C"
