/*
 * CPU oracle for the NaN-aware window-selection path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (ndfiltrs_amd/, the HIP
 * library) links, loads or calls this file; tests, the smoke check and the
 * benchmark's CPU-baseline leg use it as the checker.
 *
 * Two independent selection algorithms are restated here:
 *   direct     per-output gather + quickselect      src/filter_nd.rs:125-172,
 *                                                    src/reduction.rs:78-130
 *   histogram  block sort + sliding rank histogram   src/histogram.rs:30-631
 * They share only map_index, the key transform and the midpoint, each of which
 * is pinned by known-answer tests in tests/.
 */
#include "ndf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MSG_RANK0 "filtering requires at least one dimension"
#define MSG_SIZE0 "window sizes must be positive"
#define MSG_ORIGIN "origin must satisfy -(size // 2) <= origin <= (size - 1) // 2"
#define MSG_EMPTY_FOOTPRINT "footprint must select at least one element"
#define MSG_MIN_COUNT "min_count must be between 1 and the number of window elements"
#define MSG_HIST_FOOTPRINT "histogram reductions require rectangular size, not footprint"
#define MSG_EMPTY_AXIS "cannot map an empty axis"
#define MSG_MEMORY "unable to allocate filtering workspace"
#define MSG_OVERFLOW "shape or index arithmetic overflowed"

static int fail(int code, const char *message, char *err, size_t err_len)
{
    if (err && err_len) {
        snprintf(err, err_len, "%s", message);
    }
    return code;
}

static int fail_code(int code, char *err, size_t err_len)
{
    switch (code) {
    case NDFO_EMPTY_AXIS: return fail(code, MSG_EMPTY_AXIS, err, err_len);
    case NDFO_MEMORY: return fail(code, MSG_MEMORY, err, err_len);
    case NDFO_OVERFLOW: return fail(code, MSG_OVERFLOW, err, err_len);
    default: return fail(code, "invalid configuration", err, err_len);
    }
}

/* ------------------------------------------------------------------ */
/* Boundary mapping (src/boundary.rs:29-71)                            */
/* ------------------------------------------------------------------ */

static int64_t rem_euclid(int64_t value, int64_t modulus)
{
    int64_t r = value % modulus;
    return r < 0 ? r + modulus : r;
}

int ndfo_map_index(int64_t index, uint64_t length, int32_t mode, int64_t *mapped)
{
    if (length == 0) {
        return NDFO_EMPTY_AXIS;
    }
    if (length > (uint64_t)INT64_MAX / 4) {
        return NDFO_OVERFLOW;
    }
    const int64_t n = (int64_t)length;
    if (index >= 0 && index < n) {
        *mapped = index;
        return NDFO_OK;
    }
    switch (mode) {
    case NDFO_CONSTANT:
        *mapped = -1;
        return NDFO_OK;
    case NDFO_NEAREST:
        *mapped = index < 0 ? 0 : n - 1;
        return NDFO_OK;
    case NDFO_WRAP:
        *mapped = rem_euclid(index, n);
        return NDFO_OK;
    case NDFO_REFLECT: {
        if (n == 1) {
            *mapped = 0;
            return NDFO_OK;
        }
        const int64_t period = 2 * n;
        const int64_t reduced = rem_euclid(index, period);
        *mapped = reduced < n ? reduced : period - 1 - reduced;
        return NDFO_OK;
    }
    case NDFO_MIRROR: {
        if (n == 1) {
            *mapped = 0;
            return NDFO_OK;
        }
        const int64_t period = 2 * (n - 1);
        const int64_t reduced = rem_euclid(index, period);
        *mapped = reduced < n ? reduced : period - reduced;
        return NDFO_OK;
    }
    default:
        return NDFO_INVALID;
    }
}

/* ------------------------------------------------------------------ */
/* Keys and midpoint                                                    */
/* ------------------------------------------------------------------ */

/* Total order as unsigned integers: flip every bit of negatives, only the
 * sign bit of positives (equivalent to src/median_network.rs:21-24 plus a
 * sign-bit flip that turns its signed order into unsigned order). */
uint32_t ndfo_key32(uint32_t bits) { return bits ^ ((bits >> 31) ? 0xFFFFFFFFu : 0x80000000u); }
uint64_t ndfo_key64(uint64_t bits)
{
    return bits ^ ((bits >> 63) ? 0xFFFFFFFFFFFFFFFFull : 0x8000000000000000ull);
}
uint32_t ndfo_unkey32(uint32_t key) { return key ^ ((key >> 31) ? 0x80000000u : 0xFFFFFFFFu); }
uint64_t ndfo_unkey64(uint64_t key)
{
    return key ^ ((key >> 63) ? 0x8000000000000000ull : 0xFFFFFFFFFFFFFFFFull);
}

/* docs/semantics.md:74-85; tests/_reference.py:36-47. */
double ndfo_midpoint(double lower, double upper)
{
    const int sl = signbit(lower) != 0;
    const int su = signbit(upper) != 0;
    if (sl == su && lower == upper) {
        return lower;
    }
    if (lower == 0.0 && upper == 0.0) {
        return 0.0;
    }
    if (isinf(lower) || isinf(upper)) {
        if (isinf(lower) && isinf(upper)) {
            return NAN;
        }
        return isinf(lower) ? lower : upper;
    }
    if (sl == su) {
        return lower + (upper - lower) / 2.0;
    }
    return (lower + upper) / 2.0;
}

static const uint32_t CANONICAL_NAN32 = 0x7FC00000u;
static const uint64_t CANONICAL_NAN64 = 0x7FF8000000000000ull;

static uint64_t load_bits(const void *input, int dbytes, size_t index)
{
    if (dbytes == 4) {
        uint32_t v;
        memcpy(&v, (const char *)input + 4 * index, 4);
        return v;
    }
    uint64_t v;
    memcpy(&v, (const char *)input + 8 * index, 8);
    return v;
}

static void store_bits(void *output, int dbytes, size_t index, uint64_t bits)
{
    if (dbytes == 4) {
        uint32_t v = (uint32_t)bits;
        memcpy((char *)output + 4 * index, &v, 4);
    } else {
        memcpy((char *)output + 8 * index, &bits, 8);
    }
}

static int bits_is_nan(uint64_t bits, int dbytes)
{
    if (dbytes == 4) {
        return (bits & 0x7FFFFFFFu) > 0x7F800000u;
    }
    return (bits & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull;
}

static int bits_is_inf(uint64_t bits, int dbytes)
{
    if (dbytes == 4) {
        return (bits & 0x7FFFFFFFu) == 0x7F800000u;
    }
    return (bits & 0x7FFFFFFFFFFFFFFFull) == 0x7FF0000000000000ull;
}

static uint64_t key_of(uint64_t bits, int dbytes)
{
    return dbytes == 4 ? ndfo_key32((uint32_t)bits) : ndfo_key64(bits);
}

static uint64_t unkey(uint64_t key, int dbytes)
{
    return dbytes == 4 ? ndfo_unkey32((uint32_t)key) : ndfo_unkey64(key);
}

static double bits_to_double(uint64_t bits, int dbytes)
{
    if (dbytes == 4) {
        float f;
        uint32_t b = (uint32_t)bits;
        memcpy(&f, &b, 4);
        return (double)f;
    }
    double d;
    memcpy(&d, &bits, 8);
    return d;
}

static uint64_t canonical_nan(int dbytes) { return dbytes == 4 ? CANONICAL_NAN32 : CANONICAL_NAN64; }

/* Median of the two middle values: f64 midpoint, cast once to the dtype. */
static uint64_t midpoint_bits(uint64_t lo_bits, uint64_t hi_bits, int dbytes)
{
    const double mid = ndfo_midpoint(bits_to_double(lo_bits, dbytes), bits_to_double(hi_bits, dbytes));
    if (isnan(mid)) {
        return canonical_nan(dbytes);
    }
    if (dbytes == 4) {
        float f = (float)mid;
        uint32_t b;
        memcpy(&b, &f, 4);
        return b;
    }
    uint64_t b;
    memcpy(&b, &mid, 8);
    return b;
}

/* ------------------------------------------------------------------ */
/* Percentile interpolation (src/robust.rs:320-356), with the two-term     */
/* finite weighted sum it calls for opposite signs (src/convolution.rs:    */
/* 203-384, normalize = false) and Neumaier summation (src/mean.rs:6-26).  */
/* Every step is one IEEE f64 operation; the Makefile forbids contraction.  */
/* ------------------------------------------------------------------ */

static void split_f64(double value, double *mantissa, int *exponent)
{
    uint64_t bits;
    memcpy(&bits, &value, 8);
    const int field = (int)((bits >> 52) & 0x7FF);
    const uint64_t fraction = bits & ((1ull << 52) - 1);
    if (field == 0) {
        const int leading = 63 - __builtin_clzll(fraction);
        *mantissa = copysign((double)fraction / (double)(1ull << leading), value);
        *exponent = leading - 1074;
    } else {
        const uint64_t m = (bits & ((1ull << 63) | ((1ull << 52) - 1))) | (1023ull << 52);
        memcpy(mantissa, &m, 8);
        *exponent = field - 1023;
    }
}

static double pow2_bits(int biased)
{
    const uint64_t b = (uint64_t)biased << 52;
    double d;
    memcpy(&d, &b, 8);
    return d;
}

static double scale_f64(double value, int exponent)
{
    if (value == 0.0) return value;
    double mantissa;
    int own;
    split_f64(value, &mantissa, &own);
    const int target = own + exponent;
    if (target > 1023) return copysign(INFINITY, value);
    if (target < -1075) return copysign(0.0, value);
    if (target >= -1022) return mantissa * pow2_bits(target + 1023);
    return (mantissa * 2.2250738585072014e-308) * pow2_bits(target + 1022 + 1023);
}

typedef struct {
    double value, correction;
} neumaier_t;

static void neumaier_add(neumaier_t *s, double x)
{
    const double next = s->value + x;
    if (fabs(s->value) >= fabs(x)) {
        s->correction += (s->value - next) + x;
    } else {
        s->correction += (x - next) + s->value;
    }
    s->value = next;
}

static double finite_weighted_pair(double v0, double w0, double v1, double w1)
{
    const double values[2] = {v0, v1}, weights[2] = {w0, w1};
    neumaier_t sum = {0.0, 0.0};
    int fallback = 0;
    for (int i = 0; i < 2; ++i) {
        const double product = values[i] * weights[i];
        if (!isfinite(product) || (values[i] != 0.0 && fabs(product) < 2.2250738585072014e-308)) {
            fallback = 1;
            break;
        }
        neumaier_add(&sum, product);
    }
    const double total = sum.value + sum.correction;
    if (!fallback && isfinite(total)) return total;
    int have = 0, maximum = 0;
    for (int i = 0; i < 2; ++i) {
        if (values[i] == 0.0) continue;
        double m;
        int ev, ew;
        split_f64(values[i], &m, &ev);
        split_f64(weights[i], &m, &ew);
        if (!have || ev + ew > maximum) maximum = ev + ew;
        have = 1;
    }
    if (!have) return 0.0;
    const int exponent = maximum + 1 + 2; /* bit length of (pairs - 1) = 1, plus 2 */
    neumaier_t scaled = {0.0, 0.0};
    for (int i = 0; i < 2; ++i) {
        if (values[i] == 0.0) continue;
        double mv, mw;
        int ev, ew;
        split_f64(values[i], &mv, &ev);
        split_f64(weights[i], &mw, &ew);
        neumaier_add(&scaled, scale_f64(mv * mw, ev + ew - exponent));
    }
    return scale_f64(scaled.value + scaled.correction, exponent);
}

static double clamp_f64(double value, double lower, double upper)
{
    double x = value;
    if (x < lower) x = lower;
    if (x > upper) x = upper;
    return x;
}

double ndfo_interpolate(double lower, double upper, double fraction)
{
    uint64_t lb, ub;
    memcpy(&lb, &lower, 8);
    memcpy(&ub, &upper, 8);
    if (fraction == 0.0 || lb == ub) return lower;
    if (fraction == 1.0) return upper;
    if (isinf(lower) || isinf(upper)) {
        if (lower == -INFINITY && upper == INFINITY) return NAN;
        return isinf(lower) ? lower : upper;
    }
    if ((signbit(lower) != 0) != (signbit(upper) != 0)) {
        if (!isfinite(upper - lower)) {
            const double half_lower = lower / 2.0;
            const double half_upper = upper / 2.0;
            return clamp_f64((half_lower + (half_upper - half_lower) * fraction) * 2.0, lower, upper);
        }
        return finite_weighted_pair(lower, 1.0 - fraction, upper, fraction);
    }
    const double difference = upper - lower;
    const double result = fraction <= 0.5 ? lower + difference * fraction : upper - difference * (1.0 - fraction);
    return clamp_f64(result, lower, upper);
}

/* ------------------------------------------------------------------ */
/* Geometry (src/window.rs:44-137)                                      */
/* ------------------------------------------------------------------ */

typedef struct {
    size_t rank;
    const size_t *shape;
    const int32_t *modes;
    size_t strides[64];
    int64_t low[64];
    int64_t high[64];
    size_t count;
    int64_t *taps; /* count * rank */
    int64_t *flat; /* count */
    size_t len;
} window_t;

static int checked_mul(size_t a, size_t b, size_t *out)
{
    if (a != 0 && b > SIZE_MAX / a) {
        return 0;
    }
    *out = a * b;
    return 1;
}

static int window_build(window_t *w, size_t rank, const size_t *shape, const size_t *size,
                        const uint8_t *footprint, const ptrdiff_t *origin, const int32_t *modes,
                        char *err, size_t err_len)
{
    memset(w, 0, sizeof(*w));
    if (rank == 0) {
        return fail(NDFO_INVALID, MSG_RANK0, err, err_len);
    }
    if (rank > 64) {
        return fail(NDFO_INVALID, "array rank exceeds 64", err, err_len);
    }
    for (size_t a = 0; a < rank; ++a) {
        if (size[a] == 0) {
            return fail(NDFO_INVALID, MSG_SIZE0, err, err_len);
        }
    }
    size_t volume = 1;
    for (size_t a = 0; a < rank; ++a) {
        if (!checked_mul(volume, size[a], &volume)) {
            return fail_code(NDFO_OVERFLOW, err, err_len);
        }
    }
    int64_t anchors[64];
    for (size_t a = 0; a < rank; ++a) {
        if (size[a] > (size_t)INT64_MAX) {
            return fail_code(NDFO_OVERFLOW, err, err_len);
        }
        const int64_t length = (int64_t)size[a];
        const int64_t shift = (int64_t)origin[a];
        if (shift < -(length / 2) || shift > (length - 1) / 2) {
            return fail(NDFO_INVALID, MSG_ORIGIN, err, err_len);
        }
        anchors[a] = length / 2 + shift;
    }
    size_t len = 1;
    for (size_t a = 0; a < rank; ++a) {
        if (!checked_mul(len, shape[a], &len)) {
            return fail_code(NDFO_OVERFLOW, err, err_len);
        }
    }
    w->strides[rank - 1] = 1;
    for (size_t a = rank - 1; a-- > 0;) {
        if (!checked_mul(w->strides[a + 1], shape[a + 1], &w->strides[a])) {
            return fail_code(NDFO_OVERFLOW, err, err_len);
        }
    }
    size_t count = 0;
    if (footprint) {
        for (size_t i = 0; i < volume; ++i) {
            count += footprint[i] != 0;
        }
    } else {
        count = volume;
    }
    if (count == 0) {
        return fail(NDFO_INVALID, MSG_EMPTY_FOOTPRINT, err, err_len);
    }
    w->taps = (int64_t *)malloc(sizeof(int64_t) * count * rank);
    w->flat = (int64_t *)malloc(sizeof(int64_t) * count);
    if (!w->taps || !w->flat) {
        free(w->taps);
        free(w->flat);
        w->taps = w->flat = NULL;
        return fail_code(NDFO_MEMORY, err, err_len);
    }
    for (size_t a = 0; a < rank; ++a) {
        w->low[a] = INT64_MAX;
        w->high[a] = INT64_MIN;
    }
    size_t index[64] = {0};
    size_t t = 0;
    for (size_t element = 0; element < volume; ++element) {
        if (!footprint || footprint[element]) {
            int64_t flat = 0;
            for (size_t a = 0; a < rank; ++a) {
                const int64_t offset = (int64_t)index[a] - anchors[a];
                flat += offset * (int64_t)w->strides[a];
                if (offset < w->low[a]) w->low[a] = offset;
                if (offset > w->high[a]) w->high[a] = offset;
                w->taps[t * rank + a] = offset;
            }
            w->flat[t] = flat;
            ++t;
        }
        for (size_t a = rank; a-- > 0;) {
            if (++index[a] < size[a]) {
                break;
            }
            index[a] = 0;
        }
    }
    w->rank = rank;
    w->shape = shape;
    w->modes = modes;
    w->count = count;
    w->len = len;
    return NDFO_OK;
}

static void window_free(window_t *w)
{
    free(w->taps);
    free(w->flat);
    w->taps = w->flat = NULL;
}

/* ------------------------------------------------------------------ */
/* Direct selection                                                      */
/* ------------------------------------------------------------------ */

/* k-th smallest (0-based) of a[0..n); afterwards a[0..k) <= a[k] <= a[k+1..n). */
static uint64_t select_kth(uint64_t *a, ptrdiff_t n, ptrdiff_t k)
{
    ptrdiff_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const ptrdiff_t mid = lo + (hi - lo) / 2;
        uint64_t x = a[lo], y = a[mid], z = a[hi];
        uint64_t pivot = x < y ? (y < z ? y : (x < z ? z : x)) : (x < z ? x : (y < z ? z : y));
        ptrdiff_t i = lo, j = hi;
        while (i <= j) {
            while (a[i] < pivot) ++i;
            while (a[j] > pivot) --j;
            if (i <= j) {
                const uint64_t tmp = a[i];
                a[i] = a[j];
                a[j] = tmp;
                ++i;
                --j;
            }
        }
        if (k <= j) {
            hi = j;
        } else if (k >= i) {
            lo = i;
        } else {
            return a[k];
        }
    }
    return a[k];
}

typedef struct {
    int dbytes;
    const void *input;
    void *output;
    const window_t *w;
    uint64_t cval_bits;
    int statistic;
    size_t parameter;
    int omit;
    int ignore_inf;
    size_t min_count;
    const size_t *positions; /* NULL: every output; else outputs positions[begin..end) only */
    double q;                /* percentile in [0, 100] (NDFO_STAT_PERCENTILE) */
} direct_job_t;

typedef struct {
    const direct_job_t *job;
    size_t begin, end;
    int status;
} direct_part_t;

static int compare_keys(const void *a, const void *b)
{
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

static uint64_t double_to_bits(double value, int dbytes)
{
    if (isnan(value)) return canonical_nan(dbytes);
    if (dbytes == 4) {
        float f = (float)value;
        uint32_t b;
        memcpy(&b, &f, 4);
        return b;
    }
    uint64_t b;
    memcpy(&b, &value, 8);
    return b;
}

/* percentile() of src/robust.rs:358-367 over the retained keys: sort, rank
 * (q / 100) * (n - 1), interpolate the floor and the next sample (clamped to
 * the last), cast the f64 result once. */
static uint64_t percentile_keys(uint64_t *keys, size_t n, double q, int dbytes)
{
    qsort(keys, n, sizeof(uint64_t), compare_keys);
    if (q == 100.0) return unkey(keys[n - 1], dbytes);
    const double rank = (q / 100.0) * (double)(n - 1);
    const size_t lower = (size_t)floor(rank);
    const size_t upper = lower + 1 < n - 1 ? lower + 1 : n - 1;
    const double value = ndfo_interpolate(bits_to_double(unkey(keys[lower], dbytes), dbytes),
                                          bits_to_double(unkey(keys[upper], dbytes), dbytes), rank - (double)lower);
    return double_to_bits(value, dbytes);
}

static uint64_t finish_keys(uint64_t *keys, size_t n, int nan_seen, const direct_job_t *job)
{
    const int db = job->dbytes;
    if (!job->omit) {
        if (nan_seen) return canonical_nan(db);
    } else if (n < job->min_count) {
        return canonical_nan(db);
    }
    if (job->statistic == NDFO_STAT_PERCENTILE) {
        if (n == 0) return canonical_nan(db);
        return percentile_keys(keys, n, job->q, db);
    }
    if (job->statistic == NDFO_STAT_MIN || job->statistic == NDFO_STAT_MAX) {
        /* Minimum / Maximum: total-order extreme of the retained samples
         * (src/reduction.rs:110-119); none retained gives NaN. */
        if (n == 0) return canonical_nan(db);
        uint64_t best = keys[0];
        for (size_t i = 1; i < n; ++i) {
            if (job->statistic == NDFO_STAT_MIN ? keys[i] < best : keys[i] > best) best = keys[i];
        }
        return unkey(best, db);
    }
    if (job->statistic == NDFO_STAT_RANK) {
        if (job->parameter >= n) return canonical_nan(db);
        return unkey(select_kth(keys, (ptrdiff_t)n, (ptrdiff_t)job->parameter), db);
    }
    if (n == 0) return canonical_nan(db);
    if (n % 2 == 1) {
        return unkey(select_kth(keys, (ptrdiff_t)n, (ptrdiff_t)(n / 2)), db);
    }
    const uint64_t lower = select_kth(keys, (ptrdiff_t)n, (ptrdiff_t)(n / 2 - 1));
    uint64_t upper = keys[n / 2];
    for (size_t i = n / 2 + 1; i < n; ++i) {
        if (keys[i] < upper) upper = keys[i];
    }
    return midpoint_bits(unkey(lower, db), unkey(upper, db), db);
}

static void *direct_worker(void *arg)
{
    direct_part_t *part = (direct_part_t *)arg;
    const direct_job_t *job = part->job;
    const window_t *w = job->w;
    const size_t rank = w->rank;
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * w->count);
    if (!keys) {
        part->status = NDFO_MEMORY;
        return NULL;
    }
    int64_t coord[64];
    for (size_t item = part->begin; item < part->end; ++item) {
        const size_t center = job->positions ? job->positions[item] : item;
        size_t rem = center;
        int interior = 1;
        for (size_t a = 0; a < rank; ++a) {
            coord[a] = (int64_t)(rem / w->strides[a]);
            rem %= w->strides[a];
            interior &= coord[a] + w->low[a] >= 0 && coord[a] + w->high[a] < (int64_t)w->shape[a];
        }
        size_t n = 0;
        int nan_seen = 0;
        for (size_t t = 0; t < w->count; ++t) {
            uint64_t bits;
            if (interior) {
                bits = load_bits(job->input, job->dbytes, (size_t)((int64_t)center + w->flat[t]));
            } else {
                size_t source = 0;
                int outside = 0;
                for (size_t a = 0; a < rank; ++a) {
                    int64_t mapped;
                    const int st = ndfo_map_index(coord[a] + w->taps[t * rank + a], w->shape[a], w->modes[a], &mapped);
                    if (st != NDFO_OK) {
                        part->status = st;
                        free(keys);
                        return NULL;
                    }
                    if (mapped < 0) {
                        outside = 1;
                        break;
                    }
                    source += (size_t)mapped * w->strides[a];
                }
                bits = outside ? job->cval_bits : load_bits(job->input, job->dbytes, source);
            }
            if (bits_is_nan(bits, job->dbytes)) {
                nan_seen = 1;
                continue;
            }
            if (job->omit && job->ignore_inf && bits_is_inf(bits, job->dbytes)) {
                continue;
            }
            keys[n++] = key_of(bits, job->dbytes);
        }
        store_bits(job->output, job->dbytes, center, finish_keys(keys, n, nan_seen, job));
    }
    free(keys);
    part->status = NDFO_OK;
    return NULL;
}

static int run_parts(void *(*worker)(void *), void *parts, size_t part_size, int nthreads)
{
    pthread_t threads[256];
    int started = 0;
    for (int t = 1; t < nthreads; ++t) {
        if (pthread_create(&threads[t], NULL, worker, (char *)parts + part_size * (size_t)t) != 0) {
            break;
        }
        started = t;
    }
    worker(parts);
    for (int t = started + 1; t < nthreads; ++t) {
        worker((char *)parts + part_size * (size_t)t);
    }
    for (int t = 1; t <= started; ++t) {
        pthread_join(threads[t], NULL);
    }
    return NDFO_OK;
}

/* ------------------------------------------------------------------ */
/* Histogram method (src/histogram.rs)                                   */
/* ------------------------------------------------------------------ */

#define CODE_OMITTED 0xFFFFFFFFu
#define CODE_NAN 0xFFFFFFFEu
#define GROUP_SHIFT 12
#define MIN_BLOCK 32
#define OUTSIDE SIZE_MAX

static void block_lengths(size_t rank, const size_t *shape, const size_t *size, size_t *block)
{
    for (size_t a = 0; a < rank; ++a) {
        const size_t window = size[a] > MIN_BLOCK ? size[a] : MIN_BLOCK;
        size_t blocks = shape[a] / window;
        if (blocks < 1) blocks = 1;
        size_t len = (shape[a] + blocks - 1) / blocks;
        block[a] = len < 1 ? 1 : len;
    }
}

static int validate_support(size_t rank, const size_t *shape, const size_t *size)
{
    size_t block[64];
    block_lengths(rank, shape, size, block);
    size_t support = 1;
    for (size_t a = 0; a < rank; ++a) {
        const size_t length = block[a] + size[a] - 1;
        if (length < block[a] || !checked_mul(support, length, &support)) {
            return NDFO_OVERFLOW;
        }
    }
    return support > (size_t)CODE_NAN ? NDFO_OVERFLOW : NDFO_OK;
}

typedef struct {
    uint64_t key;
    uint32_t index;
} pair_t;

/* LSD radix sort on the key, skipping bytes that every key shares. */
static void sort_pairs(pair_t *a, pair_t *tmp, size_t n)
{
    if (n < 2) return;
    uint64_t all_or = 0, all_and = ~0ull;
    for (size_t i = 0; i < n; ++i) {
        all_or |= a[i].key;
        all_and &= a[i].key;
    }
    const uint64_t varying = all_or ^ all_and;
    for (int shift = 0; shift < 64; shift += 8) {
        if (((varying >> shift) & 0xFF) == 0) continue;
        size_t count[257] = {0};
        for (size_t i = 0; i < n; ++i) count[((a[i].key >> shift) & 0xFF) + 1]++;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (size_t i = 0; i < n; ++i) tmp[count[(a[i].key >> shift) & 0xFF]++] = a[i];
        memcpy(a, tmp, n * sizeof(pair_t));
    }
}

typedef struct {
    uint64_t *words;
    uint32_t *groups;
    size_t count;
    size_t nan;
} histo_t;

static inline void histo_add(histo_t *h, uint32_t code)
{
    if (code < CODE_NAN) {
        h->words[code >> 6] |= 1ull << (code & 63);
        h->groups[code >> GROUP_SHIFT] += 1;
        h->count += 1;
    } else if (code == CODE_NAN) {
        h->nan += 1;
    }
}

static inline void histo_remove(histo_t *h, uint32_t code)
{
    if (code < CODE_NAN) {
        h->words[code >> 6] &= ~(1ull << (code & 63));
        h->groups[code >> GROUP_SHIFT] -= 1;
        h->count -= 1;
    } else if (code == CODE_NAN) {
        h->nan -= 1;
    }
}

static size_t histo_select(const histo_t *h, size_t k)
{
    size_t group = 0;
    while (k >= h->groups[group]) {
        k -= h->groups[group];
        ++group;
    }
    size_t word = group << (GROUP_SHIFT - 6);
    for (;;) {
        const size_t c = (size_t)__builtin_popcountll(h->words[word]);
        if (k < c) break;
        k -= c;
        ++word;
    }
    uint64_t bits = h->words[word];
    for (size_t i = 0; i < k; ++i) bits &= bits - 1;
    return (word << 6) + (size_t)__builtin_ctzll(bits);
}

typedef struct {
    int dbytes;
    const void *input;
    void *output;
    size_t rank;
    const size_t *shape;
    const size_t *size;
    const int32_t *modes;
    int64_t anchors[64];
    size_t strides[64];
    size_t block[64];
    size_t nblocks[64];
    size_t order[64];
    size_t total_blocks;
    uint64_t cval_bits;
    int omit;
    int ignore_inf;
    size_t min_count;
    int statistic;
    size_t parameter;
    double q; /* NDFO_STAT_PERCENTILE */
} hjob_t;

typedef struct {
    const hjob_t *job;
    pthread_mutex_t *lock;
    size_t *next;
    int status;
} hpart_t;

/* Visit every support index of a box (low, lengths) in C order. */
static void box_apply(histo_t *h, const uint32_t *codes, size_t rank, const size_t *low, const size_t *lengths,
                      const size_t *sstrides, int add)
{
    size_t counter[64] = {0};
    for (;;) {
        size_t index = 0;
        for (size_t a = 0; a < rank; ++a) index += (low[a] + counter[a]) * sstrides[a];
        if (add) {
            histo_add(h, codes[index]);
        } else {
            histo_remove(h, codes[index]);
        }
        size_t a = rank;
        for (;;) {
            if (a == 0) return;
            --a;
            if (++counter[a] < lengths[a]) break;
            counter[a] = 0;
        }
    }
}

static uint64_t histo_finish(const histo_t *h, const uint64_t *value_bits, const hjob_t *job)
{
    const int db = job->dbytes;
    if (!job->omit) {
        if (h->nan > 0) return canonical_nan(db);
    } else if (h->count < job->min_count) {
        return canonical_nan(db);
    }
    if (job->statistic == NDFO_STAT_RANK) {
        if (job->parameter >= h->count) return canonical_nan(db);
        return value_bits[histo_select(h, job->parameter)];
    }
    if (h->count == 0) return canonical_nan(db);
    if (job->statistic == NDFO_STAT_MIN) return value_bits[histo_select(h, 0)];
    if (job->statistic == NDFO_STAT_MAX) return value_bits[histo_select(h, h->count - 1)];
    if (job->statistic == NDFO_STAT_PERCENTILE) {
        /* percentile(), src/robust.rs:358-367, on the ranked window. */
        const size_t n = h->count;
        if (job->q == 100.0) return value_bits[histo_select(h, n - 1)];
        const double rank = (job->q / 100.0) * (double)(n - 1);
        const size_t lower = (size_t)floor(rank);
        const size_t upper = lower + 1 < n - 1 ? lower + 1 : n - 1;
        const double value = ndfo_interpolate(bits_to_double(value_bits[histo_select(h, lower)], db),
                                              bits_to_double(value_bits[histo_select(h, upper)], db),
                                              rank - (double)lower);
        return double_to_bits(value, db);
    }
    if (h->count % 2 == 1) {
        return value_bits[histo_select(h, h->count / 2)];
    }
    return midpoint_bits(value_bits[histo_select(h, h->count / 2 - 1)], value_bits[histo_select(h, h->count / 2)],
                         db);
}

static int hist_block(const hjob_t *job, const size_t *start, const size_t *extent, size_t **maps, uint32_t **codes,
                      pair_t **pairs, pair_t **tmp, uint64_t **values, uint64_t **words, uint32_t **groups,
                      size_t *capacity)
{
    const size_t rank = job->rank;
    size_t lengths[64], sstrides[64];
    for (size_t a = 0; a < rank; ++a) lengths[a] = extent[a] + job->size[a] - 1;
    sstrides[rank - 1] = 1;
    for (size_t a = rank - 1; a-- > 0;) sstrides[a] = sstrides[a + 1] * lengths[a + 1];
    const size_t total = sstrides[0] * lengths[0];
    if (total > *capacity) {
        free(*codes);
        free(*pairs);
        free(*tmp);
        free(*values);
        free(*words);
        free(*groups);
        *codes = (uint32_t *)malloc(total * sizeof(uint32_t));
        *pairs = (pair_t *)malloc(total * sizeof(pair_t));
        *tmp = (pair_t *)malloc(total * sizeof(pair_t));
        *values = (uint64_t *)malloc(total * sizeof(uint64_t));
        *words = (uint64_t *)malloc(((total + 63) / 64 + 1) * sizeof(uint64_t));
        *groups = (uint32_t *)malloc(((total >> GROUP_SHIFT) + 2) * sizeof(uint32_t));
        if (!*codes || !*pairs || !*tmp || !*values || !*words || !*groups) {
            *capacity = 0;
            return NDFO_MEMORY;
        }
        *capacity = total;
    }
    for (size_t a = 0; a < rank; ++a) {
        const int64_t first = (int64_t)start[a] - job->anchors[a];
        for (size_t s = 0; s < lengths[a]; ++s) {
            int64_t mapped;
            const int st = ndfo_map_index(first + (int64_t)s, job->shape[a], job->modes[a], &mapped);
            if (st != NDFO_OK) return st;
            maps[a][s] = mapped < 0 ? OUTSIDE : (size_t)mapped * job->strides[a];
        }
    }
    size_t npairs = 0;
    size_t counter[64] = {0};
    for (size_t si = 0; si < total; ++si) {
        size_t base = 0;
        int outside = 0;
        for (size_t a = 0; a < rank; ++a) {
            const size_t off = maps[a][counter[a]];
            if (off == OUTSIDE) {
                outside = 1;
            } else {
                base += off;
            }
        }
        const uint64_t bits = outside ? job->cval_bits : load_bits(job->input, job->dbytes, base);
        const int nan = bits_is_nan(bits, job->dbytes);
        const int retained = !nan && !(job->omit && job->ignore_inf && bits_is_inf(bits, job->dbytes));
        (*codes)[si] = CODE_OMITTED;
        if (retained) {
            (*pairs)[npairs].key = key_of(bits, job->dbytes);
            (*pairs)[npairs].index = (uint32_t)si;
            ++npairs;
        } else if (!job->omit) {
            (*codes)[si] = CODE_NAN;
        }
        for (size_t a = rank; a-- > 0;) {
            if (++counter[a] < lengths[a]) break;
            counter[a] = 0;
        }
    }
    sort_pairs(*pairs, *tmp, npairs);
    for (size_t r = 0; r < npairs; ++r) {
        (*codes)[(*pairs)[r].index] = (uint32_t)r;
        (*values)[r] = unkey((*pairs)[r].key, job->dbytes);
    }
    memset(*words, 0, ((npairs + 63) / 64 + 1) * sizeof(uint64_t));
    memset(*groups, 0, ((npairs >> GROUP_SHIFT) + 2) * sizeof(uint32_t));
    histo_t h = {*words, *groups, 0, 0};

    size_t position[64] = {0}, low[64], box[64];
    int forward[64];
    for (size_t a = 0; a < rank; ++a) {
        forward[a] = 1;
        low[a] = 0;
        box[a] = job->size[a];
    }
    box_apply(&h, *codes, rank, low, box, sstrides, 1);
    for (;;) {
        size_t out = 0;
        for (size_t a = 0; a < rank; ++a) out += (start[a] + position[a]) * job->strides[a];
        store_bits(job->output, job->dbytes, out, histo_finish(&h, *values, job));
        /* Move along the fastest axis that can still move, reversing the
         * direction of every faster axis that cannot. */
        size_t moved = SIZE_MAX;
        for (size_t i = rank; i-- > 0;) {
            const size_t axis = job->order[i];
            const int can = forward[axis] ? position[axis] + 1 < extent[axis] : position[axis] > 0;
            if (can) {
                moved = axis;
                break;
            }
            forward[axis] = !forward[axis];
        }
        if (moved == SIZE_MAX) break;
        size_t removed, added;
        if (forward[moved]) {
            removed = position[moved];
            added = position[moved] + job->size[moved];
        } else {
            removed = position[moved] + job->size[moved] - 1;
            added = position[moved] - 1;
        }
        for (size_t a = 0; a < rank; ++a) {
            low[a] = position[a];
            box[a] = job->size[a];
        }
        box[moved] = 1;
        low[moved] = removed;
        box_apply(&h, *codes, rank, low, box, sstrides, 0);
        low[moved] = added;
        box_apply(&h, *codes, rank, low, box, sstrides, 1);
        if (forward[moved]) {
            position[moved] += 1;
        } else {
            position[moved] -= 1;
        }
    }
    return NDFO_OK;
}

static void *hist_worker(void *arg)
{
    hpart_t *part = (hpart_t *)arg;
    const hjob_t *job = part->job;
    const size_t rank = job->rank;
    size_t *maps[64];
    for (size_t a = 0; a < rank; ++a) {
        maps[a] = (size_t *)malloc(sizeof(size_t) * (job->block[a] + job->size[a]));
        if (!maps[a]) {
            part->status = NDFO_MEMORY;
            for (size_t b = 0; b < a; ++b) free(maps[b]);
            return NULL;
        }
    }
    uint32_t *codes = NULL, *groups = NULL;
    pair_t *pairs = NULL, *tmp = NULL;
    uint64_t *values = NULL, *words = NULL;
    size_t capacity = 0;
    part->status = NDFO_OK;
    for (;;) {
        pthread_mutex_lock(part->lock);
        const size_t b = (*part->next)++;
        pthread_mutex_unlock(part->lock);
        if (b >= job->total_blocks) break;
        size_t start[64], extent[64];
        size_t rem = b;
        for (size_t a = rank; a-- > 0;) {
            const size_t idx = rem % job->nblocks[a];
            rem /= job->nblocks[a];
            start[a] = idx * job->block[a];
            extent[a] = job->block[a] < job->shape[a] - start[a] ? job->block[a] : job->shape[a] - start[a];
        }
        const int st = hist_block(job, start, extent, maps, &codes, &pairs, &tmp, &values, &words, &groups, &capacity);
        if (st != NDFO_OK) {
            part->status = st;
            break;
        }
    }
    for (size_t a = 0; a < rank; ++a) free(maps[a]);
    free(codes);
    free(pairs);
    free(tmp);
    free(values);
    free(words);
    free(groups);
    return NULL;
}

/* ------------------------------------------------------------------ */
/* Entry point                                                           */
/* ------------------------------------------------------------------ */

static uint64_t cval_to_bits(double cval, int dbytes)
{
    if (dbytes == 4) {
        float f = (float)cval;
        uint32_t b;
        memcpy(&b, &f, 4);
        return b;
    }
    uint64_t b;
    memcpy(&b, &cval, 8);
    return b;
}

static int filter_histogram(int dbytes, const void *input, void *output, size_t rank, const size_t *shape,
                            const size_t *size, const uint8_t *footprint, const ptrdiff_t *origin,
                            const int32_t *modes, double cval, int32_t statistic, size_t parameter, double q,
                            int omit, int ignore_inf, size_t min_count, int nthreads, char *err, size_t err_len)
{
    /* src/plan.rs:294-312 then validate_rectangular (:154-197). */
    if (footprint) return fail(NDFO_INVALID, MSG_HIST_FOOTPRINT, err, err_len);
    if (rank == 0) return fail(NDFO_INVALID, MSG_RANK0, err, err_len);
    if (rank > 64) return fail(NDFO_INVALID, "array rank exceeds 64", err, err_len);
    for (size_t a = 0; a < rank; ++a) {
        if (size[a] == 0) return fail(NDFO_INVALID, MSG_SIZE0, err, err_len);
    }
    size_t len = 1, volume = 1;
    for (size_t a = 0; a < rank; ++a) {
        if (!checked_mul(len, shape[a], &len)) return fail_code(NDFO_OVERFLOW, err, err_len);
    }
    for (size_t a = 0; a < rank; ++a) {
        if (!checked_mul(volume, size[a], &volume)) return fail_code(NDFO_OVERFLOW, err, err_len);
    }
    hjob_t job;
    memset(&job, 0, sizeof(job));
    for (size_t a = 0; a < rank; ++a) {
        if (size[a] > (size_t)INT64_MAX) return fail_code(NDFO_OVERFLOW, err, err_len);
        const int64_t length = (int64_t)size[a];
        const int64_t shift = (int64_t)origin[a];
        if (shift < -(length / 2) || shift > (length - 1) / 2) return fail(NDFO_INVALID, MSG_ORIGIN, err, err_len);
        job.anchors[a] = length / 2 + shift;
    }
    if (omit && (min_count == 0 || min_count > volume)) return fail(NDFO_INVALID, MSG_MIN_COUNT, err, err_len);
    {
        const int st = validate_support(rank, shape, size);
        if (st != NDFO_OK) return fail_code(st, err, err_len);
    }
    if (len == 0) return NDFO_OK;
    job.dbytes = dbytes;
    job.input = input;
    job.output = output;
    job.rank = rank;
    job.shape = shape;
    job.size = size;
    job.modes = modes;
    job.cval_bits = cval_to_bits(cval, dbytes);
    job.omit = omit;
    job.ignore_inf = ignore_inf;
    job.min_count = min_count;
    job.statistic = statistic;
    job.parameter = parameter;
    job.q = q;
    job.strides[rank - 1] = 1;
    for (size_t a = rank - 1; a-- > 0;) job.strides[a] = job.strides[a + 1] * shape[a + 1];
    block_lengths(rank, shape, size, job.block);
    job.total_blocks = 1;
    for (size_t a = 0; a < rank; ++a) {
        job.nblocks[a] = (shape[a] + job.block[a] - 1) / job.block[a];
        job.total_blocks *= job.nblocks[a];
    }
    /* Fastest axis: largest window size, ties keep the last axis. */
    size_t fastest = 0;
    for (size_t a = 0; a < rank; ++a) {
        if (size[a] >= size[fastest]) fastest = a;
    }
    size_t o = 0;
    for (size_t a = 0; a < rank; ++a) {
        if (a != fastest) job.order[o++] = a;
    }
    job.order[o] = fastest;

    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_mutex_t lock = PTHREAD_MUTEX_INITIALIZER;
    size_t next = 0;
    hpart_t parts[256];
    for (int t = 0; t < nthreads; ++t) {
        parts[t].job = &job;
        parts[t].lock = &lock;
        parts[t].next = &next;
        parts[t].status = NDFO_OK;
    }
    run_parts(hist_worker, parts, sizeof(hpart_t), nthreads);
    for (int t = 0; t < nthreads; ++t) {
        if (parts[t].status != NDFO_OK) return fail_code(parts[t].status, err, err_len);
    }
    return NDFO_OK;
}

int ndfo_filter(int dtype_bytes, const void *input, void *output, size_t rank, const size_t *shape,
                const size_t *size, const uint8_t *footprint, const ptrdiff_t *origin, const int32_t *modes,
                double cval, int32_t statistic, size_t parameter, int32_t omit_nan, int32_t ignore_inf,
                size_t min_count, int32_t method, int32_t nthreads, char *err, size_t err_len)
{
    return ndfo_filter_q(dtype_bytes, input, output, rank, shape, size, footprint, origin, modes, cval, statistic,
                         parameter, 0.0, omit_nan, ignore_inf, min_count, method, nthreads, err, err_len);
}

int ndfo_filter_q(int dtype_bytes, const void *input, void *output, size_t rank, const size_t *shape,
                  const size_t *size, const uint8_t *footprint, const ptrdiff_t *origin, const int32_t *modes,
                  double cval, int32_t statistic, size_t parameter, double q, int32_t omit_nan, int32_t ignore_inf,
                  size_t min_count, int32_t method, int32_t nthreads, char *err, size_t err_len)
{
    if (dtype_bytes != 4 && dtype_bytes != 8) return fail(NDFO_INVALID, "dtype must be float32 or float64", err, err_len);
    for (size_t a = 0; a < rank; ++a) {
        if (modes[a] < 0 || modes[a] > 4) return fail(NDFO_INVALID, "unsupported boundary mode", err, err_len);
    }
    if (method == NDFO_METHOD_HISTOGRAM) {
        if (statistic == NDFO_STAT_PERCENTILE && (!isfinite(q) || q < 0.0 || q > 100.0)) {
            return fail(NDFO_INVALID, "percentile must be finite and between zero and 100", err, err_len);
        }
        return filter_histogram(dtype_bytes, input, output, rank, shape, size, footprint, origin, modes, cval,
                                statistic, parameter, q, omit_nan != 0, ignore_inf != 0, min_count, nthreads, err,
                                err_len);
    }
    window_t w;
    int st = window_build(&w, rank, shape, size, footprint, origin, modes, err, err_len);
    if (st != NDFO_OK) return st;
    if (statistic == NDFO_STAT_PERCENTILE) {
        /* RobustReduction::validate (src/robust.rs:189-241). */
        if (!isfinite(q) || q < 0.0 || q > 100.0) {
            window_free(&w);
            return fail(NDFO_INVALID, "percentile must be finite and between zero and 100", err, err_len);
        }
        if (omit_nan && (min_count == 0 || min_count > w.count)) {
            window_free(&w);
            return fail(NDFO_INVALID, "min_count must be between one and the number of window elements", err,
                        err_len);
        }
    }
    if (omit_nan && (min_count == 0 || min_count > w.count)) {
        window_free(&w);
        return fail(NDFO_INVALID, MSG_MIN_COUNT, err, err_len);
    }
    if (w.len == 0) {
        window_free(&w);
        return NDFO_OK;
    }
    direct_job_t job = {dtype_bytes, input, output, &w, cval_to_bits(cval, dtype_bytes), statistic, parameter,
                        omit_nan != 0, ignore_inf != 0, min_count, NULL, q};
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((size_t)nthreads > w.len) nthreads = (int)w.len;
    direct_part_t parts[256];
    const size_t chunk = (w.len + (size_t)nthreads - 1) / (size_t)nthreads;
    for (int t = 0; t < nthreads; ++t) {
        parts[t].job = &job;
        parts[t].begin = chunk * (size_t)t < w.len ? chunk * (size_t)t : w.len;
        parts[t].end = parts[t].begin + chunk < w.len ? parts[t].begin + chunk : w.len;
        parts[t].status = NDFO_OK;
    }
    run_parts(direct_worker, parts, sizeof(direct_part_t), nthreads);
    st = NDFO_OK;
    for (int t = 0; t < nthreads; ++t) {
        if (parts[t].status != NDFO_OK) st = parts[t].status;
    }
    window_free(&w);
    return st == NDFO_OK ? NDFO_OK : fail_code(st, err, err_len);
}

/* ------------------------------------------------------------------ */
/* Separable extrema (src/extrema.rs:125-436)                             */
/* ------------------------------------------------------------------ */

/* A window summary: the extreme key so far, a propagated NaN, or nothing
 * retained yet (src/extrema.rs:11-62). Equal keys are equal bits, so which
 * of two equal extremes survives does not matter. */
typedef struct {
    uint64_t key;
    unsigned char nan;
    unsigned char empty;
} summary_t;

static summary_t summary_of(uint64_t bits, int dbytes, int omit, int ignore_inf)
{
    summary_t out = {0, 0, 0};
    if (!omit) {
        if (bits_is_nan(bits, dbytes)) {
            out.nan = 1;
        } else {
            out.key = key_of(bits, dbytes);
        }
        return out;
    }
    if (bits_is_nan(bits, dbytes) || (ignore_inf && bits_is_inf(bits, dbytes))) {
        out.empty = 1;
    } else {
        out.key = key_of(bits, dbytes);
    }
    return out;
}

static summary_t summary_combine(summary_t a, summary_t b, int maximum)
{
    if (a.nan || b.nan) {
        summary_t n = {0, 1, 0};
        return n;
    }
    if (a.empty) return b;
    if (b.empty) return a;
    if (maximum ? a.key < b.key : a.key > b.key) return b;
    return a;
}

static int retained_bits(uint64_t bits, int dbytes, int omit, int ignore_inf)
{
    if (!omit) return 1;
    return !bits_is_nan(bits, dbytes) && !(ignore_inf && bits_is_inf(bits, dbytes));
}

int ndfo_extrema(int dtype_bytes, const void *input, void *output, size_t rank, const size_t *shape,
                 const size_t *size, const ptrdiff_t *origin, const int32_t *modes, double cval, int32_t maximum,
                 int32_t omit_nan, int32_t ignore_inf, size_t min_count, char *err, size_t err_len)
{
    const int db = dtype_bytes;
    if (db != 4 && db != 8) return fail(NDFO_INVALID, "dtype must be float32 or float64", err, err_len);
    if (rank == 0) return fail(NDFO_INVALID, MSG_RANK0, err, err_len);
    if (rank > 64) return fail(NDFO_INVALID, "array rank exceeds 64", err, err_len);
    for (size_t a = 0; a < rank; ++a) {
        if (size[a] == 0) return fail(NDFO_INVALID, MSG_SIZE0, err, err_len);
        if (modes[a] < 0 || modes[a] > 4) return fail(NDFO_INVALID, "unsupported boundary mode", err, err_len);
    }
    size_t volume = 1, len = 1;
    for (size_t a = 0; a < rank; ++a) {
        if (!checked_mul(volume, size[a], &volume)) return fail_code(NDFO_OVERFLOW, err, err_len);
    }
    int64_t anchors[64];
    for (size_t a = 0; a < rank; ++a) {
        const int64_t length = (int64_t)size[a], shift = (int64_t)origin[a];
        if (shift < -(length / 2) || shift > (length - 1) / 2) return fail(NDFO_INVALID, MSG_ORIGIN, err, err_len);
        anchors[a] = length / 2 + shift;
    }
    if (omit_nan && (min_count == 0 || min_count > volume)) return fail(NDFO_INVALID, MSG_MIN_COUNT, err, err_len);
    for (size_t a = 0; a < rank; ++a) {
        if (!checked_mul(len, shape[a], &len)) return fail_code(NDFO_OVERFLOW, err, err_len);
    }
    if (len == 0) return NDFO_OK;
    size_t strides[64];
    strides[rank - 1] = 1;
    for (size_t a = rank - 1; a-- > 0;) strides[a] = strides[a + 1] * shape[a + 1];
    const uint64_t cbits = cval_to_bits(cval, db);
    const summary_t csum = summary_of(cbits, db, omit_nan, ignore_inf);
    const int cret = retained_bits(cbits, db, omit_nan, ignore_inf);
    size_t longest = 0;
    for (size_t a = 0; a < rank; ++a) {
        if (shape[a] + size[a] - 1 > longest) longest = shape[a] + size[a] - 1;
    }
    summary_t *prev = malloc(sizeof(summary_t) * len), *next = malloc(sizeof(summary_t) * len);
    size_t *pc = malloc(sizeof(size_t) * len), *nc = malloc(sizeof(size_t) * len);
    summary_t *pad = malloc(sizeof(summary_t) * longest), *pre = malloc(sizeof(summary_t) * longest),
              *suf = malloc(sizeof(summary_t) * longest);
    size_t *padc = malloc(sizeof(size_t) * longest), *prec = malloc(sizeof(size_t) * longest),
           *sufc = malloc(sizeof(size_t) * longest);
    int status = NDFO_OK;
    if (!prev || !next || !pc || !nc || !pad || !pre || !suf || !padc || !prec || !sufc) {
        status = NDFO_MEMORY;
        goto done;
    }
    for (size_t i = 0; i < len; ++i) {
        const uint64_t bits = load_bits(input, db, i);
        prev[i] = summary_of(bits, db, omit_nan, ignore_inf);
        pc[i] = (size_t)retained_bits(bits, db, omit_nan, ignore_inf);
    }
    size_t prior = 1;
    for (size_t axis = 0; axis < rank; ++axis) {
        const size_t n = shape[axis], k = size[axis], stride = strides[axis];
        const size_t padded = n + k - 1, lines = len / n, block = n * stride;
        for (size_t line = 0; line < lines; ++line) {
            const size_t base = (line / stride) * block + line % stride;
            for (size_t p = 0; p < padded; ++p) {
                int64_t mapped;
                status = ndfo_map_index((int64_t)p - anchors[axis], n, modes[axis], &mapped);
                if (status != NDFO_OK) goto done;
                if (mapped >= 0) {
                    pad[p] = prev[base + (size_t)mapped * stride];
                    padc[p] = pc[base + (size_t)mapped * stride];
                } else {
                    pad[p] = csum;
                    padc[p] = prior * (size_t)cret;
                }
            }
            for (size_t p = 0; p < padded; ++p) {
                const int start = p % k == 0;
                pre[p] = start ? pad[p] : summary_combine(pre[p - 1], pad[p], maximum);
                prec[p] = start ? padc[p] : prec[p - 1] + padc[p];
            }
            for (size_t p = padded; p-- > 0;) {
                const int start = p == padded - 1 || p % k == k - 1;
                suf[p] = start ? pad[p] : summary_combine(pad[p], suf[p + 1], maximum);
                sufc[p] = start ? padc[p] : sufc[p + 1] + padc[p];
            }
            for (size_t x = 0; x < n; ++x) {
                const size_t end = x + k - 1;
                next[base + x * stride] = summary_combine(suf[x], pre[end], maximum);
                nc[base + x * stride] = x / k == end / k ? prec[end] - prec[x] + padc[x] : sufc[x] + prec[end];
            }
        }
        summary_t *ts = prev;
        prev = next;
        next = ts;
        size_t *tc = pc;
        pc = nc;
        nc = tc;
        prior *= k;
    }
    for (size_t i = 0; i < len; ++i) {
        const int none = prev[i].nan || prev[i].empty || (omit_nan && pc[i] < min_count);
        store_bits(output, db, i, none ? canonical_nan(db) : unkey(prev[i].key, db));
    }
done:
    free(prev);
    free(next);
    free(pc);
    free(nc);
    free(pad);
    free(pre);
    free(suf);
    free(padc);
    free(prec);
    free(sufc);
    return status == NDFO_OK ? NDFO_OK : fail_code(status, err, err_len);
}

/* ------------------------------------------------------------------ */
/* Mesh-grid median filter (MeshBackground::filter, src/background.rs:    */
/* 797-855, with median_of / median_sorted, :940-963)                     */
/* ------------------------------------------------------------------ */

static int compare_f64_total(const void *a, const void *b)
{
    uint64_t x, y;
    memcpy(&x, a, 8);
    memcpy(&y, b, 8);
    x = ndfo_key64(x);
    y = ndfo_key64(y);
    return x < y ? -1 : x > y;
}

/* Median of f64 values, NaN when empty or when any is NaN; an even count
 * takes (a + b) / 2, or a / 2 + b / 2 when that overflows. Reorders. */
static double mesh_median(double *values, size_t n)
{
    if (n == 0) return NAN;
    for (size_t i = 0; i < n; ++i) {
        if (isnan(values[i])) return NAN;
    }
    qsort(values, n, sizeof(double), compare_f64_total);
    if (n % 2 == 1) return values[n / 2];
    const double a = values[n / 2 - 1], b = values[n / 2];
    const double mid = (a + b) / 2.0;
    return isfinite(mid) ? mid : a / 2.0 + b / 2.0;
}

int ndfo_mesh_filter(const double *background, const double *rms, double *out_background, double *out_rms,
                     size_t ndim, const size_t *mesh_shape, const size_t *filter_size, int32_t edge,
                     int32_t threshold_kind, double threshold, char *err, size_t err_len)
{
    if (ndim == 0 || ndim > 64) return fail(NDFO_INVALID, "mesh grids need 1 to 64 dimensions", err, err_len);
    if ((threshold_kind == 1 || threshold_kind == 2) && !isfinite(threshold)) {
        return fail(NDFO_INVALID, "filter threshold must be finite", err, err_len);
    }
    size_t len = 1, volume = 1, all_one = 1;
    for (size_t a = 0; a < ndim; ++a) {
        if (filter_size[a] == 0) return fail(NDFO_INVALID, "filter sizes must be positive", err, err_len);
        if (!checked_mul(len, mesh_shape[a], &len) || !checked_mul(volume, filter_size[a], &volume)) {
            return fail_code(NDFO_OVERFLOW, err, err_len);
        }
        all_one &= filter_size[a] == 1;
    }
    memcpy(out_background, background, len * sizeof(double));
    memcpy(out_rms, rms, len * sizeof(double));
    if (all_one || len == 0) return NDFO_OK;
    double *wb = malloc(sizeof(double) * volume), *wr = malloc(sizeof(double) * volume);
    if (!wb || !wr) {
        free(wb);
        free(wr);
        return fail_code(NDFO_MEMORY, err, err_len);
    }
    size_t index[64], low[64], high[64];
    for (size_t flat = 0; flat < len; ++flat) {
        size_t rest = flat;
        for (size_t a = ndim; a-- > 0;) {
            index[a] = rest % mesh_shape[a];
            rest /= mesh_shape[a];
        }
        for (size_t a = 0; a < ndim; ++a) {
            const size_t k = index[a], n = mesh_shape[a];
            size_t half = filter_size[a] / 2;
            if (edge == 0) { /* shrink: the window stays centred */
                if (half > k) half = k;
                if (half > n - 1 - k) half = n - 1 - k;
                low[a] = k - half;
                high[a] = k + half;
            } else { /* truncate: drop positions outside the grid */
                low[a] = k > half ? k - half : 0;
                high[a] = k + half < n - 1 ? k + half : n - 1;
            }
        }
        size_t count = 1;
        for (size_t a = 0; a < ndim; ++a) count *= high[a] - low[a] + 1;
        for (size_t position = 0; position < count; ++position) {
            size_t r = position, offset = 0, stride = 1;
            for (size_t a = ndim; a-- > 0;) {
                const size_t width = high[a] - low[a] + 1;
                offset += (low[a] + r % width) * stride;
                r /= width;
                stride *= mesh_shape[a];
            }
            wb[position] = background[offset];
            wr[position] = rms[offset];
        }
        const double median = mesh_median(wb, count);
        int replace = 1;
        if (threshold_kind == 1) replace = fabs(median - background[flat]) >= threshold;
        if (threshold_kind == 2) replace = background[flat] > threshold;
        if (replace) {
            out_background[flat] = median;
            out_rms[flat] = mesh_median(wr, count);
        }
    }
    free(wb);
    free(wr);
    return NDFO_OK;
}

/* fixpix(method="median"): src/repair.rs:107-126 (fixpix_median) with
 * positions() at src/repair.rs:14-29 and repair_median at
 * src/filter_nd.rs:174-215. The output starts as a copy of the input; every
 * masked element of a snapshot becomes NaN; each masked output is the
 * direct-method Omit median of its window over that snapshot (all estimates
 * are computed before any is stored, so repaired values never feed another
 * window). */
int ndfo_fixpix_median(int dtype_bytes, const void *input, void *output, const uint8_t *mask, size_t mask_len,
                       size_t rank, const size_t *shape, const size_t *size, const uint8_t *footprint,
                       const ptrdiff_t *origin, const int32_t *modes, double cval, int32_t ignore_inf,
                       size_t min_count, int32_t nthreads, char *err, size_t err_len)
{
    if (dtype_bytes != 4 && dtype_bytes != 8) return fail(NDFO_INVALID, "dtype must be float32 or float64", err, err_len);
    for (size_t a = 0; a < rank; ++a) {
        if (modes[a] < 0 || modes[a] > 4) return fail(NDFO_INVALID, "unsupported boundary mode", err, err_len);
    }
    window_t w;
    int st = window_build(&w, rank, shape, size, footprint, origin, modes, err, err_len);
    if (st != NDFO_OK) return st;
    if (min_count == 0 || min_count > w.count) {
        window_free(&w);
        return fail(NDFO_INVALID, MSG_MIN_COUNT, err, err_len);
    }
    if (mask_len != w.len) {
        window_free(&w);
        return fail_code(NDFO_LENGTH_MISMATCH, err, err_len);
    }
    const size_t bytes = w.len * (size_t)dtype_bytes;
    if (bytes) memcpy(output, input, bytes);
    size_t count = 0;
    for (size_t i = 0; i < w.len; ++i) count += mask[i] != 0;
    if (count == 0) {
        window_free(&w);
        return NDFO_OK;
    }
    size_t *positions = (size_t *)malloc(sizeof(size_t) * count);
    void *snapshot = malloc(bytes);
    if (!positions || !snapshot) {
        free(positions);
        free(snapshot);
        window_free(&w);
        return fail_code(NDFO_MEMORY, err, err_len);
    }
    memcpy(snapshot, input, bytes);
    count = 0;
    for (size_t i = 0; i < w.len; ++i) {
        if (mask[i]) {
            positions[count++] = i;
            store_bits(snapshot, dtype_bytes, i, canonical_nan(dtype_bytes));
        }
    }
    direct_job_t job = {dtype_bytes, snapshot, output, &w, cval_to_bits(cval, dtype_bytes), NDFO_STAT_MEDIAN, 0,
                        1, ignore_inf != 0, min_count, positions, 0.0};
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((size_t)nthreads > count) nthreads = (int)count;
    direct_part_t parts[256];
    const size_t chunk = (count + (size_t)nthreads - 1) / (size_t)nthreads;
    for (int t = 0; t < nthreads; ++t) {
        parts[t].job = &job;
        parts[t].begin = chunk * (size_t)t < count ? chunk * (size_t)t : count;
        parts[t].end = parts[t].begin + chunk < count ? parts[t].begin + chunk : count;
        parts[t].status = NDFO_OK;
    }
    run_parts(direct_worker, parts, sizeof(direct_part_t), nthreads);
    st = NDFO_OK;
    for (int t = 0; t < nthreads; ++t) {
        if (parts[t].status != NDFO_OK) st = parts[t].status;
    }
    free(positions);
    free(snapshot);
    window_free(&w);
    return st == NDFO_OK ? NDFO_OK : fail_code(st, err, err_len);
}
