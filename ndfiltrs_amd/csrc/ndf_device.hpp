// Device-side helpers shared by the selection kernels: total-order keys,
// retention policy, canonical NaN and the even-count midpoint.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ndf {

// Key traits: an unsigned integer whose order is IEEE total order
// (src/median_network.rs:9-37, with the sign bit flipped so the order is
// unsigned). KeyT::SENTINEL (all ones) sorts above every retained value:
// retained values are never NaN, and +inf maps to 0xFF80... / 0xFFF0....
template <typename T>
struct KeyTraits;

template <>
struct KeyTraits<float> {
    using Key = uint32_t;
    static constexpr uint32_t SENTINEL = 0xFFFFFFFFu;
    static constexpr uint32_t CANONICAL_NAN = 0x7FC00000u;
    __device__ __host__ static inline uint32_t bits(float v) { return __builtin_bit_cast(uint32_t, v); }
    __device__ __host__ static inline float from_bits(uint32_t b) { return __builtin_bit_cast(float, b); }
    __device__ __host__ static inline uint32_t key(uint32_t b) { return b ^ ((b >> 31) ? 0xFFFFFFFFu : 0x80000000u); }
    __device__ __host__ static inline uint32_t unkey(uint32_t k) { return k ^ ((k >> 31) ? 0x80000000u : 0xFFFFFFFFu); }
    __device__ __host__ static inline bool is_nan(uint32_t b) { return (b & 0x7FFFFFFFu) > 0x7F800000u; }
    __device__ __host__ static inline bool is_inf(uint32_t b) { return (b & 0x7FFFFFFFu) == 0x7F800000u; }
    __device__ __host__ static inline double to_double(uint32_t b) { return (double)from_bits(b); }
    __device__ __host__ static inline uint32_t from_double(double d)
    {
        return bits((float)d);
    }
};

template <>
struct KeyTraits<double> {
    using Key = uint64_t;
    static constexpr uint64_t SENTINEL = 0xFFFFFFFFFFFFFFFFull;
    static constexpr uint64_t CANONICAL_NAN = 0x7FF8000000000000ull;
    __device__ __host__ static inline uint64_t bits(double v) { return __builtin_bit_cast(uint64_t, v); }
    __device__ __host__ static inline double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }
    __device__ __host__ static inline uint64_t key(uint64_t b)
    {
        return b ^ ((b >> 63) ? 0xFFFFFFFFFFFFFFFFull : 0x8000000000000000ull);
    }
    __device__ __host__ static inline uint64_t unkey(uint64_t k)
    {
        return k ^ ((k >> 63) ? 0x8000000000000000ull : 0xFFFFFFFFFFFFFFFFull);
    }
    __device__ __host__ static inline bool is_nan(uint64_t b)
    {
        return (b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull;
    }
    __device__ __host__ static inline bool is_inf(uint64_t b)
    {
        return (b & 0x7FFFFFFFFFFFFFFFull) == 0x7FF0000000000000ull;
    }
    __device__ __host__ static inline double to_double(uint64_t b) { return from_bits(b); }
    __device__ __host__ static inline uint64_t from_double(double d) { return bits(d); }
};

// Entry of a read-only table (the per-axis maps) at a wave-uniform index,
// read through the constant address space so the compiler issues scalar
// loads that batch under one wait (a generic pointer gives a vector load and
// a wait per entry). Tables are never written while a kernel runs.
template <typename V>
__device__ inline V uniform_entry(const V *table, int64_t index)
{
    return reinterpret_cast<const __attribute__((address_space(4))) V *>(
        reinterpret_cast<uintptr_t>(table))[index];
}

// Retention under InvalidPolicy::Omit (src/reduction.rs:85-101). The
// propagate policy is expressed as Omit{ignore_inf=false, min_count=count}:
// a window with a NaN then has fewer retained samples than taps.
template <typename T>
__device__ __host__ inline typename KeyTraits<T>::Key classify(typename KeyTraits<T>::Key bits, int ignore_inf)
{
    using KT = KeyTraits<T>;
    // Both arms are computed and selected (no branch), so a caller's loads
    // feeding this are not sunk into a conditional block.
    const typename KT::Key key = KT::key(bits);
    const bool omit = KT::is_nan(bits) || (ignore_inf && KT::is_inf(bits));
    return omit ? KT::SENTINEL : key;
}

// Even-count midpoint in f64 (docs/semantics.md:74-85,
// tests/_reference.py:36-47). Compiled with -ffp-contract=off: the
// same-sign branch must round the subtraction, the halving and the addition
// separately.
__device__ __host__ inline double midpoint(double lower, double upper)
{
    const bool sl = __builtin_signbit(lower);
    const bool su = __builtin_signbit(upper);
    if (sl == su && lower == upper) {
        return lower;
    }
    if (lower == 0.0 && upper == 0.0) {
        return 0.0;
    }
    const bool il = __builtin_isinf(lower);
    const bool iu = __builtin_isinf(upper);
    if (il || iu) {
        if (il && iu) {
            return __builtin_nan("");
        }
        return il ? lower : upper;
    }
    if (sl == su) {
        return lower + (upper - lower) / 2.0;
    }
    return (lower + upper) / 2.0;
}

// Output bits for the median of two middle keys: decode, midpoint in f64,
// cast once to T, canonical NaN.
template <typename T>
__device__ __host__ inline typename KeyTraits<T>::Key median_pair_bits(typename KeyTraits<T>::Key lower_key,
                                                                       typename KeyTraits<T>::Key upper_key)
{
    using KT = KeyTraits<T>;
    const double mid = midpoint(KT::to_double(KT::unkey(lower_key)), KT::to_double(KT::unkey(upper_key)));
    if (mid != mid) {
        return KT::CANONICAL_NAN;
    }
    return KT::from_double(mid);
}

// Statistic selector shared by every engine. STAT_PERCENTILE reads the
// fraction q100 = q / 100 of the retained population (the reference's
// linear-interpolated percentile, src/robust.rs:358-367); the minimum and
// the maximum are its q100 = 0 and q100 = 1 cases, whose rank is exact and
// whose fraction is zero, so they return a sample's bits unchanged.
enum Statistic : int32_t { STAT_MEDIAN = 0, STAT_RANK = 1, STAT_PERCENTILE = 2 };

// The order statistics a result reads, for n retained samples: the key at
// zero-based rank `hi`, and with `pair` also the key at rank hi - 1;
// `frac` weights a percentile pair.
struct Pick {
    int64_t hi;
    bool pair;
    double frac;
};

// Fills `pick` and returns true when the window has a result other than the
// canonical NaN: n >= min_count retained samples, and for a fixed rank,
// rank < n.
__device__ __host__ inline bool pick_ranks(int32_t stat, int64_t n, int64_t min_count, int64_t rank_param,
                                           double q100, Pick &pick)
{
    if (n < min_count || n <= 0) return false;
    if (stat == STAT_RANK) {
        if (rank_param >= n) return false;
        pick = Pick{rank_param, false, 0.0};
        return true;
    }
    if (stat == STAT_MEDIAN) {
        pick = Pick{n >> 1, (n & 1) == 0, 0.5};
        return true;
    }
    // rank = q * (n - 1) in f64; lower = floor(rank), fraction = rank - lower
    // (src/robust.rs:363-366). A zero fraction needs the lower sample only;
    // otherwise lower + 1 <= n - 1, because q <= 1 keeps rank <= n - 1.
    const double rank = q100 * (double)(n - 1);
    const double lower = __builtin_floor(rank);
    const double frac = rank - lower;
    const int64_t lo = (int64_t)lower;
    pick = frac == 0.0 ? Pick{lo, false, 0.0} : Pick{lo + 1, true, frac};
    return true;
}

// Rank a walk targets for n retained samples (-1: no result).
__device__ __host__ inline int64_t target_rank(int32_t stat, int64_t n, int64_t min_count, int64_t rank_param,
                                               double q100)
{
    Pick pick;
    return pick_ranks(stat, n, min_count, rank_param, q100, pick) ? pick.hi : -1;
}

// ---- percentile interpolation (src/robust.rs:320-356) ----------------------
// Every step is one IEEE f64 operation (compiled with -ffp-contract=off), in
// the reference's order, so the rounding is the reference's.

// Signed mantissa in [1, 2) and binary exponent of a finite nonzero value
// (src/convolution.rs:203-218).
__device__ __host__ inline void split_binary(double value, double &mantissa, int &exponent)
{
    const uint64_t bits = __builtin_bit_cast(uint64_t, value);
    const int field = (int)((bits >> 52) & 0x7FF);
    const uint64_t fraction = bits & ((1ull << 52) - 1);
    if (field == 0) {
        const int leading = 63 - __builtin_clzll(fraction);
        const double m = (double)fraction / (double)(1ull << leading);
        mantissa = __builtin_copysign(m, value);
        exponent = leading - 1074;
    } else {
        mantissa = __builtin_bit_cast(double, (bits & ((1ull << 63) | ((1ull << 52) - 1))) | (1023ull << 52));
        exponent = field - 1023;
    }
}

// value * 2^exponent without an overflowing or zero scale factor
// (src/convolution.rs:221-240).
__device__ __host__ inline double scale_binary(double value, int exponent)
{
    if (value == 0.0) return value;
    double mantissa;
    int value_exponent;
    split_binary(value, mantissa, value_exponent);
    const int target = value_exponent + exponent;
    if (target > 1023) return __builtin_copysign(__builtin_inf(), value);
    if (target < -1075) return __builtin_copysign(0.0, value);
    if (target >= -1022) return mantissa * __builtin_bit_cast(double, (uint64_t)(target + 1023) << 52);
    const double remaining = __builtin_bit_cast(double, (uint64_t)(target + 1022 + 1023) << 52);
    return (mantissa * 2.2250738585072014e-308) * remaining;
}

// Neumaier accumulation in input order (src/mean.rs:6-26).
struct NeumaierSum {
    double value = 0.0, correction = 0.0;
    __device__ __host__ inline void add(double x)
    {
        const double next = value + x;
        correction += __builtin_fabs(value) >= __builtin_fabs(x) ? (value - next) + x : (x - next) + value;
        value = next;
    }
    __device__ __host__ inline double total() const { return value + correction; }
};

// lower * (1 - fraction) + upper * fraction for finite samples of opposite
// signs: the compensated two-term sum, replayed with binary scaling when a
// product leaves the normal range (src/convolution.rs:243-269, 310-384 with
// normalize = false).
__device__ __host__ inline double weighted_pair(double lower, double upper, double fraction)
{
    const double values[2] = {lower, upper};
    const double weights[2] = {1.0 - fraction, fraction};
    NeumaierSum sum;
    bool fallback = false;
    for (int i = 0; i < 2 && !fallback; ++i) {
        const double product = values[i] * weights[i];
        if (!__builtin_isfinite(product) || (values[i] != 0.0 && __builtin_fabs(product) < 2.2250738585072014e-308)) {
            fallback = true;
        } else {
            sum.add(product);
        }
    }
    const double total = sum.total();
    if (!fallback && __builtin_isfinite(total)) return total;
    // Scaled replay: all terms below 1 after removing a common exponent.
    bool any = false;
    int maximum = 0;
    for (int i = 0; i < 2; ++i) {
        if (values[i] == 0.0) continue;
        double vm, wm;
        int ve, we;
        split_binary(values[i], vm, ve);
        split_binary(weights[i], wm, we);
        maximum = any ? (ve + we > maximum ? ve + we : maximum) : ve + we;
        any = true;
    }
    if (!any) return scale_binary(0.0, 0);
    const int exponent = maximum + 1 + 2;  // count bits of 2 pairs (1) + 2
    NeumaierSum scaled;
    for (int i = 0; i < 2; ++i) {
        if (values[i] == 0.0) continue;
        double vm, wm;
        int ve, we;
        split_binary(values[i], vm, ve);
        split_binary(weights[i], wm, we);
        scaled.add(scale_binary(vm * wm, ve + we - exponent));
    }
    return scale_binary(scaled.total(), exponent);
}

// f64::clamp(value, lower, upper) for lower <= upper.
__device__ __host__ inline double clamp_to(double value, double lower, double upper)
{
    double x = value;
    if (x < lower) x = lower;
    if (x > upper) x = upper;
    return x;
}

// Linear interpolation between two ordered samples (src/robust.rs:320-356).
__device__ __host__ inline double interpolate(double lower, double upper, double fraction)
{
    if (fraction == 0.0 || __builtin_bit_cast(uint64_t, lower) == __builtin_bit_cast(uint64_t, upper)) return lower;
    if (fraction == 1.0) return upper;
    const bool il = __builtin_isinf(lower), iu = __builtin_isinf(upper);
    if (il || iu) {
        if (lower == -__builtin_inf() && upper == __builtin_inf()) return __builtin_nan("");
        return il ? lower : upper;
    }
    if (__builtin_signbit(lower) != __builtin_signbit(upper)) {
        if (!__builtin_isfinite(upper - lower)) {
            const double half_lower = lower / 2.0;
            const double half_upper = upper / 2.0;
            return clamp_to((half_lower + (half_upper - half_lower) * fraction) * 2.0, lower, upper);
        }
        return weighted_pair(lower, upper, fraction);
    }
    const double difference = upper - lower;
    const double result = fraction <= 0.5 ? lower + difference * fraction : upper - difference * (1.0 - fraction);
    return clamp_to(result, lower, upper);
}

// Output bits of a percentile pair: decode, interpolate in f64, cast once to
// T, canonical NaN.
template <typename T>
__device__ __host__ inline typename KeyTraits<T>::Key percentile_pair_bits(typename KeyTraits<T>::Key lower_key,
                                                                           typename KeyTraits<T>::Key upper_key,
                                                                           double fraction)
{
    using KT = KeyTraits<T>;
    const double value = interpolate(KT::to_double(KT::unkey(lower_key)), KT::to_double(KT::unkey(upper_key)), fraction);
    if (value != value) return KT::CANONICAL_NAN;
    return KT::from_double(value);
}

// Output bits of a pair of adjacent order statistics.
template <typename T>
__device__ __host__ inline typename KeyTraits<T>::Key pair_bits(int32_t stat, typename KeyTraits<T>::Key lower_key,
                                                                typename KeyTraits<T>::Key upper_key, double fraction)
{
    return stat == STAT_PERCENTILE ? percentile_pair_bits<T>(lower_key, upper_key, fraction)
                                   : median_pair_bits<T>(lower_key, upper_key);
}

}  // namespace ndf
