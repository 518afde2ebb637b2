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

// Statistic selector shared by every engine.
enum Statistic : int32_t { STAT_MEDIAN = 0, STAT_RANK = 1 };

}  // namespace ndf
