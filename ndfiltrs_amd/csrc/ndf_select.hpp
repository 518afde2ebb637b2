// Register-array helpers shared by the register engines (network, stack):
// reading a sorted register array at a per-lane index, and a wave-level
// LDS ordering point.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ndf {

// a when the selector is 0x07060504 (every byte from the first operand),
// b when it is 0x03020100. A byte permute instead of a conditional keeps the
// compiler from merging the two selected operands into one indexed access.
__device__ __forceinline__ uint32_t pick_word(uint32_t a, uint32_t b, uint32_t selector)
{
    return __builtin_amdgcn_perm(a, b, selector);
}
__device__ __forceinline__ uint64_t pick_word(uint64_t a, uint64_t b, uint32_t selector)
{
    const uint32_t lo = __builtin_amdgcn_perm((uint32_t)a, (uint32_t)b, selector);
    const uint32_t hi = __builtin_amdgcn_perm((uint32_t)(a >> 32), (uint32_t)(b >> 32), selector);
    return ((uint64_t)hi << 32) | lo;
}

// k[index] for a per-lane index below N: a binary tree of selects on the
// index bits, lowest first (N - 1 selects instead of the N compares and N
// selects of a linear scan).
template <typename Key, int N>
__device__ __forceinline__ Key select_index(const Key (&k)[N], uint32_t index)
{
    if constexpr (N == 1) {
        return k[0];
    } else {
        constexpr int M = (N + 1) / 2;
        Key t[M];
        const uint32_t selector = (index & 1u) != 0 ? 0x07060504u : 0x03020100u;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            t[j] = 2 * j + 1 < N ? pick_word(k[2 * j + 1 < N ? 2 * j + 1 : 2 * j], k[2 * j], selector) : k[2 * j];
        }
        return select_index<Key, M>(t, index >> 1);
    }
}

// Orders the LDS accesses of one wave around this point: a wave's LDS
// operations execute in program order, so a compiler barrier (and the
// wavefront-scope fences) is all a wave-private exchange needs.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace ndf
