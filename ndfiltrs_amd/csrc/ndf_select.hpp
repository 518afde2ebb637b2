// Register-array helpers shared by the register engines (network, stack):
// applying a compile-time comparator network, reading a sorted register
// array at a per-lane index (and a median's two ranks from a lower half),
// and a wave-level LDS ordering point.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <utility>

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

// Applies the comparator network P (P::count comparators; comparator i puts
// the smaller key on wire P::lo(i) and the larger on P::hi(i)) to the
// register array w. Comparators are applied in chunks of 512 so no fold
// expression exceeds the compiler's nesting limit.
constexpr int NET_CHUNK = 512;

template <class P, typename Key, int N, int BASE, size_t... I>
__device__ __forceinline__ void apply_chunk(Key (&w)[N], std::index_sequence<I...>)
{
    (
        [&] {
            constexpr int a = P::lo(BASE + (int)I);
            constexpr int b = P::hi(BASE + (int)I);
            static_assert(a < N && b < N, "comparator outside the register array");
            const Key x = w[a], y = w[b];
            w[a] = x < y ? x : y;
            w[b] = x < y ? y : x;
        }(),
        ...);
}

template <class P, typename Key, int N, int BASE = 0>
__device__ __forceinline__ void apply_net(Key (&w)[N])
{
    if constexpr (BASE < P::count) {
        constexpr int len = P::count - BASE < NET_CHUNK ? P::count - BASE : NET_CHUNK;
        apply_chunk<P, Key, N, BASE>(w, std::make_index_sequence<(size_t)len>{});
        apply_net<P, Key, N, BASE + NET_CHUNK>(w);
    }
}

// Reads ranks lo <= hi of the lower half h (both below H; `need` false:
// any ranks). Few omitted samples put both in the top ranks of the half, so
// the wave reads from the top 4, else the top 16, when every lane qualifies,
// and from all H otherwise (wave-uniform branches).
template <typename Key, int H, int F>
__device__ __forceinline__ bool top_ranks(const Key (&h)[H], bool need, int lo, int hi, Key &klo, Key &khi)
{
    const bool fits = !need || lo >= H - F;
    if (__builtin_amdgcn_ballot_w64(!fits) != 0) return false;
    Key t[F];
#pragma unroll
    for (int j = 0; j < F; ++j) t[j] = h[H - F + j];
    khi = select_index<Key, F>(t, need ? (uint32_t)(hi - (H - F)) : 0u);
    klo = select_index<Key, F>(t, need ? (uint32_t)(lo - (H - F)) : 0u);
    return true;
}

template <typename Key, int H>
__device__ __forceinline__ void median_ranks(const Key (&h)[H], bool need, int lo, int hi, Key &klo, Key &khi)
{
    if constexpr (H > 4) {
        if (top_ranks<Key, H, 4>(h, need, lo, hi, klo, khi)) return;
    }
    if constexpr (H > 16) {
        if (top_ranks<Key, H, 16>(h, need, lo, hi, klo, khi)) return;
    }
    khi = select_index<Key, H>(h, need ? (uint32_t)hi : 0u);
    klo = select_index<Key, H>(h, need ? (uint32_t)lo : 0u);
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
