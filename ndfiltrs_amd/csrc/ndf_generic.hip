// Generic engine: one thread per output, any rank, any footprint.
//
// Gather follows Window::visit_mapped (src/window.rs:271-294) through
// precomputed per-axis maps (the maps restate map_index,
// src/boundary.rs:29-71). Selection is exact: the k-th smallest retained key
// is found bit by bit, counting keys below each candidate prefix, so no
// per-thread buffer is needed for any window size.
#include "ndf_device.hpp"
#include "ndf_internal.hpp"

namespace ndf {
namespace {

// R > 0: the array rank is the compile-time R, so the loops (1 to 3 trips)
// unroll fully and the coordinates stay in registers; R = 0: any rank up to MAX_RANK (the
// coordinate array is indexed at run time and lives in scratch memory).
template <typename T, int R = 0>
struct Gather {
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    static constexpr int SLOTS = R > 0 ? R : MAX_RANK;
    const GenericArgs &a;
    int64_t coord[SLOTS];

    __device__ int rank() const { return R > 0 ? R : a.rank; }

    __device__ Gather(const GenericArgs &args, int64_t index) : a(args)
    {
        for (int ax = 0; ax < rank(); ++ax) {
            coord[ax] = index / a.stride[ax];
            index -= coord[ax] * a.stride[ax];
        }
    }

    __device__ Key key(int t) const
    {
        int64_t offset = 0;
        const int32_t *tap = a.taps + (int64_t)t * rank();
        for (int ax = 0; ax < rank(); ++ax) {
            const int64_t m = a.maps[a.map_base[ax] + coord[ax] + tap[ax]];
            if (m < 0) {
                return classify<T>((Key)a.cval_bits, a.ignore_inf);
            }
            offset += m;
        }
        return classify<T>(static_cast<const Key *>(a.input)[offset], a.ignore_inf);
    }
};

// k-th smallest (0-based) retained key; requires k < retained count.
template <typename T, int R>
__device__ typename KeyTraits<T>::Key select_kth(const Gather<T, R> &g, int count, int64_t k)
{
    using Key = typename KeyTraits<T>::Key;
    constexpr int BITS = sizeof(Key) * 8;
    Key answer = 0;
    for (int b = BITS - 1; b >= 0; --b) {
        const Key candidate = answer | (Key(1) << b);
        int64_t below = 0;
        for (int t = 0; t < count; ++t) {
            below += g.key(t) < candidate;
        }
        if (below <= k) {
            answer = candidate;
        }
    }
    return answer;
}

template <typename T, int R>
__global__ void __launch_bounds__(256) k_generic(GenericArgs a)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.len; i += step) {
        int64_t index = i;
        if (a.positions) {
            index = a.positions[i];
            if (index < 0 || index >= a.extent) {
                atomicOr(a.error_flag, 1);
                continue;
            }
        }
        Gather<T, R> g(a, index);
        int64_t n = 0;
        for (int t = 0; t < a.count; ++t) {
            n += g.key(t) != KT::SENTINEL;
        }
        Key out = KT::CANONICAL_NAN;
        Pick pick;
        if (pick_ranks(a.stat, n, a.min_count, a.rank_param, a.q100, pick)) {
            if (!pick.pair) {
                out = KT::unkey(select_kth(g, a.count, pick.hi));
            } else {
                const Key lower = select_kth(g, a.count, pick.hi - 1);
                const Key upper = select_kth(g, a.count, pick.hi);
                out = pair_bits<T>(a.stat, lower, upper, pick.frac);
            }
        }
        static_cast<Key *>(a.output)[index] = out;
    }
}

// Sparse repair, one wave per listed position (windows of at most
// REPAIR_WAVE_MAX_TAPS taps). The wave gathers the window once: with KPL > 0
// tap lane + 64 q goes straight into register q of the lane (wide 2-D
// rectangles: window row q, column lane), and omitted taps keep the
// sentinel key; with KPL = 0 the retained keys are compacted into the
// wave's LDS region (ballot + prefix count). Selection is the same
// bitwise search as k_generic, but each pass counts the keys below the
// candidate across the wave: every lane scans its keys and the counts are
// summed across lanes. An even median takes the (n/2 - 1)-th key and then
// the smallest key above it, unless that key repeats.
constexpr int REPAIR_WAVES = 4;
constexpr int REPAIR_GATHER = 8;

// Sum over the 64 lanes, returned wave-uniform. DPP steps add lane ^ 1 and
// lane ^ 2 (quad_perm [1,0,3,2], [2,3,0,1]), then the other quad of each 8
// (row_half_mirror) and the other half of each 16-lane row (row_mirror);
// the four row sums are read out with v_readlane. VALU lane movement only:
// the reduction sits on every select pass's dependency chain.
__device__ __forceinline__ int wave_sum(int v)
{
    constexpr int ALL = 0xF;
    v += __builtin_amdgcn_update_dpp(0, v, 0xB1, ALL, ALL, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x4E, ALL, ALL, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x141, ALL, ALL, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x140, ALL, ALL, true);
    return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) + __builtin_amdgcn_readlane(v, 32) +
           __builtin_amdgcn_readlane(v, 48);
}

template <typename Key>
__device__ inline Key wave_min(Key v)
{
    for (int d = 32; d >= 1; d >>= 1) {
        const Key o = __shfl_xor(v, d);
        v = o < v ? o : v;
    }
    return v;
}

__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The n retained keys of one wave. KPL > 0: lane l holds keys l + 64 q
// (q < KPL) in registers, with the all-ones sentinel past n (it is above
// every retained key, so it never counts below a candidate); KPL = 0: the
// keys stay in the wave's LDS region and each pass reads them there.
template <typename Key, int KPL>
struct WaveKeys {
    static constexpr Key SENTINEL = static_cast<Key>(~Key(0));
    Key r[KPL > 0 ? KPL : 1];
    const Key *lds = nullptr;
    int n = 0;

    // Wave-wide count of keys below c (uniform). Up to 4 register keys per
    // lane: one vector compare per slot whose lane mask is counted on the
    // scalar unit, with no cross-lane reduction (5x5 repair 0.79 -> 0.62 ms);
    // more register keys: per-lane counts and one wave_sum (the scalar
    // popcount chain measured slower there: 31x31 2.9 -> 3.3 ms); LDS keys:
    // per-lane counts and wave_sum.
    __device__ int count_below(Key c, int lane) const
    {
        int below = 0;
        if constexpr (KPL > 0 && KPL <= 4) {
#pragma unroll
            for (int q = 0; q < KPL; ++q) below += __popcll(__ballot(r[q] < c));
            return below;
        } else if constexpr (KPL > 0) {
#pragma unroll
            for (int q = 0; q < KPL; ++q) below += r[q] < c;
            return wave_sum(below);
        } else {
            for (int j = lane; j < n; j += 64) below += lds[j] < c;
            return wave_sum(below);
        }
    }

    template <typename F>
    __device__ void each(int lane, F &&f) const
    {
        if constexpr (KPL > 0) {
#pragma unroll
            for (int q = 0; q < KPL; ++q) f(r[q]);
        } else {
            for (int j = lane; j < n; j += 64) f(lds[j]);
        }
    }
};

template <typename Key>
__device__ inline Key wave_and(Key v)
{
    for (int d = 32; d >= 1; d >>= 1) v &= __shfl_xor(v, d);
    return v;
}

template <typename Key>
__device__ inline Key wave_or(Key v)
{
    for (int d = 32; d >= 1; d >>= 1) v |= __shfl_xor(v, d);
    return v;
}

// k-th smallest (0-based) of the wave's keys; requires k < n. The bits that
// every retained key shares (AND == OR across the wave) are the answer's
// prefix, so the search starts at the highest bit where the keys differ.
template <typename Key, int KPL>
__device__ Key wave_select(const WaveKeys<Key, KPL> &keys, int64_t k, int lane)
{
    constexpr int BITS = sizeof(Key) * 8;
    constexpr Key SENTINEL = WaveKeys<Key, KPL>::SENTINEL;
    Key all = SENTINEL, any = 0;
    keys.each(lane, [&](Key v) {
        if (v != SENTINEL) {
            all &= v;
            any |= v;
        }
    });
    all = wave_and(all);
    any = wave_or(any);
    const Key differ = all ^ any;
    if (differ == 0) return all;
    const int top = BITS - 1 - (sizeof(Key) == 8 ? __clzll((unsigned long long)differ) : __clz((unsigned)differ));
    Key answer = all & ~((Key(2) << top) - 1);
    for (int b = top; b >= 0; --b) {
        const Key candidate = answer | (Key(1) << b);
        if (keys.count_below(candidate, lane) <= k) answer = candidate;
    }
    return answer;
}

template <typename T, int R, int KPL>
__global__ void __launch_bounds__(64 * REPAIR_WAVES) k_repair_wave(GenericArgs a)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    Key *keys = reinterpret_cast<Key *>(lds_raw) + (size_t)wave * a.count;
    const int64_t step = (int64_t)gridDim.x * REPAIR_WAVES;
    for (int64_t i = (int64_t)blockIdx.x * REPAIR_WAVES + wave; i < a.len; i += step) {
        const int64_t index = a.positions[i];
        if (index < 0 || index >= a.extent) {
            if (lane == 0) atomicOr(a.error_flag, 1);
            continue;
        }
        Gather<T, R> g(a, index);
        WaveKeys<Key, KPL> held;
        int n = 0;
        if constexpr (KPL > 0) {
            int kept = 0;
            if (R == 2 && KPL == 64 && a.rect_kx > 32 && a.rect_kx <= 64 && a.rect_ky <= 64) {
                // Wide 2-D rectangles: register q holds window row q and lane
                // l column l, so the column map entry is read once per lane,
                // the row entry is uniform and each key is one coalesced load.
                const int64_t mx = lane < a.rect_kx ? a.maps[a.map_base[1] + g.coord[1] + lane] : MAP_CVAL;
#pragma unroll
                for (int q = 0; q < KPL; ++q) {
                    Key key = KT::SENTINEL;
                    if (q < a.rect_ky && lane < a.rect_kx) {
                        const int64_t my = a.maps[a.map_base[0] + g.coord[0] + q];
                        key = (my < 0 || mx < 0) ? classify<T>((Key)a.cval_bits, a.ignore_inf)
                                                 : classify<T>(static_cast<const Key *>(a.input)[my + mx], a.ignore_inf);
                    }
                    held.r[q] = key;
                    kept += key != KT::SENTINEL;
                }
            } else {
                // Tap t = lane + 64 q lands in register q: every gather of the
                // window is in flight at once, and omitted taps keep the
                // sentinel key, which no select pass counts.
#pragma unroll
                for (int q = 0; q < KPL; ++q) {
                    const int t = lane + 64 * q;
                    held.r[q] = t < a.count ? g.key(t) : KT::SENTINEL;
                    kept += held.r[q] != KT::SENTINEL;
                }
            }
            n = wave_sum(kept);
        } else {
            // REPAIR_GATHER taps per lane in flight (each a tap -> map ->
            // input chain of dependent loads), compacted into LDS.
            for (int t0 = 0; t0 < a.count; t0 += 64 * REPAIR_GATHER) {
                Key gathered[REPAIR_GATHER];
#pragma unroll
                for (int u = 0; u < REPAIR_GATHER; ++u) {
                    const int t = t0 + u * 64 + lane;
                    gathered[u] = t < a.count ? g.key(t) : KT::SENTINEL;
                }
#pragma unroll
                for (int u = 0; u < REPAIR_GATHER; ++u) {
                    const bool keep = gathered[u] != KT::SENTINEL;
                    const uint64_t ballot = __ballot(keep);
                    if (keep) keys[n + __popcll(ballot & ((1ull << lane) - 1))] = gathered[u];
                    n += __popcll(ballot);
                }
            }
            wave_lds_sync();
            held.lds = keys;
            held.n = n;
        }
        Key out = KT::CANONICAL_NAN;
        Pick pick;
        if (pick_ranks(a.stat, n, a.min_count, a.rank_param, a.q100, pick)) {
            if (!pick.pair) {
                out = KT::unkey(wave_select(held, (int)pick.hi, lane));
            } else {
                const Key lower = wave_select(held, (int)pick.hi - 1, lane);
                // Keys <= lower are those below lower + 1 (lower is a
                // retained key, so lower + 1 does not wrap); when more than
                // hi of them exist, the key at rank hi equals lower.
                const int at_most = held.count_below(lower + 1, lane);
                Key above = KT::SENTINEL;
                held.each(lane, [&](Key v) {
                    if (v > lower && v < above) above = v;
                });
                above = wave_min(above);
                out = pair_bits<T>(a.stat, lower, at_most > pick.hi ? lower : above, pick.frac);
            }
        }
        if (lane == 0) static_cast<Key *>(a.output)[index] = out;
        if constexpr (KPL == 0) wave_lds_sync();
    }
}

// Mask bytes a thread reads per step (one 16-byte load for full chunks).
constexpr int MASK_CHUNK = 16;

__device__ inline int chunk_flags(const uint8_t *mask, int64_t len, int64_t chunk, uint32_t &bits)
{
    const int64_t begin = chunk * MASK_CHUNK;
    bits = 0;
    if (begin + MASK_CHUNK <= len) {
        const uint4 v = *reinterpret_cast<const uint4 *>(mask + begin);
        const uint32_t words[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int w = 0; w < 4; ++w) {
#pragma unroll
            for (int b = 0; b < 4; ++b) bits |= (uint32_t)(((words[w] >> (8 * b)) & 0xFFu) != 0) << (4 * w + b);
        }
    } else {
        for (int b = 0; begin + b < len; ++b) bits |= (uint32_t)(mask[begin + b] != 0) << b;
    }
    return __popc(bits);
}

// Number of nonzero mask bytes (one atomic per block).
__global__ void __launch_bounds__(256) k_mask_count(const uint8_t *mask, int64_t len, unsigned long long *count)
{
    __shared__ unsigned long long partial[4];
    const int64_t chunks = (len + MASK_CHUNK - 1) / MASK_CHUNK;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    unsigned long long n = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < chunks; c += step) {
        uint32_t bits;
        n += (unsigned long long)chunk_flags(mask, len, c, bits);
    }
    for (int off = 32; off > 0; off >>= 1) n += __shfl_down(n, off);
    if ((threadIdx.x & 63) == 0) partial[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(count, partial[0] + partial[1] + partial[2] + partial[3]);
}

// Flat indices of the nonzero mask bytes (positions() of src/repair.rs:14-
// 29, in an order the repair does not depend on): each wave reserves room
// for its chunks with one atomic, each thread writes its chunk's indices.
__global__ void __launch_bounds__(256) k_mask_positions(const uint8_t *mask, int64_t len, int64_t *positions,
                                                        unsigned long long *cursor)
{
    const int64_t chunks = (len + MASK_CHUNK - 1) / MASK_CHUNK;
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    for (int64_t c0 = wave * 64; c0 < chunks; c0 += waves * 64) {
        const int64_t c = c0 + lane;
        uint32_t bits = 0;
        const int n = c < chunks ? chunk_flags(mask, len, c, bits) : 0;
        int prefix = n;
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(prefix, off);
            if (lane >= off) prefix += v;
        }
        const int total = __shfl(prefix, 63);
        unsigned long long base = 0;
        if (lane == 63 && total > 0) base = atomicAdd(cursor, (unsigned long long)total);
        base = __shfl(base, 63);
        int64_t at = (int64_t)base + prefix - n;
        while (bits) {
            const int b = __builtin_ctz(bits);
            bits &= bits - 1;
            positions[at++] = c * MASK_CHUNK + b;
        }
    }
}

template <typename Key>
__global__ void __launch_bounds__(256) k_fill(Key *out, int64_t len, Key bits)
{
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += step) {
        out[i] = bits;
    }
}

template <typename Key>
__global__ void __launch_bounds__(256)
    k_scatter_fill(Key *out, const int64_t *positions, int64_t n, int64_t extent, Key bits, int32_t *error_flag)
{
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const int64_t p = positions[i];
        if (p < 0 || p >= extent) {
            atomicOr(error_flag, 1);
        } else {
            out[p] = bits;
        }
    }
}

template <typename Key>
__global__ void __launch_bounds__(256) k_scatter_copy(Key *dst, const Key *src, const int64_t *positions, int64_t n,
                                                      int64_t extent, int32_t *error_flag)
{
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const int64_t p = positions[i];
        if (p < 0 || p >= extent) {
            atomicOr(error_flag, 1);
        } else {
            dst[p] = src[p];
        }
    }
}

int64_t grid_for(int64_t len)
{
    const int64_t blocks = (len + 255) / 256;
    return blocks < 65536 ? (blocks < 1 ? 1 : blocks) : 65536;
}

}  // namespace

hipError_t launch_generic(int dtype_bytes, const GenericArgs &args, hipStream_t stream)
{
    const dim3 grid((unsigned)grid_for(args.len));
    const int r = args.rank;
    void (*fn)(GenericArgs) = nullptr;
    if (dtype_bytes == 4) {
        fn = r == 1 ? k_generic<float, 1> : r == 2 ? k_generic<float, 2> : r == 3 ? k_generic<float, 3>
                                                                          : k_generic<float, 0>;
    } else {
        fn = r == 1 ? k_generic<double, 1> : r == 2 ? k_generic<double, 2> : r == 3 ? k_generic<double, 3>
                                                                           : k_generic<double, 0>;
    }
    hipLaunchKernelGGL(fn, grid, dim3(256), 0, stream, args);
    return hipGetLastError();
}

template <typename T, int R>
void (*repair_kernel_r(int kpl))(GenericArgs)
{
    return kpl == 1 ? k_repair_wave<T, R, 1> : kpl == 4 ? k_repair_wave<T, R, 4>
           : kpl == 16 ? k_repair_wave<T, R, 16> : kpl == 64 ? k_repair_wave<T, R, 64> : k_repair_wave<T, R, 0>;
}

template <typename T>
void (*repair_kernel(int rank, int kpl))(GenericArgs)
{
    return rank == 1 ? repair_kernel_r<T, 1>(kpl) : rank == 2 ? repair_kernel_r<T, 2>(kpl)
           : rank == 3 ? repair_kernel_r<T, 3>(kpl) : repair_kernel_r<T, 0>(kpl);
}

size_t repair_wave_lds_bytes(int dtype_bytes, int count) { return (size_t)REPAIR_WAVES * count * dtype_bytes; }

hipError_t launch_repair_wave(int dtype_bytes, const GenericArgs &args, hipStream_t stream)
{
    int64_t blocks = (args.len + REPAIR_WAVES - 1) / REPAIR_WAVES;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    // Keys per lane in registers: 1 (<= 64 taps), 4 (<= 256), 16 (<= 1024),
    // 64 (<= 4096 = REPAIR_WAVE_MAX_TAPS), else 0 (keys compacted into LDS
    // and read there on every pass; only that form uses LDS).
    const int kpl = args.count <= 64 ? 1 : args.count <= 256 ? 4 : args.count <= 1024 ? 16
                    : args.count <= 4096 ? 64 : 0;
    const size_t lds = kpl == 0 ? repair_wave_lds_bytes(dtype_bytes, args.count) : 0;
    void (*fn)(GenericArgs) = dtype_bytes == 4 ? repair_kernel<float>(args.rank, kpl)
                                               : repair_kernel<double>(args.rank, kpl);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(64 * REPAIR_WAVES), lds, stream, args);
    return hipGetLastError();
}

hipError_t launch_mask_count(const uint8_t *mask, int64_t len, unsigned long long *count, hipStream_t stream)
{
    const int64_t chunks = (len + MASK_CHUNK - 1) / MASK_CHUNK;
    hipLaunchKernelGGL(k_mask_count, dim3((unsigned)grid_for(chunks)), dim3(256), 0, stream, mask, len, count);
    return hipGetLastError();
}

hipError_t launch_mask_positions(const uint8_t *mask, int64_t len, int64_t *positions, unsigned long long *cursor,
                                 hipStream_t stream)
{
    const int64_t chunks = (len + MASK_CHUNK - 1) / MASK_CHUNK;
    hipLaunchKernelGGL(k_mask_positions, dim3((unsigned)grid_for(chunks)), dim3(256), 0, stream, mask, len,
                       positions, cursor);
    return hipGetLastError();
}

hipError_t launch_fill(int dtype_bytes, void *output, int64_t len, uint64_t bits, hipStream_t stream)
{
    const dim3 grid((unsigned)grid_for(len));
    if (dtype_bytes == 4) {
        hipLaunchKernelGGL(k_fill<uint32_t>, grid, dim3(256), 0, stream, static_cast<uint32_t *>(output), len,
                           (uint32_t)bits);
    } else {
        hipLaunchKernelGGL(k_fill<uint64_t>, grid, dim3(256), 0, stream, static_cast<uint64_t *>(output), len,
                           bits);
    }
    return hipGetLastError();
}

hipError_t launch_scatter_fill(int dtype_bytes, void *output, const int64_t *positions, int64_t n, int64_t extent,
                               uint64_t bits, int32_t *error_flag, hipStream_t stream)
{
    const dim3 grid((unsigned)grid_for(n));
    if (dtype_bytes == 4) {
        hipLaunchKernelGGL(k_scatter_fill<uint32_t>, grid, dim3(256), 0, stream, static_cast<uint32_t *>(output),
                           positions, n, extent, (uint32_t)bits, error_flag);
    } else {
        hipLaunchKernelGGL(k_scatter_fill<uint64_t>, grid, dim3(256), 0, stream, static_cast<uint64_t *>(output),
                           positions, n, extent, bits, error_flag);
    }
    return hipGetLastError();
}

hipError_t launch_scatter_copy(int dtype_bytes, void *dst, const void *src, const int64_t *positions, int64_t n,
                               int64_t extent, int32_t *error_flag, hipStream_t stream)
{
    const dim3 grid((unsigned)grid_for(n));
    if (dtype_bytes == 4) {
        hipLaunchKernelGGL(k_scatter_copy<uint32_t>, grid, dim3(256), 0, stream, static_cast<uint32_t *>(dst),
                           static_cast<const uint32_t *>(src), positions, n, extent, error_flag);
    } else {
        hipLaunchKernelGGL(k_scatter_copy<uint64_t>, grid, dim3(256), 0, stream, static_cast<uint64_t *>(dst),
                           static_cast<const uint64_t *>(src), positions, n, extent, error_flag);
    }
    return hipGetLastError();
}

}  // namespace ndf
