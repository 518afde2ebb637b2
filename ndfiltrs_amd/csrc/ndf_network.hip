// Network engine: small windows (up to 160 taps), one output per thread, the
// whole window sorted in registers by a compile-time sorting network.
//
// The reference's median_network (src/median_network.rs:65-209) applies a
// Batcher network column-wise over many outputs on the CPU. Here the same
// idea is MI355X-shaped:
//   * a workgroup (4 waves) stages the support of a 64-column x TILE_ROWS
//     output tile (all window planes) into LDS through the boundary maps,
//     classifying each sample once (omitted samples become the sentinel key,
//     constant-mode samples the cval key);
//   * each thread gathers its window's keys from LDS into registers (taps
//     beyond the window are sentinels) and sorts them with a network of
//     Batcher odd-even merges (a balanced merge tree of single wires,
//     generated with its 0-1 proof by gen_stack_networks.py);
//   * sentinels sort last, so the retained count n is the number of
//     non-sentinel keys and the statistic is read at a data-dependent index
//     with a binary tree of selects. A median never reads past rank LIVE/2,
//     so its kernels keep only the comparators the lower half depends on.
// The result is the exact sample (or the f64 midpoint of two), bit-identical
// to the slide and generic engines.
#include <utility>

#include "ndf_device.hpp"
#include "ndf_internal.hpp"
#include "ndf_select.hpp"
#include "ndf_stack_networks.inc"

namespace ndf {
namespace {

constexpr int NETWORK_THREADS = 256;

// The window's sorting network (ndf_stack_networks.inc): a balanced merge
// tree of single wires; the median applies only the comparators that ranks
// 0 .. LIVE/2 depend on (omitted samples carry the all-ones key and sort
// last, so a median's ranks never pass LIVE/2).
template <int LIVE>
struct SortFull {
    static constexpr int count = stack_net::Sort<LIVE>::full_count;
    static constexpr int lo(int i) { return stack_net::Sort<LIVE>::full[i][0]; }
    static constexpr int hi(int i) { return stack_net::Sort<LIVE>::full[i][1]; }
};
template <int LIVE>
struct SortHalf {
    static constexpr int count = stack_net::Sort<LIVE>::half_count;
    static constexpr int lo(int i) { return stack_net::Sort<LIVE>::half[i][0]; }
    static constexpr int hi(int i) { return stack_net::Sort<LIVE>::half[i][1]; }
};

// Which statistics a kernel instance computes: the median only, the others
// only, or either (chosen per image at run time).
enum StatSet { STATS_MEDIAN, STATS_OTHER, STATS_ANY };

// Output bits of a median from the sorted lower half h (H ranks), n retained.
template <typename T, int H>
__device__ __forceinline__ typename KeyTraits<T>::Key median_from_half(const NetworkArgs &a,
                                                                       const typename KeyTraits<T>::Key (&h)[H], int n)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    const bool ok = n >= a.min_count && n > 0;
    const int hi = n >> 1;
    const bool pair = (n & 1) == 0;
    const int lo = pair ? hi - 1 : hi;
    Key klo, khi;
    median_ranks<Key, H>(h, ok, lo, hi, klo, khi);
    return ok ? (pair ? median_pair_bits<T>(klo, khi) : KT::unkey(khi)) : KT::CANONICAL_NAN;
}

// Output bits of a rank or percentile from the sorted keys f, n retained.
template <typename T, int W>
__device__ __forceinline__ typename KeyTraits<T>::Key statistic_from_sorted(const NetworkArgs &a,
                                                                            const typename KeyTraits<T>::Key (&f)[W],
                                                                            int n)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    Pick pick;
    const bool ok = pick_ranks(a.stat, n, a.min_count, a.rank_param, a.q100, pick);
    const uint32_t hi_index = ok ? (uint32_t)pick.hi : 0u;
    const Key hi = select_index<Key, W>(f, hi_index);
    Key lo = hi;
    if (a.stat != STAT_RANK) lo = select_index<Key, W>(f, ok && pick.pair ? hi_index - 1u : hi_index);
    return ok ? (pick.pair ? pair_bits<T>(a.stat, lo, hi, pick.frac) : KT::unkey(hi)) : KT::CANONICAL_NAN;
}

// The output bits of one window (k: its LIVE keys, n of them retained).
template <typename T, int LIVE, int STATS>
__device__ __forceinline__ typename KeyTraits<T>::Key window_result(const NetworkArgs &a,
                                                                     typename KeyTraits<T>::Key (&k)[LIVE], int n)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    using Sort = stack_net::Sort<LIVE>;
    Key out = KT::CANONICAL_NAN;
    if (STATS == STATS_MEDIAN || (STATS == STATS_ANY && a.stat == STAT_MEDIAN)) {
        if constexpr (STATS != STATS_OTHER) {
            apply_net<SortHalf<LIVE>>(k);
            Key h[Sort::H];
#pragma unroll
            for (int i = 0; i < Sort::H; ++i) h[i] = k[Sort::half_out[i]];
            out = median_from_half<T, Sort::H>(a, h, n);
        }
    } else if constexpr (STATS != STATS_MEDIAN) {
        apply_net<SortFull<LIVE>>(k);
        Key f[LIVE];
#pragma unroll
        for (int i = 0; i < LIVE; ++i) f[i] = k[Sort::full_out[i]];
        out = statistic_from_sorted<T, LIVE>(a, f, n);
    }
    return out;
}

// Stages one tile's support (classified keys) into LDS: (z0, y0, x0) is the
// tile's first output.
template <typename T>
__device__ __forceinline__ void stage_network_tile(const NetworkArgs &a, int tile,
                                                   typename KeyTraits<T>::Key *support, int64_t &z0, int64_t &y0,
                                                   int64_t &x0)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const Key *input = static_cast<const Key *>(a.input);
    const Key cval_key = classify<T>((Key)a.cval_bits, a.ignore_inf);
    int rest = tile;
    const int tx = rest % a.tiles_x;
    rest /= a.tiles_x;
    const int tyi = rest % a.tiles_y;
    z0 = rest / a.tiles_y;
    x0 = (int64_t)tx * 64;
    y0 = (int64_t)tyi * a.ty;
    __syncthreads();  // the previous tile's reads of `support` are done
    for (int row = wave; row < a.rows; row += 4) {
        const int z = row / a.sy;
        const int y = row - z * a.sy;
        const int64_t mz = uniform_entry(a.mapz, z0 + z);
        const int64_t my = uniform_entry(a.mapy, y0 + y);
        for (int x = lane; x < a.sx; x += 64) {
            const int64_t mx = a.mapx[x0 + x];
            const bool pad = mz == MAP_PAD || my == MAP_PAD || mx == MAP_PAD;
            const bool cval = !pad && (mz < 0 || my < 0 || mx < 0);
            const Key raw = input[pad || cval ? 0 : mz + my + mx];
            support[row * a.sx + x] = pad ? KT::SENTINEL : (cval ? cval_key : classify<T>(raw, a.ignore_inf));
        }
    }
    // Taps past the window's count read offset a.s: every window base
    // (r * sx + lane, r < ty, lane < 64) lands in this sentinel run.
    for (int i = tid; i < a.pad; i += NETWORK_THREADS) {
        support[a.s + i] = KT::SENTINEL;
    }
    __syncthreads();
}

// One tile: stage the support (classified keys) into LDS, then one output
// per thread and row step. `taps` holds the window's support offsets.
template <typename T, int LIVE, int STATS>
__device__ __forceinline__ void network_tile(const NetworkArgs &a, int tile,
                                             typename KeyTraits<T>::Key *support, const uint32_t *taps)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    Key *output = static_cast<Key *>(a.output);
    int64_t z0, y0, x0;
    stage_network_tile<T>(a, tile, support, z0, y0, x0);
    const int64_t x = x0 + lane;
    for (int r = wave; r < a.ty; r += 4) {
        const int64_t y = y0 + r;
        if (y >= a.ny || x >= a.nx) {
            continue;
        }
        // Every tap table holds LIVE entries (those past the window's count
        // point into the sentinel run), so the loads are unconditional.
        Key k[LIVE];
        const Key *window = support + r * a.sx + lane;
        int n = 0;
#pragma unroll
        for (int i = 0; i < LIVE; ++i) {
            k[i] = window[taps[i]];
            n += k[i] != KT::SENTINEL;
        }
        const Key out = window_result<T, LIVE, STATS>(a, k, n);
        output[(z0 * a.ny + y) * a.nx + x] = out;
    }
}

// Parts of the two-row networks (PairSort<C, E> in ndf_stack_networks.inc).
enum PairNet { PAIR_CORE_FULL, PAIR_CORE_HALF, PAIR_EXTRA_FULL, PAIR_EXTRA_HALF, PAIR_FIN_FULL, PAIR_FIN_HALF };
template <int C, int E, int PART>
struct PairPart {
    using PS = stack_net::PairSort<C, E>;
    static constexpr int count = PART == PAIR_CORE_FULL    ? PS::core_full_count
                                 : PART == PAIR_CORE_HALF  ? PS::core_half_count
                                 : PART == PAIR_EXTRA_FULL ? PS::extra_full_count
                                 : PART == PAIR_EXTRA_HALF ? PS::extra_half_count
                                 : PART == PAIR_FIN_FULL   ? PS::fin_full_count
                                                           : PS::fin_half_count;
    static constexpr int wire(int i, int side)
    {
        return PART == PAIR_CORE_FULL    ? PS::core_full[i][side]
               : PART == PAIR_CORE_HALF  ? PS::core_half[i][side]
               : PART == PAIR_EXTRA_FULL ? PS::extra_full[i][side]
               : PART == PAIR_EXTRA_HALF ? PS::extra_half[i][side]
               : PART == PAIR_FIN_FULL   ? PS::fin_full[i][side]
                                         : PS::fin_half[i][side];
    }
    static constexpr int lo(int i) { return wire(i, 0); }
    static constexpr int hi(int i) { return wire(i, 1); }
};

// One tile with two outputs per thread and step: the windows at rows r and
// r + 1 share the C core taps (taps[0 .. C)); each has E taps of its own
// (taps[C .. C + E) for row r, taps[C + E .. C + 2E) for row r + 1, all
// relative to row r; entries past the real counts read the sentinel run).
// The core and the two extras are sorted once each, then each window merges
// the sorted core with its sorted extra.
template <typename T, int C, int E, int STATS>
__device__ __forceinline__ void network_pair_tile(const NetworkArgs &a, int tile,
                                                  typename KeyTraits<T>::Key *support, const uint32_t *taps)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    using PS = stack_net::PairSort<C, E>;
    constexpr int W = C + E;
    constexpr bool MEDIAN = STATS == STATS_MEDIAN;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    Key *output = static_cast<Key *>(a.output);
    int64_t z0, y0, x0;
    stage_network_tile<T>(a, tile, support, z0, y0, x0);
    const int64_t x = x0 + lane;
    for (int r = 2 * wave; r < a.ty; r += 8) {
        const int64_t y = y0 + r;
        if (y >= a.ny || x >= a.nx) {
            continue;
        }
        const Key *window = support + r * a.sx + lane;
        Key core[C];
        int ncore = 0;
#pragma unroll
        for (int i = 0; i < C; ++i) {
            core[i] = window[taps[i]];
            ncore += core[i] != KT::SENTINEL;
        }
        Key extra[2][E];
        int nextra[2] = {0, 0};
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int i = 0; i < E; ++i) {
                extra[j][i] = window[taps[C + j * E + i]];
                nextra[j] += extra[j][i] != KT::SENTINEL;
            }
        }
        apply_net<PairPart<C, E, MEDIAN ? PAIR_CORE_HALF : PAIR_CORE_FULL>>(core);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            apply_net<PairPart<C, E, MEDIAN ? PAIR_EXTRA_HALF : PAIR_EXTRA_FULL>>(extra[j]);
            Key u[W];
#pragma unroll
            for (int i = 0; i < C; ++i) u[i] = core[PS::core_out[i]];
#pragma unroll
            for (int i = 0; i < E; ++i) u[C + i] = extra[j][PS::extra_out[i]];
            const int n = ncore + nextra[j];
            Key out;
            if constexpr (MEDIAN) {
                apply_net<PairPart<C, E, PAIR_FIN_HALF>>(u);
                Key h[PS::H];
#pragma unroll
                for (int i = 0; i < PS::H; ++i) h[i] = u[PS::fin_half_out[i]];
                out = median_from_half<T, PS::H>(a, h, n);
            } else {
                apply_net<PairPart<C, E, PAIR_FIN_FULL>>(u);
                Key f[W];
#pragma unroll
                for (int i = 0; i < W; ++i) f[i] = u[PS::fin_full_out[i]];
                out = statistic_from_sorted<T, W>(a, f, n);
            }
            if (y + j < a.ny) output[(z0 * a.ny + y + j) * a.nx + x] = out;
        }
    }
}

template <typename T, int C, int E, int STATS>
__global__ void __launch_bounds__(NETWORK_THREADS) k_network_pair(NetworkArgs a)
{
    using Key = typename KeyTraits<T>::Key;
    extern __shared__ __align__(16) unsigned char lds[];
    Key *support = reinterpret_cast<Key *>(lds);
    uint32_t *taps = reinterpret_cast<uint32_t *>(support + a.s + a.pad);
    for (int t = threadIdx.x; t < C + 2 * E; t += NETWORK_THREADS) {
        taps[t] = a.taps[t];
    }
    const int tiles = a.tiles_x * a.tiles_y * (int)a.nz;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        network_pair_tile<T, C, E, STATS>(a, tile, support, taps);
    }
}

template <typename T, int C, int E, int STATS>
hipError_t launch_pair_stats(const NetworkArgs &args, size_t lds_bytes, hipStream_t stream)
{
    const void *fn = reinterpret_cast<const void *>(&k_network_pair<T, C, E, STATS>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) {
        return e;
    }
    const int64_t tiles = (int64_t)args.tiles_x * args.tiles_y * args.nz;
    const int64_t blocks = tiles < ((int64_t)1 << 24) ? tiles : ((int64_t)1 << 24);
    hipLaunchKernelGGL((k_network_pair<T, C, E, STATS>), dim3((unsigned)(blocks < 1 ? 1 : blocks)),
                       dim3(NETWORK_THREADS), lds_bytes, stream, args);
    return hipGetLastError();
}

template <int C, int E>
hipError_t launch_pair(const NetworkArgs &args, size_t lds_bytes, hipStream_t stream)
{
    return args.stat == STAT_MEDIAN ? launch_pair_stats<float, C, E, STATS_MEDIAN>(args, lds_bytes, stream)
                                    : launch_pair_stats<float, C, E, STATS_OTHER>(args, lds_bytes, stream);
}

// One image: workgroups stride over the tiles.
template <typename T, int LIVE, int STATS>
__global__ void __launch_bounds__(NETWORK_THREADS) k_network(NetworkArgs a)
{
    using Key = typename KeyTraits<T>::Key;
    extern __shared__ __align__(16) unsigned char lds[];
    Key *support = reinterpret_cast<Key *>(lds);
    uint32_t *taps = reinterpret_cast<uint32_t *>(support + a.s + a.pad);
    for (int t = threadIdx.x; t < LIVE; t += NETWORK_THREADS) {
        taps[t] = a.taps[t];
    }
    const int tiles = a.tiles_x * a.tiles_y * (int)a.nz;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        network_tile<T, LIVE, STATS>(a, tile, support, taps);
    }
}

// A batch of images with windows of at most LIVE taps: (image, tile) pairs
// from a work queue; taps are read from global memory (uniform loads).
template <typename T, int LIVE>
__global__ void __launch_bounds__(NETWORK_THREADS) k_network_batch(const NetworkArgs *images, const int64_t *offsets,
                                                                   const int32_t *work, int nwork, int32_t *counter,
                                                                   const void *input, void *output)
{
    using Key = typename KeyTraits<T>::Key;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int next_shared;
    Key *support = reinterpret_cast<Key *>(lds);
    for (;;) {
        if (threadIdx.x == 0) {
            next_shared = atomicAdd(counter, 1);
        }
        __syncthreads();
        // Uniform across the workgroup: keeps the descriptor loads scalar.
        const int w = __builtin_amdgcn_readfirstlane(next_shared);
        __syncthreads();
        if (w >= nwork) {
            break;
        }
        const int image = work[2 * w];
        NetworkArgs a = images[image];
        a.input = static_cast<const Key *>(input) + offsets[image];
        a.output = static_cast<Key *>(output) + offsets[image];
        network_tile<T, LIVE, STATS_ANY>(a, work[2 * w + 1], support, a.taps);
    }
}

template <typename T, int LIVE>
hipError_t launch_batch_live(const NetworkArgs *images, const int64_t *offsets, const int32_t *work, int nwork,
                             int32_t *counter, const void *input, void *output, size_t lds_bytes, int grid,
                             hipStream_t stream)
{
    hipLaunchKernelGGL((k_network_batch<T, LIVE>), dim3((unsigned)grid), dim3(NETWORK_THREADS), lds_bytes, stream,
                       images, offsets, work, nwork, counter, input, output);
    return hipGetLastError();
}

template <typename T, int LIVE>
hipError_t batch_grid_live(size_t lds_bytes, int *grid)
{
    const void *fn = reinterpret_cast<const void *>(&k_network_batch<T, LIVE>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) {
        return e;
    }
    int per_cu = 0, device = 0, cus = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NETWORK_THREADS, lds_bytes)) != hipSuccess ||
        (e = hipGetDevice(&device)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) {
        return e;
    }
    *grid = (per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
    return hipSuccess;
}

template <typename T, int LIVE, int STATS>
hipError_t launch_stats(const NetworkArgs &args, size_t lds_bytes, hipStream_t stream)
{
    const void *fn = reinterpret_cast<const void *>(&k_network<T, LIVE, STATS>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) {
        return e;
    }
    // One workgroup per tile (tiles cost the same, and the dispatcher keeps
    // every CU at the occupancy registers and LDS allow); the tile loop
    // strides only past 2^24 tiles.
    const int64_t tiles = (int64_t)args.tiles_x * args.tiles_y * args.nz;
    const int64_t blocks = tiles < ((int64_t)1 << 24) ? tiles : ((int64_t)1 << 24);
    hipLaunchKernelGGL((k_network<T, LIVE, STATS>), dim3((unsigned)(blocks < 1 ? 1 : blocks)),
                       dim3(NETWORK_THREADS), lds_bytes, stream, args);
    return hipGetLastError();
}

template <typename T, int LIVE>
hipError_t launch_live(const NetworkArgs &args, size_t lds_bytes, hipStream_t stream)
{
    return args.stat == STAT_MEDIAN ? launch_stats<T, LIVE, STATS_MEDIAN>(args, lds_bytes, stream)
                                    : launch_stats<T, LIVE, STATS_OTHER>(args, lds_bytes, stream);
}

}  // namespace

int network_live(int dtype_bytes, int count)
{
    const int limit = dtype_bytes == 4 ? 160 : 64;
    if (count > limit) {
        return 0;
    }
    for (int live : {9, 16, 25, 32, 49, 64, 72, 96, 128, 160}) {
        if (count <= live) {
            return live;
        }
    }
    return 0;
}

namespace {

struct PairCost {
    int c, e, half, full;
};

// Comparators per output of each compiled two-row class (median / full).
template <int C, int E>
constexpr PairCost pair_cost()
{
    using PS = stack_net::PairSort<C, E>;
    return PairCost{C, E, (PS::core_half_count + 2 * PS::extra_half_count + 2 * PS::fin_half_count) / 2,
                    (PS::core_full_count + 2 * PS::extra_full_count + 2 * PS::fin_full_count) / 2};
}

constexpr PairCost PAIR_COSTS[] = {pair_cost<16, 8>(),  pair_cost<32, 12>(), pair_cost<48, 20>(),
                                   pair_cost<64, 24>(), pair_cost<96, 32>(), pair_cost<128, 32>()};

template <int LIVE>
constexpr int sort_cost(bool median)
{
    return median ? stack_net::Sort<LIVE>::half_count : stack_net::Sort<LIVE>::full_count;
}

int single_cost(int live, bool median)
{
    switch (live) {
    case 9: return sort_cost<9>(median);
    case 16: return sort_cost<16>(median);
    case 25: return sort_cost<25>(median);
    case 32: return sort_cost<32>(median);
    case 49: return sort_cost<49>(median);
    case 64: return sort_cost<64>(median);
    case 72: return sort_cost<72>(median);
    case 96: return sort_cost<96>(median);
    case 128: return sort_cost<128>(median);
    case 160: return sort_cost<160>(median);
    default: return 1 << 30;
    }
}

}  // namespace

bool network_pair_class(int dtype_bytes, int core, int extra, int live, bool median, int *c, int *e)
{
    if (dtype_bytes != 4) return false;
    for (const PairCost &k : PAIR_COSTS) {
        if (core <= k.c && extra <= k.e) {
            // Fewer comparators per output than one output per step, with a
            // tenth to spare for the second extra's loads and merge set-up.
            const int pair = median ? k.half : k.full;
            if (10 * pair < 9 * single_cost(live, median)) {
                *c = k.c;
                *e = k.e;
                return true;
            }
            return false;
        }
    }
    return false;
}

size_t network_lds_bytes(int dtype_bytes, int s, int pad, int live)
{
    return ((size_t)s + (size_t)pad) * (size_t)dtype_bytes + (size_t)live * 4;
}

hipError_t network_batch_grid(int dtype_bytes, int live, size_t lds_bytes, int *grid)
{
#define NDF_NETWORK_GRID(LIVE)                                                                     \
    case LIVE:                                                                                     \
        return dtype_bytes == 4 ? batch_grid_live<float, LIVE>(lds_bytes, grid)                    \
                                : batch_grid_live<double, LIVE>(lds_bytes, grid);
    switch (live) {
        NDF_NETWORK_GRID(9)
        NDF_NETWORK_GRID(16)
        NDF_NETWORK_GRID(25)
        NDF_NETWORK_GRID(32)
        NDF_NETWORK_GRID(49)
        NDF_NETWORK_GRID(64)
    case 72:
        return dtype_bytes == 4 ? batch_grid_live<float, 72>(lds_bytes, grid) : hipErrorInvalidValue;
    case 96:
        return dtype_bytes == 4 ? batch_grid_live<float, 96>(lds_bytes, grid) : hipErrorInvalidValue;
    case 128:
        return dtype_bytes == 4 ? batch_grid_live<float, 128>(lds_bytes, grid) : hipErrorInvalidValue;
    case 160:
        return dtype_bytes == 4 ? batch_grid_live<float, 160>(lds_bytes, grid) : hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
#undef NDF_NETWORK_GRID
}

hipError_t launch_network_batch(int dtype_bytes, int live, const NetworkArgs *images, const int64_t *offsets,
                                const int32_t *work, int nwork, int32_t *counter, const void *input, void *output,
                                size_t lds_bytes, int grid, hipStream_t stream)
{
#define NDF_NETWORK_BATCH(LIVE)                                                                                  \
    case LIVE:                                                                                                   \
        return dtype_bytes == 4 ? launch_batch_live<float, LIVE>(images, offsets, work, nwork, counter, input, \
                                                                 output, lds_bytes, grid, stream)               \
                                : launch_batch_live<double, LIVE>(images, offsets, work, nwork, counter, input, \
                                                                  output, lds_bytes, grid, stream);
    switch (live) {
        NDF_NETWORK_BATCH(9)
        NDF_NETWORK_BATCH(16)
        NDF_NETWORK_BATCH(25)
        NDF_NETWORK_BATCH(32)
        NDF_NETWORK_BATCH(49)
        NDF_NETWORK_BATCH(64)
    case 72:
        return dtype_bytes == 4 ? launch_batch_live<float, 72>(images, offsets, work, nwork, counter, input, output,
                                                               lds_bytes, grid, stream)
                                : hipErrorInvalidValue;
    case 96:
        return dtype_bytes == 4 ? launch_batch_live<float, 96>(images, offsets, work, nwork, counter, input, output,
                                                               lds_bytes, grid, stream)
                                : hipErrorInvalidValue;
    case 128:
        return dtype_bytes == 4 ? launch_batch_live<float, 128>(images, offsets, work, nwork, counter, input, output,
                                                                lds_bytes, grid, stream)
                                : hipErrorInvalidValue;
    case 160:
        return dtype_bytes == 4 ? launch_batch_live<float, 160>(images, offsets, work, nwork, counter, input, output,
                                                                lds_bytes, grid, stream)
                                : hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
#undef NDF_NETWORK_BATCH
}

hipError_t launch_network(int dtype_bytes, const NetworkArgs &args, size_t lds_bytes, hipStream_t stream)
{
    if (args.pair_c > 0) {
        if (dtype_bytes != 4) return hipErrorInvalidValue;
#define NDF_PAIR_CASE(C, E) \
    if (args.pair_c == C && args.pair_e == E) return launch_pair<C, E>(args, lds_bytes, stream);
        NDF_PAIR_CASE(16, 8)
        NDF_PAIR_CASE(32, 12)
        NDF_PAIR_CASE(48, 20)
        NDF_PAIR_CASE(64, 24)
        NDF_PAIR_CASE(96, 32)
        NDF_PAIR_CASE(128, 32)
#undef NDF_PAIR_CASE
        return hipErrorInvalidValue;
    }
#define NDF_NETWORK_CASE(LIVE)                                                                     \
    case LIVE:                                                                                     \
        return dtype_bytes == 4 ? launch_live<float, LIVE>(args, lds_bytes, stream)                \
                                : launch_live<double, LIVE>(args, lds_bytes, stream);
    switch (args.live) {
        NDF_NETWORK_CASE(9)
        NDF_NETWORK_CASE(16)
        NDF_NETWORK_CASE(25)
        NDF_NETWORK_CASE(32)
        NDF_NETWORK_CASE(49)
        NDF_NETWORK_CASE(64)
    case 72:
        return dtype_bytes == 4 ? launch_live<float, 72>(args, lds_bytes, stream) : hipErrorInvalidValue;
    case 96:
        return dtype_bytes == 4 ? launch_live<float, 96>(args, lds_bytes, stream) : hipErrorInvalidValue;
    case 128:
        return dtype_bytes == 4 ? launch_live<float, 128>(args, lds_bytes, stream) : hipErrorInvalidValue;
    case 160:
        return dtype_bytes == 4 ? launch_live<float, 160>(args, lds_bytes, stream) : hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
#undef NDF_NETWORK_CASE
}

}  // namespace ndf
