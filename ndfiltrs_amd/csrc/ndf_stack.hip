// Stack engine: box windows, sorted slices merged per output.
//
// A box window of K0 x SH x SW taps (3-D: kz planes of ky x kx; 2-D: ky rows
// of 1 x kx) is K0 slices along its slowest axis. Neighbouring outputs along
// that axis share K0 - 1 slices, so each wave slides along it:
//   * per step it builds ONE new sorted slice: the lanes stage the slice's
//     columns (SH samples at one x, classified through the boundary maps),
//     sort each column in registers, exchange them through a wave-private LDS
//     row, and each lane merges the SW sorted columns of its own window;
//   * the output merges the K0 sorted slices with a merge network. A median
//     only reads ranks 0 .. W/2 (omitted samples carry the all-ones key and
//     sort last), so its network keeps only the comparators those ranks
//     depend on; rank and percentile statistics merge the full order.
// Per output this is one slice build plus one K0-way merge instead of a
// W-wire sort: for 3 x 7 x 7 about 880 comparators against 2,257, for 5 x 5
// about 94 against 140. The networks (and their 0-1 proofs) come from
// gen_stack_networks.py. The reference computes the same statistic per
// window with its median network or histogram (src/median_network.rs:65-209,
// src/filter_nd.rs); results are bit-identical to every other engine.
#include <utility>

#include "ndf_device.hpp"
#include "ndf_internal.hpp"
#include "ndf_select.hpp"
#include "ndf_stack_networks.inc"

namespace ndf {
namespace {

constexpr int STACK_THREADS = 256;
constexpr int STACK_WAVES = STACK_THREADS / 64;

enum NetPart { NET_COL, NET_SLICE, NET_HALF, NET_FULL, NET_CORE, NET_CORE_HALF, NET_FIN_HALF, NET_FIN_FULL };

template <class Net, int PART>
struct Part;
template <class Net>
struct Part<Net, NET_COL> {
    static constexpr int count = Net::col_count;
    static constexpr int lo(int i) { return Net::col[i][0]; }
    static constexpr int hi(int i) { return Net::col[i][1]; }
};
template <class Net>
struct Part<Net, NET_SLICE> {
    static constexpr int count = Net::slice_count;
    static constexpr int lo(int i) { return Net::slice[i][0]; }
    static constexpr int hi(int i) { return Net::slice[i][1]; }
};
template <class Net>
struct Part<Net, NET_HALF> {
    static constexpr int count = Net::half_count;
    static constexpr int lo(int i) { return Net::half[i][0]; }
    static constexpr int hi(int i) { return Net::half[i][1]; }
};
template <class Net>
struct Part<Net, NET_FULL> {
    static constexpr int count = Net::full_count;
    static constexpr int lo(int i) { return Net::full[i][0]; }
    static constexpr int hi(int i) { return Net::full[i][1]; }
};

template <class Net>
struct Part<Net, NET_CORE> {
    static constexpr int count = Net::core_count;
    static constexpr int lo(int i) { return Net::core[i][0]; }
    static constexpr int hi(int i) { return Net::core[i][1]; }
};
template <class Net>
struct Part<Net, NET_CORE_HALF> {
    static constexpr int count = Net::core_half_count;
    static constexpr int lo(int i) { return Net::core_half[i][0]; }
    static constexpr int hi(int i) { return Net::core_half[i][1]; }
};
template <class Net>
struct Part<Net, NET_FIN_HALF> {
    static constexpr int count = Net::fin_half_count;
    static constexpr int lo(int i) { return Net::fin_half[i][0]; }
    static constexpr int hi(int i) { return Net::fin_half[i][1]; }
};
template <class Net>
struct Part<Net, NET_FIN_FULL> {
    static constexpr int count = Net::fin_full_count;
    static constexpr int lo(int i) { return Net::fin_full[i][0]; }
    static constexpr int hi(int i) { return Net::fin_full[i][1]; }
};

// Combined offset of a slide-axis and a row-axis map entry: MAP_PAD when
// either is padding, MAP_CVAL when either is outside under constant mode.
__device__ __forceinline__ int64_t combine(int64_t a, int64_t b)
{
    if (a == MAP_PAD || b == MAP_PAD) return MAP_PAD;
    if (a < 0 || b < 0) return MAP_CVAL;
    return a + b;
}

// One unit: 64 columns of one row of the row axis, a run along the slide
// axis. MEDIAN: the statistic is the median (the lower-half network);
// otherwise rank or percentile (the full network). Separate kernels keep
// the median's register count to what its network needs. `lds` is the
// workgroup's dynamic LDS (one region per wave).
template <typename T, int K0, int SH, int SW, bool MEDIAN>
__device__ __forceinline__ void stack_unit(const StackArgs &a, int64_t unit, unsigned char *lds)
{
    using Net = stack_net::Net<K0, SH, SW>;
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    constexpr int L = Net::L, W = Net::W, H = Net::H;
    static_assert(L == SH * SW && W == K0 * L && H == W / 2 + 1, "network shape");
    // LDS rows per wave: the SH sorted keys of each column, and its retained
    // count (one row of single keys when SH == 1).
    constexpr int ROWS = SH > 1 ? SH + 1 : 1;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int sx = a.sx;
    Key *col = reinterpret_cast<Key *>(lds) + (size_t)wave * ROWS * sx;
    const int64_t xc = unit % a.tiles_x;
    const int64_t rest = unit / a.tiles_x;
    const int64_t r = rest % a.nr;
    const int64_t p0 = (rest / a.nr) * a.run;
    const int64_t p1 = p0 + a.run < a.ns ? p0 + a.run : a.ns;
    const int64_t x0 = xc * 64;

    const Key *input = static_cast<const Key *>(a.input);
    Key *output = static_cast<Key *>(a.output);
    const Key cval_key = classify<T>((Key)a.cval_bits, a.ignore_inf);
    const int64_t mxa = a.mapx[x0 + lane];
    const bool second = lane + 64 < sx;
    const int64_t mxb = second ? a.mapx[x0 + lane + 64] : MAP_PAD;
    int64_t mr[SH];
#pragma unroll
    for (int dy = 0; dy < SH; ++dy) mr[dy] = uniform_entry(a.mapr, r + dy);

    // Raw bits of one sample (a harmless in-bounds read where the sample is
    // padding or cval), and its key once the bits have arrived.
    auto fetch = [&](int64_t base, int64_t mx) -> Key { return input[base < 0 || mx < 0 ? 0 : base + mx]; };
    auto keyed = [&](Key bits, int64_t base, int64_t mx) -> Key {
        const bool pad = base == MAP_PAD || mx == MAP_PAD;
        const bool cv = !pad && (base == MAP_CVAL || mx < 0);
        return pad ? KT::SENTINEL : (cv ? cval_key : classify<T>(bits, a.ignore_inf));
    };

    // The lane's raw samples of one slide position (map entry ms): column
    // `lane` in raw[0 .. SH), column `lane + 64` (halo lanes only) in
    // raw[SH .. 2 SH). Loaded one step ahead, so the loads are in flight
    // during the output merge; keys are formed when the slice is built.
    auto load = [&](Key (&raw)[2 * SH], int64_t ms) {
#pragma unroll
        for (int dy = 0; dy < SH; ++dy) raw[dy] = fetch(combine(ms, mr[dy]), mxa);
        if (second) {
#pragma unroll
            for (int dy = 0; dy < SH; ++dy) raw[SH + dy] = fetch(combine(ms, mr[dy]), mxb);
        }
    };

    // Stages column c (its SH samples in ck) into LDS.
    auto stage = [&](Key (&ck)[SH], int c) {
        if constexpr (SH > 1) {
            int n = 0;
#pragma unroll
            for (int dy = 0; dy < SH; ++dy) n += ck[dy] != KT::SENTINEL;
            apply_net<Part<Net, NET_COL>>(ck);
#pragma unroll
            for (int dy = 0; dy < SH; ++dy) col[dy * sx + c] = ck[Net::col_out[dy]];
            col[SH * sx + c] = (Key)n;
        } else {
            col[c] = ck[0];
        }
    };

    // Builds the sorted slice of the loaded samples `raw` for this lane's
    // window; `count` receives its retained samples.
    auto build = [&](Key (&dst)[L], int &count, const Key (&raw)[2 * SH], int64_t ms) {
        wave_lds_sync();  // the previous slice's reads of `col` are done
        {
            Key ck[SH];
#pragma unroll
            for (int dy = 0; dy < SH; ++dy) ck[dy] = keyed(raw[dy], combine(ms, mr[dy]), mxa);
            stage(ck, lane);
        }
        if (second) {
            Key ck[SH];
#pragma unroll
            for (int dy = 0; dy < SH; ++dy) ck[dy] = keyed(raw[SH + dy], combine(ms, mr[dy]), mxb);
            stage(ck, lane + 64);
        }
        wave_lds_sync();
        Key v[L];
#pragma unroll
        for (int dx = 0; dx < SW; ++dx) {
#pragma unroll
            for (int dy = 0; dy < SH; ++dy) v[dx * SH + dy] = col[dy * sx + lane + dx];
        }
        int n = 0;
        if constexpr (SH > 1) {
#pragma unroll
            for (int dx = 0; dx < SW; ++dx) n += (int)col[SH * sx + lane + dx];
        } else {
#pragma unroll
            for (int i = 0; i < L; ++i) n += v[i] != KT::SENTINEL;
        }
        apply_net<Part<Net, NET_SLICE>>(v);
#pragma unroll
        for (int i = 0; i < L; ++i) dst[i] = v[Net::slice_out[i]];
        count = n;
    };

    // The statistic of one window from its sorted lower half h (median) or
    // its sorted keys f (other statistics), n retained.
    auto median_of = [&](const Key (&h)[H], int n) -> Key {
        const bool ok = n >= a.min_count && n > 0;
        const int hi = n >> 1;
        const bool pair = (n & 1) == 0;
        const int lo = pair ? hi - 1 : hi;
        Key klo, khi;
        median_ranks<Key, H>(h, ok, lo, hi, klo, khi);
        return ok ? (pair ? median_pair_bits<T>(klo, khi) : KT::unkey(khi)) : KT::CANONICAL_NAN;
    };
    auto statistic_of = [&](const Key (&f)[W], int n) -> Key {
        Pick pick;
        const bool ok = pick_ranks(a.stat, n, a.min_count, a.rank_param, a.q100, pick);
        const uint32_t hi_index = ok ? (uint32_t)pick.hi : 0u;
        const Key khi = select_index<Key, W>(f, hi_index);
        Key klo = khi;
        if (a.stat != STAT_RANK) klo = select_index<Key, W>(f, ok && pick.pair ? hi_index - 1u : hi_index);
        return ok ? (pick.pair ? pair_bits<T>(a.stat, klo, khi, pick.frac) : KT::unkey(khi)) : KT::CANONICAL_NAN;
    };

    // Slots: K0 slices, and one more for the two-output step of 2-D windows.
    constexpr bool PAIRS = SH == 1;
    constexpr int SLOTS = PAIRS ? K0 + 1 : K0;
    Key S[SLOTS][L];
    int cnt[SLOTS];
    // One window's result from slots [first, first + K0).
    auto emit_one = [&](int first) -> Key {
        Key w[W];
        int n = 0;
#pragma unroll
        for (int s = 0; s < K0; ++s) {
            n += cnt[first + s];
#pragma unroll
            for (int i = 0; i < L; ++i) w[s * L + i] = S[first + s][i];
        }
        if constexpr (MEDIAN) {
            apply_net<Part<Net, NET_HALF>>(w);
            Key h[H];
#pragma unroll
            for (int i = 0; i < H; ++i) h[i] = w[Net::half_out[i]];
            return median_of(h, n);
        } else {
            apply_net<Part<Net, NET_FULL>>(w);
            Key f[W];
#pragma unroll
            for (int i = 0; i < W; ++i) f[i] = w[Net::full_out[i]];
            return statistic_of(f, n);
        }
    };

    const int64_t last_q = a.ns + K0 - 2;  // last map entry of the slide axis
    const int64_t x = x0 + lane;
    Key raw[2 * SH];
#pragma unroll
    for (int s = 0; s + 1 < K0; ++s) {
        const int64_t ms = uniform_entry(a.maps, p0 + s);
        load(raw, ms);
        build(S[s], cnt[s], raw, ms);
    }
    int64_t ms = uniform_entry(a.maps, p0 + K0 - 1);
    load(raw, ms);
    int64_t p = p0;
    if constexpr (PAIRS) {
        // Two outputs per step: rows p and p + 1 share the K0 - 1 slices in
        // slots 1 .. K0 - 1 (the core), merged once; each output merges the
        // sorted core with its own last slice (slot 0, slot K0).
        Key raw1[2 * SH];
        int64_t ms1 = uniform_entry(a.maps, p0 + K0 <= last_q ? p0 + K0 : last_q);
        load(raw1, ms1);
#pragma unroll 1
        for (; p + 1 < p1; p += 2) {
            build(S[K0 - 1], cnt[K0 - 1], raw, ms);
            build(S[K0], cnt[K0], raw1, ms1);
            // The next step's samples (clamped to the slide axis's map).
            ms = uniform_entry(a.maps, p + K0 + 1 <= last_q ? p + K0 + 1 : last_q);
            load(raw, ms);
            ms1 = uniform_entry(a.maps, p + K0 + 2 <= last_q ? p + K0 + 2 : last_q);
            load(raw1, ms1);
            constexpr int C = Net::C;
            Key core[C];
            int ncore = 0;
#pragma unroll
            for (int s = 1; s < K0; ++s) {
                ncore += cnt[s];
#pragma unroll
                for (int i = 0; i < L; ++i) core[(s - 1) * L + i] = S[s][i];
            }
            if constexpr (MEDIAN) {
                apply_net<Part<Net, NET_CORE_HALF>>(core);
            } else {
                apply_net<Part<Net, NET_CORE>>(core);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int extra = j == 0 ? 0 : K0;
                Key u[W];
#pragma unroll
                for (int i = 0; i < C; ++i) u[i] = core[Net::core_out[i]];
#pragma unroll
                for (int i = 0; i < L; ++i) u[C + i] = S[extra][i];
                const int n = ncore + cnt[extra];
                Key out;
                if constexpr (MEDIAN) {
                    apply_net<Part<Net, NET_FIN_HALF>>(u);
                    Key h[H];
#pragma unroll
                    for (int i = 0; i < H; ++i) h[i] = u[Net::fin_half_out[i]];
                    out = median_of(h, n);
                } else {
                    apply_net<Part<Net, NET_FIN_FULL>>(u);
                    Key f[W];
#pragma unroll
                    for (int i = 0; i < W; ++i) f[i] = u[Net::fin_full_out[i]];
                    out = statistic_of(f, n);
                }
                if (x < a.nx) output[(p + j) * a.out_ss + r * a.out_sr + x] = out;
            }
#pragma unroll
            for (int s = 0; s + 2 < SLOTS; ++s) {
                cnt[s] = cnt[s + 2];
#pragma unroll
                for (int i = 0; i < L; ++i) S[s][i] = S[s + 2][i];
            }
        }
    }
    // One output per step (3-D windows; the odd last row of a 2-D run).
#pragma unroll 1
    for (; p < p1; ++p) {
        build(S[K0 - 1], cnt[K0 - 1], raw, ms);
        ms = uniform_entry(a.maps, p + K0 <= last_q ? p + K0 : last_q);
        load(raw, ms);
        const Key out = emit_one(0);
        if (x < a.nx) output[p * a.out_ss + r * a.out_sr + x] = out;
#pragma unroll
        for (int s = 0; s + 1 < K0; ++s) {
            cnt[s] = cnt[s + 1];
#pragma unroll
            for (int i = 0; i < L; ++i) S[s][i] = S[s + 1][i];
        }
    }
}

template <typename T, int K0, int SH, int SW, bool MEDIAN>
__global__ void __launch_bounds__(STACK_THREADS) k_stack(StackArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // XCD-aware order: the grid is a multiple of 8 blocks, dispatched round
    // robin over the 8 XCDs; XCD j runs the contiguous block range
    // [j * nb / 8, (j + 1) * nb / 8), whose units share input rows in its L2.
    const int nb8 = (int)(gridDim.x >> 3);
    const int64_t block = (int64_t)(blockIdx.x & 7) * nb8 + (blockIdx.x >> 3);
    const int64_t unit = block * STACK_WAVES + wave;
    if (unit >= a.units) return;  // wave-uniform; the kernel has no barrier
    stack_unit<T, K0, SH, SW, MEDIAN>(a, unit, lds);
}

// A batch of images whose windows share the box: workgroups take (image,
// block of 4 units) pairs from a work queue; images[i] holds image i's
// arguments (input / output unused: its data starts at offsets[i]).
template <typename T, int K0, int SH, int SW, bool MEDIAN>
__global__ void __launch_bounds__(STACK_THREADS) k_stack_batch(const StackArgs *images, const int64_t *offsets,
                                                               const int32_t *work, int nwork, int32_t *counter,
                                                               const void *input, void *output)
{
    using Key = typename KeyTraits<T>::Key;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int next_shared;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (;;) {
        if (threadIdx.x == 0) next_shared = atomicAdd(counter, 1);
        __syncthreads();
        const int w = __builtin_amdgcn_readfirstlane(next_shared);
        __syncthreads();
        if (w >= nwork) break;
        const int image = work[2 * w];
        StackArgs a = images[image];
        a.input = static_cast<const Key *>(input) + offsets[image];
        a.output = static_cast<Key *>(output) + offsets[image];
        const int64_t unit = (int64_t)work[2 * w + 1] * STACK_WAVES + wave;
        if (unit < a.units) stack_unit<T, K0, SH, SW, MEDIAN>(a, unit, lds);
    }
}

template <typename T, int K0, int SH, int SW>
hipError_t launch_shape(const StackArgs &args, hipStream_t stream)
{
    const size_t lds = stack_lds_bytes((int)sizeof(T), SH, args.sx);
    const int64_t blocks = (args.units + STACK_WAVES - 1) / STACK_WAVES;
    const int64_t grid = (blocks + 7) / 8 * 8;
    if (args.stat == STAT_MEDIAN) {
        hipLaunchKernelGGL((k_stack<T, K0, SH, SW, true>), dim3((unsigned)grid), dim3(STACK_THREADS), lds, stream,
                           args);
    } else {
        hipLaunchKernelGGL((k_stack<T, K0, SH, SW, false>), dim3((unsigned)grid), dim3(STACK_THREADS), lds, stream,
                           args);
    }
    return hipGetLastError();
}

template <typename T, int K0, int SH, int SW>
hipError_t batch_shape(const StackArgs *images, const int64_t *offsets, const int32_t *work, int nwork,
                       int32_t *counter, const void *input, void *output, bool median, int sx, int grid,
                       hipStream_t stream)
{
    const size_t lds = stack_lds_bytes((int)sizeof(T), SH, sx);
    if (median) {
        hipLaunchKernelGGL((k_stack_batch<T, K0, SH, SW, true>), dim3((unsigned)grid), dim3(STACK_THREADS), lds,
                           stream, images, offsets, work, nwork, counter, input, output);
    } else {
        hipLaunchKernelGGL((k_stack_batch<T, K0, SH, SW, false>), dim3((unsigned)grid), dim3(STACK_THREADS), lds,
                           stream, images, offsets, work, nwork, counter, input, output);
    }
    return hipGetLastError();
}

template <typename T, int K0, int SH, int SW>
hipError_t batch_grid_shape(bool median, int sx, int *grid)
{
    const void *fn = median ? reinterpret_cast<const void *>(&k_stack_batch<T, K0, SH, SW, true>)
                            : reinterpret_cast<const void *>(&k_stack_batch<T, K0, SH, SW, false>);
    const size_t lds = stack_lds_bytes((int)sizeof(T), SH, sx);
    int per_cu = 0, device = 0, cus = 0;
    hipError_t e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, STACK_THREADS, lds)) != hipSuccess ||
        (e = hipGetDevice(&device)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) {
        return e;
    }
    *grid = (per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
    return hipSuccess;
}

}  // namespace

size_t stack_lds_bytes(int dtype_bytes, int sh, int sx)
{
    return (size_t)STACK_WAVES * (size_t)(sh > 1 ? sh + 1 : 1) * (size_t)sx * (size_t)dtype_bytes;
}

bool stack_supported(int dtype_bytes, int k0, int sh, int sw)
{
    const int w = k0 * sh * sw;
    if (dtype_bytes == 8 && w > 49) return false;
    return (sh == 1 && k0 == sw && (k0 == 3 || k0 == 5 || k0 == 7 || k0 == 9)) ||
           (k0 == 3 && sh == sw && (sh == 3 || sh == 5 || sh == 7));
}

int64_t stack_max_units() { return ((int64_t)1 << 31) - 8 * STACK_WAVES; }

hipError_t launch_stack(int dtype_bytes, const StackArgs &args, hipStream_t stream)
{
#define NDF_STACK_CASE(K0, SH, SW)                                                              \
    if (args.k0 == K0 && args.sh == SH && args.sw == SW) {                                      \
        return dtype_bytes == 4 ? launch_shape<float, K0, SH, SW>(args, stream)                 \
                                : launch_shape<double, K0, SH, SW>(args, stream);               \
    }
#define NDF_STACK_CASE_F32(K0, SH, SW)                                                          \
    if (args.k0 == K0 && args.sh == SH && args.sw == SW) {                                      \
        return dtype_bytes == 4 ? launch_shape<float, K0, SH, SW>(args, stream) : hipErrorInvalidValue; \
    }
    NDF_STACK_CASE(3, 1, 3)
    NDF_STACK_CASE(5, 1, 5)
    NDF_STACK_CASE(7, 1, 7)
    NDF_STACK_CASE_F32(9, 1, 9)
    NDF_STACK_CASE(3, 3, 3)
    NDF_STACK_CASE_F32(3, 5, 5)
    NDF_STACK_CASE_F32(3, 7, 7)
#undef NDF_STACK_CASE
#undef NDF_STACK_CASE_F32
    return hipErrorInvalidValue;
}

// Shapes compiled for both dtypes (B) and for float32 only (F).
#define NDF_STACK_SHAPES(B, F) B(3, 1, 3) B(5, 1, 5) B(7, 1, 7) F(9, 1, 9) B(3, 3, 3) F(3, 5, 5) F(3, 7, 7)

hipError_t stack_batch_grid(int dtype_bytes, const StackArgs &shape, int *grid)
{
    if (!stack_supported(dtype_bytes, shape.k0, shape.sh, shape.sw)) return hipErrorInvalidValue;
    const bool median = shape.stat == STAT_MEDIAN;
#define NDF_GRID_BOTH(K0, SH, SW)                                                                        \
    if (shape.k0 == K0 && shape.sh == SH && shape.sw == SW) {                                           \
        return dtype_bytes == 4 ? batch_grid_shape<float, K0, SH, SW>(median, shape.sx, grid)            \
                                : batch_grid_shape<double, K0, SH, SW>(median, shape.sx, grid);          \
    }
#define NDF_GRID_F32(K0, SH, SW)                                                                         \
    if (shape.k0 == K0 && shape.sh == SH && shape.sw == SW) {                                           \
        return batch_grid_shape<float, K0, SH, SW>(median, shape.sx, grid);                               \
    }
    NDF_STACK_SHAPES(NDF_GRID_BOTH, NDF_GRID_F32)
#undef NDF_GRID_BOTH
#undef NDF_GRID_F32
    return hipErrorInvalidValue;
}

hipError_t launch_stack_batch(int dtype_bytes, const StackArgs &shape, const StackArgs *images,
                              const int64_t *offsets, const int32_t *work, int nwork, int32_t *counter,
                              const void *input, void *output, int grid, hipStream_t stream)
{
    if (!stack_supported(dtype_bytes, shape.k0, shape.sh, shape.sw)) return hipErrorInvalidValue;
    const bool median = shape.stat == STAT_MEDIAN;
#define NDF_BATCH_BOTH(K0, SH, SW)                                                                        \
    if (shape.k0 == K0 && shape.sh == SH && shape.sw == SW) {                                            \
        return dtype_bytes == 4 ? batch_shape<float, K0, SH, SW>(images, offsets, work, nwork, counter,   \
                                                                 input, output, median, shape.sx, grid,   \
                                                                 stream)                                  \
                                : batch_shape<double, K0, SH, SW>(images, offsets, work, nwork, counter,  \
                                                                  input, output, median, shape.sx, grid,  \
                                                                  stream);                                \
    }
#define NDF_BATCH_F32(K0, SH, SW)                                                                         \
    if (shape.k0 == K0 && shape.sh == SH && shape.sw == SW) {                                            \
        return batch_shape<float, K0, SH, SW>(images, offsets, work, nwork, counter, input, output, median, \
                                              shape.sx, grid, stream);                                    \
    }
    NDF_STACK_SHAPES(NDF_BATCH_BOTH, NDF_BATCH_F32)
#undef NDF_BATCH_BOTH
#undef NDF_BATCH_F32
    return hipErrorInvalidValue;
}

}  // namespace ndf
