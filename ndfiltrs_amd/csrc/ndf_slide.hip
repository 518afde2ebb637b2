// Slide engine: LDS-staged tile + block sort + per-lane column slide.
//
// This is the reference's block rank histogram (src/histogram.rs:1-16,
// 302-485) re-designed for a 64-wide wavefront:
//   * a workgroup (4 waves) stages the support of a 64 x ty output tile (tile
//     plus window halo, mapped through the boundary maps) into LDS with
//     coalesced row loads; retained samples are compacted, omitted ones
//     (NaN, and +-inf under ignore_inf) never enter the sort;
//   * the retained samples are sorted once per tile by a register-resident
//     bitonic network (each thread owns E consecutive elements; strides
//     below E stay in registers, larger strides exchange through LDS), which
//     gives every support sample a tile rank ("code") and every rank its
//     position;
//   * each lane owns one output column and slides its window down the rows
//     of its wave. Instead of a bit histogram per window it keeps a cursor:
//     a rank m and c = number of in-window ranks below m. A step updates c
//     from the removed and added edge taps only, then walks m to the target
//     rank, testing window membership of 16 ranks per LDS batch;
//   * a wave starts from a two-level rank histogram of its first window
//     (column counts for rectangles, per-lane counts for footprints), so the
//     first walk is short.
// The selected rank maps back to the original key, so the output is the
// exact sample (or the f64 midpoint of two), independent of the tile shape.
#include <type_traits>

#include "ndf_device.hpp"
#include "ndf_internal.hpp"

namespace ndf {
namespace {

// Codes of support samples in the plain kernel: 1 + tile rank for retained
// samples, CODE_OMIT for omitted samples; CODE_PENDING marks retained
// samples between staging and ranking. Every rank code is below
// CODE_PENDING. (The bracket kernel keeps codes only for inside samples.)
constexpr uint16_t CODE_PENDING = 0xFFFD;
constexpr uint16_t CODE_OMIT = 0xFFFF;
constexpr uint16_t POS_NONE = 0xFFFF;
constexpr int NT = SLIDE_WAVES * 64;
constexpr int BINS = 16;
constexpr int BATCH = 16;
constexpr int START = 4;  // start-cursor candidates per wave in the bracket kernel
constexpr int MAX_PASSES = 8;  // key ranges the bracket kernel may run per tile
constexpr int GUARD = BATCH;

// Support positions are packed as (row << 8) | x with row = z * zrows + y
// (row <= 254, x <= 254); POS_NONE (row 255, x 255) is outside every window.
struct Geometry {
    int kz, ky, kx;
    int zshift;
    uint32_t vmask;  // keeps (y within its plane, x) of a packed position
    const uint8_t *mask;
    // Byte-wise membership applies: kx, ky < 128 and a row difference
    // y - row (at most ty - 1) never wraps into [256 - ky, 256).
    bool bytes_ok;

    // Window membership of the sample at packed position v for the window
    // whose top-left support sample is `base` = (y << 8) | lane. Branch-free:
    // d = v - base holds dx in its low byte when x >= lane (otherwise the low
    // byte is >= 256 - 63 > kx) and dy in the bits above it (huge when
    // negative).
    __device__ __forceinline__ uint32_t contains(uint32_t v, uint32_t base) const
    {
        const uint32_t d = (v & vmask) - base;
        const uint32_t dx = d & 0xFFu;
        const uint32_t dy = d >> 8;
        const uint32_t box = (uint32_t)(dx < (uint32_t)kx) & (uint32_t)(dy < (uint32_t)ky);
        if (mask == nullptr) {
            return box;
        }
        const uint32_t dz = v >> (8 + zshift);
        const uint32_t index = box ? (dz * (uint32_t)ky + dy) * (uint32_t)kx + dx : 0u;
        return box & (uint32_t)(mask[index] != 0);
    }
};

// Position of the `need`-th set bit (0-based) of a 16-bit mask holding more
// than `need` set bits.
__device__ __forceinline__ int select_bit16(uint32_t mask, int need)
{
    int base = 0;
#pragma unroll
    for (int width = 8; width >= 1; width >>= 1) {
        const int low = __builtin_popcount(mask & ((1u << width) - 1u));
        const bool up = need >= low;
        need -= up ? low : 0;
        mask >>= up ? width : 0;
        base += up ? width : 0;
    }
    return base;
}

// Membership bits of ranks [first, first + BATCH) (bit i <-> rank first + i).
// `pos` has BATCH guard entries (POS_NONE) on both sides, and ranks in
// [retained, N) hold POS_NONE, so first in [-BATCH, N) needs no clamping.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Rectangle membership of two packed positions at once with 16-bit lane
// arithmetic: for each half, dy = row - y and dx = x - lane (both wrap when
// negative), and d < K holds exactly when min(d, K) - K wraps below zero, so
// the result's sign bits (15 and 31) are the two membership bits.
__device__ __forceinline__ uint32_t contains2(u16x2 v, u16x2 vmask, u16x2 y, u16x2 lane, u16x2 ky, u16x2 kx)
{
    v &= vmask;
    const u16x2 dy = (u16x2)(v >> (u16x2)8) - y;
    const u16x2 dx = (u16x2)(v & (u16x2)0xFF) - lane;
    const u16x2 iny = __builtin_elementwise_min(dy, ky) - ky;
    const u16x2 inx = __builtin_elementwise_min(dx, kx) - kx;
    const u16x2 both = iny & inx;
    return __builtin_bit_cast(uint32_t, both) & 0x80008000u;
}

// Byte-wise rectangle membership of the two packed positions in `w` (one per
// 16-bit half) for windows with kx, ky < 128: u = w - base per half gives the
// bytes (dx, dy) of each position (dx >= 193 when x < lane, dy >= 128 when
// row < y); a byte b is inside its bound K exactly when b < 128 and
// (b | 128) - K < 128, which one 32-bit subtraction tests for all four bytes
// without borrows between them. Returns bit 7 (position 0) and bit 23
// (position 1); bit 15 (dy of position 0 with dx of position 1) is cleared.
__device__ __forceinline__ uint32_t contains2_bytes(uint32_t w, uint32_t base2, uint32_t bounds)
{
    const uint32_t u = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, w) - __builtin_bit_cast(u16x2, base2));
    const uint32_t s = (u | 0x80808080u) - bounds;
    const uint32_t ok = ~(s | u) & 0x80808080u;
    return ok & (ok >> 8) & 0x00800080u;
}

// Byte-wise membership bits of the BATCH packed positions in w (rectangles
// with g.bytes_ok).
__device__ __forceinline__ uint32_t bytes_mask(const uint32_t w[BATCH / 2], const Geometry &g, uint32_t base)
{
    const uint32_t base2 = base * 0x10001u;
    const uint32_t vmask2 = g.vmask * 0x10001u;
    const uint32_t bounds = ((uint32_t)g.kx | ((uint32_t)g.ky << 8)) * 0x10001u;
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < BATCH / 2; ++j) {
        const uint32_t both = contains2_bytes(w[j] & vmask2, base2, bounds);
        acc |= 2 * j >= 7 ? both << (2 * j - 7) : both >> (7 - 2 * j);
    }
    return (acc & 0x5555u) | ((acc >> 15) & 0xAAAAu);
}

__device__ __forceinline__ uint32_t batch_mask(const uint16_t *pos, int first, const Geometry &g, uint32_t base)
{
    if (g.mask == nullptr && g.bytes_ok) {
        uint32_t w[BATCH / 2];
        __builtin_memcpy(w, pos + first, sizeof(w));
        return bytes_mask(w, g, base);
    }
    uint32_t v[BATCH];
#pragma unroll
    for (int i = 0; i < BATCH; ++i) {
        v[i] = pos[first + i];
    }
    uint32_t mask = 0;
    if (g.mask == nullptr) {
        // Rectangles: two positions per 16-bit-pair operation. Bit 15 of a
        // pair result lands on bit 2j of `acc`, bit 31 on bit 16 + 2j.
        const u16x2 vmask = (u16x2)(unsigned short)g.vmask;
        const u16x2 y = (u16x2)(unsigned short)(base >> 8);
        const u16x2 lane = (u16x2)(unsigned short)(base & 0xFFu);
        const u16x2 ky = (u16x2)(unsigned short)g.ky;
        const u16x2 kx = (u16x2)(unsigned short)g.kx;
        uint32_t acc = 0;
#pragma unroll
        for (int j = 0; j < BATCH / 2; ++j) {
            const u16x2 pair = {(unsigned short)v[2 * j], (unsigned short)v[2 * j + 1]};
            acc |= contains2(pair, vmask, y, lane, ky, kx) >> (15 - 2 * j);
        }
        mask = (acc & 0x5555u) | ((acc >> 15) & 0xAAAAu);
    } else {
#pragma unroll
        for (int i = 0; i < BATCH; ++i) {
            mask |= g.contains(v[i], base) << i;
        }
    }
    return mask;
}

// Moves the cursor (m, c) to the rank m with c == k in-window samples below it
// and m itself in the window. Ranks are 1-based (pos[r] for r in
// [1, retained]); rank 0 stands for the samples below the bracket, which are
// counted in c but never visited. Returns false if the target is not a
// ranked sample (below or above the bracket in bracket mode; never in plain
// mode for a consistent cursor).
__device__ inline bool walk_to(int &m, int &c, int k, const uint16_t *pos, int retained, const Geometry &g,
                               uint32_t base)
{
    const int limit = retained / BATCH + 4;
    // Upward (c <= k): the (k - c)-th (0-based) in-window rank at or above m.
    // Downward (c > k): the (c - k)-th in-window rank below m, counting down.
    // One loop serves both directions, so a wave whose lanes walk both ways
    // runs its longest walk, not the sum of the longest walk each way.
    const bool up = c <= k;
    int need = up ? k - c : c - k;
    int first = up ? m : m - BATCH;
    const int advance = up ? BATCH : -BATCH;
    if (g.mask == nullptr && g.bytes_ok) {
        // Rectangles: the next batch's positions are read while this one is
        // tested. A read past the last batch stays inside LDS (keys lie below
        // the positions, the codes above) and is never used.
        uint32_t cur[BATCH / 2];
        __builtin_memcpy(cur, pos + first, sizeof(cur));
        for (int step = 0; step <= limit; ++step) {
            uint32_t next[BATCH / 2];
            __builtin_memcpy(next, pos + first + advance, sizeof(next));
            const uint32_t mask = bytes_mask(cur, g, base);
            const int count = __builtin_popcount(mask);
            if (up ? count > need : count >= need) {
                m = first + select_bit16(mask, up ? need : count - need);
                c = k;
                return true;
            }
            need -= count;
            first += advance;
            if (up ? first > retained : first + BATCH <= 1) {
                return false;
            }
#pragma unroll
            for (int j = 0; j < BATCH / 2; ++j) {
                cur[j] = next[j];
            }
        }
        return false;
    }
    for (int step = 0; step <= limit; ++step) {
        const uint32_t mask = batch_mask(pos, first, g, base);
        const int count = __builtin_popcount(mask);
        if (up ? count > need : count >= need) {
            m = first + select_bit16(mask, up ? need : count - need);
            c = k;
            return true;
        }
        need -= count;
        first += advance;
        if (up ? first > retained : first + BATCH <= 1) {
            return false;
        }
    }
    return false;
}

// Largest in-window rank below m, or a value below 1 when none is ranked.
__device__ inline int previous_member(int m, const uint16_t *pos, const Geometry &g, uint32_t base)
{
    for (int first = m - BATCH; first > -BATCH; first -= BATCH) {
        const uint32_t mask = batch_mask(pos, first, g, base) & ((1u << (m - first)) - 1u);
        if (mask) {
            return first + 31 - __builtin_clz(mask);
        }
        m = first;
    }
    return -1;
}

// Packed 16-bin counters: 8 bits per bin while counts stay below 256,
// otherwise 16 bits per bin.
// add(b) counts bin b when b < 16 and nothing otherwise, without branches.
struct Bins8 {
    uint64_t lo = 0, hi = 0;
    __device__ __forceinline__ void add(uint32_t b)
    {
        const uint64_t one = 1ull << ((b & 7) * 8);
        lo += b < 8 ? one : 0ull;
        hi += b - 8u < 8u ? one : 0ull;
    }
    __device__ inline int get(int b) const { return (int)(((b < 8 ? lo : hi) >> ((b & 7) * 8)) & 0xFF); }
};

struct Bins16 {
    uint64_t q[4] = {0, 0, 0, 0};
    __device__ __forceinline__ void add(uint32_t b)
    {
        const uint64_t one = 1ull << ((b & 3) * 16);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            q[i] += (b >> 2) == i ? one : 0ull;
        }
    }
    __device__ inline int get(int b) const { return (int)((q[b >> 2] >> ((b & 3) * 16)) & 0xFFFF); }
};

// Counts, for one lane's window at row y, the codes in [lo, lo + 16 << shift)
// per bin, in `below` the codes under lo and in `above` the retained codes at
// or over the range. Rectangles use per-column counts shared through LDS
// (hist, BINS + 2 rows of sx + 1 entries); other footprints scan their taps.
template <typename Bins>
__device__ __forceinline__ void window_bins(int counts[BINS], int &below, int &above, int lo, int shift, const uint16_t *codes,
                            uint16_t *hist, const SlideArgs &a, int lane, int y)
{
    const int hi = lo + (BINS << shift);
    if (a.mask == nullptr) {
        const int stride = a.sx + 1;
        for (int x = lane; x < a.sx; x += 64) {
            Bins bins;
            int under = 0, over = 0;
            auto count = [&](int code) {
                under += code < lo;
                over += code >= hi && code != CODE_OMIT;
                bins.add((uint32_t)(code - lo) >> shift);  // no bin below lo or at hi and above
            };
            for (int z = 0; z < a.kz; ++z) {
                const uint16_t *column = codes + (z * a.zrows + y) * a.sx + x;
                int dy = 0;
                // Eight reads in flight per step.
                for (; dy + 8 <= a.ky; dy += 8) {
                    int code[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        code[i] = column[(dy + i) * a.sx];
                    }
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        count(code[i]);
                    }
                }
                for (; dy < a.ky; ++dy) {
                    count(column[dy * a.sx]);
                }
            }
#pragma unroll
            for (int b = 0; b < BINS; ++b) {
                hist[b * stride + x] = (uint16_t)bins.get(b);
            }
            hist[BINS * stride + x] = (uint16_t)under;
            hist[(BINS + 1) * stride + x] = (uint16_t)over;
        }
        // Exclusive prefix over columns, one bin per lane. LDS accesses of one
        // wave complete in order; the wave barrier keeps the compiler from
        // moving them across the phases.
        __builtin_amdgcn_wave_barrier();
        if (lane <= BINS + 1) {
            uint16_t *row = hist + lane * stride;
            int run = 0;
            for (int x0 = 0; x0 < a.sx; x0 += 8) {
                int v[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    v[i] = x0 + i < a.sx ? row[x0 + i] : 0;
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if (x0 + i < a.sx) {
                        row[x0 + i] = (uint16_t)run;
                        run += v[i];
                    }
                }
            }
            row[a.sx] = (uint16_t)run;
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int b = 0; b < BINS; ++b) {
            counts[b] = (int)hist[b * stride + lane + a.kx] - (int)hist[b * stride + lane];
        }
        below = (int)hist[BINS * stride + lane + a.kx] - (int)hist[BINS * stride + lane];
        above = (int)hist[(BINS + 1) * stride + lane + a.kx] - (int)hist[(BINS + 1) * stride + lane];
        __builtin_amdgcn_wave_barrier();
    } else {
        Bins bins;
        int under = 0, over = 0;
        const int base = y * a.sx + lane;
        for (int t = 0; t < a.count; ++t) {
            const int code = codes[base + a.taps[t]];
            under += code < lo;
            over += code >= hi && code != CODE_OMIT;
            if (code >= lo && code < hi) {
                bins.add((code - lo) >> shift);
            }
        }
#pragma unroll
        for (int b = 0; b < BINS; ++b) {
            counts[b] = bins.get(b);
        }
        below = under;
        above = over;
    }
}

// Compile-time loop: calls f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Value of lane ^ MASK (MASK < 64) within the wave, through VALU lane
// movement only (no LDS crossbar round trip):
//   1, 2: DPP quad_perm [1,0,3,2] / [2,3,0,1];
//   4:    DPP row_half_mirror (lane ^ 7 within 8) then quad_perm [3,2,1,0]
//         (lane ^ 3);
//   8:    DPP row_ror:8 (a half rotation of a 16-lane row is lane ^ 8);
//   16:   v_permlane16_swap of v with itself: the first result holds rows
//         (0, 0, 2, 2), the second rows (1, 1, 3, 3);
//   32:   v_permlane32_swap of v with itself: the first result holds halves
//         (0, 0), the second (1, 1).
template <int MASK>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v)
{
    constexpr int ALL = 0xF;
    if constexpr (MASK == 1) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, ALL, ALL, true);
    } else if constexpr (MASK == 2) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, ALL, ALL, true);
    } else if constexpr (MASK == 4) {
        const int mirrored = __builtin_amdgcn_update_dpp(0, (int)v, 0x141, ALL, ALL, true);
        return (uint32_t)__builtin_amdgcn_update_dpp(0, mirrored, 0x1B, ALL, ALL, true);
    } else if constexpr (MASK == 8) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, ALL, ALL, true);
    } else if constexpr (MASK == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (threadIdx.x & 16) ? r[0] : r[1];
    } else {
        static_assert(MASK == 32, "lane masks are powers of two below 64");
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (threadIdx.x & 32) ? r[0] : r[1];
    }
}

template <int MASK>
__device__ __forceinline__ uint64_t xor_lane(uint64_t v)
{
    const uint32_t lo = xor_lane<MASK>((uint32_t)v);
    const uint32_t hi = xor_lane<MASK>((uint32_t)(v >> 32));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

template <typename Key, int E>
struct Chunk {
    Key k[E];
    uint32_t p[E];

    // Compare-exchange of local elements I < J: ascending puts the smaller
    // key at I. A descending pair of equal keys is swapped, which only
    // permutes two equal elements of this thread. The swap condition is one
    // compare and a lane-mask xor (no boolean select, which the compiler
    // would materialise in vector registers).
    template <int I, int J>
    __device__ __forceinline__ void ce(bool asc)
    {
        const Key ki = k[I], kj = k[J];
        const bool swap = (kj < ki) ^ !asc;
        const uint32_t pi = p[I], pj = p[J];
        k[I] = swap ? kj : ki;
        k[J] = swap ? ki : kj;
        p[I] = swap ? pj : pi;
        p[J] = swap ? pi : pj;
    }

    // Element I against the partner's element `o` (position `op`): the lane
    // that keeps the minimum takes `o` only when it is smaller, the other
    // lane only when it is larger, so equal keys stay where they are on both
    // sides and no position is duplicated.
    __device__ __forceinline__ static void take_partner(Key &k_i, uint32_t &p_i, Key o, uint32_t op, bool keep_min)
    {
        const Key lo = o < k_i ? o : k_i;
        const Key hi = o < k_i ? k_i : o;
        const Key kept = keep_min ? lo : hi;
        p_i = kept != k_i ? op : p_i;
        k_i = kept;
    }

    // One in-register bitonic stage of stride STRIDE; the direction of the
    // pair (I, I + STRIDE) is ascending when bit SIZE of its global index
    // (tbase | I) is clear.
    template <int SIZE, int STRIDE>
    __device__ __forceinline__ void stage(int tbase)
    {
        static_for<0, E>([&](auto ic) {
            constexpr int I = decltype(ic)::value;
            if constexpr ((I & STRIDE) == 0) {
                ce<I, I + STRIDE>(((tbase | I) & SIZE) == 0);
            }
        });
    }

    template <int SIZE, int STRIDE>
    __device__ __forceinline__ void stages_down(int tbase)
    {
        if constexpr (STRIDE > 0) {
            stage<SIZE, STRIDE>(tbase);
            stages_down<SIZE, STRIDE / 2>(tbase);
        }
    }

    // Sizes 2 .. E: a full bitonic sort of the chunk (direction per the
    // global index, so neighbouring chunks form bitonic sequences).
    template <int SIZE>
    __device__ __forceinline__ void local_sort(int tbase)
    {
        if constexpr (SIZE <= E) {
            stages_down<SIZE, SIZE / 2>(tbase);
            local_sort<SIZE * 2>(tbase);
        }
    }

    // Compare-exchange of the whole chunk with the chunk of lane ^ MASK (same
    // wave): keep_min keeps the smaller of each pair of elements.
    template <int MASK>
    __device__ __forceinline__ void exchange_xor(bool keep_min)
    {
        if constexpr (MASK <= 8 && sizeof(Key) == 4) {
            // DPP lane movement folds into the min / max / select that
            // consume it, so keys and positions move one by one.
            static_for<0, E>([&](auto ic) {
                constexpr int I = decltype(ic)::value;
                take_partner(k[I], p[I], xor_lane<MASK>(k[I]), xor_lane<MASK>(p[I]), keep_min);
                if constexpr (I % 8 == 7) {
                    __builtin_amdgcn_sched_barrier(0);
                }
            });
            return;
        }
        static_for<0, E / 2>([&](auto ic) {
            constexpr int I = 2 * decltype(ic)::value;
            const uint32_t packed = xor_lane<MASK>((uint32_t)(p[I] | (p[I + 1] << 16)));
            const Key o0 = xor_lane<MASK>(k[I]);
            const Key o1 = xor_lane<MASK>(k[I + 1]);
            take_partner(k[I], p[I], o0, packed & 0xFFFFu, keep_min);
            take_partner(k[I + 1], p[I + 1], o1, packed >> 16, keep_min);
            if constexpr (I % 8 == 6) {
                __builtin_amdgcn_sched_barrier(0);  // bounds the swizzles in flight
            }
        });
    }

    // Strides E/2 .. 1 of a merge level larger than the chunk: one direction
    // for the whole chunk.
    __device__ __forceinline__ void merge_local(bool asc)
    {
        merge_down<E / 2>(asc);
    }

    template <int STRIDE>
    __device__ __forceinline__ void merge_down(bool asc)
    {
        if constexpr (STRIDE > 0) {
            static_for<0, E>([&](auto ic) {
                constexpr int I = decltype(ic)::value;
                if constexpr ((I & STRIDE) == 0) {
                    ce<I, I + STRIDE>(asc);
                }
            });
            merge_down<STRIDE / 2>(asc);
        }
    }
};

// Exchange layout of the bitonic sort's cross-thread stages. For E >= 8 a
// thread's chunk is stored as 16-byte blocks (keys: 16 / sizeof(Key) per
// block, positions: 8 u16 per block) at row t, with the block index XOR-
// swizzled by the thread index so that the 8 lanes of one 128-byte LDS pass
// hit 8 distinct bank groups; every access is an aligned b128. Small chunks
// (E = 4) use rows of E + 1 scalar entries.
template <typename Key, int E>
struct SortLayout {
    static constexpr bool VEC = E >= 8;
    static constexpr int KPB = 16 / (int)sizeof(Key);  // keys per block
    static constexpr int KB = E / KPB;                  // key blocks per row
    static constexpr int PB = E / 8;                    // position blocks per row
    static constexpr int log2c(int v) { return v <= 1 ? 0 : 1 + log2c(v / 2); }
    static constexpr int KSHIFT = KB >= 8 ? 0 : log2c(8 / KB);
    static constexpr int KMASK = (KB >= 8 ? 8 : KB) - 1;
    static constexpr int PSHIFT = PB >= 8 ? 0 : log2c(8 / (PB > 0 ? PB : 1));
    static constexpr int PMASK = (PB >= 8 ? 8 : (PB > 0 ? PB : 1)) - 1;
    __device__ static int key_block(int t, int b) { return t * KB + (b ^ ((t >> KSHIFT) & KMASK)); }
    __device__ static int pos_block(int t, int b) { return t * PB + (b ^ ((t >> PSHIFT) & PMASK)); }
};

// Bitonic sort of NT * E (key, pos) pairs; thread t holds elements
// [t*E, t*E + E) in registers. Strides below E run in registers; larger
// strides exchange whole chunks with the partner thread through LDS (xk:
// NT * E keys, xp: NT * E u16 positions, both 16-byte aligned).
template <typename Key, int E>
__device__ __forceinline__ void block_sort(Chunk<Key, E> &c, Key *xk, uint16_t *xp, int t)
{
    using L = SortLayout<Key, E>;
    constexpr int N = NT * E;
    const int tbase = t * E;
    c.template local_sort<2>(tbase);
    for (int size = 2 * E; size <= N; size <<= 1) {
        const bool asc = (tbase & size) == 0;
        for (int stride = size >> 1; stride >= E; stride >>= 1) {
            const int partner = t ^ (stride / E);
            const bool keep_min = ((t & (stride / E)) == 0) == asc;
            if (stride / E < 64) {
                // Partner in the same wave: exchange through cross-lane
                // swizzles, no LDS round trip and no workgroup barrier.
                switch (stride / E) {
                case 1: c.template exchange_xor<1>(keep_min); break;
                case 2: c.template exchange_xor<2>(keep_min); break;
                case 4: c.template exchange_xor<4>(keep_min); break;
                case 8: c.template exchange_xor<8>(keep_min); break;
                case 16: c.template exchange_xor<16>(keep_min); break;
                default: c.template exchange_xor<32>(keep_min); break;
                }
                continue;
            }
            if constexpr (L::VEC) {
                uint4 *kb = reinterpret_cast<uint4 *>(xk);
                uint4 *pb = reinterpret_cast<uint4 *>(xp);
                static_for<0, L::KB>([&](auto bc) {
                    constexpr int B = decltype(bc)::value;
                    uint4 v;
                    if constexpr (sizeof(Key) == 4) {
                        v = make_uint4((uint32_t)c.k[4 * B], (uint32_t)c.k[4 * B + 1], (uint32_t)c.k[4 * B + 2],
                                       (uint32_t)c.k[4 * B + 3]);
                    } else {
                        v = make_uint4((uint32_t)c.k[2 * B], (uint32_t)((uint64_t)c.k[2 * B] >> 32),
                                       (uint32_t)c.k[2 * B + 1], (uint32_t)((uint64_t)c.k[2 * B + 1] >> 32));
                    }
                    kb[L::key_block(t, B)] = v;
                });
                static_for<0, L::PB>([&](auto bc) {
                    constexpr int B = decltype(bc)::value;
                    pb[L::pos_block(t, B)] =
                        make_uint4(c.p[8 * B] | (c.p[8 * B + 1] << 16), c.p[8 * B + 2] | (c.p[8 * B + 3] << 16),
                                   c.p[8 * B + 4] | (c.p[8 * B + 5] << 16), c.p[8 * B + 6] | (c.p[8 * B + 7] << 16));
                });
                __syncthreads();
                // Eight elements at a time: one position block and the key
                // blocks holding the same elements, then compare-and-take.
                static_for<0, L::PB>([&](auto bc) {
                    constexpr int B = decltype(bc)::value;
                    const uint4 pv = pb[L::pos_block(partner, B)];
                    const uint32_t w[4] = {pv.x, pv.y, pv.z, pv.w};
                    Key other[8];
                    static_for<0, 8 / L::KPB>([&](auto qc) {
                        constexpr int Q = decltype(qc)::value;
                        const uint4 v = kb[L::key_block(partner, B * (8 / L::KPB) + Q)];
                        if constexpr (sizeof(Key) == 4) {
                            other[4 * Q] = v.x;
                            other[4 * Q + 1] = v.y;
                            other[4 * Q + 2] = v.z;
                            other[4 * Q + 3] = v.w;
                        } else {
                            other[2 * Q] = (Key)v.x | ((Key)v.y << 32);
                            other[2 * Q + 1] = (Key)v.z | ((Key)v.w << 32);
                        }
                    });
                    static_for<0, 8>([&](auto ic) {
                        constexpr int J = decltype(ic)::value;
                        constexpr int I = 8 * B + J;
                        const uint32_t op = (J & 1) ? (w[J / 2] >> 16) : (w[J / 2] & 0xFFFFu);
                        Chunk<Key, E>::take_partner(c.k[I], c.p[I], other[J], op, keep_min);
                    });
                });
            } else {
                static_for<0, E>([&](auto ic) {
                    constexpr int I = decltype(ic)::value;
                    xk[t * (E + 1) + I] = c.k[I];
                    xp[t * (E + 1) + I] = (uint16_t)c.p[I];
                });
                __syncthreads();
                static_for<0, E>([&](auto ic) {
                    constexpr int I = decltype(ic)::value;
                    const Key other = xk[partner * (E + 1) + I];
                    const uint32_t op = xp[partner * (E + 1) + I];
                    Chunk<Key, E>::take_partner(c.k[I], c.p[I], other, op, keep_min);
                });
            }
            __syncthreads();
        }
        c.merge_local(asc);
    }
}

// Layout of the merge sort's exchange. Neighbouring threads access their
// own chunks (E elements) at nearby offsets, which without a remedy land on
// one LDS bank group. Chunks of E >= 32 elements are each followed by 16
// bytes of padding (keys and positions), so thread t's chunk starts 16 * t
// bytes further into the bank rotation and the 16-byte chunk stores stay
// aligned. Smaller chunks (the bracket kernel's sorts, whose LDS budget has
// no room for padding) keep the flat layout; an XOR swizzle there measured
// slower (address arithmetic under that kernel's register cap).
template <typename Key, int E>
struct MergeLayout {
    static constexpr int KPAD = E >= 32 ? 16 / (int)sizeof(Key) : 0;  // keys of padding per chunk
    static constexpr int PPAD = E >= 32 ? 8 : 0;                       // u16 positions of padding per chunk
    static constexpr int log2c(int v) { return v <= 1 ? 0 : 1 + log2c(v / 2); }
    static constexpr int SHIFT = log2c(E);
    __device__ static int key(int i) { return i + KPAD * (i >> SHIFT); }
    __device__ static int pos(int i) { return i + PPAD * (i >> SHIFT); }
    // Start of the position array (16-byte aligned) and the bytes of both.
    static constexpr size_t key_bytes() { return (size_t)NT * (E + KPAD) * sizeof(Key); }
    static constexpr size_t bytes() { return key_bytes() + (size_t)NT * (E + PPAD) * 2; }
};

// Stable merge sort of NT * E (key, pos) pairs; thread t ends with the
// sorted elements [t*E, t*E + E) in registers. Each thread first sorts its
// own E elements (a register bitonic network with compile-time directions);
// then log2(NT) merge levels double the sorted run length. At each level a
// thread produces the E outputs [t*E, t*E + E) of its pair of runs (A, B):
// a binary search on the merge path finds how many of them come from A, and
// a sequential merge reads one key per output from LDS (ties take A first).
// Outputs carry their source index, and positions are gathered afterwards.
// The levels exchange through `region` in the MergeLayout (the caller's
// chunk has been read into registers before the call).
template <typename Key, int E>
__device__ __forceinline__ void merge_sort(Chunk<Key, E> &c, unsigned char *region, int t)
{
    using ML = MergeLayout<Key, E>;
    constexpr int N = NT * E;
    Key *mk = reinterpret_cast<Key *>(region);
    uint16_t *mp = reinterpret_cast<uint16_t *>(region + ML::key_bytes());
    // Logical element i of the exchange.
    auto xk = [&](int i) -> Key { return mk[ML::key(i)]; };
    c.template local_sort<2>(0);
    const int first = t * E;
    for (int run = E; run < N; run <<= 1) {
        static_for<0, E>([&](auto ic) {
            constexpr int I = decltype(ic)::value;
            mk[ML::key(first + I)] = c.k[I];
            mp[ML::pos(first + I)] = (uint16_t)c.p[I];
        });
        __syncthreads();
        const int a0 = first & ~(2 * run - 1);
        const int b0 = a0 + run;
        const int b_end = b0 + run;
        const int diag = first - a0;
        int lo = diag > run ? diag - run : 0;
        int hi = diag < run ? diag : run;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (xk(a0 + mid) <= xk(b0 + diag - mid - 1)) {
                lo = mid + 1;
            } else {
                hi = mid;
            }
        }
        int ia = a0 + lo, ib = b0 + diag - lo;
        Key ka = xk(ia < b0 ? ia : b0 - 1);
        Key kb = xk(ib < b_end ? ib : b_end - 1);
        static_for<0, E>([&](auto ic) {
            constexpr int I = decltype(ic)::value;
            const bool take_b = ib < b_end && (ia >= b0 || kb < ka);
            c.k[I] = take_b ? kb : ka;
            c.p[I] = (uint32_t)(take_b ? ib : ia);
            if constexpr (I + 1 < E) {
                const int next = (take_b ? ib : ia) + 1;
                const Key loaded = xk(next < N ? next : N - 1);
                ia += take_b ? 0 : 1;
                ib += take_b ? 1 : 0;
                ka = take_b ? ka : loaded;
                kb = take_b ? loaded : kb;
            }
        });
        static_for<0, E>([&](auto ic) {
            constexpr int I = decltype(ic)::value;
            c.p[I] = mp[ML::pos((int)c.p[I])];
        });
        __syncthreads();
    }
}

// Bytes of the exchange region: the sort's exchange (xk then xp) and, after
// the sort, keys[N] | guard | pos[N] | guard. 16-byte aligned.
template <typename Key, int E>
constexpr size_t exchange_bytes()
{
    const size_t rows = SortLayout<Key, E>::VEC ? (size_t)NT * E * (sizeof(Key) + 2)
                                                : (size_t)NT * (E + 1) * (sizeof(Key) + 2);
    const size_t sorted = (size_t)NT * E * (sizeof(Key) + 2) + 4 * GUARD;
    const size_t merge = MergeLayout<Key, E>::bytes();
    const size_t most = rows > sorted ? rows : sorted;
    return ((most > merge ? most : merge) + 15) & ~(size_t)15;
}

static_assert(exchange_bytes<uint32_t, 64>() == (size_t)NT * 64 * 6 + (size_t)NT * 32, "exchange layout");
static_assert(exchange_bytes<uint32_t, 16>() == (size_t)NT * 16 * 6 + 4 * GUARD, "exchange layout");
static_assert(exchange_bytes<uint64_t, 4>() == (((size_t)NT * 5 * 10 + 15) & ~(size_t)15), "exchange layout");

// Start of the position half of the exchange (16-byte aligned).
template <typename Key, int E>
__device__ __forceinline__ uint16_t *exchange_positions(Key *xk)
{
    return SortLayout<Key, E>::VEC ? reinterpret_cast<uint16_t *>(xk + NT * E)
                                   : reinterpret_cast<uint16_t *>(xk + NT * (E + 1));
}

// Stage of the plain kernel: loads the support of the tile at (x0, y0, z0)
// through the boundary maps, writes every support sample's code (CODE_OMIT or
// CODE_PENDING) and compacts the retained samples' keys and packed positions
// into xk / xp (count in count_shared; at most n_sort are stored). Each wave takes R support rows per
// batch (H 64-column chunks each): the rows' map entries are read first
// (wave-uniform, scalar loads), then every lane issues all R * H sample
// loads before the first wait, and one LDS atomic reserves the batch's
// slots.
template <typename T, int R, int H>
__device__ __forceinline__ void stage_support(const SlideArgs &a, typename KeyTraits<T>::Key *xk, uint16_t *xp,
                                              uint16_t *codes, int &count_shared, int64_t x0, int64_t y0, int64_t z0,
                                              typename KeyTraits<T>::Key cval_key, int n_sort)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const Key *input = static_cast<const Key *>(a.input);
    int64_t mx[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        mx[h] = h * 64 + lane < a.sx ? a.mapx[x0 + h * 64 + lane] : MAP_PAD;
    }
    for (int row0 = wave * R; row0 < a.rows; row0 += SLIDE_WAVES * R) {
        // Row offsets (mz + my) and row kinds: 0 loaded, 1 constant, 2 padding.
        int64_t roff[R];
        int rkind[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = row0 + r;
            const int z = row >> a.zshift;
            const int y = row & ((1 << a.zshift) - 1);
            const bool live = row < a.rows && y < a.sy;  // padding rows between z planes are skipped
            // Unconditional uniform loads (a dead row reads the tile's first
            // entries, which exist).
            const int64_t mz_at = uniform_entry(a.mapz, z0 + (live ? z : 0));
            const int64_t my_at = uniform_entry(a.mapy, y0 + (live ? y : 0));
            const int64_t mz = live ? mz_at : MAP_PAD;
            const int64_t my = live ? my_at : MAP_PAD;
            const bool pad = mz == MAP_PAD || my == MAP_PAD;
            rkind[r] = pad ? 2 : (mz < 0 || my < 0 ? 1 : 0);
            roff[r] = pad ? 0 : mz + my;
        }
        Key kv[R][H];
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int h = 0; h < H; ++h) {
                // Every lane loads (element 0 stands in for padding and
                // constant samples).
                const bool load = rkind[r] == 0 && mx[h] >= 0;
                kv[r][h] = input[load ? roff[r] + mx[h] : 0];
            }
        }
        int total = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = row0 + r;
            const int y = row & ((1 << a.zshift) - 1);
            const bool live = row < a.rows && y < a.sy;
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const bool pad = rkind[r] == 2 || mx[h] == MAP_PAD;
                const bool cval = !pad && (rkind[r] == 1 || mx[h] < 0);
                kv[r][h] = pad ? KT::SENTINEL : (cval ? cval_key : classify<T>(kv[r][h], a.ignore_inf));
                const int x = h * 64 + lane;
                if (live && x < a.sx) {
                    codes[row * a.sx + x] = kv[r][h] == KT::SENTINEL ? CODE_OMIT : CODE_PENDING;
                }
                total += (int)__popcll(__ballot(kv[r][h] != KT::SENTINEL));
            }
        }
        int base = 0;
        if (lane == 0 && total) {
            base = atomicAdd(&count_shared, total);
        }
        base = __shfl(base, 0);
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const uint64_t bits = __ballot(kv[r][h] != KT::SENTINEL);
                if ((bits >> lane) & 1ull) {
                    const int dst = base + (int)__popcll(bits & ((1ull << lane) - 1ull));
                    if (dst < n_sort) {
                        xk[dst] = kv[r][h];
                        xp[dst] = (uint16_t)(((row0 + r) << 8) | (h * 64 + lane));
                    }
                }
                base += (int)__popcll(bits);
            }
        }
    }
}

// One tile of the plain kernel: stage, sort, rank and slide. Every retained
// sample of the support is ranked; returns false when a rank walk left its
// bounds (never for a consistent tile; reported through the error flag).
// `mask` is the footprint in LDS (nullptr for a rectangle).
template <typename T, int E>
__device__ __forceinline__ bool slide_tile(const SlideArgs &a, int tile, unsigned char *lds, const uint8_t *mask,
                                           int &count_shared, int &failed_shared)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    constexpr int N = NT * E;
    // Layout: [exchange rows NT x (E + 1) keys + u16 positions; after the
    // sort: keys[N] | guard | pos[N] | guard] | codes [s] | hist | mask.
    Key *xk = reinterpret_cast<Key *>(lds);
    uint16_t *xp = exchange_positions<Key, E>(xk);
    Key *keys = reinterpret_cast<Key *>(lds);
    uint16_t *pos0 = reinterpret_cast<uint16_t *>(keys + N) + GUARD;
    uint16_t *pos1 = pos0 - 1;  // 1-based ranks
    uint16_t *codes = reinterpret_cast<uint16_t *>(lds + exchange_bytes<Key, E>());
    uint16_t *hist_all = codes + ((a.s + 7) & ~7);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    Key *output = static_cast<Key *>(a.output);
    const Key cval_key = classify<T>((Key)a.cval_bits, a.ignore_inf);
    const Geometry g{a.kz, a.ky, a.kx, a.zshift, ((((uint32_t)1 << a.zshift) - 1u) << 8) | 0xFFu, mask,
                     a.ky < 128 && a.kx < 128 && a.ty + a.ky <= 257};
    uint16_t *hist = hist_all + wave * a.hist_stride;
    const bool wide = a.kz * a.ky > 255 || (a.mask != nullptr && a.count > 255);
    bool ok = true;
    {
        int rest = tile;
        const int tyi = rest % a.tiles_y;  // tiles run down a column first
        rest /= a.tiles_y;
        const int tx = rest % a.tiles_x;
        const int64_t z0 = rest / a.tiles_x;
        const int64_t x0 = (int64_t)tx * SLIDE_TILE_X;
        const int64_t y0 = (int64_t)tyi * a.ty;
        if (tid == 0) {
            count_shared = 0;
            failed_shared = 0;
        }
        __syncthreads();

        // ---- stage: classify the support, compact the retained samples ----
        if (a.sx <= 128) {
            stage_support<T, 16, 2>(a, xk, xp, codes, count_shared, x0, y0, z0, cval_key, N);
        } else {
            stage_support<T, 4, 4>(a, xk, xp, codes, count_shared, x0, y0, z0, cval_key, N);
        }
        __syncthreads();
        const int retained = count_shared;
        if (retained > N) {
            // Cannot happen: the planner sizes the sort for the whole support.
            __syncthreads();
            return false;
        }

        // ---- count gate: no window with a result, no sort ----
        // Each lane sums the retained samples of its two support columns over
        // the window rows (every plane), slides the sums down its wave's rows
        // and adds the kx columns of its window with a wave prefix scan.
        if (a.count_gate) {
            const int y_first = wave * a.tw;
            const bool has1 = lane + 64 < a.sx;
            auto kept_at = [&](int row, int col) -> int {
                return (int)(codes[row * a.sx + col] != CODE_OMIT);
            };
            int col0 = 0, col1 = 0;
            for (int z = 0; z < a.kz; ++z) {
                for (int dy = 0; dy < a.ky; ++dy) {
                    const int row = z * a.zrows + y_first + dy;
                    col0 += kept_at(row, lane);
                    col1 += has1 ? kept_at(row, lane + 64) : 0;
                }
            }
            const int last = lane + a.kx - 1;  // last support column of this lane's window
            bool any = false;
            for (int y = y_first; y < y_first + a.tw; ++y) {
                if (y > y_first) {
                    for (int z = 0; z < a.kz; ++z) {
                        const int leave = z * a.zrows + y - 1;
                        const int enter = leave + a.ky;
                        col0 += kept_at(enter, lane) - kept_at(leave, lane);
                        col1 += has1 ? kept_at(enter, lane + 64) - kept_at(leave, lane + 64) : 0;
                    }
                }
                int p0 = col0, p1 = col1;
                for (int off = 1; off < 64; off <<= 1) {
                    const int u0 = __shfl_up(p0, off), u1 = __shfl_up(p1, off);
                    p0 += lane >= off ? u0 : 0;
                    p1 += lane >= off ? u1 : 0;
                }
                const int total0 = __shfl(p0, 63);
                const int hi0 = __shfl(p0, last & 63);
                const int hi1 = __shfl(p1, (last - 64) & 63);
                const int before = __shfl(p0, (lane + 63) & 63);
                const int n = (last < 64 ? hi0 : total0 + hi1) - (lane ? before : 0);
                Pick pick;
                any = any || pick_ranks(a.stat, n, a.min_count, a.rank_param, a.q100, pick);
            }
            if (!__syncthreads_or(any ? 1 : 0)) {
                // Every window of the tile lacks a result: canonical NaN.
                const bool col_valid = x0 + lane < a.nx;
                for (int y = y_first; y < y_first + a.tw && y0 + y < a.ny; ++y) {
                    if (col_valid) {
                        output[z0 * a.ny * a.nx + (y0 + y) * a.nx + x0 + lane] = KT::CANONICAL_NAN;
                    }
                }
                __syncthreads();
                return true;
            }
        }

        // ---- sort the retained samples (padded with sentinels to N) ----
        {
            Chunk<Key, E> chunk;
            // Thread tid starts with elements I * NT + tid: lanes read
            // consecutive entries (no LDS bank conflicts), and the sorted
            // result does not depend on the initial assignment.
            static_for<0, E>([&](auto ic) {
                constexpr int I = decltype(ic)::value;
                const int gi = I * NT + tid;
                chunk.k[I] = gi < retained ? xk[gi] : KT::SENTINEL;
                chunk.p[I] = gi < retained ? xp[gi] : POS_NONE;
            });
            __syncthreads();
            merge_sort<Key, E>(chunk, reinterpret_cast<unsigned char *>(xk), tid);
            __syncthreads();
            static_for<0, E>([&](auto ic) {
                constexpr int I = decltype(ic)::value;
                keys[tid * E + I] = chunk.k[I];
                pos0[tid * E + I] = (uint16_t)chunk.p[I];
                if (chunk.p[I] != POS_NONE) {
                    codes[(chunk.p[I] >> 8) * a.sx + (chunk.p[I] & 0xFFu)] = (uint16_t)(tid * E + I + 1);
                }
            });
            if (tid < GUARD) {
                pos0[tid - GUARD] = POS_NONE;
                pos0[N + tid] = POS_NONE;
            }
        }
        __syncthreads();

        // ---- per-lane slide ----
        const int y_first = wave * a.tw;
        if (y0 + y_first < a.ny) {
            const bool col_valid = x0 + lane < a.nx;
            auto target_of = [&](int n) { return (int)target_rank(a.stat, n, a.min_count, a.rank_param, a.q100); };

            // Initial cursor: two levels of 16-bin histograms over the ranks.
            int m = 1, c = 0, n = 0;
            bool failed = false;
            {
                int shift1 = 0;
                while ((BINS << shift1) <= retained) {
                    ++shift1;
                }
                int counts[BINS], below = 0, above = 0;
                if (wide) {
                    window_bins<Bins16>(counts, below, above, 1, shift1, codes, hist, a, lane, y_first);
                } else {
                    window_bins<Bins8>(counts, below, above, 1, shift1, codes, hist, a, lane, y_first);
                }
                n = below + above;
#pragma unroll
                for (int b = 0; b < BINS; ++b) {
                    n += counts[b];
                }
                const int k = target_of(n);
                int bin = -1, base1 = below;
                if (k >= 0) {
                    if (k < below || k >= n - above) {
                        failed = true;  // the target lies outside the bracket
                    } else {
                        int run = below;
#pragma unroll
                        for (int b = 0; b < BINS; ++b) {
                            if (bin < 0 && k < run + counts[b]) {
                                bin = b;
                                base1 = run;
                            }
                            run += counts[b];
                        }
                    }
                }
                c = below;  // consistent with m = 1 for lanes without a target
                int bmin = BINS, bmax = -1;
                for (int b = 0; b < BINS; ++b) {
                    if (__ballot(bin == b)) {
                        bmin = b < bmin ? b : bmin;
                        bmax = b;
                    }
                }
                if (bmax >= 0 && shift1 > 2) {
                    const int lo2 = 1 + (bmin << shift1);
                    const int span = (bmax + 1 - bmin) << shift1;
                    int shift2 = 0;
                    while ((BINS << shift2) < span) {
                        ++shift2;
                    }
                    int counts2[BINS], below2 = 0, above2 = 0;
                    if (wide) {
                        window_bins<Bins16>(counts2, below2, above2, lo2, shift2, codes, hist, a, lane, y_first);
                    } else {
                        window_bins<Bins8>(counts2, below2, above2, lo2, shift2, codes, hist, a, lane, y_first);
                    }
                    if (bin >= 0) {
                        int run = below2, bin2 = -1, base2 = below2, in2 = 0;
#pragma unroll
                        for (int b = 0; b < BINS; ++b) {
                            if (bin2 < 0 && k < run + counts2[b]) {
                                bin2 = b;
                                base2 = run;
                                in2 = counts2[b];
                            }
                            run += counts2[b];
                        }
                        if (bin2 < 0) {
                            bin2 = 0;
                            base2 = below2;
                            in2 = 0;
                        }
                        // Start from the bin's end when the target is nearer to
                        // it (c there counts the whole bin); the end must stay
                        // within the ranked samples plus one.
                        const int m_end = lo2 + ((bin2 + 1) << shift2);
                        const bool from_end = in2 > 0 && m_end <= retained + 1 && base2 + in2 - k < k - base2;
                        m = from_end ? m_end : lo2 + (bin2 << shift2);
                        c = from_end ? base2 + in2 : base2;
                    }
                } else if (bin >= 0) {
                    m = 1 + (bin << shift1);
                    c = base1;
                }
            }

            for (int y = y_first; y < y_first + a.tw && y0 + y < a.ny && !failed; ++y) {
                if (y > y_first) {
                    int dn = 0, dc = 0;
                    if (a.mask == nullptr) {
                        // Rectangle: the top row of the previous window leaves,
                        // the bottom row of this one enters (every z plane).
                        for (int z = 0; z < a.kz; ++z) {
                            const uint16_t *top = codes + (z * a.zrows + y - 1) * a.sx + lane;
                            const uint16_t *bottom = top + a.ky * a.sx;
#pragma unroll 8
                            for (int dx = 0; dx < a.kx; ++dx) {
                                const int cr = top[dx];
                                const int ca = bottom[dx];
                                dn += (int)(ca != CODE_OMIT) - (int)(cr != CODE_OMIT);
                                dc += (int)(ca < m) - (int)(cr < m);
                            }
                        }
                    } else {
                        const int rbase = (y - 1) * a.sx + lane;
                        const int abase = y * a.sx + lane;
#pragma unroll 8
                        for (int e = 0; e < a.nedge; ++e) {
                            const int cr = codes[rbase + a.edge_remove[e]];
                            const int ca = codes[abase + a.edge_add[e]];
                            dn += (int)(ca != CODE_OMIT) - (int)(cr != CODE_OMIT);
                            dc += (int)(ca < m) - (int)(cr < m);
                        }
                    }
                    n += dn;
                    c += dc;
                }
                Key out = KT::CANONICAL_NAN;
                Pick pick;
                if (pick_ranks(a.stat, n, a.min_count, a.rank_param, a.q100, pick)) {
                    const int k = (int)pick.hi;
                    const uint32_t base = ((uint32_t)y << 8) | (uint32_t)lane;
                    if (!walk_to(m, c, k, pos1, retained, g, base)) {
                        failed = true;
                        break;
                    }
                    if (pick.pair) {
                        const int m2 = previous_member(m, pos1, g, base);
                        if (m2 < 1) {
                            failed = true;
                            break;
                        }
                        out = pair_bits<T>(a.stat, keys[m2 - 1], keys[m - 1], pick.frac);
                    } else {
                        out = KT::unkey(keys[m - 1]);
                    }
                }
                if (col_valid) {
                    output[z0 * a.ny * a.nx + (y0 + y) * a.nx + x0 + lane] = out;
                }
            }
            if (failed) {
                failed_shared = 1;
            }
        }
        __syncthreads();
        // Only thread 0 reads the flag: it is the one that resets it for the
        // next tile, so no other thread can race with that reset.
        ok = tid != 0 || failed_shared == 0;
    }
    return ok;
}

// Next tile index in [0, ntiles) for a persistent bracket workgroup, or -1
// when every chunk is drained (called by one thread). The tiles are split into
// BRACKET_QUEUES contiguous chunks with a counter each: workgroup b starts on
// chunk b % 8 (the XCD it runs on, so vertical neighbours share their halo
// rows in that XCD's L2) and moves on to the next chunk when its own is
// drained. Taking tiles on demand keeps slow tiles from setting the kernel's
// end.
__device__ inline int take_tile(int32_t *queue, int ntiles)
{
    const int chunk = (ntiles + BRACKET_QUEUES - 1) / BRACKET_QUEUES;
    for (int probe = 0; probe < BRACKET_QUEUES; ++probe) {
        const int q = (int)((blockIdx.x + probe) % BRACKET_QUEUES);
        const int size = min(chunk, ntiles - q * chunk);
        if (size > 0 && __hip_atomic_load(&queue[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < size) {
            const int i = atomicAdd(&queue[q], 1);
            if (i < size) {
                return q * chunk + i;
            }
        }
    }
    return -1;
}

// Plain kernel on one image: one workgroup per tile, persistent over tiles
// (all tiles, or tile_list[0 .. *tile_count) for fallback tiles).
template <typename T, int E>
__global__ void __launch_bounds__(NT, 1) k_slide(SlideArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int count_shared;
    __shared__ int failed_shared;
    using Key = typename KeyTraits<T>::Key;
    const int tid = threadIdx.x;
    uint8_t *mask = nullptr;
    if (a.mask) {
        uint16_t *hist_all = reinterpret_cast<uint16_t *>(lds + exchange_bytes<Key, E>()) + ((a.s + 7) & ~7);
        mask = reinterpret_cast<uint8_t *>(hist_all + SLIDE_WAVES * a.hist_stride);
        const int mbytes = a.kz * a.ky * a.kx;
        for (int q = tid; q < mbytes; q += NT) {
            mask[q] = a.mask[q];
        }
    }
    const int ntiles = a.tile_count ? *a.tile_count : a.tiles_x * a.tiles_y * (int)a.nz;
    // XCD-aware order: block b runs on XCD b % 8, so slot L = b + i * grid
    // maps to tile (L % 8) * chunk + L / 8 and each XCD walks a contiguous
    // run of tiles (vertical neighbours share their halo rows in its L2). A
    // fixed order measured faster than on-demand tiles for this kernel.
    const int chunk = (ntiles + 7) / 8;
    for (int slot = blockIdx.x; slot < 8 * chunk; slot += gridDim.x) {
        const int t = (slot % 8) * chunk + slot / 8;
        if (t >= ntiles) {
            continue;
        }
        const int tile = a.tile_list ? a.tile_list[t] : t;
        if (!slide_tile<T, E>(a, tile, lds, mask, count_shared, failed_shared) && tid == 0) {
            atomicOr(a.error_flag, 1);
        }
    }
}

// Plain kernel on a batch of 2-D images (rectangular windows): tiles come
// from a work queue of (image, tile) pairs ordered by decreasing cost; each
// image has its own geometry, maps and window. Inputs and outputs of image i
// start at offsets[i] elements from the batch buffers.
template <typename T, int E>
__global__ void __launch_bounds__(NT, 1) k_slide_batch(const SlideArgs *images, const int64_t *offsets,
                                                       const int32_t *work, int nwork, int32_t *counter,
                                                       const void *input, void *output, int32_t *error_flag)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int count_shared;
    __shared__ int failed_shared;
    __shared__ int next_shared;
    using Key = typename KeyTraits<T>::Key;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) {
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
        SlideArgs a = images[image];
        a.input = static_cast<const Key *>(input) + offsets[image];
        a.output = static_cast<Key *>(output) + offsets[image];
        if (!slide_tile<T, E>(a, work[2 * w + 1], lds, nullptr, count_shared, failed_shared) && tid == 0) {
            atomicOr(error_flag, 1);
        }
    }
}

// Bit masks of one support row of a bracket tile (bit x <-> column x of the
// support; rows are at most 128 samples wide).
struct RowMasks {
    uint64_t kept[2];    // retained samples
    uint64_t below[2];   // retained samples below the tile's bracket
    uint64_t inside[2];  // samples inside the bracket (ranked)
};

// Set bits of a 128-bit row mask below column `col` (col < 128).
__device__ __forceinline__ int inside_before(const uint64_t w[2], uint32_t col)
{
    const uint64_t low = col >= 64 ? w[0] : w[0] & ((1ull << col) - 1ull);
    const uint64_t high = col > 64 ? w[1] & ((1ull << (col - 64)) - 1ull) : 0ull;
    return (int)__popcll(low) + (int)__popcll(high);
}

// Sorts the first `retained` staged (key, position) pairs (xk/xp, flat) with
// a network of NT * E entries padded with sentinels, then writes the sorted
// keys and positions from the start of the region (keys[NT*E], guard,
// pos[NT*E], guard) and the 1-based rank of every sorted sample into codes,
// at the sample's compact index (row_offset[row] + the row's inside samples
// left of its column).
template <typename Key, int E>
__device__ __forceinline__ void sort_inside(Key *xk, uint16_t *xp, uint16_t *codes, const uint16_t *row_offset,
                                            const RowMasks *rowmasks, int retained, int tid)
{
    using KT = KeyTraits<std::conditional_t<sizeof(Key) == 4, float, double>>;
    constexpr int N = NT * E;
    Key *keys = xk;
    uint16_t *pos0 = reinterpret_cast<uint16_t *>(keys + N) + GUARD;
    Chunk<Key, E> chunk;
    // Thread tid starts with elements I * NT + tid (consecutive entries per
    // lane; the sorted result does not depend on the initial assignment).
    static_for<0, E>([&](auto ic) {
        constexpr int I = decltype(ic)::value;
        const int gi = I * NT + tid;
        chunk.k[I] = gi < retained ? xk[gi] : KT::SENTINEL;
        chunk.p[I] = gi < retained ? xp[gi] : POS_NONE;
    });
    __syncthreads();
    merge_sort<Key, E>(chunk, reinterpret_cast<unsigned char *>(xk), tid);
    __syncthreads();
    static_for<0, E>([&](auto ic) {
        constexpr int I = decltype(ic)::value;
        keys[tid * E + I] = chunk.k[I];
        pos0[tid * E + I] = (uint16_t)chunk.p[I];
        if (chunk.p[I] != POS_NONE) {
            const uint32_t row = chunk.p[I] >> 8, col = chunk.p[I] & 0xFFu;
            codes[row_offset[row] + inside_before(rowmasks[row].inside, col)] = (uint16_t)(tid * E + I + 1);
        }
    });
    if (tid < GUARD) {
        pos0[tid - GUARD] = POS_NONE;
        pos0[N + tid] = POS_NONE;
    }
}

// Bits [lane, lane + kx) of a 128-bit row mask, shifted down to bit 0.
__device__ __forceinline__ uint64_t window_bits(const uint64_t w[2], int lane, uint64_t kxmask)
{
    const uint64_t low = w[0] >> lane;
    const uint64_t high = lane ? (w[1] << (64 - lane)) : 0ull;
    return (low | high) & kxmask;
}

// Set bits of a 64-bit wave mask below this lane.
__device__ __forceinline__ int lanes_below(uint64_t w)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(w >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)w, 0u));
}

// (entering samples with a code below m) - (leaving samples with a code
// below m), over the n_in codes at code_in and the n_out codes at code_out.
// Each round reads up to EDGE_BATCH codes of both runs before the first
// compare, so the LDS latency is paid once per round rather than once per
// sample.
constexpr int EDGE_BATCH = 4;
__device__ __forceinline__ int edge_codes_below(const uint16_t *code_in, int n_in, const uint16_t *code_out,
                                                int n_out, int m)
{
    int delta = 0;
    const int n = n_in > n_out ? n_in : n_out;
    for (int k = 0; k < n; k += EDGE_BATCH) {
        int v_in[EDGE_BATCH], v_out[EDGE_BATCH];
#pragma unroll
        for (int q = 0; q < EDGE_BATCH; ++q) {
            v_in[q] = code_in[k + q < n_in ? k + q : 0];
            v_out[q] = code_out[k + q < n_out ? k + q : 0];
        }
#pragma unroll
        for (int q = 0; q < EDGE_BATCH; ++q) {
            delta += (k + q < n_in && v_in[q] < m) ? 1 : 0;
            delta -= (k + q < n_out && v_out[q] < m) ? 1 : 0;
        }
    }
    return delta;
}

// Bracket kernel for 2-D rectangular windows (support rows <= 128 samples).
// One workgroup per tile, persistent over tiles. A tile runs one or more
// passes, each over one key range (the first is the tile's learned bracket
// [lo, hi]). In a pass, support samples are classified against the range:
// only inside samples are sorted and ranked; per support row, bit masks
// record the retained, below and inside samples. Each lane keeps the
// retained / below / inside counts of its window as popcounts of its
// window's bits, and a cursor (m, c) over the inside ranks where c counts the
// window's below samples plus its inside samples ranked under m. A row step
// therefore touches only the inside samples of the leaving and entering
// rows. A window whose result ranks lie in the range is computed; one whose
// ranks lie in other ranges waits for their passes (a median or percentile
// pair split between two ranges parks its first element in the output).
// More passes follow when the range holds more samples than the sort (the
// range is split at its key midpoint) or when a target lies below or above
// every range queued so far (an adjacent range doubles their span). A tile
// that would exceed MAX_PASSES passes, or whose walk fails, goes to the
// fallback list, where the plain kernel recomputes it.
template <typename T, int E>
// Workgroups per CU: at a 4096-entry sort four for float32 (37 KB of LDS each
// with compact codes; 128 VGPRs) and three for float64 (52 KB; 168 VGPRs);
// at 2048 entries or fewer six for float32 and four for float64. The
// register budgets spill a few sort registers (96 B/lane for float32 and 312
// for float64 at 4096) but measured faster than one workgroup fewer: C5
// median 5.06 -> 4.46 ms, 31x31 float64 median 1.71 -> 1.48 ms, C5 rank 397
// at 2048 entries 3.20 -> 3.04 ms, 31x31 float64 rank 154 1.14 -> 1.07 ms.
__global__ void __launch_bounds__(NT, E >= 32 ? 1 : (sizeof(T) == 4 ? (E <= 8 ? 6 : 4) : (E <= 8 ? 4 : 3)))
    k_bracket(SlideArgs a)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    constexpr int N = NT * E;
    constexpr int R = 8;  // support rows a wave loads per batch
    extern __shared__ __align__(16) unsigned char lds[];
    Key *xk = reinterpret_cast<Key *>(lds);
    uint16_t *xp = exchange_positions<Key, E>(xk);
    Key *keys = reinterpret_cast<Key *>(lds);
    // Rank codes of the inside samples only, row after row in column order:
    // row r's start at row_offset[r] (N entries hold every pass's inside
    // samples, since a pass never sorts more than N).
    uint16_t *codes = reinterpret_cast<uint16_t *>(lds + exchange_bytes<Key, E>());
    uint16_t *row_offset = codes + N;
    RowMasks *rowmasks = reinterpret_cast<RowMasks *>(row_offset + ((a.rows + 1 + 7) & ~7));
    __shared__ int count_shared;
    // Pass bookkeeping of the current tile (written by thread 0 between
    // barriers): pending key ranges (a stack, processed last-in first-out),
    // the union of every range queued so far, passes started, the OR of the
    // waves' pass flags and the tile's failure.
    __shared__ Key pend_lo[MAX_PASSES], pend_hi[MAX_PASSES];
    __shared__ Key union_lo, union_hi;
    __shared__ int npend_shared, passes_shared, flags_shared, tile_failed;
    __shared__ int next_tile;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wave_u = __builtin_amdgcn_readfirstlane(wave);
    const int ntiles = a.tile_count ? *a.tile_count : a.tiles_x * a.tiles_y;
    const Key *input = static_cast<const Key *>(a.input);
    Key *output = static_cast<Key *>(a.output);
    const Key cval_key = classify<T>((Key)a.cval_bits, a.ignore_inf);
    const Geometry g{1, a.ky, a.kx, 8, 0xFFFFu, nullptr, a.ky < 128 && a.kx < 128 && a.ty + a.ky <= 257};
    const uint64_t kxmask = a.kx >= 64 ? ~0ull : ((1ull << a.kx) - 1ull);
    const int64_t mz = uniform_entry(a.mapz, 0);
    const uint64_t lanemask_lt = (1ull << lane) - 1ull;

    // Tiles are taken on demand (take_tile): a tile that needs several
    // passes delays only the workgroup that took it.
    for (;;) {
        if (tid == 0) {
            next_tile = take_tile(a.tile_queue, ntiles);
        }
        __syncthreads();
        const int t = next_tile;
        if (t < 0) {
            break;
        }
        const int tile = a.tile_list ? a.tile_list[t] : t;
        const int tyi = tile % a.tiles_y;
        const int tx = tile / a.tiles_y;
        const int64_t x0 = (int64_t)tx * SLIDE_TILE_X;
        const int64_t y0 = (int64_t)tyi * a.ty;
        if (tid == 0) {
            pend_lo[0] = static_cast<const Key *>(a.bracket_lo)[tile];
            pend_hi[0] = static_cast<const Key *>(a.bracket_hi)[tile];
            union_lo = pend_lo[0];
            union_hi = pend_hi[0];
            npend_shared = 1;
            passes_shared = 0;
            tile_failed = 0;
        }
        __syncthreads();
        const int64_t mx0 = a.mapx[x0 + lane];
        const int64_t mx1 = lane + 64 < a.sx ? a.mapx[x0 + 64 + lane] : MAP_PAD;
        // Bit r: the output of this lane's row y_first + r holds the key of
        // one element of a pair whose other element lies in another range.
        uint64_t halves = 0;

        // ---- passes: one key range each, until none is pending ----
        for (;;) {
        const int np = npend_shared;
        if (np == 0 || tile_failed) {
            break;
        }
        const Key lo_key = pend_lo[np - 1], hi_key = pend_hi[np - 1];
        const Key ulo = union_lo, uhi = union_hi;
        const int pass = passes_shared;
        __syncthreads();
        if (tid == 0) {
            npend_shared = np - 1;
            passes_shared = pass + 1;
            count_shared = 0;
            flags_shared = 0;
        }
        __syncthreads();

        // ---- stage: R rows per wave per batch, loads first ----
        for (int row0 = wave_u * R; row0 < a.rows; row0 += SLIDE_WAVES * R) {
            // Row map entries first (wave-uniform, scalar loads), then all
            // R * 2 sample loads of the batch before the first wait.
            int64_t my[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int64_t at = uniform_entry(a.mapy, y0 + (row0 + r < a.rows ? row0 + r : 0));
                my[r] = row0 + r < a.rows ? at : MAP_PAD;
            }
            Key kv[R][2];
#pragma unroll
            for (int r = 0; r < R; ++r) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int64_t mx = h ? mx1 : mx0;
                    const bool load = my[r] >= 0 && mx >= 0;  // every lane loads; element 0 stands in
                    kv[r][h] = input[load ? mz + my[r] + mx : 0];
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int64_t mx = h ? mx1 : mx0;
                    const bool pad = my[r] == MAP_PAD || mx == MAP_PAD;
                    const bool cval = !pad && (my[r] < 0 || mx < 0);
                    kv[r][h] = pad ? KT::SENTINEL : (cval ? cval_key : classify<T>(kv[r][h], a.ignore_inf));
                }
            }
            uint64_t in_bits[R][2];
            int total = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const Key key = kv[r][h];
                    const bool kept = key != KT::SENTINEL;
                    const uint64_t bk = __ballot(kept);
                    const uint64_t bb = __ballot(kept && key < lo_key);
                    const uint64_t bi = __ballot(kept && key >= lo_key && key <= hi_key);
                    if (lane == 0 && row0 + r < a.rows) {
                        rowmasks[row0 + r].kept[h] = bk;
                        rowmasks[row0 + r].below[h] = bb;
                        rowmasks[row0 + r].inside[h] = bi;
                    }
                    in_bits[r][h] = bi;
                    total += (int)__popcll(bi);
                }
            }
            int base = 0;
            if (lane == 0 && total) {
                base = atomicAdd(&count_shared, total);
            }
            base = __shfl(base, 0);
#pragma unroll
            for (int r = 0; r < R; ++r) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const uint64_t bi = in_bits[r][h];
                    if ((bi >> lane) & 1ull) {
                        const int dst = base + (int)__popcll(bi & lanemask_lt);
                        if (dst < N) {
                            xk[dst] = kv[r][h];
                            xp[dst] = (uint16_t)(((row0 + r) << 8) | (h * 64 + lane));
                        }
                    }
                    base += (int)__popcll(bi);
                }
            }
        }
        __syncthreads();
        const int retained = count_shared;
        if (retained > N) {
            // More samples in the range than the sort holds: split the range
            // at its key midpoint (lower half first), within the pass budget.
            if (tid == 0) {
                // Passes started (pass + 1) plus pending after the split
                // (np + 1) stay within the budget.
                if (pass + np + 2 <= MAX_PASSES && hi_key > lo_key) {
                    const Key mid = lo_key + (hi_key - lo_key) / 2;
                    pend_lo[np - 1] = mid + 1;
                    pend_hi[np - 1] = hi_key;
                    pend_lo[np] = lo_key;
                    pend_hi[np] = mid;
                    npend_shared = np + 1;
                } else {
                    tile_failed = 1;
                }
            }
            __syncthreads();
            continue;
        }

        // ---- row offsets of the compact codes (wave 0; read after the sort's barriers) ----
        if (wave == 0) {
            // Lane l sums the inside samples of rows 4l .. 4l + 3, then a wave
            // prefix scan gives every row's start.
            int counts[4], total = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = 4 * lane + j;
                counts[j] = row < a.rows ? (int)__popcll(rowmasks[row].inside[0]) + (int)__popcll(rowmasks[row].inside[1])
                                         : 0;
                total += counts[j];
            }
            int scan = total;
            for (int off = 1; off < 64; off <<= 1) {
                const int u = __shfl_up(scan, off);
                scan += lane >= off ? u : 0;
            }
            int start = scan - total;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = 4 * lane + j;
                if (row <= a.rows) {
                    row_offset[row] = (uint16_t)start;
                }
                start += counts[j];
            }
        }

        // ---- sort and rank the inside samples (smallest network that holds them) ----
        int nsort = N;
        if constexpr (E >= 16) {
            if (retained <= NT * (E / 4)) {
                nsort = NT * (E / 4);
            }
        }
        if constexpr (E >= 8) {
            if (nsort == N && retained <= NT * (E / 2)) {
                nsort = NT * (E / 2);
            }
        }
        if constexpr (E >= 16) {
            if (nsort == NT * (E / 4)) {
                sort_inside<Key, E / 4>(xk, xp, codes, row_offset, rowmasks, retained, tid);
            }
        }
        if constexpr (E >= 8) {
            if (nsort == NT * (E / 2)) {
                sort_inside<Key, E / 2>(xk, xp, codes, row_offset, rowmasks, retained, tid);
            }
        }
        if (nsort == N) {
            sort_inside<Key, E>(xk, xp, codes, row_offset, rowmasks, retained, tid);
        }
        __syncthreads();
        const uint16_t *pos1 = reinterpret_cast<uint16_t *>(keys + nsort) + GUARD - 1;  // 1-based ranks

        // ---- per-lane slide ----
        const int y_first = wave * a.tw;
        if (y0 + y_first < a.ny) {
            const bool col_valid = x0 + lane < a.nx;
            int n = 0, below = 0, inside = 0;
            for (int dy = 0; dy < a.ky; ++dy) {
                const RowMasks &rm = rowmasks[y_first + dy];
                n += (int)__popcll(window_bits(rm.kept, lane, kxmask));
                below += (int)__popcll(window_bits(rm.below, lane, kxmask));
                inside += (int)__popcll(window_bits(rm.inside, lane, kxmask));
            }
            // Start cursor: START candidate ranks spread over the lanes'
            // estimated targets (each lane's target position relative to its
            // bracket, applied to the tile's ranks). For every candidate m_j,
            // per-column counts of inside ranks below m_j over the window rows
            // (two candidates per 16-bit pair), summed over each lane's kx
            // columns with a wave prefix scan, give the exact cursor count
            // c_j; each lane starts from the candidate nearest its target.
            int m = 1, c = below;
            {
                int est = 0;
                bool have = false;
                if (n >= a.min_count && n > 0 && inside > 0) {
                    const int k0 = (int)target_rank(a.stat, n, a.min_count, a.rank_param, a.q100);
                    if (k0 >= 0 && k0 < n && k0 >= below && k0 < below + inside) {
                        const float rel = ((float)(k0 - below) + 0.5f) / (float)inside;
                        est = 1 + (int)(rel * (float)retained);
                        est = est < 1 ? 1 : (est > retained ? retained : est);
                        have = true;
                    }
                }
                int lo_est = have ? est : retained, hi_est = have ? est : 1;
                for (int off = 32; off > 0; off >>= 1) {
                    const int l = __shfl_xor(lo_est, off), h = __shfl_xor(hi_est, off);
                    lo_est = l < lo_est ? l : lo_est;
                    hi_est = h > hi_est ? h : hi_est;
                }
                if (lo_est <= hi_est) {
                    // Candidates: midpoints of START equal parts of [lo_est, hi_est].
                    u16x2 thr[START / 2];
#pragma unroll
                    for (int q = 0; q < START / 2; ++q) {
                        const int span = hi_est - lo_est;
                        const int t0 = lo_est + (span * (4 * q + 1)) / (2 * START);
                        const int t1 = lo_est + (span * (4 * q + 3)) / (2 * START);
                        thr[q] = u16x2{(unsigned short)t0, (unsigned short)t1};
                    }
                    u16x2 col0[START / 2], col1[START / 2];
#pragma unroll
                    for (int q = 0; q < START / 2; ++q) {
                        col0[q] = (u16x2)(unsigned short)0;
                        col1[q] = (u16x2)(unsigned short)0;
                    }
                    const bool has1 = lane + 64 < a.sx;
                    // Samples that are not inside get 0x7FFF, above every rank,
                    // so a 16-bit difference's sign is `code < m_j`. The masks
                    // and row offsets of four rows are read before any is
                    // used, then their codes (columns lane and lane + 64 sit
                    // at the row offset plus the row's inside samples left of
                    // them; for a sample that is not inside that index reads
                    // a neighbouring code, which is replaced).
                    auto add_row = [&](uint64_t in0, uint64_t in1, unsigned short c0, unsigned short c1) {
                        const unsigned short e0 = ((in0 >> lane) & 1ull) ? c0 : (unsigned short)0x7FFF;
                        const unsigned short e1 = (has1 && ((in1 >> lane) & 1ull)) ? c1 : (unsigned short)0x7FFF;
                        const u16x2 v0 = (u16x2)e0, v1 = (u16x2)e1;
#pragma unroll
                        for (int q = 0; q < START / 2; ++q) {
                            col0[q] += (u16x2)(v0 - thr[q]) >> (u16x2)15;
                            col1[q] += (u16x2)(v1 - thr[q]) >> (u16x2)15;
                        }
                    };
                    int dy = 0;
                    for (; dy + 4 <= a.ky; dy += 4) {
                        uint64_t in0[4], in1[4];
                        int off[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = y_first + dy + r;
                            in0[r] = rowmasks[row].inside[0];
                            in1[r] = rowmasks[row].inside[1];
                            off[r] = row_offset[row];
                        }
                        unsigned short c0[4], c1[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int i0 = off[r] + lanes_below(in0[r]);
                            c0[r] = codes[i0];
                            c1[r] = codes[off[r] + (int)__popcll(in0[r]) + lanes_below(in1[r])];
                        }
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            add_row(in0[r], in1[r], c0[r], c1[r]);
                        }
                    }
                    for (; dy < a.ky; ++dy) {
                        const int row = y_first + dy;
                        const uint64_t w0 = rowmasks[row].inside[0], w1 = rowmasks[row].inside[1];
                        const int off = row_offset[row];
                        add_row(w0, w1, codes[off + lanes_below(w0)],
                                codes[off + (int)__popcll(w0) + lanes_below(w1)]);
                    }
                    // Inclusive prefix over columns 0..127 (two candidates per
                    // register; counts stay below 2^16).
                    const int last = lane + a.kx - 1;  // last column of this lane's window
                    int cj[START];
#pragma unroll
                    for (int q = 0; q < START / 2; ++q) {
                        uint32_t p0 = __builtin_bit_cast(uint32_t, col0[q]);
                        uint32_t p1 = __builtin_bit_cast(uint32_t, col1[q]);
                        for (int off = 1; off < 64; off <<= 1) {
                            const uint32_t u0 = (uint32_t)__shfl_up((int)p0, off);
                            const uint32_t u1 = (uint32_t)__shfl_up((int)p1, off);
                            p0 += lane >= off ? u0 : 0u;  // no carry: both halves stay below 2^16
                            p1 += lane >= off ? u1 : 0u;
                        }
                        const uint32_t total0 = (uint32_t)__shfl((int)p0, 63);
                        const uint32_t hi0 = (uint32_t)__shfl((int)p0, last & 63);
                        const uint32_t hi1 = (uint32_t)__shfl((int)p1, (last - 64) & 63);
                        const uint32_t before = (uint32_t)__shfl((int)p0, (lane + 63) & 63);
                        const uint32_t upto = last < 64 ? hi0 : total0 + hi1;
                        const uint32_t win = upto - (lane ? before : 0u);
                        cj[2 * q] = (int)(win & 0xFFFFu);
                        cj[2 * q + 1] = (int)(win >> 16);
                    }
                    // Nearest candidate by window count: the last one at or
                    // below the target, or the next one if that is closer.
                    const int k0 = have ? (int)target_rank(a.stat, n, a.min_count, a.rank_param, a.q100) : 0;
                    int pick = 0;
#pragma unroll
                    for (int j = 1; j < START; ++j) {
                        pick = below + cj[j] <= k0 ? j : pick;
                    }
                    int best_c = cj[0], best_j = 0;
#pragma unroll
                    for (int j = 0; j < START; ++j) {
                        const bool take = j == pick || (j == pick + 1 && below + cj[j] - k0 < k0 - below - cj[pick]);
                        best_c = take ? cj[j] : best_c;
                        best_j = take ? j : best_j;
                    }
                    const u16x2 tq = thr[0];
                    int mj = tq.x;
#pragma unroll
                    for (int q = 0; q < START / 2; ++q) {
                        mj = best_j == 2 * q ? (int)thr[q].x : mj;
                        mj = best_j == 2 * q + 1 ? (int)thr[q].y : mj;
                    }
                    m = mj;
                    c = below + best_c;
                }
            }
            bool failed = false, need_low = false, need_high = false;
            for (int y = y_first; y < y_first + a.tw && y0 + y < a.ny && !failed; ++y) {
                if (y > y_first) {
                    const RowMasks &leave = rowmasks[y - 1];
                    const RowMasks &enter = rowmasks[y - 1 + a.ky];
                    const int db = (int)__popcll(window_bits(enter.below, lane, kxmask)) -
                                   (int)__popcll(window_bits(leave.below, lane, kxmask));
                    n += (int)__popcll(window_bits(enter.kept, lane, kxmask)) -
                         (int)__popcll(window_bits(leave.kept, lane, kxmask));
                    below += db;
                    c += db;
                    // The window's inside samples of a row are consecutive
                    // in the compact codes: from the row offset plus the
                    // row's inside samples left of the lane's column.
                    const int n_out = (int)__popcll(window_bits(leave.inside, lane, kxmask));
                    const int n_in = (int)__popcll(window_bits(enter.inside, lane, kxmask));
                    inside += n_in - n_out;
                    const uint16_t *code_out = codes + row_offset[y - 1] + lanes_below(leave.inside[0]);
                    const uint16_t *code_in = codes + row_offset[y - 1 + a.ky] + lanes_below(enter.inside[0]);
                    c += edge_codes_below(code_in, n_in, code_out, n_out, m);
                }
                Pick pick;
                if (!pick_ranks(a.stat, n, a.min_count, a.rank_param, a.q100, pick)) {
                    if (col_valid) {
                        output[(y0 + y) * a.nx + x0 + lane] = KT::CANONICAL_NAN;
                    }
                    continue;
                }
                // Ranks the result reads: k, and k - 1 for a pair. Each lies
                // below this pass's range (rank < below), in it, or above it;
                // a rank outside the union of the tile's ranges asks for an
                // adjacent range.
                const int k = (int)pick.hi;
                const int first = pick.pair ? k - 1 : k;
                const bool low_in = first >= below && first < below + inside;
                const bool high_in = k >= below && k < below + inside;
                need_low = need_low || (first < below && lo_key == ulo);
                need_high = need_high || (k >= below + inside && hi_key == uhi);
                if (!low_in && !high_in) {
                    continue;
                }
                const uint32_t base = ((uint32_t)y << 8) | (uint32_t)lane;
                if (!walk_to(m, c, high_in ? k : first, pos1, retained, g, base)) {
                    failed = true;
                    break;
                }
                Key out;
                if (!pick.pair) {
                    out = KT::unkey(keys[m - 1]);
                } else if (low_in && high_in) {
                    const int m2 = previous_member(m, pos1, g, base);
                    if (m2 < 1) {
                        failed = true;
                        break;
                    }
                    out = pair_bits<T>(a.stat, keys[m2 - 1], keys[m - 1], pick.frac);
                } else {
                    // The pair straddles two ranges: the first pass to find
                    // its element parks that key in the output, the second
                    // combines both.
                    if (!col_valid) {
                        continue;
                    }
                    Key *slot = output + (y0 + y) * a.nx + x0 + lane;
                    const uint64_t bit = 1ull << (y - y_first);
                    if (!(halves & bit)) {
                        *slot = keys[m - 1];
                        halves |= bit;
                        continue;
                    }
                    const Key other = *slot;
                    halves &= ~bit;
                    out = high_in ? pair_bits<T>(a.stat, other, keys[m - 1], pick.frac)
                                  : pair_bits<T>(a.stat, keys[m - 1], other, pick.frac);
                }
                if (col_valid) {
                    output[(y0 + y) * a.nx + x0 + lane] = out;
                }
            }
            const int wave_flags =
                (__ballot(failed) ? 1 : 0) | (__ballot(need_low) ? 2 : 0) | (__ballot(need_high) ? 4 : 0);
            if (lane == 0 && wave_flags) {
                atomicOr(&flags_shared, wave_flags);
            }
        }
        __syncthreads();
        // Queue the adjacent ranges the waves asked for (each doubles the
        // union's key span), within the pass budget; a failed walk, or a
        // range that cannot be added, sends the tile to the fallback list.
        if (tid == 0) {
            const int f = flags_shared;
            bool fail = (f & 1) != 0;
            int n2 = npend_shared;
            const int started = passes_shared;
            const Key span = uhi - ulo;
            const Key width = span < (Key)15 ? (Key)16 : (Key)(span + 1);
            if (!fail && (f & 2)) {
                if (ulo == 0 || started + n2 + 1 > MAX_PASSES) {
                    fail = true;
                } else {
                    const Key nlo = ulo > width ? ulo - width : (Key)0;
                    pend_lo[n2] = nlo;
                    pend_hi[n2] = ulo - 1;
                    ++n2;
                    union_lo = nlo;
                }
            }
            if (!fail && (f & 4)) {
                const Key top = KT::SENTINEL - 1;  // the sentinel is never a retained key
                if (uhi >= top || started + n2 + 1 > MAX_PASSES) {
                    fail = true;
                } else {
                    const Key nhi = top - uhi > width ? uhi + width : top;
                    pend_lo[n2] = uhi + 1;
                    pend_hi[n2] = nhi;
                    ++n2;
                    union_hi = nhi;
                }
            }
            npend_shared = n2;
            tile_failed = fail ? 1 : 0;
        }
        __syncthreads();
        }  // passes
        if (tid == 0 && tile_failed) {
            a.fallback_tiles[atomicAdd(a.fallback_count, 1)] = tile;
        }
        // Thread 0 rewrites the pass state at the next tile.
        __syncthreads();
    }
}

template <typename T>
hipError_t launch_bracket_typed(const SlideArgs &args, size_t lds_bytes, int max_tiles, hipStream_t stream)
{
    const dim3 block(NT);
    const void *fn = nullptr;
    switch (args.sort_e) {
    case 4: fn = reinterpret_cast<const void *>(&k_bracket<T, 4>); break;
    case 8: fn = reinterpret_cast<const void *>(&k_bracket<T, 8>); break;
    case 16: fn = reinterpret_cast<const void *>(&k_bracket<T, 16>); break;
    case 32: fn = reinterpret_cast<const void *>(&k_bracket<T, 32>); break;
    default: return hipErrorInvalidValue;
    }
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) {
        return e;
    }
    int per_cu = 0, device = 0, cus = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NT, lds_bytes)) != hipSuccess ||
        (e = hipGetDevice(&device)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) {
        return e;
    }
    int64_t blocks = (int64_t)(per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
    blocks = blocks < max_tiles ? blocks : max_tiles;
    const dim3 grid((unsigned)(blocks < 1 ? 1 : blocks));
    switch (args.sort_e) {
    case 4: hipLaunchKernelGGL((k_bracket<T, 4>), grid, block, lds_bytes, stream, args); break;
    case 8: hipLaunchKernelGGL((k_bracket<T, 8>), grid, block, lds_bytes, stream, args); break;
    case 16: hipLaunchKernelGGL((k_bracket<T, 16>), grid, block, lds_bytes, stream, args); break;
    default: hipLaunchKernelGGL((k_bracket<T, 32>), grid, block, lds_bytes, stream, args); break;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_typed(const SlideArgs &args, size_t lds_bytes, int max_tiles, hipStream_t stream)
{
    const dim3 block(NT);
    const void *fn = nullptr;
    switch (args.sort_e) {
    case 4: fn = reinterpret_cast<const void *>(&k_slide<T, 4>); break;
    case 8: fn = reinterpret_cast<const void *>(&k_slide<T, 8>); break;
    case 16: fn = reinterpret_cast<const void *>(&k_slide<T, 16>); break;
    case 32: fn = reinterpret_cast<const void *>(&k_slide<T, 32>); break;
    case 64: fn = reinterpret_cast<const void *>(&k_slide<T, 64>); break;
    default: return hipErrorInvalidValue;
    }
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) {
        return e;
    }
    // Persistent grid: as many workgroups as can be resident at once.
    int per_cu = 0, device = 0, cus = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NT, lds_bytes)) != hipSuccess ||
        (e = hipGetDevice(&device)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) {
        return e;
    }
    int64_t blocks = (int64_t)(per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
    blocks = blocks < max_tiles ? blocks : max_tiles;
    const dim3 grid((unsigned)(blocks < 1 ? 1 : blocks));
    switch (args.sort_e) {
    case 4: hipLaunchKernelGGL((k_slide<T, 4>), grid, block, lds_bytes, stream, args); break;
    case 8: hipLaunchKernelGGL((k_slide<T, 8>), grid, block, lds_bytes, stream, args); break;
    case 16: hipLaunchKernelGGL((k_slide<T, 16>), grid, block, lds_bytes, stream, args); break;
    case 32: hipLaunchKernelGGL((k_slide<T, 32>), grid, block, lds_bytes, stream, args); break;
    default: hipLaunchKernelGGL((k_slide<T, 64>), grid, block, lds_bytes, stream, args); break;
    }
    return hipGetLastError();
}

template <typename T, int E>
hipError_t launch_batch_typed(const SlideArgs *images, const int64_t *offsets, const int32_t *work, int nwork,
                              int32_t *counter, const void *input, void *output, int32_t *error_flag,
                              size_t lds_bytes, int grid, hipStream_t stream)
{
    hipLaunchKernelGGL((k_slide_batch<T, E>), dim3((unsigned)grid), dim3(NT), lds_bytes, stream, images, offsets,
                       work, nwork, counter, input, output, error_flag);
    return hipGetLastError();
}

template <typename T, int E>
hipError_t batch_grid_typed(size_t lds_bytes, int *grid)
{
    const void *fn = reinterpret_cast<const void *>(&k_slide_batch<T, E>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) {
        return e;
    }
    int per_cu = 0, device = 0, cus = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NT, lds_bytes)) != hipSuccess ||
        (e = hipGetDevice(&device)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) {
        return e;
    }
    *grid = (per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
    return hipSuccess;
}

}  // namespace

hipError_t slide_batch_grid(int dtype_bytes, int sort_e, size_t lds_bytes, int *grid)
{
#define NDF_BATCH_GRID(E)                                                                          \
    case E:                                                                                        \
        return dtype_bytes == 4 ? batch_grid_typed<float, E>(lds_bytes, grid)                      \
                                : batch_grid_typed<double, E>(lds_bytes, grid);
    switch (sort_e) {
        NDF_BATCH_GRID(4)
        NDF_BATCH_GRID(8)
        NDF_BATCH_GRID(16)
        NDF_BATCH_GRID(32)
        NDF_BATCH_GRID(64)
    default: return hipErrorInvalidValue;
    }
#undef NDF_BATCH_GRID
}

hipError_t launch_slide_batch(int dtype_bytes, int sort_e, const SlideArgs *images, const int64_t *offsets,
                              const int32_t *work, int nwork, int32_t *counter, const void *input, void *output,
                              int32_t *error_flag, size_t lds_bytes, int grid, hipStream_t stream)
{
#define NDF_BATCH_LAUNCH(E)                                                                            \
    case E:                                                                                            \
        return dtype_bytes == 4 ? launch_batch_typed<float, E>(images, offsets, work, nwork, counter, input, \
                                                               output, error_flag, lds_bytes, grid, stream) \
                                : launch_batch_typed<double, E>(images, offsets, work, nwork, counter, input, \
                                                                output, error_flag, lds_bytes, grid, stream);
    switch (sort_e) {
        NDF_BATCH_LAUNCH(4)
        NDF_BATCH_LAUNCH(8)
        NDF_BATCH_LAUNCH(16)
        NDF_BATCH_LAUNCH(32)
        NDF_BATCH_LAUNCH(64)
    default: return hipErrorInvalidValue;
    }
#undef NDF_BATCH_LAUNCH
}

// Host-side twin of exchange_bytes<Key, E>().
static size_t exchange_region(int dtype_bytes, int sort_e)
{
    const size_t per = (size_t)dtype_bytes + 2;
    const size_t rows = sort_e >= 8 ? (size_t)NT * sort_e * per : (size_t)NT * (sort_e + 1) * per;
    const size_t sorted = (size_t)NT * sort_e * per + 4 * GUARD;
    const size_t merge = (size_t)NT * sort_e * per + (sort_e >= 32 ? (size_t)NT * 32 : 0);  // MergeLayout padding
    const size_t most = rows > sorted ? rows : sorted;
    return ((most > merge ? most : merge) + 15) & ~(size_t)15;
}

size_t slide_lds_bytes(int dtype_bytes, int s, int sort_e, int hist_stride, int mask_bytes)
{
    // Exchange region, then codes[s], the per-wave histograms and the mask.
    return exchange_region(dtype_bytes, sort_e) + (size_t)((s + 7) & ~7) * 2 +
           (size_t)SLIDE_WAVES * (size_t)hist_stride * 2 + (size_t)mask_bytes;
}

size_t bracket_lds_bytes(int dtype_bytes, int s, int rows, int sort_e)
{
    // Exchange, compact codes (one per sort entry), row offsets, row masks.
    (void)s;
    return exchange_region(dtype_bytes, sort_e) + (size_t)NT * sort_e * 2 + (size_t)((rows + 1 + 7) & ~7) * 2 +
           (size_t)rows * sizeof(RowMasks);
}

hipError_t launch_bracket_slide(int dtype_bytes, const SlideArgs &args, size_t lds_bytes, int max_tiles,
                                hipStream_t stream)
{
    if (max_tiles <= 0) {
        return hipSuccess;
    }
    return dtype_bytes == 4 ? launch_bracket_typed<float>(args, lds_bytes, max_tiles, stream)
                            : launch_bracket_typed<double>(args, lds_bytes, max_tiles, stream);
}

hipError_t launch_slide(int dtype_bytes, const SlideArgs &args, size_t lds_bytes, int max_tiles, hipStream_t stream)
{
    if (max_tiles <= 0) {
        return hipSuccess;
    }
    return dtype_bytes == 4 ? launch_typed<float>(args, lds_bytes, max_tiles, stream)
                            : launch_typed<double>(args, lds_bytes, max_tiles, stream);
}

}  // namespace ndf
