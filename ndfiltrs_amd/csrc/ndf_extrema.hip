// Extrema engine: rectangular minimum / maximum filters (rank <= 3) by
// separable passes over an LDS-staged tile.
//
// The reference computes rectangular extrema with exact separable van Herk /
// Gil-Werman passes over whole lines (src/extrema.rs:125-436), one axis at a
// time through image-sized buffers. Here one workgroup owns an output tile
// of tx columns x ty rows of one plane and stages its support (kz planes of
// ty + ky - 1 rows of tx + kx - 1 samples) in LDS once, so each input sample
// crosses HBM about once and the intermediate results never leave the CU:
//   * every sample becomes an unsigned total-order key; a propagated NaN
//     becomes the key that dominates the reduction (0 for a minimum, all
//     ones for a maximum) and an omitted sample the one that loses, so the
//     policy rides inside the reduction;
//   * the x pass and the y pass each take the extreme over k samples in
//     radix-4 steps: m_4w[i] = op(m_w[i], m_w[i + w], m_w[i + 2w],
//     m_w[i + 3w]) while 4w <= k, then op over the ceil(k / w) <= 4 windows
//     of width w at i, i + w, ..., i + k - w (they overlap, which an
//     idempotent op allows), about log4(k) steps in all; every step runs
//     in place, the x steps on wave-private rows and the y steps on
//     wave-private columns, so one barrier per tile separates the passes,
//     and each wave reads a batch of rows before writing them, so it waits
//     on LDS once per batch;
//   * the z pass folds the kz planes per output.
// The extreme key is a retained sample's key (its bits unchanged), or one of
// the two marks (canonical NaN). Equal keys are equal bits, so which of two
// equal extremes wins cannot change a result: the output is bit-identical to
// the direct total-order extreme of the retained samples. Windows that need
// a retained count (min_count > 1) run on the selection engines instead.
#include "ndf_device.hpp"
#include "ndf_internal.hpp"

namespace ndf {
namespace {

// Waves per workgroup: several waves per SIMD share one staged tile (the
// passes have no barriers of their own, so they hide each other's LDS
// waits). A tile above half of the CU's 160 KiB of LDS leaves one workgroup
// per CU, which then runs 16 waves; smaller tiles run 8 and share the CU
// with other workgroups (8192^2 float32: 63x63 0.68 ms at 16 waves against
// 0.95 at 8; 15x15 0.46 ms at 8 against 0.63 at 16; 9x9 0.39 against 0.43).
constexpr int EXTREMA_WIDE_WAVES = 16;
constexpr int EXTREMA_NARROW_WAVES = 8;
constexpr size_t EXTREMA_WIDE_LDS = 80 * 1024;

template <typename Key, bool MAXIMUM>
__device__ __forceinline__ Key pick_extreme(Key a, Key b)
{
    if constexpr (MAXIMUM) {
        return a > b ? a : b;
    } else {
        return a < b ? a : b;
    }
}

__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Support columns one lane stages per row (sx <= 64 * STAGE_COLUMNS).
constexpr int STAGE_COLUMNS = 8;
// Rows a wave stages per batch (all loads of a batch are issued first).
constexpr int STAGE_ROWS = 4;
// Rows a wave reads before writing, per step of the x and y passes.
constexpr int X_ROWS = 8;
constexpr int Y_ROWS = 8;

// One radix-4 step along an axis with window k: from extremes over w
// consecutive samples to extremes over 4w (reads at +0, w, 2w, 3w) while
// 4w <= k; else the final combine of the n = ceil(k / w) <= 4 windows of
// width w that cover k (reads at +0, w, ..., the last at k - w; reads past
// the n-th repeat k - w, which an idempotent op allows).
struct Radix {
    int o1, o2, o3;
    bool final;
};
__device__ inline Radix radix_step(int w, int k)
{
    if (4 * w <= k) return {w, 2 * w, 3 * w, false};
    const int last = k - w;
    const int n = (k + w - 1) / w;
    return {n > 2 ? w : last, n > 3 ? 2 * w : last, last, true};
}

template <typename Key, bool MAXIMUM>
__device__ __forceinline__ Key pick4(Key a, Key b, Key c, Key d)
{
    return pick_extreme<Key, MAXIMUM>(pick_extreme<Key, MAXIMUM>(a, b), pick_extreme<Key, MAXIMUM>(c, d));
}

template <typename T, bool MAXIMUM, int WAVES>
__global__ void __launch_bounds__(WAVES * 64) k_extrema(ExtremaArgs a)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    extern __shared__ __align__(16) unsigned char lds[];
    const int plane = a.sy * a.sx;  // staging buffer: kz planes of sy x sx
    Key *stage = reinterpret_cast<Key *>(lds);
    const Key *input = static_cast<const Key *>(a.input);
    Key *output = static_cast<Key *>(a.output);
    // Marks: an omitted (or padding) sample never wins; a propagated NaN always does.
    const Key lose = MAXIMUM ? (Key)0 : KT::SENTINEL;
    const Key win = MAXIMUM ? KT::SENTINEL : (Key)0;
    auto classify_key = [&](Key bits) -> Key {
        if (KT::is_nan(bits)) return a.omit ? lose : win;
        if (a.omit && a.ignore_inf && KT::is_inf(bits)) return lose;
        return KT::key(bits);
    };
    const Key cval_key = classify_key((Key)a.cval_bits);
    const int tiles = a.tiles_x * a.tiles_y * (int)a.nz;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int rows = a.kz * a.sy;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        int rest = tile;
        const int txi = rest % a.tiles_x;
        rest /= a.tiles_x;
        const int tyi = rest % a.tiles_y;
        const int64_t z0 = rest / a.tiles_y;
        const int64_t x0 = (int64_t)txi * a.tx, y0 = (int64_t)tyi * a.ty;
        // Column maps of this lane, shared by every row of the tile.
        int64_t mx[STAGE_COLUMNS];
#pragma unroll
        for (int j = 0; j < STAGE_COLUMNS; ++j) {
            const int x = lane + 64 * j;
            mx[j] = x < a.sx ? a.mapx[x0 + x] : MAP_PAD;
        }
        __syncthreads();  // every wave of the previous tile is done with the buffer
        // ---- stage the support as keys: each wave owns rows wave + WAVES i ----
        for (int row0 = wave; row0 < rows; row0 += STAGE_ROWS * WAVES) {
            Key raw[STAGE_ROWS][STAGE_COLUMNS];
            int64_t my[STAGE_ROWS];
#pragma unroll
            for (int u = 0; u < STAGE_ROWS; ++u) {
                const int row = row0 + u * WAVES;
                const int z = row / a.sy;
                const int y = row - z * a.sy;
                const int64_t mz_at = uniform_entry(a.mapz, z0 + (row < rows ? z : 0));
                const int64_t my_at = uniform_entry(a.mapy, y0 + (row < rows ? y : 0));
                const int64_t mz = row < rows ? mz_at : MAP_PAD;
                const int64_t myy = row < rows ? my_at : MAP_PAD;
                my[u] = (mz == MAP_PAD || myy == MAP_PAD) ? MAP_PAD : ((mz < 0 || myy < 0) ? MAP_CVAL : mz + myy);
#pragma unroll
                for (int j = 0; j < STAGE_COLUMNS; ++j) {
                    const bool real = my[u] >= 0 && mx[j] >= 0;
                    raw[u][j] = input[real ? my[u] + mx[j] : 0];
                }
            }
#pragma unroll
            for (int u = 0; u < STAGE_ROWS; ++u) {
                const int row = row0 + u * WAVES;
                if (row >= rows) continue;
#pragma unroll
                for (int j = 0; j < STAGE_COLUMNS; ++j) {
                    const int x = lane + 64 * j;
                    if (x >= a.sx) continue;
                    const bool pad = my[u] == MAP_PAD || mx[j] == MAP_PAD;
                    const bool cval = !pad && (my[u] < 0 || mx[j] < 0);
                    stage[row * a.sx + x] = pad ? lose : (cval ? cval_key : classify_key(raw[u][j]));
                }
            }
        }
        // ---- x pass, in place: each wave reduces the rows it staged ----
        // Batches of X_ROWS rows, one 64-column chunk at a time: a wave reads
        // the chunk's values of every row of the batch before it writes any,
        // so the LDS latency is paid once per batch. Rows are wave-private
        // and a wave's LDS operations complete in order, so each read sees
        // the previous step's value; chunks go in increasing order, so a
        // chunk never reads a value a lower chunk already overwrote. Indices
        // are clamped instead of branched on (the loop stays one basic
        // block): a clamped lane or row repeats an in-range element's read
        // and writes back the value that element's own lane writes.
        wave_lds_sync();
        for (int w = 1; w < a.kx; w *= 4) {
            const Radix st = radix_step(w, a.kx);
            const int columns = st.final ? a.tx : a.sx;
            for (int r0 = wave; r0 < rows; r0 += X_ROWS * WAVES) {
                Key *row[X_ROWS];
#pragma unroll
                for (int u = 0; u < X_ROWS; ++u) {
                    const int r = r0 + u * WAVES;
                    row[u] = stage + (r < rows ? r : r0) * a.sx;
                }
                for (int c = 0; c < columns; c += 64) {
                    const int i = min(c + lane, columns - 1);
                    const int i1 = i + st.o1 < a.sx ? i + st.o1 : i;
                    const int i2 = i + st.o2 < a.sx ? i + st.o2 : i;
                    const int i3 = i + st.o3 < a.sx ? i + st.o3 : i;
                    Key v0[X_ROWS], v1[X_ROWS], v2[X_ROWS], v3[X_ROWS];
#pragma unroll
                    for (int u = 0; u < X_ROWS; ++u) {
                        v0[u] = row[u][i];
                        v1[u] = row[u][i1];
                        v2[u] = row[u][i2];
                        v3[u] = row[u][i3];
                    }
#pragma unroll
                    for (int u = 0; u < X_ROWS; ++u) row[u][i] = pick4<Key, MAXIMUM>(v0[u], v1[u], v2[u], v3[u]);
                }
            }
            wave_lds_sync();
            if (st.final) break;
        }
        __syncthreads();  // the y pass reads rows the other waves staged
        // ---- y pass, in place on wave-private columns ----
        // Wave w owns columns [w cw, (w + 1) cw) of the tile (cw = tx / WAVES);
        // a lane covers column lane % cw of rows lane / cw + rp i. Rows go in
        // increasing order, and a batch reads all of its rows (and the rows
        // w, 2w, 3w below them, not yet written this step) before writing
        // any, so every read sees the previous step's value; only wave-level
        // ordering is needed. A read below the plane falls back to the row
        // itself, so the plane's last row never changes, and a clamped row
        // repeats it. The final combine happens in the output step.
        const int cw = a.tx / WAVES;
        const int rp = 64 / cw;
        const int col = wave * cw + (lane & (cw - 1));
        const int ro = lane / cw;
        int wy = 1;
        for (; 4 * wy <= a.ky; wy *= 4) {
            for (int z = 0; z < a.kz; ++z) {
                Key *base = stage + z * plane + col;
                for (int r0 = 0; r0 < a.sy; r0 += Y_ROWS * rp) {
                    int at[Y_ROWS];
                    Key v0[Y_ROWS], v1[Y_ROWS], v2[Y_ROWS], v3[Y_ROWS];
#pragma unroll
                    for (int u = 0; u < Y_ROWS; ++u) {
                        const int r = min(r0 + u * rp + ro, a.sy - 1);
                        at[u] = r * a.sx;
                        v0[u] = base[at[u]];
                        v1[u] = base[(r + wy < a.sy ? r + wy : r) * a.sx];
                        v2[u] = base[(r + 2 * wy < a.sy ? r + 2 * wy : r) * a.sx];
                        v3[u] = base[(r + 3 * wy < a.sy ? r + 3 * wy : r) * a.sx];
                    }
#pragma unroll
                    for (int u = 0; u < Y_ROWS; ++u) base[at[u]] = pick4<Key, MAXIMUM>(v0[u], v1[u], v2[u], v3[u]);
                }
            }
            wave_lds_sync();
        }
        // ---- final y step, z fold and output (the same wave-private columns) ----
        const Radix fy = radix_step(wy, a.ky);
        const int64_t x = x0 + col;
        for (int r0 = 0; r0 < a.ty; r0 += Y_ROWS * rp) {
#pragma unroll
            for (int u = 0; u < Y_ROWS; ++u) {
                const int r = r0 + u * rp + ro;
                const int rc = min(r, a.ty - 1);
                Key best = lose;
                for (int z = 0; z < a.kz; ++z) {
                    const Key *at = stage + z * plane + rc * a.sx + col;
                    best = pick_extreme<Key, MAXIMUM>(
                        best, pick4<Key, MAXIMUM>(at[0], at[fy.o1 * a.sx], at[fy.o2 * a.sx], at[fy.o3 * a.sx]));
                }
                const int64_t y = y0 + r;
                if (r < a.ty && y < a.ny && x < a.nx) {
                    output[(z0 * a.ny + y) * a.nx + x] = (best == win || best == lose) ? KT::CANONICAL_NAN : KT::unkey(best);
                }
            }
        }
    }
}

template <typename T, bool MAXIMUM, int WAVES>
hipError_t launch_waves(const ExtremaArgs &args, size_t lds_bytes, hipStream_t stream)
{
    constexpr int threads = WAVES * 64;
    const void *fn = reinterpret_cast<const void *>(&k_extrema<T, MAXIMUM, WAVES>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    int per_cu = 0, device = 0, cus = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds_bytes)) != hipSuccess ||
        (e = hipGetDevice(&device)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) {
        return e;
    }
    const int64_t tiles = (int64_t)args.tiles_x * args.tiles_y * args.nz;
    int64_t blocks = (int64_t)(per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus) * 2;
    blocks = blocks < tiles ? blocks : tiles;
    hipLaunchKernelGGL((k_extrema<T, MAXIMUM, WAVES>), dim3((unsigned)(blocks < 1 ? 1 : blocks)), dim3(threads),
                       lds_bytes, stream, args);
    return hipGetLastError();
}

template <typename T, bool MAXIMUM>
hipError_t launch_typed(const ExtremaArgs &args, size_t lds_bytes, hipStream_t stream)
{
    return lds_bytes > EXTREMA_WIDE_LDS ? launch_waves<T, MAXIMUM, EXTREMA_WIDE_WAVES>(args, lds_bytes, stream)
                                        : launch_waves<T, MAXIMUM, EXTREMA_NARROW_WAVES>(args, lds_bytes, stream);
}

}  // namespace

size_t extrema_lds_bytes(int dtype_bytes, int kz, int sy, int sx, int tx)
{
    (void)tx;  // every pass works in place in the staging buffer
    return (size_t)kz * (size_t)sy * (size_t)sx * (size_t)dtype_bytes;
}

int extrema_max_support_columns() { return 64 * STAGE_COLUMNS; }

hipError_t launch_extrema(int dtype_bytes, const ExtremaArgs &args, size_t lds_bytes, hipStream_t stream)
{
    if (dtype_bytes == 4) {
        return args.maximum ? launch_typed<float, true>(args, lds_bytes, stream)
                            : launch_typed<float, false>(args, lds_bytes, stream);
    }
    return args.maximum ? launch_typed<double, true>(args, lds_bytes, stream)
                        : launch_typed<double, false>(args, lds_bytes, stream);
}

}  // namespace ndf
