// Slide engine: LDS-staged tile + block sort + per-lane column slide.
//
// This is the reference's block rank histogram (src/histogram.rs:1-16,
// 302-485) re-designed for a 64-wide wavefront:
//   * a workgroup stages the support of a 64 x ty output tile (tile + window
//     halo, mapped through the boundary maps) into LDS with coalesced row
//     loads, classifying each sample under the policy;
//   * the retained samples are sorted once (bitonic, in LDS), giving every
//     support sample a tile rank (codes) and every rank its position;
//   * each lane owns one output column and slides its window down the tile.
//     Instead of a bit histogram per window it keeps one cursor: a rank m and
//     c = number of in-window ranks below m. A step updates c from the removed
//     and added edge taps only, then walks m to the target rank, testing
//     window membership of each rank it passes from its position.
// The selected rank maps back to the original key bits, so the output is the
// exact sample (or the f64 midpoint of two), independent of the tile shape.
#include "ndf_device.hpp"
#include "ndf_internal.hpp"

namespace ndf {
namespace {

constexpr uint16_t CODE_OMIT = 0xFFFF;

__device__ inline bool in_window(uint32_t p, int lane, int y, int kz, int ky, int kx, const uint8_t *mask)
{
    const int x = (int)(p & 0x3FFu);
    const int yy = (int)((p >> 10) & 0x3FFu);
    const int zz = (int)(p >> 20);
    const unsigned dx = (unsigned)(x - lane);
    const unsigned dy = (unsigned)(yy - y);
    if (dx >= (unsigned)kx || dy >= (unsigned)ky || (unsigned)zz >= (unsigned)kz) {
        return false;
    }
    return mask == nullptr || mask[(zz * ky + (int)dy) * kx + (int)dx] != 0;
}

template <typename T>
__global__ void __launch_bounds__(SLIDE_WAVES * 64) k_slide(SlideArgs a)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    extern __shared__ __align__(16) unsigned char lds[];
    Key *keys = reinterpret_cast<Key *>(lds);
    uint32_t *pos = reinterpret_cast<uint32_t *>(keys + a.npad);
    uint16_t *codes = reinterpret_cast<uint16_t *>(pos + a.npad);
    uint8_t *mask = a.mask ? reinterpret_cast<uint8_t *>(codes + ((a.s + 1) & ~1)) : nullptr;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    constexpr int NT = SLIDE_WAVES * 64;

    int tile = blockIdx.x;
    const int tx = tile % a.tiles_x;
    tile /= a.tiles_x;
    const int tyi = tile % a.tiles_y;
    const int64_t z0 = tile / a.tiles_y;
    const int64_t x0 = (int64_t)tx * SLIDE_TILE_X;
    const int64_t y0 = (int64_t)tyi * a.ty;

    // ---- stage: one support row per wave iteration, lanes along x ----
    const Key *input = static_cast<const Key *>(a.input);
    const Key cval_key = classify<T>((Key)a.cval_bits, a.ignore_inf);
    const int rows = a.kz * a.sy;
    for (int row = wave; row < rows; row += SLIDE_WAVES) {
        const int z = row / a.sy;
        const int y = row - z * a.sy;
        const int64_t mz = a.mapz[z0 + z];
        const int64_t my = a.mapy[y0 + y];
        for (int x = lane; x < a.sx; x += 64) {
            const int64_t mx = a.mapx[x0 + x];
            Key key;
            if (mz == MAP_PAD || my == MAP_PAD || mx == MAP_PAD) {
                key = KT::SENTINEL;
            } else if (mz < 0 || my < 0 || mx < 0) {
                key = cval_key;
            } else {
                key = classify<T>(input[mz + my + mx], a.ignore_inf);
            }
            const int p = row * a.sx + x;
            keys[p] = key;
            pos[p] = key == KT::SENTINEL ? SLIDE_NO_POS : ((uint32_t)z << 20) | ((uint32_t)y << 10) | (uint32_t)x;
            codes[p] = CODE_OMIT;
        }
    }
    for (int p = a.s + tid; p < a.npad; p += NT) {
        keys[p] = KT::SENTINEL;
        pos[p] = SLIDE_NO_POS;
    }
    if (mask) {
        const int mbytes = a.kz * a.ky * a.kx;
        for (int i = tid; i < mbytes; i += NT) {
            mask[i] = a.mask[i];
        }
    }
    __syncthreads();

    // ---- bitonic sort of (key, pos) over npad entries ----
    const int half = a.npad >> 1;
    for (int k = 2; k <= a.npad; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int q = tid; q < half; q += NT) {
                const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                const int l = i + j;
                const bool up = (i & k) == 0;
                const Key ki = keys[i];
                const Key kl = keys[l];
                if ((ki > kl) == up) {
                    keys[i] = kl;
                    keys[l] = ki;
                    const uint32_t pi = pos[i];
                    pos[i] = pos[l];
                    pos[l] = pi;
                }
            }
            __syncthreads();
        }
    }

    // ---- ranks: codes[support position] = rank; R = retained count ----
    __shared__ int retained_shared;
    if (tid == 0) {
        retained_shared = 0;
    }
    __syncthreads();
    for (int r = tid; r < a.npad; r += NT) {
        const uint32_t p = pos[r];
        if (p != SLIDE_NO_POS) {
            const int x = (int)(p & 0x3FFu);
            const int yy = (int)((p >> 10) & 0x3FFu);
            const int zz = (int)(p >> 20);
            codes[(zz * a.sy + yy) * a.sx + x] = (uint16_t)r;
            if (r + 1 == a.npad || pos[r + 1] == SLIDE_NO_POS) {
                retained_shared = r + 1;
            }
        }
    }
    __syncthreads();
    const int retained = retained_shared;

    // ---- per-lane slide ----
    const int y_first = wave * a.tw;
    if (y0 + y_first >= a.ny) {
        return;
    }
    const bool col_valid = x0 + lane < a.nx;
    Key *output = static_cast<Key *>(a.output);
    int m = retained >> 1;
    int c = 0;
    int n = 0;
    {
        const int base = y_first * a.sx + lane;
        for (int t = 0; t < a.count; ++t) {
            const uint16_t code = codes[base + a.taps[t]];
            n += code != CODE_OMIT;
            c += code < m;
        }
    }
    const int walk_limit = 2 * a.npad + 4;
    for (int y = y_first; y < y_first + a.tw && y0 + y < a.ny; ++y) {
        if (y > y_first) {
            const int rbase = (y - 1) * a.sx + lane;
            const int abase = y * a.sx + lane;
            for (int e = 0; e < a.nedge; ++e) {
                const uint16_t cr = codes[rbase + a.edge_remove[e]];
                const uint16_t ca = codes[abase + a.edge_add[e]];
                n += (int)(ca != CODE_OMIT) - (int)(cr != CODE_OMIT);
                c += (int)(ca < m) - (int)(cr < m);
            }
        }
        Key out = KT::CANONICAL_NAN;
        int k = -1;
        if (n >= a.min_count && n > 0) {
            if (a.stat == STAT_RANK) {
                k = a.rank_param < n ? a.rank_param : -1;
            } else {
                k = n >> 1;
            }
        }
        if (k >= 0) {
            // Walk m until it is an in-window rank with exactly k in-window
            // ranks below it.
            // Every access is bounds-checked; a cursor that leaves [0, retained)
            // or a walk longer than the tile raises the error flag.
            int steps = 0;
            bool ok = true;
            for (;;) {
                if (++steps > walk_limit) {
                    ok = false;
                    break;
                }
                if (c > k) {
                    if (m <= 0) {
                        ok = false;
                        break;
                    }
                    --m;
                    c -= in_window(pos[m], lane, y, a.kz, a.ky, a.kx, mask);
                } else {
                    if (m >= retained) {
                        ok = false;
                        break;
                    }
                    const bool inside = in_window(pos[m], lane, y, a.kz, a.ky, a.kx, mask);
                    if (c == k && inside) {
                        break;
                    }
                    c += inside;
                    ++m;
                }
            }
            if (ok && a.stat == STAT_MEDIAN && (n & 1) == 0) {
                int m2 = m - 1;
                while (m2 >= 0 && !in_window(pos[m2], lane, y, a.kz, a.ky, a.kx, mask)) {
                    --m2;
                }
                if (m2 >= 0) {
                    out = median_pair_bits<T>(keys[m2], keys[m]);
                } else {
                    ok = false;
                }
            } else if (ok) {
                out = KT::unkey(keys[m]);
            }
            if (!ok) {
                atomicOr(a.error_flag, 1);
                m = retained >> 1;
                k = -1;
            }
            if (k < 0) {
                // Restore a consistent cursor after a failed walk.
                c = 0;
                const int base = y * a.sx + lane;
                for (int t = 0; t < a.count; ++t) {
                    c += codes[base + a.taps[t]] < m;
                }
            }
        }
        if (col_valid) {
            output[z0 * a.ny * a.nx + (y0 + y) * a.nx + x0 + lane] = out;
        }
    }
}

}  // namespace

size_t slide_lds_bytes(int dtype_bytes, int s, int npad, int mask_bytes)
{
    return (size_t)npad * (size_t)(dtype_bytes + 4) + (size_t)((s + 1) & ~1) * 2 + (size_t)mask_bytes;
}

hipError_t launch_slide(int dtype_bytes, const SlideArgs &args, size_t lds_bytes, hipStream_t stream)
{
    const int64_t blocks = (int64_t)args.tiles_x * args.tiles_y * args.nz;
    if (blocks <= 0) {
        return hipSuccess;
    }
    const dim3 grid((unsigned)blocks);
    const dim3 block(SLIDE_WAVES * 64);
    if (dtype_bytes == 4) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_slide<float>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) {
            return e;
        }
        hipLaunchKernelGGL(k_slide<float>, grid, block, lds_bytes, stream, args);
    } else {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_slide<double>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) {
            return e;
        }
        hipLaunchKernelGGL(k_slide<double>, grid, block, lds_bytes, stream, args);
    }
    return hipGetLastError();
}

}  // namespace ndf
