// Learned bracket proposal for the slide engine (2-D images).
//
// A tile's in-bracket samples are the only ones the bracket kernel sorts and
// walks. The bracket comes from a tiny MLP (6 -> 32 -> 32 -> 2, ReLU, fp32)
// that reads cheap box-sum features of the tile's support:
//   1. k_block_stats: one pass over the image, per 8x8 block: retained
//      count, mean and standard deviation (f64);
//   2. k_tile_features: per tile, robust level c and noise scale sigma
//      (medians of the block means and block std devs), a plane slope of
//      the block means, the retained fraction and the fraction of bright
//      blocks, standardised so that the features are scale free;
//   3. the MLP (run by the same wave as the tile's features): its two outputs
//      are standardised offsets: bracket = [c + sigma*o0, c + sigma*o1],
//      converted to keys.
// Training (k_tile_labels, k_mlp_grad, k_mlp_adam) fits the offsets of the
// smallest and largest window result per tile with pinball losses, so that
// the bracket contains every target of a tile with high probability.
// Exactness never depends on the model: the bracket kernel sends any tile
// whose targets leave the bracket to the plain kernel.
#include "ndf_bracket.hpp"
#include "ndf_device.hpp"

namespace ndf {
namespace {

constexpr int BLOCK = 8;
// Blocks per tile support the features read (at most 4 per lane of the
// wave); wider supports are subsampled by every 2nd, 4th, ... block row and
// column.
constexpr int MAX_TILE_BLOCKS = 256;

// One thread per 8x8 block; consecutive threads take consecutive blocks of
// a block row, so each image row segment is read by 64 lanes as 2 KB of
// contiguous memory. The block's retained count, sum and sum of squares are
// accumulated in f64 in row-major order within the block. A whole block
// whose rows start on 16-byte boundaries reads each row segment with 16-byte
// vector loads, all eight rows before the first use; the sums are the same
// additions in the same order either way.
template <typename T>
__global__ void __launch_bounds__(256) k_block_stats(const T *image, int64_t ny, int64_t nx, int64_t by, int64_t bx,
                                                     int ignore_inf, double *out)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    const int64_t block = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (block >= by * bx) {
        return;
    }
    const int64_t y0 = (block / bx) * BLOCK, x0 = (block % bx) * BLOCK;
    double n = 0.0, s = 0.0, q = 0.0;
    auto add = [&](Key bits) {
        // Omitted samples add +0.0, which leaves every sum unchanged (the
        // sums start at +0.0 and never become -0.0).
        const bool keep = classify<T>(bits, ignore_inf) != KT::SENTINEL;
        const double v = KT::to_double(bits);
        n += keep ? 1.0 : 0.0;
        s += keep ? v : 0.0;
        q += keep ? v * v : 0.0;
    };
    constexpr int PER16 = 16 / (int)sizeof(T);  // elements per 16-byte load
    const bool whole = y0 + BLOCK <= ny && x0 + BLOCK <= nx && ((nx * (int64_t)sizeof(T)) % 16) == 0 &&
                       (reinterpret_cast<uintptr_t>(image) % 16) == 0;
    if (whole) {
        uint4 raw[BLOCK][BLOCK / PER16];
#pragma unroll
        for (int dy = 0; dy < BLOCK; ++dy) {
            const uint4 *row = reinterpret_cast<const uint4 *>(image + (y0 + dy) * nx + x0);
#pragma unroll
            for (int c = 0; c < BLOCK / PER16; ++c) raw[dy][c] = row[c];
        }
#pragma unroll
        for (int dy = 0; dy < BLOCK; ++dy) {
#pragma unroll
            for (int c = 0; c < BLOCK / PER16; ++c) {
                const Key *keys = reinterpret_cast<const Key *>(&raw[dy][c]);
#pragma unroll
                for (int e = 0; e < PER16; ++e) add(keys[e]);
            }
        }
    } else {
        for (int dy = 0; dy < BLOCK; ++dy) {
            const int64_t y = y0 + dy;
            if (y >= ny) {
                break;
            }
            const T *row = image + y * nx;
#pragma unroll
            for (int dx = 0; dx < BLOCK; ++dx) {
                if (x0 + dx < nx) add(KT::bits(row[x0 + dx]));
            }
        }
    }
    // Block moments for the tile features: retained count, mean and
    // standard deviation (f64, computed once per block here rather than per
    // tile that reads the block). A block with n == 0 stores a NaN mean,
    // which no tile reads (blocks with fewer than 16 retained samples only
    // contribute their count).
    const double mu = s / n;
    const double var = q / n - mu * mu;
    double *o = out + block * 3;
    o[0] = n;
    o[1] = mu;
    o[2] = var > 0.0 ? __builtin_sqrt(var) : 0.0;
}

// asinh(v) = sign(v) * log(|v| + sqrt(v^2 + 1)); log(2|v|) once v^2 would overflow.
__device__ inline float feature_asinh(double v)
{
    const double a = v < 0.0 ? -v : v;
    const double r = a > 1e150 ? log(2.0 * a) : log(a + sqrt(a * a + 1.0));
    return (float)(v < 0.0 ? -r : r);
}

// Order-preserving 32-bit key of a float (total order: -NaN < -inf < ... <
// -0 < +0 < ... < +inf < +NaN) and its inverse.
__device__ inline uint32_t order_key(float v)
{
    const uint32_t u = __builtin_bit_cast(uint32_t, v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline float order_value(uint32_t k)
{
    return __builtin_bit_cast(float, (k & 0x80000000u) ? (k ^ 0x80000000u) : ~k);
}

// v broadcast from lane i (i wave-uniform).
__device__ inline float lane_value(float v, int i)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), i));
}

// Wave-wide minimum and maximum of the PER keys per lane.
template <int PER>
__device__ inline void wave_key_range(const uint32_t (&key)[PER], uint32_t &lo, uint32_t &hi)
{
    lo = key[0];
    hi = key[0];
#pragma unroll
    for (int j = 1; j < PER; ++j) {
        lo = key[j] < lo ? key[j] : lo;
        hi = key[j] > hi ? key[j] : hi;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t l = (uint32_t)__shfl_xor((int)lo, off), h = (uint32_t)__shfl_xor((int)hi, off);
        lo = l < lo ? l : lo;
        hi = h > hi ? h : hi;
    }
}

// The ka-th smallest of the wave's keys a[] and the kb-th smallest of b[]
// (0-based; PER keys per lane each), by bitwise searches from the highest
// bit in which a key range's minimum and maximum differ (the bits above are
// common to every key in the range): each answer is the largest v with at
// most k keys below it. A step counts the keys below each candidate with one
// wave ballot per register; the two searches share the steps.
template <int PER>
__device__ inline void wave_select2(const uint32_t (&a)[PER], int ka, const uint32_t (&b)[PER], int kb, uint32_t &ra,
                                    uint32_t &rb)
{
    uint32_t alo, ahi, blo, bhi;
    wave_key_range<PER>(a, alo, ahi);
    wave_key_range<PER>(b, blo, bhi);
    const uint32_t adiff = alo ^ ahi, bdiff = blo ^ bhi;
    const int abit = adiff ? 31 - __builtin_clz(adiff) : -1;
    const int bbit = bdiff ? 31 - __builtin_clz(bdiff) : -1;
    // Common high bits (all of them when every key is equal).
    uint32_t pa = abit < 0 ? alo : abit >= 31 ? 0u : alo & ~((2u << abit) - 1u);
    uint32_t pb = bbit < 0 ? blo : bbit >= 31 ? 0u : blo & ~((2u << bbit) - 1u);
    for (int bit = abit > bbit ? abit : bbit; bit >= 0; --bit) {
        const uint32_t ca = pa | (bit <= abit ? (1u << bit) : 0u);
        const uint32_t cb = pb | (bit <= bbit ? (1u << bit) : 0u);
        int na = 0, nb = 0;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            na += (int)__popcll(__ballot(a[j] < ca));
            nb += (int)__popcll(__ballot(b[j] < cb));
        }
        pa = na <= ka ? ca : pa;
        pb = nb <= kb ? cb : pb;
    }
    ra = pa;
    rb = pb;
}

// The support's 8x8 blocks as one row-major list of `count` entries
// (w_used per row, every step-th block row and column).
struct BlockList {
    int64_t br0, bc0, bx;
    int w_used, step, count;
    __device__ const double *at(const double *blocks, int i) const
    {
        const int64_t br = br0 + (int64_t)(i / w_used) * step, bc = bc0 + (int64_t)(i % w_used) * step;
        return blocks + (br * bx + bc) * 3;
    }
};

// A tile's block statistics, summed over the wave (every lane holds the
// totals): retained count n over `cells` pixels, `usable` blocks (at least
// 16 retained samples), level c and scale sigma (medians of the usable
// blocks' means and standard deviations), the plane-fit sums over the blocks
// within 0.75 sigma of c, and the bright-block sums.
struct TileSums {
    double n, cells;
    int usable;
    float c, sigma;
    double sw, sr, sq, sm, srr, sqq, srq, srm, sqm, extended, spiky;
};

// Block i = j * 64 + lane of the list sits in register j of its lane; the
// three moments of all PER blocks are loaded before the first use.
template <int PER>
__device__ inline TileSums tile_sums(const double *blocks, const BlockList &list, int lane)
{
    double o0[PER], o1[PER], o2[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = j * 64 + lane;
        const double *o = list.at(blocks, i < list.count ? i : 0);
        o0[j] = o[0];
        o1[j] = o[1];
        o2[j] = o[2];
    }
    TileSums t{};
    uint32_t mean_key[PER], sd_key[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const bool valid = j * 64 + lane < list.count;
        const bool usable = valid && o0[j] >= 16.0;
        t.n += valid ? o0[j] : 0.0;
        t.cells += valid ? 64.0 : 0.0;
        t.usable += usable ? 1 : 0;
        // Unusable blocks and padding sort last (above every float key).
        mean_key[j] = usable ? order_key((float)o1[j]) : 0xFFFFFFFFu;
        sd_key[j] = usable ? order_key((float)o2[j]) : 0xFFFFFFFFu;
    }
    for (int off = 32; off > 0; off >>= 1) {
        t.n += __shfl_xor(t.n, off);
        t.cells += __shfl_xor(t.cells, off);
        t.usable += __shfl_xor(t.usable, off);
    }
    t.c = 0.0f;
    t.sigma = 1.0f;
    if (t.usable > 0) {
        uint32_t mk, sk;
        wave_select2<PER>(mean_key, (t.usable - 1) / 2, sd_key, (t.usable - 1) / 2, mk, sk);
        t.c = order_value(mk);
        t.sigma = order_value(sk);
    }
    const float floor_sigma = 1e-6f * (__builtin_fabsf(t.c) + 1e-30f);
    t.sigma = t.sigma > floor_sigma ? t.sigma : floor_sigma;
    const double c = (double)t.c, sigma = (double)t.sigma;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = j * 64 + lane;
        if (i < list.count && o0[j] >= 16.0) {
            const double mu = o1[j] - c;
            // A bright block is "spiky" when its spread dwarfs its mean
            // shift (one hot pixel: sd ~ 8 x shift) and "extended" otherwise
            // (a source covering the block); only extended brightness moves
            // window ranks much.
            if (mu > 0.5 * sigma) {
                if (o2[j] < 4.0 * mu) {
                    t.extended += mu;
                } else {
                    t.spiky += 1.0;
                }
            }
            if (__builtin_fabs(mu) <= 0.75 * sigma) {
                const double r = (double)((i / list.w_used) * list.step * BLOCK);
                const double q = (double)((i % list.w_used) * list.step * BLOCK);
                t.sw += 1.0;
                t.sr += r;
                t.sq += q;
                t.sm += mu;
                t.srr += r * r;
                t.sqq += q * q;
                t.srq += r * q;
                t.srm += r * mu;
                t.sqm += q * mu;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        t.sw += __shfl_xor(t.sw, off);
        t.sr += __shfl_xor(t.sr, off);
        t.sq += __shfl_xor(t.sq, off);
        t.sm += __shfl_xor(t.sm, off);
        t.srr += __shfl_xor(t.srr, off);
        t.sqq += __shfl_xor(t.sqq, off);
        t.srq += __shfl_xor(t.srq, off);
        t.srm += __shfl_xor(t.srm, off);
        t.sqm += __shfl_xor(t.sqm, off);
        t.extended += __shfl_xor(t.extended, off);
        t.spiky += __shfl_xor(t.spiky, off);
    }
    return t;
}

// The tile's key bracket [c + sigma * o0, c + sigma * o1] (offsets ordered),
// and the standardised offsets themselves.
template <typename T>
__device__ inline void store_bracket(const BracketArgs &b, int64_t tile, double mu, double sigma, float out0,
                                     float out1)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    const float o0 = out0 < out1 ? out0 : out1;
    const float o1 = out0 < out1 ? out1 : out0;
    const double lo = mu + sigma * (double)o0;
    const double hi = mu + sigma * (double)o1;
    // The bracket is a key range of the dtype; rounding lo down and hi up
    // is not needed for exactness (a miss only costs a fallback).
    Key lo_key = KT::key(KT::from_double(lo));
    Key hi_key = KT::key(KT::from_double(hi));
    if (!(lo == lo) || !(hi == hi)) {
        lo_key = 0;
        hi_key = KT::SENTINEL;
    }
    static_cast<Key *>(b.lo_keys)[tile] = lo_key;
    static_cast<Key *>(b.hi_keys)[tile] = hi_key;
    b.predicted[tile * 2 + 0] = o0;
    b.predicted[tile * 2 + 1] = o1;
}

// Per tile, from the 8x8 blocks covering its support (clipped to the image):
//   c     = median of the block means       (robust level: sources move few blocks)
//   sigma = median of the block std devs    (robust noise scale)
//   (gy, gx) = least-squares plane slope of the block means over the blocks
//           within 0.75 sigma of c (sources excluded), per pixel
// Features (all scale free): asinh of the standardised excess of extended
// bright blocks per block (mean above c + sigma / 2, spread below 4 x the
// shift: sources), asinh(gy * ty / sigma),
// asinh(gx * 64 / sigma), retained fraction, target quantile, fraction of
// spiky bright blocks (spread at least 4 x the shift: isolated hot pixels).
// Targets are standardised as (value - c) / sigma.
//
// One wave per tile, four tiles per workgroup. With PREDICT the wave also
// runs the MLP (lane j < 32 computes hidden unit j; the sums run in the same
// order as mlp_forward's, so the outputs are the same bits) and writes the
// tile's key bracket; workgroup 0 also clears the counters at `reset` (the
// fallback count the bracket kernel appends to and its tile queues).
template <typename T, bool PREDICT>
__global__ void __launch_bounds__(256) k_tile_features(BracketArgs b, const double *blocks, const float *weights,
                                                       int64_t tiles)
{
    __shared__ float w[PREDICT ? PARAMS : 1];
    if (PREDICT && blockIdx.x == 0 && b.reset && (int)threadIdx.x < b.reset_count) b.reset[threadIdx.x] = 0;
    // Waves past the last tile compute the last tile again and store nothing
    // (every wave reaches the barrier before the MLP).
    const int64_t wave_tile = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const bool active = wave_tile < tiles;
    const int64_t tile = active ? wave_tile : tiles - 1;
    const int lane = threadIdx.x % 64;
    const int tyi = (int)(tile % b.tiles_y);
    const int tx = (int)(tile / b.tiles_y);
    int64_t r0 = (int64_t)tyi * b.ty + b.low_y, r1 = (int64_t)tyi * b.ty + b.ty - 1 + b.high_y;
    int64_t c0 = (int64_t)tx * 64 + b.low_x, c1 = (int64_t)tx * 64 + 63 + b.high_x;
    r0 = r0 < 0 ? 0 : r0;
    c0 = c0 < 0 ? 0 : c0;
    r1 = r1 >= b.ny ? b.ny - 1 : r1;
    c1 = c1 >= b.nx ? b.nx - 1 : c1;
    BlockList list;
    list.br0 = r0 / BLOCK;
    list.bc0 = c0 / BLOCK;
    list.bx = b.bx;
    const int height = (int)(r1 / BLOCK - list.br0 + 1), width = (int)(c1 / BLOCK - list.bc0 + 1);
    // Subsample block rows/columns when a support spans more blocks than
    // MAX_TILE_BLOCKS.
    int step = 1;
    while (((height + step - 1) / step) * ((width + step - 1) / step) > MAX_TILE_BLOCKS) {
        step <<= 1;
    }
    list.step = step;
    list.w_used = (width + step - 1) / step;
    list.count = ((height + step - 1) / step) * list.w_used;
    static_assert(MAX_TILE_BLOCKS == 4 * 64, "tile_sums holds at most 4 blocks per lane");
    TileSums t;
    if (list.count <= 64) {
        t = tile_sums<1>(blocks, list, lane);
    } else if (list.count <= 128) {
        t = tile_sums<2>(blocks, list, lane);
    } else {
        t = tile_sums<4>(blocks, list, lane);
    }
    const float c = t.c, sigma = t.sigma;
    const double n = t.n, cells = t.cells;
    const int usable = t.usable;
    const double sw = t.sw, sr = t.sr, sq = t.sq, sm = t.sm, srr = t.srr, sqq = t.sqq, srq = t.srq, srm = t.srm,
                 sqm = t.sqm, extended = t.extended, spiky = t.spiky;
    // Every lane holds the totals; each computes the features.
    double gy = 0.0, gx = 0.0;
    if (sw >= 3.0) {
        const double vrr = srr - sr * sr / sw, vqq = sqq - sq * sq / sw, vrq = srq - sr * sq / sw;
        const double vrm = srm - sr * sm / sw, vqm = sqm - sq * sm / sw;
        const double det = vrr * vqq - vrq * vrq;
        if (det > 1e-9 * (vrr * vqq + 1e-30)) {
            gy = (vrm * vqq - vqm * vrq) / det;
            gx = (vqm * vrr - vrm * vrq) / det;
        } else if (vrr > 0.0) {
            gy = vrm / vrr;
        } else if (vqq > 0.0) {
            gx = vqm / vqq;
        }
    }
    const double sg = (double)sigma;
    // The three asinh features run on lanes 0, 1, 2 (one evaluation per
    // wave instead of three), then are broadcast.
    const double arg = lane == 0   ? (usable > 0 ? extended / sg / (double)usable : 0.0)
                       : lane == 1 ? gy * (double)b.ty / sg
                                   : gx * 64.0 / sg;
    const float shaped = feature_asinh(arg);
    float x[FEATURES];
    x[0] = __shfl(shaped, 0);
    x[1] = __shfl(shaped, 1);
    x[2] = __shfl(shaped, 2);
    x[3] = cells > 0.0 ? (float)(n / cells) : 0.0f;
    x[4] = b.quantile;
    x[5] = usable > 0 ? (float)(spiky / (double)usable) : 0.0f;
    if (active && lane == 0) {
        float *f = b.features + tile * FEATURES;
        for (int k = 0; k < FEATURES; ++k) f[k] = x[k];
        double *m = b.moments + tile * 3;
        m[0] = (double)c;
        m[1] = sg;
        m[2] = n;
    }
    if (!PREDICT) {
        return;
    }
    // Weights staged after the features, so their loads overlap the block
    // loads of the other waves.
    for (int i = threadIdx.x; i < PARAMS; i += blockDim.x) w[i] = weights[i];
    __syncthreads();
    const int j = lane % HIDDEN;
    float acc = w[B1_OFF + j];
    for (int i = 0; i < FEATURES; ++i) acc += x[i] * w[W1_OFF + i * HIDDEN + j];
    const float h1 = acc > 0.0f ? acc : 0.0f;
    acc = w[B2_OFF + j];
#pragma unroll
    for (int i = 0; i < HIDDEN; ++i) acc += lane_value(h1, i) * w[W2_OFF + i * HIDDEN + j];
    const float h2 = acc > 0.0f ? acc : 0.0f;
    const int k = lane % 2;
    acc = w[B3_OFF + k];
#pragma unroll
    for (int i = 0; i < HIDDEN; ++i) acc += lane_value(h2, i) * w[W3_OFF + i * 2 + k];
    const float out0 = lane_value(acc, 0), out1 = lane_value(acc, 1);
    if (active && lane == 0) {
        store_bracket<T>(b, tile, (double)c, sg, out0, out1);
    }
}

// MLP forward for one sample; weights in the documented flat layout.
__device__ inline void mlp_forward(const float *w, const float *x, float h1[HIDDEN], float h2[HIDDEN], float out[2])
{
    const float *w1 = w + W1_OFF, *b1 = w + B1_OFF, *w2 = w + W2_OFF, *b2 = w + B2_OFF, *w3 = w + W3_OFF,
                *b3 = w + B3_OFF;
    for (int j = 0; j < HIDDEN; ++j) {
        float acc = b1[j];
        for (int i = 0; i < FEATURES; ++i) {
            acc += x[i] * w1[i * HIDDEN + j];
        }
        h1[j] = acc > 0.0f ? acc : 0.0f;
    }
    for (int j = 0; j < HIDDEN; ++j) {
        float acc = b2[j];
        for (int i = 0; i < HIDDEN; ++i) {
            acc += h1[i] * w2[i * HIDDEN + j];
        }
        h2[j] = acc > 0.0f ? acc : 0.0f;
    }
    for (int j = 0; j < 2; ++j) {
        float acc = b3[j];
        for (int i = 0; i < HIDDEN; ++i) {
            acc += h2[i] * w3[i * 2 + j];
        }
        out[j] = acc;
    }
}

// Per tile: standardised smallest and largest non-NaN result of the tile's
// outputs; tiles without a result get a zero weight.
template <typename T>
__global__ void __launch_bounds__(64) k_tile_labels(BracketArgs b, const T *output, float *targets, float *weight)
{
    using KT = KeyTraits<T>;
    const int tile = blockIdx.x;
    const int lane = threadIdx.x;
    const int tyi = tile % b.tiles_y;
    const int tx = tile / b.tiles_y;
    const int64_t x = (int64_t)tx * 64 + lane;
    double mn = __builtin_huge_val(), mx = -__builtin_huge_val();
    if (x < b.nx) {
        for (int r = 0; r < b.ty; ++r) {
            const int64_t y = (int64_t)tyi * b.ty + r;
            if (y >= b.ny) {
                break;
            }
            const double v = KT::to_double(KT::bits(output[y * b.nx + x]));
            if (v == v) {
                mn = v < mn ? v : mn;
                mx = v > mx ? v : mx;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double a = __shfl_xor(mn, off), c = __shfl_xor(mx, off);
        mn = a < mn ? a : mn;
        mx = c > mx ? c : mx;
    }
    if (lane == 0) {
        const double mu = b.moments[tile * 3 + 0];
        const double sigma = b.moments[tile * 3 + 1];
        const bool ok = mn <= mx && __builtin_isfinite(mn) && __builtin_isfinite(mx);
        targets[tile * 2 + 0] = ok ? (float)((mn - mu) / sigma) : 0.0f;
        targets[tile * 2 + 1] = ok ? (float)((mx - mu) / sigma) : 0.0f;
        weight[tile] = ok ? 1.0f : 0.0f;
    }
}

// One thread per sample: forward, pinball-loss gradient, backward; the
// gradient is summed per block in LDS (float atomics) and written to
// partial[block][PARAMS].
__global__ void __launch_bounds__(256) k_mlp_grad(const float *weights, const float *features, const float *targets,
                                                 const float *weight, const int32_t *index, int64_t count,
                                                 float tau_lo, float tau_hi, float *partial, float *loss_partial)
{
    __shared__ float w[PARAMS];
    __shared__ float g[PARAMS];
    __shared__ float loss_sum;
    for (int i = threadIdx.x; i < PARAMS; i += blockDim.x) {
        w[i] = weights[i];
        g[i] = 0.0f;
    }
    if (threadIdx.x == 0) {
        loss_sum = 0.0f;
    }
    __syncthreads();
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) {
        const int64_t s = index ? index[k] : k;
        const float *x = features + s * FEATURES;
        const float wt = weight[s];
        float h1[HIDDEN], h2[HIDDEN], out[2];
        mlp_forward(w, x, h1, h2, out);
        // Pinball loss: quantile tau_lo of the lower target, tau_hi of the upper.
        float dout[2], loss = 0.0f;
        const float taus[2] = {tau_lo, tau_hi};
        for (int j = 0; j < 2; ++j) {
            const float r = targets[s * 2 + j] - out[j];
            loss += wt * (r >= 0.0f ? taus[j] * r : (taus[j] - 1.0f) * r);
            dout[j] = wt * (r >= 0.0f ? -taus[j] : 1.0f - taus[j]);
        }
        atomicAdd(&loss_sum, loss);
        float dh2[HIDDEN];
        for (int i = 0; i < HIDDEN; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < 2; ++j) {
                atomicAdd(&g[W3_OFF + i * 2 + j], h2[i] * dout[j]);
                acc += w[W3_OFF + i * 2 + j] * dout[j];
            }
            dh2[i] = h2[i] > 0.0f ? acc : 0.0f;
        }
        for (int j = 0; j < 2; ++j) {
            atomicAdd(&g[B3_OFF + j], dout[j]);
        }
        float dh1[HIDDEN];
        for (int i = 0; i < HIDDEN; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < HIDDEN; ++j) {
                if (dh2[j] != 0.0f) {
                    atomicAdd(&g[W2_OFF + i * HIDDEN + j], h1[i] * dh2[j]);
                }
                acc += w[W2_OFF + i * HIDDEN + j] * dh2[j];
            }
            dh1[i] = h1[i] > 0.0f ? acc : 0.0f;
        }
        for (int j = 0; j < HIDDEN; ++j) {
            if (dh2[j] != 0.0f) {
                atomicAdd(&g[B2_OFF + j], dh2[j]);
            }
            if (dh1[j] != 0.0f) {
                atomicAdd(&g[B1_OFF + j], dh1[j]);
                for (int i = 0; i < FEATURES; ++i) {
                    atomicAdd(&g[W1_OFF + i * HIDDEN + j], x[i] * dh1[j]);
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < PARAMS; i += blockDim.x) {
        partial[(int64_t)blockIdx.x * PARAMS + i] = g[i];
    }
    if (threadIdx.x == 0) {
        loss_partial[blockIdx.x] = loss_sum;
    }
}

// One thread per parameter: sum the partial gradients (fixed order), then an
// Adam step; thread 0 also sums the loss.
__global__ void __launch_bounds__(256) k_mlp_adam(float *weights, float *m1, float *m2, const float *partial,
                                                 const float *loss_partial, int blocks, float inv_count, float lr,
                                                 float beta1, float beta2, float eps, int step, float *loss_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < PARAMS) {
        float grad = 0.0f;
        for (int b = 0; b < blocks; ++b) {
            grad += partial[(int64_t)b * PARAMS + i];
        }
        grad *= inv_count;
        m1[i] = beta1 * m1[i] + (1.0f - beta1) * grad;
        m2[i] = beta2 * m2[i] + (1.0f - beta2) * grad * grad;
        const float c1 = 1.0f - __builtin_powf(beta1, (float)step);
        const float c2 = 1.0f - __builtin_powf(beta2, (float)step);
        weights[i] -= lr * (m1[i] / c1) / (__builtin_sqrtf(m2[i] / c2) + eps);
    }
    if (i == 0 && loss_out) {
        float total = 0.0f;
        for (int b = 0; b < blocks; ++b) {
            total += loss_partial[b];
        }
        *loss_out = total * inv_count;
    }
}

}  // namespace

hipError_t launch_block_stats(int dtype_bytes, const void *image, int64_t ny, int64_t nx, int ignore_inf,
                              double *blocks, hipStream_t stream)
{
    const int64_t by = (ny + BLOCK - 1) / BLOCK, bx = (nx + BLOCK - 1) / BLOCK;
    const dim3 grid((unsigned)((by * bx + 255) / 256));
    if (dtype_bytes == 4) {
        hipLaunchKernelGGL(k_block_stats<float>, grid, dim3(256), 0, stream, static_cast<const float *>(image), ny, nx,
                           by, bx, ignore_inf, blocks);
    } else {
        hipLaunchKernelGGL(k_block_stats<double>, grid, dim3(256), 0, stream, static_cast<const double *>(image), ny,
                           nx, by, bx, ignore_inf, blocks);
    }
    return hipGetLastError();
}

hipError_t launch_tile_features(const BracketArgs &args, const double *blocks, int64_t tiles, hipStream_t stream)
{
    hipLaunchKernelGGL((k_tile_features<float, false>), dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, stream, args,
                       blocks, nullptr, tiles);
    return hipGetLastError();
}

hipError_t launch_bracket_proposal(int dtype_bytes, const BracketArgs &args, const double *blocks,
                                   const float *weights, int64_t tiles, hipStream_t stream)
{
    const dim3 grid((unsigned)((tiles + 3) / 4));
    if (dtype_bytes == 4) {
        hipLaunchKernelGGL((k_tile_features<float, true>), grid, dim3(256), 0, stream, args, blocks, weights, tiles);
    } else {
        hipLaunchKernelGGL((k_tile_features<double, true>), grid, dim3(256), 0, stream, args, blocks, weights, tiles);
    }
    return hipGetLastError();
}

hipError_t launch_tile_labels(int dtype_bytes, const BracketArgs &args, const void *output, float *targets,
                              float *weight, int64_t tiles, hipStream_t stream)
{
    if (dtype_bytes == 4) {
        hipLaunchKernelGGL(k_tile_labels<float>, dim3((unsigned)tiles), dim3(64), 0, stream, args,
                           static_cast<const float *>(output), targets, weight);
    } else {
        hipLaunchKernelGGL(k_tile_labels<double>, dim3((unsigned)tiles), dim3(64), 0, stream, args,
                           static_cast<const double *>(output), targets, weight);
    }
    return hipGetLastError();
}

int mlp_grad_blocks(int64_t count) { return (int)((count + 255) / 256); }

hipError_t launch_mlp_step(float *weights, float *m1, float *m2, const float *features, const float *targets,
                           const float *weight, const int32_t *index, int64_t count, float tau_lo, float tau_hi,
                           float lr, int step, float *partial, float *loss_partial, float *loss_out,
                           hipStream_t stream)
{
    const int blocks = mlp_grad_blocks(count);
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)blocks), dim3(256), 0, stream, weights, features, targets, weight,
                       index, count, tau_lo, tau_hi, partial, loss_partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        return e;
    }
    hipLaunchKernelGGL(k_mlp_adam, dim3((PARAMS + 255) / 256), dim3(256), 0, stream, weights, m1, m2, partial,
                       loss_partial, blocks, 1.0f / (float)(count > 0 ? count : 1), lr, 0.9f, 0.999f, 1e-8f, step,
                       loss_out);
    return hipGetLastError();
}

}  // namespace ndf
