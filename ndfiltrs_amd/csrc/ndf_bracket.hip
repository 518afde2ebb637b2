// Learned bracket proposal for the slide engine (2-D images).
//
// A tile's in-bracket samples are the only ones the bracket kernel sorts and
// walks. The bracket comes from a tiny MLP (6 -> 32 -> 32 -> 2, ReLU, fp32)
// that reads cheap box-sum features of the tile's support:
//   1. k_block_stats: one pass over the image, per 8x8 block: retained
//      count, sum, sum of squares (f64), min, max;
//   2. k_tile_features: per tile, robust level c and noise scale sigma
//      (medians of the block means and block std devs), a plane slope of
//      the block means, the retained fraction and the fraction of bright
//      blocks, standardised so that the features are scale free;
//   3. k_bracket_predict: the MLP's two outputs are standardised offsets:
//      bracket = [c + sigma*o0, c + sigma*o1], converted to keys.
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
constexpr int MAX_TILE_BLOCKS = 1024;

// One thread per 8x8 block; consecutive threads take consecutive blocks of
// a block row, so each image row segment is read by 64 lanes as 2 KB of
// contiguous memory. Output per block: retained count, sum, sum of squares
// (f64), accumulated in row-major order within the block. A whole block
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
    double *o = out + block * 3;
    o[0] = n;
    o[1] = s;
    o[2] = q;
}

// asinh(v) = sign(v) * log(|v| + sqrt(v^2 + 1)); log(2|v|) once v^2 would overflow.
__device__ inline float feature_asinh(double v)
{
    const double a = v < 0.0 ? -v : v;
    const double r = a > 1e150 ? log(2.0 * a) : log(a + sqrt(a * a + 1.0));
    return (float)(v < 0.0 ? -r : r);
}

// Ascending bitonic sort of v[0 .. n) in LDS by one 64-lane workgroup
// (n a power of two, 64 <= n <= MAX_TILE_BLOCKS).
__device__ void wave_sort(float *v, int n, int lane)
{
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = lane; i < n / 2; i += 64) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool asc = (lo & size) == 0;
                const float a = v[lo], b = v[hi];
                const bool swap = asc ? (b < a) : (a < b);
                v[lo] = swap ? b : a;
                v[hi] = swap ? a : b;
            }
            __syncthreads();
        }
    }
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
__global__ void __launch_bounds__(64) k_tile_features(BracketArgs b, const double *blocks)
{
    __shared__ float means[MAX_TILE_BLOCKS];
    __shared__ float sds[MAX_TILE_BLOCKS];
    const int tile = blockIdx.x;
    const int lane = threadIdx.x;
    const int tyi = tile % b.tiles_y;
    const int tx = tile / b.tiles_y;
    int64_t r0 = (int64_t)tyi * b.ty + b.low_y, r1 = (int64_t)tyi * b.ty + b.ty - 1 + b.high_y;
    int64_t c0 = (int64_t)tx * 64 + b.low_x, c1 = (int64_t)tx * 64 + 63 + b.high_x;
    r0 = r0 < 0 ? 0 : r0;
    c0 = c0 < 0 ? 0 : c0;
    r1 = r1 >= b.ny ? b.ny - 1 : r1;
    c1 = c1 >= b.nx ? b.nx - 1 : c1;
    const int64_t br0 = r0 / BLOCK, bc0 = c0 / BLOCK;
    const int height = (int)(r1 / BLOCK - br0 + 1), width = (int)(c1 / BLOCK - bc0 + 1);
    // Subsample block rows/columns when a support spans more blocks than the
    // sort holds.
    int step = 1;
    while (((height + step - 1) / step) * ((width + step - 1) / step) > MAX_TILE_BLOCKS) {
        step <<= 1;
    }
    const int h_used = (height + step - 1) / step, w_used = (width + step - 1) / step;
    const int count = h_used * w_used;
    int n_sort = 64;
    while (n_sort < count) {
        n_sort <<= 1;
    }
    double n = 0.0, cells = 0.0;
    int usable = 0;
    for (int i = lane; i < n_sort; i += 64) {
        float mean = __builtin_huge_valf(), sd = __builtin_huge_valf();
        if (i < count) {
            const int64_t br = br0 + (int64_t)(i / w_used) * step, bc = bc0 + (int64_t)(i % w_used) * step;
            const double *o = blocks + (br * b.bx + bc) * 3;
            n += o[0];
            cells += 64.0;
            if (o[0] >= 16.0) {
                const double mu = o[1] / o[0];
                const double var = o[2] / o[0] - mu * mu;
                mean = (float)mu;
                sd = (float)(var > 0.0 ? __builtin_sqrt(var) : 0.0);
                ++usable;
            }
        }
        means[i] = mean;
        sds[i] = sd;
    }
    for (int off = 32; off > 0; off >>= 1) {
        n += __shfl_xor(n, off);
        cells += __shfl_xor(cells, off);
        usable += __shfl_xor(usable, off);
    }
    __syncthreads();
    wave_sort(means, n_sort, lane);
    wave_sort(sds, n_sort, lane);
    float c = 0.0f, sigma = 1.0f;
    if (usable > 0) {
        c = means[(usable - 1) / 2];
        sigma = sds[(usable - 1) / 2];
    }
    const float floor_sigma = 1e-6f * (__builtin_fabsf(c) + 1e-30f);
    sigma = sigma > floor_sigma ? sigma : floor_sigma;
    // Plane fit over unsorted block means (re-read from the block table).
    double sw = 0, sr = 0, sq = 0, sm = 0, srr = 0, sqq = 0, srq = 0, srm = 0, sqm = 0, extended = 0, spiky = 0;
    for (int i = lane; i < count; i += 64) {
        const int64_t br = br0 + (int64_t)(i / w_used) * step, bc = bc0 + (int64_t)(i % w_used) * step;
        const double *o = blocks + (br * b.bx + bc) * 3;
        if (o[0] >= 16.0) {
            const double mu = o[1] / o[0] - (double)c;
            // A bright block is "spiky" when its spread dwarfs its mean
            // shift (one hot pixel: sd ~ 8 x shift) and "extended" otherwise
            // (a source covering the block); only extended brightness moves
            // window ranks much.
            if (mu > 0.5 * (double)sigma) {
                const double var = o[2] / o[0] - (o[1] / o[0]) * (o[1] / o[0]);
                const double sd = var > 0.0 ? __builtin_sqrt(var) : 0.0;
                if (sd < 4.0 * mu) {
                    extended += mu / (double)sigma;
                } else {
                    spiky += 1.0;
                }
            }
            if (__builtin_fabs(mu) <= 0.75 * (double)sigma) {
                const double r = (double)((i / w_used) * step * BLOCK), q = (double)((i % w_used) * step * BLOCK);
                sw += 1.0;
                sr += r;
                sq += q;
                sm += mu;
                srr += r * r;
                sqq += q * q;
                srq += r * q;
                srm += r * mu;
                sqm += q * mu;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        sw += __shfl_xor(sw, off);
        sr += __shfl_xor(sr, off);
        sq += __shfl_xor(sq, off);
        sm += __shfl_xor(sm, off);
        srr += __shfl_xor(srr, off);
        sqq += __shfl_xor(sqq, off);
        srq += __shfl_xor(srq, off);
        srm += __shfl_xor(srm, off);
        sqm += __shfl_xor(sqm, off);
        extended += __shfl_xor(extended, off);
        spiky += __shfl_xor(spiky, off);
    }
    if (lane == 0) {
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
        float *f = b.features + (int64_t)tile * FEATURES;
        f[0] = usable > 0 ? feature_asinh(extended / (double)usable) : 0.0f;
        f[1] = feature_asinh(gy * (double)b.ty / sg);
        f[2] = feature_asinh(gx * 64.0 / sg);
        f[3] = cells > 0.0 ? (float)(n / cells) : 0.0f;
        f[4] = b.quantile;
        f[5] = usable > 0 ? (float)(spiky / (double)usable) : 0.0f;
        double *m = b.moments + (int64_t)tile * 3;
        m[0] = (double)c;
        m[1] = sg;
        m[2] = n;
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

template <typename T>
__global__ void __launch_bounds__(64) k_bracket_predict(BracketArgs b, const float *weights, int64_t tiles)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    __shared__ float w[PARAMS];
    for (int i = threadIdx.x; i < PARAMS; i += blockDim.x) {
        w[i] = weights[i];
    }
    __syncthreads();
    const int64_t tile = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tile >= tiles) {
        return;
    }
    float h1[HIDDEN], h2[HIDDEN], out[2];
    mlp_forward(w, b.features + tile * FEATURES, h1, h2, out);
    const double mu = b.moments[tile * 3 + 0];
    const double sigma = b.moments[tile * 3 + 1];
    const float o0 = out[0] < out[1] ? out[0] : out[1];
    const float o1 = out[0] < out[1] ? out[1] : out[0];
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
    hipLaunchKernelGGL(k_tile_features, dim3((unsigned)tiles), dim3(64), 0, stream, args, blocks);
    return hipGetLastError();
}

hipError_t launch_bracket_predict(int dtype_bytes, const BracketArgs &args, const float *weights, int64_t tiles,
                                  hipStream_t stream)
{
    const dim3 grid((unsigned)((tiles + 63) / 64));
    if (dtype_bytes == 4) {
        hipLaunchKernelGGL(k_bracket_predict<float>, grid, dim3(64), 0, stream, args, weights, tiles);
    } else {
        hipLaunchKernelGGL(k_bracket_predict<double>, grid, dim3(64), 0, stream, args, weights, tiles);
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
