// Median filter of mesh-background grids (the grid stage of the reference's
// mesh background, MeshBackground::filter, src/background.rs:797-855).
//
// A mesh grid holds one float64 background and one RMS value per mesh. Each
// mesh takes the median of the meshes in its window: the window is centred
// and shrinks at the grid edge (SExtractor) or is truncated there
// (photutils); a window holding a NaN gives NaN; an even count averages the
// middle two, halving first when the sum overflows (median_of /
// median_sorted, src/background.rs:940-963). A threshold can restrict the
// replacement to meshes whose median deviates enough, or whose background
// is above a value. Grids are small (one value per mesh), so one thread
// evaluates one mesh: it selects the middle ranks by a bitwise search over
// the total-order keys of its window, which needs no per-thread array.
#include <stdio.h>

#include "../../include/ndfgpu.h"
#include "ndf_device.hpp"
#include "ndf_internal.hpp"

namespace ndf {
namespace {

struct MeshArgs {
    const double *background, *rms;
    double *out_background, *out_rms;
    int64_t len;
    int32_t ndim;
    int64_t shape[MAX_RANK];
    int64_t stride[MAX_RANK];
    int64_t half[MAX_RANK];
    int32_t edge;
    int32_t threshold_kind;
    double threshold;
};

struct MeshWindow {
    int64_t low[MAX_RANK], width[MAX_RANK];
    int64_t count;
};

__device__ inline int64_t window_offset(const MeshArgs &a, const MeshWindow &w, int64_t position)
{
    int64_t offset = 0;
    for (int d = a.ndim; d-- > 0;) {
        const int64_t q = position / w.width[d];
        offset += (w.low[d] + (position - q * w.width[d])) * a.stride[d];
        position = q;
    }
    return offset;
}

// Zero-based rank k of the window's keys (no NaN present).
__device__ inline uint64_t window_kth(const MeshArgs &a, const MeshWindow &w, const double *grid, int64_t k)
{
    uint64_t answer = 0;
    for (int bit = 63; bit >= 0; --bit) {
        const uint64_t candidate = answer | (1ull << bit);
        int64_t below = 0;
        for (int64_t p = 0; p < w.count; ++p) {
            below += KeyTraits<double>::key(__builtin_bit_cast(uint64_t, grid[window_offset(a, w, p)])) < candidate;
        }
        if (below <= k) answer = candidate;
    }
    return answer;
}

__device__ inline double window_median(const MeshArgs &a, const MeshWindow &w, const double *grid)
{
    for (int64_t p = 0; p < w.count; ++p) {
        const double v = grid[window_offset(a, w, p)];
        if (v != v) return __builtin_nan("");
    }
    using KT = KeyTraits<double>;
    const int64_t n = w.count;
    if (n & 1) return KT::from_bits(KT::unkey(window_kth(a, w, grid, n / 2)));
    const double lo = KT::from_bits(KT::unkey(window_kth(a, w, grid, n / 2 - 1)));
    const double hi = KT::from_bits(KT::unkey(window_kth(a, w, grid, n / 2)));
    const double mid = (lo + hi) / 2.0;
    return __builtin_isfinite(mid) ? mid : lo / 2.0 + hi / 2.0;
}

__global__ void __launch_bounds__(256) k_mesh_filter(MeshArgs a)
{
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t flat = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; flat < a.len; flat += step) {
        MeshWindow w;
        w.count = 1;
        int64_t rest = flat;
        for (int d = a.ndim; d-- > 0;) {
            const int64_t n = a.shape[d];
            const int64_t k = rest % n;
            rest /= n;
            int64_t half = a.half[d];
            int64_t low, high;
            if (a.edge == NDF_MESH_SHRINK) {
                half = half < k ? half : k;
                half = half < n - 1 - k ? half : n - 1 - k;
                low = k - half;
                high = k + half;
            } else {
                low = k > half ? k - half : 0;
                high = k + half < n - 1 ? k + half : n - 1;
            }
            w.low[d] = low;
            w.width[d] = high - low + 1;
            w.count *= high - low + 1;
        }
        const double own = a.background[flat];
        const double median = window_median(a, w, a.background);
        bool replace = true;
        if (a.threshold_kind == NDF_MESH_DEVIATION) replace = __builtin_fabs(median - own) >= a.threshold;
        if (a.threshold_kind == NDF_MESH_ABOVE) replace = own > a.threshold;
        a.out_background[flat] = replace ? median : own;
        a.out_rms[flat] = replace ? window_median(a, w, a.rms) : a.rms[flat];
    }
}

ndf_status mesh_fail(ndf_status code, const char *message, char *err, size_t err_len)
{
    if (err && err_len) snprintf(err, err_len, "%s", message);
    return code;
}

}  // namespace
}  // namespace ndf

extern "C" ndf_status ndf_mesh_median_filter(const double *background, const double *rms, double *out_background,
                                             double *out_rms, size_t ndim, const size_t *mesh_shape,
                                             const size_t *filter_size, int32_t edge, int32_t threshold_kind,
                                             double threshold, char *err, size_t err_len)
{
    using namespace ndf;
    if (ndim == 0 || ndim > (size_t)MAX_RANK) {
        return mesh_fail(NDF_ERR_INVALID, "mesh grids need 1 to 32 dimensions", err, err_len);
    }
    if (edge != NDF_MESH_SHRINK && edge != NDF_MESH_TRUNCATE) {
        return mesh_fail(NDF_ERR_INVALID, "filter_edge must be 'shrink' or 'truncate'", err, err_len);
    }
    if (threshold_kind < NDF_MESH_ALL || threshold_kind > NDF_MESH_ABOVE) {
        return mesh_fail(NDF_ERR_INVALID, "unknown filter threshold kind", err, err_len);
    }
    // MeshConfig::validate (src/background.rs:271-277).
    if (threshold_kind != NDF_MESH_ALL && !__builtin_isfinite(threshold)) {
        return mesh_fail(NDF_ERR_INVALID, "filter threshold must be finite", err, err_len);
    }
    MeshArgs a{};
    a.ndim = (int32_t)ndim;
    a.edge = edge;
    a.threshold_kind = threshold_kind;
    a.threshold = threshold;
    size_t len = 1;
    bool identity = true;
    for (size_t d = 0; d < ndim; ++d) {
        if (filter_size[d] == 0) return mesh_fail(NDF_ERR_INVALID, "filter sizes must be positive", err, err_len);
        if (mesh_shape[d] != 0 && len > SIZE_MAX / mesh_shape[d]) {
            return mesh_fail(NDF_ERR_OVERFLOW, "shape or index arithmetic overflowed", err, err_len);
        }
        len *= mesh_shape[d];
        a.shape[d] = (int64_t)mesh_shape[d];
        a.half[d] = (int64_t)(filter_size[d] / 2);
        identity &= filter_size[d] == 1;
    }
    int64_t stride = 1;
    for (size_t d = ndim; d-- > 0;) {
        a.stride[d] = stride;
        stride *= a.shape[d];
    }
    a.len = (int64_t)len;
    // An all-ones filter, or an empty grid, leaves the grids unchanged
    // (src/background.rs:798-800).
    if (identity || len == 0) {
        for (size_t i = 0; i < len; ++i) {
            out_background[i] = background[i];
            out_rms[i] = rms[i];
        }
        return NDF_OK;
    }
    const size_t bytes = len * sizeof(double);
    double *buffers[4] = {nullptr, nullptr, nullptr, nullptr};
    hipError_t e = hipSuccess;
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipMalloc(&buffers[i], bytes);
    if (e == hipSuccess) e = hipMemcpy(buffers[0], background, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(buffers[1], rms, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        a.background = buffers[0];
        a.rms = buffers[1];
        a.out_background = buffers[2];
        a.out_rms = buffers[3];
        const int64_t blocks = ((int64_t)len + 255) / 256;
        hipLaunchKernelGGL(k_mesh_filter, dim3((unsigned)(blocks < 65535 ? blocks : 65535)), dim3(256), 0, nullptr, a);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out_background, buffers[2], bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out_rms, buffers[3], bytes, hipMemcpyDeviceToHost);
    for (double *b : buffers) {
        if (b) (void)hipFree(b);
    }
    if (e != hipSuccess) {
        char message[160];
        snprintf(message, sizeof message, "HIP error %d: %s", (int)e, hipGetErrorString(e));
        return mesh_fail(NDF_ERR_DEVICE, message, err, err_len);
    }
    return NDF_OK;
}
