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

template <typename T>
struct Gather {
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    const GenericArgs &a;
    int64_t coord[MAX_RANK];

    __device__ Gather(const GenericArgs &args, int64_t index) : a(args)
    {
        for (int ax = 0; ax < a.rank; ++ax) {
            coord[ax] = index / a.stride[ax];
            index -= coord[ax] * a.stride[ax];
        }
    }

    __device__ Key key(int t) const
    {
        int64_t offset = 0;
        const int32_t *tap = a.taps + (int64_t)t * a.rank;
        for (int ax = 0; ax < a.rank; ++ax) {
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
template <typename T>
__device__ typename KeyTraits<T>::Key select_kth(const Gather<T> &g, int count, int64_t k)
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

template <typename T>
__global__ void __launch_bounds__(256) k_generic(GenericArgs a)
{
    using KT = KeyTraits<T>;
    using Key = typename KT::Key;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t index = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; index < a.len; index += step) {
        Gather<T> g(a, index);
        int64_t n = 0;
        for (int t = 0; t < a.count; ++t) {
            n += g.key(t) != KT::SENTINEL;
        }
        Key out = KT::CANONICAL_NAN;
        if (n >= a.min_count && n > 0) {
            if (a.stat == STAT_RANK) {
                if (a.rank_param < n) {
                    out = KT::unkey(select_kth<T>(g, a.count, a.rank_param));
                }
            } else if (n & 1) {
                out = KT::unkey(select_kth<T>(g, a.count, n / 2));
            } else {
                const Key lower = select_kth<T>(g, a.count, n / 2 - 1);
                const Key upper = select_kth<T>(g, a.count, n / 2);
                out = median_pair_bits<T>(lower, upper);
            }
        }
        static_cast<Key *>(a.output)[index] = out;
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

int64_t grid_for(int64_t len)
{
    const int64_t blocks = (len + 255) / 256;
    return blocks < 65536 ? (blocks < 1 ? 1 : blocks) : 65536;
}

}  // namespace

hipError_t launch_generic(int dtype_bytes, const GenericArgs &args, hipStream_t stream)
{
    const dim3 grid((unsigned)grid_for(args.len));
    if (dtype_bytes == 4) {
        hipLaunchKernelGGL(k_generic<float>, grid, dim3(256), 0, stream, args);
    } else {
        hipLaunchKernelGGL(k_generic<double>, grid, dim3(256), 0, stream, args);
    }
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

}  // namespace ndf
