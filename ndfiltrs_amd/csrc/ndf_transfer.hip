// Host <-> device copies of large arrays through pinned staging buffers.
//
// hipMemcpy from pageable memory moves about 25-30 GB/s on the MI355X box
// and D2H into pageable memory about 16 GB/s. Here an array is cut into
// chunks that alternate between two pinned buffers: the host threads copy
// chunk k between the caller's memory and one buffer while the DMA engine
// moves chunk k - 1 (upload) or k + 1 (download) through the other, and
// the host copy of each chunk is split across a small persistent thread
// pool, so first-touch page faults of a fresh output array are taken in
// parallel as well.
#include <hip/hip_runtime.h>
#include <string.h>

#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "ndf_internal.hpp"

namespace ndf {
namespace {

// A fixed set of worker threads that run one parallel copy at a time.
class CopyPool {
public:
    static CopyPool &get()
    {
        static CopyPool pool;
        return pool;
    }

    // Copies `bytes` from src to dst with every worker and the caller.
    void copy(void *dst, const void *src, size_t bytes)
    {
        const size_t parts = workers_.size() + 1;
        if (bytes < ((size_t)1 << 20) || parts == 1) {
            memcpy(dst, src, bytes);
            return;
        }
        std::unique_lock<std::mutex> lock(run_mutex_);  // one parallel copy at a time
        const size_t step = ((bytes + parts - 1) / parts + 4095) & ~(size_t)4095;
        {
            std::lock_guard<std::mutex> guard(mutex_);
            dst_ = static_cast<char *>(dst);
            src_ = static_cast<const char *>(src);
            bytes_ = bytes;
            step_ = step;
            pending_ = workers_.size();
            ++generation_;
        }
        wake_.notify_all();
        copy_part(0);
        std::unique_lock<std::mutex> wait(mutex_);
        idle_.wait(wait, [&] { return pending_ == 0; });
    }

private:
    CopyPool()
    {
        unsigned n = std::thread::hardware_concurrency();
        n = n == 0 ? 1 : (n > 16 ? 16 : n);
        for (unsigned i = 1; i < n; ++i) workers_.emplace_back([this, i] { loop(i); });
    }

    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> guard(mutex_);
            stop_ = true;
        }
        wake_.notify_all();
        for (std::thread &t : workers_) t.join();
    }

    void copy_part(size_t index)
    {
        const size_t begin = index * step_;
        if (begin >= bytes_) return;
        const size_t end = begin + step_ < bytes_ ? begin + step_ : bytes_;
        memcpy(dst_ + begin, src_ + begin, end - begin);
    }

    void loop(size_t index)
    {
        size_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lock(mutex_);
                wake_.wait(lock, [&] { return stop_ || generation_ != seen; });
                if (stop_) return;
                seen = generation_;
            }
            copy_part(index);
            std::lock_guard<std::mutex> guard(mutex_);
            if (--pending_ == 0) idle_.notify_one();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex mutex_, run_mutex_;
    std::condition_variable wake_, idle_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0, step_ = 0, pending_ = 0, generation_ = 0;
    bool stop_ = false;
};

}  // namespace

Staging::~Staging()
{
    for (int i = 0; i < 2; ++i) {
        if (done[i]) (void)hipEventDestroy(done[i]);
        if (pinned[i]) (void)hipHostFree(pinned[i]);
    }
}

hipError_t Staging::ensure()
{
    for (int i = 0; i < 2; ++i) {
        if (!pinned[i]) {
            hipError_t e = hipHostMalloc(&pinned[i], STAGING_CHUNK, hipHostMallocDefault);
            if (e != hipSuccess) {
                pinned[i] = nullptr;
                return e;
            }
        }
        if (!done[i]) {
            hipError_t e = hipEventCreateWithFlags(&done[i], hipEventDisableTiming);
            if (e != hipSuccess) {
                done[i] = nullptr;
                return e;
            }
        }
    }
    return hipSuccess;
}

hipError_t staged_upload(Staging &s, void *d_dst, const void *h_src, size_t bytes, hipStream_t stream)
{
    if (bytes < STAGING_MIN_BYTES) return hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice);
    hipError_t e = s.ensure();
    if (e != hipSuccess) return e;
    const char *src = static_cast<const char *>(h_src);
    char *dst = static_cast<char *>(d_dst);
    size_t k = 0;
    for (size_t off = 0; off < bytes && e == hipSuccess; off += STAGING_CHUNK, ++k) {
        const int b = (int)(k & 1);
        const size_t n = bytes - off < STAGING_CHUNK ? bytes - off : STAGING_CHUNK;
        if (k >= 2) e = hipEventSynchronize(s.done[b]);  // the DMA out of this buffer is done
        if (e != hipSuccess) break;
        CopyPool::get().copy(s.pinned[b], src + off, n);
        e = hipMemcpyAsync(dst + off, s.pinned[b], n, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipEventRecord(s.done[b], stream);
    }
    // The caller's memory is free to change once this returns (snapshot).
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e;
}

hipError_t staged_download(Staging &s, void *h_dst, const void *d_src, size_t bytes, hipStream_t stream)
{
    if (bytes < STAGING_MIN_BYTES) {
        hipError_t e = hipStreamSynchronize(stream);
        return e == hipSuccess ? hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost) : e;
    }
    hipError_t e = s.ensure();
    if (e != hipSuccess) return e;
    const char *src = static_cast<const char *>(d_src);
    char *dst = static_cast<char *>(h_dst);
    const size_t chunks = (bytes + STAGING_CHUNK - 1) / STAGING_CHUNK;
    auto issue = [&](size_t k) {
        const size_t off = k * STAGING_CHUNK;
        const size_t n = bytes - off < STAGING_CHUNK ? bytes - off : STAGING_CHUNK;
        hipError_t r = hipMemcpyAsync(s.pinned[k & 1], src + off, n, hipMemcpyDeviceToHost, stream);
        return r == hipSuccess ? hipEventRecord(s.done[k & 1], stream) : r;
    };
    for (size_t k = 0; k < 2 && k < chunks && e == hipSuccess; ++k) e = issue(k);
    for (size_t k = 0; k < chunks && e == hipSuccess; ++k) {
        const size_t off = k * STAGING_CHUNK;
        const size_t n = bytes - off < STAGING_CHUNK ? bytes - off : STAGING_CHUNK;
        e = hipEventSynchronize(s.done[k & 1]);
        if (e != hipSuccess) break;
        CopyPool::get().copy(dst + off, s.pinned[k & 1], n);
        if (k + 2 < chunks) e = issue(k + 2);
    }
    return e;
}

}  // namespace ndf
