// hostcopy.hip -- large copies between pageable host memory and the device.
//
// hipMemcpyAsync from or to pageable memory stages the data through the
// runtime's own small pinned buffers on the calling thread (about 8 GB/s down
// and 18 GB/s up for the 2048^2 frames and result arrays).  Here a context
// keeps a pinned double buffer of two 4 MiB chunks; the host side of a chunk
// (pageable <-> pinned) is spread over a few threads while the other chunk's
// DMA runs, so host copies and PCIe transfers overlap.  Copies below 2 MiB keep
// the runtime's path.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <thread>
#include <vector>

#include "dic_internal.hpp"

namespace dic {
namespace {

// The process's copy threads: a copy job is split in kPiece tasks that the
// calling thread and the workers take until none is left; the caller returns
// once every worker has left the job.  One job runs at a time (contexts on
// several threads queue on `jobs`).
struct CopyPool {
    static constexpr size_t kPiece = (size_t)256 << 10;
    std::mutex jobs;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_job, cv_done;
    unsigned char *dst = nullptr;
    const unsigned char *src = nullptr;
    size_t bytes = 0;
    std::atomic<size_t> next{0};
    size_t busy = 0;
    uint64_t job = 0;
    bool stop = false;

    CopyPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const unsigned n = hw > 2 ? std::min(7u, hw / 2) : 0u; // workers beside the calling thread
        for (unsigned i = 0; i < n; ++i) {
            try {
                workers.emplace_back([this] { worker(); });
            } catch (...) {
                break; // fewer workers
            }
        }
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_job.notify_all();
        for (auto &t : workers) t.join();
    }
    void drain()
    {
        for (size_t o; (o = next.fetch_add(kPiece)) < bytes;) std::memcpy(dst + o, src + o, std::min(kPiece, bytes - o));
    }
    void worker()
    {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_job.wait(lk, [&] { return stop || job != seen; });
            if (stop) return;
            seen = job;
            lk.unlock();
            drain();
            lk.lock();
            if (--busy == 0) cv_done.notify_one();
        }
    }
    void copy(void *d, const void *s, size_t n)
    {
        if (workers.empty() || n < 2 * kPiece) {
            std::memcpy(d, s, n);
            return;
        }
        std::lock_guard<std::mutex> one(jobs);
        {
            std::lock_guard<std::mutex> lk(mu);
            dst = (unsigned char *)d;
            src = (const unsigned char *)s;
            bytes = n;
            next = 0;
            busy = workers.size();
            ++job;
        }
        cv_job.notify_all();
        drain();
        std::unique_lock<std::mutex> lk(mu);
        cv_done.wait(lk, [&] { return busy == 0; });
    }
};

CopyPool &copy_pool()
{
    static CopyPool pool;
    return pool;
}

} // namespace

// A context's pinned double buffer.
struct HostStager {
    static constexpr size_t kChunk = (size_t)4 << 20; // bytes per pinned slot
    unsigned char *pin[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr}; // the last DMA that touched each slot
    bool used[2] = {false, false};
    std::mutex use; // one copy at a time per context

    void copy(void *d, const void *s, size_t n) { copy_pool().copy(d, s, n); }
    hipError_t ready(int k) { return used[k] ? hipEventSynchronize(done[k]) : hipSuccess; }
    hipError_t mark(int k, hipStream_t s)
    {
        used[k] = true;
        return hipEventRecord(done[k], s);
    }
    ~HostStager()
    {
        for (int k = 0; k < 2; ++k) {
            if (done[k]) (void)hipEventSynchronize(done[k]);
            if (done[k]) (void)hipEventDestroy(done[k]);
            if (pin[k]) (void)hipHostFree(pin[k]);
        }
    }
};

namespace {

constexpr size_t kStagedMin = (size_t)2 << 20;

// The context's stager, made on first use; nullptr if its pinned memory or
// events cannot be had (the copies then take the runtime's path).
HostStager *stager(dic_ctx *ctx)
{
    if (ctx->stager) return ctx->stager;
    if (ctx->stager_failed) return nullptr;
    HostStager *h = new (std::nothrow) HostStager();
    bool ok = h != nullptr;
    for (int k = 0; ok && k < 2; ++k)
        ok = hipHostMalloc((void **)&h->pin[k], HostStager::kChunk, hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&h->done[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        delete h;
        ctx->stager_failed = true;
        return nullptr;
    }
    ctx->stager = h;
    return h;
}

} // namespace

hipError_t upload_host(dic_ctx *ctx, void *device, const void *host, size_t bytes, hipStream_t s)
{
    HostStager *h = bytes >= kStagedMin ? stager(ctx) : nullptr;
    if (!h) return hipMemcpyAsync(device, host, bytes, hipMemcpyHostToDevice, s);
    std::lock_guard<std::mutex> lk(h->use);
    const unsigned char *src = (const unsigned char *)host;
    unsigned char *dst = (unsigned char *)device;
    for (size_t off = 0, i = 0; off < bytes; off += HostStager::kChunk, ++i) {
        const int k = (int)(i & 1);
        const size_t len = std::min(HostStager::kChunk, bytes - off);
        hipError_t e = h->ready(k); // the slot's previous DMA has finished with it
        if (e != hipSuccess) return e;
        h->copy(h->pin[k], src + off, len);
        e = hipMemcpyAsync(dst + off, h->pin[k], len, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = h->mark(k, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t download_host(dic_ctx *ctx, void *host, const void *device, size_t bytes, hipStream_t s)
{
    HostStager *h = bytes >= kStagedMin ? stager(ctx) : nullptr;
    if (!h) {
        hipError_t e = hipMemcpyAsync(host, device, bytes, hipMemcpyDeviceToHost, s);
        return e == hipSuccess ? hipStreamSynchronize(s) : e;
    }
    std::lock_guard<std::mutex> lk(h->use);
    const unsigned char *src = (const unsigned char *)device;
    unsigned char *dst = (unsigned char *)host;
    const size_t chunks = (bytes + HostStager::kChunk - 1) / HostStager::kChunk;
    auto issue = [&](size_t i) -> hipError_t { // chunk i's DMA into slot i mod 2
        const int k = (int)(i & 1);
        const size_t off = i * HostStager::kChunk, len = std::min(HostStager::kChunk, bytes - off);
        hipError_t e = h->ready(k);
        if (e == hipSuccess) e = hipMemcpyAsync(h->pin[k], src + off, len, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = h->mark(k, s);
        return e;
    };
    hipError_t e = issue(0);
    for (size_t i = 0; e == hipSuccess && i < chunks; ++i) {
        if (i + 1 < chunks) e = issue(i + 1); // the next DMA runs while this chunk is copied out
        const int k = (int)(i & 1);
        if (e == hipSuccess) e = hipEventSynchronize(h->done[k]);
        if (e != hipSuccess) break;
        const size_t off = i * HostStager::kChunk;
        h->copy(dst + off, h->pin[k], std::min(HostStager::kChunk, bytes - off));
    }
    return e;
}

void destroy_stager(dic_ctx *ctx)
{
    delete ctx->stager;
    ctx->stager = nullptr;
}

} // namespace dic
