// Host-to-device copies from ordinary (pageable) host memory through a ring of
// page-locked slots.
//
// A DMA engine reads only page-locked memory, so a copy from pageable memory
// goes through a bounce buffer either way; the runtime's own path does it with
// one thread.  Here a pool of worker threads copies each piece of the source
// into a free slot of the ring in parallel, then the slot is sent with an
// asynchronous copy on the caller's stream, and an event recorded after it
// marks when the slot may be refilled.  The CPU copy of piece i + 1 overlaps
// the DMA of piece i.  The call returns once every piece has been queued: the
// source may then be reused, and work queued later on the stream sees the data.
//
// One ring per device (kStageSlots x kStageSlotBytes page-locked), created on
// first use; calls on one device take turns on it (a mutex).
#include <condition_variable>
#include <cstring>
#include <functional>
#include <thread>

#include "tvh_internal.hpp"

namespace tvh {

namespace {

constexpr size_t kStageSlotBytes = 32ull << 20;  // 32 MiB per slot
constexpr int kStageSlots = 8;                    // 256 MiB page-locked per device
constexpr int kStageThreads = 8;                  // host copy workers
constexpr size_t kStageMinBytes = 4ull << 20;     // smaller copies go straight to the runtime

// Fixed pool of threads that run one parallel job at a time: job(i) for
// i in [0, kStageThreads).
class CopyPool {
public:
    CopyPool() {
        for (int i = 0; i < kStageThreads; i++) th_.emplace_back([this, i] { run(i); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    void parallel(const std::function<void(int)>& job) {
        std::unique_lock<std::mutex> lk(mu_);
        job_ = &job;
        pending_ = kStageThreads;
        gen_++;
        cv_.notify_all();
        done_cv_.wait(lk, [&] { return pending_ == 0; });
        job_ = nullptr;
    }

private:
    void run(int i) {
        unsigned long seen = 0;
        for (;;) {
            const std::function<void(int)>* job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                job = job_;
            }
            (*job)(i);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--pending_ == 0) done_cv_.notify_all();
            }
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)>* job_ = nullptr;
    int pending_ = 0;
    unsigned long gen_ = 0;
    bool stop_ = false;
};

struct Ring {
    std::mutex mu;
    bool ready = false;
    void* slot[kStageSlots] = {};
    hipEvent_t ev[kStageSlots] = {};
    bool in_flight[kStageSlots] = {};
    CopyPool* pool = nullptr;
};

Ring g_rings[64];

void ring_release(Ring& r) {
    for (int s = 0; s < kStageSlots; s++) {
        if (r.slot[s]) (void)hipHostFree(r.slot[s]);
        if (r.ev[s]) (void)hipEventDestroy(r.ev[s]);
        r.slot[s] = nullptr;
        r.ev[s] = nullptr;
    }
}

int ring_setup(Ring& r) {
    for (int s = 0; s < kStageSlots; s++) {
        if (hipHostMalloc(&r.slot[s], kStageSlotBytes, hipHostMallocDefault) != hipSuccess) {
            r.slot[s] = nullptr;
            ring_release(r);
            return set_error(TVH_ERR_OOM, "page-locked staging slot of %zu bytes", kStageSlotBytes);
        }
        if (hipEventCreateWithFlags(&r.ev[s], hipEventDisableTiming) != hipSuccess) {
            r.ev[s] = nullptr;
            ring_release(r);
            return set_error(TVH_ERR_DEVICE, "staging event");
        }
    }
    // the pool's threads live as long as the process (the ring is never torn down)
    r.pool = new CopyPool();
    r.ready = true;
    return TVH_OK;
}

}  // namespace

int staged_h2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (!bytes) return TVH_OK;
    if (bytes < kStageMinBytes) {
        if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
            return set_error(TVH_ERR_DEVICE, "host-to-device copy of %zu bytes", bytes);
        return TVH_OK;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return set_error(TVH_ERR_DEVICE, "no current HIP device");
    Ring& r = g_rings[dev];
    std::lock_guard<std::mutex> lk(r.mu);
    if (!r.ready) {
        int rc = ring_setup(r);
        if (rc) return rc;
    }
    const char* s8 = (const char*)src;
    char* d8 = (char*)dst;
    const size_t pieces = (bytes + kStageSlotBytes - 1) / kStageSlotBytes;
    for (size_t i = 0; i < pieces; i++) {
        const int s = (int)(i % kStageSlots);
        const size_t off = i * kStageSlotBytes;
        const size_t len = bytes - off < kStageSlotBytes ? bytes - off : kStageSlotBytes;
        if (r.in_flight[s] && hipEventSynchronize(r.ev[s]) != hipSuccess)
            return set_error(TVH_ERR_DEVICE, "staging slot wait");
        char* slot = (char*)r.slot[s];
        const char* from = s8 + off;
        std::function<void(int)> job = [&](int t) {
            // 4 KiB-aligned shares so that workers do not split pages
            size_t share = ((len + kStageThreads - 1) / kStageThreads + 4095) & ~(size_t)4095;
            size_t a = (size_t)t * share, b = a + share < len ? a + share : len;
            if (a < b) std::memcpy(slot + a, from + a, b - a);
        };
        r.pool->parallel(job);
        if (hipMemcpyAsync(d8 + off, slot, len, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipEventRecord(r.ev[s], st) != hipSuccess)
            return set_error(TVH_ERR_DEVICE, "staged host-to-device copy");
        r.in_flight[s] = true;
    }
    return TVH_OK;
}

}  // namespace tvh

using namespace tvh;

int tvh_upload_dev(uint64_t* d_dst, const uint64_t* h_src, size_t count, void* stream) {
    if (count && (!d_dst || !h_src)) return set_error(TVH_ERR_INVALID_ARG, "null buffer");
    if (!current_device()) return TVH_ERR_DEVICE;
    return staged_h2d(d_dst, h_src, count * sizeof(uint64_t), (hipStream_t)stream);
}
