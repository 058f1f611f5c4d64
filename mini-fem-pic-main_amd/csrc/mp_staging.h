// Host<->device copies of large host arrays the library does not own (the drop-in path's
// Particle vector): the host side is copied by several threads into two pinned chunks that
// alternate, each chunk's DMA overlapping the next chunk's host copy.  Pageable
// hipMemcpyAsync runs at a few GB/s; pinned DMA at PCIe speed.  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <mutex>
#include <thread>
#include <vector>

namespace mpstage {

// memcpy split over a fixed set of worker threads plus the caller; copy() returns when done.
class CopyPool {
public:
    explicit CopyPool(int threads);
    ~CopyPool();
    void copy(void* dst, const void* src, size_t bytes);

private:
    void work();
    void run_parts();
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable go_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0, part_ = 0;
    int parts_ = 0;
    std::atomic<int> next_{0};
    int finished_ = 0;
    unsigned long long gen_ = 0;
    bool stop_ = false;
};

// Two pinned chunks, their events, and the pool.
struct Stager {
    static constexpr size_t kChunk = size_t(64) << 20;
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    CopyPool* pool = nullptr;
    bool ready = false;
};

// Arrays at least this large go through the stager; smaller ones are copied directly.
constexpr size_t kStageMin = size_t(32) << 20;

hipError_t stager_init(Stager& s);
void stager_free(Stager& s);
// host -> device, stream-ordered: returns once every chunk's DMA is queued (the host array is
// no longer read after the call returns)
hipError_t upload(Stager& s, void* dev, const void* host, size_t bytes, hipStream_t st);
// device -> host after the work queued on st: returns when the host array holds the data
hipError_t download(Stager& s, void* host, const void* dev, size_t bytes, hipStream_t st);

}  // namespace mpstage
