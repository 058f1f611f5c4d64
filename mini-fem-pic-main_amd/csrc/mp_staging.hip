// Pinned, double-buffered host<->device copies with a small copy thread pool (mp_staging.h);
// host code only.
#include "mp_staging.h"

#include <algorithm>
#include <cstring>

namespace mpstage {

CopyPool::CopyPool(int threads) {
    for (int k = 0; k < threads; k++) th_.emplace_back([this] { work(); });
}

CopyPool::~CopyPool() {
    {
        std::lock_guard<std::mutex> l(m_);
        stop_ = true;
    }
    go_.notify_all();
    for (auto& t : th_) t.join();
}

void CopyPool::run_parts() {
    for (int p = next_.fetch_add(1); p < parts_; p = next_.fetch_add(1)) {
        const size_t off = (size_t)p * part_;
        std::memcpy(dst_ + off, src_ + off, std::min(part_, bytes_ - off));
    }
}

void CopyPool::work() {
    unsigned long long seen = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> l(m_);
            go_.wait(l, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
        }
        run_parts();
        {
            std::lock_guard<std::mutex> l(m_);
            finished_++;
        }
        done_.notify_one();
    }
}

void CopyPool::copy(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    const int workers = (int)th_.size();
    const size_t min_part = size_t(1) << 20;
    int parts = (int)std::min<size_t>((size_t)(workers + 1) * 2, (bytes + min_part - 1) / min_part);
    if (parts < 1) parts = 1;
    {
        std::lock_guard<std::mutex> l(m_);
        dst_ = (char*)dst;
        src_ = (const char*)src;
        bytes_ = bytes;
        part_ = (bytes + parts - 1) / parts;
        parts_ = parts;
        next_.store(0);
        finished_ = 0;
        gen_++;
    }
    go_.notify_all();
    run_parts();
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [&] { return finished_ == workers; });
}

hipError_t stager_init(Stager& s) {
    if (s.ready) return hipSuccess;
    for (int k = 0; k < 2; k++) {
        hipError_t e = hipHostMalloc(&s.buf[k], Stager::kChunk, 0);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.ev[k], hipEventDisableTiming);
        if (e != hipSuccess) {
            stager_free(s);
            return e;
        }
    }
    const unsigned hc = std::thread::hardware_concurrency();
    s.pool = new CopyPool((int)std::max(1u, std::min(16u, hc ? hc : 1u)) - 1);
    s.ready = true;
    return hipSuccess;
}

void stager_free(Stager& s) {
    for (int k = 0; k < 2; k++) {
        if (s.ev[k]) (void)hipEventDestroy(s.ev[k]);
        if (s.buf[k]) (void)hipHostFree(s.buf[k]);
        s.ev[k] = nullptr;
        s.buf[k] = nullptr;
    }
    delete s.pool;
    s.pool = nullptr;
    s.ready = false;
}

hipError_t upload(Stager& s, void* dev, const void* host, size_t bytes, hipStream_t st) {
    const char* src = (const char*)host;
    char* dst = (char*)dev;
    int k = 0;
    for (size_t off = 0; off < bytes; off += Stager::kChunk, k++) {
        const size_t len = std::min(Stager::kChunk, bytes - off);
        const int b = k & 1;
        hipError_t e = hipEventSynchronize(s.ev[b]);  // the DMA that last used this chunk is done
        if (e != hipSuccess) return e;
        s.pool->copy(s.buf[b], src + off, len);
        e = hipMemcpyAsync(dst + off, s.buf[b], len, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipEventRecord(s.ev[b], st);
        if (e != hipSuccess) return e;
    }
    // the pinned chunks may be reused as soon as their DMAs finish; the caller's array is free
    return hipSuccess;
}

hipError_t download(Stager& s, void* host, const void* dev, size_t bytes, hipStream_t st) {
    char* dst = (char*)host;
    const char* src = (const char*)dev;
    const size_t n_chunks = (bytes + Stager::kChunk - 1) / Stager::kChunk;
    auto issue = [&](size_t k) -> hipError_t {
        const size_t off = k * Stager::kChunk;
        const size_t len = std::min(Stager::kChunk, bytes - off);
        hipError_t e = hipMemcpyAsync(s.buf[k & 1], src + off, len, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(s.ev[k & 1], st);
        return e;
    };
    if (n_chunks == 0) return hipSuccess;
    hipError_t e = issue(0);
    for (size_t k = 0; e == hipSuccess && k < n_chunks; k++) {
        // chunk k+1's DMA goes into the other buffer, whose host copy (chunk k-1) is done
        if (k + 1 < n_chunks) e = issue(k + 1);
        if (e == hipSuccess) e = hipEventSynchronize(s.ev[k & 1]);
        if (e == hipSuccess) {
            const size_t off = k * Stager::kChunk;
            s.pool->copy(dst + off, s.buf[k & 1], std::min(Stager::kChunk, bytes - off));
        }
    }
    return e;
}

}  // namespace mpstage
