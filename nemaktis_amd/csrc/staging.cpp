#include "staging.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace nmk {

namespace {

#define STAGECHK(x)                                                                                       \
    do {                                                                                                  \
        hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess)                                                                             \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

constexpr size_t kChunk = size_t(32) << 20;      // bytes per pinned chunk
constexpr size_t kMinStaged = size_t(16) << 20;  // smaller copies go straight through

// memcpy split over up to 8 host threads (the calling thread takes a share)
void par_memcpy(void *dst, const void *src, size_t n) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t T = std::min<size_t>({8, hw, std::max<size_t>(1, n >> 22)});
    if (T <= 1) {
        std::memcpy(dst, src, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T - 1);
    auto part = [&](size_t i) {
        const size_t a = n * i / T, b = n * (i + 1) / T;
        std::memcpy(static_cast<char *>(dst) + a, static_cast<const char *>(src) + a, b - a);
    };
    for (size_t i = 1; i < T; i++) th.emplace_back(part, i);
    part(0);
    for (auto &t : th) t.join();
}

// two pinned chunks and their DMA events, allocated on first use and kept for
// the life of the process; one copy at a time uses them (others go straight
// through rather than wait)
struct Stager {
    std::mutex mu;
    bool tried = false, ok = false;
    int dev = -1;  // the device the events belong to: copies on other devices go straight through
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};

    bool init() {  // under mu
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess) return false;
        if (tried) return ok && cur == dev;
        tried = true;
        dev = cur;
        const char *off = std::getenv("NMK_NO_STAGING");
        if (off && *off && *off != '0') return ok = false;
        for (int b = 0; b < 2; b++) {
            if (hipHostMalloc(&buf[b], kChunk, hipHostMallocDefault) != hipSuccess) buf[b] = nullptr;
            if (hipEventCreateWithFlags(&ev[b], hipEventDisableTiming) != hipSuccess) ev[b] = nullptr;
        }
        ok = buf[0] && buf[1] && ev[0] && ev[1];
        return ok;
    }
};

Stager &stager() {
    static Stager *s = new Stager;  // never destroyed: its buffers live until the process ends
    return *s;
}

}  // namespace

void staged_h2d(void *dst, const void *src, size_t bytes, hipStream_t st) {
    Stager &s = stager();
    std::unique_lock<std::mutex> lk(s.mu, std::defer_lock);
    if (bytes < kMinStaged || !lk.try_lock() || !s.init()) {
        STAGECHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        return;
    }
    bool pending[2] = {false, false};
    try {
        for (size_t off = 0, c = 0; off < bytes; off += kChunk, c++) {
            const int b = int(c & 1);
            const size_t n = std::min(kChunk, bytes - off);
            if (pending[b]) STAGECHK(hipEventSynchronize(s.ev[b]));  // the chunk's previous DMA is done
            par_memcpy(s.buf[b], static_cast<const char *>(src) + off, n);
            STAGECHK(hipMemcpyAsync(static_cast<char *>(dst) + off, s.buf[b], n, hipMemcpyHostToDevice, st));
            STAGECHK(hipEventRecord(s.ev[b], st));
            pending[b] = true;
        }
    } catch (...) {
        // the buffers go to the next copy only once their DMAs are over
        (void)hipStreamSynchronize(st);
        throw;
    }
    for (int b = 0; b < 2; b++)
        if (pending[b]) STAGECHK(hipEventSynchronize(s.ev[b]));
}

void staged_d2h(void *dst, const void *src, size_t bytes, hipStream_t st) {
    Stager &s = stager();
    std::unique_lock<std::mutex> lk(s.mu, std::defer_lock);
    if (bytes < kMinStaged || !lk.try_lock() || !s.init()) {
        STAGECHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        STAGECHK(hipStreamSynchronize(st));
        return;
    }
    const size_t chunks = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](size_t c) {
        const size_t off = c * kChunk, n = std::min(kChunk, bytes - off);
        const int b = int(c & 1);
        STAGECHK(hipMemcpyAsync(s.buf[b], static_cast<const char *>(src) + off, n, hipMemcpyDeviceToHost, st));
        STAGECHK(hipEventRecord(s.ev[b], st));
    };
    try {
        issue(0);
        for (size_t c = 0; c < chunks; c++) {
            // chunk c+1 moves while chunk c is copied out (its buffer held
            // chunk c-1, already copied out)
            if (c + 1 < chunks) issue(c + 1);
            const int b = int(c & 1);
            STAGECHK(hipEventSynchronize(s.ev[b]));
            const size_t off = c * kChunk;
            par_memcpy(static_cast<char *>(dst) + off, s.buf[b], std::min(kChunk, bytes - off));
        }
    } catch (...) {
        (void)hipStreamSynchronize(st);
        throw;
    }
}

}  // namespace nmk
