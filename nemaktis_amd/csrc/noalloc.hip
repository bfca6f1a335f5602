// No-allocation variant of the bpm propagation: every device buffer beyond
// the caller's inputs and output is carved from one static device array per
// GPU, so that callers that cannot allocate device memory (code under graph
// capture, tight serving loops, harnesses that forbid hipMalloc) can run it.
//
// The array `scratch` exists once per GPU the library runs on and stays
// until the process ends.  Host code never uses its name as an address: the
// device address is looked up with hipGetSymbolAddress on the first call on
// each device (with that device current) and kept per device.  Every call
// on a GPU reuses that GPU's array; the march writes every scratch element it
// reads before reading it, so nothing needs clearing between calls, and calls
// that overlap on one GPU race (the caller serialises them).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/nemaktis_bpm.h"
#include "engine.h"
#include "noalloc.h"

namespace nmk {

namespace {

constexpr size_t kScratchBytes = NMK_BPM_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % kAllocAlign == 0, "scratch size");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
std::map<int, unsigned char *> g_scratch;  // device ordinal -> address of that device's `scratch`

// address of the current device's scratch array (looked up once per device)
unsigned char *scratch_base() {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e) + " at hipGetDevice");
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_scratch.find(dev);
    if (it != g_scratch.end()) return it->second;
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e) +
                                 " at hipGetSymbolAddress(scratch)");
    g_scratch[dev] = static_cast<unsigned char *>(ptr);
    return static_cast<unsigned char *>(ptr);
}

// bump allocator over the scratch array; a run's buffers are all released
// together when the run ends
class ScratchAllocator final : public DeviceAllocator {
public:
    explicit ScratchAllocator(unsigned char *base) : base_(base) {}
    void *allocate(size_t bytes) override {
        const size_t at = (used_ + kAllocAlign - 1) / kAllocAlign * kAllocAlign;
        if (at + bytes > kScratchBytes)
            throw ScratchExhausted("the no-allocation run exceeded its device scratch of " +
                                   std::to_string(kScratchBytes) + " bytes");
        used_ = at + bytes;
        return base_ + at;
    }
    void release(void *) override {}

private:
    unsigned char *base_;
    size_t used_ = 0;
};

}  // namespace

size_t noalloc_scratch_bytes() { return kScratchBytes; }

void run_noalloc(const Problem &p, int precision, const DeviceIO &io, hipStream_t stream) {
    Options o;
    o.precision = precision;
    o.wl_batch = 1;
    o.use_graph = 0;
    o.profile = 0;
    o.dev = io;
    o.no_timing = true;
    o.no_side_stream = true;
    o.coef_batch = 1;  // two slab buffers of coefficients: the smallest footprint
    const int64_t need = estimate_footprint(p, o);
    if (need > (int64_t)kScratchBytes)
        throw ScratchExhausted("the run needs " + std::to_string(need) +
                               " bytes of device scratch; the no-allocation entry point holds " +
                               std::to_string(kScratchBytes) + " (NMK_BPM_NOALLOC_SCRATCH_BYTES)");
    ScratchAllocator scratch_alloc(scratch_base());
    o.allocator = &scratch_alloc;
    auto r = make_runner(p, o, stream ? stream : hipStreamPerThread);
    r->propagate(nullptr);
}

}  // namespace nmk
