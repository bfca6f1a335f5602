// Host interface of the no-allocation run (noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "engine.h"
#include "settings.h"

namespace nmk {

// a no-allocation run whose buffers do not fit the static scratch array
struct ScratchExhausted : std::runtime_error {
    explicit ScratchExhausted(const std::string &m) : std::runtime_error(m) {}
};

// bytes of the per-GPU scratch array
size_t noalloc_scratch_bytes();

// The propagation of p (one wavelength at a time, eager launches, no HIP
// events or extra streams) with the caller's device arrays io and every
// other buffer carved from the current device's scratch array; returns once
// the output is in io.E.  Throws ScratchExhausted before any device work when
// the run does not fit.
void run_noalloc(const Problem &p, int precision, const DeviceIO &io, hipStream_t stream);

}  // namespace nmk
