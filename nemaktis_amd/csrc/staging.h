// Host <-> device copies of large caller arrays through pinned staging
// buffers: the caller's (pageable) memory is copied into one of two pinned
// chunks by several host threads while the DMA engine moves the other chunk,
// so a copy runs at the pinned DMA rate instead of the pageable one.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace nmk {

// dst (device) <- src (host); queued on st, src may be reused on return
void staged_h2d(void *dst, const void *src, size_t bytes, hipStream_t st);
// dst (host) <- src (device) after the work queued on st; complete on return
void staged_d2h(void *dst, const void *src, size_t bytes, hipStream_t st);

}  // namespace nmk
