// Host ingest and egress. Copies between pageable host memory (numpy arrays, callers' buffers)
// and the device go through two pinned staging buffers of the calling thread, chunk by chunk:
// the host side of each chunk is spread over a small pool of copy threads and overlaps the DMA
// of the neighbouring chunk. A plain hipMemcpy from pageable memory stages through the
// runtime's own buffers with one host thread and reaches about a fifth of the PCIe rate.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <functional>

namespace qh {

// Threads a host copy is spread over (the caller included): half the CPUs this process may
// run on, at most 8.
int host_copy_threads();

// dst[0, bytes) = src[0, bytes), split over host_copy_threads() threads.
void host_copy(void* dst, const void* src, std::size_t bytes);

// part(0) .. part(parts - 1) on the copy threads and the calling thread; returns when all are
// done.
void parallel_for(int parts, const std::function<void(int)>& part);

// Device dst <- pageable host src. Returns once src may be reused; the DMA of the last chunks
// may still be in flight on s, ahead of any later work on s.
void upload(void* dst, const void* src, std::size_t bytes, hipStream_t s);

// Pageable host dst <- device src, after the work already queued on s. Returns with dst
// complete.
void download(void* dst, const void* src, std::size_t bytes, hipStream_t s);

}  // namespace qh
