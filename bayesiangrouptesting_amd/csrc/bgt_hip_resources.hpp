// Process-wide device and host resources of the provider: stream-ordered device memory,
// pinned host blocks and streams, each recycled rather than released. Releasing device or
// pinned memory with hipFree / hipHostFree synchronises the whole device, and creating a
// stream or a pinned buffer costs tens to hundreds of microseconds; a serving loop or a
// tree creates and retires lattices continually, so they come from here.
#pragma once

#include <mutex>

#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace bgt_hip {

// Device memory ordered on `stream` (hipMallocAsync on the device's default pool, which
// keeps freed memory for re-use; plain hipMalloc where stream-ordered allocation is not
// supported). The allocation is usable by work on `stream`; other streams must wait for it.
// Held while a thread captures a batched step into a graph, and by the allocation
// helpers below: the runtime refuses stream-ordered allocation on any stream while a
// capture is in progress, so allocations on other threads wait for the capture to end.
std::mutex &capture_gate();

void *device_alloc(size_t bytes, hipStream_t stream);
// Release after the work already queued on `stream`. Whoever frees must make sure no
// other stream still uses the memory.
void device_free(void *p, hipStream_t stream);

// Pinned host memory in power-of-two blocks (at least 256 bytes), never returned to the
// system; pinned_release takes the block back for re-use.
void *pinned_alloc(size_t bytes);
void pinned_release(void *p, size_t bytes);

// A stream of default priority for a lattice, one of a bounded set that lattices share (a
// lattice's work stays ordered; other lattices' work may queue on the same stream);
// stream_release gives up the caller's use of it.
hipStream_t stream_acquire();
void stream_release(hipStream_t s);

} // namespace bgt_hip
