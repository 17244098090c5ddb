// lapgpu_mem.hpp — device and pinned host memory without device-wide waits.
//
// hipFree and hipHostFree each wait for every stream on the device.  A
// service that drops one graph while other graphs' solves run (seconds for
// a large multi-RHS solve) cannot afford that, so:
//  * device buffers are stream-ordered: hipMallocAsync / hipFreeAsync on the
//    owner's stream, from the device's default pool, which is set once per
//    device to keep freed memory for reuse instead of returning it;
//  * the small pinned host buffers (solver state mirrors, stop-flag rings,
//    request tables, staging) come from a process-wide free list in power-of-
//    two size classes and are never handed back to the system.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace lapgpu {

// Once per device (later calls return at once): the default memory pool
// keeps freed memory (release threshold = max).
hipError_t pool_setup(int device);

// Pinned host memory of at least `bytes` bytes; *cap gets the usable size.
hipError_t pinned_get(void** p, size_t bytes, size_t* cap = nullptr);
// Return a block from pinned_get (no copy may still be reading or writing it).
void pinned_put(void* p, size_t bytes);

}  // namespace lapgpu
