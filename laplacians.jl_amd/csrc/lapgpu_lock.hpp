// lapgpu_lock.hpp — one process-wide lock around the HIP calls that must
// not overlap a stream capture running in another thread (allocations,
// frees, synchronous copies, which synchronise with the null stream) and
// around the captures themselves.  Solves on different contexts otherwise
// run concurrently from different host threads (the serving loop does).
#pragma once

#include <mutex>

namespace lapgpu {
std::recursive_mutex& device_lock();
}

#define LAPGPU_LOCK() \
  std::lock_guard<std::recursive_mutex> lapgpu_lock_guard_(::lapgpu::device_lock())
