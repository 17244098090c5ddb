// lapgpu_lock.hpp — one process-wide lock around the stream captures and
// the few HIP calls that may touch the null stream or device-wide state
// (stream and event creation and destruction, pinned host allocation), so
// that none of them overlaps a capture running in another thread.  Device
// memory is stream-ordered on each owner's own non-blocking stream and takes
// no lock; solves on different contexts run concurrently from different
// host threads (the serving loop does).
#pragma once

#include <mutex>

namespace lapgpu {
std::recursive_mutex& device_lock();
}

#define LAPGPU_LOCK() \
  std::lock_guard<std::recursive_mutex> lapgpu_lock_guard_(::lapgpu::device_lock())
