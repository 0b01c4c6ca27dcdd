// launch.h -- the kernel launch of the library's .hip files.  While the
// profiler times a launch group on the calling thread, launches bind its
// events to the dispatches themselves (hipExtLaunchKernelGGL): the start
// event to the group's first kernel, the stop event to each kernel, so it
// ends up at the last one's end.  Events recorded around the launches
// instead would add the host's launch latency (device idle) and the event
// packets' own cost to every interval.  Outside a timed group it is
// hipLaunchKernelGGL.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

namespace dfsph {
struct PendingTiming {
    hipEvent_t start = nullptr;   // taken by the next launch
    hipEvent_t stop = nullptr;    // given to every launch of the group
    bool used = false;            // a launch took the events
};
// The calling thread's timed launch group (stop == nullptr: none).
PendingTiming& pending_timing();
}  // namespace dfsph

#define DFSPH_LAUNCH(kernel, grid, block, shmem, stream, ...)                                         \
    do {                                                                                             \
        ::dfsph::PendingTiming& dfsph_pt_ = ::dfsph::pending_timing();                               \
        if (dfsph_pt_.stop) {                                                                        \
            const hipEvent_t dfsph_start_ = dfsph_pt_.start;                                         \
            dfsph_pt_.start = nullptr;                                                               \
            dfsph_pt_.used = true;                                                                   \
            hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, dfsph_start_, dfsph_pt_.stop, 0u, \
                                  __VA_ARGS__);                                                      \
        } else {                                                                                     \
            hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                     \
        }                                                                                            \
    } while (0)
