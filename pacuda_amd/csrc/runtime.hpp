// SPDX-License-Identifier: 0BSD
// Host helpers shared by the v1 runner and the fused path.
#pragma once
#include <hip/hip_runtime.h>

namespace pacuda {
// prints "pacman recurrence runner: <step>: <error>" and returns true on a HIP error
bool hip_failed(hipError_t e, const char *step);
// the library's own stream (created once, default priority, non-blocking)
hipStream_t runner_stream();
}  // namespace pacuda
