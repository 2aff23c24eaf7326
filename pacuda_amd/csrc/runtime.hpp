// SPDX-License-Identifier: 0BSD
// Host helpers shared by the v1 runner and the fused path.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "pacman_recurrence.h"

struct pacuda_plan;

namespace pacuda {
// prints "pacman recurrence runner: <step>: <error>" and returns true on a HIP error
bool hip_failed(hipError_t e, const char *step);
// the library's own stream (created once, default priority, non-blocking)
hipStream_t runner_stream();

// v1 runs on the fused kernels. The fused plan of a v1 call's words and tables, or nullptr
// (nothing printed) when its program cannot run there with the v1 kernels' results: a program
// error, a current plane read before the program writes it (the v1 kernels would read what an
// earlier call left in the arena), no closure. `amplitudes` receives the amplitude planes the
// closures write, ascending. Destroy with pacuda_plan_destroy.
pacuda_plan *v1_fused_plan(const pacman_recurrence_call &c, bool sort, std::vector<uint32_t> &amplitudes);
// One launch of the fused kernels (the default launch of `kernel`) for `points` points of the
// external momenta on the device (planes ext_stride values apart): the plan's amplitude planes,
// each from zero, into are / aim (planes `stride` values apart) in the kernel's amplitude layout,
// queued on `s`. 0 OK, else a pacman_recurrence status.
int v1_fused_run(const pacuda_plan *plan, uint32_t kernel, const double *external, uint32_t points, uint32_t ext_stride,
                 double *are, double *aim, size_t stride, hipStream_t s);
}  // namespace pacuda
