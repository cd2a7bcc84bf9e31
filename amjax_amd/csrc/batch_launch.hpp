// Launch of one merged phase of a batch of solves (batch_launch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "batch.hpp"

namespace amg {

// Runs kb_phase<kid, T, kpc> (kb_finalize for K_FINALIZE) over `total` virtual
// blocks on a grid of `grid` blocks.  Returns 0, or -1 for an unknown kid.
template <typename T>
int kb_launch(int kid, int kpc, int grid, const Step* steps, const int* gmap, int total, BatchFlag bf,
              hipStream_t s);

}  // namespace amg
