// Launch of one merged phase of a batch of solves (batch_launch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "batch.hpp"

namespace amg {

// The batched phase kernels are instantiated in KB_NPART translation units per
// value type (batch_launch.hip compiled once per (type, part)), so that they
// compile in parallel.  kb_part names the unit that holds a kernel id; the
// row kernels, which dominate compile time, are spread over all units.
constexpr int KB_NPART = 4;
constexpr int kb_part(int kid) {
    switch (kid) {
        case K_SPMV: case K_RESIDUAL_ZERO: case K_POLY_STEP: case K_COPY: case K_COARSE_PINV:
        case K_COARSE_LU: case K_FINALIZE:
            return 0;
        case K_SPMV_ADD: case K_PCG_SPMV: case K_TRUE_RESNORM: case K_COPY64: case K_COARSE_PINV64:
        case K_COARSE_LU64:
            return 1;
        case K_JACOBI: case K_RESIDUAL: case K_JACOBI_ZERO: case K_PCG_INIT: case K_PCG_START: case K_SET_PARAMS:
        case K_COARSE_QR:
            return 2;
        default:  // K_PROLONG_X1, K_JACOBI_GAMMA, K_POLY_FIRST(_ZERO), K_PCG_P/UPDATE/GAMMA, K_REPORT, K_COARSE_QR64
            return 3;
    }
}

// Runs kb_phase<kid, T, kpc> (kb_finalize for K_FINALIZE) over `total` virtual
// blocks on a grid of `grid` blocks, for a kid of unit PART.  Returns 0, or -1
// for an unknown kid.
template <typename T, int PART>
int kb_launch_part(int kid, int kpc, int grid, const Step* steps, const int* gmap, int total, BatchFlag bf,
                   hipStream_t s);

template <typename T>
inline int kb_launch(int kid, int kpc, int grid, const Step* steps, const int* gmap, int total, BatchFlag bf,
                     hipStream_t s) {
    if (kid < 0 || kid >= K_NKID) return -1;
    switch (kb_part(kid)) {
        case 0: return kb_launch_part<T, 0>(kid, kpc, grid, steps, gmap, total, bf, s);
        case 1: return kb_launch_part<T, 1>(kid, kpc, grid, steps, gmap, total, bf, s);
        case 2: return kb_launch_part<T, 2>(kid, kpc, grid, steps, gmap, total, bf, s);
        default: return kb_launch_part<T, 3>(kid, kpc, grid, steps, gmap, total, bf, s);
    }
}

}  // namespace amg
