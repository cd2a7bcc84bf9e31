// Batched phase kernels (kb_phase / kb_finalize instantiations) for one value
// type AMG_KB_T and one unit AMG_KB_PART of the kernel ids (kb_part): the build
// compiles this file once per (type, unit), so the instantiations compile in
// parallel with each other and with the rest of the library.
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "batch_launch.hpp"
#include "kernels.hpp"

#ifndef AMG_KB_T
#error "define AMG_KB_T (float or double)"
#endif
#ifndef AMG_KB_PART
#error "define AMG_KB_PART (0 .. KB_NPART - 1)"
#endif

namespace amg {

template <typename T, int PART>
int kb_launch_part(int kid, int kpc, int grid, const Step* steps, const int* gmap, int total, BatchFlag bf,
                   hipStream_t s) {
#define AMG_KB_CASE0(K)                                                                       \
    case K:                                                                                   \
        if constexpr (kb_part(K) == PART) {                                                   \
            kb_phase<K, T, 0><<<grid, BLOCK, 0, s>>>(steps, gmap, total, bf);                 \
            break;                                                                            \
        } else {                                                                              \
            return -1;                                                                        \
        }
#define AMG_KB_CASE(K)                                                                        \
    case K:                                                                                   \
        if constexpr (kb_part(K) == PART) {                                                   \
            switch (kpc) {                                                                    \
                case 1: kb_phase<K, T, 1><<<grid, BLOCK, 0, s>>>(steps, gmap, total, bf); break;   \
                case 2: kb_phase<K, T, 2><<<grid, BLOCK, 0, s>>>(steps, gmap, total, bf); break;   \
                case 4: kb_phase<K, T, 4><<<grid, BLOCK, 0, s>>>(steps, gmap, total, bf); break;   \
                case 8: kb_phase<K, T, 8><<<grid, BLOCK, 0, s>>>(steps, gmap, total, bf); break;   \
                case 16: kb_phase<K, T, 16><<<grid, BLOCK, 0, s>>>(steps, gmap, total, bf); break; \
                case 32: kb_phase<K, T, 32><<<grid, BLOCK, 0, s>>>(steps, gmap, total, bf); break; \
                case 64: kb_phase<K, T, 64><<<grid, BLOCK, 0, s>>>(steps, gmap, total, bf); break; \
                default: kb_phase<K, T, 0><<<grid, BLOCK, 0, s>>>(steps, gmap, total, bf); break;  \
            }                                                                                 \
            break;                                                                            \
        } else {                                                                              \
            return -1;                                                                        \
        }
    switch (kid) {
        AMG_KB_CASE(K_SPMV)
        AMG_KB_CASE(K_SPMV_ADD)
        AMG_KB_CASE(K_JACOBI)
        AMG_KB_CASE(K_JACOBI_ZERO)
        AMG_KB_CASE(K_RESIDUAL)
        AMG_KB_CASE(K_RESIDUAL_ZERO)
        AMG_KB_CASE(K_PROLONG_X1)
        AMG_KB_CASE(K_COPY)
        AMG_KB_CASE0(K_COARSE_PINV)
        AMG_KB_CASE0(K_COARSE_PINV64)
        AMG_KB_CASE(K_PCG_INIT)
        AMG_KB_CASE(K_PCG_START)
        AMG_KB_CASE(K_PCG_P)
        AMG_KB_CASE(K_PCG_SPMV)
        AMG_KB_CASE(K_PCG_UPDATE)
        AMG_KB_CASE(K_PCG_GAMMA)
        AMG_KB_CASE(K_TRUE_RESNORM)
        case K_FINALIZE:
            if constexpr (kb_part(K_FINALIZE) == PART) {
                kb_finalize<<<grid, FIN_BLOCK, 0, s>>>(steps, gmap, total, bf);
                break;
            } else {
                return -1;
            }
        AMG_KB_CASE0(K_SET_PARAMS)
        AMG_KB_CASE(K_COPY64)
        AMG_KB_CASE0(K_REPORT)
        AMG_KB_CASE(K_JACOBI_GAMMA)
        AMG_KB_CASE(K_POLY_FIRST)
        AMG_KB_CASE(K_POLY_FIRST_ZERO)
        AMG_KB_CASE(K_POLY_STEP)
        AMG_KB_CASE0(K_COARSE_LU)
        AMG_KB_CASE0(K_COARSE_LU64)
        AMG_KB_CASE0(K_COARSE_QR)
        AMG_KB_CASE0(K_COARSE_QR64)
        default:
            return -1;
    }
#undef AMG_KB_CASE0
#undef AMG_KB_CASE
    return 0;
}

template int kb_launch_part<AMG_KB_T, AMG_KB_PART>(int, int, int, const Step*, const int*, int, BatchFlag,
                                                   hipStream_t);

}  // namespace amg
