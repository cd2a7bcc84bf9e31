// Batched launches: one kernel launch carries the same phase of many
// independent solves (different hierarchies, sizes and block widths k).
//
// Each problem's share of a launch is a Step: the arguments of the
// single-problem kernel plus the problem's own virtual grid (nbx x nby blocks,
// exactly the grid the single-problem launch would use).  The steps of a launch
// are laid end to end in one virtual block range, each starting at a multiple
// of 8 so that a problem's blocks keep the XCD placement of a plain launch; a
// map from every group of 8 virtual blocks to its step makes the lookup one
// load.  The launch is persistent: its grid may be smaller than the virtual
// range and every block walks the range with a stride of gridDim.x (a multiple
// of 8), running the same device bodies as the plain kernels, so a batched
// solve computes bit for bit what the single solves compute.  A launch either
// holds steps of one column-lane count KP (kb_phase<.., KP>: the register
// budget of the plain kernel) or of several (kb_phase<.., 0>: KP chosen per
// step, the register budget of the widest variant).
#pragma once
#include "kernels.hpp"

namespace amg {

enum Kid : int {
    K_SPMV,
    K_SPMV_ADD,
    K_JACOBI,
    K_JACOBI_ZERO,
    K_RESIDUAL,
    K_RESIDUAL_ZERO,
    K_PROLONG_X1,
    K_COPY,
    K_COARSE_PINV,
    K_COARSE_PINV64,
    K_PCG_INIT,
    K_PCG_START,
    K_PCG_P,
    K_PCG_SPMV,
    K_PCG_UPDATE,
    K_PCG_GAMMA,
    K_TRUE_RESNORM,
    K_FINALIZE,
    K_SET_PARAMS,
    K_COPY64,
    K_REPORT,
    K_JACOBI_GAMMA,
    K_POLY_FIRST,
    K_POLY_FIRST_ZERO,
    K_POLY_STEP,
    K_COARSE_LU,
    K_COARSE_LU64,
    K_COARSE_QR,
    K_COARSE_QR64,
    K_NKID
};

struct SellRaw {
    const int* sptr;
    const int* col;
    const void* val;
    int nrows, ncols, maxw, npr;
};

template <typename T>
__host__ __device__ inline SellRaw raw_of(const Sell<T>& s) {
    return SellRaw{s.sptr, s.col, s.val, s.nrows, s.ncols, s.maxw, s.npr};
}
template <typename T>
__device__ __forceinline__ Sell<T> sell_of(const SellRaw& r) {
    return Sell<T>{r.sptr, r.col, static_cast<const T*>(r.val), r.nrows, r.ncols, r.maxw, r.npr};
}

// One problem's share of a batched launch.  Field use per kernel is listed in
// step_body; unused fields are zero.
struct Step {
    int blk0, nbx, nby, kp;  // first virtual block, virtual grid, column lanes
    int wide, k, n, op;      // wide-row variant, block width, rows, finaliser op
    int nv, nb, flag, honor; // finaliser: sums, partial blocks, flag, honour `done`
    int maxiter, ival;       // set_params: maxiter, reset the batch count; report: output offset
    int subl, pad;           // row kernels: log2 of the rows per unit
    SellRaw M;
    const void* dinv;
    const void* in0;
    const void* in1;
    const void* in2;
    void* out0;
    void* out1;
    void* out2;
    const int* mask;
    const int* done;
    const int* piv;  // LU coarse solve: row interchanges
    double w, tol, atol;
    Ctl ctl;
};

#define AMG_KP_SWITCH(kp, ...)                                  \
    switch (kp) {                                               \
        case 1: { constexpr int KP = 1; __VA_ARGS__; } break;   \
        case 2: { constexpr int KP = 2; __VA_ARGS__; } break;   \
        case 4: { constexpr int KP = 4; __VA_ARGS__; } break;   \
        case 8: { constexpr int KP = 8; __VA_ARGS__; } break;   \
        case 16: { constexpr int KP = 16; __VA_ARGS__; } break; \
        case 32: { constexpr int KP = 32; __VA_ARGS__; } break; \
        default: { constexpr int KP = 64; __VA_ARGS__; } break; \
    }
// KPC > 0: the launch holds steps of one column-lane count only (compile-time KP)
#define AMG_KP(KPC, kp, ...)                                          \
    if constexpr (KPC > 0) { constexpr int KP = KPC; __VA_ARGS__; } \
    else { AMG_KP_SWITCH(kp, __VA_ARGS__) }
#define AMG_WIDE_SWITCH(wd, ...)                                 \
    if (wd) { constexpr bool WD = true; __VA_ARGS__; }           \
    else { constexpr bool WD = false; __VA_ARGS__; }
// vector-mode steps (wide == VMODE) have a column chunk of 16, 32 or 64
#define AMG_VKP(KPC, kp, ...)                                                  \
    if constexpr (KPC > 0) {                                                   \
        if constexpr (KPC >= VKP_MIN) { constexpr int KP = KPC; __VA_ARGS__; } \
    } else {                                                                   \
        switch (kp) {                                                          \
            case 16: { constexpr int KP = 16; __VA_ARGS__; } break;            \
            case 32: { constexpr int KP = 32; __VA_ARGS__; } break;            \
            case 64: { constexpr int KP = 64; __VA_ARGS__; } break;            \
            default: break;                                                    \
        }                                                                      \
    }
// a row kernel step: the vector body for wide == VMODE, else the lane-per-column body
#define AMG_ROW(KPC, st, VEC, ...)                                          \
    if ((st).wide == VMODE) { AMG_VKP(KPC, (st).kp, VEC); }                 \
    else { AMG_KP(KPC, (st).kp, AMG_WIDE_SWITCH((st).wide, __VA_ARGS__)); }

template <int KID, typename T, int KPC>
__device__ __forceinline__ void step_body(const Step& st, const VB& vb, const BatchFlag& bf) {
    if constexpr (KID == K_SPMV || KID == K_SPMV_ADD) {
        // y (out0) = [y +] M x (in0)
        const Sell<T> M = sell_of<T>(st.M);
        AMG_ROW(KPC, st,
                (spmv_vbody<T, KP, KID == K_SPMV_ADD>(M, (const T*)st.in0, (T*)st.out0, st.k, st.done, vb)),
                (spmv_body<T, KP, WD, KID == K_SPMV_ADD>(M, (const T*)st.in0, (T*)st.out0, st.k, st.subl, st.done,
                                                         vb)));
    } else if constexpr (KID == K_JACOBI) {
        // xo (out0) = sweep of x (in0) for b (in1)
        const Sell<T> A = sell_of<T>(st.M);
        AMG_ROW(KPC, st,
                (jacobi_vbody<T, KP>(A, (const T*)st.dinv, (const T*)st.in0, (const T*)st.in1, (T*)st.out0,
                                     (T)st.w, st.k, st.mask, st.done, vb)),
                (jacobi_body<T, KP, WD>(A, (const T*)st.dinv, (const T*)st.in0, (const T*)st.in1, (T*)st.out0,
                                        (T)st.w, st.k, st.subl, st.mask, st.done, vb)));
    } else if constexpr (KID == K_JACOBI_ZERO) {
        // xo (out0) = w Dinv b (in0)
        AMG_KP(KPC, st.kp, jacobi_zero_body<T, KP>((const T*)st.dinv, (const T*)st.in0, (T*)st.out0, (T)st.w, st.n,
                                                      st.k, st.done, vb));
    } else if constexpr (KID == K_RESIDUAL) {
        // r (out0) = b (in1) - A x (in0)
        const Sell<T> A = sell_of<T>(st.M);
        AMG_ROW(KPC, st,
                (residual_vbody<T, KP>(A, (const T*)st.in0, (const T*)st.in1, (T*)st.out0, st.k, st.done, vb)),
                (residual_body<T, KP, WD>(A, (const T*)st.in0, (const T*)st.in1, (T*)st.out0, st.k, st.subl,
                                          st.done, vb)));
    } else if constexpr (KID == K_RESIDUAL_ZERO) {
        // x1 (out0) = w Dinv b (in0), r (out1) = b - A x1
        const Sell<T> A = sell_of<T>(st.M);
        AMG_ROW(KPC, st,
                (residual_zero_vbody<T, KP>(A, (const T*)st.dinv, (const T*)st.in0, (T*)st.out0, (T*)st.out1,
                                            (T)st.w, st.k, st.done, vb)),
                (residual_zero_body<T, KP, WD>(A, (const T*)st.dinv, (const T*)st.in0, (T*)st.out0, (T*)st.out1,
                                               (T)st.w, st.k, st.subl, st.done, vb)));
    } else if constexpr (KID == K_PROLONG_X1) {
        // x (out0) = w Dinv b (in0) + P e (in1)
        const Sell<T> P = sell_of<T>(st.M);
        AMG_ROW(KPC, st,
                (prolong_x1_vbody<T, KP>(P, (const T*)st.dinv, (const T*)st.in0, (const T*)st.in1, (T*)st.out0,
                                         (T)st.w, st.k, st.done, vb)),
                (prolong_x1_body<T, KP, WD>(P, (const T*)st.dinv, (const T*)st.in0, (const T*)st.in1,
                                            (T*)st.out0, (T)st.w, st.k, st.subl, st.done, vb)));
    } else if constexpr (KID == K_COPY) {
        AMG_KP(KPC, st.kp, copy_body<T, T, KP>((const T*)st.in0, (T*)st.out0, st.n, st.k, st.mask, st.done, vb));
    } else if constexpr (KID == K_COPY64) {
        AMG_KP(KPC, st.kp, copy_body<double, double, KP>((const double*)st.in0, (double*)st.out0, st.n, st.k,
                                                            st.mask, st.done, vb));
    } else if constexpr (KID == K_COARSE_PINV) {
        // x (out0) = Ainv (in2) b (in0)
        coarse_pinv_body<T, T>((const T*)st.in2, (const T*)st.in0, (T*)st.out0, st.n, st.k, st.done, vb);
    } else if constexpr (KID == K_COARSE_PINV64) {
        coarse_pinv_body<T, double>((const double*)st.in2, (const T*)st.in0, (T*)st.out0, st.n, st.k, st.done, vb);
    } else if constexpr (KID == K_COARSE_LU) {
        // x (out0) = LU (in2, piv) solve of b (in0); work vector out1, one thread per column
        coarse_lu_body<T, T>((const T*)st.in2, st.piv, (const T*)st.in0, (T*)st.out0, (T*)st.out1, st.n, st.k,
                             st.done, vb);
    } else if constexpr (KID == K_COARSE_LU64) {
        coarse_lu_body<T, double>((const double*)st.in2, st.piv, (const T*)st.in0, (T*)st.out0, (double*)st.out1,
                                  st.n, st.k, st.done, vb);
    } else if constexpr (KID == K_COARSE_QR) {
        // x (out0) = R^-1 Q^T b: Q (in2), R (in1), b (in0); work vector out1
        coarse_qr_body<T, T>((const T*)st.in2, (const T*)st.in1, (const T*)st.in0, (T*)st.out0, (T*)st.out1, st.n,
                             st.k, st.done, vb);
    } else if constexpr (KID == K_COARSE_QR64) {
        coarse_qr_body<T, double>((const double*)st.in2, (const double*)st.in1, (const T*)st.in0, (T*)st.out0,
                                  (double*)st.out1, st.n, st.k, st.done, vb);
    } else if constexpr (KID == K_PCG_INIT) {
        // b (in0); x (out0), r (out1), rc (out2)
        AMG_KP(KPC, st.kp, pcg_init_body<T, KP>((const double*)st.in0, (double*)st.out0, (double*)st.out1,
                                                   (T*)st.out2, st.n, st.k, st.ctl, vb));
    } else if constexpr (KID == K_PCG_START) {
        // r (in0), z (in1); p_prev (out0)
        AMG_KP(KPC, st.kp, pcg_start_body<T, KP>((const double*)st.in0, (const T*)st.in1, (double*)st.out0, st.n,
                                                    st.k, st.ctl, st.flag, vb));
    } else if constexpr (KID == K_PCG_P) {
        // p (out0) = z (in0) + beta p
        AMG_KP(KPC, st.kp, pcg_p_body<T, KP>((const T*)st.in0, (double*)st.out0, st.n, st.k, st.ctl, vb));
    } else if constexpr (KID == K_PCG_SPMV) {
        // Ap (out0) = A p (in0)
        const Sell<double> A = sell_of<double>(st.M);
        AMG_ROW(KPC, st, (pcg_spmv_vbody<KP>(A, (const double*)st.in0, (double*)st.out0, st.k, st.ctl, vb)),
                (pcg_spmv_body<KP, WD>(A, (const double*)st.in0, (double*)st.out0, st.k, st.ctl, vb)));
    } else if constexpr (KID == K_PCG_UPDATE) {
        // p (in0), Ap (in1); x (out0), r (out1), rc (out2)
        AMG_KP(KPC, st.kp, pcg_update_body<T, KP>((const double*)st.in0, (const double*)st.in1, (double*)st.out0,
                                                     (double*)st.out1, (T*)st.out2, st.n, st.k, st.ctl, vb));
    } else if constexpr (KID == K_PCG_GAMMA) {
        // r (in0), z (in1), Ap (in2)
        AMG_KP(KPC, st.kp, pcg_gamma_body<T, KP>((const double*)st.in0, (const T*)st.in1, (const double*)st.in2,
                                                    st.n, st.k, st.ctl, st.flag, vb));
    } else if constexpr (KID == K_JACOBI_GAMMA) {
        // z (out0) = last sweep of x (in0) for (T) r (in1); r.z, z.Ap (in2) partials (vector mode)
        const Sell<T> A = sell_of<T>(st.M);
        AMG_VKP(KPC, st.kp, (jacobi_gamma_vbody<T, KP>(A, (const T*)st.dinv, (const T*)st.in0,
                                                          (const double*)st.in1, (const double*)st.in2,
                                                          (T*)st.out0, (T)st.w, st.k, st.flag, st.ctl, vb)));
    } else if constexpr (KID == K_POLY_FIRST) {
        // polynomial sweep, first term: x (in0), b (in1); h (out0), r (out1), or xo (out2)
        const Sell<T> A = sell_of<T>(st.M);
        AMG_KP(KPC, st.kp, AMG_WIDE_SWITCH(st.wide, poly_first_body<T, KP, WD>(
                                                          A, (const T*)st.in0, (const T*)st.in1, (T*)st.out1,
                                                          (T*)st.out0, (T*)st.out2, (T)st.w, st.k, st.subl, st.mask,
                                                          st.done, vb)));
    } else if constexpr (KID == K_POLY_FIRST_ZERO) {
        // zero guess: out0 = c0 b (in0)
        AMG_KP(KPC, st.kp, poly_first_zero_body<T, KP>((const T*)st.in0, (T*)st.out0, (T)st.w, st.n, st.k, st.mask,
                                                          st.done, vb));
    } else if constexpr (KID == K_POLY_STEP) {
        // h (in0), r (in1), x (in2); out0 = c r + A h, or x + (c r + A h) on the last (flag)
        const Sell<T> A = sell_of<T>(st.M);
        AMG_KP(KPC, st.kp, AMG_WIDE_SWITCH(st.wide, poly_step_body<T, KP, WD>(
                                                          A, (const T*)st.in0, (const T*)st.in1, (const T*)st.in2,
                                                          (T*)st.out0, (T)st.w, st.flag, st.k, st.subl, st.mask,
                                                          st.done, vb)));
    } else if constexpr (KID == K_TRUE_RESNORM) {
        // ||b (in1) - A x (in0)||
        const Sell<double> A = sell_of<double>(st.M);
        AMG_KP(KPC, st.kp, AMG_WIDE_SWITCH(st.wide, true_resnorm_body<double, KP, WD>(
                                                          A, (const double*)st.in0, (const double*)st.in1, st.k,
                                                          st.ctl, vb)));
    } else if constexpr (KID == K_FINALIZE) {
        // launched by kb_finalize (FIN_BLOCK threads)
    } else if constexpr (KID == K_SET_PARAMS) {
        set_params_body(st.ctl, st.tol, st.atol, st.maxiter, 0);
        if (st.ival && threadIdx.x == 0) {
            *bf.count = 0;
            *bf.done = 0;
            __hip_atomic_store(bf.hdone, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    } else if constexpr (KID == K_REPORT) {
        // iterations (out0, int) and relative residuals (out1) from offset ival
        for (int c = threadIdx.x; c < st.k; c += BLOCK) {
            ((int*)st.out0)[st.ival + c] = st.ctl.it[c];
            ((double*)st.out1)[st.ival + c] = st.ctl.normr[c] / st.ctl.normb[c];
        }
    }
}

template <int KID, typename T, int KPC>
__global__ __launch_bounds__(BLOCK) void kb_phase(const Step* __restrict__ steps, const int* __restrict__ gmap,
                                                  int total, BatchFlag bf) {
    for (int v = blockIdx.x; v < total; v += gridDim.x) {
        const Step& st = steps[gmap[v >> 3]];
        const int local = v - st.blk0;
        if (local < st.nbx * st.nby) {
            const VB vb{local % st.nbx, local / st.nbx, st.nbx};
            step_body<KID, T, KPC>(st, vb, bf);
        }
        __syncthreads();
    }
}

// batched finaliser: FIN_BLOCK threads per (problem, column)
static __global__ __launch_bounds__(FIN_BLOCK) void kb_finalize(const Step* __restrict__ steps, const int* __restrict__ gmap,
                                                         int total, BatchFlag bf) {
    for (int v = blockIdx.x; v < total; v += gridDim.x) {
        const Step& st = steps[gmap[v >> 3]];
        const int local = v - st.blk0;
        if (local < st.nbx * st.nby) {
            const VB vb{local % st.nbx, local / st.nbx, st.nbx};
            finalize_body(st.ctl, st.op, st.nv, st.nb, st.k, st.flag, st.honor, bf, vb);
        }
        __syncthreads();
    }
}

}  // namespace amg
