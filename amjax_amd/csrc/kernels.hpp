// gfx950 kernels of the multigrid solve phase.
//
// Matrices are sliced ELLPACK with 64-row slices (one slice = one wavefront
// of rows); vectors holding k right-hand sides are row-major (n, k).  A
// launch maps KP lanes to one row (KP = next power of two >= min(k, 64)) and
// gridDim.y to chunks of KP columns, so k = 1 runs one row per lane with fully
// coalesced matrix loads and k = 64 runs one row per wavefront with coalesced
// 512-byte loads of each neighbour row of x.  Every kernel first reads the
// device "done" flag and exits when it is set, so a captured sequence of
// cycles turns into no-ops once the solve has converged.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amg {

constexpr int BLOCK = 256;       // threads of a row-parallel launch
constexpr int RBLOCK = 1024;     // threads of a reducing launch
constexpr int RED_MAXB = 256;    // max blocks of a reducing launch
constexpr int NXCD = 8;          // MI355X accelerator complex dies

template <typename T>
struct Sell {
    const int* __restrict__ sptr;
    const int* __restrict__ col;
    const T* __restrict__ val;
    int nrows, ncols;
};

// Per-solve device state (one instance per workspace).
struct Ctl {
    int* done;      // 1 when no column is active
    int* ticket;    // last-block counter of the reducing kernels (left at 0)
    int* active;    // [k] column still iterating
    int* it;        // [k] iterations done
    int* maxiter;   // [1]
    double* tol;    // [2] tol, atol
    double* normb;  // [k]
    double* normr;  // [k]
    double* gamma;  // [k]
    double* alpha;  // [k]
    double* beta;   // [k]
    double* atol2;  // [k]
    double* part;   // [RED_MAXB * 2 * k] per-block partial sums
    double* red;    // [2 * k] reduced sums
};

__device__ __forceinline__ bool stopped(const int* done) { return done != nullptr && *done != 0; }

// Map hardware block order (round-robin over the 8 XCDs) to a logical order in
// which each XCD owns one contiguous range of rows, so neighbour rows read by
// the sparse products hit the same XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b % NXCD, i = b / NXCD, q = nb / NXCD, r = nb % NXCD;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

template <int KP>
struct RowMap {
    static constexpr int RPB = BLOCK / KP;
    int row, c;
    __device__ __forceinline__ RowMap() {
        const int lb = xcd_block(blockIdx.x, gridDim.x);
        row = lb * RPB + threadIdx.x / KP;
        c = blockIdx.y * KP + (threadIdx.x % KP);
    }
};

// sum_j M[row, j] * x[j, c] in stored (ascending column) order
template <typename T, typename TX>
__device__ __forceinline__ T sell_dot(const Sell<T>& M, int row, const TX* __restrict__ x, int k, int c) {
    const int s = row >> 6;
    const int beg = M.sptr[s] + (row & 63), end = M.sptr[s + 1];
    T acc = T(0);
#pragma unroll 4
    for (int p = beg; p < end; p += 64) {
        const int j = M.col[p];
        if (j >= 0) acc += M.val[p] * static_cast<T>(x[(size_t)j * k + c]);
    }
    return acc;
}

// ------------------------------------------------------------ row-parallel kernels

// y = (ADD ? y : 0) + M x
template <typename T, int KP, bool ADD>
__global__ __launch_bounds__(BLOCK) void k_spmv(Sell<T> M, const T* __restrict__ x, T* y, int k,
                                                const int* done) {
    if (stopped(done)) return;
    RowMap<KP> m;
    if (m.row >= M.nrows || m.c >= k) return;
    const T acc = sell_dot(M, m.row, x, k, m.c);
    const size_t o = (size_t)m.row * k + m.c;
    y[o] = ADD ? y[o] + acc : acc;
}

// one damped Jacobi sweep: xo = (1-w) x + w (Dinv (b - A x) + x); mask: store only active columns
template <typename T, int KP>
__global__ __launch_bounds__(BLOCK) void k_jacobi(Sell<T> A, const T* __restrict__ dinv,
                                                  const T* __restrict__ x, const T* __restrict__ b,
                                                  T* __restrict__ xo, T w, int k, const int* mask,
                                                  const int* done) {
    if (stopped(done)) return;
    RowMap<KP> m;
    if (m.row >= A.nrows || m.c >= k) return;
    if (mask != nullptr && mask[m.c] == 0) return;
    const size_t o = (size_t)m.row * k + m.c;
    const T ax = sell_dot(A, m.row, x, k, m.c);
    const T xv = x[o];
    const T temp = dinv[m.row] * (b[o] - ax) + xv;
    xo[o] = (T(1) - w) * xv + w * temp;
}

// first sweep from a zero guess: xo = w (Dinv b)
template <typename T, int KP>
__global__ __launch_bounds__(BLOCK) void k_jacobi_zero(const T* __restrict__ dinv, const T* __restrict__ b,
                                                       T* __restrict__ xo, T w, int n, int k,
                                                       const int* done) {
    if (stopped(done)) return;
    RowMap<KP> m;
    if (m.row >= n || m.c >= k) return;
    const size_t o = (size_t)m.row * k + m.c;
    xo[o] = w * (dinv[m.row] * b[o]);
}

// r = b - A x
template <typename T, int KP>
__global__ __launch_bounds__(BLOCK) void k_residual(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                                    T* __restrict__ r, int k, const int* done) {
    if (stopped(done)) return;
    RowMap<KP> m;
    if (m.row >= A.nrows || m.c >= k) return;
    const size_t o = (size_t)m.row * k + m.c;
    r[o] = b[o] - sell_dot(A, m.row, x, k, m.c);
}

// Zero-guess pre-smoothing sweep fused with the residual:
//   x1 = w (Dinv b) (stored for the own row, recomputed for neighbours), r = b - A x1
template <typename T, int KP>
__global__ __launch_bounds__(BLOCK) void k_residual_zero(Sell<T> A, const T* __restrict__ dinv,
                                                         const T* __restrict__ b, T* __restrict__ x1,
                                                         T* __restrict__ r, T w, int k, const int* done) {
    if (stopped(done)) return;
    RowMap<KP> m;
    if (m.row >= A.nrows || m.c >= k) return;
    const int s = m.row >> 6;
    const int beg = A.sptr[s] + (m.row & 63), end = A.sptr[s + 1];
    T acc = T(0);
#pragma unroll 4
    for (int p = beg; p < end; p += 64) {
        const int j = A.col[p];
        if (j >= 0) acc += A.val[p] * (w * (dinv[j] * b[(size_t)j * k + m.c]));
    }
    const size_t o = (size_t)m.row * k + m.c;
    const T bv = b[o];
    x1[o] = w * (dinv[m.row] * bv);
    r[o] = bv - acc;
}

template <typename TD, typename TS, int KP>
__global__ __launch_bounds__(BLOCK) void k_copy(const TS* __restrict__ src, TD* __restrict__ dst, int n, int k,
                                                const int* mask, const int* done) {
    if (stopped(done)) return;
    RowMap<KP> m;
    if (m.row >= n || m.c >= k) return;
    if (mask != nullptr && mask[m.c] == 0) return;
    const size_t o = (size_t)m.row * k + m.c;
    dst[o] = static_cast<TD>(src[o]);
}

// x += e for active columns (iterative refinement update)
template <typename TX, typename TE, int KP>
__global__ __launch_bounds__(BLOCK) void k_add_masked(TX* __restrict__ x, const TE* __restrict__ e, int n, int k,
                                                      const int* active, const int* done) {
    if (stopped(done)) return;
    RowMap<KP> m;
    if (m.row >= n || m.c >= k) return;
    if (active[m.c] == 0) return;
    const size_t o = (size_t)m.row * k + m.c;
    x[o] = x[o] + static_cast<TX>(e[o]);
}

// p = z + beta p for active columns
template <typename TZ, int KP>
__global__ __launch_bounds__(BLOCK) void k_pcg_p(const TZ* __restrict__ z, double* __restrict__ p, int n, int k,
                                                 Ctl ctl, const int* done) {
    if (stopped(done)) return;
    RowMap<KP> m;
    if (m.row >= n || m.c >= k) return;
    if (ctl.active[m.c] == 0) return;
    const size_t o = (size_t)m.row * k + m.c;
    p[o] = static_cast<double>(z[o]) + ctl.beta[m.c] * p[o];
}

// ------------------------------------------------------------ coarse-grid solvers

// x = A_inv b (dense, row-major n x n)
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_coarse_pinv(const T* __restrict__ Ainv, const T* __restrict__ b,
                                                       T* __restrict__ x, int n, int k, const int* done) {
    if (stopped(done)) return;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= n * k) return;
    const int i = t / k, c = t % k;
    T acc = T(0);
    for (int j = 0; j < n; ++j) acc += Ainv[(size_t)i * n + j] * b[(size_t)j * k + c];
    x[(size_t)i * k + c] = acc;
}

// LAPACK-style LU solve with 0-based row interchanges, one thread per column
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_coarse_lu(const T* __restrict__ LU, const int* __restrict__ piv,
                                                     const T* __restrict__ b, T* __restrict__ x, int n, int k,
                                                     const int* done) {
    if (stopped(done)) return;
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= k) return;
    for (int i = 0; i < n; ++i) x[(size_t)i * k + c] = b[(size_t)i * k + c];
    for (int i = 0; i < n; ++i) {
        const int pi = piv[i];
        if (pi != i) {
            const T t = x[(size_t)i * k + c];
            x[(size_t)i * k + c] = x[(size_t)pi * k + c];
            x[(size_t)pi * k + c] = t;
        }
    }
    for (int i = 0; i < n; ++i) {
        T acc = x[(size_t)i * k + c];
        for (int j = 0; j < i; ++j) acc -= LU[(size_t)i * n + j] * x[(size_t)j * k + c];
        x[(size_t)i * k + c] = acc;
    }
    for (int i = n - 1; i >= 0; --i) {
        T acc = x[(size_t)i * k + c];
        for (int j = i + 1; j < n; ++j) acc -= LU[(size_t)i * n + j] * x[(size_t)j * k + c];
        x[(size_t)i * k + c] = acc / LU[(size_t)i * n + i];
    }
}

// x = R^-1 (Q^T b), one thread per column; y is scratch (n, k)
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_coarse_qr(const T* __restrict__ Q, const T* __restrict__ Rm,
                                                     const T* __restrict__ b, T* __restrict__ x, int n, int k,
                                                     const int* done) {
    if (stopped(done)) return;
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= k) return;
    for (int i = 0; i < n; ++i) {
        T acc = T(0);
        for (int j = 0; j < n; ++j) acc += Q[(size_t)j * n + i] * b[(size_t)j * k + c];
        x[(size_t)i * k + c] = acc;
    }
    for (int i = n - 1; i >= 0; --i) {
        T acc = x[(size_t)i * k + c];
        for (int j = i + 1; j < n; ++j) acc -= Rm[(size_t)i * n + j] * x[(size_t)j * k + c];
        x[(size_t)i * k + c] = acc / Rm[(size_t)i * n + i];
    }
}

// ------------------------------------------------------------ deterministic reductions
//
// A reducing kernel covers the rows with a grid-stride loop over at most
// RED_MAXB blocks of RBLOCK threads.  Each block reduces its NV sums per column
// in a fixed order (wave shuffles, then LDS), writes them to ctl.part, and the
// last block to finish (device-scope ticket) adds the per-block sums in block
// order and runs the finaliser.  No floating-point atomics: results are
// bit-reproducible run to run.

template <int KP>
struct RedMap {
    static constexpr int RPB = RBLOCK / KP;
    int lrow, c, lb;
    __device__ __forceinline__ RedMap() {
        lb = xcd_block(blockIdx.x, gridDim.x);
        lrow = threadIdx.x / KP;
        c = blockIdx.y * KP + (threadIdx.x % KP);
    }
};

template <int KP, int NV>
__device__ __forceinline__ bool block_reduce_and_ticket(double (&v)[NV], int k, const Ctl& ctl) {
    __shared__ double sh[RBLOCK / 64][64 * NV];
    __shared__ int last;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
#pragma unroll
        for (int off = 32; off >= KP; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
    }
    if (lane < KP) {
#pragma unroll
        for (int q = 0; q < NV; ++q) sh[wave][q * 64 + lane] = v[q];
    }
    __syncthreads();
    const int nbx = gridDim.x;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    if (threadIdx.x < KP) {
        const int c = blockIdx.y * KP + threadIdx.x;
        if (c < k) {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                double s = 0.0;
                for (int w = 0; w < RBLOCK / 64; ++w) s += sh[w][q * 64 + threadIdx.x];
                ctl.part[((size_t)lb * NV + q) * k + c] = s;
            }
        }
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const int total = nbx * gridDim.y;
        const int t = __hip_atomic_fetch_add(ctl.ticket, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = (t == total - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    // last block: reduce ctl.part over blocks in block order for every (q, c)
    const int P = NV * k;
    for (int q0 = 0; q0 < P; q0 += RBLOCK) {
        const int pc = min(RBLOCK, P - q0);
        const int S = RBLOCK / pc;
        const int qi = threadIdx.x % pc, s = threadIdx.x / pc;
        double acc = 0.0;
        if (s < S) {
            const int q = q0 + qi;
            for (int bx = s; bx < nbx; bx += S)
                acc += __hip_atomic_load(&ctl.part[(size_t)bx * P + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        double* flat = &sh[0][0];
        // (RBLOCK/64)*(64*NV) >= RBLOCK doubles of scratch
        flat[threadIdx.x] = acc;
        __syncthreads();
        if (threadIdx.x < pc) {
            double t = 0.0;
            for (int s2 = 0; s2 < S; ++s2) t += flat[s2 * pc + threadIdx.x];
            ctl.red[q0 + threadIdx.x] = t;
        }
        __syncthreads();
    }
    __threadfence_block();
    if (threadIdx.x == 0) *ctl.ticket = 0;
    __syncthreads();
    return true;
}

// solver parameters (written by a kernel so that no host buffer has to outlive the call)
__global__ void k_set_params(Ctl ctl, double tol, double atol, int maxiter, int done) {
    ctl.tol[0] = tol;
    ctl.tol[1] = atol;
    *ctl.maxiter = maxiter;
    *ctl.done = done;
}

// any column still active -> done = 0 (called by the whole last block)
__device__ __forceinline__ void update_done(const Ctl& ctl, int k) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < k; c += RBLOCK)
        if (ctl.active[c]) any = 1;
    __syncthreads();
    if (threadIdx.x == 0) *ctl.done = any ? 0 : 1;
}

// solve start: x = 0, normb = ||b|| (0 -> 1), normr = ||b||, it = 0,
// active = 0 < maxiter && normr >= tol normb
template <typename T, int KP>
__global__ __launch_bounds__(RBLOCK) void k_solve_init(const T* __restrict__ b, T* __restrict__ x, int n, int k,
                                                       Ctl ctl) {
    RedMap<KP> m;
    double v[1] = {0.0};
    if (m.c < k) {
        for (int r0 = m.lb * RedMap<KP>::RPB; r0 < n; r0 += gridDim.x * RedMap<KP>::RPB) {
            const int row = r0 + m.lrow;
            if (row < n) {
                const size_t o = (size_t)row * k + m.c;
                const double bv = static_cast<double>(b[o]);
                v[0] += bv * bv;
                if (x != nullptr) x[o] = T(0);
            }
        }
    }
    if (!block_reduce_and_ticket<KP, 1>(v, k, ctl)) return;
    const double tol = ctl.tol[0];
    const int maxit = *ctl.maxiter;
    for (int c = threadIdx.x; c < k; c += RBLOCK) {
        const double nb = sqrt(ctl.red[c]);
        ctl.normb[c] = nb == 0.0 ? 1.0 : nb;
        ctl.normr[c] = nb;
        ctl.it[c] = 0;
        ctl.active[c] = (0 < maxit) && (nb >= tol * ctl.normb[c]);
    }
    update_done(ctl, k);
}

// after a cycle: r = b - A x (optionally stored as TR), normr = ||r||,
// it += active, active = it < maxiter && normr >= tol normb
template <typename T, typename TR, int KP>
__global__ __launch_bounds__(RBLOCK) void k_resnorm(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                                    TR* __restrict__ r_out, int k, Ctl ctl, int count_iter) {
    if (stopped(ctl.done)) return;
    RedMap<KP> m;
    const int n = A.nrows;
    double v[1] = {0.0};
    if (m.c < k) {
        for (int r0 = m.lb * RedMap<KP>::RPB; r0 < n; r0 += gridDim.x * RedMap<KP>::RPB) {
            const int row = r0 + m.lrow;
            if (row < n) {
                const size_t o = (size_t)row * k + m.c;
                const T rv = b[o] - sell_dot(A, row, x, k, m.c);
                if (r_out != nullptr) r_out[o] = static_cast<TR>(rv);
                v[0] += static_cast<double>(rv) * static_cast<double>(rv);
            }
        }
    }
    if (!block_reduce_and_ticket<KP, 1>(v, k, ctl)) return;
    const double tol = ctl.tol[0];
    const int maxit = *ctl.maxiter;
    for (int c = threadIdx.x; c < k; c += RBLOCK) {
        const double nr = sqrt(ctl.red[c]);
        ctl.normr[c] = nr;
        if (ctl.active[c]) {
            const int it = ctl.it[c] + count_iter;
            ctl.it[c] = it;
            ctl.active[c] = (it < maxit) && (nr >= tol * ctl.normb[c]);
        }
    }
    update_done(ctl, k);
}

// ------------------------------------------------------------ preconditioned CG (float64 outer)

// x = 0, r = b, rc = (TC) b, atol2 = max(tol^2 b.b, atol^2), it = 0
template <typename TC, int KP>
__global__ __launch_bounds__(RBLOCK) void k_pcg_init(const double* __restrict__ b, double* __restrict__ x,
                                                     double* __restrict__ r, TC* __restrict__ rc, int n, int k,
                                                     Ctl ctl) {
    RedMap<KP> m;
    double v[1] = {0.0};
    if (m.c < k) {
        for (int r0 = m.lb * RedMap<KP>::RPB; r0 < n; r0 += gridDim.x * RedMap<KP>::RPB) {
            const int row = r0 + m.lrow;
            if (row < n) {
                const size_t o = (size_t)row * k + m.c;
                const double bv = b[o];
                v[0] += bv * bv;
                x[o] = 0.0;
                r[o] = bv;
                if (rc != nullptr) rc[o] = static_cast<TC>(bv);
            }
        }
    }
    if (!block_reduce_and_ticket<KP, 1>(v, k, ctl)) return;
    const double tol = ctl.tol[0], atol = ctl.tol[1];
    for (int c = threadIdx.x; c < k; c += RBLOCK) {
        const double bb = ctl.red[c];
        ctl.atol2[c] = fmax(tol * tol * bb, atol * atol);
        ctl.normb[c] = sqrt(bb) == 0.0 ? 1.0 : sqrt(bb);
        ctl.it[c] = 0;
        ctl.active[c] = 1;
    }
}

// p = z, gamma = r.z, active = (precond ? r.r : gamma) > atol2 && 0 < maxiter
template <typename TZ, int KP>
__global__ __launch_bounds__(RBLOCK) void k_pcg_start(const double* __restrict__ r, const TZ* __restrict__ z,
                                                      double* __restrict__ p, int n, int k, Ctl ctl, int precond) {
    RedMap<KP> m;
    double v[2] = {0.0, 0.0};
    if (m.c < k) {
        for (int r0 = m.lb * RedMap<KP>::RPB; r0 < n; r0 += gridDim.x * RedMap<KP>::RPB) {
            const int row = r0 + m.lrow;
            if (row < n) {
                const size_t o = (size_t)row * k + m.c;
                const double rv = r[o], zv = static_cast<double>(z[o]);
                v[0] += rv * zv;
                v[1] += rv * rv;
                p[o] = zv;
            }
        }
    }
    if (!block_reduce_and_ticket<KP, 2>(v, k, ctl)) return;
    const int maxit = *ctl.maxiter;
    for (int c = threadIdx.x; c < k; c += RBLOCK) {
        const double g = ctl.red[c], rr = ctl.red[k + c];
        ctl.gamma[c] = g;
        const double rs = precond ? rr : g;
        ctl.active[c] = (rs > ctl.atol2[c]) && (0 < maxit);
    }
    update_done(ctl, k);
}

// Ap = A p, alpha = gamma / (p.Ap) for active columns (0 otherwise)
template <int KP>
__global__ __launch_bounds__(RBLOCK) void k_pcg_spmv(Sell<double> A, const double* __restrict__ p,
                                                     double* __restrict__ Ap, int k, Ctl ctl) {
    if (stopped(ctl.done)) return;
    RedMap<KP> m;
    const int n = A.nrows;
    double v[1] = {0.0};
    if (m.c < k) {
        for (int r0 = m.lb * RedMap<KP>::RPB; r0 < n; r0 += gridDim.x * RedMap<KP>::RPB) {
            const int row = r0 + m.lrow;
            if (row < n) {
                const size_t o = (size_t)row * k + m.c;
                const double ap = sell_dot(A, row, p, k, m.c);
                Ap[o] = ap;
                v[0] += p[o] * ap;
            }
        }
    }
    if (!block_reduce_and_ticket<KP, 1>(v, k, ctl)) return;
    for (int c = threadIdx.x; c < k; c += RBLOCK)
        ctl.alpha[c] = ctl.active[c] ? ctl.gamma[c] / ctl.red[c] : 0.0;
}

// active columns: x += alpha p, r -= alpha Ap, rc = (TC) r; then
// it += 1, active = r.r > atol2 && it < maxiter
template <typename TC, int KP>
__global__ __launch_bounds__(RBLOCK) void k_pcg_update(const double* __restrict__ p, const double* __restrict__ Ap,
                                                       double* __restrict__ x, double* __restrict__ r,
                                                       TC* __restrict__ rc, int n, int k, Ctl ctl) {
    if (stopped(ctl.done)) return;
    RedMap<KP> m;
    double v[1] = {0.0};
    if (m.c < k) {
        const bool act = ctl.active[m.c] != 0;
        const double a = ctl.alpha[m.c];
        for (int r0 = m.lb * RedMap<KP>::RPB; r0 < n; r0 += gridDim.x * RedMap<KP>::RPB) {
            const int row = r0 + m.lrow;
            if (row < n) {
                const size_t o = (size_t)row * k + m.c;
                double rv = r[o];
                if (act) {
                    x[o] = x[o] + a * p[o];
                    rv = rv - a * Ap[o];
                    r[o] = rv;
                    if (rc != nullptr) rc[o] = static_cast<TC>(rv);
                }
                v[0] += rv * rv;
            }
        }
    }
    if (!block_reduce_and_ticket<KP, 1>(v, k, ctl)) return;
    const int maxit = *ctl.maxiter;
    for (int c = threadIdx.x; c < k; c += RBLOCK) {
        ctl.normr[c] = sqrt(ctl.red[c]);
        if (ctl.active[c]) {
            const int it = ctl.it[c] + 1;
            ctl.it[c] = it;
            ctl.active[c] = (ctl.red[c] > ctl.atol2[c]) && (it < maxit);
        }
    }
    update_done(ctl, k);
}

// gamma' = r.z; beta = gamma'/gamma (Hestenes-Stiefel) or -alpha (z.Ap)/gamma
// (Polak-Ribiere, flexible); gamma = gamma' - for active columns
template <typename TZ, int KP>
__global__ __launch_bounds__(RBLOCK) void k_pcg_gamma(const double* __restrict__ r, const TZ* __restrict__ z,
                                                      const double* __restrict__ Ap, int n, int k, Ctl ctl,
                                                      int flexible) {
    if (stopped(ctl.done)) return;
    RedMap<KP> m;
    double v[2] = {0.0, 0.0};
    if (m.c < k) {
        for (int r0 = m.lb * RedMap<KP>::RPB; r0 < n; r0 += gridDim.x * RedMap<KP>::RPB) {
            const int row = r0 + m.lrow;
            if (row < n) {
                const size_t o = (size_t)row * k + m.c;
                const double zv = static_cast<double>(z[o]);
                v[0] += r[o] * zv;
                v[1] += zv * Ap[o];
            }
        }
    }
    if (!block_reduce_and_ticket<KP, 2>(v, k, ctl)) return;
    for (int c = threadIdx.x; c < k; c += RBLOCK) {
        if (!ctl.active[c]) continue;
        const double g = ctl.red[c], zap = ctl.red[k + c];
        ctl.beta[c] = flexible ? -ctl.alpha[c] * zap / ctl.gamma[c] : g / ctl.gamma[c];
        ctl.gamma[c] = g;
    }
}

// ||b - A x|| per column into ctl.normr (no state update)
template <typename T, int KP>
__global__ __launch_bounds__(RBLOCK) void k_true_resnorm(Sell<T> A, const T* __restrict__ x,
                                                         const T* __restrict__ b, int k, Ctl ctl) {
    RedMap<KP> m;
    const int n = A.nrows;
    double v[1] = {0.0};
    if (m.c < k) {
        for (int r0 = m.lb * RedMap<KP>::RPB; r0 < n; r0 += gridDim.x * RedMap<KP>::RPB) {
            const int row = r0 + m.lrow;
            if (row < n) {
                const size_t o = (size_t)row * k + m.c;
                const double rv = static_cast<double>(b[o] - sell_dot(A, row, x, k, m.c));
                v[0] += rv * rv;
            }
        }
    }
    if (!block_reduce_and_ticket<KP, 1>(v, k, ctl)) return;
    for (int c = threadIdx.x; c < k; c += RBLOCK) ctl.normr[c] = sqrt(ctl.red[c]);
}

}  // namespace amg
