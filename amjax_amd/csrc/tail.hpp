// Coarse tail of a V-cycle in one launch.
//
// Below some level the hierarchy is small (at 500 x 500: 532, 65 and 9
// unknowns) and a cycle's visits of those levels are a chain of short
// dependent launches (down-sweep products, coarse solve, prolongations,
// post-sweeps) whose cost is launch and memory latency, not work.  k_tail runs
// the whole chain from level l0 down to the coarsest and back in one launch:
// one workgroup per column of the block.  At set-up the handle packs every
// operator the tail reads (D, P, A, D^-1 per level, the coarse A_inv) into one
// contiguous device image; each workgroup copies the image into LDS with one
// round of 16-byte loads, together with its column of b, so that every later
// phase reads LDS only, separated by workgroup barriers.  Every phase computes
// the same row sums as the per-level kernel it replaces (products in slot
// order, or the wide-row lane split where that launch uses it) with the same
// epilogue expressions, so the tail's output equals the per-level launches bit
// for bit.  It covers the levels' visits of a V, W, F cycle (or the F-cycle's
// transpose 'T') from a zero or a nonzero guess, when every level takes one
// Jacobi pre-sweep with a fused down-sweep operator D for its zero-guess
// visits, one Jacobi post-sweep, and a pseudo-inverse coarse solve.  The host
// writes the visit sequence as a program of phases (TailOp) that mirrors the
// per-level emission step for step; the kernel runs it with a workgroup
// barrier after each phase.
#pragma once
#include "kernels.hpp"

namespace amg {

constexpr int TAIL_BLOCK = 512;   // threads per workgroup (one column)
constexpr int TAIL_MAXL = 8;      // levels above the coarsest one launch may cover
constexpr int TAIL_COPY_UNROLL = 20;  // 16-byte image loads per thread: images up to 160 KiB
constexpr int TAIL_B_UNROLL = 8;      // rows of b per thread and round
constexpr int TAIL_MAXOPS = 96;       // phases of one launch's program

// A sliced-ELL matrix in the image, compacted: slice s holds v_s = min(64,
// nrows - 64 s) rows, slot t of its row q at sp[s] + t v_s + q (the global
// layout, minus the padding rows of a partial last slice).  Byte offsets from
// the image start.
struct TailMat {
    int nrows, sp, col, val;
};

// one level above the coarsest.  tD, tP, tA, tR: the per-level launch over
// that matrix splits rows over WIDE_LANES lanes (one column per lane, wide
// rows), whose sums the tail then reproduces.  ob: element offset, in the
// vector region, of the level's right-hand side (the program names every
// other vector itself).  R is in the image only when nonzero-guess visits can
// occur (R.nrows = 0 otherwise).
template <typename T>
struct TailLv {
    TailMat D, P, A, R;
    int dinv;  // byte offset of D^-1 (nrows values)
    T w_pre, w_post;
    int tD, tP, tA, tR;
    int ob;      // right-hand side
    int oy, ox;  // the zero-guess V visit's prolonged and post-swept iterates (k_tail_v)
};

// One phase of the program: level `lvl`, vector operands as element offsets
// in the vector region (out = -1: the launch's global output).
enum TailOpKind {
    TOP_DOWN_D = 0,  // out = D b                      (zero-guess pre-sweep, residual, restriction)
    TOP_PRE = 1,     // out = one Jacobi sweep of in (w_pre)
    TOP_RES = 2,     // out = b - A in
    TOP_RESTRICT = 3,  // out = R in
    TOP_COARSE = 4,  // x_c = A_inv b_c
    TOP_UP_X1 = 5,   // out = w_pre (D^-1 b) + P in     (prolongation after a fused down-sweep)
    TOP_UP_ADD = 6,  // out = aux + P in                (prolongation added to the iterate aux)
    TOP_POST = 7     // out = one Jacobi sweep of in (w_post)
};
struct TailOp {
    int kind, lvl, in, out, aux;
};

template <typename T, typename TF>
struct TailArgs {
    TailLv<T> L[TAIL_MAXL];  // levels l0 .. l0 + nl - 1
    int nl;
    const uint4* img;  // the image, img_n 16-byte units
    int img_n;
    int ainv;          // byte offset of the coarsest level's A_inv (nc x nc, TF)
    int nc, obc, oxc;  // its size; element offsets of its b and x
    int vec;           // byte offset of the vector region in LDS
    const T* b;        // right-hand side of level l0 (n x k, row-major)
    const T* xin;      // iterate of level l0 on entry (nullptr: zero guess)
    T* x;              // result on level l0 (n x k)
    int oxin;          // element offset where xin is staged
    int k;
    const int* done;
    int nops;
    TailOp ops[TAIL_MAXOPS];
};
static_assert(sizeof(TailArgs<double, double>) <= 4096, "the tail's arguments must fit the 4 KiB kernel-argument segment");

// C slots of one row, slot i at p0 + i v: every load of the chunk, then every
// gather (a padding slot, col -1, gathers entry 0 and contributes 0), then the
// products, added into acc in slot order (TREE: slot i into acc[i], the sums
// of the WIDE_LANES lanes of a wide-row launch, whose slot t goes to lane
// t mod WIDE_LANES; a chunk starts at a multiple of WIDE_LANES).  Straight-line
// code: LDS reads need no guards inside a slice.
template <typename T, int C, bool TREE, typename G>
__device__ __forceinline__ void tail_chunk(const int* col, const T* val, int p0, int v, G&& gather,
                                           T (&acc)[TREE ? WIDE_LANES : 1]) {
    int jj[C];
    T vv[C], xv[C];
#pragma unroll
    for (int i = 0; i < C; ++i) {
        jj[i] = col[p0 + i * v];
        vv[i] = val[p0 + i * v];
    }
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const T g = gather(jj[i] < 0 ? 0 : jj[i]);
        xv[i] = jj[i] < 0 ? T(0) : g;
    }
#pragma unroll
    for (int i = 0; i < C; ++i) acc[TREE ? i : 0] += vv[i] * xv[i];
}

// Row r of a staged matrix: sum_t M[r, col_t] gather(col_t), products added
// in slot order (TREE: as the wide-row kernels do, lane g of a WIDE_LANES
// group summing slots g, g + WIDE_LANES, ... in order, the lane sums combined
// by the xor tree ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7))).
template <typename T, bool TREE, typename G>
__device__ __forceinline__ T tail_dot(const unsigned char* lds, const TailMat& m, int r, G&& gather) {
    static_assert(WIDE_LANES == 8, "chunks of 8 slots: one slot per lane of a wide-row group");
    const int* sp = reinterpret_cast<const int*>(lds + m.sp);
    const int* col = reinterpret_cast<const int*>(lds + m.col);
    const T* val = reinterpret_cast<const T*>(lds + m.val);
    const int s = r >> 6, q = r & 63;
    const int v = min(64, m.nrows - (s << 6));
    const int base = sp[s] + q;
    const int w = (sp[s + 1] - sp[s]) / v;
    T acc[TREE ? WIDE_LANES : 1];
#pragma unroll
    for (int i = 0; i < (TREE ? WIDE_LANES : 1); ++i) acc[i] = T(0);
    int t0 = 0;
    for (; t0 + 8 <= w; t0 += 8) tail_chunk<T, 8, TREE>(col, val, base + t0 * v, v, gather, acc);
    switch (w - t0) {
#define AMG_TAIL_REM(C) \
    case C: tail_chunk<T, C, TREE>(col, val, base + t0 * v, v, gather, acc); break;
        AMG_TAIL_REM(1)
        AMG_TAIL_REM(2)
        AMG_TAIL_REM(3)
        AMG_TAIL_REM(4)
        AMG_TAIL_REM(5)
        AMG_TAIL_REM(6)
        AMG_TAIL_REM(7)
#undef AMG_TAIL_REM
        default: break;
    }
    if constexpr (TREE) return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    else return acc[0];
}

// both orders, chosen per matrix
template <typename T, typename G>
__device__ __forceinline__ T tail_dot(const unsigned char* lds, const TailMat& m, int r, bool tree, G&& gather) {
    return tree ? tail_dot<T, true>(lds, m, r, gather) : tail_dot<T, false>(lds, m, r, gather);
}

// stage this workgroup's column of b_l0 (and of the iterate xin, when given)
// and the image.  Out-of-range indices are clamped to the last element (the
// same value written again), so every load of a round is issued before the
// first store.
template <typename T, typename TF>
__device__ __forceinline__ void tail_stage(const TailArgs<T, TF>& a, unsigned char* lds, uint4* lds4, T* s) {
    const int c = blockIdx.x, k = a.k, tid = threadIdx.x;
    const int n0 = a.L[0].A.nrows;
    for (int r0 = 0; r0 < n0; r0 += TAIL_B_UNROLL * TAIL_BLOCK) {
        T bv[TAIL_B_UNROLL];
#pragma unroll
        for (int u = 0; u < TAIL_B_UNROLL; ++u) {
            const int r = min(r0 + u * TAIL_BLOCK + tid, n0 - 1);
            bv[u] = a.b[(size_t)r * k + c];
        }
#pragma unroll
        for (int u = 0; u < TAIL_B_UNROLL; ++u) s[a.L[0].ob + min(r0 + u * TAIL_BLOCK + tid, n0 - 1)] = bv[u];
    }
    {
        uint4 v[TAIL_COPY_UNROLL];  // the image is at most TAIL_COPY_UNROLL * TAIL_BLOCK units
#pragma unroll
        for (int u = 0; u < TAIL_COPY_UNROLL; ++u) v[u] = a.img[min(u * TAIL_BLOCK + tid, a.img_n - 1)];
#pragma unroll
        for (int u = 0; u < TAIL_COPY_UNROLL; ++u) lds4[min(u * TAIL_BLOCK + tid, a.img_n - 1)] = v[u];
    }
    if (a.xin != nullptr) {
        for (int r0 = 0; r0 < n0; r0 += TAIL_B_UNROLL * TAIL_BLOCK) {
            T xv[TAIL_B_UNROLL];
#pragma unroll
            for (int u = 0; u < TAIL_B_UNROLL; ++u) {
                const int r = min(r0 + u * TAIL_BLOCK + tid, n0 - 1);
                xv[u] = a.xin[(size_t)r * k + c];
            }
#pragma unroll
            for (int u = 0; u < TAIL_B_UNROLL; ++u) s[a.oxin + min(r0 + u * TAIL_BLOCK + tid, n0 - 1)] = xv[u];
        }
    }
    (void)lds;
    __syncthreads();
}

// The zero-guess V visit, written out (the phases of its program in a fixed
// order, each level's fields read once per phase): the common case, without
// the interpreter's per-phase loads of the program.
template <typename T, typename TF>
__global__ __launch_bounds__(TAIL_BLOCK) void k_tail_v(const TailArgs<T, TF> a) {
    extern __shared__ uint4 tail_lds4[];
    unsigned char* lds = reinterpret_cast<unsigned char*>(tail_lds4);
    if (stopped(a.done)) return;
    const int c = blockIdx.x, k = a.k, tid = threadIdx.x;
    T* s = reinterpret_cast<T*>(lds + a.vec);
    tail_stage(a, lds, tail_lds4, s);
    // down: b_{l+1} = D_l b_l  (pre-sweep from zero, residual and restriction)
    for (int l = 0; l < a.nl; ++l) {
        const TailMat D = a.L[l].D;
        const bool tree = a.L[l].tD != 0;
        const T* sb = s + a.L[l].ob;
        T* sb1 = s + (l + 1 < a.nl ? a.L[l + 1].ob : a.obc);
        for (int r = tid; r < D.nrows; r += TAIL_BLOCK) sb1[r] = tail_dot<T>(lds, D, r, tree, [&](int j) { return sb[j]; });
        __syncthreads();
    }
    // coarsest level: x = A_inv b, summed in TF and rounded once
    {
        const TF* ainv = reinterpret_cast<const TF*>(lds + a.ainv);
        const T* sbc = s + a.obc;
        T* sxc = s + a.oxc;
        const int nc = a.nc;
        for (int r = tid; r < nc; r += TAIL_BLOCK) {
            TF acc = TF(0);
            for (int j = 0; j < nc; ++j) acc += ainv[r * nc + j] * static_cast<TF>(sbc[j]);
            sxc[r] = static_cast<T>(acc);
        }
    }
    __syncthreads();
    // up: x = w_pre (Dinv b) + P e, then one damped Jacobi sweep
    for (int l = a.nl - 1; l >= 0; --l) {
        const TailMat P = a.L[l].P, A = a.L[l].A;
        const bool tP = a.L[l].tP != 0, tA = a.L[l].tA != 0;
        const T w_pre = a.L[l].w_pre, w = a.L[l].w_post;
        const T* dinv = reinterpret_cast<const T*>(lds + a.L[l].dinv);
        const T* sb = s + a.L[l].ob;
        T* st = s + a.L[l].oy;
        T* sx = l == 0 ? nullptr : s + a.L[l].ox;
        const T* se = s + (l + 1 < a.nl ? a.L[l + 1].ox : a.oxc);
        for (int r = tid; r < P.nrows; r += TAIL_BLOCK) {
            const T pe = tail_dot<T>(lds, P, r, tP, [&](int m) { return se[m]; });
            st[r] = w_pre * (dinv[r] * sb[r]) + pe;
        }
        __syncthreads();
        for (int r = tid; r < A.nrows; r += TAIL_BLOCK) {
            const T ax = tail_dot<T>(lds, A, r, tA, [&](int j) { return st[j]; });
            const T xv = st[r];
            const T temp = dinv[r] * (sb[r] - ax) + xv;
            const T o = (T(1) - w) * xv + w * temp;
            if (l == 0) a.x[(size_t)r * k + c] = o;
            else sx[r] = o;
        }
        __syncthreads();
    }
}

// The program interpreter (every visit kind).  grid: k workgroups of
// TAIL_BLOCK threads; dynamic LDS: vec + the vector region's bytes
template <typename T, typename TF>
__global__ __launch_bounds__(TAIL_BLOCK) void k_tail(const TailArgs<T, TF> a) {
    extern __shared__ uint4 tail_lds4[];
    unsigned char* lds = reinterpret_cast<unsigned char*>(tail_lds4);
    if (stopped(a.done)) return;
    const int c = blockIdx.x, k = a.k, tid = threadIdx.x;
    T* s = reinterpret_cast<T*>(lds + a.vec);
    tail_stage(a, lds, tail_lds4, s);
    // the program: each phase copies its level's fields out of the kernel
    // arguments first (a field read inside a row loop would be a scalar load,
    // and its wait, between every two gathers)
    for (int o = 0; o < a.nops; ++o) {
        const TailOp op = a.ops[o];
        if (op.kind == TOP_COARSE) {
            // x = A_inv b, summed in TF and rounded once
            const TF* ainv = reinterpret_cast<const TF*>(lds + a.ainv);
            const T* sbc = s + a.obc;
            T* sxc = s + a.oxc;
            const int nc = a.nc;
            for (int r = tid; r < nc; r += TAIL_BLOCK) {
                TF acc = TF(0);
                for (int j = 0; j < nc; ++j) acc += ainv[r * nc + j] * static_cast<TF>(sbc[j]);
                sxc[r] = static_cast<T>(acc);
            }
        } else if (op.kind == TOP_DOWN_D || op.kind == TOP_RESTRICT) {
            const TailMat M = op.kind == TOP_DOWN_D ? a.L[op.lvl].D : a.L[op.lvl].R;
            const bool tree = (op.kind == TOP_DOWN_D ? a.L[op.lvl].tD : a.L[op.lvl].tR) != 0;
            const T* si = s + op.in;
            T* so = s + op.out;
            for (int r = tid; r < M.nrows; r += TAIL_BLOCK) so[r] = tail_dot<T>(lds, M, r, tree, [&](int j) { return si[j]; });
        } else if (op.kind == TOP_UP_X1 || op.kind == TOP_UP_ADD) {
            const TailMat P = a.L[op.lvl].P;
            const bool tP = a.L[op.lvl].tP != 0;
            const T w_pre = a.L[op.lvl].w_pre;
            const T* dinv = reinterpret_cast<const T*>(lds + a.L[op.lvl].dinv);
            const T* sb = s + a.L[op.lvl].ob;
            const T* se = s + op.in;
            const T* sx = s + op.aux;
            T* st = s + op.out;
            // two loops: a select between the two epilogues would let the
            // compiler split w_pre (D^-1 b) + P e into a product and an add,
            // where the per-level kernel fuses them
            if (op.kind == TOP_UP_X1) {
                for (int r = tid; r < P.nrows; r += TAIL_BLOCK) {
                    const T pe = tail_dot<T>(lds, P, r, tP, [&](int m) { return se[m]; });
                    st[r] = w_pre * (dinv[r] * sb[r]) + pe;
                }
            } else {
                for (int r = tid; r < P.nrows; r += TAIL_BLOCK) {
                    const T pe = tail_dot<T>(lds, P, r, tP, [&](int m) { return se[m]; });
                    st[r] = sx[r] + pe;
                }
            }
        } else if (op.kind == TOP_RES) {
            const TailMat A = a.L[op.lvl].A;
            const bool tA = a.L[op.lvl].tA != 0;
            const T* sb = s + a.L[op.lvl].ob;
            const T* sx = s + op.in;
            T* sr = s + op.out;
            for (int r = tid; r < A.nrows; r += TAIL_BLOCK) {
                const T ax = tail_dot<T>(lds, A, r, tA, [&](int j) { return sx[j]; });
                sr[r] = sb[r] - ax;
            }
        } else {
            // TOP_PRE / TOP_POST: one damped Jacobi sweep
            const TailMat A = a.L[op.lvl].A;
            const bool tA = a.L[op.lvl].tA != 0;
            const T w = op.kind == TOP_PRE ? a.L[op.lvl].w_pre : a.L[op.lvl].w_post;
            const T* dinv = reinterpret_cast<const T*>(lds + a.L[op.lvl].dinv);
            const T* sb = s + a.L[op.lvl].ob;
            const T* st = s + op.in;
            T* sx = op.out >= 0 ? s + op.out : nullptr;
            for (int r = tid; r < A.nrows; r += TAIL_BLOCK) {
                const T ax = tail_dot<T>(lds, A, r, tA, [&](int j) { return st[j]; });
                const T xv = st[r];
                const T temp = dinv[r] * (sb[r] - ax) + xv;
                const T ov = (T(1) - w) * xv + w * temp;
                if (sx == nullptr) a.x[(size_t)r * k + c] = ov;
                else sx[r] = ov;
            }
        }
        __syncthreads();
    }
}

}  // namespace amg
