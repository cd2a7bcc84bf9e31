// gfx950 kernels of the multigrid solve phase.
//
// Layout.  Matrices are sliced ELLPACK with 64-row slices: slice s holds rows
// 64s..64s+63 and entry t of row 64s+r sits at sptr[s] + 64 t + r (col = -1
// marks padding).  A block of k right-hand sides is row-major (n, k).
//
// Work mapping.  One wavefront owns one 64-row slice.  With k = 1 a lane owns
// a row and reads its slots with fully coalesced loads.  With k >= 2 the lanes
// are split into 64/KP row groups of KP = next_pow2(min(k, 64)) column lanes;
// the wavefront first loads the whole slice's column indices and values into
// registers (coalesced, once), then walks the rows and broadcasts each row's
// (col, val) pairs with lane shuffles, so every gather of a neighbour row of x
// is a coalesced k-wide load and all gathers of a row are in flight together.
// Slices wider than WMAX slots (coarse levels, restriction) read their slots
// directly instead.  gridDim.y splits k > 64 into 64-column chunks.
//
// Every kernel first reads the device "done" flag and exits when it is set,
// so a captured sequence of cycles turns into no-ops once the solve has
// converged.  Blocks are remapped so each XCD (8 on MI355X, round-robin block
// dispatch) owns one contiguous range of slices and neighbour rows of x are
// found in its own L2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amg {

constexpr int BLOCK = 256;     // every launch: 4 wavefronts = 4 work units
constexpr int NXCD = 8;        // MI355X accelerator complex dies
constexpr int WMAX = 8;        // slice width held in registers
constexpr int WIDE_LANES = 8;  // k = 1: lanes that share one wide row

template <typename T>
struct Sell {
    const int* __restrict__ sptr;
    const int* __restrict__ col;
    const T* __restrict__ val;
    int nrows, ncols, maxw;
    int npr;  // stored entries per row, rounded (host-side tuning key)
};

__host__ __device__ inline int nslices(int n) { return (n + 63) >> 6; }

// A wavefront's unit of work is SUB consecutive rows of one slice.  k = 1: a
// whole slice, one row per lane (wide slices: WIDE_LANES lanes per row).  Slices held in registers: at least 16 rows,
// so the slots are loaded once per 16 rows.  Wide slices: one pass of 64/KP
// rows, so short matrices with long rows still spread over many wavefronts.
__host__ __device__ constexpr int sub_rows(int kp, bool wide) {
    return kp == 1 ? (wide ? 64 / WIDE_LANES : 64) : (wide ? 64 / kp : (64 / kp > 16 ? 64 / kp : 16));
}
__host__ __device__ inline int nunits(int n, int kp, bool wide) { return nslices(n) * (64 / sub_rows(kp, wide)); }

__host__ __device__ constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }
// Rows per unit of the multi-column row kernels are a launch parameter (2^subl,
// from 64/KP (one pass of the row groups) to 64); the k = 1 paths keep theirs.
__host__ __device__ constexpr int default_subl(int kp, bool wide) { return ilog2(sub_rows(kp, wide)); }
__host__ __device__ inline int nunits_l(int n, int kp, bool wide, int subl) {
    if (kp == 1) return nslices(n) * (wide ? WIDE_LANES : 1);
    return nslices(n) << (6 - subl);
}

// Reducing kernels (per-column sums) and the CG vector kernels use short units:
// 4 rows of the slice (at least one pass of the row groups), so waves are
// short-lived and the loads of all passes of a unit are in flight together.
__host__ __device__ constexpr int red_sub(int kp) { return 64 / kp > 4 ? 64 / kp : 4; }
__host__ __device__ inline int nunits_red(int n, int kp) { return nslices(n) * (64 / red_sub(kp)); }

// Per-solve device state (one instance per workspace).
struct Ctl {
    int* done;      // 1 when no column is active
    int* hdone;     // host-mapped mirror of `done` polled by the host loop (may be null)
    int* ticket;    // [2] finaliser block counter, active-column counter (left at 0)
    int* active;    // [k] column still iterating
    int* it;        // [k] iterations done
    int* maxiter;   // [1]
    double* tol;    // [3] tol, atol, reliable-update threshold delta
    double* normb;  // [k]
    double* normr;  // [k]
    double* gamma;  // [k]
    double* alpha;  // [k]
    double* beta;   // [k]
    double* atol2;  // [k]
    double* omega;  // [k] BiCGSTAB omega; reliable-update CG: r.r at the last update
    int* exitf;     // [k] BiCGSTAB early exit of the current iteration; reliable-update CG:
                    //     an update of the column is pending
    int* ruflag;    // [1] reliable-update CG: some column has an update pending
    int* count;     // [k] reliable-update CG: updates done
    double* part;   // [2 * k * blocks] per-block partial sums, part[(q * k + c) * blocks + b]
};

__device__ __forceinline__ bool stopped(const int* done) { return done != nullptr && *done != 0; }

// broadcast lane l's value (l wave-uniform) through a scalar register
__device__ __forceinline__ int bcast(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float bcast(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double bcast(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b % NXCD, i = b / NXCD, q = nb / NXCD, r = nb % NXCD;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// A block's place in its launch: x, y and the x extent.  A plain launch uses
// blockIdx/gridDim; a batched launch hands each problem a virtual grid of its own.
struct VB {
    int x, y, gx;
};
__device__ __forceinline__ VB hw_vb() { return VB{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x}; }

// wavefront index inside the block, held in a scalar register: unit indices
// derived from it are wave-uniform, so the slice pointers of a unit are read
// with scalar loads and do not wait for the previous unit's stores
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// unit of a row-parallel launch (one per wavefront)
__device__ __forceinline__ int row_unit(const VB& vb) {
    return xcd_block(vb.x, vb.gx) * (BLOCK / 64) + wave_id();
}

// Persistent, XCD-aware walk over nu units.  Blocks are dispatched round-robin
// over the 8 XCDs, so the blocks of a launch form G = min(8, gx) groups by x
// mod G; group x owns a contiguous range of units in proportion to its block
// count and its waves stride through that range together, so the neighbour
// rows a unit gathers were brought into the same XCD's L2 by waves running at
// the same time.  With one block per 4 units every wave takes one unit; a
// smaller grid makes the waves loop and removes the partly filled last round.
template <typename F>
__device__ __forceinline__ void for_units(int nu, const VB& vb, F&& f) {
    const int G = vb.gx < NXCD ? vb.gx : NXCD;
    const int x = vb.x % G, lb = vb.x / G;
    const int q = vb.gx / G, r = vb.gx % G;
    const int mx = q + (x < r ? 1 : 0);
    const int before = x * q + (x < r ? x : r);
    const int u0 = (int)((long long)nu * before / vb.gx);
    const int u1 = (int)((long long)nu * (before + mx) / vb.gx);
    const int W = mx * (BLOCK / 64);
    for (int u = u0 + lb * (BLOCK / 64) + wave_id(); u < u1; u += W) f(u);
}

template <int KP>
struct Lane {
    int lane, g, c;
    bool cok;
    __device__ __forceinline__ Lane(int k, const VB& vb) {
        lane = threadIdx.x & 63;
        g = lane / KP;
        c = vb.y * KP + (lane % KP);
        cok = c < k;
    }
};

// sum over the w slots p0, p0 + 64, ... of one row: slot indices and values are
// loaded WMAX at a time, then the WMAX gathers, so a long row costs two
// dependent memory latencies per WMAX slots.  Products are added in slot order.
template <typename T, typename G>
__device__ __forceinline__ T chunked_row_dot(const Sell<T>& M, int p0, int w, bool live, G&& gather,
                                             int stride = 64) {
    T acc = T(0);
    for (int t0 = 0; t0 < w; t0 += WMAX) {
        int jj[WMAX];
        T vv[WMAX], xv[WMAX];
#pragma unroll
        for (int i = 0; i < WMAX; ++i) {
            jj[i] = -1;
            vv[i] = T(0);
            if (t0 + i < w) {
                const int p = p0 + stride * (t0 + i);
                jj[i] = M.col[p];
                vv[i] = M.val[p];
            }
        }
#pragma unroll
        for (int i = 0; i < WMAX; ++i) {
            xv[i] = T(0);
            if (jj[i] >= 0 && live) xv[i] = gather(jj[i]);
        }
#pragma unroll
        for (int i = 0; i < WMAX; ++i)
            if (t0 + i < w) acc += vv[i] * xv[i];
    }
    return acc;
}

// Sum over the w slots of one row (slot t at p0 + 64 t) for a row group of KP
// >= 8 lanes: the group's lanes load KP consecutive slots at once, then each
// slot's (col, val) is broadcast within the group and the gathers are issued
// GB at a time.  Products are added in slot order, as in chunked_row_dot.
template <typename T, int KP, typename G>
__device__ __forceinline__ T coop_row_dot(const Sell<T>& M, int p0, int w, bool live, G&& gather) {
    constexpr int GB = sizeof(T) == 4 ? 16 : 8;
    const int lane = threadIdx.x & 63;
    const int l = lane & (KP - 1), gbase = lane & ~(KP - 1);
    T acc = T(0);
    for (int t0 = 0; t0 < w; t0 += KP) {
        int cj = -1;
        T cv = T(0);
        if (t0 + l < w) {
            cj = M.col[p0 + 64 * (t0 + l)];
            cv = M.val[p0 + 64 * (t0 + l)];
        }
        const int nt = w - t0 < KP ? w - t0 : KP;
        for (int i0 = 0; i0 < nt; i0 += GB) {
            T xv[GB], mv[GB];
#pragma unroll
            for (int i = 0; i < GB; ++i) {
                const int src = gbase + i0 + i;
                const int j = KP == 64 ? bcast(cj, i0 + i < 64 ? i0 + i : 63) : __shfl(cj, src, 64);
                mv[i] = KP == 64 ? bcast(cv, i0 + i < 64 ? i0 + i : 63) : __shfl(cv, src, 64);
                xv[i] = T(0);
                if (i0 + i < nt && j >= 0 && live) xv[i] = gather(j);
            }
#pragma unroll
            for (int i = 0; i < GB; ++i)
                if (i0 + i < nt) acc += mv[i] * xv[i];
        }
    }
    return acc;
}

// Calls epi(row, acc) with acc = sum_t M[row, col_t] * gather(col_t) (slots in
// stored order) for every row of work unit `unit` owned by this lane.  `cok` =
// this lane's column is live (dead lanes still take part in the shuffles).
// Multi-column units hold 2^subl rows of a slice.
template <typename T, int KP, bool WIDE, typename G, typename E>
__device__ __forceinline__ void unit_spmv(const Sell<T>& M, int unit, int subl, bool cok, G&& gather, E&& epi) {
    const int lane = threadIdx.x & 63;
    if constexpr (KP == 1 && WIDE) {
        // WIDE_LANES lanes per row: lane g takes slots g, g + WIDE_LANES, ...;
        // the lane sums are combined by a fixed xor tree
        constexpr int SUB = 64 / WIDE_LANES, U = 64 / SUB;
        const int slice = unit / U, q0 = (unit % U) * SUB;
        const int base = M.sptr[slice];
        const int w = (M.sptr[slice + 1] - base) >> 6;
        const int row0 = (slice << 6) + q0;
        const int rr = lane / WIDE_LANES, g = lane % WIDE_LANES;
        const int ns = w > g ? (w - g + WIDE_LANES - 1) / WIDE_LANES : 0;
        T acc = chunked_row_dot<T>(M, base + 64 * g + q0 + rr, ns, true, gather, 64 * WIDE_LANES);
#pragma unroll
        for (int off = 1; off < WIDE_LANES; off <<= 1) acc += __shfl_xor(acc, off, 64);
        if (g == 0) epi(row0 + rr, acc);
    } else if constexpr (KP == 1) {
        // one row per lane, the whole slice
        const int base = M.sptr[unit];
        const int w = (M.sptr[unit + 1] - base) >> 6;
        epi((unit << 6) + lane, chunked_row_dot<T>(M, base + lane, w, true, gather));
    } else {
        constexpr int RG = 64 / KP;
        const int SUB = 1 << subl;
        const int sh = 6 - subl;
        const int slice = unit >> sh;
        const int q0 = (unit & ((1 << sh) - 1)) << subl;  // first row of the unit inside the slice
        const int base = M.sptr[slice];
        const int w = (M.sptr[slice + 1] - base) >> 6;
        const int row0 = (slice << 6) + q0;
        const int g = lane / KP;
        if constexpr (!WIDE) {
            int cr[WMAX];
            T vr[WMAX];
#pragma unroll
            for (int t = 0; t < WMAX; ++t) {
                cr[t] = -1;
                vr[t] = T(0);
                if (t < w && lane < SUB) {
                    cr[t] = M.col[base + 64 * t + q0 + lane];
                    vr[t] = M.val[base + 64 * t + q0 + lane];
                }
            }
            // one pass = one row per row group; the gathers of four passes are
            // issued together (explicit unroll: the shuffles are convergent, so
            // the compiler does not unroll a loop of run-time length itself)
            auto pass = [&](int r0) {
                const int r = r0 + g;
                T xv[WMAX];
                T mv[WMAX];
#pragma unroll
                for (int t = 0; t < WMAX; ++t) {
                    xv[t] = T(0);
                    mv[t] = T(0);
                    if (t < w) {
                        // KP == 64: one row per pass, r is wave-uniform -> scalar broadcast
                        const int j = KP == 64 ? bcast(cr[t], r0) : __shfl(cr[t], r, 64);
                        mv[t] = KP == 64 ? bcast(vr[t], r0) : __shfl(vr[t], r, 64);
                        if (j >= 0 && cok) xv[t] = gather(j);
                    }
                }
                T acc = T(0);
#pragma unroll
                for (int t = 0; t < WMAX; ++t)
                    if (t < w) acc += mv[t] * xv[t];
                epi(row0 + r, acc);
            };
            if (SUB >= 4 * RG) {
                for (int r0 = 0; r0 < SUB; r0 += 4 * RG) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) pass(r0 + q * RG);
                }
            } else {
                for (int r0 = 0; r0 < SUB; r0 += RG) pass(r0);
            }
        } else if constexpr (KP >= 8) {
            for (int r0 = 0; r0 < SUB; r0 += RG) {
                const int r = r0 + g;
                epi(row0 + r, coop_row_dot<T, KP>(M, base + q0 + r, w, cok, gather));
            }
        } else {
            for (int r0 = 0; r0 < SUB; r0 += RG) {
                const int r = r0 + g;
                epi(row0 + r, chunked_row_dot<T>(M, base + q0 + r, w, cok, gather));
            }
        }
    }
}

// Calls f(row) for every row of work unit `unit` (matrix-free, SUB of a
// register-held slice) owned by this lane.
template <int KP, typename F>
__device__ __forceinline__ void unit_rows(int unit, F&& f) {
    constexpr int SUB = sub_rows(KP, false);
    constexpr int U = 64 / SUB;
    const int lane = threadIdx.x & 63;
    const int row0 = ((unit / U) << 6) + (unit % U) * SUB;
    if constexpr (KP == 1) {
        f(row0 + lane);
    } else {
        constexpr int RG = 64 / KP;
        const int g = lane / KP;
#pragma unroll 4
        for (int r0 = 0; r0 < SUB; r0 += RG) f(row0 + r0 + g);
    }
}

// Rows of reducing unit u owned by this lane, in two phases: ld(q, row, ok) for
// every pass q first (ok = row in range and column live), so the loads of all
// passes are issued together, then st(q, row) for the rows in range.
template <int KP>
__host__ __device__ constexpr int red_passes() {
    return red_sub(KP) / (64 / KP);
}
template <int KP, typename L, typename S>
__device__ __forceinline__ void red_rows(int u, int n, bool cok, L&& ld, S&& st) {
    constexpr int SUB = red_sub(KP), RG = 64 / KP, NP = SUB / RG, U = 64 / SUB;
    const int row0 = ((u / U) << 6) + (u % U) * SUB + (threadIdx.x & 63) / KP;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const int row = row0 + q * RG;
        ld(q, row, row < n && cok);
    }
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const int row = row0 + q * RG;
        if (row < n && cok) st(q, row);
    }
}

// ------------------------------------------------------------ vector rows
//
// Vector mode of the row kernels (VMODE; needs k a multiple of VW = 16 /
// sizeof(T) and a column chunk KP >= VKP_MIN): a lane holds VW consecutive
// columns of one row, so LPR = KP / VW lanes cover a row's column chunk and one
// 16-byte-per-lane load covers RPI = 64 / LPR rows.  A wavefront's unit is
// vrows(KP, VW) = min(64, 2 RPI) rows of one slice (two passes).  For slices of
// at most WVMAX slots the slice width is a template constant, so every slot
// load, gather and epilogue load of the unit is one straight-line sequence
// issued before the unit's first store; wider slices run in chunks of 8 slots.
// Each row still adds its products in slot order, so the results equal the
// one-column-per-lane paths bit for bit.
constexpr int VMODE = 2;
constexpr int WVMAX = 9;
constexpr int VKP_MIN = 16;
template <typename T>
__host__ __device__ constexpr int vec_width() {
    return 16 / (int)sizeof(T);
}
__host__ __device__ constexpr int vrows(int kp, int vw) {
    return 2 * (64 / (kp / vw)) < 64 ? 2 * (64 / (kp / vw)) : 64;
}
__host__ __device__ inline int nunits_v(int n, int kp, int vw) { return nslices(n) * (64 / vrows(kp, vw)); }

template <typename T, int VW>
struct Vec {
    T v[VW];
};
template <typename T, int VW>
__device__ __forceinline__ Vec<T, VW> vzero() {
    Vec<T, VW> r;
#pragma unroll
    for (int e = 0; e < VW; ++e) r.v[e] = T(0);
    return r;
}
template <typename T, int VW>
__device__ __forceinline__ Vec<T, VW> vload(const T* p) {
    typedef T VT __attribute__((ext_vector_type(VW)));
    const VT t = *reinterpret_cast<const VT*>(p);
    Vec<T, VW> r;
#pragma unroll
    for (int e = 0; e < VW; ++e) r.v[e] = t[e];
    return r;
}
template <typename T, int VW>
__device__ __forceinline__ void vstore(T* p, const Vec<T, VW>& a) {
    typedef T VT __attribute__((ext_vector_type(VW)));
    VT t;
#pragma unroll
    for (int e = 0; e < VW; ++e) t[e] = a.v[e];
    *reinterpret_cast<VT*>(p) = t;
}

// zero of a gathered vector type (a gather may return a narrower element
// type than the matrix's; products promote it)
template <typename V>
__device__ __forceinline__ V vzero_of() {
    V r;
#pragma unroll
    for (auto& e : r.v) e = 0;
    return r;
}

template <int KP, int VW>
struct LaneV {
    int sub, c0;  // row within a pass, first of the lane's VW columns
    bool cok;
    __device__ __forceinline__ LaneV(int k, const VB& vb) {
        constexpr int LPR = KP / VW;
        const int lane = threadIdx.x & 63;
        sub = lane / LPR;
        c0 = vb.y * KP + (lane % LPR) * VW;
        cok = c0 < k;
    }
};

// v as held by lane p * RPI + sub (p wave-uniform)
template <int RPI, typename V>
__device__ __forceinline__ V from_row(V v, int p, int sub) {
    if constexpr (RPI == 1) {
        return bcast(v, p);
    } else if constexpr (RPI <= 4) {
        V r = bcast(v, p * RPI);
#pragma unroll
        for (int q = 1; q < RPI; ++q) {
            const V t = bcast(v, p * RPI + q);
            r = sub == q ? t : r;
        }
        return r;
    } else {
        return __shfl(v, p * RPI + sub, 64);
    }
}

// The (slot, row) entries of a unit are loaded packed: entry e = t NR + r
// (slot t, row r of the unit's NR rows) by lane e % 64 of load e / 64, so a
// unit of W slots costs ceil(W NR / 64) index loads and as many value loads
// instead of W each.  from_row_p returns the entry of slot t for the row of
// pass p held by this lane (row p RPI + sub); every lane index is a
// compile-time constant of the unrolled loops, so each is a readlane.
template <int NR, int RPI, int NLD, typename V>
__device__ __forceinline__ V from_row_p(const V (&v)[NLD], int t, int p, int sub) {
    const int e0 = t * NR + p * RPI;
    V r = bcast(v[e0 / 64], e0 % 64);
#pragma unroll
    for (int q = 1; q < RPI; ++q) {
        const int e = e0 + q;
        const V tq = bcast(v[e / 64], e % 64);
        r = sub == q ? tq : r;
    }
    return r;
}
template <int NR, int NLD, typename T>
__device__ __forceinline__ void load_slots(const Sell<T>& M, int base, int q0, int t0, int ns, int (&cr)[NLD],
                                           T (&vr)[NLD]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
        const int e = i * 64 + lane;
        cr[i] = -1;
        vr[i] = T(0);
        if (e < ns) {
            const int o = base + 64 * (t0 + e / NR) + q0 + e % NR;
            cr[i] = M.col[o];
            vr[i] = M.val[o];
        }
    }
}

// One unit of W-slot slices: gather(j) -> Vec (x row j at this lane's
// columns), load(row, ok) -> In (epilogue inputs; ok = row in range and columns
// live), store(row, acc, in) for the rows in range.
template <typename T, int KP, int VW, int W, int NR, typename G, typename L, typename S>
__device__ __forceinline__ void vec_unit_w(const Sell<T>& M, int base, int q0, int row0, bool cok, G&& gather,
                                           L&& load, S&& store) {
    constexpr int LPR = KP / VW, RPI = 64 / LPR, NPASS = NR / RPI;
    constexpr int WA = W > 0 ? W : 1;
    constexpr bool PK = RPI <= 4;  // packed slot loads (readlane broadcasts)
    constexpr int NLD = PK ? (WA * NR + 63) / 64 : WA;
    const int lane = threadIdx.x & 63;
    const int sub = lane / LPR;
    const int n = M.nrows;
    int cr[NLD];
    T vr[NLD];
    if constexpr (PK) {
        load_slots<NR>(M, base, q0, 0, W * NR, cr, vr);
    } else {
#pragma unroll
        for (int t = 0; t < W; ++t) {
            cr[t] = -1;
            vr[t] = T(0);
            if (lane < NR) {
                cr[t] = M.col[base + 64 * t + q0 + lane];
                vr[t] = M.val[base + 64 * t + q0 + lane];
            }
        }
    }
    using In = decltype(load(0, false));
    using GV = decltype(gather(0));
    GV xg[NPASS][WA];
    In in[NPASS];
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
#pragma unroll
        for (int t = 0; t < W; ++t) {
            int j;
            if constexpr (PK) j = from_row_p<NR, RPI, NLD>(cr, t, p, sub);
            else j = from_row<RPI>(cr[t], p, sub);
            xg[p][t] = (j >= 0 && cok) ? gather(j) : vzero_of<GV>();
        }
        const int row = row0 + p * RPI + sub;
        in[p] = load(row, row < n && cok);
    }
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
        Vec<T, VW> acc = vzero<T, VW>();
#pragma unroll
        for (int t = 0; t < W; ++t) {
            T m;
            if constexpr (PK) m = from_row_p<NR, RPI, NLD>(vr, t, p, sub);
            else m = from_row<RPI>(vr[t], p, sub);
#pragma unroll
            for (int e = 0; e < VW; ++e) acc.v[e] += m * xg[p][t].v[e];
        }
        const int row = row0 + p * RPI + sub;
        if (row < n && cok) store(row, acc, in[p]);
    }
}

// W slots starting at slot t0, added into acc (the rows' epilogue inputs were
// loaded before the first chunk)
template <typename T, int KP, int VW, int W, int NPASS, typename G>
__device__ __forceinline__ void vec_chunk(const Sell<T>& M, int base, int q0, int t0, bool cok, G&& gather,
                                          Vec<T, VW> (&acc)[NPASS]) {
    constexpr int LPR = KP / VW, RPI = 64 / LPR, NR = NPASS * RPI;
    constexpr bool PK = RPI <= 4;  // packed slot loads (readlane broadcasts)
    constexpr int NLD = PK ? (W * NR + 63) / 64 : W;
    const int lane = threadIdx.x & 63;
    const int sub = lane / LPR;
    int cr[NLD];
    T vr[NLD];
    if constexpr (PK) {
        load_slots<NR>(M, base, q0, t0, W * NR, cr, vr);
    } else {
#pragma unroll
        for (int t = 0; t < W; ++t) {
            cr[t] = -1;
            vr[t] = T(0);
            if (lane < NR) {
                cr[t] = M.col[base + 64 * (t0 + t) + q0 + lane];
                vr[t] = M.val[base + 64 * (t0 + t) + q0 + lane];
            }
        }
    }
    using GV = decltype(gather(0));
    GV xg[NPASS][W];
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
#pragma unroll
        for (int t = 0; t < W; ++t) {
            int j;
            if constexpr (PK) j = from_row_p<NR, RPI, NLD>(cr, t, p, sub);
            else j = from_row<RPI>(cr[t], p, sub);
            xg[p][t] = (j >= 0 && cok) ? gather(j) : vzero_of<GV>();
        }
    }
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
#pragma unroll
        for (int t = 0; t < W; ++t) {
            T m;
            if constexpr (PK) m = from_row_p<NR, RPI, NLD>(vr, t, p, sub);
            else m = from_row<RPI>(vr[t], p, sub);
#pragma unroll
            for (int e = 0; e < VW; ++e) acc[p].v[e] += m * xg[p][t].v[e];
        }
    }
}

// slices wider than WVMAX: the epilogue loads first, then the slots in chunks
// of CH (each chunk's slot loads and gathers issued together), the remainder
// last; products are still added in slot order
template <typename T, int KP, int VW, int NR, int CH, typename G, typename L, typename S>
__device__ __forceinline__ void vec_unit_any(const Sell<T>& M, int base, int q0, int row0, int w, bool cok,
                                             G&& gather, L&& load, S&& store) {
    constexpr int LPR = KP / VW, RPI = 64 / LPR, NPASS = NR / RPI;
    const int sub = (threadIdx.x & 63) / LPR;
    const int n = M.nrows;
    using In = decltype(load(0, false));
    In in[NPASS];
    Vec<T, VW> acc[NPASS];
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
        const int row = row0 + p * RPI + sub;
        in[p] = load(row, row < n && cok);
        acc[p] = vzero<T, VW>();
    }
    int t0 = 0;
    for (; t0 + CH <= w; t0 += CH) vec_chunk<T, KP, VW, CH, NPASS>(M, base, q0, t0, cok, gather, acc);
    switch (w - t0) {
#define AMG_VEC_TAIL(WC) \
    case WC:                                                                              \
        if constexpr (WC < CH) vec_chunk<T, KP, VW, WC, NPASS>(M, base, q0, t0, cok, gather, acc); \
        break;
        AMG_VEC_TAIL(1)
        AMG_VEC_TAIL(2)
        AMG_VEC_TAIL(3)
        AMG_VEC_TAIL(4)
        AMG_VEC_TAIL(5)
        AMG_VEC_TAIL(6)
        AMG_VEC_TAIL(7)
        AMG_VEC_TAIL(8)
        AMG_VEC_TAIL(9)
        AMG_VEC_TAIL(10)
        AMG_VEC_TAIL(11)
        AMG_VEC_TAIL(12)
        AMG_VEC_TAIL(13)
        AMG_VEC_TAIL(14)
        AMG_VEC_TAIL(15)
#undef AMG_VEC_TAIL
        default: break;
    }
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
        const int row = row0 + p * RPI + sub;
        if (row < n && cok) store(row, acc[p], in[p]);
    }
}

// Walk the vector units of M (persistent, XCD-aware) and run each one.  NRU
// rows per unit (0: vrows(KP, VW); one pass = 64 / (KP / VW) rows); slices
// wider than WVMAX run in chunks of CH slots (at most 16).  VWO: columns per
// lane (0: vec_width<T>(), 16 bytes of T).
template <typename T, int KP, int NRU = 0, int CH = 8, int VWO = 0, int WCAP = 64, typename G, typename L,
          typename S>
__device__ __forceinline__ void vec_rows(const Sell<T>& M, bool cok, const VB& vb, G&& gather, L&& load,
                                         S&& store) {
    constexpr int VW = VWO > 0 ? VWO : vec_width<T>(), NR = NRU > 0 ? NRU : vrows(KP, VW), U = 64 / NR;
    for_units(nslices(M.nrows) * U, vb, [&](int unit) {
        const int slice = unit / U, q0 = (unit % U) * NR;
        const int base = M.sptr[slice];
        const int w = (M.sptr[slice + 1] - base) >> 6;
        const int row0 = (slice << 6) + q0;
        switch (w) {
#define AMG_VEC_W(WC)                                                                              \
    case WC:                                                                                       \
        if constexpr (WC <= WCAP) vec_unit_w<T, KP, VW, WC, NR>(M, base, q0, row0, cok, gather, load, store); \
        break;
            AMG_VEC_W(0)
            AMG_VEC_W(1)
            AMG_VEC_W(2)
            AMG_VEC_W(3)
            AMG_VEC_W(4)
            AMG_VEC_W(5)
            AMG_VEC_W(6)
            AMG_VEC_W(7)
            AMG_VEC_W(8)
            AMG_VEC_W(9)
#undef AMG_VEC_W
            default:
                if constexpr (WCAP > WVMAX) vec_unit_any<T, KP, VW, NR, CH>(M, base, q0, row0, w, cok, gather, load, store);
                break;
        }
    });
}

// ------------------------------------------------------------ row-parallel kernels

// y = (ADD ? y : 0) + M x
template <typename T, int KP, bool WIDE, bool ADD>
__device__ __forceinline__ void spmv_body(Sell<T> M, const T* __restrict__ x, T* __restrict__ y, int k, int subl,
                                                const int* done, const VB& vb) {
    if (stopped(done)) return;
    const Lane<KP> ln(k, vb);
    for_units(nunits_l(M.nrows, KP, WIDE, subl), vb, [&](int unit) {
    unit_spmv<T, KP, WIDE>(
        M, unit, subl, ln.cok, [&](int j) { return x[(size_t)j * k + ln.c]; },
        [&](int row, T acc) {
            if (row < M.nrows && ln.cok) {
                const size_t o = (size_t)row * k + ln.c;
                y[o] = ADD ? y[o] + acc : acc;
            }
        });
    });
}
template <typename T, int KP, bool WIDE, bool ADD>
__global__ __launch_bounds__(BLOCK) void k_spmv(Sell<T> M, const T* __restrict__ x, T* __restrict__ y, int k, int subl,
                                                const int* done) {
    spmv_body<T, KP, WIDE, ADD>(M, x, y, k, subl, done, hw_vb());
}

// one damped Jacobi sweep: xo = (1-w) x + w (Dinv (b - A x) + x); mask: only active columns
template <typename T, int KP, bool WIDE>
__device__ __forceinline__ void jacobi_body(Sell<T> A, const T* __restrict__ dinv, const T* __restrict__ x,
                                                  const T* __restrict__ b, T* __restrict__ xo, T w, int k, int subl,
                                                  const int* __restrict__ mask, const int* done, const VB& vb) {
    if (stopped(done)) return;
    const Lane<KP> ln(k, vb);
    for_units(nunits_l(A.nrows, KP, WIDE, subl), vb, [&](int unit) {
    const bool live = ln.cok && (mask == nullptr || mask[ln.c] != 0);
    unit_spmv<T, KP, WIDE>(
        A, unit, subl, live, [&](int j) { return x[(size_t)j * k + ln.c]; },
        [&](int row, T ax) {
            if (row < A.nrows && live) {
                const size_t o = (size_t)row * k + ln.c;
                const T xv = x[o];
                const T temp = dinv[row] * (b[o] - ax) + xv;
                xo[o] = (T(1) - w) * xv + w * temp;
            }
        });
    });
}
template <typename T, int KP, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_jacobi(Sell<T> A, const T* __restrict__ dinv, const T* __restrict__ x,
                                                  const T* __restrict__ b, T* __restrict__ xo, T w, int k, int subl,
                                                  const int* __restrict__ mask, const int* done) {
    jacobi_body<T, KP, WIDE>(A, dinv, x, b, xo, w, k, subl, mask, done, hw_vb());
}

// Prolongation and correction of a level whose zero-guess pre-sweep was
// folded into the fused down-sweep and never stored:
//   x = w_pre (Dinv b) + P e   (the pre-swept iterate, recomputed per row, plus P e)
template <typename T, int KP, bool WIDE>
__device__ __forceinline__ void prolong_x1_body(Sell<T> P, const T* __restrict__ dinv,
                                                      const T* __restrict__ b, const T* __restrict__ e,
                                                      T* __restrict__ x, T w_pre, int k, int subl, const int* done, const VB& vb) {
    if (stopped(done)) return;
    const Lane<KP> ln(k, vb);
    for_units(nunits_l(P.nrows, KP, WIDE, subl), vb, [&](int unit) {
    unit_spmv<T, KP, WIDE>(
        P, unit, subl, ln.cok, [&](int m) { return e[(size_t)m * k + ln.c]; },
        [&](int row, T pe) {
            if (row < P.nrows && ln.cok) {
                const size_t o = (size_t)row * k + ln.c;
                x[o] = w_pre * (dinv[row] * b[o]) + pe;
            }
        });
    });
}
template <typename T, int KP, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_prolong_x1(Sell<T> P, const T* __restrict__ dinv,
                                                      const T* __restrict__ b, const T* __restrict__ e,
                                                      T* __restrict__ x, T w_pre, int k, int subl, const int* done) {
    prolong_x1_body<T, KP, WIDE>(P, dinv, b, e, x, w_pre, k, subl, done, hw_vb());
}

// first sweep from a zero guess: xo = w (Dinv b)
template <typename T, int KP>
__device__ __forceinline__ void jacobi_zero_body(const T* __restrict__ dinv, const T* __restrict__ b,
                                                       T* __restrict__ xo, T w, int n, int k, const int* done, const VB& vb) {
    if (stopped(done)) return;
    const int unit = row_unit(vb);
    if (unit >= nunits(n, KP, false)) return;
    const Lane<KP> ln(k, vb);
    unit_rows<KP>(unit, [&](int row) {
        if (row < n && ln.cok) {
            const size_t o = (size_t)row * k + ln.c;
            xo[o] = w * (dinv[row] * b[o]);
        }
    });
}
template <typename T, int KP>
__global__ __launch_bounds__(BLOCK) void k_jacobi_zero(const T* __restrict__ dinv, const T* __restrict__ b,
                                                       T* __restrict__ xo, T w, int n, int k, const int* done) {
    jacobi_zero_body<T, KP>(dinv, b, xo, w, n, k, done, hw_vb());
}

// r = b - A x
template <typename T, int KP, bool WIDE>
__device__ __forceinline__ void residual_body(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                                    T* __restrict__ r, int k, int subl, const int* done, const VB& vb) {
    if (stopped(done)) return;
    const Lane<KP> ln(k, vb);
    for_units(nunits_l(A.nrows, KP, WIDE, subl), vb, [&](int unit) {
    unit_spmv<T, KP, WIDE>(
        A, unit, subl, ln.cok, [&](int j) { return x[(size_t)j * k + ln.c]; },
        [&](int row, T ax) {
            if (row < A.nrows && ln.cok) {
                const size_t o = (size_t)row * k + ln.c;
                r[o] = b[o] - ax;
            }
        });
    });
}
template <typename T, int KP, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_residual(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                                    T* __restrict__ r, int k, int subl, const int* done) {
    residual_body<T, KP, WIDE>(A, x, b, r, k, subl, done, hw_vb());
}

// Zero-guess pre-smoothing sweep fused with the residual:
//   x1 = w (Dinv b) (stored for the own row, recomputed for neighbours), r = b - A x1
template <typename T, int KP, bool WIDE>
__device__ __forceinline__ void residual_zero_body(Sell<T> A, const T* __restrict__ dinv,
                                                         const T* __restrict__ b, T* __restrict__ x1,
                                                         T* __restrict__ r, T w, int k, int subl, const int* done, const VB& vb) {
    if (stopped(done)) return;
    const Lane<KP> ln(k, vb);
    for_units(nunits_l(A.nrows, KP, WIDE, subl), vb, [&](int unit) {
    unit_spmv<T, KP, WIDE>(
        A, unit, subl, ln.cok, [&](int j) { return w * (dinv[j] * b[(size_t)j * k + ln.c]); },
        [&](int row, T ax) {
            if (row < A.nrows && ln.cok) {
                const size_t o = (size_t)row * k + ln.c;
                const T bv = b[o];
                x1[o] = w * (dinv[row] * bv);
                r[o] = bv - ax;
            }
        });
    });
}
template <typename T, int KP, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_residual_zero(Sell<T> A, const T* __restrict__ dinv,
                                                         const T* __restrict__ b, T* __restrict__ x1,
                                                         T* __restrict__ r, T w, int k, int subl, const int* done) {
    residual_zero_body<T, KP, WIDE>(A, dinv, b, x1, r, w, k, subl, done, hw_vb());
}

// ---- vector-mode bodies of the row kernels (same results as the bodies above)

struct NoIn {};

template <typename T, int KP, bool ADD, int CH = 8, int WC = 64, int NRU = 0>
__device__ __forceinline__ void spmv_vbody(Sell<T> M, const T* __restrict__ x, T* __restrict__ y, int k,
                                           const int* done, const VB& vb) {
    if (stopped(done)) return;
    constexpr int VW = vec_width<T>();
    const LaneV<KP, VW> ln(k, vb);
    struct In {
        Vec<T, VW> y;
    };
    vec_rows<T, KP, NRU, CH, 0, WC>(
        M, ln.cok, vb, [&](int j) { return vload<T, VW>(x + (size_t)j * k + ln.c0); },
        [&](int row, bool ok) {
            In in;
            in.y = (ADD && ok) ? vload<T, VW>(y + (size_t)row * k + ln.c0) : vzero<T, VW>();
            return in;
        },
        [&](int row, const Vec<T, VW>& acc, const In& in) {
            Vec<T, VW> o = acc;
            if (ADD) {
#pragma unroll
                for (int e = 0; e < VW; ++e) o.v[e] = in.y.v[e] + acc.v[e];
            }
            vstore<T, VW>(y + (size_t)row * k + ln.c0, o);
        });
}
// CH = 16 for matrices wider than WVMAX (the fused down-sweep operators and
// restrictions: half the chunks, twice the gathers in flight per chunk), with
// one pass of rows per unit (NRU = 64 / (KP / VW): half the registers, twice
// the waves).  WC: the widest slice the kernel handles (registers sized for that
// width; 64: any)
template <typename T, int KP, bool ADD, int CH = 8, int WC = 64, int NRU = 0>
__global__ __launch_bounds__(BLOCK) void k_spmv_v(Sell<T> M, const T* __restrict__ x, T* __restrict__ y, int k,
                                                  const int* done) {
    spmv_vbody<T, KP, ADD, CH, WC, NRU>(M, x, y, k, done, hw_vb());
}
// rows per unit of the wide products: one pass
template <typename T, int KP>
constexpr int wide_nru() {
    return 64 / (KP / vec_width<T>());
}

// ACC: instead of writing xo, add the sweep's result to the float64 iterate
// xacc (xacc += (double) sweep, on the columns mask selects), the update
// k_add_masked makes from a stored result, with the same sums
template <typename T, int KP, bool ACC = false, int WC = 64>
__device__ __forceinline__ void jacobi_vbody(Sell<T> A, const T* __restrict__ dinv, const T* __restrict__ x,
                                             const T* __restrict__ b, T* __restrict__ xo, T w, int k,
                                             const int* __restrict__ mask, const int* done, const VB& vb,
                                             double* __restrict__ xacc = nullptr) {
    if (stopped(done)) return;
    constexpr int VW = vec_width<T>();
    const LaneV<KP, VW> ln(k, vb);
    bool lv[VW];
    bool all = true;
#pragma unroll
    for (int e = 0; e < VW; ++e) {
        lv[e] = ln.cok && (mask == nullptr || mask[ln.c0 + e] != 0);
        all = all && lv[e];
    }
    struct In {
        Vec<T, VW> x, b;
        T d;
    };
    vec_rows<T, KP, 0, 8, 0, WC>(
        A, ln.cok, vb, [&](int j) { return vload<T, VW>(x + (size_t)j * k + ln.c0); },
        [&](int row, bool ok) {
            In in;
            if (ok) {
                in.x = vload<T, VW>(x + (size_t)row * k + ln.c0);
                in.b = vload<T, VW>(b + (size_t)row * k + ln.c0);
                in.d = dinv[row];
            } else {
                in.x = vzero<T, VW>();
                in.b = vzero<T, VW>();
                in.d = T(0);
            }
            return in;
        },
        [&](int row, const Vec<T, VW>& ax, const In& in) {
            Vec<T, VW> o;
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                const T temp = in.d * (in.b.v[e] - ax.v[e]) + in.x.v[e];
                o.v[e] = (T(1) - w) * in.x.v[e] + w * temp;
            }
            if constexpr (ACC) {
                double* dst = xacc + (size_t)row * k + ln.c0;
                if (all) {
                    Vec<double, VW> cur = vload<double, VW>(dst);
#pragma unroll
                    for (int e = 0; e < VW; ++e) cur.v[e] = cur.v[e] + static_cast<double>(o.v[e]);
                    vstore<double, VW>(dst, cur);
                } else {
#pragma unroll
                    for (int e = 0; e < VW; ++e)
                        if (lv[e]) dst[e] = dst[e] + static_cast<double>(o.v[e]);
                }
            } else {
                T* dst = xo + (size_t)row * k + ln.c0;
                if (all) {
                    vstore<T, VW>(dst, o);
                } else {
#pragma unroll
                    for (int e = 0; e < VW; ++e)
                        if (lv[e]) dst[e] = o.v[e];
                }
            }
        });
}
// the last top-level sweep of a refinement step added straight into the float64
// iterate (x64 += sweep on the active columns)
template <typename T, int KP, int WC = 64>
__global__ __launch_bounds__(BLOCK) void k_jacobi_acc_v(Sell<T> A, const T* __restrict__ dinv,
                                                        const T* __restrict__ x, const T* __restrict__ b,
                                                        double* __restrict__ xacc, T w, int k,
                                                        const int* __restrict__ active, const int* done) {
    jacobi_vbody<T, KP, true, WC>(A, dinv, x, b, nullptr, w, k, active, done, hw_vb(), xacc);
}
template <typename T, int KP, int WC = 64>
__global__ __launch_bounds__(BLOCK) void k_jacobi_v(Sell<T> A, const T* __restrict__ dinv, const T* __restrict__ x,
                                                    const T* __restrict__ b, T* __restrict__ xo, T w, int k,
                                                    const int* __restrict__ mask, const int* done) {
    jacobi_vbody<T, KP, false, WC>(A, dinv, x, b, xo, w, k, mask, done, hw_vb());
}

template <typename T, int KP, int WC = 64>
__device__ __forceinline__ void prolong_x1_vbody(Sell<T> P, const T* __restrict__ dinv, const T* __restrict__ b,
                                                 const T* __restrict__ e, T* __restrict__ x, T w_pre, int k,
                                                 const int* done, const VB& vb) {
    if (stopped(done)) return;
    constexpr int VW = vec_width<T>();
    const LaneV<KP, VW> ln(k, vb);
    struct In {
        Vec<T, VW> b;
        T d;
    };
    vec_rows<T, KP, 0, 8, 0, WC>(
        P, ln.cok, vb, [&](int m) { return vload<T, VW>(e + (size_t)m * k + ln.c0); },
        [&](int row, bool ok) {
            In in;
            in.b = ok ? vload<T, VW>(b + (size_t)row * k + ln.c0) : vzero<T, VW>();
            in.d = ok ? dinv[row] : T(0);
            return in;
        },
        [&](int row, const Vec<T, VW>& pe, const In& in) {
            Vec<T, VW> o;
#pragma unroll
            for (int q = 0; q < VW; ++q) o.v[q] = w_pre * (in.d * in.b.v[q]) + pe.v[q];
            vstore<T, VW>(x + (size_t)row * k + ln.c0, o);
        });
}
template <typename T, int KP, int WC = 64>
__global__ __launch_bounds__(BLOCK) void k_prolong_x1_v(Sell<T> P, const T* __restrict__ dinv,
                                                        const T* __restrict__ b, const T* __restrict__ e,
                                                        T* __restrict__ x, T w_pre, int k, const int* done) {
    prolong_x1_vbody<T, KP, WC>(P, dinv, b, e, x, w_pre, k, done, hw_vb());
}

template <typename T, int KP, int WC = 64>
__device__ __forceinline__ void residual_vbody(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                               T* __restrict__ r, int k, const int* done, const VB& vb) {
    if (stopped(done)) return;
    constexpr int VW = vec_width<T>();
    const LaneV<KP, VW> ln(k, vb);
    struct In {
        Vec<T, VW> b;
    };
    vec_rows<T, KP, 0, 8, 0, WC>(
        A, ln.cok, vb, [&](int j) { return vload<T, VW>(x + (size_t)j * k + ln.c0); },
        [&](int row, bool ok) {
            In in;
            in.b = ok ? vload<T, VW>(b + (size_t)row * k + ln.c0) : vzero<T, VW>();
            return in;
        },
        [&](int row, const Vec<T, VW>& ax, const In& in) {
            Vec<T, VW> o;
#pragma unroll
            for (int e = 0; e < VW; ++e) o.v[e] = in.b.v[e] - ax.v[e];
            vstore<T, VW>(r + (size_t)row * k + ln.c0, o);
        });
}
template <typename T, int KP, int WC = 64>
__global__ __launch_bounds__(BLOCK) void k_residual_v(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                                      T* __restrict__ r, int k, const int* done) {
    residual_vbody<T, KP, WC>(A, x, b, r, k, done, hw_vb());
}

template <typename T, int KP, int WC = 64>
__device__ __forceinline__ void residual_zero_vbody(Sell<T> A, const T* __restrict__ dinv, const T* __restrict__ b,
                                                    T* __restrict__ x1, T* __restrict__ r, T w, int k,
                                                    const int* done, const VB& vb) {
    if (stopped(done)) return;
    constexpr int VW = vec_width<T>();
    const LaneV<KP, VW> ln(k, vb);
    struct In {
        Vec<T, VW> b;
        T d;
    };
    vec_rows<T, KP, 0, 8, 0, WC>(
        A, ln.cok, vb,
        [&](int j) {
            const Vec<T, VW> bj = vload<T, VW>(b + (size_t)j * k + ln.c0);
            const T dj = dinv[j];
            Vec<T, VW> o;
#pragma unroll
            for (int e = 0; e < VW; ++e) o.v[e] = w * (dj * bj.v[e]);
            return o;
        },
        [&](int row, bool ok) {
            In in;
            in.b = ok ? vload<T, VW>(b + (size_t)row * k + ln.c0) : vzero<T, VW>();
            in.d = ok ? dinv[row] : T(0);
            return in;
        },
        [&](int row, const Vec<T, VW>& ax, const In& in) {
            Vec<T, VW> xo, ro;
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                xo.v[e] = w * (in.d * in.b.v[e]);
                ro.v[e] = in.b.v[e] - ax.v[e];
            }
            vstore<T, VW>(x1 + (size_t)row * k + ln.c0, xo);
            vstore<T, VW>(r + (size_t)row * k + ln.c0, ro);
        });
}
template <typename T, int KP, int WC = 64>
__global__ __launch_bounds__(BLOCK) void k_residual_zero_v(Sell<T> A, const T* __restrict__ dinv,
                                                           const T* __restrict__ b, T* __restrict__ x1,
                                                           T* __restrict__ r, T w, int k, const int* done) {
    residual_zero_vbody<T, KP, WC>(A, dinv, b, x1, r, w, k, done, hw_vb());
}

template <typename TD, typename TS, int KP>
__device__ __forceinline__ void copy_body(const TS* __restrict__ src, TD* __restrict__ dst, int n, int k,
                                                const int* __restrict__ mask, const int* done, const VB& vb) {
    if (stopped(done)) return;
    const int unit = row_unit(vb);
    if (unit >= nunits(n, KP, false)) return;
    const Lane<KP> ln(k, vb);
    const bool live = ln.cok && (mask == nullptr || mask[ln.c] != 0);
    unit_rows<KP>(unit, [&](int row) {
        if (row < n && live) {
            const size_t o = (size_t)row * k + ln.c;
            dst[o] = static_cast<TD>(src[o]);
        }
    });
}
template <typename TD, typename TS, int KP>
__global__ __launch_bounds__(BLOCK) void k_copy(const TS* __restrict__ src, TD* __restrict__ dst, int n, int k,
                                                const int* __restrict__ mask, const int* done) {
    copy_body<TD, TS, KP>(src, dst, n, k, mask, done, hw_vb());
}

// x += e for active columns (iterative refinement update)
template <typename TX, typename TE, int KP>
__device__ __forceinline__ void add_masked_body(TX* __restrict__ x, const TE* __restrict__ e, int n, int k,
                                                      const int* __restrict__ active, const int* done, const VB& vb) {
    if (stopped(done)) return;
    const int unit = row_unit(vb);
    if (unit >= nunits(n, KP, false)) return;
    const Lane<KP> ln(k, vb);
    const bool live = ln.cok && active[ln.c] != 0;
    unit_rows<KP>(unit, [&](int row) {
        if (row < n && live) {
            const size_t o = (size_t)row * k + ln.c;
            x[o] = x[o] + static_cast<TX>(e[o]);
        }
    });
}
template <typename TX, typename TE, int KP>
__global__ __launch_bounds__(BLOCK) void k_add_masked(TX* __restrict__ x, const TE* __restrict__ e, int n, int k,
                                                      const int* __restrict__ active, const int* done) {
    add_masked_body<TX, TE, KP>(x, e, n, k, active, done, hw_vb());
}

// x[i] = v for i < n
// Cache flush for cache-cold timings: every 16-byte vector of a buffer read
// once (clean lines only, so the timed kernel that follows evicts them
// without write-back); sink is written only if the sum takes a value the
// caller never lets it take, which keeps the loads.
static __global__ __launch_bounds__(BLOCK) void k_flush_read(const float4* __restrict__ p, size_t n, float* sink) {
    float a = 0.f;
    for (size_t i = blockIdx.x * (size_t)BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) {
        const float4 v = p[i];
        a += v.x + v.y + v.z + v.w;
    }
    if (a == 1.0f) sink[0] = a;
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_fill(T* __restrict__ x, size_t n, T v) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) x[i] = v;
}

// ------------------------------------------------------------ polynomial smoother
//
// One sweep of PyAMG's polynomial smoother (the 'chebyshev' and 'richardson'
// smoothers): r = b - A x, h = c_0 r, h = c_i r + A h (i = 1..), x += h.  One
// kernel per coefficient: poly_first forms r and c_0 r (with a single
// coefficient it writes x + c_0 r), each poly_step forms c r + A h, and the last
// step writes x + h.  A zero guess skips the first product (r = b).  `mask`
// (nullable) restricts all kernels to active columns.

// r = b - A x, then h = c0 r (xo == nullptr) or xo = x + c0 r
template <typename T, int KP, bool WIDE>
__device__ __forceinline__ void poly_first_body(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                                T* __restrict__ r, T* __restrict__ h, T* __restrict__ xo, T c0,
                                                int k, int subl, const int* __restrict__ mask, const int* done,
                                                const VB& vb) {
    if (stopped(done)) return;
    const Lane<KP> ln(k, vb);
    const bool live = ln.cok && (mask == nullptr || mask[ln.c] != 0);
    for_units(nunits_l(A.nrows, KP, WIDE, subl), vb, [&](int unit) {
        unit_spmv<T, KP, WIDE>(
            A, unit, subl, live, [&](int j) { return x[(size_t)j * k + ln.c]; },
            [&](int row, T ax) {
                if (row < A.nrows && live) {
                    const size_t o = (size_t)row * k + ln.c;
                    const T rv = b[o] - ax;
                    if (xo != nullptr) {
                        xo[o] = x[o] + c0 * rv;
                    } else {
                        r[o] = rv;
                        h[o] = c0 * rv;
                    }
                }
            });
    });
}
template <typename T, int KP, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_poly_first(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                                      T* __restrict__ r, T* __restrict__ h, T* __restrict__ xo, T c0,
                                                      int k, int subl, const int* __restrict__ mask, const int* done) {
    poly_first_body<T, KP, WIDE>(A, x, b, r, h, xo, c0, k, subl, mask, done, hw_vb());
}

// zero guess (r = b): h = c0 b, or xo = c0 b with a single coefficient
template <typename T, int KP>
__device__ __forceinline__ void poly_first_zero_body(const T* __restrict__ b, T* __restrict__ out, T c0, int n, int k,
                                                     const int* __restrict__ mask, const int* done, const VB& vb) {
    if (stopped(done)) return;
    const int unit = row_unit(vb);
    if (unit >= nunits(n, KP, false)) return;
    const Lane<KP> ln(k, vb);
    const bool live = ln.cok && (mask == nullptr || mask[ln.c] != 0);
    unit_rows<KP>(unit, [&](int row) {
        if (row < n && live) {
            const size_t o = (size_t)row * k + ln.c;
            out[o] = c0 * b[o];
        }
    });
}
template <typename T, int KP>
__global__ __launch_bounds__(BLOCK) void k_poly_first_zero(const T* __restrict__ b, T* __restrict__ out, T c0, int n,
                                                           int k, const int* __restrict__ mask, const int* done) {
    poly_first_zero_body<T, KP>(b, out, c0, n, k, mask, done, hw_vb());
}

// v = c r + A h; h_out = v, or (last) xo = x + v (x == nullptr: zero guess, xo = v)
template <typename T, int KP, bool WIDE>
__device__ __forceinline__ void poly_step_body(Sell<T> A, const T* __restrict__ h, const T* __restrict__ r,
                                               const T* __restrict__ x, T* __restrict__ out, T c, int last, int k,
                                               int subl, const int* __restrict__ mask, const int* done,
                                               const VB& vb) {
    if (stopped(done)) return;
    const Lane<KP> ln(k, vb);
    const bool live = ln.cok && (mask == nullptr || mask[ln.c] != 0);
    for_units(nunits_l(A.nrows, KP, WIDE, subl), vb, [&](int unit) {
        unit_spmv<T, KP, WIDE>(
            A, unit, subl, live, [&](int j) { return h[(size_t)j * k + ln.c]; },
            [&](int row, T ah) {
                if (row < A.nrows && live) {
                    const size_t o = (size_t)row * k + ln.c;
                    const T v = c * r[o] + ah;
                    out[o] = (last && x != nullptr) ? x[o] + v : v;
                }
            });
    });
}
template <typename T, int KP, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_poly_step(Sell<T> A, const T* __restrict__ h, const T* __restrict__ r,
                                                     const T* __restrict__ x, T* __restrict__ out, T c, int last,
                                                     int k, int subl, const int* __restrict__ mask, const int* done) {
    poly_step_body<T, KP, WIDE>(A, h, r, x, out, c, last, k, subl, mask, done, hw_vb());
}

// ------------------------------------------------------------ adjoint helpers

// g[q] = scale * sum_c lam[row_q, c] x[col_q, c]  over the nnz entries of a pattern
// (the pattern restriction of -outer(lam, x)); one thread per entry, c in order
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_grad_values(const int* __restrict__ rows, const int* __restrict__ cols,
                                                       long long nnz, const T* __restrict__ lam,
                                                       const T* __restrict__ x, int k, T scale, T* __restrict__ g) {
    const long long q = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (q >= nnz) return;
    const size_t ro = (size_t)rows[q] * k, co = (size_t)cols[q] * k;
    T acc = T(0);
    for (int c = 0; c < k; ++c) acc += lam[ro + c] * x[co + c];
    g[q] = scale * acc;
}

// dst[dest[q]] = src[q]: values in CSR order into their sliced-ELL slots
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_scatter_values(const int* __restrict__ dest, long long nnz,
                                                          const T* __restrict__ src, T* __restrict__ dst) {
    const long long q = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (q < nnz) dst[dest[q]] = src[q];
}

// ------------------------------------------------------------ coarse-grid solvers

// Dense coarse solves.  TF is the factor type (float64 factors with a float32
// cycle keep the coarse solve in float64); sums run in TF, x is rounded once.

// x = A_inv b (dense, row-major n x n)
template <typename T, typename TF>
__device__ __forceinline__ void coarse_pinv_body(const TF* __restrict__ Ainv, const T* __restrict__ b,
                                                       T* __restrict__ x, int n, int k, const int* done, const VB& vb) {
    if (stopped(done)) return;
    const int t = vb.x * BLOCK + threadIdx.x;
    if (t >= n * k) return;
    const int i = t / k, c = t % k;
    TF acc = TF(0);
    for (int j = 0; j < n; ++j) acc += Ainv[(size_t)i * n + j] * static_cast<TF>(b[(size_t)j * k + c]);
    x[(size_t)i * k + c] = static_cast<T>(acc);
}
template <typename T, typename TF>
__global__ __launch_bounds__(BLOCK) void k_coarse_pinv(const TF* __restrict__ Ainv, const T* __restrict__ b,
                                                       T* __restrict__ x, int n, int k, const int* done) {
    coarse_pinv_body<T, TF>(Ainv, b, x, n, k, done, hw_vb());
}

// LAPACK-style LU solve with 0-based row interchanges, one thread per column;
// the work vector lives in y (n x k, TF)
template <typename T, typename TF>
__device__ __forceinline__ void coarse_lu_body(const TF* __restrict__ LU, const int* __restrict__ piv,
                                                     const T* __restrict__ b, T* __restrict__ x, TF* __restrict__ y,
                                                     int n, int k, const int* done, const VB& vb) {
    if (stopped(done)) return;
    const int c = vb.x * BLOCK + threadIdx.x;
    if (c >= k) return;
    for (int i = 0; i < n; ++i) y[(size_t)i * k + c] = static_cast<TF>(b[(size_t)i * k + c]);
    for (int i = 0; i < n; ++i) {
        const int pi = piv[i];
        if (pi != i) {
            const TF t = y[(size_t)i * k + c];
            y[(size_t)i * k + c] = y[(size_t)pi * k + c];
            y[(size_t)pi * k + c] = t;
        }
    }
    for (int i = 0; i < n; ++i) {
        TF acc = y[(size_t)i * k + c];
        for (int j = 0; j < i; ++j) acc -= LU[(size_t)i * n + j] * y[(size_t)j * k + c];
        y[(size_t)i * k + c] = acc;
    }
    for (int i = n - 1; i >= 0; --i) {
        TF acc = y[(size_t)i * k + c];
        for (int j = i + 1; j < n; ++j) acc -= LU[(size_t)i * n + j] * y[(size_t)j * k + c];
        y[(size_t)i * k + c] = acc / LU[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) x[(size_t)i * k + c] = static_cast<T>(y[(size_t)i * k + c]);
}
template <typename T, typename TF>
__global__ __launch_bounds__(BLOCK) void k_coarse_lu(const TF* __restrict__ LU, const int* __restrict__ piv,
                                                     const T* __restrict__ b, T* __restrict__ x, TF* __restrict__ y,
                                                     int n, int k, const int* done) {
    coarse_lu_body<T, TF>(LU, piv, b, x, y, n, k, done, hw_vb());
}

// x = R^-1 (Q^T b), one thread per column; work vector in y (n x k, TF)
template <typename T, typename TF>
__device__ __forceinline__ void coarse_qr_body(const TF* __restrict__ Q, const TF* __restrict__ Rm,
                                                     const T* __restrict__ b, T* __restrict__ x, TF* __restrict__ y,
                                                     int n, int k, const int* done, const VB& vb) {
    if (stopped(done)) return;
    const int c = vb.x * BLOCK + threadIdx.x;
    if (c >= k) return;
    for (int i = 0; i < n; ++i) {
        TF acc = TF(0);
        for (int j = 0; j < n; ++j) acc += Q[(size_t)j * n + i] * static_cast<TF>(b[(size_t)j * k + c]);
        y[(size_t)i * k + c] = acc;
    }
    for (int i = n - 1; i >= 0; --i) {
        TF acc = y[(size_t)i * k + c];
        for (int j = i + 1; j < n; ++j) acc -= Rm[(size_t)i * n + j] * y[(size_t)j * k + c];
        y[(size_t)i * k + c] = acc / Rm[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) x[(size_t)i * k + c] = static_cast<T>(y[(size_t)i * k + c]);
}
template <typename T, typename TF>
__global__ __launch_bounds__(BLOCK) void k_coarse_qr(const TF* __restrict__ Q, const TF* __restrict__ Rm,
                                                     const T* __restrict__ b, T* __restrict__ x, TF* __restrict__ y,
                                                     int n, int k, const int* done) {
    coarse_qr_body<T, TF>(Q, Rm, b, x, y, n, k, done, hw_vb());
}

// ------------------------------------------------------------ deterministic reductions
//
// A reducing kernel walks short units on a capped grid.  Each
// block reduces its NV sums per column in a fixed order (wave shuffles, then
// LDS) and stores them with plain stores into ctl.part.  A following
// k_finalize launch (one block per column) adds the per-block sums in block
// order and applies the solver's per-column update; the kernel boundary makes
// the partial sums visible, so no device-scope fences are needed.  No
// floating-point atomics: results are bit-reproducible run to run.

// a reducing kernel's wavefronts walk the short units (the register-slice
// unit size, also for wide matrices) as the persistent row kernels do,
// accumulating their sums across units
template <typename F>
__device__ __forceinline__ void red_unit(int nu, const VB& vb, F&& f) {
    for_units(nu, vb, f);
}

template <int KP, int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], int k, const Ctl& ctl, const VB& vb) {
    constexpr int NW = BLOCK / 64;
    __shared__ double sh[NW * 64 * NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
#pragma unroll
        for (int off = 32; off >= KP; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
    }
    if (lane < KP) {
#pragma unroll
        for (int q = 0; q < NV; ++q) sh[wave * 64 * NV + q * 64 + lane] = v[q];
    }
    __syncthreads();
    if (threadIdx.x < KP) {
        const int c = vb.y * KP + threadIdx.x;
        if (c < k) {
            const int nb = vb.gx;
            const int lb = xcd_block(vb.x, nb);
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                double s = 0.0;
                for (int w = 0; w < NW; ++w) s += sh[w * 64 * NV + q * 64 + threadIdx.x];
                ctl.part[((size_t)q * k + c) * nb + lb] = s;
            }
        }
    }
}

// block_partials for vector lanes: a lane holds the sums of VW consecutive
// columns; lanes with the same columns are combined by an xor tree, then the
// block's waves in LDS.
template <int KP, int VW, int NV>
__device__ __forceinline__ void block_partials_v(Vec<double, VW> (&v)[NV], int k, const Ctl& ctl, const VB& vb) {
    constexpr int NW = BLOCK / 64, LPR = KP / VW;
    __shared__ double sh[NW * 64 * NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
#pragma unroll
        for (int e = 0; e < VW; ++e) {
#pragma unroll
            for (int off = 32; off >= LPR; off >>= 1) v[q].v[e] += __shfl_xor(v[q].v[e], off, 64);
        }
    }
    if (lane < LPR) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
#pragma unroll
            for (int e = 0; e < VW; ++e) sh[wave * 64 * NV + q * 64 + lane * VW + e] = v[q].v[e];
        }
    }
    __syncthreads();
    if (threadIdx.x < KP) {
        const int c = vb.y * KP + threadIdx.x;
        if (c < k) {
            const int nb = vb.gx;
            const int lb = xcd_block(vb.x, nb);
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                double s = 0.0;
                for (int w = 0; w < NW; ++w) s += sh[w * 64 * NV + q * 64 + threadIdx.x];
                ctl.part[((size_t)q * k + c) * nb + lb] = s;
            }
        }
    }
}

enum FinOp {
    FIN_SOLVE_INIT,  // s0 = b.b: normb (0 -> 1), normr, it = 0, active = 0 < maxiter && normr >= tol normb
    FIN_RESNORM,     // s0 = r.r: normr; active columns: it += 1, active = it < maxiter && normr >= tol normb
    FIN_PCG_INIT,    // s0 = b.b: atol2 = max(tol^2 b.b, atol^2[, omega with flag]), normb, it = 0, active = 1
    FIN_PCG_START,   // s0 = r.z, s1 = r.r: gamma, beta = 0, active = (precond ? s1 : s0) > atol2 && 0 < maxiter
    FIN_ALPHA,       // s0 = p.Ap: alpha = active ? gamma / s0 : 0
    FIN_UPDATE,      // s0 = r.r: normr; active columns: it += 1, active = s0 > atol2 && it < maxiter
    FIN_GAMMA,       // s0 = r.z, s1 = z.Ap: active columns: beta (HS or flexible PR), gamma = s0
    FIN_TRUE,        // s0 = r.r: normr
    // BiCGSTAB (jax.scipy.sparse.linalg.bicgstab semantics; gamma holds rho_ = rhat.r)
    FIN_BICG_INIT,   // s0 = b.b: atol2, normb, it = 0, rho_ = s0, alpha = omega = 1, beta = s0,
                     //   active = s0 > atol2 && 0 < maxiter
    FIN_BICG_ALPHA,  // s0 = rhat.q: active columns: alpha = rho_ / s0
    FIN_BICG_EXIT,   // s0 = s.s: active columns: exit = s0 < atol2
    FIN_BICG_OMEGA,  // s0 = t.s, s1 = t.t: active columns: omega = s0 / s1
    FIN_BICG_RHO,    // s0 = rhat.r, s1 = r.r: normr; active columns: it += 1, breakdown if omega, alpha
                     //   or rho_ is 0, active = !breakdown && s1 > atol2 && it < maxiter,
                     //   beta = s0 / rho_ * alpha / omega, rho_ = s0
    // CG with reliable updates (omega holds r.r at the last update; exitf marks a column
    // whose update is pending and ruflag that some column's is: both set by DECIDE and
    // cleared by ALPHA of the next iteration)
    FIN_RU_START,    // s0 = r.z, s1 = r.r: gamma = s0, beta = 0, omega = s1,
                     //   active = s1 > atol2 && 0 < maxiter
    FIN_RU_ALPHA,    // s0 = p.Ap: alpha = active ? gamma / s0 : 0; no update pending
    FIN_RU_DECIDE,   // s0 = r.r of the recurrence: normr; active columns: it += 1, an update is
                     //   pending when s0 <= delta^2 omega, s0 <= atol2 or it >= maxiter
    FIN_RU_GAMMA     // s0 = r.z, s1 = z.Ap, s2 = r.r of the true residual (read only while an
                     //   update is pending): columns with a pending update: normr, omega = s2,
                     //   count += 1, active = s2 > atol2 && it < maxiter; then as FIN_GAMMA on
                     //   the active columns
};

// Completion count of a batch of solves: the finaliser that finds a problem
// newly done adds one to `count`; the one that completes the batch sets `done`
// and its host-mapped mirror.  A single solve passes count = nullptr.
struct BatchFlag {
    int* count;
    int* done;
    int* hdone;
    int n;
};

// The finaliser runs FIN_BLOCK threads per column; each thread adds its
// strided share of the per-block sums (up to 16 loads in flight), then a fixed
// xor tree inside each wavefront and the 16 wavefront sums in order.
constexpr int FIN_BLOCK = 1024;

// One block of FIN_BLOCK threads per column (block x of a k-wide grid): add
// the nb per-block sums (fixed order: strided per thread, then a tree), then
// apply `op`.  Ops that change `active` also recompute `done` (last block,
// integer counters).
__device__ __forceinline__ void finalize_body(const Ctl& ctl, int op, int nv, int nb, int k, int flag,
                                              int honor_done, const BatchFlag& bf, const VB& vb) {
    if (honor_done && stopped(ctl.done)) return;
    __shared__ double sh[FIN_BLOCK / 64];
    __shared__ double tot[3];
    const int c = vb.x;
    // without a pending update FIN_RU_GAMMA is FIN_GAMMA (the third sums are stale)
    const bool ru_pending = op == FIN_RU_GAMMA && *ctl.ruflag != 0;
    if (op == FIN_RU_GAMMA && !ru_pending) {
        op = FIN_GAMMA;
        nv = 2;
    }
    for (int q = 0; q < nv; ++q) {
        const double* src = ctl.part + ((size_t)q * k + c) * nb;
        double acc = 0.0;
#pragma unroll 16
        for (int b = threadIdx.x; b < nb; b += FIN_BLOCK) acc += src[b];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = sh[0];
#pragma unroll
            for (int wv = 1; wv < FIN_BLOCK / 64; ++wv) t += sh[wv];
            tot[q] = t;
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double s0 = tot[0], s1 = nv > 1 ? tot[1] : 0.0, s2 = nv > 2 ? tot[2] : 0.0;
    const double tol = ctl.tol[0], atol = ctl.tol[1];
    const int maxit = *ctl.maxiter;
    bool changes_active = true;
    switch (op) {
        case FIN_SOLVE_INIT: {
            const double nb2 = sqrt(s0);
            ctl.normb[c] = nb2 == 0.0 ? 1.0 : nb2;
            ctl.normr[c] = nb2;
            ctl.it[c] = 0;
            ctl.active[c] = (0 < maxit) && (nb2 >= tol * ctl.normb[c]);
            break;
        }
        case FIN_RESNORM: {
            const double nr = sqrt(s0);
            ctl.normr[c] = nr;
            if (ctl.active[c]) {
                const int it = ctl.it[c] + 1;
                ctl.it[c] = it;
                ctl.active[c] = (it < maxit) && (nr >= tol * ctl.normb[c]);
            }
            break;
        }
        case FIN_PCG_INIT:
            // flag: a per-column floor on the squared stopping norm, held in omega
            ctl.atol2[c] = fmax(fmax(tol * tol * s0, atol * atol), flag ? ctl.omega[c] : 0.0);
            ctl.normb[c] = s0 == 0.0 ? 1.0 : sqrt(s0);
            ctl.normr[c] = sqrt(s0);  // the residual of x = 0
            ctl.it[c] = 0;
            ctl.active[c] = 1;
            break;
        case FIN_PCG_START:
            ctl.gamma[c] = s0;
            ctl.beta[c] = 0.0;
            ctl.active[c] = ((flag ? s1 : s0) > ctl.atol2[c]) && (0 < maxit);
            break;
        case FIN_ALPHA:
            ctl.alpha[c] = ctl.active[c] ? ctl.gamma[c] / s0 : 0.0;
            changes_active = false;
            break;
        case FIN_UPDATE:
            ctl.normr[c] = sqrt(s0);
            if (ctl.active[c]) {
                const int it = ctl.it[c] + 1;
                ctl.it[c] = it;
                ctl.active[c] = (s0 > ctl.atol2[c]) && (it < maxit);
            }
            break;
        case FIN_GAMMA:
            if (ctl.active[c]) {
                ctl.beta[c] = flag ? -ctl.alpha[c] * s1 / ctl.gamma[c] : s0 / ctl.gamma[c];
                ctl.gamma[c] = s0;
            }
            changes_active = false;
            break;
        case FIN_BICG_INIT:
            ctl.atol2[c] = fmax(tol * tol * s0, atol * atol);
            ctl.normb[c] = s0 == 0.0 ? 1.0 : sqrt(s0);
            ctl.normr[c] = sqrt(s0);
            ctl.it[c] = 0;
            ctl.gamma[c] = s0;
            ctl.alpha[c] = 1.0;
            ctl.omega[c] = 1.0;
            ctl.beta[c] = s0;
            ctl.exitf[c] = 0;
            ctl.active[c] = (s0 > ctl.atol2[c]) && (0 < maxit);
            break;
        case FIN_BICG_ALPHA:
            if (ctl.active[c]) ctl.alpha[c] = ctl.gamma[c] / s0;
            changes_active = false;
            break;
        case FIN_BICG_EXIT:
            if (ctl.active[c]) ctl.exitf[c] = s0 < ctl.atol2[c] ? 1 : 0;
            changes_active = false;
            break;
        case FIN_BICG_OMEGA:
            if (ctl.active[c]) ctl.omega[c] = s0 / s1;
            changes_active = false;
            break;
        case FIN_BICG_RHO:
            ctl.normr[c] = sqrt(s1);
            if (ctl.active[c]) {
                const int it = ctl.it[c] + 1;
                ctl.it[c] = it;
                const double a = ctl.alpha[c], w = ctl.omega[c], rho = ctl.gamma[c];
                const bool breakdown = (w == 0.0) || (a == 0.0) || (rho == 0.0);
                ctl.active[c] = !breakdown && (s1 > ctl.atol2[c]) && (it < maxit);
                ctl.beta[c] = s0 / rho * a / w;
                ctl.gamma[c] = s0;
            }
            break;
        case FIN_RU_START:
            ctl.gamma[c] = s0;
            ctl.beta[c] = 0.0;
            ctl.omega[c] = s1;
            ctl.active[c] = (s1 > ctl.atol2[c]) && (0 < maxit);
            break;
        case FIN_RU_ALPHA:
            ctl.alpha[c] = ctl.active[c] ? ctl.gamma[c] / s0 : 0.0;
            ctl.exitf[c] = 0;
            *ctl.ruflag = 0;  // every block stores 0; read after the kernel boundary
            changes_active = false;
            break;
        case FIN_RU_DECIDE:
            ctl.normr[c] = sqrt(s0);
            if (ctl.active[c]) {
                const int it = ctl.it[c] + 1;
                ctl.it[c] = it;
                const double d = ctl.tol[2];
                const bool due = (s0 <= d * d * ctl.omega[c]) || (s0 <= ctl.atol2[c]) || (it >= maxit);
                ctl.exitf[c] = due ? 1 : 0;
                if (due) *ctl.ruflag = 1;  // every writer stores 1; read after the kernel boundary
            }
            changes_active = false;
            break;
        case FIN_RU_GAMMA:
            if (ctl.exitf[c]) {
                ctl.normr[c] = sqrt(s2);
                ctl.omega[c] = s2;
                ctl.count[c] += 1;
                ctl.active[c] = (s2 > ctl.atol2[c]) && (ctl.it[c] < maxit);
            }
            if (ctl.active[c]) {
                ctl.beta[c] = flag ? -ctl.alpha[c] * s1 / ctl.gamma[c] : s0 / ctl.gamma[c];
                ctl.gamma[c] = s0;
            }
            break;
        default:
            ctl.normr[c] = sqrt(s0);
            changes_active = false;
            break;
    }
    if (!changes_active) return;
    // done = no active column: integer counters, the last finaliser block decides
    __hip_atomic_fetch_add(&ctl.ticket[1], ctl.active[c] ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int t = __hip_atomic_fetch_add(&ctl.ticket[0], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (t == vb.gx - 1) {
        const int na = __hip_atomic_load(&ctl.ticket[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int was = *ctl.done;
        *ctl.done = na == 0 ? 1 : 0;
        if (ctl.hdone != nullptr)
            __hip_atomic_store(ctl.hdone, na == 0 ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&ctl.ticket[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl.ticket[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (bf.count != nullptr && na == 0 && was == 0) {
            const int d = __hip_atomic_fetch_add(bf.count, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (d == bf.n - 1) {
                *bf.done = 1;
                __hip_atomic_store(bf.hdone, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
}
static __global__ __launch_bounds__(FIN_BLOCK) void k_finalize(Ctl ctl, int op, int nv, int nb, int k, int flag,
                                                         int honor_done) {
    finalize_body(ctl, op, nv, nb, k, flag, honor_done, BatchFlag{nullptr, nullptr, nullptr, 0}, hw_vb());
}

// solver parameters (written by a kernel so that no host buffer has to outlive the call)
__device__ __forceinline__ void set_params_body(const Ctl& ctl, double tol, double atol, int maxiter, int done) {
    if (threadIdx.x != 0) return;
    ctl.tol[0] = tol;
    ctl.tol[1] = atol;
    *ctl.maxiter = maxiter;
    *ctl.done = done;
    if (ctl.hdone != nullptr) __hip_atomic_store(ctl.hdone, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// done = 0 (and its host mirror), the solver parameters left as they are
static __global__ void k_reset_done(Ctl ctl) {
    if (threadIdx.x != 0) return;
    *ctl.done = 0;
    if (ctl.hdone != nullptr) __hip_atomic_store(ctl.hdone, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// refinement: the inner CG of column c may stop once its residual norm is
// below half the outer target, inner.omega[c] = (tol normb[c] / 2)^2
static __global__ void k_ir_floor(Ctl inner, Ctl outer, double tol, int k) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < k) {
        const double f = 0.5 * tol * outer.normb[c];
        inner.omega[c] = f * f;
    }
}

// total[c] += it[c] on the columns still active (refinement: inner CG
// iterations summed over the steps that ran them)
static __global__ void k_add_iters(int* __restrict__ total, const int* __restrict__ it,
                                   const int* __restrict__ active, int k) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < k && active[c]) total[c] += it[c];
}
static __global__ void k_set_params(Ctl ctl, double tol, double atol, int maxiter, int done) {
    set_params_body(ctl, tol, atol, maxiter, done);
}

// CG with reliable updates: threshold delta, no update pending, counters at 0
static __global__ void k_ru_params(Ctl ctl, double delta, int k) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) {
        ctl.tol[2] = delta;
        *ctl.ruflag = 0;
    }
    if (c < k) {
        ctl.count[c] = 0;
        ctl.exitf[c] = 0;
    }
}

// The CG vector updates of the reliable-update CG with 16 bytes per lane, for
// k = KP in {16, 32, 64} (a multiple of the vector width VW): the (n, k) block
// is walked as one flat array of VW-column vectors with a stride of the grid, so
// a lane keeps the same VW columns through the launch and the LPR = KP / VW
// lanes of a row read it contiguously.  Same values as pcg_p_body /
// pcg_update_body bit for bit (the products are formed per element in float64);
// the r.r partials are grouped per lane, as block_partials_v expects.
//
// The p update also completes the previous iteration's reliable update on the
// columns that had one (exitf, still set: FIN_RU_ALPHA clears it after this
// kernel): x += (double) xh in float64 and xh = 0.  The true residual of that
// update was formed from the same sum x + (double) xh, so deferring the fold
// to here changes no value.
template <typename T, int KP>
__device__ __forceinline__ void ru_p_vbody(const T* __restrict__ z, T* __restrict__ p, double* __restrict__ x,
                                           T* __restrict__ xh, int n, Ctl ctl, const VB& vb) {
    if (stopped(ctl.done)) return;
    constexpr int VW = vec_width<T>(), LPR = KP / VW;
    const int c0 = (threadIdx.x % LPR) * VW;
    double beta[VW];
    bool live[VW], fold[VW];
    bool any = false;
    const bool pending = *ctl.ruflag != 0;
#pragma unroll
    for (int e = 0; e < VW; ++e) {
        live[e] = ctl.active[c0 + e] != 0;
        beta[e] = live[e] ? ctl.beta[c0 + e] : 0.0;
        fold[e] = pending && ctl.exitf[c0 + e] != 0;
        any = any || fold[e];
    }
    const size_t nv = (size_t)n * LPR, stride = (size_t)vb.gx * BLOCK;
    for (size_t i = (size_t)vb.x * BLOCK + threadIdx.x; i < nv; i += stride) {
        const Vec<T, VW> zv = vload<T, VW>(z + i * VW);
        Vec<T, VW> pv = vload<T, VW>(p + i * VW);
#pragma unroll
        for (int e = 0; e < VW; ++e)
            if (live[e]) pv.v[e] = static_cast<T>(static_cast<double>(zv.v[e]) + beta[e] * static_cast<double>(pv.v[e]));
        vstore<T, VW>(p + i * VW, pv);
        if (any) {
            // k = KP: the flat vector i holds elements i VW .. i VW + VW - 1
            Vec<double, VW> xv = vload<double, VW>(x + i * VW);
            Vec<T, VW> ev = vload<T, VW>(xh + i * VW);
#pragma unroll
            for (int e = 0; e < VW; ++e)
                if (fold[e]) {
                    xv.v[e] = xv.v[e] + static_cast<double>(ev.v[e]);
                    ev.v[e] = T(0);
                }
            vstore<double, VW>(x + i * VW, xv);
            vstore<T, VW>(xh + i * VW, ev);
        }
    }
}
template <typename T, int KP>
__global__ __launch_bounds__(BLOCK) void k_ru_p_v(const T* __restrict__ z, T* __restrict__ p, double* __restrict__ x,
                                                  T* __restrict__ xh, int n, Ctl ctl) {
    ru_p_vbody<T, KP>(z, p, x, xh, n, ctl, hw_vb());
}

// p update of the reliable-update CG, one column per lane (any k): as
// pcg_p_body, plus the deferred fold of ru_p_vbody
template <typename T, int KP>
__device__ __forceinline__ void ru_p_body(const T* __restrict__ z, T* __restrict__ p, double* __restrict__ x,
                                          T* __restrict__ xh, int n, int k, Ctl ctl, const VB& vb) {
    if (stopped(ctl.done)) return;
    const Lane<KP> ln(k, vb);
    const bool live = ln.cok && ctl.active[ln.c] != 0;
    const bool fold = ln.cok && *ctl.ruflag != 0 && ctl.exitf[ln.c] != 0;
    const double beta = live ? ctl.beta[ln.c] : 0.0;
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double zv[NP], pv[NP];
        red_rows<KP>(
            u, n, live || fold,
            [&](int q, int row, bool ok) {
                const size_t o = (size_t)row * k + ln.c;
                zv[q] = ok && live ? static_cast<double>(z[o]) : 0.0;
                pv[q] = ok && live ? static_cast<double>(p[o]) : 0.0;
            },
            [&](int q, int row) {
                const size_t o = (size_t)row * k + ln.c;
                if (live) p[o] = static_cast<T>(zv[q] + beta * pv[q]);
                if (fold) {
                    x[o] = x[o] + static_cast<double>(xh[o]);
                    xh[o] = T(0);
                }
            });
    });
}
template <typename T, int KP>
__global__ __launch_bounds__(BLOCK) void k_ru_p(const T* __restrict__ z, T* __restrict__ p, double* __restrict__ x,
                                                T* __restrict__ xh, int n, int k, Ctl ctl) {
    ru_p_body<T, KP>(z, p, x, xh, n, k, ctl, hw_vb());
}

// End of the reliable-update CG: x_out = x + (double) xh on every column (the
// last update's fold, and xh = 0 on the columns folded before).
template <typename TE, int KP>
__global__ __launch_bounds__(BLOCK) void k_ru_out(const double* __restrict__ x, const TE* __restrict__ xh,
                                                  double* __restrict__ out, int n, int k) {
    const VB vb = hw_vb();
    const Lane<KP> ln(k, vb);
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double xv[NP], ev[NP];
        red_rows<KP>(
            u, n, ln.cok,
            [&](int q, int row, bool ok) {
                const size_t o = (size_t)row * k + ln.c;
                xv[q] = ok ? x[o] : 0.0;
                ev[q] = ok ? static_cast<double>(xh[o]) : 0.0;
            },
            [&](int q, int row) { out[(size_t)row * k + ln.c] = xv[q] + ev[q]; });
    });
}

template <typename T, int KP>
__device__ __forceinline__ void ru_update_vbody(const T* __restrict__ p, const T* __restrict__ Ap,
                                                T* __restrict__ x, T* __restrict__ r, int n, int k, Ctl ctl,
                                                const VB& vb) {
    if (stopped(ctl.done)) return;
    constexpr int VW = vec_width<T>(), LPR = KP / VW;
    const int c0 = (threadIdx.x % LPR) * VW;
    double a[VW];
    bool act[VW];
#pragma unroll
    for (int e = 0; e < VW; ++e) {
        act[e] = ctl.active[c0 + e] != 0;
        a[e] = act[e] ? ctl.alpha[c0 + e] : 0.0;
    }
    Vec<double, VW> v[1] = {vzero<double, VW>()};
    const size_t nv = (size_t)n * LPR, stride = (size_t)vb.gx * BLOCK;
    for (size_t i = (size_t)vb.x * BLOCK + threadIdx.x; i < nv; i += stride) {
        const Vec<T, VW> pv = vload<T, VW>(p + i * VW), av = vload<T, VW>(Ap + i * VW);
        Vec<T, VW> xv = vload<T, VW>(x + i * VW), rv = vload<T, VW>(r + i * VW);
#pragma unroll
        for (int e = 0; e < VW; ++e) {
            double rr = static_cast<double>(rv.v[e]);
            if (act[e]) {
                xv.v[e] = static_cast<T>(static_cast<double>(xv.v[e]) + a[e] * static_cast<double>(pv.v[e]));
                rr = static_cast<double>(static_cast<T>(rr - a[e] * static_cast<double>(av.v[e])));
                rv.v[e] = static_cast<T>(rr);
            }
            v[0].v[e] += rr * rr;
        }
        vstore<T, VW>(x + i * VW, xv);
        vstore<T, VW>(r + i * VW, rv);
    }
    block_partials_v<KP, VW, 1>(v, k, ctl, vb);
}
template <typename T, int KP>
__global__ __launch_bounds__(BLOCK) void k_ru_update_v(const T* __restrict__ p, const T* __restrict__ Ap,
                                                       T* __restrict__ x, T* __restrict__ r, int n, int k, Ctl ctl) {
    ru_update_vbody<T, KP>(p, Ap, x, r, n, k, ctl, hw_vb());
}

// solve start: x = 0, normb = ||b|| (0 -> 1), normr = ||b||, it = 0,
// active = 0 < maxiter && normr >= tol normb
template <typename T, int KP>
__device__ __forceinline__ void solve_init_body(const T* __restrict__ b, T* __restrict__ x, int n, int k,
                                                       Ctl ctl, const VB& vb) {
    const Lane<KP> ln(k, vb);
    double v[1] = {0.0};
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double bv[NP];
        red_rows<KP>(
            u, n, ln.cok,
            [&](int q, int row, bool ok) { bv[q] = ok ? static_cast<double>(b[(size_t)row * k + ln.c]) : 0.0; },
            [&](int q, int row) {
                v[0] += bv[q] * bv[q];
                if (x != nullptr) x[(size_t)row * k + ln.c] = T(0);
            });
    });
    block_partials<KP, 1>(v, k, ctl, vb);
}
template <typename T, int KP>
__global__ __launch_bounds__(BLOCK) void k_solve_init(const T* __restrict__ b, T* __restrict__ x, int n, int k,
                                                       Ctl ctl) {
    solve_init_body<T, KP>(b, x, n, k, ctl, hw_vb());
}

// after a cycle: r = b - A x (optionally stored as TR), normr = ||r||,
// it += active, active = it < maxiter && normr >= tol normb
template <typename T, typename TR, int KP, bool WIDE>
__device__ __forceinline__ void resnorm_body(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                                    TR* __restrict__ r_out, int k, Ctl ctl, const VB& vb) {
    if (stopped(ctl.done)) return;
    const Lane<KP> ln(k, vb);
    const int n = A.nrows;
    double v[1] = {0.0};
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        unit_spmv<T, KP, (KP > 1 && WIDE)>(
            A, u, ilog2(red_sub(KP)), ln.cok, [&](int j) { return x[(size_t)j * k + ln.c]; },
            [&](int row, T ax) {
                if (row < n && ln.cok) {
                    const size_t o = (size_t)row * k + ln.c;
                    const T rv = b[o] - ax;
                    if (r_out != nullptr) r_out[o] = static_cast<TR>(rv);
                    v[0] += static_cast<double>(rv) * static_cast<double>(rv);
                }
            });
    });
    block_partials<KP, 1>(v, k, ctl, vb);
}
template <typename T, typename TR, int KP, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_resnorm(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                                    TR* __restrict__ r_out, int k, Ctl ctl) {
    resnorm_body<T, TR, KP, WIDE>(A, x, b, r_out, k, ctl, hw_vb());
}
// Reliable update of CG with reliable updates.  s holds b - A x in float64 for
// the folded iterate x, xh the correction not yet folded (CG precision).  On
// the columns with a pending update r = s - A xh (equal to b - A (x + xh) in
// exact arithmetic) becomes the new s and, rounded to the CG precision, the new
// recurrence residual; the partials sum r.r.  Only xh is gathered; s is read and
// written row by row by the lane that owns the row.  The kernel returns at once
// when no update is pending.
template <typename T, typename TR, int KP, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_ru_resnorm(Sell<T> A, const TR* __restrict__ xh, T* s,
                                                      TR* __restrict__ r_out, int k, Ctl ctl) {
    if (stopped(ctl.done)) return;
    if (*ctl.ruflag == 0) return;
    const VB vb = hw_vb();
    const Lane<KP> ln(k, vb);
    const int n = A.nrows;
    const bool wr = ln.cok && ctl.exitf[ln.c] != 0;
    double v[1] = {0.0};
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        unit_spmv<T, KP, (KP > 1 && WIDE)>(
            A, u, ilog2(red_sub(KP)), ln.cok, [&](int j) { return static_cast<T>(xh[(size_t)j * k + ln.c]); },
            [&](int row, T ax) {
                if (row < n && ln.cok) {
                    const size_t o = (size_t)row * k + ln.c;
                    const T rv = s[o] - ax;
                    if (wr) {
                        s[o] = rv;
                        r_out[o] = static_cast<TR>(rv);
                    }
                    v[0] += static_cast<double>(rv) * static_cast<double>(rv);
                }
            });
    });
    block_partials<KP, 1>(v, k, ctl, vb);
}

// vector mode of resnorm_body (k a multiple of vec_width<T>())
template <typename T, typename TR, int KP, int WC = 64>
__device__ __forceinline__ void resnorm_vbody(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                              TR* __restrict__ r_out, int k, Ctl ctl, const VB& vb) {
    if (stopped(ctl.done)) return;
    constexpr int VW = vec_width<T>();
    const LaneV<KP, VW> ln(k, vb);
    Vec<double, VW> v[1] = {vzero<double, VW>()};
    struct In {
        Vec<T, VW> b;
    };
    vec_rows<T, KP, 0, 8, 0, WC>(
        A, ln.cok, vb, [&](int j) { return vload<T, VW>(x + (size_t)j * k + ln.c0); },
        [&](int row, bool ok) {
            In in;
            in.b = ok ? vload<T, VW>(b + (size_t)row * k + ln.c0) : vzero<T, VW>();
            return in;
        },
        [&](int row, const Vec<T, VW>& ax, const In& in) {
            Vec<TR, VW> ro;
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                const T rv = in.b.v[e] - ax.v[e];
                ro.v[e] = static_cast<TR>(rv);
                v[0].v[e] += static_cast<double>(rv) * static_cast<double>(rv);
            }
            if (r_out != nullptr) vstore<TR, VW>(r_out + (size_t)row * k + ln.c0, ro);
        });
    block_partials_v<KP, VW, 1>(v, k, ctl, vb);
}
template <typename T, typename TR, int KP, int WC = 64>
__global__ __launch_bounds__(BLOCK) void k_resnorm_v(Sell<T> A, const T* __restrict__ x, const T* __restrict__ b,
                                                     TR* __restrict__ r_out, int k, Ctl ctl) {
    resnorm_vbody<T, TR, KP, WC>(A, x, b, r_out, k, ctl, hw_vb());
}
// reliable update (vector mode of k_ru_resnorm).  The lanes are laid out for
// the CG precision TR (16 bytes of xh per lane, VW = vec_width<TR>() columns),
// so the gathers of xh are full-width; s, the matrix and the sums are float64.
template <typename T, typename TR, int KP, int WC = 64>
__global__ __launch_bounds__(BLOCK) void k_ru_resnorm_v(Sell<T> A, const TR* __restrict__ xh, T* s,
                                                        TR* __restrict__ r_out, int k, Ctl ctl) {
    if (stopped(ctl.done)) return;
    if (*ctl.ruflag == 0) return;
    const VB vb = hw_vb();
    constexpr int VW = vec_width<TR>();
    const LaneV<KP, VW> ln(k, vb);
    // the lane's columns with a pending update (all of them: vector stores)
    bool wr[VW];
    bool all = true;
#pragma unroll
    for (int e = 0; e < VW; ++e) {
        wr[e] = ln.cok && ctl.exitf[ln.c0 + e] != 0;
        all = all && wr[e];
    }
    Vec<double, VW> v[1] = {vzero<double, VW>()};
    struct In {
        Vec<T, VW> s;
    };
    vec_rows<T, KP, 0, 8, VW, WC>(
        A, ln.cok, vb, [&](int j) { return vload<TR, VW>(xh + (size_t)j * k + ln.c0); },
        [&](int row, bool ok) {
            In in;
            in.s = ok ? vload<T, VW>(s + (size_t)row * k + ln.c0) : vzero<T, VW>();
            return in;
        },
        [&](int row, const Vec<T, VW>& ax, const In& in) {
            Vec<T, VW> so;
            Vec<TR, VW> ro;
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                so.v[e] = in.s.v[e] - ax.v[e];
                ro.v[e] = static_cast<TR>(so.v[e]);
                v[0].v[e] += static_cast<double>(so.v[e]) * static_cast<double>(so.v[e]);
            }
            const size_t o = (size_t)row * k + ln.c0;
            if (all) {
                vstore<T, VW>(s + o, so);
                vstore<TR, VW>(r_out + o, ro);
            } else {
#pragma unroll
                for (int e = 0; e < VW; ++e)
                    if (wr[e]) {
                        s[o + e] = so.v[e];
                        r_out[o + e] = ro.v[e];
                    }
            }
        });
    block_partials_v<KP, VW, 1>(v, k, ctl, vb);
}

// ------------------------------------------------------------ preconditioned CG (float64 outer)

// x = 0, r = b, rc = (TC) b, atol2 = max(tol^2 b.b, atol^2), it = 0, active = 1;
// optionally bcopy = b (the solve's own copy of the caller's b) and zero = 0
template <typename TC, int KP>
__device__ __forceinline__ void pcg_init_body(const double* __restrict__ b, double* __restrict__ x,
                                                     double* __restrict__ r, TC* __restrict__ rc, int n, int k,
                                                     Ctl ctl, const VB& vb, double* __restrict__ bcopy = nullptr,
                                                     TC* __restrict__ zero = nullptr) {
    const Lane<KP> ln(k, vb);
    double v[1] = {0.0};
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double bv[NP];
        red_rows<KP>(
            u, n, ln.cok, [&](int q, int row, bool ok) { bv[q] = ok ? b[(size_t)row * k + ln.c] : 0.0; },
            [&](int q, int row) {
                const size_t o = (size_t)row * k + ln.c;
                v[0] += bv[q] * bv[q];
                x[o] = 0.0;
                if (r != nullptr) r[o] = bv[q];
                if (rc != nullptr) rc[o] = static_cast<TC>(bv[q]);
                if (bcopy != nullptr) bcopy[o] = bv[q];
                if (zero != nullptr) zero[o] = TC(0);
            });
    });
    block_partials<KP, 1>(v, k, ctl, vb);
}
template <typename TC, int KP>
__global__ __launch_bounds__(BLOCK) void k_pcg_init(const double* __restrict__ b, double* __restrict__ x,
                                                     double* __restrict__ r, TC* __restrict__ rc, int n, int k,
                                                     Ctl ctl, double* __restrict__ bcopy, TC* __restrict__ zero) {
    pcg_init_body<TC, KP>(b, x, r, rc, n, k, ctl, hw_vb(), bcopy, zero);
}

// gamma = r.z, active = (precond ? r.r : gamma) > atol2 && 0 < maxiter;
// p_prev = 0 and beta = 0, so the first fused update p = z + beta p_prev gives p = z
template <typename TZ, int KP, typename TV = double>
__device__ __forceinline__ void pcg_start_body(const TV* __restrict__ r, const TZ* __restrict__ z,
                                                      TV* __restrict__ p_prev, int n, int k, Ctl ctl,
                                                      int precond, const VB& vb) {
    const Lane<KP> ln(k, vb);
    double v[2] = {0.0, 0.0};
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double rv[NP], zv[NP];
        red_rows<KP>(
            u, n, ln.cok,
            [&](int q, int row, bool ok) {
                const size_t o = (size_t)row * k + ln.c;
                rv[q] = ok ? static_cast<double>(r[o]) : 0.0;
                zv[q] = ok ? static_cast<double>(z[o]) : 0.0;
            },
            [&](int q, int row) {
                v[0] += rv[q] * zv[q];
                v[1] += rv[q] * rv[q];
                p_prev[(size_t)row * k + ln.c] = TV(0);
            });
    });
    block_partials<KP, 2>(v, k, ctl, vb);
}
template <typename TZ, int KP, typename TV = double>
__global__ __launch_bounds__(BLOCK) void k_pcg_start(const TV* __restrict__ r, const TZ* __restrict__ z,
                                                      TV* __restrict__ p_prev, int n, int k, Ctl ctl,
                                                      int precond) {
    pcg_start_body<TZ, KP, TV>(r, z, p_prev, n, k, ctl, precond, hw_vb());
}

// Ap = A p and the partial sums of p.Ap (finaliser: alpha = gamma / p.Ap)
template <int KP, bool WIDE, typename TV = double>
__device__ __forceinline__ void pcg_spmv_body(Sell<TV> A, const TV* __restrict__ p,
                                                     TV* __restrict__ Ap, int k, Ctl ctl, const VB& vb) {
    if (stopped(ctl.done)) return;
    const Lane<KP> ln(k, vb);
    const int n = A.nrows;
    double v[1] = {0.0};
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        unit_spmv<TV, KP, (KP > 1 && WIDE)>(
            A, u, ilog2(red_sub(KP)), ln.cok, [&](int j) { return p[(size_t)j * k + ln.c]; },
            [&](int row, TV ap) {
                if (row < n && ln.cok) {
                    const size_t o = (size_t)row * k + ln.c;
                    Ap[o] = ap;
                    v[0] += static_cast<double>(p[o]) * static_cast<double>(ap);
                }
            });
    });
    block_partials<KP, 1>(v, k, ctl, vb);
}
template <int KP, bool WIDE, typename TV = double>
__global__ __launch_bounds__(BLOCK) void k_pcg_spmv(Sell<TV> A, const TV* __restrict__ p,
                                                     TV* __restrict__ Ap, int k, Ctl ctl) {
    pcg_spmv_body<KP, WIDE, TV>(A, p, Ap, k, ctl, hw_vb());
}

// The last post-smoothing sweep of the preconditioner's top level fused with
// the partial sums of pcg_gamma_body (vector mode): z = sweep of x for the
// right-hand side (T) r, then r.z and, for the flexible beta, z.Ap.  r and Ap
// are the float64 CG vectors; the cycle's own right-hand side is (T) r, so the
// sweep reads r instead of its rounded copy and gives the same z bit for bit.
template <typename T, int KP, typename TV = double, int WC = 64>
__device__ __forceinline__ void jacobi_gamma_vbody(Sell<T> A, const T* __restrict__ dinv, const T* __restrict__ x,
                                                   const TV* __restrict__ r, const TV* __restrict__ Ap,
                                                   T* __restrict__ z, T w, int k, int flexible, Ctl ctl,
                                                   const VB& vb) {
    if (stopped(ctl.done)) return;
    constexpr int VW = vec_width<T>();
    const LaneV<KP, VW> ln(k, vb);
    Vec<double, VW> v[2] = {vzero<double, VW>(), vzero<double, VW>()};
    struct In {
        Vec<T, VW> x;
        Vec<TV, VW> r, a;
        T d;
    };
    // float32 sweeps: one pass per unit (the float64 r and Ap of four columns
    // per lane for a second pass would cost occupancy); float64: two passes
    vec_rows<T, KP, sizeof(T) == 4 ? 64 / (KP / VW) : 0, 8, 0, WC>(
        A, ln.cok, vb, [&](int j) { return vload<T, VW>(x + (size_t)j * k + ln.c0); },
        [&](int row, bool ok) {
            In in;
            const size_t o = (size_t)row * k + ln.c0;
            in.x = ok ? vload<T, VW>(x + o) : vzero<T, VW>();
            in.r = ok ? vload<TV, VW>(r + o) : vzero<TV, VW>();
            in.a = (ok && flexible) ? vload<TV, VW>(Ap + o) : vzero<TV, VW>();
            in.d = ok ? dinv[row] : T(0);
            return in;
        },
        [&](int row, const Vec<T, VW>& ax, const In& in) {
            Vec<T, VW> o;
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                const T bv = static_cast<T>(in.r.v[e]);
                const T temp = in.d * (bv - ax.v[e]) + in.x.v[e];
                o.v[e] = (T(1) - w) * in.x.v[e] + w * temp;
                const double zz = static_cast<double>(o.v[e]);
                v[0].v[e] += static_cast<double>(in.r.v[e]) * zz;
                if (flexible) v[1].v[e] += zz * static_cast<double>(in.a.v[e]);
            }
            vstore<T, VW>(z + (size_t)row * k + ln.c0, o);
        });
    block_partials_v<KP, VW, 2>(v, k, ctl, vb);
}
template <typename T, int KP, typename TV = double, int WC = 64>
__global__ __launch_bounds__(BLOCK) void k_jacobi_gamma_v(Sell<T> A, const T* __restrict__ dinv,
                                                          const T* __restrict__ x, const TV* __restrict__ r,
                                                          const TV* __restrict__ Ap, T* __restrict__ z, T w,
                                                          int k, int flexible, Ctl ctl) {
    jacobi_gamma_vbody<T, KP, TV, WC>(A, dinv, x, r, Ap, z, w, k, flexible, ctl, hw_vb());
}

// vector mode of pcg_spmv_body (k a multiple of vec_width<TV>())
template <int KP, typename TV = double, int WC = 64>
__device__ __forceinline__ void pcg_spmv_vbody(Sell<TV> A, const TV* __restrict__ p, TV* __restrict__ Ap,
                                               int k, Ctl ctl, const VB& vb) {
    if (stopped(ctl.done)) return;
    constexpr int VW = vec_width<TV>();
    const LaneV<KP, VW> ln(k, vb);
    Vec<double, VW> v[1] = {vzero<double, VW>()};
    struct In {
        Vec<TV, VW> p;
    };
    vec_rows<TV, KP, 0, 8, 0, WC>(
        A, ln.cok, vb, [&](int j) { return vload<TV, VW>(p + (size_t)j * k + ln.c0); },
        [&](int row, bool ok) {
            In in;
            in.p = ok ? vload<TV, VW>(p + (size_t)row * k + ln.c0) : vzero<TV, VW>();
            return in;
        },
        [&](int row, const Vec<TV, VW>& ap, const In& in) {
            vstore<TV, VW>(Ap + (size_t)row * k + ln.c0, ap);
#pragma unroll
            for (int e = 0; e < VW; ++e)
                v[0].v[e] += static_cast<double>(in.p.v[e]) * static_cast<double>(ap.v[e]);
        });
    block_partials_v<KP, VW, 1>(v, k, ctl, vb);
}
template <int KP, typename TV = double, int WC = 64>
__global__ __launch_bounds__(BLOCK) void k_pcg_spmv_v(Sell<TV> A, const TV* __restrict__ p,
                                                      TV* __restrict__ Ap, int k, Ctl ctl) {
    pcg_spmv_vbody<KP, TV, WC>(A, p, Ap, k, ctl, hw_vb());
}

// p = z + beta p for active columns (beta = 0 on the first iteration: p = z)
template <typename TZ, int KP, typename TV = double>
__device__ __forceinline__ void pcg_p_body(const TZ* __restrict__ z, TV* __restrict__ p, int n, int k,
                                                 Ctl ctl, const VB& vb) {
    if (stopped(ctl.done)) return;
    const Lane<KP> ln(k, vb);
    const bool live = ln.cok && ctl.active[ln.c] != 0;
    const double beta = live ? ctl.beta[ln.c] : 0.0;
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double zv[NP], pv[NP];
        red_rows<KP>(
            u, n, live,
            [&](int q, int row, bool ok) {
                const size_t o = (size_t)row * k + ln.c;
                zv[q] = ok ? static_cast<double>(z[o]) : 0.0;
                pv[q] = ok ? static_cast<double>(p[o]) : 0.0;
            },
            [&](int q, int row) { p[(size_t)row * k + ln.c] = static_cast<TV>(zv[q] + beta * pv[q]); });
    });
}
template <typename TZ, int KP, typename TV = double>
__global__ __launch_bounds__(BLOCK) void k_pcg_p(const TZ* __restrict__ z, TV* __restrict__ p, int n, int k,
                                                 Ctl ctl) {
    pcg_p_body<TZ, KP, TV>(z, p, n, k, ctl, hw_vb());
}

// active columns: x += alpha p, r -= alpha Ap, rc = (TC) r; then
// it += 1, active = r.r > atol2 && it < maxiter
template <typename TC, int KP, typename TV = double>
__device__ __forceinline__ void pcg_update_body(const TV* __restrict__ p, const TV* __restrict__ Ap,
                                                       TV* __restrict__ x, TV* __restrict__ r,
                                                       TC* __restrict__ rc, int n, int k, Ctl ctl, const VB& vb) {
    if (stopped(ctl.done)) return;
    const Lane<KP> ln(k, vb);
    const bool act = ln.cok && ctl.active[ln.c] != 0;
    const double a = act ? ctl.alpha[ln.c] : 0.0;
    double v[1] = {0.0};
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double rv[NP], xv[NP], pv[NP], av[NP];
        red_rows<KP>(
            u, n, ln.cok,
            [&](int q, int row, bool ok) {
                const size_t o = (size_t)row * k + ln.c;
                rv[q] = ok ? static_cast<double>(r[o]) : 0.0;
                const bool upd = ok && act;
                xv[q] = upd ? static_cast<double>(x[o]) : 0.0;
                pv[q] = upd ? static_cast<double>(p[o]) : 0.0;
                av[q] = upd ? static_cast<double>(Ap[o]) : 0.0;
            },
            [&](int q, int row) {
                const size_t o = (size_t)row * k + ln.c;
                double rr = rv[q];
                if (act) {
                    x[o] = static_cast<TV>(xv[q] + a * pv[q]);
                    rr = static_cast<double>(static_cast<TV>(rr - a * av[q]));
                    r[o] = static_cast<TV>(rr);
                    if (rc != nullptr) rc[o] = static_cast<TC>(rr);
                }
                v[0] += rr * rr;
            });
    });
    block_partials<KP, 1>(v, k, ctl, vb);
}
template <typename TC, int KP, typename TV = double>
__global__ __launch_bounds__(BLOCK) void k_pcg_update(const TV* __restrict__ p, const TV* __restrict__ Ap,
                                                       TV* __restrict__ x, TV* __restrict__ r,
                                                       TC* __restrict__ rc, int n, int k, Ctl ctl) {
    pcg_update_body<TC, KP, TV>(p, Ap, x, r, rc, n, k, ctl, hw_vb());
}

// gamma' = r.z; beta = gamma'/gamma (Hestenes-Stiefel) or -alpha (z.Ap)/gamma
// (Polak-Ribiere, flexible); gamma = gamma' - for active columns
template <typename TZ, int KP, typename TV = double>
__device__ __forceinline__ void pcg_gamma_body(const TV* __restrict__ r, const TZ* __restrict__ z,
                                                      const TV* __restrict__ Ap, int n, int k, Ctl ctl,
                                                      int flexible, const VB& vb) {
    if (stopped(ctl.done)) return;
    const Lane<KP> ln(k, vb);
    double v[2] = {0.0, 0.0};
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double rv[NP], zv[NP], av[NP];
        red_rows<KP>(
            u, n, ln.cok,
            [&](int q, int row, bool ok) {
                const size_t o = (size_t)row * k + ln.c;
                rv[q] = ok ? static_cast<double>(r[o]) : 0.0;
                zv[q] = ok ? static_cast<double>(z[o]) : 0.0;
                av[q] = ok && flexible ? static_cast<double>(Ap[o]) : 0.0;
            },
            [&](int q, int) {
                v[0] += rv[q] * zv[q];
                if (flexible) v[1] += zv[q] * av[q];
            });
    });
    block_partials<KP, 2>(v, k, ctl, vb);
}
template <typename TZ, int KP, typename TV = double>
__global__ __launch_bounds__(BLOCK) void k_pcg_gamma(const TV* __restrict__ r, const TZ* __restrict__ z,
                                                      const TV* __restrict__ Ap, int n, int k, Ctl ctl,
                                                      int flexible) {
    pcg_gamma_body<TZ, KP, TV>(r, z, Ap, n, k, ctl, flexible, hw_vb());
}

// ------------------------------------------------------------ restarted GMRES (float64 vectors)
//
// The Krylov basis V_0..V_m (n x k each, V_q at V + q ldv) is float64.  Dot
// products of several basis vectors with one vector share one pass and one
// deterministic two-step reduction (per-block partials, then per (column,
// vector) sums in block order).
constexpr int GMRES_MMAX = 32;  // restart length limit (basis vectors m + 1 <= 33)

// part[(q k + c) nb + block] = sum over the block's rows of V_q[row, c] w[row, c], q < nv
template <int KP>
__global__ __launch_bounds__(BLOCK) void k_gmres_dots(const double* __restrict__ V, size_t ldv, int nv,
                                                      const double* __restrict__ w, int n, int k,
                                                      double* __restrict__ part) {
    const VB vb = hw_vb();
    const Lane<KP> ln(k, vb);
    double v[GMRES_MMAX + 1];
#pragma unroll
    for (int q = 0; q <= GMRES_MMAX; ++q) v[q] = 0.0;
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double wv[NP];
        red_rows<KP>(
            u, n, ln.cok, [&](int q, int row, bool ok) { wv[q] = ok ? w[(size_t)row * k + ln.c] : 0.0; },
            [&](int q, int row) {
                const size_t o = (size_t)row * k + ln.c;
#pragma unroll
                for (int j = 0; j <= GMRES_MMAX; ++j)
                    if (j < nv) v[j] += V[(size_t)j * ldv + o] * wv[q];
            });
    });
    constexpr int NW = BLOCK / 64;
    __shared__ double sh[NW * 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nb = vb.gx, lb = xcd_block(vb.x, nb);
    for (int j = 0; j < nv; ++j) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q <= GMRES_MMAX; ++q)
            if (q == j) t = v[q];
#pragma unroll
        for (int off = 32; off >= KP; off >>= 1) t += __shfl_xor(t, off, 64);
        if (lane < KP) sh[wave * 64 + lane] = t;
        __syncthreads();
        if (threadIdx.x < KP) {
            const int c = vb.y * KP + threadIdx.x;
            if (c < k) {
                double acc = 0.0;
                for (int ww = 0; ww < NW; ++ww) acc += sh[ww * 64 + threadIdx.x];
                part[((size_t)j * k + c) * nb + lb] = acc;
            }
        }
        __syncthreads();
    }
}

// out[q k + c] = sum of part[(q k + c) nb + b] over b in block order; one block per (c, q)
static __global__ __launch_bounds__(BLOCK) void k_gmres_sum(const double* __restrict__ part, int nb, int k,
                                                     double* __restrict__ out) {
    const int c = blockIdx.x, q = blockIdx.y;
    const double* src = part + ((size_t)q * k + c) * nb;
    __shared__ double sh[BLOCK];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += BLOCK) acc += src[b];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)q * k + c] = sh[0];
}

// w -= sum_q h[q k + c] V_q (q ascending), elementwise over the n x k block
static __global__ __launch_bounds__(BLOCK) void k_gmres_orth(const double* __restrict__ V, size_t ldv, int nv,
                                                      const double* __restrict__ h, double* __restrict__ w,
                                                      size_t nk, int k) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nk) return;
    const int c = (int)(i % (size_t)k);
    double acc = w[i];
    for (int q = 0; q < nv; ++q) acc -= h[(size_t)q * k + c] * V[(size_t)q * ldv + i];
    w[i] = acc;
}

// v = w / sqrt(ww[c]) (0 where ww[c] = 0: a breakdown), elementwise
static __global__ __launch_bounds__(BLOCK) void k_gmres_scale(const double* __restrict__ w, const double* __restrict__ ww,
                                                       double* __restrict__ v, size_t nk, int k) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nk) return;
    const double nrm = sqrt(ww[i % (size_t)k]);
    v[i] = nrm > 0.0 ? w[i] / nrm : 0.0;
}

// column j of every column's Hessenberg matrix: H[c][q][j] = h1 + h2 (q < nv),
// H[c][nv][j] = sqrt(ww[c]); H is laid out [k][m + 1][m]
static __global__ void k_gmres_hcol(double* __restrict__ H, int m, int j, int nv, const double* __restrict__ h1,
                             const double* __restrict__ h2, const double* __restrict__ ww, int k) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k) return;
    double* Hc = H + (size_t)c * (m + 1) * m;
    for (int q = 0; q < nv; ++q) Hc[(size_t)q * m + j] = h1[(size_t)q * k + c] + h2[(size_t)q * k + c];
    Hc[(size_t)nv * m + j] = sqrt(ww[c]);
}

// x += sum_j y[c][j] V_j, j < m (y laid out [k][m]; zero rows leave x as it is)
static __global__ __launch_bounds__(BLOCK) void k_gmres_update(double* __restrict__ x, const double* __restrict__ V,
                                                        size_t ldv, int m, const double* __restrict__ y,
                                                        size_t nk, int k) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nk) return;
    const int c = (int)(i % (size_t)k);
    const double* yc = y + (size_t)c * m;
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += yc[j] * V[(size_t)j * ldv + i];
    x[i] += acc;
}

// ------------------------------------------------------------ BiCGSTAB (float64 vectors)
//
// jax.scipy.sparse.linalg.bicgstab, right-preconditioned, rhat = r0, per column:
//   p = r + beta (p - omega q); phat = M p; q = A phat; alpha = rho_ / rhat.q;
//   s = r - alpha q; exit = s.s < atol2; shat = M s; t = A shat;
//   omega = t.s / t.t; x += exit ? alpha phat : alpha phat + omega shat;
//   r = exit ? s : s - omega t; rho_ = rhat.r.
// phat and shat are the preconditioner outputs (cycle precision TZ, or the
// float64 vectors themselves without a preconditioner).

// x = 0, r = rhat = p = q = b; partial b.b
template <int KP>
__device__ __forceinline__ void bicg_init_body(const double* __restrict__ b, double* __restrict__ x,
                                               double* __restrict__ r, double* __restrict__ rh,
                                               double* __restrict__ p, double* __restrict__ q, int n, int k,
                                               Ctl ctl, const VB& vb) {
    const Lane<KP> ln(k, vb);
    double v[1] = {0.0};
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double bv[NP];
        red_rows<KP>(
            u, n, ln.cok, [&](int qq, int row, bool ok) { bv[qq] = ok ? b[(size_t)row * k + ln.c] : 0.0; },
            [&](int qq, int row) {
                const size_t o = (size_t)row * k + ln.c;
                v[0] += bv[qq] * bv[qq];
                x[o] = 0.0;
                r[o] = bv[qq];
                rh[o] = bv[qq];
                p[o] = bv[qq];
                q[o] = bv[qq];
            });
    });
    block_partials<KP, 1>(v, k, ctl, vb);
}
template <int KP>
__global__ __launch_bounds__(BLOCK) void k_bicg_init(const double* __restrict__ b, double* __restrict__ x,
                                                     double* __restrict__ r, double* __restrict__ rh,
                                                     double* __restrict__ p, double* __restrict__ q, int n, int k,
                                                     Ctl ctl) {
    bicg_init_body<KP>(b, x, r, rh, p, q, n, k, ctl, hw_vb());
}

// active columns: p = r + beta (p - omega q); pc = (TC) p (nullable)
template <typename TC, int KP>
__global__ __launch_bounds__(BLOCK) void k_bicg_p(const double* __restrict__ r, const double* __restrict__ q,
                                                  double* __restrict__ p, TC* __restrict__ pc, int n, int k,
                                                  Ctl ctl) {
    if (stopped(ctl.done)) return;
    const VB vb = hw_vb();
    const Lane<KP> ln(k, vb);
    const bool live = ln.cok && ctl.active[ln.c] != 0;
    const double beta = live ? ctl.beta[ln.c] : 0.0, omega = live ? ctl.omega[ln.c] : 0.0;
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double rv[NP], pv[NP], qv[NP];
        red_rows<KP>(
            u, n, live,
            [&](int qq, int row, bool ok) {
                const size_t o = (size_t)row * k + ln.c;
                rv[qq] = ok ? r[o] : 0.0;
                pv[qq] = ok ? p[o] : 0.0;
                qv[qq] = ok ? q[o] : 0.0;
            },
            [&](int qq, int row) {
                const size_t o = (size_t)row * k + ln.c;
                const double pn = rv[qq] + beta * (pv[qq] - omega * qv[qq]);
                p[o] = pn;
                if (pc != nullptr) pc[o] = static_cast<TC>(pn);
            });
    });
}

// y = A zhat (zhat in TZ, products in float64); MODE 0: partial w.y (w = rhat);
// MODE 1: partials y.w and y.y (w = s)
template <typename TZ, int KP, bool WIDE, int MODE>
__global__ __launch_bounds__(BLOCK) void k_bicg_spmv(Sell<double> A, const TZ* __restrict__ zh,
                                                     const double* __restrict__ w, double* __restrict__ y, int k,
                                                     Ctl ctl) {
    if (stopped(ctl.done)) return;
    const VB vb = hw_vb();
    const Lane<KP> ln(k, vb);
    const int n = A.nrows;
    double v[2] = {0.0, 0.0};
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        unit_spmv<double, KP, (KP > 1 && WIDE)>(
            A, u, ilog2(red_sub(KP)), ln.cok, [&](int j) { return static_cast<double>(zh[(size_t)j * k + ln.c]); },
            [&](int row, double yv) {
                if (row < n && ln.cok) {
                    const size_t o = (size_t)row * k + ln.c;
                    y[o] = yv;
                    v[0] += MODE == 0 ? w[o] * yv : yv * w[o];
                    if (MODE == 1) v[1] += yv * yv;
                }
            });
    });
    block_partials<KP, 2>(v, k, ctl, vb);
}

// active columns: s = r - alpha q; sc = (TC) s (nullable); partial s.s
template <typename TC, int KP>
__global__ __launch_bounds__(BLOCK) void k_bicg_s(const double* __restrict__ r, const double* __restrict__ q,
                                                  double* __restrict__ sv, TC* __restrict__ sc, int n, int k,
                                                  Ctl ctl) {
    if (stopped(ctl.done)) return;
    const VB vb = hw_vb();
    const Lane<KP> ln(k, vb);
    const bool act = ln.cok && ctl.active[ln.c] != 0;
    const double a = act ? ctl.alpha[ln.c] : 0.0;
    double v[1] = {0.0};
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double rv[NP], qv[NP];
        red_rows<KP>(
            u, n, act,
            [&](int qq, int row, bool ok) {
                const size_t o = (size_t)row * k + ln.c;
                rv[qq] = ok ? r[o] : 0.0;
                qv[qq] = ok ? q[o] : 0.0;
            },
            [&](int qq, int row) {
                const size_t o = (size_t)row * k + ln.c;
                const double s_ = rv[qq] - a * qv[qq];
                sv[o] = s_;
                if (sc != nullptr) sc[o] = static_cast<TC>(s_);
                v[0] += s_ * s_;
            });
    });
    block_partials<KP, 1>(v, k, ctl, vb);
}

// active columns: x += exit ? alpha phat : alpha phat + omega shat; r = exit ? s : s - omega t;
// partials rhat.r and r.r (all columns)
template <typename TZ, int KP>
__global__ __launch_bounds__(BLOCK) void k_bicg_update(const TZ* __restrict__ ph, const TZ* __restrict__ sh,
                                                       const double* __restrict__ sv, const double* __restrict__ t,
                                                       const double* __restrict__ rh, double* __restrict__ x,
                                                       double* __restrict__ r, int n, int k, Ctl ctl) {
    if (stopped(ctl.done)) return;
    const VB vb = hw_vb();
    const Lane<KP> ln(k, vb);
    const bool act = ln.cok && ctl.active[ln.c] != 0;
    const double a = act ? ctl.alpha[ln.c] : 0.0, om = act ? ctl.omega[ln.c] : 0.0;
    const bool ex = act && ctl.exitf[ln.c] != 0;
    double v[2] = {0.0, 0.0};
    constexpr int NP = red_passes<KP>();
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        double xv[NP], pv[NP], shv[NP], sv_[NP], tv[NP], rv[NP], rhv[NP];
        red_rows<KP>(
            u, n, ln.cok,
            [&](int qq, int row, bool ok) {
                const size_t o = (size_t)row * k + ln.c;
                const bool upd = ok && act;
                rhv[qq] = ok ? rh[o] : 0.0;
                rv[qq] = ok && !act ? r[o] : 0.0;
                xv[qq] = upd ? x[o] : 0.0;
                pv[qq] = upd ? static_cast<double>(ph[o]) : 0.0;
                sv_[qq] = upd ? sv[o] : 0.0;
                shv[qq] = upd && !ex ? static_cast<double>(sh[o]) : 0.0;
                tv[qq] = upd && !ex ? t[o] : 0.0;
            },
            [&](int qq, int row) {
                const size_t o = (size_t)row * k + ln.c;
                double rr = rv[qq];
                if (act) {
                    if (ex) {
                        x[o] = xv[qq] + a * pv[qq];
                        rr = sv_[qq];
                    } else {
                        x[o] = xv[qq] + (a * pv[qq] + om * shv[qq]);
                        rr = sv_[qq] - om * tv[qq];
                    }
                    r[o] = rr;
                }
                v[0] += rhv[qq] * rr;
                v[1] += rr * rr;
            });
    });
    block_partials<KP, 2>(v, k, ctl, vb);
}

// ||b - A x|| per column into ctl.normr (no state update)
template <typename T, int KP, bool WIDE>
__device__ __forceinline__ void true_resnorm_body(Sell<T> A, const T* __restrict__ x,
                                                         const T* __restrict__ b, int k, Ctl ctl, const VB& vb) {
    const Lane<KP> ln(k, vb);
    const int n = A.nrows;
    double v[1] = {0.0};
    red_unit(nunits_red(n, KP), vb, [&](int u) {
        unit_spmv<T, KP, (KP > 1 && WIDE)>(
            A, u, ilog2(red_sub(KP)), ln.cok, [&](int j) { return x[(size_t)j * k + ln.c]; },
            [&](int row, T ax) {
                if (row < n && ln.cok) {
                    const double rv = static_cast<double>(b[(size_t)row * k + ln.c] - ax);
                    v[0] += rv * rv;
                }
            });
    });
    block_partials<KP, 1>(v, k, ctl, vb);
}
template <typename T, int KP, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_true_resnorm(Sell<T> A, const T* __restrict__ x,
                                                         const T* __restrict__ b, int k, Ctl ctl) {
    true_resnorm_body<T, KP, WIDE>(A, x, b, k, ctl, hw_vb());
}

}  // namespace amg
