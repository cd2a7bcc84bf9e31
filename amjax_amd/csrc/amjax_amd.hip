// C-ABI and host engine of the gfx950 multigrid solve phase.
//
// A handle owns no hierarchy memory: the caller uploads the level-packed
// arrays once (two device buffers, int32 and values) and passes offsets.  The
// handle owns per-k workspaces (level vectors, reduction scratch, device
// control block) and the HIP graphs captured on them.  A cycle is emitted as a
// static kernel sequence from the V/W/F recursion; the standalone solve and
// PCG loops are captured as graphs of several iterations each whose kernels
// turn into no-ops once the device-side convergence flag is set, and the host
// keeps one graph in flight while it checks the flag of the previous one.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "amjax_amd.h"
#include "batch.hpp"
#include "batch_launch.hpp"
#include "kernels.hpp"
#include "tail.hpp"

using namespace amg;

// Device scratch of the allocation-free cycle (amg_cycle_noalloc): every GPU the
// library is loaded on holds its own copy until the process ends.  Host code
// uses the array's name only to ask hipGetSymbolAddress for the current GPU's
// copy (scratch_ptr below); kernels get that address as an argument.
static constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(AMG_NOALLOC_SCRATCH_MB) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

struct Error {
    int code;
    std::string msg;
};

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            throw Error{AMG_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)};       \
    } while (0)

#define ARGCHK(cond, msg)                                       \
    do {                                                        \
        if (!(cond)) throw Error{AMG_E_ARG, std::string(msg)};  \
    } while (0)

int kp_of(int k) {
    int kp = 1;
    while (kp < k && kp < 64) kp <<= 1;
    return kp;
}

#define KP_DISPATCH(kp, ...)                          \
    switch (kp) {                                     \
        case 1: { constexpr int KP = 1; __VA_ARGS__; } break;   \
        case 2: { constexpr int KP = 2; __VA_ARGS__; } break;   \
        case 4: { constexpr int KP = 4; __VA_ARGS__; } break;   \
        case 8: { constexpr int KP = 8; __VA_ARGS__; } break;   \
        case 16: { constexpr int KP = 16; __VA_ARGS__; } break; \
        case 32: { constexpr int KP = 32; __VA_ARGS__; } break; \
        default: { constexpr int KP = 64; __VA_ARGS__; } break; \
    }

#define WIDE_IF(wd, ...)                                       \
    if (wd) { constexpr bool WD = true; __VA_ARGS__; }         \
    else { constexpr bool WD = false; __VA_ARGS__; }

#define WIDE_DISPATCH(M, ...)                                  \
    if ((M).maxw > WMAX) { constexpr bool WD = true; __VA_ARGS__; } \
    else { constexpr bool WD = false; __VA_ARGS__; }

// vector-mode launches: column chunk of 16, 32 or 64
#define VKP_DISPATCH(kp, ...)                                               \
    switch (kp) {                                                           \
        case 16: { constexpr int KP = 16; __VA_ARGS__; } break;             \
        case 32: { constexpr int KP = 32; __VA_ARGS__; } break;             \
        case 64: { constexpr int KP = 64; __VA_ARGS__; } break;             \
        default: throw Error{AMG_E_STATE, "vector mode needs a column chunk of 16, 32 or 64"}; \
    }
// Vector row kernels over a matrix whose widest slice has maxw slots: with the
// full column chunk (KP = 64) the kernel is also specialised on the widest slice
// it handles, WC (5: five-point stencils, 7: seven-point, WVMAX, 64: any), so its
// registers are sized for that width only (more waves per SIMD for narrow rows).
#define VKPW_DISPATCH(kp, maxw, ...)                                                    \
    switch (kp) {                                                                       \
        case 16: { constexpr int KP = 16, WC = 64; __VA_ARGS__; } break;                \
        case 32: { constexpr int KP = 32, WC = 64; __VA_ARGS__; } break;                \
        case 64:                                                                        \
            if ((maxw) <= 5) { constexpr int KP = 64, WC = 5; __VA_ARGS__; }            \
            else if ((maxw) <= 7) { constexpr int KP = 64, WC = 7; __VA_ARGS__; }       \
            else if ((maxw) <= WVMAX) { constexpr int KP = 64, WC = WVMAX; __VA_ARGS__; } \
            else { constexpr int KP = 64, WC = 64; __VA_ARGS__; }                       \
            break;                                                                      \
        default: throw Error{AMG_E_STATE, "vector mode needs a column chunk of 16, 32 or 64"}; \
    }

// one wavefront per work unit, 4 units per block
dim3 row_grid_kp(int n, int k, int kp, bool wide) {
    const int upb = BLOCK / 64;
    int nb = (nunits(n, kp, wide) + upb - 1) / upb;
    if (nb < 1) nb = 1;
    return dim3((unsigned)nb, (unsigned)((k + kp - 1) / kp));
}
dim3 row_grid(int n, int k, bool wide = false) { return row_grid_kp(n, k, kp_of(k), wide); }

int cu_count();

// row blocks are read as 16-byte vectors (vector mode of the row kernels)
inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// Reducing launches and the CG vector kernels walk short units; the grid is
// capped at RED_BPC blocks per CU (over all column chunks), so waves loop over
// several units and a launch leaves at most that many per-block sums per
// column for the finaliser to add.
#ifndef RED_BPC
#define RED_BPC 8
#endif
// the cap, read once: AMG_RED_BPC in the environment overrides it (measurements)
int red_bpc() {
    static const int v = [] {
        const char* e = std::getenv("AMG_RED_BPC");
        return e != nullptr ? std::atoi(e) : RED_BPC;
    }();
    return v;
}
dim3 red_grid_cap(int n, int k, int bpc) {
    const int kp = kp_of(k);
    const int gy = (k + kp - 1) / kp;
    int nb = (nunits_red(n, kp) + BLOCK / 64 - 1) / (BLOCK / 64);
    if (bpc > 0) {
        const int cap = (cu_count() * bpc + gy - 1) / gy;
        if (nb > cap) nb = cap;
    }
    if (nb < 1) nb = 1;
    return dim3((unsigned)nb, (unsigned)gy);
}
// The reducing row kernels (SpMV-type, with gathers) of the reliable-update CG
// hide gather latency with more waves than the vector kernels: their own cap
// (AMG_RU_ROW_BPC in the environment overrides it; 0: uncapped).
#ifndef RU_ROW_BPC
#define RU_ROW_BPC 8
#endif
int ru_row_bpc() {
    static const int v = [] {
        const char* e = std::getenv("AMG_RU_ROW_BPC");
        return e != nullptr ? std::atoi(e) : RU_ROW_BPC;
    }();
    return v;
}
dim3 ru_row_grid(int n, int k) { return red_grid_cap(n, k, ru_row_bpc()); }
dim3 red_grid(int n, int k) {
    const int kp = kp_of(k);
    const int gy = (k + kp - 1) / kp;
    int nb = (nunits_red(n, kp) + BLOCK / 64 - 1) / (BLOCK / 64);
    if (red_bpc() > 0) {
        const int cap = (cu_count() * red_bpc() + gy - 1) / gy;
        if (nb > cap) nb = cap;
    }
    if (nb < 1) nb = 1;
    return dim3((unsigned)nb, (unsigned)gy);
}

// Address of the current GPU's `scratch`, resolved once per device with that
// device current (at handle creation, or at a device's first allocation-free call).
constexpr int SCRATCH_MAX_DEVICES = 64;
void* scratch_ptr() {
    static std::mutex mu;
    static void* ptr[SCRATCH_MAX_DEVICES] = {};
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    ARGCHK(dev >= 0 && dev < SCRATCH_MAX_DEVICES, "device ordinal out of range");
    std::lock_guard<std::mutex> g(mu);
    if (ptr[dev] == nullptr) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) throw Error{AMG_E_HIP, "hipGetSymbolAddress returned no address for the scratch array"};
        ptr[dev] = p;
    }
    return ptr[dev];
}

struct DevMem {
    void* p = nullptr;
    DevMem() = default;
    explicit DevMem(size_t bytes) {
        if (bytes == 0) bytes = 16;
        HIPCHK(hipMalloc(&p, bytes));
    }
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
};

// Launch `body` repeatedly until the host-mapped flag says done.  One launch
// stays queued ahead of the one being checked; the finaliser kernels mirror
// the device flag into mapped host memory, so checking it costs no copy.
void run_until_done(const int* hflag, hipGraphExec_t body, int max_launches, hipStream_t s) {
    hipEvent_t ev[2];
    HIPCHK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    try {
        int launched = 0, checked = 0;
        auto launch = [&]() {
            HIPCHK(hipGraphLaunch(body, s));
            HIPCHK(hipEventRecord(ev[launched & 1], s));
            ++launched;
        };
        if (max_launches > 0) launch();
        while (checked < launched) {
            if (launched < max_launches && launched - checked < 2) launch();
            HIPCHK(hipEventSynchronize(ev[checked & 1]));
            ++checked;
            if (*(volatile const int*)hflag) break;
        }
    } catch (...) {
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        throw;
    }
    HIPCHK(hipEventDestroy(ev[0]));
    HIPCHK(hipEventDestroy(ev[1]));
}

// ------------------------------------------------------------------ recording
//
// A handle can emit its kernel sequence into a Recorder instead of launching
// it; a batch merges the recorded sequences of many problems into one launch
// per phase.  Every recorded step carries a tag (stage, level, sub-step); the
// stage order PRE < DOWN (level ascending) < COARSE < UP (level descending) <
// POST is the order of a V-cycle inside a PCG iteration, so sorting the steps
// of all problems by tag keeps each problem's own order.

enum Stage { ST_PRE = 0, ST_DOWN = 1, ST_COARSE = 2, ST_UP = 3, ST_POST = 4 };

struct RecStep {
    int stage, level, sub, kid;
    Step st;
};

struct Recorder {
    std::vector<RecStep> steps;
    int stage = ST_PRE, level = 0, sub = 0;
    bool tuned = false;  // row kernels take the autotuner's plan (else the default plan)
    bool vec = true;     // vector mode of the row kernels where it applies
    void tag(int s, int l) {
        stage = s;
        level = l;
        sub = 0;
    }
};

// ------------------------------------------------------------------ handle

struct Smoother {
    int kind = AMG_SMOOTH_NONE, iters = 0;
    double omega = 1.0;
    std::vector<double> coef;  // AMG_SMOOTH_POLY coefficients
};

template <typename T>
struct LevelT {
    int n = 0;
    Sell<T> A{}, P{}, R{};
    Sell<T> D{};  // fused down-sweep operator R (I - w_pre A D^-1); nrows = 0 if absent
    int64_t nnzA = 0, nnzP = 0, nnzR = 0, padA = 0, padP = 0, padR = 0;
    const T* dinv = nullptr;
    Smoother pre, post;
};

struct HandleBase {
    virtual ~HandleBase() = default;
    int dtype = AMG_F64;
    int nlev = 0;
    virtual void cycle(char cyc, int cpl, const void* b, void* x, int k, int x_zero, hipStream_t s) = 0;
    // first level of the one-launch coarse tail of a zero-guess V-cycle (-1: none)
    virtual int tail_level() const = 0;
    // the cycle with its level vectors carved from `bytes` of device memory at `base`
    virtual void cycle_noalloc(char cyc, int cpl, const void* b, void* x, int k, int x_zero, void* base,
                               size_t bytes, hipStream_t s) = 0;
    virtual void solve(char cyc, int cpl, const void* b, void* x, int k, double tol, int maxiter, int refine,
                       int* iters, double* relres, hipStream_t s) = 0;
    virtual void pcg(char cyc, int cpl, const double* b, double* x, int k, double tol, double atol, int maxiter,
                     int flexible, int precond, int* iters, double* relres, hipStream_t s) = 0;
    virtual void bicgstab(char cyc, int cpl, const double* b, double* x, int k, double tol, double atol, int maxiter,
                          int precond, int* iters, double* relres, hipStream_t s) = 0;
    virtual void pcg_ir(char cyc, int cpl, const double* b, double* x, int k, double tol, double inner_tol,
                        int maxiter, int inner_maxiter, int flexible, int* iters, int* outer, double* relres,
                        hipStream_t s) = 0;
    virtual void gmres(char cyc, int cpl, const double* b, double* x, int k, double tol, double atol, int restart,
                       int maxiter, int precond, int method, int* iters, double* relres, hipStream_t s) = 0;
    virtual void pcg_ru(char cyc, int cpl, const double* b, double* x, int k, double tol, double atol, double delta,
                        int maxiter, int flexible, int* iters, int* updates, double* relres, hipStream_t s) = 0;
    virtual void set_level0(const amg_sell_desc* A) = 0;
    virtual void set_option(int opt, long long value) = 0;
    virtual float time_kernel(int kernel, int level, int k, int reps, hipStream_t s) = 0;
};

template <typename T>
Sell<T> make_sell(const amg_sell_desc& d) {
    if (d.nrows < 0 || d.ncols < 0) throw Error{AMG_E_ARG, "negative matrix shape"};
    if (d.nrows > 0 && (d.sptr == nullptr || (d.padded > 0 && (d.col == nullptr || d.val == nullptr))))
        throw Error{AMG_E_ARG, "null sliced-ELL array"};
    Sell<T> s;
    s.sptr = d.sptr;
    s.col = d.col;
    s.val = static_cast<const T*>(d.val);
    s.nrows = d.nrows;
    s.ncols = d.ncols;
    s.maxw = d.max_width;
    s.npr = d.nrows > 0 ? (int)((d.nnz + d.nrows / 2) / d.nrows) : 0;
    return s;
}
// A handle's matrices: the vector row kernels are specialised on the declared
// max_width (registers for that width only), so it must cover every slice; the
// slice offsets are read back once to check it (the caller's uploads may still
// be in flight, hence the device-wide wait).
template <typename T>
Sell<T> handle_sell(const amg_sell_desc& d) {
    const Sell<T> m = make_sell<T>(d);
    if (d.nrows > 0) {
        const int ns = nslices(d.nrows);
        std::vector<int32_t> h((size_t)ns + 1);
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(h.data(), d.sptr, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost));
        int w = 0;
        for (int i = 0; i < ns; ++i) w = std::max(w, (h[i + 1] - h[i]) / 64);
        if (w > d.max_width) throw Error{AMG_E_ARG, "max_width is smaller than the widest slice"};
    }
    return m;
}

// ------------------------------------------------------------------ autotuner
//
// A row kernel's launch shape is a Plan: KP column lanes per row group (the
// k-block; k > KP runs ceil(k / KP) column chunks on gridDim.y), the wide-row
// variant, the rows per wavefront unit (2^subl) and the grid (0: one block
// per four units; else a persistent grid of that many blocks whose waves loop
// over the units).  The tuner times a matrix's kernel under candidate plans in
// three stages (KP x wide, then rows per unit, then grid size in multiples of
// the kernel's resident blocks) and keeps the fastest if it beats the fixed
// shape by 3%; plans are cached process-wide per (kernel kind, dtype, kp(k),
// stored entries per row, rows <= WMAX slots, size class).  A plan changes
// only the work split: every row still adds its products in slot order, so
// all plans give the same sums, except the k = 1 wide variant (lane partial
// sums joined by an xor tree).  Matrices below 8192 rows are launch-latency
// bound and keep the fixed shape.

enum TuneKind { TK_SPMV = 0, TK_JACOBI = 1 };

struct Plan {
    int kp, wide, subl, grid;
    int vw;  // columns per lane in vector mode (wide == VMODE), else 0
};

struct TuneRec {
    Plan plan;
    float us_default, us_best;
};

struct Tuner {
    std::mutex mu;
    std::map<std::tuple<int, int, int, int, int, int>, TuneRec> cache;
};
Tuner& tuner() {
    static Tuner t;
    return t;
}

inline int size_class(int n) { return n >= 65536 ? 2 : (n >= 8192 ? 1 : 0); }
// Fixed shape: register-held slices of at least 16 rows per wave; below 8192
// rows (latency bound: few waves, each a chain of dependent loads) one pass
// of the row groups per wave.
// vector mode (vw > 0: the element type's columns per 16-byte lane) when k is
// a multiple of vw and the column chunk is at least VKP_MIN.  The fixed shape
// takes it for slices of at most WVMAX slots, and for slices of up to
// VDEF_MAXW slots (run in 8-slot chunks, one memory round per chunk) on
// matrices of at least 8192 rows; small wide matrices are latency bound and
// keep the cooperative lane-per-column rows.  The tuner also tries it on the
// rest.
constexpr int VDEF_MAXW = 64;
inline bool vec_ok(int k, int kp, int vw) { return vw > 0 && k % vw == 0 && kp >= VKP_MIN; }
inline bool vec_default(int maxw, int nrows) {
    return maxw <= WVMAX || (maxw <= VDEF_MAXW && size_class(nrows) > 0);
}
inline Plan default_plan(int maxw, int k, int nrows, int vw) {
    const int kp = kp_of(k), wide = maxw > WMAX ? 1 : 0;
    if (vec_ok(k, kp, vw) && vec_default(maxw, nrows)) return Plan{kp, VMODE, 0, 0, vw};
    const int subl = (size_class(nrows) == 0 && kp > 1) ? ilog2(64 / kp) : default_subl(kp, wide != 0);
    return Plan{kp, wide, subl, 0, 0};
}
inline int vw_of_dtype(int dtype) { return dtype == AMG_F64 ? vec_width<double>() : vec_width<float>(); }
inline std::tuple<int, int, int, int, int, int> tune_key(int kind, int dtype, int npr, int maxw, int nrows, int k) {
    return std::make_tuple(kind, dtype, kp_of(k), npr < 255 ? npr : 255, maxw <= WMAX ? 1 : 0, size_class(nrows));
}

// grid of a row kernel under a plan
dim3 row_grid_plan(int n, int k, const Plan& pl) {
    const int upb = BLOCK / 64;
    const int nu = pl.wide == VMODE ? nunits_v(n, pl.kp, pl.vw) : nunits_l(n, pl.kp, pl.wide != 0, pl.subl);
    int nb = (nu + upb - 1) / upb;
    if (nb < 1) nb = 1;
    if (pl.grid > 0 && pl.grid < nb) nb = pl.grid;
    return dim3((unsigned)nb, (unsigned)((k + pl.kp - 1) / pl.kp));
}

// cached plan for a matrix, or the fixed shape
inline Plan lookup_plan(int kind, int dtype, const SellRaw& m, int k) {
    const int vw = vw_of_dtype(dtype);
    if (size_class(m.nrows) == 0) return default_plan(m.maxw, k, m.nrows, vw);
    Tuner& t = tuner();
    std::lock_guard<std::mutex> g(t.mu);
    auto it = t.cache.find(tune_key(kind, dtype, m.npr, m.maxw, m.nrows, k));
    return it == t.cache.end() ? default_plan(m.maxw, k, m.nrows, vw) : it->second.plan;
}

int cu_count() {
    static int n = 0;
    if (n == 0) {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        if (n <= 0) n = 1;
    }
    return n;
}

// Launch the tuned kernel of `kind` over M once (kind TK_JACOBI needs a square M);
// with occ != nullptr only report its resident blocks per CU.
template <typename T>
void tune_launch(int kind, const Sell<T>& M, const T* x, const T* b, const T* dinv, T* y, int k, const Plan& pl,
                 hipStream_t s, int* occ = nullptr) {
    const dim3 g = row_grid_plan(M.nrows, k, pl);
    if (pl.wide == VMODE) {
        VKPW_DISPATCH(pl.kp, M.maxw, {
            if (kind == TK_SPMV && M.maxw > WVMAX) {
                constexpr int NRW = wide_nru<T, KP>();
                dim3 gw = g;
                if (pl.grid == 0) gw.x *= 2;  // one pass per unit: twice the units
                if (occ) HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_spmv_v<T, KP, false, 16, 64, NRW>, BLOCK, 0));
                else k_spmv_v<T, KP, false, 16, 64, NRW><<<gw, BLOCK, 0, s>>>(M, x, y, k, nullptr);
            } else if (kind == TK_SPMV) {
                if (occ) HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_spmv_v<T, KP, false, 8, WC>, BLOCK, 0));
                else k_spmv_v<T, KP, false, 8, WC><<<g, BLOCK, 0, s>>>(M, x, y, k, nullptr);
            } else {
                if (occ) HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_jacobi_v<T, KP, WC>, BLOCK, 0));
                else k_jacobi_v<T, KP, WC><<<g, BLOCK, 0, s>>>(M, dinv, x, b, y, T(0.5), k, nullptr, nullptr);
            }
        });
        HIPCHK(hipGetLastError());
        return;
    }
    KP_DISPATCH(pl.kp, WIDE_IF(pl.wide, {
        if (kind == TK_SPMV) {
            if (occ) HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_spmv<T, KP, WD, false>, BLOCK, 0));
            else k_spmv<T, KP, WD, false><<<g, BLOCK, 0, s>>>(M, x, y, k, pl.subl, nullptr);
        } else {
            if (occ) HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_jacobi<T, KP, WD>, BLOCK, 0));
            else k_jacobi<T, KP, WD><<<g, BLOCK, 0, s>>>(M, dinv, x, b, y, T(0.5), k, pl.subl, nullptr, nullptr);
        }
    }));
    HIPCHK(hipGetLastError());
}

// Time the candidate plans of M at block width k (once per key) and cache the fastest.
template <typename T>
void tune_matrix(int kind, const Sell<T>& M, int k) {
    if (M.nrows <= 0 || size_class(M.nrows) == 0) return;
    if (kind == TK_JACOBI && M.nrows != M.ncols) return;
    const int dtype = std::is_same<T, double>::value ? AMG_F64 : AMG_F32;
    const auto key = tune_key(kind, dtype, M.npr, M.maxw, M.nrows, k);
    {
        Tuner& t = tuner();
        std::lock_guard<std::mutex> g(t.mu);
        if (t.cache.count(key)) return;
    }
    const size_t nx = (size_t)M.ncols * k, ny = (size_t)M.nrows * k;
    DevMem xs(sizeof(T) * nx), ys(sizeof(T) * ny), bs(sizeof(T) * ny), ds(sizeof(T) * (size_t)M.nrows);
    HIPCHK(hipMemset(xs.p, 0, sizeof(T) * nx));
    HIPCHK(hipMemset(bs.p, 0, sizeof(T) * ny));
    HIPCHK(hipMemset(ds.p, 0, sizeof(T) * (size_t)M.nrows));
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int vw = vec_width<T>();
    const Plan dp = default_plan(M.maxw, k, M.nrows, vw);
    TuneRec best{dp, 0.f, 1e30f};
    // a candidate's time is the better of two rounds; it displaces the default
    // plan only when more than 5% faster (the timing runs on warm buffers with
    // the plain product, so near-ties do not carry over to the solve)
    const auto is_default = [&](const Plan& pl) {
        return pl.kp == dp.kp && pl.wide == dp.wide && pl.subl == dp.subl && pl.grid == dp.grid;
    };
    auto time_plan = [&](const Plan& pl) {
        constexpr int reps = 10;
        auto once = [&] {
            tune_launch<T>(kind, M, (const T*)xs.p, (const T*)bs.p, (const T*)ds.p, (T*)ys.p, k, pl, s);
        };
        once();
        once();
        float us = 1e30f;
        for (int round = 0; round < 2; ++round) {
            HIPCHK(hipEventRecord(e0, s));
            for (int i = 0; i < reps; ++i) once();
            HIPCHK(hipEventRecord(e1, s));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            us = std::min(us, ms * 1e3f / reps);
        }
        const float margin = is_default(best.plan) && !is_default(pl) ? 0.95f : 1.0f;
        if (us < best.us_best * margin) {
            best.us_best = us;
            best.plan = pl;
        }
        return us;
    };
    try {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        best.us_default = time_plan(dp);
        // stage 1: column lanes per row group and the wide-row variant
        const int kp0 = kp_of(k);
        for (int kp = kp0; kp >= 1 && kp >= kp0 / 4; kp >>= 1)
            for (int wide = 0; wide <= VMODE; ++wide) {
                if (wide == VMODE && !vec_ok(k, kp, vw)) continue;
                if (wide == 0 && !(kp == 1 || M.maxw <= WMAX)) continue;
                const Plan pl = wide == VMODE ? Plan{kp, VMODE, 0, 0, vw} : Plan{kp, wide, default_subl(kp, wide != 0), 0, 0};
                if (pl.kp != dp.kp || pl.wide != dp.wide) time_plan(pl);
            }
        // stage 2: rows per unit (multi-column paths)
        const Plan p1 = best.plan;
        if (p1.kp > 1 && p1.wide != VMODE) {
            const int lo = ilog2(64 / p1.kp);
            for (int sl = lo; sl <= 6; ++sl)
                if (sl != p1.subl && sl >= p1.subl - 2 && sl <= p1.subl + 2) time_plan(Plan{p1.kp, p1.wide, sl, 0, 0});
        }
        // stage 3: persistent grid of 1, 2 or 3 rounds of resident blocks
        const Plan p2 = best.plan;
        int occ = 0;
        tune_launch<T>(kind, M, nullptr, nullptr, nullptr, nullptr, k, p2, s, &occ);
        const int full = (int)row_grid_plan(M.nrows, k, Plan{p2.kp, p2.wide, p2.subl, 0, p2.vw}).x;
        if (occ > 0)
            for (int m = 1; m <= 3; ++m) {
                const int gsz = m * occ * cu_count();
                if (gsz < full) time_plan(Plan{p2.kp, p2.wide, p2.subl, gsz, p2.vw});
            }
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipStreamDestroy(s));
    if (best.us_best > 0.97f * best.us_default) {
        best.plan = dp;
        best.us_best = best.us_default;
    }
    Tuner& t = tuner();
    std::lock_guard<std::mutex> g(t.mu);
    t.cache[key] = best;
}

// One sweep of the polynomial smoother over A: xo = x + h with h = c_0 r,
// h = c_i r + A h, r = b - A x (x == nullptr: zero guess, r = b).  r, h0, h1 are
// n x k work vectors; xo may not alias x.
template <typename T>
void launch_poly_sweep(const Sell<T>& A, const Plan& plan, const double* coef, int nc, const T* x, const T* b, T* xo,
                       T* r, T* h0, T* h1, int k, const int* mask, const int* done, hipStream_t s) {
    // the polynomial kernels have no vector mode
    const Plan pl = plan.wide == VMODE ? default_plan(A.maxw, k, A.nrows, 0) : plan;
    T* hb[2] = {h0, h1};
    const T* rr = b;
    const T c0 = static_cast<T>(coef[0]);
    const int n = A.nrows;
    if (x == nullptr) {
        KP_DISPATCH(kp_of(k), {
            k_poly_first_zero<T, KP><<<row_grid(n, k), BLOCK, 0, s>>>(b, nc == 1 ? xo : hb[0], c0, n, k, mask, done);
        });
    } else {
        rr = r;
        KP_DISPATCH(pl.kp, WIDE_IF(pl.wide, {
            k_poly_first<T, KP, WD><<<row_grid_plan(n, k, pl), BLOCK, 0, s>>>(A, x, b, r, hb[0], nc == 1 ? xo : nullptr,
                                                                            c0, k, pl.subl, mask, done);
        }));
    }
    HIPCHK(hipGetLastError());
    int cur = 0;
    for (int i = 1; i < nc; ++i) {
        const bool last = i == nc - 1;
        T* out = last ? xo : hb[1 - cur];
        const T ci = static_cast<T>(coef[i]);
        KP_DISPATCH(pl.kp, WIDE_IF(pl.wide, {
            k_poly_step<T, KP, WD><<<row_grid_plan(n, k, pl), BLOCK, 0, s>>>(A, hb[cur], rr, x, out, ci, last ? 1 : 0, k,
                                                                           pl.subl, mask, done);
        }));
        HIPCHK(hipGetLastError());
        cur = 1 - cur;
    }
}

// Device vectors of one block size k.
template <typename T>
struct Workspace {
    int k = 0;
    uint64_t last_use = 0;  // handle-local use stamp (eviction order)
    std::vector<std::unique_ptr<DevMem>> mem;
    std::vector<T*> X, Tm, B, Rr;   // per level: iterate, temp, rhs, residual
    std::vector<T*> H0, H1;         // per level: polynomial smoother terms (levels with one)
    T* X0 = nullptr;                // level-0 iterate of the solve loop
    T* T0b = nullptr;               // second level-0 temp (masked solve)
    T* B0 = nullptr;                // level-0 rhs of the solve loop / cycle input of PCG
    T* Z = nullptr;                 // preconditioner output (PCG / refinement)
    double* cwork = nullptr;        // coarse LU/QR work vector when the factors are float64
    double *ob = nullptr, *ox = nullptr, *orr = nullptr, *oAp = nullptr;  // float64 outer
    double* op[2] = {nullptr, nullptr};  // search direction, double-buffered across iterations
    double *orh = nullptr, *os = nullptr, *ot = nullptr;  // BiCGSTAB rhat, s, t (allocated on first use)
    T* Z2 = nullptr;                     // BiCGSTAB second preconditioner output
    // refinement around an inner CG in the cycle precision (allocated on first use):
    // its control block, host mirror of its `done`, and its x, p, Ap (r is B0)
    Ctl ictl{};
    int* ihflag = nullptr;
    T *ix = nullptr, *ip = nullptr, *iap = nullptr;
    int* itot = nullptr;  // [k] inner iterations summed over the steps
    // restarted GMRES (allocated on first use, for restart length gm): basis
    // V_0..V_gm, work vectors, per-column Hessenberg matrices [k][gm + 1][gm],
    // dot products, least-squares solutions [k][gm], partial sums, and a control
    // block whose `done` stays 0 (the reducing kernels it serves always run)
    int gm = 0;
    double *gV = nullptr, *gw = nullptr, *gAv = nullptr, *gH = nullptr, *gh1 = nullptr, *gh2 = nullptr;
    double *gww = nullptr, *gy = nullptr, *gpart = nullptr, *gout = nullptr;
    Ctl gctl{};
    int* ghflag = nullptr;
    Ctl ctl{};
    int* hflag = nullptr;           // pinned, device-mapped mirror of `done`
    std::map<std::tuple<int, char, int, int, int>, hipGraphExec_t> graphs;
    std::vector<hipGraph_t> raw;

    T* alloc_t(size_t n) {
        mem.emplace_back(new DevMem(n * sizeof(T)));
        return static_cast<T*>(mem.back()->p);
    }
    double* alloc_d(size_t n) {
        mem.emplace_back(new DevMem(n * sizeof(double)));
        return static_cast<double*>(mem.back()->p);
    }
    ~Workspace() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        for (auto g : raw) (void)hipGraphDestroy(g);
        if (hflag) (void)hipHostFree(hflag);
        if (ihflag) (void)hipHostFree(ihflag);
        if (ghflag) (void)hipHostFree(ghflag);
    }
};

template <typename T>
struct Handle : HandleBase {
    std::vector<LevelT<T>> lv;
    amg_coarse_desc cd{};
    const void* cf0 = nullptr;   // coarse factors (T, or double with cf64)
    const void* cf1 = nullptr;
    const int* cpiv = nullptr;
    bool cf64 = false;
    Sell<double> outer{};
    Sell<T> A0_orig{};
    int64_t outer_nnz = 0;
    bool has_outer = false;
    bool outer_is_level0 = false;
    std::map<int, std::unique_ptr<Workspace<T>>> ws;
    hipStream_t cap = nullptr;
    bool fuse = true;  // fused down-sweep where the level allows it
    Recorder* rec_ = nullptr;  // set while a batch records this handle's sequence
    // Eager timing: events recorded around each launch of one kernel kind
    // (pairs: before, after) while a time_* helper has the hook armed.
    enum EvKind { EV_PCG_UPDATE = 0, EV_RU_UPDATE = 1, EV_RU_GAMMA = 2 };
    struct EvHook {
        hipEvent_t* ev = nullptr;
        int n = 0, cap = 0, kind = -1;
        const float4* flush = nullptr;  // cache-cold: read before each timed launch
        size_t nflush = 0;
        bool on(int k) const { return ev != nullptr && kind == k && n < cap; }
        hipEvent_t next() { return ev[n++]; }
    };
    EvHook ev_hook_;
    static constexpr int FLUSH_BLOCKS = 4096;
    // event before (begin) or after a launch of kind `kind` under timing; a
    // cache-cold timing first reads the flush buffer, outside the events
    void ev_mark(int kind, bool begin, hipStream_t s) {
        if (!ev_hook_.on(kind)) return;
        if (begin && ev_hook_.flush != nullptr) {
            k_flush_read<<<FLUSH_BLOCKS, BLOCK, 0, s>>>(ev_hook_.flush, ev_hook_.nflush,
                                                        reinterpret_cast<float*>(const_cast<float4*>(ev_hook_.flush)));
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipEventRecord(ev_hook_.next(), s));
    }

    void tag(int stage, int level) {
        if (rec_) rec_->tag(stage, level);
    }
    Step step0(int n, int k) const {
        Step st;
        std::memset(&st, 0, sizeof(st));
        st.n = n;
        st.k = k;
        st.kp = kp_of(k);
        return st;
    }
    bool tune_ = false;  // single solves use the autotuner's plans (AMG_OPT_AUTOTUNE)
    bool vec_ = true;    // vector mode of the row kernels where it applies (AMG_OPT_VECTOR)
    int tail_n_ = AMG_TAIL_ROWS_DEFAULT;  // AMG_OPT_TAIL: largest first level of a one-launch coarse tail

    // ---------------------------------------------------------------- coarse tail
    // The tail's plan: first level, LDS image of its operators (see tail.hpp)
    // and the layout, rebuilt whenever an option it depends on changes.
    struct TailPlan {
        int start = -1;  // first level of the tail, -1: none
        bool has_r = false;  // R in the image: nonzero-guess visits (W, F, 'T' cycles)
        std::unique_ptr<DevMem> img;
        int img_n = 0;   // image size in 16-byte units
        size_t lds = 0;  // LDS per workgroup: image + level vectors
        std::vector<TailMat> D, P, A, R;
        std::vector<int> dinv;
        int ainv = 0;
    };
    TailPlan tail_;   // the shallowest tail (zero-guess V visits; every visit when it has R)
    TailPlan tailr_;  // a deeper tail with R, when tail_ has none (W, F, 'T' and nonzero visits)
    size_t tail_lds_max_ = 0;  // LDS a workgroup may use on this device

    template <typename E>
    static std::vector<E> d2h(const E* p, size_t n) {
        std::vector<E> v(n);
        if (n > 0) HIPCHK(hipMemcpy(v.data(), p, n * sizeof(E), hipMemcpyDeviceToHost));
        return v;
    }
    // every level from l down takes the fused down-sweep and one Jacobi post-sweep
    bool tail_levels_ok(int l) const {
        for (int m = l; m < nlev - 1; ++m) {
            const LevelT<T>& L = lv[m];
            if (L.D.nrows == 0 || L.pre.kind != AMG_SMOOTH_JACOBI || L.pre.iters != 1 ||
                L.post.kind != AMG_SMOOTH_JACOBI || L.post.iters != 1)
                return false;
        }
        return true;
    }
    // Lay out (and with `img` != nullptr write) the image of the tail from level
    // l; returns its size in bytes.  Matrices are compacted slice by slice.
    size_t tail_image(int l, bool with_r, std::vector<unsigned char>* img, TailPlan& tp) const {
        size_t off = 0;
        auto region = [&](size_t bytes) {
            const size_t at = off;
            off += (bytes + 15) & ~size_t(15);
            if (img) img->resize(off, 0);
            return (int)at;
        };
        auto put = [&](const Sell<T>& M) {
            const int n = M.nrows, ns = nslices(n);
            const std::vector<int> sp = d2h(M.sptr, (size_t)ns + 1);
            std::vector<int> spc((size_t)ns + 1, 0);
            for (int s = 0; s < ns; ++s) {
                const int v = std::min(64, n - 64 * s);
                spc[s + 1] = spc[s] + v * ((sp[s + 1] - sp[s]) / 64);
            }
            TailMat m{n, 0, 0, 0};
            m.sp = region(sizeof(int) * ((size_t)ns + 1));
            m.col = region(sizeof(int) * (size_t)spc[ns]);
            m.val = region(sizeof(T) * (size_t)spc[ns]);
            if (img) {
                const size_t g = (size_t)(sp[ns] - sp[0]);
                const std::vector<int> col = d2h(M.col + sp[0], g);
                const std::vector<T> val = d2h(M.val + sp[0], g);
                std::memcpy(img->data() + m.sp, spc.data(), sizeof(int) * spc.size());
                int* oc = reinterpret_cast<int*>(img->data() + m.col);
                T* ov = reinterpret_cast<T*>(img->data() + m.val);
                for (int s = 0; s < ns; ++s) {
                    const int v = std::min(64, n - 64 * s), w = (sp[s + 1] - sp[s]) / 64;
                    for (int t = 0; t < w; ++t)
                        for (int q = 0; q < v; ++q) {
                            const size_t src = (size_t)(sp[s] - sp[0]) + 64 * (size_t)t + q;
                            const size_t dst = (size_t)spc[s] + (size_t)t * v + q;
                            oc[dst] = col[src];
                            ov[dst] = val[src];
                        }
                }
            }
            return m;
        };
        tp.D.clear();
        tp.P.clear();
        tp.A.clear();
        tp.R.clear();
        tp.dinv.clear();
        for (int m = l; m < nlev - 1; ++m) {
            const LevelT<T>& L = lv[m];
            tp.D.push_back(put(L.D));
            tp.P.push_back(put(L.P));
            tp.A.push_back(put(L.A));
            tp.R.push_back(with_r ? put(L.R) : TailMat{0, 0, 0, 0});
            tp.dinv.push_back(region(sizeof(T) * (size_t)L.n));
            if (img) {
                const std::vector<T> d = d2h(L.dinv, (size_t)L.n);
                std::memcpy(img->data() + tp.dinv.back(), d.data(), sizeof(T) * d.size());
            }
        }
        const size_t nc = (size_t)lv[nlev - 1].n;
        const size_t tf = cf64 ? sizeof(double) : sizeof(T);
        tp.ainv = region(tf * nc * nc);
        if (img) {
            if (cf64) {
                const std::vector<double> f = d2h(static_cast<const double*>(cf0), nc * nc);
                std::memcpy(img->data() + tp.ainv, f.data(), tf * nc * nc);
            } else {
                const std::vector<T> f = d2h(static_cast<const T*>(cf0), nc * nc);
                std::memcpy(img->data() + tp.ainv, f.data(), tf * nc * nc);
            }
        }
        return off;
    }
    // level-vector elements of the tail from l: b and two iterates on every
    // level below l, b and x on the coarsest; on level l b and one iterate,
    // plus the caller's iterate when nonzero-guess visits can occur (with_r)
    size_t tail_vec_elems(int l, bool with_r) const {
        size_t e = 2 * (size_t)lv[nlev - 1].n;
        for (int m = l; m < nlev - 1; ++m) e += (size_t)lv[m].n * (m == l && !with_r ? 2 : 3);
        return e;
    }
    template <typename TF>
    void tail_attr(size_t lds) {
        HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tail<T, TF>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tail_v<T, TF>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    // build the tail from level l (with or without R) if it fits the LDS
    bool build_tail(int l, bool with_r, TailPlan& out) {
        TailPlan tp;
        const size_t bytes = tail_image(l, with_r, nullptr, tp);
        const size_t lds = bytes + sizeof(T) * tail_vec_elems(l, with_r);
        if (lds > tail_lds_max_ || bytes > (size_t)16 * TAIL_COPY_UNROLL * TAIL_BLOCK) return false;
        std::vector<unsigned char> img;
        tail_image(l, with_r, &img, tp);
        tp.img.reset(new DevMem(img.size()));
        HIPCHK(hipMemcpy(tp.img->p, img.data(), img.size(), hipMemcpyHostToDevice));
        tp.img_n = (int)(img.size() / 16);
        tp.lds = lds;
        tp.start = l;
        tp.has_r = with_r;
        out = std::move(tp);
        return true;
    }
    void plan_tail() {
        tail_ = TailPlan{};
        tailr_ = TailPlan{};
        if (tail_n_ <= 0 || !fuse || cd.kind != AMG_COARSE_PINV || cf0 == nullptr || tail_lds_max_ == 0) return;
        // the operators may still be in flight on the caller's stream
        HIPCHK(hipDeviceSynchronize());
        for (int l = 1; l < nlev - 1; ++l) {
            if (lv[l].n > tail_n_ || nlev - 1 - l > TAIL_MAXL || !tail_levels_ok(l)) continue;
            if (tail_.start < 0) {
                // the shallowest level: with R when it fits (every visit), else without
                if ((build_tail(l, true, tail_) || build_tail(l, false, tail_)) && tail_.has_r) break;
                continue;
            }
            // a deeper level whose tail has R, for the visits tail_ cannot take
            if (build_tail(l, true, tailr_)) break;
        }
        // one kernel attribute for both: the larger LDS
        const size_t lds = std::max(tail_.lds, tailr_.lds);
        if (lds > 0) {
            if (cf64) tail_attr<double>(lds);
            else tail_attr<T>(lds);
        }
    }

    // The tail's program for one visit of its first level: the phases the
    // per-level emission (rec) would launch, in the same order, on the same
    // values.  Level i of the tail has its right-hand side at B[i] and two
    // iterate buffers X[i], Y[i]; visit() returns the buffer that holds the
    // level's new iterate (the first level's goes to the launch's global
    // output).  ok = false if the program is too long or needs R where the
    // image has none.
    struct TailProg {
        std::vector<TailOp> ops;
        std::vector<int> B, X, Y;
        int bc = 0, xc = 0;
        int nl = 0;
        bool has_r = false;
        bool ok = true;
        void push(int kind, int lvl, int in, int out, int aux = 0) {
            if ((int)ops.size() >= TAIL_MAXOPS) ok = false;
            else ops.push_back(TailOp{kind, lvl, in, out, aux});
        }
        int other(int i, int u) const { return u == X[i] ? Y[i] : X[i]; }
        // visit level i from a zero guess, or from the iterate in buffer cur
        int visit(int i, bool xzero, char ct, int cpl, int cur) {
            const int bnext = i + 1 < nl ? B[i + 1] : bc;
            int u = X[i], v = Y[i];
            if (xzero) {
                push(TOP_DOWN_D, i, B[i], bnext);  // pre-sweep from 0, residual, restriction
            } else {
                if (!has_r) ok = false;
                u = cur;
                v = other(i, cur);
                push(TOP_PRE, i, u, v);           // x = sweep(x) into v
                push(TOP_RES, i, v, u);           // r = b - A x into u
                push(TOP_RESTRICT, i, u, bnext);  // b_{i+1} = R r
            }
            int e = xc;
            if (i + 1 == nl) {
                push(TOP_COARSE, i, 0, 0);
            } else if (ct == 'V') {
                e = visit(i + 1, true, 'V', cpl, 0);
            } else if (ct == 'W') {
                e = visit(i + 1, true, 'W', cpl, 0);
                e = visit(i + 1, false, 'W', cpl, e);
            } else if (ct == 'F') {
                e = visit(i + 1, true, 'F', cpl, 0);
                for (int c = 0; c < cpl; ++c) e = visit(i + 1, false, 'V', cpl, e);
            } else {
                bool z = true;
                e = X[i + 1];
                for (int c = 0; c < cpl; ++c, z = false) e = visit(i + 1, z, 'V', cpl, e);
                e = visit(i + 1, z, 'T', cpl, e);
            }
            const int dst = i == 0 ? -1 : v;
            if (xzero) {
                push(TOP_UP_X1, i, e, v);   // t = w_pre D^-1 b + P e into v
                push(TOP_POST, i, v, i == 0 ? -1 : u);
                return u;
            }
            push(TOP_UP_ADD, i, e, u, v);   // t = x (in v) + P e into u
            push(TOP_POST, i, u, dst);      // x = sweep(t) into v
            return v;
        }
    };
    // the program of one visit of the tail's first level (the caller's iterate
    // is staged in X[0] when the guess is not zero)
    bool tail_prog(const TailPlan& tp, char ct, int cpl, bool xzero, TailProg& pg) const {
        const int l = tp.start;
        pg.nl = nlev - 1 - l;
        pg.has_r = tp.has_r;
        int off = 0;
        pg.B.resize(pg.nl);
        pg.X.resize(pg.nl);
        pg.Y.resize(pg.nl);
        for (int i = 0; i < pg.nl; ++i) {
            // B, Y, X: the first level's X (the caller's iterate) only with R
            const int n = lv[l + i].n;
            const bool x_here = i > 0 || tp.has_r;
            pg.B[i] = off;
            pg.Y[i] = off + n;
            pg.X[i] = x_here ? off + 2 * n : -1;
            off += (x_here ? 3 : 2) * n;
        }
        pg.bc = off;
        pg.xc = off + lv[nlev - 1].n;
        pg.visit(0, xzero, ct, cpl, pg.X[0]);
        return pg.ok;
    }
    // a visit of the tail's first level runs as one k_tail launch
    // the tail that runs a visit of level l as one launch, or nullptr
    const TailPlan* tail_for(int l, char ct, bool xzero, int cpl) const {
        if (rec_ != nullptr) return nullptr;
        for (const TailPlan* tp : {&tail_, &tailr_}) {
            TailProg pg;
            if (tp->start >= 1 && l == tp->start && tail_prog(*tp, ct, cpl, xzero, pg)) return tp;
        }
        return nullptr;
    }
    int tail_level() const override { return tail_.start; }
    template <typename TF>
    void tail_launch_t(const TailPlan& tp, char ct, int cpl, const T* b, const T* xin, T* x, int k,
                       const int* done, hipStream_t s) {
        const int l = tp.start;
        TailProg pg;
        if (!tail_prog(tp, ct, cpl, xin == nullptr, pg)) throw Error{AMG_E_STATE, "coarse tail program not available"};
        TailArgs<T, TF> a;
        std::memset(&a, 0, sizeof(a));
        a.nl = pg.nl;
        const auto tree = [](const Plan& pl) { return pl.kp == 1 && pl.wide == 1 ? 1 : 0; };
        for (int i = 0; i < a.nl; ++i) {
            const LevelT<T>& L = lv[l + i];
            TailLv<T>& t = a.L[i];
            t.D = tp.D[i];
            t.P = tp.P[i];
            t.A = tp.A[i];
            t.R = tp.R[i];
            t.dinv = tp.dinv[i];
            t.w_pre = static_cast<T>(L.pre.omega);
            t.w_post = static_cast<T>(L.post.omega);
            t.tD = tree(plan_of(L.D, k));
            t.tP = tree(plan_of(L.P, k));
            t.tA = tree(plan_of(L.A, k, TK_JACOBI));
            t.tR = tp.has_r ? tree(plan_of(L.R, k)) : 0;
            t.ob = pg.B[i];
            t.oy = pg.Y[i];
            t.ox = pg.X[i];
        }
        a.img = static_cast<const uint4*>(tp.img->p);
        a.img_n = tp.img_n;
        a.ainv = tp.ainv;
        a.nc = lv[nlev - 1].n;
        a.obc = pg.bc;
        a.oxc = pg.xc;
        a.vec = tp.img_n * 16;
        a.b = b;
        a.xin = xin;
        a.oxin = pg.X[0];
        a.x = x;
        a.k = k;
        a.done = done;
        a.nops = (int)pg.ops.size();
        std::copy(pg.ops.begin(), pg.ops.end(), a.ops);
        if (ct == 'V' && xin == nullptr)
            k_tail_v<T, TF><<<dim3((unsigned)k), TAIL_BLOCK, tp.lds, s>>>(a);
        else
            k_tail<T, TF><<<dim3((unsigned)k), TAIL_BLOCK, tp.lds, s>>>(a);
        HIPCHK(hipGetLastError());
    }
    // one visit (cycle ct from level tp.start, from xin or from zero) as one launch
    void tail_launch(const TailPlan& tp, char ct, int cpl, const T* b, const T* xin, T* x, int k, const int* done,
                     hipStream_t s) {
        if (cf64) tail_launch_t<double>(tp, ct, cpl, b, xin, x, k, done, s);
        else tail_launch_t<T>(tp, ct, cpl, b, xin, x, k, done, s);
    }

    // wide-row variant of the reducing kernels (fixed work split)
    template <typename TT>
    int wide_of(const Sell<TT>& M, int) const {
        return M.maxw > WMAX ? 1 : 0;
    }
    // variant of the reducing SpMV kernels (fixed work split): vector mode where it applies
    template <typename TT>
    int red_mode(const Sell<TT>& M, int k) const {
        const bool vec = rec_ ? rec_->vec : vec_;
        return vec && vec_ok(k, kp_of(k), vec_width<TT>()) && vec_default(M.maxw, M.nrows) ? VMODE : wide_of(M, k);
    }
    // launch shape of a row kernel over M
    template <typename TT>
    Plan plan_of(const Sell<TT>& M, int k, int kind = TK_SPMV) const {
        const bool tuned = rec_ ? rec_->tuned : tune_;
        const bool vec = rec_ ? rec_->vec : vec_;
        if (!tuned) return default_plan(M.maxw, k, M.nrows, vec ? vec_width<TT>() : 0);
        const Plan pl = lookup_plan(kind, std::is_same<TT, double>::value ? AMG_F64 : AMG_F32, raw_of(M), k);
        return (pl.wide == VMODE && !vec) ? default_plan(M.maxw, k, M.nrows, 0) : pl;
    }
    // measure plans for every matrix the cycle applies at block width k
    void tune_all(int k) {
        for (int l = 0; l < nlev; ++l) {
            tune_matrix<T>(TK_JACOBI, lv[l].A, k);
            if (l + 1 < nlev) {
                tune_matrix<T>(TK_SPMV, lv[l].A, k);
                tune_matrix<T>(TK_SPMV, lv[l].P, k);
                tune_matrix<T>(TK_SPMV, lv[l].R, k);
                if (lv[l].D.nrows > 0) tune_matrix<T>(TK_SPMV, lv[l].D, k);
            }
        }
    }
    void push(int kid, Step st, dim3 g) {
        st.nbx = (int)g.x;
        st.nby = (int)g.y;
        rec_->steps.push_back(RecStep{rec_->stage, rec_->level, rec_->sub++, kid, st});
    }

    ~Handle() override {
        ws.clear();
        if (cap) (void)hipStreamDestroy(cap);
    }

    // ---------------------------------------------------------------- workspace
    // One workspace per block width k, at most WS_KEEP of them: a new width
    // evicts the least recently used one (after the device has drained, since
    // queued work may still read it) together with the graphs captured on it.
    static constexpr int WS_KEEP = 8;
    uint64_t ws_clock_ = 0;
    Workspace<T>& get_ws(int k) {
        auto it = ws.find(k);
        if (it != ws.end()) {
            it->second->last_use = ++ws_clock_;
            return *it->second;
        }
        if ((int)ws.size() >= WS_KEEP) {
            auto old = ws.begin();
            for (auto j = ws.begin(); j != ws.end(); ++j)
                if (j->second->last_use < old->second->last_use) old = j;
            HIPCHK(hipDeviceSynchronize());
            ws.erase(old);
        }
        std::unique_ptr<Workspace<T>> w = make_ws(k);
        if (tune_) tune_all(k);
        w->last_use = ++ws_clock_;
        auto& ref = *w;
        ws[k] = std::move(w);
        return ref;
    }

    std::unique_ptr<Workspace<T>> make_ws(int k) {
        std::unique_ptr<Workspace<T>> w(new Workspace<T>());
        w->k = k;
        const size_t K = (size_t)k;
        for (int l = 0; l < nlev; ++l) {
            const size_t n = (size_t)lv[l].n;
            w->X.push_back(w->alloc_t(n * K));
            w->Tm.push_back(w->alloc_t(n * K));
            w->B.push_back(w->alloc_t(n * K));
            w->Rr.push_back(w->alloc_t(n * K));
            const bool poly = l + 1 < nlev && (lv[l].pre.kind == AMG_SMOOTH_POLY || lv[l].post.kind == AMG_SMOOTH_POLY);
            w->H0.push_back(poly ? w->alloc_t(n * K) : nullptr);
            w->H1.push_back(poly ? w->alloc_t(n * K) : nullptr);
        }
        const size_t n0 = (size_t)lv[0].n;
        w->X0 = w->alloc_t(n0 * K);
        w->T0b = w->alloc_t(n0 * K);
        w->B0 = w->alloc_t(n0 * K);
        w->Z = w->alloc_t(n0 * K);
        w->cwork = w->alloc_d((size_t)lv[nlev - 1].n * K);
        if (has_outer) {
            w->ob = w->alloc_d(n0 * K);
            w->ox = w->alloc_d(n0 * K);
            w->orr = w->alloc_d(n0 * K);
            w->op[0] = w->alloc_d(n0 * K);
            w->op[1] = w->alloc_d(n0 * K);
            w->oAp = w->alloc_d(n0 * K);
        }
        w->ctl = make_ctl(*w, &w->hflag);
        HIPCHK(hipDeviceSynchronize());
        return w;
    }

    // control block of a solver loop; reducing kernels run on level-0 sized data
    Ctl make_ctl(Workspace<T>& w, int** hflag) {
        const size_t K = (size_t)w.k;
        const size_t nblk = std::max(red_grid(lv[0].n, w.k).x, ru_row_grid(lv[0].n, w.k).x);
        const size_t nint = 4 + 4 * K + 2;
        w.mem.emplace_back(new DevMem(sizeof(int) * nint));
        int* ib = static_cast<int*>(w.mem.back()->p);
        HIPCHK(hipMemset(ib, 0, sizeof(int) * nint));
        Ctl c{};
        c.done = ib;
        c.ruflag = ib + 1;
        c.maxiter = ib + 2;
        c.active = ib + 4;
        c.it = ib + 4 + K;
        c.ticket = ib + 4 + 2 * K;
        c.exitf = ib + 4 + 2 * K + 2;
        c.count = ib + 4 + 3 * K + 2;
        double* db = w.alloc_d(4 + 7 * K + nblk * 3 * K);
        c.tol = db;
        c.normb = db + 4;
        c.normr = db + 4 + K;
        c.gamma = db + 4 + 2 * K;
        c.alpha = db + 4 + 3 * K;
        c.beta = db + 4 + 4 * K;
        c.atol2 = db + 4 + 5 * K;
        c.omega = db + 4 + 6 * K;
        c.part = db + 4 + 7 * K;
        HIPCHK(hipHostMalloc((void**)hflag, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
        **hflag = 0;
        HIPCHK(hipHostGetDevicePointer((void**)&c.hdone, *hflag, 0));
        return c;
    }

    // ---------------------------------------------------------------- launches
    // per-column reduction + solver update after a reducing kernel on level-0 data
    // nb0 > 0: the reducing launch ran nb0 blocks (else the red_grid of level 0)
    void finalize(Workspace<T>& w, FinOp op, int nv, int k, int flag, bool honor_done, hipStream_t s, int nb0 = 0) {
        finalize_ctl(w.ctl, op, nv, k, flag, honor_done, s, nb0);
    }
    void finalize_ctl(const Ctl& ctl, FinOp op, int nv, int k, int flag, bool honor_done, hipStream_t s,
                      int nb0 = 0) {
        const int nb = nb0 > 0 ? nb0 : (int)red_grid(lv[0].n, k).x;
        if (rec_) {
            Step st = step0(lv[0].n, k);
            st.ctl = ctl;
            st.op = (int)op;
            st.nv = nv;
            st.nb = nb;
            st.flag = flag;
            st.honor = honor_done ? 1 : 0;
            push(K_FINALIZE, st, dim3((unsigned)k, 1));
            return;
        }
        k_finalize<<<k, FIN_BLOCK, 0, s>>>(ctl, (int)op, nv, nb, k, flag, honor_done ? 1 : 0);
        HIPCHK(hipGetLastError());
    }

    void spmv_launch(const Sell<T>& M, const T* x, T* y, int k, bool add, const int* done, hipStream_t s) {
        const int kp = kp_of(k);
        if (M.nrows <= 0) return;
        const Plan pl = plan_of(M, k);
        if (rec_) {
            Step st = step0(M.nrows, k);
            st.kp = pl.kp;
            st.M = raw_of(M);
            st.wide = pl.wide;
            st.in0 = x;
            st.out0 = y;
            st.done = done;
            st.subl = pl.subl;
            push(add ? K_SPMV_ADD : K_SPMV, st, row_grid_plan(M.nrows, k, pl));
            return;
        }
        (void)kp;
        if (pl.wide == VMODE) {
            const dim3 g = row_grid_plan(M.nrows, k, pl);
            const bool wide = M.maxw > WVMAX;  // 16-slot chunks
            VKPW_DISPATCH(pl.kp, M.maxw, {
                if (wide) {
                    constexpr int NRW = wide_nru<T, KP>();
                    dim3 gw = g;
                    if (pl.grid == 0) gw.x *= 2;  // one pass per unit: twice the units
                    if (add) k_spmv_v<T, KP, true, 16, 64, NRW><<<gw, BLOCK, 0, s>>>(M, x, y, k, done);
                    else k_spmv_v<T, KP, false, 16, 64, NRW><<<gw, BLOCK, 0, s>>>(M, x, y, k, done);
                } else {
                    if (add) k_spmv_v<T, KP, true, 8, WC><<<g, BLOCK, 0, s>>>(M, x, y, k, done);
                    else k_spmv_v<T, KP, false, 8, WC><<<g, BLOCK, 0, s>>>(M, x, y, k, done);
                }
            });
        } else {
            KP_DISPATCH(pl.kp, WIDE_IF(pl.wide, {
                const dim3 g = row_grid_plan(M.nrows, k, pl);
                if (add) k_spmv<T, KP, WD, true><<<g, BLOCK, 0, s>>>(M, x, y, k, pl.subl, done);
                else k_spmv<T, KP, WD, false><<<g, BLOCK, 0, s>>>(M, x, y, k, pl.subl, done);
            }));
        }
        HIPCHK(hipGetLastError());
    }

    void jacobi_launch(int l, const T* x, const T* b, T* xo, T w, int k, const int* mask, const int* done,
                       hipStream_t s) {
        const LevelT<T>& L = lv[l];
        const int kp = kp_of(k);
        const Plan pl = plan_of(L.A, k, TK_JACOBI);
        (void)kp;
        if (rec_) {
            Step st = step0(L.n, k);
            st.kp = pl.kp;
            st.subl = pl.subl;
            st.M = raw_of(L.A);
            st.wide = pl.wide;
            st.dinv = L.dinv;
            st.in0 = x;
            st.in1 = b;
            st.out0 = xo;
            st.w = (double)w;
            st.mask = mask;
            st.done = done;
            push(K_JACOBI, st, row_grid_plan(L.n, k, pl));
            return;
        }
        if (pl.wide == VMODE)
            VKPW_DISPATCH(pl.kp, L.A.maxw, { k_jacobi_v<T, KP, WC><<<row_grid_plan(L.n, k, pl), BLOCK, 0, s>>>(L.A, L.dinv, x, b, xo, w, k, mask, done); })
        else
            KP_DISPATCH(pl.kp, WIDE_IF(pl.wide, { k_jacobi<T, KP, WD><<<row_grid_plan(L.n, k, pl), BLOCK, 0, s>>>(L.A, L.dinv, x, b, xo, w, k, pl.subl, mask, done); }));
        HIPCHK(hipGetLastError());
    }

    // Iterative refinement: the cycle's last top-level Jacobi sweep adds its result
    // straight into the float64 iterate instead of storing the correction for a
    // separate x += e (vector mode, no column mask).  The refinement loop arms rf_
    // around the cycle; post_sweeps consumes it.
    struct RefineFuse {
        bool on = false, used = false;
        double* x = nullptr;
        const int* active = nullptr;
    };
    RefineFuse rf_;
    struct RefineFuseArm {
        RefineFuse& r;
        RefineFuseArm(RefineFuse& r_, const RefineFuse& v) : r(r_) { r = v; }
        ~RefineFuseArm() { r = RefineFuse{}; }
    };
    bool jacobi_acc_launch(const T* x, const T* b, T w, int k, const int* done, hipStream_t s) {
        if (rec_) return false;
        const LevelT<T>& L = lv[0];
        const Plan pl = plan_of(L.A, k, TK_JACOBI);
        if (pl.wide != VMODE) return false;
        VKPW_DISPATCH(pl.kp, L.A.maxw, {
            k_jacobi_acc_v<T, KP, WC><<<row_grid_plan(L.n, k, pl), BLOCK, 0, s>>>(L.A, L.dinv, x, b, rf_.x, w, k,
                                                                              rf_.active, done);
        });
        HIPCHK(hipGetLastError());
        return true;
    }

    // PCG: the preconditioner's last top-level Jacobi sweep also adds the
    // sums r.z and z.Ap of the following beta (vector mode only).  emit_pcg_iter
    // arms gf_ around the cycle; post_sweeps consumes it.
    struct GammaFuse {
        bool on = false, used = false;
        const void* r = nullptr;   // CG residual and A p: float64, or T with tv_t
        const void* ap = nullptr;
        bool tv_t = false;
        int flexible = 0;
        int nb = 0;  // blocks of the reducing launch (0: red_grid)
        Ctl ctl{};
    };
    GammaFuse gf_;
    struct GammaFuseArm {
        GammaFuse& g;
        GammaFuseArm(GammaFuse& g_, const GammaFuse& v) : g(g_) { g = v; }
        ~GammaFuseArm() { g = GammaFuse{}; }
    };

    // z = last sweep of x at level 0 with the gamma sums; false if the level's
    // sweep does not run in vector mode (the caller then sweeps and reduces apart)
    bool jacobi_gamma_launch(const T* x, T* z, T w, int k, hipStream_t s) {
        const LevelT<T>& L = lv[0];
        if (red_mode(L.A, k) != VMODE) return false;
        const int kp = kp_of(k);
        const dim3 g((unsigned)(gf_.nb > 0 ? gf_.nb : (int)red_grid(L.n, k).x), (unsigned)((k + kp - 1) / kp));
        if (rec_) {
            if (gf_.tv_t) return false;
            Step st = step0(L.n, k);
            st.M = raw_of(L.A);
            st.wide = VMODE;
            st.dinv = L.dinv;
            st.in0 = x;
            st.in1 = gf_.r;
            st.in2 = gf_.ap;
            st.out0 = z;
            st.w = (double)w;
            st.flag = gf_.flexible;
            st.ctl = gf_.ctl;
            push(K_JACOBI_GAMMA, st, g);
            return true;
        }
        if (gf_.tv_t) {
            ev_mark(EV_RU_GAMMA, true, s);
            VKPW_DISPATCH(kp, L.A.maxw, {
                k_jacobi_gamma_v<T, KP, T, WC><<<g, BLOCK, 0, s>>>(L.A, L.dinv, x, static_cast<const T*>(gf_.r),
                                                               static_cast<const T*>(gf_.ap), z, w, k, gf_.flexible,
                                                               gf_.ctl);
            })
            ev_mark(EV_RU_GAMMA, false, s);
        } else
            VKPW_DISPATCH(kp, L.A.maxw, {
                k_jacobi_gamma_v<T, KP, double, WC><<<g, BLOCK, 0, s>>>(L.A, L.dinv, x, static_cast<const double*>(gf_.r),
                                                            static_cast<const double*>(gf_.ap), z, w, k,
                                                            gf_.flexible, gf_.ctl);
            })
        HIPCHK(hipGetLastError());
        return true;
    }

    void jacobi_zero_launch(int l, const T* b, T* xo, T w, int k, const int* done, hipStream_t s) {
        const LevelT<T>& L = lv[l];
        const int kp = kp_of(k);
        const dim3 g = row_grid(L.n, k);
        if (rec_) {
            Step st = step0(L.n, k);
            st.dinv = L.dinv;
            st.in0 = b;
            st.out0 = xo;
            st.w = (double)w;
            st.done = done;
            push(K_JACOBI_ZERO, st, g);
            return;
        }
        KP_DISPATCH(kp, { k_jacobi_zero<T, KP><<<g, BLOCK, 0, s>>>(L.dinv, b, xo, w, L.n, k, done); });
        HIPCHK(hipGetLastError());
    }

    void residual_launch(int l, const T* x, const T* b, T* r, int k, const int* done, hipStream_t s) {
        const LevelT<T>& L = lv[l];
        const int kp = kp_of(k);
        const Plan pl = plan_of(L.A, k, TK_JACOBI);
        (void)kp;
        if (rec_) {
            Step st = step0(L.n, k);
            st.kp = pl.kp;
            st.subl = pl.subl;
            st.M = raw_of(L.A);
            st.wide = pl.wide;
            st.in0 = x;
            st.in1 = b;
            st.out0 = r;
            st.done = done;
            push(K_RESIDUAL, st, row_grid_plan(L.n, k, pl));
            return;
        }
        if (pl.wide == VMODE)
            VKPW_DISPATCH(pl.kp, L.A.maxw, { k_residual_v<T, KP, WC><<<row_grid_plan(L.n, k, pl), BLOCK, 0, s>>>(L.A, x, b, r, k, done); })
        else
            KP_DISPATCH(pl.kp, WIDE_IF(pl.wide, { k_residual<T, KP, WD><<<row_grid_plan(L.n, k, pl), BLOCK, 0, s>>>(L.A, x, b, r, k, pl.subl, done); }));
        HIPCHK(hipGetLastError());
    }

    void residual_zero_launch(int l, const T* b, T* x1, T* r, T w, int k, const int* done, hipStream_t s) {
        const LevelT<T>& L = lv[l];
        const int kp = kp_of(k);
        const Plan pl = plan_of(L.A, k, TK_JACOBI);
        (void)kp;
        if (rec_) {
            Step st = step0(L.n, k);
            st.kp = pl.kp;
            st.subl = pl.subl;
            st.M = raw_of(L.A);
            st.wide = pl.wide;
            st.dinv = L.dinv;
            st.in0 = b;
            st.out0 = x1;
            st.out1 = r;
            st.w = (double)w;
            st.done = done;
            push(K_RESIDUAL_ZERO, st, row_grid_plan(L.n, k, pl));
            return;
        }
        if (pl.wide == VMODE)
            VKPW_DISPATCH(pl.kp, L.A.maxw, { k_residual_zero_v<T, KP, WC><<<row_grid_plan(L.n, k, pl), BLOCK, 0, s>>>(L.A, L.dinv, b, x1, r, w, k, done); })
        else
            KP_DISPATCH(pl.kp, WIDE_IF(pl.wide, { k_residual_zero<T, KP, WD><<<row_grid_plan(L.n, k, pl), BLOCK, 0, s>>>(L.A, L.dinv, b, x1, r, w, k, pl.subl, done); }));
        HIPCHK(hipGetLastError());
    }

    void prolong_x1_launch(int l, const T* b, const T* e, T* x, T w_pre, int k, const int* done, hipStream_t s) {
        const LevelT<T>& L = lv[l];
        const int kp = kp_of(k);
        const Plan pl = plan_of(L.P, k);
        (void)kp;
        if (rec_) {
            Step st = step0(L.n, k);
            st.kp = pl.kp;
            st.subl = pl.subl;
            st.M = raw_of(L.P);
            st.wide = pl.wide;
            st.dinv = L.dinv;
            st.in0 = b;
            st.in1 = e;
            st.out0 = x;
            st.w = (double)w_pre;
            st.done = done;
            push(K_PROLONG_X1, st, row_grid_plan(L.n, k, pl));
            return;
        }
        if (pl.wide == VMODE)
            VKPW_DISPATCH(pl.kp, L.P.maxw, { k_prolong_x1_v<T, KP, WC><<<row_grid_plan(L.n, k, pl), BLOCK, 0, s>>>(L.P, L.dinv, b, e, x, w_pre, k, done); })
        else
            KP_DISPATCH(pl.kp, WIDE_IF(pl.wide, {
                k_prolong_x1<T, KP, WD><<<row_grid_plan(L.n, k, pl), BLOCK, 0, s>>>(L.P, L.dinv, b, e, x, w_pre, k, pl.subl, done);
            }));
        HIPCHK(hipGetLastError());
    }

    template <typename TD, typename TS>
    void copy_launch(const TS* src, TD* dst, int n, int k, const int* mask, const int* done, hipStream_t s) {
        const int kp = kp_of(k);
        const dim3 g = row_grid(n, k);
        if (rec_) {
            if constexpr (std::is_same<TD, T>::value && std::is_same<TS, T>::value) {
                Step st = step0(n, k);
                st.in0 = src;
                st.out0 = dst;
                st.mask = mask;
                st.done = done;
                push(K_COPY, st, g);
                return;
            } else {
                throw Error{AMG_E_UNSUPPORTED, "mixed-type copy in a batched sequence"};
            }
        }
        KP_DISPATCH(kp, { k_copy<TD, TS, KP><<<g, BLOCK, 0, s>>>(src, dst, n, k, mask, done); });
        HIPCHK(hipGetLastError());
    }

    // coarse solve on the coarsest level: x = solve(b), x may not alias b
    void coarse_launch(const T* b, T* x, T* tmp, double* cwork, int k, const int* done, hipStream_t s) {
        const int lc = nlev - 1;
        const int n = lv[lc].n;
        const unsigned gk = (unsigned)((k + BLOCK - 1) / BLOCK);
        // (recording: the Jacobi coarse solver's sweeps below record themselves)
        if (rec_ && cd.kind != AMG_COARSE_JACOBI) {
            Step st = step0(n, k);
            st.in0 = b;
            st.out0 = x;
            st.in2 = cf0;
            st.done = done;
            if (cd.kind == AMG_COARSE_PINV) {
                push(cf64 ? K_COARSE_PINV64 : K_COARSE_PINV, st, dim3((unsigned)((n * k + BLOCK - 1) / BLOCK), 1));
            } else if (cd.kind == AMG_COARSE_LU || cd.kind == AMG_COARSE_QR) {
                // one thread per column, work vector as in the launches below
                st.piv = cpiv;
                st.in1 = cf1;
                st.out1 = cf64 ? static_cast<void*>(cwork) : static_cast<void*>(x);
                const int kid = cd.kind == AMG_COARSE_LU ? (cf64 ? K_COARSE_LU64 : K_COARSE_LU)
                                                         : (cf64 ? K_COARSE_QR64 : K_COARSE_QR);
                push(kid, st, dim3(gk, 1));
            } else {
                throw Error{AMG_E_ARG, "unknown coarse solver"};
            }
            return;
        }
        switch (cd.kind) {
            case AMG_COARSE_PINV: {
                const unsigned g = (unsigned)((n * k + BLOCK - 1) / BLOCK);
                if (cf64) k_coarse_pinv<T, double><<<g, BLOCK, 0, s>>>((const double*)cf0, b, x, n, k, done);
                else k_coarse_pinv<T, T><<<g, BLOCK, 0, s>>>((const T*)cf0, b, x, n, k, done);
                break;
            }
            case AMG_COARSE_LU:
                if (cf64) k_coarse_lu<T, double><<<gk, BLOCK, 0, s>>>((const double*)cf0, cpiv, b, x, cwork, n, k, done);
                else k_coarse_lu<T, T><<<gk, BLOCK, 0, s>>>((const T*)cf0, cpiv, b, x, x, n, k, done);
                break;
            case AMG_COARSE_QR:
                if (cf64)
                    k_coarse_qr<T, double><<<gk, BLOCK, 0, s>>>((const double*)cf0, (const double*)cf1, b, x, cwork,
                                                               n, k, done);
                else
                    k_coarse_qr<T, T><<<gk, BLOCK, 0, s>>>((const T*)cf0, (const T*)cf1, b, x, x, n, k, done);
                break;
            case AMG_COARSE_JACOBI: {
                const T w = static_cast<T>(cd.omega);
                if (cd.iters <= 0) {
                    // zero sweeps: x stays at the zero guess
                    jacobi_zero_launch(lc, b, x, T(0), k, done, s);
                    break;
                }
                // sweeps ping-pong between tmp and x so that the last lands in x
                T* bufs[2] = {x, tmp};
                int cur = (cd.iters % 2 == 1) ? 0 : 1;
                jacobi_zero_launch(lc, b, bufs[cur], w, k, done, s);
                for (int it = 1; it < cd.iters; ++it) {
                    jacobi_launch(lc, bufs[cur], b, bufs[1 - cur], w, k, nullptr, done, s);
                    cur = 1 - cur;
                }
                break;
            }
            default:
                throw Error{AMG_E_ARG, "unknown coarse solver"};
        }
        HIPCHK(hipGetLastError());
    }

    // ---------------------------------------------------------------- cycle emission
    //
    // Level-0 buffer pool: `xin` (read only until the final store), tA, tB.
    // Levels >= 1 ping-pong between X[l] and Tm[l].
    struct Pool {
        T* bufs[3];
        int nb;
        int cur;
    };

    T* other(Pool& p) {
        // pick a buffer different from the current one, never index 0 when nb == 3
        if (p.nb == 3) return p.bufs[p.cur == 1 ? 2 : 1];
        return p.bufs[1 - p.cur];
    }
    void set_cur(Pool& p, T* b) {
        for (int i = 0; i < p.nb; ++i)
            if (p.bufs[i] == b) p.cur = i;
    }

    // smoothing sweeps on level l; returns the buffer holding the result
    // one polynomial-smoother sweep on level l: xo = sweep(x, b) (x == nullptr: zero guess)
    void poly_sweep(int l, const Smoother& sm, const T* x, const T* b, T* xo, int k, const int* mask,
                    const int* done, hipStream_t s, Workspace<T>& w) {
        const LevelT<T>& L = lv[l];
        if (rec_) {
            // the launches of launch_poly_sweep, recorded
            const Plan p0 = plan_of(L.A, k, TK_JACOBI);
            const Plan pl = p0.wide == VMODE ? default_plan(L.A.maxw, k, L.A.nrows, 0) : p0;
            const int nc = (int)sm.coef.size(), n = L.A.nrows;
            T* hb[2] = {w.H0[l], w.H1[l]};
            const T* rr = b;
            Step st = step0(n, k);
            st.mask = mask;
            st.done = done;
            st.w = sm.coef[0];
            if (x == nullptr) {
                st.in0 = b;
                st.out0 = nc == 1 ? xo : hb[0];
                push(K_POLY_FIRST_ZERO, st, row_grid(n, k));
            } else {
                rr = w.Rr[l];
                st.M = raw_of(L.A);
                st.kp = pl.kp;
                st.wide = pl.wide;
                st.subl = pl.subl;
                st.in0 = x;
                st.in1 = b;
                st.out0 = hb[0];
                st.out1 = w.Rr[l];
                st.out2 = nc == 1 ? xo : nullptr;
                push(K_POLY_FIRST, st, row_grid_plan(n, k, pl));
            }
            int cur = 0;
            for (int i = 1; i < nc; ++i) {
                const bool last = i == nc - 1;
                Step q = step0(n, k);
                q.M = raw_of(L.A);
                q.kp = pl.kp;
                q.wide = pl.wide;
                q.subl = pl.subl;
                q.in0 = hb[cur];
                q.in1 = rr;
                q.in2 = x;
                q.out0 = last ? xo : hb[1 - cur];
                q.w = sm.coef[i];
                q.flag = last ? 1 : 0;
                q.mask = mask;
                q.done = done;
                push(K_POLY_STEP, q, row_grid_plan(n, k, pl));
                cur = 1 - cur;
            }
            return;
        }
        launch_poly_sweep<T>(L.A, plan_of(L.A, k, TK_JACOBI), sm.coef.data(), (int)sm.coef.size(), x, b, xo,
                             w.Rr[l], w.H0[l], w.H1[l], k, mask, done, s);
    }

    void smooth(int l, const Smoother& sm, Pool& pool, bool& xzero, const T* b, int k, const int* done,
                hipStream_t s, Workspace<T>& w) {
        if (sm.kind == AMG_SMOOTH_POLY) {
            for (int it = 0; it < sm.iters; ++it) {
                T* dst = other(pool);
                poly_sweep(l, sm, xzero ? nullptr : pool.bufs[pool.cur], b, dst, k, nullptr, done, s, w);
                set_cur(pool, dst);
                xzero = false;
            }
            return;
        }
        if (sm.kind != AMG_SMOOTH_JACOBI || sm.iters <= 0) return;
        const T om = static_cast<T>(sm.omega);
        int it = 0;
        if (xzero) {
            T* dst = other(pool);
            jacobi_zero_launch(l, b, dst, om, k, done, s);
            set_cur(pool, dst);
            xzero = false;
            it = 1;
        }
        for (; it < sm.iters; ++it) {
            T* dst = other(pool);
            jacobi_launch(l, pool.bufs[pool.cur], b, dst, om, k, nullptr, done, s);
            set_cur(pool, dst);
        }
    }

    // `sweeps` post-smoothing sweeps; with final_dst the last sweep (or a copy
    // when there is none) writes final_dst under `mask`.
    void post_sweeps(int l, const Smoother& sm, Pool& pool, const T* b, int k, T* final_dst, const int* mask,
                     const int* done, hipStream_t s, Workspace<T>& w) {
        const bool any = (sm.kind == AMG_SMOOTH_JACOBI || sm.kind == AMG_SMOOTH_POLY) && sm.iters >= 1;
        const int sweeps = any ? sm.iters : 0;
        const T omega = static_cast<T>(sm.omega);
        for (int i = 0; i < sweeps; ++i) {
            const bool last = (i == sweeps - 1) && final_dst != nullptr;
            T* dst = last ? final_dst : other(pool);
            if (sm.kind == AMG_SMOOTH_POLY)
                poly_sweep(l, sm, pool.bufs[pool.cur], b, dst, k, last ? mask : nullptr, done, s, w);
            else if (l == 0 && last && mask == nullptr && gf_.on && !gf_.used &&
                     jacobi_gamma_launch(pool.bufs[pool.cur], dst, omega, k, s))
                gf_.used = true;
            else if (l == 0 && last && mask == nullptr && rf_.on && !rf_.used &&
                     jacobi_acc_launch(pool.bufs[pool.cur], b, omega, k, done, s))
                rf_.used = true;
            else
                jacobi_launch(l, pool.bufs[pool.cur], b, dst, omega, k, last ? mask : nullptr, done, s);
            if (!last) set_cur(pool, dst);
        }
        if (sweeps == 0 && final_dst != nullptr && (pool.bufs[pool.cur] != final_dst || mask != nullptr))
            copy_launch<T, T>(pool.bufs[pool.cur], final_dst, lv[l].n, k, mask, done, s);
    }

    // One visit of level l (< nlev - 1).  On return pool.cur holds x_l, or, with
    // final_dst (top level), x_l has been written to final_dst under `mask`.
    void rec(int l, Pool& pool, bool xzero, const T* b, char ct, int cpl, int k, const int* done, hipStream_t s,
             Workspace<T>& w, T* final_dst = nullptr, const int* mask = nullptr) {
        LevelT<T>& L = lv[l];
        const TailPlan* tp = final_dst == nullptr ? tail_for(l, ct, xzero, cpl) : nullptr;
        if (tp != nullptr) {
            // this visit and every visit below it in one launch
            T* dst = other(pool);
            tail_launch(*tp, ct, cpl, b, xzero ? nullptr : pool.bufs[pool.cur], dst, k, done, s);
            set_cur(pool, dst);
            return;
        }
        const bool last_pair = (l == nlev - 2);
        const Smoother& pre = L.pre;
        const Smoother& post = L.post;
        bool x1v = false;  // x_l = w_pre D^-1 b, never stored (fused down-sweep)
        tag(ST_DOWN, l);
        // ---- pre-smoothing + residual + restriction
        if (xzero && pre.kind == AMG_SMOOTH_JACOBI && pre.iters == 1 && fuse && L.D.nrows > 0) {
            spmv_launch(L.D, b, w.B[l + 1], k, false, done, s);
            x1v = true;
            xzero = false;
        } else {
            const T* rsrc = nullptr;
            if (xzero && pre.kind == AMG_SMOOTH_JACOBI && pre.iters == 1) {
                T* dst = other(pool);
                residual_zero_launch(l, b, dst, w.Rr[l], static_cast<T>(pre.omega), k, done, s);
                set_cur(pool, dst);
                xzero = false;
                rsrc = w.Rr[l];
            } else {
                smooth(l, pre, pool, xzero, b, k, done, s, w);
                if (xzero) {
                    rsrc = b;  // x = 0: r = b - A 0 = b
                } else {
                    residual_launch(l, pool.bufs[pool.cur], b, w.Rr[l], k, done, s);
                    rsrc = w.Rr[l];
                }
            }
            spmv_launch(L.R, rsrc, w.B[l + 1], k, false, done, s);
        }
        // ---- coarse-grid correction
        Pool cp{{w.X[l + 1], w.Tm[l + 1], nullptr}, 2, 0};
        if (last_pair) {
            tag(ST_COARSE, 0);
            coarse_launch(w.B[l + 1], w.X[l + 1], w.Tm[l + 1], w.cwork, k, done, s);
            cp.cur = 0;
        } else if (ct == 'V') {
            rec(l + 1, cp, true, w.B[l + 1], 'V', cpl, k, done, s, w);
        } else if (ct == 'W') {
            rec(l + 1, cp, true, w.B[l + 1], 'W', cpl, k, done, s, w);
            rec(l + 1, cp, false, w.B[l + 1], 'W', cpl, k, done, s, w);
        } else if (ct == 'F') {
            rec(l + 1, cp, true, w.B[l + 1], 'F', cpl, k, done, s, w);
            for (int i = 0; i < cpl; ++i) rec(l + 1, cp, false, w.B[l + 1], 'V', cpl, k, done, s, w);
        } else {
            // 'T', the F-cycle's transpose: the same coarse visits in reverse order
            // (cpl V-cycles, the first from zero, then 'T' from their result)
            bool z = true;
            for (int i = 0; i < cpl; ++i, z = false) rec(l + 1, cp, z, w.B[l + 1], 'V', cpl, k, done, s, w);
            rec(l + 1, cp, z, w.B[l + 1], 'T', cpl, k, done, s, w);
        }
        const T* e = cp.bufs[cp.cur];
        // ---- prolongation + correction + post-smoothing
        tag(ST_UP, l);
        if (x1v) {
            T* dst = other(pool);
            prolong_x1_launch(l, b, e, dst, static_cast<T>(pre.omega), k, done, s);
            set_cur(pool, dst);
        } else if (xzero) {
            // x_l = 0 + P e: write into a fresh buffer without reading x
            T* dst = other(pool);
            spmv_launch(L.P, e, dst, k, false, done, s);
            set_cur(pool, dst);
        } else {
            T* cur = pool.bufs[pool.cur];
            if (pool.nb == 3 && pool.cur == 0) {
                // never modify the caller's iterate in place
                T* dst = other(pool);
                copy_launch<T, T>(cur, dst, L.n, k, nullptr, done, s);
                set_cur(pool, dst);
                cur = dst;
            }
            spmv_launch(L.P, e, cur, k, true, done, s);
        }
        post_sweeps(l, post, pool, b, k, final_dst, mask, done, s, w);
    }

    // Emit one cycle: xout = cycle(xin, b) (xin == nullptr: zero guess).
    // mask (nullable): only active columns of xout are written.
    void emit_cycle(char ct, int cpl, const T* b, const T* xin, T* xout, int k, const int* mask, const int* done,
                    hipStream_t s, Workspace<T>& w) {
        if (nlev == 1) {
            tag(ST_COARSE, 0);
            T* dst = (mask != nullptr) ? w.Tm[0] : xout;
            coarse_launch(b, dst, w.T0b, w.cwork, k, done, s);
            if (dst != xout) copy_launch<T, T>(dst, xout, lv[0].n, k, mask, done, s);
            return;
        }
        const bool xzero = (xin == nullptr);
        // pool slot 0 holds the caller's iterate, which is only read; with a zero
        // guess it is a placeholder that is never read
        Pool pool{{xzero ? xout : const_cast<T*>(xin), w.Tm[0], w.T0b}, 3, 0};
        rec(0, pool, xzero, b, ct, cpl, k, done, s, w, xout, mask);
    }

    // ---------------------------------------------------------------- graphs
    template <typename F>
    hipGraphExec_t graph(Workspace<T>& w, std::tuple<int, char, int, int, int> key, F&& emit) {
        auto it = w.graphs.find(key);
        if (it != w.graphs.end()) return it->second;
        if (!cap) HIPCHK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        try {
            emit(cap);
        } catch (...) {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(cap, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            throw;
        }
        HIPCHK(hipStreamEndCapture(cap, &g));
        hipGraphExec_t ex = nullptr;
        HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        w.raw.push_back(g);
        w.graphs[key] = ex;
        return ex;
    }

    static constexpr int GRAPH_ITERS = 4;

    // ---------------------------------------------------------------- entry points
    void cycle(char ct, int cpl, const void* b, void* x, int k, int x_zero, hipStream_t s) override {
        Workspace<T>& w = get_ws(k);
        const T* bb = static_cast<const T*>(b);
        T* xx = static_cast<T*>(x);
        if (x_zero) {
            emit_cycle(ct, cpl, bb, nullptr, xx, k, nullptr, nullptr, s, w);
        } else {
            // x <- cycle(x, b): read the iterate from a copy so the output can alias it
            HIPCHK(hipMemcpyAsync(w.X0, xx, sizeof(T) * (size_t)lv[0].n * k, hipMemcpyDeviceToDevice, s));
            emit_cycle(ct, cpl, bb, w.X0, xx, k, nullptr, nullptr, s, w);
        }
    }

    // ---------------------------------------------------------------- allocation-free cycle
    // The cycle of `cycle` with the vectors it uses (per level x, temp, rhs,
    // residual and polynomial terms; the level-0 pool and input copy; the coarse
    // float64 work vector) carved from caller-provided device memory, each at a
    // 256-byte boundary, into a view whose host arrays were sized at creation:
    // no device or host allocation, no memset or memcpy, no autotuning (plans
    // come from the tuner's cache or the fixed shape).  No vector needs to start
    // at zero: zero-guess kernels write before anything reads.
    Workspace<T> nw_;
    static size_t carve_size(size_t elems, size_t esz) { return (elems * esz + 255) & ~size_t(255); }
    size_t noalloc_bytes(int k) const {
        const size_t K = (size_t)k;
        size_t tot = 0;
        for (int l = 0; l < nlev; ++l) {
            const size_t n = (size_t)lv[l].n;
            const bool poly = l + 1 < nlev && (lv[l].pre.kind == AMG_SMOOTH_POLY || lv[l].post.kind == AMG_SMOOTH_POLY);
            tot += (poly ? 6 : 4) * carve_size(n * K, sizeof(T));
        }
        tot += 2 * carve_size((size_t)lv[0].n * K, sizeof(T));              // X0, T0b
        tot += carve_size((size_t)lv[nlev - 1].n * K, sizeof(double));      // cwork
        return tot;
    }
    void cycle_noalloc(char ct, int cpl, const void* b, void* x, int k, int x_zero, void* base, size_t bytes,
                       hipStream_t s) override {
        const size_t need = noalloc_bytes(k);
        if (need > bytes) {
            char msg[200];
            std::snprintf(msg, sizeof(msg),
                          "allocation-free cycle needs %zu bytes of scratch for k = %d, the limit is %zu bytes", need,
                          k, bytes);
            throw Error{AMG_E_ARG, msg};
        }
        Workspace<T>& w = nw_;
        char* p = static_cast<char*>(base);
        auto take_t = [&](size_t elems) {
            T* q = reinterpret_cast<T*>(p);
            p += carve_size(elems, sizeof(T));
            return q;
        };
        const size_t K = (size_t)k;
        w.k = k;
        for (int l = 0; l < nlev; ++l) {
            const size_t n = (size_t)lv[l].n;
            w.X[l] = take_t(n * K);
            w.Tm[l] = take_t(n * K);
            w.B[l] = take_t(n * K);
            w.Rr[l] = take_t(n * K);
            const bool poly = l + 1 < nlev && (lv[l].pre.kind == AMG_SMOOTH_POLY || lv[l].post.kind == AMG_SMOOTH_POLY);
            w.H0[l] = poly ? take_t(n * K) : nullptr;
            w.H1[l] = poly ? take_t(n * K) : nullptr;
        }
        w.X0 = take_t((size_t)lv[0].n * K);
        w.T0b = take_t((size_t)lv[0].n * K);
        w.cwork = reinterpret_cast<double*>(p);
        const T* bb = static_cast<const T*>(b);
        T* xx = static_cast<T*>(x);
        if (x_zero) {
            emit_cycle(ct, cpl, bb, nullptr, xx, k, nullptr, nullptr, s, w);
        } else {
            // the iterate is read from a copy (a copy kernel) so that the output can alias it
            copy_launch<T, T>(xx, w.X0, lv[0].n, k, nullptr, nullptr, s);
            emit_cycle(ct, cpl, bb, w.X0, xx, k, nullptr, nullptr, s, w);
        }
    }

    void solve(char ct, int cpl, const void* b, void* x, int k, double tol, int maxiter, int refine, int* iters,
               double* relres, hipStream_t s) override {
        Workspace<T>& w = get_ws(k);
        const size_t n0 = (size_t)lv[0].n;
        k_set_params<<<1, 1, 0, s>>>(w.ctl, tol, 0.0, maxiter, 0);
        HIPCHK(hipGetLastError());
        const int kp = kp_of(k);
        if (!refine) {
            HIPCHK(hipMemcpyAsync(w.B0, b, sizeof(T) * n0 * k, hipMemcpyDeviceToDevice, s));
            const dim3 g = red_grid((int)n0, k);
            KP_DISPATCH(kp, { k_solve_init<T, KP><<<g, BLOCK, 0, s>>>(w.B0, w.X0, (int)n0, k, w.ctl); });
            HIPCHK(hipGetLastError());
            finalize(w, FIN_SOLVE_INIT, 1, k, 0, false, s);
            auto key = std::make_tuple(1, ct, cpl, GRAPH_ITERS, 0);
            hipGraphExec_t ex = graph(w, key, [&](hipStream_t cs) {
                for (int i = 0; i < GRAPH_ITERS; ++i) {
                    emit_cycle(ct, cpl, w.B0, w.X0, w.X0, k, w.ctl.active, w.ctl.done, cs, w);
                    const dim3 gr = red_grid((int)n0, k);
                    if (red_mode(lv[0].A, k) == VMODE)
                        VKPW_DISPATCH(kp, lv[0].A.maxw, { k_resnorm_v<T, T, KP, WC><<<gr, BLOCK, 0, cs>>>(lv[0].A, w.X0, w.B0, (T*)nullptr, k, w.ctl); })
                    else
                        KP_DISPATCH(kp, WIDE_DISPATCH(lv[0].A, {
                            k_resnorm<T, T, KP, WD><<<gr, BLOCK, 0, cs>>>(lv[0].A, w.X0, w.B0, (T*)nullptr, k, w.ctl);
                        }));
                    HIPCHK(hipGetLastError());
                    finalize(w, FIN_RESNORM, 1, k, 0, true, cs);
                }
            });
            run_until_done(w.hflag, ex, (maxiter + GRAPH_ITERS - 1) / GRAPH_ITERS + 1, s);
            HIPCHK(hipMemcpyAsync(x, w.X0, sizeof(T) * n0 * k, hipMemcpyDeviceToDevice, s));
        } else {
            if (!has_outer) throw Error{AMG_E_ARG, "refine=1 needs the float64 outer operator"};
            // iterative refinement: x64 += cycle(0, (T) r64), r64 = b - A64 x64
            HIPCHK(hipMemcpyAsync(w.ob, b, sizeof(double) * n0 * k, hipMemcpyDeviceToDevice, s));
            const dim3 g = red_grid((int)n0, k);
            KP_DISPATCH(kp, {
                k_solve_init<double, KP><<<g, BLOCK, 0, s>>>(w.ob, w.ox, (int)n0, k, w.ctl);
            });
            HIPCHK(hipGetLastError());
            finalize(w, FIN_SOLVE_INIT, 1, k, 0, false, s);
            copy_launch<T, double>(w.ob, w.B0, (int)n0, k, nullptr, nullptr, s);
            auto key = std::make_tuple(2, ct, cpl, GRAPH_ITERS, 0);
            hipGraphExec_t ex = graph(w, key, [&](hipStream_t cs) {
                for (int i = 0; i < GRAPH_ITERS; ++i) {
                    bool added = false;
                    {
                        RefineFuse arm;
                        arm.on = fuse;
                        arm.x = w.ox;
                        arm.active = w.ctl.active;
                        RefineFuseArm guard(rf_, arm);
                        emit_cycle(ct, cpl, w.B0, nullptr, w.Z, k, nullptr, w.ctl.done, cs, w);
                        added = rf_.used;
                    }
                    if (!added)
                        KP_DISPATCH(kp, {
                            k_add_masked<double, T, KP><<<row_grid((int)n0, k), BLOCK, 0, cs>>>(
                                w.ox, w.Z, (int)n0, k, w.ctl.active, w.ctl.done);
                        });
                    const dim3 gr = red_grid((int)n0, k);
                    if (red_mode(outer, k) == VMODE)
                        VKPW_DISPATCH(kp, outer.maxw, { k_resnorm_v<double, T, KP, WC><<<gr, BLOCK, 0, cs>>>(outer, w.ox, w.ob, w.B0, k, w.ctl); })
                    else
                        KP_DISPATCH(kp, WIDE_DISPATCH(outer, {
                            k_resnorm<double, T, KP, WD><<<gr, BLOCK, 0, cs>>>(outer, w.ox, w.ob, w.B0, k, w.ctl);
                        }));
                    HIPCHK(hipGetLastError());
                    finalize(w, FIN_RESNORM, 1, k, 0, true, cs);
                }
            });
            run_until_done(w.hflag, ex, (maxiter + GRAPH_ITERS - 1) / GRAPH_ITERS + 1, s);
            HIPCHK(hipMemcpyAsync(x, w.ox, sizeof(double) * n0 * k, hipMemcpyDeviceToDevice, s));
        }
        report(w, k, iters, relres, s);
    }

    void report(Workspace<T>& w, int k, int* iters, double* relres, hipStream_t s) {
        if (!iters && !relres) return;
        std::vector<int> it(k);
        std::vector<double> nr(k), nb(k);
        HIPCHK(hipMemcpyAsync(it.data(), w.ctl.it, sizeof(int) * k, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(nr.data(), w.ctl.normr, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(nb.data(), w.ctl.normb, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int c = 0; c < k; ++c) {
            if (iters) iters[c] = it[c];
            if (relres) relres[c] = nr[c] / nb[c];
        }
    }

    // ---------------------------------------------------------------- preconditioned CG
    // cycle input: r itself (float64 handle) or its float32 copy
    const T* pcg_cyc_in(Workspace<T>& w) const {
        if constexpr (std::is_same<T, double>::value) return w.orr;
        else return w.B0;
    }
    T* pcg_rc(Workspace<T>& w) const {
        if constexpr (std::is_same<T, double>::value) return nullptr;
        else return w.B0;
    }

    // x = 0, r = b, rc = (T) b, atol2, with b read from w.ob, or (eager launches
    // only) from the caller's b, copied into w.ob on the way
    void emit_pcg_init(Workspace<T>& w, int k, hipStream_t s, const double* b = nullptr) {
        const int n0 = lv[0].n;
        const dim3 gr = red_grid(n0, k);
        tag(ST_PRE, 0);
        if (rec_) {
            Step st = step0(n0, k);
            st.in0 = w.ob;
            st.out0 = w.ox;
            st.out1 = w.orr;
            st.out2 = pcg_rc(w);
            st.ctl = w.ctl;
            push(K_PCG_INIT, st, gr);
        } else {
            const double* src = b != nullptr ? b : w.ob;
            double* bcopy = src != w.ob ? w.ob : nullptr;
            KP_DISPATCH(kp_of(k), {
                k_pcg_init<T, KP><<<gr, BLOCK, 0, s>>>(src, w.ox, w.orr, pcg_rc(w), n0, k, w.ctl, bcopy, (T*)nullptr);
            });
            HIPCHK(hipGetLastError());
        }
        finalize(w, FIN_PCG_INIT, 1, k, 0, false, s);
    }

    // z = M r, gamma = r.z, p_prev = 0
    void emit_pcg_start(Workspace<T>& w, char ct, int cpl, int k, int precond, hipStream_t s) {
        const int n0 = lv[0].n;
        const dim3 gr = red_grid(n0, k);
        if (precond) emit_cycle(ct, cpl, pcg_cyc_in(w), nullptr, w.Z, k, nullptr, nullptr, s, w);
        tag(ST_POST, 0);
        if (rec_) {
            if (!precond) throw Error{AMG_E_UNSUPPORTED, "a batched solve is preconditioned"};
            Step st = step0(n0, k);
            st.in0 = w.orr;
            st.in1 = w.Z;
            st.out0 = w.op[0];
            st.ctl = w.ctl;
            st.flag = 1;
            push(K_PCG_START, st, gr);
        } else if (precond) {
            KP_DISPATCH(kp_of(k), { k_pcg_start<T, KP><<<gr, BLOCK, 0, s>>>(w.orr, w.Z, w.op[0], n0, k, w.ctl, 1); });
        } else {
            KP_DISPATCH(kp_of(k), {
                k_pcg_start<double, KP><<<gr, BLOCK, 0, s>>>(w.orr, w.orr, w.op[0], n0, k, w.ctl, 0);
            });
        }
        HIPCHK(hipGetLastError());
        finalize(w, FIN_PCG_START, 2, k, precond, false, s);
    }

    // one iteration: p = z + beta p, Ap, alpha, x/r update, z = M r, beta
    void emit_pcg_iter(Workspace<T>& w, char ct, int cpl, int k, int precond, int flexible, hipStream_t s) {
        const int n0 = lv[0].n;
        const int kp = kp_of(k);
        const dim3 gr = red_grid(n0, k);
        double* pn = w.op[0];
        tag(ST_PRE, 0);
        if (rec_) {
            if (!precond) throw Error{AMG_E_UNSUPPORTED, "a batched solve is preconditioned"};
            Step a = step0(n0, k);
            a.in0 = w.Z;
            a.out0 = pn;
            a.ctl = w.ctl;
            push(K_PCG_P, a, red_grid(n0, k));
            Step b = step0(n0, k);
            b.M = raw_of(outer);
            b.wide = red_mode(outer, k);
            b.in0 = pn;
            b.out0 = w.oAp;
            b.ctl = w.ctl;
            push(K_PCG_SPMV, b, gr);
        } else {
            if (precond) {
                KP_DISPATCH(kp, { k_pcg_p<T, KP><<<red_grid(n0, k), BLOCK, 0, s>>>(w.Z, pn, n0, k, w.ctl); });
            } else {
                KP_DISPATCH(kp, { k_pcg_p<double, KP><<<red_grid(n0, k), BLOCK, 0, s>>>(w.orr, pn, n0, k, w.ctl); });
            }
            if (red_mode(outer, k) == VMODE)
                VKPW_DISPATCH(kp, outer.maxw, { k_pcg_spmv_v<KP, double, WC><<<gr, BLOCK, 0, s>>>(outer, pn, w.oAp, k, w.ctl); })
            else
                KP_DISPATCH(kp, WIDE_DISPATCH(outer, { k_pcg_spmv<KP, WD><<<gr, BLOCK, 0, s>>>(outer, pn, w.oAp, k, w.ctl); }));
            HIPCHK(hipGetLastError());
        }
        finalize(w, FIN_ALPHA, 1, k, 0, true, s);
        if (rec_) {
            Step u = step0(n0, k);
            u.in0 = pn;
            u.in1 = w.oAp;
            u.out0 = w.ox;
            u.out1 = w.orr;
            u.out2 = pcg_rc(w);
            u.ctl = w.ctl;
            push(K_PCG_UPDATE, u, gr);
        } else {
            ev_mark(EV_PCG_UPDATE, true, s);
            KP_DISPATCH(kp, {
                k_pcg_update<T, KP><<<gr, BLOCK, 0, s>>>(pn, w.oAp, w.ox, w.orr, pcg_rc(w), n0, k, w.ctl);
            });
            HIPCHK(hipGetLastError());
            ev_mark(EV_PCG_UPDATE, false, s);
        }
        finalize(w, FIN_UPDATE, 1, k, 0, true, s);
        bool fused = false;
        if (precond) {
            GammaFuse arm;
            arm.on = true;
            arm.r = w.orr;
            arm.ap = w.oAp;
            arm.flexible = flexible;
            arm.ctl = w.ctl;
            GammaFuseArm guard(gf_, arm);
            emit_cycle(ct, cpl, pcg_cyc_in(w), nullptr, w.Z, k, nullptr, w.ctl.done, s, w);
            fused = gf_.used;
        }
        tag(ST_POST, 0);
        if (fused) {
            // r.z and z.Ap were summed by the cycle's last sweep
        } else if (rec_) {
            Step g = step0(n0, k);
            g.in0 = w.orr;
            g.in1 = w.Z;
            g.in2 = w.oAp;
            g.ctl = w.ctl;
            g.flag = flexible;
            push(K_PCG_GAMMA, g, gr);
        } else if (precond) {
            KP_DISPATCH(kp, { k_pcg_gamma<T, KP><<<gr, BLOCK, 0, s>>>(w.orr, w.Z, w.oAp, n0, k, w.ctl, flexible); });
        } else {
            KP_DISPATCH(kp, {
                k_pcg_gamma<double, KP><<<gr, BLOCK, 0, s>>>(w.orr, w.orr, w.oAp, n0, k, w.ctl, flexible);
            });
        }
        HIPCHK(hipGetLastError());
        finalize(w, FIN_GAMMA, 2, k, flexible, true, s);
    }

    // ||b - A x|| of the float64 iterate into normr
    void emit_pcg_final(Workspace<T>& w, int k, const double* b, hipStream_t s) {
        const int n0 = lv[0].n;
        const dim3 gr = red_grid(n0, k);
        tag(ST_PRE, 0);
        if (rec_) {
            Step st = step0(n0, k);
            st.M = raw_of(outer);
            st.wide = wide_of(outer, k);
            st.in0 = w.ox;
            st.in1 = b;
            st.ctl = w.ctl;
            push(K_TRUE_RESNORM, st, gr);
        } else {
            KP_DISPATCH(kp_of(k), WIDE_DISPATCH(outer, {
                k_true_resnorm<double, KP, WD><<<gr, BLOCK, 0, s>>>(outer, w.ox, b, k, w.ctl);
            }));
            HIPCHK(hipGetLastError());
        }
        finalize(w, FIN_TRUE, 1, k, 0, false, s);
    }

    void pcg(char ct, int cpl, const double* b, double* x, int k, double tol, double atol, int maxiter, int flexible,
             int precond, int* iters, double* relres, hipStream_t s) override {
        if (!has_outer) throw Error{AMG_E_STATE, "handle has no float64 outer operator"};
        Workspace<T>& w = get_ws(k);
        const int n0 = lv[0].n;
        k_set_params<<<1, 1, 0, s>>>(w.ctl, tol, atol, maxiter, 0);
        HIPCHK(hipGetLastError());
        emit_pcg_init(w, k, s, b);
        hipGraphExec_t ex0 = graph(w, std::make_tuple(3, ct, cpl, precond, flexible),
                                   [&](hipStream_t cs) { emit_pcg_start(w, ct, cpl, k, precond, cs); });
        HIPCHK(hipGraphLaunch(ex0, s));
        hipGraphExec_t ex = graph(w, std::make_tuple(4, ct, cpl, precond, flexible), [&](hipStream_t cs) {
            for (int i = 0; i < GRAPH_ITERS; ++i) emit_pcg_iter(w, ct, cpl, k, precond, flexible, cs);
        });
        run_until_done(w.hflag, ex, (maxiter + GRAPH_ITERS - 1) / GRAPH_ITERS + 1, s);
        HIPCHK(hipMemcpyAsync(x, w.ox, sizeof(double) * (size_t)n0 * k, hipMemcpyDeviceToDevice, s));
        if (relres) emit_pcg_final(w, k, w.ob, s);
        report(w, k, iters, relres, s);
    }

    // ------------------------------------------- refinement around an inner CG
    // Outer float64 iterative refinement x += e, r = b - A64 x, with each
    // correction e from a CG whose vectors (x, r, p, Ap) are in the cycle
    // precision T, preconditioned by one cycle, run to inner_tol relative to
    // its right-hand side (T) r.  Dot products accumulate in float64.
    void ensure_ipcg(Workspace<T>& w) {
        if (w.ix) return;
        const size_t nk = (size_t)lv[0].n * w.k;
        w.ictl = make_ctl(w, &w.ihflag);
        w.ix = w.alloc_t(nk);
        w.ip = w.alloc_t(nk);
        w.iap = w.alloc_t(nk);
        w.mem.emplace_back(new DevMem(sizeof(int) * (size_t)w.k));
        w.itot = static_cast<int*>(w.mem.back()->p);
        HIPCHK(hipDeviceSynchronize());
    }

    // inner CG on r = B0 (updated in place): x = 0, z = M r, gamma = r.z
    void emit_ipcg_start(Workspace<T>& w, char ct, int cpl, int k, hipStream_t s) {
        const int n0 = lv[0].n;
        const dim3 gr = red_grid(n0, k);
        KP_DISPATCH(kp_of(k), { k_solve_init<T, KP><<<gr, BLOCK, 0, s>>>(w.B0, w.ix, n0, k, w.ictl); });
        HIPCHK(hipGetLastError());
        finalize_ctl(w.ictl, FIN_PCG_INIT, 1, k, 1, false, s);
        emit_cycle(ct, cpl, w.B0, nullptr, w.Z, k, nullptr, nullptr, s, w);
        KP_DISPATCH(kp_of(k), { k_pcg_start<T, KP, T><<<gr, BLOCK, 0, s>>>(w.B0, w.Z, w.ip, n0, k, w.ictl, 1); });
        HIPCHK(hipGetLastError());
        finalize_ctl(w.ictl, FIN_PCG_START, 2, k, 1, false, s);
    }

    void emit_ipcg_iter(Workspace<T>& w, char ct, int cpl, int k, int flexible, hipStream_t s) {
        const int n0 = lv[0].n;
        const int kp = kp_of(k);
        const dim3 gr = red_grid(n0, k);
        const Sell<T>& A = lv[0].A;
        KP_DISPATCH(kp, { k_pcg_p<T, KP, T><<<gr, BLOCK, 0, s>>>(w.Z, w.ip, n0, k, w.ictl); });
        if (red_mode(A, k) == VMODE)
            VKPW_DISPATCH(kp, A.maxw, { k_pcg_spmv_v<KP, T, WC><<<gr, BLOCK, 0, s>>>(A, w.ip, w.iap, k, w.ictl); })
        else
            KP_DISPATCH(kp, WIDE_DISPATCH(A, { k_pcg_spmv<KP, WD, T><<<gr, BLOCK, 0, s>>>(A, w.ip, w.iap, k, w.ictl); }));
        HIPCHK(hipGetLastError());
        finalize_ctl(w.ictl, FIN_ALPHA, 1, k, 0, true, s);
        KP_DISPATCH(kp, {
            k_pcg_update<T, KP, T><<<gr, BLOCK, 0, s>>>(w.ip, w.iap, w.ix, w.B0, (T*)nullptr, n0, k, w.ictl);
        });
        HIPCHK(hipGetLastError());
        finalize_ctl(w.ictl, FIN_UPDATE, 1, k, 0, true, s);
        bool fused = false;
        {
            GammaFuse arm;
            arm.on = true;
            arm.r = w.B0;
            arm.ap = w.iap;
            arm.tv_t = true;
            arm.flexible = flexible;
            arm.ctl = w.ictl;
            GammaFuseArm guard(gf_, arm);
            emit_cycle(ct, cpl, w.B0, nullptr, w.Z, k, nullptr, w.ictl.done, s, w);
            fused = gf_.used;
        }
        if (!fused) {
            KP_DISPATCH(kp, {
                k_pcg_gamma<T, KP, T><<<gr, BLOCK, 0, s>>>(w.B0, w.Z, w.iap, n0, k, w.ictl, flexible);
            });
            HIPCHK(hipGetLastError());
        }
        finalize_ctl(w.ictl, FIN_GAMMA, 2, k, flexible, true, s);
    }

    void pcg_ir(char ct, int cpl, const double* b, double* x, int k, double tol, double inner_tol, int maxiter,
                int inner_maxiter, int flexible, int* iters, int* outer_it, double* relres, hipStream_t s) override {
        if (!has_outer) throw Error{AMG_E_STATE, "handle has no float64 outer operator"};
        Workspace<T>& w = get_ws(k);
        ensure_ipcg(w);
        const int n0 = lv[0].n;
        const int kp = kp_of(k);
        const dim3 gr = red_grid(n0, k);
        k_set_params<<<1, 1, 0, s>>>(w.ctl, tol, 0.0, maxiter, 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemsetAsync(w.itot, 0, sizeof(int) * (size_t)k, s));
        HIPCHK(hipMemcpyAsync(w.ob, b, sizeof(double) * (size_t)n0 * k, hipMemcpyDeviceToDevice, s));
        // x = 0, normb, normr = ||b||, active = normr >= tol normb; B0 = (T) b
        KP_DISPATCH(kp, { k_solve_init<double, KP><<<gr, BLOCK, 0, s>>>(w.ob, w.ox, n0, k, w.ctl); });
        HIPCHK(hipGetLastError());
        finalize(w, FIN_SOLVE_INIT, 1, k, 0, false, s);
        k_ir_floor<<<(k + 255) / 256, 256, 0, s>>>(w.ictl, w.ctl, tol, k);
        HIPCHK(hipGetLastError());
        copy_launch<T, double>(w.ob, w.B0, n0, k, nullptr, nullptr, s);
        // one step = head (inner parameters, x_e = 0, z = M r, gamma), inner
        // iterations until the inner `done`, tail (x += e, float64 residual)
        k_set_params<<<1, 1, 0, s>>>(w.ictl, inner_tol, 0.0, inner_maxiter, 0);
        HIPCHK(hipGetLastError());
        hipGraphExec_t head = graph(w, std::make_tuple(7, ct, cpl, flexible, 0), [&](hipStream_t cs) {
            k_reset_done<<<1, 1, 0, cs>>>(w.ictl);
            HIPCHK(hipGetLastError());
            emit_ipcg_start(w, ct, cpl, k, cs);
        });
        hipGraphExec_t body = graph(w, std::make_tuple(8, ct, cpl, flexible, 0), [&](hipStream_t cs) {
            for (int i = 0; i < GRAPH_ITERS; ++i) emit_ipcg_iter(w, ct, cpl, k, flexible, cs);
        });
        hipGraphExec_t tail = graph(w, std::make_tuple(9, ct, cpl, flexible, 0), [&](hipStream_t cs) {
            KP_DISPATCH(kp, {
                k_add_masked<double, T, KP><<<row_grid(n0, k), BLOCK, 0, cs>>>(w.ox, w.ix, n0, k, w.ctl.active,
                                                                               w.ctl.done);
            });
            if (red_mode(outer, k) == VMODE)
                VKPW_DISPATCH(kp, outer.maxw, { k_resnorm_v<double, T, KP, WC><<<gr, BLOCK, 0, cs>>>(outer, w.ox, w.ob, w.B0, k, w.ctl); })
            else
                KP_DISPATCH(kp, WIDE_DISPATCH(outer, {
                    k_resnorm<double, T, KP, WD><<<gr, BLOCK, 0, cs>>>(outer, w.ox, w.ob, w.B0, k, w.ctl);
                }));
            k_add_iters<<<(k + 255) / 256, 256, 0, cs>>>(w.itot, w.ictl.it, w.ctl.active, k);
            HIPCHK(hipGetLastError());
            finalize(w, FIN_RESNORM, 1, k, 0, true, cs);
        });
        HIPCHK(hipStreamSynchronize(s));
        for (int o = 0; o < maxiter && !*(volatile const int*)w.hflag; ++o) {
            HIPCHK(hipGraphLaunch(head, s));
            run_until_done(w.ihflag, body, (inner_maxiter + GRAPH_ITERS - 1) / GRAPH_ITERS + 1, s);
            HIPCHK(hipGraphLaunch(tail, s));
            HIPCHK(hipStreamSynchronize(s));
        }
        HIPCHK(hipMemcpyAsync(x, w.ox, sizeof(double) * (size_t)n0 * k, hipMemcpyDeviceToDevice, s));
        std::vector<int> oi(k), tot(k);
        HIPCHK(hipMemcpyAsync(tot.data(), w.itot, sizeof(int) * (size_t)k, hipMemcpyDeviceToHost, s));
        report(w, k, oi.data(), relres, s);
        for (int c = 0; c < k; ++c) {
            if (iters) iters[c] = tot[c];
            if (outer_it) outer_it[c] = oi[c];
        }
    }

    // ------------------------------------------------ CG with reliable updates
    // Mixed precision without restarts (Strzodka & Goeddeke 2006; Clark et al.
    // 2010, "Solving lattice QCD systems of equations using mixed precision
    // solvers on GPUs"): the CG vectors - the correction xh, the recurrence
    // residual rh, p and Ap - are in the cycle precision T, the dot products
    // accumulate in float64, and one cycle preconditions.  When a column's
    // recurrence residual has fallen by delta since its last update, reaches
    // the tolerance, or the column reaches maxiter, the update folds
    // x += xh (float64), xh = 0, and replaces rh by the true residual
    // b - A64 x; the search direction is kept.  A column stops on its true
    // float64 residual, ||b - A64 x||^2 <= max(tol^2 b.b, atol^2).
    // x = 0, rh = (T) b, xh = 0, atol2, normb, normr = ||b||, it = 0; b read from the
    // caller's b (copied into w.ob on the way) or from w.ob
    void emit_ru_init(Workspace<T>& w, int k, const double* b, hipStream_t s) {
        const int n0 = lv[0].n;
        const dim3 gr = red_grid(n0, k);
        double* bcopy = b != w.ob ? w.ob : nullptr;
        KP_DISPATCH(kp_of(k), {
            k_pcg_init<T, KP><<<gr, BLOCK, 0, s>>>(b, w.ox, (double*)nullptr, w.B0, n0, k, w.ctl, bcopy, w.ix);
        });
        HIPCHK(hipGetLastError());
        finalize(w, FIN_PCG_INIT, 1, k, 0, false, s);
    }

    void emit_ru_start(Workspace<T>& w, char ct, int cpl, int k, hipStream_t s) {
        const int n0 = lv[0].n;
        const dim3 gr = red_grid(n0, k);
        // z = M rh, p_prev = 0, gamma = rh.z, r.r at the last update = rh.rh
        emit_cycle(ct, cpl, w.B0, nullptr, w.Z, k, nullptr, nullptr, s, w);
        KP_DISPATCH(kp_of(k), { k_pcg_start<T, KP, T><<<gr, BLOCK, 0, s>>>(w.B0, w.Z, w.ip, n0, k, w.ctl, 1); });
        HIPCHK(hipGetLastError());
        finalize(w, FIN_RU_START, 2, k, 0, false, s);
    }

    // xh += alpha p, rh -= alpha Ap (cycle precision), rh.rh partials
    void ru_update_launch(Workspace<T>& w, int k, hipStream_t s) {
        const int n0 = lv[0].n;
        const int kp = kp_of(k);
        const dim3 gr = red_grid(n0, k);
        const bool vk = vec_ && k == kp && vec_ok(k, kp, vec_width<T>());
        if (vk)
            VKP_DISPATCH(kp, { k_ru_update_v<T, KP><<<gr, BLOCK, 0, s>>>(w.ip, w.iap, w.ix, w.B0, n0, k, w.ctl); })
        else
            KP_DISPATCH(kp, {
                k_pcg_update<T, KP, T><<<gr, BLOCK, 0, s>>>(w.ip, w.iap, w.ix, w.B0, (T*)nullptr, n0, k, w.ctl);
            });
        HIPCHK(hipGetLastError());
    }

    void emit_ru_iter(Workspace<T>& w, char ct, int cpl, int k, int flexible, hipStream_t s) {
        const int n0 = lv[0].n;
        const int kp = kp_of(k);
        const dim3 gr = red_grid(n0, k);
        const dim3 grr = ru_row_grid(n0, k);  // the reducing row kernels
        const Sell<T>& A = lv[0].A;
        // CG vector kernels with 16 bytes per lane where k allows it
        const bool vk = vec_ && k == kp && vec_ok(k, kp, vec_width<T>());
        // p = z + beta p, with the previous iteration's pending fold x += xh, xh = 0;
        // Ap = A p, p.Ap; alpha (and no update pending)
        if (vk) VKP_DISPATCH(kp, { k_ru_p_v<T, KP><<<gr, BLOCK, 0, s>>>(w.Z, w.ip, w.ox, w.ix, n0, w.ctl); })
        else KP_DISPATCH(kp, { k_ru_p<T, KP><<<gr, BLOCK, 0, s>>>(w.Z, w.ip, w.ox, w.ix, n0, k, w.ctl); });
        if (red_mode(A, k) == VMODE)
            VKPW_DISPATCH(kp, A.maxw, { k_pcg_spmv_v<KP, T, WC><<<grr, BLOCK, 0, s>>>(A, w.ip, w.iap, k, w.ctl); })
        else
            KP_DISPATCH(kp, WIDE_DISPATCH(A, { k_pcg_spmv<KP, WD, T><<<grr, BLOCK, 0, s>>>(A, w.ip, w.iap, k, w.ctl); }));
        HIPCHK(hipGetLastError());
        finalize(w, FIN_RU_ALPHA, 1, k, 0, true, s, (int)grr.x);
        // xh += alpha p, rh -= alpha Ap, rh.rh; the update decision
        ev_mark(EV_RU_UPDATE, true, s);
        ru_update_launch(w, k, s);
        ev_mark(EV_RU_UPDATE, false, s);
        finalize(w, FIN_RU_DECIDE, 1, k, 0, true, s);
        // the reliable update on the columns with one pending (the kernel returns at once
        // when none is): s = s - A xh in float64, where s (in ob, b at the start) holds
        // b - A x, so s becomes b - A (x + xh); rh = s in the cycle precision; its sums
        // in the third partial slot, read by the finaliser after the cycle; the fold
        // x += xh, xh = 0 follows in the next iteration's p update (or in k_ru_out at the end)
        Ctl c3 = w.ctl;
        c3.part += 2 * (size_t)k * grr.x;
        // vector mode with the lanes laid out for T (the gathered xh)
        if (vec_ && vec_ok(k, kp, vec_width<T>()) && vec_default(outer.maxw, outer.nrows))
            VKPW_DISPATCH(kp, outer.maxw, {
                k_ru_resnorm_v<double, T, KP, WC><<<grr, BLOCK, 0, s>>>(outer, w.ix, w.ob, w.B0, k, c3);
            })
        else
            KP_DISPATCH(kp, WIDE_DISPATCH(outer, {
                k_ru_resnorm<double, T, KP, WD><<<grr, BLOCK, 0, s>>>(outer, w.ix, w.ob, w.B0, k, c3);
            }));
        HIPCHK(hipGetLastError());
        // z = M rh with rh.z (and z.Ap) summed by the last sweep; beta
        bool fused = false;
        {
            GammaFuse arm;
            arm.on = true;
            arm.r = w.B0;
            arm.ap = w.iap;
            arm.tv_t = true;
            arm.flexible = flexible;
            arm.ctl = w.ctl;
            arm.nb = (int)grr.x;
            GammaFuseArm guard(gf_, arm);
            emit_cycle(ct, cpl, w.B0, nullptr, w.Z, k, nullptr, w.ctl.done, s, w);
            fused = gf_.used;
        }
        if (!fused) {
            KP_DISPATCH(kp, {
                k_pcg_gamma<T, KP, T><<<grr, BLOCK, 0, s>>>(w.B0, w.Z, w.iap, n0, k, w.ctl, flexible);
            });
            HIPCHK(hipGetLastError());
        }
        // the update's outcome (normr, convergence) and beta
        finalize(w, FIN_RU_GAMMA, 3, k, flexible, true, s, (int)grr.x);
    }

    void pcg_ru(char ct, int cpl, const double* b, double* x, int k, double tol, double atol, double delta,
                int maxiter, int flexible, int* iters, int* updates, double* relres, hipStream_t s) override {
        if (!has_outer) throw Error{AMG_E_STATE, "handle has no float64 outer operator"};
        Workspace<T>& w = get_ws(k);
        ensure_ipcg(w);
        const int n0 = lv[0].n;
        k_set_params<<<1, 1, 0, s>>>(w.ctl, tol, atol, maxiter, 0);
        k_ru_params<<<(k + 255) / 256, 256, 0, s>>>(w.ctl, delta, k);
        HIPCHK(hipGetLastError());
        emit_ru_init(w, k, b, s);
        hipGraphExec_t ex0 = graph(w, std::make_tuple(13, ct, cpl, 0, 0),
                                   [&](hipStream_t cs) { emit_ru_start(w, ct, cpl, k, cs); });
        HIPCHK(hipGraphLaunch(ex0, s));
        hipGraphExec_t ex = graph(w, std::make_tuple(14, ct, cpl, flexible, 0), [&](hipStream_t cs) {
            for (int i = 0; i < GRAPH_ITERS; ++i) emit_ru_iter(w, ct, cpl, k, flexible, cs);
        });
        run_until_done(w.hflag, ex, (maxiter + GRAPH_ITERS - 1) / GRAPH_ITERS + 1, s);
        // x = x + xh (the last update's fold); every column stopped on a reliable
        // update, so normr already holds its true residual ||b - A x||
        KP_DISPATCH(kp_of(k), { k_ru_out<T, KP><<<red_grid(n0, k), BLOCK, 0, s>>>(w.ox, w.ix, x, n0, k); });
        HIPCHK(hipGetLastError());
        if (updates) {
            std::vector<int> u(k);
            HIPCHK(hipMemcpyAsync(u.data(), w.ctl.count, sizeof(int) * (size_t)k, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            std::copy(u.begin(), u.end(), updates);
        }
        report(w, k, iters, relres, s);
    }

    // ---------------------------------------------------------------- restarted GMRES
    // jax.scipy.sparse.linalg.gmres with solve_method='batched' (left
    // preconditioning): per restart, `restart` Arnoldi steps on M A with
    // classical Gram-Schmidt applied twice, the least-squares problem of the
    // Hessenberg matrix solved once on the host, x += V y, then the
    // preconditioned residual M (b - A x) restarts the basis; a column stops when
    // ||M (b - A x)|| <= ||M b|| min(1, max(tol ||b||, atol) / ||b||) or after
    // maxiter restarts.
    void ensure_gmres(Workspace<T>& w, int m) {
        if (w.gm >= m) return;
        // graphs captured for a shorter basis name the buffers replaced below
        for (auto it = w.graphs.begin(); it != w.graphs.end();) {
            const int kind = std::get<0>(it->first);
            if (kind == 10 || kind == 11 || kind == 12) {
                HIPCHK(hipGraphExecDestroy(it->second));
                it = w.graphs.erase(it);
            } else {
                ++it;
            }
        }
        const size_t nk = (size_t)lv[0].n * w.k, K = (size_t)w.k;
        const size_t nblk = red_grid(lv[0].n, w.k).x;
        w.gV = w.alloc_d(nk * (size_t)(m + 1));
        w.gw = w.alloc_d(nk);
        w.gAv = w.alloc_d(nk);
        w.gH = w.alloc_d(K * (size_t)(m + 1) * m);
        w.gh1 = w.alloc_d(K * (size_t)(m + 1));
        w.gh2 = w.alloc_d(K * (size_t)(m + 1));
        w.gww = w.alloc_d(K);
        w.gy = w.alloc_d(K * (size_t)m);
        w.gpart = w.alloc_d(K * (size_t)(m + 1) * nblk);
        w.gout = w.alloc_d(K);
        if (!w.ghflag) w.gctl = make_ctl(w, &w.ghflag);
        w.gm = m;
        HIPCHK(hipDeviceSynchronize());
    }

    // out[q k + c] = V_q . w for q < nv (one pass, deterministic sums)
    void gmres_dots(Workspace<T>& w, const double* V, size_t ldv, int nv, const double* x, int k, double* out,
                    hipStream_t s) {
        const int n0 = lv[0].n;
        const dim3 gr = red_grid(n0, k);
        KP_DISPATCH(kp_of(k), { k_gmres_dots<KP><<<gr, BLOCK, 0, s>>>(V, ldv, nv, x, n0, k, w.gpart); });
        k_gmres_sum<<<dim3((unsigned)k, (unsigned)nv), BLOCK, 0, s>>>(w.gpart, (int)gr.x, k, out);
        HIPCHK(hipGetLastError());
    }

    // gw = M (b - A x) (or b - A x without a preconditioner), V_0 = gw / ||gw||, gww = ||gw||^2
    void emit_gmres_start(Workspace<T>& w, char ct, int cpl, int k, int precond, hipStream_t s) {
        const int n0 = lv[0].n, kp = kp_of(k);
        const size_t nk = (size_t)n0 * k;
        const dim3 gr = red_grid(n0, k);
        const unsigned ge = (unsigned)((nk + BLOCK - 1) / BLOCK);
        if (precond) {
            if (red_mode(outer, k) == VMODE)
                VKPW_DISPATCH(kp, outer.maxw, { k_resnorm_v<double, T, KP, WC><<<gr, BLOCK, 0, s>>>(outer, w.ox, w.ob, w.B0, k, w.gctl); })
            else
                KP_DISPATCH(kp, WIDE_DISPATCH(outer, {
                    k_resnorm<double, T, KP, WD><<<gr, BLOCK, 0, s>>>(outer, w.ox, w.ob, w.B0, k, w.gctl);
                }));
            HIPCHK(hipGetLastError());
            emit_cycle(ct, cpl, w.B0, nullptr, w.Z, k, nullptr, nullptr, s, w);
            copy_launch<double, T>(w.Z, w.gw, n0, k, nullptr, nullptr, s);
        } else {
            KP_DISPATCH(kp, WIDE_DISPATCH(outer, {
                k_resnorm<double, double, KP, WD><<<gr, BLOCK, 0, s>>>(outer, w.ox, w.ob, w.gw, k, w.gctl);
            }));
            HIPCHK(hipGetLastError());
        }
        gmres_dots(w, w.gw, 0, 1, w.gw, k, w.gww, s);
        k_gmres_scale<<<ge, BLOCK, 0, s>>>(w.gw, w.gww, w.gV, nk, k);
        HIPCHK(hipGetLastError());
    }

    // Arnoldi step j: w = M A V_j, two classical Gram-Schmidt passes against V_0..V_j,
    // V_{j+1} = w / ||w||, column j of H
    void emit_gmres_step(Workspace<T>& w, char ct, int cpl, int k, int m, int j, int precond, hipStream_t s) {
        const int n0 = lv[0].n, kp = kp_of(k);
        const size_t nk = (size_t)n0 * k;
        const dim3 gr = red_grid(n0, k);
        const unsigned ge = (unsigned)((nk + BLOCK - 1) / BLOCK);
        const double* Vj = w.gV + (size_t)j * nk;
        if (red_mode(outer, k) == VMODE)
            VKPW_DISPATCH(kp, outer.maxw, { k_pcg_spmv_v<KP, double, WC><<<gr, BLOCK, 0, s>>>(outer, Vj, w.gAv, k, w.gctl); })
        else
            KP_DISPATCH(kp, WIDE_DISPATCH(outer, { k_pcg_spmv<KP, WD><<<gr, BLOCK, 0, s>>>(outer, Vj, w.gAv, k, w.gctl); }));
        HIPCHK(hipGetLastError());
        if (precond) {
            copy_launch<T, double>(w.gAv, w.B0, n0, k, nullptr, nullptr, s);
            emit_cycle(ct, cpl, w.B0, nullptr, w.Z, k, nullptr, nullptr, s, w);
            copy_launch<double, T>(w.Z, w.gw, n0, k, nullptr, nullptr, s);
        } else {
            // a copy kernel, not a memcpy node, inside the captured Arnoldi graphs
            copy_launch<double, double>(w.gAv, w.gw, n0, k, nullptr, nullptr, s);
        }
        gmres_dots(w, w.gV, nk, j + 1, w.gw, k, w.gh1, s);
        k_gmres_orth<<<ge, BLOCK, 0, s>>>(w.gV, nk, j + 1, w.gh1, w.gw, nk, k);
        gmres_dots(w, w.gV, nk, j + 1, w.gw, k, w.gh2, s);
        k_gmres_orth<<<ge, BLOCK, 0, s>>>(w.gV, nk, j + 1, w.gh2, w.gw, nk, k);
        gmres_dots(w, w.gw, 0, 1, w.gw, k, w.gww, s);
        k_gmres_scale<<<ge, BLOCK, 0, s>>>(w.gw, w.gww, w.gV + (size_t)(j + 1) * nk, nk, k);
        k_gmres_hcol<<<(k + 63) / 64, 64, 0, s>>>(w.gH, m, j, j + 1, w.gh1, w.gh2, w.gww, k);
        HIPCHK(hipGetLastError());
    }

    // min || beta e_1 - H y || for one column: Givens rotations on the (m + 1) x m
    // upper Hessenberg H (row-major), stopping at the first zero subdiagonal
    static void gmres_lstsq(const double* H, int m, double beta, double* y) {
        std::vector<double> R((size_t)(m + 1) * m), g(m + 1, 0.0), cs(m), sn(m);
        std::copy(H, H + (size_t)(m + 1) * m, R.begin());
        g[0] = beta;
        int mm = m;
        for (int j = 0; j < m; ++j) {
            for (int i = 0; i < j; ++i) {
                const double a = R[(size_t)i * m + j], b = R[(size_t)(i + 1) * m + j];
                R[(size_t)i * m + j] = cs[i] * a + sn[i] * b;
                R[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
            }
            const double a = R[(size_t)j * m + j], b = R[(size_t)(j + 1) * m + j];
            const double r = std::hypot(a, b);
            cs[j] = r > 0.0 ? a / r : 1.0;
            sn[j] = r > 0.0 ? b / r : 0.0;
            R[(size_t)j * m + j] = r;
            R[(size_t)(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            if (b == 0.0) {  // breakdown: V_{j+1} = 0, the space is invariant
                mm = j + 1;
                break;
            }
        }
        for (int j = 0; j < m; ++j) y[j] = 0.0;
        for (int i = mm - 1; i >= 0; --i) {
            double acc = g[i];
            for (int j = i + 1; j < mm; ++j) acc -= R[(size_t)i * m + j] * y[j];
            const double d = R[(size_t)i * m + i];
            y[i] = d != 0.0 ? acc / d : 0.0;
        }
    }

    void gmres(char ct, int cpl, const double* b, double* x, int k, double tol, double atol, int restart, int maxiter,
               int precond, int method, int* iters, double* relres, hipStream_t s) override {
        if (!has_outer) throw Error{AMG_E_STATE, "handle has no float64 outer operator"};
        Workspace<T>& w = get_ws(k);
        const int n0 = lv[0].n, kp = kp_of(k);
        const int m = std::min(restart, n0);
        if (m < 1 || m > GMRES_MMAX) throw Error{AMG_E_ARG, "restart must be 1..32"};
        ensure_gmres(w, m);
        const size_t nk = (size_t)n0 * k;
        k_set_params<<<1, 1, 0, s>>>(w.gctl, 0.0, 0.0, 0, 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(w.ob, b, sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
        k_fill<double><<<(unsigned)((nk + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(w.ox, nk, 0.0);
        HIPCHK(hipGetLastError());
        // ||b|| and ||M b||
        std::vector<double> bn(k), mbn(k), beta(k), Hh((size_t)k * (m + 1) * m), yh((size_t)k * m, 0.0);
        gmres_dots(w, w.ob, 0, 1, w.ob, k, w.gout, s);
        HIPCHK(hipMemcpyAsync(bn.data(), w.gout, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        if (precond) {
            copy_launch<T, double>(w.ob, w.B0, n0, k, nullptr, nullptr, s);
            emit_cycle(ct, cpl, w.B0, nullptr, w.Z, k, nullptr, nullptr, s, w);
            copy_launch<double, T>(w.Z, w.gw, n0, k, nullptr, nullptr, s);
            gmres_dots(w, w.gw, 0, 1, w.gw, k, w.gout, s);
            HIPCHK(hipMemcpyAsync(mbn.data(), w.gout, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        }
        HIPCHK(hipStreamSynchronize(s));
        if (!precond) mbn = bn;
        std::vector<double> ptol(k);
        std::vector<int> it(k, 0), act(k, 0);
        for (int c = 0; c < k; ++c) {
            bn[c] = std::sqrt(bn[c]);
            mbn[c] = std::sqrt(mbn[c]);
            const double at = std::max(tol * bn[c], atol);
            ptol[c] = bn[c] > 0.0 ? mbn[c] * std::min(1.0, at / bn[c]) : 0.0;
        }
        hipGraphExec_t start = graph(w, std::make_tuple(10, ct, cpl, precond, 0),
                                     [&](hipStream_t cs) { emit_gmres_start(w, ct, cpl, k, precond, cs); });
        const bool incr = method == 1;
        hipGraphExec_t arnoldi = nullptr;
        std::vector<hipGraphExec_t> step(m, nullptr);  // incremental: one graph per Arnoldi step
        if (incr) {
            for (int j = 0; j < m; ++j)
                step[j] = graph(w, std::make_tuple(12, ct, cpl, precond, m * (GMRES_MMAX + 1) + j),
                                [&](hipStream_t cs) { emit_gmres_step(w, ct, cpl, k, m, j, precond, cs); });
        } else {
            arnoldi = graph(w, std::make_tuple(11, ct, cpl, precond, m), [&](hipStream_t cs) {
                for (int j = 0; j < m; ++j) emit_gmres_step(w, ct, cpl, k, m, j, precond, cs);
            });
        }
        // incremental: per column Givens rotations (cos, sin), rotated right-hand side g
        // and rotated columns of H, basis length used (0 while still extending)
        std::vector<double> gc((size_t)k * m), gs((size_t)k * m), gg((size_t)k * (m + 1)), Rc((size_t)k * (m + 1) * m);
        std::vector<int> kk(k, 0);
        auto fetch_beta = [&]() {
            HIPCHK(hipMemcpyAsync(beta.data(), w.gww, sizeof(double) * k, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            for (int c = 0; c < k; ++c) beta[c] = std::sqrt(beta[c]);
        };
        HIPCHK(hipGraphLaunch(start, s));
        fetch_beta();
        int nact = 0;
        for (int c = 0; c < k; ++c) {
            act[c] = bn[c] > 0.0 && beta[c] > ptol[c] && 0 < maxiter;
            nact += act[c];
        }
        while (nact > 0) {
            int used = m;  // basis vectors the update reads
            if (!incr) {
                HIPCHK(hipGraphLaunch(arnoldi, s));
                HIPCHK(hipMemcpyAsync(Hh.data(), w.gH, sizeof(double) * Hh.size(), hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                for (int c = 0; c < k; ++c) {
                    if (act[c]) gmres_lstsq(Hh.data() + (size_t)c * (m + 1) * m, m, beta[c], yh.data() + (size_t)c * m);
                    else std::fill(yh.begin() + (size_t)c * m, yh.begin() + (size_t)(c + 1) * m, 0.0);
                }
            } else {
                // jax's solve_method='incremental': after each Arnoldi step rotate the new
                // column of every column's H and stop extending a column's basis once its
                // least-squares residual |g_{j+1}| <= ptol; the launches stop when no column
                // extends any more
                for (int c = 0; c < k; ++c) {
                    kk[c] = act[c] ? 0 : -1;
                    gg[(size_t)c * (m + 1)] = beta[c];
                }
                for (int j = 0; j < m; ++j) {
                    HIPCHK(hipGraphLaunch(step[j], s));
                    HIPCHK(hipMemcpyAsync(Hh.data(), w.gH, sizeof(double) * Hh.size(), hipMemcpyDeviceToHost, s));
                    HIPCHK(hipStreamSynchronize(s));
                    used = j + 1;
                    bool more = false;
                    for (int c = 0; c < k; ++c) {
                        if (kk[c] != 0) continue;
                        const double* Hc = Hh.data() + (size_t)c * (m + 1) * m;
                        double* R = Rc.data() + (size_t)c * (m + 1) * m;
                        double* cs = gc.data() + (size_t)c * m;
                        double* sn = gs.data() + (size_t)c * m;
                        double* g = gg.data() + (size_t)c * (m + 1);
                        for (int i = 0; i <= j + 1; ++i) R[(size_t)i * m + j] = Hc[(size_t)i * m + j];
                        for (int i = 0; i < j; ++i) {
                            const double a = R[(size_t)i * m + j], bb = R[(size_t)(i + 1) * m + j];
                            R[(size_t)i * m + j] = cs[i] * a + sn[i] * bb;
                            R[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * bb;
                        }
                        const double a = R[(size_t)j * m + j], bb = R[(size_t)(j + 1) * m + j];
                        const double r = std::hypot(a, bb);
                        cs[j] = r > 0.0 ? a / r : 1.0;
                        sn[j] = r > 0.0 ? bb / r : 0.0;
                        g[j + 1] = -sn[j] * g[j];
                        g[j] = cs[j] * g[j];
                        if (std::fabs(g[j + 1]) <= ptol[c] || j + 1 == m) kk[c] = j + 1;
                        else more = true;
                    }
                    if (!more) break;
                }
                // y from the leading kk x kk block (the same Givens solve as batched)
                std::vector<double> Ht;
                for (int c = 0; c < k; ++c) {
                    double* yc = yh.data() + (size_t)c * used;
                    std::fill(yc, yc + used, 0.0);
                    const int q = kk[c] > 0 ? kk[c] : (kk[c] == 0 ? used : 0);
                    if (q == 0) continue;
                    const double* Hc = Hh.data() + (size_t)c * (m + 1) * m;
                    Ht.assign((size_t)(q + 1) * q, 0.0);
                    for (int i = 0; i <= q; ++i)
                        for (int jj = 0; jj < q; ++jj) Ht[(size_t)i * q + jj] = Hc[(size_t)i * m + jj];
                    gmres_lstsq(Ht.data(), q, beta[c], yc);
                }
            }
            HIPCHK(hipMemcpyAsync(w.gy, yh.data(), sizeof(double) * (size_t)k * used, hipMemcpyHostToDevice, s));
            k_gmres_update<<<(unsigned)((nk + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(w.ox, w.gV, nk, used, w.gy, nk, k);
            HIPCHK(hipGetLastError());
            HIPCHK(hipGraphLaunch(start, s));
            fetch_beta();
            nact = 0;
            for (int c = 0; c < k; ++c) {
                if (!act[c]) continue;
                ++it[c];
                act[c] = beta[c] > ptol[c] && it[c] < maxiter;
                nact += act[c];
            }
        }
        HIPCHK(hipMemcpyAsync(x, w.ox, sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
        if (relres) {
            const dim3 gr = red_grid(n0, k);
            KP_DISPATCH(kp, WIDE_DISPATCH(outer, {
                k_resnorm<double, double, KP, WD><<<gr, BLOCK, 0, s>>>(outer, w.ox, w.ob, (double*)nullptr, k, w.gctl);
            }));
            HIPCHK(hipGetLastError());
            k_gmres_sum<<<dim3((unsigned)k, 1), BLOCK, 0, s>>>(w.gctl.part, (int)gr.x, k, w.gout);
            std::vector<double> rr(k);
            HIPCHK(hipMemcpyAsync(rr.data(), w.gout, sizeof(double) * k, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            for (int c = 0; c < k; ++c) relres[c] = std::sqrt(rr[c]) / (bn[c] > 0.0 ? bn[c] : 1.0);
        }
        if (iters)
            for (int c = 0; c < k; ++c) iters[c] = it[c];
        HIPCHK(hipStreamSynchronize(s));
    }

    // ---------------------------------------------------------------- BiCGSTAB
    void ensure_bicg(Workspace<T>& w) {
        if (w.orh) return;
        const size_t nk = (size_t)lv[0].n * w.k;
        w.orh = w.alloc_d(nk);
        w.os = w.alloc_d(nk);
        w.ot = w.alloc_d(nk);
        w.Z2 = w.alloc_t(nk);
    }

    // one BiCGSTAB iteration (precond: the cycle as right preconditioner)
    void emit_bicg_iter(Workspace<T>& w, char ct, int cpl, int k, int precond, hipStream_t s) {
        if (rec_) throw Error{AMG_E_UNSUPPORTED, "a batched solve runs preconditioned CG only"};
        const int n0 = lv[0].n;
        const int kp = kp_of(k);
        const dim3 gr = red_grid(n0, k);
        const bool same = std::is_same<T, double>::value;
        T* pc = (precond && !same) ? w.B0 : nullptr;
        KP_DISPATCH(kp, { k_bicg_p<T, KP><<<gr, BLOCK, 0, s>>>(w.orr, w.oAp, w.op[0], pc, n0, k, w.ctl); });
        HIPCHK(hipGetLastError());
        if (precond) {
            const T* cin = same ? reinterpret_cast<const T*>(w.op[0]) : w.B0;
            emit_cycle(ct, cpl, cin, nullptr, w.Z, k, nullptr, w.ctl.done, s, w);
            KP_DISPATCH(kp, WIDE_DISPATCH(outer, {
                k_bicg_spmv<T, KP, WD, 0><<<gr, BLOCK, 0, s>>>(outer, w.Z, w.orh, w.oAp, k, w.ctl);
            }));
        } else {
            KP_DISPATCH(kp, WIDE_DISPATCH(outer, {
                k_bicg_spmv<double, KP, WD, 0><<<gr, BLOCK, 0, s>>>(outer, w.op[0], w.orh, w.oAp, k, w.ctl);
            }));
        }
        HIPCHK(hipGetLastError());
        finalize(w, FIN_BICG_ALPHA, 1, k, 0, true, s);
        KP_DISPATCH(kp, { k_bicg_s<T, KP><<<gr, BLOCK, 0, s>>>(w.orr, w.oAp, w.os, pc, n0, k, w.ctl); });
        HIPCHK(hipGetLastError());
        finalize(w, FIN_BICG_EXIT, 1, k, 0, true, s);
        if (precond) {
            const T* cin = same ? reinterpret_cast<const T*>(w.os) : w.B0;
            emit_cycle(ct, cpl, cin, nullptr, w.Z2, k, nullptr, w.ctl.done, s, w);
            KP_DISPATCH(kp, WIDE_DISPATCH(outer, {
                k_bicg_spmv<T, KP, WD, 1><<<gr, BLOCK, 0, s>>>(outer, w.Z2, w.os, w.ot, k, w.ctl);
            }));
            HIPCHK(hipGetLastError());
            finalize(w, FIN_BICG_OMEGA, 2, k, 0, true, s);
            KP_DISPATCH(kp, {
                k_bicg_update<T, KP><<<gr, BLOCK, 0, s>>>(w.Z, w.Z2, w.os, w.ot, w.orh, w.ox, w.orr, n0, k, w.ctl);
            });
        } else {
            KP_DISPATCH(kp, WIDE_DISPATCH(outer, {
                k_bicg_spmv<double, KP, WD, 1><<<gr, BLOCK, 0, s>>>(outer, w.os, w.os, w.ot, k, w.ctl);
            }));
            HIPCHK(hipGetLastError());
            finalize(w, FIN_BICG_OMEGA, 2, k, 0, true, s);
            KP_DISPATCH(kp, {
                k_bicg_update<double, KP><<<gr, BLOCK, 0, s>>>(w.op[0], w.os, w.os, w.ot, w.orh, w.ox, w.orr, n0, k,
                                                              w.ctl);
            });
        }
        HIPCHK(hipGetLastError());
        finalize(w, FIN_BICG_RHO, 2, k, 0, true, s);
    }

    void bicgstab(char ct, int cpl, const double* b, double* x, int k, double tol, double atol, int maxiter,
                  int precond, int* iters, double* relres, hipStream_t s) override {
        if (!has_outer) throw Error{AMG_E_STATE, "handle has no float64 outer operator"};
        Workspace<T>& w = get_ws(k);
        ensure_bicg(w);
        const int n0 = lv[0].n;
        k_set_params<<<1, 1, 0, s>>>(w.ctl, tol, atol, maxiter, 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(w.ob, b, sizeof(double) * (size_t)n0 * k, hipMemcpyDeviceToDevice, s));
        const dim3 gr = red_grid(n0, k);
        KP_DISPATCH(kp_of(k), {
            k_bicg_init<KP><<<gr, BLOCK, 0, s>>>(w.ob, w.ox, w.orr, w.orh, w.op[0], w.oAp, n0, k, w.ctl);
        });
        HIPCHK(hipGetLastError());
        finalize(w, FIN_BICG_INIT, 1, k, 0, false, s);
        hipGraphExec_t ex = graph(w, std::make_tuple(5, ct, cpl, precond, 0), [&](hipStream_t cs) {
            for (int i = 0; i < GRAPH_ITERS; ++i) emit_bicg_iter(w, ct, cpl, k, precond, cs);
        });
        run_until_done(w.hflag, ex, (maxiter + GRAPH_ITERS - 1) / GRAPH_ITERS + 1, s);
        HIPCHK(hipMemcpyAsync(x, w.ox, sizeof(double) * (size_t)n0 * k, hipMemcpyDeviceToDevice, s));
        if (relres) emit_pcg_final(w, k, w.ob, s);
        report(w, k, iters, relres, s);
    }

    void clear_graphs() {
        for (auto& kv : ws) {
            for (auto& g : kv.second->graphs) (void)hipGraphExecDestroy(g.second);
            kv.second->graphs.clear();
        }
    }

    void set_option(int opt, long long value) override {
        if (opt == AMG_OPT_FUSE) {
            fuse = value != 0;
            plan_tail();
        } else if (opt == AMG_OPT_AUTOTUNE) {
            tune_ = value != 0;
            if (tune_)
                for (auto& kv : ws) tune_all(kv.first);
        } else if (opt == AMG_OPT_VECTOR) {
            vec_ = value != 0;
        } else if (opt == AMG_OPT_TAIL) {
            ARGCHK(value >= 0 && value <= (1LL << 30), "tail rows must be in [0, 2^30]");
            tail_n_ = (int)value;
            plan_tail();
        } else {
            throw Error{AMG_E_ARG, "unknown option"};
        }
        clear_graphs();
    }

    void set_level0(const amg_sell_desc* A) override {
        if (A == nullptr) {
            lv[0].A = A0_orig;
            if constexpr (std::is_same<T, double>::value) {
                if (outer_is_level0) outer = A0_orig;
            }
        } else {
            ARGCHK(A->nrows == lv[0].n && A->ncols == lv[0].n, "level-0 operator shape mismatch");
            lv[0].A = handle_sell<T>(*A);
            if (outer_is_level0) outer = make_sell<double>(*A);
        }
        clear_graphs();  // captured graphs hold the old operator
    }

    // Mean duration of the float64 CG update inside an eager preconditioned CG
    // solve (V-cycle, all columns active: b = 1, tol = 0) over its first `reps`
    // iterations: the kernel in its solve context, where its inputs were just
    // written by the kernels before it.
    // Mean duration of the launches of one kernel kind inside an eager solve
    // (b = 1, tol = 0, V-cycle, `reps` iterations), each bracketed by events:
    // EV_PCG_UPDATE the float64 CG update of pcg; EV_RU_UPDATE and EV_RU_GAMMA the
    // update and the last sweep with the gamma sums of the reliable-update CG.
    float time_in_solve(int kind, int k, int reps, hipStream_t s, bool cold = false) {
        ARGCHK(has_outer, "the in-solve timings need a handle with an outer operator");
        Workspace<T>& w = get_ws(k);
        const int n0 = lv[0].n;
        const size_t nk = (size_t)n0 * k;
        if (kind != EV_PCG_UPDATE) ensure_ipcg(w);
        k_fill<double><<<(unsigned)((nk + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(w.ob, nk, 1.0);
        HIPCHK(hipGetLastError());
        k_set_params<<<1, 1, 0, s>>>(w.ctl, 0.0, 0.0, reps + 1, 0);
        k_ru_params<<<(k + 255) / 256, 256, 0, s>>>(w.ctl, 1e-3, k);
        HIPCHK(hipGetLastError());
        std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
        // cache-cold: 512 MB (twice the Infinity Cache) read before each timed launch
        const size_t nflush = cold ? (size_t(512) << 20) / sizeof(float4) : 0;
        DevMem flush(nflush * sizeof(float4));
        if (cold) HIPCHK(hipMemsetAsync(flush.p, 0, nflush * sizeof(float4), s));
        float total = 0.f;
        try {
            for (auto& e : ev) HIPCHK(hipEventCreate(&e));
            if (kind == EV_PCG_UPDATE) {
                emit_pcg_init(w, k, s);
                emit_pcg_start(w, 'V', 1, k, 1, s);
            } else {
                emit_ru_init(w, k, w.ob, s);
                emit_ru_start(w, 'V', 1, k, s);
            }
            ev_hook_ = EvHook{ev.data(), 0, 2 * reps, kind};
            if (cold) {
                ev_hook_.flush = static_cast<const float4*>(flush.p);
                ev_hook_.nflush = nflush;
            }
            for (int i = 0; i < reps; ++i) {
                if (kind == EV_PCG_UPDATE) emit_pcg_iter(w, 'V', 1, k, 1, 0, s);
                else emit_ru_iter(w, 'V', 1, k, 0, s);
            }
            const int recorded = ev_hook_.n;
            ev_hook_ = EvHook{};
            if (recorded != 2 * reps) throw Error{AMG_E_UNSUPPORTED, "the timed kernel does not run in this solve"};
            HIPCHK(hipEventSynchronize(ev.back()));
            for (int i = 0; i < reps; ++i) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
                total += ms;
            }
        } catch (...) {
            ev_hook_ = EvHook{};
            (void)hipStreamSynchronize(s);
            for (auto e : ev)
                if (e) (void)hipEventDestroy(e);
            throw;
        }
        for (auto e : ev) (void)hipEventDestroy(e);
        return total / reps;
    }

    float time_kernel(int kernel, int level, int k, int reps, hipStream_t s) override {
        ARGCHK(level >= 0 && level < nlev, "level out of range");
        ARGCHK(reps > 0, "reps must be > 0");
        const bool cold = (kernel & AMG_TIME_COLD) != 0;
        kernel &= ~AMG_TIME_COLD;
        ARGCHK(kernel >= 0 && kernel <= 11, "kernel must be 0..11");
        if (kernel == 8) return time_in_solve(EV_PCG_UPDATE, k, reps, s, cold);
        if (kernel == 9) return time_in_solve(EV_RU_UPDATE, k, reps, s, cold);
        if (kernel == 10) return time_in_solve(EV_RU_GAMMA, k, reps, s, cold);
        ARGCHK(kernel <= 2 || kernel == 5 || level < nlev - 1, "the coarsest level has no transfer operators");
        ARGCHK(kernel != 6 || lv[level].D.nrows > 0, "level has no fused down-sweep operator");
        ARGCHK(kernel != 5 || (has_outer && level == 0), "the outer update runs on level 0 of a handle with an outer operator");
        ARGCHK(kernel != 11 || level == 0, "the reliable-update CG update runs on level 0");
        Workspace<T>& w = get_ws(k);
        LevelT<T>& L = lv[level];
        HIPCHK(hipMemsetAsync(w.X[level], 0, sizeof(T) * (size_t)L.n * k, s));
        HIPCHK(hipMemsetAsync(w.B[level], 0, sizeof(T) * (size_t)L.n * k, s));
        if (level + 1 < nlev) HIPCHK(hipMemsetAsync(w.X[level + 1], 0, sizeof(T) * (size_t)lv[level + 1].n * k, s));
        if (kernel == 5) {
            // every column active with alpha = 0: the full read/write traffic, values unchanged
            const size_t nk = (size_t)L.n * k;
            HIPCHK(hipMemsetAsync(w.op[0], 0, sizeof(double) * nk, s));
            HIPCHK(hipMemsetAsync(w.oAp, 0, sizeof(double) * nk, s));
            HIPCHK(hipMemsetAsync(w.ox, 0, sizeof(double) * nk, s));
            HIPCHK(hipMemsetAsync(w.orr, 0, sizeof(double) * nk, s));
            HIPCHK(hipMemsetD32Async((hipDeviceptr_t)w.ctl.active, 1, k, s));
            HIPCHK(hipMemsetAsync(w.ctl.alpha, 0, sizeof(double) * k, s));
            HIPCHK(hipMemsetAsync(w.ctl.done, 0, sizeof(int), s));
        }
        if (kernel == 11) {
            // every column active with alpha = 0: the full read/write traffic, values unchanged
            ensure_ipcg(w);
            const size_t nk = (size_t)L.n * k;
            HIPCHK(hipMemsetAsync(w.ip, 0, sizeof(T) * nk, s));
            HIPCHK(hipMemsetAsync(w.iap, 0, sizeof(T) * nk, s));
            HIPCHK(hipMemsetAsync(w.ix, 0, sizeof(T) * nk, s));
            HIPCHK(hipMemsetAsync(w.B0, 0, sizeof(T) * nk, s));
            HIPCHK(hipMemsetD32Async((hipDeviceptr_t)w.ctl.active, 1, k, s));
            HIPCHK(hipMemsetAsync(w.ctl.alpha, 0, sizeof(double) * k, s));
            HIPCHK(hipMemsetAsync(w.ctl.done, 0, sizeof(int), s));
        }
        const int kp = kp_of(k);
        T* rc = std::is_same<T, double>::value ? nullptr : w.B0;
        auto once = [&]() {
            switch (kernel) {
                case 0: spmv_launch(L.A, w.X[level], w.Tm[level], k, false, nullptr, s); break;
                case 1: jacobi_launch(level, w.X[level], w.B[level], w.Tm[level], T(0.5), k, nullptr, nullptr, s); break;
                case 2: residual_launch(level, w.X[level], w.B[level], w.Rr[level], k, nullptr, s); break;
                case 3:
                    residual_launch(level, w.X[level], w.B[level], w.Rr[level], k, nullptr, s);
                    spmv_launch(L.R, w.Rr[level], w.B[level + 1], k, false, nullptr, s);
                    break;
                case 4: spmv_launch(L.P, w.X[level + 1], w.Tm[level], k, true, nullptr, s); break;
                case 5:
                    KP_DISPATCH(kp, {
                        k_pcg_update<T, KP><<<red_grid(L.n, k), BLOCK, 0, s>>>(w.op[0], w.oAp, w.ox, w.orr, rc,
                                                                             L.n, k, w.ctl);
                    });
                    HIPCHK(hipGetLastError());
                    break;
                case 6: spmv_launch(L.D, w.B[level], w.B[level + 1], k, false, nullptr, s); break;
                case 11: ru_update_launch(w, k, s); break;
                default: prolong_x1_launch(level, w.B[level], w.X[level + 1], w.Tm[level], T(0.5), k, nullptr, s); break;
            }
        };
        once();
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        float ms = 0.f;
        if (!cold) {
            HIPCHK(hipEventRecord(e0, s));
            for (int i = 0; i < reps; ++i) once();
            HIPCHK(hipEventRecord(e1, s));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        } else {
            // cache-cold: 512 MB (twice the Infinity Cache) read before each timed
            // launch, so the launch reads its inputs from HBM (the flush leaves clean
            // lines: no write-back is charged to the timed launch)
            const size_t nflush = (size_t(512) << 20) / sizeof(float4);
            DevMem flush(nflush * sizeof(float4));
            HIPCHK(hipMemsetAsync(flush.p, 0, nflush * sizeof(float4), s));
            for (int i = 0; i < reps; ++i) {
                k_flush_read<<<FLUSH_BLOCKS, BLOCK, 0, s>>>(static_cast<const float4*>(flush.p), nflush,
                                                            static_cast<float*>(flush.p));
                HIPCHK(hipGetLastError());
                HIPCHK(hipEventRecord(e0, s));
                once();
                HIPCHK(hipEventRecord(e1, s));
                HIPCHK(hipEventSynchronize(e1));
                float one = 0.f;
                HIPCHK(hipEventElapsedTime(&one, e0, e1));
                ms += one;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return ms / reps;
    }
};

template <typename T>
HandleBase* create_typed(const amg_level_desc* levels, int nlev, const amg_coarse_desc* coarse,
                         const amg_sell_desc* outer) {
    std::unique_ptr<Handle<T>> h(new Handle<T>());
    h->dtype = std::is_same<T, double>::value ? AMG_F64 : AMG_F32;
    h->nlev = nlev;
    for (int l = 0; l < nlev; ++l) {
        const amg_level_desc& d = levels[l];
        LevelT<T> L;
        L.n = d.A.nrows;
        ARGCHK(d.A.nrows == d.A.ncols && d.A.nrows > 0, "level A must be square and non-empty");
        ARGCHK(d.dinv != nullptr, "null D^-1");
        L.A = handle_sell<T>(d.A);
        L.nnzA = d.A.nnz;
        L.padA = d.A.padded;
        L.dinv = static_cast<const T*>(d.dinv);
        if (l + 1 < nlev) {
            const int nn = levels[l + 1].A.nrows;
            ARGCHK(d.P.nrows == L.n && d.P.ncols == nn, "P shape mismatch");
            ARGCHK(d.R.nrows == nn && d.R.ncols == L.n, "R shape mismatch");
            L.P = handle_sell<T>(d.P);
            L.R = handle_sell<T>(d.R);
            L.nnzP = d.P.nnz;
            L.nnzR = d.R.nnz;
            L.padP = d.P.padded;
            L.padR = d.R.padded;
            ARGCHK(d.pre_kind >= AMG_SMOOTH_NONE && d.pre_kind <= AMG_SMOOTH_POLY, "bad pre-smoother kind");
            ARGCHK(d.post_kind >= AMG_SMOOTH_NONE && d.post_kind <= AMG_SMOOTH_POLY, "bad post-smoother kind");
            ARGCHK(d.pre_iters >= 0 && d.post_iters >= 0, "negative smoother iterations");
            L.pre = Smoother{d.pre_kind, d.pre_iters, d.pre_omega, {}};
            L.post = Smoother{d.post_kind, d.post_iters, d.post_omega, {}};
            if (d.pre_kind == AMG_SMOOTH_POLY) {
                ARGCHK(d.pre_coef != nullptr && d.pre_ncoef >= 1 && d.pre_ncoef <= AMG_POLY_MAX,
                       "polynomial pre-smoother needs 1..AMG_POLY_MAX coefficients");
                L.pre.coef.assign(d.pre_coef, d.pre_coef + d.pre_ncoef);
            }
            if (d.post_kind == AMG_SMOOTH_POLY) {
                ARGCHK(d.post_coef != nullptr && d.post_ncoef >= 1 && d.post_ncoef <= AMG_POLY_MAX,
                       "polynomial post-smoother needs 1..AMG_POLY_MAX coefficients");
                L.post.coef.assign(d.post_coef, d.post_coef + d.post_ncoef);
            }
            if (d.D.nrows > 0) {
                ARGCHK(d.D.nrows == nn && d.D.ncols == L.n, "fused down-sweep operator shape mismatch");
                L.D = handle_sell<T>(d.D);
            }
        }
        h->lv.push_back(L);
    }
    h->A0_orig = h->lv[0].A;
    h->cd = *coarse;
    ARGCHK(coarse->n == h->lv.back().n, "coarse solver size mismatch");
    ARGCHK(coarse->kind >= AMG_COARSE_PINV && coarse->kind <= AMG_COARSE_JACOBI, "unknown coarse solver");
    if (coarse->kind != AMG_COARSE_JACOBI) ARGCHK(coarse->f0 != nullptr, "null coarse factor");
    if (coarse->kind == AMG_COARSE_LU || coarse->kind == AMG_COARSE_QR) ARGCHK(coarse->f1 != nullptr, "null coarse factor");
    h->cf0 = coarse->f0;
    h->cf64 = coarse->factor_f64 != 0;
    if (coarse->kind == AMG_COARSE_QR) h->cf1 = coarse->f1;
    if (coarse->kind == AMG_COARSE_LU) h->cpiv = static_cast<const int*>(coarse->f1);
    if (outer != nullptr) {
        ARGCHK(outer->nrows == h->lv[0].n && outer->ncols == h->lv[0].n, "outer operator shape mismatch");
        h->outer = handle_sell<double>(*outer);
        h->outer_nnz = outer->nnz;
        h->has_outer = true;
    } else if (std::is_same<T, double>::value) {
        h->outer = make_sell<double>(levels[0].A);
        h->outer_nnz = levels[0].A.nnz;
        h->has_outer = true;
        h->outer_is_level0 = true;
    }
    // the allocation-free cycle's view: host arrays sized once, here
    for (auto* v : {&h->nw_.X, &h->nw_.Tm, &h->nw_.B, &h->nw_.Rr, &h->nw_.H0, &h->nw_.H1}) v->assign(nlev, nullptr);
    (void)scratch_ptr();  // resolve this device's scratch address at set-up
    {
        int dev = 0, lds = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        h->tail_lds_max_ = (size_t)std::max(lds, 0);
    }
    h->plan_tail();
    return h.release();
}

// ------------------------------------------------------------------ batch
//
// A batch of independent AMG-preconditioned CG solves (one hierarchy and one
// block width k per problem).  At creation every problem's PCG sequence is
// recorded once, the steps of all problems are merged by tag, and each merged
// phase becomes one kb_phase launch whose step table lives in device memory;
// the start and iteration sequences are captured as HIP graphs.  A problem's
// kernels still stop on its own `done` flag; the batch loop stops when the
// finalisers have counted every problem done.

struct BatchBase {
    virtual ~BatchBase() = default;
    int nprob = 0;
    int dtype = AMG_F64;
    virtual void pcg(const double* const* b, double* const* x, double tol, double atol, int maxiter, int flexible,
                     int* iters, double* relres, hipStream_t s) = 0;
    virtual void set_option(int opt, long long value) = 0;
    virtual void info(int* nlaunch_iter, int* total_blocks_iter) = 0;
};

// one merged launch: the steps of one phase of every problem that has it
struct Launch {
    int kid = 0;
    int kpc = 0;    // column lanes shared by every step, 0: mixed
    int total = 0;  // virtual blocks, a multiple of 8
    const Step* steps = nullptr;
    const int* gmap = nullptr;
};

inline int round8(int v) { return (v + 7) & ~7; }

template <typename T>
struct Batch : BatchBase {
    std::vector<Handle<T>*> hs;
    std::vector<int> ks;
    std::vector<std::unique_ptr<Workspace<T>>> ws;
    std::vector<std::unique_ptr<DevMem>> tables;      // flags and per-call tables (kept)
    std::vector<std::unique_ptr<DevMem>> seg_tables;  // step tables of the recorded sequences
    std::vector<Launch> seg_init, seg_final, seg_dyn;
    std::map<int, std::vector<Launch>> seg_start, seg_iter;  // by flexible
    std::map<int, hipGraphExec_t> g_start, g_iter;
    std::vector<hipGraph_t> raw;
    hipStream_t cap = nullptr;
    int grid_cap = 0;  // persistent grid size (0: one block per virtual block)
    bool split_kp = true;  // one launch per column-lane count within a phase
    bool autotune = true;  // row kernels take the autotuner's plans
    char cyc = 'V';        // cycle of the preconditioner (AMG_BATCH_OPT_CYCLE)
    int cpl = 1;           // F-cycle: V-cycles per level (AMG_BATCH_OPT_CPL)

    static bool kp_dependent(int kid) {
        return !(kid == K_COARSE_PINV || kid == K_COARSE_PINV64 || kid == K_COARSE_LU || kid == K_COARSE_LU64 ||
                 kid == K_COARSE_QR || kid == K_COARSE_QR64 || kid == K_FINALIZE || kid == K_SET_PARAMS ||
                 kid == K_REPORT);
    }
    // batch completion flag
    DevMem* flags = nullptr;
    int* hflag = nullptr;
    BatchFlag bf{};
    // per-call steps (copy-in, parameters, copy-out, report) and the report buffers
    std::vector<Step> dyn_host;
    Step* dyn_pin = nullptr;
    int* rep_it = nullptr;
    double* rep_rr = nullptr;
    int* rep_it_h = nullptr;
    double* rep_rr_h = nullptr;
    int ktot = 0;

    ~Batch() override {
        for (auto& kv : g_start) (void)hipGraphExecDestroy(kv.second);
        for (auto& kv : g_iter) (void)hipGraphExecDestroy(kv.second);
        for (auto g : raw) (void)hipGraphDestroy(g);
        if (cap) (void)hipStreamDestroy(cap);
        if (hflag) (void)hipHostFree(hflag);
        if (dyn_pin) (void)hipHostFree(dyn_pin);
        if (rep_it_h) (void)hipHostFree(rep_it_h);
        if (rep_rr_h) (void)hipHostFree(rep_rr_h);
    }

    // Record `emit(p, handle, workspace)` for every problem and merge the steps.
    template <typename E>
    std::vector<Launch> build(E&& emit) {
        struct Item {
            int stage, lkey, sub, kid, kpc, prob;
            Step st;
        };
        std::vector<Item> items;
        for (int p = 0; p < nprob; ++p) {
            Recorder r;
            r.tuned = autotune;
            r.vec = hs[p]->vec_;
            hs[p]->rec_ = &r;
            try {
                emit(p, *hs[p], *ws[p]);
            } catch (...) {
                hs[p]->rec_ = nullptr;
                throw;
            }
            hs[p]->rec_ = nullptr;
            // V-cycles merge by tag (stage, level, sub-step), which lines the levels of
            // problems of different depth up; W and F cycles revisit levels, so their
            // steps merge by position in each problem's own sequence (any merge that
            // keeps every problem's order is valid)
            int pos = 0;
            for (const RecStep& rs : r.steps) {
                const int kpc = split_kp && kp_dependent(rs.kid) ? rs.st.kp : 0;
                if (cyc == 'V')
                    items.push_back(Item{rs.stage, rs.stage == ST_UP ? -rs.level : rs.level, rs.sub, rs.kid, kpc, p,
                                         rs.st});
                else
                    items.push_back(Item{0, pos, 0, rs.kid, kpc, p, rs.st});
                ++pos;
            }
        }
        std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
            return std::tie(a.stage, a.lkey, a.sub, a.kid, a.kpc, a.prob) <
                   std::tie(b.stage, b.lkey, b.sub, b.kid, b.kpc, b.prob);
        });
        std::vector<Launch> out;
        size_t i = 0;
        while (i < items.size()) {
            size_t j = i;
            while (j < items.size() && items[j].stage == items[i].stage && items[j].lkey == items[i].lkey &&
                   items[j].sub == items[i].sub && items[j].kid == items[i].kid && items[j].kpc == items[i].kpc)
                ++j;
            std::vector<Step> steps;
            std::vector<int> gmap;
            int total = 0;
            for (size_t q = i; q < j; ++q) {
                Step st = items[q].st;
                st.blk0 = total;
                const int nb = round8(st.nbx * st.nby);
                for (int g = 0; g < nb / 8; ++g) gmap.push_back((int)steps.size());
                total += nb;
                steps.push_back(st);
            }
            const size_t sb = sizeof(Step) * steps.size(), mb = sizeof(int) * gmap.size();
            seg_tables.emplace_back(new DevMem(sb + mb));
            char* base = static_cast<char*>(seg_tables.back()->p);
            HIPCHK(hipMemcpy(base, steps.data(), sb, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(base + sb, gmap.data(), mb, hipMemcpyHostToDevice));
            Launch L;
            L.kid = items[i].kid;
            L.total = total;
            // a launch whose steps share one column-lane count runs the specialised kernel
            bool one_kp = kp_dependent(L.kid);
            for (const Step& st : steps) one_kp = one_kp && st.kp == steps[0].kp;
            L.kpc = one_kp ? steps[0].kp : 0;
            L.steps = reinterpret_cast<const Step*>(base);
            L.gmap = reinterpret_cast<const int*>(base + sb);
            out.push_back(L);
            i = j;
        }
        return out;
    }

    void run(const Launch& L, hipStream_t s) {
        int grid = L.total;
        if (grid_cap > 0 && grid > grid_cap) grid = grid_cap;
        if (grid <= 0) return;
        const int rc = kb_launch<T>(L.kid, L.kpc, grid, L.steps, L.gmap, L.total, bf, s);
        if (rc != 0) throw Error{AMG_E_STATE, "unknown batched kernel"};
        HIPCHK(hipGetLastError());
    }
    void run_all(const std::vector<Launch>& v, hipStream_t s) {
        for (const Launch& L : v) run(L, s);
    }

    hipGraphExec_t capture(std::map<int, hipGraphExec_t>& cache, int key, const std::vector<Launch>& v, int reps) {
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
        if (!cap) HIPCHK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        try {
            for (int r = 0; r < reps; ++r) run_all(v, cap);
        } catch (...) {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(cap, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            throw;
        }
        HIPCHK(hipStreamEndCapture(cap, &g));
        hipGraphExec_t ex = nullptr;
        HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        raw.push_back(g);
        cache[key] = ex;
        return ex;
    }

    // Drop the captured graphs (queued launches may still run them: drain first).
    void clear_graphs() {
        if (!g_start.empty() || !g_iter.empty() || !raw.empty()) HIPCHK(hipDeviceSynchronize());
        for (auto& kv : g_start) (void)hipGraphExecDestroy(kv.second);
        for (auto& kv : g_iter) (void)hipGraphExecDestroy(kv.second);
        for (auto g : raw) (void)hipGraphDestroy(g);
        g_start.clear();
        g_iter.clear();
        raw.clear();
    }

    const std::vector<Launch>& start_seq(int flexible) {
        (void)flexible;  // the start sequence does not depend on it
        auto it = seg_start.find(0);
        if (it != seg_start.end()) return it->second;
        seg_start[0] = build([&](int, Handle<T>& h, Workspace<T>& w) { h.emit_pcg_start(w, cyc, cpl, w.k, 1, nullptr); });
        return seg_start[0];
    }
    const std::vector<Launch>& iter_seq(int flexible) {
        auto it = seg_iter.find(flexible);
        if (it != seg_iter.end()) return it->second;
        seg_iter[flexible] = build([&](int, Handle<T>& h, Workspace<T>& w) {
            h.emit_pcg_iter(w, cyc, cpl, w.k, 1, flexible, nullptr);
        });
        return seg_iter[flexible];
    }

    // Re-record every sequence; the old step tables are released (after the
    // device has drained, in clear_graphs) instead of accumulating.
    void rebuild() {
        clear_graphs();
        seg_start.clear();
        seg_iter.clear();
        seg_tables.clear();
        seg_init = build([&](int, Handle<T>& h, Workspace<T>& w) { h.emit_pcg_init(w, w.k, nullptr); });
        seg_final = build([&](int, Handle<T>& h, Workspace<T>& w) { h.emit_pcg_final(w, w.k, w.ob, nullptr); });
    }

    void setup() {
        for (int p = 0; p < nprob; ++p) ws.push_back(hs[p]->make_ws(ks[p]));
        if (autotune)
            for (int p = 0; p < nprob; ++p) hs[p]->tune_all(ks[p]);
        seg_init = build([&](int, Handle<T>& h, Workspace<T>& w) { h.emit_pcg_init(w, w.k, nullptr); });
        seg_final = build([&](int, Handle<T>& h, Workspace<T>& w) { h.emit_pcg_final(w, w.k, w.ob, nullptr); });
        start_seq(0);
        iter_seq(0);
        // batch flag: count, done (device), mirror (mapped host)
        flags = new DevMem(sizeof(int) * 4);
        tables.emplace_back(flags);
        HIPCHK(hipMemset(flags->p, 0, sizeof(int) * 4));
        HIPCHK(hipHostMalloc((void**)&hflag, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
        *hflag = 0;
        bf.count = static_cast<int*>(flags->p);
        bf.done = bf.count + 1;
        HIPCHK(hipHostGetDevicePointer((void**)&bf.hdone, hflag, 0));
        bf.n = nprob;
        // per-call launches: copy-in, parameters, copy-out, report (steps patched per call)
        for (int p = 0; p < nprob; ++p) ktot += ks[p];
        tables.emplace_back(new DevMem(sizeof(int) * (size_t)(ktot + 1) + sizeof(double) * (size_t)ktot));
        rep_it = static_cast<int*>(tables.back()->p);
        rep_rr = reinterpret_cast<double*>(rep_it + ktot + (ktot & 1));
        HIPCHK(hipHostMalloc((void**)&rep_it_h, sizeof(int) * (size_t)ktot, 0));
        HIPCHK(hipHostMalloc((void**)&rep_rr_h, sizeof(double) * (size_t)ktot, 0));
        const int kids[4] = {K_COPY64, K_SET_PARAMS, K_COPY64, K_REPORT};
        std::vector<std::vector<int>> gm(4);
        int off = 0;
        for (int q = 0; q < 4; ++q) {
            int total = 0;
            for (int p = 0; p < nprob; ++p) {
                const int n0 = hs[p]->lv[0].n, k = ks[p];
                Step st = hs[p]->step0(n0, k);
                const dim3 g = (kids[q] == K_COPY64) ? row_grid(n0, k) : dim3(1, 1);
                st.nbx = (int)g.x;
                st.nby = (int)g.y;
                st.blk0 = total;
                st.ctl = ws[p]->ctl;
                const int nb = round8(st.nbx * st.nby);
                for (int i = 0; i < nb / 8; ++i) gm[q].push_back(p);
                total += nb;
                dyn_host.push_back(st);
            }
            Launch L;
            L.kid = kids[q];
            L.total = total;
            seg_dyn.push_back(L);
            off += nprob;
        }
        (void)off;
        size_t mb = 0;
        for (auto& g : gm) mb += sizeof(int) * g.size();
        const size_t sb = sizeof(Step) * dyn_host.size();
        tables.emplace_back(new DevMem(sb + mb));
        char* base = static_cast<char*>(tables.back()->p);
        size_t mo = sb;
        for (int q = 0; q < 4; ++q) {
            seg_dyn[q].steps = reinterpret_cast<const Step*>(base) + (size_t)q * nprob;
            HIPCHK(hipMemcpy(base + mo, gm[q].data(), sizeof(int) * gm[q].size(), hipMemcpyHostToDevice));
            seg_dyn[q].gmap = reinterpret_cast<const int*>(base + mo);
            mo += sizeof(int) * gm[q].size();
        }
        HIPCHK(hipHostMalloc((void**)&dyn_pin, sb, 0));
        HIPCHK(hipDeviceSynchronize());
    }

    void pcg(const double* const* b, double* const* x, double tol, double atol, int maxiter, int flexible, int* iters,
             double* relres, hipStream_t s) override {
        // patch the per-call steps and upload them in one copy
        int off = 0, max_it = 0;
        for (int p = 0; p < nprob; ++p) {
            Step& cin = dyn_host[p];
            cin.in0 = b[p];
            cin.out0 = ws[p]->ob;
            Step& prm = dyn_host[nprob + p];
            prm.tol = tol;
            prm.atol = atol;
            // maxiter < 0: each problem's own default, 10 n_p (that of a single pcg)
            prm.maxiter = maxiter >= 0 ? maxiter : 10 * hs[p]->lv[0].n;
            max_it = std::max(max_it, prm.maxiter);
            prm.ival = p == 0 ? 1 : 0;
            Step& cout = dyn_host[2 * nprob + p];
            cout.in0 = ws[p]->ox;
            cout.out0 = x[p];
            Step& rep = dyn_host[3 * nprob + p];
            rep.out0 = rep_it;
            rep.out1 = rep_rr;
            rep.ival = off;
            off += ks[p];
        }
        std::memcpy(dyn_pin, dyn_host.data(), sizeof(Step) * dyn_host.size());
        HIPCHK(hipMemcpyAsync(const_cast<Step*>(seg_dyn[0].steps), dyn_pin, sizeof(Step) * dyn_host.size(),
                              hipMemcpyHostToDevice, s));
        run(seg_dyn[0], s);
        run(seg_dyn[1], s);
        run_all(seg_init, s);
        HIPCHK(hipGraphLaunch(capture(g_start, 0, start_seq(flexible), 1), s));
        hipGraphExec_t ex = capture(g_iter, flexible, iter_seq(flexible), Handle<T>::GRAPH_ITERS);
        run_until_done(hflag, ex, (max_it + Handle<T>::GRAPH_ITERS - 1) / Handle<T>::GRAPH_ITERS + 1, s);
        run_all(seg_final, s);
        run(seg_dyn[2], s);
        run(seg_dyn[3], s);
        HIPCHK(hipMemcpyAsync(rep_it_h, rep_it, sizeof(int) * (size_t)ktot, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(rep_rr_h, rep_rr, sizeof(double) * (size_t)ktot, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (iters) std::memcpy(iters, rep_it_h, sizeof(int) * (size_t)ktot);
        if (relres) std::memcpy(relres, rep_rr_h, sizeof(double) * (size_t)ktot);
    }

    void set_option(int opt, long long value) override {
        if (opt == AMG_BATCH_OPT_GRID) {
            ARGCHK(value >= 0 && value % 8 == 0, "persistent grid size must be a non-negative multiple of 8");
            grid_cap = (int)value;
            clear_graphs();
        } else if (opt == AMG_BATCH_OPT_SPLIT_KP) {
            split_kp = value != 0;
            rebuild();
        } else if (opt == AMG_BATCH_OPT_AUTOTUNE) {
            autotune = value != 0;
            if (autotune)
                for (int p = 0; p < nprob; ++p) hs[p]->tune_all(ks[p]);
            rebuild();
        } else if (opt == AMG_BATCH_OPT_CYCLE) {
            ARGCHK(value == 'V' || value == 'W' || value == 'F', "cycle must be 'V', 'W' or 'F'");
            cyc = (char)value;
            rebuild();
        } else if (opt == AMG_BATCH_OPT_CPL) {
            ARGCHK(value >= 0, "cycles_per_level must be >= 0");
            cpl = (int)value;
            rebuild();
        } else {
            throw Error{AMG_E_ARG, "unknown batch option"};
        }
    }

    void info(int* nlaunch_iter, int* total_blocks_iter) override {
        const std::vector<Launch>& v = iter_seq(0);
        if (nlaunch_iter) *nlaunch_iter = (int)v.size();
        if (total_blocks_iter) {
            long long t = 0;
            for (const Launch& L : v) t += L.total;
            *total_blocks_iter = (int)t;
        }
    }
};

template <typename T>
BatchBase* create_batch(const amg_handle* handles, const int* k, int nprob) {
    std::unique_ptr<Batch<T>> b(new Batch<T>());
    b->nprob = nprob;
    b->dtype = std::is_same<T, double>::value ? AMG_F64 : AMG_F32;
    for (int p = 0; p < nprob; ++p) {
        Handle<T>* h = static_cast<Handle<T>*>(reinterpret_cast<HandleBase*>(handles[p]));
        ARGCHK(h->has_outer, "every handle of a batch needs the float64 outer operator");
        ARGCHK(k[p] > 0, "k must be > 0");
        b->hs.push_back(h);
        b->ks.push_back(k[p]);
    }
    b->setup();
    return b.release();
}

template <typename F>
int guard(F&& f) {
    try {
        f();
        return AMG_OK;
    } catch (const Error& e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return AMG_E_STATE;
    } catch (...) {
        g_err = "unknown error";
        return AMG_E_STATE;
    }
}

HandleBase* H(amg_handle h) {
    if (h == nullptr) throw Error{AMG_E_ARG, "null handle"};
    return reinterpret_cast<HandleBase*>(h);
}

void check_cycle(char c, int cpl) {
    ARGCHK(c == 'V' || c == 'W' || c == 'F', std::string("Cycle '") + c + "' not supported. Use 'V', 'W', or 'F'.");
    ARGCHK(cpl >= 0, "cycles_per_level must be >= 0");
}

// One cycle may also be 'T', the transposed F-cycle (the adjoint of fixed F-cycles).
void check_cycle_t(char c, int cpl) {
    if (c != 'T') check_cycle(c, cpl);
    ARGCHK(cpl >= 0, "cycles_per_level must be >= 0");
}

}  // namespace

extern "C" {

int amg_create(const amg_level_desc* levels, int nlev, const amg_coarse_desc* coarse, int dtype,
               const amg_sell_desc* outer, amg_handle* out) {
    return guard([&] {
        ARGCHK(levels != nullptr && coarse != nullptr && out != nullptr, "null argument");
        ARGCHK(nlev >= 1, "need at least one level");
        HandleBase* h = nullptr;
        if (dtype == AMG_F64) h = create_typed<double>(levels, nlev, coarse, outer);
        else if (dtype == AMG_F32) h = create_typed<float>(levels, nlev, coarse, outer);
        else throw Error{AMG_E_ARG, "dtype must be AMG_F64 or AMG_F32"};
        *out = reinterpret_cast<amg_handle>(h);
    });
}

int amg_destroy(amg_handle h) {
    return guard([&] { delete H(h); });
}

int amg_cycle(amg_handle h, char cycle, int cpl, const void* b, void* x, int k, int x_zero, void* stream) {
    return guard([&] {
        check_cycle_t(cycle, cpl);
        ARGCHK(b && x && k > 0, "bad b/x/k");
        ARGCHK(aligned16(b) && aligned16(x), "b and x must be 16-byte aligned");
        H(h)->cycle(cycle, cpl, b, x, k, x_zero, static_cast<hipStream_t>(stream));
    });
}

int amg_cycle_noalloc(amg_handle h, char cycle, int cpl, const void* b, void* x, int k, int x_zero, void* stream) {
    return guard([&] {
        check_cycle_t(cycle, cpl);
        ARGCHK(b && x && k > 0, "bad b/x/k");
        ARGCHK(aligned16(b) && aligned16(x), "b and x must be 16-byte aligned");
        HandleBase* hb = H(h);
        hb->cycle_noalloc(cycle, cpl, b, x, k, x_zero, scratch_ptr(), NOALLOC_SCRATCH_BYTES,
                          static_cast<hipStream_t>(stream));
    });
}

int amg_solve(amg_handle h, char cycle, int cpl, const void* b, void* x, int k, double tol, int maxiter, int refine,
              int* iters_out, double* relres_out, void* stream) {
    return guard([&] {
        check_cycle(cycle, cpl);
        ARGCHK(b && x && k > 0, "bad b/x/k");
        ARGCHK(aligned16(b) && aligned16(x), "b and x must be 16-byte aligned");
        ARGCHK(maxiter >= 0, "maxiter must be >= 0");
        H(h)->solve(cycle, cpl, b, x, k, tol, maxiter, refine, iters_out, relres_out,
                    static_cast<hipStream_t>(stream));
    });
}

int amg_pcg(amg_handle h, char cycle, int cpl, const double* b, double* x, int k, double tol, double atol,
            int maxiter, int flexible, int precond, int* iters_out, double* relres_out, void* stream) {
    return guard([&] {
        check_cycle(cycle, cpl);
        ARGCHK(b && x && k > 0, "bad b/x/k");
        ARGCHK(aligned16(b) && aligned16(x), "b and x must be 16-byte aligned");
        ARGCHK(maxiter >= 0, "maxiter must be >= 0");
        H(h)->pcg(cycle, cpl, b, x, k, tol, atol, maxiter, flexible, precond, iters_out, relres_out,
                  static_cast<hipStream_t>(stream));
    });
}

int amg_bicgstab(amg_handle h, char cycle, int cpl, const double* b, double* x, int k, double tol, double atol,
                 int maxiter, int precond, int* iters_out, double* relres_out, void* stream) {
    return guard([&] {
        check_cycle(cycle, cpl);
        ARGCHK(b && x && k > 0, "bad b/x/k");
        ARGCHK(aligned16(b) && aligned16(x), "b and x must be 16-byte aligned");
        ARGCHK(maxiter >= 0, "maxiter must be >= 0");
        H(h)->bicgstab(cycle, cpl, b, x, k, tol, atol, maxiter, precond, iters_out, relres_out,
                       static_cast<hipStream_t>(stream));
    });
}

int amg_pcg_ir(amg_handle h, char cycle, int cpl, const double* b, double* x, int k, double tol, double inner_tol,
               int maxiter, int inner_maxiter, int flexible, int* iters_out, int* outer_out, double* relres_out,
               void* stream) {
    return guard([&] {
        check_cycle(cycle, cpl);
        ARGCHK(b && x && k > 0, "bad b/x/k");
        ARGCHK(aligned16(b) && aligned16(x), "b and x must be 16-byte aligned");
        ARGCHK(maxiter >= 0 && inner_maxiter >= 1, "maxiter must be >= 0 and inner_maxiter >= 1");
        ARGCHK(inner_tol > 0.0 && inner_tol < 1.0, "inner_tol must be in (0, 1)");
        H(h)->pcg_ir(cycle, cpl, b, x, k, tol, inner_tol, maxiter, inner_maxiter, flexible, iters_out, outer_out,
                     relres_out, static_cast<hipStream_t>(stream));
    });
}

int amg_pcg_ru(amg_handle h, char cycle, int cpl, const double* b, double* x, int k, double tol, double atol,
               double delta, int maxiter, int flexible, int* iters_out, int* updates_out, double* relres_out,
               void* stream) {
    return guard([&] {
        check_cycle(cycle, cpl);
        ARGCHK(b && x && k > 0, "bad b/x/k");
        ARGCHK(aligned16(b) && aligned16(x), "b and x must be 16-byte aligned");
        ARGCHK(maxiter >= 0, "maxiter must be >= 0");
        ARGCHK(delta > 0.0 && delta < 1.0, "delta must be in (0, 1)");
        H(h)->pcg_ru(cycle, cpl, b, x, k, tol, atol, delta, maxiter, flexible ? 1 : 0, iters_out, updates_out,
                     relres_out, static_cast<hipStream_t>(stream));
    });
}

int amg_gmres(amg_handle h, char cycle, int cpl, const double* b, double* x, int k, double tol, double atol,
              int restart, int maxiter, int precond, int method, int* iters_out, double* relres_out, void* stream) {
    return guard([&] {
        check_cycle(cycle, cpl);
        ARGCHK(b && x && k > 0, "bad b/x/k");
        ARGCHK(aligned16(b) && aligned16(x), "b and x must be 16-byte aligned");
        ARGCHK(maxiter >= 0, "maxiter must be >= 0");
        ARGCHK(restart >= 1 && restart <= GMRES_MMAX, "restart must be 1..32");
        ARGCHK(method == 0 || method == 1, "method must be 0 (batched) or 1 (incremental)");
        H(h)->gmres(cycle, cpl, b, x, k, tol, atol, restart, maxiter, precond, method, iters_out, relres_out,
                    static_cast<hipStream_t>(stream));
    });
}

int amg_sell_spmv(const amg_sell_desc* M, int dtype, const void* x, void* y, int k, int add, void* stream) {
    return guard([&] {
        ARGCHK(M && x && y && k > 0, "bad arguments");
        ARGCHK(x != y, "x and y may not alias");
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (M->nrows <= 0) return;
        const int kp = kp_of(k);
        if (dtype == AMG_F64) {
            const Sell<double> S = make_sell<double>(*M);
            KP_DISPATCH(kp, WIDE_DISPATCH(S, {
                const dim3 g = row_grid(M->nrows, k, WD);
                if (add) k_spmv<double, KP, WD, true><<<g, BLOCK, 0, s>>>(S, (const double*)x, (double*)y, k, default_subl(KP, WD), nullptr);
                else k_spmv<double, KP, WD, false><<<g, BLOCK, 0, s>>>(S, (const double*)x, (double*)y, k, default_subl(KP, WD), nullptr);
            }));
        } else if (dtype == AMG_F32) {
            const Sell<float> S = make_sell<float>(*M);
            KP_DISPATCH(kp, WIDE_DISPATCH(S, {
                const dim3 g = row_grid(M->nrows, k, WD);
                if (add) k_spmv<float, KP, WD, true><<<g, BLOCK, 0, s>>>(S, (const float*)x, (float*)y, k, default_subl(KP, WD), nullptr);
                else k_spmv<float, KP, WD, false><<<g, BLOCK, 0, s>>>(S, (const float*)x, (float*)y, k, default_subl(KP, WD), nullptr);
            }));
        } else {
            throw Error{AMG_E_ARG, "dtype must be AMG_F64 or AMG_F32"};
        }
        HIPCHK(hipGetLastError());
    });
}

int amg_sell_jacobi(const amg_sell_desc* A, const void* dinv, int dtype, const void* x, const void* b, void* x_out,
                    int k, double omega, void* stream) {
    return guard([&] {
        ARGCHK(A && dinv && x && b && x_out && k > 0, "bad arguments");
        ARGCHK(x != x_out, "x and x_out may not alias");
        ARGCHK(A->nrows == A->ncols, "A must be square");
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (A->nrows <= 0) return;
        const int kp = kp_of(k);
        if (dtype == AMG_F64) {
            const Sell<double> S = make_sell<double>(*A);
            KP_DISPATCH(kp, WIDE_DISPATCH(S, {
                k_jacobi<double, KP, WD><<<row_grid(A->nrows, k, WD), BLOCK, 0, s>>>(S, (const double*)dinv, (const double*)x, (const double*)b,
                                                              (double*)x_out, omega, k, default_subl(KP, WD), nullptr, nullptr);
            }));
        } else if (dtype == AMG_F32) {
            const Sell<float> S = make_sell<float>(*A);
            KP_DISPATCH(kp, WIDE_DISPATCH(S, {
                k_jacobi<float, KP, WD><<<row_grid(A->nrows, k, WD), BLOCK, 0, s>>>(S, (const float*)dinv, (const float*)x, (const float*)b,
                                                             (float*)x_out, (float)omega, k, default_subl(KP, WD), nullptr, nullptr);
            }));
        } else {
            throw Error{AMG_E_ARG, "dtype must be AMG_F64 or AMG_F32"};
        }
        HIPCHK(hipGetLastError());
    });
}

int amg_sell_poly(const amg_sell_desc* A, int dtype, const double* coef, int ncoef, int iters, const void* x,
                  const void* b, void* x_out, void* work, int k, int x_zero, void* stream) {
    return guard([&] {
        ARGCHK(A && coef && x && b && x_out && work && k > 0, "bad arguments");
        ARGCHK(ncoef >= 1 && ncoef <= AMG_POLY_MAX, "ncoef must be 1..AMG_POLY_MAX");
        ARGCHK(iters >= 0, "iters must be >= 0");
        ARGCHK(x != x_out, "x and x_out may not alias");
        ARGCHK(A->nrows == A->ncols, "A must be square");
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (A->nrows <= 0) return;
        auto run = [&](auto tag) {
            using T = decltype(tag);
            const Sell<T> S = make_sell<T>(*A);
            const size_t nk = (size_t)A->nrows * k;
            if (iters == 0) {
                HIPCHK(hipMemcpyAsync(x_out, x, sizeof(T) * nk, hipMemcpyDeviceToDevice, s));
                return;
            }
            T* r = static_cast<T*>(work);
            T* h0 = r + nk;
            T* h1 = h0 + nk;
            T* tmp = h1 + nk;
            const Plan pl = default_plan(S.maxw, k, S.nrows, 0);
            const T* cur = static_cast<const T*>(x);
            bool zero = x_zero != 0;
            for (int it = 0; it < iters; ++it) {
                // the sweeps alternate between tmp and x_out so that the last lands in x_out
                T* dst = ((iters - 1 - it) % 2 == 0) ? static_cast<T*>(x_out) : tmp;
                launch_poly_sweep<T>(S, pl, coef, ncoef, zero ? nullptr : cur, static_cast<const T*>(b), dst, r, h0, h1,
                                     k, nullptr, nullptr, s);
                cur = dst;
                zero = false;
            }
        };
        if (dtype == AMG_F64) run(double{});
        else if (dtype == AMG_F32) run(float{});
        else throw Error{AMG_E_ARG, "dtype must be AMG_F64 or AMG_F32"};
    });
}

int amg_set_level0(amg_handle h, const amg_sell_desc* A) {
    return guard([&] { H(h)->set_level0(A); });
}

int amg_set_option(amg_handle h, int option, long long value) {
    return guard([&] { H(h)->set_option(option, value); });
}

int amg_tail_level(amg_handle h, int* level) {
    return guard([&] {
        ARGCHK(level != nullptr, "bad arguments");
        *level = H(h)->tail_level();
    });
}

int amg_dense_solve(const amg_coarse_desc* c, int dtype, const void* b, void* x, int k, void* stream) {
    return guard([&] {
        ARGCHK(c && b && x && k > 0 && b != x, "bad arguments");
        ARGCHK(c->n > 0 && c->f0 != nullptr, "bad coarse factor");
        ARGCHK(c->kind == AMG_COARSE_PINV || c->f1 != nullptr, "bad coarse factor");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const int n = c->n;
        auto run = [&](auto tag) {
            using T = decltype(tag);
            const T* f0 = static_cast<const T*>(c->f0);
            switch (c->kind) {
                case AMG_COARSE_PINV:
                    k_coarse_pinv<T, T><<<(n * k + BLOCK - 1) / BLOCK, BLOCK, 0, s>>>(f0, (const T*)b, (T*)x, n, k, nullptr);
                    break;
                case AMG_COARSE_LU:
                    k_coarse_lu<T, T><<<(k + BLOCK - 1) / BLOCK, BLOCK, 0, s>>>(f0, (const int*)c->f1, (const T*)b, (T*)x, (T*)x,
                                                                            n, k, nullptr);
                    break;
                case AMG_COARSE_QR:
                    k_coarse_qr<T, T><<<(k + BLOCK - 1) / BLOCK, BLOCK, 0, s>>>(f0, (const T*)c->f1, (const T*)b, (T*)x, (T*)x,
                                                                            n, k, nullptr);
                    break;
                default:
                    throw Error{AMG_E_ARG, "amg_dense_solve supports pinv, lu and qr"};
            }
        };
        if (dtype == AMG_F64) run(double{});
        else if (dtype == AMG_F32) run(float{});
        else throw Error{AMG_E_ARG, "dtype must be AMG_F64 or AMG_F32"};
        HIPCHK(hipGetLastError());
    });
}

int amg_grad_values(const int32_t* rows, const int32_t* cols, long long nnz, int dtype, const void* lam,
                    const void* x, int k, double scale, void* g, void* stream) {
    return guard([&] {
        ARGCHK(nnz >= 0 && k > 0, "bad sizes");
        if (nnz == 0) return;
        ARGCHK(rows && cols && lam && x && g, "null argument");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const unsigned nb = (unsigned)((nnz + BLOCK - 1) / BLOCK);
        if (dtype == AMG_F64)
            k_grad_values<double><<<nb, BLOCK, 0, s>>>(rows, cols, nnz, (const double*)lam, (const double*)x, k,
                                                       scale, (double*)g);
        else if (dtype == AMG_F32)
            k_grad_values<float><<<nb, BLOCK, 0, s>>>(rows, cols, nnz, (const float*)lam, (const float*)x, k,
                                                      (float)scale, (float*)g);
        else
            throw Error{AMG_E_ARG, "dtype must be AMG_F64 or AMG_F32"};
        HIPCHK(hipGetLastError());
    });
}

int amg_scatter_values(const int32_t* dest, long long nnz, int dtype, const void* src, void* dst, void* stream) {
    return guard([&] {
        ARGCHK(nnz >= 0, "bad size");
        if (nnz == 0) return;
        ARGCHK(dest && src && dst, "null argument");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const unsigned nb = (unsigned)((nnz + BLOCK - 1) / BLOCK);
        if (dtype == AMG_F64)
            k_scatter_values<double><<<nb, BLOCK, 0, s>>>(dest, nnz, (const double*)src, (double*)dst);
        else if (dtype == AMG_F32)
            k_scatter_values<float><<<nb, BLOCK, 0, s>>>(dest, nnz, (const float*)src, (float*)dst);
        else
            throw Error{AMG_E_ARG, "dtype must be AMG_F64 or AMG_F32"};
        HIPCHK(hipGetLastError());
    });
}

int amg_time_kernel(amg_handle h, int kernel, int level, int k, int reps, float* ms_out, void* stream) {
    return guard([&] {
        ARGCHK(ms_out && k > 0, "bad arguments");
        *ms_out = H(h)->time_kernel(kernel, level, k, reps, static_cast<hipStream_t>(stream));
    });
}

int amg_info(amg_handle h, int* nlev, int* dtype) {
    return guard([&] {
        HandleBase* hb = H(h);
        if (nlev) *nlev = hb->nlev;
        if (dtype) *dtype = hb->dtype;
    });
}

int amg_batch_create(const amg_handle* handles, const int* k, int nprob, amg_batch* out) {
    return guard([&] {
        ARGCHK(handles != nullptr && k != nullptr && out != nullptr, "null argument");
        ARGCHK(nprob >= 1, "need at least one problem");
        int dt = -1;
        for (int p = 0; p < nprob; ++p) {
            HandleBase* h = H(handles[p]);
            ARGCHK(dt < 0 || h->dtype == dt, "every handle of a batch must have the same dtype");
            dt = h->dtype;
        }
        BatchBase* b = dt == AMG_F64 ? create_batch<double>(handles, k, nprob) : create_batch<float>(handles, k, nprob);
        *out = reinterpret_cast<amg_batch>(b);
    });
}

int amg_batch_pcg(amg_batch b, const double* const* B, double* const* X, double tol, double atol, int maxiter,
                  int flexible, int* iters_out, double* relres_out, void* stream) {
    return guard([&] {
        ARGCHK(b != nullptr && B != nullptr && X != nullptr, "null argument");
        BatchBase* bb = reinterpret_cast<BatchBase*>(b);
        for (int p = 0; p < bb->nprob; ++p) ARGCHK(B[p] != nullptr && X[p] != nullptr, "null right-hand side or solution");
        for (int p = 0; p < bb->nprob; ++p) ARGCHK(aligned16(B[p]) && aligned16(X[p]), "b and x must be 16-byte aligned");
        bb->pcg(B, X, tol, atol, maxiter, flexible ? 1 : 0, iters_out, relres_out, static_cast<hipStream_t>(stream));
    });
}

int amg_batch_set_option(amg_batch b, int option, long long value) {
    return guard([&] {
        ARGCHK(b != nullptr, "null batch");
        reinterpret_cast<BatchBase*>(b)->set_option(option, value);
    });
}

int amg_batch_info(amg_batch b, int* nlaunch_iter, int* blocks_iter) {
    return guard([&] {
        ARGCHK(b != nullptr, "null batch");
        reinterpret_cast<BatchBase*>(b)->info(nlaunch_iter, blocks_iter);
    });
}

int amg_batch_destroy(amg_batch b) {
    return guard([&] {
        ARGCHK(b != nullptr, "null batch");
        delete reinterpret_cast<BatchBase*>(b);
    });
}

int amg_tune_report(char* buf, int len) {
    std::string out;
    {
        Tuner& t = tuner();
        std::lock_guard<std::mutex> g(t.mu);
        for (const auto& kv : t.cache) {
            char line[200];
            std::snprintf(line, sizeof(line), "%s %s %d %d %d %d -> %d %d %d %d %.2f %.2f\n",
                          std::get<0>(kv.first) == TK_SPMV ? "spmv" : "jacobi",
                          std::get<1>(kv.first) == AMG_F64 ? "f64" : "f32", std::get<2>(kv.first),
                          std::get<3>(kv.first), std::get<4>(kv.first), std::get<5>(kv.first), kv.second.plan.kp,
                          kv.second.plan.wide, kv.second.plan.subl, kv.second.plan.grid, kv.second.us_default,
                          kv.second.us_best);
            out += line;
        }
    }
    if (buf && len > 0) {
        std::strncpy(buf, out.c_str(), (size_t)len - 1);
        buf[len - 1] = '\0';
    }
    return (int)out.size();
}

int amg_last_error(char* buf, int len) {
    if (buf && len > 0) {
        std::strncpy(buf, g_err.c_str(), (size_t)len - 1);
        buf[len - 1] = '\0';
    }
    return (int)g_err.size();
}

}  // extern "C"
