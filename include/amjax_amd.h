/*
 * amjax_amd.h - C-ABI of the gfx950 multigrid solve-phase library.
 *
 * Plain C: device pointers, sizes and status codes; no C++ or torch types.
 * All work is enqueued on the caller's HIP stream (passed as void*; NULL is
 * the default stream).  A handle is single-stream and not thread-safe.
 * Every call returns 0 on success and a negative AMG_E* code otherwise;
 * amg_last_error() gives the message of the last failure on this thread.
 *
 * Data layout (shared by all entry points):
 *   - sparse matrices are sliced ELLPACK with 64-row slices (one slice per
 *     wavefront): slice s holds rows 64s..64s+63, entry j of row 64s+r sits at
 *     sptr[s] + 64*j + r; unused slots have col = -1 and val = 0.  The host
 *     layer packs a whole hierarchy into one int32 and one value allocation;
 *     descriptors hold absolute device pointers into them.
 *   - a block of k right-hand sides is stored row-major (n, k): entry (i, c)
 *     at i*k + c ("k-contiguous").  The solver entry points (amg_cycle,
 *     amg_solve, amg_pcg, amg_bicgstab, amg_batch_pcg) take b and x at
 *     16-byte-aligned addresses (AMG_E_ARG otherwise): when k is a multiple
 *     of 16 / sizeof(element) the row kernels read them as 16-byte vectors.
 */
#ifndef AMJAX_AMD_H
#define AMJAX_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    AMG_OK = 0,
    AMG_E_ARG = -1,       /* invalid argument */
    AMG_E_HIP = -2,       /* HIP runtime error */
    AMG_E_STATE = -3,     /* handle in an unusable state */
    AMG_E_UNSUPPORTED = -4
};

enum { AMG_F64 = 0, AMG_F32 = 1 };
enum { AMG_SMOOTH_NONE = 0, AMG_SMOOTH_JACOBI = 1, AMG_SMOOTH_POLY = 2 };
enum { AMG_POLY_MAX = 8 }; /* coefficients of a polynomial smoother */
enum { AMG_COARSE_PINV = 0, AMG_COARSE_LU = 1, AMG_COARSE_QR = 2, AMG_COARSE_JACOBI = 3 };

typedef struct amg_sell_desc {
    int32_t nrows, ncols;
    const int32_t *sptr; /* nslices + 1 slice offsets (elements) */
    const int32_t *col;  /* column indices, padded length */
    const void *val;     /* values (float64 or float32), padded length */
    int64_t nnz;         /* stored nonzeros without padding */
    int64_t padded;      /* stored slots including padding */
    int32_t max_width;   /* widest slice (slots per row); a handle's row kernels are specialised on it,
                            so amg_create / amg_set_level0 refuse one below the widest slice the
                            offsets give (AMG_E_ARG) */
    int32_t reserved;    /* 0 */
} amg_sell_desc;

typedef struct amg_level_desc {
    amg_sell_desc A, P, R; /* P (n x n_next), R (n_next x n); unused on the coarsest level */
    amg_sell_desc D;       /* optional fused down-sweep operator R (I - pre_omega A D^-1)
                              (n_next x n): restriction of the residual after one
                              zero-guess Jacobi pre-sweep; nrows = 0 if absent */
    const void *dinv;      /* D^-1 (n entries) */
    int32_t pre_kind, pre_iters;
    double pre_omega;
    int32_t post_kind, post_iters;
    double post_omega;
    /* AMG_SMOOTH_POLY (PyAMG's polynomial smoother, used by its 'chebyshev' and
       'richardson'): per sweep r = b - A x, h = c_0 r, h = c_i r + A h, x += h.
       1..AMG_POLY_MAX float64 coefficients in host memory, copied at creation. */
    const double *pre_coef;
    int32_t pre_ncoef, reserved0;
    const double *post_coef;
    int32_t post_ncoef, reserved1;
} amg_level_desc;

typedef struct amg_coarse_desc {
    int32_t kind; /* AMG_COARSE_* */
    int32_t n;
    const void *f0; /* pinv: A_inv (n x n, row-major); lu: packed LU; qr: Q */
    const void *f1; /* lu: int32 pivots (LAPACK convention, 0-based); qr: R (upper) */
    int32_t iters;  /* jacobi */
    double omega;   /* jacobi */
    int32_t factor_f64; /* 1: the factors are float64 whatever the cycle dtype; the
                           coarse solve then accumulates in float64 and rounds only x */
    int32_t reserved;
} amg_coarse_desc;

typedef struct amg_handle_s *amg_handle;

/*
 * Create a solve-phase handle for a hierarchy already resident on the device.
 * Replaces MultilevelSolver.from_pyamg's device upload
 * (amjax/multilevel.py:125-186, 557-590): the caller keeps the arrays alive.
 *   dtype  AMG_F64 or AMG_F32: arithmetic type of the cycle.
 *   outer  level-0 operator with float64 values used by the outer Krylov /
 *          refinement residual (may be NULL for AMG_F64: level 0's A is used).
 */
int amg_create(const amg_level_desc *levels, int nlev, const amg_coarse_desc *coarse, int dtype,
               const amg_sell_desc *outer, amg_handle *out);

/*
 * Replace the level-0 operator used by the cycle's smoothing and residuals and
 * by the outer iteration (the `A` argument of MultilevelSolver.solve,
 * amjax/multilevel.py:318,387,391): the rest of the hierarchy, including
 * level 0's D^-1, stays fixed.  A == NULL restores the original operator.
 * Same dtype as the handle (float64 handles only for the outer operator).
 */
int amg_set_level0(amg_handle h, const amg_sell_desc *A);
int amg_destroy(amg_handle h);

/*
 * One multigrid cycle x <- cycle(x, b).  Replaces MultilevelSolver._cycle
 * (amjax/multilevel.py:290-339) and, with x_zero = 1, the matvec returned by
 * aspreconditioner (amjax/multilevel.py:265-288).
 *   cycle 'V', 'W' or 'F'; cpl = cycles_per_level (F only).  Also 'T'
 *          (no counterpart in the reference; for the adjoint of fixed
 *          F-cycles): the F-cycle's visits in reverse order, each level's
 *          coarse correction running cpl V-cycles (the first from zero) and
 *          then 'T' from their result.  On the transposed hierarchy (A_l^T,
 *          R_l^T as P, P_l^T as R, pre- and post-smoothers exchanged) it
 *          applies B^T where the F-cycle applies B.
 *   b, x   (n0, k) in the handle's dtype.  x_zero = 1: x is treated as zero on
 *          entry (its contents are ignored).
 */
int amg_cycle(amg_handle h, char cycle, int cpl, const void *b, void *x, int k, int x_zero,
              void *stream);

/*
 * amg_cycle without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and the
 * same result, bit for bit, as amg_cycle; the cycle's level vectors live in a
 * device array of AMG_NOALLOC_SCRATCH_MB MiB that each GPU holds from the
 * library's load to the end of the process (its address is resolved once per
 * device, at amg_create or at the first call on a device).  It needs
 * (4 sum_l n_l + 2 n_0) k elements of the handle's dtype, plus 2 n_l k more on
 * levels with a polynomial smoother and n_c k doubles, each vector rounded up
 * to 256 bytes; a larger request fails with AMG_E_ARG.  It calls no hipMalloc,
 * hipFree, memset or memcpy and no autotuning (row kernels take the tuner's
 * cached plans or the fixed shape).  Every call on a GPU uses that GPU's one
 * array: calls that overlap on one GPU (other streams, other threads, other
 * handles) race and must be serialised by the caller.
 */
enum { AMG_NOALLOC_SCRATCH_MB = 64 };
int amg_cycle_noalloc(amg_handle h, char cycle, int cpl, const void *b, void *x, int k, int x_zero,
                      void *stream);

/*
 * Standalone solve (amjax/multilevel.py:341-425, _solve_loop): x = 0, then
 * repeat x <- cycle(x, b) while it < maxiter and ||b - A x|| >= tol ||b||,
 * independently per column (jax.vmap semantics: a converged column is frozen).
 * b, x are (n0, k) in the handle's dtype, except for a float32 handle with
 * refine = 1, where they are float64 and the loop is iterative refinement
 * x <- x + cycle_f32(0, b - A x) with the residual in float64.
 * iters_out (k ints) and relres_out (k doubles, ||b - A x|| / ||b||) are host
 * pointers and may be NULL.
 */
int amg_solve(amg_handle h, char cycle, int cpl, const void *b, void *x, int k, double tol,
              int maxiter, int refine, int *iters_out, double *relres_out, void *stream);

/*
 * Preconditioned CG on the float64 outer operator with one cycle as M
 * (the jax.scipy.sparse.linalg.cg call of benchmarks/benchmark.ipynb cell 7 with
 * M = ml.aspreconditioner()).  x0 = 0; stops per column when r.r <=
 * max(tol^2 b.b, atol^2) or after maxiter iterations.  b, x are float64
 * (n0, k).  flexible = 1 uses the Polak-Ribiere beta (for an inexact,
 * e.g. float32, preconditioner); 0 is the textbook Hestenes-Stiefel beta.
 * precond = 0 runs plain CG (M = identity).
 */
int amg_pcg(amg_handle h, char cycle, int cpl, const double *b, double *x, int k, double tol,
            double atol, int maxiter, int flexible, int precond, int *iters_out,
            double *relres_out, void *stream);

/*
 * Right-preconditioned BiCGSTAB in float64 with jax.scipy.sparse.linalg.bicgstab
 * semantics (the reference's README roadmap pairs AMJax preconditioning with JAX
 * GMRES/BiCGSTAB, README.md:140-149): x0 = 0, rhat = r0, per column stop when
 * r.r <= max(tol^2 b.b, atol^2), at maxiter, or on breakdown (alpha, omega or
 * rhat.r zero); early exit of an iteration when s.s < atol^2.  precond = 1 uses
 * one cycle of the handle (its dtype: float32 handles give the mixed form) as M.
 * Same outputs as amg_pcg.
 */
int amg_bicgstab(amg_handle h, char cycle, int cpl, const double *b, double *x, int k, double tol,
                 double atol, int maxiter, int precond, int *iters_out, double *relres_out, void *stream);

/*
 * Mixed precision by iterative refinement around an inner CG (no reference
 * counterpart: an FP32 cycle inside FP64 iterative refinement, with a CG in
 * the cycle's precision between them): x = 0, then per
 * column, while ||b - A64 x|| >= tol ||b|| and fewer than maxiter refinement
 * steps: e = CG(A, (T) r) in the handle's dtype T (x, r, p, Ap in T, dot
 * products accumulated in float64, one cycle as preconditioner, stopped at
 * ||r_e|| <= inner_tol ||(T) r|| or inner_maxiter iterations), x += e, and
 * r = b - A64 x in float64 with the float64 outer operator.  b, x are float64
 * (n0, k).  iters_out (k ints): inner CG iterations summed over the steps;
 * outer_out (k ints): refinement steps; relres_out: ||b - A64 x|| / ||b||.
 * Host pointers, each may be NULL.
 */
int amg_pcg_ir(amg_handle h, char cycle, int cpl, const double *b, double *x, int k, double tol,
               double inner_tol, int maxiter, int inner_maxiter, int flexible, int *iters_out,
               int *outer_out, double *relres_out, void *stream);

/*
 * Mixed precision by preconditioned CG with reliable updates (no reference
 * counterpart: the north star's FP32 cycle inside an FP64 outer iteration,
 * here without restarts; Strzodka & Goeddeke 2006, Clark et al. 2010).  The CG
 * vectors (the correction xh, the recurrence residual rh, p, Ap) are in the
 * handle's dtype with dot products accumulated in float64, one cycle is M,
 * and x = 0.  When a column's rh.rh falls to delta^2 times its value at the
 * last update, or to the stopping threshold, or the column reaches maxiter,
 * a reliable update replaces rh by the float64 residual b - A64 x of the
 * folded iterate, keeping p: the residual s = b - A64 x is kept in float64
 * (s = b at the start) and each update sets s -= A64 xh (A64 the float64
 * outer operator; equal to b - A64 (x + xh) in exact arithmetic), then folds
 * x += xh in float64 and sets xh = 0.  A column stops when this residual
 * satisfies ||s||^2 <= max(tol^2 b.b, atol^2), or after maxiter CG
 * iterations.  b, x are float64 (n0, k).  iters_out (k ints): CG iterations;
 * updates_out (k ints): reliable updates; relres_out: ||s|| / ||b|| at the
 * last update.  Host pointers, each may be NULL.
 * delta in (0, 1).
 */
int amg_pcg_ru(amg_handle h, char cycle, int cpl, const double *b, double *x, int k, double tol,
               double atol, double delta, int maxiter, int flexible, int *iters_out, int *updates_out,
               double *relres_out, void *stream);

/*
 * Restarted GMRES in float64 with jax.scipy.sparse.linalg.gmres semantics for its
 * default solve_method='batched' (the README roadmap's nonsymmetric Krylov caller,
 * README.md:140-149): left preconditioning by one cycle (precond = 1; its dtype:
 * float32 handles give the mixed form), x0 = 0; per restart `restart` (1..32)
 * Arnoldi steps on M A with classical Gram-Schmidt applied twice, the Hessenberg
 * least-squares problem solved once, x += V y; per column stop when
 * ||M (b - A x)|| <= ||M b|| min(1, max(tol ||b||, atol) / ||b||) or after maxiter
 * restarts.  method 0 is jax's solve_method='batched' (all `restart` steps, one
 * least-squares solve); method 1 is 'incremental': after each Arnoldi step the
 * column's Givens factorisation is updated and a column stops extending its
 * basis once the least-squares residual is <= that threshold (x then takes the
 * steps built so far).  iters_out: restarts per column; relres_out:
 * ||b - A x|| / ||b||.
 */
int amg_gmres(amg_handle h, char cycle, int cpl, const double *b, double *x, int k, double tol,
              double atol, int restart, int maxiter, int precond, int method, int *iters_out,
              double *relres_out, void *stream);

/*
 * y = (add ? y : 0) + M x for one sliced-ELL matrix (the BCOO products
 * A @ x, R @ r, P @ e of amjax/multilevel.py:321,335,387 and
 * amjax/relaxation/relaxation.py:95).  x is (ncols, k), y is (nrows, k),
 * both in `dtype`; x and y may not alias.
 */
int amg_sell_spmv(const amg_sell_desc *M, int dtype, const void *x, void *y, int k, int add, void *stream);

/*
 * One damped Jacobi sweep (amjax/relaxation/relaxation.py:64-98, body 94-96):
 *   temp = Dinv*(b - A x) + x;  x_out = (1 - omega) x + omega temp.
 * x and x_out may not alias.
 */
int amg_sell_jacobi(const amg_sell_desc *A, const void *dinv, int dtype, const void *x, const void *b,
                    void *x_out, int k, double omega, void *stream);

/*
 * Polynomial smoother (PyAMG relaxation.polynomial, the sweep of its 'chebyshev'
 * and 'richardson' smoothers, listed as not implemented by the reference,
 * amjax/relaxation/smoothing.py:120-127): `iters` times r = b - A x (b when
 * x_zero), h = coef[0] r, h = coef[i] r + A h (i = 1..ncoef-1), x += h; one fused
 * kernel per coefficient.  work: 4 * n * k elements of dtype (no allocation
 * inside); x and x_out may not alias.
 */
int amg_sell_poly(const amg_sell_desc *A, int dtype, const double *coef, int ncoef, int iters,
                  const void *x, const void *b, void *x_out, void *work, int k, int x_zero, void *stream);

/*
 * Dense coarse-grid solve x = solve(b) with a precomputed factor
 * (amjax/multilevel.py:509-536: pinv -> A_inv @ b, lu -> lu_solve,
 * qr -> solve_triangular(R, Q^T b)).  kind pinv, lu or qr; b and x (n, k)
 * may not alias.
 */
int amg_dense_solve(const amg_coarse_desc *c, int dtype, const void *b, void *x, int k, void *stream);

/*
 * Gradient of a solve with respect to the stored values of A on its fixed
 * pattern: g[q] = scale * sum_c lam[rows[q], c] * x[cols[q], c] for the nnz
 * pattern entries (scale = -1 gives the pattern restriction of the reference's
 * -outer(lam, x), amjax/multilevel.py:441).  lam, x are (n, k) row-major.
 */
int amg_grad_values(const int32_t *rows, const int32_t *cols, long long nnz, int dtype, const void *lam,
                    const void *x, int k, double scale, void *g, void *stream);

/*
 * New values on a fixed pattern: dst[dest[q]] = src[q], where dest maps the
 * q-th stored entry in CSR order to its sliced-ELL slot (the device side of
 * replacing A's values, as `A=` in MultilevelSolver.solve does).
 */
int amg_scatter_values(const int32_t *dest, long long nnz, int dtype, const void *src, void *dst, void *stream);

/*
 * Launch the fused kernel `kernel` of level `level` `reps` times between two
 * HIP events on `stream` and return the average duration in ms (roofline
 * measurement).  kernel: 0 = SpMV of A, 1 = Jacobi sweep, 2 = residual,
 * 3 = residual + restriction, 4 = prolongation + correction, 5 = float64 outer
 * CG update (x += alpha p, r -= alpha Ap, r.r partials; level 0), 6 = fused
 * down-sweep product with D, 7 = prolongation rebuilding the pre-swept iterate,
 * 8 = the float64 CG update inside an eager preconditioned CG solve (b = 1,
 * tol = 0, V-cycle): mean over the update launches of its first `reps`
 * iterations, each bracketed by events (the kernel in its solve context);
 * 9 and 10 = likewise inside an eager reliable-update CG solve (amg_pcg_ru):
 * 9 its update xh += alpha p, rh -= alpha Ap, 10 its last level-0 sweep with
 * the rh.z sums; 11 = that update alone (level 0, every column active with
 * alpha = 0).  kernel | AMG_TIME_COLD: cache-cold, 512 MB (twice the Infinity
 * Cache) are read before each timed launch (outside its events, leaving only
 * clean lines in the caches) and the launches are timed one by one, so their
 * inputs come from HBM.
 */
enum { AMG_TIME_COLD = 256 };
int amg_time_kernel(amg_handle h, int kernel, int level, int k, int reps, float *ms_out,
                    void *stream);

/*
 * Handle options (no reference counterpart; execution strategy only):
 *   AMG_OPT_FUSE  1 (default): on a zero-guess level with one Jacobi
 *                 pre-sweep, replace pre-sweep + residual + restriction by one
 *                 product with the level's D operator and rebuild the pre-swept
 *                 iterate inside the prolongation; in iterative refinement
 *                 (amg_solve refine = 1) the cycle's last top-level sweep adds
 *                 its result into the float64 iterate (vector mode; the same
 *                 sums as a stored correction); 0: one kernel per reference
 *                 operation, in the reference's operation order.
 *   AMG_OPT_AUTOTUNE  0 (default): row kernels use the fixed launch shape;
 *                 1: they use the autotuner's plan (column lanes per row group,
 *                 wide-row variant, rows per wavefront and persistent grid size,
 *                 timed once per (kernel kind, dtype, k rounded up to a power of
 *                 two, stored entries per row, narrow/wide rows, size class) and
 *                 cached for the process; see amg_tune_report).
 *                 Plans change the work split only: sums are the same except
 *                 for k = 1 rows split over several lanes.
 *   AMG_OPT_VECTOR  1 (default): row kernels over matrices with at most 64
 *                 stored entries per row run in vector mode when k is a
 *                 multiple of 16 / sizeof(element) and the column chunk is
 *                 16, 32 or 64 (each lane holds 16 bytes of a row; the same
 *                 sums, bit for bit, except the per-column dot products of the
 *                 reducing kernels, which are grouped differently);
 *                 0: one column per lane.
 *   AMG_OPT_TAIL  AMG_TAIL_ROWS_DEFAULT (default): a zero-guess V-cycle visit
 *                 of the first level l >= 1 with at most this many rows runs,
 *                 with every visit below it, as one launch: one workgroup per
 *                 column, which stages the levels' operators (an image packed
 *                 at set-up, about 100-150 KiB at 500 x 500) and its level
 *                 vectors in LDS; l is the first such level whose image and
 *                 vectors fit the device's LDS per workgroup (160 KiB on
 *                 MI355X).  When the restrictions fit in the image too, W, F
 *                 and 'T' visits and nonzero-guess visits of that level run the
 *                 same way (else from the first deeper level where they fit).
 *                 It applies when every level from l down takes the fused
 *                 down-sweep (AMG_OPT_FUSE) and one Jacobi post-sweep and the
 *                 coarse solver is the pseudo-inverse; the same sums, bit for
 *                 bit; 0: one launch per level operation.
 */
enum { AMG_OPT_FUSE = 1, AMG_OPT_AUTOTUNE = 2, AMG_OPT_VECTOR = 3, AMG_OPT_TAIL = 4 };
enum { AMG_TAIL_ROWS_DEFAULT = 8192 };
int amg_set_option(amg_handle h, int option, long long value);
/* First level of the one-launch coarse tail (AMG_OPT_TAIL) of a zero-guess
 * V-cycle, or -1 when every level runs its own launches. */
int amg_tail_level(amg_handle h, int *level);

/*
 * Batch of independent AMG-preconditioned CG solves: problem p runs
 * amg_pcg(handles[p], cycle, cpl, B[p], X[p], k[p], ...) (cycle 'V' and cpl 1
 * unless set by AMG_BATCH_OPT_CYCLE / AMG_BATCH_OPT_CPL) with flexible as given and
 * precond = 1; all problems advance together, one kernel launch per phase of
 * the PCG iteration across all problems.  Replaces the per-problem loop
 * of the reference's benchmark, which calls jax.scipy.sparse.linalg.cg(A, b,
 * M=ml.aspreconditioner()) once per problem (benchmarks/benchmark.ipynb, JSON
 * lines 148-177 and 255-319); each problem's iterates equal those of amg_pcg
 * bit for bit.  Handles must share one dtype, carry the float64 outer operator
 * (float64 handles, or float32 handles created with `outer`); any smoother and
 * coarse solver of a handle runs in a batch; the batch records their levels,
 * smoothers and AMG_OPT_FUSE setting at creation and must not outlive them.
 * Results: iters_out / relres_out (nullable) hold sum_p k[p] entries, problem
 * by problem.  maxiter < 0 gives every problem its own limit of 10 n_p
 * iterations (n_p its number of rows, the default of a single solve).
 * Single-stream, not thread-safe, like a handle.
 */
typedef struct amg_batch_s *amg_batch;
int amg_batch_create(const amg_handle *handles, const int *k, int nprob, amg_batch *out);
int amg_batch_pcg(amg_batch b, const double *const *B, double *const *X, double tol, double atol,
                  int maxiter, int flexible, int *iters_out, double *relres_out, void *stream);
/*
 * Batch options (execution strategy only):
 *   AMG_BATCH_OPT_GRID  persistent grid: a launch uses at most this many
 *                       blocks (a multiple of 8), each walking the virtual
 *                       blocks with a stride of the grid; 0 (default): one
 *                       block per virtual block.
 *   AMG_BATCH_OPT_SPLIT_KP  1 (default): a phase whose problems use different
 *                       column-lane counts (k rounded up to a power of two,
 *                       at most 64) runs as one launch per count, each with
 *                       the register budget of the plain kernel; 0: one
 *                       launch per phase with the lane count chosen per step.
 *   AMG_BATCH_OPT_AUTOTUNE  1 (default): the problems' row kernels use the
 *                       autotuner's plans (as AMG_OPT_AUTOTUNE); 0: the
 *                       fixed launch shape.
 *   AMG_BATCH_OPT_CYCLE  'V' (default), 'W' or 'F' (the character code):
 *                       the preconditioner's cycle.  W and F cycles revisit
 *                       levels, so their steps merge by position in each
 *                       problem's sequence rather than by level.
 *   AMG_BATCH_OPT_CPL   V-cycles per level of an F-cycle (default 1, >= 0).
 */
enum {
    AMG_BATCH_OPT_GRID = 1,
    AMG_BATCH_OPT_SPLIT_KP = 2,
    AMG_BATCH_OPT_AUTOTUNE = 3,
    AMG_BATCH_OPT_CYCLE = 4,
    AMG_BATCH_OPT_CPL = 5
};
int amg_batch_set_option(amg_batch b, int option, long long value);
/* Kernel launches and virtual blocks of one batched PCG iteration. */
int amg_batch_info(amg_batch b, int *nlaunch_iter, int *blocks_iter);
int amg_batch_destroy(amg_batch b);

/*
 * The autotuner's cache as text, one line per key:
 *   "kind dtype kp(k) entries/row narrow size_class -> kp wide log2(rows/unit) grid
 *    us_default us_best" (kind spmv or jacobi; grid 0 = one block per 4 units).
 * Returns the full length (the text is truncated to len - 1 characters).
 */
int amg_tune_report(char *buf, int len);

/* Number of levels and dtype (AMG_F64 / AMG_F32) of a handle. */
int amg_info(amg_handle h, int *nlev, int *dtype);

/* Message of the last failure on the calling thread (truncated to len). */
int amg_last_error(char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif
