/*
 * Host-side loops of the hierarchy setup (the step before the solve phase).
 *
 * The reference delegates setup to PyAMG (amjax/multilevel.py:557-590 reads
 * pyamg_solver.levels[i].A/.P/.R).  PyAMG is not available here, so the
 * in-repo smoothed-aggregation and classical builders call these loops,
 * which are too slow as Python loops at 10^6 unknowns.  Plain C, built with
 * gcc, no GPU code.
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <math.h>

/*
 * Standard (greedy, three-pass) aggregation on the pattern of a strength
 * graph given in CSR form.  Diagonal entries are ignored.
 *   pass 1: a node whose neighbours are all free seeds a new aggregate made
 *           of itself and its neighbours;
 *   pass 2: a remaining node joins the aggregate of its first neighbour that
 *           was aggregated in pass 1;
 *   pass 3: each node still free seeds an aggregate with its free neighbours.
 * Nodes without neighbours stay unaggregated (agg = -1).
 * Returns the number of aggregates; root[a] is the seed node of aggregate a.
 */
int amg_setup_standard_aggregation(int n, const int *indptr, const int *indices,
                                   int *agg, int *root)
{
    int nagg = 0;
    for (int i = 0; i < n; i++) agg[i] = -2; /* -2: free */
    /* pass 1 */
    for (int i = 0; i < n; i++) {
        if (agg[i] != -2) continue;
        int has_nb = 0, nb_taken = 0;
        for (int jj = indptr[i]; jj < indptr[i + 1]; jj++) {
            int j = indices[jj];
            if (j == i) continue;
            has_nb = 1;
            if (agg[j] != -2) { nb_taken = 1; break; }
        }
        if (!has_nb) { agg[i] = -1; continue; }
        if (nb_taken) continue;
        agg[i] = nagg;
        for (int jj = indptr[i]; jj < indptr[i + 1]; jj++) agg[indices[jj]] = nagg;
        root[nagg] = i;
        nagg++;
    }
    int npass1 = nagg;
    /* pass 2: attach to a pass-1 aggregate; mark with -(a+3) so a node attached
       in this pass does not act as an anchor for later nodes of the pass */
    for (int i = 0; i < n; i++) {
        if (agg[i] != -2) continue;
        for (int jj = indptr[i]; jj < indptr[i + 1]; jj++) {
            int j = indices[jj];
            int a = agg[j];
            if (j != i && a >= 0 && a < npass1) { agg[i] = -(a + 3); break; }
        }
    }
    for (int i = 0; i < n; i++)
        if (agg[i] <= -3) agg[i] = -agg[i] - 3;
    /* pass 3 */
    for (int i = 0; i < n; i++) {
        if (agg[i] != -2) continue;
        agg[i] = nagg;
        for (int jj = indptr[i]; jj < indptr[i + 1]; jj++) {
            int j = indices[jj];
            if (agg[j] == -2) agg[j] = nagg;
        }
        root[nagg] = i;
        nagg++;
    }
    return nagg;
}

/*
 * One Gauss-Seidel sweep over rows start, start+step, ... (stop exclusive)
 * on A x = b for ncol right-hand sides stored row-major (n, ncol).
 * Rows with a zero diagonal are left unchanged.
 */
void amg_setup_gauss_seidel(int n, const int *indptr, const int *indices, const double *data,
                            double *x, const double *b, int ncol,
                            int start, int stop, int step)
{
    (void)n;
    for (int i = start; i != stop; i += step) {
        for (int c = 0; c < ncol; c++) {
            double diag = 0.0, rsum = 0.0;
            for (int jj = indptr[i]; jj < indptr[i + 1]; jj++) {
                int j = indices[jj];
                if (j == i) diag += data[jj];
                else rsum += data[jj] * x[(int64_t)j * ncol + c];
            }
            if (diag != 0.0)
                x[(int64_t)i * ncol + c] = (b[(int64_t)i * ncol + c] - rsum) / diag;
        }
    }
}

/*
 * Classical (Ruge-Stueben) C/F splitting, first pass: greedy selection of C
 * points by largest measure lambda_i = |{j != i : j strongly depends on i}|
 * (off-diagonal entries in column i of the strength matrix S), optionally
 * followed by a second pass that makes every strong F-F connection have a
 * common C point.  S is given in CSR (row i lists the points i strongly
 * depends on) and its transpose ST likewise.  splitting[i] = 1 for C, 0 for F.
 *
 * The order of selection follows PyAMG's rs_cf_splitting (pyamg >= 4, its
 * amg_core ruge_stuben routines; PyAMG is not vendored here), because on
 * structured grids the tie-breaking decides the coarse grid:
 *   - the nodes sit in one array sorted by measure, in contiguous intervals of
 *     equal measure, each interval in ascending node order at the start;
 *   - the walk takes the node at the highest position (largest measure, and
 *     among equals the highest position), once; an undecided node becomes C,
 *     a decided one is passed over; the walk ends at measure 0;
 *   - the undecided points that strongly depend on a new C point become F,
 *     and each undecided point such an F point depends on gains one unit of
 *     measure: it moves to the lowest position of the next interval;
 *   - each undecided point the new C point depends on loses one unit: it
 *     moves to the highest position of the interval below;
 *   - nodes of measure 0 and nodes left undecided become F.
 */
void amg_setup_rs_cf_splitting(int n, const int *Sp, const int *Sj, const int *Tp, const int *Tj,
                               int second_pass, int *splitting)
{
    enum { U = -1, F = 0, C = 1, PRE_F = 2 };
    const size_t nn = (size_t)(n > 0 ? n : 1);
    int *lam = (int *)malloc(sizeof(int) * nn);
    int *state = (int *)malloc(sizeof(int) * nn);
    int *pos = (int *)malloc(sizeof(int) * nn);     /* node -> position */
    int *at = (int *)malloc(sizeof(int) * nn);      /* position -> node */
    int maxl = 0;
    for (int i = 0; i < n; i++) {
        int l = 0;
        for (int jj = Tp[i]; jj < Tp[i + 1]; jj++) if (Tj[jj] != i) l++;
        lam[i] = l;
        if (l > maxl) maxl = l;
    }
    /* measures only grow while below n - 1 */
    const int nint = (2 * maxl > n + 1 ? 2 * maxl : n + 1) + 1;
    int *start = (int *)calloc((size_t)nint, sizeof(int));  /* first position of an interval */
    int *cnt = (int *)calloc((size_t)nint, sizeof(int));    /* nodes in an interval */
    for (int i = 0; i < n; i++) cnt[lam[i]]++;
    for (int l = 0, sum = 0; l < nint; l++) {
        start[l] = sum;
        sum += cnt[l];
        cnt[l] = 0;
    }
    for (int i = 0; i < n; i++) {
        const int p = start[lam[i]] + cnt[lam[i]]++;
        at[p] = i;
        pos[i] = p;
        state[i] = lam[i] == 0 ? F : U;
    }
#define RS_SWAP(p, q) do { int _a = at[p], _b = at[q]; at[p] = _b; at[q] = _a; pos[_b] = (p); pos[_a] = (q); } while (0)
    for (int top = n - 1; top >= 0; top--) {
        const int i = at[top];
        cnt[lam[i]]--;
        if (lam[i] <= 0) break;
        if (state[i] != U) continue;
        state[i] = C;
        for (int jj = Tp[i]; jj < Tp[i + 1]; jj++)
            if (state[Tj[jj]] == U) state[Tj[jj]] = PRE_F;
        for (int jj = Tp[i]; jj < Tp[i + 1]; jj++) {
            const int j = Tj[jj];
            if (state[j] != PRE_F) continue;
            state[j] = F;
            for (int kk = Sp[j]; kk < Sp[j + 1]; kk++) {
                const int k = Sj[kk];
                if (state[k] != U || lam[k] >= n - 1) continue;
                const int l = lam[k], np = start[l] + cnt[l] - 1;
                RS_SWAP(pos[k], np);
                cnt[l]--;
                cnt[l + 1]++;
                start[l + 1] = np;
                lam[k]++;
            }
        }
        for (int jj = Sp[i]; jj < Sp[i + 1]; jj++) {
            const int j = Sj[jj];
            if (state[j] != U || lam[j] == 0) continue;
            const int l = lam[j], np = start[l];
            RS_SWAP(pos[j], np);
            cnt[l]--;
            cnt[l - 1]++;
            start[l]++;
            start[l - 1] = start[l] - cnt[l - 1];
            lam[j]--;
        }
    }
#undef RS_SWAP
    for (int i = 0; i < n; i++) splitting[i] = state[i] == C ? 1 : 0;

    if (second_pass) {
        /* every pair of strongly connected F points needs a common C point */
        int *mark = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        for (int i = 0; i < n; i++) mark[i] = -1;
        for (int i = 0; i < n; i++) {
            if (splitting[i] != 0) continue;
            for (int jj = Sp[i]; jj < Sp[i + 1]; jj++) {
                int j = Sj[jj];
                if (j != i && splitting[j] == 1) mark[j] = i;
            }
            for (int jj = Sp[i]; jj < Sp[i + 1]; jj++) {
                int j = Sj[jj];
                if (j == i || splitting[j] != 0) continue;
                int common = 0;
                for (int kk = Sp[j]; kk < Sp[j + 1]; kk++) {
                    int k = Sj[kk];
                    if (k != j && splitting[k] == 1 && mark[k] == i) { common = 1; break; }
                }
                if (!common) {
                    splitting[j] = 1;
                    mark[j] = i;
                }
            }
        }
        free(mark);
    }
    free(lam); free(state); free(pos); free(at); free(start); free(cnt);
}
