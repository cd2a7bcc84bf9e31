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
 * points by largest measure lambda_i = |{j : j strongly depends on i}|
 * (number of entries in column i of the strength matrix S), followed by the
 * second pass that makes every strong F-F connection have a common C point.
 * S is given in CSR (row i lists the points i strongly depends on) and its
 * transpose ST likewise.  splitting[i] = 1 for C, 0 for F.
 */
void amg_setup_rs_cf_splitting(int n, const int *Sp, const int *Sj, const int *Tp, const int *Tj,
                               int second_pass, int *splitting)
{
    int *lambda = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *state = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1)); /* -1 undecided, 0 F, 1 C */
    int maxl = 0;
    for (int i = 0; i < n; i++) {
        int l = 0;
        for (int jj = Tp[i]; jj < Tp[i + 1]; jj++) if (Tj[jj] != i) l++;
        lambda[i] = l;
        if (l > maxl) maxl = l;
        state[i] = -1;
    }
    /* bucket lists by lambda (values grow by at most n) */
    int cap = maxl + n + 2;
    int *head = (int *)malloc(sizeof(int) * (size_t)cap);
    int *next = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *prev = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int v = 0; v < cap; v++) head[v] = -1;
    /* insert in reverse so that equal-lambda ties pop the lowest index first */
    for (int i = n - 1; i >= 0; i--) {
        int l = lambda[i];
        next[i] = head[l]; prev[i] = -1;
        if (head[l] >= 0) prev[head[l]] = i;
        head[l] = i;
    }
    int top = maxl;
#define RS_REMOVE(node) do { int _v = (node), _l = lambda[_v]; \
        if (prev[_v] >= 0) next[prev[_v]] = next[_v]; else head[_l] = next[_v]; \
        if (next[_v] >= 0) prev[next[_v]] = prev[_v]; } while (0)
#define RS_INSERT(node) do { int _v = (node), _l = lambda[_v]; \
        next[_v] = head[_l]; prev[_v] = -1; if (head[_l] >= 0) prev[head[_l]] = _v; head[_l] = _v; \
        if (_l > top) top = _l; } while (0)
    for (;;) {
        while (top >= 0 && head[top] < 0) top--;
        if (top < 0) break;
        int i = head[top];
        RS_REMOVE(i);
        if (lambda[i] == 0) {
            /* no point depends on i: it becomes F unless it is isolated, then C */
            int has = 0;
            for (int jj = Sp[i]; jj < Sp[i + 1]; jj++) if (Sj[jj] != i) { has = 1; break; }
            state[i] = has ? 0 : 1;
            continue;
        }
        state[i] = 1;
        /* points that strongly depend on i become F */
        for (int jj = Tp[i]; jj < Tp[i + 1]; jj++) {
            int j = Tj[jj];
            if (j == i || state[j] != -1) continue;
            RS_REMOVE(j);
            state[j] = 0;
            /* the points j depends on gain measure */
            for (int kk = Sp[j]; kk < Sp[j + 1]; kk++) {
                int k = Sj[kk];
                if (k == j || state[k] != -1) continue;
                RS_REMOVE(k);
                lambda[k]++;
                if (lambda[k] >= cap) lambda[k] = cap - 1;
                RS_INSERT(k);
            }
        }
        /* the points i depends on lose measure */
        for (int jj = Sp[i]; jj < Sp[i + 1]; jj++) {
            int j = Sj[jj];
            if (j == i || state[j] != -1) continue;
            RS_REMOVE(j);
            if (lambda[j] > 0) lambda[j]--;
            RS_INSERT(j);
        }
    }
#undef RS_REMOVE
#undef RS_INSERT
    for (int i = 0; i < n; i++) splitting[i] = state[i] == 1 ? 1 : 0;

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
    free(lambda); free(state); free(head); free(next); free(prev);
}
