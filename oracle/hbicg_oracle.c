/*
 * hbicg_oracle.c -- CPU restatement of the reference shifted BiCG ("lopbicg")
 * with seed switching.  TEST INFRASTRUCTURE ONLY (the checker and the timed
 * CPU baseline); product code never links it.
 *
 * Followed, op for op:
 *   src/solver_dynamic.c:101-123  initialisation
 *   src/solver_dynamic.c:156-210  seed BiCG step
 *   src/solver_dynamic.c:239-256  per-shift recurrences and vector updates
 *   src/solver_dynamic.c:271-303  convergence check
 *   src/solver_dynamic.c:309-359  seed switch (archive rescale + pi/zeta replay)
 *   src/openmp_matrix.c:38-84     row-sequential CSR SpMV
 *   src/openmp_vector.c:3-84      parallel BLAS-1, serial-master dot (ddot_v3)
 *   src/vector.c:3-27             serial BLAS-1
 *   src/matrix.c:268-419          row blocks of a P-rank run (orc_shifted_lopbicg_ranks)
 * Parallelism mirrors the reference's single-rank OpenMP structure (rows in
 * parallel for SpMV and BLAS-1, dot products summed serially by one thread,
 * shifts distributed with schedule(dynamic,1)); none of it changes a rounding,
 * so the result is independent of the thread count.
 * Build with -ffp-contract=off so that no multiply-add is fused.
 */
#include "hbicg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double wtime(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- serial BLAS-1 (src/vector.c:3-27) ---------------------------------- */
void orc_daxpy(int n, double a, const double *x, double *y)
{
    for (int i = 0; i < n; ++i) y[i] += a * x[i];
}

double orc_ddot(int n, const double *x, const double *y)
{
    double acc = 0.0;
    for (int i = 0; i < n; ++i) acc += x[i] * y[i];
    return acc;
}

void orc_dscal(int n, double a, double *x)
{
    for (int i = 0; i < n; ++i) x[i] *= a;
}

void orc_dcopy(int n, const double *x, double *y)
{
    for (int i = 0; i < n; ++i) y[i] = x[i];
}

/* ---- parallel BLAS-1 (src/openmp_vector.c) ------------------------------ */
static void par_axpy(int n, double a, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) y[i] += a * x[i];
}

static void par_scal(int n, double a, double *x)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) x[i] *= a;
}

static void par_copy(int n, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) y[i] = x[i];
}

static void par_zero(int n, double *x)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) x[i] = 0.0;
}

/* Products formed in parallel into tmp, then one thread sums them from 0.0 in
 * ascending order (src/openmp_vector.c:60-76). */
double orc_ddot_v3(int n, const double *x, const double *y, double *tmp)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) tmp[i] = x[i] * y[i];
    double acc = 0.0;
    for (int i = 0; i < n; ++i) acc += tmp[i];
    return acc;
}

/* ---- SpMV ---------------------------------------------------------------- */
static inline double row_dot(const orc_csr *A, unsigned i, const double *x)
{
    double t = 0.0;
    for (unsigned e = A->ptr[i]; e < A->ptr[i + 1]; ++e) t += A->val[e] * x[A->col[e]];
    return t;
}

void orc_mult(const orc_csr *A, const double *x, double *y)
{
    for (unsigned i = 0; i < A->n; ++i) y[i] += row_dot(A, i, x);
}

/* src/openmp_matrix.c:38-60 (static rows) */
static void par_mult(const orc_csr *A, const double *x, double *y)
{
    const int n = (int)A->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) y[i] += row_dot(A, (unsigned)i, x);
}

/* src/openmp_matrix.c:62-84 (dynamic,1 rows); same arithmetic */
static void par_mult_dynamic(const orc_csr *A, const double *x, double *y)
{
    const int n = (int)A->n;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) y[i] += row_dot(A, (unsigned)i, x);
}

/* ---- the reference on P MPI ranks -------------------------------------------
 * With P ranks (src/matrix.c:296-306) rank p owns the contiguous rows
 * [start(p), start(p+1)), the first m mod P ranks one row more; a row's
 * entries whose columns fall in its own rank's range form the diagonal block,
 * the others the off-diagonal block, both in the row's CSR order
 * (src/matrix.c:378-392).  The SpMV sums each row's diagonal-block entries,
 * then adds the sum of its off-diagonal-block entries
 * (src/solver_dynamic.c:158-160, src/openmp_matrix.c:39-60); a dot is each
 * rank's serial sum of its own rows (src/openmp_vector.c:60-76) combined by
 * MPI_Allreduce, here recursive doubling, the pairwise tree of MPICH's
 * allreduce for one double on a power-of-two communicator.  P = 1 is the
 * one-rank run (its empty off-diagonal block adds 0.0 to a row sum that is
 * never -0.0, so nothing changes). */
typedef struct {
    int P;
    unsigned start[ORC_MAX_RANKS + 1];
    double *tmp;   /* [n] products, as dot_temp_vec */
} orc_ranks;

static unsigned rank_start(unsigned m, int P, int p)
{
    const unsigned per = m / (unsigned)P, extra = m % (unsigned)P;
    return (unsigned)p * per + ((unsigned)p < extra ? (unsigned)p : extra);
}

static void ranks_init(orc_ranks *R, unsigned n, int P, double *tmp)
{
    R->P = P;
    for (int p = 0; p <= P; ++p) R->start[p] = p == P ? n : rank_start(n, P, p);
    R->tmp = tmp;
}

static int rank_of(const orc_ranks *R, unsigned i)
{
    int p = 0;
    while (i >= R->start[p + 1]) ++p;
    return p;
}

/* y[i] = 0 + (diagonal-block sum) + (off-diagonal-block sum) */
static void par_mult_ranks(const orc_csr *A, const orc_ranks *R, const double *x, double *y)
{
    const int n = (int)A->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const int p = rank_of(R, (unsigned)i);
        const unsigned lo = R->start[p], hi = R->start[p + 1];
        double td = 0.0, to = 0.0;
        for (unsigned e = A->ptr[i]; e < A->ptr[i + 1]; ++e) {
            const unsigned c = A->col[e];
            if (c >= lo && c < hi) td += A->val[e] * x[c];
            else to += A->val[e] * x[c];
        }
        double t = y[i];
        t += td;
        t += to;
        y[i] = t;
    }
}

/* MPI_Allreduce(MPI_SUM) of one double per rank by recursive doubling */
static double allreduce(double *v, int P)
{
    for (int d = 1; d < P; d *= 2)
        for (int p = 0; p < P; p += 2 * d) v[p] = v[p] + v[p + d];
    return v[0];
}

/* each rank's my_openmp_ddot_v3 (products, then a serial sum of its rows),
 * then MPI_Allreduce */
static double dot_ranks(const orc_ranks *R, int n, const double *x, const double *y)
{
    if (R->P == 1) return orc_ddot_v3(n, x, y, R->tmp);
    double part[ORC_MAX_RANKS];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) R->tmp[i] = x[i] * y[i];
    for (int p = 0; p < R->P; ++p) {
        double acc = 0.0;
        for (unsigned i = R->start[p]; i < R->start[p + 1]; ++i) acc += R->tmp[i];
        part[p] = acc;
    }
    return allreduce(part, R->P);
}

/* each rank's serial my_ddot, then MPI_Allreduce (src/solver_dynamic.c:
 * 100-103 and 323-324) */
static double ddot_ranks(const orc_ranks *R, const double *x, const double *y)
{
    double part[ORC_MAX_RANKS];
    for (int p = 0; p < R->P; ++p) {
        double acc = 0.0;
        for (unsigned i = R->start[p]; i < R->start[p + 1]; ++i) acc += x[i] * y[i];
        part[p] = acc;
    }
    return allreduce(part, R->P);
}

void orc_mult_ranks(const orc_csr *A, int ranks, const double *x, double *y)
{
    orc_ranks R;
    ranks_init(&R, A->n, ranks, NULL);
    par_mult_ranks(A, &R, x, y);
}

double orc_dot_ranks(int n, int ranks, const double *x, const double *y)
{
    double *tmp = malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    orc_ranks R;
    ranks_init(&R, (unsigned)n, ranks, tmp);
    const double v = dot_ranks(&R, n, x, y);
    free(tmp);
    return v;
}

/* ---- solver -------------------------------------------------------------- */
typedef struct {
    int n, M, cap;              /* cap = max_iter + 1 archive entries */
    double *r_old, *r_hat, *s, *y, *q, *tmp;
    double *p_set;              /* [M][n] */
    double *a_j, *b_j, *w_j, *eta, *zeta;
    double *aa, *ba, *wa;       /* seed archives alpha~, beta~, omega~ [cap] */
    double *pi;                 /* [M][cap] */
    char   *stop;
} orc_ws;

static int ws_alloc(orc_ws *w, int n, int M, int cap)
{
    memset(w, 0, sizeof(*w));
    w->n = n; w->M = M; w->cap = cap;
    size_t nb = (size_t)n * sizeof(double);
    w->r_old = malloc(nb); w->r_hat = malloc(nb); w->s = malloc(nb);
    w->y = malloc(nb); w->q = malloc(nb); w->tmp = malloc(nb);
    w->p_set = calloc((size_t)n * (size_t)M, sizeof(double));
    w->a_j = malloc(M * sizeof(double)); w->b_j = malloc(M * sizeof(double));
    w->w_j = malloc(M * sizeof(double)); w->eta = malloc(M * sizeof(double));
    w->zeta = malloc(M * sizeof(double));
    w->aa = malloc(cap * sizeof(double)); w->ba = malloc(cap * sizeof(double));
    w->wa = malloc(cap * sizeof(double));
    w->pi = malloc((size_t)cap * (size_t)M * sizeof(double));
    w->stop = calloc(M, 1);
    return (w->r_old && w->r_hat && w->s && w->y && w->q && w->tmp && w->p_set &&
            w->a_j && w->b_j && w->w_j && w->eta && w->zeta && w->aa && w->ba &&
            w->wa && w->pi && w->stop) ? 0 : -1;
}

static void ws_free(orc_ws *w)
{
    free(w->r_old); free(w->r_hat); free(w->s); free(w->y); free(w->q); free(w->tmp);
    free(w->p_set); free(w->a_j); free(w->b_j); free(w->w_j); free(w->eta);
    free(w->zeta); free(w->aa); free(w->ba); free(w->wa); free(w->pi); free(w->stop);
}

#define PI(w, j, i) ((w)->pi[(size_t)(j) * (size_t)(w)->cap + (size_t)(i)])

/* Per-shift recurrences and vector updates of one non-seed shift j at
 * iteration k (src/solver_dynamic.c:243-255).  The six vector passes are
 * serial, one thread per shift, like the reference's my_daxpy/my_dscal. */
static void shift_update(orc_ws *w, const double *sigma, int seed, int j, int k,
                         double *x_j, const double *r)
{
    const int n = w->n;
    double *p_j = w->p_set + (size_t)j * (size_t)n;
    const double d = sigma[seed] - sigma[j];
    const double pim1 = PI(w, j, k - 1);

    w->eta[j] = (w->ba[k - 1] / w->aa[k - 1]) * w->aa[k] * w->eta[j] - d * w->aa[k] * pim1;
    PI(w, j, k) = w->eta[j] + pim1;
    const double pik = PI(w, j, k);
    w->a_j[j] = (pim1 / pik) * w->aa[k];
    w->w_j[j] = w->wa[k] / (1.0 - w->wa[k] * d);

    orc_daxpy(n, w->w_j[j] / (pik * w->zeta[j]), w->q, x_j);
    orc_daxpy(n, w->a_j[j], p_j, x_j);
    orc_daxpy(n, w->w_j[j] / (w->a_j[j] * w->zeta[j] * pik), w->q, p_j);
    orc_daxpy(n, -w->w_j[j] / (w->a_j[j] * w->zeta[j] * pim1), w->r_old, p_j);
    w->zeta[j] = (1.0 - w->wa[k] * d) * w->zeta[j];
    w->b_j[j] = (pim1 / pik) * (pim1 / pik) * w->ba[k];
    orc_dscal(n, w->b_j[j], p_j);
    orc_daxpy(n, 1.0 / (pik * w->zeta[j]), r, p_j);
}

int orc_shifted_lopbicg(const orc_csr *A, double *x_set, double *r,
                        const double *sigma, int M, int seed,
                        double tol, int max_iter, orc_trace *tr)
{
    return orc_shifted_lopbicg_ranks(A, x_set, r, sigma, M, seed, tol, max_iter, 1, tr);
}

int orc_shifted_lopbicg_ranks(const orc_csr *A, double *x_set, double *r,
                              const double *sigma, int M, int seed,
                              double tol, int max_iter, int ranks, orc_trace *tr)
{
    if (!A || !x_set || !r || !sigma || M <= 0 || seed < 0 || seed >= M || max_iter < 1 ||
        ranks < 1 || ranks > ORC_MAX_RANKS || (ranks & (ranks - 1)) || (unsigned)ranks > A->n)
        return -1;
    const int n = (int)A->n;
    const int cap = max_iter + 1;
    orc_ws w;
    if (ws_alloc(&w, n, M, cap) != 0) { ws_free(&w); return -1; }
    orc_ranks R;
    ranks_init(&R, A->n, ranks, w.tmp);
    if (tr) {
        tr->n_switch = 0;
        tr->status = 0;
        tr->t_seed = tr->t_shift = tr->t_check = 0.0;
        if (tr->stop_iter) memset(tr->stop_iter, 0, (size_t)M * sizeof(int));
        if (tr->res_hist) for (size_t t = 0; t < (size_t)cap * M; ++t) tr->res_hist[t] = -1.0;
        if (tr->seed_hist) memset(tr->seed_hist, 0xff, (size_t)cap * sizeof(int));
        if (tr->dots) memset(tr->dots, 0, (size_t)cap * 5 * sizeof(double));
    }

    /* initialisation (src/solver_dynamic.c:101-123) */
    double rTr = ranks == 1 ? orc_ddot(n, r, r) : ddot_ranks(&R, r, r);
    orc_dcopy(n, r, w.r_hat);
    for (int j = 0; j < M; ++j) {
        orc_dcopy(n, r, w.p_set + (size_t)j * n);
        w.a_j[j] = 1.0; w.b_j[j] = 0.0; w.eta[j] = 0.0;
        PI(&w, j, 0) = 1.0; PI(&w, j, 1) = 1.0;
        w.zeta[j] = 1.0;
    }
    double dot_r = rTr, dot_zero = rTr;
    w.aa[0] = 1.0; w.ba[0] = 0.0;
    int k = 1, stop_count = 0, max_sigma = -1;

    while (stop_count < M && k < cap) {
        double *p_s = w.p_set + (size_t)seed * n;
        double *x_s = x_set + (size_t)seed * n;
        if (tr && tr->seed_hist) tr->seed_hist[k] = seed;
        const double t0 = wtime();

        /* seed step (src/solver_dynamic.c:156-210) */
        par_copy(n, r, w.r_old);
        par_zero(n, w.s);
        if (ranks == 1) par_mult(A, p_s, w.s);
        else par_mult_ranks(A, &R, p_s, w.s);
        par_axpy(n, sigma[seed], p_s, w.s);
        const double rTs = dot_ranks(&R, n, w.r_hat, w.s);
        w.aa[k] = rTr / rTs;

        par_axpy(n, -w.aa[k], w.s, r);          /* r now holds q */
        par_copy(n, r, w.q);
        par_zero(n, w.y);
        if (ranks == 1) par_mult_dynamic(A, r, w.y);
        else par_mult_ranks(A, &R, r, w.y);
        par_axpy(n, sigma[seed], r, w.y);
        const double qTq = dot_ranks(&R, n, r, r);
        const double qTy = dot_ranks(&R, n, r, w.y);
        w.wa[k] = qTq / qTy;
        const double rTr_old = rTr;

        par_axpy(n, w.aa[k], p_s, x_s);
        par_axpy(n, w.wa[k], r, x_s);
        par_axpy(n, -w.wa[k], w.y, r);          /* r now holds the new residual */
        dot_r = dot_ranks(&R, n, r, r);
        rTr = dot_ranks(&R, n, w.r_hat, r);
        w.ba[k] = (w.aa[k] / w.wa[k]) * (rTr / rTr_old);

        par_scal(n, w.ba[k], p_s);
        par_axpy(n, 1.0, r, p_s);
        par_axpy(n, -w.ba[k] * w.wa[k], w.s, p_s);
        if (tr && tr->dots) {
            double *dd = tr->dots + (size_t)k * 5;
            dd[0] = rTs; dd[1] = qTq; dd[2] = qTy; dd[3] = dot_r; dd[4] = rTr;
        }

        const double t1 = wtime();
        /* shifts (src/solver_dynamic.c:239-256) */
#pragma omp parallel for schedule(dynamic, 1)
        for (int j = 0; j < M; ++j) {
            if (j == seed || w.stop[j]) continue;
            shift_update(&w, sigma, seed, j, k, x_set + (size_t)j * n, r);
        }

        const double t2 = wtime();
        /* convergence check (src/solver_dynamic.c:278-303) */
        double max_zp = 1.0;
        for (int j = 0; j < M; ++j) {
            if (w.stop[j]) continue;
            const double zp = (j == seed) ? 1.0 : fabs(1.0 / (w.zeta[j] * PI(&w, j, k)));
            if (tr && tr->res_hist) tr->res_hist[(size_t)k * M + j] = zp * sqrt(dot_r / dot_zero);
            if (zp * zp * dot_r <= tol * tol * dot_zero) {
                w.stop[j] = 1;
                ++stop_count;
                if (tr && tr->stop_iter) tr->stop_iter[j] = k;
            } else if (zp > max_zp) {
                max_zp = zp;
                max_sigma = j;
            }
        }

        /* seed switch (src/solver_dynamic.c:312-341).  The reference never
         * initialises max_sigma (src/solver_dynamic.c:25), so a seed that
         * stops before any running shift ever had an estimate above 1 --
         * only possible when those shifts' estimates are NaN -- switches to
         * an undefined index there.  Here that case ends the solve after
         * this iteration with ORC_NO_SWITCH_TARGET in the trace status. */
        if (w.stop[seed] && stop_count < M && max_sigma < 0) {
            if (tr) tr->status |= ORC_NO_SWITCH_TARGET;
            if (tr) {
                const double t3 = wtime();
                tr->t_seed += t1 - t0;
                tr->t_shift += t2 - t1;
                tr->t_check += t3 - t2;
            }
            ++k;
            break;
        }
        if (w.stop[seed] && stop_count < M) {
            const int m = max_sigma;
            const double ds = sigma[seed] - sigma[m];
            for (int i = 1; i <= k; ++i) {
                const double t0 = PI(&w, m, i - 1), t1 = PI(&w, m, i);
                w.aa[i] = (t0 / t1) * w.aa[i];
                w.ba[i] = (t0 / t1) * (t0 / t1) * w.ba[i];
                w.wa[i] = w.wa[i] / (1.0 - w.wa[i] * ds);
            }
            orc_dscal(n, 1.0 / (w.zeta[m] * PI(&w, m, k)), r);
            rTr = ranks == 1 ? orc_ddot(n, w.r_hat, r) : ddot_ranks(&R, w.r_hat, r);
            for (int j = 0; j < M; ++j) { w.eta[j] = 0.0; w.zeta[j] = 1.0; }
            for (int i = 1; i <= k; ++i) {
                for (int j = 0; j < M; ++j) {
                    if (w.stop[j] || j == m) continue;
                    const double dm = sigma[m] - sigma[j];
                    w.eta[j] = (w.ba[i - 1] / w.aa[i - 1]) * w.aa[i] * w.eta[j] - dm * w.aa[i] * PI(&w, j, i - 1);
                    PI(&w, j, i) = w.eta[j] + PI(&w, j, i - 1);
                    w.zeta[j] = (1.0 - w.wa[i] * dm) * w.zeta[j];
                }
            }
            if (tr && tr->switch_k && tr->switch_to) {
                tr->switch_k[tr->n_switch] = k;
                tr->switch_to[tr->n_switch] = m;
            }
            if (tr) tr->n_switch++;
            seed = m;
        }
        if (tr) {
            const double t3 = wtime();
            tr->t_seed += t1 - t0;
            tr->t_shift += t2 - t1;
            tr->t_check += t3 - t2;
        }
        ++k;
    }

    if (tr) {
        if (tr->alpha) memcpy(tr->alpha, w.aa, (size_t)cap * sizeof(double));
        if (tr->beta) memcpy(tr->beta, w.ba, (size_t)cap * sizeof(double));
        if (tr->omega) memcpy(tr->omega, w.wa, (size_t)cap * sizeof(double));
        if (tr->zeta) memcpy(tr->zeta, w.zeta, (size_t)M * sizeof(double));
        if (tr->pi_k) for (int j = 0; j < M; ++j) tr->pi_k[j] = PI(&w, j, k - 1);
    }
    ws_free(&w);
    return k;
}

void orc_true_residuals(const orc_csr *A, const double *x_set, const double *b,
                        const double *sigma, int M, double *out)
{
    const int n = (int)A->n;
    double *t = malloc((size_t)n * sizeof(double));
    double bb = 0.0;
    for (int i = 0; i < n; ++i) bb += b[i] * b[i];
    for (int j = 0; j < M; ++j) {
        const double *x = x_set + (size_t)j * n;
        for (int i = 0; i < n; ++i) t[i] = 0.0;
        orc_mult(A, x, t);
        orc_daxpy(n, sigma[j], x, t);
        double dd = 0.0;
        for (int i = 0; i < n; ++i) {
            const double e = t[i] - b[i];
            dd += e * e;
        }
        out[j] = sqrt(dd) / sqrt(bb);
    }
    free(t);
}

/* ---- COO -> CSR (src/matrix.c:125-187, 206-232) ------------------------------
 * Bottom-up merge sort of the entry indices by row.  The merge takes the left
 * run's entry when the rows are equal (src/matrix.c:139, A[i] <= A[j]), so the
 * sort is stable, and every stable sort yields this one order: rows ascending,
 * input order inside a row.  The reference sorts row/col/val together; sorting
 * the indices and gathering once yields the same arrays. */
int orc_coo2csr(unsigned rows, unsigned nz, const unsigned *row, const unsigned *col,
                const double *val, unsigned *ptr, unsigned *ocol, double *oval)
{
    for (unsigned e = 0; e < nz; ++e)
        if (row[e] >= rows) return -1;
    unsigned *a = malloc(((size_t)nz + 1) * sizeof(unsigned));
    unsigned *t = malloc(((size_t)nz + 1) * sizeof(unsigned));
    if (!a || !t) { free(a); free(t); return -1; }
    for (unsigned e = 0; e < nz; ++e) a[e] = e;
    for (size_t w = 1; w < nz; w *= 2) {
        for (size_t lo = 0; lo < nz; lo += 2 * w) {
            const size_t mid = lo + w < nz ? lo + w : nz;
            const size_t hi = lo + 2 * w < nz ? lo + 2 * w : nz;
            size_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) t[k++] = (row[a[i]] <= row[a[j]]) ? a[i++] : a[j++];
            while (i < mid) t[k++] = a[i++];
            while (j < hi) t[k++] = a[j++];
        }
        unsigned *s = a; a = t; t = s;
    }
    /* ptr as src/matrix.c:221-229 fills it: row i ends where the sorted rows
     * pass i */
    size_t e = 0;
    ptr[0] = 0;
    for (unsigned i = 0; i < rows; ++i) {
        while (e < nz && row[a[e]] == i) {
            ocol[e] = col[a[e]];
            oval[e] = val[a[e]];
            ++e;
        }
        ptr[i + 1] = (unsigned)e;
    }
    free(a);
    free(t);
    return 0;
}
