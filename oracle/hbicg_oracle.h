/*
 * hbicg_oracle.h -- CPU restatement of the reference shifted BiCG solver with
 * seed switching.  TEST INFRASTRUCTURE ONLY: this library is the checker that
 * the HIP product path is compared against (and the timed CPU baseline).  No
 * product code may link, load or call it.
 *
 * The arithmetic follows src/solver_dynamic.c of the reference op for op
 * (same operation order, every axpy a separate rounding, dot products summed
 * serially in ascending index order), so that a build with
 * -ffp-contract=off reproduces the reference's FP64 results bit for bit.
 * The three reference variants (normal/static/dynamic) differ only in MPI
 * placement and OpenMP schedules and share this one restatement.
 */
#ifndef HBICG_ORACLE_H
#define HBICG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Plain CSR view, 0-based, 32-bit indices: the reference CSR_Matrix fields
 * (src/matrix.h:19-26) without ownership. */
typedef struct {
    const double   *val;
    const unsigned *col;
    const unsigned *ptr;
    unsigned        n;
    unsigned        nz;
} orc_csr;

/* Optional per-solve trace.  Any pointer may be NULL. */
typedef struct {
    int    *stop_iter;    /* [M]   iteration k at which shift j was stopped (0: never)      */
    int    *switch_k;     /* [M]   iteration at which the i-th seed switch fired             */
    int    *switch_to;    /* [M]   new seed chosen by the i-th switch                        */
    int     n_switch;     /* out:  number of switches                                        */
    double *alpha;        /* [max_iter+1] seed archive alpha~ as it stands on exit           */
    double *beta;         /* [max_iter+1] seed archive beta~                                 */
    double *omega;        /* [max_iter+1] seed archive omega~                                */
    double *res_hist;     /* [(max_iter+1)*M] estimated residual |1/(zeta pi)|*sqrt(dot_r/dot_zero)
                             of every unstopped shift at iteration k (row k), -1 when stopped */
    int    *seed_hist;    /* [max_iter+1] seed index used by iteration k                     */
    double *zeta;         /* [M]   zeta_j on exit                                            */
    double *pi_k;         /* [M]   pi_j[k_last] on exit                                      */
    double *dots;         /* [(max_iter+1)*5] rTs,qTq,qTy,dot_r,rTr of iteration k           */
    double  t_seed;       /* out: seconds in seed steps (src/solver_dynamic.c:212-221 buckets) */
    double  t_shift;      /* out: seconds in the shift loop                                  */
    double  t_check;      /* out: seconds in convergence check and seed switches             */
    int     status;       /* out: ORC_NO_SWITCH_TARGET when the solve ended for that reason   */
} orc_trace;

/* The seed stopped while other shifts still ran, none of which ever had a
 * residual estimate above the seed's (their estimates are NaN): there is no
 * shift to switch to.  The reference leaves this undefined (it switches to
 * an uninitialised index); the restatement ends the solve after that
 * iteration. */
#define ORC_NO_SWITCH_TARGET 4

/* One solve of (A + sigma_j I) x_j = b, j = 0..M-1.
 *   x_set : [M][n] shift-major, accumulated in place (zero on entry for x0 = 0)
 *   r     : b on entry, the last seed's residual on exit (reference r_loc)
 *   tol   : relative-residual stopping rule (reference EPS)
 *   max_iter : iteration cap (reference MAX_ITER); the loop runs while k < max_iter+1
 * Returns k = iterations + 1 like the reference, or -1 on bad arguments. */
int orc_shifted_lopbicg(const orc_csr *A, double *x_set, double *r,
                        const double *sigma, int M, int seed,
                        double tol, int max_iter, orc_trace *tr);

/* The same solve as the reference computes it on `ranks` MPI ranks (a power
 * of two up to ORC_MAX_RANKS, at most n): contiguous row blocks as
 * src/matrix.c:296-306 assigns them, each row's SpMV sum split into its
 * diagonal-block and off-diagonal-block parts (src/openmp_matrix.c:39-60,
 * src/solver_dynamic.c:158-160, 179-181), dots as per-rank serial sums
 * combined by MPI_Allreduce (recursive doubling).  Only these roundings
 * change with the rank count; ranks = 1 is orc_shifted_lopbicg. */
#define ORC_MAX_RANKS 64
int orc_shifted_lopbicg_ranks(const orc_csr *A, double *x_set, double *r,
                              const double *sigma, int M, int seed,
                              double tol, int max_iter, int ranks, orc_trace *tr);

/* The P-rank SpMV (y[i] += diagonal-block sum, then += off-diagonal-block
 * sum) and dot (per-rank serial sums, recursive doubling) of
 * orc_shifted_lopbicg_ranks, exported for the tests. */
void   orc_mult_ranks(const orc_csr *A, int ranks, const double *x, double *y);
double orc_dot_ranks(int n, int ranks, const double *x, const double *y);

/* y[i] += sum_j val*x[col], rows summed sequentially from 0.0 (reference
 * src/matrix.c:498-516 mult). */
void orc_mult(const orc_csr *A, const double *x, double *y);

/* True relative residual ||b - (A + sigma_j I) x_j|| / ||b|| of every shift,
 * computed in FP64 the way the reference's DISPLAY_ERROR block does
 * (src/solver_dynamic.c:405-421) but with the caller's b. */
void orc_true_residuals(const orc_csr *A, const double *x_set, const double *b,
                        const double *sigma, int M, double *out);

/* Serial BLAS-1 of src/vector.c:3-27 and the product-vector dot of
 * src/openmp_vector.c:60-76, exported so the tests can hold them against the
 * reference sources compiled under oracle/_ref. */
void   orc_daxpy(int n, double a, const double *x, double *y);
double orc_ddot(int n, const double *x, const double *y);
void   orc_dscal(int n, double a, double *x);
void   orc_dcopy(int n, const double *x, double *y);
double orc_ddot_v3(int n, const double *x, const double *y, double *tmp);

/* Threads the OpenMP build uses (omp_get_max_threads), 1 without OpenMP. */
int orc_num_threads(void);

/* COO -> CSR as src/matrix.c:206-232 does it: the entries are ordered by row
 * with a stable merge sort (src/matrix.c:125-187: on equal rows the left run
 * goes first, so entries of one row keep their input order), then ptr counts
 * them.  row, col, val: nz 0-based entries; ptr[rows+1], ocol[nz], oval[nz]
 * are written.  Returns 0, or -1 when an allocation fails or a row index is
 * >= rows (the reference does not check; its loop would drop such entries
 * from ptr). */
int orc_coo2csr(unsigned rows, unsigned nz, const unsigned *row, const unsigned *col,
                const double *val, unsigned *ptr, unsigned *ocol, double *oval);

#ifdef __cplusplus
}
#endif
#endif
