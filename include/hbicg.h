/*
 * hbicg.h -- extended C API of the MI355X shifted BiCG library: runtime
 * parameters (the reference fixes them at compile time, src/solver.h:13-31),
 * the per-solve report the reference only prints (src/solver_dynamic.c:
 * 274-306, 397-452), and a device-resident plan used by the benchmarks and
 * the PyTorch binding.  No function here has a reference counterpart except
 * where a "Replaces" line says so.
 */
#ifndef HBICG_HBICG_H
#define HBICG_HBICG_H

#include "solver.h"

#ifdef __cplusplus
extern "C" {
#endif

/* negative return codes */
#define HBICG_E_ARG        (-1)  /* bad argument (null pointer, M <= 0, seed out of range) */
#define HBICG_E_NOTSQUARE  (-2)  /* A_info->rows != A_info->cols                        */
#define HBICG_E_OFFD       (-3)  /* non-empty off-process block / row split             */
#define HBICG_E_HIP        (-4)  /* HIP runtime error                                   */
#define HBICG_E_NOMEM      (-5)  /* host or device allocation failed                    */
#define HBICG_E_NODEVICE   (-6)  /* no HIP device                                       */
#define HBICG_E_SCRATCH    (-7)  /* input larger than a no-allocation entry point's scratch */

/* status bits of hbicg_report.status (device-detected conditions).  The
 * reference does not trap breakdowns (src/solver_dynamic.c:167,189,245-253):
 * the arithmetic goes on with the resulting NaN or infinity, and so does
 * this solver, to the stopping rule or the iteration cap; these bits say
 * that it happened. */
#define HBICG_S_BREAKDOWN_RTS    1  /* (r^, s) = 0 or non-finite alpha~ */
#define HBICG_S_BREAKDOWN_QTY    2  /* (q, y) = 0 or non-finite omega~  */
#define HBICG_S_NO_SWITCH_TARGET 4  /* the seed stopped while shifts still ran, none of which
                                       ever had a residual estimate above the seed's (only
                                       NaN estimates do that): the solve ends after that
                                       iteration (the reference switches to an
                                       uninitialised index, src/solver_dynamic.c:25,318) */
#define HBICG_S_BREAKDOWN_SHIFT  8  /* a running shift's 1 - omega~ (sigma_s - sigma_j) or
                                       pi_j[k] was 0 or not finite (a NaN shift, or a shift
                                       that makes the recurrence divide by zero) */
#define HBICG_S_SYNC_TIMEOUT    16  /* a fused launch's workgroup gave up waiting for a value
                                       its first workgroups publish; the results are not
                                       valid (never expected: those are dispatched first) */

typedef struct hbicg_params {
    double tol;          /* stopping rule, default EPS (1e-12)                           */
    int    max_iter;     /* iteration cap, default MAX_ITER (2000)                       */
    int    window;       /* shift-update window 1..1024 (exact_order: capped at 16); 0 = variant
                            default: normal 1, static/dynamic min(512, max_iter), halved
                            while its 2W+1 ring vectors would take over a quarter of the
                            free device memory (exact_order 16).  x_j and p_j are written
                            at a window's end; the solve's last window writes x_j only  */
    int    exact_order;  /* 1: serial dots, row-sequential SpMV and per-iteration replay of
                            the shift updates (reference order and roundings, any window);
                            0: tree reductions and the combined-coefficient window flush
                            (fused multiply-adds; x_j depend on the window by roundings) */
    int    graph;        /* -1 variant default (dynamic on), 0 off, 1 on                  */
    int    chunk;        /* iterations between host checks of the done flag (default 16) */
    int    device;       /* HIP device ordinal                                            */
    int    exit_on_error;/* 1 (default): non-square input prints and exit(1) like the reference */
    int    upload_x;     /* 1 (default): x_loc_set copied to the GPU (honours x0 != 0)   */
    int    profile;      /* 1: time every kernel with HIP events (direct launches)       */
    double growth_limit; /* default mode, > 0: a window closes early, its last iteration
                            applied on its own, when the size bound of a shift's regrouped
                            p exceeds this multiple of one iteration's own terms, which
                            bounds what regrouping can add to x's rounding error at about
                            this multiple of the unit roundoff.  Default 1e8; 0: off    */
    int    trace;        /* 1: record the per-iteration residual estimate of every shift  */
    int    value_index;  /* 1 (default): a matrix with at most 256 distinct values is read
                            by the SpMV as one byte per entry into a table of its values,
                            and its tiles whose entries lie on at most 16 diagonals as one
                            byte per row and diagonal (no columns, no row pointers); the
                            products and their order are those of the CSR read, so results
                            are bitwise unchanged.  0: 8-byte CSR values                   */
    int    q_mode;       /* SpMV 2 input q = r - alpha~ s (default mode only): 0 (default)
                            auto: formed at every gathered column when most entries lie
                            within 16-bit column offsets of their tile (stencils), else
                            formed for every row first and gathered; 1: always at the
                            gathered columns; 2: always beforehand.  Same results.        */
    int    fused;        /* default mode: 1 (default) four launches per iteration --
                            the alpha~ finalize runs as the first workgroup of SpMV
                            2's launch (where SpMV 2 forms q at its gathered columns
                            in one pass; else it keeps its own launch), the omega~
                            finalize as the first workgroup of the seed update's, the
                            scalar stage as the first workgroup of the shift kernel's
                            (beta~ handed to the seed p update as soon as it is
                            formed, the window decision after), each handing its
                            values to the launch's other workgroups, and the combined
                            window coefficients run in the shift kernel's next
                            workgroups (with too many shifts for its tiles to stay
                            resident beside them, about 256 and more, the stage runs
                            as in 3); 3: the same with the scalar stage in a launch
                            of its own (five launches); 2: also alpha~ on its own
                            (six); 0: eight launches.  Same results bit for bit; the
                            batched solver always uses eight, exact order seven (no
                            coefficient kernel)                                         */
    int    panel_cols;   /* seed SpMVs in column panels of this many columns: one pass
                            per panel over the entries whose columns lie in it, row sums
                            carried between passes, so the gathers of a pass stay within
                            one XCD's L2.  0 (default) auto: 262144 columns (2 MiB of x)
                            for a matrix read as 8-byte CSR whose columns are scattered
                            beyond 16-bit tile offsets, with x over 4 MiB and every row's
                            columns ascending; -1 off.  Needs ascending columns (else
                            off).  Exact order is bitwise unchanged (one lane sums a row
                            in CSR order across the passes); the default order sums a
                            row's panels in turn, a deterministic regrouping           */
    int    x_zero;       /* 1: x_set's incoming contents are never read -- every x_j
                            starts from 0, so the caller need not clear it (the default
                            mode's first window flush writes x_j without reading it,
                            exact order clears x_set first; a host x_loc_set is neither
                            uploaded nor cleared).  0 (default): device plans accumulate
                            onto x_set; the drop-ins upload x_loc_set (upload_x)      */
} hbicg_params;

typedef struct hbicg_report {
    int    k;            /* returned value, iterations + 1 */
    int    stop_count;   /* shifts that met the stopping rule */
    int    n_switch;     /* seed switches */
    int    seed_final;   /* seed of the last iteration */
    int    status;       /* HBICG_S_* bits */
    int    M, n;
    double dot_r, dot_zero, rTr;
    double t_setup_ms;   /* plan creation (allocation, upload, SpMV binning) */
    double t_solve_ms;   /* device time of init + iteration loop              */
    double t_total_ms;   /* wall time of the whole call                        */
    int    window;       /* shift-update window the solve ran with             */
    int    value_table;  /* distinct values the SpMV read through its table (0: none) */
    int    diag_tiles;   /* SpMV tiles stored by diagonals (value table only)         */
    int    q_unfused;    /* 1: most entries' columns lie beyond 16-bit offsets of their
                            tile, so SpMV 2 gathered q = r - alpha s formed beforehand
                            (one gather per entry instead of two); results unchanged */
    int    panels;       /* column panels of the seed SpMVs (1: one pass)             */
    int    const_tiles;  /* of diag_tiles, those whose every diagonal holds one value:
                            read as presence bits per row and the diagonals' values */
} hbicg_report;

void hbicg_default_params(hbicg_params *p);
/* Process-wide parameters used by the shifted_lopbicg_* entry points. */
int  hbicg_set_params(const hbicg_params *p);
int  hbicg_get_params(hbicg_params *p);

/* Residual trace of the last drop-in solve (params.trace = 1): row k, shift
 * j = |1/(zeta_j pi_j[k])| sqrt((r,r)/(b,b)), the estimate the reference
 * prints under DISPLAY_SIGMA_RESIDUAL (src/solver_dynamic.c:274-292), for the
 * shifts still running at iteration k (the one that stops included); -1 for
 * shifts stopped earlier and for row 0.  out is [rows][M] row-major.
 * Returns the rows written (the solve's k) or a negative code. */
int  hbicg_last_residual_trace(double *out, int M, int rows);

/* Report of the last solve on this process.  stop_iter[M] gets the
 * iteration at which each shift stopped (0 = never); switch_k / switch_to
 * (capacity max_switch) the switch log.  Any array may be NULL.
 * Returns the number of switches or a negative code. */
int  hbicg_last_report(hbicg_report *rep, int *stop_iter, int M,
                       int *switch_k, int *switch_to, int max_switch);

/* ---- device-resident plan ------------------------------------------------ */
typedef struct hbicg_csr {
    const double   *val;
    const unsigned *col;
    const unsigned *ptr;
    unsigned        n;
    unsigned        nz;
} hbicg_csr;

typedef struct hbicg_plan hbicg_plan;

/* Uploads A (a_on_device = 0) or references caller device memory (1),
 * bins the rows for the SpMV and allocates the workspace for M shifts.
 * With params.value_index = 1 and a matrix that qualifies, the solver's
 * SpMV reads the value table built here, so values the caller changes in
 * its device memory afterwards need a new plan (or value_index = 0, which
 * reads A in place at every solve).
 * stream: a hipStream_t, or NULL for a stream of the plan's own.
 * Returns NULL on failure (hbicg_errno() tells why). */
hbicg_plan *hbicg_plan_create(const hbicg_csr *A, int a_on_device, int M,
                              const hbicg_params *p, void *stream);
void        hbicg_plan_destroy(hbicg_plan *pl);
int         hbicg_errno(void);

/* One solve with the plan.  on_device = 1: b, x_set and r_out are device
 * pointers (x_set accumulated in place, r_out may be NULL); 0: host
 * pointers (x_set uploaded when params.upload_x, else zeroed).  sigma is a
 * host array of M shifts.  Returns k or a negative code; rep may be NULL. */
int hbicg_plan_solve(hbicg_plan *pl, const double *b, const double *sigma, int seed,
                     double *x_set, double *r_out, int on_device, hbicg_report *rep);

/* Per-shift true residual ||b - (A + sigma_j I) x_j|| / ||b|| in FP64 on the
 * GPU (pointers as in hbicg_plan_solve; out is a host array of M). */
int hbicg_plan_true_residuals(hbicg_plan *pl, const double *b, const double *sigma,
                              const double *x_set, int on_device, double *out);

/* hbicg_plan_true_residuals without allocating: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops).  b, sigma (M shifts), x_set
 * ([M][n]) and out (M doubles) are device pointers; the kernels' partial sums
 * live in a static 1 MiB scratch array of each GPU, so at most 65536 shifts
 * (HBICG_E_SCRATCH beyond).  The work is enqueued on the plan's stream and
 * the call returns without waiting: out holds the residuals once that stream
 * gets there.  Results equal hbicg_plan_true_residuals' bit for bit.  Calls
 * that overlap on one GPU (other streams, other threads) share the scratch
 * and race. */
int hbicg_plan_true_residuals_noalloc(hbicg_plan *pl, const double *b, const double *sigma,
                                      const double *x_set, double *out);

/* y = (A + shift I) x with the plan's SpMV (device pointers). */
int hbicg_plan_spmv(hbicg_plan *pl, const double *x, double *y, double shift);

/* Per-kernel statistics of the plan's last solve run with params.profile:
 * one entry per launch of an iteration: [0] seed SpMV 1, [1] its finalize
 * (alpha~), [2] fused q + SpMV 2, [3] its finalize (omega~), [4] seed
 * update, [5] scalar stage, [6] combined window coefficients, [7] the shift
 * kernel's launches that ran the seed p update alone, [8] those that also
 * applied a window flush.  launches counts the launches (executed
 * iterations), total_ms the summed HIP-event durations, bytes the
 * algorithmic HBM bytes those launches had to move, flops ([8] only) twice
 * the flush's multiply-adds.  With the fused launches (params.fused, the
 * default order) [3] and [6] launch nothing, and [1] too where alpha~ joins
 * SpMV 2: the alpha~ finalize is then timed in [2], omega~'s in [4], the
 * coefficients in [7] and [8].  Returns the number of entries filled (at
 * most 9). */
typedef struct hbicg_kernel_stat {
    int    launches;
    double total_ms;
    double bytes;
    int    replays;      /* [2] only: extra back-to-back launches timed on their own */
    double replay_ms;    /*   (same inputs and outputs, no event gaps between them)  */
    double flops;
} hbicg_kernel_stat;
int hbicg_plan_kernel_stats(hbicg_plan *pl, hbicg_kernel_stat *stats, int nmax);

/* The plan's residual trace (plan created with params.trace = 1), laid out
 * as for hbicg_last_residual_trace: min(rows, k) rows of M doubles. */
int hbicg_plan_residual_trace(hbicg_plan *pl, double *out, int rows);

/* Stop iterations / switch log of the plan's last solve (host arrays). */
int hbicg_plan_report(hbicg_plan *pl, hbicg_report *rep, int *stop_iter,
                      int *switch_k, int *switch_to, int max_switch);

/* ---- batched solver ----------------------------------------------------------
 * Many independent shifted problems of different n and shift counts solved
 * together: every phase of the iteration is one launch over the tiles of all
 * problems (no counterpart in the reference, which solves one family per
 * call). */
typedef struct hbicg_batch hbicg_batch;

typedef struct hbicg_tuning {
    int    window;       /* shift-update window chosen                          */
    int    group_cap;    /* exact order: shifts per shift-kernel workgroup (0: all
                            of a problem); always 0 in the default mode         */
    double ms_per_iter;  /* measured time per batch iteration                   */
} hbicg_tuning;

/* A[p] (host, a_on_device = 0, or device memory), M[p] shifts each.
 * params: tol, max_iter, window (0 or above 64: 64, exact_order 16), exact_order, graph
 * (0 off), chunk.  Ring storage is sized for the largest window the mode
 * allows, so autotuning can change it. */
hbicg_batch *hbicg_batch_create(const hbicg_csr *A, int a_on_device, int nprob, const int *M,
                                const hbicg_params *p, void *stream);
void hbicg_batch_destroy(hbicg_batch *bt);

/* b[p] device vectors, x[p] device [M_p][n_p] arrays (accumulated in place),
 * sigma[p] host arrays, seed[p] first seeds.  k_out[p] (optional) receives
 * each problem's k; solve_ms (optional) the device time of the batch. */
int hbicg_batch_solve(hbicg_batch *bt, const double *const *b, const double *const *sigma,
                      const int *seed, double *const *x, int *k_out, double *solve_ms);
int hbicg_batch_report(hbicg_batch *bt, int p, hbicg_report *rep, int *stop_iter,
                       int *switch_k, int *switch_to, int max_switch);
int hbicg_batch_true_residuals(hbicg_batch *bt, int p, const double *b, const double *sigma,
                               const double *x_set, double *out);

/* ---- adjoint support (device pointers) ------------------------------------- */
/* A^T of an n x n device CSR into tptr[n+1], tcol[nz], tval[nz] (device),
 * rows listing their entries in ascending column order and duplicate entries
 * in their order in A (a stable transpose: the result does not depend on the
 * order the device's atomics place entries in).  stream: hipStream_t or
 * NULL.  Synchronises the stream. */
int hbicg_csr_transpose(const hbicg_csr *A, unsigned *tptr, unsigned *tcol, double *tval, void *stream);

/* The same transpose in two steps, for a pattern that stays fixed while its
 * values change (an optimiser updating A in place): the pattern tptr[n+1],
 * tcol[nz] and tsrc[nz], the entry of A that each entry of A^T holds (A->val
 * may be NULL; synchronises the stream), then, for every new set of values,
 * tval[e] = val[tsrc[e]] with hbicg_gather_values (only enqueued on the
 * stream; allocates nothing).  Together they give hbicg_csr_transpose's
 * tptr, tcol and tval bit for bit. */
int hbicg_csr_transpose_pattern(const hbicg_csr *A, unsigned *tptr, unsigned *tcol, unsigned *tsrc,
                                void *stream);
int hbicg_gather_values(const double *val, const unsigned *src, unsigned nz, double *out, void *stream);

/* Gradients of a loss through x_k = (A + sigma_k I)^{-1} b from the adjoint
 * solutions lam_k = w_k mu_k (w NULL: w_k = 1), all [M][n] shift-major:
 *   grad_b = sum_k lam_k,  grad_sigma_k = -lam_k . x_k,
 *   grad_val[e] = -sum_k lam_k[row(e)] x_k[col(e)] on A's pattern.
 * Any output may be NULL.  Synchronises the stream. */
int hbicg_adjoint_gradients(const hbicg_csr *A, int M, const double *mu, const double *w,
                            const double *x, double *grad_b, double *grad_sigma, double *grad_val,
                            void *stream);

/* hbicg_adjoint_gradients with a caller's device workspace of work_bytes
 * (at least hbicg_adjoint_work_bytes(n, M) when grad_sigma is wanted, else
 * work may be NULL): allocates nothing and only enqueues on the stream. */
size_t hbicg_adjoint_work_bytes(unsigned n, int M);
int hbicg_adjoint_gradients_ws(const hbicg_csr *A, int M, const double *mu, const double *w,
                               const double *x, double *grad_b, double *grad_sigma, double *grad_val,
                               void *work, size_t work_bytes, void *stream);

/* ---- Matrix Market ingest on the GPU ------------------------------------------ */
/* COO -> CSR on the device: the arrays the reference's coo2csr builds
 * (Replaces src/matrix.c:206-232 with its stable merge sort by row,
 * src/matrix.c:125-187): rows ascending, the entries of a row in input
 * order, duplicates kept.  row, col, val: nz entries, 0-based, device
 * memory (may be NULL when nz = 0); ptr[rows+1], ocol[nz], oval[nz]: device
 * outputs.  Every row index
 * must be < rows and every column < cols (HBICG_E_ARG otherwise, outputs
 * untouched).  stream: hipStream_t or NULL; synchronises it.  Returns 0 or a
 * negative code. */
int hbicg_coo2csr_device(const unsigned *row, const unsigned *col, const double *val, unsigned nz,
                         unsigned rows, unsigned cols, unsigned *ptr, unsigned *ocol, double *oval,
                         void *stream);

/* Reads a Matrix Market coordinate file on the host (real, integer or
 * pattern; symmetric storage is not expanded, as in src/matrix.c:26-94) and
 * builds its CSR on the device with hbicg_coo2csr_device.  A receives device
 * arrays owned by the caller (release with hbicg_csr_free_device); info
 * (optional) gets nz, rows, cols, the banner code and the single-rank
 * recvcounts/displs, like src/matrix.c:402-419 on one rank.  A non-square
 * matrix returns HBICG_E_NOTSQUARE. */
int  hbicg_csr_load_device(const char *filename, hbicg_csr *A, INFO_Matrix *info, void *stream);
void hbicg_csr_free_device(hbicg_csr *A);

/* Times candidate layouts on the batch for `iters` iterations each and keeps
 * the fastest (x is overwritten): windows 32/64/128 in the default mode
 * (those the batch's rings hold: at creation the largest window is halved
 * from 128 while the rings of all problems would take over a quarter of the
 * free device memory; the window before tuning is 64; a window's flushes
 * only count once a trial reaches them, so give `iters` at least the solve's
 * length), (window 4/8/16) x (shifts per workgroup all/16/4) in exact order. */
int hbicg_batch_autotune(hbicg_batch *bt, const double *const *b, const double *const *sigma,
                         const int *seed, double *const *x, int iters, hbicg_tuning *best);

#ifdef __cplusplus
}
#endif
#endif
