/*
 * hb_launch.h -- the thin launch layer between the C host driver and the
 * gfx950 kernels in hb_kernels.hip.  Every function enqueues work on the
 * given stream and returns the hipError_t of the launch (0 on success).
 */
#ifndef HB_LAUNCH_H
#define HB_LAUNCH_H

#include <hip/hip_runtime_api.h>
#include "hb_shared.h"

#ifdef __cplusplus
extern "C" {
#endif

/* r^ = r_0 = p_j = b (d->b_in) for every shift and the reference's scalar
 * initialisation with seed d->seed_in (src/solver_dynamic.c:101-123);
 * (b,b) reduced on the device. */
int hb_k_init(const hb_dev *d, hipStream_t s);

/* One iteration: seed SpMV 1, fused q + SpMV 2, seed update, each with its
 * one-workgroup finalize (the last one is the scalar stage: shift
 * recurrences, convergence check, seed switch), then the window flush
 * (combined coefficients and the shift kernel).  With fused launches the
 * alpha~ and omega~ finalizes run as the first workgroups of SpMV 2 and the
 * seed update, and the scalar stage and the combined coefficients as the
 * first workgroups of the shift kernel.  Each kernel returns at once when
 * the solve is done. */
int hb_k_iteration(const hb_dev *d, hipStream_t s);
/* 1 if hb_k_stage launches a kernel for `stage` (fused launches run some
 * stages inside a neighbouring one) */
int hb_k_stage_runs(const hb_dev *d, int stage);
/* Tile workgroups of the shift kernel that runs the scalar stage (d->shift_tw)
 * for n_tiles flush tiles and M shifts on `device`; 0 when M leaves too few
 * resident tiles (the stage then keeps a launch of its own); < 0 on a HIP
 * error. */
int hb_k_shift_stage_tiles(int n_tiles, int M, int device);

/* The kernels of an iteration, one launch at a time (per-kernel timing). */
#define HB_STAGE_SPMV1  0   /* SpMV s = (A + sigma_s I) p_s, (r^, s) partials   */
#define HB_STAGE_FIN1   1   /* alpha~                                           */
#define HB_STAGE_SPMV2  2   /* q = r - alpha~ s fused into y = (A + sigma_s I) q */
#define HB_STAGE_FIN2   3   /* omega~                                           */
#define HB_STAGE_SEED   4   /* x_s, r updates, (r,r), (r^,r) partials           */
#define HB_STAGE_SCALAR 5   /* beta~ and the scalar stage                       */
#define HB_STAGE_COEF   6   /* combined window coefficients (default mode)      */
#define HB_STAGE_SHIFT  7   /* seed p update and the window flush               */
#define HB_NSTAGE       8
int hb_k_stage(const hb_dev *d, int stage, hipStream_t s);

/* Copies {done, k} of the device state into host-visible memory. */
int hb_k_publish(const hb_dev *d, int *flag_dev, hipStream_t s);

/* After the iteration loop: the solve's last, x-only window flush when the
 * shift kernel runs the scalar stage (d->fused_s), then a seed-switch rescale
 * of r left pending by the last iteration. */
int hb_k_finalize(const hb_dev *d, hipStream_t s);

/* out[j] = ||b - (A + sigma_j I) x_j|| / ||b|| for every shift j (sigma, x,
 * b, out device pointers).  part must hold 2*M*hb_k_resid_blocks(n, M)
 * doubles; the grid, and so the roundings, depend on (n, M) only. */
int hb_k_resid_blocks(int n, int M);
int hb_k_true_residual(const hb_dev *d, const double *b, const double *sigma, const double *x,
                       double *part, double *out, hipStream_t s);

/* p[0 .. count) = v. */
int hb_k_fill(double *p, size_t count, double v, hipStream_t s);

/* col16 / cflag of the plan's SpMV tiles (see hb_dev); once per plan. */
int hb_k_compress_cols(const hb_dev *d, short *col16, unsigned *cflag, hipStream_t s);

/* y = A x + shift*x through the row-binned SpMV (exact: row-sequential sums). */
int hb_k_spmv(const hb_dev *d, const double *x, double *y, double shift, hipStream_t s);

/* Batched counterparts: each launch covers the tiles of every problem of
 * the batch; done problems return at once. */
int hb_kb_init(const hb_batch *B, hipStream_t s);
int hb_kb_stage(const hb_batch *B, int stage, hipStream_t s);
int hb_kb_iteration(const hb_batch *B, hipStream_t s);
int hb_kb_publish(const hb_batch *B, int *flag_dev, hipStream_t s);
int hb_kb_compact(const hb_batch *B, hipStream_t s);
int hb_kb_finalize(const hb_batch *B, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
