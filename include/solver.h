/*
 * solver.h -- C-ABI drop-in for the reference's shifted BiCG entry points
 * (src/solver.h:33-35), backed by gfx950 HIP kernels on one GPU.
 *
 * Contract, as in the reference (src/solver_dynamic.c:3-463):
 *   A_loc_diag : the n x n matrix in CSR (host memory, read only)
 *   A_loc_offd : the off-process block; must be empty (nz == 0) on one GPU
 *   A_info     : rows == cols == A_loc_diag->rows
 *   x_loc_set  : [sigma_len][rows] shift-major, accumulated in place
 *                (zero on entry for the usual x0 = 0)
 *   r_loc      : b on entry; on exit the last seed's residual (rescaled at
 *                every seed switch), as in the reference
 *   sigma      : sigma_len shifts; seed: index of the first seed system
 * Returns k = iterations + 1 (the reference's convention), or a negative
 * HBICG_E* code (see hbicg.h) for input the GPU path cannot take.  A
 * non-square A_info prints the reference's message and calls exit(1), like
 * the reference, unless hbicg_params.exit_on_error is 0.
 *
 * The three variants differ only in when the per-shift vector updates are
 * applied (the reference's differ only in MPI placement and OpenMP schedule):
 *   normal  - shift vectors updated every iteration (window 1)
 *   static  - shift vectors updated in windows of hbicg_params.window
 *             iterations, on the solver stream
 *   dynamic - windowed shift updates, iteration replayed as hipGraph chunks
 * With hbicg_params.exact_order = 1 all three are bit-identical to each
 * other and to the reference's arithmetic.  In the default mode static and
 * dynamic are bit-identical to each other; normal (window 1) agrees with
 * them to rounding: x within 1e-10 relative without a seed switch, and
 * after a switch k, stop and switch iterations within 2 and x within 1e-9.
 * Stopping rule and iteration cap come from hbicg_set_params (defaults
 * EPS = 1e-12 and MAX_ITER = 2000, src/solver.h:13-14).
 */
#ifndef HBICG_SOLVER_H
#define HBICG_SOLVER_H

#include "matrix.h"
#include "vector.h"

#ifdef __cplusplus
extern "C" {
#endif

#define EPS      1.0e-12   /* default relative-residual stopping rule */
#define MAX_ITER 2000      /* default iteration cap                   */

/* Replaces src/solver_dynamic.c:3 (src/solver.h:33) */
int shifted_lopbicg_dynamic(CSR_Matrix *A_loc_diag, CSR_Matrix *A_loc_offd, INFO_Matrix *A_info,
                            double *x_loc_set, double *r_loc, double *sigma, int sigma_len, int seed);
/* Replaces src/solver_static.c:3 (src/solver.h:34) */
int shifted_lopbicg_static(CSR_Matrix *A_loc_diag, CSR_Matrix *A_loc_offd, INFO_Matrix *A_info,
                           double *x_loc_set, double *r_loc, double *sigma, int sigma_len, int seed);
/* Replaces src/solver_normal.c:3 (src/solver.h:35) */
int shifted_lopbicg_normal(CSR_Matrix *A_loc_diag, CSR_Matrix *A_loc_offd, INFO_Matrix *A_info,
                           double *x_loc_set, double *r_loc, double *sigma, int sigma_len, int seed);

#ifdef __cplusplus
}
#endif
#endif
