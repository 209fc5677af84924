/*
 * hb_adjoint.h -- launch layer of hb_adjoint.hip (device CSR transpose and
 * adjoint gradient assembly).  Returns the hipError_t of the launches.
 */
#ifndef HB_ADJOINT_H
#define HB_ADJOINT_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A^T of an n x n CSR on the device; rows of A^T list their entries in
 * ascending column order, duplicate entries in their order in A (a stable
 * transpose).  tsrc[nnz]: the entry of A each entry of A^T holds.  tval
 * NULL: the pattern and tsrc only (val unused).  work: 2*n + ceil(n/256)
 * unsigned. */
int hb_transpose(int n, unsigned nnz, const unsigned *ptr, const unsigned *col, const double *val,
                 unsigned *tptr, unsigned *tcol, double *tval, unsigned *tsrc, unsigned *work,
                 hipStream_t s);

/* out[e] = val[src[e]] for e < nnz */
int hb_gather_vals(size_t nnz, const double *val, const unsigned *src, double *out, hipStream_t s);

/* lam_k = w_k mu_k (w NULL: 1).  grad_b = sum_k lam_k, grad_sigma_k =
 * -lam_k . x_k, grad_val[e] = -sum_k lam_k[row(e)] x_k[col(e)]; any output
 * may be NULL.  work: M * min(256, ceil(n/256)) doubles. */
int hb_adjoint_grads(int n, int M, const unsigned *ptr, const unsigned *col, const double *mu,
                     const double *w, const double *x, double *grad_b, double *grad_sigma,
                     double *grad_val, double *work, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
