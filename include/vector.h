/*
 * vector.h -- serial host BLAS-1 re-exported by the drop-in library so that
 * drivers written against the reference link unchanged.
 * Replaces src/vector.h:4-7 (implementations src/vector.c:3-27).
 */
#ifndef HBICG_VECTOR_H
#define HBICG_VECTOR_H

#ifdef __cplusplus
extern "C" {
#endif

void   my_daxpy(int n, double alpha, const double *x, double *y); /* y += alpha x      */
double my_ddot(int n, const double *x, const double *y);          /* serial, ascending */
void   my_dscal(int n, double alpha, double *x);                  /* x *= alpha        */
void   my_dcopy(int n, const double *x, double *y);               /* y = x             */

#ifdef __cplusplus
}
#endif
#endif
