/*
 * matrix.h -- sparse-matrix types of the drop-in library, layout-identical
 * to the reference's (src/matrix.h:10-33, MM_typecode from src/mmio.h:16),
 * without the reference's <mpi.h> dependency: the library runs on one GPU.
 *
 * All index arrays are 0-based unsigned 32-bit, values FP64.
 */
#ifndef HBICG_MATRIX_H
#define HBICG_MATRIX_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef MM_IO_H
/* Matrix Market type code: 4 characters (src/mmio.h:16). */
typedef char MM_typecode[4];
#endif

/* src/matrix.h:10-17 */
typedef struct {
    double       *val;
    unsigned int *row;
    unsigned int *col;
    unsigned int  nz;
    unsigned int  rows;
    unsigned int  cols;
} COO_Matrix;

/* src/matrix.h:19-26 */
typedef struct {
    double       *val;
    unsigned int *col;
    unsigned int *ptr;
    unsigned int  nz;
    unsigned int  rows;
    unsigned int  cols;
} CSR_Matrix;

/* src/matrix.h:28-33.  On one GPU: recvcounts = {rows}, displs = {0}. */
typedef struct {
    unsigned int nz, rows, cols;
    MM_typecode  code;
    int         *recvcounts;
    int         *displs;
} INFO_Matrix;

/* Host helpers of src/matrix.h:35-56 kept so existing drivers link.
 * Replaces src/matrix.c:9-24 (init/free) */
void coo_init_matrix(COO_Matrix *m);
void coo_free_matrix(COO_Matrix *m);
void csr_init_matrix(CSR_Matrix *m);
void csr_free_matrix(CSR_Matrix *m);

/* Replaces src/matrix.c:26-94 (Matrix Market -> COO, 1- to 0-based).
 * Returns 1 for a symmetric file, 0 otherwise, -1 on error (the reference
 * exits instead).  Real, integer and pattern files are read. */
int  coo_load_matrix(char *filename, COO_Matrix *coo);
/* Replaces src/matrix.c:206-232: stable row sort, then CSR. */
void coo2csr(COO_Matrix *in, CSR_Matrix *out);
/* Replaces src/matrix.c:234-242. */
int  csr_load_matrix(char *filename, CSR_Matrix *m);
/* Single-GPU replacement of src/matrix.c:402-419: the whole matrix becomes
 * the diagonal block, the off-diagonal block is empty, recvcounts = {rows},
 * displs = {0} (info->recvcounts/displs must point at one int each). */
void MPI_csr_load_matrix_block(char *filename, CSR_Matrix *diag, CSR_Matrix *offd, INFO_Matrix *info);
/* Replaces src/matrix.c:498-516: y[i] += sum_j val*x[col], rows in order. */
void mult(CSR_Matrix *A_loc, double *x, double *y_loc);

#ifdef __cplusplus
}
#endif
#endif
