/*
 * driver.c -- a C program that uses libhbicg.so the way the reference's own
 * driver uses its solver (src/main.c:87-150): INFO_Matrix with per-rank
 * recvcounts/displs, csr_init_matrix, MPI_csr_load_matrix_block from a
 * Matrix Market file, the square check, x_loc_set zeroed, r_loc = b, one
 * shifted_lopbicg_* call, then the frees.  It includes only the headers under
 * include/ and links -lhbicg; no MPI.
 *
 * usage: driver A.mtx b.bin sigma.bin M seed variant exact out.bin
 *   b.bin: n doubles; sigma.bin: M doubles; variant: normal|static|dynamic;
 *   exact: 1 for the reference's summation order.
 * out.bin: int32 k, int32 n_switch, int32 stop_iter[M], int32 switch_k[M],
 *          int32 switch_to[M], double x_loc_set[M*n], double r_loc[n].
 * Prints "k=<k>" and returns 0, or prints the error and returns 1.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "solver.h"
#include "hbicg.h"

static double *read_doubles(const char *path, size_t count)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    double *a = (double *)malloc(count * sizeof(double));
    if (a && fread(a, sizeof(double), count, f) != count) {
        free(a);
        a = NULL;
    }
    fclose(f);
    return a;
}

int main(int argc, char *argv[])
{
    if (argc != 9) {
        fprintf(stderr, "usage: %s A.mtx b.bin sigma.bin M seed variant exact out.bin\n", argv[0]);
        return 1;
    }
    const int numprocs = 1;
    const int sigma_len = atoi(argv[4]);
    const int seed = atoi(argv[5]);
    const char *variant = argv[6];

    /* src/main.c:87-97 */
    INFO_Matrix A_info;
    A_info.recvcounts = (int *)malloc(numprocs * sizeof(int));
    A_info.displs = (int *)malloc(numprocs * sizeof(int));
    CSR_Matrix *A_loc_diag = (CSR_Matrix *)malloc(sizeof(CSR_Matrix));
    CSR_Matrix *A_loc_offd = (CSR_Matrix *)malloc(sizeof(CSR_Matrix));
    csr_init_matrix(A_loc_diag);
    csr_init_matrix(A_loc_offd);
    MPI_csr_load_matrix_block(argv[1], A_loc_diag, A_loc_offd, &A_info);
    if (A_info.cols != A_info.rows) {
        printf("Error: matrix is not square.\n");
        exit(1);
    }

    /* src/main.c:110-146, with b and sigma from files */
    const int vec_loc_size = A_loc_diag->rows;
    double *sigma = read_doubles(argv[3], (size_t)sigma_len);
    double *r_loc = read_doubles(argv[2], (size_t)vec_loc_size);
    double *x_loc_set = (double *)malloc((size_t)vec_loc_size * sigma_len * sizeof(double));
    if (!sigma || !r_loc || !x_loc_set) {
        fprintf(stderr, "cannot read the inputs\n");
        return 1;
    }
    for (size_t i = 0; i < (size_t)vec_loc_size * sigma_len; i++) x_loc_set[i] = 0;

    hbicg_params p;
    hbicg_default_params(&p);
    p.tol = 1e-10;
    p.exact_order = atoi(argv[7]);
    if (hbicg_set_params(&p) != 0) {
        fprintf(stderr, "hbicg_set_params failed\n");
        return 1;
    }

    int total_iter;
    if (strcmp(variant, "dynamic") == 0)
        total_iter = shifted_lopbicg_dynamic(A_loc_diag, A_loc_offd, &A_info, x_loc_set, r_loc, sigma, sigma_len, seed);
    else if (strcmp(variant, "static") == 0)
        total_iter = shifted_lopbicg_static(A_loc_diag, A_loc_offd, &A_info, x_loc_set, r_loc, sigma, sigma_len, seed);
    else
        total_iter = shifted_lopbicg_normal(A_loc_diag, A_loc_offd, &A_info, x_loc_set, r_loc, sigma, sigma_len, seed);
    if (total_iter < 0) {
        fprintf(stderr, "shifted_lopbicg_%s returned %d\n", variant, total_iter);
        return 1;
    }

    int *stop = (int *)calloc((size_t)sigma_len, sizeof(int));
    int *swk = (int *)calloc((size_t)sigma_len, sizeof(int));
    int *swt = (int *)calloc((size_t)sigma_len, sizeof(int));
    hbicg_report rep;
    const int nsw = hbicg_last_report(&rep, stop, sigma_len, swk, swt, sigma_len);
    FILE *f = fopen(argv[8], "wb");
    if (!f || nsw < 0) {
        fprintf(stderr, "cannot write the results\n");
        return 1;
    }
    fwrite(&total_iter, sizeof(int), 1, f);
    fwrite(&nsw, sizeof(int), 1, f);
    fwrite(stop, sizeof(int), (size_t)sigma_len, f);
    fwrite(swk, sizeof(int), (size_t)sigma_len, f);
    fwrite(swt, sizeof(int), (size_t)sigma_len, f);
    fwrite(x_loc_set, sizeof(double), (size_t)vec_loc_size * sigma_len, f);
    fwrite(r_loc, sizeof(double), (size_t)vec_loc_size, f);
    fclose(f);
    printf("k=%d\n", total_iter);

    /* src/main.c:166-169 */
    csr_free_matrix(A_loc_diag); free(A_loc_diag);
    csr_free_matrix(A_loc_offd); free(A_loc_offd);
    free(x_loc_set); free(r_loc); free(sigma);
    free(A_info.recvcounts); free(A_info.displs);
    free(stop); free(swk); free(swt);
    return 0;
}
