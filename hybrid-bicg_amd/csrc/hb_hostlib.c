/*
 * hb_hostlib.c -- host helpers the reference declares in src/vector.h and
 * src/matrix.h, re-exported so that drivers written against the reference
 * link against the drop-in unchanged.  Matrix Market input is read here on
 * the host (the file format is text); the CSR it yields is what the GPU
 * solver consumes.
 */
#define _POSIX_C_SOURCE 200809L
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include "../../include/matrix.h"
#include "../../include/vector.h"

/* ---- src/vector.c:3-27 ------------------------------------------------------ */
void my_daxpy(int n, double alpha, const double *x, double *y)
{
    for (int i = 0; i < n; ++i) y[i] += alpha * x[i];
}

double my_ddot(int n, const double *x, const double *y)
{
    double acc = 0.0;
    for (int i = 0; i < n; ++i) acc += x[i] * y[i];
    return acc;
}

void my_dscal(int n, double alpha, double *x)
{
    for (int i = 0; i < n; ++i) x[i] *= alpha;
}

void my_dcopy(int n, const double *x, double *y)
{
    for (int i = 0; i < n; ++i) y[i] = x[i];
}

/* ---- src/matrix.c:9-24, 189-204 --------------------------------------------- */
void coo_init_matrix(COO_Matrix *m)
{
    m->val = NULL; m->row = NULL; m->col = NULL;
    m->nz = m->rows = m->cols = 0;
}

void coo_free_matrix(COO_Matrix *m)
{
    free(m->val); free(m->row); free(m->col);
    coo_init_matrix(m);
}

void csr_init_matrix(CSR_Matrix *m)
{
    m->val = NULL; m->col = NULL; m->ptr = NULL;
    m->nz = m->rows = m->cols = 0;
}

void csr_free_matrix(CSR_Matrix *m)
{
    free(m->val); free(m->col); free(m->ptr);
    csr_init_matrix(m);
}

/* ---- Matrix Market coordinate reader ---------------------------------------
 * Banner "%%MatrixMarket matrix coordinate <real|integer|pattern> <symmetry>",
 * '%' comment lines, a "rows cols nz" line, then nz entries "i j [v]" with
 * 1-based indices.  The type code follows src/mmio.h: [0] 'M', [1] 'C',
 * [2] 'R'/'I'/'P', [3] 'G'/'S'/'H'/'K'.  As in the reference
 * (src/matrix.c:93) a symmetric file is reported, not expanded. */
static int read_banner(FILE *f, MM_typecode code)
{
    char line[1100], w[5][64];
    if (!fgets(line, sizeof line, f)) return -1;
    if (sscanf(line, "%63s %63s %63s %63s %63s", w[0], w[1], w[2], w[3], w[4]) != 5) return -1;
    if (strcmp(w[0], "%%MatrixMarket") != 0) return -1;
    if (strcasecmp(w[1], "matrix") != 0 || strcasecmp(w[2], "coordinate") != 0) return -1;
    code[0] = 'M';
    code[1] = 'C';
    if (!strcasecmp(w[3], "real")) code[2] = 'R';
    else if (!strcasecmp(w[3], "integer")) code[2] = 'I';
    else if (!strcasecmp(w[3], "pattern")) code[2] = 'P';
    else return -1;
    if (!strcasecmp(w[4], "general")) code[3] = 'G';
    else if (!strcasecmp(w[4], "symmetric")) code[3] = 'S';
    else if (!strcasecmp(w[4], "hermitian")) code[3] = 'H';
    else if (!strcasecmp(w[4], "skew-symmetric")) code[3] = 'K';
    else return -1;
    return 0;
}

static int read_size(FILE *f, int *m, int *n, int *nz)
{
    char line[1100];
    do {
        if (!fgets(line, sizeof line, f)) return -1;
    } while (line[0] == '%');
    return sscanf(line, "%d %d %d", m, n, nz) == 3 ? 0 : -1;
}

static int load_coo(const char *filename, COO_Matrix *coo, MM_typecode code)
{
    FILE *f = fopen(filename, "r");
    if (!f) {
        fprintf(stderr, "ERROR: can't open file \"%s\"\n", filename);
        return -1;
    }
    int m, n, nz;
    if (read_banner(f, code) != 0 || read_size(f, &m, &n, &nz) != 0 || m < 0 || n < 0 || nz < 0) {
        fprintf(stderr, "ERROR: Could not process Matrix Market header.\n");
        fclose(f);
        return -1;
    }
    coo->rows = (unsigned)m;
    coo->cols = (unsigned)n;
    coo->nz = (unsigned)nz;
    coo->val = (double *)malloc(((size_t)nz + 1) * sizeof(double));
    coo->row = (unsigned *)malloc(((size_t)nz + 1) * sizeof(unsigned));
    coo->col = (unsigned *)malloc(((size_t)nz + 1) * sizeof(unsigned));
    if (!coo->val || !coo->row || !coo->col) { fclose(f); coo_free_matrix(coo); return -1; }
    for (int e = 0; e < nz; ++e) {
        long i, j;
        double v = 1.0;
        int got;
        if (code[2] == 'P') got = fscanf(f, "%ld %ld", &i, &j) == 2;
        else if (code[2] == 'I') { long iv; got = fscanf(f, "%ld %ld %ld", &i, &j, &iv) == 3; v = (double)iv; }
        else got = fscanf(f, "%ld %ld %lg", &i, &j, &v) == 3;
        if (!got || i < 1 || j < 1 || i > m || j > n) {
            fprintf(stderr, "ERROR: reading matrix data.\n");
            fclose(f);
            coo_free_matrix(coo);
            return -1;
        }
        coo->row[e] = (unsigned)(i - 1);
        coo->col[e] = (unsigned)(j - 1);
        coo->val[e] = v;
    }
    fclose(f);
    return 0;
}

/* The same reader for the device loader (hbicg_csr_load_device); code gets
 * the banner's MM_typecode. */
int hb_mm_read_coo(const char *filename, COO_Matrix *coo, char *code)
{
    return load_coo(filename, coo, code);
}

/* Replaces src/matrix.c:26-94. */
int coo_load_matrix(char *filename, COO_Matrix *coo)
{
    MM_typecode code;
    if (load_coo(filename, coo, code) != 0) return -1;
    return code[3] == 'S';
}

/* Replaces src/matrix.c:206-232: entries grouped by row keeping file order
 * inside a row (the reference's merge sort is stable), by a counting sort. */
void coo2csr(COO_Matrix *in, CSR_Matrix *out)
{
    const unsigned n = in->rows, nz = in->nz;
    out->rows = n;
    out->cols = in->cols;
    out->nz = nz;
    out->ptr = (unsigned *)calloc((size_t)n + 1, sizeof(unsigned));
    out->col = (unsigned *)malloc(((size_t)nz + 1) * sizeof(unsigned));
    out->val = (double *)malloc(((size_t)nz + 1) * sizeof(double));
    if (!out->ptr || !out->col || !out->val) { csr_free_matrix(out); return; }
    for (unsigned e = 0; e < nz; ++e) out->ptr[in->row[e] + 1]++;
    for (unsigned i = 0; i < n; ++i) out->ptr[i + 1] += out->ptr[i];
    unsigned *fill = (unsigned *)malloc(((size_t)n + 1) * sizeof(unsigned));
    if (!fill) { csr_free_matrix(out); return; }
    memcpy(fill, out->ptr, ((size_t)n + 1) * sizeof(unsigned));
    for (unsigned e = 0; e < nz; ++e) {
        const unsigned d = fill[in->row[e]]++;
        out->col[d] = in->col[e];
        out->val[d] = in->val[e];
    }
    free(fill);
}

/* Replaces src/matrix.c:234-242. */
int csr_load_matrix(char *filename, CSR_Matrix *m)
{
    COO_Matrix t;
    coo_init_matrix(&t);
    const int sym = coo_load_matrix(filename, &t);
    if (sym < 0) return -1;
    coo2csr(&t, m);
    coo_free_matrix(&t);
    return sym;
}

/* Single-GPU replacement of src/matrix.c:402-419 (one "rank" owns every
 * row, so every entry lands in the diagonal block). */
void MPI_csr_load_matrix_block(char *filename, CSR_Matrix *diag, CSR_Matrix *offd, INFO_Matrix *info)
{
    COO_Matrix t;
    MM_typecode code;
    coo_init_matrix(&t);
    csr_init_matrix(diag);
    csr_init_matrix(offd);
    if (load_coo(filename, &t, code) != 0) {
        fprintf(stderr, "ERROR: Could not load \"%s\".\n", filename);
        exit(EXIT_FAILURE);
    }
    info->nz = t.nz;
    info->rows = t.rows;
    info->cols = t.cols;
    memcpy(info->code, code, sizeof(MM_typecode));
    if (info->recvcounts) info->recvcounts[0] = (int)t.rows;
    if (info->displs) info->displs[0] = 0;
    coo2csr(&t, diag);
    offd->rows = t.rows;
    offd->cols = t.cols;
    offd->ptr = (unsigned *)calloc((size_t)t.rows + 1, sizeof(unsigned));
    coo_free_matrix(&t);
}

/* Replaces src/matrix.c:498-516. */
void mult(CSR_Matrix *A_loc, double *x, double *y_loc)
{
    for (unsigned i = 0; i < A_loc->rows; ++i) {
        double t = 0.0;
        for (unsigned e = A_loc->ptr[i]; e < A_loc->ptr[i + 1]; ++e) t += A_loc->val[e] * x[A_loc->col[e]];
        y_loc[i] += t;
    }
}
