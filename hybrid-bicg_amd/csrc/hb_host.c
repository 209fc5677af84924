/*
 * hb_host.c -- C host side of the MI355X shifted BiCG drop-in: argument
 * checks of the reference entry points, device plans (uploads, SpMV row
 * binning, workspace), the iteration driver (direct launches or hipGraph
 * chunks, one host check of a device "done" flag per chunk) and the report.
 *
 * The reference's shifted_lopbicg_* (src/solver_dynamic.c:3-463,
 * src/solver_static.c:3-468, src/solver_normal.c:3-432) malloc their
 * workspace per call and free it on return; so does the drop-in here.
 */
#define _POSIX_C_SOURCE 199309L
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <limits.h>
#include <string.h>
#include <time.h>

#include "hb_adjoint.h"
#include "hb_launch.h"
#include "hb_shared.h"
#include "../../include/hbicg.h"

/* ---- parameters and errors ----------------------------------------------- */
static hbicg_params g_params;
static int g_params_set = 0;
static int g_errno = 0;

void hbicg_default_params(hbicg_params *p)
{
    memset(p, 0, sizeof(*p));
    p->tol = EPS;
    p->max_iter = MAX_ITER;
    p->window = 0;
    p->exact_order = 0;
    p->graph = -1;
    p->chunk = 16;
    p->device = 0;
    p->exit_on_error = 1;
    p->upload_x = 1;
    p->profile = 0;
    p->growth_limit = 1e8;
    p->trace = 0;
    p->value_index = 1;
    p->q_mode = 0;
    p->fused = 1;
    p->panel_cols = 0;
}

static const hbicg_params *params(void)
{
    if (!g_params_set) {
        hbicg_default_params(&g_params);
        g_params_set = 1;
    }
    return &g_params;
}

int hbicg_set_params(const hbicg_params *p)
{
    if (!p || !(p->tol > 0.0) || p->max_iter < 1 || p->window < 0 || p->window > HB_WMAX ||
        p->chunk < 0 || p->q_mode < 0 || p->q_mode > 2)
        return HBICG_E_ARG;
    g_params = *p;
    g_params_set = 1;
    return 0;
}

int hbicg_get_params(hbicg_params *p)
{
    if (!p) return HBICG_E_ARG;
    *p = *params();
    return 0;
}

int hbicg_errno(void) { return g_errno; }

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

#define HIPCK(call)                                                                     \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "hbicg: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                \
            rc = HBICG_E_HIP;                                                           \
            goto fail;                                                                  \
        }                                                                               \
    } while (0)

/* ---- plan ------------------------------------------------------------------ */
struct hbicg_plan {
    int n, M, W, cap, exact, graph, chunk, upload_x, x_zero, device, profile, vec2, want_trace;
    double tol, growth_limit;
    unsigned nz;
    hipStream_t stream;
    int own_stream;
    /* matrix */
    double *val;
    unsigned *col, *ptr, *blk;
    short *col16;
    unsigned *cflag;
    double col_bytes;    /* column-index bytes per SpMV (16-bit tiles count 2 per entry) */
    double val_bytes;    /* value bytes per SpMV (8 per entry, 1 through a value table) */
    double ptr_bytes;    /* row-pointer bytes per SpMV (diagonal tiles read none)    */
    unsigned char *vidx; /* value table of the SpMV, when the matrix has one, with */
    double *vdict;       /* col16 in the same tile-padded layout (col16p, tpad) */
    short *col16p;
    unsigned *tpad;
    unsigned *dtile;     /* diagonal tiles of the value-table SpMV: records, offsets, */
    int *doff;           /* index bytes                                               */
    unsigned char *didx;
    double *dval;        /* constant diagonal tiles: their diagonals' values        */
    int n_diag_tiles, n_const_tiles;
    double *pval;        /* column panels: the entries panel by panel, each      */
    unsigned *pcol, *pptr, *pblk; /* panel's row starts and tiles                */
    int own_matrix, nblk;
    int maxparts;        /* partial-sum pairs per phase the part buffer holds */
    /* workspace */
    double *rhat, *rring, *qring, *s, *y, *p_set, *x_host_buf, *b_buf, *sigma;
    double *aa, *ba, *wa, *pi, *eta, *zeta, *coef, *ccoef, *grow, *part, *trace, *p_snap;
    int *stop, *stop_iter, *act_len, *flist, *fpos, *switch_k, *switch_to;
    hb_state *st;
    int *flag_host;      /* pinned: {done, k} */
    hipEvent_t ev[2], t0, t1;
    /* graph of one chunk, captured for a given x_set pointer */
    hipGraphExec_t gexec;
    hipGraph_t graph_obj;
    double *graph_x;
    int graph_vec2;
    hb_dev dev;
    /* per-kernel timing (params.profile) */
    hipEvent_t *pev;          /* [2][chunk][HB_NSTAGE + 3] */
    double *ptimes;           /* [iterations launched][HB_NSTAGE + 1] */
    int ptimes_cap;
    hbicg_kernel_stat kstat[HB_NSTAGE + 1];
    int *flush_log;           /* device [cap], params.profile */
    int *flush_host;          /* host copy */
    /* last report */
    hbicg_report rep;
    int *h_stop_iter, *h_switch_k, *h_switch_to;
};

static int build_blocks(const unsigned *ptr, int n, unsigned **out, int *nblk)
{
    unsigned *blk = (unsigned *)malloc(((size_t)n + 1) * sizeof(unsigned));
    if (!blk) return HBICG_E_NOMEM;
    int b = 0, r = 0;
    while (r < n) {
        blk[b++] = (unsigned)r;
        const unsigned len = ptr[r + 1] - ptr[r];
        if (len > HB_NNZ_CAP) { ++r; continue; }
        unsigned nnz = 0;
        const int start = r;
        while (r < n && r - start < HB_ROWS_CAP) {
            const unsigned l = ptr[r + 1] - ptr[r];
            if (nnz + l > HB_NNZ_CAP) break;
            nnz += l;
            ++r;
        }
    }
    blk[b] = (unsigned)n;
    *out = blk;
    *nblk = b;
    return 0;
}

static void *dalloc(size_t bytes, int *rc)
{
    void *p = NULL;
    if (bytes == 0) bytes = 8;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        *rc = HBICG_E_NOMEM;
        return NULL;
    }
    return p;
}

void hbicg_plan_destroy(hbicg_plan *pl)
{
    if (!pl) return;
    hipSetDevice(pl->device);
    if (pl->stream) hipStreamSynchronize(pl->stream);
    if (pl->gexec) hipGraphExecDestroy(pl->gexec);
    if (pl->graph_obj) hipGraphDestroy(pl->graph_obj);
    if (pl->own_matrix) { hipFree(pl->val); hipFree(pl->col); hipFree(pl->ptr); }
    void *bufs[] = {pl->blk, pl->col16, pl->cflag, pl->rhat, pl->rring, pl->qring, pl->s, pl->y, pl->p_set, pl->x_host_buf,
                    pl->b_buf, pl->sigma, pl->aa, pl->ba, pl->wa, pl->pi, pl->eta, pl->zeta,
                    pl->coef, pl->ccoef, pl->grow, pl->part, pl->trace, pl->p_snap, pl->vidx, pl->vdict, pl->col16p, pl->tpad, pl->dtile, pl->doff, pl->didx, pl->dval, pl->pval, pl->pcol, pl->pptr, pl->pblk, pl->flush_log, pl->stop, pl->stop_iter, pl->act_len, pl->flist, pl->fpos,
                    pl->switch_k, pl->switch_to, pl->st};
    for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); ++i)
        if (bufs[i]) hipFree(bufs[i]);
    if (pl->flag_host) hipHostFree(pl->flag_host);
    for (int i = 0; i < 2; ++i) if (pl->ev[i]) hipEventDestroy(pl->ev[i]);
    if (pl->t0) hipEventDestroy(pl->t0);
    if (pl->t1) hipEventDestroy(pl->t1);
    if (pl->pev) {
        for (int i = 0; i < 2 * pl->chunk * (HB_NSTAGE + 3); ++i)
            if (pl->pev[i]) hipEventDestroy(pl->pev[i]);
        free(pl->pev);
    }
    free(pl->ptimes);
    free(pl->flush_host);
    if (pl->own_stream && pl->stream) hipStreamDestroy(pl->stream);
    free(pl->h_stop_iter); free(pl->h_switch_k); free(pl->h_switch_to);
    free(pl);
}

/* Window of a plan: the caller's, else 1 for the normal variant (variant 0,
 * per-iteration updates) and the largest the arithmetic mode supports for
 * the others.  Exact-order results do not depend on the window, so an exact
 * plan clamps it to the exact replay's limit. */
/* Shift-update window of a plan.  Exact order: at most HB_WMAX_EXACT.  A
 * window the caller sets: as given, at most cap.  Default: normal 1;
 * static/dynamic the smaller of HB_WAUTO and max_iter (a solve that ends
 * inside its first window flushes once, at the end), halved while its ring
 * buffers (2W+1 n-vectors) would take more than a quarter of the free
 * device memory (at least 16). */
static int window_for(const hbicg_params *p, int variant, int cap, size_t n, size_t free_bytes)
{
    const int wmax = p->exact_order ? HB_WMAX_EXACT : cap;
    if (p->window > 0) return p->window < wmax ? p->window : wmax;
    if (variant == 0) return 1;
    if (p->exact_order) return wmax;
    int w = HB_WAUTO < wmax ? HB_WAUTO : wmax;
    if (p->max_iter < w) w = p->max_iter;
    while (w > 16 && (2.0 * w + 1.0) * (double)n * 8.0 > 0.25 * (double)free_bytes) w /= 2;
    return w;
}

/* Value table of a matrix with at most HB_VDICT distinct values (compared
 * as bit patterns, so -0.0 and 0.0 stay apart): dict gets them in order of
 * first appearance, idx[e] the table position of val[e].  Returns the table
 * size, or 0 when the matrix has more distinct values. */
static int build_value_table(const double *val, size_t nz, double *dict, unsigned char *idx)
{
    enum { SLOTS = 1024 };   /* open addressing, at most a quarter full */
    uint64_t key[SLOTS];
    short pos[SLOTS];
    for (int h = 0; h < SLOTS; ++h) pos[h] = -1;
    int nd = 0;
    for (size_t e = 0; e < nz; ++e) {
        uint64_t b;
        memcpy(&b, &val[e], sizeof b);
        unsigned h = (unsigned)((b * 0x9E3779B97F4A7C15ull) >> 54);
        while (pos[h] >= 0 && key[h] != b) h = (h + 1) & (SLOTS - 1);
        if (pos[h] < 0) {
            if (nd == HB_VDICT) return 0;
            key[h] = b;
            pos[h] = (short)nd;
            dict[nd++] = val[e];
        }
        idx[e] = (unsigned char)pos[h];
    }
    return nd;
}

/* Diagonals of tile [r0, r1): a tile of more than HB_BLOCK/2 rows whose rows
 * list their columns in strictly ascending order and whose entries lie on at
 * most HB_DIAG_K diagonals (column - row) is stored by diagonals.  off gets
 * the diagonals in ascending order; returns their number, or 0 where the
 * tile does not qualify.  A row's entries then come in the order of the
 * diagonals, which is the row's CSR order. */
static int tile_diagonals(const unsigned *ptr, const unsigned *col, unsigned r0, unsigned r1, int *off)
{
    int k = 0;
    if (r1 - r0 <= HB_BLOCK / 2 || ptr[r1] == ptr[r0]) return 0;
    for (unsigned i = r0; i < r1; ++i) {
        for (unsigned e = ptr[i]; e < ptr[i + 1]; ++e) {
            if (e > ptr[i] && col[e] <= col[e - 1]) return 0;
            const long long o = (long long)col[e] - (long long)i;
            int s = 0;
            while (s < k && off[s] < o) ++s;
            if (s < k && off[s] == o) continue;
            if (k == HB_DIAG_K) return 0;
            for (int t = k; t > s; --t) off[t] = off[t - 1];
            off[s] = (int)o;
            ++k;
        }
    }
    return k;
}

/* A diagonal tile whose every diagonal holds one value (the constant-
 * coefficient stencils): returns 1 and that value's table index per
 * diagonal in cidx (0 for a diagonal no row has an entry on). */
static int tile_constant(const unsigned *ptr, const unsigned *col, const unsigned char *idx, unsigned r0,
                         unsigned r1, const int *off, int k, unsigned char *cidx)
{
    int seen[HB_DIAG_K] = {0};
    for (int s = 0; s < k; ++s) cidx[s] = 0;
    for (unsigned i = r0; i < r1; ++i) {
        int s = 0;
        for (unsigned e = ptr[i]; e < ptr[i + 1]; ++e, ++s) {
            const long long o = (long long)col[e] - (long long)i;
            while (off[s] != o) ++s;
            if (!seen[s]) {
                seen[s] = 1;
                cidx[s] = idx[e];
            } else if (cidx[s] != idx[e]) {
                return 0;
            }
        }
    }
    return 1;
}

/* Matrix layouts of the value-table SpMV.  A matrix with at most HB_VDICT
 * distinct values (device matrices: a prefix is checked first) is read
 * through a table of them, one index byte per entry.  Tiles that qualify
 * (tile_diagonals, and at most HB_VDICT - 1 values, HB_DIAG_NONE marking a
 * row's missing diagonal) are stored by diagonals: row i of tile t has its
 * KP = 8 or 16 index bytes at dtile[2t] + KP*i, K = dtile[2t+1] > 0 diagonals
 * at doff[HB_DIAG_K*t ...]; the kernel reads no columns and no row pointers
 * for them, unless KP*rows exceeds the 3 bytes per entry of the CSR form.
 * When every diagonal tile holds one value per diagonal (tile_constant: the
 * constant-coefficient stencils), they store, instead of index bytes, two
 * bytes of presence bits per row (bit s: the row has an entry on diagonal
 * s) at dtile[2t] + 2i, and their diagonals' values at dval[HB_DIAG_K*t ...];
 * dev.cdia is then set and the SpMVs read that form.
 * The other tiles keep CSR order, their index bytes and 16-bit columns in a
 * tile-padded copy: tile t at tpad[t], a multiple of 8.  Leaves the plan
 * unchanged where the matrix has more values. */
static int build_value_layout(hbicg_plan *pl, const hbicg_csr *A, int a_on_device, const unsigned *hptr,
                              const unsigned *hblk, int allow_const)
{
    const size_t nzs = pl->nz;
    const int nb = pl->nblk;
    hb_dev *d = &pl->dev;
    double dict[HB_VDICT];
    int rc = 0, nd = 0;
    double *tmp = NULL;
    unsigned *tcol = NULL, *htp = NULL, *hdt = NULL, *hf = NULL;
    int *hoff = NULL;
    double *hdv = NULL;
    unsigned char *hcidx = NULL;
    int *hk = NULL;
    unsigned char *hidx = (unsigned char *)malloc(nzs), *hvp = NULL, *hdi = NULL;
    short *hc16 = NULL, *hcp = NULL;
    if (!hidx) return HBICG_E_NOMEM;
    if (a_on_device) {
        const size_t pre = nzs < 65536 ? nzs : 65536;
        tmp = (double *)malloc(nzs * sizeof(double));
        if (!tmp) { rc = HBICG_E_NOMEM; goto out; }
        if (hipMemcpy(tmp, A->val, pre * 8, hipMemcpyDeviceToHost) != hipSuccess) { rc = HBICG_E_HIP; goto out; }
        if (build_value_table(tmp, pre, dict, hidx) > 0) {
            if (pre < nzs && hipMemcpy(tmp + pre, A->val + pre, (nzs - pre) * 8, hipMemcpyDeviceToHost) != hipSuccess) {
                rc = HBICG_E_HIP;
                goto out;
            }
            nd = build_value_table(tmp, nzs, dict, hidx);
        }
    } else {
        nd = build_value_table(A->val, nzs, dict, hidx);
    }
    if (nd == 0) goto out;

    const unsigned *hcol = A->col;
    if (a_on_device) {
        tcol = (unsigned *)malloc(nzs * 4);
        if (!tcol) { rc = HBICG_E_NOMEM; goto out; }
        if (hipMemcpy(tcol, A->col, nzs * 4, hipMemcpyDeviceToHost) != hipSuccess) { rc = HBICG_E_HIP; goto out; }
        hcol = tcol;
    }
    htp = (unsigned *)malloc(((size_t)nb + 1) * 4);
    hdt = (unsigned *)calloc((size_t)nb * 2, 4);
    hoff = (int *)calloc((size_t)nb * HB_DIAG_K, sizeof(int));
    hf = (unsigned *)malloc((size_t)nb * 4);
    hc16 = (short *)malloc(nzs * 2);
    hdv = (double *)calloc((size_t)nb * HB_DIAG_K, sizeof(double));
    hcidx = (unsigned char *)calloc((size_t)nb * HB_DIAG_K, 1);
    hk = (int *)calloc((size_t)nb, sizeof(int));
    if (!htp || !hdt || !hoff || !hf || !hc16 || !hdv || !hcidx || !hk) { rc = HBICG_E_NOMEM; goto out; }
    if (hipMemcpy(hc16, pl->col16, nzs * 2, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hf, pl->cflag, (size_t)nb * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        rc = HBICG_E_HIP;
        goto out;
    }
    size_t P = 0, D = 0;
    double vb = 0.0, cb = 0.0, pb = 0.0;
    int ndiag = 0;
    /* which tiles go by diagonals (their index bytes would not exceed the
     * CSR form's 3 bytes per entry), and whether every one of them holds a
     * single value per diagonal: then all of them store presence bits */
    int all_const = allow_const;
    for (int t = 0; t < nb; ++t) {
        const unsigned r0 = hblk[t], r1 = hblk[t + 1];
        const size_t R = r1 - r0, N = hptr[r1] - hptr[r0];
        /* row + diagonal of a missing entry stays within HB_ROWS_CAP of
         * [0, n): the kernel forms it in int */
        const int dia_ok = nd < HB_VDICT && (long long)pl->n + HB_ROWS_CAP < (long long)INT_MAX;
        int k = dia_ok ? tile_diagonals(hptr, hcol, r0, r1, hoff + (size_t)HB_DIAG_K * t) : 0;
        if (k > 0 && (size_t)(k <= 8 ? 8 : 16) * R > 3 * N) k = 0;
        hk[t] = k;
        if (k > 0) {
            ++ndiag;
            if (all_const && !tile_constant(hptr, hcol, hidx, r0, r1, hoff + (size_t)HB_DIAG_K * t, k,
                                            hcidx + (size_t)HB_DIAG_K * t))
                all_const = 0;
        }
    }
    if (!ndiag) all_const = 0;
    for (int t = 0; t < nb; ++t) {
        const unsigned r0 = hblk[t], r1 = hblk[t + 1];
        const size_t R = r1 - r0, N = hptr[r1] - hptr[r0];
        const int k = hk[t];
        /* bytes per row: index bytes, or two bytes of presence bits */
        const size_t KP = all_const ? 2 : (k <= 8 ? 8 : 16);
        htp[t] = (unsigned)P;
        if (k > 0) {
            D = (D + 15) & ~(size_t)15;
            hdt[2 * t] = (unsigned)D;
            hdt[2 * t + 1] = (unsigned)k;
            D += KP * R;
            vb += (double)(KP * R) + 4.0 * k + 8.0;
            if (all_const) {
                for (int s2 = 0; s2 < k; ++s2)
                    hdv[(size_t)HB_DIAG_K * t + s2] = dict[hcidx[(size_t)HB_DIAG_K * t + s2]];
                vb += 8.0 * k;
            }
            pb += 8.0;
        } else {
            memset(hoff + (size_t)HB_DIAG_K * t, 0, HB_DIAG_K * sizeof(int));
            P += (N + 7) & ~(size_t)7;
            vb += (double)N;
            cb += (hf[t] ? 2.0 : 4.0) * (double)N;
            pb += 4.0 * ((double)R + 1.0);
        }
        if (P > 0xffffffffu || D > 0xffffffffu) { rc = HBICG_E_ARG; goto out; }
    }
    htp[nb] = (unsigned)P;
    hvp = (unsigned char *)calloc(P ? P : 1, 1);
    hcp = (short *)calloc(P ? P : 1, 2);
    hdi = (unsigned char *)malloc(D ? D : 1);
    if (!hvp || !hcp || !hdi) { rc = HBICG_E_NOMEM; goto out; }
    memset(hdi, HB_DIAG_NONE, D ? D : 1);
    for (int t = 0; t < nb; ++t) {
        const unsigned r0 = hblk[t], r1 = hblk[t + 1];
        if (hdt[2 * t + 1] && all_const) {
            const int *off = hoff + (size_t)HB_DIAG_K * t;
            for (unsigned i = r0; i < r1; ++i) {
                unsigned m = 0;
                int s = 0;
                for (unsigned e = hptr[i]; e < hptr[i + 1]; ++e, ++s) {
                    const long long o = (long long)hcol[e] - (long long)i;
                    while (off[s] != o) ++s;
                    m |= 1u << s;
                }
                unsigned char *dst = hdi + hdt[2 * t] + 2 * (size_t)(i - r0);
                dst[0] = (unsigned char)(m & 255u);
                dst[1] = (unsigned char)(m >> 8);
            }
        } else if (hdt[2 * t + 1]) {
            const int *off = hoff + (size_t)HB_DIAG_K * t;
            const size_t KP = hdt[2 * t + 1] <= 8 ? 8 : 16;
            for (unsigned i = r0; i < r1; ++i) {
                unsigned char *dst = hdi + hdt[2 * t] + KP * (i - r0);
                int s = 0;
                for (unsigned e = hptr[i]; e < hptr[i + 1]; ++e, ++s) {
                    const long long o = (long long)hcol[e] - (long long)i;
                    while (off[s] != o) ++s;
                    dst[s] = hidx[e];
                }
            }
        } else {
            const size_t a = hptr[r0], len = hptr[r1] - a;
            memcpy(hvp + htp[t], hidx + a, len);
            memcpy(hcp + htp[t], hc16 + a, len * 2);
        }
    }
    pl->vidx = (unsigned char *)dalloc(P, &rc);
    pl->col16p = (short *)dalloc(P * 2, &rc);
    pl->tpad = (unsigned *)dalloc(((size_t)nb + 1) * 4, &rc);
    pl->vdict = (double *)dalloc(HB_VDICT * sizeof(double), &rc);
    if (ndiag) {
        pl->dtile = (unsigned *)dalloc((size_t)nb * 8, &rc);
        pl->doff = (int *)dalloc((size_t)nb * HB_DIAG_K * sizeof(int), &rc);
        pl->didx = (unsigned char *)dalloc(D, &rc);
    }
    if (all_const) pl->dval = (double *)dalloc((size_t)nb * HB_DIAG_K * sizeof(double), &rc);
    if (rc) goto out;
    if (hipMemcpy(pl->vidx, hvp, P, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->col16p, hcp, P * 2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->tpad, htp, ((size_t)nb + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->vdict, dict, (size_t)nd * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        (ndiag && (hipMemcpy(pl->dtile, hdt, (size_t)nb * 8, hipMemcpyHostToDevice) != hipSuccess ||
                   hipMemcpy(pl->doff, hoff, (size_t)nb * HB_DIAG_K * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
                   hipMemcpy(pl->didx, hdi, D, hipMemcpyHostToDevice) != hipSuccess)) ||
        (all_const && hipMemcpy(pl->dval, hdv, (size_t)nb * HB_DIAG_K * sizeof(double), hipMemcpyHostToDevice) !=
                       hipSuccess)) {
        rc = HBICG_E_HIP;
        goto out;
    }
    d->vidx = pl->vidx;
    d->col16p = pl->col16p;
    d->tpad = pl->tpad;
    d->vdict = pl->vdict;
    d->nvdict = nd;
    if (ndiag) {
        d->dtile = pl->dtile;
        d->doff = pl->doff;
        d->didx = pl->didx;
        d->dval = pl->dval;
        d->cdia = all_const;
    }
    pl->n_diag_tiles = ndiag;
    pl->n_const_tiles = all_const ? ndiag : 0;
    pl->val_bytes = vb;
    pl->col_bytes = cb + 4.0 * (double)nb;
    pl->ptr_bytes = pb;
out:
    free(tmp);
    free(tcol);
    free(htp);
    free(hdt);
    free(hf);
    free(hoff);
    free(hdv);
    free(hcidx);
    free(hk);
    free(hidx);
    free(hvp);
    free(hdi);
    free(hc16);
    free(hcp);
    return rc;
}

/* Column panels of `width` columns (params.panel_cols): the entries of
 * every row split by column into panels, stored panel by panel (each
 * panel's rows in order, a row's entries in CSR order) with one row-pointer
 * array per panel.  Needs every row's columns strictly ascending, so that
 * the panels take a row's entries in CSR order; returns 1 without building
 * anything when one is not, or when there would be a single panel. */
static int build_panels(hbicg_plan *pl, const hbicg_csr *A, int a_on_device, const unsigned *hptr, int width)
{
    const int n = pl->n;
    const size_t nz = pl->nz;
    if ((long long)width * HB_MAXPANEL < (long long)n) width = (int)(((long long)n + HB_MAXPANEL - 1) / HB_MAXPANEL);
    const int P = (int)(((long long)n + width - 1) / width);
    if (P < 2 || nz == 0) return 1;
    int rc = 0, grow = 0;
    unsigned *qb[HB_MAXPANEL] = {0}, *hbt = NULL;
    const unsigned *hc = (const unsigned *)A->col;
    const double *hv = (const double *)A->val;
    unsigned *tc = NULL, *hpp = NULL, *pc = NULL;
    double *tv = NULL, *pv = NULL;
    size_t *cur = NULL;
    if (a_on_device) {
        tc = (unsigned *)malloc(nz * 4);
        tv = (double *)malloc(nz * 8);
        if (!tc || !tv) { rc = HBICG_E_NOMEM; goto done; }
        if (hipMemcpy(tc, A->col, nz * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(tv, A->val, nz * 8, hipMemcpyDeviceToHost) != hipSuccess) { rc = HBICG_E_HIP; goto done; }
        hc = tc;
        hv = tv;
    }
    for (int i = 0; i < n; ++i)
        for (unsigned e = hptr[i] + 1; e < hptr[i + 1]; ++e)
            if (hc[e] <= hc[e - 1]) { rc = 1; goto done; }
    hpp = (unsigned *)malloc((size_t)P * ((size_t)n + 1) * 4);
    pc = (unsigned *)malloc(nz * 4);
    pv = (double *)malloc(nz * 8);
    cur = (size_t *)calloc((size_t)P + 1, sizeof(size_t));
    if (!hpp || !pc || !pv || !cur) { rc = HBICG_E_NOMEM; goto done; }
    for (size_t e = 0; e < nz; ++e) {
        if (hc[e] >= (unsigned)n) { rc = HBICG_E_ARG; goto done; }
        cur[hc[e] / (unsigned)width + 1]++;
    }
    for (int q = 0; q < P; ++q) cur[q + 1] += cur[q];   /* cur[q]: start of panel q */
    for (int i = 0; i < n; ++i) {
        for (int q = 0; q < P; ++q) hpp[(size_t)q * ((size_t)n + 1) + i] = (unsigned)cur[q];
        for (unsigned e = hptr[i]; e < hptr[i + 1]; ++e) {
            const size_t at = cur[hc[e] / (unsigned)width]++;
            pc[at] = hc[e];
            pv[at] = hv[e];
        }
    }
    for (int q = 0; q < P; ++q) hpp[(size_t)q * ((size_t)n + 1) + n] = (unsigned)cur[q];
    /* each panel's tiles, binned on its own entries */
    {
        size_t tot = 0;
        for (int q = 0; q < P; ++q) {
            if ((rc = build_blocks(hpp + (size_t)q * ((size_t)n + 1), n, &qb[q], &pl->dev.pnblk[q])) != 0)
                goto done;
            pl->dev.pboff[q] = (unsigned)tot;
            tot += (size_t)pl->dev.pnblk[q] + 1;
            if (pl->dev.pnblk[q] > pl->maxparts && pl->dev.pnblk[q] > grow) grow = pl->dev.pnblk[q];
        }
        hbt = (unsigned *)malloc(tot * 4);
        if (!hbt) { rc = HBICG_E_NOMEM; goto done; }
        for (int q = 0; q < P; ++q)
            memcpy(hbt + pl->dev.pboff[q], qb[q], ((size_t)pl->dev.pnblk[q] + 1) * 4);
        pl->pblk = (unsigned *)dalloc(tot * 4, &rc);
        if (rc) goto done;
        if (hipMemcpy(pl->pblk, hbt, tot * 4, hipMemcpyHostToDevice) != hipSuccess) { rc = HBICG_E_HIP; goto done; }
    }
    if (grow) {
        /* the last pass's tiles write the partial sums: room for the most
         * tiles any pass has, in each phase's slice */
        double *np = (double *)dalloc((size_t)grow * 6 * 8, &rc);
        if (rc) goto done;
        hipFree(pl->part);
        pl->part = np;
        pl->maxparts = grow;
        pl->dev.part = np;
        pl->dev.part2 = np + (size_t)grow * 2;
        pl->dev.part3 = np + (size_t)grow * 4;
    }
    pl->pval = (double *)dalloc(nz * 8, &rc);
    pl->pcol = (unsigned *)dalloc(nz * 4, &rc);
    pl->pptr = (unsigned *)dalloc((size_t)P * ((size_t)n + 1) * 4, &rc);
    if (rc) goto done;
    if (hipMemcpy(pl->pval, pv, nz * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->pcol, pc, nz * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->pptr, hpp, (size_t)P * ((size_t)n + 1) * 4, hipMemcpyHostToDevice) != hipSuccess) {
        rc = HBICG_E_HIP;
        goto done;
    }
    pl->dev.npanel = P;
    pl->dev.pval = pl->pval;
    pl->dev.pcol = pl->pcol;
    pl->dev.pptr = pl->pptr;
    pl->dev.pblk = pl->pblk;
    pl->dev.npart = pl->dev.pnblk[P - 1];
done:
    for (int q = 0; q < P && q < HB_MAXPANEL; ++q) free(qb[q]);
    free(tc); free(tv); free(hpp); free(pc); free(pv); free(cur); free(hbt);
    return rc;
}

/* w_alloc: ring storage for windows up to this (the batched solver's
 * autotuner); w_cap: the largest window the plan may choose; allow_vi: the
 * SpMV may read the matrix through a value table (params.value_index);
 * batched: the plan is one problem of a batch, whose launches run its tiles
 * (no column panels, no fused launches of its own); allow_const: diagonal
 * tiles may take the constant-diagonal form. */
static hbicg_plan *plan_create_v(const hbicg_csr *A, int a_on_device, int M,
                                 const hbicg_params *p, void *stream, int variant, int w_alloc,
                                 int w_cap, int allow_vi, int batched, int allow_const)
{
    int rc = 0;
    unsigned *hptr = NULL, *hblk = NULL;
    hbicg_plan *pl = NULL;
    g_errno = 0;
    if (!A || !A->ptr || !p || M <= 0 || A->n == 0 || p->window < 0 || p->window > HB_WMAX ||
        p->q_mode < 0 || p->q_mode > 2) {
        g_errno = HBICG_E_ARG;
        return NULL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { g_errno = HBICG_E_NODEVICE; return NULL; }
    pl = (hbicg_plan *)calloc(1, sizeof(*pl));
    if (!pl) { g_errno = HBICG_E_NOMEM; return NULL; }
    pl->device = p->device;
    HIPCK(hipSetDevice(p->device));
    const int n = (int)A->n;
    pl->n = n;
    pl->M = M;
    pl->nz = A->nz;
    {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
        pl->W = window_for(p, variant, w_cap, (size_t)n, free_b);
    }
    pl->cap = p->max_iter + 1;
    pl->tol = p->tol;
    pl->growth_limit = p->growth_limit > 0.0 ? p->growth_limit : 0.0;
    pl->want_trace = p->trace ? 1 : 0;
    pl->exact = p->exact_order ? 1 : 0;
    pl->graph = p->graph < 0 ? (variant == 2) : (p->graph ? 1 : 0);
    pl->chunk = p->chunk > 0 ? p->chunk : 16;
    pl->upload_x = p->upload_x;
    pl->x_zero = p->x_zero != 0;
    pl->profile = p->profile ? 1 : 0;
    if (pl->profile) pl->graph = 0;
    if (stream) {
        pl->stream = (hipStream_t)stream;
    } else {
        HIPCK(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking));
        pl->own_stream = 1;
    }

    /* matrix and row binning (needs the row pointers on the host) */
    if (a_on_device) {
        pl->val = (double *)A->val; pl->col = (unsigned *)A->col; pl->ptr = (unsigned *)A->ptr;
        hptr = (unsigned *)malloc(((size_t)n + 1) * sizeof(unsigned));
        if (!hptr) { rc = HBICG_E_NOMEM; goto fail; }
        HIPCK(hipMemcpy(hptr, A->ptr, ((size_t)n + 1) * sizeof(unsigned), hipMemcpyDeviceToHost));
    } else {
        pl->own_matrix = 1;
        hptr = (unsigned *)A->ptr;
        if (hptr[n] != A->nz) { rc = HBICG_E_ARG; goto fail; }
        pl->val = (double *)dalloc((size_t)A->nz * 8, &rc);
        pl->col = (unsigned *)dalloc((size_t)A->nz * 4, &rc);
        pl->ptr = (unsigned *)dalloc(((size_t)n + 1) * 4, &rc);
        if (rc) goto fail;
        HIPCK(hipMemcpy(pl->val, A->val, (size_t)A->nz * 8, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(pl->col, A->col, (size_t)A->nz * 4, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(pl->ptr, A->ptr, ((size_t)n + 1) * 4, hipMemcpyHostToDevice));
    }
    /* validate the row pointers once on the host: every SpMV workgroup
     * indexes val/col inside [0, nz) only through them */
    for (int i = 0; i < n; ++i)
        if (hptr[i + 1] < hptr[i]) { rc = HBICG_E_ARG; goto fail; }
    if (hptr[0] != 0 || hptr[n] != A->nz) { rc = HBICG_E_ARG; goto fail; }
    if ((rc = build_blocks(hptr, n, &hblk, &pl->nblk)) != 0) goto fail;
    pl->blk = (unsigned *)dalloc(((size_t)pl->nblk + 1) * 4, &rc);
    if (rc) goto fail;
    HIPCK(hipMemcpy(pl->blk, hblk, ((size_t)pl->nblk + 1) * 4, hipMemcpyHostToDevice));

    /* workspace */
    const size_t nb = (size_t)n * 8;
    /* ring and coefficient storage for the largest window this plan may run */
    const int W = w_alloc > pl->W ? w_alloc : pl->W;
    int vec_grid = (n + HB_BLOCK - 1) / HB_BLOCK;
    if (vec_grid > 2048) vec_grid = 2048;
    const int maxparts = pl->nblk > vec_grid ? pl->nblk : vec_grid;
    const int vec2 = (n % 2) == 0;
    const int gx = (n + HB_BLOCK - 1) / HB_BLOCK;
    int per = 4096 / gx;
    if (per < 1) per = 1;
    int G = (M + per - 1) / per;
    if (G < 1) G = 1;
    const int gy = (M + G - 1) / G;
    const size_t cap = (size_t)pl->cap;
    pl->rhat = (double *)dalloc(nb, &rc);
    pl->rring = (double *)dalloc(nb * (W + 1), &rc);
    pl->qring = (double *)dalloc(nb * W, &rc);
    pl->s = (double *)dalloc(nb, &rc);
    pl->y = (double *)dalloc(nb, &rc);
    pl->b_buf = (double *)dalloc(nb, &rc);
    pl->p_set = (double *)dalloc(nb * (size_t)M, &rc);
    pl->sigma = (double *)dalloc((size_t)M * 8, &rc);
    pl->aa = (double *)dalloc(cap * 8, &rc);
    pl->ba = (double *)dalloc(cap * 8, &rc);
    pl->wa = (double *)dalloc(cap * 8, &rc);
    pl->pi = (double *)dalloc(cap * (size_t)M * 8, &rc);
    pl->eta = (double *)dalloc((size_t)M * 8, &rc);
    pl->zeta = (double *)dalloc((size_t)M * 8, &rc);
    pl->coef = (double *)dalloc((size_t)W * M * 6 * 8, &rc);
    const size_t ccn = (size_t)((M + HB_UG - 1) / HB_UG) * HB_CC_ROWS(W) * (2 * HB_UG);
    pl->ccoef = (double *)dalloc(ccn * 8, &rc);
    pl->grow = (double *)dalloc((size_t)2 * M * 8, &rc);
    if (pl->want_trace) pl->trace = (double *)dalloc(cap * (size_t)M * 8, &rc);
    if (pl->profile) {
        pl->flush_log = (int *)dalloc(cap * sizeof(int), &rc);
        pl->flush_host = (int *)calloc(cap, sizeof(int));
        if (!pl->flush_host) rc = HBICG_E_NOMEM;
    }
    pl->part = (double *)dalloc((size_t)maxparts * 6 * 8, &rc);   /* three phases' partials */
    pl->p_snap = (double *)dalloc(nb, &rc);
    pl->stop = (int *)dalloc((size_t)M * 4, &rc);
    pl->stop_iter = (int *)dalloc((size_t)M * 4, &rc);
    pl->act_len = (int *)dalloc((size_t)M * 4, &rc);
    pl->flist = (int *)dalloc((size_t)M * 4, &rc);
    pl->fpos = (int *)dalloc((size_t)M * 4, &rc);
    pl->switch_k = (int *)dalloc((size_t)M * 4, &rc);
    pl->switch_to = (int *)dalloc((size_t)M * 4, &rc);
    pl->st = (hb_state *)dalloc(sizeof(hb_state), &rc);
    if (rc) goto fail;
    /* unused combined-coefficient columns stay finite */
    HIPCK(hipMemset(pl->ccoef, 0, ccn * 8));
    HIPCK(hipHostMalloc((void **)&pl->flag_host, 2 * sizeof(int), hipHostMallocDefault));
    HIPCK(hipEventCreateWithFlags(&pl->ev[0], hipEventDisableTiming));
    HIPCK(hipEventCreateWithFlags(&pl->ev[1], hipEventDisableTiming));
    HIPCK(hipEventCreate(&pl->t0));
    HIPCK(hipEventCreate(&pl->t1));
    pl->h_stop_iter = (int *)calloc((size_t)M, sizeof(int));
    pl->h_switch_k = (int *)calloc((size_t)M, sizeof(int));
    pl->h_switch_to = (int *)calloc((size_t)M, sizeof(int));
    if (!pl->h_stop_iter || !pl->h_switch_k || !pl->h_switch_to) { rc = HBICG_E_NOMEM; goto fail; }

    hb_dev *d = &pl->dev;
    memset(d, 0, sizeof(*d));
    d->n = n; d->M = M; d->W = pl->W; d->cap = pl->cap; d->nblk = pl->nblk;
    d->exact = pl->exact;
    d->vec_grid = vec_grid;
    d->shift_gx = gx; d->shift_gy = gy; d->shift_group = G; d->vec2 = vec2;
    d->x_zero = p->x_zero != 0;
    pl->vec2 = vec2;
    d->tol = pl->tol;
    d->growth_limit = pl->growth_limit;
    d->val = pl->val; d->col = pl->col; d->ptr = pl->ptr; d->blk = pl->blk;
    d->sigma = pl->sigma; d->rhat = pl->rhat; d->rring = pl->rring; d->qring = pl->qring;
    d->s = pl->s; d->y = pl->y; d->p_set = pl->p_set; d->x_set = NULL;
    d->aa = pl->aa; d->ba = pl->ba; d->wa = pl->wa; d->pi = pl->pi; d->eta = pl->eta;
    d->zeta = pl->zeta; d->coef = pl->coef; d->ccoef = pl->ccoef; d->grow = pl->grow; d->trace = pl->trace;
    d->flush_log = pl->flush_log;
    d->stop = pl->stop;
    d->stop_iter = pl->stop_iter;
    d->act_len = pl->act_len; d->flist = pl->flist; d->fpos = pl->fpos; d->switch_k = pl->switch_k;
    d->switch_to = pl->switch_to; d->part = pl->part; d->st = pl->st; d->p_snap = pl->p_snap;
    d->part2 = pl->part + (size_t)maxparts * 2;
    d->part3 = pl->part + (size_t)maxparts * 4;
    d->npart = pl->nblk;
    pl->maxparts = maxparts;
    /* batches launch each phase over all their problems' tiles unfused */
    d->fused = !pl->exact && p->fused && !batched;

    int scattered = 0;   /* most entries beyond 16-bit column offsets of their tile */
    /* 16-bit column offsets for the tiles whose entries stay within 32767
     * columns of the tile's first row */
    pl->col16 = (short *)dalloc((size_t)pl->nz * 2, &rc);
    pl->cflag = (unsigned *)dalloc((size_t)pl->nblk * 4, &rc);
    if (rc) goto fail;
    if (hb_k_compress_cols(d, pl->col16, pl->cflag, pl->stream) != 0) { rc = HBICG_E_HIP; goto fail; }
    HIPCK(hipStreamSynchronize(pl->stream));
    d->col16 = pl->col16;
    d->cflag = pl->cflag;
    {
        /* column-index bytes one SpMV reads, for the kernel statistics */
        unsigned *hf = (unsigned *)malloc((size_t)pl->nblk * 4);
        if (!hf) { rc = HBICG_E_NOMEM; goto fail; }
        if (hipMemcpy(hf, pl->cflag, (size_t)pl->nblk * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            free(hf);
            rc = HBICG_E_HIP;
            goto fail;
        }
        double cb = 4.0 * (double)pl->nblk, narrow_nz = 0.0;
        for (int t = 0; t < pl->nblk; ++t) {
            const double len = (double)(hptr[hblk[t + 1]] - hptr[hblk[t]]);
            cb += (hf[t] ? 2.0 : 4.0) * len;
            if (hf[t]) narrow_nz += len;
        }
        pl->col_bytes = cb;
        free(hf);
        /* columns scattered beyond 16-bit offsets for most entries: SpMV 2
         * gathers q formed beforehand rather than r and s at every entry */
        scattered = 2.0 * narrow_nz < (double)pl->nz;
        d->q_unfused = !pl->exact && (p->q_mode == 2 || (p->q_mode == 0 && scattered));
    }

    pl->val_bytes = 8.0 * (double)pl->nz;
    pl->ptr_bytes = 4.0 * ((double)n + 1.0);
    if (allow_vi && pl->nz > 0 && (rc = build_value_layout(pl, A, a_on_device, hptr, hblk, allow_const)) != 0)
        goto fail;
    /* column panels: 8-byte CSR reads whose gathers would otherwise range
     * over an x larger than one XCD's L2 (4 MiB) */
    if (!batched && !d->vidx && p->panel_cols >= 0) {
        const int width = p->panel_cols > 0 ? p->panel_cols : HB_PANEL_COLS;
        if (p->panel_cols > 0 || (scattered && (size_t)n * 8 > (size_t)2 * HB_PANEL_COLS * 8)) {
            rc = build_panels(pl, A, a_on_device, hptr, width);
            if (rc < 0) goto fail;
            rc = 0;
        }
    }

    /* alpha~'s finalize joins SpMV 2 where that launch forms q at its
     * gathered columns in one pass */
    d->fused_a = d->fused && !d->q_unfused && d->npanel <= 1 && (p->fused == 1 || p->fused == 3);
    /* the scalar stage joins the shift kernel (level 1), whose tiles then
     * walk the seed p update and the flush tiles grid-stride */
    d->fused_s = d->fused && p->fused == 1;
    if (d->fused_s) {
        d->shift_tw = hb_k_shift_stage_tiles(d->shift_gx, M, pl->device);
        if (d->shift_tw < 0) { rc = HBICG_E_HIP; goto fail; }
        /* too many shifts for the stage's launch: the stage runs in a
         * launch of its own (the five-launch form, bitwise the same) */
        if (d->shift_tw == 0) d->fused_s = 0;
    }
    if (a_on_device) free(hptr);
    free(hblk);
    return pl;
fail:
    if (a_on_device) free(hptr);
    free(hblk);
    hbicg_plan_destroy(pl);
    g_errno = rc ? rc : HBICG_E_HIP;
    return NULL;
}

hbicg_plan *hbicg_plan_create(const hbicg_csr *A, int a_on_device, int M,
                              const hbicg_params *p, void *stream)
{
    const hbicg_params *pp = p ? p : params();
    return plan_create_v(A, a_on_device, M, pp, stream, 1, 0, HB_WMAX, pp->value_index, 0, 1);
}

/* Enqueues one chunk of iterations followed by the done-flag publication. */
static int enqueue_chunk(hbicg_plan *pl)
{
    for (int c = 0; c < pl->chunk; ++c) {
        int e = hb_k_iteration(&pl->dev, pl->stream);
        if (e) return e;
    }
    return hb_k_publish(&pl->dev, pl->flag_host, pl->stream);
}

static int capture_graph(hbicg_plan *pl)
{
    int rc = 0;
    if (pl->gexec && pl->graph_x == pl->dev.x_set && pl->graph_vec2 == pl->dev.vec2) return 0;
    if (pl->gexec) { hipGraphExecDestroy(pl->gexec); pl->gexec = NULL; }
    if (pl->graph_obj) { hipGraphDestroy(pl->graph_obj); pl->graph_obj = NULL; }
    HIPCK(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal));
    if (enqueue_chunk(pl) != 0) {
        hipGraph_t g;
        hipStreamEndCapture(pl->stream, &g);
        if (g) hipGraphDestroy(g);
        return HBICG_E_HIP;
    }
    HIPCK(hipStreamEndCapture(pl->stream, &pl->graph_obj));
    HIPCK(hipGraphInstantiate(&pl->gexec, pl->graph_obj, NULL, NULL, 0));
    pl->graph_x = pl->dev.x_set;
    pl->graph_vec2 = pl->dev.vec2;
    return 0;
fail:
    return rc;
}

/* Direct launches with an event around every kernel; per-iteration
 * durations land in pl->ptimes, summed once the final k is known.  Every
 * HB_REPLAY_EVERY-th iteration the fused q + SpMV 2 kernel is launched
 * HB_REPLAY more times back to back between two extra events: it rewrites
 * the same q, y and partials from the same inputs, so the replays time the
 * kernel without the event and dispatch gaps that bracket a single launch. */
#define HB_REPLAY       4
#define HB_REPLAY_EVERY 4
#define HB_PT           (HB_NSTAGE + 1)   /* ptimes per iteration: stages + replay */

static int profile_times(hbicg_plan *pl, int c)
{
    int rc = 0;
    const int C = pl->chunk, E = HB_NSTAGE + 3;
    hipEvent_t *pe = pl->pev + (size_t)(c & 1) * C * E;
    for (int it = 0; it < C; ++it) {
        const int rep = (it % HB_REPLAY_EVERY) == 0;
        double *t = pl->ptimes + ((size_t)c * C + it) * HB_PT;
        for (int sg = 0; sg < HB_NSTAGE; ++sg) {
            float ms = 0.0f;
            const hipEvent_t e0 = (rep && sg == HB_STAGE_FIN2) ? pe[it * E + HB_NSTAGE + 2]
                                                               : pe[it * E + sg];
            HIPCK(hipEventElapsedTime(&ms, e0, pe[it * E + sg + 1]));
            t[sg] = ms;
        }
        t[HB_NSTAGE] = 0.0;
        if (rep) {
            float ms = 0.0f;
            HIPCK(hipEventElapsedTime(&ms, pe[it * E + HB_NSTAGE + 1], pe[it * E + HB_NSTAGE + 2]));
            t[HB_NSTAGE] = ms;
        }
    }
fail:
    return rc;
}

static int run_loop_profiled(hbicg_plan *pl)
{
    int rc = 0;
    const int C = pl->chunk, E = HB_NSTAGE + 3;
    const int max_chunks = pl->cap / C + 3;
    if (!pl->pev) {
        pl->pev = (hipEvent_t *)calloc((size_t)2 * C * E, sizeof(hipEvent_t));
        if (!pl->pev) return HBICG_E_NOMEM;
        for (int i = 0; i < 2 * C * E; ++i) HIPCK(hipEventCreate(&pl->pev[i]));
    }
    if (pl->ptimes_cap < (max_chunks + 1) * C) {
        free(pl->ptimes);
        pl->ptimes_cap = (max_chunks + 1) * C;
        pl->ptimes = (double *)calloc((size_t)pl->ptimes_cap * HB_PT, sizeof(double));
        if (!pl->ptimes) { pl->ptimes_cap = 0; return HBICG_E_NOMEM; }
    }
    pl->flag_host[0] = 0;
    int c = 0, stop = 0;
    for (; c <= max_chunks && !stop; ++c) {
        hipEvent_t *ev = pl->pev + (size_t)(c & 1) * C * E;
        for (int it = 0; it < C; ++it) {
            HIPCK(hipEventRecord(ev[it * E], pl->stream));
            for (int sg = 0; sg < HB_NSTAGE; ++sg) {
                if (hb_k_stage(&pl->dev, sg, pl->stream) != 0) return HBICG_E_HIP;
                HIPCK(hipEventRecord(ev[it * E + sg + 1], pl->stream));
                if (sg == HB_STAGE_SPMV2 && (it % HB_REPLAY_EVERY) == 0) {
                    HIPCK(hipEventRecord(ev[it * E + HB_NSTAGE + 1], pl->stream));
                    for (int r = 0; r < HB_REPLAY; ++r)
                        if (hb_k_stage(&pl->dev, sg, pl->stream) != 0) return HBICG_E_HIP;
                    HIPCK(hipEventRecord(ev[it * E + HB_NSTAGE + 2], pl->stream));
                }
            }
        }
        if (hb_k_publish(&pl->dev, pl->flag_host, pl->stream) != 0) return HBICG_E_HIP;
        HIPCK(hipEventRecord(pl->ev[c & 1], pl->stream));
        if (c > 0) {
            HIPCK(hipEventSynchronize(pl->ev[(c - 1) & 1]));
            if ((rc = profile_times(pl, c - 1)) != 0) return rc;
            stop = __atomic_load_n(&pl->flag_host[0], __ATOMIC_ACQUIRE);
        }
    }
    HIPCK(hipStreamSynchronize(pl->stream));
    if ((rc = profile_times(pl, c - 1)) != 0) return rc;
    return 0;
fail:
    return rc;
}

static int run_loop(hbicg_plan *pl)
{
    int rc = 0;
    if (pl->profile) return run_loop_profiled(pl);
    if (pl->graph && (rc = capture_graph(pl)) != 0) return rc;
    pl->flag_host[0] = 0;
    pl->flag_host[1] = 0;
    const int max_chunks = pl->cap / pl->chunk + 3;
    int launched = 0, prev = -1;
    for (;;) {
        if (pl->graph) {
            HIPCK(hipGraphLaunch(pl->gexec, pl->stream));
        } else if (enqueue_chunk(pl) != 0) {
            return HBICG_E_HIP;
        }
        const int cur = launched & 1;
        HIPCK(hipEventRecord(pl->ev[cur], pl->stream));
        ++launched;
        if (prev >= 0) {
            HIPCK(hipEventSynchronize(pl->ev[prev]));
            if (__atomic_load_n(&pl->flag_host[0], __ATOMIC_ACQUIRE)) break;
        }
        prev = cur;
        if (launched > max_chunks) break;   /* every kernel returns at once past the cap */
    }
    HIPCK(hipStreamSynchronize(pl->stream));
    return 0;
fail:
    return rc;
}

int hbicg_plan_solve(hbicg_plan *pl, const double *b, const double *sigma, int seed,
                     double *x_set, double *r_out, int on_device, hbicg_report *rep)
{
    int rc = 0;
    if (!pl || !b || !sigma || !x_set || seed < 0 || seed >= pl->M) return HBICG_E_ARG;
    const double t_start = now_ms();
    HIPCK(hipSetDevice(pl->device));
    const size_t nb = (size_t)pl->n * 8, xb = nb * (size_t)pl->M;
    double *dx = x_set;
    const double *db = b;
    if (!on_device) {
        if (!pl->x_host_buf) {
            pl->x_host_buf = (double *)dalloc(xb, &rc);
            if (rc) return rc;
        }
        dx = pl->x_host_buf;
        if (pl->x_zero) {
            /* nothing to move: the solve writes every x_j before reading it */
        } else if (pl->upload_x) {
            HIPCK(hipMemcpyAsync(dx, x_set, xb, hipMemcpyHostToDevice, pl->stream));
        } else {
            HIPCK(hipMemsetAsync(dx, 0, xb, pl->stream));
        }
        HIPCK(hipMemcpyAsync(pl->b_buf, b, nb, hipMemcpyHostToDevice, pl->stream));
        db = pl->b_buf;
    }
    pl->dev.x_set = dx;
    /* 16-byte accesses in the shift kernel need an aligned x_set */
    pl->dev.vec2 = pl->vec2 && ((((size_t)dx) & 15) == 0);
    HIPCK(hipMemcpyAsync(pl->sigma, sigma, (size_t)pl->M * 8, hipMemcpyHostToDevice, pl->stream));
    HIPCK(hipMemsetAsync(pl->st, 0, sizeof(hb_state), pl->stream));
    HIPCK(hipEventRecord(pl->t0, pl->stream));
    pl->dev.b_in = db;
    pl->dev.seed_in = seed;
    if (pl->trace && hb_k_fill(pl->trace, (size_t)pl->cap * pl->M, -1.0, pl->stream) != 0) {
        rc = HBICG_E_HIP;
        goto fail;
    }
    if (pl->flush_log) HIPCK(hipMemsetAsync(pl->flush_log, 0, (size_t)pl->cap * sizeof(int), pl->stream));
    if (hb_k_init(&pl->dev, pl->stream) != 0) { rc = HBICG_E_HIP; goto fail; }
    if ((rc = run_loop(pl)) != 0) goto fail;
    /* profiled solves time the finalize: it holds the last window flush
     * when the shift kernel runs the scalar stage */
    if (pl->profile) HIPCK(hipEventRecord(pl->pev[0], pl->stream));
    if (hb_k_finalize(&pl->dev, pl->stream) != 0) { rc = HBICG_E_HIP; goto fail; }
    if (pl->profile) HIPCK(hipEventRecord(pl->pev[1], pl->stream));
    HIPCK(hipEventRecord(pl->t1, pl->stream));

    hb_state st;
    HIPCK(hipMemcpyAsync(&st, pl->st, sizeof(st), hipMemcpyDeviceToHost, pl->stream));
    HIPCK(hipStreamSynchronize(pl->stream));
    const double *r_dev = pl->rring + (size_t)(st.kprev % (pl->W + 1)) * pl->n;
    if (r_out) {
        if (on_device) HIPCK(hipMemcpyAsync(r_out, r_dev, nb, hipMemcpyDeviceToDevice, pl->stream));
        else HIPCK(hipMemcpyAsync(r_out, r_dev, nb, hipMemcpyDeviceToHost, pl->stream));
    }
    if (!on_device) HIPCK(hipMemcpyAsync(x_set, dx, xb, hipMemcpyDeviceToHost, pl->stream));
    HIPCK(hipMemcpyAsync(pl->h_stop_iter, pl->stop_iter, (size_t)pl->M * 4, hipMemcpyDeviceToHost, pl->stream));
    HIPCK(hipMemcpyAsync(pl->h_switch_k, pl->switch_k, (size_t)pl->M * 4, hipMemcpyDeviceToHost, pl->stream));
    HIPCK(hipMemcpyAsync(pl->h_switch_to, pl->switch_to, (size_t)pl->M * 4, hipMemcpyDeviceToHost, pl->stream));
    HIPCK(hipStreamSynchronize(pl->stream));
    float ms = 0.0f;
    HIPCK(hipEventElapsedTime(&ms, pl->t0, pl->t1));

    hbicg_report *R = &pl->rep;
    memset(R, 0, sizeof(*R));
    R->k = st.k;
    R->stop_count = st.stop_count;
    R->n_switch = st.n_switch;
    R->seed_final = st.seed;
    R->status = st.status;
    R->M = pl->M;
    R->n = pl->n;
    R->window = pl->W;
    R->value_table = pl->dev.nvdict;
    R->diag_tiles = pl->n_diag_tiles;
    R->q_unfused = pl->dev.q_unfused;
    R->panels = pl->dev.npanel > 1 ? pl->dev.npanel : 1;
    R->const_tiles = pl->n_const_tiles;
    R->dot_r = st.dot_r;
    R->dot_zero = st.dot_zero;
    R->rTr = st.rTr;
    R->t_solve_ms = ms;
    R->t_total_ms = now_ms() - t_start;
    if (pl->profile) {
        const double n = (double)pl->n;
        const double spmv = pl->val_bytes + pl->col_bytes + pl->ptr_bytes + 4.0 * (pl->nblk + 1.0);
        /* algorithmic bytes per launch: SpMV 1 gathers p_s and reads r^
         * (writes s), SpMV 2 reads r and s (writes q and y), the seed update
         * reads q, y, r^ and writes r (exact order: also p_s and x_s in,
         * x_s out; the default mode defers x_s to the window flush); the
         * finalizes read the workgroup partial pairs; the shift kernel's
         * bytes are counted on the device (seed p update plus the window
         * flushes) */
        const double fin_spmv = 16.0 * pl->nblk, fin_seed = 16.0 * pl->dev.vec_grid;
        const double seed_upd = (pl->exact ? 56.0 : 32.0) * n;
        /* SpMV 2 with q formed beforehand: r, s in and q out (k_qform), q
         * gathered, y out */
        const double spmv2 = spmv + (pl->dev.q_unfused ? 40.0 : 32.0) * n;
        const double per_launch[HB_NSTAGE] = {spmv + 24.0 * n, fin_spmv, spmv2, fin_spmv,
                                              seed_upd, fin_seed, 0.0, 0.0};
        const int iters = st.k - 1;
        for (int sg = 0; sg < HB_NSTAGE; ++sg) {
            hbicg_kernel_stat *K = &pl->kstat[sg];
            K->launches = iters;
            K->total_ms = 0.0;
            K->replays = 0;
            K->replay_ms = 0.0;
            for (int it = 0; it < iters && it < pl->ptimes_cap; ++it) {
                K->total_ms += pl->ptimes[(size_t)it * HB_PT + sg];
                if (sg == HB_STAGE_SPMV2 && (it % pl->chunk) % HB_REPLAY_EVERY == 0) {
                    K->replays += HB_REPLAY;
                    K->replay_ms += pl->ptimes[(size_t)it * HB_PT + HB_NSTAGE];
                }
            }
            K->bytes = per_launch[sg] * iters;
            K->flops = 0.0;
        }
        /* stages that run inside a neighbouring launch (fused): no launches
         * of their own; their bytes go to the launch that runs them */
        static const int host_of[HB_NSTAGE] = {-1, HB_STAGE_SPMV2, -1, HB_STAGE_SEED, -1, HB_STAGE_SHIFT,
                                               HB_STAGE_SHIFT, -1};
        for (int sg = 0; sg < HB_NSTAGE; ++sg) {
            if (hb_k_stage_runs(&pl->dev, sg)) continue;
            hbicg_kernel_stat *K = &pl->kstat[sg];
            if (host_of[sg] >= 0) pl->kstat[host_of[sg]].bytes += K->bytes;
            K->launches = 0;
            K->total_ms = 0.0;
            K->bytes = 0.0;
        }
        /* the shift kernel's launches split in two: [7] the seed p update
         * alone (p, r, s in, p out), [8] the launches that also ran a window
         * flush (the device-counted rest of the bytes, and its multiply-adds) */
        HIPCK(hipMemcpy(pl->flush_host, pl->flush_log, (size_t)pl->cap * sizeof(int), hipMemcpyDeviceToHost));
        hbicg_kernel_stat *Kp = &pl->kstat[HB_STAGE_SHIFT], *Kf = &pl->kstat[HB_NSTAGE];
        memset(Kf, 0, sizeof(*Kf));
        Kp->total_ms = 0.0;
        Kp->launches = 0;
        for (int it = 0; it < iters && it < pl->ptimes_cap; ++it) {
            const double t = pl->ptimes[(size_t)it * HB_PT + HB_STAGE_SHIFT];
            if (it + 1 < pl->cap && pl->flush_host[it + 1]) {
                Kf->launches += 1;
                Kf->total_ms += t;
            } else {
                Kp->launches += 1;
                Kp->total_ms += t;
            }
        }
        if (pl->dev.fused_s && Kf->launches > 0) {
            float fin = 0.0f;
            HIPCK(hipEventElapsedTime(&fin, pl->pev[0], pl->pev[1]));
            Kf->total_ms += fin;
        }
        Kp->bytes = 32.0 * n * Kp->launches;
        Kf->bytes = st.shift_bytes - Kp->bytes;
        Kf->flops = 2.0 * st.flush_fma * n;
    }
    if (rep) *rep = *R;
    return st.k;
fail:
    hipStreamSynchronize(pl->stream);
    return rc;
}

int hbicg_plan_residual_trace(hbicg_plan *pl, double *out, int rows)
{
    int rc = 0;
    if (!pl || !out || rows < 0 || !pl->trace) return HBICG_E_ARG;
    const int m = rows < pl->rep.k ? rows : pl->rep.k;
    if (m > 0) {
        HIPCK(hipSetDevice(pl->device));
        HIPCK(hipMemcpy(out, pl->trace, (size_t)m * pl->M * 8, hipMemcpyDeviceToHost));
    }
    return m;
fail:
    return rc;
}

int hbicg_plan_kernel_stats(hbicg_plan *pl, hbicg_kernel_stat *stats, int nmax)
{
    if (!pl || !stats) return HBICG_E_ARG;
    const int m = nmax < HB_NSTAGE + 1 ? nmax : HB_NSTAGE + 1;
    for (int i = 0; i < m; ++i) stats[i] = pl->kstat[i];
    return m;
}

int hbicg_plan_report(hbicg_plan *pl, hbicg_report *rep, int *stop_iter,
                      int *switch_k, int *switch_to, int max_switch)
{
    if (!pl) return HBICG_E_ARG;
    if (rep) *rep = pl->rep;
    if (stop_iter) memcpy(stop_iter, pl->h_stop_iter, (size_t)pl->M * sizeof(int));
    const int ns = pl->rep.n_switch < max_switch ? pl->rep.n_switch : max_switch;
    if (switch_k) memcpy(switch_k, pl->h_switch_k, (size_t)(ns > 0 ? ns : 0) * sizeof(int));
    if (switch_to) memcpy(switch_to, pl->h_switch_to, (size_t)(ns > 0 ? ns : 0) * sizeof(int));
    return pl->rep.n_switch;
}

int hbicg_plan_true_residuals(hbicg_plan *pl, const double *b, const double *sigma,
                              const double *x_set, int on_device, double *out)
{
    int rc = 0;
    double *part = NULL, *dout = NULL;
    if (!pl || !b || !sigma || !x_set || !out) return HBICG_E_ARG;
    HIPCK(hipSetDevice(pl->device));
    const size_t nb = (size_t)pl->n * 8, xb = nb * (size_t)pl->M;
    const double *dx = x_set, *db = b;
    if (!on_device) {
        if (!pl->x_host_buf) {
            pl->x_host_buf = (double *)dalloc(xb, &rc);
            if (rc) return rc;
        }
        HIPCK(hipMemcpyAsync(pl->x_host_buf, x_set, xb, hipMemcpyHostToDevice, pl->stream));
        HIPCK(hipMemcpyAsync(pl->b_buf, b, nb, hipMemcpyHostToDevice, pl->stream));
        dx = pl->x_host_buf;
        db = pl->b_buf;
    }
    HIPCK(hipMemcpyAsync(pl->sigma, sigma, (size_t)pl->M * 8, hipMemcpyHostToDevice, pl->stream));
    const int R = hb_k_resid_blocks(pl->n, pl->M);
    part = (double *)dalloc((size_t)R * pl->M * 2 * 8, &rc);
    dout = (double *)dalloc((size_t)pl->M * 8, &rc);
    if (rc) goto fail;
    if (hb_k_true_residual(&pl->dev, db, pl->sigma, dx, part, dout, pl->stream) != 0) {
        rc = HBICG_E_HIP;
        goto fail;
    }
    HIPCK(hipMemcpyAsync(out, dout, (size_t)pl->M * 8, hipMemcpyDeviceToHost, pl->stream));
    HIPCK(hipStreamSynchronize(pl->stream));
fail:
    if (part) hipFree(part);
    if (dout) hipFree(dout);
    return rc;
}

int hbicg_plan_spmv(hbicg_plan *pl, const double *x, double *y, double shift)
{
    if (!pl || !x || !y) return HBICG_E_ARG;
    if (hipSetDevice(pl->device) != hipSuccess) return HBICG_E_HIP;
    if (hb_k_spmv(&pl->dev, x, y, shift, pl->stream) != 0) return HBICG_E_HIP;
    if (hipStreamSynchronize(pl->stream) != hipSuccess) return HBICG_E_HIP;
    return 0;
}

/* ---- reference entry points ------------------------------------------------ */
static hbicg_report g_last;
static int *g_stop_iter = NULL, *g_switch_k = NULL, *g_switch_to = NULL, g_last_M = 0;
static double *g_trace = NULL;
static int g_trace_rows = 0;

int hbicg_last_residual_trace(double *out, int M, int rows)
{
    if (!out || rows < 0 || M != g_last_M) return HBICG_E_ARG;
    if (!g_trace) return 0;
    const int m = rows < g_trace_rows ? rows : g_trace_rows;
    memcpy(out, g_trace, (size_t)m * M * sizeof(double));
    return m;
}

int hbicg_last_report(hbicg_report *rep, int *stop_iter, int M,
                      int *switch_k, int *switch_to, int max_switch)
{
    if (rep) *rep = g_last;
    if (stop_iter && g_stop_iter) memcpy(stop_iter, g_stop_iter, (size_t)(M < g_last_M ? M : g_last_M) * sizeof(int));
    const int ns = g_last.n_switch < max_switch ? g_last.n_switch : max_switch;
    if (ns > 0 && switch_k && g_switch_k) memcpy(switch_k, g_switch_k, (size_t)ns * sizeof(int));
    if (ns > 0 && switch_to && g_switch_to) memcpy(switch_to, g_switch_to, (size_t)ns * sizeof(int));
    return g_last.n_switch;
}

/* variant: 0 normal, 1 static, 2 dynamic */
static int dropin(CSR_Matrix *A_diag, CSR_Matrix *A_offd, INFO_Matrix *A_info, double *x_set,
                  double *r, double *sigma, int M, int seed, int variant)
{
    const hbicg_params *p = params();
    if (!A_diag || !A_info || !x_set || !r || !sigma) return HBICG_E_ARG;
    /* src/solver_dynamic.c:11-14 */
    if (A_info->cols != A_info->rows) {
        if (p->exit_on_error) {
            printf("Error: matrix is not square.\n");
            exit(1);
        }
        return HBICG_E_NOTSQUARE;
    }
    if (A_offd && A_offd->nz != 0) return HBICG_E_OFFD;
    if (A_diag->rows != A_info->rows || A_diag->cols != A_info->cols) return HBICG_E_OFFD;
    if (M <= 0 || seed < 0 || seed >= M) return HBICG_E_ARG;
    if (!A_diag->ptr || (A_diag->nz > 0 && (!A_diag->val || !A_diag->col))) return HBICG_E_ARG;
    const double t0 = now_ms();
    hbicg_csr A = {A_diag->val, A_diag->col, A_diag->ptr, A_diag->rows, A_diag->nz};
    hbicg_plan *pl = plan_create_v(&A, 0, M, p, NULL, variant, 0, HB_WMAX, p->value_index, 0, 1);
    if (!pl) return g_errno ? g_errno : HBICG_E_HIP;
    const double t1 = now_ms();
    const int k = hbicg_plan_solve(pl, r, sigma, seed, x_set, r, 0, NULL);
    if (k > 0) {
        g_last = pl->rep;
        g_last.t_setup_ms = t1 - t0;
        g_last.t_total_ms = now_ms() - t0;
        if (g_last_M < M) {
            free(g_stop_iter); free(g_switch_k); free(g_switch_to);
            g_stop_iter = (int *)calloc((size_t)M, sizeof(int));
            g_switch_k = (int *)calloc((size_t)M, sizeof(int));
            g_switch_to = (int *)calloc((size_t)M, sizeof(int));
        }
        g_last_M = M;
        if (g_stop_iter) memcpy(g_stop_iter, pl->h_stop_iter, (size_t)M * sizeof(int));
        if (g_switch_k) memcpy(g_switch_k, pl->h_switch_k, (size_t)M * sizeof(int));
        if (g_switch_to) memcpy(g_switch_to, pl->h_switch_to, (size_t)M * sizeof(int));
        free(g_trace);
        g_trace = NULL;
        g_trace_rows = 0;
        if (pl->trace) {
            g_trace = (double *)malloc((size_t)k * M * sizeof(double));
            if (g_trace && hbicg_plan_residual_trace(pl, g_trace, k) == k) g_trace_rows = k;
        }
    }
    hbicg_plan_destroy(pl);
    return k;
}

int shifted_lopbicg_dynamic(CSR_Matrix *A_loc_diag, CSR_Matrix *A_loc_offd, INFO_Matrix *A_info,
                            double *x_loc_set, double *r_loc, double *sigma, int sigma_len, int seed)
{
    return dropin(A_loc_diag, A_loc_offd, A_info, x_loc_set, r_loc, sigma, sigma_len, seed, 2);
}

int shifted_lopbicg_static(CSR_Matrix *A_loc_diag, CSR_Matrix *A_loc_offd, INFO_Matrix *A_info,
                           double *x_loc_set, double *r_loc, double *sigma, int sigma_len, int seed)
{
    return dropin(A_loc_diag, A_loc_offd, A_info, x_loc_set, r_loc, sigma, sigma_len, seed, 1);
}

int shifted_lopbicg_normal(CSR_Matrix *A_loc_diag, CSR_Matrix *A_loc_offd, INFO_Matrix *A_info,
                           double *x_loc_set, double *r_loc, double *sigma, int sigma_len, int seed)
{
    return dropin(A_loc_diag, A_loc_offd, A_info, x_loc_set, r_loc, sigma, sigma_len, seed, 0);
}

/* ---- batched solver -----------------------------------------------------------
 * Many independent problems (different matrices, sizes, shift counts) solved
 * together: each phase of the iteration is one launch whose workgroups are
 * the tiles of every problem, found through (problem, tile) tables; done
 * problems' tiles return at once.  One chunk of iterations is replayed as a
 * hipGraph, and one device flag reports when every problem is done. */
/* chunk graphs by launch width: level l gives the phase launches
 * ceil(full / 2^l) workgroups */
#define HB_GLEVELS 7

struct hbicg_batch {
    int nprob, device, exact, chunk, graph, W, group_cap;
    int wmax;                  /* the largest window the rings hold */
    hipStream_t stream;
    int own_stream;
    hbicg_plan **plans;
    hb_dev *hdevs;             /* host copies of the per-problem records */
    hb_dev *ddevs;             /* device array                           */
    int *t_spmv, *t_vec, *t_shift, *t_coef;
    int *a_tab[HB_NPH], *act;  /* device: active lists and their lengths */
    int *hpofs, *hpcnt;        /* host [HB_NPH][nprob]: each problem's range in the full tables */
    int *pofs, *pcnt;          /* device copies                          */
    hb_batch B;
    int *flag_host;            /* pinned: done, problems running, active entries per phase */
    hipEvent_t ev[2], t0, t1;
    hipGraphExec_t gexec[HB_GLEVELS];
    hipGraph_t graph_obj[HB_GLEVELS];
    int graph_dirty;
    int levels_used;           /* bit l: the last run launched level l */
    double last_ms;
};

void hbicg_batch_destroy(hbicg_batch *bt)
{
    if (!bt) return;
    hipSetDevice(bt->device);
    if (bt->stream) hipStreamSynchronize(bt->stream);
    for (int l = 0; l < HB_GLEVELS; ++l) {
        if (bt->gexec[l]) hipGraphExecDestroy(bt->gexec[l]);
        if (bt->graph_obj[l]) hipGraphDestroy(bt->graph_obj[l]);
    }
    for (int ph = 0; ph < HB_NPH; ++ph) if (bt->a_tab[ph]) hipFree(bt->a_tab[ph]);
    if (bt->act) hipFree(bt->act);
    if (bt->pofs) hipFree(bt->pofs);
    if (bt->pcnt) hipFree(bt->pcnt);
    free(bt->hpofs);
    free(bt->hpcnt);
    if (bt->plans)
        for (int p = 0; p < bt->nprob; ++p) hbicg_plan_destroy(bt->plans[p]);
    free(bt->plans);
    free(bt->hdevs);
    if (bt->ddevs) hipFree(bt->ddevs);
    if (bt->t_spmv) hipFree(bt->t_spmv);
    if (bt->t_vec) hipFree(bt->t_vec);
    if (bt->t_shift) hipFree(bt->t_shift);
    if (bt->t_coef) hipFree(bt->t_coef);
    if (bt->flag_host) hipHostFree(bt->flag_host);
    for (int i = 0; i < 2; ++i) if (bt->ev[i]) hipEventDestroy(bt->ev[i]);
    if (bt->t0) hipEventDestroy(bt->t0);
    if (bt->t1) hipEventDestroy(bt->t1);
    if (bt->own_stream && bt->stream) hipStreamDestroy(bt->stream);
    free(bt);
}

/* shift tiles of problem p with at most G shifts per workgroup row */
static void shift_grid(const hbicg_plan *pl, int G, int *gx, int *gy)
{
    *gx = (pl->n + HB_BLOCK - 1) / HB_BLOCK;
    *gy = (pl->M + G - 1) / G;
}

/* (Re)builds the shift tile table for group cap G (0: all shifts of a
 * problem in one workgroup row) and uploads the per-problem records.  The
 * default mode's combined-coefficient flush takes every shift in one row. */
static int batch_layout(hbicg_batch *bt, int W, int group_cap)
{
    int rc = 0, *tab = NULL;
    size_t tiles = 0;
    if (!bt->exact) group_cap = 0;
    if (W < 1 || W > bt->wmax) return HBICG_E_ARG;
    for (int p = 0; p < bt->nprob; ++p) {
        const hbicg_plan *pl = bt->plans[p];
        const int G = group_cap > 0 && group_cap < pl->M ? group_cap : pl->M;
        int gx, gy;
        shift_grid(pl, G, &gx, &gy);
        tiles += (size_t)gx * gy;
        hb_dev *d = &bt->hdevs[p];
        d->W = W;
        d->shift_group = G;
        d->shift_gx = gx;
        d->shift_gy = gy;
        d->vec2 = 0;
    }
    tab = (int *)malloc(tiles * 2 * sizeof(int));
    if (!tab) return HBICG_E_NOMEM;
    size_t t = 0;
    for (int p = 0; p < bt->nprob; ++p) {
        const hb_dev *d = &bt->hdevs[p];
        for (int i = 0; i < d->shift_gx * d->shift_gy; ++i) { tab[2 * t] = p; tab[2 * t + 1] = i; ++t; }
    }
    if (bt->t_shift) hipFree(bt->t_shift);
    bt->t_shift = (int *)dalloc(tiles * 2 * sizeof(int), &rc);
    if (bt->a_tab[HB_PH_SHIFT]) hipFree(bt->a_tab[HB_PH_SHIFT]);
    bt->a_tab[HB_PH_SHIFT] = (int *)dalloc(tiles * 2 * sizeof(int), &rc);
    if (rc) { free(tab); return rc; }
    HIPCK(hipMemcpy(bt->t_shift, tab, tiles * 2 * sizeof(int), hipMemcpyHostToDevice));
    {
        int o = 0;
        for (int p = 0; p < bt->nprob; ++p) {
            const int c = bt->hdevs[p].shift_gx * bt->hdevs[p].shift_gy;
            bt->hpofs[HB_PH_SHIFT * bt->nprob + p] = o;
            bt->hpcnt[HB_PH_SHIFT * bt->nprob + p] = c;
            o += c;
        }
    }
    HIPCK(hipMemcpy(bt->pofs, bt->hpofs, (size_t)HB_NPH * bt->nprob * sizeof(int), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(bt->pcnt, bt->hpcnt, (size_t)HB_NPH * bt->nprob * sizeof(int), hipMemcpyHostToDevice));
    bt->B.t_shift = bt->t_shift;
    bt->B.a_tab[HB_PH_SHIFT] = bt->a_tab[HB_PH_SHIFT];
    bt->B.n_shift_tiles = (int)tiles;
    bt->B.W = W;
    bt->W = W;
    bt->group_cap = group_cap;
    bt->graph_dirty = 1;
    for (int p = 0; p < bt->nprob; ++p) bt->plans[p]->W = W;
fail:
    free(tab);
    return rc;
}

hbicg_batch *hbicg_batch_create(const hbicg_csr *A, int a_on_device, int nprob, const int *M,
                                const hbicg_params *p, void *stream)
{
    int rc = 0, *tab = NULL;
    g_errno = 0;
    if (!A || !M || nprob <= 0) { g_errno = HBICG_E_ARG; return NULL; }
    if (!p) p = params();
    hbicg_batch *bt = (hbicg_batch *)calloc(1, sizeof(*bt));
    if (!bt) { g_errno = HBICG_E_NOMEM; return NULL; }
    bt->nprob = nprob;
    bt->device = p->device;
    bt->exact = p->exact_order ? 1 : 0;
    bt->chunk = p->chunk > 0 ? p->chunk : 16;
    bt->graph = p->graph != 0;
    HIPCK(hipSetDevice(p->device));
    if (stream) {
        bt->stream = (hipStream_t)stream;
    } else {
        HIPCK(hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking));
        bt->own_stream = 1;
    }
    bt->plans = (hbicg_plan **)calloc((size_t)nprob, sizeof(hbicg_plan *));
    bt->hdevs = (hb_dev *)calloc((size_t)nprob, sizeof(hb_dev));
    if (!bt->plans || !bt->hdevs) { rc = HBICG_E_NOMEM; goto fail; }
    size_t nsp = 0, nvec = 0, ncoef = 0;
    /* ring storage for the largest window the autotuner may pick: halved
     * while the rings of all problems (2W+1 n-vectors each) would take more
     * than a quarter of the free device memory */
    int wmax = bt->exact ? HB_WMAX_EXACT : HB_WBATCH;
    if (!bt->exact) {
        size_t free_b = 0, total_b = 0;
        double rows = 0.0;
        for (int q = 0; q < nprob; ++q) rows += (double)A[q].n;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
        while (wmax > 16 && (2.0 * wmax + 1.0) * rows * 8.0 > 0.25 * (double)free_b) wmax /= 2;
    }
    bt->wmax = wmax;
    for (int q = 0; q < nprob; ++q) {
        /* batched plans keep index bytes in their diagonal tiles: the
         * constant-diagonal form measured 199 against 192 ms on the
         * 256-problem batch (its tiles read each problem's record from
         * memory before the diagonal values) */
        bt->plans[q] = plan_create_v(&A[q], a_on_device, M[q], p, bt->stream, 1, wmax, wmax, p->value_index, 1, 0);
        if (!bt->plans[q]) { rc = g_errno ? g_errno : HBICG_E_HIP; goto fail; }
        bt->hdevs[q] = bt->plans[q]->dev;
        nsp += (size_t)bt->plans[q]->nblk;
        nvec += (size_t)bt->plans[q]->dev.vec_grid;
        ncoef += (size_t)M[q];
    }
    {
        size_t most = nsp > nvec ? nsp : nvec;
        if (ncoef > most) most = ncoef;
        tab = (int *)malloc(most * 2 * sizeof(int));
    }
    if (!tab) { rc = HBICG_E_NOMEM; goto fail; }
    {
        size_t t = 0;
        for (int q = 0; q < nprob; ++q)
            for (int j = 0; j < M[q]; ++j) { tab[2 * t] = q; tab[2 * t + 1] = j; ++t; }
        bt->t_coef = (int *)dalloc(ncoef * 2 * sizeof(int), &rc);
        if (rc) goto fail;
        HIPCK(hipMemcpy(bt->t_coef, tab, ncoef * 2 * sizeof(int), hipMemcpyHostToDevice));
        t = 0;
        for (int q = 0; q < nprob; ++q)
            for (int b = 0; b < bt->plans[q]->nblk; ++b) { tab[2 * t] = q; tab[2 * t + 1] = b; ++t; }
        bt->t_spmv = (int *)dalloc(nsp * 2 * sizeof(int), &rc);
        if (rc) goto fail;
        HIPCK(hipMemcpy(bt->t_spmv, tab, nsp * 2 * sizeof(int), hipMemcpyHostToDevice));
        t = 0;
        for (int q = 0; q < nprob; ++q)
            for (int b = 0; b < bt->plans[q]->dev.vec_grid; ++b) { tab[2 * t] = q; tab[2 * t + 1] = b; ++t; }
        bt->t_vec = (int *)dalloc(nvec * 2 * sizeof(int), &rc);
        if (rc) goto fail;
        HIPCK(hipMemcpy(bt->t_vec, tab, nvec * 2 * sizeof(int), hipMemcpyHostToDevice));
    }
    bt->ddevs = (hb_dev *)dalloc((size_t)nprob * sizeof(hb_dev), &rc);
    bt->hpofs = (int *)calloc((size_t)HB_NPH * nprob, sizeof(int));
    bt->hpcnt = (int *)calloc((size_t)HB_NPH * nprob, sizeof(int));
    if (!bt->hpofs || !bt->hpcnt) rc = HBICG_E_NOMEM;
    bt->pofs = (int *)dalloc((size_t)HB_NPH * nprob * sizeof(int), &rc);
    bt->pcnt = (int *)dalloc((size_t)HB_NPH * nprob * sizeof(int), &rc);
    bt->act = (int *)dalloc(HB_NPH * sizeof(int), &rc);
    bt->a_tab[HB_PH_SPMV] = (int *)dalloc(nsp * 2 * sizeof(int), &rc);
    bt->a_tab[HB_PH_VEC] = (int *)dalloc(nvec * 2 * sizeof(int), &rc);
    bt->a_tab[HB_PH_COEF] = (int *)dalloc(ncoef * 2 * sizeof(int), &rc);
    if (rc) goto fail;
    {
        int os = 0, ov = 0, oc = 0;
        for (int q = 0; q < nprob; ++q) {
            bt->hpofs[HB_PH_SPMV * nprob + q] = os;
            bt->hpcnt[HB_PH_SPMV * nprob + q] = bt->plans[q]->nblk;
            os += bt->plans[q]->nblk;
            bt->hpofs[HB_PH_VEC * nprob + q] = ov;
            bt->hpcnt[HB_PH_VEC * nprob + q] = bt->plans[q]->dev.vec_grid;
            ov += bt->plans[q]->dev.vec_grid;
            bt->hpofs[HB_PH_COEF * nprob + q] = oc;
            bt->hpcnt[HB_PH_COEF * nprob + q] = M[q];
            oc += M[q];
        }
    }
    bt->B.a_tab[HB_PH_SPMV] = bt->a_tab[HB_PH_SPMV];
    bt->B.a_tab[HB_PH_VEC] = bt->a_tab[HB_PH_VEC];
    bt->B.a_tab[HB_PH_COEF] = bt->a_tab[HB_PH_COEF];
    bt->B.act = bt->act;
    bt->B.pofs = bt->pofs;
    bt->B.pcnt = bt->pcnt;
    bt->B.div = 1;
    HIPCK(hipHostMalloc((void **)&bt->flag_host, (2 + HB_NPH) * sizeof(int), hipHostMallocDefault));
    HIPCK(hipEventCreateWithFlags(&bt->ev[0], hipEventDisableTiming));
    HIPCK(hipEventCreateWithFlags(&bt->ev[1], hipEventDisableTiming));
    HIPCK(hipEventCreate(&bt->t0));
    HIPCK(hipEventCreate(&bt->t1));
    bt->B.nprob = nprob;
    bt->B.exact = bt->exact;
    /* the value-table SpMV when every problem has a table (one kernel per
     * launch serves all problems) */
    bt->B.vi = 1;
    for (int q = 0; q < nprob; ++q)
        if (!bt->plans[q]->dev.vidx) bt->B.vi = 0;
    bt->B.devs = bt->ddevs;
    bt->B.t_spmv = bt->t_spmv;
    bt->B.t_vec = bt->t_vec;
    bt->B.n_spmv_tiles = (int)nsp;
    bt->B.n_vec_tiles = (int)nvec;
    bt->B.t_coef = bt->t_coef;
    bt->B.n_coef_tiles = (int)ncoef;
    {
        const int wdef = bt->exact || wmax < HB_WBATCH_DEF ? wmax : HB_WBATCH_DEF;
        if ((rc = batch_layout(bt, p->window > 0 && p->window < wmax ? p->window : wdef, 0)) != 0)
            goto fail;
    }
    free(tab);
    return bt;
fail:
    free(tab);
    hbicg_batch_destroy(bt);
    g_errno = rc ? rc : HBICG_E_HIP;
    return NULL;
}

static int batch_enqueue_chunk(hbicg_batch *bt)
{
    if (hb_kb_compact(&bt->B, bt->stream) != 0) return HBICG_E_HIP;
    for (int c = 0; c < bt->chunk; ++c) {
        const int e = hb_kb_iteration(&bt->B, bt->stream);
        if (e) return e;
    }
    return hb_kb_publish(&bt->B, bt->flag_host, bt->stream);
}

/* the chunk graph of width level l (captured on first use) */
static int batch_capture(hbicg_batch *bt, int l)
{
    int rc = 0;
    if (bt->graph_dirty) {
        for (int i = 0; i < HB_GLEVELS; ++i) {
            if (bt->gexec[i]) { hipGraphExecDestroy(bt->gexec[i]); bt->gexec[i] = NULL; }
            if (bt->graph_obj[i]) { hipGraphDestroy(bt->graph_obj[i]); bt->graph_obj[i] = NULL; }
        }
        bt->graph_dirty = 0;
    }
    if (bt->gexec[l]) return 0;
    bt->B.div = 1 << l;
    HIPCK(hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
    if (batch_enqueue_chunk(bt) != 0) {
        hipGraph_t g = NULL;
        hipStreamEndCapture(bt->stream, &g);
        if (g) hipGraphDestroy(g);
        return HBICG_E_HIP;
    }
    HIPCK(hipStreamEndCapture(bt->stream, &bt->graph_obj[l]));
    HIPCK(hipGraphInstantiate(&bt->gexec[l], bt->graph_obj[l], NULL, NULL, 0));
    return 0;
fail:
    return rc;
}

/* Launch width for the next chunk: the largest level whose phase launches
 * still give every active SpMV tile of the last published count a
 * workgroup of its own (the lists only shrink, so an older count errs
 * wide). */
static int batch_level(const hbicg_batch *bt, int active_spmv)
{
    int l = 0;
    while (l + 1 < HB_GLEVELS && ((bt->B.n_spmv_tiles + (2 << l) - 1) >> (l + 1)) >= active_spmv) ++l;
    return l;
}

/* Runs the batch with the current layout, max_iter_override > 0 capping
 * every problem's iterations (tuning runs).  b, x: per-problem pointers,
 * device memory; sigma: host arrays. */
static int batch_run(hbicg_batch *bt, const double *const *b, const double *const *sigma,
                     const int *seed, double *const *x, int max_iter_override)
{
    int rc = 0;
    HIPCK(hipSetDevice(bt->device));
    int max_cap = 0;
    for (int p = 0; p < bt->nprob; ++p) {
        hbicg_plan *pl = bt->plans[p];
        hb_dev *d = &bt->hdevs[p];
        if (!b[p] || !sigma[p] || !x[p] || seed[p] < 0 || seed[p] >= pl->M) return HBICG_E_ARG;
        d->x_set = x[p];
        d->b_in = b[p];
        d->seed_in = seed[p];
        d->cap = max_iter_override > 0 ? max_iter_override + 1 : pl->cap;
        if (d->cap > pl->cap) d->cap = pl->cap;
        if (d->cap > max_cap) max_cap = d->cap;
        HIPCK(hipMemcpyAsync(pl->sigma, sigma[p], (size_t)pl->M * 8, hipMemcpyHostToDevice, bt->stream));
        HIPCK(hipMemsetAsync(pl->st, 0, sizeof(hb_state), bt->stream));
    }
    HIPCK(hipMemcpyAsync(bt->ddevs, bt->hdevs, (size_t)bt->nprob * sizeof(hb_dev), hipMemcpyHostToDevice,
                         bt->stream));
    HIPCK(hipEventRecord(bt->t0, bt->stream));
    if (hb_kb_init(&bt->B, bt->stream) != 0) return HBICG_E_HIP;
    bt->flag_host[0] = 0;
    bt->levels_used = 0;
    const int max_chunks = max_cap / bt->chunk + 3;
    int launched = 0, prev = -1, level = 0;
    for (;;) {
        bt->levels_used |= 1 << level;
        if (bt->graph) {
            if ((rc = batch_capture(bt, level)) != 0) return rc;
            HIPCK(hipGraphLaunch(bt->gexec[level], bt->stream));
        } else {
            bt->B.div = 1 << level;
            if (batch_enqueue_chunk(bt) != 0) return HBICG_E_HIP;
        }
        const int cur = launched & 1;
        HIPCK(hipEventRecord(bt->ev[cur], bt->stream));
        ++launched;
        if (prev >= 0) {
            HIPCK(hipEventSynchronize(bt->ev[prev]));
            if (__atomic_load_n(&bt->flag_host[0], __ATOMIC_ACQUIRE)) break;
            level = batch_level(bt, __atomic_load_n(&bt->flag_host[2 + HB_PH_SPMV], __ATOMIC_ACQUIRE));
        }
        prev = cur;
        if (launched > max_chunks) break;
    }
    if (hb_kb_finalize(&bt->B, bt->stream) != 0) return HBICG_E_HIP;
    HIPCK(hipEventRecord(bt->t1, bt->stream));
    HIPCK(hipStreamSynchronize(bt->stream));
    float ms = 0.0f;
    HIPCK(hipEventElapsedTime(&ms, bt->t0, bt->t1));
    bt->last_ms = ms;
    return 0;
fail:
    hipStreamSynchronize(bt->stream);
    return rc;
}

int hbicg_batch_solve(hbicg_batch *bt, const double *const *b, const double *const *sigma,
                      const int *seed, double *const *x, int *k_out, double *solve_ms)
{
    int rc = 0;
    if (!bt || !b || !sigma || !seed || !x) return HBICG_E_ARG;
    if ((rc = batch_run(bt, b, sigma, seed, x, 0)) != 0) return rc;
    for (int p = 0; p < bt->nprob; ++p) {
        hbicg_plan *pl = bt->plans[p];
        hb_state st;
        HIPCK(hipMemcpyAsync(&st, pl->st, sizeof(st), hipMemcpyDeviceToHost, bt->stream));
        HIPCK(hipMemcpyAsync(pl->h_stop_iter, pl->stop_iter, (size_t)pl->M * 4, hipMemcpyDeviceToHost, bt->stream));
        HIPCK(hipMemcpyAsync(pl->h_switch_k, pl->switch_k, (size_t)pl->M * 4, hipMemcpyDeviceToHost, bt->stream));
        HIPCK(hipMemcpyAsync(pl->h_switch_to, pl->switch_to, (size_t)pl->M * 4, hipMemcpyDeviceToHost, bt->stream));
        HIPCK(hipStreamSynchronize(bt->stream));
        hbicg_report *R = &pl->rep;
        memset(R, 0, sizeof(*R));
        R->k = st.k;
        R->stop_count = st.stop_count;
        R->n_switch = st.n_switch;
        R->seed_final = st.seed;
        R->status = st.status;
        R->M = pl->M;
        R->n = pl->n;
        R->window = bt->hdevs[p].W;
        /* the SpMV form the batch's launches read for this problem (they
         * form q at the gathered columns, one pass) */
        R->value_table = bt->B.vi ? pl->dev.nvdict : 0;
        R->diag_tiles = bt->B.vi ? pl->n_diag_tiles : 0;
        R->panels = 1;
        R->dot_r = st.dot_r;
        R->dot_zero = st.dot_zero;
        R->rTr = st.rTr;
        R->t_solve_ms = bt->last_ms;
        if (k_out) k_out[p] = st.k;
    }
    if (solve_ms) *solve_ms = bt->last_ms;
    return 0;
fail:
    return rc;
}

int hbicg_batch_report(hbicg_batch *bt, int p, hbicg_report *rep, int *stop_iter,
                       int *switch_k, int *switch_to, int max_switch)
{
    if (!bt || p < 0 || p >= bt->nprob) return HBICG_E_ARG;
    return hbicg_plan_report(bt->plans[p], rep, stop_iter, switch_k, switch_to, max_switch);
}

int hbicg_batch_true_residuals(hbicg_batch *bt, int p, const double *b, const double *sigma,
                               const double *x_set, double *out)
{
    if (!bt || p < 0 || p >= bt->nprob) return HBICG_E_ARG;
    return hbicg_plan_true_residuals(bt->plans[p], b, sigma, x_set, 1, out);
}

/* Candidate layouts are timed on the batch itself for `iters` iterations
 * (x is overwritten); the fastest per-iteration layout is kept. */
int hbicg_batch_autotune(hbicg_batch *bt, const double *const *b, const double *const *sigma,
                         const int *seed, double *const *x, int iters, hbicg_tuning *best)
{
    /* exact order: window replay with (window, shifts per workgroup row);
     * default mode: the combined-coefficient flush, window only (those the
     * rings hold).  A window's flushes cost its trial only once it reaches
     * them: time the default mode's candidates over whole solves (iters at
     * least the solve's length) */
    static const int windows_exact[] = {4, 8, 16}, groups_exact[] = {0, 16, 4};
    static const int windows_cc[] = {32, 64, 128}, groups_cc[] = {0};
    const int *windows = bt->exact ? windows_exact : windows_cc;
    const int *groups = bt->exact ? groups_exact : groups_cc;
    const size_t nw = 3, ng = bt->exact ? 3 : 1;
    int rc = 0;
    if (!bt || iters < 1) return HBICG_E_ARG;
    double best_ms = 1e300;
    int bw = bt->W, bg = bt->group_cap;
    for (size_t wi = 0; wi < nw; ++wi)
        for (size_t gi = 0; gi < ng; ++gi) {
            if (windows[wi] > bt->wmax) continue;
            if ((rc = batch_layout(bt, windows[wi], groups[gi])) != 0) return rc;
            double ms = 0.0;
            for (int rep = 0; rep < 2; ++rep) {   /* first run warms the graph */
                if ((rc = batch_run(bt, b, sigma, seed, x, iters)) != 0) return rc;
                ms = bt->last_ms;
            }
            if (ms < best_ms) { best_ms = ms; bw = windows[wi]; bg = groups[gi]; }
        }
    if ((rc = batch_layout(bt, bw, bg)) != 0) return rc;
    if (best) {
        best->window = bw;
        best->group_cap = bg;
        best->ms_per_iter = best_ms / iters;
    }
    return 0;
}

/* ---- adjoint support ---------------------------------------------------------- */
int hbicg_csr_transpose(const hbicg_csr *A, unsigned *tptr, unsigned *tcol, double *tval, void *stream)
{
    int rc = 0;
    unsigned *work = NULL;
    if (!A || !A->ptr || !tptr || (A->nz && (!A->col || !A->val || !tcol || !tval)) || A->n == 0)
        return HBICG_E_ARG;
    const hipStream_t s = (hipStream_t)stream;
    const size_t n = A->n, nb = (n + HB_BLOCK - 1) / HB_BLOCK;
    /* cnt | fill | bsum | the source index of every entry */
    work = (unsigned *)dalloc((2 * n + nb + (size_t)A->nz + 1) * sizeof(unsigned), &rc);
    if (rc) return rc;
    if (hb_transpose((int)n, A->nz, A->ptr, A->col, A->val, tptr, tcol, tval, work + 2 * n + nb, work, s) != 0) {
        rc = HBICG_E_HIP;
        goto fail;
    }
    HIPCK(hipStreamSynchronize(s));
fail:
    hipFree(work);
    return rc;
}

int hbicg_csr_transpose_pattern(const hbicg_csr *A, unsigned *tptr, unsigned *tcol, unsigned *tsrc,
                                void *stream)
{
    int rc = 0;
    unsigned *work = NULL;
    if (!A || !A->ptr || !tptr || (A->nz && (!A->col || !tcol || !tsrc)) || A->n == 0) return HBICG_E_ARG;
    const hipStream_t s = (hipStream_t)stream;
    const size_t n = A->n, nb = (n + HB_BLOCK - 1) / HB_BLOCK;
    work = (unsigned *)dalloc((2 * n + nb) * sizeof(unsigned), &rc);
    if (rc) return rc;
    if (hb_transpose((int)n, A->nz, A->ptr, A->col, NULL, tptr, tcol, NULL, tsrc, work, s) != 0) {
        rc = HBICG_E_HIP;
        goto fail;
    }
    HIPCK(hipStreamSynchronize(s));
fail:
    hipFree(work);
    return rc;
}

int hbicg_gather_values(const double *val, const unsigned *src, unsigned nz, double *out, void *stream)
{
    if (nz == 0) return 0;
    if (!val || !src || !out) return HBICG_E_ARG;
    return hb_gather_vals(nz, val, src, out, (hipStream_t)stream) != 0 ? HBICG_E_HIP : 0;
}

size_t hbicg_adjoint_work_bytes(unsigned n, int M)
{
    const size_t nb = ((size_t)n + HB_BLOCK - 1) / HB_BLOCK;
    return (M > 0 ? (size_t)M : 0) * (nb < 256 ? nb : 256) * sizeof(double);
}

int hbicg_adjoint_gradients_ws(const hbicg_csr *A, int M, const double *mu, const double *w,
                               const double *x, double *grad_b, double *grad_sigma, double *grad_val,
                               void *work, size_t work_bytes, void *stream)
{
    if (!A || !A->ptr || M <= 0 || !mu || !x || A->n == 0 || (grad_val && A->nz && !A->col)) return HBICG_E_ARG;
    if (grad_sigma && (!work || work_bytes < hbicg_adjoint_work_bytes(A->n, M))) return HBICG_E_ARG;
    return hb_adjoint_grads((int)A->n, M, A->ptr, A->col, mu, w, x, grad_b, grad_sigma, grad_val,
                            (double *)work, (hipStream_t)stream) != 0 ? HBICG_E_HIP : 0;
}

int hbicg_adjoint_gradients(const hbicg_csr *A, int M, const double *mu, const double *w,
                            const double *x, double *grad_b, double *grad_sigma, double *grad_val,
                            void *stream)
{
    int rc = 0;
    void *work = NULL;
    if (!A || !A->ptr || M <= 0 || !mu || !x || A->n == 0) return HBICG_E_ARG;
    const hipStream_t s = (hipStream_t)stream;
    const size_t bytes = hbicg_adjoint_work_bytes(A->n, M);
    work = dalloc(bytes, &rc);
    if (rc) return rc;
    rc = hbicg_adjoint_gradients_ws(A, M, mu, w, x, grad_b, grad_sigma, grad_val, work, bytes, stream);
    if (rc) goto fail;
    HIPCK(hipStreamSynchronize(s));
fail:
    hipFree(work);
    return rc;
}


/* ---- plan internals for the library's other translation units (hb_plan.h) ---- */
const hb_dev *hb_plan_dev(const hbicg_plan *pl) { return &pl->dev; }
hipStream_t hb_plan_stream(const hbicg_plan *pl) { return pl->stream; }
int hb_plan_device(const hbicg_plan *pl) { return pl->device; }
