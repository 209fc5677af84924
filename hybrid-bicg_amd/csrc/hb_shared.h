/*
 * hb_shared.h -- data shared between the C host driver and the HIP kernels:
 * the device-resident solver state and the bundle of device pointers one
 * solve works on.  Plain C so both gcc (host) and hipcc (kernels) read it.
 */
#ifndef HB_SHARED_H
#define HB_SHARED_H

#ifdef __cplusplus
extern "C" {
#endif

#define HB_BLOCK        256   /* threads per workgroup of every kernel        */
#define HB_NNZ_CAP      2048  /* nonzeros staged in LDS by one SpMV workgroup */
#define HB_ROWS_CAP     256   /* rows per SpMV workgroup                       */
#define HB_WMAX         1024  /* largest shift-update window                   */
#define HB_WAUTO        512   /* default window of static/dynamic, memory permitting */
#define HB_WBATCH       128   /* largest window of the batched solver (less where the
                                 rings of all its problems would not fit)      */
#define HB_WBATCH_DEF   64    /* the batched solver's window before autotuning    */
#define HB_WMAX_EXACT   16    /* largest window of the exact-order replay      */
#define HB_UG           16    /* shifts per chunk of the combined-coefficient flush */
#define HB_VDICT        256   /* largest value table of the value-indexed SpMV */
#define HB_DIAG_K       16    /* most diagonals of a diagonal tile             */
#define HB_DIAG_NONE    255   /* index byte of a diagonal a row has no entry on */
#define HB_CC_ROWS(W)   (2 * (W) + 5) /* rows of one chunk's coefficient block  */
/* default column-panel width of the seed SpMVs: 2 MiB of x, half an XCD's L2 */
#define HB_PANEL_COLS 262144
#define HB_MAXPANEL   32      /* most panels; wider panels beyond               */
#define HB_RESID_PAIRS  65536 /* most partial pairs of one true-residual evaluation */

/* Error bits in hb_state.status */
#define HB_ST_BREAKDOWN_RTS   1   /* (r^,s) == 0 or non-finite alpha          */
#define HB_ST_BREAKDOWN_QTY   2   /* (q,y) == 0 or non-finite omega           */
#define HB_ST_NO_SWITCH_TARGET 4  /* seed stopped but no shift qualified      */
#define HB_ST_BREAKDOWN_SHIFT  8  /* a shift's 1 - omega~ d or pi_k is 0 or non-finite */
#define HB_ST_SYNC_TIMEOUT    16  /* a workgroup gave up waiting for a value another
                                     workgroup of its launch publishes (never expected:
                                     the publishing workgroups are dispatched first) */

/* A mailbox word that holds no value yet: a signalling-NaN bit pattern, which
 * no arithmetic produces (NaN results are quiet). */
#define HB_MAIL_EMPTY 0x7ff4dead00c0ffeeull
/* beta~'s mailbox when the scalar stage did not run (the solve is done) */
#define HB_MAIL_SKIP 0x7ff4dead00005c1full
/* bits of the window decision the fused scalar stage publishes (mail_f);
 * 0: no flush, or the stage did not run */
#define HB_FL_FLUSH 1
#define HB_FL_XONLY 2

/*
 * Device-resident state of one shifted solve.  Every scalar the reference
 * keeps in host variables of shifted_lopbicg_* (src/solver_dynamic.c:22-46)
 * lives here, written only by the "last" workgroup of a kernel, so no value
 * crosses to the host during the iteration loop.
 */
typedef struct hb_state {
    double rTr;        /* (r^, r)                                  */
    double rTr_old;    /* (r^, r) of the previous iteration        */
    double dot_r;      /* (r, r)                                   */
    double dot_zero;   /* (b, b)                                   */
    double rTs, qTq, qTy;
    double zscale;     /* 1/(zeta_m pi_m[k]) to apply to r after a switch */
    double pb, pbw;    /* beta~[k] and -beta~[k]*omega~[k] for the seed p update */
    double shift_bytes;/* algorithmic HBM bytes the shift kernel moved (summed) */
    double flush_fma;  /* multiply-adds of the window flushes, per row of x_j (summed) */
    int k;             /* iteration about to run (1-based, reference k)   */
    int kprev;         /* iteration the shift kernel works for            */
    int seed;          /* current seed index                              */
    int prev_seed;     /* seed of the iteration just finished             */
    int stop_count;
    int max_sigma;     /* arg-max shift, persistent like the reference's  */
    int done;          /* 1 once stop_count == M or the cap is reached    */
    int switched;      /* a switch fired: r must be rescaled, rTr recomputed */
    int n_switch;
    int win_k0;        /* first iteration of the pending shift window     */
    int flush;         /* shift kernel applies the window this iteration  */
    int flush_k0;      /* window the flush covers: iterations flush_k0 .. */
    int flush_len;     /*   .. flush_k0 + flush_len - 1                   */
    int flush_cnt;     /* shifts in the flush list                        */
    int k4_epoch;      /* k when the shift kernel last ran (runs while < k) */
    int status;        /* HB_ST_* bits                                    */
    int flush_tail;    /* 1: the flush applies its last iteration on its own */
    int flush_xonly;   /* 1: last flush of the solve: x_j only, p_j not needed */
    int win_seedpos;   /* default mode: list position of the seed in the pending
                          window (its x is deferred like a shift's), else -1 */
    int coef_ready;    /* fused shift kernel: list positions whose combined
                          coefficients are stored (reset by SpMV 1) */
    /* Mailbox of the fused seed update: omega~ of the running iteration,
     * published by the launch's first workgroup to its other workgroups;
     * HB_MAIL_EMPTY until then (the scalar stage empties it). */
    unsigned long long mail_w;
    unsigned long long mail_a;   /* alpha~ handed from SpMV 2's first workgroup */
    /* The shift kernel that runs the scalar stage in its first workgroup:
     * beta~ (HB_MAIL_SKIP when the stage did not run), published as soon as
     * it is formed, then the window decision (HB_FL_* bits); both emptied by
     * SpMV 1. */
    unsigned long long mail_b;
    unsigned long long mail_f;
    double pu_w;       /* omega~ of the running iteration, and its iteration and */
    int pu_k, pu_seed; /* seed (written by omega~'s finalize, read by the seed p
                          update of the shift kernel that runs the scalar stage:
                          the stage rescales the archives at a switch)         */
    int pu_live;       /* 1 once omega~'s finalize ran this iteration (SpMV 1
                          clears it): the shift kernel's tiles have work      */
} hb_state;

/* Every pointer below addresses global device memory.  Device code sees
 * that in the pointee type (address space 1), so kernels that read a record
 * from device memory (the batched ones) still get global and scalar memory
 * instructions rather than flat ones; host code sees plain pointers of the
 * same size. */
#if defined(__HIP_DEVICE_COMPILE__)
#define HB_G __attribute__((address_space(1)))
#else
#define HB_G
#endif

/* Device pointers and sizes of one solve (passed by value to kernels). */
typedef struct hb_dev {
    int n, M, W, cap, nblk;    /* rows, shifts, window, archive length, SpMV workgroups */
    int exact;                 /* 1: reference summation order everywhere  */
    int fused;                 /* default mode: the omega~ finalize runs as the first
                                  workgroup of the seed update's launch, the combined
                                  coefficients as the first workgroups of the shift kernel */
    int q_unfused;             /* 1: SpMV 2 gathers q, formed beforehand by k_qform
                                  (matrices with scattered columns); 0: it forms
                                  q = r - alpha s at every gathered column       */
    int vec_grid;              /* workgroups of the elementwise kernels    */
    int shift_gx, shift_gy;    /* grid of the shift kernel                 */
    int shift_group;           /* shifts per shift-kernel workgroup row    */
    int vec2;                  /* 1: shift kernel moves 2 doubles per access (n even) */
    int x_zero;                /* 1: x_set starts from 0 and is never read before its
                                  first write (params.x_zero)                */
    double tol;
    double growth_limit;       /* > 0: combined-window growth that closes a window (tail) */
    HB_G const double   *val;
    HB_G const unsigned *col;
    HB_G const unsigned *ptr;
    HB_G const unsigned *blk;  /* SpMV workgroup row starts [nblk+1]       */
    HB_G const short    *col16;/* [nz] col - (first row of the entry's tile), where it fits */
    HB_G const unsigned *cflag;      /* [nblk] 1: the tile's col16 entries are exact */
    HB_G const unsigned char *vidx;  /* value table indices, tile-padded (tpad), or NULL */
    HB_G const short    *col16p;     /* col16 in the same tile-padded layout (with vidx) */
    HB_G const unsigned *tpad;       /* [nblk+1] tile b's entries start at tpad[b], a
                                        multiple of 8, in vidx and col16p            */
    HB_G const unsigned *dtile;      /* [2*nblk] diagonal tiles: {first index byte, K};
                                        K = 0: CSR tile.  NULL: no diagonal tiles    */
    HB_G const int      *doff;       /* [HB_DIAG_K*nblk] diagonals (column - row)    */
    HB_G const unsigned char *didx;  /* index bytes of the diagonal tiles            */
    HB_G const double   *dval;       /* cdia: [HB_DIAG_K*nblk] each diagonal's value  */
    int cdia;                        /* 1: every diagonal tile holds one value per
                                        diagonal; didx holds 2 bytes of presence bits
                                        per row instead of index bytes               */
    HB_G const double   *vdict;      /* [nvdict <= HB_VDICT] the matrix's distinct values */
    int nvdict;
    int npanel;                      /* > 1: the seed SpMVs run as this many column-panel
                                        passes over pval/pcol/pptr                    */
    HB_G const double   *pval;       /* [nz] the entries panel by panel, each panel's  */
    HB_G const unsigned *pcol;       /*      rows in order, a row's entries in CSR order */
    HB_G const unsigned *pptr;       /* [npanel][n+1] row starts of each panel in pval */
    HB_G const unsigned *pblk;       /* each panel's tiles (row starts), binned on that
                                        panel's entries: panel p's [pnblk[p]+1] start
                                        at pblk + pboff[p]                             */
    int pnblk[HB_MAXPANEL];
    unsigned pboff[HB_MAXPANEL];
    int npart;                       /* partial-sum pairs the seed SpMVs write: nblk, or
                                        the last panel's tile count                    */
    int fused_a;                     /* the alpha~ finalize runs as SpMV 2's first
                                        workgroup (fused launches, q formed at the
                                        gathered columns, one pass)                    */
    int fused_s;                     /* the scalar stage runs as the shift kernel's
                                        first workgroup                                */
    int shift_tw;                    /* tile workgroups of that kernel: each walks the
                                        seed p update and the flush tiles grid-stride */
    HB_G const double   *sigma;/* [M]                                      */
    HB_G const double   *rhat; /* [n]   shadow residual r^ = b             */
    HB_G double *rring;        /* [(W+1)][n] residual of iteration k at slot k mod (W+1) */
    HB_G double *qring;        /* [W][n]     q of iteration k at slot k mod W            */
    HB_G double *s, *y;        /* [n] seed SpMV results                    */
    HB_G double *x_set, *p_set;/* [M][n] shift-major                       */
    HB_G double *aa, *ba, *wa; /* seed archives alpha~, beta~, omega~ [cap] */
    HB_G double *pi;           /* [M][cap]                                 */
    HB_G double *eta, *zeta;   /* [M]                                      */
    HB_G double *coef;         /* [W][M][6] vector coefficients by window iteration and list position */
    HB_G double *ccoef;        /* [ceil(M/HB_UG)][2W+5][2*HB_UG] combined window coefficients */
    HB_G double *grow;         /* [2][M] size bound of each shift's combined p, and its
                                  last iteration's own terms */
    HB_G int    *stop;         /* [M]                                      */
    HB_G int    *stop_iter;    /* [M]                                      */
    HB_G int    *act_len;      /* [M] by list position: window iterations that shift was updated in */
    HB_G int    *flist;        /* [M] the window's shift list, ascending   */
    HB_G int    *fpos;         /* [M] position of shift j in the list      */
    HB_G int    *switch_k, *switch_to; /* [M] switch log                   */
    HB_G double *part;         /* workgroup partial sums [max(nblk,vec_grid)*2] of the
                                  initialisation and SpMV 1 */
    HB_G double *part2;        /* the same for SpMV 2 */
    HB_G double *part3;        /* the same for the seed update */
    HB_G double *p_snap;       /* [n] the seed's p at the pending window's start */
    HB_G int *flush_log;       /* [cap] params.profile: 1 where iteration k's shift
                                  kernel applied a window flush, else NULL       */
    HB_G double *trace;        /* [cap][M] estimated residual per iteration and shift,
                                  -1 once stopped (NULL: not recorded)     */
    HB_G hb_state *st;
    HB_G const double *b_in;   /* right-hand side read by the initialisation */
    int seed_in;               /* first seed                               */
    int pad_;
} hb_dev;

/* A batch of independent problems: per-problem hb_dev records in device
 * memory and, per phase, a table of (problem, tile) pairs so that one
 * launch covers the tiles of every problem. */
/* the phases whose launches run over (problem, tile) tables */
enum { HB_PH_SPMV = 0, HB_PH_VEC = 1, HB_PH_SHIFT = 2, HB_PH_COEF = 3, HB_NPH = 4 };

typedef struct hb_batch {
    int nprob, exact, W;
    int vi;                    /* 1: every problem's SpMV reads its value table  */
    int n_spmv_tiles, n_vec_tiles, n_shift_tiles, n_coef_tiles;
    const hb_dev *devs;        /* [nprob] device array                     */
    const int *t_spmv;         /* [2*n_spmv_tiles]  (problem, SpMV tile)   */
    const int *t_vec;          /* [2*n_vec_tiles]   (problem, vector tile) */
    const int *t_shift;        /* [2*n_shift_tiles] (problem, by*gx + bx)  */
    const int *t_coef;         /* [2*n_coef_tiles]  (problem, list position) */
    /* active lists: each table restricted to the problems still running,
     * rebuilt on the device at the start of every chunk (kb_compact); the
     * phase launches walk them grid-stride */
    int *a_tab[HB_NPH];        /* [2 * entries of the full table]          */
    int *act;                  /* [HB_NPH] entries in each active list     */
    const int *pofs;           /* [HB_NPH][nprob] problem p's first entry in each full table */
    const int *pcnt;           /* [HB_NPH][nprob] and its entries          */
    int div;                   /* the phase launches take ceil(full / div) workgroups */
} hb_batch;

#ifdef __cplusplus
}
#endif
#endif
