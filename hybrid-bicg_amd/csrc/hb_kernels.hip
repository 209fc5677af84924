/*
 * hb_kernels.hip -- gfx950 kernels of the shifted BiCG solver with seed
 * switching, and the C-callable launch layer declared in hb_launch.h.
 *
 * One iteration of the reference loop (src/solver_dynamic.c:147-379) is
 * eight launches that keep every scalar on the device:
 *   spmv<SEED1> + fin   s = (A + sigma_s I) p_s; (r^,s) -> alpha~[k]   (:156-167)
 *   spmv<SEED2> + fin   q = r - alpha~ s fused into y = (A + sigma_s I) q;
 *                       (q,q), (q,y) -> omega~[k]                      (:171-190)
 *   seed_update + fin   x_s and r updates, (r,r), (r^,r) -> beta~[k];
 *                       then the scalar stage: per-shift recurrences,
 *                       convergence check, seed switch with replay     (:194-204,
 *                                                                       :239-359)
 *   flush_coef          at a window's end, each shift's combined
 *                       coefficients over the window (double-double)
 *   shift               seed p update, and at a window's end the
 *                       deferred x_j / p_j updates of all shifts       (:208-210,
 *                                                                       :248-255)
 * Per-shift vector updates are deferred over a window of up to W iterations
 * (the window's r and q stay in ring buffers) and applied in one pass that
 * reads and writes x_j and p_j once.  Default mode: x_j and p_j are formed
 * from combined coefficients with fused multiply-adds (they move with the
 * window by roundings only).  Exact order: the reference's six vector
 * operations of every iteration are replayed in their original order with
 * separate roundings, bit-identical to per-iteration updates.
 *
 * Every kernel body is a device function of (problem, tile).  The plain
 * solver launches one problem's tiles; the batched solver launches the
 * tiles of all its problems in the same launch, looked up through a
 * tile -> (problem, tile) table, so each phase is one launch per batch.
 *
 * Every floating-point operation is rounded separately (the library is
 * built with -ffp-contract=off) so that each elementwise step reproduces the
 * reference's own sequence of roundings; the default mode's window flush
 * and double-double coefficients ask for fused multiply-adds explicitly.  Reductions: workgroup partials,
 * combined in a fixed order by a one-workgroup finalize (deterministic).
 * With d.exact set, every dot is summed serially from 0.0 in ascending index
 * order and every SpMV row sequentially, the reference's order
 * (src/openmp_vector.c:68-74, src/openmp_matrix.c:47-58).
 */
#include <hip/hip_runtime.h>
#include <type_traits>
#include <math.h>
#include "hb_shared.h"
#include "hb_launch.h"

#define DEV __device__ __forceinline__

typedef double dbl2 __attribute__((ext_vector_type(2)));

/* Constant address space: the compiler takes what is read through it as
 * unchanged for the rest of the launch, and a wave-uniform address there
 * becomes a scalar load. */
#define HB_K __attribute__((address_space(4)))

/* p, from here to the end of the launch read-only, as constant memory.  The
 * empty asm makes the returned pointer unknown to the compiler until this
 * point, so no read through it moves above what comes before (a wait for
 * the data in the same launch).  Without it a kernel with a barrier or an
 * atomic anywhere before the reads loses their scalar loads: the compiler
 * cannot show that nothing in between stored to them, and makes every
 * uniform read a vector load. */
template <typename T>
DEV const HB_K T *read_only(const T *p)
{
    asm volatile("" : "+s"(p));
    return (const HB_K T *)p;
}

/* ---- reductions ----------------------------------------------------------- */
DEV double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

/* Sums a and b over the workgroup; the result is valid in thread 0. */
DEV void block_sum2(double &a, double &b)
{
    __shared__ double sh[2][HB_BLOCK / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[0][w] = a; sh[1][w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = sh[0][0]; b = sh[1][0];
#pragma unroll
        for (int t = 1; t < HB_BLOCK / 64; ++t) { a += sh[0][t]; b += sh[1][t]; }
    }
    __syncthreads();
}

/* Combines per-workgroup partial pairs in a fixed order (thread 0 result):
 * thread t sums partials t, t + 256, t + 512, ... in ascending order, then
 * the workgroup sum.  32 16-byte pairs are in flight per thread before any
 * is added (the combining workgroup is alone on its CU, so registers are
 * free), so 8192 partials cost one memory latency, not four. */
template <bool DEEP = true>
DEV void reduce_partials(const double *part, int nb, double &a, double &b)
{
    const dbl2 *p2 = (const dbl2 *)part;
    double sa = 0.0, sb = 0.0;
    int i = threadIdx.x;
    /* DEEP: 32 loads in flight per thread (128 registers).  Every depth
     * adds a thread's pairs in ascending order: the sums are the same */
    for (; DEEP && i + 31 * HB_BLOCK < nb; i += 32 * HB_BLOCK) {
        dbl2 v[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) v[u] = p2[i + u * HB_BLOCK];
#pragma unroll
        for (int u = 0; u < 32; ++u) { sa += v[u].x; sb += v[u].y; }
    }
    for (; i + 7 * HB_BLOCK < nb; i += 8 * HB_BLOCK) {
        dbl2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p2[i + u * HB_BLOCK];
#pragma unroll
        for (int u = 0; u < 8; ++u) { sa += v[u].x; sb += v[u].y; }
    }
    for (; i < nb; i += HB_BLOCK) {
        const dbl2 v = p2[i];
        sa += v.x;
        sb += v.y;
    }
    block_sum2(sa, sb);
    a = sa; b = sb;
}

/* sum_i x[i]*y[i] from 0.0 in ascending i (the reference's serial master
 * sum); products staged in LDS, summed by thread 0; result in thread 0. */
DEV double serial_dot(const double *x, const double *y, int n, double *lds)
{
    double acc = 0.0;
    for (int base = 0; base < n; base += HB_NNZ_CAP) {
        const int m = min(HB_NNZ_CAP, n - base);
        for (int t = threadIdx.x; t < m; t += HB_BLOCK) lds[t] = x[base + t] * y[base + t];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int t = 0; t < m; ++t) acc += lds[t];
        __syncthreads();
    }
    return acc;
}

/* ---- values handed between the workgroups of one launch ---------------------
 * A workgroup dispatched first computes a value the launch's other
 * workgroups need.  It stores it with an agent-scope relaxed atomic store (a
 * write-through store, sc1, on gfx950) into a mailbox that holds
 * HB_MAIL_EMPTY until then; the others poll it with agent-scope relaxed
 * atomic loads (sc1), which read past any stale cache line.  Every word
 * that crosses workgroups inside a launch moves this way, so no L2
 * write-back or invalidate (what an agent-scope release or acquire costs)
 * is needed.  The first workgroups of a grid are dispatched before the
 * later ones (in order on each XCD), so the publisher is always running
 * when a waiter spins; the spin is bounded all the same (about 0.5 s), and
 * a waiter that gives up sets HB_ST_SYNC_TIMEOUT and goes on instead of
 * hanging; every later waiter sees the bit and does not wait at all. */
#define HB_SPIN_LIMIT (1 << 17)

/* a waiter polls the status word too: once one has timed out, none waits */
DEV bool sync_timed_out(hb_state *st)
{
    return (__hip_atomic_load(&st->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & HB_ST_SYNC_TIMEOUT) != 0;
}

DEV void mail_put(unsigned long long *slot, unsigned long long v)
{
    __hip_atomic_store(slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

DEV void mail_put_d(unsigned long long *slot, double v)
{
    mail_put(slot, (unsigned long long)__double_as_longlong(v));
}

/* the published word, polled by the calling thread alone (HB_MAIL_EMPTY
 * after a time-out) */
DEV unsigned long long mail_poll_bits(unsigned long long *slot, hb_state *st)
{
    unsigned long long v;
    for (int spins = 0;; ++spins) {
        v = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v != HB_MAIL_EMPTY) break;
        if (spins == HB_SPIN_LIMIT || ((spins & 255) == 255 && sync_timed_out(st))) {
            atomicOr(&st->status, HB_ST_SYNC_TIMEOUT);
            break;
        }
        /* a few hundred cycles between polls: thousands of resident
         * workgroups poll one word, and the publisher needs the memory
         * system more than they do */
        __builtin_amdgcn_s_sleep(8);
    }
    return v;
}

DEV double mail_poll(unsigned long long *slot, hb_state *st)
{
    return __longlong_as_double((long long)mail_poll_bits(slot, st));
}

/* the published double; thread 0 of the workgroup waits, every thread gets it */
DEV double mail_wait(unsigned long long *slot, hb_state *st)
{
    __shared__ double sh_v;
    if (threadIdx.x == 0) sh_v = mail_poll(slot, st);
    __syncthreads();
    return sh_v;
}

/* the published word, to every thread of the workgroup */
DEV unsigned long long mail_wait_bits(unsigned long long *slot, hb_state *st)
{
    __shared__ unsigned long long sh_u;
    if (threadIdx.x == 0) sh_u = mail_poll_bits(slot, st);
    __syncthreads();
    return sh_u;
}

/* a value the kernel's other workgroups read before the launch ends */
template <bool PUB>
DEV void store_pub(double *p, double v)
{
    if (PUB) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

template <bool PUB>
DEV void store_pub(int *p, int v)
{
    if (PUB) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

/* a word another workgroup of the launch stored with store_pub<true> */
DEV int load_pub(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

DEV size_t off(int j, int n) { return (size_t)j * (size_t)n; }
#define PI(d, j, i) ((d).pi[(size_t)(j) * (size_t)(d).cap + (size_t)(i)])

/* ---- SpMV ------------------------------------------------------------------
 * Tile b owns rows [blk[b], blk[b+1]).  When their nonzeros fit in LDS
 * (<= HB_NNZ_CAP, <= HB_ROWS_CAP rows), the workgroup loads val/col
 * coalesced, gathers x, stores the products in LDS and reduces each row with
 * T lanes, T the largest power of two with T*rows <= 256 (at most 64).
 * Stencil rows give 256 rows per tile and T = 1: each row is summed
 * sequentially in CSR order.  A row longer than HB_NNZ_CAP gets the whole
 * workgroup.  exact forces T = 1 and sequential long rows.
 * The arrays come in as __restrict__ parameters (kept as noalias scopes when
 * inlined): the compiler keeps several val/col/x loads in flight and reads
 * uniform values through scalar loads.  SEED1 gathers p_set[seed], SEED2
 * gathers r_{k-1} (in rring) and s, PLAIN gathers x. */
enum { SP_SEED1 = 0, SP_SEED2 = 1, SP_PLAIN = 2 };

/* Column panels (matrices whose columns are scattered over an x larger than
 * one XCD's L2): the SpMV runs as one launch per panel of columns, each over
 * the entries of every row whose columns lie in that panel, so the gathers
 * of all the workgroups on an XCD hit the same few MB of x.  Each row's sum
 * is carried from pass to pass in the output vector: the first pass stores
 * it, the middle ones read and store it, the last reads it and runs the
 * epilogue.  A row's entries are column-ascending, so the panels take them
 * in CSR order, and a row summed by one lane (exact order) is summed in
 * exactly the order of one pass. */
enum { PASS_ONE = 0, PASS_FIRST = 1, PASS_MID = 2, PASS_LAST = 3 };

template <int MODE, bool EXACT, bool VI = false, bool UNF = false, int PASS = PASS_ONE, bool FA = false,
          bool CD = false>
DEV void spmv_body(const hb_dev &d, unsigned b, const double *__restrict__ val,
                   const unsigned *__restrict__ col, const short *__restrict__ col16,
                   const unsigned *__restrict__ cflag, const unsigned *__restrict__ ptr,
                   const unsigned *__restrict__ blk, const double *__restrict__ xbase,
                   double *__restrict__ rring, const double *__restrict__ svec,
                   const double *__restrict__ rhat, double *__restrict__ out1,
                   double *__restrict__ qring, double *__restrict__ part, double shift,
                   const unsigned char *__restrict__ vidx = nullptr, const short *__restrict__ col16p = nullptr,
                   const unsigned *__restrict__ tpad = nullptr, const double *__restrict__ vdict = nullptr,
                   const unsigned *__restrict__ dtile = nullptr, const int *__restrict__ doff = nullptr,
                   const unsigned char *__restrict__ didx = nullptr, const double *__restrict__ dval = nullptr)
{
    __shared__ double prod[VI ? 1 : HB_NNZ_CAP + 1];
    __shared__ __attribute__((aligned(16))) short scol[HB_NNZ_CAP + 8];
    __shared__ __attribute__((aligned(16))) unsigned char sidx[VI ? HB_NNZ_CAP + 8 : 1];
    __shared__ unsigned rp[HB_ROWS_CAP + 1];
    const int n = d.n;
    int switched = 0;
    bool snap = false;   /* SEED1: keep p_s as the window-start p of the seed */
    double sig, zscale = 1.0, nalpha = 0.0;
    const double *xa;
    double *rprev = nullptr, *qout = nullptr;
    if (MODE != SP_PLAIN) {
        hb_state *st = d.st;
        /* the shift kernel runs once per scalar stage: it compares k with
         * the epoch written here, after the previous shift kernel finished */
        /* every state word is read before that store, so the loads issue
         * together (a store through st would order the reads after it) */
        const int done = st->done, k = st->k, seed = st->seed;
        const int sw = MODE == SP_SEED1 ? st->switched : 0;
        const double zs = MODE == SP_SEED1 ? st->zscale : 1.0;
        const int wk0 = MODE == SP_SEED1 ? st->win_k0 : 0;
        if (MODE == SP_SEED1 && b == 0 && threadIdx.x == 0) {
            st->k4_epoch = k;
            /* the shift kernel's mailboxes and coefficient count: its
             * readers finished with the last launch of it */
            st->mail_b = HB_MAIL_EMPTY;
            st->mail_f = HB_MAIL_EMPTY;
            st->coef_ready = 0;
            st->pu_live = 0;
        }
        /* k >= 1, seed >= 0 and switched >= 0 always hold; testing them
         * keeps their loads ahead of the branch, in one round trip with done */
        if (done | (k < 1) | (seed < 0) | (sw < 0)) return;
        sig = d.sigma[seed];
        rprev = rring + off((k - 1) % (d.W + 1), n);
        if (MODE == SP_SEED1) {
            xa = xbase + off(seed, n);
            switched = sw;
            zscale = zs;
            snap = !d.exact && k == wk0;
        } else {
            /* FA: alpha~ comes from the launch's first workgroup, read once
             * the tile's own loads are in flight */
            if (!FA) nalpha = -d.aa[k];
            qout = qring + off(k % d.W, n);
            /* UNF: q was formed for every row by k_qform; gather it */
            xa = UNF ? qout : rprev;
        }
    } else {
        sig = shift;
        xa = xbase;
    }
    auto xval = [&](unsigned c) -> double {
        if (MODE == SP_SEED2 && !UNF) return xa[c] + nalpha * svec[c];
        return xa[c];
    };

    const unsigned r0 = blk[b], r1 = blk[b + 1];
    const unsigned narrow = cflag ? __builtin_amdgcn_readfirstlane(cflag[b]) : 0u;
    /* the row sums carried in from the previous panels, and whether this
     * pass ends them (epilogue) or hands them on */
    constexpr bool CARRY_IN = PASS == PASS_MID || PASS == PASS_LAST;
    constexpr bool ENDS = PASS == PASS_ONE || PASS == PASS_LAST;
    const int R = (int)(r1 - r0);
    const unsigned e0 = ptr[r0];
    const int N = (int)(ptr[r1] - e0);
    const int tid = threadIdx.x;
    double part0 = 0.0, part1 = 0.0;

    /* Row i's result t: the operands a, bb, c of the row's own entries
     * (SEED1: x_i, r^_i and, after a switch, r_i; SEED2: r_i, s_i; PLAIN:
     * x_i) are loaded before the tile's reduction, so the epilogue waits on
     * no memory. */
    auto epilogue = [&](unsigned i, double t, double a, double bb, double c) {
        if (MODE == SP_SEED1) {
            const double si = t + sig * a;
            out1[i] = si;
            if (snap) d.p_snap[i] = a;
            part0 += bb * si;
            if (switched) {
                const double rr = c * zscale;
                rprev[i] = rr;
                part1 += bb * rr;
            }
        } else if (MODE == SP_SEED2) {
            /* UNF: a is q_i, formed by k_qform as r_i + (-alpha) s_i */
            const double qi = UNF ? a : a + nalpha * bb;
            if (!UNF) qout[i] = qi;
            const double yi = t + sig * qi;
            out1[i] = yi;
            part0 += qi * qi;
            part1 += qi * yi;
        } else {
            out1[i] = t + sig * a;
        }
    };
    auto finish = [&](unsigned i, double t, double a, double bb, double c) {
        if (ENDS) epilogue(i, t, a, bb, c);
        else out1[i] = t;
    };
    auto operands = [&](unsigned i, double &a, double &bb, double &c) {
        if (!ENDS) return;
        if (MODE == SP_SEED1) {
            a = xa[i];
            bb = rhat[i];
            c = switched ? rprev[i] : 0.0;
        } else if (MODE == SP_SEED2) {
            a = xa[i];
            bb = UNF ? 0.0 : svec[i];
        } else {
            a = xa[i];
        }
    };

    /* VI: the matrix is read as one byte per entry into its table of
     * distinct values (at most HB_VDICT, staged in LDS) instead of eight
     * bytes per entry; the products are the same doubles */
    __shared__ double vd[VI ? HB_VDICT : 1];
    /* VI layout: tile b's indices and 16-bit columns start at tpad[b] (a
     * multiple of 8), so a thread fetches 8 entries of each with one 8- and
     * one 16-byte load.  Stencil tiles (one row per thread) stage them in LDS
     * and walk their rows there; other tiles read them where they lie, T
     * lanes per row.  Products and their summation order are those of the
     * 8-byte path, so results are bitwise the same. */
    /* Diagonal tile (more than HB_BLOCK/2 rows, so T = 1 as in the CSR
     * form): thread tid owns row r0 + tid, whose index bytes (one per
     * diagonal, HB_DIAG_NONE where the row has no entry) come in one 8- or
     * 16-byte load; column = row + diagonal.  The row sum skips the missing
     * diagonals and runs over the others in ascending order, i.e. in CSR
     * order.  Eight diagonals at a time, every gather is issued before the
     * value table is staged (its barrier then overlaps their latency), and
     * the sum selects rather than branches, so no load waits inside a branch.
     * vdv: this thread's value-table entry, loaded at kernel start. */
    auto tile_dia = [&](unsigned dk, unsigned dbase, double vdv) {
        const unsigned i_me = r0 + (unsigned)min(tid, R - 1);
        double oa = 0.0, ob = 0.0, oc = 0.0;
        operands(i_me, oa, ob, oc);
        const int *__restrict__ offs = doff + (size_t)HB_DIAG_K * b;
        int o[HB_DIAG_K];
#pragma unroll
        for (int j = 0; j < HB_DIAG_K; ++j) o[j] = offs[j];
        /* CD (every diagonal tile holds one value per diagonal): the
         * diagonals' values come by scalar loads issued with the offsets (the
         * same round trip), the row's two bytes of presence bits replace its
         * index bytes, and there is no value table stage */
        const double *__restrict__ dv = CD ? dval + (size_t)HB_DIAG_K * b : nullptr;
        double dvr[CD ? 8 : 1];
        if constexpr (CD) {
#pragma unroll
            for (int j = 0; j < 8; ++j) dvr[j] = dv[j];
        }
        unsigned w[4];
        if constexpr (CD) {
            w[0] = *(const unsigned short *)(didx + dbase + 2u * (unsigned)min(tid, R - 1));
            w[1] = w[2] = w[3] = 0u;
        } else if (dk <= 8) {
            typedef unsigned u2 __attribute__((ext_vector_type(2)));
            const u2 v = *(const u2 *)(didx + dbase + 8u * (unsigned)min(tid, R - 1));
            w[0] = v.x; w[1] = v.y; w[2] = 0xffffffffu; w[3] = 0xffffffffu;
        } else {
            typedef unsigned u4 __attribute__((ext_vector_type(4)));
            const u4 v = *(const u4 *)(didx + dbase + 16u * (unsigned)min(tid, R - 1));
            w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
        }
        double xv[8];
        /* the gather address does not wait for the index bytes: row +
         * diagonal, clamped into [0, n) for a missing diagonal or a padding
         * slot (diagonal 0); rows of a tile lie within HB_ROWS_CAP of each
         * other, so the sum cannot overflow */
        auto gather8 = [&](int j0) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int cs = (int)i_me + o[j0 + j];
                xv[j] = xval((unsigned)min(max(cs, 0), n - 1));
            }
        };
        /* FA: the first eight gathers load r and s apart; alpha~ arrives
         * while they are in flight, and q = r - alpha~ s is formed after */
        double sv8[FA ? 8 : 1];
        auto gather8_rs = [&]() {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const unsigned c = (unsigned)min(max((int)i_me + o[j], 0), n - 1);
                xv[j] = xa[c];
                sv8[FA ? j : 0] = svec[c];
            }
        };
        double acc = 0.0;
        auto sum8 = [&](int h) {
            if constexpr (CD) {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    /* diagonals 8..15 (tiles of more than 8) load theirs here */
                    const double t = acc + (h ? dv[8 + j] : dvr[j]) * xv[j];
                    acc = (w[0] >> (8 * h + j)) & 1u ? t : acc;
                }
                return;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const unsigned ix = (w[2 * h + j / 4] >> (8 * (j % 4))) & 255u;
                const double t = acc + vd[ix] * xv[j];
                acc = ix != HB_DIAG_NONE ? t : acc;
            }
        };
        if constexpr (FA) {
            __shared__ double sh_na;
            gather8_rs();
            if (!CD && tid < d.nvdict) vd[tid] = vdv;
            if (tid == 0) sh_na = -mail_poll(&d.st->mail_a, d.st);
            __syncthreads();
            nalpha = sh_na;
#pragma unroll
            for (int j = 0; j < 8; ++j) xv[j] = xv[j] + nalpha * sv8[j];
        } else {
            gather8(0);
            if constexpr (!CD) {
                if (tid < d.nvdict) vd[tid] = vdv;
                __syncthreads();
            }
        }
        sum8(0);
        if (dk > 8) {
            gather8(8);
            sum8(1);
        }
        if (tid < R) epilogue(i_me, acc, oa, ob, oc);
    };
    auto tile_vi = [&]() {
        const unsigned char *__restrict__ vidxp = vidx;
        const unsigned pb = tpad[b];
        if (N <= HB_NNZ_CAP) {
            int T = 1;
            if (!EXACT) {
                const int q = HB_BLOCK / R;
                while ((T << 1) <= q && (T << 1) <= 64) T <<= 1;
            }
            const int row = tid / T, lane = tid & (T - 1);
            const unsigned i_me = r0 + (unsigned)min(row, R - 1);
            double oa = 0.0, ob = 0.0, oc = 0.0;
            operands(i_me, oa, ob, oc);
            if (T == 1 && narrow) {
                if (8 * tid < N) {
                    typedef unsigned u2 __attribute__((ext_vector_type(2)));
                    typedef int i4 __attribute__((ext_vector_type(4)));
                    const u2 v8 = *(const u2 *)(vidxp + pb + 8 * tid);
                    const i4 c8 = *(const i4 *)(col16p + pb + 8 * tid);
                    *(u2 *)(sidx + 8 * tid) = v8;
                    *(i4 *)(scol + 8 * tid) = c8;
                }
                for (int t = tid; t <= R; t += HB_BLOCK) rp[t] = ptr[r0 + t] - e0;
                __syncthreads();
                if (row < R) {
                    constexpr int CH = 4;
                    const int lo = (int)rp[row], hi = (int)rp[row + 1];
                    double acc = 0.0;
                    for (int base = lo; base < hi; base += CH) {
                        double vv[CH], xv[CH];
#pragma unroll
                        for (int j = 0; j < CH; ++j) {
                            const int e = min(base + j, hi - 1);
                            vv[j] = vd[sidx[e]];
                            xv[j] = xval(r0 + (unsigned)(int)scol[e]);
                        }
#pragma unroll
                        for (int j = 0; j < CH; ++j)
                            if (base + j < hi) acc += vv[j] * xv[j];
                    }
                    epilogue(i_me, acc, oa, ob, oc);
                }
            } else if (row < R) {
                const int lo = (int)(ptr[i_me] - e0), hi = (int)(ptr[i_me + 1] - e0);
                double acc = 0.0;
                for (int e = lo + lane; e < hi; e += T) {
                    const unsigned cc = narrow ? r0 + (unsigned)(int)col16p[pb + e] : col[e0 + e];
                    acc += vd[vidxp[pb + e]] * xval(cc);
                }
                for (int o = T >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
                if (lane == 0) epilogue(i_me, acc, oa, ob, oc);
            }
        } else {
            /* one row longer than the LDS stage */
            double acc = 0.0;
            if (EXACT) {
                if (tid == 0)
                    for (int e = 0; e < N; ++e) acc += vd[vidxp[pb + e]] * xval(col[e0 + e]);
            } else {
                for (int e = tid; e < N; e += HB_BLOCK) acc += vd[vidxp[pb + e]] * xval(col[e0 + e]);
                double dummy = 0.0;
                block_sum2(acc, dummy);
            }
            if (tid == 0) {
                double oa = 0.0, ob = 0.0, oc = 0.0;
                operands(r0, oa, ob, oc);
                epilogue(r0, acc, oa, ob, oc);
            }
        }
    };
    auto tile = [&]() {
        auto vat = [&](unsigned e) -> double { return val[e]; };
        if (N <= HB_NNZ_CAP) {
            /* rows reduced by T lanes each, T the largest power of two with
             * T*R <= HB_BLOCK (at most 64; 1 in exact order): every row has one
             * group of lanes, and lane 0 of the group owns its epilogue */
            int T = 1;
            if (!EXACT) {
                const int q = HB_BLOCK / R;
                while ((T << 1) <= q && (T << 1) <= 64) T <<= 1;
            }
            const int row = tid / T, lane = tid & (T - 1);
            const unsigned i_me = r0 + (unsigned)min(row, R - 1);
            double oa = 0.0, ob = 0.0, oc = 0.0;
            operands(i_me, oa, ob, oc);
            const double cin = CARRY_IN ? out1[i_me] : 0.0;

            constexpr int U = HB_NNZ_CAP / HB_BLOCK;
            static_assert(U * HB_BLOCK == HB_NNZ_CAP, "tile capacity");
            if (MODE == SP_SEED2 && !UNF && T == 1 && narrow && N > 0) {
                /* One row per thread.  Values and 16-bit column offsets are
                 * staged in LDS, then each thread walks its own row: the lanes of
                 * a wave gather the same stencil slot of consecutive rows, i.e.
                 * consecutive addresses, where the entry-per-lane order touches
                 * several cache lines per instruction.  Two gathers per entry
                 * (r and s), four entries in flight.  The row sum runs in CSR
                 * order with the products rounded on their own, as in the LDS
                 * products path below. */
    #pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int e = tid + u * HB_BLOCK;
                    const unsigned ec = e0 + (unsigned)min(e, N - 1);
                    const double vv = vat(ec);
                    const short cc = col16[ec];
                    prod[e < N ? e : HB_NNZ_CAP] = vv;
                    scol[e < N ? e : HB_NNZ_CAP] = cc;
                }
                for (int t = tid; t <= R; t += HB_BLOCK) rp[t] = ptr[r0 + t] - e0;
                __syncthreads();
                if (row < R) {
                    constexpr int CH = 4;
                    const int lo = (int)rp[row], hi = (int)rp[row + 1];
                    double acc = cin;
                    for (int base = lo; base < hi; base += CH) {
                        double vv[CH], xv[CH];
    #pragma unroll
                        for (int j = 0; j < CH; ++j) {
                            const int e = min(base + j, hi - 1);
                            vv[j] = prod[e];
                            xv[j] = xval(r0 + (unsigned)(int)scol[e]);
                        }
    #pragma unroll
                        for (int j = 0; j < CH; ++j)
                            if (base + j < hi) acc += vv[j] * xv[j];
                    }
                    finish(i_me, acc, oa, ob, oc);
                }
                return;
            }
            if (MODE == SP_SEED2 && !UNF) {
                /* two gathers per entry (r and s): entries in rounds of four */
                if (narrow) {
    #pragma unroll 4
                    for (int e = tid; e < N; e += HB_BLOCK)
                        prod[e] = vat(e0 + e) * xval(r0 + (unsigned)(int)col16[e0 + e]);
                } else {
    #pragma unroll 4
                    for (int e = tid; e < N; e += HB_BLOCK) prod[e] = vat(e0 + e) * xval(col[e0 + e]);
                }
            } else if (N > 0) {
                /* A thread owns entries tid + u*HB_BLOCK, u < U.  All of its
                 * column and value loads are issued together (indices clamped
                 * into the tile, so no load is predicated), then all of its
                 * gathers: two memory latencies per tile, not two per entry. */
                double v[U];
                unsigned c[U];
                if (narrow) {
                    /* 16-bit column offsets from the tile's first row: 2 bytes
                     * per entry instead of 4 */
    #pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const unsigned e = e0 + (unsigned)min(tid + u * HB_BLOCK, N - 1);
                        c[u] = r0 + (unsigned)(int)col16[e];
                        v[u] = vat(e);
                    }
                } else {
    #pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const unsigned e = e0 + (unsigned)min(tid + u * HB_BLOCK, N - 1);
                        c[u] = col[e];
                        v[u] = vat(e);
                    }
                }
                /* entries past the tile end go to the spare slot prod[HB_NNZ_CAP]
                 * (no branch, so the gathers are not serialised behind one) */
    #pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int e = tid + u * HB_BLOCK;
                    prod[e < N ? e : HB_NNZ_CAP] = v[u] * xval(c[u]);
                }
            }
            for (int t = tid; t <= R; t += HB_BLOCK) rp[t] = ptr[r0 + t] - e0;
            __syncthreads();
            if (row < R) {
                /* lane 0 starts from the sum carried in: with one lane per row
                 * the row is summed in CSR order across the panels */
                double acc = lane == 0 ? cin : 0.0;
                const int hi = (int)rp[row + 1];
                for (int e = (int)rp[row] + lane; e < hi; e += T) acc += prod[e];
                for (int o = T >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
                if (lane == 0) finish(i_me, acc, oa, ob, oc);
            }
        } else {
            /* one row longer than the LDS stage */
            double acc = 0.0;
            if (EXACT) {
                if (tid == 0) {
                    if (CARRY_IN) acc = out1[r0];
                    for (int e = 0; e < N; ++e) acc += vat(e0 + e) * xval(col[e0 + e]);
                }
            } else {
                for (int e = tid; e < N; e += HB_BLOCK) acc += vat(e0 + e) * xval(col[e0 + e]);
                double dummy = 0.0;
                block_sum2(acc, dummy);
                if (CARRY_IN && tid == 0) acc = out1[r0] + acc;
            }
            if (tid == 0) {
                double oa = 0.0, ob = 0.0, oc = 0.0;
                operands(r0, oa, ob, oc);
                finish(r0, acc, oa, ob, oc);
            }
        }
    };
    static_assert(!VI || PASS == PASS_ONE, "column panels read the 8-byte values");
    if constexpr (VI) {
        const double vdv = tid < d.nvdict ? vdict[tid] : 0.0;
        unsigned dk = 0, dbase = 0;
        if (dtile) {
            dk = dtile[2 * b + 1];
            dbase = dtile[2 * b];
        }
        if (dk) {
            tile_dia(dk, dbase, vdv);
        } else {
            if (FA) nalpha = -mail_wait(&d.st->mail_a, d.st);
            if (tid < d.nvdict) vd[tid] = vdv;
            __syncthreads();
            tile_vi();
        }
    } else {
        if (FA) nalpha = -mail_wait(&d.st->mail_a, d.st);
        tile();
    }

    if (MODE == SP_PLAIN || !ENDS) return;

    block_sum2(part0, part1);
    if (tid == 0) { part[2 * b] = part0; part[2 * b + 1] = part1; }
}

template <int MODE, bool EXACT, bool VI = false, bool UNF = false, bool CD = false>
DEV void spmv_tile(const hb_dev &d, unsigned b)
{
    if (MODE == SP_SEED1)
        spmv_body<MODE, EXACT, VI, false, PASS_ONE, false, CD>(d, b, d.val, d.col, d.col16, d.cflag, d.ptr, d.blk, d.p_set, d.rring, nullptr,
                                   d.rhat, d.s, nullptr, d.part, 0.0, d.vidx, d.col16p, d.tpad, d.vdict, d.dtile,
                                   d.doff, d.didx, d.dval);
    else
        spmv_body<MODE, EXACT, VI, UNF, PASS_ONE, false, CD>(d, b, d.val, d.col, d.col16, d.cflag, d.ptr, d.blk,
                                                             nullptr, d.rring, d.s, d.rhat, d.y, d.qring, d.part2,
                                                             0.0, d.vidx, d.col16p, d.tpad, d.vdict, d.dtile,
                                                             d.doff, d.didx, d.dval);
}

/* Workgroup b runs on XCD b % 8 (the dispatcher deals workgroups round
 * robin over the eight XCDs, each with its own L2).  Tile order: XCD x gets
 * the contiguous tiles [start(x), start(x+1)), so a tile's column halo (the
 * rows around it, and the rows a plane away for stencils) was mostly
 * fetched into the same L2 by its neighbours.  A bijection on 0 .. nb-1. */
DEV unsigned xcd_tile(unsigned b, unsigned nb)
{
    const unsigned q = nb / 8, r = nb % 8, x = b % 8;
    return x * q + min(x, r) + b / 8;
}

template <int MODE, bool EXACT, bool VI, bool UNF = false, bool CD = false>
__global__ __launch_bounds__(HB_BLOCK) void k_spmv(hb_dev d)
{
    spmv_tile<MODE, EXACT, VI, UNF, CD>(d, xcd_tile(blockIdx.x, gridDim.x));
}

/* Pass of column panel `panel` (of d.npanel): the panel's entries of every
 * row, in the panel-major copy of the matrix (pval, pcol; row pointers
 * pptr + panel (n + 1)), over the panel's own tiles (up to HB_ROWS_CAP rows
 * and HB_NNZ_CAP of the panel's entries each).  The last pass's tiles write
 * the partial sums (d.npart of them). */
template <int MODE, bool EXACT, bool UNF, int PASS>
__global__ __launch_bounds__(HB_BLOCK) void k_spmv_panel(hb_dev d, int panel)
{
    const unsigned b = xcd_tile(blockIdx.x, gridDim.x);
    const unsigned *pp = d.pptr + (size_t)panel * ((size_t)d.n + 1);
    const unsigned *pb = d.pblk + d.pboff[panel];
    if (MODE == SP_SEED1)
        spmv_body<MODE, EXACT, false, false, PASS>(d, b, d.pval, d.pcol, nullptr, nullptr, pp, pb, d.p_set,
                                                   d.rring, nullptr, d.rhat, d.s, nullptr, d.part, 0.0);
    else
        spmv_body<MODE, EXACT, false, UNF, PASS>(d, b, d.pval, d.pcol, nullptr, nullptr, pp, pb, nullptr,
                                                 d.rring, d.s, d.rhat, d.y, d.qring, d.part2, 0.0);
}

/* q = r_{k-1} - alpha~ s for every row (matrices whose SpMV 2 would gather r
 * and s at scattered columns: it then gathers q alone).  The same expression
 * as the fused form, so q and everything after it are bitwise the same.
 * src/solver_dynamic.c:171-176. */
__global__ __launch_bounds__(HB_BLOCK) void k_qform(hb_dev d)
{
    const hb_state *st = d.st;
    const int done = st->done, k = st->k;
    if (done | (k < 1)) return;
    const int n = d.n;
    const double nalpha = -d.aa[k];
    const double *__restrict__ r = d.rring + off((k - 1) % (d.W + 1), n);
    const double *__restrict__ s = d.s;
    double *__restrict__ q = d.qring + off(k % d.W, n);
    for (int i = blockIdx.x * HB_BLOCK + threadIdx.x; i < n; i += gridDim.x * HB_BLOCK)
        q[i] = r[i] + nalpha * s[i];
}

/* p[0 .. count) = v (a fill kernel: usable where a memset cannot express v) */
__global__ __launch_bounds__(HB_BLOCK) void k_fill(double *p, size_t count, double v)
{
    for (size_t i = (size_t)blockIdx.x * HB_BLOCK + threadIdx.x; i < count;
         i += (size_t)gridDim.x * HB_BLOCK)
        p[i] = v;
}

/* Tile b of the SpMV: col16[e] = col[e] - (tile's first row) for its
 * entries, and cflag[b] = 1 when every such offset fits in 16 bits (banded
 * and stencil matrices; the SpMV then reads 2 bytes per column index).
 * Built once per plan. */
__global__ __launch_bounds__(HB_BLOCK) void k_compress_cols(const unsigned *__restrict__ ptr,
                                                            const unsigned *__restrict__ col,
                                                            const unsigned *__restrict__ blk,
                                                            short *__restrict__ col16,
                                                            unsigned *__restrict__ cflag)
{
    const unsigned b = blockIdx.x;
    const unsigned r0 = blk[b], e0 = ptr[r0], e1 = ptr[blk[b + 1]];
    int ok = 1;
    for (unsigned e = e0 + threadIdx.x; e < e1; e += HB_BLOCK) {
        const long long o = (long long)col[e] - (long long)r0;
        if (o < -32768 || o > 32767) ok = 0;
        else col16[e] = (short)o;
    }
    ok = __syncthreads_and(ok);
    if (threadIdx.x == 0) cflag[b] = (unsigned)ok;
}

template <bool EXACT>
__global__ __launch_bounds__(HB_BLOCK) void k_spmv_plain(hb_dev d, const double *x, double *y,
                                                         double shift)
{
    spmv_body<SP_PLAIN, EXACT>(d, blockIdx.x, d.val, d.col, d.col16, d.cflag, d.ptr, d.blk, x, nullptr, nullptr,
                               nullptr, y, nullptr, nullptr, shift);
}

/* Combines the SpMV partials and forms alpha~[k] (after SpMV 1,
 * src/solver_dynamic.c:163-167, with the post-switch (r^,r) of :323) or
 * omega~[k] (after SpMV 2, :183-190).  One workgroup per problem: on gfx950
 * an agent-scope release is an L2 write-back, far too costly to pay once per
 * SpMV workgroup for a last-to-arrive tail. */
template <int MODE, bool EXACT, bool PUB = false>
DEV void fin_spmv_body(const hb_dev &d)
{
    __shared__ double lds[EXACT ? HB_NNZ_CAP : 1];
    hb_state *st = d.st;
    if (st->done) return;
    /* PUB: this workgroup runs first in the next phase's launch (SpMV 2
     * for alpha~, the seed update for omega~) and hands the value to that
     * launch's other workgroups.  The mailbox is empty already (the scalar
     * stage empties it); emptying it again here keeps a back-to-back repeat
     * of the launch waiting too. */
    unsigned long long *mail = MODE == SP_SEED1 ? &st->mail_a : &st->mail_w;
    if (PUB && threadIdx.x == 0) mail_put(mail, HB_MAIL_EMPTY);
    const int k = st->k, tid = threadIdx.x;
    const int switched = (MODE == SP_SEED1) ? st->switched : 0;
    const double rTr = st->rTr;
    double a0, a1;
    if (EXACT) {
        const double *rprev = d.rring + off((k - 1) % (d.W + 1), d.n);
        const double *q = d.qring + off(k % d.W, d.n);
        if (MODE == SP_SEED1) {
            a0 = serial_dot(d.rhat, d.s, d.n, lds);
            a1 = switched ? serial_dot(d.rhat, rprev, d.n, lds) : 0.0;
        } else {
            a0 = serial_dot(q, q, d.n, lds);
            a1 = serial_dot(q, d.y, d.n, lds);
        }
    } else {
        reduce_partials(MODE == SP_SEED1 ? d.part : d.part2, d.npart, a0, a1);
    }
    if (tid == 0) {
        if (MODE == SP_SEED1) {
            const double rtr = switched ? a1 : rTr;
            const double alpha = rtr / a0;
            if (switched) { st->rTr = a1; st->switched = 0; }
            st->rTs = a0;
            d.aa[k] = alpha;
            if (a0 == 0.0 || !isfinite(alpha)) st->status |= HB_ST_BREAKDOWN_RTS;
            if (PUB) mail_put_d(mail, alpha);
        } else {
            const double omega = a0 / a1;
            st->qTq = a0;
            st->qTy = a1;
            st->pu_w = omega;
            st->pu_k = k;
            st->pu_seed = st->seed;
            st->pu_live = 1;
            d.wa[k] = omega;
            st->rTr_old = rTr;
            if (a1 == 0.0 || !isfinite(omega)) st->status |= HB_ST_BREAKDOWN_QTY;
            if (PUB) mail_put_d(mail, omega);
        }
    }
}

template <int MODE, bool EXACT>
__global__ __launch_bounds__(HB_BLOCK) void k_fin_spmv(hb_dev d)
{
    fin_spmv_body<MODE, EXACT>(d);
}

/* Fused SpMV 2 (default mode, q formed at the gathered columns): workgroup
 * 0 finalizes alpha~ from SpMV 1's partials and publishes it; the others
 * are the tiles, which issue their loads and gathers first and read alpha~
 * after (src/solver_dynamic.c:163-184). */
template <bool VI, bool CD = false>
__global__ __launch_bounds__(HB_BLOCK) void k_spmv2_fa(hb_dev d)
{
    if (blockIdx.x == 0) {
        fin_spmv_body<SP_SEED1, false, true>(d);
        return;
    }
    const unsigned b = xcd_tile(blockIdx.x - 1, gridDim.x - 1);
    spmv_body<SP_SEED2, false, VI, false, PASS_ONE, true, CD>(d, b, d.val, d.col, d.col16, d.cflag, d.ptr, d.blk,
                                                         nullptr, d.rring, d.s, d.rhat, d.y, d.qring, d.part2,
                                                         0.0, d.vidx, d.col16p, d.tpad, d.vdict, d.dtile,
                                                         d.doff, d.didx, d.dval);
}

/* ---- scalar stage (one workgroup per problem, after the seed update) -------
 * src/solver_dynamic.c:204 (beta~), :243-253 (per-shift recurrences),
 * :278-303 (convergence check), :312-341 (seed switch).
 * PUB: the stage runs as the first workgroup of the shift kernel.  It hands
 * beta~ to the seed p update at once; at its end it stores again, with
 * write-through stores, what the launch's flush reads (this iteration's
 * coefficients and update counts, the list and the window's extent), then
 * publishes the window decision.  (Atomic stores inside the stage's body
 * would cost it about 30 registers: the compiler stops keeping its uniform
 * loads in scalar registers.) */
/* The stage's inputs from earlier launches.  The state words are read with
 * the done test, the archive entries they index next, both before the
 * seed-update partials are reduced, so their round trips overlap. */
struct hb_stage_in {
    int k, s, win_k0, win_seedpos;
    int stop_count, max_sigma, status, n_switch, flush_cnt;
    double rTr_old, dot_zero, dot_r_prev, qTq, shift_bytes;
    double ak, wk, bkm1, akm1, sig_s;
    int stop_s;   /* the seed's stop flag (0 while it runs) */
};

template <bool PUB = false>
DEV void scalar_stage(const hb_dev &d, const hb_stage_in &in, double dot_r, double rTr)
{
    __shared__ double sh_beta;
    __shared__ double sh_zp[HB_BLOCK];
    __shared__ int sh_j[HB_BLOCK];
    __shared__ int sh_nc[HB_BLOCK];
    __shared__ int sh_switch, sh_spike, sh_flush, sh_xonly, sh_bad, sh_upd, sh_sc, sh_ms, sh_nst;
    hb_state *st = d.st;
    const int tid = threadIdx.x;
    const int k = in.k, s = in.s, M = d.M;
    const int win_k0 = in.win_k0, win_seedpos = in.win_seedpos;
    const double rTr_old = in.rTr_old, dot_zero = in.dot_zero;
    const double ak = in.ak, wk = in.wk, bkm1 = in.bkm1, akm1 = in.akm1;
    const double sig_s = in.sig_s;

    if (tid == 0) {
        const double beta = (ak / wk) * (rTr / rTr_old);
        if (PUB) mail_put_d(&st->mail_b, beta);
        sh_spike = 0;
        sh_bad = 0;
        st->dot_r = dot_r;
        st->rTr = rTr;
        d.ba[k] = beta;
        st->pb = beta;
        st->pbw = -beta * wk;
        sh_beta = beta;
    }
    __syncthreads();
    const double beta = sh_beta;
    const double ratio = bkm1 / akm1;
    const double thr = (d.tol * d.tol) * dot_zero;
    const int slot = k - win_k0;
    double *coef = d.coef + (size_t)slot * (size_t)M * 6;
    int seedpos = slot == 0 ? -1 : win_seedpos;
    int cnt_now = in.flush_cnt;   /* the window's list length */

    if (slot == 0) {
        /* A window opens: its shift list is every shift updated in the
         * window's first iteration (not stopped, not the seed), ascending.
         * The set only shrinks inside a window (a seed switch closes it), so
         * list positions stay valid until the flush; coefficients and
         * update counts are stored by position, contiguous for the shift
         * kernel's shifts in flight.  The default mode lists the seed first
         * (below), the shifts after it. */
        int base = d.exact ? 0 : 1;
        for (int c0 = 0; c0 < M; c0 += HB_BLOCK) {
            const int j = c0 + tid;
            const int f = (j < M && !d.stop[j] && j != s) ? 1 : 0;
            sh_nc[tid] = f;
            __syncthreads();
            for (int o = 1; o < HB_BLOCK; o <<= 1) {
                const int v = tid >= o ? sh_nc[tid - o] : 0;
                __syncthreads();
                sh_nc[tid] += v;
                __syncthreads();
            }
            if (f) {
                d.flist[base + sh_nc[tid] - 1] = j;
                d.fpos[j] = base + sh_nc[tid] - 1;
            }
            base += sh_nc[HB_BLOCK - 1];
            __syncthreads();
        }
        /* default mode: the seed's own x update (x_s += alpha~ p_s +
         * omega~ q, src/solver_dynamic.c:194-195) is deferred too, the seed
         * listed first with the coefficients of its recurrence.  It is
         * updated in every window step, like the slowest shifts (the
         * smallest shifts, first in the list): grouped with them, the flush
         * chunks of shifts that stopped early end their loops early. */
        seedpos = d.exact ? -1 : 0;
        if (tid == 0) {
            if (seedpos >= 0) {
                d.flist[seedpos] = s;
                d.fpos[s] = seedpos;
            }
            st->win_seedpos = seedpos;
            st->flush_cnt = base;
        }
        cnt_now = base;
    }

    /* Opt-in guard of the default mode (growth_limit > 0): how large the
     * window's combined representation of p_j grows.  grow[j] bounds
     * sum_k |KP_k| |g_k| (plus the window-start p_j term); an iteration adds
     * |c4| (grow + |c2| |q| + |c3| |r_old|) + |c5| |r_new|, while its own
     * terms weigh |c4| (|c2| |q| + |c3| |r_old|) + |c5| |r_new|.  Past
     * growth_limit times the iteration's own terms the window closes, with
     * this iteration applied on its own (the per-iteration order).  Off by
     * default: with the coefficients carried in double-double the regrouped
     * flush already matches the per-iteration accuracy, and closing a window
     * where the bound peaks lands on the iterations of heaviest
     * cancellation. */
    const double nq = sqrt(in.qTq), nro = sqrt(in.dot_r_prev), nrn = sqrt(dot_r);
    double best_zp = 1.0;
    int best_j = -1, nconv = 0;
    for (int j = tid; j < M; j += HB_BLOCK) {
        /* every operand is loaded ahead of the stop test: one round trip */
        const int stopped = d.stop[j];
        const double sgj = d.sigma[j], pim1 = PI(d, j, k - 1), etaj = d.eta[j], zj = d.zeta[j];
        const int posj = d.fpos[j];
        const double growj = d.growth_limit > 0.0 ? d.grow[j] : 0.0;
        if (stopped) continue;
        double zp;
        if (j != s) {
            const double dj = sig_s - sgj;
            const double eta = ratio * ak * etaj - dj * ak * pim1;
            const double pik = eta + pim1;
            const double aj = (pim1 / pik) * ak;
            const double den = 1.0 - wk * dj;
            /* a shift whose 1 - omega~ (sigma_s - sigma_j) or pi_j[k] is zero
             * or not finite divides by it below (src/solver_dynamic.c:245-253);
             * the arithmetic goes on as in the reference and the status says so */
            if (!(den != 0.0 && isfinite(den) && pik != 0.0 && isfinite(pik))) sh_bad = 1;
            const double wj = wk / den;
            const double z = zj;
            const double az = aj * z;
            const double c1 = wj / (pik * z);
            const double c2 = wj / (az * pik);
            const double c3 = -wj / (az * pim1);
            const double znew = den * z;
            const double t = pim1 / pik;
            const double bj = t * t * beta;
            const double c4 = 1.0 / (pik * znew);
            d.eta[j] = eta;
            PI(d, j, k) = pik;
            d.zeta[j] = znew;
            const int pos = posj;
            double *c = coef + (size_t)pos * 6;
            c[0] = c1; c[1] = aj; c[2] = c2; c[3] = c3; c[4] = bj; c[5] = c4;
            d.act_len[pos] = slot + 1;
            if (d.growth_limit > 0.0) {
                const double own = fabs(bj) * (fabs(c2) * nq + fabs(c3) * nro) + fabs(c4) * nrn;
                const double prev = k == 1 ? sqrt(dot_zero) : growj;   /* p_j = b at k = 1 */
                const double bound = fabs(bj) * prev + own;
                if (!(bound <= d.growth_limit * own)) sh_spike = 1;
                d.grow[j] = bound;
                d.grow[M + j] = own;
            }
            zp = fabs(1.0 / (znew * pik));
        } else {
            zp = 1.0;
            if (seedpos >= 0) {
                /* the shift recurrence at sigma_j = sigma_s (pi = zeta = 1):
                 * x += omega~ q + alpha~ p,
                 * p = beta~ (p + (omega~/alpha~) q - (omega~/alpha~) r_old) + r */
                double *c = coef + (size_t)seedpos * 6;
                const double wa = wk / ak;
                c[0] = wk; c[1] = ak; c[2] = wa; c[3] = -wa; c[4] = beta; c[5] = 1.0;
                d.act_len[seedpos] = slot + 1;
            }
        }
        /* the residual estimate the reference prints under
         * DISPLAY_SIGMA_RESIDUAL (src/solver_dynamic.c:286-292) */
        if (d.trace) d.trace[(size_t)k * M + j] = zp * sqrt(dot_r / dot_zero);
        if (zp * zp * dot_r <= thr) {
            d.stop[j] = 1;
            d.stop_iter[j] = k;
            ++nconv;
        } else if (zp > best_zp) {
            best_zp = zp;
            best_j = j;
        }
    }
    sh_zp[tid] = best_zp;
    sh_j[tid] = best_j;
    sh_nc[tid] = nconv;
    __syncthreads();
    /* A window closes only at its full length, at the iteration cap, when
     * the seed converged (a switch follows, or the end) or at a growth-guard
     * tail: the decision below is a flush exactly when one of these holds
     * (the seed's own test is zp = 1, i.e. dot_r <= thr).  Otherwise the
     * launch's other workgroups learn here that no flush is due, without
     * waiting for the rest of the stage. */
    const bool may_close = slot + 1 >= d.W || k + 1 >= d.cap || dot_r <= thr || sh_spike;
    if (PUB && !may_close && tid == 0) mail_put(&st->mail_f, 0ull);
    for (int w = HB_BLOCK / 2; w > 0; w >>= 1) {
        if (tid < w) {
            const double z2 = sh_zp[tid + w];
            const int j2 = sh_j[tid + w];
            if (j2 >= 0 && (sh_j[tid] < 0 || z2 > sh_zp[tid] || (z2 == sh_zp[tid] && j2 < sh_j[tid]))) {
                sh_zp[tid] = z2;
                sh_j[tid] = j2;
            }
            sh_nc[tid] += sh_nc[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        /* the status word takes bits from other workgroups too: atomically */
        const int sc = in.stop_count + sh_nc[0];
        st->stop_count = sc;
        if (sh_bad) atomicOr(&st->status, HB_ST_BREAKDOWN_SHIFT);
        const int ms = sh_j[0] >= 0 ? sh_j[0] : in.max_sigma;
        st->max_sigma = ms;
        int sw = 0, nst = 0;
        /* the seed stopped: before (never, while it is the seed), or in the
         * loop above, whose test for it is dot_r <= thr (zp = 1) */
        if ((in.stop_s || dot_r <= thr) && sc < M) {
            if (ms >= 0) {
                sw = 1;
            } else {
                atomicOr(&st->status, HB_ST_NO_SWITCH_TARGET);
                nst = 1;
            }
        }
        sh_switch = sw;
        sh_sc = sc;
        sh_ms = ms;
        sh_nst = nst | ((in.status & HB_ST_NO_SWITCH_TARGET) != 0);
    }
    __syncthreads();

    const int do_switch = sh_switch;
    if (do_switch) {
        const int m = sh_ms;
        const double ds = sig_s - d.sigma[m];
        if (tid == 0) st->zscale = 1.0 / (d.zeta[m] * PI(d, m, k));
        for (int i = 1 + tid; i <= k; i += HB_BLOCK) {
            const double t0 = PI(d, m, i - 1), t1 = PI(d, m, i);
            d.aa[i] = (t0 / t1) * d.aa[i];
            d.ba[i] = (t0 / t1) * (t0 / t1) * d.ba[i];
            d.wa[i] = d.wa[i] / (1.0 - d.wa[i] * ds);
        }
        __syncthreads();
        const double sig_m = d.sigma[m];
        for (int j = tid; j < M; j += HB_BLOCK) {
            double e = 0.0, z = 1.0;
            if (!d.stop[j] && j != m) {
                const double dm = sig_m - d.sigma[j];
                for (int i = 1; i <= k; ++i) {
                    const double pm1 = PI(d, j, i - 1);
                    e = (d.ba[i - 1] / d.aa[i - 1]) * d.aa[i] * e - dm * d.aa[i] * pm1;
                    PI(d, j, i) = e + pm1;
                    z = (1.0 - d.wa[i] * dm) * z;
                }
            }
            d.eta[j] = e;
            d.zeta[j] = z;
        }
        __syncthreads();
    }

    const int len = slot + 1;
    if (d.flush_log) {
        /* profiled solves: window steps the listed shifts were updated in,
         * for the flop count of a flush (the flush skips each chunk's steps
         * after its last update) */
        const int Lc = len - sh_spike, cnt = cnt_now;
        int u = 0;
        for (int pos = tid; pos < cnt; pos += HB_BLOCK) u += min(d.act_len[pos], Lc);
        sh_nc[tid] = u;
        __syncthreads();
        for (int w = HB_BLOCK / 2; w > 0; w >>= 1) {
            if (tid < w) sh_nc[tid] += sh_nc[tid + w];
            __syncthreads();
        }
        if (tid == 0) sh_upd = sh_nc[0];
    }
    if (tid == 0) {
        const int done_next = (sh_sc >= M) || (k + 1 >= d.cap) || sh_nst;
        st->kprev = k;
        st->prev_seed = s;
        if (do_switch) {
            d.switch_k[in.n_switch] = k;
            d.switch_to[in.n_switch] = sh_ms;
            st->n_switch = in.n_switch + 1;
            st->seed = sh_ms;
            st->switched = 1;
        }
        const int spike = sh_spike;
        const int fl = do_switch || done_next || len >= d.W || spike;
        st->flush = fl;
        /* the solve's last flush leaves p_j unused: x_j only */
        const int xonly = fl && done_next && !do_switch && !spike;
        st->flush_xonly = xonly;
        if (fl) {
            st->flush_k0 = win_k0;
            st->flush_len = len;
            st->flush_tail = spike;
            st->win_k0 = k + 1;
        }
        sh_flush = fl;
        sh_xonly = xonly;
        st->k = k + 1;
        st->done = done_next;
        /* seed p update: p, r, s in, p out; at a flush also the window
         * vectors r_{k0-1..k0+L-1}, q_{k0..k0+L-1} and, per listed shift,
         * x_j out and in (not read in the first window with x_zero), p_j in
         * (b once instead in the first window) and p_j out (not in the
         * last, x-only flush) */
        double bytes = 32.0;
        if (fl) {
            const bool first = win_k0 == 1;
            const double per = 8.0 + (first && d.x_zero ? 0.0 : 8.0) + (first ? 0.0 : 8.0) + (xonly ? 0.0 : 8.0);
            bytes += 8.0 * (2 * len + 1) + (first ? 8.0 : 0.0) + per * cnt_now;
        }
        st->shift_bytes = in.shift_bytes + bytes * (double)d.n;
        /* multiply-adds per row and listed shift: x_j takes p_j, r_{k0-1}
         * and two window vectors per iteration it was updated in; p_j as
         * many again; a tail iteration applied on its own six more */
        if (fl && d.flush_log)
            st->flush_fma += (2.0 * (double)sh_upd + 2.0 * cnt_now) * (xonly ? 1.0 : 2.0) +
                             (spike ? 6.0 * cnt_now : 0.0);
        if (d.flush_log && k < d.cap) d.flush_log[k] = fl;
    }
    if (d.growth_limit > 0.0) {
        /* a closing window leaves p_j materialised: the next window's bound
         * starts from this iteration's own terms */
        __syncthreads();
        if (sh_flush)
            for (int j = tid; j < M; j += HB_BLOCK) d.grow[j] = d.grow[M + j];
    }
    if (PUB && may_close) {
        __syncthreads();
        if (sh_flush) {
            /* the list covers every position this iteration wrote (those of
             * shifts that stopped in the window before are stored again
             * unchanged) */
            const int cnt = cnt_now;
            for (int pos = tid; pos < cnt; pos += HB_BLOCK) {
                double *c = coef + (size_t)pos * 6;
#pragma unroll
                for (int e = 0; e < 6; ++e) store_pub<true>(&c[e], c[e]);
                store_pub<true>(&d.act_len[pos], d.act_len[pos]);
                if (slot == 0) store_pub<true>(&d.flist[pos], d.flist[pos]);
            }
            if (tid == 0) {
                store_pub<true>(&st->flush_cnt, cnt);
                store_pub<true>(&st->win_seedpos, slot == 0 ? seedpos : win_seedpos);
                store_pub<true>(&st->flush_k0, win_k0);
                store_pub<true>(&st->flush_len, len);
                store_pub<true>(&st->flush_tail, sh_spike);
            }
        }
        /* every store the launch's other workgroups read has landed before
         * the decision moves */
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0)
            mail_put(&st->mail_f, sh_flush ? (unsigned long long)(HB_FL_FLUSH | (sh_xonly ? HB_FL_XONLY : 0)) : 0ull);
    }
}

/* ---- seed update: x_s, r, (r,r), (r^,r) (src/solver_dynamic.c:194-204) --- */
template <bool MAIL = false>
DEV void seed_update_body(const hb_dev &d, int b, const double *__restrict__ qring,
                          const double *__restrict__ y, const double *__restrict__ pset,
                          double *__restrict__ xset, const double *__restrict__ rhat,
                          double *__restrict__ rring, double *__restrict__ part)
{
    hb_state *st = d.st;
    if (st->done) return;
    const int k = st->k, s = st->seed, n = d.n, nb = d.vec_grid;
    const double *q = qring + off(k % d.W, n);
    double *rn = rring + off(k % (d.W + 1), n);
    const double *ps = pset + off(s, n);
    double *xs = xset + off(s, n);
    double part0 = 0.0, part1 = 0.0;
    int i = b * HB_BLOCK + threadIdx.x;
    const int stride = nb * HB_BLOCK;
    double w;
    if (MAIL) {
        /* omega~ comes from the launch's first workgroup: the first four
         * elements' q, y and r^ are loaded while it is on its way, and
         * summed in the same order as the loop below */
        constexpr int E = 4;
        double qv[E], yv[E], hv[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int ie = min(i + e * stride, n - 1);
            qv[e] = q[ie];
            yv[e] = y[ie];
            hv[e] = rhat[ie];
        }
        w = mail_wait(&st->mail_w, st);
        const double nw = -w;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (i < n) {
                const double ri = qv[e] + nw * yv[e];
                rn[i] = ri;
                part0 += ri * ri;
                part1 += hv[e] * ri;
            }
            i += stride;
        }
    } else {
        w = d.wa[k];
    }
    const double a = d.aa[k], nw = -w;
    /* exact order updates x_s here, in the reference's order; the default
     * mode defers it to the window flush (the seed is listed there) */
    const bool upd_x = d.exact;
#pragma unroll 2
    for (; i < n; i += stride) {
        const double qi = q[i];
        if (upd_x) {
            double xi = xs[i];
            xi = xi + a * ps[i];
            xi = xi + w * qi;
            xs[i] = xi;
        }
        const double ri = qi + nw * y[i];
        rn[i] = ri;
        part0 += ri * ri;
        part1 += rhat[i] * ri;
    }
    block_sum2(part0, part1);
    if (threadIdx.x == 0) { part[2 * b] = part0; part[2 * b + 1] = part1; }
}

template <bool MAIL = false>
DEV void seed_update_tile(const hb_dev &d, int b)
{
    seed_update_body<MAIL>(d, b, d.qring, d.y, d.p_set, d.x_set, d.rhat, d.rring, d.part3);
}

__global__ __launch_bounds__(HB_BLOCK) void k_seed_update(hb_dev d)
{
    seed_update_tile(d, blockIdx.x);
}

/* Fused seed update (default mode): workgroup 0 finalizes omega~ from SpMV
 * 2's partials and publishes it; the others run the update once it is
 * there. */
__global__ __launch_bounds__(HB_BLOCK) void k_seed_update_fin(hb_dev d)
{
    if (blockIdx.x == 0) {
        fin_spmv_body<SP_SEED2, false, true>(d);
        return;
    }
    seed_update_tile<true>(d, blockIdx.x - 1);
}

/* (r,r), (r^,r) from the seed-update partials, then the scalar stage. */
/* false when the solve is done (nothing ran) */
template <bool EXACT, bool PUB = false>
DEV bool fin_seed_body(const hb_dev &d)
{
    __shared__ double lds[EXACT ? HB_NNZ_CAP : 1];
    __shared__ double sh[2];
    hb_state *st = d.st;
    hb_stage_in in;
    const int done = st->done;
    in.k = st->k;
    in.s = st->seed;
    in.win_k0 = st->win_k0;
    in.win_seedpos = st->win_seedpos;
    in.rTr_old = st->rTr_old;
    in.dot_zero = st->dot_zero;
    in.dot_r_prev = st->dot_r;
    in.qTq = st->qTq;
    in.stop_count = st->stop_count;
    in.max_sigma = st->max_sigma;
    in.status = st->status;
    in.n_switch = st->n_switch;
    in.flush_cnt = st->flush_cnt;
    in.shift_bytes = st->shift_bytes;
    if (done) return false;
    /* 1 <= k < cap while the solve runs */
    in.ak = d.aa[in.k];
    in.wk = d.wa[in.k];
    in.bkm1 = d.ba[in.k - 1];
    in.akm1 = d.aa[in.k - 1];
    in.sig_s = d.sigma[in.s];
    in.stop_s = d.stop[in.s];
    double dr, rr;
    if (EXACT) {
        const double *rn = d.rring + off(in.k % (d.W + 1), d.n);
        dr = serial_dot(rn, rn, d.n, lds);
        rr = serial_dot(d.rhat, rn, d.n, lds);
    } else {
        /* at most 2048 pairs: eight loads in flight per thread cover them,
         * and the shift kernel this stage may run in keeps its registers */
        reduce_partials<!PUB>(d.part3, d.vec_grid, dr, rr);
    }
    if (threadIdx.x == 0) {
        sh[0] = dr;
        sh[1] = rr;
        /* the fused launches' readers of omega~ finished with the last
         * launches */
        st->mail_w = HB_MAIL_EMPTY;
        st->mail_a = HB_MAIL_EMPTY;
    }
    __syncthreads();
    scalar_stage<PUB>(d, in, sh[0], sh[1]);
    return true;
}

template <bool EXACT>
__global__ __launch_bounds__(HB_BLOCK) void k_fin_seed(hb_dev d)
{
    fin_seed_body<EXACT>(d);
}

/* ---- shift window ------------------------------------------------------------ */
/* V consecutive doubles; V = 2 moves them with one 16-byte access. */
template <int V>
struct Vd {
    double v[V];
};

template <int V, bool NT>
DEV Vd<V> vload(const double *p)
{
    Vd<V> r;
    if (V == 2) {
        const dbl2 t = NT ? __builtin_nontemporal_load((const dbl2 *)p) : *(const dbl2 *)p;
        r.v[0] = t.x;
        r.v[V - 1] = t.y;
    } else {
        r.v[0] = NT ? __builtin_nontemporal_load(p) : *p;
    }
    return r;
}

template <int V, bool NT>
DEV void vstore(double *p, const Vd<V> &r)
{
    if (V == 2) {
        dbl2 t;
        t.x = r.v[0];
        t.y = r.v[V - 1];
        if (NT) __builtin_nontemporal_store(t, (dbl2 *)p);
        else *(dbl2 *)p = t;
    } else {
        if (NT) __builtin_nontemporal_store(r.v[0], p);
        else *p = r.v[0];
    }
}

/* The six vector operations of one shift for window iteration b, in the
 * reference's order and roundings (src/solver_dynamic.c:248-255). */
template <int V>
DEV void shift_step(Vd<V> &x, Vd<V> &p, const double *c, const Vd<V> &q, const Vd<V> &ro,
                    const Vd<V> &rn)
{
#pragma unroll
    for (int e = 0; e < V; ++e) {
        double xe = x.v[e], pe = p.v[e];
        xe = xe + c[0] * q.v[e];
        xe = xe + c[1] * pe;
        pe = pe + c[2] * q.v[e];
        pe = pe + c[3] * ro.v[e];
        pe = c[4] * pe;
        pe = pe + c[5] * rn.v[e];
        x.v[e] = xe;
        p.v[e] = pe;
    }
}

/* Exact-order window replay.  Tile (bx, by): each thread owns V consecutive
 * elements.  Row by == 0
 * applies the seed p update of the finished iteration
 * (src/solver_dynamic.c:208-210).  When a flush is due, row by applies the
 * pending window to the listed shifts by*G .. by*G+G-1: the window's r and q
 * values of the thread's elements stay in registers, x_j and p_j are read
 * and written once per flush (U shifts in flight per thread), and every
 * window iteration replays the six vector operations of
 * src/solver_dynamic.c:248-255 in their order.  The pointer parameters are
 * __restrict__ so that the coefficients, uniform across a wavefront, come
 * through scalar loads. */
template <int WM, int V, int U>
DEV void shift_body(const hb_dev &d, int bx, int by, double *__restrict__ xset,
                    double *__restrict__ pset, const double *__restrict__ rring,
                    const double *__restrict__ qring, const double *__restrict__ svec,
                    const double *__restrict__ coef, const int *__restrict__ flist,
                    const int *__restrict__ act_len)
{
    const hb_state *st = d.st;
    if (st->k4_epoch >= st->k) return;   /* no scalar stage since the last run */
    const int n = d.n, W = d.W, M = d.M;
    const int i = (bx * HB_BLOCK + threadIdx.x) * V;
    if (i >= n) return;
    const int kp = st->kprev;
    if (by == 0) {
        double *ps = pset + off(st->prev_seed, n) + i;
        const double pb = st->pb, pbw = st->pbw;
        Vd<V> p = vload<V, false>(ps);
        const Vd<V> r = vload<V, false>(rring + off(kp % (W + 1), n) + i);
        const Vd<V> sv = vload<V, false>(svec + i);
#pragma unroll
        for (int e = 0; e < V; ++e) {
            double pe = pb * p.v[e];
            pe = pe + r.v[e];
            p.v[e] = pe + pbw * sv.v[e];
        }
        vstore<V, false>(ps, p);
    }
    if (!st->flush) return;
    const int L = st->flush_len, k0 = st->flush_k0, cnt = st->flush_cnt;
    Vd<V> R[WM + 1], Q[WM];
    R[0] = vload<V, false>(rring + off((k0 - 1) % (W + 1), n) + i);
#pragma unroll
    for (int b = 0; b < WM; ++b) {
        if (b < L) {
            Q[b] = vload<V, false>(qring + off((k0 + b) % W, n) + i);
            R[b + 1] = vload<V, false>(rring + off((k0 + b) % (W + 1), n) + i);
        }
    }
    const int G = d.shift_group;
    const int j0 = by * G;
    const int j1 = min(cnt, j0 + G);
    int jj = j0;
    for (; jj + U <= j1; jj += U) {
        int jv[U], lv[U];
        Vd<V> x[U], p[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            jv[u] = flist[jj + u];
            lv[u] = act_len[jj + u];
            x[u] = vload<V, true>(xset + off(jv[u], n) + i);
            p[u] = vload<V, true>(pset + off(jv[u], n) + i);
        }
        bool whole = true;   /* every shift in flight active for the whole window */
#pragma unroll
        for (int u = 0; u < U; ++u) whole = whole && (lv[u] == L);
        if (whole) {
#pragma unroll
            for (int b = 0; b < WM; ++b) {
                if (b < L) {
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        shift_step<V>(x[u], p[u], coef + ((size_t)b * M + jj + u) * 6, Q[b], R[b], R[b + 1]);
                }
            }
        } else {
#pragma unroll
            for (int b = 0; b < WM; ++b) {
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (b < lv[u]) shift_step<V>(x[u], p[u], coef + ((size_t)b * M + jj + u) * 6, Q[b], R[b], R[b + 1]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            vstore<V, true>(xset + off(jv[u], n) + i, x[u]);
            vstore<V, true>(pset + off(jv[u], n) + i, p[u]);
        }
    }
    for (; jj < j1; ++jj) {
        const int ja = flist[jj];
        const int la = act_len[jj];
        double *xa = xset + off(ja, n) + i, *pa = pset + off(ja, n) + i;
        Vd<V> x0 = vload<V, true>(xa), p0 = vload<V, true>(pa);
#pragma unroll
        for (int b = 0; b < WM; ++b)
            if (b < la) shift_step<V>(x0, p0, coef + ((size_t)b * M + jj) * 6, Q[b], R[b], R[b + 1]);
        vstore<V, true>(xa, x0);
        vstore<V, true>(pa, p0);
    }
}

template <int WM, int V, int U>
DEV void shift_tile(const hb_dev &d, int bx, int by)
{
    shift_body<WM, V, U>(d, bx, by, d.x_set, d.p_set, d.rring, d.qring, d.s, d.coef, d.flist,
                         d.act_len);
}

template <int WM, int V, int U>
__global__ __launch_bounds__(HB_BLOCK) void k_shift(hb_dev d)
{
    shift_tile<WM, V, U>(d, blockIdx.x, blockIdx.y);
}

/* ---- combined-coefficient window flush (default mode) ----------------------
 * Applying the window's L iterations to shift j is linear in the window's
 * vectors g_0 = r_{k0-1}, g_{2b+1} = q_{k0+b}, g_{2b+2} = r_{k0+b}
 * (b = 0 .. L-1):
 *     x_j <- x_j + Xp p_j + sum_k KX_k g_k,     p_j <- Pp p_j + sum_k KP_k g_k
 * with x_j, p_j taken at the window start.  flush_coef_body forms
 * (Xp, Pp, KX, KP) of one listed shift by running the six operations of
 * src/solver_dynamic.c:248-255 on coefficient vectors (lane k of the
 * workgroup carries component k, the last lane the p_j component) for the
 * act_len iterations the shift was updated in; components of later
 * iterations stay 0.  shift_cc_body then streams every g_k once per tile
 * and accumulates HB_UG shifts per wavefront with fused multiply-adds, so
 * a flush costs 2L+1 vector reads and one read and write of x_j, p_j, and
 * its arithmetic does not grow with the window's per-step bookkeeping.
 *
 * Regrouping moves where cancellation happens.  When one iteration's
 * coefficients are huge (a pi_k or 1 - omega~ d of the shift near zero) and
 * the next cancel them, the per-iteration order cancels whole vectors, but
 * the regrouped form cancels in the coefficients: KX_k collects
 * c1 * KP_k over iterations, and two nearly opposite large terms leave it
 * with few correct digits in double precision.  The coefficients are
 * therefore carried in double-double arithmetic (an unevaluated sum
 * hi + lo, with Dekker's and Knuth's error-free product and sum -- the
 * scheme of Hida, Li and Bailey's QD library) and rounded to double once;
 * the flush sums then lose no more than the per-iteration order does.  As a
 * second guard the scalar stage watches a bound on the growth of the
 * regrouped p_j and, past growth_limit, closes the window with a tail: the
 * first L-1 iterations regrouped, the last applied on its own with its six
 * operations.
 *
 * Combined coefficients of list position pos live in chunk pos / HB_UG,
 * column pos % HB_UG of a [2W+5][2*HB_UG] block: row 0 holds (Xp | Pp),
 * row 1+k holds (KX_k | KP_k), x columns first; rows 2W+2 .. 2W+4 hold the
 * tail iteration's (c0 | c1), (c2 | c3), (c4 | c5), the identity
 * (0, 0, 0, 0, 1, 0) for a shift that was not updated in it. */
struct dd {
    double hi, lo;
};

/* s + e = a + b exactly (Knuth's two-sum) */
DEV dd two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

/* s + e = a + b exactly, given |a| >= |b| (Dekker's fast two-sum) */
DEV dd quick_two_sum(double a, double b)
{
    const double s = a + b;
    return {s, b - (s - a)};
}

/* a + c */
DEV dd dd_add(dd a, double c)
{
    const dd s = two_sum(a.hi, c);
    return quick_two_sum(s.hi, s.lo + a.lo);
}

/* a + b */
DEV dd dd_add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s = quick_two_sum(s.hi, s.lo + t.hi);
    return quick_two_sum(s.hi, s.lo + t.lo);
}

/* a * c: the product's rounding error from one fused multiply-add */
DEV dd dd_mul(dd a, double c)
{
    const double p = a.hi * c;
    const double e = __builtin_fma(a.hi, c, -p);
    return quick_two_sum(p, __builtin_fma(a.lo, c, e));
}
/* PUB: the block is read by the flush workgroups of the same launch.  MAIL:
 * the scalar stage runs in the launch's first workgroup; the decision and the
 * window's extent come from it. */
template <bool PUB = false, bool MAIL = false>
DEV void flush_coef_body(const hb_dev &d, int pos, const double *__restrict__ coef,
                         double *__restrict__ ccoef)
{
    hb_state *st = d.st;
    int L, cnt, ntail, lv;
    if (MAIL) {
        const unsigned long long f = mail_wait_bits(&st->mail_f, st);
        if (f == HB_MAIL_EMPTY || !(f & HB_FL_FLUSH)) return;
        cnt = load_pub(&st->flush_cnt);
        if (pos >= cnt) return;
        L = load_pub(&st->flush_len);
        ntail = load_pub(&st->flush_tail);
        lv = load_pub(&d.act_len[pos]);
    } else {
        if (st->k4_epoch >= st->k) return;   /* no scalar stage since the last run */
        if (!st->flush || pos >= st->flush_cnt) return;
        L = st->flush_len;
        ntail = st->flush_tail;
        lv = d.act_len[pos];
    }
    const int M = d.M, W = d.W;
    const int Lc = L - ntail;   /* regrouped iterations */
    const int t = threadIdx.x;
    const bool pch = t == HB_BLOCK - 1;   /* the p_j component */
    double *blk = ccoef + (size_t)(pos / HB_UG) * (size_t)HB_CC_ROWS(W) * (2 * HB_UG);
    if (pch && Lc < L) {
        const double *c = coef + ((size_t)(L - 1) * M + pos) * 6;
        const bool on = lv == L;
        const int base = (2 * W + 2) * (2 * HB_UG) + pos % HB_UG;
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            const double id = e == 4 ? 1.0 : 0.0;
            store_pub<PUB>(&blk[base + (e / 2) * (2 * HB_UG) + (e % 2) * HB_UG], on ? c[e] : id);
        }
    }
    const int nb = min(lv, Lc);
    /* thread HB_BLOCK-1 carries the p_j component; the others carry the
     * g_k components k = t, t + HB_BLOCK-1, ... (k <= 2Lc).  Component k
     * is zero before step (k-1)/2, so its walk starts there. */
    for (int kc = pch ? -1 : t; kc <= 2 * Lc; kc += HB_BLOCK - 1) {
        dd xk = {0.0, 0.0}, pk = {pch ? 1.0 : 0.0, 0.0};
        for (int b = kc > 0 ? (kc - 1) / 2 : 0; b < nb; ++b) {
            const double *c = coef + ((size_t)b * M + pos) * 6;
            const double c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3], c4 = c[4], c5 = c[5];
            /* x += c0 g_{2b+1} + c1 p ;  p = c4 (p + c2 g_{2b+1} + c3 g_{2b}) + c5 g_{2b+2} */
            dd xn = dd_add(xk, dd_mul(pk, c1));
            dd tp = pk;
            if (kc == 2 * b + 1) { xn = dd_add(xn, c0); tp = dd_add(tp, c2); }
            if (kc == 2 * b) tp = dd_add(tp, c3);
            dd pn = dd_mul(tp, c4);
            if (kc == 2 * b + 2) pn = dd_add(pn, c5);
            xk = xn;
            pk = pn;
        }
        const int row = pch ? 0 : 1 + kc;
        store_pub<PUB>(&blk[(size_t)row * (2 * HB_UG) + pos % HB_UG], xk.hi + xk.lo);
        store_pub<PUB>(&blk[(size_t)row * (2 * HB_UG) + HB_UG + pos % HB_UG], pk.hi + pk.lo);
        if (pch) break;
    }
    if (PUB) {
        /* every store of this workgroup is done before the count moves (the
         * flush workgroups of this launch read the block once the count
         * reaches the list length) */
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_fetch_add(&st->coef_ready, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ __launch_bounds__(HB_BLOCK) void k_flush_coef(hb_dev d)
{
    flush_coef_body(d, blockIdx.x, d.coef, d.ccoef);
}

/* The window a flush applies: its first iteration, list length, regrouped
 * iterations, tail flag and the seed's list position, uniform across the
 * workgroup (read once by the caller, kept in scalar registers). */
struct hb_fl {
    int k0, cnt, Lc, tail, sp;
};

DEV hb_fl flush_window(int k0, int cnt, int len, int tail, int sp)
{
    return {__builtin_amdgcn_readfirstlane(k0), __builtin_amdgcn_readfirstlane(cnt),
            __builtin_amdgcn_readfirstlane(len - tail), __builtin_amdgcn_readfirstlane(tail),
            __builtin_amdgcn_readfirstlane(sp)};
}

/* The window flush of tile bx (256 elements): the four wavefronts walk the
 * tile's four 64-element sub-tiles together, wavefront w taking shift
 * chunks w, w+4, ...: the wavefronts read the same g_k slices at about the
 * same time, so each g_k comes from HBM once and from cache for the other
 * chunks.  The chunk's coefficients are uniform across the wavefront and
 * arrive through scalar loads; the wavefront index goes through
 * readfirstlane so the compiler knows it is uniform.  The solve's last
 * flush, which needs x_j only, is flush_apply_x2 below. */
template <int UG>
DEV void flush_apply(const hb_dev &d, int bx, const hb_fl &fw, double *__restrict__ xset,
                     double *__restrict__ pset, const double *__restrict__ rring,
                     const double *__restrict__ qring, const double *__restrict__ ccoef,
                     const int *__restrict__ flist)
{
    const int n = d.n, W = d.W;
    const int k0 = fw.k0, cnt = fw.cnt, Lc = fw.Lc, tail = fw.tail, sp = fw.sp;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nch = (cnt + UG - 1) / UG;
    const HB_K double *cc = read_only(ccoef);
    const HB_K int *fl = read_only(flist), *al = read_only((const int *)d.act_len);
    const size_t cstride = (size_t)HB_CC_ROWS(W) * (2 * UG);
    for (int sub = 0; sub < HB_BLOCK / 64; ++sub) {
        if (bx * HB_BLOCK + sub * 64 >= n) break;   /* the tile ends: uniform */
        /* the wavefronts meet before each sub-tile, so none runs ahead into
         * elements whose window vectors the others would fetch again */
        if (sub) __syncthreads();
        const int ir = bx * HB_BLOCK + sub * 64 + lane;
        const bool has = ir < n;
        const int i = has ? ir : n - 1;       /* a clamped duplicate past the end */
        /* chunk residue rotated by tile: the list runs from the slowest
         * shifts (the longest loops) to the fastest, so the long chunks land
         * on different wavefronts, and SIMDs, from tile to tile */
        for (int c = (wave + bx) % (HB_BLOCK / 64); c < nch; c += HB_BLOCK / 64) {
            const HB_K double *cb = cc + (size_t)c * cstride;
            const int nu = min(UG, cnt - c * UG);
            /* a shift's coefficients of window steps after the last one it
             * was updated in are exactly zero: the chunk's loop stops at its
             * latest such step (shifts stopping at similar iterations sit in
             * the same chunk, since the list is ascending and the stop
             * iteration follows the shift) */
            int nbc = 0;
            for (int u = 0; u < nu; ++u) nbc = max(nbc, min(al[c * UG + u], Lc));
            /* uniform across the wavefront: the loop index, and with it the
             * coefficient addresses, stay scalar */
            nbc = __builtin_amdgcn_readfirstlane(nbc);
            if (tail) nbc = Lc;   /* the tail starts from the ring slots the loop ends on */
            double x[UG], p[UG];
            /* the first window starts from p_j = b for every shift (the
             * initialisation stores only the seed's p), and with x_zero from
             * x_j = 0 */
            const double pb = k0 == 1 ? d.b_in[i] : 0.0;
            const bool xz = k0 == 1 && d.x_zero;
#pragma unroll
            for (int u = 0; u < UG; ++u) {
                if (u < nu) {
                    const int j = fl[c * UG + u];
                    x[u] = xz ? 0.0 : __builtin_nontemporal_load(xset + off(j, n) + i);
                    /* the seed's window-start p is the snapshot (its p_s
                     * has moved on every iteration) */
                    p[u] = k0 == 1 ? pb
                                   : __builtin_nontemporal_load(c * UG + u == sp ? d.p_snap + i : pset + off(j, n) + i);
                } else {
                    x[u] = 0.0;
                    p[u] = 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < UG; ++u) {
                x[u] = __builtin_fma(cb[u], p[u], x[u]);
                p[u] = cb[UG + u] * p[u];
            }
            int rs = (k0 - 1) % (W + 1), qs = k0 % W;
            double gr = rring[off(rs, n) + i];
            {
                const HB_K double *ck = cb + 2 * UG;
#pragma unroll
                for (int u = 0; u < UG; ++u) {
                    x[u] = __builtin_fma(ck[u], gr, x[u]);
                    p[u] = __builtin_fma(ck[UG + u], gr, p[u]);
                }
            }
            for (int b = 0; b < nbc; ++b) {
                rs = rs + 1 == W + 1 ? 0 : rs + 1;
                const double gq = qring[off(qs, n) + i];
                gr = rring[off(rs, n) + i];
                qs = qs + 1 == W ? 0 : qs + 1;
                const HB_K double *ck = cb + (size_t)(2 * b + 2) * (2 * UG);
#pragma unroll
                for (int u = 0; u < UG; ++u) {
                    x[u] = __builtin_fma(ck[u], gq, x[u]);
                    p[u] = __builtin_fma(ck[UG + u], gq, p[u]);
                }
#pragma unroll
                for (int u = 0; u < UG; ++u) {
                    x[u] = __builtin_fma(ck[2 * UG + u], gr, x[u]);
                    p[u] = __builtin_fma(ck[3 * UG + u], gr, p[u]);
                }
            }
            if (tail) {
                /* the last iteration on its own: gr is r_{k0+Lc-1} */
                rs = rs + 1 == W + 1 ? 0 : rs + 1;
                const double gq = qring[off(qs, n) + i];
                const double gn = rring[off(rs, n) + i];
                const HB_K double *ct = cb + (size_t)(2 * W + 2) * (2 * UG);
#pragma unroll
                for (int u = 0; u < UG; ++u) {
                    double xe = __builtin_fma(ct[u], gq, x[u]);
                    xe = __builtin_fma(ct[UG + u], p[u], xe);
                    double pe = __builtin_fma(ct[2 * UG + u], gq, p[u]);
                    pe = __builtin_fma(ct[3 * UG + u], gr, pe);
                    pe = ct[4 * UG + u] * pe;
                    p[u] = __builtin_fma(ct[5 * UG + u], gn, pe);
                    x[u] = xe;
                }
            }
#pragma unroll
            for (int u = 0; u < UG; ++u) {
                if (u < nu && has) {
                    const int j = fl[c * UG + u];
                    __builtin_nontemporal_store(x[u], xset + off(j, n) + i);
                    /* the seed keeps its own p_s (the regrouped one is
                     * only its x's companion) */
                    if (c * UG + u != sp) __builtin_nontemporal_store(p[u], pset + off(j, n) + i);
                }
            }
        }
    }
}

/* The solve's last flush (x_j only) of tile bx.  Each step of the window
 * loop waits for its g_k loads (cache hits, but hundreds of cycles), so a
 * lane carries two elements, i and i + 64: every wait then covers twice
 * the multiply-adds, and the loop runs half as many steps. */
template <int UG>
DEV void flush_apply_x2(const hb_dev &d, int bx, const hb_fl &fw, double *__restrict__ xset,
                        const double *__restrict__ pset, const double *__restrict__ rring,
                        const double *__restrict__ qring, const double *__restrict__ ccoef,
                        const int *__restrict__ flist)
{
    const int n = d.n, W = d.W;
    const int k0 = fw.k0, cnt = fw.cnt, Lc = fw.Lc, sp = fw.sp;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nch = (cnt + UG - 1) / UG;
    const HB_K double *cc = read_only(ccoef);
    const HB_K int *fl = read_only(flist), *al = read_only((const int *)d.act_len);
    const size_t cstride = (size_t)HB_CC_ROWS(W) * (2 * UG);
    for (int sub = 0; sub < HB_BLOCK / 64; sub += 2) {
        if (bx * HB_BLOCK + sub * 64 >= n) break;   /* the tile ends: uniform */
        /* the four wavefronts walk the same elements through the window in
         * step; they meet at a barrier before the next pair of sub-tiles, so
         * the ones whose chunks end early do not run ahead into elements
         * whose window vectors the others would fetch again */
        if (sub) __syncthreads();
        const int ir = bx * HB_BLOCK + sub * 64 + lane;
        const bool has0 = ir < n, has1 = ir + 64 < n;
        const int i0 = has0 ? ir : n - 1;     /* clamped duplicates past the end */
        const int i1 = has1 ? ir + 64 : i0;
        /* chunk residue rotated by tile: the list runs from the slowest
         * shifts (the longest loops) to the fastest, so the long chunks land
         * on different wavefronts, and SIMDs, from tile to tile */
        for (int c = (wave + bx) % (HB_BLOCK / 64); c < nch; c += HB_BLOCK / 64) {
            const HB_K double *cb = cc + (size_t)c * cstride;
            const int nu = min(UG, cnt - c * UG);
            /* a shift's coefficients of window steps after the last one it
             * was updated in are exactly zero: the chunk's loop stops at its
             * latest such step (shifts stopping at similar iterations sit in
             * the same chunk, since the list is ascending and the stop
             * iteration follows the shift) */
            int nbc = 0;
            for (int u = 0; u < nu; ++u) nbc = max(nbc, min(al[c * UG + u], Lc));
            /* uniform across the wavefront: the loop index, and with it the
             * coefficient addresses, stay scalar */
            nbc = __builtin_amdgcn_readfirstlane(nbc);
            double x0[UG], x1[UG];
            if (k0 == 1) {
                /* the first window starts from p_j = b for every shift (the
                 * initialisation stores only the seed's p): one load; and
                 * with x_zero from x_j = 0, not read */
                const double b0 = d.b_in[i0], b1 = d.b_in[i1];
                const bool xz = d.x_zero;
#pragma unroll
                for (int u = 0; u < UG; ++u) {
                    if (u < nu) {
                        const int j = fl[c * UG + u];
                        x0[u] = __builtin_fma(cb[u], b0, xz ? 0.0 : __builtin_nontemporal_load(xset + off(j, n) + i0));
                        x1[u] = __builtin_fma(cb[u], b1, xz ? 0.0 : __builtin_nontemporal_load(xset + off(j, n) + i1));
                    } else {
                        x0[u] = 0.0;
                        x1[u] = 0.0;
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < UG; ++u) {
                    if (u < nu) {
                        const int j = fl[c * UG + u];
                        const double *pj = c * UG + u == sp ? d.p_snap : pset + off(j, n);
                        x0[u] = __builtin_fma(cb[u], __builtin_nontemporal_load(pj + i0),
                                              __builtin_nontemporal_load(xset + off(j, n) + i0));
                        x1[u] = __builtin_fma(cb[u], __builtin_nontemporal_load(pj + i1),
                                              __builtin_nontemporal_load(xset + off(j, n) + i1));
                    } else {
                        x0[u] = 0.0;
                        x1[u] = 0.0;
                    }
                }
            }
            int rs = (k0 - 1) % (W + 1), qs = k0 % W;
            {
                const double *gp = rring + off(rs, n);
                const double g0 = gp[i0], g1 = gp[i1];
                const HB_K double *ck = cb + 2 * UG;
#pragma unroll
                for (int u = 0; u < UG; ++u) {
                    x0[u] = __builtin_fma(ck[u], g0, x0[u]);
                    x1[u] = __builtin_fma(ck[u], g1, x1[u]);
                }
            }
            for (int b = 0; b < nbc; ++b) {
                rs = rs + 1 == W + 1 ? 0 : rs + 1;
                const double *qp = qring + off(qs, n), *rp = rring + off(rs, n);
                const double q0 = qp[i0], q1 = qp[i1], r0 = rp[i0], r1 = rp[i1];
                qs = qs + 1 == W ? 0 : qs + 1;
                const HB_K double *ck = cb + (size_t)(2 * b + 2) * (2 * UG);
#pragma unroll
                for (int u = 0; u < UG; ++u) {
                    x0[u] = __builtin_fma(ck[u], q0, x0[u]);
                    x1[u] = __builtin_fma(ck[u], q1, x1[u]);
                }
#pragma unroll
                for (int u = 0; u < UG; ++u) {
                    x0[u] = __builtin_fma(ck[2 * UG + u], r0, x0[u]);
                    x1[u] = __builtin_fma(ck[2 * UG + u], r1, x1[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < UG; ++u) {
                if (u < nu) {
                    const int j = fl[c * UG + u];
                    if (has0) __builtin_nontemporal_store(x0[u], xset + off(j, n) + i0);
                    if (has1) __builtin_nontemporal_store(x1[u], xset + off(j, n) + i1);
                }
            }
        }
    }
}

/* Tile bx: every thread first applies the seed p update of the finished
 * iteration to its element (src/solver_dynamic.c:208-210), then the tile
 * takes part in the window flush when one is due. */
template <int UG, bool FUSED = false>
DEV void shift_cc_body(const hb_dev &d, int bx, double *__restrict__ xset,
                       double *__restrict__ pset, const double *__restrict__ rring,
                       const double *__restrict__ qring, const double *__restrict__ svec,
                       const double *__restrict__ ccoef, const int *__restrict__ flist)
{
    hb_state *st = d.st;
    if (st->k4_epoch >= st->k) return;   /* no scalar stage since the last run */
    const int n = d.n, W = d.W;
    const double pb = st->pb, pbw = st->pbw;
    if (!(n & 1)) {
        /* even n (every vector 16-byte aligned): workgroup bx updates rows
         * 512 bx .. 512 bx + 511, two per thread with 16-byte accesses; the
         * flush below never reads this p (the seed's window-start p is the
         * snapshot) */
        const int i2 = 2 * (bx * HB_BLOCK + (int)threadIdx.x);
        if (i2 < n) {
            dbl2 *ps = (dbl2 *)(pset + off(st->prev_seed, n) + i2);
            const dbl2 r = *(const dbl2 *)(rring + off(st->kprev % (W + 1), n) + i2);
            const dbl2 sv = *(const dbl2 *)(svec + i2);
            dbl2 p = *ps;
            double pe = pb * p.x;
            pe = pe + r.x;
            p.x = pe + pbw * sv.x;
            pe = pb * p.y;
            pe = pe + r.y;
            p.y = pe + pbw * sv.y;
            *ps = p;
        }
    } else {
        const int i0 = bx * HB_BLOCK + threadIdx.x;
        if (i0 < n) {
            double *ps = pset + off(st->prev_seed, n) + i0;
            const double r = rring[off(st->kprev % (W + 1), n) + i0];
            double pe = pb * *ps;
            pe = pe + r;
            *ps = pe + pbw * svec[i0];
        }
    }
    if (!st->flush) return;
    if (FUSED) {
        /* the combined coefficients come from the first workgroups of this
         * launch: wait until every listed position has stored its block.
         * This workgroup reads no coefficient before, so no cache holds a
         * stale copy of one */
        if (threadIdx.x == 0) {
            const int want = st->flush_cnt;
            for (int spins = 0;; ++spins) {
                if (__hip_atomic_load(&st->coef_ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= want) break;
                if (spins == HB_SPIN_LIMIT / 16 || ((spins & 15) == 15 && sync_timed_out(st))) {
                    atomicOr(&st->status, HB_ST_SYNC_TIMEOUT);
                    break;
                }
                /* the coefficients take tens to hundreds of microseconds (a
                 * long window's last flush): every resident tile waits here,
                 * and polls that come every few microseconds leave the memory
                 * system to the workgroups forming them */
                __builtin_amdgcn_s_sleep(127);
            }
        }
        __syncthreads();
    }
    const hb_fl fw = flush_window(st->flush_k0, st->flush_cnt, st->flush_len, st->flush_tail, st->win_seedpos);
    if (st->flush_xonly) flush_apply_x2<UG>(d, bx, fw, xset, pset, rring, qring, ccoef, flist);
    else flush_apply<UG>(d, bx, fw, xset, pset, rring, qring, ccoef, flist);
}

template <bool FUSED = false>
DEV void shift_cc_tile(const hb_dev &d, int bx)
{
    shift_cc_body<HB_UG, FUSED>(d, bx, d.x_set, d.p_set, d.rring, d.qring, d.s, d.ccoef, d.flist);
}

__global__ __launch_bounds__(HB_BLOCK) void k_shift_cc(hb_dev d)
{
    shift_cc_tile(d, blockIdx.x);
}

/* Fused shift kernel (default mode): workgroups 0..M-1 form the combined
 * coefficients of list positions 0..M-1 when a flush is due (at once
 * otherwise); the rest are the 256-element tiles of the seed p update and
 * the flush. */
__global__ __launch_bounds__(HB_BLOCK) void k_shift_fused(hb_dev d)
{
    if ((int)blockIdx.x < d.M) {
        flush_coef_body<true>(d, blockIdx.x, d.coef, d.ccoef);
        return;
    }
    shift_cc_tile<true>(d, blockIdx.x - d.M);
}

/* p = beta~ p + r + (-beta~ omega~) s in the order of the shift kernel's
 * seed p update above (src/solver_dynamic.c:208-210) */
DEV double p_step(double p, double r, double s, double pb, double pbw)
{
    double pe = pb * p;
    pe = pe + r;
    return pe + pbw * s;
}

DEV dbl2 p_step2(dbl2 p, dbl2 r, dbl2 s, double pb, double pbw)
{
    p.x = p_step(p.x, r.x, s.x, pb, pbw);
    p.y = p_step(p.y, r.y, s.y, pb, pbw);
    return p;
}

/* Tile workgroup t of T in the shift kernel that runs the scalar stage.  The
 * seed p update of the finished iteration takes elements in 256-thread
 * chunks t, t + T, ... (pairs of elements when n is even, so every access is
 * 16 bytes); each thread loads its first four chunks' p, r and s before
 * beta~ arrives from the stage.  T is sized so that the tiles stay resident:
 * their loads run while the stage does.  Then, when the stage has decided a
 * flush and the combined coefficients are in, the workgroup applies flush
 * tiles t, t + T, ...  The solve's last flush (x_j only, the longest) is left
 * to k_flush_final after the loop: in this kernel the compiler keeps one of
 * a window step's four coefficient blocks in flight instead of four, and
 * that flush ran 6.9 ms here against 4.3 ms in a kernel of its own. */
DEV void shift_stage_tiles(const hb_dev &d, int t, int T)
{
    hb_state *st = d.st;
    const int n = d.n, W = d.W;
    /* written by omega~'s finalize in the seed update's launch: the stage
     * moves k and the seed, and rescales the archives at a switch.  A
     * launch after the solve ended (no finalize ran) returns before loading
     * anything */
    const int live = st->pu_live, k = st->pu_k, s = min(max(st->pu_seed, 0), d.M - 1);
    const double w = st->pu_w;
    if (!live) return;
    double *ps = d.p_set + off(s, n);
    const double *rk = d.rring + off(k % (W + 1), n);
    const double *sv = d.s;
    constexpr int E = 4;
    const int stride = T * HB_BLOCK;
    int q = t * HB_BLOCK + (int)threadIdx.x;
    if (!(n & 1)) {
        const int np = n >> 1;
        dbl2 pv[E], rv[E], sw[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            /* clamped duplicates past the end (n >= 2 here) */
            const int qe = min(q + e * stride, np - 1);
            pv[e] = ((const dbl2 *)ps)[qe];
            rv[e] = ((const dbl2 *)rk)[qe];
            sw[e] = ((const dbl2 *)sv)[qe];
        }
        const unsigned long long bb = mail_wait_bits(&st->mail_b, st);
        if (bb == HB_MAIL_SKIP || bb == HB_MAIL_EMPTY) return;
        const double pb = __longlong_as_double((long long)bb), pbw = -pb * w;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (q < np) ((dbl2 *)ps)[q] = p_step2(pv[e], rv[e], sw[e], pb, pbw);
            q += stride;
        }
        for (; q < np; q += stride)
            ((dbl2 *)ps)[q] = p_step2(((const dbl2 *)ps)[q], ((const dbl2 *)rk)[q], ((const dbl2 *)sv)[q], pb, pbw);
    } else {
        double pv[E], rv[E], sw[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int ie = min(q + e * stride, n - 1);
            pv[e] = ps[ie];
            rv[e] = rk[ie];
            sw[e] = sv[ie];
        }
        const unsigned long long bb = mail_wait_bits(&st->mail_b, st);
        if (bb == HB_MAIL_SKIP || bb == HB_MAIL_EMPTY) return;
        const double pb = __longlong_as_double((long long)bb), pbw = -pb * w;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (q < n) ps[q] = p_step(pv[e], rv[e], sw[e], pb, pbw);
            q += stride;
        }
        for (; q < n; q += stride) ps[q] = p_step(ps[q], rk[q], sv[q], pb, pbw);
    }

    const unsigned long long f = mail_wait_bits(&st->mail_f, st);
    if (f == HB_MAIL_EMPTY || !(f & HB_FL_FLUSH) || (f & HB_FL_XONLY)) return;
    __shared__ int sh_w[5];
    if (threadIdx.x == 0) {
        /* the combined coefficients come from workgroups 1..M: wait until
         * every listed position has stored its block (as shift_cc_body) */
        const int want = load_pub(&st->flush_cnt);
        for (int spins = 0;; ++spins) {
            if (__hip_atomic_load(&st->coef_ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= want) break;
            if (spins == HB_SPIN_LIMIT / 16 || ((spins & 15) == 15 && sync_timed_out(st))) {
                atomicOr(&st->status, HB_ST_SYNC_TIMEOUT);
                break;
            }
            __builtin_amdgcn_s_sleep(127);
        }
        sh_w[0] = load_pub(&st->flush_k0);
        sh_w[1] = want;
        sh_w[2] = load_pub(&st->flush_len);
        sh_w[3] = load_pub(&st->flush_tail);
        sh_w[4] = load_pub(&st->win_seedpos);
    }
    __syncthreads();
    const hb_fl fw = flush_window(sh_w[0], sh_w[1], sh_w[2], sh_w[3], sh_w[4]);
    for (int bx = t; bx < d.shift_gx; bx += T) {
        /* the wavefronts start each tile together, as they walk its
         * sub-tiles together */
        if (bx != t) __syncthreads();
        flush_apply<HB_UG>(d, bx, fw, d.x_set, d.p_set, d.rring, d.qring, d.ccoef, d.flist);
    }
}

/* The solve's last window flush (x_j only) after the iteration loop, when the
 * shift kernel runs the scalar stage: that iteration's stage marked it
 * (flush, flush_xonly) and its coefficient workgroups formed the combined
 * coefficients.  One workgroup per 256-element tile. */
__global__ __launch_bounds__(HB_BLOCK) void k_flush_final(hb_dev d)
{
    const hb_state *st = d.st;
    if (!st->flush || !st->flush_xonly) return;
    const hb_fl fw = flush_window(st->flush_k0, st->flush_cnt, st->flush_len, st->flush_tail, st->win_seedpos);
    flush_apply_x2<HB_UG>(d, blockIdx.x, fw, d.x_set, d.p_set, d.rring, d.qring, d.ccoef, d.flist);
}

/* Shift kernel with the scalar stage (default mode, fused level 1):
 * workgroup 0 runs the stage, workgroups 1..M form the combined coefficients
 * when it decides a flush, the rest are the shift_tw tiles above.  It takes
 * about 106 registers (four waves per SIMD); held to five waves it spills
 * in the stage and measured 0.4 ms per 128^3 solve slower. */
__global__ __launch_bounds__(HB_BLOCK) void k_shift_stage(hb_dev d)
{
    if (blockIdx.x == 0) {
        /* a done solve: the launch's seed p update and flush have nothing
         * to do */
        if (!fin_seed_body<false, true>(d) && threadIdx.x == 0) {
            mail_put(&d.st->mail_b, HB_MAIL_SKIP);
            mail_put(&d.st->mail_f, 0ull);
        }
        return;
    }
    if ((int)blockIdx.x <= d.M) {
        flush_coef_body<true, true>(d, blockIdx.x - 1, d.coef, d.ccoef);
        return;
    }
    shift_stage_tiles(d, blockIdx.x - 1 - d.M, gridDim.x - 1 - d.M);
}

/* ---- initialisation (src/solver_dynamic.c:101-123) ----------------------- */
DEV void init_body(const hb_dev &d, int blk)
{
    const int n = d.n, M = d.M, nb = d.vec_grid;
    const double *b = d.b_in;
    double *rhat = (double *)d.rhat;
    double part0 = 0.0, part1 = 0.0;
    /* p_j = b: exact order replays every shift's update from p_set; the
     * default mode reads b as the first window's p_j in its flush, so only
     * the seed's p (which its own update moves every iteration) is stored */
    const int j0 = d.exact ? 0 : d.seed_in, j1 = d.exact ? M : d.seed_in + 1;
    /* x_zero: exact order reads x_j every iteration, so it starts them at 0
     * here; the default mode's first flush writes them without reading */
    const bool xz = d.exact && d.x_zero;
    for (int i = blk * HB_BLOCK + threadIdx.x; i < n; i += nb * HB_BLOCK) {
        const double bi = b[i];
        rhat[i] = bi;
        d.rring[i] = bi;
        for (int j = j0; j < j1; ++j) d.p_set[off(j, n) + i] = bi;
        if (xz)
            for (int j = 0; j < M; ++j) d.x_set[off(j, n) + i] = 0.0;
        part0 += bi * bi;
    }
    block_sum2(part0, part1);
    if (threadIdx.x == 0) { d.part[2 * blk] = part0; d.part[2 * blk + 1] = 0.0; }
}

__global__ __launch_bounds__(HB_BLOCK) void k_init(hb_dev d)
{
    init_body(d, blockIdx.x);
}

/* (b,b) and the scalar initialisation. */
template <bool EXACT>
DEV void init_fin_body(const hb_dev &d)
{
    __shared__ double lds[EXACT ? HB_NNZ_CAP : 1];
    const int M = d.M;
    double bb, unused;
    if (EXACT) bb = serial_dot(d.b_in, d.b_in, d.n, lds);
    else reduce_partials(d.part, d.vec_grid, bb, unused);
    for (int j = threadIdx.x; j < M; j += HB_BLOCK) {
        d.eta[j] = 0.0;
        d.zeta[j] = 1.0;
        d.stop[j] = 0;
        d.stop_iter[j] = 0;
        d.act_len[j] = 0;
        d.fpos[j] = -1;
        PI(d, j, 0) = 1.0;
        PI(d, j, 1) = 1.0;
    }
    if (threadIdx.x == 0) {
        hb_state *st = d.st;
        d.aa[0] = 1.0;
        d.ba[0] = 0.0;
        st->rTr = bb;
        st->rTr_old = bb;
        st->dot_r = bb;
        st->dot_zero = bb;
        st->k = 1;
        st->kprev = 0;
        st->seed = d.seed_in;
        st->prev_seed = d.seed_in;
        st->stop_count = 0;
        st->max_sigma = -1;
        st->done = (d.cap <= 1);
        st->switched = 0;
        st->n_switch = 0;
        st->win_k0 = 1;
        st->flush = 0;
        st->k4_epoch = 1;
        st->status = 0;
        st->zscale = 1.0;
        st->shift_bytes = 0.0;
        st->flush_fma = 0.0;
        st->coef_ready = 0;
        st->mail_w = HB_MAIL_EMPTY;
        st->mail_a = HB_MAIL_EMPTY;
    }
}

template <bool EXACT>
__global__ __launch_bounds__(HB_BLOCK) void k_init_fin(hb_dev d)
{
    init_fin_body<EXACT>(d);
}

/* A switch in the last iteration leaves r unscaled; apply it as the
 * reference does before returning r (src/solver_dynamic.c:322). */
DEV void finalize_body(const hb_dev &d, int blk)
{
    const hb_state *st = d.st;
    if (!st->switched) return;
    double *r = d.rring + off(st->kprev % (d.W + 1), d.n);
    const double c = st->zscale;
    for (int i = blk * HB_BLOCK + threadIdx.x; i < d.n; i += d.vec_grid * HB_BLOCK) r[i] = r[i] * c;
}

__global__ __launch_bounds__(HB_BLOCK) void k_finalize(hb_dev d)
{
    finalize_body(d, blockIdx.x);
}

__global__ void k_publish(const hb_state *st, int *flag)
{
    if (threadIdx.x == 0) {
        flag[0] = st->done;
        flag[1] = st->k;
        __threadfence_system();
    }
}

/* ---- true residual of every shift ----------------------------------------
 * out[j] = ||b - (A + sigma_j I) x_j|| / ||b|| in FP64: grid (R, M), the R
 * workgroups of shift j striding over the rows; row i's product summed in
 * CSR order, then + sigma_j x_i and - b_i as the reference forms it
 * (src/solver_dynamic.c:405-421, src/matrix.c:498-516); the squared entries
 * go to one partial pair per workgroup, combined in a fixed order per shift.
 * R depends only on (n, M), so every caller gets the same roundings. */
__global__ __launch_bounds__(HB_BLOCK) void k_resid_part(hb_dev d, const double *__restrict__ b,
                                                         const double *__restrict__ sigma,
                                                         const double *__restrict__ x_set,
                                                         double *__restrict__ part)
{
    const int j = blockIdx.y;
    const double *__restrict__ x = x_set + off(j, d.n);
    const double sj = sigma[j];
    double e2 = 0.0, b2 = 0.0;
    for (int i = blockIdx.x * HB_BLOCK + threadIdx.x; i < d.n; i += gridDim.x * HB_BLOCK) {
        double t = 0.0;
        for (unsigned e = d.ptr[i]; e < d.ptr[i + 1]; ++e) t += d.val[e] * x[d.col[e]];
        t = t + sj * x[i];
        const double bi = b[i];
        const double e1 = t - bi;
        e2 += e1 * e1;
        b2 += bi * bi;
    }
    block_sum2(e2, b2);
    if (threadIdx.x == 0) {
        part[((size_t)j * gridDim.x + blockIdx.x) * 2] = e2;
        part[((size_t)j * gridDim.x + blockIdx.x) * 2 + 1] = b2;
    }
}

__global__ __launch_bounds__(HB_BLOCK) void k_resid_final(const double *part, int nb, double *out)
{
    double a, b;
    reduce_partials(part + (size_t)blockIdx.x * nb * 2, nb, a, b);
    if (threadIdx.x == 0) out[blockIdx.x] = sqrt(a) / sqrt(b);
}

/* ---- batched launches --------------------------------------------------------
 * A phase launch walks its active list grid-stride: entry t works on
 * problem tab[2t], tile tab[2t+1], for t < *cnt (the list holds the tiles
 * of the problems still running at the chunk's start, kb_compact).
 * Finalize launches use workgroup p for problem p.  Problems that finish
 * inside a chunk return at once inside the bodies. */
template <int MODE, bool EXACT, bool VI = false>
__global__ __launch_bounds__(HB_BLOCK) void kb_spmv(const hb_dev *__restrict__ devs,
                                                    const int *__restrict__ tab, const int *__restrict__ cnt)
{
    /* hardware order: problems finish at different iterations, and the
     * tiles still working must stay spread over all XCDs (an XCD-contiguous
     * order measured 14% slower on the 256-problem batch) */
    const int nt = *cnt;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const hb_dev &d = devs[tab[2 * t]];
        spmv_tile<MODE, EXACT, VI>(d, tab[2 * t + 1]);
    }
}

template <int MODE, bool EXACT>
__global__ __launch_bounds__(HB_BLOCK) void kb_fin_spmv(const hb_dev *__restrict__ devs)
{
    const hb_dev &d = devs[blockIdx.x];
    fin_spmv_body<MODE, EXACT>(d);
}

__global__ __launch_bounds__(HB_BLOCK) void kb_seed_update(const hb_dev *__restrict__ devs,
                                                           const int *__restrict__ tab,
                                                           const int *__restrict__ cnt)
{
    const int nt = *cnt;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) seed_update_tile(devs[tab[2 * t]], tab[2 * t + 1]);
}

template <bool EXACT>
__global__ __launch_bounds__(HB_BLOCK) void kb_fin_seed(const hb_dev *__restrict__ devs)
{
    const hb_dev &d = devs[blockIdx.x];
    fin_seed_body<EXACT>(d);
}

/* shift tiles: tab[2t+1] = by * gx + bx of the problem's (gx, gy) grid */
template <int WM, int V, int U>
__global__ __launch_bounds__(HB_BLOCK) void kb_shift(const hb_dev *__restrict__ devs,
                                                     const int *__restrict__ tab, const int *__restrict__ cnt)
{
    const int nt = *cnt;
    for (int e = blockIdx.x; e < nt; e += gridDim.x) {
        const hb_dev &d = devs[tab[2 * e]];
        const int t = tab[2 * e + 1];
        shift_tile<WM, V, U>(d, t % d.shift_gx, t / d.shift_gx);
    }
}

/* combined-coefficient flush: tab[2t+1] = list position */
__global__ __launch_bounds__(HB_BLOCK) void kb_flush_coef(const hb_dev *__restrict__ devs,
                                                          const int *__restrict__ tab,
                                                          const int *__restrict__ cnt)
{
    const int nt = *cnt;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const hb_dev &d = devs[tab[2 * t]];
        flush_coef_body(d, tab[2 * t + 1], d.coef, d.ccoef);
    }
}

/* tab[2t+1] = 256-element tile (the problem's shift grid has one row) */
__global__ __launch_bounds__(HB_BLOCK) void kb_shift_cc(const hb_dev *__restrict__ devs,
                                                        const int *__restrict__ tab, const int *__restrict__ cnt)
{
    const int nt = *cnt;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) shift_cc_tile(devs[tab[2 * t]], tab[2 * t + 1]);
}

/* Active sums over the problems still running: workgroup-wide, every
 * thread gets the sums of problems before `p` (pre) and of all (tot). */
DEV void active_sums(const hb_batch &B, int p, int pre[HB_NPH], int tot[HB_NPH], bool &p_active)
{
    __shared__ int sh[2 * HB_NPH][HB_BLOCK / 64];
    __shared__ int sh_act;
    const int tid = threadIdx.x;
    int a[2 * HB_NPH] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (tid == 0) sh_act = 0;
    __syncthreads();
    for (int q = tid; q < B.nprob; q += HB_BLOCK) {
        if (B.devs[q].st->done) continue;
        if (q == p) sh_act = 1;
#pragma unroll
        for (int ph = 0; ph < HB_NPH; ++ph) {
            const int c = B.pcnt[ph * B.nprob + q];
            a[ph] += q < p ? c : 0;
            a[HB_NPH + ph] += c;
        }
    }
#pragma unroll
    for (int i = 0; i < 2 * HB_NPH; ++i) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a[i] += __shfl_xor(a[i], o, 64);
        if ((tid & 63) == 0) sh[i][tid >> 6] = a[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < HB_NPH; ++i) {
        pre[i] = 0;
        tot[i] = 0;
#pragma unroll
        for (int w = 0; w < HB_BLOCK / 64; ++w) {
            pre[i] += sh[i][w];
            tot[i] += sh[HB_NPH + i][w];
        }
    }
    p_active = sh_act != 0;
    __syncthreads();
}

/* Workgroup p copies problem p's entries of every table into the active
 * lists, after those of the running problems before it; workgroup 0 stores
 * the lists' lengths.  The done flags do not change during the launch. */
__global__ __launch_bounds__(HB_BLOCK) void kb_compact(hb_batch B)
{
    const int p = blockIdx.x;
    int pre[HB_NPH], tot[HB_NPH];
    bool act;
    active_sums(B, p, pre, tot, act);
    if (p == 0 && threadIdx.x < HB_NPH) B.act[threadIdx.x] = tot[threadIdx.x];
    if (!act) return;
    const int *full[HB_NPH] = {B.t_spmv, B.t_vec, B.t_shift, B.t_coef};
#pragma unroll
    for (int ph = 0; ph < HB_NPH; ++ph) {
        const int *src = full[ph] + 2 * (size_t)B.pofs[ph * B.nprob + p];
        int *dst = B.a_tab[ph] + 2 * (size_t)pre[ph];
        const int m = 2 * B.pcnt[ph * B.nprob + p];
        for (int i = threadIdx.x; i < m; i += HB_BLOCK) dst[i] = src[i];
    }
}

__global__ __launch_bounds__(HB_BLOCK) void kb_init(const hb_dev *__restrict__ devs,
                                                    const int *__restrict__ tab)
{
    const hb_dev &d = devs[tab[2 * blockIdx.x]];
    init_body(d, tab[2 * blockIdx.x + 1]);
}

template <bool EXACT>
__global__ __launch_bounds__(HB_BLOCK) void kb_init_fin(const hb_dev *__restrict__ devs)
{
    const hb_dev &d = devs[blockIdx.x];
    init_fin_body<EXACT>(d);
}

__global__ __launch_bounds__(HB_BLOCK) void kb_finalize(const hb_dev *__restrict__ devs,
                                                        const int *__restrict__ tab)
{
    const hb_dev &d = devs[tab[2 * blockIdx.x]];
    finalize_body(d, tab[2 * blockIdx.x + 1]);
}

/* flag[0] = 1 once every problem is done; flag[1] = problems still running;
 * flag[2 + ph] = entries of phase ph's table those problems hold */
__global__ __launch_bounds__(HB_BLOCK) void kb_publish(hb_batch B, int *flag)
{
    __shared__ int cnt[HB_BLOCK];
    int c = 0;
    for (int q = threadIdx.x; q < B.nprob; q += HB_BLOCK) c += B.devs[q].st->done ? 0 : 1;
    cnt[threadIdx.x] = c;
    __syncthreads();
    for (int w = HB_BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) cnt[threadIdx.x] += cnt[threadIdx.x + w];
        __syncthreads();
    }
    int pre[HB_NPH], tot[HB_NPH];
    bool act;
    active_sums(B, 0, pre, tot, act);
    if (threadIdx.x == 0) {
        flag[0] = cnt[0] == 0;
        flag[1] = cnt[0];
        for (int ph = 0; ph < HB_NPH; ++ph) flag[2 + ph] = tot[ph];
        __threadfence_system();
    }
}

/* ---- launch layer --------------------------------------------------------- */
/* One SpMV as d->npanel column-panel passes (k_spmv_panel). */
template <int MODE, bool EXACT, bool UNF>
static void spmv_panels(const hb_dev *d, hipStream_t s)
{
    const dim3 bl(HB_BLOCK);
    const int P = d->npanel;
    hipLaunchKernelGGL((k_spmv_panel<MODE, EXACT, UNF, PASS_FIRST>), dim3(d->pnblk[0]), bl, 0, s, *d, 0);
    for (int p = 1; p + 1 < P; ++p)
        hipLaunchKernelGGL((k_spmv_panel<MODE, EXACT, UNF, PASS_MID>), dim3(d->pnblk[p]), bl, 0, s, *d, p);
    hipLaunchKernelGGL((k_spmv_panel<MODE, EXACT, UNF, PASS_LAST>), dim3(d->pnblk[P - 1]), bl, 0, s, *d,
                       P - 1);
}

template <int WM, int V, int U>
static void launch_shift(const hb_dev *d, hipStream_t s)
{
    const int per = HB_BLOCK * V;
    const dim3 g((d->n + per - 1) / per, d->shift_gy), b(HB_BLOCK);
    hipLaunchKernelGGL((k_shift<WM, V, U>), g, b, 0, s, *d);
}

/* workgroups of a phase launch over a table of `full` entries */
static unsigned kb_grid(const hb_batch *B, int full)
{
    const int d = B->div > 1 ? B->div : 1;
    const int g = (full + d - 1) / d;
    return (unsigned)(g > 0 ? g : 1);
}

template <int WM>
static void launch_shift_b(const hb_batch *B, hipStream_t s)
{
    constexpr int U = WM >= 8 ? 2 : 4;
    hipLaunchKernelGGL((kb_shift<WM, 1, U>), dim3(kb_grid(B, B->n_shift_tiles)), dim3(HB_BLOCK), 0, s,
                       B->devs, B->a_tab[HB_PH_SHIFT], B->act + HB_PH_SHIFT);
}

extern "C" {

int hb_k_init(const hb_dev *d, hipStream_t s)
{
    hipLaunchKernelGGL(k_init, dim3(d->vec_grid), dim3(HB_BLOCK), 0, s, *d);
    if (d->exact) hipLaunchKernelGGL(k_init_fin<true>, dim3(1), dim3(HB_BLOCK), 0, s, *d);
    else hipLaunchKernelGGL(k_init_fin<false>, dim3(1), dim3(HB_BLOCK), 0, s, *d);
    return (int)hipGetLastError();
}

/* The fused default-mode iteration: SpMV 1; SpMV 2 with the alpha~
 * finalize as its first workgroup (fused_a; else the finalize keeps a launch
 * of its own); the seed update with the omega~ finalize as its first
 * workgroup; the shift kernel with the scalar stage as its first workgroup
 * (fused_s; else the stage keeps a launch of its own) and the combined
 * coefficients in the next ones.  The stages folded into another launch
 * enqueue nothing. */
static int stage_fused(const hb_dev *d, int stage, hipStream_t s)
{
    const dim3 bl(HB_BLOCK);
    switch (stage) {
    case HB_STAGE_FIN1:
        break;
    case HB_STAGE_SCALAR:
        if (!d->fused_s) hipLaunchKernelGGL(k_fin_seed<false>, dim3(1), bl, 0, s, *d);
        break;
    case HB_STAGE_SPMV2:
        if (d->cdia) hipLaunchKernelGGL((k_spmv2_fa<true, true>), dim3(d->nblk + 1), bl, 0, s, *d);
        else if (d->vidx) hipLaunchKernelGGL(k_spmv2_fa<true>, dim3(d->nblk + 1), bl, 0, s, *d);
        else hipLaunchKernelGGL(k_spmv2_fa<false>, dim3(d->nblk + 1), bl, 0, s, *d);
        break;
    case HB_STAGE_SEED:
        hipLaunchKernelGGL(k_seed_update_fin, dim3(d->vec_grid + 1), bl, 0, s, *d);
        break;
    case HB_STAGE_SHIFT:
        if (d->fused_s) hipLaunchKernelGGL(k_shift_stage, dim3(1 + d->M + d->shift_tw), bl, 0, s, *d);
        else hipLaunchKernelGGL(k_shift_fused, dim3(d->M + d->shift_gx), bl, 0, s, *d);
        break;
    case HB_STAGE_FIN2:
    case HB_STAGE_COEF:
        break;
    default:
        return -1;
    }
    return (int)hipGetLastError();
}

int hb_k_stage(const hb_dev *d, int stage, hipStream_t s)
{
    const dim3 gs(d->nblk), gv(d->vec_grid), bl(HB_BLOCK), one(1);
    if (d->fused && !d->exact &&
        (stage == HB_STAGE_FIN2 || stage == HB_STAGE_SEED || stage == HB_STAGE_COEF || stage == HB_STAGE_SHIFT ||
         stage == HB_STAGE_SCALAR || (d->fused_a && (stage == HB_STAGE_FIN1 || stage == HB_STAGE_SPMV2))))
        return stage_fused(d, stage, s);
    if (d->npanel > 1 && (stage == HB_STAGE_SPMV1 || stage == HB_STAGE_SPMV2)) {
        if (stage == HB_STAGE_SPMV2 && d->q_unfused) hipLaunchKernelGGL(k_qform, gv, bl, 0, s, *d);
        if (stage == HB_STAGE_SPMV1) {
            if (d->exact) spmv_panels<SP_SEED1, true, false>(d, s);
            else spmv_panels<SP_SEED1, false, false>(d, s);
        } else if (d->exact) {
            spmv_panels<SP_SEED2, true, false>(d, s);
        } else if (d->q_unfused) {
            spmv_panels<SP_SEED2, false, true>(d, s);
        } else {
            spmv_panels<SP_SEED2, false, false>(d, s);
        }
        return (int)hipGetLastError();
    }
    switch (stage) {
    case HB_STAGE_SPMV1:
        if (d->exact) {
            if (d->cdia) hipLaunchKernelGGL((k_spmv<SP_SEED1, true, true, false, true>), gs, bl, 0, s, *d);
            else if (d->vidx) hipLaunchKernelGGL((k_spmv<SP_SEED1, true, true>), gs, bl, 0, s, *d);
            else hipLaunchKernelGGL((k_spmv<SP_SEED1, true, false>), gs, bl, 0, s, *d);
        } else {
            if (d->cdia) hipLaunchKernelGGL((k_spmv<SP_SEED1, false, true, false, true>), gs, bl, 0, s, *d);
            else if (d->vidx) hipLaunchKernelGGL((k_spmv<SP_SEED1, false, true>), gs, bl, 0, s, *d);
            else hipLaunchKernelGGL((k_spmv<SP_SEED1, false, false>), gs, bl, 0, s, *d);
        }
        break;
    case HB_STAGE_FIN1:
        if (d->exact) hipLaunchKernelGGL((k_fin_spmv<SP_SEED1, true>), one, bl, 0, s, *d);
        else hipLaunchKernelGGL((k_fin_spmv<SP_SEED1, false>), one, bl, 0, s, *d);
        break;
    case HB_STAGE_SPMV2:
        if (d->q_unfused && !d->exact) {
            hipLaunchKernelGGL(k_qform, gv, bl, 0, s, *d);
            if (d->cdia) hipLaunchKernelGGL((k_spmv<SP_SEED2, false, true, true, true>), gs, bl, 0, s, *d);
            else if (d->vidx) hipLaunchKernelGGL((k_spmv<SP_SEED2, false, true, true>), gs, bl, 0, s, *d);
            else hipLaunchKernelGGL((k_spmv<SP_SEED2, false, false, true>), gs, bl, 0, s, *d);
        } else if (d->exact) {
            if (d->cdia) hipLaunchKernelGGL((k_spmv<SP_SEED2, true, true, false, true>), gs, bl, 0, s, *d);
            else if (d->vidx) hipLaunchKernelGGL((k_spmv<SP_SEED2, true, true>), gs, bl, 0, s, *d);
            else hipLaunchKernelGGL((k_spmv<SP_SEED2, true, false>), gs, bl, 0, s, *d);
        } else {
            if (d->cdia) hipLaunchKernelGGL((k_spmv<SP_SEED2, false, true, false, true>), gs, bl, 0, s, *d);
            else if (d->vidx) hipLaunchKernelGGL((k_spmv<SP_SEED2, false, true>), gs, bl, 0, s, *d);
            else hipLaunchKernelGGL((k_spmv<SP_SEED2, false, false>), gs, bl, 0, s, *d);
        }
        break;
    case HB_STAGE_FIN2:
        if (d->exact) hipLaunchKernelGGL((k_fin_spmv<SP_SEED2, true>), one, bl, 0, s, *d);
        else hipLaunchKernelGGL((k_fin_spmv<SP_SEED2, false>), one, bl, 0, s, *d);
        break;
    case HB_STAGE_SEED:
        hipLaunchKernelGGL(k_seed_update, gv, bl, 0, s, *d);
        break;
    case HB_STAGE_SCALAR:
        if (d->exact) hipLaunchKernelGGL(k_fin_seed<true>, one, bl, 0, s, *d);
        else hipLaunchKernelGGL(k_fin_seed<false>, one, bl, 0, s, *d);
        break;
    case HB_STAGE_COEF:
        /* combined coefficients: one workgroup per list position (the exact
         * replay needs none) */
        if (!d->exact) hipLaunchKernelGGL(k_flush_coef, dim3(d->M), bl, 0, s, *d);
        break;
    default:
        if (!d->exact) {
            /* the streaming flush over 256-element tiles */
            hipLaunchKernelGGL(k_shift_cc, dim3(d->shift_gx), bl, 0, s, *d);
            break;
        }
        /* exact order: windows up to 4 move 16-byte pairs; longer windows
         * hold one element per thread (the window's r and q stay in
         * registers) with two shifts in flight: four cost registers
         * (occupancy 4 instead of 5 waves per SIMD) and measured 5% slower
         * at 128^3, 64 shifts */
        if (d->W <= 1) {
            if (d->vec2) launch_shift<1, 2, 2>(d, s);
            else launch_shift<1, 1, 4>(d, s);
        } else if (d->W <= 4) {
            if (d->vec2) launch_shift<4, 2, 2>(d, s);
            else launch_shift<4, 1, 4>(d, s);
        } else if (d->W <= 8) {
            launch_shift<8, 1, 2>(d, s);
        } else {
            launch_shift<16, 1, 2>(d, s);
        }
        break;
    }
    return (int)hipGetLastError();
}

int hb_k_stage_runs(const hb_dev *d, int stage)
{
    if (stage == HB_STAGE_COEF) return !d->exact && !d->fused;
    if (d->exact || !d->fused) return 1;
    if (stage == HB_STAGE_FIN2) return 0;
    if (stage == HB_STAGE_FIN1) return !d->fused_a;
    if (stage == HB_STAGE_SCALAR) return !d->fused_s;
    return 1;
}

int hb_k_shift_stage_tiles(int n_tiles, int M, int device)
{
    /* as many tile workgroups as stay resident next to the stage and the
     * coefficient workgroups, at most one per flush tile; 0 when the M
     * coefficient workgroups would leave fewer than three tiles per compute
     * unit (the flushes inside the kernel then lack the wavefronts to hide
     * their loads: 1024 shifts at 128^3 took 670 against 275 ms) */
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_shift_stage, HB_BLOCK, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return -1;
    long long t = (long long)per_cu * cus - 1 - M;
    if (t < n_tiles && t < 3LL * cus) return 0;
    if (t > n_tiles) t = n_tiles;
    return t > 0 ? (int)t : 1;
}

int hb_k_iteration(const hb_dev *d, hipStream_t s)
{
    for (int st = 0; st < HB_NSTAGE; ++st) {
        const int e = hb_k_stage(d, st, s);
        if (e) return e;
    }
    return 0;
}

int hb_k_publish(const hb_dev *d, int *flag_dev, hipStream_t s)
{
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, s, d->st, flag_dev);
    return (int)hipGetLastError();
}

int hb_k_finalize(const hb_dev *d, hipStream_t s)
{
    if (d->fused_s) hipLaunchKernelGGL(k_flush_final, dim3(d->shift_gx), dim3(HB_BLOCK), 0, s, *d);
    hipLaunchKernelGGL(k_finalize, dim3(d->vec_grid), dim3(HB_BLOCK), 0, s, *d);
    return (int)hipGetLastError();
}

int hb_k_resid_blocks(int n, int M)
{
    const int nb = (n + HB_BLOCK - 1) / HB_BLOCK;
    const int cap = M > 0 && M <= HB_RESID_PAIRS ? HB_RESID_PAIRS / M : 1;
    return nb < cap ? nb : cap;
}

int hb_k_true_residual(const hb_dev *d, const double *b, const double *sigma, const double *x,
                       double *part, double *out, hipStream_t s)
{
    const int R = hb_k_resid_blocks(d->n, d->M);
    hipLaunchKernelGGL(k_resid_part, dim3(R, d->M), dim3(HB_BLOCK), 0, s, *d, b, sigma, x, part);
    hipLaunchKernelGGL(k_resid_final, dim3(d->M), dim3(HB_BLOCK), 0, s, part, R, out);
    return (int)hipGetLastError();
}

int hb_k_fill(double *p, size_t count, double v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(256), dim3(HB_BLOCK), 0, s, p, count, v);
    return (int)hipGetLastError();
}

int hb_k_compress_cols(const hb_dev *d, short *col16, unsigned *cflag, hipStream_t s)
{
    hipLaunchKernelGGL(k_compress_cols, dim3(d->nblk), dim3(HB_BLOCK), 0, s, d->ptr, d->col, d->blk,
                       col16, cflag);
    return (int)hipGetLastError();
}

int hb_k_spmv(const hb_dev *d, const double *x, double *y, double shift, hipStream_t s)
{
    if (d->exact) hipLaunchKernelGGL(k_spmv_plain<true>, dim3(d->nblk), dim3(HB_BLOCK), 0, s, *d, x, y, shift);
    else hipLaunchKernelGGL(k_spmv_plain<false>, dim3(d->nblk), dim3(HB_BLOCK), 0, s, *d, x, y, shift);
    return (int)hipGetLastError();
}

/* ---- batched ---- */
int hb_kb_init(const hb_batch *B, hipStream_t s)
{
    hipLaunchKernelGGL(kb_init, dim3(B->n_vec_tiles), dim3(HB_BLOCK), 0, s, B->devs, B->t_vec);
    if (B->exact) hipLaunchKernelGGL(kb_init_fin<true>, dim3(B->nprob), dim3(HB_BLOCK), 0, s, B->devs);
    else hipLaunchKernelGGL(kb_init_fin<false>, dim3(B->nprob), dim3(HB_BLOCK), 0, s, B->devs);
    return (int)hipGetLastError();
}

int hb_kb_stage(const hb_batch *B, int stage, hipStream_t s)
{
    const dim3 bl(HB_BLOCK), gp(B->nprob);
    const dim3 gs(kb_grid(B, B->n_spmv_tiles)), gv(kb_grid(B, B->n_vec_tiles));
    const int *ts = B->a_tab[HB_PH_SPMV], *cs = B->act + HB_PH_SPMV;
    switch (stage) {
    case HB_STAGE_SPMV1:
        if (B->exact && B->vi) hipLaunchKernelGGL((kb_spmv<SP_SEED1, true, true>), gs, bl, 0, s, B->devs, ts, cs);
        else if (B->exact) hipLaunchKernelGGL((kb_spmv<SP_SEED1, true, false>), gs, bl, 0, s, B->devs, ts, cs);
        else if (B->vi) hipLaunchKernelGGL((kb_spmv<SP_SEED1, false, true>), gs, bl, 0, s, B->devs, ts, cs);
        else hipLaunchKernelGGL((kb_spmv<SP_SEED1, false, false>), gs, bl, 0, s, B->devs, ts, cs);
        break;
    case HB_STAGE_FIN1:
        if (B->exact) hipLaunchKernelGGL((kb_fin_spmv<SP_SEED1, true>), gp, bl, 0, s, B->devs);
        else hipLaunchKernelGGL((kb_fin_spmv<SP_SEED1, false>), gp, bl, 0, s, B->devs);
        break;
    case HB_STAGE_SPMV2:
        if (B->exact && B->vi) hipLaunchKernelGGL((kb_spmv<SP_SEED2, true, true>), gs, bl, 0, s, B->devs, ts, cs);
        else if (B->exact) hipLaunchKernelGGL((kb_spmv<SP_SEED2, true, false>), gs, bl, 0, s, B->devs, ts, cs);
        else if (B->vi) hipLaunchKernelGGL((kb_spmv<SP_SEED2, false, true>), gs, bl, 0, s, B->devs, ts, cs);
        else hipLaunchKernelGGL((kb_spmv<SP_SEED2, false, false>), gs, bl, 0, s, B->devs, ts, cs);
        break;
    case HB_STAGE_FIN2:
        if (B->exact) hipLaunchKernelGGL((kb_fin_spmv<SP_SEED2, true>), gp, bl, 0, s, B->devs);
        else hipLaunchKernelGGL((kb_fin_spmv<SP_SEED2, false>), gp, bl, 0, s, B->devs);
        break;
    case HB_STAGE_SEED:
        hipLaunchKernelGGL(kb_seed_update, gv, bl, 0, s, B->devs, B->a_tab[HB_PH_VEC], B->act + HB_PH_VEC);
        break;
    case HB_STAGE_SCALAR:
        if (B->exact) hipLaunchKernelGGL(kb_fin_seed<true>, gp, bl, 0, s, B->devs);
        else hipLaunchKernelGGL(kb_fin_seed<false>, gp, bl, 0, s, B->devs);
        break;
    case HB_STAGE_COEF:
        if (!B->exact)
            hipLaunchKernelGGL(kb_flush_coef, dim3(kb_grid(B, B->n_coef_tiles)), bl, 0, s, B->devs,
                               B->a_tab[HB_PH_COEF], B->act + HB_PH_COEF);
        break;
    default:
        if (!B->exact) {
            hipLaunchKernelGGL(kb_shift_cc, dim3(kb_grid(B, B->n_shift_tiles)), bl, 0, s, B->devs,
                               B->a_tab[HB_PH_SHIFT], B->act + HB_PH_SHIFT);
            break;
        }
        if (B->W <= 1) launch_shift_b<1>(B, s);
        else if (B->W <= 4) launch_shift_b<4>(B, s);
        else if (B->W <= 8) launch_shift_b<8>(B, s);
        else launch_shift_b<16>(B, s);
        break;
    }
    return (int)hipGetLastError();
}

int hb_kb_iteration(const hb_batch *B, hipStream_t s)
{
    for (int st = 0; st < HB_NSTAGE; ++st) {
        const int e = hb_kb_stage(B, st, s);
        if (e) return e;
    }
    return 0;
}

int hb_kb_publish(const hb_batch *B, int *flag_dev, hipStream_t s)
{
    hipLaunchKernelGGL(kb_publish, dim3(1), dim3(HB_BLOCK), 0, s, *B, flag_dev);
    return (int)hipGetLastError();
}

int hb_kb_compact(const hb_batch *B, hipStream_t s)
{
    hipLaunchKernelGGL(kb_compact, dim3(B->nprob), dim3(HB_BLOCK), 0, s, *B);
    return (int)hipGetLastError();
}

int hb_kb_finalize(const hb_batch *B, hipStream_t s)
{
    hipLaunchKernelGGL(kb_finalize, dim3(B->n_vec_tiles), dim3(HB_BLOCK), 0, s, B->devs, B->t_vec);
    return (int)hipGetLastError();
}

} /* extern "C" */
