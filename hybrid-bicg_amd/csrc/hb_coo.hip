/*
 * hb_coo.hip -- COO -> CSR on the GPU, for Matrix Market input.
 *
 * The reference orders the entries by row with a stable merge sort and then
 * counts rows (src/matrix.c:125-187, 206-232): rows ascending, entries of one
 * row in file order.  Here the same order comes from an LSD radix sort of the
 * row indices (8-bit digits, as many passes as the row count needs), whose
 * passes are stable, carrying each entry's position; col and val are then
 * gathered through the sorted positions, and ptr[i] is the first sorted
 * position whose row is >= i (a binary search per row).
 *
 * One radix pass = three launches:
 *   rs_hist     per tile of RS_TILE entries, the count of each digit, stored
 *               digit-major: hist[d * ntiles + tile]
 *   scan        exclusive scan of hist (3 launches, below): the first output
 *               position of digit d within tile t
 *   rs_scatter  per tile, every entry's stable rank among the tile's entries
 *               of its digit: each wavefront owns RS_TILE/4 consecutive
 *               entries and walks them 64 at a time; lanes with the same
 *               digit find each other with 8 ballots, the lowest such lane
 *               adds the group to the wavefront's LDS counter for the digit
 *               and hands the old count to its group.
 * All counters are 32-bit: nz < 2^32.
 */
#include <hip/hip_runtime.h>
#include <stdlib.h>

#include "hb_shared.h"
#include "../../include/hbicg.h"

#define RS_BITS   8
#define RS_DIGITS (1 << RS_BITS)
#define RS_WAVES  (HB_BLOCK / 64)
#define RS_TILE   (HB_BLOCK * 16)          /* entries per workgroup      */
#define RS_WITEMS (RS_TILE / RS_WAVES)      /* entries per wavefront       */
#define SC_TILE   (HB_BLOCK * 16)          /* scan: entries per workgroup */

static_assert(RS_DIGITS == HB_BLOCK, "one thread per digit");

typedef unsigned long long u64;

/* lanes of the wavefront whose valid flag is set and whose digit equals
 * this lane's */
__device__ __forceinline__ u64 same_digit(unsigned d, bool valid)
{
    u64 peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < RS_BITS; ++b) {
        const bool bit = (d >> b) & 1u;
        const u64 m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

__global__ __launch_bounds__(HB_BLOCK) void k_rs_hist(const unsigned *__restrict__ keys, unsigned nz,
                                                      int shift, unsigned *__restrict__ hist,
                                                      unsigned ntiles)
{
    __shared__ unsigned cnt[RS_DIGITS];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * RS_TILE;
    const size_t end = base + RS_TILE < nz ? base + RS_TILE : nz;
    for (size_t i = base + threadIdx.x; i < end; i += HB_BLOCK)
        atomicAdd(&cnt[(keys[i] >> shift) & (RS_DIGITS - 1)], 1u);
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

/* vals_in == NULL: the values are the entry positions 0 .. nz-1 */
__global__ __launch_bounds__(HB_BLOCK) void k_rs_scatter(const unsigned *__restrict__ keys_in,
                                                         const unsigned *__restrict__ vals_in,
                                                         unsigned *__restrict__ keys_out,
                                                         unsigned *__restrict__ vals_out,
                                                         unsigned nz, int shift,
                                                         const unsigned *__restrict__ offs,
                                                         unsigned ntiles)
{
    __shared__ unsigned wcnt[RS_WAVES][RS_DIGITS];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int v = 0; v < RS_WAVES; ++v) wcnt[v][tid] = 0;
    __syncthreads();
    const size_t wbase = (size_t)blockIdx.x * RS_TILE + (size_t)w * RS_WITEMS;
    const u64 lt = (1ull << lane) - 1ull;

    /* count each digit per wavefront */
    for (int r = 0; r < RS_WITEMS / 64; ++r) {
        const size_t i = wbase + (size_t)r * 64 + lane;
        const bool valid = i < nz;
        const unsigned d = valid ? (keys_in[i] >> shift) & (RS_DIGITS - 1) : 0u;
        const u64 peers = same_digit(d, valid);
        if (valid && (peers & lt) == 0) atomicAdd(&wcnt[w][d], (unsigned)__popcll(peers));
    }
    __syncthreads();
    /* thread = digit: the tile's first position for the digit, then the
     * wavefronts' shares in order */
    {
        unsigned run = offs[(size_t)tid * ntiles + blockIdx.x];
#pragma unroll
        for (int v = 0; v < RS_WAVES; ++v) {
            const unsigned c = wcnt[v][tid];
            wcnt[v][tid] = run;
            run += c;
        }
    }
    __syncthreads();
    /* scatter in order: the group leader takes the group's positions from the
     * wavefront's counter and passes the first one to its group */
    for (int r = 0; r < RS_WITEMS / 64; ++r) {
        const size_t i = wbase + (size_t)r * 64 + lane;
        const bool valid = i < nz;
        const unsigned key = valid ? keys_in[i] : 0u;
        const unsigned d = (key >> shift) & (RS_DIGITS - 1);
        const u64 peers = same_digit(d, valid);
        const int leader = valid ? __ffsll((long long)peers) - 1 : lane;
        unsigned first = 0;
        if (valid && leader == lane) first = atomicAdd(&wcnt[w][d], (unsigned)__popcll(peers));
        first = __shfl(first, leader, 64);
        if (valid) {
            const unsigned dst = first + (unsigned)__popcll(peers & lt);
            keys_out[dst] = key;
            vals_out[dst] = vals_in ? vals_in[i] : (unsigned)i;
        }
    }
}

/* ---- exclusive scan of u32 ---------------------------------------------------- */
/* exclusive prefix of v over the workgroup; total in *tot (all threads) */
__device__ __forceinline__ unsigned block_exscan(unsigned v, unsigned *tot)
{
    __shared__ unsigned ws[RS_WAVES];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    unsigned before = 0, all = 0;
#pragma unroll
    for (int v2 = 0; v2 < RS_WAVES; ++v2) {
        const unsigned s = ws[v2];
        if (v2 < w) before += s;
        all += s;
    }
    __syncthreads();
    *tot = all;
    return before + inc - v;
}

/* bsum[b] = sum of in[b*SC_TILE .. ) */
__global__ __launch_bounds__(HB_BLOCK) void k_scan_sums(const unsigned *__restrict__ in, size_t len,
                                                        unsigned *__restrict__ bsum)
{
    const size_t base = (size_t)blockIdx.x * SC_TILE;
    unsigned s = 0;
    for (size_t i = base + threadIdx.x; i < base + SC_TILE && i < len; i += HB_BLOCK) s += in[i];
    unsigned tot;
    block_exscan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

/* exclusive scan of bsum[0 .. nb) in place, one workgroup */
__global__ __launch_bounds__(HB_BLOCK) void k_scan_top(unsigned *bsum, unsigned nb)
{
    unsigned carry = 0;
    for (unsigned base = 0; base < nb; base += HB_BLOCK) {
        const unsigned i = base + threadIdx.x;
        const unsigned v = i < nb ? bsum[i] : 0u;
        unsigned tot;
        const unsigned ex = block_exscan(v, &tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
    }
}

/* out[i] = bsum[tile] + exclusive prefix of in within the tile; thread t owns
 * 16 consecutive entries of the tile */
/* in and out may be the same array: each thread reads its entries before it
 * writes them */
__global__ __launch_bounds__(HB_BLOCK) void k_scan_apply(const unsigned *in, size_t len,
                                                         const unsigned *__restrict__ bsum,
                                                         unsigned *out)
{
    constexpr int PER = SC_TILE / HB_BLOCK;
    const size_t base = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * PER;
    unsigned v[PER];
    unsigned s = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        v[u] = base + u < len ? in[base + u] : 0u;
        s += v[u];
    }
    unsigned tot;
    unsigned run = bsum[blockIdx.x] + block_exscan(s, &tot);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        if (base + u < len) out[base + u] = run;
        run += v[u];
    }
}

/* ---- after the sort ------------------------------------------------------------ */
__global__ __launch_bounds__(HB_BLOCK) void k_coo_check(const unsigned *__restrict__ row,
                                                        const unsigned *__restrict__ col, unsigned nz,
                                                        unsigned rows, unsigned cols,
                                                        unsigned *__restrict__ bad)
{
    int b = 0;
    for (size_t i = (size_t)blockIdx.x * HB_BLOCK + threadIdx.x; i < nz; i += (size_t)gridDim.x * HB_BLOCK)
        if (row[i] >= rows || col[i] >= cols) b = 1;
    b = __syncthreads_or(b);
    if (b && threadIdx.x == 0) atomicOr(bad, 1u);
}

/* ocol[e] = col[perm[e]], oval[e] = val[perm[e]] (perm NULL: identity) */
__global__ __launch_bounds__(HB_BLOCK) void k_coo_gather(const unsigned *__restrict__ perm,
                                                         const unsigned *__restrict__ col,
                                                         const double *__restrict__ val, unsigned nz,
                                                         unsigned *__restrict__ ocol,
                                                         double *__restrict__ oval)
{
    for (size_t e = (size_t)blockIdx.x * HB_BLOCK + threadIdx.x; e < nz; e += (size_t)gridDim.x * HB_BLOCK) {
        const unsigned s = perm ? perm[e] : (unsigned)e;
        ocol[e] = col[s];
        oval[e] = val[s];
    }
}

/* ptr[i] = first position of skeys with a row >= i, i = 0 .. rows */
__global__ __launch_bounds__(HB_BLOCK) void k_coo_ptr(const unsigned *__restrict__ skeys, unsigned nz,
                                                      unsigned rows, unsigned *__restrict__ ptr)
{
    for (size_t i = (size_t)blockIdx.x * HB_BLOCK + threadIdx.x; i <= rows; i += (size_t)gridDim.x * HB_BLOCK) {
        size_t lo = 0, hi = nz;
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (skeys[mid] < i) lo = mid + 1;
            else hi = mid;
        }
        ptr[i] = (unsigned)lo;
    }
}

/* ---- host side -------------------------------------------------------------------- */
static unsigned grid_for(size_t items)
{
    size_t g = (items + HB_BLOCK - 1) / HB_BLOCK;
    if (g > 8192) g = 8192;
    return g ? (unsigned)g : 1u;
}

/* exclusive scan of len u32 from in to out (out may equal in); tmp holds
 * ceil(len / SC_TILE) words */
static hipError_t scan_u32(const unsigned *in, unsigned *out, size_t len, unsigned *tmp, hipStream_t s)
{
    const unsigned nb = (unsigned)((len + SC_TILE - 1) / SC_TILE);
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scan_sums, dim3(nb), dim3(HB_BLOCK), 0, s, in, len, tmp);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(HB_BLOCK), 0, s, tmp, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(HB_BLOCK), 0, s, in, len, (const unsigned *)tmp, out);
    return hipGetLastError();
}

extern "C" int hbicg_coo2csr_device(const unsigned *row, const unsigned *col, const double *val,
                                    unsigned nz, unsigned rows, unsigned cols, unsigned *ptr,
                                    unsigned *ocol, double *oval, void *stream)
{
    if (!ptr || rows == 0 || (nz > 0 && (!row || !col || !val || !ocol || !oval))) return HBICG_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    int passes = 0;
    for (unsigned top = rows - 1; top; top >>= RS_BITS) ++passes;
    const unsigned ntiles = (unsigned)(((size_t)nz + RS_TILE - 1) / RS_TILE);
    const size_t hlen = (size_t)RS_DIGITS * ntiles;
    const size_t scan_tmp = (hlen + SC_TILE - 1) / SC_TILE + 1;
    /* one allocation: flag, keys x2, positions x2, histogram, scan partials */
    const size_t words = 64 + (passes ? 4 * (size_t)nz + hlen + scan_tmp : 0);
    unsigned *buf = nullptr;
    if (hipMalloc((void **)&buf, words * sizeof(unsigned)) != hipSuccess) return HBICG_E_NOMEM;
    int rc = 0;
    unsigned bad = 0;
    unsigned *flag = buf;
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(unsigned), s);
    if (e == hipSuccess && nz > 0)
        hipLaunchKernelGGL(k_coo_check, dim3(grid_for(nz)), dim3(HB_BLOCK), 0, s, row, col, nz, rows, cols, flag);
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, flag, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { rc = HBICG_E_HIP; goto done; }
    if (bad) { rc = HBICG_E_ARG; goto done; }

    if (passes == 0 || nz == 0) {
        /* one row (every index 0, already in order) or no entries: the input
         * order is the CSR order */
        hipLaunchKernelGGL(k_coo_ptr, dim3(grid_for((size_t)rows + 1)), dim3(HB_BLOCK), 0, s, row, nz, rows, ptr);
        if (nz > 0)
            hipLaunchKernelGGL(k_coo_gather, dim3(grid_for(nz)), dim3(HB_BLOCK), 0, s,
                               (const unsigned *)nullptr, col, val, nz, ocol, oval);
    } else {
        unsigned *k0 = buf + 64, *k1 = k0 + nz, *v0 = k1 + nz, *v1 = v0 + nz;
        unsigned *hist = v1 + nz, *tmp = hist + hlen;
        const unsigned *kin = row, *vin = nullptr;
        unsigned *kout = k0, *vout = v0;
        for (int p = 0; p < passes && e == hipSuccess; ++p) {
            const int shift = p * RS_BITS;
            hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(HB_BLOCK), 0, s, kin, nz, shift, hist, ntiles);
            e = scan_u32(hist, hist, hlen, tmp, s);
            if (e != hipSuccess) break;
            hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(HB_BLOCK), 0, s, kin, vin, kout, vout, nz,
                               shift, (const unsigned *)hist, ntiles);
            kin = kout;
            vin = vout;
            kout = (kout == k0) ? k1 : k0;
            vout = (vout == v0) ? v1 : v0;
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_coo_gather, dim3(grid_for(nz)), dim3(HB_BLOCK), 0, s, vin, col, val, nz, ocol, oval);
            hipLaunchKernelGGL(k_coo_ptr, dim3(grid_for((size_t)rows + 1)), dim3(HB_BLOCK), 0, s, kin, nz, rows, ptr);
        }
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = HBICG_E_HIP;
done:
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    return rc;
}

/* Matrix Market reader of the host library (src/matrix.c:26-94 semantics) */
extern "C" int hb_mm_read_coo(const char *filename, COO_Matrix *coo, char *code);
extern "C" void coo_free_matrix(COO_Matrix *m);

extern "C" void hbicg_csr_free_device(hbicg_csr *A)
{
    if (!A) return;
    (void)hipFree((void *)A->val);
    (void)hipFree((void *)A->col);
    (void)hipFree((void *)A->ptr);
    A->val = nullptr;
    A->col = nullptr;
    A->ptr = nullptr;
    A->n = A->nz = 0;
}

extern "C" int hbicg_csr_load_device(const char *filename, hbicg_csr *A, INFO_Matrix *info, void *stream)
{
    if (!filename || !A) return HBICG_E_ARG;
    A->val = nullptr;
    A->col = nullptr;
    A->ptr = nullptr;
    A->n = A->nz = 0;
    COO_Matrix coo;
    coo.val = nullptr;
    coo.row = coo.col = nullptr;
    coo.nz = coo.rows = coo.cols = 0;
    char code[4];
    if (hb_mm_read_coo(filename, &coo, code) != 0) return HBICG_E_ARG;
    if (coo.rows != coo.cols) { coo_free_matrix(&coo); return HBICG_E_NOTSQUARE; }
    hipStream_t s = (hipStream_t)stream;
    const size_t nz = coo.nz, n = coo.rows;
    unsigned *drow = nullptr, *dcol = nullptr, *ptr = nullptr, *col = nullptr;
    double *dval = nullptr, *val = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&drow, (nz + 1) * 4) != hipSuccess || hipMalloc((void **)&dcol, (nz + 1) * 4) != hipSuccess ||
        hipMalloc((void **)&dval, (nz + 1) * 8) != hipSuccess || hipMalloc((void **)&ptr, (n + 1) * 4) != hipSuccess ||
        hipMalloc((void **)&col, (nz + 1) * 4) != hipSuccess || hipMalloc((void **)&val, (nz + 1) * 8) != hipSuccess) {
        rc = HBICG_E_NOMEM;
    } else if (hipMemcpy(drow, coo.row, nz * 4, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(dcol, coo.col, nz * 4, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(dval, coo.val, nz * 8, hipMemcpyHostToDevice) != hipSuccess) {
        rc = HBICG_E_HIP;
    } else {
        rc = hbicg_coo2csr_device(drow, dcol, dval, (unsigned)nz, (unsigned)n, (unsigned)n, ptr, col, val, s);
    }
    (void)hipFree(drow);
    (void)hipFree(dcol);
    (void)hipFree(dval);
    if (rc == 0) {
        A->val = val;
        A->col = col;
        A->ptr = ptr;
        A->n = (unsigned)n;
        A->nz = (unsigned)nz;
        if (info) {
            info->nz = (unsigned)nz;
            info->rows = (unsigned)n;
            info->cols = (unsigned)n;
            for (int c = 0; c < 4; ++c) info->code[c] = code[c];
            if (info->recvcounts) info->recvcounts[0] = (int)n;
            if (info->displs) info->displs[0] = 0;
        }
    } else {
        (void)hipFree(ptr);
        (void)hipFree(col);
        (void)hipFree(val);
    }
    coo_free_matrix(&coo);
    return rc;
}
