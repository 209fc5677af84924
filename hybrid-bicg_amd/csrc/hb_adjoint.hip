/*
 * hb_adjoint.hip -- device CSR transpose and the gradient assembly of the
 * adjoint of a shifted solve (no counterpart in the reference, which has no
 * differentiation).
 *
 * For x_k = (A + sigma_k I)^{-1} b and a loss L(x_1..x_M), the adjoint
 * solutions lam_k = (A + sigma_k I)^{-T} dL/dx_k give
 *   dL/db        = sum_k lam_k
 *   dL/dsigma_k  = -lam_k . x_k
 *   dL/dA_ij     = -sum_k lam_k[i] x_k[j]     (on the sparsity pattern)
 * The adjoint solutions come from the shifted solver run on A^T.
 */
#include <hip/hip_runtime.h>
#include "hb_adjoint.h"

#define TB 256

/* ---- transpose: count, scan, scatter, per-row sort -------------------------- */
__global__ __launch_bounds__(TB) void k_zero_u(unsigned *__restrict__ a, int n)
{
    for (int i = blockIdx.x * TB + threadIdx.x; i < n; i += gridDim.x * TB) a[i] = 0u;
}

__global__ __launch_bounds__(TB) void k_count_cols(const unsigned *__restrict__ col, unsigned nnz,
                                                   unsigned *__restrict__ cnt)
{
    for (unsigned e = blockIdx.x * TB + threadIdx.x; e < nnz; e += gridDim.x * TB)
        atomicAdd(&cnt[col[e]], 1u);
}

/* Exclusive scan of cnt[0..n) into tptr[0..n], three passes: per-block
 * sums, a one-workgroup scan of the block sums, then the per-element pass. */
__global__ __launch_bounds__(TB) void k_scan_blocks(const unsigned *__restrict__ cnt, int n,
                                                    unsigned *__restrict__ bsum)
{
    __shared__ unsigned sh[TB];
    const int i = blockIdx.x * TB + threadIdx.x;
    sh[threadIdx.x] = i < n ? cnt[i] : 0u;
    __syncthreads();
    for (int w = TB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(TB) void k_scan_sums(unsigned *__restrict__ bsum, int nb)
{
    __shared__ unsigned sh[TB];
    unsigned carry = 0;
    for (int base = 0; base < nb; base += TB) {
        const int i = base + threadIdx.x;
        const unsigned v = i < nb ? bsum[i] : 0u;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < TB; o <<= 1) {
            const unsigned t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + sh[threadIdx.x] - v;   /* exclusive */
        const unsigned tot = sh[TB - 1];
        __syncthreads();
        carry += tot;
    }
}

__global__ __launch_bounds__(TB) void k_scan_final(const unsigned *__restrict__ cnt, int n,
                                                   const unsigned *__restrict__ bsum,
                                                   unsigned *__restrict__ tptr,
                                                   unsigned *__restrict__ fill)
{
    __shared__ unsigned sh[TB];
    const int i = blockIdx.x * TB + threadIdx.x;
    const unsigned v = i < n ? cnt[i] : 0u;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < TB; o <<= 1) {
        const unsigned t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) {
        const unsigned ex = bsum[blockIdx.x] + sh[threadIdx.x] - v;
        tptr[i] = ex;
        fill[i] = ex;
        if (i == n - 1) tptr[n] = ex + v;
    }
}

/* one thread per row of A; entries land in their column's segment of A^T
 * in arbitrary order, each with its source entry index e (sorted
 * afterwards).  VALS false: the pattern only (tval, val unused). */
template <bool VALS>
__global__ __launch_bounds__(TB) void k_scatter(const unsigned *__restrict__ ptr,
                                                const unsigned *__restrict__ col,
                                                const double *__restrict__ val, int n,
                                                unsigned *__restrict__ fill,
                                                unsigned *__restrict__ tcol,
                                                double *__restrict__ tval,
                                                unsigned *__restrict__ tsrc)
{
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    for (unsigned e = ptr[i]; e < ptr[i + 1]; ++e) {
        const unsigned pos = atomicAdd(&fill[col[e]], 1u);
        tcol[pos] = (unsigned)i;
        if (VALS) tval[pos] = val[e];
        tsrc[pos] = e;
    }
}

/* insertion sort of each row of A^T by source entry index e: entries of A
 * are numbered row by row, so the row of A^T lists its entries by ascending
 * source row and, for duplicate entries of one row, in their CSR order --
 * the order a stable transpose produces, whatever order the scatter's
 * atomics placed them in.  tsrc[e'] is then the entry of A that entry e' of
 * A^T holds. */
template <bool VALS>
__global__ __launch_bounds__(TB) void k_sort_rows(const unsigned *__restrict__ tptr, int n,
                                                  unsigned *__restrict__ tcol,
                                                  double *__restrict__ tval,
                                                  unsigned *__restrict__ tsrc)
{
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const unsigned lo = tptr[i], hi = tptr[i + 1];
    for (unsigned a = lo + 1; a < hi; ++a) {
        const unsigned c = tcol[a], e = tsrc[a];
        const double v = VALS ? tval[a] : 0.0;
        unsigned b = a;
        while (b > lo && tsrc[b - 1] > e) {
            tcol[b] = tcol[b - 1];
            if (VALS) tval[b] = tval[b - 1];
            tsrc[b] = tsrc[b - 1];
            --b;
        }
        tcol[b] = c;
        if (VALS) tval[b] = v;
        tsrc[b] = e;
    }
}

/* out[e] = val[src[e]]: A^T's values from A's through the source indices
 * of a pattern-only transpose */
__global__ __launch_bounds__(TB) void k_gather_vals(const double *__restrict__ val,
                                                    const unsigned *__restrict__ src, size_t nnz,
                                                    double *__restrict__ out)
{
    for (size_t e = (size_t)blockIdx.x * TB + threadIdx.x; e < nnz; e += (size_t)gridDim.x * TB)
        out[e] = val[src[e]];
}

/* ---- gradients -------------------------------------------------------------- */
/* lam_k = w_k * mu_k (w == nullptr: w_k = 1).  grad_b[i] = sum_k lam_k[i],
 * summed over k in ascending order. */
__global__ __launch_bounds__(TB) void k_grad_b(const double *__restrict__ mu,
                                               const double *__restrict__ w, int n, int M,
                                               double *__restrict__ gb)
{
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int k = 0; k < M; ++k) {
        const double l = mu[(size_t)k * n + i];
        acc += w ? w[k] * l : l;
    }
    gb[i] = acc;
}

/* partial dots lam_k . x_k per workgroup, then one workgroup per shift sums
 * them in order: grad_sigma[k] = -sum */
__global__ __launch_bounds__(TB) void k_grad_sigma_part(const double *__restrict__ mu,
                                                        const double *__restrict__ x, int n,
                                                        double *__restrict__ part)
{
    __shared__ double sh[TB];
    const int k = blockIdx.y;
    double acc = 0.0;
    for (int i = blockIdx.x * TB + threadIdx.x; i < n; i += gridDim.x * TB)
        acc += mu[(size_t)k * n + i] * x[(size_t)k * n + i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = TB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)k * gridDim.x + blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(TB) void k_grad_sigma_fin(const double *__restrict__ part, int nb,
                                                       const double *__restrict__ w,
                                                       double *__restrict__ gs)
{
    __shared__ double sh[TB];
    const int k = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += TB) acc += part[(size_t)k * nb + b];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = TB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) gs[k] = -(w ? w[k] * sh[0] : sh[0]);
}

/* grad_val[e] = -sum_k lam_k[row(e)] * x_k[col[e]], shifts in ascending
 * order.  One workgroup per 256 rows of A; its entries are taken 256 at a
 * time, one per thread, each thread finding its entry's row among the
 * tile's row pointers (LDS binary search), so the work per thread does not
 * follow the row lengths.  KC shifts' loads of lam and x are issued
 * together; the shifts past M in the last round are loaded clamped and
 * left out of the sum. */
__global__ __launch_bounds__(TB) void k_grad_vals(const unsigned *__restrict__ ptr,
                                                  const unsigned *__restrict__ col, int n, int M,
                                                  const double *__restrict__ mu,
                                                  const double *__restrict__ w,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ gv)
{
    constexpr int KC = 16;
    __shared__ unsigned rp[TB + 1];
    const int r0 = blockIdx.x * TB;
    const int R = min(TB, n - r0);
    for (int t = threadIdx.x; t <= R; t += TB) rp[t] = ptr[r0 + t];
    __syncthreads();
    const unsigned e1 = rp[R];
    for (unsigned e = rp[0] + threadIdx.x; e < e1; e += TB) {
        /* row: the largest r with rp[r] <= e (empty rows share a pointer
         * with the next row and lose to it) */
        int lo = 0, hi = R;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (rp[mid] <= e) lo = mid;
            else hi = mid;
        }
        const size_t i = (size_t)(r0 + lo), c = col[e];
        double acc = 0.0;
        for (int k0 = 0; k0 < M; k0 += KC) {
            double l[KC], xv[KC];
#pragma unroll
            for (int j = 0; j < KC; ++j) {
                const int k = min(k0 + j, M - 1);
                const double m = mu[(size_t)k * n + i];
                l[j] = w ? w[k] * m : m;
                xv[j] = x[(size_t)k * n + c];
            }
#pragma unroll
            for (int j = 0; j < KC; ++j) {
                const double t = acc - l[j] * xv[j];
                acc = k0 + j < M ? t : acc;
            }
        }
        gv[e] = acc;
    }
}

extern "C" {

int hb_transpose(int n, unsigned nnz, const unsigned *ptr, const unsigned *col, const double *val,
                 unsigned *tptr, unsigned *tcol, double *tval, unsigned *tsrc, unsigned *work,
                 hipStream_t s)
{
    /* work: cnt[n] | fill[n] | bsum[ceil(n/TB)] */
    const int nb = (n + TB - 1) / TB;
    unsigned *cnt = work, *fill = work + n, *bsum = work + 2 * (size_t)n;
    hipLaunchKernelGGL(k_zero_u, dim3(nb < 4096 ? nb : 4096), dim3(TB), 0, s, cnt, n);
    const int g = nnz ? (int)((nnz + TB - 1) / TB < 65536 ? (nnz + TB - 1) / TB : 65536) : 1;
    hipLaunchKernelGGL(k_count_cols, dim3(g), dim3(TB), 0, s, col, nnz, cnt);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(TB), 0, s, cnt, n, bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(TB), 0, s, bsum, nb);
    hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(TB), 0, s, cnt, n, bsum, tptr, fill);
    if (tval) {
        hipLaunchKernelGGL(k_scatter<true>, dim3(nb), dim3(TB), 0, s, ptr, col, val, n, fill, tcol, tval, tsrc);
        hipLaunchKernelGGL(k_sort_rows<true>, dim3(nb), dim3(TB), 0, s, tptr, n, tcol, tval, tsrc);
    } else {
        hipLaunchKernelGGL(k_scatter<false>, dim3(nb), dim3(TB), 0, s, ptr, col, val, n, fill, tcol, tval, tsrc);
        hipLaunchKernelGGL(k_sort_rows<false>, dim3(nb), dim3(TB), 0, s, tptr, n, tcol, tval, tsrc);
    }
    return (int)hipGetLastError();
}

int hb_gather_vals(size_t nnz, const double *val, const unsigned *src, double *out, hipStream_t s)
{
    if (nnz) {
        const size_t b = (nnz + TB - 1) / TB;
        hipLaunchKernelGGL(k_gather_vals, dim3(b < 65536 ? (unsigned)b : 65536u), dim3(TB), 0, s, val, src,
                           nnz, out);
    }
    return (int)hipGetLastError();
}

int hb_adjoint_grads(int n, int M, const unsigned *ptr, const unsigned *col, const double *mu,
                     const double *w, const double *x, double *grad_b, double *grad_sigma,
                     double *grad_val, double *work, hipStream_t s)
{
    const int nb = (n + TB - 1) / TB;
    if (grad_b) hipLaunchKernelGGL(k_grad_b, dim3(nb), dim3(TB), 0, s, mu, w, n, M, grad_b);
    if (grad_sigma) {
        const int pb = nb < 256 ? nb : 256;
        hipLaunchKernelGGL(k_grad_sigma_part, dim3(pb, M), dim3(TB), 0, s, mu, x, n, work);
        hipLaunchKernelGGL(k_grad_sigma_fin, dim3(M), dim3(TB), 0, s, work, pb, w, grad_sigma);
    }
    if (grad_val) hipLaunchKernelGGL(k_grad_vals, dim3(nb), dim3(TB), 0, s, ptr, col, n, M, mu, w, x, grad_val);
    return (int)hipGetLastError();
}

} /* extern "C" */
