// SpMV device code shared by the single-case and the per-case (batched)
// solvers: row products in the reference's order and the canonical chunk
// reduction of staged per-row terms.
#pragma once

#include "k_util.h"

namespace fvh {

// ---------------------------------------------------------------------------
// SpMV  y = sigma * A x  (optionally with the partial sums of w.y)
// ---------------------------------------------------------------------------
// Row i of A x, summed from the first entry in ascending column order.  The
// entries are taken in groups of kRowBatch (one group for the rows of 2-D
// quad and triangle meshes) with every load of a group issued before the
// first use, so a row costs two dependent memory latencies (col/val, then x)
// instead of two per entry.  Slots past the row's end re-read its last entry
// (in bounds, same cache line) and are dropped by a select.
constexpr int kRowBatch = 5;
// The row products read x through `xat(j)`: a plain array, or values formed
// on the fly from other arrays (CG's p = z + beta p).
template <typename V, class X>
__device__ inline V row_ax_at(const int *__restrict__ rowptr, const int *__restrict__ col, const V *__restrict__ val,
                              X xat, int i) {
    V s = 0;
    const int p1 = rowptr[i + 1];
    for (int p0 = rowptr[i]; p0 < p1; p0 += kRowBatch) {
        int c[kRowBatch];
        V a[kRowBatch], xv[kRowBatch];
#pragma unroll
        for (int k = 0; k < kRowBatch; ++k) {
            const int p = min(p0 + k, p1 - 1);
            c[k] = col[p];
            a[k] = val[p];
        }
#pragma unroll
        for (int k = 0; k < kRowBatch; ++k) xv[k] = xat(c[k]);
#pragma unroll
        for (int k = 0; k < kRowBatch; ++k) {
            const V t = s + a[k] * xv[k];
            s = (p0 + k < p1) ? t : s;
        }
    }
    return s;
}
template <typename V>
__device__ inline V row_ax(const int *__restrict__ rowptr, const int *__restrict__ col, const V *__restrict__ val,
                           const V *__restrict__ x, int i) {
    return row_ax_at(rowptr, col, val, [=](int j) { return x[j]; }, i);
}

// Row i of a banded grid's A x: entries among i-nx, i-1, i, i+1, i+nx (bit k
// of mask[i]), values C[k n + i]; the present ones summed from 0 in
// ascending column order, as row_ax does, so the two agree bit for bit.
template <typename V, class X>
__device__ inline V band_ax_at(int n, int nx, const unsigned char *__restrict__ mask, const V *__restrict__ C, X xat,
                               int i) {
    const int mk = mask[i];
    const int off[5] = {-nx, -1, 0, 1, nx};
    V a[5], xv[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) a[k] = C[(size_t)k * n + i];
#pragma unroll
    for (int k = 0; k < 5; ++k) xv[k] = xat(((mk >> k) & 1) ? i + off[k] : i);
    V s = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const V t = s + a[k] * xv[k];
        s = ((mk >> k) & 1) ? t : s;
    }
    return s;
}
template <typename V>
__device__ inline V band_ax(int n, int nx, const unsigned char *__restrict__ mask, const V *__restrict__ C,
                            const V *__restrict__ x, int i) {
    return band_ax_at(n, nx, mask, C, [=](int j) { return x[j]; }, i);
}

template <typename V>
__global__ void __launch_bounds__(256) k_spmv_band(int n, int nx, const unsigned char *__restrict__ mask,
                                                   const V *__restrict__ C, const V *__restrict__ x,
                                                   V *__restrict__ y, V sigma, const KScal *ks) {
    if (ks && ks->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = sigma * band_ax(n, nx, mask, C, x, i);
}

// One thread per row.
template <typename V>
__global__ void __launch_bounds__(256) k_spmv(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                                              const V *__restrict__ val, const V *__restrict__ x, V *__restrict__ y,
                                              V sigma, const KScal *ks) {
    if (ks && ks->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = sigma * row_ax(rowptr, col, val, x, i);
}

// Canonical partial sums of one 2048-row chunk from per-row terms staged in
// LDS (p[v][j] = term of row chunk_base + j, 0 for rows past n): lane t < 256
// adds rows t + 256 k in ascending k, then the same tree as chunk_reduce.
// All kWide threads of the block call it.
constexpr int kWide = 1024;
template <int NV>
__device__ inline void chunk_reduce_staged(double (&p)[NV][kChunk], double *__restrict__ part, int nch) {
    const int t = threadIdx.x;
    __syncthreads();
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    if (t < kBlock) {
#pragma unroll
        for (int k = 0; k < kPer; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] = acc[v] + p[v][t + k * kBlock];
    }
    __syncthreads();
    if (t < kBlock) {
#pragma unroll
        for (int v = 0; v < NV; ++v) p[v][t] = acc[v];
    }
    __syncthreads();
    if (t < 128) {
#pragma unroll
        for (int v = 0; v < NV; ++v) p[v][t] = p[v][t] + p[v][t + 128];
    }
    __syncthreads();
    if (t < 64) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double a = p[v][t] + p[v][t + 64];
            a = a + __shfl_down(a, 32);
            a = a + __shfl_down(a, 16);
            a = a + __shfl_down(a, 8);
            a = a + __shfl_down(a, 4);
            a = a + __shfl_down(a, 2);
            a = a + __shfl_down(a, 1);
            if (t == 0) part[v * nch + blockIdx.x] = a;
        }
    }
}

// y = sigma A x and the partial sums of w.y: one block of kWide threads per
// chunk, two rows per thread.
template <typename V>
__global__ void __launch_bounds__(kWide) k_spmv_dot(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                                                    const V *__restrict__ val, const V *__restrict__ x,
                                                    V *__restrict__ y, V sigma, const KScal *ks,
                                                    const V *__restrict__ w, double *part, int nch) {
    if (ks && ks->done) return;
    __shared__ double p[1][kChunk];
    const int c0 = blockIdx.x * kChunk;
#pragma unroll
    for (int h = 0; h < kChunk / kWide; ++h) {
        const int j = threadIdx.x + h * kWide, i = c0 + j;
        double pr = 0.0;
        if (i < n) {
            const V s = sigma * row_ax(rowptr, col, val, x, i);
            y[i] = s;
            pr = (double)w[i] * (double)s;
        }
        p[0][j] = pr;
    }
    chunk_reduce_staged<1>(p, part, nch);
}

// The banded k_spmv_dot.
template <typename V>
__global__ void __launch_bounds__(kWide) k_spmv_dot_band(int n, int nx, const unsigned char *__restrict__ mask,
                                                         const V *__restrict__ C, const V *__restrict__ x,
                                                         V *__restrict__ y, V sigma, const KScal *ks,
                                                         const V *__restrict__ w, double *part, int nch) {
    if (ks && ks->done) return;
    __shared__ double p[1][kChunk];
    const int c0 = blockIdx.x * kChunk;
#pragma unroll
    for (int h = 0; h < kChunk / kWide; ++h) {
        const int j = threadIdx.x + h * kWide, i = c0 + j;
        double pr = 0.0;
        if (i < n) {
            const V s = sigma * band_ax(n, nx, mask, C, x, i);
            y[i] = s;
            pr = (double)w[i] * (double)s;
        }
        p[0][j] = pr;
    }
    chunk_reduce_staged<1>(p, part, nch);
}

// A's values by band position (see band_ax) from the CSR values.
template <typename V>
__global__ void k_band_values(int n, const int *__restrict__ rowptr, const unsigned char *__restrict__ mask,
                              const double *__restrict__ val, V *__restrict__ C) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int mk = mask[i];
    int p = rowptr[i];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const bool in = (mk >> k) & 1;
        C[(size_t)k * n + i] = in ? (V)val[p] : V(0);
        p += in ? 1 : 0;
    }
}

}  // namespace fvh
