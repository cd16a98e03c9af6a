// Generic device helpers: elementwise maps, canonical-order reductions and
// single-thread scalar updates, with the early exit on a finished Krylov
// solve (ks->done).
#pragma once

#include "k_calc.h"

namespace fvh {

// ---------------------------------------------------------------------------
// Canonical reductions
// ---------------------------------------------------------------------------
template <int NV>
__device__ inline void chunk_reduce(double (&acc)[NV], double *__restrict__ part, int nch) {
    __shared__ double sh[NV][kBlock];
    const int t = threadIdx.x;
#pragma unroll
    for (int v = 0; v < NV; ++v) sh[v][t] = acc[v];
    __syncthreads();
    if (t < 128) {
#pragma unroll
        for (int v = 0; v < NV; ++v) sh[v][t] = sh[v][t] + sh[v][t + 128];
    }
    __syncthreads();
    if (t < 64) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double a = sh[v][t] + sh[v][t + 64];
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

// The first kBlock threads of a block reduce NV partial arrays of `cnt`
// entries (array v at part + v*stride + base); every thread of the block
// (kBlock or more) calls it and gets r.
template <int NV>
__device__ inline void final_reduce_at(const double *__restrict__ part, int stride, int base, int cnt,
                                       double (&r)[NV]) {
    __shared__ double sh[NV][kBlock];
    const int t = threadIdx.x;
    if (t < kBlock) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double a = 0.0;
            for (int j = t; j < cnt; j += kBlock) a = a + part[v * stride + base + j];
            sh[v][t] = a;
        }
    }
    __syncthreads();
    if (t < 128) {
#pragma unroll
        for (int v = 0; v < NV; ++v) sh[v][t] = sh[v][t] + sh[v][t + 128];
    }
    __syncthreads();
    if (t < 64) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double a = sh[v][t] + sh[v][t + 64];
            a = a + __shfl_down(a, 32);
            a = a + __shfl_down(a, 16);
            a = a + __shfl_down(a, 8);
            a = a + __shfl_down(a, 4);
            a = a + __shfl_down(a, 2);
            a = a + __shfl_down(a, 1);
            if (t == 0) sh[v][0] = a;
        }
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < NV; ++v) r[v] = sh[v][0];
}

template <int NV>
__device__ inline void final_reduce(const double *__restrict__ part, int nch, double (&r)[NV]) {
    final_reduce_at<NV>(part, nch, 0, nch, r);
}

// Elementwise work + partial sums over the canonical chunk mapping.
template <int NV, class Op>
__global__ void __launch_bounds__(kBlock) k_mapred(int n, const KScal *ks, double *part, int nch, Op op) {
    if (ks && ks->done) return;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    const int base = blockIdx.x * kChunk + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        int i = base + k * kBlock;
        if (i < n) op(i, acc);
    }
    chunk_reduce<NV>(acc, part, nch);
}

template <class Op>
__global__ void __launch_bounds__(kBlock) k_map(int n, const KScal *ks, Op op) {
    if (ks && ks->done) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) op(i);
}

template <int NV, class Fin>
__global__ void __launch_bounds__(kBlock) k_final(const double *part, int nch, KScal *ks, bool skip_if_done, Fin fin) {
    if (skip_if_done && ks->done) return;
    double r[NV];
    final_reduce<NV>(part, nch, r);
    if (threadIdx.x == 0) fin(ks, r);
}

template <class Fn>
__global__ void k_single(KScal *ks, bool skip_if_done, Fn fn) {
    if (skip_if_done && ks->done) return;
    fn(ks);
}

template <int NV, class Op>
inline int mapred(int64_t n, const KScal *ks, double *part, Op op) {
    const int nch = nchunks(n);
    hipLaunchKernelGGL((k_mapred<NV, Op>), dim3(nch > 0 ? nch : 1), dim3(kBlock), 0, stream(), (int)n, ks, part, nch, op);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
template <class Op>
inline int map(int64_t n, const KScal *ks, Op op) {
    hipLaunchKernelGGL((k_map<Op>), grid1d(n), dim3(kBlock), 0, stream(), (int)n, ks, op);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
template <int NV, class Fin>
inline int finalize(int64_t n, const double *part, KScal *ks, bool skip, Fin fin) {
    hipLaunchKernelGGL((k_final<NV, Fin>), dim3(1), dim3(kBlock), 0, stream(), part, nchunks(n), ks, skip, fin);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
template <class Fn>
inline int single(KScal *ks, bool skip, Fn fn) {
    hipLaunchKernelGGL((k_single<Fn>), dim3(1), dim3(1), 0, stream(), ks, skip, fn);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

}  // namespace fvh
