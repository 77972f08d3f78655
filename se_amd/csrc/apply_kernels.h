// apply_kernels.h -- helpers of the operator-apply and CG kernels: global
// memory views, fixed-order block partial sums and their consumer-side
// reduction, the batched apply's member descriptor.  The apply itself is
// apply_lines.h.
#pragma once
#include "helm_kernels.h"

namespace se {

// Global-memory view of a pointer: accesses through it are global_*
// instructions.  Through a generic pointer (one read from a struct in memory)
// they would be flat_*, which share the LDS counter, so waits for LDS reads
// would also wait for them.
#define SE_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const SE_GLOBAL T* gview(const T* p) {
    return (const SE_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ SE_GLOBAL T* gview(T* p) {
    return (SE_GLOBAL T*)p;
}

// block-sum K values (blockDim BS) and store them to partials[k*np + idx]
template <int K, int BS>
__device__ __forceinline__ void block_store_partials_at(const double (&v)[K], double* __restrict__ partials, int idx,
                                                        int np) {
    constexpr int NW = BS / 64;
    __shared__ double s_red[K][NW];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0) s_red[k][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; w++) s += s_red[threadIdx.x][w];
        partials[(size_t)threadIdx.x * np + idx] = s;
    }
}
// one partial set per workgroup of the grid: partials[k*gridDim.x + blockIdx.x]
template <int K, int BS>
__device__ __forceinline__ void block_store_partials(const double (&v)[K], double* __restrict__ partials) {
    block_store_partials_at<K, BS>(v, partials, blockIdx.x, gridDim.x);
}

// Every block reduces the np partials of the previous kernel in the same fixed
// order; all threads receive the totals.  No atomics and no cross-block
// communication inside a kernel: the kernel boundary orders the data.
template <int K, int BS>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ partials, int np, double (&tot)[K]) {
    constexpr int NW = BS / 64;
    __shared__ double s_red2[K][NW];
    __shared__ double s_tot[K];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < np; b += BS) s += partials[(size_t)k * np + b];
        s = wave_sum(s);
        if (lane == 0) s_red2[k][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; w++) s += s_red2[threadIdx.x][w];
        s_tot[threadIdx.x] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] = s_tot[k];
}

// Batched apply: each member brings its own operator, input and output.
struct ApplyMember {
    OpParams P;
    const double* in;
    double* out;
    double* part = nullptr;      // AP_DOT / AP_TRUE partials, np of them per sum
    const double* b = nullptr;   // AP_TRUE: the right-hand side
    const CgState* gate = nullptr;  // gated launches: only while this CG awaits its true residual
    int np = 0;
};

}  // namespace se
