// sep_kernels.h -- separable form of the operator, on the unknown grid, for
// the low-precision inner solves of iterative refinement.
//
// The unknowns of every numbering the reference builds are a row-major grid
// (nU1 rows x nU0 columns, BCs.cpp:91-172: Dirichlet lines removed, periodic
// duplicate lines folded onto line 0).  With the metric factored per element
// column / row (op_metric_tables), the assembled operator without its
// round-off-sized Jacobian cross terms (relative size ~1e-12, below FP32
// resolution) is a Kronecker sum of 1-D banded operators:
//   (A u)(I,J) = sy[J] sum_s gx[I][s] u(I+s-N, J) + sx[I] sum_s gy[J][s] u(I, J+s-N)
//                + mx[I] my[J] u(I,J)
// with 2N+1 band coefficients per line (periodic lines wrap, band entries
// beyond a non-periodic edge are zero).  One workgroup owns a 64 x 32 tile;
// the x sweep keeps a column's band in registers, the y sweep a column window
// in registers and the row's band in wave-uniform (scalar) loads.
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include "apply_kernels.h"

namespace se {

template <typename T> __device__ __forceinline__ float to_f(T v);
template <> __device__ __forceinline__ float to_f<float>(float v) { return v; }
template <> __device__ __forceinline__ float to_f<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }
template <typename T> __device__ __forceinline__ T from_f(float v);
template <> __device__ __forceinline__ float from_f<float>(float v) { return v; }
template <> __device__ __forceinline__ __hip_bfloat16 from_f<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

struct SepParams {
    int nU0, nU1, per0, per1;
    const float* gx;  // [nU0][2N+1]
    const float* gy;  // [nU1][2N+1]
    const float* sx;  // [nU0]  scale of the y sweep
    const float* sy;  // [nU1]  scale of the x sweep
    const float* mx;  // [nU0]  mass factor
    const float* my;  // [nU1]
};

constexpr int SEP_TX = 64, SEP_TY = 32;

template <int N, typename T, typename TQ, bool DOT>
__global__ __launch_bounds__(256) void k_sep_apply(SepParams S, const T* __restrict__ in, TQ* __restrict__ out,
                                                    double* __restrict__ partials, const CgState* __restrict__ st) {
    constexpr int B = 2 * N + 1;
    constexpr int W = SEP_TX + 2 * N, H = SEP_TY + 2 * N;
    constexpr int WP = W + 1;
    constexpr int RPT = SEP_TY / 4;  // rows per thread
    if (DOT && __hip_atomic_load(&st->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    __shared__ float s_u[H * WP];
    __shared__ float s_x[SEP_TY * SEP_TX];

    const int tid = threadIdx.x;
    const int tilesX = (S.nU0 + SEP_TX - 1) / SEP_TX;
    const int I0 = (blockIdx.x % tilesX) * SEP_TX;
    const int J0 = (blockIdx.x / tilesX) * SEP_TY;

    // tile + halo; periodic lines wrap, other out-of-grid values are zero
    for (int i = tid; i < H * W; i += 256) {
        const int a = i / W, b = i % W;
        int I = I0 - N + b, J = J0 - N + a;
        bool ok = true;
        if (I < 0 || I >= S.nU0) {
            if (S.per0) I = (I % S.nU0 + S.nU0) % S.nU0;
            else ok = false;
        }
        if (J < 0 || J >= S.nU1) {
            if (S.per1) J = (J % S.nU1 + S.nU1) % S.nU1;
            else ok = false;
        }
        s_u[a * WP + b] = ok ? to_f<T>(in[(size_t)J * S.nU0 + I]) : 0.0f;
    }
    const int j = tid & 63, g = tid >> 6;
    const int I = I0 + j;
    const bool colok = I < S.nU0;
    float gxr[B];
#pragma unroll
    for (int s = 0; s < B; s++) gxr[s] = colok ? S.gx[(size_t)I * B + s] : 0.0f;
    const float sxI = colok ? S.sx[I] : 0.0f, mxI = colok ? S.mx[I] : 0.0f;
    __syncthreads();

    // x sweep: row r of the tile, band of column j
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int r = g * RPT + k;
        const float* row = &s_u[(r + N) * WP + j];
        float acc = 0.0f;
#pragma unroll
        for (int s = 0; s < B; s++) acc += gxr[s] * row[s];
        s_x[r * SEP_TX + j] = acc;
    }
    // y sweep: column window of rows g*RPT - N .. g*RPT + RPT - 1 + N
    float win[RPT + 2 * N];
#pragma unroll
    for (int t = 0; t < RPT + 2 * N; t++) win[t] = s_u[(g * RPT + t) * WP + j + N];
    double dot = 0.0;
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int r = g * RPT + k;
        const int J = J0 + r;
        if (J >= S.nU1) break;
        const float* gy = &S.gy[(size_t)J * B];  // wave-uniform row
        float acc = 0.0f;
#pragma unroll
        for (int s = 0; s < B; s++) acc += gy[s] * win[k + s];
        const float u = win[k + N];
        const float y = S.sy[J] * s_x[r * SEP_TX + j] + sxI * acc + (mxI * S.my[J]) * u;
        if (colok) {
            out[(size_t)J * S.nU0 + I] = from_f<TQ>(y);
            if (DOT) dot += (double)u * (double)y;
        }
    }
    if (DOT) {
        double v[1] = {dot};
        block_store_partials<1, 256>(v, partials);
    }
}

}  // namespace se
