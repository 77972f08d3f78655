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
// in registers and reads the row's band from LDS (one address per wave).
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

constexpr int SEP_TX = 64, SEP_TY = 32, SEP_NT = 256;

// Every global load (tile + halo, both band tables, the line scales) is issued
// before the first barrier, so a workgroup waits for memory once.
template <int N, typename T, typename TQ, bool DOT, int TY = SEP_TY, int NT = SEP_NT>
__global__ __launch_bounds__(NT) void k_sep_apply(SepParams S, const T* __restrict__ in, TQ* __restrict__ out,
                                                   double* __restrict__ partials, const CgState* __restrict__ st) {
    constexpr int B = 2 * N + 1;
    constexpr int W = SEP_TX + 2 * N, H = TY + 2 * N;
    constexpr int WP = W + 1;
    constexpr int RPT = TY / (NT / 64);  // rows per thread
    constexpr int NG = (TY * B + NT - 1) / NT;
    __shared__ float s_u[H * WP];
    __shared__ float s_gy[TY * B];
    __shared__ float s_sy[TY], s_my[TY];

    const int tid = threadIdx.x;
    const int tilesX = (S.nU0 + SEP_TX - 1) / SEP_TX;
    const int I0 = (blockIdx.x % tilesX) * SEP_TX;
    const int J0 = (blockIdx.x / tilesX) * TY;

    // tile + halo; periodic lines wrap, other out-of-grid values are zero.
    // Wave wv loads rows wv, wv + NW, ... (row index wave-uniform); lane l
    // holds tile columns l and 64 + l (the latter for l < 2N).
    constexpr int NW = NT / 64;
    constexpr int RL = (H + NW - 1) / NW;  // rows loaded per wave
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    int Ic0 = I0 - N + lane, Ic1 = I0 - N + 64 + lane;
    bool c0 = true, c1 = lane < 2 * N;
    if (S.per0) {
        Ic0 = Ic0 < 0 ? Ic0 + S.nU0 : (Ic0 >= S.nU0 ? Ic0 - S.nU0 : Ic0);
        Ic1 = Ic1 >= S.nU0 ? Ic1 - S.nU0 : Ic1;
    }
    // loads are unconditional from clamped (in-range) addresses, the value
    // is then zeroed where the node is outside the grid: no divergent branches
    c0 = c0 && Ic0 >= 0 && Ic0 < S.nU0;
    c1 = c1 && Ic1 >= 0 && Ic1 < S.nU0;
    Ic0 = c0 ? Ic0 : 0;
    Ic1 = c1 ? Ic1 : 0;
    float u0[RL], u1[RL];
#pragma unroll
    for (int k = 0; k < RL; k++) {
        const int a = wv + NW * k;
        int J = J0 - N + a;
        if (S.per1) J = J < 0 ? J + S.nU1 : (J >= S.nU1 ? J - S.nU1 : J);
        const bool rok = a < H && J >= 0 && J < S.nU1;
        const T* row = in + (size_t)(rok ? J : 0) * S.nU0;
        const float v0 = to_f<T>(row[Ic0]), v1 = to_f<T>(row[Ic1]);
        u0[k] = (rok && c0) ? v0 : 0.0f;
        u1[k] = (rok && c1) ? v1 : 0.0f;
    }
    float gyv[NG];
#pragma unroll
    for (int k = 0; k < NG; k++) {
        const int i = tid + NT * k;
        const bool ok = i < TY * B && J0 + i / B < S.nU1;
        const float v = S.gy[ok ? (size_t)J0 * B + i : 0];
        gyv[k] = ok ? v : 0.0f;
    }
    const bool lok = tid < TY && J0 + tid < S.nU1;
    float syv = S.sy[lok ? J0 + tid : 0], myv = S.my[lok ? J0 + tid : 0];
    syv = lok ? syv : 0.0f;
    myv = lok ? myv : 0.0f;
    const int j = lane, g = wv;
    const int I = I0 + j;
    const bool colok = I < S.nU0;
    const int Ic = colok ? I : 0;
    float gxr[B];
#pragma unroll
    for (int s = 0; s < B; s++) gxr[s] = S.gx[(size_t)Ic * B + s];
    const float sxI = S.sx[Ic], mxI = S.mx[Ic];
    // a stopped solve skips the work (checked after the loads are in flight;
    // the flag is the same for every thread of the workgroup)
    if (DOT && cg_stopped(st)) return;

#pragma unroll
    for (int k = 0; k < RL; k++) {
        const int a = wv + NW * k;
        if (a < H) {
            s_u[a * WP + lane] = u0[k];
            if (lane < 2 * N) s_u[a * WP + 64 + lane] = u1[k];
        }
    }
#pragma unroll
    for (int k = 0; k < NG; k++) {
        const int i = tid + NT * k;
        if (i < TY * B) s_gy[i] = gyv[k];
    }
    if (tid < TY) {
        s_sy[tid] = syv;
        s_my[tid] = myv;
    }
    __syncthreads();

    // x sweep of rows g*RPT .. g*RPT+RPT-1 with column j's band
    float ax[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const float* row = &s_u[(g * RPT + k + N) * WP + j];
        float acc = 0.0f;
#pragma unroll
        for (int s = 0; s < B; s++) acc += gxr[s] * row[s];
        ax[k] = acc;
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
        const float* gy = &s_gy[r * B];  // same address in every lane: broadcast
        float acc = 0.0f;
#pragma unroll
        for (int s = 0; s < B; s++) acc += gy[s] * win[k + s];
        const float u = win[k + N];
        const float y = s_sy[r] * ax[k] + sxI * acc + (mxI * s_my[r]) * u;
        if (colok && J < S.nU1) {
            out[(size_t)J * S.nU0 + I] = from_f<TQ>(y);
            if (DOT) dot += (double)u * (double)y;
        }
    }
    if (DOT) {
        double v[1] = {dot};
        block_store_partials<1, NT>(v, partials);
    }
}

}  // namespace se
