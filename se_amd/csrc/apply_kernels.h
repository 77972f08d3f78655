// apply_kernels.h -- the hot matrix-free operator apply (FP64), y = s*A x.
//
// Same tile / owner-computes structure as k_helm_tile (one workgroup owns the
// nodes of TX x TY elements and also evaluates the element column to its left
// and the row below, so every owned node sums all its element contributions
// without atomics), but organised for LDS bandwidth:
//   1. gather the node block through the numbering map into LDS;
//   2. forward contractions u_xi, u_eta by element lines: one thread holds a
//      line of N+1 values in registers and produces N+1 outputs, D is read
//      with wave-uniform (broadcast) LDS addresses;
//   3. per quadrature point: exact reference Jacobian from the per-column /
//      per-row tables, one division, symmetric 2x2 metric, fluxes in place;
//   4. owned nodes: back-contractions of the up to four containing elements
//      with the needed columns of D held in registers, plus the GLL-diagonal
//      mass term; the block's partial of p.Ap is stored for the CG update.
#pragma once
#include "helm_kernels.h"

namespace se {

// block-sum K values (blockDim BS) and store them to partials[k*gridDim.x + b]
template <int K, int BS>
__device__ __forceinline__ void block_store_partials(const double (&v)[K], double* __restrict__ partials) {
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
        partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
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

template <int N, int TX, int TY, bool DOT>
__global__ __launch_bounds__(256) void k_apply(OpParams P, const double* __restrict__ in, double* __restrict__ out,
                                                double* __restrict__ partials, const CgState* __restrict__ st) {
    constexpr int N1 = N + 1;
    constexpr int NP = N1 * N1;
    constexpr int BW = ((TX + 1) * N + 2) | 1;  // odd row pitch (in doubles): conflict-free column access
    constexpr int BH = (TY + 1) * N + 2;
    constexpr int NE = (TX + 1) * (TY + 1);
    constexpr int NC = N * N;                   // (rx, ry) combinations of an owned element
    constexpr int NG = 256 / NC;                // thread groups in the owned-node phase

    if (DOT && __hip_atomic_load(&st->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;

    __shared__ double s_u[BH * BW];
    __shared__ double s_a[NE * NP];  // u_xi, then f_xi     [e][qy][qx]
    __shared__ double s_b[NE * NP];  // u_eta, then f_eta   [e][qy][qx]
    __shared__ double s_D[NP];
    __shared__ double s_w[N1];
    __shared__ double s_J00[(TX + 1) * N1], s_J10[(TX + 1) * NP];
    __shared__ double s_J11[(TY + 1) * N1], s_J01[(TY + 1) * NP];
    __shared__ int s_exw[TX + 1], s_eyw[TY + 1];

    const int tid = threadIdx.x;
    const int tilesX = (P.nEls0 + TX - 1) / TX;
    const int ex0 = (blockIdx.x % tilesX) * TX;
    const int ey0 = (blockIdx.x / tilesX) * TY;

    if (tid <= TX) {
        int ex = ex0 - 1 + tid;
        if (ex < 0) ex = P.per0 ? ex + P.nEls0 : -1;
        if (ex >= P.nEls0) ex = -1;
        s_exw[tid] = ex;
    }
    if (tid >= 64 && tid <= 64 + TY) {
        int ey = ey0 - 1 + (tid - 64);
        if (ey < 0) ey = P.per1 ? ey + P.nEls1 : -1;
        if (ey >= P.nEls1) ey = -1;
        s_eyw[tid - 64] = ey;
    }
    for (int i = tid; i < NP; i += 256) s_D[i] = P.D[i];
    if (tid < N1) s_w[tid] = P.w[tid];
    __syncthreads();
    for (int i = tid; i < (TX + 1) * N1; i += 256) {
        int c = i / N1, e = s_exw[c];
        s_J00[i] = e >= 0 ? P.J00[(size_t)e * N1 + i % N1] : 1.0;
    }
    for (int i = tid; i < (TX + 1) * NP; i += 256) {
        int c = i / NP, e = s_exw[c];
        s_J10[i] = e >= 0 ? P.J10[(size_t)e * NP + i % NP] : 0.0;
    }
    for (int i = tid; i < (TY + 1) * N1; i += 256) {
        int r = i / N1, e = s_eyw[r];
        s_J11[i] = e >= 0 ? P.J11[(size_t)e * N1 + i % N1] : 1.0;
    }
    for (int i = tid; i < (TY + 1) * NP; i += 256) {
        int r = i / NP, e = s_eyw[r];
        s_J01[i] = e >= 0 ? P.J01[(size_t)e * NP + i % NP] : 0.0;
    }

    // ---- 1. gather ---------------------------------------------------------
    constexpr int BWU = (TX + 1) * N + 2;  // used columns
    for (int i = tid; i < BH * BWU; i += 256) {
        int a = i / BWU, b = i % BWU;
        int c, ix, r, iy;
        if (b <= N) { c = 0; ix = b; }
        else { int bb = b - N1; c = 1 + bb / N; ix = bb % N; if (ix == 0 && c >= 2) { c -= 1; ix = N; } }
        if (a <= N) { r = 0; iy = a; }
        else { int aa = a - N1; r = 1 + aa / N; iy = aa % N; if (iy == 0 && r >= 2) { r -= 1; iy = N; } }
        int ex = s_exw[c], ey = s_eyw[r];
        double val = 0.0;
        if (ex >= 0 && ey >= 0) {
            int u = P.f2v[(ey * N + iy) * P.nv0 + ex * N + ix];
            val = u >= 0 ? in[u] : 0.0;
        }
        s_u[a * BW + b] = val;
    }
    __syncthreads();

    // ---- 2. forward contractions by element lines -----------------------------
    for (int l = tid; l < 2 * NE * N1; l += 256) {
        const int e = l / (2 * N1), rem = l % (2 * N1), dir = rem / N1, idx = rem % N1;
        const int c = e % (TX + 1), r = e / (TX + 1);
        if (s_exw[c] < 0 || s_eyw[r] < 0) continue;
        double u[N1];
        if (dir == 0) {
            const double* row = &s_u[slot<N>(r, idx) * BW];
#pragma unroll
            for (int k = 0; k < N1; k++) u[k] = row[slot<N>(c, k)];
        } else {
            const int col = slot<N>(c, idx);
#pragma unroll
            for (int k = 0; k < N1; k++) u[k] = s_u[slot<N>(r, k) * BW + col];
        }
#pragma unroll
        for (int q = 0; q < N1; q++) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < N1; k++) s += s_D[q * N1 + k] * u[k];
            if (dir == 0) s_a[e * NP + idx * N1 + q] = s;  // u_xi(qx = q, qy = idx)
            else s_b[e * NP + q * N1 + idx] = s;           // u_eta(qx = idx, qy = q)
        }
    }
    __syncthreads();

    // ---- 3. metric and fluxes per quadrature point ----------------------------
    for (int i = tid; i < NE * NP; i += 256) {
        const int e = i / NP, q = i % NP;
        const int c = e % (TX + 1), r = e / (TX + 1);
        if (s_exw[c] < 0 || s_eyw[r] < 0) continue;
        const int qx = q % N1, qy = q / N1;
        const double j00 = s_J00[c * N1 + qx], j01 = s_J01[r * NP + qy * N1 + qx];
        const double j10 = s_J10[c * NP + qx * N1 + qy], j11 = s_J11[r * N1 + qy];
        double det, W;
        {
#pragma clang fp contract(off)
            det = j00 * j11 - j01 * j10;
            const double wq = s_w[qx] * s_w[qy];
            W = (P.detj ? det * wq : 1.0 * wq) * P.c_lap;
        }
        const double inv = 1.0 / det;
        const double i00 = j11 * inv, i11 = j00 * inv, i01 = -j01 * inv, i10 = -j10 * inv;
        const double g00 = W * (i00 * i00 + i10 * i10);
        const double g01 = W * (i00 * i01 + i10 * i11);
        const double g11 = W * (i01 * i01 + i11 * i11);
        const double ux = s_a[i], uy = s_b[i];
        s_a[i] = g00 * ux + g01 * uy;
        s_b[i] = g01 * ux + g11 * uy;
    }
    __syncthreads();

    // ---- 4. owned nodes -------------------------------------------------------
    const int txo = min(TX, P.nEls0 - ex0), tyo = min(TY, P.nEls1 - ey0);
    double dot = 0.0;

    // contribution of tile element (c, r) at its local node (lx, ly)
    auto contrib = [&](int c, int r, int lx, int ly, const double* Dx, const double* Dy) -> double {
        const int e = r * (TX + 1) + c;
        const double* fx = &s_a[e * NP + ly * N1];
        const double* fy = &s_b[e * NP + lx];
        double k = 0.0;
#pragma unroll
        for (int q = 0; q < N1; q++) k += Dx[q] * fx[q];
#pragma unroll
        for (int q = 0; q < N1; q++) k += Dy[q] * fy[q * N1];
        double mass;
        {
#pragma clang fp contract(off)
            const double det = s_J00[c * N1 + lx] * s_J11[r * N1 + ly] -
                               s_J01[r * NP + ly * N1 + lx] * s_J10[c * NP + lx * N1 + ly];
            mass = (det * (s_w[lx] * s_w[ly])) * P.c_mass;
        }
        return k + mass * s_u[slot<N>(r, ly) * BW + slot<N>(c, lx)];
    };

    if (tid < NG * NC) {
        const int cmb = tid % NC, g = tid / NC;
        const int rx = cmb % N, ry = cmb / N;
        double Dcx[N1], Dcy[N1], DcN[N1];
#pragma unroll
        for (int q = 0; q < N1; q++) {
            Dcx[q] = s_D[q * N1 + rx];
            Dcy[q] = s_D[q * N1 + ry];
            DcN[q] = s_D[q * N1 + N];
        }
        for (int oe = g; oe < txo * tyo; oe += NG) {
            const int ce = 1 + oe % txo, re = 1 + oe / txo;
            const int ox = (ce - 1) * N + rx, oy = (re - 1) * N + ry;
            const int u = P.f2v[(ey0 * N + oy) * P.nv0 + ex0 * N + ox];
            if (u < 0) continue;
            double sum = 0.0;
            if (ry == 0 && s_eyw[re - 1] >= 0) {
                if (rx == 0 && s_exw[ce - 1] >= 0) sum += contrib(ce - 1, re - 1, N, N, DcN, DcN);
                sum += contrib(ce, re - 1, rx, N, Dcx, DcN);
            }
            if (rx == 0 && s_exw[ce - 1] >= 0) sum += contrib(ce - 1, re, N, ry, DcN, Dcy);
            sum += contrib(ce, re, rx, ry, Dcx, Dcy);
            const double yv = P.scale * sum;
            out[u] = yv;
            if (DOT) dot += s_u[slot<N>(re, ry) * BW + slot<N>(ce, rx)] * yv;
        }
    }
    // closing lines of a non-periodic domain edge (x = max or y = max)
    {
        const bool cx = !P.per0 && ex0 + txo == P.nEls0, cy = !P.per1 && ey0 + tyo == P.nEls1;
        const int nx = cx ? tyo * N + (cy ? 1 : 0) : 0;   // nodes on the closing column
        const int ny = cy ? txo * N : 0;                  // nodes on the closing row (corner counted above)
        for (int i = tid; i < nx + ny; i += 256) {
            int ox, oy;
            if (i < nx) { ox = txo * N; oy = i; }
            else { ox = i - nx; oy = tyo * N; }
            const int u = P.f2v[(ey0 * N + oy) * P.nv0 + ex0 * N + ox];
            if (u < 0) continue;
            int cs[2], xs[2], nc = 0, rs[2], ys[2], nr = 0;
            if (ox % N == 0) { cs[nc] = ox / N; xs[nc] = N; nc++; }
            if (ox / N < txo) { cs[nc] = 1 + ox / N; xs[nc] = ox % N; nc++; }
            if (oy % N == 0) { rs[nr] = oy / N; ys[nr] = N; nr++; }
            if (oy / N < tyo) { rs[nr] = 1 + oy / N; ys[nr] = oy % N; nr++; }
            double sum = 0.0, own = 0.0;
            for (int jr = 0; jr < nr; jr++) {
                if (s_eyw[rs[jr]] < 0) continue;
                for (int jc = 0; jc < nc; jc++) {
                    if (s_exw[cs[jc]] < 0) continue;
                    double Dx[N1], Dy[N1];
#pragma unroll
                    for (int q = 0; q < N1; q++) {
                        Dx[q] = s_D[q * N1 + xs[jc]];
                        Dy[q] = s_D[q * N1 + ys[jr]];
                    }
                    sum += contrib(cs[jc], rs[jr], xs[jc], ys[jr], Dx, Dy);
                    own = s_u[slot<N>(rs[jr], ys[jr]) * BW + slot<N>(cs[jc], xs[jc])];
                }
            }
            const double yv = P.scale * sum;
            out[u] = yv;
            if (DOT) dot += own * yv;
        }
    }

    if (DOT) {
        double v[1] = {dot};
        block_store_partials<1, 256>(v, partials);
    }
}

}  // namespace se
