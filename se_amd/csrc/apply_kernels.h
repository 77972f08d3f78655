// apply_kernels.h -- the hot matrix-free operator apply (FP64), y = s*A x.
//
// Same tile / owner-computes structure as k_helm_tile (one workgroup owns the
// nodes of TX x TY elements and also evaluates the element column to its left
// and the row below, so every owned node sums all its element contributions
// without atomics), but organised for LDS bandwidth:
//   1. gather the node block through the numbering map into LDS;
//   2. forward contractions u_xi, u_eta by element lines: one thread holds a
//      line of N+1 values in registers and produces N+1 outputs; D comes in
//      the kernel arguments (DMat), so with compile-time indices every D entry
//      is a scalar load and an SGPR operand of the FMA, not an LDS read (as a
//      wave-uniform LDS read it cost the full LDS return bandwidth: 2 (N+1)^3
//      reads per element, most of the kernel's LDS traffic);
//   3. per quadrature point: symmetric 2x2 metric of the reference's exact
//      per-point Jacobian, assembled from per-element-column and per-element-row
//      factor tables (see op_metric_tables), fluxes in place;
//   4. owned nodes: back-contractions of the up to four containing elements
//      with the needed columns of D held in registers, plus the GLL-diagonal
//      mass term; the block's partial of p.Ap is stored for the CG update.
#pragma once
#include <vector>

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

// The GLL derivative matrix D[q][k] = l_k'(x_q) of order N, passed by value.
template <int N>
struct DMat {
    double d[(N + 1) * (N + 1)];
};

// element tile of the batched apply (work items are (member, tile))
constexpr int kBatchApplyTile = 4;
inline int batch_apply_tiles(int nEls0, int nEls1) {
    return ((nEls0 + kBatchApplyTile - 1) / kBatchApplyTile) * ((nEls1 + kBatchApplyTile - 1) / kBatchApplyTile);
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

// One tile of TX x TY elements (tile index `tile`, row-major over the element
// grid).  With DOT, the workgroup's partial of in.out goes to
// partials[tile] (np = number of tiles).
template <int N, int TX, int TY, bool DOT>
__device__ __forceinline__ void apply_tile(const OpParams& P, const DMat<N>& dm, int tile,
                                           const double* __restrict__ in, double* __restrict__ out,
                                           double* __restrict__ partials, int np) {
    constexpr int N1 = N + 1;
    constexpr int NP = N1 * N1;
    constexpr int BW = ((TX + 1) * N + 2) | 1;  // odd row pitch (in doubles): conflict-free column access
    constexpr int BH = (TY + 1) * N + 2;
    constexpr int BWU = (TX + 1) * N + 2;       // used columns
    constexpr int NE = (TX + 1) * (TY + 1);
    constexpr int NC = N * N;                   // (rx, ry) combinations of an owned element
    constexpr int NG = 256 / NC;                // thread groups in the owned-node phase
    constexpr int NOWN = (TX * TY + NG - 1) / NG;
    constexpr int CS = 4 * N1 + NP;             // metric table stride per element column / row
    constexpr int NG1 = (BH * BWU + 255) / 256;

    __shared__ double s_u[BH * BW];
    __shared__ double s_a[NE * NP];  // u_xi, then f_xi     [e][qy][qx]
    __shared__ double s_b[NE * NP];  // u_eta, then f_eta   [e][qy][qx]
    __shared__ double s_D[NP];
    __shared__ double s_col[(TX + 1) * CS], s_row[(TY + 1) * CS];
    __shared__ int s_exw[TX + 1], s_eyw[TY + 1];

    const int tid = threadIdx.x;
    const int tilesX = (P.nEls0 + TX - 1) / TX;
    const int ex0 = (tile % tilesX) * TX;
    const int ey0 = (tile / tilesX) * TY;
    const int txo = min(TX, P.nEls0 - ex0), tyo = min(TY, P.nEls1 - ey0);
    const SE_GLOBAL int32_t* f2v = gview(P.f2v);
    const SE_GLOBAL double* gin = gview(in);
    SE_GLOBAL double* gout = gview(out);
    auto exw = [&](int c) {  // wrapped element index of tile column c (-1: none)
        int ex = ex0 - 1 + c;
        if (ex < 0) ex = P.per0 ? ex + P.nEls0 : -1;
        return ex >= P.nEls0 ? -1 : ex;
    };
    auto eyw = [&](int r) {
        int ey = ey0 - 1 + r;
        if (ey < 0) ey = P.per1 ? ey + P.nEls1 : -1;
        return ey >= P.nEls1 ? -1 : ey;
    };
    const int cmb = tid % NC, grp4 = tid / NC;
    const int rx = cmb % N, ry = cmb / N;

    // ---- 1. loads: map entries first, then everything independent of them,
    // then the values they index; a single barrier ---------------------------
    int um[NG1];
#pragma unroll
    for (int g = 0; g < NG1; g++) {
        const int i = tid + g * 256;
        um[g] = -1;
        if (i < BH * BWU) {
            const int a = i / BWU, b = i % BWU;
            int c, ix, r, iy;
            if (b <= N) { c = 0; ix = b; }
            else { int bb = b - N1; c = 1 + bb / N; ix = bb % N; if (ix == 0 && c >= 2) { c -= 1; ix = N; } }
            if (a <= N) { r = 0; iy = a; }
            else { int aa = a - N1; r = 1 + aa / N; iy = aa % N; if (iy == 0 && r >= 2) { r -= 1; iy = N; } }
            const int ex = exw(c), ey = eyw(r);
            if (ex >= 0 && ey >= 0) um[g] = f2v[(ey * N + iy) * P.nv0 + ex * N + ix];
        }
    }
    int ou[NOWN];  // unknowns of this thread's owned nodes (phase 4)
#pragma unroll
    for (int j = 0; j < NOWN; j++) {
        const int oe = grp4 + j * NG;
        ou[j] = -1;
        if (grp4 < NG && oe < txo * tyo) {
            const int ox = (oe % txo) * N + rx, oy = (oe / txo) * N + ry;
            ou[j] = f2v[(ey0 * N + oy) * P.nv0 + ex0 * N + ox];
        }
    }
    if (tid <= TX) s_exw[tid] = exw(tid);
    if (tid >= 64 && tid <= 64 + TY) s_eyw[tid - 64] = eyw(tid - 64);
    for (int i = tid; i < NP; i += 256) s_D[i] = dm.d[i];
    for (int i = tid; i < (TX + 1) * CS; i += 256) {
        const int e = exw(i / CS);
        s_col[i] = e >= 0 ? gview(P.colTab)[(size_t)e * CS + i % CS] : 0.0;
    }
    for (int i = tid; i < (TY + 1) * CS; i += 256) {
        const int e = eyw(i / CS);
        s_row[i] = e >= 0 ? gview(P.rowTab)[(size_t)e * CS + i % CS] : 0.0;
    }
    double uv[NG1];
#pragma unroll
    for (int g = 0; g < NG1; g++) uv[g] = um[g] >= 0 ? gin[um[g]] : 0.0;
#pragma unroll
    for (int g = 0; g < NG1; g++) {
        const int i = tid + g * 256;
        if (i < BH * BWU) s_u[(i / BWU) * BW + i % BWU] = uv[g];
    }
    __syncthreads();

    // ---- 2. forward contractions by element lines (waves 0-1: x lines,
    // waves 2-3: y lines; a thread keeps its line index across elements) -------
    {
        constexpr int LG = 128 / N1;  // line groups per direction
        const int dir = tid >> 7, tt = tid & 127;
        const int idx = tt % N1, lg = tt / N1;
        if (lg < LG) {
            for (int e = lg; e < NE; e += LG) {
                const int c = e % (TX + 1), r = e / (TX + 1);
                if (s_exw[c] < 0 || s_eyw[r] < 0) continue;
                double u[N1];
                if (dir == 0) {
                    const double* row = &s_u[slot<N>(r, idx) * BW];
#pragma unroll
                    for (int k = 0; k < N1; k++) u[k] = row[slot<N>(c, k)];
#pragma unroll
                    for (int q = 0; q < N1; q++) {
                        double s = 0.0;
#pragma unroll
                        for (int k = 0; k < N1; k++) s += dm.d[q * N1 + k] * u[k];
                        s_a[e * NP + idx * N1 + q] = s;  // u_xi(qx = q, qy = idx)
                    }
                } else {
                    const int col = slot<N>(c, idx);
#pragma unroll
                    for (int k = 0; k < N1; k++) u[k] = s_u[slot<N>(r, k) * BW + col];
#pragma unroll
                    for (int q = 0; q < N1; q++) {
                        double s = 0.0;
#pragma unroll
                        for (int k = 0; k < N1; k++) s += dm.d[q * N1 + k] * u[k];
                        s_b[e * NP + q * N1 + idx] = s;  // u_eta(qx = idx, qy = q)
                    }
                }
            }
        }
    }
    __syncthreads();

    // ---- 3. metric and fluxes per quadrature point (fixed point per thread) -----
    {
        constexpr int QG = 256 / NP;
        const int q = tid % NP, qg = tid / NP;
        const int qx = q % N1, qy = q / N1;
        const double c = P.c_lap;
        const double wq = gview(P.w)[qx] * gview(P.w)[qy];
        if (qg < QG) {
            for (int e = qg; e < NE; e += QG) {
                const int ce = e % (TX + 1), re = e / (TX + 1);
                if (s_exw[ce] < 0 || s_eyw[re] < 0) continue;
                const double* cx = &s_col[ce * CS];
                const double* ry_ = &s_row[re * CS];
                const double g00 = c * cx[qx] * ry_[qy];
                const double g11 = c * cx[N1 + qx] * ry_[N1 + qy];
                const double g01 = -c * wq * (cx[2 * N1 + qx] * ry_[4 * N1 + qy * N1 + qx] +
                                              cx[4 * N1 + qx * N1 + qy] * ry_[2 * N1 + qy]);
                const int i = e * NP + q;
                const double ux = s_a[i], uy = s_b[i];
                s_a[i] = g00 * ux + g01 * uy;
                s_b[i] = g01 * ux + g11 * uy;
            }
        }
    }
    __syncthreads();

    // ---- 4. owned nodes -------------------------------------------------------
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
        const double mass = P.c_mass * s_col[c * CS + 3 * N1 + lx] * s_row[r * CS + 3 * N1 + ly];
        return k + mass * s_u[slot<N>(r, ly) * BW + slot<N>(c, lx)];
    };

    if (grp4 < NG) {
        double Dcx[N1], Dcy[N1], DcN[N1];
#pragma unroll
        for (int q = 0; q < N1; q++) {
            Dcx[q] = s_D[q * N1 + rx];
            Dcy[q] = s_D[q * N1 + ry];
            DcN[q] = s_D[q * N1 + N];
        }
#pragma unroll
        for (int j = 0; j < NOWN; j++) {
            const int oe = grp4 + j * NG;
            if (oe >= txo * tyo) break;
            const int u = ou[j];
            if (u < 0) continue;
            const int ce = 1 + oe % txo, re = 1 + oe / txo;
            double sum = 0.0;
            if (ry == 0 && s_eyw[re - 1] >= 0) {
                if (rx == 0 && s_exw[ce - 1] >= 0) sum += contrib(ce - 1, re - 1, N, N, DcN, DcN);
                sum += contrib(ce, re - 1, rx, N, Dcx, DcN);
            }
            if (rx == 0 && s_exw[ce - 1] >= 0) sum += contrib(ce - 1, re, N, ry, DcN, Dcy);
            sum += contrib(ce, re, rx, ry, Dcx, Dcy);
            const double yv = P.scale * sum;
            gout[u] = yv;
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
            const int u = f2v[(ey0 * N + oy) * P.nv0 + ex0 * N + ox];
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
            gout[u] = yv;
            if (DOT) dot += own * yv;
        }
    }

    if (DOT) {
        double v[1] = {dot};
        block_store_partials_at<1, 256>(v, partials, tile, np);
    }
}

template <int N, int TX, int TY, bool DOT>
__global__ __launch_bounds__(256) void k_apply(OpParams P, DMat<N> dm, const double* __restrict__ in,
                                                double* __restrict__ out, double* __restrict__ partials,
                                                const CgState* __restrict__ st) {
    if (DOT && cg_stopped(st)) return;
    apply_tile<N, TX, TY, DOT>(P, dm, blockIdx.x, in, out, partials, gridDim.x);
}

// Batched apply: work items (member, tile) over members of one order (they
// share D); the grid loops over the items.  Each member brings its own
// operator, input and output.
struct ApplyMember {
    OpParams P;
    const double* in;
    double* out;
};

template <int N, int TX, int TY>
__global__ __launch_bounds__(256) void k_apply_batch(DMat<N> dm, const ApplyMember* __restrict__ members,
                                                      const int2* __restrict__ work, int nwork) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        const ApplyMember& a = members[it.x];
        __syncthreads();  // the previous item's LDS reads are done
        apply_tile<N, TX, TY, false>(a.P, dm, it.y, a.in, a.out, nullptr, 0);
    }
}

// D of a mesh's order as a kernel argument
template <int N>
inline DMat<N> dmat_of(const std::vector<double>& D) {
    DMat<N> m;
    for (int i = 0; i < (N + 1) * (N + 1); i++) m.d[i] = D[i];
    return m;
}

}  // namespace se
