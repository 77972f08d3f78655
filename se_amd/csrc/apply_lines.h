// apply_lines.h -- the hot matrix-free FP64 operator apply, y = s*A x, on the
// unknown grid, organised by element lines.
//
// One workgroup owns a tile of TX x TY elements (shape per order, so that the
// tile and its halo fill one pass of 256 line threads) and produces the
// assembled result on the tile's nodes: the half-open node block of its
// elements plus, on a non-periodic domain edge, the closing line.  It also
// evaluates the element column to its left and the row below, so every owned
// node sums all its element contributions in a fixed order without atomics.
//
// Per element (N+1 = N1 nodes a side) the reference's element operator
// (Laplacian + mass, src/base/Operator.cpp:63-84,157-177, with the exact
// per-point Jacobian of Mesh::ShapeFuncDerivs, src/base/Mesh.cpp:217-250) is
//   Y = Dx^T (g00 o Dx U + g01 o Dy U) + Dy^T (g01 o Dx U + g11 o Dy U) + M o U
// with the metric factored per element column / row and the operator's
// coefficients folded in (op_metric_tables):
//   g00 = A0[qx] Y0[qy],  g11 = A1[qx] Y1[qy],  M = AM[qx] MY[qy],
//   g01 = A2[qx] QYw[qy][qx] + QXw[qx][qy] Y2w[qy].
// Thread t of an element holds row t and column t of U in registers:
//   1. row t: u_xi(q, t) = sum_k D[q][k] U[t][k];  column t: u_eta(t, q)
//      likewise.  D is read through constant-address-space pointers with
//      compile-time indices: each entry is a scalar load next to its use and
//      an SGPR operand of the FMA.  (As a by-value kernel argument all of D is
//      loaded at kernel entry and overflows the scalar registers into VGPR
//      lanes, one v_readlane per half at every use; the back contraction
//      reads its own copy, D^T, for the same reason.)
//   2. g01 is round-off-sized (the off-diagonal Jacobian entries come from
//      cancellation in sum_i D[q][i] x_i; about 1e-11..1e-10 of the operator
//      at these meshes), so the values that meet it across lines (u_eta on
//      row t, u_xi on column t) are exchanged through LDS in FP32 and the QX,
//      QY tables are FP32: both change y by ~1e-17 relative.
//   3. row t: f_xi(q, t), then Y_x(l, t) = sum_q D[q][l] f_xi(q, t) plus the
//      mass term; column t: f_eta(t, q), then Y_y(t, l).  Both sums of an
//      element node meet in one FP64 LDS array (Y_y first, then + Y_x).
//   4. node phase: each owned node adds the contributions of the up to four
//      elements containing it (bottom-left, bottom, left, own), scales, and
//      stores y (AP_PLAIN); with AP_DOT also the workgroup's partial of x.y
//      to partials[tile]; with AP_TRUE it stores t = b - y instead and the
//      partials of sum t and sum t^2 (a CG's true residual).  A batch member
//      without partials (or, for AP_TRUE, without b) gets plain y.
// LDS stays under 40 KB at every order (four workgroups per CU).
#pragma once
#include "apply_kernels.h"

namespace se {

enum { AP_PLAIN = 0, AP_DOT = 1, AP_TRUE = 2 };  // epilogues of apply_lines_tile

// owned elements per tile: (TX+1)(TY+1)(N+1) <= 256 line threads
template <int N> struct LineShape;
template <> struct LineShape<2> { static constexpr int TX = 8, TY = 8; };
template <> struct LineShape<3> { static constexpr int TX = 7, TY = 7; };
template <> struct LineShape<4> { static constexpr int TX = 6, TY = 6; };
template <> struct LineShape<5> { static constexpr int TX = 5, TY = 6; };
template <> struct LineShape<6> { static constexpr int TX = 5, TY = 5; };
template <> struct LineShape<7> { static constexpr int TX = 4, TY = 5; };
template <> struct LineShape<8> { static constexpr int TX = 4, TY = 4; };

// tiles of the line apply for order N (host and device)
__host__ __device__ inline int line_tiles(int N, int nEls0, int nEls1) {
    int tx = 4, ty = 4;
    switch (N) {
        case 2: tx = 8, ty = 8; break;
        case 3: tx = 7, ty = 7; break;
        case 4: tx = 6, ty = 6; break;
        case 5: tx = 5, ty = 6; break;
        case 6: tx = 5, ty = 5; break;
        case 7: tx = 4, ty = 5; break;
        default: break;
    }
    return ((nEls0 + tx - 1) / tx) * ((nEls1 + ty - 1) / ty);
}

// unknown-grid column of field column X (rows alike), -1 where there is no
// unknown (a Dirichlet line, or outside the domain); a periodic direction
// folds X = M (the duplicate line) and X < 0 (the wrapped left neighbour)
// read-only, kernel-invariant view: uniform loads through it are scalar loads
typedef const __attribute__((address_space(4))) double cdouble;

__device__ __forceinline__ int unk_line(int X, int M, int per, int off, int nU) {
    if (per) {
        if (X < 0) X += M;
        if (X == M) X = 0;
        return (X >= 0 && X < M) ? X : -1;
    }
    if (X < 0 || X > M) return -1;
    const int I = X - off;
    return (I >= 0 && I < nU) ? I : -1;
}

template <int N, int MODE>
__device__ __forceinline__ void apply_lines_tile(const OpParams& P, int tile,
                                                 const double* __restrict__ in, double* __restrict__ out,
                                                 double* __restrict__ partials, int np,
                                                 const double* __restrict__ b = nullptr) {
    constexpr int TX = LineShape<N>::TX, TY = LineShape<N>::TY;
    constexpr int N1 = N + 1, NP = N1 * N1;
    constexpr int BWU = (TX + 1) * N + 1, BH = (TY + 1) * N + 1;
    constexpr int BW = BWU | 1;                 // odd row pitch (doubles)
    constexpr int NE = (TX + 1) * (TY + 1);
    static_assert(NE * N1 <= 256, "one pass of line threads");
    constexpr int F = 4 * N1;                   // line factors A0 A1 A2 AM (Y0 Y1 Y2w MY)
    constexpr int QS = 2 * NP;                  // FP32 cross tables: Q, Q transposed

    __shared__ double s_u[BH * BW];
    __shared__ double s_o[NE * NP];              // FP32 exchange, then the element outputs
    __shared__ double s_cf[(TX + 1) * F], s_rf[(TY + 1) * F];
    __shared__ float s_cq[(TX + 1) * QS], s_rq[(TY + 1) * QS];
    __shared__ int s_I[BWU], s_J[BH];            // unknown column / row of each block column / row (-1: none)
    float* s_xf = (float*)s_o;                   // u_xi  [e][qx][qy]
    float* s_ef = s_xf + NE * NP;                // u_eta [e][qy][qx]

    const int tid = threadIdx.x;
    const int tilesX = (P.nEls0 + TX - 1) / TX;
    const int ex0 = (tile % tilesX) * TX, ey0 = (tile / tilesX) * TY;
    const int txo = min(TX, P.nEls0 - ex0), tyo = min(TY, P.nEls1 - ey0);
    const int M0 = P.nEls0 * N, M1 = P.nEls1 * N;
    const bool halo_c = ex0 > 0 || P.per0, halo_r = ey0 > 0 || P.per1;
    const SE_GLOBAL double* gin = gview(in);
    SE_GLOBAL double* gout = gview(out);
    cdouble* Dc = (cdouble*)P.D;
    cdouble* DTc = (cdouble*)P.DT;

    // ---- 1. the tile's node block and its tables; two barriers ----------------
    for (int i = tid; i < BWU + BH; i += 256) {
        if (i < BWU) s_I[i] = unk_line((ex0 - 1) * N + i, M0, P.per0, P.off0, P.nU0);
        else s_J[i - BWU] = unk_line((ey0 - 1) * N + i - BWU, M1, P.per1, P.off1, P.nU1);
    }
    {
        for (int i = tid; i < (TX + 1) * F; i += 256) {
            const int c = i / F, ex = ex0 - 1 + c;
            const int e = ex < 0 ? ex + P.nEls0 : ex;
            s_cf[i] = (e >= 0 && e < P.nEls0) ? gview(P.colTab)[(size_t)e * F + i % F] : 0.0;
        }
        for (int i = tid; i < (TY + 1) * F; i += 256) {
            const int r = i / F, ey = ey0 - 1 + r;
            const int e = ey < 0 ? ey + P.nEls1 : ey;
            s_rf[i] = (e >= 0 && e < P.nEls1) ? gview(P.rowTab)[(size_t)e * F + i % F] : 0.0;
        }
        for (int i = tid; i < (TX + 1) * QS; i += 256) {
            const int c = i / QS, ex = ex0 - 1 + c;
            const int e = ex < 0 ? ex + P.nEls0 : ex;
            s_cq[i] = (e >= 0 && e < P.nEls0) ? gview(P.colQ)[(size_t)e * QS + i % QS] : 0.0f;
        }
        for (int i = tid; i < (TY + 1) * QS; i += 256) {
            const int r = i / QS, ey = ey0 - 1 + r;
            const int e = ey < 0 ? ey + P.nEls1 : ey;
            s_rq[i] = (e >= 0 && e < P.nEls1) ? gview(P.rowQ)[(size_t)e * QS + i % QS] : 0.0f;
        }
    }
    __syncthreads();  // the index tables
    {
        constexpr int NL = (BH * BWU + 255) / 256;
        double v[NL];
#pragma unroll
        for (int g = 0; g < NL; g++) {
            const int i = tid + 256 * g;
            v[g] = 0.0;
            if (i < BH * BWU) {
                const int a = i / BWU, b = i - a * BWU;
                const int I = s_I[b], J = s_J[a];
                if (I >= 0 && J >= 0) v[g] = gin[(size_t)J * P.nU0 + I];
            }
        }
#pragma unroll
        for (int g = 0; g < NL; g++) {
            const int i = tid + 256 * g;
            if (i < BH * BWU) {
                const int a = i / BWU, b = i - a * BWU;
                s_u[a * BW + b] = v[g];
            }
        }
    }
    __syncthreads();

    // ---- 2.-3. element lines ---------------------------------------------------
    const int e = tid / N1, t = tid - (tid / N1) * N1;
    const int c = e % (TX + 1), r = e / (TX + 1);
    const bool live = e < NE && (c > 0 || halo_c) && (r > 0 || halo_r) && c <= txo && r <= tyo;
    double uxi[N1], uet[N1];
    if (live) {
        double ur[N1], uc[N1];
#pragma unroll
        for (int k = 0; k < N1; k++) {
            ur[k] = s_u[(r * N + t) * BW + c * N + k];
            uc[k] = s_u[(r * N + k) * BW + c * N + t];
        }
#pragma unroll
        for (int q = 0; q < N1; q++) {
            double sx = 0.0, sy = 0.0;
#pragma unroll
            for (int k = 0; k < N1; k++) {
                sx += Dc[q * N1 + k] * ur[k];
                sy += Dc[q * N1 + k] * uc[k];
            }
            uxi[q] = sx;  // u_xi at (qx = q, qy = t)
            uet[q] = sy;  // u_eta at (qx = t, qy = q)
            __builtin_amdgcn_sched_barrier(0);  // one row of D live at a time
        }
#pragma unroll
        for (int q = 0; q < N1; q++) {
            s_xf[e * NP + q * N1 + t] = (float)uxi[q];  // [qx = q][qy = t]
            s_ef[e * NP + q * N1 + t] = (float)uet[q];  // [qy = q][qx = t]
        }
    }
    __syncthreads();
    double yx[N1], yy[N1];
    if (live) {
        const double* cf = &s_cf[c * F];
        const double* rf = &s_rf[r * F];
        const float* cq = &s_cq[c * QS];
        const float* rq = &s_rq[r * QS];
        double fx[N1], fy[N1];
        // row t (qy = t, qx = q): g01 = A2[q] QYw[t][q] + QXw[q][t] Y2w[t]
        const double y0 = rf[t], y2w = rf[2 * N1 + t];
#pragma unroll
        for (int q = 0; q < N1; q++) {
            const double ue = (double)s_ef[e * NP + t * N1 + q];  // u_eta at (q, t)
            const double g01 = cf[2 * N1 + q] * (double)rq[t * N1 + q] + (double)cq[NP + t * N1 + q] * y2w;
            fx[q] = cf[q] * y0 * uxi[q] + g01 * ue;
        }
        // column t (qx = t, qy = q): g01 = A2[t] QYw[q][t] + QXw[t][q] Y2w[q]
        const double a1 = cf[N1 + t], a2 = cf[2 * N1 + t];
#pragma unroll
        for (int q = 0; q < N1; q++) {
            const double ux = (double)s_xf[e * NP + t * N1 + q];  // u_xi at (t, q)
            const double g01 = a2 * (double)rq[NP + t * N1 + q] + (double)cq[t * N1 + q] * rf[2 * N1 + q];
            fy[q] = g01 * ux + a1 * rf[N1 + q] * uet[q];
        }
        const double my = rf[3 * N1 + t];
#pragma unroll
        for (int l = 0; l < N1; l++) {
            double sx = 0.0, sy = 0.0;
#pragma unroll
            for (int q = 0; q < N1; q++) {
                sx += DTc[l * N1 + q] * fx[q];
                sy += DTc[l * N1 + q] * fy[q];
            }
            // node (l, t); the mass term re-reads u (fewer live registers)
            yx[l] = sx + my * cf[3 * N1 + l] * s_u[(r * N + t) * BW + c * N + l];
            yy[l] = sy;                                // node (t, l)
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    __syncthreads();  // the exchange arrays are read; s_o takes the outputs
    if (live) {
#pragma unroll
        for (int l = 0; l < N1; l++) s_o[e * NP + l * N1 + t] = yy[l];
    }
    __syncthreads();
    if (live) {
#pragma unroll
        for (int l = 0; l < N1; l++) s_o[e * NP + t * N1 + l] += yx[l];
    }
    __syncthreads();

    // ---- 4. owned nodes ----------------------------------------------------------
    const bool close_x = !P.per0 && ex0 + txo == P.nEls0, close_y = !P.per1 && ey0 + tyo == P.nEls1;
    const int NOX = txo * N + (close_x ? 1 : 0), NOY = tyo * N + (close_y ? 1 : 0);
    double dot = 0.0, st = 0.0, st2 = 0.0;
    for (int i = tid; i < NOX * NOY; i += 256) {
        const int oy = i / NOX, ox = i - oy * NOX;
        const int I = s_I[N + ox], J = s_J[N + oy];
        if (I < 0 || J < 0) continue;
        const int cq = ox / N, lx = ox - cq * N, rq = oy / N, ly = oy - rq * N;
        // element columns / rows holding the node: (c, local index) pairs
        int cs[2], ls[2], nc = 0, rs[2], ms[2], nr = 0;
        if (lx == 0 && (cq > 0 || halo_c)) cs[nc] = cq, ls[nc] = N, nc++;
        if (cq < txo) cs[nc] = cq + 1, ls[nc] = lx, nc++;
        if (ly == 0 && (rq > 0 || halo_r)) rs[nr] = rq, ms[nr] = N, nr++;
        if (rq < tyo) rs[nr] = rq + 1, ms[nr] = ly, nr++;
        double sum = 0.0;
        for (int jr = 0; jr < nr; jr++)
            for (int jc = 0; jc < nc; jc++) sum += s_o[(rs[jr] * (TX + 1) + cs[jc]) * NP + ms[jr] * N1 + ls[jc]];
        const double yv = P.scale * sum;
        const size_t gi = (size_t)J * P.nU0 + I;
        if (MODE == AP_TRUE && b) {
            const double ti = b[gi] - yv;
            gout[gi] = ti;
            st += ti;
            st2 += ti * ti;
        } else {
            gout[gi] = yv;
            if (MODE == AP_DOT) dot += s_u[(N + oy) * BW + N + ox] * yv;
        }
    }
    if (!partials) return;
    if (MODE == AP_DOT) {
        double v[1] = {dot};
        block_store_partials_at<1, 256>(v, partials, tile, np);
    } else if (MODE == AP_TRUE) {
        double v[2] = {st, st2};
        block_store_partials_at<2, 256>(v, partials, tile, np);
    }
}

template <int N, bool DOT>
__global__ __launch_bounds__(256) void k_apply_lines(OpParams P, const double* __restrict__ in,
                                                      double* __restrict__ out, double* __restrict__ partials,
                                                      const CgState* __restrict__ st) {
    if (DOT && cg_stopped(st)) return;
    // workgroup b runs on XCD b mod 8: give each XCD a contiguous run of
    // tiles, so that neighbouring tiles' shared halo lines meet in one L2
    const int T = gridDim.x, b = blockIdx.x, q = T / 8, r = T % 8, x = b % 8;
    const int tile = x * q + min(x, r) + b / 8;
    apply_lines_tile<N, DOT ? AP_DOT : AP_PLAIN>(P, tile, in, out, partials, T);
}

// Batched: work items (member, tile) over members of one order; MODE's
// partials go to the member's `part` (line_tiles of its mesh entries).
// gated: a member whose `gate` CG is not awaiting its true-residual check
// (the direction kernel before this launch marks it) has nothing to do.
// Four workgroups per CU (at most 128 VGPRs; at N = 8 a few values spill to
// scratch): 4% faster on the ensemble than three at 145 VGPRs.
template <int N, int MODE>
__global__ __launch_bounds__(256, 4) void k_apply_lines_batch(const ApplyMember* __restrict__ members,
                                                            const int2* __restrict__ work, int nwork, int gated) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        if (it.x < 0) continue;  // a padding item of an XCD-ordered list
        const ApplyMember& a = members[it.x];
        if (gated && a.gate && a.gate->status != CG_PENDING) continue;
        __syncthreads();  // the previous item's LDS reads are done
        apply_lines_tile<N, MODE>(a.P, it.y, a.in, a.out, a.part, a.np, a.b);
    }
}

}  // namespace se
