// helm_kernels.h -- device code of the matrix-free GLL Helmholtz operator.
//
// Setup-time modes (lifting, right-hand side, diagonal) and a reference form
// of the apply; the hot apply is k_apply_lines in apply_lines.h.
//
// One workgroup owns a tile of TX x TY elements and produces the assembled
// result on the tile's nodes (the half-open node block of its elements plus,
// on a non-periodic domain edge, the closing line).  Each node value is the
// sum of the contributions of every element containing the node (up to four,
// including the periodic wrap), so the tile also evaluates the element column
// to its left and the element row below it: no atomics, no second pass, and
// the result is bitwise deterministic.
//
// Per element the operator is applied by sum factorisation with the exact
// per-point 2x2 Jacobian of the reference (Mesh::ShapeFuncDerivs,
// src/base/Mesh.cpp:217-250), rebuilt from small per-column / per-row tables:
//   u_xi(q)  = sum_i D[qx][i] u(i,qy),   u_eta(q) = sum_j D[qy][j] u(qx,j)
//   g = Jinv (u_xi, u_eta),  f = W_q Jinv^T g
//   (K u)(r) = sum_qx D[qx][rx] f_xi(qx,ry) + sum_qy D[qy][ry] f_eta(rx,qy)
// with W_q = w_qx w_qy c_lap (literal mode, detJ := 1 as in
// Laplacian::AssembleElement, src/base/Operator.cpp:72) or det_q w_qx w_qy c_lap
// (physical), plus the GLL-diagonal mass det_q w_qx w_qy c_mass.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace se {

enum TileMode { MODE_APPLY = 0, MODE_LIFT = 1, MODE_RHS = 2, MODE_DIAG = 3 };

struct OpParams {
    int nEls0, nEls1, nv0, nv1, per0, per1;
    int detj;              // 0 literal, 1 physical
    double c_mass, c_lap;  // operator coefficients (reference sign)
    double scale;          // output scale (+1, or -1 for the sign-normalised solve)
    const int32_t* f2v;
    const double* J00;  // [nEls0][N+1]
    const double* J10;  // [nEls0][N+1 (qx)][N+1 (qy)]
    const double* J11;  // [nEls1][N+1]
    const double* J01;  // [nEls1][N+1 (qy)][N+1 (qx)]
    const double* w;    // [N+1]
    const double* D;    // [(N+1)^2]
    const double* DT;   // [(N+1)^2], D transposed
    // the line apply's metric tables per element column / row, with the
    // operator's coefficients folded in (op_metric_tables): FP64 line factors
    // (stride 4(N+1)) and FP32 cross-term tables, each followed by its
    // transpose (stride 2 (N+1)^2)
    const double* colTab;
    const double* rowTab;
    const float* colQ;
    const float* rowQ;
    // the unknown grid: nU0 x nU1, first unknown field column / row
    int nU0, nU1, off0, off1;
};

// Scalars of the CG recurrence, in device memory.  Written by block 0 of the
// kernel that finalises them; every block of a consuming kernel recomputes the
// same totals from the producer's per-block partials in a fixed order.
struct CgState {
    double rho, rho_prev, pq, alpha, beta, rr, bb, tol2, mu, mu_b, mu_t, tt;
    int iter, stop, status, maxit, null_space, n;
    unsigned counter;
    int pad;
};

enum CgStatus { CG_RUNNING = -1, CG_PENDING = -2 };

// The stop flag as a kernel sees it at its start: the value earlier kernels
// (or a stream-ordered host copy) left, which the kernel boundary makes
// visible to an ordinary load.  A write by the lead workgroup of the same
// kernel may or may not be seen, and either is fine: every workgroup takes
// the same decision from the same totals.  A plain read on purpose: read
// before the kernel stores anything, it compiles to one scalar load per wave.
// Any vector load of the one word (atomic at agent, workgroup or wavefront
// scope alike) has every wave of hundreds of resident workgroups ask for
// the same line, and those requests queue: about 12 us per launch of 1024
// workgroups, a third of a p = 8 apply.
__device__ __forceinline__ int cg_stopped(const CgState* st) { return st->stop; }

// Sum over the 64 lanes of a wave, returned to every lane.  Inside each row of
// 16 lanes a shift-and-add scan with DPP row shifts (lanes shifted in from
// outside the row read 0) leaves the row total in its last lane; the four row
// totals are then read as wave-uniform values and added in lane order.  The
// summation order is fixed, so the result is deterministic.
__device__ __forceinline__ double dpp_shr(double v, int sh) {
    const long long b = __double_as_longlong(v);
    int lo = (int)b, hi = (int)(b >> 32);
    switch (sh) {
        case 1:
            lo = __builtin_amdgcn_update_dpp(0, lo, 0x111, 0xf, 0xf, true);
            hi = __builtin_amdgcn_update_dpp(0, hi, 0x111, 0xf, 0xf, true);
            break;
        case 2:
            lo = __builtin_amdgcn_update_dpp(0, lo, 0x112, 0xf, 0xf, true);
            hi = __builtin_amdgcn_update_dpp(0, hi, 0x112, 0xf, 0xf, true);
            break;
        case 4:
            lo = __builtin_amdgcn_update_dpp(0, lo, 0x114, 0xf, 0xf, true);
            hi = __builtin_amdgcn_update_dpp(0, hi, 0x114, 0xf, 0xf, true);
            break;
        default:
            lo = __builtin_amdgcn_update_dpp(0, lo, 0x118, 0xf, 0xf, true);
            hi = __builtin_amdgcn_update_dpp(0, hi, 0x118, 0xf, 0xf, true);
            break;
    }
    return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_value(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane), hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_shr(v, 1);
    v += dpp_shr(v, 2);
    v += dpp_shr(v, 4);
    v += dpp_shr(v, 8);
    return ((lane_value(v, 15) + lane_value(v, 31)) + lane_value(v, 47)) + lane_value(v, 63);
}

// Per-point inverse Jacobian and weights, in the reference's operation order
// (Mesh.cpp:227-236, Operator.cpp:69-73, 165-167).
struct PointGeom {
    double i00, i01, i10, i11, det;
};

__device__ __forceinline__ PointGeom point_geom(double j00, double j01, double j10, double j11) {
#pragma clang fp contract(off)
    PointGeom g;
    g.det = j00 * j11 - j01 * j10;
    g.i00 = j11 / g.det;
    g.i11 = j00 / g.det;
    g.i01 = -j01 / g.det;
    g.i10 = -j10 / g.det;
    return g;
}

// LDS column / row slot of local node index i of tile element column c.  The
// line shared by element columns 0 and 1 is stored twice, so that element 0
// (the left neighbour, possibly wrapped across a periodic seam) keeps its own
// view of that line; every other shared line is stored once.
template <int N>
__device__ __forceinline__ int slot(int c, int i) {
    return c == 0 ? i : (N + 1) + (c - 1) * N + i;
}

template <int N, int TX, int TY, int MODE, bool DOT>
__global__ __launch_bounds__(256) void k_helm_tile(OpParams P, const double* __restrict__ in,
                                                    double* __restrict__ out, const double* __restrict__ aux,
                                                    double* __restrict__ partials, CgState* __restrict__ st) {
    constexpr int N1 = N + 1;
    constexpr int NP = N1 * N1;
    constexpr int BW = (TX + 1) * N + 2;
    constexpr int BH = (TY + 1) * N + 2;
    constexpr int NE = (TX + 1) * (TY + 1);

    __shared__ double s_u[MODE == MODE_DIAG ? 1 : BH * BW];
    __shared__ double s_fx[MODE == MODE_APPLY || MODE == MODE_LIFT ? NE * NP : 1];
    __shared__ double s_fy[MODE == MODE_APPLY || MODE == MODE_LIFT ? NE * NP : 1];
    __shared__ double s_D[NP];
    __shared__ double s_w[N1];
    __shared__ double s_J00[(TX + 1) * N1], s_J10[(TX + 1) * NP];
    __shared__ double s_J11[(TY + 1) * N1], s_J01[(TY + 1) * NP];
    __shared__ int s_exw[TX + 1], s_eyw[TY + 1];

    const int tid = threadIdx.x;
    const int tilesX = (P.nEls0 + TX - 1) / TX;
    const int ex0 = (blockIdx.x % tilesX) * TX;
    const int ey0 = (blockIdx.x / tilesX) * TY;

    // wrapped element indices of the tile columns / rows (-1: none)
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
        int c = i / N1, q = i % N1, e = s_exw[c];
        s_J00[i] = e >= 0 ? P.J00[(size_t)e * N1 + q] : 1.0;
    }
    for (int i = tid; i < (TX + 1) * NP; i += 256) {
        int c = i / NP, q = i % NP, e = s_exw[c];
        s_J10[i] = e >= 0 ? P.J10[(size_t)e * NP + q] : 0.0;
    }
    for (int i = tid; i < (TY + 1) * N1; i += 256) {
        int r = i / N1, q = i % N1, e = s_eyw[r];
        s_J11[i] = e >= 0 ? P.J11[(size_t)e * N1 + q] : 1.0;
    }
    for (int i = tid; i < (TY + 1) * NP; i += 256) {
        int r = i / NP, q = i % NP, e = s_eyw[r];
        s_J01[i] = e >= 0 ? P.J01[(size_t)e * NP + q] : 0.0;
    }

    // ---- gather the tile's node block through the numbering map ----------
    if (MODE != MODE_DIAG) {
        for (int i = tid; i < BH * BW; i += 256) {
            int a = i / BW, b = i % BW;
            int c, ix, r, iy;
            if (b <= N) { c = 0; ix = b; }
            else { int bb = b - N1; c = 1 + bb / N; ix = bb % N; if (ix == 0 && c >= 2) { c -= 1; ix = N; } }
            if (a <= N) { r = 0; iy = a; }
            else { int aa = a - N1; r = 1 + aa / N; iy = aa % N; if (iy == 0 && r >= 2) { r -= 1; iy = N; } }
            int ex = s_exw[c], ey = s_eyw[r];
            double val = 0.0;
            if (ex >= 0 && ey >= 0) {
                int v = (ey * N + iy) * P.nv0 + ex * N + ix;
                if (MODE == MODE_APPLY) {
                    int u = P.f2v[v];
                    val = u >= 0 ? in[u] : 0.0;
                } else if (MODE == MODE_LIFT) {
                    val = P.f2v[v] < 0 ? in[v] : 0.0;
                } else {
                    val = in[v];
                }
            }
            s_u[i] = val;
        }
    }
    __syncthreads();

    // ---- per-point gradient, metric, flux --------------------------------
    if (MODE == MODE_APPLY || MODE == MODE_LIFT) {
        for (int i = tid; i < NE * NP; i += 256) {
            int e = i / NP, q = i % NP;
            int c = e % (TX + 1), r = e / (TX + 1);
            int qx = q % N1, qy = q / N1;
            double fx = 0.0, fy = 0.0;
            if (s_exw[c] >= 0 && s_eyw[r] >= 0) {
                const double* row = &s_u[slot<N>(r, qy) * BW];
                double ux = 0.0, uy = 0.0;
#pragma unroll
                for (int k = 0; k < N1; k++) {
                    ux += s_D[qx * N1 + k] * row[slot<N>(c, k)];
                    uy += s_D[qy * N1 + k] * s_u[slot<N>(r, k) * BW + slot<N>(c, qx)];
                }
                PointGeom g = point_geom(s_J00[c * N1 + qx], s_J01[r * NP + qy * N1 + qx], s_J10[c * NP + qx * N1 + qy],
                                         s_J11[r * N1 + qy]);
                double W;
                {
#pragma clang fp contract(off)
                    double wq = s_w[qx] * s_w[qy];
                    W = (P.detj ? (g.det * wq) : (1.0 * wq)) * P.c_lap;
                }
                double gx = g.i00 * ux + g.i01 * uy;
                double gy = g.i10 * ux + g.i11 * uy;
                fx = W * (g.i00 * gx + g.i10 * gy);
                fy = W * (g.i01 * gx + g.i11 * gy);
            }
            s_fx[i] = fx;
            s_fy[i] = fy;
        }
        __syncthreads();
    }

    // ---- owned nodes: sum element contributions --------------------------
    const int txo = min(TX, P.nEls0 - ex0), tyo = min(TY, P.nEls1 - ey0);
    const int ownX = txo * N + ((!P.per0 && ex0 + txo == P.nEls0) ? 1 : 0);
    const int ownY = tyo * N + ((!P.per1 && ey0 + tyo == P.nEls1) ? 1 : 0);
    for (int i = tid; i < ownX * ownY; i += 256) {
        int ox = i % ownX, oy = i / ownX;
        int gI = ex0 * N + ox, gJ = ey0 * N + oy;
        int u = P.f2v[gJ * P.nv0 + gI];
        if (u < 0) continue;
        // containing elements: (tile column, local index) pairs
        int cs[2], xs[2], nc = 0, rs[2], ys[2], nr = 0;
        if (ox % N == 0) { cs[nc] = ox / N; xs[nc] = N; nc++; }
        if (ox / N < txo) { cs[nc] = 1 + ox / N; xs[nc] = ox % N; nc++; }
        if (oy % N == 0) { rs[nr] = oy / N; ys[nr] = N; nr++; }
        if (oy / N < tyo) { rs[nr] = 1 + oy / N; ys[nr] = oy % N; nr++; }
        double sum = 0.0, own = 0.0;
        for (int jr = 0; jr < nr; jr++) {
            const int r = rs[jr], ry = ys[jr];
            if (s_eyw[r] < 0) continue;
            for (int jc = 0; jc < nc; jc++) {
                const int c = cs[jc], rx = xs[jc];
                if (s_exw[c] < 0) continue;
                PointGeom g = point_geom(s_J00[c * N1 + rx], s_J01[r * NP + ry * N1 + rx], s_J10[c * NP + rx * N1 + ry],
                                         s_J11[r * N1 + ry]);
                double wq;
                {
#pragma clang fp contract(off)
                    wq = s_w[rx] * s_w[ry];
                }
                if (MODE == MODE_APPLY || MODE == MODE_LIFT) {
                    const int e = r * (TX + 1) + c;
                    const double* fx = &s_fx[e * NP];
                    const double* fy = &s_fy[e * NP];
                    double k = 0.0;
#pragma unroll
                    for (int q = 0; q < N1; q++) {
                        k += s_D[q * N1 + rx] * fx[ry * N1 + q];
                        k += s_D[q * N1 + ry] * fy[q * N1 + rx];
                    }
                    double uv = s_u[slot<N>(r, ry) * BW + slot<N>(c, rx)];
                    double mass;
                    {
#pragma clang fp contract(off)
                        mass = (g.det * wq) * P.c_mass;
                    }
                    sum += k + mass * uv;
                    own = uv;
                } else if (MODE == MODE_RHS) {
                    double fv = s_u[slot<N>(r, ry) * BW + slot<N>(c, rx)];
                    double m;
                    {
#pragma clang fp contract(off)
                        m = ((g.det * wq) * 1.0) * fv;
                    }
                    sum += m;
                } else {  // MODE_DIAG: element diagonal K~[r][r] + M[r][r]
                    double d = 0.0;
                    // quadrature points on row ry: q = (qx, ry)
                    for (int qx = 0; qx < N1; qx++) {
                        PointGeom h = point_geom(s_J00[c * N1 + qx], s_J01[r * NP + ry * N1 + qx],
                                                 s_J10[c * NP + qx * N1 + ry], s_J11[r * N1 + ry]);
                        double a0 = s_D[qx * N1 + rx], a1 = (qx == rx) ? s_D[ry * N1 + ry] : 0.0;
                        double g0 = a0 * h.i00 + a1 * h.i01, g1 = a0 * h.i10 + a1 * h.i11;
                        double wqq = s_w[qx] * s_w[ry];
                        double W = (P.detj ? h.det * wqq : wqq) * P.c_lap;
                        d += W * (g0 * g0 + g1 * g1);
                    }
                    // points on column rx other than the node's own row
                    for (int qy = 0; qy < N1; qy++) {
                        if (qy == ry) continue;
                        PointGeom h = point_geom(s_J00[c * N1 + rx], s_J01[r * NP + qy * N1 + rx],
                                                 s_J10[c * NP + rx * N1 + qy], s_J11[r * N1 + qy]);
                        double a1 = s_D[qy * N1 + ry];
                        double g0 = a1 * h.i01, g1 = a1 * h.i11;
                        double wqq = s_w[rx] * s_w[qy];
                        double W = (P.detj ? h.det * wqq : wqq) * P.c_lap;
                        d += W * (g0 * g0 + g1 * g1);
                    }
                    sum += d + (g.det * wq) * P.c_mass;
                }
            }
        }
        if (MODE == MODE_APPLY) {
            out[u] = P.scale * sum;
            (void)own;
        } else if (MODE == MODE_LIFT) {
            out[u] = -sum;
        } else if (MODE == MODE_RHS) {
            out[u] = aux ? sum + 1.0 * aux[u] : sum;
        } else {
            out[u] = 1.0 / (P.scale * sum);
        }
    }

}

}  // namespace se
