// helm.hip -- Helmholtz/Poisson operator objects, the device-resident FP64
// Jacobi-PCG and the C ABI for them.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "apply_kernels.h"
#include "apply_lines.h"
#include "helm_kernels.h"
#include "se_common.h"
#include "se_op.h"

using namespace se;

namespace {

constexpr int kBlock = 256;   // setup / tile kernels
constexpr int kVB = 1024;     // vector kernels of the CG loop

// ---------------------------------------------------------------------------
// CG vector kernels.  Each kernel stores per-block partial sums with plain
// stores; every block of the next kernel reduces them in the same fixed order
// (reduce_partials), so the recurrence scalars are identical in all blocks and
// deterministic, with no atomics or fences in the iteration.  Block 0 of the
// consuming kernel also records the scalars in CgState for the host.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kVB) void k_sum(int n, const double* __restrict__ a, double scale,
                                             double* __restrict__ part) {
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += scale * a[i];
    double v[1] = {s};
    block_store_partials<1, kVB>(v, part);
}

// b = sgn*f (- mean of sgn*f), x = 0, r = b; partials of (r.r, r.z, sum z, sum r)
__global__ __launch_bounds__(kVB) void k_cg_start(int n, double sgn, const double* __restrict__ f,
                                                  const double* __restrict__ dinv, double* __restrict__ b,
                                                  double* __restrict__ x, double* __restrict__ r,
                                                  const double* __restrict__ part_s, int np_s,
                                                  double* __restrict__ part_b, CgState* __restrict__ st) {
    double mu_b = 0.0;
    if (st->null_space) {
        double t[1];
        reduce_partials<1, kVB>(part_s, np_s, t);
        mu_b = t[0] / n;
        if (blockIdx.x == 0 && threadIdx.x == 0) st->mu_b = mu_b;
    }
    double rr = 0.0, rz = 0.0, sz = 0.0, sr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double bi = sgn * f[i] - mu_b;
        b[i] = bi;
        r[i] = bi;
        x[i] = 0.0;
        double z = dinv[i] * bi;
        rr += bi * bi;
        rz += bi * z;
        sz += z;
        sr += bi;
    }
    double v[4] = {rr, rz, sz, sr};
    block_store_partials<4, kVB>(v, part_b);
}

enum DirMode { DIR_FIRST = 0, DIR_NORMAL = 1, DIR_RESTART = 2 };

// Reduce (r.r, r.z, sum z, sum r); decide convergence; p = (D^-1 r - mu) + beta p.
__global__ __launch_bounds__(kVB) void k_cg_dir(int mode, int n, double rtol, const double* __restrict__ dinv,
                                                const double* __restrict__ r, double* __restrict__ p,
                                                const double* __restrict__ part_b, int np_b,
                                                double* __restrict__ hist, CgState* __restrict__ st) {
    if (mode == DIR_NORMAL && cg_stopped(st)) return;
    double t[4];
    reduce_partials<4, kVB>(part_b, np_b, t);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const double rr = t[0];
    const bool finite = (rr - rr == 0.0) && (t[1] - t[1] == 0.0);
    double mu = st->null_space ? t[2] / n : 0.0;
    double rho_new = t[1] - mu * t[3];
    double beta = 0.0;
    if (mode == DIR_FIRST) {
        if (lead) {
            st->bb = rr;
            st->tol2 = rtol * rtol * rr;
            st->rr = rr;
            st->rho = rho_new;
            st->mu = mu;
            st->iter = 0;
            st->stop = (rr == 0.0 || !finite) ? 1 : 0;
            st->status = rr == 0.0 ? 0 : (finite ? CG_RUNNING : 3);
            if (hist) hist[0] = 1.0;
        }
        if (rr == 0.0 || !finite) return;
    } else if (mode == DIR_NORMAL) {
        const bool conv = finite && rr <= st->tol2;
        if (lead) {
            int it = st->iter + 1;
            st->iter = it;
            st->rr = rr;
            if (hist) hist[it] = sqrt(rr / st->bb);
            if (!finite) {
                st->stop = 1;
                st->status = 3;
            } else if (conv) {
                st->stop = 1;
                st->status = CG_PENDING;
            } else {
                st->rho = rho_new;
                st->mu = mu;
                st->beta = rho_new / st->rho_prev;
                if (it >= st->maxit) {
                    st->stop = 1;
                    st->status = 1;
                }
            }
        }
        if (!finite || conv) return;
        beta = rho_new / st->rho_prev;
    } else {  // DIR_RESTART: continue from the true residual, no convergence test
        // rho_prev still holds the rho of the candidate iteration (its k_cg_dir
        // stopped before updating rho)
        beta = rho_new / st->rho_prev;
        if (lead) {
            st->rr = rr;
            st->rho = rho_new;
            st->mu = mu;
            st->beta = beta;
        }
    }
    // unknowns in pairs, one 16-byte access per vector and pair (the operator's
    // own buffers, 256-byte aligned); the odd last unknown goes to thread 0
    const int tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    const double2* __restrict__ d2 = reinterpret_cast<const double2*>(dinv);
    const double2* __restrict__ r2 = reinterpret_cast<const double2*>(r);
    double2* __restrict__ p2 = reinterpret_cast<double2*>(p);
    for (int i = tid; i < (n >> 1); i += stride) {
        const double2 dv = d2[i], rv = r2[i];
        double2 pv;
        pv.x = dv.x * rv.x - mu;
        pv.y = dv.y * rv.y - mu;
        if (mode != DIR_FIRST) {
            const double2 po = p2[i];
            pv.x += beta * po.x;
            pv.y += beta * po.y;
        }
        p2[i] = pv;
    }
    if ((n & 1) && tid == 0) {
        const int i = n - 1;
        double z = dinv[i] * r[i] - mu;
        p[i] = mode == DIR_FIRST ? z : z + beta * p[i];
    }
}

// alpha = rho / (p.Ap); x += alpha p, r -= alpha q; partials of (r.r, r.z, sum z, sum r)
__global__ __launch_bounds__(kVB) void k_cg_update(int n, const double* __restrict__ p, const double* __restrict__ q,
                                                   const double* __restrict__ dinv, double* __restrict__ x,
                                                   double* __restrict__ r, const double* __restrict__ part_a, int np_a,
                                                   double* __restrict__ part_b, CgState* __restrict__ st) {
    if (cg_stopped(st)) return;
    double t[1];
    reduce_partials<1, kVB>(part_a, np_a, t);
    const double pq = t[0];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (!(pq > 0.0)) {
        if (lead) {
            st->pq = pq;
            st->stop = 1;
            st->status = (pq - pq == 0.0) ? 2 : 3;
        }
        return;
    }
    const double rho = st->rho;
    const double alpha = rho / pq;
    if (lead) {
        st->pq = pq;
        st->alpha = alpha;
        st->rho_prev = rho;
    }
    double rr = 0.0, rz = 0.0, sz = 0.0, sr = 0.0;
    auto acc = [&](double ri, double di) {
        const double z = di * ri;
        rr += ri * ri;
        rz += ri * z;
        sz += z;
        sr += ri;
    };
    // unknowns in pairs, one 16-byte access per vector and pair (the operator's
    // own buffers, 256-byte aligned); the odd last unknown goes to thread 0
    const int tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    const double2* __restrict__ p2 = reinterpret_cast<const double2*>(p);
    const double2* __restrict__ q2 = reinterpret_cast<const double2*>(q);
    const double2* __restrict__ d2 = reinterpret_cast<const double2*>(dinv);
    double2* __restrict__ x2 = reinterpret_cast<double2*>(x);
    double2* __restrict__ r2 = reinterpret_cast<double2*>(r);
    for (int i = tid; i < (n >> 1); i += stride) {
        const double2 pv = p2[i], qv = q2[i], dv = d2[i];
        double2 xv = x2[i], rv = r2[i];
        xv.x += alpha * pv.x;
        xv.y += alpha * pv.y;
        rv.x -= alpha * qv.x;
        rv.y -= alpha * qv.y;
        x2[i] = xv;
        r2[i] = rv;
        acc(rv.x, dv.x);
        acc(rv.y, dv.y);
    }
    if ((n & 1) && tid == 0) {
        const int i = n - 1;
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        acc(ri, dinv[i]);
    }
    double v[4] = {rr, rz, sz, sr};
    block_store_partials<4, kVB>(v, part_b);
}

// t = b - q (q = A x); r = t; partials of (sum t, sum t^2)
__global__ __launch_bounds__(kVB) void k_true_resid(int n, const double* __restrict__ b, const double* __restrict__ q,
                                                    double* __restrict__ r, double* __restrict__ part_t) {
    double s = 0.0, s2 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double t = b[i] - q[i];
        r[i] = t;
        s += t;
        s2 += t * t;
    }
    double v[2] = {s, s2};
    block_store_partials<2, kVB>(v, part_t);
}

// single block: ||t - mean t||^2 and the mean, into the state
__global__ __launch_bounds__(kVB) void k_true_fin(int n, const double* __restrict__ part_t, int np_t,
                                                  CgState* __restrict__ st) {
    double t[2];
    reduce_partials<2, kVB>(part_t, np_t, t);
    if (threadIdx.x == 0) {
        double mt = st->null_space ? t[0] / n : 0.0;
        st->mu_t = mt;
        st->tt = st->null_space ? t[1] - t[0] * mt : t[1];
    }
}

// continue from the true residual: r -= mean (null space); partials as k_cg_update
__global__ __launch_bounds__(kVB) void k_cg_restart(int n, const double* __restrict__ dinv, double* __restrict__ r,
                                                    double* __restrict__ part_b, const CgState* __restrict__ st) {
    const double mt = st->null_space ? st->mu_t : 0.0;
    double rr = 0.0, rz = 0.0, sz = 0.0, sr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double ri = r[i] - mt;
        r[i] = ri;
        double z = dinv[i] * ri;
        rr += ri * ri;
        rz += ri * z;
        sz += z;
        sr += ri;
    }
    double v[4] = {rr, rz, sz, sr};
    block_store_partials<4, kVB>(v, part_b);
}

__global__ void k_scale_copy(int n, double s, const double* __restrict__ a, double* __restrict__ b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = s * a[i];
}

// ---------------------------------------------------------------------------
// tile-kernel dispatch over (N, tile, mode)
// ---------------------------------------------------------------------------
template <int N, int TX, int TY, int MODE, bool DOT>
void launch_tile_t(const OpParams& P, const double* in, double* out, const double* aux, double* partials,
                   CgState* st, hipStream_t s) {
    int tiles = ((P.nEls0 + TX - 1) / TX) * ((P.nEls1 + TY - 1) / TY);
    hipLaunchKernelGGL((k_helm_tile<N, TX, TY, MODE, DOT>), dim3(tiles), dim3(kBlock), 0, s, P, in, out, aux,
                       partials, st);
}

template <int N, int MODE, bool DOT>
void launch_tile_n(int tile, const OpParams& P, const double* in, double* out, const double* aux, double* partials,
                   CgState* st, hipStream_t s) {
    if (tile == 4) launch_tile_t<N, 4, 4, MODE, DOT>(P, in, out, aux, partials, st, s);
    else launch_tile_t<N, 2, 2, MODE, DOT>(P, in, out, aux, partials, st, s);
}

template <int N, bool DOT>
void launch_apply_t(const OpParams& P, const GLL& g, const double* in, double* out, double* partials,
                    const CgState* st, hipStream_t s) {
    hipLaunchKernelGGL((k_apply_lines<N, DOT>), dim3(line_tiles(N, P.nEls0, P.nEls1)), dim3(kBlock), 0, s, P, in,
                       out, partials, st);
}

template <bool DOT>
int launch_apply(int N, const OpParams& P, const GLL& g, const double* in, double* out, double* partials,
                 const CgState* st, hipStream_t s) {
#define SE_APPLY_CASE(NN)                                              \
    case NN: launch_apply_t<NN, DOT>(P, g, in, out, partials, st, s); break;
    switch (N) {
        SE_APPLY_CASE(2)
        SE_APPLY_CASE(3)
        SE_APPLY_CASE(4)
        SE_APPLY_CASE(5)
        SE_APPLY_CASE(6)
        SE_APPLY_CASE(7)
        SE_APPLY_CASE(8)
        default: return fail(SE_ERR_UNSUPPORTED, "operator kernels are built for orders 2..8");
    }
#undef SE_APPLY_CASE
    return SE_OK;
}

template <int MODE, bool DOT>
int launch_tile(int N, int tile, const OpParams& P, const double* in, double* out, const double* aux,
                double* partials, CgState* st, hipStream_t s) {
    switch (N) {
        case 2: launch_tile_n<2, MODE, DOT>(tile, P, in, out, aux, partials, st, s); break;
        case 3: launch_tile_n<3, MODE, DOT>(tile, P, in, out, aux, partials, st, s); break;
        case 4: launch_tile_n<4, MODE, DOT>(tile, P, in, out, aux, partials, st, s); break;
        case 5: launch_tile_n<5, MODE, DOT>(tile, P, in, out, aux, partials, st, s); break;
        case 6: launch_tile_n<6, MODE, DOT>(tile, P, in, out, aux, partials, st, s); break;
        case 7: launch_tile_n<7, MODE, DOT>(tile, P, in, out, aux, partials, st, s); break;
        case 8: launch_tile_n<8, MODE, DOT>(tile, P, in, out, aux, partials, st, s); break;
        default: return fail(SE_ERR_UNSUPPORTED, "operator kernels are built for orders 2..8");
    }
    return SE_OK;
}

int vec_blocks(int n) {
    int b = (n + kVB * 4 - 1) / (kVB * 4);
    return std::max(1, std::min(b, 256));
}

int apply_tiles(const se_op* op) {
    const se_mesh* m = op->mesh;
    return line_tiles(m->N, m->nEls[0], m->nEls[1]);
}

}  // namespace

// ---------------------------------------------------------------------------
// operator object
// ---------------------------------------------------------------------------
namespace se {

OpParams op_params(const se_op* op, double scale) {
    const se_mesh* m = op->mesh;
    OpParams P;
    P.nEls0 = m->nEls[0];
    P.nEls1 = m->nEls[1];
    P.nv0 = m->nVerts[0];
    P.nv1 = m->nVerts[1];
    P.per0 = m->periodic[0];
    P.per1 = m->periodic[1];
    P.detj = op->detj;
    P.c_mass = op->c_mass;
    P.c_lap = op->c_lap;
    P.scale = scale;
    P.f2v = m->d_f2v.as<int32_t>();
    const int n1 = m->N + 1;
    P.J00 = m->d_J00.as<double>();
    P.J10 = m->d_J10.as<double>();
    P.J11 = m->d_J11.as<double>();
    P.J01 = m->d_J01.as<double>();
    const double* basis = m->d_basis.as<double>();
    P.w = basis + n1;
    P.D = basis + 2 * n1;
    P.DT = basis + 4 * n1 + n1 * n1;
    P.colTab = op->tabs.as<double>();
    P.rowTab = op->tabs.as<double>() + (size_t)m->nEls[0] * 4 * n1;
    P.colQ = op->qtabs.as<float>();
    P.rowQ = op->qtabs.as<float>() + (size_t)m->nEls[0] * 2 * n1 * n1;
    P.nU0 = m->nU[0];
    P.nU1 = m->nU[1];
    P.off0 = m->off[0];
    P.off1 = m->off[1];
    return P;
}

// Metric factor tables of the hot apply.  At quadrature point (qx,qy) of
// element (ex,ey) the reference's Jacobian (Mesh.cpp:217-236) has diagonal
// entries j00(ex,qx), j11(ey,qy) and round-off-sized off-diagonal entries
// j01(ey,qy,qx), j10(ex,qx,qy) (their size is that of the cancellation in
// sum_i D[q][i] * coordinate).  Keeping every term of the metric
// W Jinv Jinv^T down to relative order (j01 j10)/(j00 j11), about 1e-26:
//   g00 = c PX0(ex,qx) PY0(ey,qy),   g11 = c PX1(ex,qx) PY1(ey,qy),
//   g01 = -c w_qx w_qy [PX2(ex,qx) QY(ey,qy,qx) + QX(ex,qx,qy) PY2(ey,qy)],
//   mass = c_mass MX(ex,qx) MY(ey,qy)
// literal  (W = w c):      PX0 = w/j00^2, PX1 = w, PX2 = 1/j00^2, QX = j10/j00,
//                          PY0 = w, PY1 = w/j11^2, PY2 = 1/j11^2, QY = j01/j11
// physical (W = det w c):  PX0 = w/j00, PX1 = w j00, PX2 = 1/j00, QX = j10,
//                          PY0 = w j11, PY1 = w/j11, PY2 = 1/j11, QY = j01
// and MX = w j00, MY = w j11.  Per column / row the layout is
// [X0(N+1) X1(N+1) X2(N+1) M(N+1) Q((N+1)^2)].
static int op_metric_tables(se_op* op) {
    const se_mesh* m = op->mesh;
    const int n1 = m->N + 1, CS = 4 * n1 + n1 * n1;
    const std::vector<double>& w = m->gll.w;
    std::vector<double> t((size_t)(m->nEls[0] + m->nEls[1]) * CS, 0.0);
    const bool phys = op->detj == SE_DETJ_PHYSICAL;
    for (int ex = 0; ex < m->nEls[0]; ex++) {
        double* T = &t[(size_t)ex * CS];
        for (int q = 0; q < n1; q++) {
            const double j = m->J00[(size_t)ex * n1 + q];
            T[q] = phys ? w[q] / j : w[q] / (j * j);
            T[n1 + q] = phys ? w[q] * j : w[q];
            T[2 * n1 + q] = phys ? 1.0 / j : 1.0 / (j * j);
            T[3 * n1 + q] = w[q] * j;
            for (int qy = 0; qy < n1; qy++) {
                const double j10 = m->J10[((size_t)ex * n1 + q) * n1 + qy];
                T[4 * n1 + q * n1 + qy] = phys ? j10 : j10 / j;
            }
        }
    }
    for (int ey = 0; ey < m->nEls[1]; ey++) {
        double* T = &t[(size_t)(m->nEls[0] + ey) * CS];
        for (int q = 0; q < n1; q++) {
            const double j = m->J11[(size_t)ey * n1 + q];
            T[q] = phys ? w[q] * j : w[q];
            T[n1 + q] = phys ? w[q] / j : w[q] / (j * j);
            T[2 * n1 + q] = phys ? 1.0 / j : 1.0 / (j * j);
            T[3 * n1 + q] = w[q] * j;
            for (int qx = 0; qx < n1; qx++) {
                const double j01 = m->J01[((size_t)ey * n1 + q) * n1 + qx];
                T[4 * n1 + q * n1 + qx] = phys ? j01 : j01 / j;
            }
        }
    }
    // The line apply's tables, with the operator's coefficients folded in
    // (c = c_lap):
    //   per column  A0 = c X0, A1 = c X1, A2 = -c w X2, AM = c_mass M      (FP64)
    //   per row     Y0,        Y1,       Y2w = w Y2,    MY = M             (FP64)
    //   per column  QXw[qx][qy] = -c w[qx] QX[qx][qy], then its transpose   (FP32)
    //   per row     QYw[qy][qx] =    w[qy] QY[qy][qx], then its transpose   (FP32)
    // so that g00 = A0[qx] Y0[qy], g11 = A1[qx] Y1[qy] and
    //   g01 = A2[qx] QYw[qy][qx] + QXw[qx][qy] Y2w[qy],  mass = AM[qx] MY[qy].
    // The cross-term tables are round-off-sized, so FP32 holds them to ~1e-17
    // of the operator.
    const int np = n1 * n1, F = 4 * n1;
    const int ncol = m->nEls[0], nrow = m->nEls[1];
    std::vector<double> lf((size_t)(ncol + nrow) * F);
    std::vector<float> qt((size_t)(ncol + nrow) * 2 * np);
    const double c = op->c_lap;
    for (int e = 0; e < ncol + nrow; e++) {
        const double* T = &t[(size_t)e * CS];
        const double* Q = T + 4 * n1;
        double* L = &lf[(size_t)e * F];
        float* o = &qt[(size_t)e * 2 * np];
        const bool col = e < ncol;
        for (int q = 0; q < n1; q++) {
            L[q] = col ? c * T[q] : T[q];
            L[n1 + q] = col ? c * T[n1 + q] : T[n1 + q];
            L[2 * n1 + q] = col ? -c * w[q] * T[2 * n1 + q] : w[q] * T[2 * n1 + q];
            L[3 * n1 + q] = col ? op->c_mass * T[3 * n1 + q] : T[3 * n1 + q];
        }
        for (int a = 0; a < n1; a++)  // Q[a][b]: column table [qx][qy], row table [qy][qx]; w of the first index
            for (int b = 0; b < n1; b++) {
                const double v = col ? -c * w[a] * Q[a * n1 + b] : w[a] * Q[a * n1 + b];
                o[a * n1 + b] = (float)v;
                o[np + b * n1 + a] = (float)v;
            }
    }
    int rc = op->tabs.alloc(sizeof(double) * lf.size());
    if (!rc) rc = op->qtabs.alloc(sizeof(float) * qt.size());
    if (rc) return rc;
    SE_HIP(hipMemcpy(op->tabs.ptr, lf.data(), sizeof(double) * lf.size(), hipMemcpyHostToDevice));
    SE_HIP(hipMemcpy(op->qtabs.ptr, qt.data(), sizeof(float) * qt.size(), hipMemcpyHostToDevice));
    op->tabs_host = std::move(t);
    return SE_OK;
}

int op_tile(const se_op* op) {
    // 4x4 tiles when they still give every CU two workgroups, else 2x2
    const se_mesh* m = op->mesh;
    long t4 = (long)((m->nEls[0] + 3) / 4) * ((m->nEls[1] + 3) / 4);
    return (t4 >= 2L * m->ctx->cu_count) ? 4 : 2;
}

// partial-sum areas in op->partials (doubles): A apply (one per tile),
// B CG vectors (4 per block), S mean (1 per block), T true residual (2 per block)
struct PartLayout {
    size_t a, b, s, t, total;
};

static PartLayout part_layout(const se_mesh* m) {
    PartLayout L;
    size_t tiles = (size_t)((m->nEls[0] + 1) / 2) * ((m->nEls[1] + 1) / 2);
    L.a = 0;
    L.b = tiles;
    L.s = L.b + 4 * 256;
    L.t = L.s + 256;
    L.total = L.t + 2 * 256;
    return L;
}

int op_apply_dev(se_op* op, const double* d_x, double* d_y, double scale, bool dot) {
    OpParams P = op_params(op, scale);
    hipStream_t s = op->mesh->ctx->stream;
    double* pa = op->partials.as<double>() + part_layout(op->mesh).a;
    if (op->svv_mN > 0) {
        // mass + Laplacian, the SVV term, then the partials of x.y the CG
        // update reads (as many as the apply kernel would write)
        int rc = launch_apply<false>(op->mesh->N, P, op->mesh->gll, d_x, d_y, pa, op->state.as<CgState>(), s);
        if (rc) return rc;
        SE_HIP(hipGetLastError());
        rc = svv_apply(op, d_x, d_y, scale);
        if (!rc && dot) rc = svv_dot(op, d_x, d_y, pa, apply_tiles(op));
        return rc;
    }
    int rc = dot ? launch_apply<true>(op->mesh->N, P, op->mesh->gll, d_x, d_y, pa, op->state.as<CgState>(), s)
                 : launch_apply<false>(op->mesh->N, P, op->mesh->gll, d_x, d_y, pa, op->state.as<CgState>(), s);
    if (rc) return rc;
    SE_HIP(hipGetLastError());
    return SE_OK;
}

int op_ensure_work(se_op* op) {
    const se_mesh* m = op->mesh;
    size_t nb = sizeof(double) * (size_t)std::max(m->n, 1);
    int rc = SE_OK;
    if (!rc) rc = op->b.alloc(nb);
    if (!rc) rc = op->x.alloc(nb);
    if (!rc) rc = op->r.alloc(nb);
    if (!rc) rc = op->p.alloc(nb);
    if (!rc) rc = op->q.alloc(nb);
    if (!rc) rc = op->f.alloc(nb);
    if (!rc) rc = op->partials.alloc(sizeof(double) * part_layout(m).total);
    return rc;
}

}  // namespace se

extern "C" void se_solve_opts_default(se_solve_opts* o) {
    if (!o) return;
    o->rtol = 1e-10;
    o->maxit = 100000;
    o->precision = SE_PREC_FP64;
    o->null_space = 0;
    o->check_every = 32;
    o->inner_rtol = 0.0;
    o->max_outer = 1000;
    o->history = nullptr;
    o->preconditioner = SE_PC_DEFAULT;
}

extern "C" int se_helm_create(se_mesh* m, double c_mass, double c_lap, int detj_mode, se_op** out) {
    SE_CHECK(m && out, SE_ERR_ARG, "se_helm_create: NULL argument");
    SE_CHECK(detj_mode == SE_DETJ_LITERAL || detj_mode == SE_DETJ_PHYSICAL, SE_ERR_ARG, "se_helm_create: bad detj_mode");
    SE_CHECK(m->N >= 2 && m->N <= 8, SE_ERR_UNSUPPORTED, "se_helm_create: operator kernels cover orders 2..8");
    SE_HIP(hipSetDevice(m->ctx->device));
    se_op* op = new se_op;
    op->mesh = m;
    op->c_mass = c_mass;
    op->c_lap = c_lap;
    op->detj = detj_mode;
    op->tile = op_tile(op);
    size_t nb = sizeof(double) * (size_t)std::max(m->n, 1);
    int rc = op->dinv.alloc(nb);
    if (!rc) rc = op->lift.alloc(nb);
    if (!rc) rc = op->state.alloc(sizeof(CgState));
    if (!rc) rc = op_ensure_work(op);
    if (!rc) rc = op_metric_tables(op);
    if (rc) {
        se_helm_destroy(op);
        return rc;
    }
    hipStream_t s = m->ctx->stream;
    (void)hipMemsetAsync(op->state.ptr, 0, sizeof(CgState), s);
    (void)hipMemsetAsync(op->lift.ptr, 0, nb, s);
    // inverse diagonal of the assembled operator (Matrix::InverseDiagonal)
    OpParams P = op_params(op, 1.0);
    rc = launch_tile<MODE_DIAG, false>(m->N, 2, P, nullptr, op->dinv.as<double>(), nullptr, nullptr, nullptr, s);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(s);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) rc = fail(SE_ERR_HIP, std::string("diagonal kernel: ") + hipGetErrorString(e));
    }
    if (rc) {
        se_helm_destroy(op);
        return rc;
    }
    // sign of the operator: all diagonal entries must share it (K6 of the
    // reference QG operators: negative definite)
    std::vector<double> d(m->n);
    if (m->n > 0) (void)hipMemcpy(d.data(), op->dinv.ptr, nb, hipMemcpyDeviceToHost);
    int pos = 0, neg = 0;
    for (double v : d) {
        if (v > 0) pos++;
        else if (v < 0) neg++;
    }
    op->sign = (neg > 0 && pos == 0) ? -1 : 1;
    op->definite = (pos == m->n) || (neg == m->n);
    *out = op;
    return SE_OK;
}

extern "C" int se_helm_destroy(se_op* op) {
    if (!op) return SE_OK;
    (void)hipSetDevice(op->mesh->ctx->device);
    if (op->graph_exec) (void)hipGraphExecDestroy(op->graph_exec);
    op->release_all();
    delete op;
    return SE_OK;
}

extern "C" int se_helm_apply_dev(se_op* op, const double* d_x, double* d_y) {
    SE_CHECK(op && d_x && d_y, SE_ERR_ARG, "se_helm_apply_dev: NULL argument");
    return op_apply_dev(op, d_x, d_y, 1.0, false);
}

static int stage_in(se_ctx* c, DevBuf& dst, const double* host, size_t count) {
    int rc = dst.alloc(sizeof(double) * std::max<size_t>(count, 1));
    if (rc) return rc;
    if (count) SE_HIP(hipMemcpyAsync(dst.ptr, host, sizeof(double) * count, hipMemcpyHostToDevice, c->stream));
    return SE_OK;
}

static int stage_out(se_ctx* c, double* host, const DevBuf& src, size_t count) {
    if (count) SE_HIP(hipMemcpyAsync(host, src.ptr, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    SE_HIP(hipStreamSynchronize(c->stream));
    return SE_OK;
}

extern "C" int se_helm_apply(se_op* op, const double* x, double* y) {
    SE_CHECK(op && x && y, SE_ERR_ARG, "se_helm_apply: NULL argument");
    se_ctx* c = op->mesh->ctx;
    SE_HIP(hipSetDevice(c->device));
    int n = op->mesh->n;
    int rc = stage_in(c, op->sx, x, n);
    if (!rc) rc = op->sy.alloc(sizeof(double) * std::max(n, 1));
    if (!rc) rc = op_apply_dev(op, op->sx.as<double>(), op->sy.as<double>(), 1.0, false);
    if (!rc) rc = stage_out(c, y, op->sy, n);
    return rc;
}

extern "C" int se_helm_inv_diag(se_op* op, double* dinv) {
    SE_CHECK(op && dinv, SE_ERR_ARG, "se_helm_inv_diag: NULL argument");
    SE_HIP(hipSetDevice(op->mesh->ctx->device));
    SE_HIP(hipMemcpy(dinv, op->dinv.ptr, sizeof(double) * op->mesh->n, hipMemcpyDeviceToHost));
    return SE_OK;
}

extern "C" int se_helm_set_dirichlet_dev(se_op* op, const double* d_sol_field) {
    SE_CHECK(op && d_sol_field, SE_ERR_ARG, "se_helm_set_dirichlet_dev: NULL argument");
    OpParams P = op_params(op, 1.0);
    int rc = launch_tile<MODE_LIFT, false>(op->mesh->N, op->tile, P, d_sol_field, op->lift.as<double>(), nullptr,
                                           nullptr, nullptr, op->mesh->ctx->stream);
    if (rc) return rc;
    SE_HIP(hipGetLastError());
    if (op->svv_mN > 0) return svv_lift(op, d_sol_field, op->lift.as<double>());
    return SE_OK;
}

extern "C" int se_helm_set_dirichlet(se_op* op, const double* sol_field) {
    SE_CHECK(op && sol_field, SE_ERR_ARG, "se_helm_set_dirichlet: NULL argument");
    se_ctx* c = op->mesh->ctx;
    SE_HIP(hipSetDevice(c->device));
    int rc = stage_in(c, op->sfield, sol_field, op->mesh->nV);
    if (!rc) rc = se_helm_set_dirichlet_dev(op, op->sfield.as<double>());
    if (!rc) SE_HIP(hipStreamSynchronize(c->stream));
    return rc;
}

extern "C" int se_helm_lift(se_op* op, double* b) {
    SE_CHECK(op && b, SE_ERR_ARG, "se_helm_lift: NULL argument");
    SE_HIP(hipSetDevice(op->mesh->ctx->device));
    SE_HIP(hipStreamSynchronize(op->mesh->ctx->stream));
    SE_HIP(hipMemcpy(b, op->lift.ptr, sizeof(double) * op->mesh->n, hipMemcpyDeviceToHost));
    return SE_OK;
}

extern "C" int se_helm_rhs_dev(se_op* op, const double* d_rhs_field, double* d_f) {
    SE_CHECK(op && d_rhs_field && d_f, SE_ERR_ARG, "se_helm_rhs_dev: NULL argument");
    OpParams P = op_params(op, 1.0);
    int rc = launch_tile<MODE_RHS, false>(op->mesh->N, 2, P, d_rhs_field, d_f, op->lift.as<double>(), nullptr,
                                          nullptr, op->mesh->ctx->stream);
    if (rc) return rc;
    SE_HIP(hipGetLastError());
    return SE_OK;
}

extern "C" int se_helm_rhs(se_op* op, const double* rhs_field, double* f) {
    SE_CHECK(op && rhs_field && f, SE_ERR_ARG, "se_helm_rhs: NULL argument");
    se_ctx* c = op->mesh->ctx;
    SE_HIP(hipSetDevice(c->device));
    int rc = stage_in(c, op->sfield, rhs_field, op->mesh->nV);
    if (!rc) rc = op->sy.alloc(sizeof(double) * std::max(op->mesh->n, 1));
    if (!rc) rc = se_helm_rhs_dev(op, op->sfield.as<double>(), op->sy.as<double>());
    if (!rc) rc = stage_out(c, f, op->sy, op->mesh->n);
    return rc;
}

// ---------------------------------------------------------------------------
// FP64 Jacobi-PCG
// ---------------------------------------------------------------------------
namespace se {

// one CG iteration: q = A p (+ block partials of p.q), update, direction
static int cg_iteration(se_op* op, double* hist) {
    const se_mesh* m = op->mesh;
    hipStream_t s = m->ctx->stream;
    const int n = m->n, vb = vec_blocks(n);
    const PartLayout L = part_layout(m);
    CgState* st = op->state.as<CgState>();
    double* part = op->partials.as<double>();
    int rc = op_apply_dev(op, op->p.as<double>(), op->q.as<double>(), (double)op->sign, true);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cg_update, dim3(vb), dim3(kVB), 0, s, n, op->p.as<double>(), op->q.as<double>(),
                       op->dinv_s.as<double>(), op->x.as<double>(), op->r.as<double>(), part + L.a, apply_tiles(op),
                       part + L.b, st);
    hipLaunchKernelGGL(k_cg_dir, dim3(vb), dim3(kVB), 0, s, (int)DIR_NORMAL, n, 0.0, op->dinv_s.as<double>(),
                       op->r.as<double>(), op->p.as<double>(), part + L.b, vb, hist, st);
    return SE_OK;
}

static int cg_batch(se_op* op, int k, double* hist) {
    hipStream_t s = op->mesh->ctx->stream;
    static const bool use_graph = getenv("SE_NO_GRAPH") == nullptr;
    if (use_graph && hist == nullptr) {
        if (!op->graph_exec || op->graph_iters != k) {
            if (op->graph_exec) {
                (void)hipGraphExecDestroy(op->graph_exec);
                op->graph_exec = nullptr;
            }
            hipGraph_t g = nullptr;
            SE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = SE_OK;
            for (int i = 0; i < k && !rc; i++) rc = cg_iteration(op, nullptr);
            hipError_t e = hipStreamEndCapture(s, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(SE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&op->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                op->graph_exec = nullptr;
                return fail(SE_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            }
            op->graph_iters = k;
        }
        SE_HIP(hipGraphLaunch(op->graph_exec, s));
        return SE_OK;
    }
    for (int i = 0; i < k; i++) {
        int rc = cg_iteration(op, hist);
        if (rc) return rc;
    }
    return SE_OK;
}

// true residual of the current x: st->tt = ||b - sA x||^2 (mean removed with a null space)
static int true_residual(se_op* op) {
    const se_mesh* m = op->mesh;
    hipStream_t s = m->ctx->stream;
    const int n = m->n, vb = vec_blocks(n);
    const PartLayout L = part_layout(m);
    double* part = op->partials.as<double>();
    int rc = op_apply_dev(op, op->x.as<double>(), op->q.as<double>(), (double)op->sign, false);
    if (rc) return rc;
    hipLaunchKernelGGL(k_true_resid, dim3(vb), dim3(kVB), 0, s, n, op->b.as<double>(), op->q.as<double>(),
                       op->r.as<double>(), part + L.t);
    hipLaunchKernelGGL(k_true_fin, dim3(1), dim3(kVB), 0, s, n, part + L.t, vb, op->state.as<CgState>());
    SE_HIP(hipGetLastError());
    return SE_OK;
}

// Solve sign*A x = sign*f on device.  f, x: device vectors (n).
int pcg_fp64(se_op* op, const double* d_f, double* d_x, const se_solve_opts& o, se_solve_info& info) {
    se_mesh* m = op->mesh;
    se_ctx* c = m->ctx;
    hipStream_t s = c->stream;
    const int n = m->n, vb = vec_blocks(n);
    SE_CHECK(op->definite, SE_ERR_UNSUPPORTED, "operator diagonal changes sign: not a definite system");
    int rc = op_ensure_work(op);
    if (rc) return rc;
    const PartLayout L = part_layout(m);
    double* part = op->partials.as<double>();
    if (!op->dinv_s.ptr) {
        rc = op->dinv_s.alloc(sizeof(double) * std::max(n, 1));
        if (rc) return rc;
        // inverse diagonal of sign*A
        hipLaunchKernelGGL(k_scale_copy, dim3(vb), dim3(kBlock), 0, s, n, (double)op->sign, op->dinv.as<double>(),
                           op->dinv_s.as<double>());
    }
    CgState h{};
    h.maxit = std::max(o.maxit, 0);
    h.null_space = o.null_space ? 1 : 0;
    h.n = n;
    h.status = CG_RUNNING;
    CgState* st = op->state.as<CgState>();
    DevBuf d_hist;
    if (o.history) {
        rc = d_hist.alloc(sizeof(double) * (size_t)(h.maxit + 1));
        if (rc) return rc;
    }
    double* dh = o.history ? d_hist.as<double>() : nullptr;
    if (!op->h_state) SE_HIP(hipHostMalloc(&op->h_state, sizeof(CgState), hipHostMallocDefault));
    CgState* hs = (CgState*)op->h_state;
    SE_HIP(hipEventRecord(c->ev0, s));
    *hs = h;  // uploads go through the pinned host copy
    SE_HIP(hipMemcpyAsync(st, hs, sizeof(CgState), hipMemcpyHostToDevice, s));
    if (h.null_space)
        hipLaunchKernelGGL(k_sum, dim3(vb), dim3(kVB), 0, s, n, d_f, (double)op->sign, part + L.s);
    hipLaunchKernelGGL(k_cg_start, dim3(vb), dim3(kVB), 0, s, n, (double)op->sign, d_f, op->dinv_s.as<double>(),
                       op->b.as<double>(), op->x.as<double>(), op->r.as<double>(), part + L.s, vb, part + L.b, st);
    hipLaunchKernelGGL(k_cg_dir, dim3(vb), dim3(kVB), 0, s, (int)DIR_FIRST, n, o.rtol, op->dinv_s.as<double>(),
                       op->r.as<double>(), op->p.as<double>(), part + L.b, vb, dh, st);
    SE_HIP(hipGetLastError());
    SE_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    SE_HIP(hipStreamSynchronize(s));
    const int k = std::max(1, o.check_every);
    int guard = 0;
    while (!hs->stop) {
        // batches overrun harmlessly: every kernel returns once stop is set
        rc = cg_batch(op, k, dh);
        if (rc) return rc;
        SE_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        SE_HIP(hipStreamSynchronize(s));
        if (!hs->stop || hs->status != CG_PENDING) continue;
        // recursive residual below target: confirm with the true residual
        rc = true_residual(op);
        if (rc) return rc;
        SE_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        SE_HIP(hipStreamSynchronize(s));
        if (hs->tt <= hs->tol2) {
            hs->status = 0;
            break;
        }
        if (hs->iter >= h.maxit || ++guard > 1000) {
            hs->status = 1;
            break;
        }
        // continue from the true residual (now in r)
        hs->stop = 0;
        hs->status = CG_RUNNING;
        SE_HIP(hipMemcpyAsync(&st->stop, &hs->stop, 2 * sizeof(int), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_cg_restart, dim3(vb), dim3(kVB), 0, s, n, op->dinv_s.as<double>(), op->r.as<double>(),
                           part + L.b, st);
        hipLaunchKernelGGL(k_cg_dir, dim3(vb), dim3(kVB), 0, s, (int)DIR_RESTART, n, 0.0, op->dinv_s.as<double>(),
                           op->r.as<double>(), op->p.as<double>(), part + L.b, vb, dh, st);
        SE_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        SE_HIP(hipStreamSynchronize(s));
    }
    int status = hs->status;
    const int iters = hs->iter;
    const double bb = hs->bb;
    double tt = hs->tt;
    if (status != 0 && bb > 0.0 && bb - bb == 0.0) {
        rc = true_residual(op);
        if (rc) return rc;
        SE_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        SE_HIP(hipStreamSynchronize(s));
        tt = hs->tt;
    }
    if (d_x != op->x.as<double>())
        SE_HIP(hipMemcpyAsync(d_x, op->x.ptr, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    SE_HIP(hipEventRecord(c->ev1, s));
    SE_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    if (o.history) {
        o.history[0] = 1.0;
        if (iters > 0)
            SE_HIP(hipMemcpy(o.history + 1, d_hist.as<double>() + 1, sizeof(double) * iters, hipMemcpyDeviceToHost));
    }
    info.status = status == CG_RUNNING ? 1 : status;
    info.iters = iters;
    info.outer_iters = 1;
    info.true_rel_residual = bb > 0.0 ? std::sqrt(std::max(tt, 0.0) / bb) : 0.0;
    info.seconds = ms * 1e-3;
    return SE_OK;
}

}  // namespace se

extern "C" int se_helm_pcg_dev(se_op* op, const double* d_f, double* d_x, const se_solve_opts* opts,
                               se_solve_info* info) {
    SE_CHECK(op && d_f && d_x, SE_ERR_ARG, "se_helm_pcg_dev: NULL argument");
    se_solve_opts o;
    se_solve_opts_default(&o);
    if (opts) o = *opts;
    se_solve_info inf{};
    SE_HIP(hipSetDevice(op->mesh->ctx->device));
    int rc;
    SE_CHECK(o.precision == SE_PREC_FP64 || o.precision == SE_PREC_FP32_IR || o.precision == SE_PREC_BF16_IR ||
                 o.precision == SE_PREC_FP64_FP32C || o.precision == SE_PREC_FP64_BF16C,
             SE_ERR_ARG, "se_helm_pcg_dev: bad precision");
    SE_CHECK(o.preconditioner >= SE_PC_DEFAULT && o.preconditioner <= SE_PC_FDM, SE_ERR_ARG,
             "se_helm_pcg_dev: bad preconditioner");
    SE_CHECK((o.precision != SE_PREC_FP64_FP32C && o.precision != SE_PREC_FP64_BF16C) || solve_pc(o) == SE_PC_FDM,
             SE_ERR_UNSUPPORTED, "se_helm_pcg_dev: SE_PREC_FP64_FP32C / FP64_BF16C take the FDM preconditioner");
    if (solve_pc(o) == SE_PC_FDM) rc = pcg_fdm(op, d_f, d_x, o, inf);
    else if (o.precision == SE_PREC_FP64) rc = pcg_fp64(op, d_f, d_x, o, inf);
    else rc = pcg_mixed(op, d_f, d_x, o, inf);
    if (info) *info = inf;
    return rc;
}

extern "C" int se_helm_pcg(se_op* op, const double* f, double* x, const se_solve_opts* opts, se_solve_info* info) {
    SE_CHECK(op && f && x, SE_ERR_ARG, "se_helm_pcg: NULL argument");
    se_ctx* c = op->mesh->ctx;
    SE_HIP(hipSetDevice(c->device));
    int n = op->mesh->n;
    int rc = stage_in(c, op->f, f, n);
    if (rc) return rc;
    rc = se_helm_pcg_dev(op, op->f.as<double>(), op->x.as<double>(), opts, info);
    if (!rc) rc = stage_out(c, x, op->x, n);
    return rc;
}

extern "C" int se_helm_solve_dev(se_op* op, const double* d_rhs_field, double* d_sol_field, const se_solve_opts* opts,
                                 se_solve_info* info) {
    SE_CHECK(op && d_rhs_field && d_sol_field, SE_ERR_ARG, "se_helm_solve_dev: NULL argument");
    int rc = op_ensure_work(op);
    if (!rc) rc = se_helm_rhs_dev(op, d_rhs_field, op->f.as<double>());
    if (!rc) rc = se_helm_pcg_dev(op, op->f.as<double>(), op->x.as<double>(), opts, info);
    if (!rc) rc = se_vec_to_field_dev(op->mesh, op->x.as<double>(), d_sol_field);
    return rc;
}

extern "C" int se_helm_solve(se_op* op, const double* rhs_field, double* sol_field, const se_solve_opts* opts,
                             se_solve_info* info) {
    SE_CHECK(op && rhs_field && sol_field, SE_ERR_ARG, "se_helm_solve: NULL argument");
    se_ctx* c = op->mesh->ctx;
    SE_HIP(hipSetDevice(c->device));
    size_t nV = op->mesh->nV;
    int rc = stage_in(c, op->sfield, rhs_field, nV);
    if (!rc) rc = stage_in(c, op->sfield2, sol_field, nV);
    if (!rc) rc = se_helm_solve_dev(op, op->sfield.as<double>(), op->sfield2.as<double>(), opts, info);
    if (!rc) rc = stage_out(c, sol_field, op->sfield2, nV);
    return rc;
}
