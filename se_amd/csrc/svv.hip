// svv.hip -- the spectral vanishing viscosity term of HelmholtzEqn
// (HelmholtzEqn(sol, rhs, c, nu, dt, svvCutoff), src/linsys/HelmholtzEqn.cpp:21-81;
// SVV::AssembleElement, src/base/SVV.cpp:120-176), FP64.
//
// SVV's element matrix c (Lxx + Lyy), c = dt, is built from the nodal values
// B of the hierarchical modes, the modal filter Q (modes above the cutoff mN
// in both directions, weight sqrt(exp(-(N-k)^2/(mN-k)^2)) per direction), the
// reference element's x-derivative and the quadrature weights; Lyy is built
// from the x terms like Lxx (SVV.cpp:165).  On a tensor-product element all of
// it separates: B = Phi (x) Phi, Q = diag(q) (x) diag(q), so the filter is
// F (x) F with F = Phi diag(q) Phi^-1, and
//     c (Lxx + Lyy) = 2 c detJ (F^T W F) (x) ((F D)^T W (F D))      (y (x) x)
// with W the 1-D GLL weights and D the GLL derivative matrix.  The two
// (N+1) x (N+1) factors Mf (y) and Kf (x) are computed once on the host; the
// device gathers, for each unknown, the element rows of every element that
// holds one of its node copies (periodic copies add into one unknown, as the
// assembly does), with the element's nodal values taken through the numbering:
// unknowns from the vector, Dirichlet nodes from the field (lifting) or zero
// (apply), or the unit vector of the unknown itself (diagonal).
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "se_op.h"

using namespace se;

namespace {

enum { SVV_APPLY = 0, SVV_LIFT = 1, SVV_DIAG = 2 };

struct SvvGeom {
    int N, nv0, nv1, per0, per1, n;
    const int* v2f;
    const int* f2v;
    const double* Mf;  // (N+1)^2, row-major: the y factor
    const double* Kf;  // (N+1)^2: the x factor
};

// the (element, local index) pairs at which field column/row c is an element node
__device__ __forceinline__ int svv_incidences(int c, int N, int nverts, int* e, int* i) {
    int k = 0;
    if (c % N == 0) {
        if (c > 0) e[k] = c / N - 1, i[k] = N, k++;
        if (c < nverts - 1) e[k] = c / N, i[k] = 0, k++;
    } else {
        e[k] = c / N, i[k] = c % N, k++;
    }
    return k;
}

// mode APPLY: y[u] += coef (S x)[u]; LIFT: y[u] -= coef (S x_D)[u] with x_D the
// field's Dirichlet values; DIAG: y[u] = coef S[u][u]
// (gate: when given, nothing to do unless that CG awaits its true-residual check)
__global__ __launch_bounds__(256) void k_svv(SvvGeom g, double coef, int mode, const double* __restrict__ x,
                                             const double* __restrict__ field, double* __restrict__ y,
                                             const CgState* __restrict__ gate) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= g.n) return;
    if (gate && gate->status != CG_PENDING) return;
    const int N = g.N, n1 = N + 1, nv0 = g.nv0;
    const int v = g.v2f[u];
    const int gx = v % nv0, gy = v / nv0;
    int cxs[2] = {gx, -1}, cys[2] = {gy, -1};
    if (g.per0 && gx == 0) cxs[1] = nv0 - 1;
    if (g.per1 && gy == 0) cys[1] = g.nv1 - 1;
    double acc = 0.0;
    for (int sx = 0; sx < 2; sx++) {
        if (cxs[sx] < 0) continue;
        int ex[2], ix[2];
        const int nx = svv_incidences(cxs[sx], N, nv0, ex, ix);
        for (int sy = 0; sy < 2; sy++) {
            if (cys[sy] < 0) continue;
            int ey[2], iy[2];
            const int ny = svv_incidences(cys[sy], N, g.nv1, ey, iy);
            for (int kx = 0; kx < nx; kx++)
                for (int ky = 0; ky < ny; ky++) {
                    const int e0 = ex[kx], i1 = ix[kx], e1 = ey[ky], i2 = iy[ky];
                    double s = 0.0;
                    for (int j2 = 0; j2 <= N; j2++) {
                        const int row = (e1 * N + j2) * nv0 + e0 * N;
                        double t = 0.0;
                        for (int j1 = 0; j1 <= N; j1++) {
                            const int node = row + j1;
                            const int k = g.f2v[node];
                            double val;
                            if (mode == SVV_APPLY) val = k >= 0 ? x[k] : 0.0;
                            else if (mode == SVV_LIFT) val = k >= 0 ? 0.0 : field[node];
                            else val = k == u ? 1.0 : 0.0;
                            t += g.Kf[i1 * n1 + j1] * val;
                        }
                        s += g.Mf[i2 * n1 + j2] * t;
                    }
                    acc += s;
                }
        }
    }
    if (mode == SVV_APPLY) y[u] += coef * acc;
    else if (mode == SVV_LIFT) y[u] -= coef * acc;
    else y[u] = coef * acc;
}

// dinv := 1 / (1/dinv + d): the SVV diagonal added to the operator's
__global__ void k_svv_add_diag(int n, const double* __restrict__ d, double* __restrict__ dinv) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dinv[i] = 1.0 / (1.0 / dinv[i] + d[i]);
}

// partials[b] = sum over the b-th stripe of x.y (np blocks, fixed order)
__global__ __launch_bounds__(256) void k_svv_dot(int n, const double* __restrict__ x, const double* __restrict__ y,
                                                 double* __restrict__ partials) {
    __shared__ double s_red[4];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += x[i] * y[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
}

SvvGeom svv_geom(const se_op* op) {
    const se_mesh* m = op->mesh;
    const int n1 = m->N + 1;
    SvvGeom g;
    g.N = m->N;
    g.nv0 = m->nVerts[0];
    g.nv1 = m->nVerts[1];
    g.per0 = m->periodic[0];
    g.per1 = m->periodic[1];
    g.n = m->n;
    g.v2f = m->d_v2f.as<int>();
    g.f2v = m->d_f2v.as<int>();
    g.Mf = op->svv_mk.as<double>();
    g.Kf = op->svv_mk.as<double>() + n1 * n1;
    return g;
}

double svv_coef(const se_op* op, double scale) {
    // 2 dt detJ: the element's constant, Lxx twice, the (uniform) Jacobian
    const se_mesh* m = op->mesh;
    return scale * 2.0 * op->svv_dt * (0.5 * m->dx[0]) * (0.5 * m->dx[1]);
}

// inverse of a small dense matrix (Gauss-Jordan, partial pivoting)
bool invert(std::vector<double> A, int n, std::vector<double>& R) {
    R.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) R[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)p * n + c])) p = r;
        if (A[(size_t)p * n + c] == 0.0) return false;
        for (int k = 0; k < n; k++) {
            std::swap(A[(size_t)p * n + k], A[(size_t)c * n + k]);
            std::swap(R[(size_t)p * n + k], R[(size_t)c * n + k]);
        }
        const double inv = 1.0 / A[(size_t)c * n + c];
        for (int k = 0; k < n; k++) {
            A[(size_t)c * n + k] *= inv;
            R[(size_t)c * n + k] *= inv;
        }
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = A[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; k++) {
                A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
                R[(size_t)r * n + k] -= f * R[(size_t)c * n + k];
            }
        }
    }
    return true;
}

// The 1-D factors Mf = F^T W F and Kf = (F D)^T W (F D) of the filter
// F = Phi diag(q) Phi^-1 (Phi[i][k] = mode k at GLL point i).
bool svv_factors(const GLL& gll, int mN, std::vector<double>& mk) {
    const int N = gll.N, n1 = N + 1;
    std::vector<double> Phi((size_t)n1 * n1), Pinv, F((size_t)n1 * n1, 0.0), G((size_t)n1 * n1, 0.0);
    for (int i = 0; i < n1; i++) {
        const double x = gll.x[i];
        // Jacobi P^(1,1)_j(x) by the three-term recurrence
        std::vector<double> P(std::max(n1 - 1, 1));
        for (int j = 0; j + 2 <= N; j++) {
            if (j == 0) P[0] = 1.0;
            else if (j == 1) P[1] = 2.0 * x;
            else {
                const double a = 1.0;
                const double c1 = 2.0 * j * (j + 2.0 * a) * (2.0 * j + 2.0 * a - 2.0);
                const double c2 = (2.0 * j + 2.0 * a - 1.0) * (2.0 * j + 2.0 * a) * (2.0 * j + 2.0 * a - 2.0);
                const double c3 = 2.0 * (j + a - 1.0) * (j + a - 1.0) * (2.0 * j + 2.0 * a);
                P[j] = (c2 * x * P[j - 1] - c3 * P[j - 2]) / c1;
            }
        }
        Phi[(size_t)i * n1 + 0] = 0.5 * (1.0 + x);
        Phi[(size_t)i * n1 + 1] = 0.5 * (1.0 - x);
        for (int k = 2; k <= N; k++) Phi[(size_t)i * n1 + k] = 0.25 * (1.0 - x * x) * P[k - 2];
    }
    if (!invert(Phi, n1, Pinv)) return false;
    std::vector<double> q(n1, 0.0);
    for (int k = mN + 1; k <= N; k++) {
        const double r = (double)(N - k) / (double)(mN - k);
        q[k] = std::sqrt(std::exp(-r * r));
    }
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n1; j++) {
            double s = 0.0;
            for (int k = 0; k < n1; k++) s += Phi[(size_t)i * n1 + k] * q[k] * Pinv[(size_t)k * n1 + j];
            F[(size_t)i * n1 + j] = s;
        }
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n1; j++) {
            double s = 0.0;
            for (int k = 0; k < n1; k++) s += F[(size_t)i * n1 + k] * gll.D[(size_t)k * n1 + j];
            G[(size_t)i * n1 + j] = s;
        }
    mk.assign(2 * (size_t)n1 * n1, 0.0);
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n1; j++) {
            double m = 0.0, k = 0.0;
            for (int p = 0; p < n1; p++) {
                m += F[(size_t)p * n1 + i] * gll.w[p] * F[(size_t)p * n1 + j];
                k += G[(size_t)p * n1 + i] * gll.w[p] * G[(size_t)p * n1 + j];
            }
            mk[(size_t)i * n1 + j] = m;
            mk[(size_t)n1 * n1 + (size_t)i * n1 + j] = k;
        }
    return true;
}

int blocks_for_n(int n) { return std::max(1, (n + 255) / 256); }

}  // namespace

namespace se {

int svv_apply(se_op* op, const double* d_x, double* d_y, double scale, const CgState* gate) {
    if (op->mesh->n == 0) return SE_OK;
    hipLaunchKernelGGL(k_svv, dim3(blocks_for_n(op->mesh->n)), dim3(256), 0, op->mesh->ctx->stream, svv_geom(op),
                       svv_coef(op, scale), (int)SVV_APPLY, d_x, (const double*)nullptr, d_y, gate);
    SE_HIP(hipGetLastError());
    return SE_OK;
}

int svv_dot(se_op* op, const double* d_x, const double* d_y, double* partials, int np) {
    hipLaunchKernelGGL(k_svv_dot, dim3(std::max(np, 1)), dim3(256), 0, op->mesh->ctx->stream, op->mesh->n, d_x, d_y,
                       partials);
    SE_HIP(hipGetLastError());
    return SE_OK;
}

int svv_lift(se_op* op, const double* d_field, double* d_lift) {
    if (op->mesh->n == 0) return SE_OK;
    hipLaunchKernelGGL(k_svv, dim3(blocks_for_n(op->mesh->n)), dim3(256), 0, op->mesh->ctx->stream, svv_geom(op),
                       svv_coef(op, 1.0), (int)SVV_LIFT, (const double*)nullptr, d_field, d_lift,
                       (const CgState*)nullptr);
    SE_HIP(hipGetLastError());
    return SE_OK;
}

}  // namespace se

extern "C" int se_helm_set_svv(se_op* op, double dt, int cutoff) {
    SE_CHECK(op, SE_ERR_ARG, "se_helm_set_svv: NULL argument");
    se_mesh* m = op->mesh;
    SE_CHECK(op->svv_mN == 0, SE_ERR_STATE, "se_helm_set_svv: the operator already has an SVV term");
    SE_CHECK(cutoff >= 1 && cutoff < m->N, SE_ERR_ARG, "se_helm_set_svv: the cutoff must be in 1 .. N-1");
    SE_CHECK(std::isfinite(dt), SE_ERR_ARG, "se_helm_set_svv: non-finite dt");
    SE_CHECK(m->nEls[0] >= 1 && m->nEls[1] >= 1, SE_ERR_ARG, "se_helm_set_svv: empty mesh");
    SE_HIP(hipSetDevice(m->ctx->device));
    std::vector<double> mk;
    SE_CHECK(svv_factors(m->gll, cutoff, mk), SE_ERR_STATE, "se_helm_set_svv: singular modal basis");
    int rc = op->svv_mk.alloc(sizeof(double) * mk.size());
    if (rc) return rc;
    SE_HIP(hipMemcpy(op->svv_mk.ptr, mk.data(), sizeof(double) * mk.size(), hipMemcpyHostToDevice));
    op->svv_dt = dt;
    op->svv_mN = cutoff;
    // the Jacobi diagonal gains the SVV diagonal; the sign-scaled copy and the
    // mixed-precision tables are rebuilt (the latter refuse SVV)
    const int n = m->n;
    hipStream_t s = m->ctx->stream;
    if (n > 0) {
        DevBuf d;
        rc = d.alloc(sizeof(double) * n);
        if (rc) return rc;
        hipLaunchKernelGGL(k_svv, dim3(blocks_for_n(n)), dim3(256), 0, s, svv_geom(op), svv_coef(op, 1.0),
                           (int)SVV_DIAG, (const double*)nullptr, (const double*)nullptr, d.as<double>(),
                           (const CgState*)nullptr);
        hipLaunchKernelGGL(k_svv_add_diag, dim3(blocks_for_n(n)), dim3(256), 0, s, n, d.as<double>(),
                           op->dinv.as<double>());
        SE_HIP(hipGetLastError());
        SE_HIP(hipStreamSynchronize(s));
        std::vector<double> h(n);
        SE_HIP(hipMemcpy(h.data(), op->dinv.ptr, sizeof(double) * n, hipMemcpyDeviceToHost));
        int pos = 0, neg = 0;
        for (double v : h) pos += v > 0, neg += v < 0;
        op->sign = (neg > 0 && pos == 0) ? -1 : 1;
        op->definite = (pos == n) || (neg == n);
    }
    op->dinv_s.release();
    if (op->graph_exec) (void)hipGraphExecDestroy(op->graph_exec);
    op->graph_exec = nullptr;
    op->graph_iters = 0;
    return SE_OK;
}
