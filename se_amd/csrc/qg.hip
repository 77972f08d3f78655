// qg.hip -- time steps of the quasi-geostrophic models on the device.
//
// Two layers (TwoLayerQGEqn::FirstOrder / SecondOrder,
// src/qg/TwoLayerQGEqn.cpp:174-324, GenVel :335-352, GenPV :326-333,
// SolveQ :396-416, DissipationRHS :419-462), with a = 1 (first order) or 1.5
// (second order):
//   1. u_i = GenVel(psi_i); alpha_i = omega_i + F_i (psi_j - psi_i) + beta y
//      (GenPV), advected semi-Lagrangian: alpha_i^ = alpha_i(departure points)
//      (second order: 2 alpha_i(x1) - alpha_i,prev(x2) / 2 with the previous
//      state's PV and extrapolated velocities)
//   2-4. with dissipation: (lap - (F1+F2)) tau^ = alpha1^ - alpha2^,
//      lap chi^ = alpha1^ + alpha2^ + (F1-F2) tau^ - 2 a beta y,
//      omega_i^ = alpha_i^ - F_i (psi_j^ - psi_i^) - a beta y
//   5. SolveQ: a M alpha_i = the assembled DissipationRHS of alpha_i^ (layer
//      1: relaxation +r F1 of tau^ toward tau_e; layer 2: -r F2 and Ekman drag
//      E2 of omega2^; both: viscosity nu on omega_i^).  M is the diagonal GLL
//      mass matrix, so each unknown is its assembled right-hand side over its
//      mass; Dirichlet nodes keep their GenPV value (UpdateField writes only
//      the unknowns and the periodic copies).  The first-order step solves
//      layer 1 only (TwoLayerQGEqn.cpp:214): layer 2 keeps its GenPV alpha.
//   6. baroclinic tau = psi1 - psi2:  (lap - (F1+F2)) tau = alpha1 - alpha2
//   7. barotropic chi = psi1 + psi2:  lap chi = alpha1 + alpha2 + (F1-F2) tau - 2 beta y
//   8. psi1 = (tau+chi)/2, psi2 = (chi-tau)/2,
//      omega_i = alpha_i - F_i (psi_j - psi_i) - beta y
// One layer (QuasiGeostrophicEqn::FirstOrder / SecondOrder,
// src/qg/QuasiGeostrophicEqn.cpp:83-205, SolveQ :277-301, DisRHS :303-337):
// alpha = omega - F psi + beta y + topo advected; with dissipation
// (lap - F) psi^ = alpha^ - a beta y - a topo (he1, whose Dirichlet lifting is
// zero: the constructor assembles it while psi^ is zero, :50-53; psi^'s wall
// values are zero, or 1.5 psi once the second-order path has run, :86-93) and
// omega^ = alpha^ + F psi^ - a beta y - a topo; SolveQ with linear drag r and
// viscosity nu on omega^; (lap - F) psi = alpha - beta y - topo,
// omega = alpha + F psi - beta y - topo.
// The first inversion (pair) only feeds the dissipation terms: without
// dissipation it is skipped, its values would be multiplied by zero.  A
// model's first step after its state is set is first order, the later ones
// second order from the state before the previous step, as the reference's
// drivers call them.  The inversions use HelmholtzEqn / PoissonEqn semantics
// (rhs field assembled with the mass matrix plus the Dirichlet lifting of the
// current tau / chi / psi, UpdateField after the solve).  The two-layer step
// is split into phases so that a server can run the inversions of many models
// as one batched solve.
#include <algorithm>
#include <cmath>
#include <string>
#include <utility>

#include "qg.h"

using namespace se;

namespace {

int blocks(int n) { return std::max(1, std::min((n + 255) / 256, 4096)); }

// GenPV for both layers, tau and chi
__global__ void k_qg2_pre(int nV, const double* __restrict__ verts, double F1, double F2, double beta,
                          const double* __restrict__ psi1, const double* __restrict__ om1,
                          const double* __restrict__ psi2, const double* __restrict__ om2, double* __restrict__ a1,
                          double* __restrict__ a2, double* __restrict__ tau, double* __restrict__ chi) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
        const double y = verts[2 * v + 1];
        const double p1 = psi1[v], p2 = psi2[v];
        a1[v] = om1[v] + F1 * (p2 - p1) + beta * y;
        a2[v] = om2[v] + F2 * (p1 - p2) + beta * y;
        tau[v] = p1 - p2;
        chi[v] = p1 + p2;
    }
}

// rhs = h1 - h2 (baroclinic) or h1 + h2 + (F1-F2) tau - 2 a beta y (barotropic)
__global__ void k_qg2_rhs(int nV, int barotropic, const double* __restrict__ verts, double F1, double F2,
                          double abeta, const double* __restrict__ h1, const double* __restrict__ h2,
                          const double* __restrict__ tau, double* __restrict__ rhs) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
        if (barotropic) {
            const double y = verts[2 * v + 1];
            rhs[v] = h1[v] + h2[v] + (F1 - F2) * tau[v] - abeta * 2.0 * y;
        } else {
            rhs[v] = h1[v] - h2[v];
        }
    }
}

// the previous state's PV of both layers
__global__ void k_qg2_pv(int nV, const double* __restrict__ verts, double F1, double F2, double beta,
                         const double* __restrict__ psi1, const double* __restrict__ om1,
                         const double* __restrict__ psi2, const double* __restrict__ om2, double* __restrict__ a1,
                         double* __restrict__ a2) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
        const double y = verts[2 * v + 1];
        const double p1 = psi1[v], p2 = psi2[v];
        a1[v] = om1[v] + F1 * (p2 - p1) + beta * y;
        a2[v] = om2[v] + F2 * (p1 - p2) + beta * y;
    }
}

// omega_i^ from the first inversion pair (tau^, chi^) and alpha_i^
__global__ void k_qg2_hat(int nV, const double* __restrict__ verts, double F1, double F2, double abeta,
                          const double* __restrict__ h1, const double* __restrict__ h2,
                          const double* __restrict__ tau, const double* __restrict__ chi,
                          double* __restrict__ om1h, double* __restrict__ om2h) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
        const double y = verts[2 * v + 1];
        const double p1 = 0.5 * (tau[v] + chi[v]);
        const double p2 = 0.5 * (chi[v] - tau[v]);
        om1h[v] = h1[v] - F1 * (p2 - p1) - abeta * y;
        om2h[v] = h2[v] - F2 * (p1 - p2) - abeta * y;
    }
}

// mesh description for the SolveQ gather
struct QGeom {
    int N, nv0, nv1, per0, per1, n;
    const int* v2f;
    const double* w;    // GLL weights
    const double* D;    // D[i (N+1) + j] = l_j'(x_i)
    const double* J00;  // exact reference Jacobian entries (mesh.hip)
    const double* J10;
    const double* J11;
    const double* J01;
};

// the (element, local index) pairs at which field column/row c is an element node
__device__ __forceinline__ int incidences(int c, int N, int nverts, int* e, int* i) {
    int k = 0;
    if (c % N == 0) {
        if (c > 0) e[k] = c / N - 1, i[k] = N, k++;
        if (c < nverts - 1) e[k] = c / N, i[k] = 0, k++;
    } else {
        e[k] = c / N, i[k] = c % N, k++;
    }
    return k;
}

// SolveQ (TwoLayerQGEqn.cpp:396-416 / QuasiGeostrophicEqn.cpp:277-301) under
// the diagonal GLL mass matrix: one thread per unknown gathers the
// DissipationRHS (TwoLayerQGEqn.cpp:437-462; DisRHS is the same with rF = 0,
// Ei = r) of every element node that maps to it -- the node's periodic copies
// add into one unknown, as Vector::Assemble adds them through fieldToVecMap --
// and divides by its assembled a x mass.  At the quadrature points of an
// element (the GLL nodes) the shape functions are Kronecker deltas, so the
// pointwise terms are detJ w (qhat + dt rF (tau - tau_e) - dt Ei omega^) at
// the node itself; the viscous term -dt nu sum int grad N_i . grad omega^
// runs over the points of the element row and column through the node, with
// grad omega^ at each point from grad (GenVel of omega^: the gradient in the
// element Mesh::GlobalToLocal picks at the point, as
// Field::InterpDerivsGlobal is evaluated there) and
// detJ grad N = adj(J) grad_local N from the exact Jacobian tables.
// tau, omh, tau_e, grad may be null (read as zero).
__global__ __launch_bounds__(256) void k_qg_solveq(QGeom g, double a, double dt, const double* __restrict__ qhat,
                                                   double rF, double Ei, double nu, const double* __restrict__ tau,
                                                   const double* __restrict__ omh,
                                                   const double* __restrict__ tau_e,
                                                   const double* __restrict__ grad, double* __restrict__ x) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= g.n) return;
    const int N = g.N, n1 = N + 1, nv0 = g.nv0;
    const int v = g.v2f[u];
    const int gx = v % nv0, gy = v / nv0;
    int cxs[2] = {gx, -1}, cys[2] = {gy, -1};
    if (g.per0 && gx == 0) cxs[1] = nv0 - 1;
    if (g.per1 && gy == 0) cys[1] = g.nv1 - 1;
    double f = 0.0, dg = 0.0;
    for (int sx = 0; sx < 2; sx++) {
        const int cx = cxs[sx];
        if (cx < 0) continue;
        int ex[2], ix[2];
        const int nx = incidences(cx, N, nv0, ex, ix);
        for (int sy = 0; sy < 2; sy++) {
            const int cy = cys[sy];
            if (cy < 0) continue;
            int ey[2], iy[2];
            const int ny = incidences(cy, N, g.nv1, ey, iy);
            const int node = cy * nv0 + cx;
            for (int kx = 0; kx < nx; kx++) {
                const int e0 = ex[kx], i1 = ix[kx];
                for (int ky = 0; ky < ny; ky++) {
                    const int e1 = ey[ky], i2 = iy[ky];
                    const double det = g.J00[e0 * n1 + i1] * g.J11[e1 * n1 + i2] -
                                       g.J01[(e1 * n1 + i2) * n1 + i1] * g.J10[(e0 * n1 + i1) * n1 + i2];
                    const double wgt = det * (g.w[i1] * g.w[i2]);
                    const double c = wgt * dt;
                    dg += wgt * a;
                    double t = wgt * qhat[node];
                    if (tau) t += c * rF * tau[node];
                    if (omh) t -= c * Ei * omh[node];
                    if (tau_e) t -= c * rF * tau_e[node];
                    if (grad) {
                        double s = 0.0;
                        // points (p, i2) of the element row: d/dxi of N
                        const double j11 = g.J11[e1 * n1 + i2];
                        for (int p = 0; p <= N; p++) {
                            const int pn = cy * nv0 + e0 * N + p;
                            const double gxv = grad[2 * pn + 1], gyv = -grad[2 * pn];
                            const double j10 = g.J10[(e0 * n1 + p) * n1 + i2];
                            s += (g.w[p] * g.w[i2]) * g.D[p * n1 + i1] * (j11 * gxv - j10 * gyv);
                        }
                        // points (i1, q) of the element column: d/deta of N
                        const double j00 = g.J00[e0 * n1 + i1];
                        for (int q = 0; q <= N; q++) {
                            const int qn = (e1 * N + q) * nv0 + cx;
                            const double gxv = grad[2 * qn + 1], gyv = -grad[2 * qn];
                            const double j01 = g.J01[(e1 * n1 + q) * n1 + i1];
                            s += (g.w[i1] * g.w[q]) * g.D[q * n1 + i2] * (j00 * gyv - j01 * gxv);
                        }
                        t -= dt * nu * s;
                    }
                    f += t;
                }
            }
        }
    }
    x[u] = f / dg;
}

// the new state from tau, chi and alpha (h); the old state moves to prev
__global__ void k_qg2_post(int nV, const double* __restrict__ verts, double F1, double F2, double beta,
                           const double* __restrict__ tau, const double* __restrict__ chi,
                           const double* __restrict__ h1, const double* __restrict__ h2, double* __restrict__ psi1,
                           double* __restrict__ om1, double* __restrict__ psi2, double* __restrict__ om2,
                           double* __restrict__ pp1, double* __restrict__ po1, double* __restrict__ pp2,
                           double* __restrict__ po2) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
        const double y = verts[2 * v + 1];
        pp1[v] = psi1[v];
        po1[v] = om1[v];
        pp2[v] = psi2[v];
        po2[v] = om2[v];
        const double p1 = 0.5 * (tau[v] + chi[v]);
        const double p2 = 0.5 * (chi[v] - tau[v]);
        psi1[v] = p1;
        psi2[v] = p2;
        om1[v] = h1[v] - F1 * (p2 - p1) - beta * y;
        om2[v] = h2[v] - F2 * (p1 - p2) - beta * y;
    }
}

// one layer: GenPV of the current (a) or previous (ap) state
__global__ void k_qg_pv(int nV, const double* __restrict__ verts, double F, double beta, const double* __restrict__ psi,
                        const double* __restrict__ om, const double* __restrict__ topo, double* __restrict__ a) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x)
        a[v] = om[v] - F * psi[v] + beta * verts[2 * v + 1] + (topo ? topo[v] : 0.0);
}

// one layer: rhs = alpha - a beta y - a topo
__global__ void k_qg_rhs(int nV, const double* __restrict__ verts, double abeta, const double* __restrict__ h,
                         const double* __restrict__ topo, double atopo, double* __restrict__ rhs) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x)
        rhs[v] = h[v] - abeta * verts[2 * v + 1] - (topo ? atopo * topo[v] : 0.0);
}

// one layer: psi^'s field before he1 solves it (zero, or 1.5 psi on the walls
// in the second-order step); with lift, zero everywhere (he1's wall data)
__global__ void k_qg_hat_walls(int nV, const int* __restrict__ f2v, int lift, int second,
                               const double* __restrict__ psi, double* __restrict__ ph) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x)
        ph[v] = (!lift && second && f2v[v] < 0) ? 1.5 * psi[v] : 0.0;
}

// one layer: omega^ = alpha^ + F psi^ - a beta y - a topo
__global__ void k_qg_omh(int nV, const double* __restrict__ verts, double F, double abeta, const double* __restrict__ h,
                         const double* __restrict__ ph, const double* __restrict__ topo, double atopo,
                         double* __restrict__ omh) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x)
        omh[v] = h[v] + F * ph[v] - abeta * verts[2 * v + 1] - (topo ? atopo * topo[v] : 0.0);
}

// one layer: omega = alpha + F psi - beta y - topo (psi already updated); prev := old
__global__ void k_qg_post(int nV, const double* __restrict__ verts, double F, double beta, const double* __restrict__ h,
                          const double* __restrict__ topo, const double* __restrict__ psi_new,
                          double* __restrict__ psi, double* __restrict__ om, double* __restrict__ pp,
                          double* __restrict__ po) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
        pp[v] = psi[v];
        po[v] = om[v];
        const double p = psi_new[v];
        psi[v] = p;
        om[v] = h[v] + F * p - beta * verts[2 * v + 1] - (topo ? topo[v] : 0.0);
    }
}

QGeom qgeom(se_mesh* m) {
    const double* b = m->d_basis.as<double>();
    const int n1 = m->N + 1;
    QGeom g;
    g.N = m->N;
    g.nv0 = m->nVerts[0];
    g.nv1 = m->nVerts[1];
    g.per0 = m->periodic[0];
    g.per1 = m->periodic[1];
    g.n = m->n;
    g.v2f = m->d_v2f.as<int>();
    g.w = b + n1;
    g.D = b + 2 * n1;
    g.J00 = m->d_J00.as<double>();
    g.J10 = m->d_J10.as<double>();
    g.J11 = m->d_J11.as<double>();
    g.J01 = m->d_J01.as<double>();
    return g;
}

// SolveQ of one layer: alpha's unknowns (and their periodic copies) from the
// assembled right-hand side; `alpha` holds the GenPV values, which its
// Dirichlet nodes keep.  grad (the gradient of omh) is needed when nu != 0.
int solveq(se_mesh* m, double a, double dt, const double* qhat, double rF, double Ei, double nu, const double* tau,
           const double* omh, const double* tau_e, const double* grad, double* xq, double* alpha) {
    if (m->n > 0) {
        hipLaunchKernelGGL(k_qg_solveq, dim3((m->n + 255) / 256), dim3(256), 0, m->ctx->stream, qgeom(m), a, dt, qhat,
                           rF, Ei, nu, tau, omh, tau_e, nu != 0.0 ? grad : nullptr, xq);
        SE_HIP(hipGetLastError());
    }
    return se_vec_to_field_dev(m, xq, alpha);
}

bool qg2_dissipative(const se_qg2* q) { return q->r != 0.0 || q->E2 != 0.0 || q->nu != 0.0; }

// the two-layer SolveQ of step 5: layer 1, and layer 2 in the second-order
// step; the new alpha_i end up in h_i
int qg2_solveq(se_qg2* q) {
    se_mesh* m = q->mesh;
    const bool diss = qg2_dissipative(q);
    const double a = q->second ? 1.5 : 1.0;
    const double* tau = diss ? q->tau.as<double>() : nullptr;
    const double* tau_e = diss && q->has_tau_e ? q->tau_e.as<double>() : nullptr;
    int rc = SE_OK;
    if (diss && q->nu != 0.0) {
        q->vel_prev_ok = false;  // vel1, vel2 now hold the viscous term's velocities
        rc = se_genvel_dev(m, q->om1h.as<double>(), q->vel1.as<double>());
        if (!rc && q->second) rc = se_genvel_dev(m, q->om2h.as<double>(), q->vel2.as<double>());
        if (rc) return rc;
    }
    rc = solveq(m, a, q->dt, q->h1.as<double>(), q->r * q->F1, 0.0, q->nu, tau, diss ? q->om1h.as<double>() : nullptr,
                tau_e, q->vel1.as<double>(), q->xq.as<double>(), q->a1.as<double>());
    if (!rc && q->second)
        rc = solveq(m, a, q->dt, q->h2.as<double>(), -q->r * q->F2, q->E2, q->nu, tau,
                    diss ? q->om2h.as<double>() : nullptr, tau_e, q->vel2.as<double>(), q->xq.as<double>(),
                    q->a2.as<double>());
    if (rc) return rc;
    // a_i now hold the new PVs (layer 2 of the first-order step: its GenPV values)
    std::swap(q->a1, q->h1);
    std::swap(q->a2, q->h2);
    return SE_OK;
}

}  // namespace

namespace se {

size_t qg2_bytes(int n, int nV) {
    // device fields (28 nV doubles: state, previous state, PVs, work, omega^,
    // tau_e -- 20 fields -- and four velocity pairs), SolveQ's unknowns and
    // the two operators' work (an estimate for admission)
    return (size_t)nV * 8 * 28 + (size_t)n * (2 * (8 * 10 + 4 * 4) + 8 * 5);
}

int qg2_passes(const se_qg2* q) { return qg2_dissipative(q) ? 4 : 2; }

se_op* qg2_pass_op(se_qg2* q, int k) { return (k & 1) ? q->hp : q->hh; }

int qg2_phase_advect(se_qg2* q, double dt) {
    se_mesh* m = q->mesh;
    const int nV = m->nV;
    hipStream_t s = m->ctx->stream;
    q->second = q->nstep > 0;
    q->dt = dt;
    // the previous velocities: last step's current ones when nothing has
    // overwritten them since (bit for bit GenVel of pp1, pp2)
    const bool reuse = q->second && q->vel_prev_ok;
    if (reuse) {
        std::swap(q->vp1, q->vel1);
        std::swap(q->vp2, q->vel2);
    }
    q->vel_prev_ok = false;
    int rc = se_genvel_dev(m, q->psi1.as<double>(), q->vel1.as<double>());
    if (!rc) rc = se_genvel_dev(m, q->psi2.as<double>(), q->vel2.as<double>());
    if (rc) return rc;
    q->vel_prev_ok = true;
    hipLaunchKernelGGL(k_qg2_pre, dim3(blocks(nV)), dim3(256), 0, s, nV, m->d_verts.as<double>(), q->F1, q->F2,
                       q->beta, q->psi1.as<double>(), q->om1.as<double>(), q->psi2.as<double>(), q->om2.as<double>(),
                       q->a1.as<double>(), q->a2.as<double>(), q->tau.as<double>(), q->chi.as<double>());
    SE_HIP(hipGetLastError());
    if (q->second) {
        if (!reuse) {
            rc = se_genvel_dev(m, q->pp1.as<double>(), q->vp1.as<double>());
            if (!rc) rc = se_genvel_dev(m, q->pp2.as<double>(), q->vp2.as<double>());
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_qg2_pv, dim3(blocks(nV)), dim3(256), 0, s, nV, m->d_verts.as<double>(), q->F1, q->F2,
                           q->beta, q->pp1.as<double>(), q->po1.as<double>(), q->pp2.as<double>(),
                           q->po2.as<double>(), q->ap1.as<double>(), q->ap2.as<double>());
        SE_HIP(hipGetLastError());
        rc = se_sl_advect2_dev(m, q->a1.as<double>(), q->ap1.as<double>(), 1, q->vel1.as<double>(),
                               q->vp1.as<double>(), dt, q->h1.as<double>());
        if (!rc)
            rc = se_sl_advect2_dev(m, q->a2.as<double>(), q->ap2.as<double>(), 1, q->vel2.as<double>(),
                                   q->vp2.as<double>(), dt, q->h2.as<double>());
    } else {
        rc = se_sl_advect_dev(m, q->a1.as<double>(), 1, q->vel1.as<double>(), dt, q->h1.as<double>(), nullptr);
        if (!rc)
            rc = se_sl_advect_dev(m, q->a2.as<double>(), 1, q->vel2.as<double>(), dt, q->h2.as<double>(), nullptr);
    }
    if (rc) return rc;
    // without dissipation SolveQ needs only alpha^: the PVs are final here
    return qg2_dissipative(q) ? SE_OK : qg2_solveq(q);
}

int qg2_phase_rhs(se_qg2* q, int k) {
    se_mesh* m = q->mesh;
    const int nV = m->nV;
    const bool first_pair = qg2_passes(q) == 4 && k < 2;
    const double abeta = (first_pair && q->second ? 1.5 : 1.0) * q->beta;
    se_op* op = qg2_pass_op(q, k);
    // h_i hold alpha_i^ during the first pair, the new alpha_i afterwards
    hipLaunchKernelGGL(k_qg2_rhs, dim3(blocks(nV)), dim3(256), 0, m->ctx->stream, nV, k & 1, m->d_verts.as<double>(),
                       q->F1, q->F2, abeta, q->h1.as<double>(), q->h2.as<double>(), q->tau.as<double>(),
                       q->rhs.as<double>());
    SE_HIP(hipGetLastError());
    double* sol = (k & 1) ? q->chi.as<double>() : q->tau.as<double>();
    int rc = op_ensure_work(op);
    if (!rc) rc = se_helm_set_dirichlet_dev(op, sol);
    if (!rc) rc = se_helm_rhs_dev(op, q->rhs.as<double>(), op->f.as<double>());
    return rc;
}

int qg2_phase_update(se_qg2* q, int k) {
    se_mesh* m = q->mesh;
    se_op* op = qg2_pass_op(q, k);
    int rc = se_vec_to_field_dev(m, op->x.as<double>(), (k & 1) ? q->chi.as<double>() : q->tau.as<double>());
    if (rc || !(qg2_passes(q) == 4 && k == 1)) return rc;
    // end of the first pair: omega^ of both layers, then the dissipative SolveQ
    const int nV = m->nV;
    const double abeta = (q->second ? 1.5 : 1.0) * q->beta;
    hipLaunchKernelGGL(k_qg2_hat, dim3(blocks(nV)), dim3(256), 0, m->ctx->stream, nV, m->d_verts.as<double>(), q->F1,
                       q->F2, abeta, q->h1.as<double>(), q->h2.as<double>(), q->tau.as<double>(),
                       q->chi.as<double>(), q->om1h.as<double>(), q->om2h.as<double>());
    SE_HIP(hipGetLastError());
    return qg2_solveq(q);
}

int qg2_phase_post(se_qg2* q) {
    se_mesh* m = q->mesh;
    const int nV = m->nV;
    hipLaunchKernelGGL(k_qg2_post, dim3(blocks(nV)), dim3(256), 0, m->ctx->stream, nV, m->d_verts.as<double>(), q->F1,
                       q->F2, q->beta, q->tau.as<double>(), q->chi.as<double>(), q->h1.as<double>(),
                       q->h2.as<double>(), q->psi1.as<double>(), q->om1.as<double>(), q->psi2.as<double>(),
                       q->om2.as<double>(), q->pp1.as<double>(), q->po1.as<double>(), q->pp2.as<double>(),
                       q->po2.as<double>());
    SE_HIP(hipGetLastError());
    q->nstep++;
    return SE_OK;
}

}  // namespace se

extern "C" int se_qg2_create(se_mesh* m, double F1, double F2, double beta, int detj_mode, se_qg2** out) {
    SE_CHECK(m && out, SE_ERR_ARG, "se_qg2_create: NULL argument");
    SE_HIP(hipSetDevice(m->ctx->device));
    se_qg2* q = new se_qg2;
    q->mesh = m;
    q->F1 = F1;
    q->F2 = F2;
    q->beta = beta;
    // baroclinic: (lap - (F1+F2)) tau, barotropic: lap chi, with the
    // reference's sign convention (lap = -K)
    int rc = se_helm_create(m, -(F1 + F2), -1.0, detj_mode, &q->hh);
    if (!rc) rc = se_helm_create(m, 0.0, -1.0, detj_mode, &q->hp);
    const size_t nV = (size_t)std::max(m->nV, 1);
    DevBuf* f[] = {&q->psi1, &q->om1, &q->psi2, &q->om2, &q->a1, &q->a2, &q->h1, &q->h2, &q->rhs, &q->tau, &q->chi,
                   &q->pp1, &q->po1, &q->pp2, &q->po2, &q->ap1, &q->ap2, &q->om1h, &q->om2h, &q->tau_e};
    for (DevBuf* b : f)
        if (!rc) rc = b->alloc(sizeof(double) * nV);
    for (DevBuf* b : {&q->vel1, &q->vel2, &q->vp1, &q->vp2})
        if (!rc) rc = b->alloc(sizeof(double) * 2 * nV);
    if (!rc) rc = q->xq.alloc(sizeof(double) * (size_t)std::max(m->n, 1));
    if (rc) {
        se_qg2_destroy(q);
        return rc;
    }
    *out = q;
    return SE_OK;
}

extern "C" int se_qg2_destroy(se_qg2* q) {
    if (!q) return SE_OK;
    (void)hipSetDevice(q->mesh->ctx->device);
    if (q->hh) se_helm_destroy(q->hh);
    if (q->hp) se_helm_destroy(q->hp);
    delete q;
    return SE_OK;
}

extern "C" int se_qg2_set_dissipation(se_qg2* q, double E2, double r, double nu, const double* tau_e) {
    SE_CHECK(q, SE_ERR_ARG, "se_qg2_set_dissipation: NULL argument");
    SE_CHECK(std::isfinite(E2) && std::isfinite(r) && std::isfinite(nu), SE_ERR_ARG,
             "se_qg2_set_dissipation: non-finite coefficient");
    SE_HIP(hipSetDevice(q->mesh->ctx->device));
    // the context's stream is non-blocking: a step still queued on it reads tau_e
    SE_HIP(hipStreamSynchronize(q->mesh->ctx->stream));
    if (tau_e) SE_HIP(hipMemcpy(q->tau_e.ptr, tau_e, sizeof(double) * q->mesh->nV, hipMemcpyHostToDevice));
    q->E2 = E2;
    q->r = r;
    q->nu = nu;
    q->has_tau_e = tau_e != nullptr;
    return SE_OK;
}

extern "C" int se_qg2_state_dev(se_qg2* q, double** psi1, double** om1, double** psi2, double** om2) {
    SE_CHECK(q, SE_ERR_ARG, "se_qg2_state_dev: NULL argument");
    if (psi1) *psi1 = q->psi1.as<double>();
    if (om1) *om1 = q->om1.as<double>();
    if (psi2) *psi2 = q->psi2.as<double>();
    if (om2) *om2 = q->om2.as<double>();
    return SE_OK;
}

extern "C" int se_qg2_set_state(se_qg2* q, const double* psi1, const double* om1, const double* psi2,
                                const double* om2) {
    SE_CHECK(q && psi1 && om1 && psi2 && om2, SE_ERR_ARG, "se_qg2_set_state: NULL argument");
    SE_HIP(hipSetDevice(q->mesh->ctx->device));
    // a step may still be writing the state on the (non-blocking) context stream
    SE_HIP(hipStreamSynchronize(q->mesh->ctx->stream));
    const size_t b = sizeof(double) * q->mesh->nV;
    SE_HIP(hipMemcpy(q->psi1.ptr, psi1, b, hipMemcpyHostToDevice));
    SE_HIP(hipMemcpy(q->om1.ptr, om1, b, hipMemcpyHostToDevice));
    SE_HIP(hipMemcpy(q->psi2.ptr, psi2, b, hipMemcpyHostToDevice));
    SE_HIP(hipMemcpy(q->om2.ptr, om2, b, hipMemcpyHostToDevice));
    q->nstep = 0;  // the next step is first order
    q->vel_prev_ok = false;
    return SE_OK;
}

extern "C" int se_qg2_get_state(se_qg2* q, double* psi1, double* om1, double* psi2, double* om2) {
    SE_CHECK(q && psi1 && om1 && psi2 && om2, SE_ERR_ARG, "se_qg2_get_state: NULL argument");
    SE_HIP(hipSetDevice(q->mesh->ctx->device));
    SE_HIP(hipStreamSynchronize(q->mesh->ctx->stream));
    const size_t b = sizeof(double) * q->mesh->nV;
    SE_HIP(hipMemcpy(psi1, q->psi1.ptr, b, hipMemcpyDeviceToHost));
    SE_HIP(hipMemcpy(om1, q->om1.ptr, b, hipMemcpyDeviceToHost));
    SE_HIP(hipMemcpy(psi2, q->psi2.ptr, b, hipMemcpyDeviceToHost));
    SE_HIP(hipMemcpy(om2, q->om2.ptr, b, hipMemcpyDeviceToHost));
    return SE_OK;
}

extern "C" int se_qg2_timestep(se_qg2* q, double courant, double* dt) {
    SE_CHECK(q && dt, SE_ERR_ARG, "se_qg2_timestep: NULL argument");
    se_mesh* m = q->mesh;
    SE_HIP(hipSetDevice(m->ctx->device));
    double d1 = 0.0, d2 = 0.0;
    q->vel_prev_ok = false;  // the velocities of the current state
    int rc = se_genvel_dev(m, q->psi1.as<double>(), q->vel1.as<double>());
    if (!rc) rc = se_genvel_dev(m, q->psi2.as<double>(), q->vel2.as<double>());
    if (!rc) rc = se_calc_timestep_dev(m, q->vel1.as<double>(), courant, &d1);
    if (!rc) rc = se_calc_timestep_dev(m, q->vel2.as<double>(), courant, &d2);
    if (rc) return rc;
    *dt = std::min(d1, d2);
    return SE_OK;
}

extern "C" int se_qg2_step(se_qg2* q, double dt, const se_solve_opts* opts, se_solve_info* infos) {
    SE_CHECK(q, SE_ERR_ARG, "se_qg2_step: NULL argument");
    SE_HIP(hipSetDevice(q->mesh->ctx->device));
    se_solve_info inf[2] = {};
    int rc = qg2_phase_advect(q, dt);
    const int passes = qg2_passes(q);
    if (!rc) rc = qg2_phase_rhs(q, 0);
    for (int k = 0; k < passes && !rc; k++) {
        se_op* op = qg2_pass_op(q, k);
        se_solve_info one = {};
        rc = se_helm_pcg_dev(op, op->f.as<double>(), op->x.as<double>(), opts, &one);
        if (!rc && one.status != SE_SOLVE_CONVERGED)
            rc = fail(SE_ERR_STATE, "se_qg2_step: an inversion did not converge (status " +
                                        std::to_string(one.status) + ")");
        // the operator's entry describes its last solve, with the iterations of both passes
        const int before = inf[k & 1].iters, before_outer = inf[k & 1].outer_iters;
        inf[k & 1] = one;
        inf[k & 1].iters += before;
        inf[k & 1].outer_iters += before_outer;
        if (!rc) rc = qg2_phase_update(q, k);
        if (!rc && k + 1 < passes) rc = qg2_phase_rhs(q, k + 1);
    }
    if (!rc) rc = qg2_phase_post(q);
    if (infos) infos[0] = inf[0], infos[1] = inf[1];
    return rc;
}

// ---------------------------------------------------------------------------
// one-layer QG channel model (QuasiGeostrophicEqn, src/qg/QuasiGeostrophicEqn.cpp:28-205)
// ---------------------------------------------------------------------------
extern "C" int se_qg_create(se_mesh* m, double F, double beta, int detj_mode, se_qg** out) {
    SE_CHECK(m && out, SE_ERR_ARG, "se_qg_create: NULL argument");
    SE_HIP(hipSetDevice(m->ctx->device));
    se_qg* q = new se_qg;
    q->mesh = m;
    q->F = F;
    q->beta = beta;
    // (lap - F) psi with the reference's sign convention (lap = -K):
    // HelmholtzEqn(phi, rhs, -F, -1.0, 1.0, 0)
    int rc = se_helm_create(m, -F, -1.0, detj_mode, &q->h);
    const size_t nV = (size_t)std::max(m->nV, 1);
    for (DevBuf* b : {&q->psi, &q->om, &q->a, &q->hsl, &q->rhs, &q->psin, &q->pp, &q->po, &q->ap, &q->omh, &q->topo})
        if (!rc) rc = b->alloc(sizeof(double) * nV);
    for (DevBuf* b : {&q->vel, &q->vp})
        if (!rc) rc = b->alloc(sizeof(double) * 2 * nV);
    if (!rc) rc = q->xq.alloc(sizeof(double) * (size_t)std::max(m->n, 1));
    if (rc) {
        se_qg_destroy(q);
        return rc;
    }
    *out = q;
    return SE_OK;
}

extern "C" int se_qg_destroy(se_qg* q) {
    if (!q) return SE_OK;
    (void)hipSetDevice(q->mesh->ctx->device);
    if (q->h) se_helm_destroy(q->h);
    delete q;
    return SE_OK;
}

extern "C" int se_qg_set_dissipation(se_qg* q, double nu, double r) {
    SE_CHECK(q, SE_ERR_ARG, "se_qg_set_dissipation: NULL argument");
    SE_CHECK(std::isfinite(nu) && std::isfinite(r), SE_ERR_ARG, "se_qg_set_dissipation: non-finite coefficient");
    q->nu = nu;
    q->r = r;
    return SE_OK;
}

extern "C" int se_qg_set_topography(se_qg* q, const double* topo) {
    SE_CHECK(q, SE_ERR_ARG, "se_qg_set_topography: NULL argument");
    SE_HIP(hipSetDevice(q->mesh->ctx->device));
    SE_HIP(hipStreamSynchronize(q->mesh->ctx->stream));  // a queued step reads the topography
    if (topo) SE_HIP(hipMemcpy(q->topo.ptr, topo, sizeof(double) * q->mesh->nV, hipMemcpyHostToDevice));
    q->has_topo = topo != nullptr;
    return SE_OK;
}

extern "C" int se_qg_set_state(se_qg* q, const double* psi, const double* omega) {
    SE_CHECK(q && psi && omega, SE_ERR_ARG, "se_qg_set_state: NULL argument");
    SE_HIP(hipSetDevice(q->mesh->ctx->device));
    // a step may still be writing the state on the (non-blocking) context stream
    SE_HIP(hipStreamSynchronize(q->mesh->ctx->stream));
    const size_t b = sizeof(double) * q->mesh->nV;
    SE_HIP(hipMemcpy(q->psi.ptr, psi, b, hipMemcpyHostToDevice));
    SE_HIP(hipMemcpy(q->om.ptr, omega, b, hipMemcpyHostToDevice));
    q->nstep = 0;
    return SE_OK;
}

extern "C" int se_qg_get_state(se_qg* q, double* psi, double* omega) {
    SE_CHECK(q && psi && omega, SE_ERR_ARG, "se_qg_get_state: NULL argument");
    SE_HIP(hipSetDevice(q->mesh->ctx->device));
    SE_HIP(hipStreamSynchronize(q->mesh->ctx->stream));
    const size_t b = sizeof(double) * q->mesh->nV;
    SE_HIP(hipMemcpy(psi, q->psi.ptr, b, hipMemcpyDeviceToHost));
    SE_HIP(hipMemcpy(omega, q->om.ptr, b, hipMemcpyDeviceToHost));
    return SE_OK;
}

extern "C" int se_qg_timestep(se_qg* q, double courant, double* dt) {
    SE_CHECK(q && dt, SE_ERR_ARG, "se_qg_timestep: NULL argument");
    se_mesh* m = q->mesh;
    SE_HIP(hipSetDevice(m->ctx->device));
    int rc = se_genvel_dev(m, q->psi.as<double>(), q->vel.as<double>());
    if (!rc) rc = se_calc_timestep_dev(m, q->vel.as<double>(), courant, dt);
    return rc;
}

namespace {

// one inversion of the one-layer model: (lap - F) sol = rhs with the wall data
// of `walls` (its Dirichlet lifting), the unknowns written into sol
int qg_invert(se_qg* q, const double* walls, double* sol, const se_solve_opts* opts, se_solve_info* inf) {
    se_op* op = q->h;
    int rc = op_ensure_work(op);
    if (!rc) rc = se_helm_set_dirichlet_dev(op, walls);
    if (!rc) rc = se_helm_rhs_dev(op, q->rhs.as<double>(), op->f.as<double>());
    if (!rc) rc = se_helm_pcg_dev(op, op->f.as<double>(), op->x.as<double>(), opts, inf);
    if (!rc && inf->status != SE_SOLVE_CONVERGED)
        rc = fail(SE_ERR_STATE, "se_qg_step: an inversion did not converge (status " + std::to_string(inf->status) + ")");
    if (!rc) rc = se_vec_to_field_dev(q->mesh, op->x.as<double>(), sol);
    return rc;
}

}  // namespace

extern "C" int se_qg_step(se_qg* q, double dt, const se_solve_opts* opts, se_solve_info* info) {
    SE_CHECK(q, SE_ERR_ARG, "se_qg_step: NULL argument");
    se_mesh* m = q->mesh;
    SE_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    const int nV = m->nV;
    const double* verts = m->d_verts.as<double>();
    const int second = q->nstep > 0;
    const double a = second ? 1.5 : 1.0;
    const bool diss = q->nu != 0.0 || q->r != 0.0;
    const double* topo = q->has_topo ? q->topo.as<double>() : nullptr;
    // GenVel, GenPV, SL advection of the PV (second order from the previous state)
    int rc = se_genvel_dev(m, q->psi.as<double>(), q->vel.as<double>());
    if (rc) return rc;
    hipLaunchKernelGGL(k_qg_pv, dim3(blocks(nV)), dim3(256), 0, s, nV, verts, q->F, q->beta, q->psi.as<double>(),
                       q->om.as<double>(), topo, q->a.as<double>());
    SE_HIP(hipGetLastError());
    if (second) {
        rc = se_genvel_dev(m, q->pp.as<double>(), q->vp.as<double>());
        if (rc) return rc;
        hipLaunchKernelGGL(k_qg_pv, dim3(blocks(nV)), dim3(256), 0, s, nV, verts, q->F, q->beta, q->pp.as<double>(),
                           q->po.as<double>(), topo, q->ap.as<double>());
        SE_HIP(hipGetLastError());
        rc = se_sl_advect2_dev(m, q->a.as<double>(), q->ap.as<double>(), 1, q->vel.as<double>(), q->vp.as<double>(), dt,
                               q->hsl.as<double>());
    } else {
        rc = se_sl_advect_dev(m, q->a.as<double>(), 1, q->vel.as<double>(), dt, q->hsl.as<double>(), nullptr);
    }
    if (rc) return rc;
    se_solve_info inf = {}, inf1 = {};
    if (diss) {
        // he1: psi^ with zero lifting, then omega^ and its gradient
        hipLaunchKernelGGL(k_qg_hat_walls, dim3(blocks(nV)), dim3(256), 0, s, nV, m->d_f2v.as<int>(), 1, second,
                           q->psi.as<double>(), q->psin.as<double>());
        SE_HIP(hipGetLastError());
        rc = op_ensure_work(q->h);
        if (!rc) rc = se_helm_set_dirichlet_dev(q->h, q->psin.as<double>());
        if (rc) return rc;
        hipLaunchKernelGGL(k_qg_hat_walls, dim3(blocks(nV)), dim3(256), 0, s, nV, m->d_f2v.as<int>(), 0, second,
                           q->psi.as<double>(), q->psin.as<double>());
        hipLaunchKernelGGL(k_qg_rhs, dim3(blocks(nV)), dim3(256), 0, s, nV, verts, a * q->beta, q->hsl.as<double>(),
                           topo, a, q->rhs.as<double>());
        SE_HIP(hipGetLastError());
        se_op* op = q->h;
        rc = se_helm_rhs_dev(op, q->rhs.as<double>(), op->f.as<double>());
        if (!rc) rc = se_helm_pcg_dev(op, op->f.as<double>(), op->x.as<double>(), opts, &inf1);
        if (!rc && inf1.status != SE_SOLVE_CONVERGED)
            rc = fail(SE_ERR_STATE,
                      "se_qg_step: the first inversion did not converge (status " + std::to_string(inf1.status) + ")");
        if (!rc) rc = se_vec_to_field_dev(m, op->x.as<double>(), q->psin.as<double>());
        if (rc) return rc;
        hipLaunchKernelGGL(k_qg_omh, dim3(blocks(nV)), dim3(256), 0, s, nV, verts, q->F, a * q->beta,
                           q->hsl.as<double>(), q->psin.as<double>(), topo, a, q->omh.as<double>());
        SE_HIP(hipGetLastError());
        if (q->nu != 0.0) rc = se_genvel_dev(m, q->omh.as<double>(), q->vel.as<double>());
        if (rc) return rc;
    }
    // SolveQ: q->a (GenPV) := the new alpha
    rc = solveq(m, a, dt, q->hsl.as<double>(), 0.0, q->r, q->nu, nullptr, diss ? q->omh.as<double>() : nullptr,
                nullptr, q->vel.as<double>(), q->xq.as<double>(), q->a.as<double>());
    if (rc) return rc;
    // (lap - F) psi = alpha - beta y - topo with the wall data of psi
    hipLaunchKernelGGL(k_qg_rhs, dim3(blocks(nV)), dim3(256), 0, s, nV, verts, q->beta, q->a.as<double>(), topo, 1.0,
                       q->rhs.as<double>());
    SE_HIP(hipGetLastError());
    SE_HIP(hipMemcpyAsync(q->psin.ptr, q->psi.ptr, sizeof(double) * nV, hipMemcpyDeviceToDevice, s));
    rc = qg_invert(q, q->psi.as<double>(), q->psin.as<double>(), opts, &inf);
    if (!rc) {
        hipLaunchKernelGGL(k_qg_post, dim3(blocks(nV)), dim3(256), 0, s, nV, verts, q->F, q->beta, q->a.as<double>(),
                           topo, q->psin.as<double>(), q->psi.as<double>(), q->om.as<double>(), q->pp.as<double>(),
                           q->po.as<double>());
        SE_HIP(hipGetLastError());
        q->nstep++;
    }
    inf.iters += inf1.iters;
    if (info) *info = inf;
    return rc;
}
