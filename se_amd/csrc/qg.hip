// qg.hip -- time steps of the quasi-geostrophic models on the device, with
// the dissipation off (nu = r = 0) and no topography.
//
// Two layers (TwoLayerQGEqn::FirstOrder / SecondOrder,
// src/qg/TwoLayerQGEqn.cpp:174-333, GenVel :335-352, GenPV :326-333):
//   1. u_i = GenVel(psi_i); alpha_i = omega_i + F_i (psi_j - psi_i) + beta y
//      (GenPV), advected semi-Lagrangian: alpha_i^ = alpha_i(departure points)
//      (second order: 2 alpha_i(x1) - alpha_i,prev(x2) / 2 with the previous
//      state's PV and extrapolated velocities)
//   2. SolveQ under the diagonal GLL mass: alpha_i = alpha_i^ (first order)
//      or alpha_i^ / 1.5 (second order) on the unknowns, the GenPV value on
//      Dirichlet nodes (UpdateField writes only the unknowns).  The
//      reference's first-order step calls SolveQ for layer 1 only
//      (TwoLayerQGEqn.cpp:214), so there layer 2 keeps its GenPV alpha.
//   3. baroclinic tau = psi1 - psi2:  (lap - (F1+F2)) tau = alpha1 - alpha2
//   4. barotropic chi = psi1 + psi2:  lap chi = alpha1 + alpha2 + (F1-F2) tau - 2 beta y
//   5. psi1 = (tau+chi)/2, psi2 = (chi-tau)/2,
//      omega_i = alpha_i - F_i (psi_j - psi_i) - beta y
// One layer (QuasiGeostrophicEqn::FirstOrder / SecondOrder,
// src/qg/QuasiGeostrophicEqn.cpp:83-205): alpha = omega - F psi + beta y,
// advected and SolveQ'd as above, (lap - F) psi = alpha - beta y,
// omega = alpha + F psi - beta y.
// The first inversions of the reference's steps only feed the dissipation
// terms and are not repeated.  A model's first step after its state is set is
// first order, the later ones second order from the state before the previous
// step, as the reference's drivers call them.  The inversions use
// HelmholtzEqn / PoissonEqn semantics (rhs field assembled with the mass
// matrix plus the Dirichlet lifting of the current tau / chi / psi,
// UpdateField after the solve).  The two-layer step is split into phases so
// that a server can run the inversions of many models as one batched solve.
#include <algorithm>
#include <cmath>
#include <string>

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

// rhs = h1 - h2 (baroclinic) or h1 + h2 + (F1-F2) tau - 2 beta y (barotropic)
__global__ void k_qg2_rhs(int nV, int barotropic, const double* __restrict__ verts, double F1, double F2,
                          double beta, const double* __restrict__ h1, const double* __restrict__ h2,
                          const double* __restrict__ tau, double* __restrict__ rhs) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
        if (barotropic) {
            const double y = verts[2 * v + 1];
            rhs[v] = h1[v] + h2[v] + (F1 - F2) * tau[v] - 2.0 * beta * y;
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

// SolveQ + UpdateField: h_i := the new alpha_i (the unknowns take the advected
// value, divided by 1.5 in the second-order step; Dirichlet nodes keep the
// GenPV value a_i).  n2 = 0: layer 2 is not solved (the first-order step).
__global__ void k_qg_alpha(int nV, const int* __restrict__ f2v, int second, int n2, const double* __restrict__ a1,
                           double* __restrict__ h1, const double* __restrict__ a2, double* __restrict__ h2) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
        const bool dir = f2v[v] < 0;
        h1[v] = dir ? a1[v] : (second ? h1[v] / 1.5 : h1[v]);
        if (h2) h2[v] = (dir || !n2) ? a2[v] : (second ? h2[v] / 1.5 : h2[v]);
    }
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
                        const double* __restrict__ om, double* __restrict__ a) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x)
        a[v] = om[v] - F * psi[v] + beta * verts[2 * v + 1];
}

// one layer: rhs = alpha - beta y
__global__ void k_qg_rhs(int nV, const double* __restrict__ verts, double beta, const double* __restrict__ h,
                         double* __restrict__ rhs) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x)
        rhs[v] = h[v] - beta * verts[2 * v + 1];
}

// one layer: omega = alpha + F psi - beta y (psi already updated); prev := old
__global__ void k_qg_post(int nV, const double* __restrict__ verts, double F, double beta, const double* __restrict__ h,
                          const double* __restrict__ psi_new, double* __restrict__ psi, double* __restrict__ om,
                          double* __restrict__ pp, double* __restrict__ po) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
        pp[v] = psi[v];
        po[v] = om[v];
        const double p = psi_new[v];
        psi[v] = p;
        om[v] = h[v] + F * p - beta * verts[2 * v + 1];
    }
}

}  // namespace

namespace se {

size_t qg2_bytes(int n, int nV) {
    // device fields (26 nV doubles: state, previous state, PVs, velocities,
    // work) and the two operators' work (an estimate for admission)
    return (size_t)nV * 8 * 26 + (size_t)n * (2 * (8 * 10 + 4 * 4) + 8 * 4);
}

int qg2_phase_advect(se_qg2* q, double dt) {
    se_mesh* m = q->mesh;
    const int nV = m->nV;
    hipStream_t s = m->ctx->stream;
    int rc = se_genvel_dev(m, q->psi1.as<double>(), q->vel1.as<double>());
    if (!rc) rc = se_genvel_dev(m, q->psi2.as<double>(), q->vel2.as<double>());
    if (rc) return rc;
    hipLaunchKernelGGL(k_qg2_pre, dim3(blocks(nV)), dim3(256), 0, s, nV, m->d_verts.as<double>(), q->F1, q->F2,
                       q->beta, q->psi1.as<double>(), q->om1.as<double>(), q->psi2.as<double>(), q->om2.as<double>(),
                       q->a1.as<double>(), q->a2.as<double>(), q->tau.as<double>(), q->chi.as<double>());
    SE_HIP(hipGetLastError());
    const int second = q->nstep > 0;
    if (second) {
        rc = se_genvel_dev(m, q->pp1.as<double>(), q->vp1.as<double>());
        if (!rc) rc = se_genvel_dev(m, q->pp2.as<double>(), q->vp2.as<double>());
        if (rc) return rc;
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
    // SolveQ: layer 2 is solved only in the second-order step
    hipLaunchKernelGGL(k_qg_alpha, dim3(blocks(nV)), dim3(256), 0, s, nV, m->d_f2v.as<int>(), second, second,
                       q->a1.as<double>(), q->h1.as<double>(), q->a2.as<double>(), q->h2.as<double>());
    SE_HIP(hipGetLastError());
    return SE_OK;
}

// assemble the right-hand side of inversion `which` (0 baroclinic, 1 barotropic) into op->f
int qg2_phase_rhs(se_qg2* q, int which) {
    se_mesh* m = q->mesh;
    const int nV = m->nV;
    se_op* op = which ? q->hp : q->hh;
    hipLaunchKernelGGL(k_qg2_rhs, dim3(blocks(nV)), dim3(256), 0, m->ctx->stream, nV, which, m->d_verts.as<double>(),
                       q->F1, q->F2, q->beta, q->h1.as<double>(), q->h2.as<double>(), q->tau.as<double>(),
                       q->rhs.as<double>());
    SE_HIP(hipGetLastError());
    double* sol = which ? q->chi.as<double>() : q->tau.as<double>();
    int rc = op_ensure_work(op);
    if (!rc) rc = se_helm_set_dirichlet_dev(op, sol);
    if (!rc) rc = se_helm_rhs_dev(op, q->rhs.as<double>(), op->f.as<double>());
    return rc;
}

// write the solved unknowns (op->x) of inversion `which` into tau / chi
int qg2_phase_update(se_qg2* q, int which) {
    se_op* op = which ? q->hp : q->hh;
    return se_vec_to_field_dev(q->mesh, op->x.as<double>(), which ? q->chi.as<double>() : q->tau.as<double>());
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
    DevBuf* f[] = {&q->psi1, &q->om1, &q->psi2, &q->om2, &q->a1, &q->a2, &q->h1, &q->h2, &q->rhs, &q->tau,
                   &q->chi, &q->pp1, &q->po1, &q->pp2, &q->po2, &q->ap1, &q->ap2};
    for (DevBuf* b : f)
        if (!rc) rc = b->alloc(sizeof(double) * nV);
    for (DevBuf* b : {&q->vel1, &q->vel2, &q->vp1, &q->vp2})
        if (!rc) rc = b->alloc(sizeof(double) * 2 * nV);
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
    const size_t b = sizeof(double) * q->mesh->nV;
    SE_HIP(hipMemcpy(q->psi1.ptr, psi1, b, hipMemcpyHostToDevice));
    SE_HIP(hipMemcpy(q->om1.ptr, om1, b, hipMemcpyHostToDevice));
    SE_HIP(hipMemcpy(q->psi2.ptr, psi2, b, hipMemcpyHostToDevice));
    SE_HIP(hipMemcpy(q->om2.ptr, om2, b, hipMemcpyHostToDevice));
    q->nstep = 0;  // the next step is first order
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
    for (int which = 0; which < 2 && !rc; which++) {
        se_op* op = which ? q->hp : q->hh;
        rc = qg2_phase_rhs(q, which);
        if (!rc) rc = se_helm_pcg_dev(op, op->f.as<double>(), op->x.as<double>(), opts, &inf[which]);
        if (!rc && inf[which].status != SE_SOLVE_CONVERGED)
            rc = fail(SE_ERR_STATE, "se_qg2_step: an inversion did not converge (status " +
                                        std::to_string(inf[which].status) + ")");
        if (!rc) rc = qg2_phase_update(q, which);
    }
    if (!rc) rc = qg2_phase_post(q);
    if (infos) infos[0] = inf[0], infos[1] = inf[1];
    return rc;
}

// ---------------------------------------------------------------------------
// one-layer QG channel model (QuasiGeostrophicEqn, src/qg/QuasiGeostrophicEqn.cpp:83-205)
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
    for (DevBuf* b : {&q->psi, &q->om, &q->a, &q->hsl, &q->rhs, &q->psin, &q->pp, &q->po, &q->ap})
        if (!rc) rc = b->alloc(sizeof(double) * nV);
    for (DevBuf* b : {&q->vel, &q->vp})
        if (!rc) rc = b->alloc(sizeof(double) * 2 * nV);
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

extern "C" int se_qg_set_state(se_qg* q, const double* psi, const double* omega) {
    SE_CHECK(q && psi && omega, SE_ERR_ARG, "se_qg_set_state: NULL argument");
    SE_HIP(hipSetDevice(q->mesh->ctx->device));
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

extern "C" int se_qg_step(se_qg* q, double dt, const se_solve_opts* opts, se_solve_info* info) {
    SE_CHECK(q, SE_ERR_ARG, "se_qg_step: NULL argument");
    se_mesh* m = q->mesh;
    SE_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    const int nV = m->nV;
    const double* verts = m->d_verts.as<double>();
    const int second = q->nstep > 0;
    // GenVel, GenPV, SL advection of the PV (second order from the previous state)
    int rc = se_genvel_dev(m, q->psi.as<double>(), q->vel.as<double>());
    if (rc) return rc;
    hipLaunchKernelGGL(k_qg_pv, dim3(blocks(nV)), dim3(256), 0, s, nV, verts, q->F, q->beta, q->psi.as<double>(),
                       q->om.as<double>(), q->a.as<double>());
    SE_HIP(hipGetLastError());
    if (second) {
        rc = se_genvel_dev(m, q->pp.as<double>(), q->vp.as<double>());
        if (rc) return rc;
        hipLaunchKernelGGL(k_qg_pv, dim3(blocks(nV)), dim3(256), 0, s, nV, verts, q->F, q->beta, q->pp.as<double>(),
                           q->po.as<double>(), q->ap.as<double>());
        SE_HIP(hipGetLastError());
        rc = se_sl_advect2_dev(m, q->a.as<double>(), q->ap.as<double>(), 1, q->vel.as<double>(), q->vp.as<double>(), dt,
                               q->hsl.as<double>());
    } else {
        rc = se_sl_advect_dev(m, q->a.as<double>(), 1, q->vel.as<double>(), dt, q->hsl.as<double>(), nullptr);
    }
    if (rc) return rc;
    // SolveQ: hsl := alpha
    hipLaunchKernelGGL(k_qg_alpha, dim3(blocks(nV)), dim3(256), 0, s, nV, m->d_f2v.as<int>(), second, 0,
                       q->a.as<double>(), q->hsl.as<double>(), (const double*)nullptr, (double*)nullptr);
    SE_HIP(hipGetLastError());
    // (lap - F) psi = alpha - beta y with the wall data of psi
    hipLaunchKernelGGL(k_qg_rhs, dim3(blocks(nV)), dim3(256), 0, s, nV, verts, q->beta, q->hsl.as<double>(),
                       q->rhs.as<double>());
    SE_HIP(hipGetLastError());
    se_op* op = q->h;
    rc = op_ensure_work(op);
    if (!rc) rc = se_helm_set_dirichlet_dev(op, q->psi.as<double>());
    if (!rc) rc = se_helm_rhs_dev(op, q->rhs.as<double>(), op->f.as<double>());
    se_solve_info inf = {};
    if (!rc) rc = se_helm_pcg_dev(op, op->f.as<double>(), op->x.as<double>(), opts, &inf);
    if (!rc && inf.status != SE_SOLVE_CONVERGED)
        rc = fail(SE_ERR_STATE, "se_qg_step: the inversion did not converge (status " + std::to_string(inf.status) + ")");
    if (!rc) SE_HIP(hipMemcpyAsync(q->psin.ptr, q->psi.ptr, sizeof(double) * nV, hipMemcpyDeviceToDevice, s));
    if (!rc) rc = se_vec_to_field_dev(m, op->x.as<double>(), q->psin.as<double>());
    if (!rc) {
        hipLaunchKernelGGL(k_qg_post, dim3(blocks(nV)), dim3(256), 0, s, nV, verts, q->F, q->beta, q->hsl.as<double>(),
                           q->psin.as<double>(), q->psi.as<double>(), q->om.as<double>(), q->pp.as<double>(),
                           q->po.as<double>());
        SE_HIP(hipGetLastError());
        q->nstep++;
    }
    if (info) *info = inf;
    return rc;
}
