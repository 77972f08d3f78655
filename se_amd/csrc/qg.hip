// qg.hip -- one first-order step of the two-layer quasi-geostrophic model on
// the device (TwoLayerQGEqn::FirstOrder, src/qg/TwoLayerQGEqn.cpp:174-243, with
// GenVel :335-352 and GenPV):
//   1. u_i = GenVel(psi_i); alpha_i = omega_i + F_i (psi_j - psi_i) + beta y
//      (GenPV), advected semi-Lagrangian: alpha_i^ = alpha_i(departure points)
//   2. baroclinic tau = psi1 - psi2:  (lap - (F1+F2)) tau = alpha1^ - alpha2^
//   3. barotropic chi = psi1 + psi2:  lap chi = alpha1^ + alpha2^ + (F1-F2) tau - 2 beta y
//   4. psi1 = (tau+chi)/2, psi2 = (chi-tau)/2,
//      omega_i = alpha_i^ - F_i (psi_j - psi_i) - beta y
// The inversions use HelmholtzEqn / PoissonEqn semantics (rhs field assembled
// with the mass matrix plus the Dirichlet lifting of the current tau / chi,
// UpdateField after the solve).  The step is split into phases so that a
// server can run the inversions of many models as one batched solve.
#include <algorithm>
#include <cmath>

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

__global__ void k_qg2_post(int nV, const double* __restrict__ verts, double F1, double F2, double beta,
                           const double* __restrict__ tau, const double* __restrict__ chi,
                           const double* __restrict__ h1, const double* __restrict__ h2, double* __restrict__ psi1,
                           double* __restrict__ om1, double* __restrict__ psi2, double* __restrict__ om2) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
        const double y = verts[2 * v + 1];
        const double p1 = 0.5 * (tau[v] + chi[v]);
        const double p2 = 0.5 * (chi[v] - tau[v]);
        psi1[v] = p1;
        psi2[v] = p2;
        om1[v] = h1[v] - F1 * (p2 - p1) - beta * y;
        om2[v] = h2[v] - F2 * (p1 - p2) - beta * y;
    }
}

}  // namespace

namespace se {

size_t qg2_bytes(int n, int nV) {
    // device fields and the two operators' work (an estimate for admission)
    return (size_t)nV * 8 * 16 + (size_t)n * (2 * (8 * 10 + 4 * 4) + 8 * 4);
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
    rc = se_sl_advect_dev(m, q->a1.as<double>(), 1, q->vel1.as<double>(), dt, q->h1.as<double>(), nullptr);
    if (!rc) rc = se_sl_advect_dev(m, q->a2.as<double>(), 1, q->vel2.as<double>(), dt, q->h2.as<double>(), nullptr);
    return rc;
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
                       q->om2.as<double>());
    SE_HIP(hipGetLastError());
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
    DevBuf* f[] = {&q->psi1, &q->om1, &q->psi2, &q->om2, &q->a1, &q->a2, &q->h1, &q->h2, &q->rhs, &q->tau, &q->chi};
    for (DevBuf* b : f)
        if (!rc) rc = b->alloc(sizeof(double) * nV);
    if (!rc) rc = q->vel1.alloc(sizeof(double) * 2 * nV);
    if (!rc) rc = q->vel2.alloc(sizeof(double) * 2 * nV);
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
