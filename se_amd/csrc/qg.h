// qg.h -- two-layer QG model object and its step phases (qg.hip), shared with
// the request server.
#pragma once
#include "se_common.h"
#include "se_op.h"

struct se_qg2 {
    se_mesh* mesh = nullptr;
    se_op* hh = nullptr;  // baroclinic Helmholtz (tau)
    se_op* hp = nullptr;  // barotropic Poisson (chi)
    double F1 = 0.0, F2 = 0.0, beta = 0.0;
    se::DevBuf psi1, om1, psi2, om2, vel1, vel2, a1, a2, h1, h2, rhs, tau, chi;
};

namespace se {
size_t qg2_bytes(int n, int nV);
int qg2_phase_advect(se_qg2* q, double dt);
int qg2_phase_rhs(se_qg2* q, int which);
int qg2_phase_update(se_qg2* q, int which);
int qg2_phase_post(se_qg2* q);
}  // namespace se
