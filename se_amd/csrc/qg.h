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
    int nstep = 0;  // steps since the state was set (0: the next step is first order)
    se::DevBuf psi1, om1, psi2, om2, vel1, vel2, a1, a2, h1, h2, rhs, tau, chi;
    se::DevBuf pp1, po1, pp2, po2, ap1, ap2, vp1, vp2;  // the state before the last step, its PV, velocities
};

// one-layer QG channel model
struct se_qg {
    se_mesh* mesh = nullptr;
    se_op* h = nullptr;  // (lap - F) psi
    double F = 0.0, beta = 0.0;
    int nstep = 0;
    se::DevBuf psi, om, vel, a, hsl, rhs, psin, pp, po, ap, vp;
};

namespace se {
size_t qg2_bytes(int n, int nV);
int qg2_phase_advect(se_qg2* q, double dt);
int qg2_phase_rhs(se_qg2* q, int which);
int qg2_phase_update(se_qg2* q, int which);
int qg2_phase_post(se_qg2* q);
}  // namespace se
