// qg.h -- QG model objects and the two-layer step's phases (qg.hip), shared
// with the request server.
#pragma once
#include "se_common.h"
#include "se_op.h"

struct se_qg2 {
    se_mesh* mesh = nullptr;
    se_op* hh = nullptr;  // baroclinic Helmholtz (tau)
    se_op* hp = nullptr;  // barotropic Poisson (chi)
    double F1 = 0.0, F2 = 0.0, beta = 0.0;
    double E2 = 0.0, r = 0.0, nu = 0.0;  // Ekman drag (layer 2), thermal relaxation, viscosity
    bool has_tau_e = false;              // relaxation target set (otherwise zero)
    int nstep = 0;  // steps since the state was set (0: the next step is first order)
    // vel1, vel2 still hold GenVel of the streamfunctions the last step began
    // with (which its end copied into pp1, pp2): the next step takes them as
    // the previous velocities instead of computing GenVel of pp1, pp2 again
    bool vel_prev_ok = false;
    // the step in flight (set by the advection phase)
    int second = 0;
    double dt = 0.0;
    se::DevBuf psi1, om1, psi2, om2, vel1, vel2, a1, a2, h1, h2, rhs, tau, chi;
    se::DevBuf pp1, po1, pp2, po2, ap1, ap2, vp1, vp2;  // the state before the last step, its PV, velocities
    se::DevBuf om1h, om2h, tau_e;                        // omega^ of the first inversion pair, tau_e
    se::DevBuf xq;                                       // SolveQ's unknowns
};

// one-layer QG channel model
struct se_qg {
    se_mesh* mesh = nullptr;
    se_op* h = nullptr;  // (lap - F) psi
    double F = 0.0, beta = 0.0;
    double nu = 0.0, r = 0.0;  // viscosity, linear drag
    bool has_topo = false;
    int nstep = 0;
    se::DevBuf psi, om, vel, a, hsl, rhs, psin, pp, po, ap, vp;
    se::DevBuf omh, topo, xq;
};

namespace se {
size_t qg2_bytes(int n, int nV);
// inversions per step: 2 (tau, chi), or 4 with dissipation (the first pair on
// the advected PVs feeds the dissipation terms)
int qg2_passes(const se_qg2* q);
se_op* qg2_pass_op(se_qg2* q, int k);
// GenVel, GenPV, SL advection; without dissipation also SolveQ
int qg2_phase_advect(se_qg2* q, double dt);
// right-hand side of inversion k into its operator's f (and its Dirichlet lifting)
int qg2_phase_rhs(se_qg2* q, int k);
// inversion k's solution (op->x) into tau / chi; after the first pair of a
// dissipative step, omega^ and the dissipative SolveQ
int qg2_phase_update(se_qg2* q, int k);
int qg2_phase_post(se_qg2* q);
}  // namespace se
