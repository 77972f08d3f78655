// Drop-in check: the reference's own Mesh / BCs / Field objects (compiled from
// the reference tree) as the caller, the drop-in HelmholtzEqn / PoissonEqn /
// Advector (se_amd/dropin) as the solver, the C oracle as the checker.  One
// QuasiGeostrophicEqn-style sequence on an x-periodic channel
// (src/qg/QuasiGeostrophicEqn.cpp:48-52,117-147 pattern): Helmholtz inversion
// with Dirichlet walls, first- and second-order advection, a Poisson solve;
// then a NavierStokesEqn-style velocity solve (src/ns/NavierStokesEqn.cpp:90
// pattern): HelmholtzEqn on a 2-dof field with per-dof walls (the reference's
// 2-dof BCs) and the SVV operator.  Prints "dropin ok" and exits 0 when every
// comparison passes.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "QuadPoint.h"
#include "LinAlg.h"
#include "Jacobi.h"
#include "Element.h"
#include "Mesh.h"
#include "BCs.h"
#include "Field.h"

#include "Advector.h"
#include "HelmholtzEqn.h"
#include "PoissonEqn.h"
#include "se_oracle.h"

static int failures = 0;

static void compare(const char* what, const std::vector<double>& a, const std::vector<double>& b, double tol) {
    double d = 0.0, nb = 0.0;
    for (size_t i = 0; i < a.size(); i++) {
        d += (a[i] - b[i]) * (a[i] - b[i]);
        nb += b[i] * b[i];
    }
    const double rel = std::sqrt(d / (nb > 0 ? nb : 1.0));
    const bool ok = rel <= tol;
    std::printf("%-28s rel diff %.3e (bar %.0e) %s\n", what, rel, tol, ok ? "ok" : "FAIL");
    if (!ok) failures++;
}

static std::vector<double> values(const Field* f) {
    std::vector<double> v((size_t)f->mesh->nVertsTotal * f->nDofs);
    for (int i = 0; i < f->mesh->nVertsTotal; i++)
        for (int d = 0; d < f->nDofs; d++) v[(size_t)i * f->nDofs + d] = f->vals[i][d];
    return v;
}

int main() {
    int nEls[2] = {8, 6};
    double lo[2] = {0.0, 0.0}, hi[2] = {1.0, 1.0};
    bool per[2] = {true, false};
    const int N = 6;
    Mesh* mesh = new Mesh("mesh", nEls, "legendre", N, lo, hi, per);
    BCs* bcs = new BCs(true, true, false, false, mesh, 0);
    Field* psi = new Field("psi", mesh, 1, bcs);
    Field* rhs = new Field("rhs", mesh, 1, NULL);
    Field* q = new Field("q", mesh, 1, NULL);
    Field* qPrev = new Field("qPrev", mesh, 1, NULL);
    Field* vel = new Field("vel", mesh, 2, NULL);
    Field* velPrev = new Field("velPrev", mesh, 2, NULL);
    const int nV = mesh->nVertsTotal;
    for (int i = 0; i < nV; i++) {
        const double x = mesh->verts[i][0], y = mesh->verts[i][1];
        psi->vals[i][0] = 0.0;  // walls: psi = 0
        rhs->vals[i][0] = std::cos(2.0 * M_PI * x) * std::sin(M_PI * y) + 0.3 * std::sin(4.0 * M_PI * x);
        q->vals[i][0] = std::exp(-40.0 * ((x - 0.4) * (x - 0.4) + (y - 0.5) * (y - 0.5)));
        qPrev->vals[i][0] = 0.9 * q->vals[i][0];
        vel->vals[i][0] = 0.3 * std::sin(M_PI * y);
        vel->vals[i][1] = 0.1 * std::sin(2.0 * M_PI * x) * std::sin(M_PI * y);
        velPrev->vals[i][0] = 0.95 * vel->vals[i][0];
        velPrev->vals[i][1] = 0.95 * vel->vals[i][1];
    }
    const int nE[2] = {nEls[0], nEls[1]}, pi[2] = {1, 0}, bc[4] = {1, 1, 0, 0};
    int st = 0;
    orc_mesh* om = orc_mesh_create(nE, N, lo, hi, pi, bc, &st);
    std::vector<double> verts(2 * (size_t)nV), r = values(rhs), s0 = values(psi);

    // Helmholtz inversion (QG: c = -F, nu = -1, dt = 1)
    HelmholtzEqn* he = new HelmholtzEqn(psi, rhs, -100.0, -1.0, 1.0, 0);
    he->Assemble();
    he->Solve("helm_");
    orc_helm* oh = orc_helm_create(om, -100.0, -1.0, 0);
    orc_helm_assemble(oh, s0.data());
    int it = 0;
    double rr = 0.0;
    std::vector<double> s = s0;
    orc_helm_solve(oh, r.data(), s.data(), 1e-10, 100000, 0, &it, &rr);
    std::printf("HelmholtzEqn: status %d, %d iterations (oracle %d)\n", he->info.status, he->info.iters, it);
    if (he->info.status != 0) failures++;
    compare("HelmholtzEqn::Solve", values(psi), s, 1e-8);

    // first-order advection
    Advector* a1 = new Advector(q, vel);
    a1->Advect(0.01);
    std::vector<double> qs(nV), qv = values(q), vv = values(vel);
    orc_advect_first(om, qv.data(), 1, vv.data(), 0.01, qs.data(), NULL, NULL);
    compare("Advector::Advect (1st)", values(a1->fieldSL), qs, 1e-12);

    // second-order advection
    Advector* a2 = new Advector(q, vel, qPrev, velPrev);
    a2->Advect(0.01);
    std::vector<double> qp = values(qPrev), vp = values(velPrev);
    orc_advect_second(om, qv.data(), qp.data(), 1, vv.data(), vp.data(), 0.01, qs.data());
    compare("Advector::Advect (2nd)", values(a2->fieldSL), qs, 1e-12);
    {
        // the extrapolated velocities left in the public fields (Advector.cpp:339-349)
        std::vector<double> e1(2 * (size_t)nV), eh(2 * (size_t)nV), em(2 * (size_t)nV);
        for (size_t i = 0; i < e1.size(); i++) {
            e1[i] = 2.0 * vv[i] - 1.0 * vp[i];
            eh[i] = 1.5 * vv[i] - 0.5 * vp[i];
            em[i] = 0.5 * vv[i] + 0.5 * vp[i];
        }
        compare("Advector::velPlusOne", values(a2->velPlusOne), e1, 0.0);
        compare("Advector::velPlusHalf", values(a2->velPlusHalf), eh, 0.0);
        compare("Advector::velMinusHalf", values(a2->velMinusHalf), em, 0.0);
        if (a2->fieldSL->name != "q-SL") {
            std::printf("FAIL  fieldSL name %s (expected q-SL)\n", a2->fieldSL->name.c_str());
            failures++;
        }
    }

    // cubicInterp: the reference's default (N < 3) is off here; a caller may
    // set it (QuasiGeostrophicSpectral.cpp:84) -> bicubic field interpolation
    {
        Advector* a3 = new Advector(q, vel);
        if (a3->cubicInterp) {
            std::printf("FAIL  cubicInterp defaults to true at N = %d\n", N);
            failures++;
        }
        a3->cubicInterp = true;
        a3->Advect(0.01);
        orc_advect_first_interp(om, qv.data(), 1, vv.data(), 0.01, qs.data(), 2);
        compare("Advector::Advect (1st, cubicInterp)", values(a3->fieldSL), qs, 1e-12);
        // Advector::Interpolate at scattered points, against the oracle's bicubic
        std::vector<double> got, want;
        for (int k = 0; k < 40; k++) {
            double c[2] = {lo[0] + (hi[0] - lo[0]) * (0.013 + 0.0247 * k), lo[1] + (hi[1] - lo[1]) * (0.97 - 0.0233 * k)};
            double p1 = 0.0, p0 = 0.0;
            a3->Interpolate(q, c, &p1);
            orc_interp_bicubic(om, qv.data(), 1, c, &p0);
            got.push_back(p1);
            want.push_back(p0);
        }
        compare("Advector::Interpolate", got, want, 1e-13);
        delete a3;
    }

    // Poisson (barotropic: c = -1, Dirichlet walls)
    Field* chi = new Field("chi", mesh, 1, bcs);
    PoissonEqn* pe = new PoissonEqn(chi, rhs, -1.0);
    pe->Solve(false);
    orc_helm* op = orc_helm_create(om, 0.0, -1.0, 0);
    std::vector<double> c0(nV, 0.0);
    orc_helm_assemble(op, c0.data());
    orc_helm_solve(op, r.data(), c0.data(), 1e-10, 100000, 0, &it, &rr);
    std::printf("PoissonEqn: status %d, %d iterations (oracle %d)\n", pe->info.status, pe->info.iters, it);
    if (pe->info.status != 0) failures++;
    compare("PoissonEqn::Solve", values(chi), c0, 1e-8);

    // velocity Helmholtz with SVV: u walls bottom/top, v walls on all sides,
    // non-zero wall data, rhs a 2-dof field
    {
        int nE2[2] = {5, 4};
        double lo2[2] = {0.0, 0.0}, hi2[2] = {2.0, 1.0};
        bool per2[2] = {false, false};
        Mesh* mesh2 = new Mesh("box", nE2, "legendre", N, lo2, hi2, per2);
        bool bb[2] = {true, true}, bt[2] = {true, true}, bl[2] = {false, true}, br[2] = {false, true};
        BCs* bcs2 = new BCs(bb, bt, bl, br, mesh2, 0);
        Field* u = new Field("u", mesh2, 2, bcs2);
        Field* f2 = new Field("f", mesh2, 2, NULL);
        const int nV2 = mesh2->nVertsTotal;
        for (int i = 0; i < nV2; i++) {
            const double x = mesh2->verts[i][0], y = mesh2->verts[i][1];
            u->vals[i][0] = y * (1.0 - y) + 0.5 * y;  // walls: 0 at the bottom, 0.5 at the top
            u->vals[i][1] = 0.1 * x * y;
            f2->vals[i][0] = std::sin(3.0 * x) * std::cos(2.0 * y) + 1.0;
            f2->vals[i][1] = std::exp(-((x - 1.0) * (x - 1.0) + (y - 0.5) * (y - 0.5)) * 8.0);
        }
        const double dt2 = 0.01, nu2 = 1e-3;
        const int cutoff = 3;
        std::vector<double> u0 = values(u), fv = values(f2);
        HelmholtzEqn* hv = new HelmholtzEqn(u, f2, 1.0, nu2, dt2, cutoff);
        hv->Solve("vel_");
        std::printf("HelmholtzEqn (2 dofs, SVV): status %d, %d iterations\n", hv->info.status, hv->info.iters);
        if (hv->info.status != 0) failures++;
        std::vector<double> got = values(u);
        const int sides[2][4] = {{1, 1, 0, 0}, {1, 1, 1, 1}};
        const int np2[2] = {0, 0};
        for (int d = 0; d < 2; d++) {
            orc_mesh* om2 = orc_mesh_create(nE2, N, lo2, hi2, np2, sides[d], &st);
            orc_helm* ov = orc_helm_create(om2, 1.0, nu2 * dt2, 0);
            orc_helm_set_svv(ov, dt2, cutoff);
            std::vector<double> sd(nV2), rd(nV2), gd(nV2);
            for (int i = 0; i < nV2; i++) sd[i] = u0[2 * (size_t)i + d], rd[i] = fv[2 * (size_t)i + d],
                                        gd[i] = got[2 * (size_t)i + d];
            orc_helm_assemble(ov, sd.data());
            orc_helm_solve(ov, rd.data(), sd.data(), 1e-12, 100000, 0, &it, &rr);
            compare(d ? "HelmholtzEqn SVV, dof 1" : "HelmholtzEqn SVV, dof 0", gd, sd, 1e-8);
            orc_helm_destroy(ov);
            orc_mesh_destroy(om2);
        }
        delete hv;
    }

    delete he;
    delete pe;
    delete a1;
    delete a2;
    orc_helm_destroy(oh);
    orc_helm_destroy(op);
    orc_mesh_destroy(om);
    if (failures) {
        std::printf("dropin FAILED (%d)\n", failures);
        return 1;
    }
    std::printf("dropin ok\n");
    return 0;
}
