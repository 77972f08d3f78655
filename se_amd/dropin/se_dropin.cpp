// se_dropin.cpp -- HelmholtzEqn, PoissonEqn and Advector with the reference's
// interface, on libse_amd.  Compiled inside the reference tree (its
// src/base headers on the include path) in place of src/linsys/HelmholtzEqn.cpp,
// src/linsys/PoissonEqn.cpp and src/base/Advector.cpp; links libse_amd.so.
//
// The reference's Mesh / BCs / Field stay the callers' objects.  Each distinct
// (Mesh, Dirichlet sides) pair gets one device mesh (numbering maps and
// geometry built on the GPU, kept for the process), and the field values are
// gathered from / scattered to Field::vals around each call.  Errors print to
// cerr and exit(1), as the reference does on misuse.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <mutex>
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
#include "se_amd.h"

namespace {

std::mutex g_mu;
se_ctx* g_ctx = nullptr;

struct MeshEntry {
    const Mesh* mesh;
    int bc[4];
    se_mesh* m;
};
std::vector<MeshEntry> g_meshes;

void die(const std::string& what, int rc) {
    std::cerr << "libse_amd: " << what << " failed (" << rc << "): " << se_last_error() << std::endl;
    exit(1);
}

se_ctx* ctx() {
    if (!g_ctx) {
        const char* d = getenv("SE_DEVICE");
        int rc = se_ctx_create(d ? atoi(d) : 0, &g_ctx);
        if (rc) die("se_ctx_create", rc);
    }
    return g_ctx;
}

// the device mesh of a field's dof: geometry of its Mesh, Dirichlet sides of
// its BCs for that dof (a 2-dof BCs holds one set per dof, BCs.cpp:176-322)
se_mesh* device_mesh(const Field* f, int dof = 0) {
    std::lock_guard<std::mutex> lk(g_mu);
    const Mesh* M = f->mesh;
    int bc[4] = {0, 0, 0, 0};
    if (f->bcs) {
        // BCs marks the duplicate line of a periodic direction as a boundary
        // (BCs.cpp:38-44); those are not Dirichlet sides
        bc[0] = M->periodic[1] ? 0 : f->bcs->bcBottom[dof];
        bc[1] = M->periodic[1] ? 0 : f->bcs->bcTop[dof];
        bc[2] = M->periodic[0] ? 0 : f->bcs->bcLeft[dof];
        bc[3] = M->periodic[0] ? 0 : f->bcs->bcRight[dof];
    }
    for (const MeshEntry& e : g_meshes)
        if (e.mesh == M && e.bc[0] == bc[0] && e.bc[1] == bc[1] && e.bc[2] == bc[2] && e.bc[3] == bc[3]) return e.m;
    int nEls[2] = {M->nEls[0], M->nEls[1]};
    int per[2] = {M->periodic[0] ? 1 : 0, M->periodic[1] ? 1 : 0};
    double lo[2] = {M->min[0], M->min[1]}, hi[2] = {M->max[0], M->max[1]};
    se_mesh* m = nullptr;
    int rc = se_mesh_create(ctx(), nEls, M->N, lo, hi, per, bc, &m);
    if (rc) die("se_mesh_create", rc);
    g_meshes.push_back(MeshEntry{M, {bc[0], bc[1], bc[2], bc[3]}, m});
    return m;
}

void gather(const Field* f, std::vector<double>& out) {
    const int nV = f->mesh->nVertsTotal, nd = f->nDofs;
    out.resize((size_t)nV * nd);
    for (int v = 0; v < nV; v++)
        for (int d = 0; d < nd; d++) out[(size_t)v * nd + d] = f->vals[v][d];
}

void scatter(const std::vector<double>& in, Field* f) {
    const int nV = f->mesh->nVertsTotal, nd = f->nDofs;
    for (int v = 0; v < nV; v++)
        for (int d = 0; d < nd; d++) f->vals[v][d] = in[(size_t)v * nd + d];
}

se_solve_opts solve_opts(bool null_space) {
    se_solve_opts o;
    se_solve_opts_default(&o);
    if (const char* r = getenv("SE_RTOL")) o.rtol = atof(r);
    if (const char* p = getenv("SE_PRECISION")) {
        const std::string s(p);
        if (s == "fp32") o.precision = SE_PREC_FP32_IR;
        if (s == "bf16") o.precision = SE_PREC_BF16_IR;
        if (s == "fp64_fp32c") o.precision = SE_PREC_FP64_FP32C;
        if (s == "fp64_bf16c") o.precision = SE_PREC_FP64_BF16C;
    }
    if (const char* p = getenv("SE_PRECONDITIONER")) {
        const std::string s(p);
        if (s == "jacobi") o.preconditioner = SE_PC_JACOBI;
        if (s == "fdm") o.preconditioner = SE_PC_FDM;
    }
    o.null_space = null_space ? 1 : 0;
    return o;
}

void check_scalar(const Field* f, const char* who) {
    if (f->nDofs != 1) {
        std::cerr << who << ": only single-dof fields are solved by libse_amd" << std::endl;
        exit(1);
    }
}

// dof d of a node-major field array
std::vector<double> column(const std::vector<double>& a, int nd, int d) {
    std::vector<double> c(a.size() / nd);
    for (size_t v = 0; v < c.size(); v++) c[v] = a[v * nd + d];
    return c;
}

// Field-level solve shared by both equations: f = M rhs + b, solve, UpdateField
void field_solve(se_op* op, Field* sol, Field* rhs, bool null_space, se_solve_info* info) {
    std::vector<double> r, s;
    gather(rhs, r);
    gather(sol, s);
    se_solve_opts o = solve_opts(null_space);
    int rc = se_helm_solve(op, r.data(), s.data(), &o, info);
    if (rc) die("se_helm_solve", rc);
    if (info->status != SE_SOLVE_CONVERGED)
        std::cerr << "libse_amd: solve ended with status " << info->status << " after " << info->iters
                  << " iterations, relative residual " << info->true_rel_residual << std::endl;
    scatter(s, sol);
}

}  // namespace

// ---- HelmholtzEqn (src/linsys/HelmholtzEqn.cpp:21-136) ----------------------
// A 2-dof field (a velocity) is the reference's block-diagonal system of one
// operator per dof, each with that dof's Dirichlet sides (the element
// operators couple equal dofs only, Operator.cpp:67-80,161-175, SVV.cpp:166-175;
// BCs numbers dof 0's unknowns first): one device operator per dof.
HelmholtzEqn::HelmholtzEqn(Field* _solField, Field* _rhsField, double _c, double _nu, double _dt, int _svvCutoff)
    : solField(_solField), rhsField(_rhsField), c(_c), nu(_nu), dt(_dt), svvCutoff(_svvCutoff), info(),
      assembled(false), op(nullptr), op2(nullptr) {
    const int nd = solField->nDofs;
    if (nd != 1 && nd != 2) {
        std::cerr << "HelmholtzEqn: libse_amd solves fields of 1 or 2 dofs" << std::endl;
        exit(1);
    }
    if (rhsField && rhsField->nDofs != nd) {
        std::cerr << "HelmholtzEqn: the rhs field must have the solution's dofs" << std::endl;
        exit(1);
    }
    for (int d = 0; d < nd; d++) {
        se_op** o = d ? &op2 : &op;
        int rc = se_helm_create(device_mesh(solField, d), c, nu * dt, SE_DETJ_LITERAL, o);
        if (rc) die("se_helm_create", rc);
        if (svvCutoff) {
            rc = se_helm_set_svv(*o, dt, svvCutoff);
            if (rc) die("se_helm_set_svv", rc);
        }
    }
}

HelmholtzEqn::~HelmholtzEqn() {
    se_helm_destroy(op);
    se_helm_destroy(op2);
}

void HelmholtzEqn::Assemble() {
    std::vector<double> s;
    gather(solField, s);
    const int nd = solField->nDofs;
    for (int d = 0; d < nd; d++) {
        std::vector<double> sd = column(s, nd, d);
        int rc = se_helm_set_dirichlet(d ? op2 : op, sd.data());
        if (rc) die("se_helm_set_dirichlet", rc);
    }
    assembled = true;
}

void HelmholtzEqn::Solve(std::string name, bool removeNullSpace) {
    (void)name;  // PETSc options prefix in the reference; options come from SE_* here
    if (!assembled) Assemble();
    const int nd = solField->nDofs;
    if (nd == 1) {
        field_solve(op, solField, rhsField, removeNullSpace, &info);
        return;
    }
    // both dofs gathered before either is written (ShallowWaterEqnSC solves with rhs == sol)
    std::vector<double> r, s;
    gather(rhsField, r);
    gather(solField, s);
    se_solve_opts o = solve_opts(removeNullSpace);
    info = se_solve_info();
    for (int d = 0; d < nd; d++) {
        std::vector<double> rd = column(r, nd, d), sd = column(s, nd, d);
        se_solve_info one;
        int rc = se_helm_solve(d ? op2 : op, rd.data(), sd.data(), &o, &one);
        if (rc) die("se_helm_solve", rc);
        if (one.status != SE_SOLVE_CONVERGED)
            std::cerr << "libse_amd: solve of dof " << d << " ended with status " << one.status << " after "
                      << one.iters << " iterations" << std::endl;
        for (size_t v = 0; v < sd.size(); v++) s[v * nd + d] = sd[v];
        // the pair's outcome: the worse status, all iterations, the larger residual and both times
        if (one.status > info.status) info.status = one.status;
        info.iters += one.iters;
        info.outer_iters += one.outer_iters;
        info.true_rel_residual = std::max(info.true_rel_residual, one.true_rel_residual);
        info.seconds += one.seconds;
    }
    scatter(s, solField);
}

// ---- PoissonEqn (src/linsys/PoissonEqn.cpp:30-134) ----------------------------
PoissonEqn::PoissonEqn(Field* _solField, Field* _rhsField, double _c)
    : solField(_solField), rhsField(_rhsField), c(_c), info(), assembled(false), op(nullptr) {
    check_scalar(solField, "PoissonEqn");
    int rc = se_helm_create(device_mesh(solField), 0.0, c, SE_DETJ_LITERAL, &op);
    if (rc) die("se_helm_create", rc);
}

PoissonEqn::~PoissonEqn() { se_helm_destroy(op); }

void PoissonEqn::Assemble() {
    std::vector<double> s;
    gather(solField, s);
    int rc = se_helm_set_dirichlet(op, s.data());
    if (rc) die("se_helm_set_dirichlet", rc);
    assembled = true;
}

void PoissonEqn::Solve(bool removeNullSpace) {
    if (!assembled) Assemble();
    field_solve(op, solField, rhsField, removeNullSpace, &info);
}

// ---- Advector (src/base/Advector.cpp:26-349) ----------------------------------
// names of the reference's fields: "<field>-SL", "<vel>_n+1", "<vel>_n+1/2", "<vel>_n-1/2"
Advector::Advector(Field* _field, Field* _velocity)
    : field(_field), velocity(_velocity), fieldMinusOne(nullptr), velMinusOne(nullptr), velPlusOne(nullptr),
      velPlusHalf(nullptr), velMinusHalf(nullptr), secondOrder(false), cubicInterp(_field->mesh->N < 3) {
    fieldSL = new Field(field->name + "-SL", field->mesh, field->nDofs, field->bcs);
    fieldSL->Copy(field);  // the boundary values, as the reference
}

Advector::Advector(Field* _field, Field* _velocity, Field* _fieldMinusOne, Field* _velMinusOne)
    : field(_field), velocity(_velocity), fieldMinusOne(_fieldMinusOne), velMinusOne(_velMinusOne),
      secondOrder(true), cubicInterp(_field->mesh->N < 3) {
    fieldSL = new Field(field->name + "-SL", field->mesh, field->nDofs, field->bcs);
    fieldSL->Copy(field);
    velPlusOne = new Field(velocity->name + "_n+1", velocity->mesh, velocity->nDofs, velocity->bcs);
    velPlusHalf = new Field(velocity->name + "_n+1/2", velocity->mesh, velocity->nDofs, velocity->bcs);
    velMinusHalf = new Field(velocity->name + "_n-1/2", velocity->mesh, velocity->nDofs, velocity->bcs);
}

Advector::~Advector() {
    delete fieldSL;
    if (secondOrder) {
        delete velPlusOne;
        delete velPlusHalf;
        delete velMinusHalf;
    }
}

void Advector::Advect(double dt) {
    se_mesh* m = device_mesh(field);
    int rc = se_mesh_set_sl_interp(m, cubicInterp ? SE_INTERP_BICUBIC : SE_INTERP_GLL);
    if (rc) die("se_mesh_set_sl_interp", rc);
    std::vector<double> q, v, out((size_t)field->mesh->nVertsTotal * field->nDofs);
    gather(field, q);
    gather(velocity, v);
    if (secondOrder) {
        std::vector<double> qp, vp;
        gather(fieldMinusOne, qp);
        gather(velMinusOne, vp);
        // the extrapolated velocities the reference leaves in its fields
        // (ExtrapolateVelocities, Advector.cpp:339-349); the device step
        // forms the same values for itself
        for (int i = 0; i < velocity->mesh->nVertsTotal; i++)
            for (int d = 0; d < velocity->nDofs; d++) {
                const double a = velocity->vals[i][d], b = velMinusOne->vals[i][d];
                velPlusOne->vals[i][d] = 2.0 * a - 1.0 * b;
                velPlusHalf->vals[i][d] = 1.5 * a - 0.5 * b;
                velMinusHalf->vals[i][d] = 0.5 * a + 0.5 * b;
            }
        rc = se_sl_advect2(m, q.data(), qp.data(), field->nDofs, v.data(), vp.data(), dt, out.data());
    } else {
        rc = se_sl_advect(m, q.data(), field->nDofs, v.data(), dt, out.data(), nullptr);
    }
    if (rc) die("se_sl_advect", rc);
    scatter(out, fieldSL);
}

void Advector::Interpolate(Field* f, double* coord, double* psi) {
    se_mesh* m = device_mesh(f);
    std::vector<double> vals;
    gather(f, vals);
    const double xy[2] = {coord[0], coord[1]};
    int rc = se_interp_points(m, vals.data(), f->nDofs, 1, xy, SE_INTERP_BICUBIC, psi);
    if (rc) die("se_interp_points", rc);
}
