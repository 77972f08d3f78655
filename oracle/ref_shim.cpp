// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Thin extern "C" entry points over the reference's own C++ classes
// (davelee2804/SE src/base: Legendre, Mesh, BCs, Field, Operator, RHSOp,
// Advector), compiled in place from the reference tree by build_ref.sh into
// oracle/_ref/libseref.so.  Used only to write golden vectors
// (tests/golden/make_golden.py) and to cross-check the C oracle.  Nothing here
// restates an algorithm: every value comes from a reference method call.
//
// Access: the reference keeps Legendre::dCij and Advector::IntegrateRK2
// private; `private` is widened to `public` for this translation unit only so
// that the shim can read them.  Layout and code of the classes are unchanged.
#include <cstring>
#include <string>
#include <iostream>
#define private public
#include "QuadPoint.h"
#include "LinAlg.h"
#include "Jacobi.h"
#include "Element.h"
#include "Legendre.h"
#include "Mesh.h"
#include "BCs.h"
#include "Field.h"
#include "Operator.h"
#include "RHSOp.h"
#include "Advector.h"
#undef private

namespace {
struct RefMesh {
    Mesh* mesh;
    BCs* bcs;
};

void load_field(Field* f, const double* vals) {
    for (int v = 0; v < f->mesh->nVertsTotal; v++)
        for (int d = 0; d < f->nDofs; d++) f->vals[v][d] = vals[(size_t)v * f->nDofs + d];
}

void store_field(const Field* f, double* vals) {
    for (int v = 0; v < f->mesh->nVertsTotal; v++)
        for (int d = 0; d < f->nDofs; d++) vals[(size_t)v * f->nDofs + d] = f->vals[v][d];
}

// the reference prints progress lines; keep stdout clean for the caller
struct Quiet {
    std::streambuf* old;
    Quiet() : old(std::cout.rdbuf(nullptr)) {}
    ~Quiet() { std::cout.rdbuf(old); }
};
}  // namespace

extern "C" {

void* ref_mesh_create(const int* nEls, int N, const double* min, const double* max,
                      const int* periodic, const int* bc) {
    int ne[2] = {nEls[0], nEls[1]};
    double mn[2] = {min[0], min[1]}, mx[2] = {max[0], max[1]};
    bool per[2] = {periodic[0] != 0, periodic[1] != 0};
    RefMesh* r = new RefMesh;
    r->mesh = new Mesh("mesh", ne, "legendre", N, mn, mx, per);
    r->bcs = new BCs(bc[0] != 0, bc[1] != 0, bc[2] != 0, bc[3] != 0, r->mesh, 0);
    return r;
}

void ref_mesh_destroy(void* h) {
    RefMesh* r = (RefMesh*)h;
    delete r->bcs;
    delete r->mesh;
    delete r;
}

void ref_mesh_sizes(void* h, int* out) {
    RefMesh* r = (RefMesh*)h;
    out[0] = r->mesh->nVerts[0];
    out[1] = r->mesh->nVerts[1];
    out[2] = r->mesh->nVertsTotal;
    out[3] = r->mesh->nElsTotal;
    out[4] = r->mesh->nVertsTotal - r->bcs->size[0];
    out[5] = r->bcs->size[0];
    out[6] = r->mesh->N;
}

void ref_mesh_verts(void* h, double* verts) {
    RefMesh* r = (RefMesh*)h;
    for (int v = 0; v < r->mesh->nVertsTotal; v++) {
        verts[2 * (size_t)v] = r->mesh->verts[v][0];
        verts[2 * (size_t)v + 1] = r->mesh->verts[v][1];
    }
}

void ref_mesh_maps(void* h, int* f2v, int* v2f) {
    RefMesh* r = (RefMesh*)h;
    int n = r->mesh->nVertsTotal - r->bcs->size[0];
    if (f2v) std::memcpy(f2v, r->bcs->fieldToVecMap, sizeof(int) * (size_t)r->mesh->nVertsTotal);
    if (v2f) std::memcpy(v2f, r->bcs->vecToFieldMap, sizeof(int) * (size_t)n);
}

void ref_mesh_gll(void* h, double* abcissa, double* weights, double* dCij) {
    RefMesh* r = (RefMesh*)h;
    Legendre* l = (Legendre*)r->mesh->el;
    int N = r->mesh->N;
    if (abcissa) std::memcpy(abcissa, l->abcissa, sizeof(double) * (size_t)(N + 1));
    if (weights) std::memcpy(weights, l->weights, sizeof(double) * (size_t)(N + 1));
    if (dCij) std::memcpy(dCij, l->dCij, sizeof(double) * (size_t)(N + 1) * (N + 1));
}

void ref_el_nodes(void* h, int el, int* nodes) {
    RefMesh* r = (RefMesh*)h;
    int* en = r->mesh->ElNodes(el);
    std::memcpy(nodes, en, sizeof(int) * (size_t)r->mesh->el->nNodes);
}

void ref_get_el(void* h, const double* coord, int* el) {
    RefMesh* r = (RefMesh*)h;
    double c[2] = {coord[0], coord[1]};
    r->mesh->GetEl(c, el);
}

void ref_global_to_local(void* h, const double* g, int* el, double* l) {
    RefMesh* r = (RefMesh*)h;
    double c[2] = {g[0], g[1]};
    r->mesh->GlobalToLocal(c, el, l);
}

void ref_shape_funcs_at_coord(void* h, const double* l, double* Ni) {
    RefMesh* r = (RefMesh*)h;
    double c[2] = {l[0], l[1]};
    double* ni = r->mesh->el->ShapeFuncsAtCoord(c);
    std::memcpy(Ni, ni, sizeof(double) * (size_t)r->mesh->el->nNodes);
}

void ref_interp_global(void* h, const double* vals, int nDofs, const double* g, double* out) {
    RefMesh* r = (RefMesh*)h;
    Field f("f", r->mesh, nDofs, r->bcs);
    load_field(&f, vals);
    double c[3] = {g[0], g[1], 0.0};
    f.InterpGlobal(c, out);
}

double ref_det_jac(void* h, int el, int pt) {
    RefMesh* r = (RefMesh*)h;
    return r->mesh->DetJac(el, pt);
}

void ref_shape_func_derivs(void* h, int el, int pt, double* gGNix, double* detJac) {
    RefMesh* r = (RefMesh*)h;
    double** g = r->mesh->ShapeFuncDerivs(el, pt, detJac);
    int nl = r->mesh->el->nNodes;
    std::memcpy(gGNix, g[0], sizeof(double) * (size_t)nl);
    std::memcpy(gGNix + nl, g[1], sizeof(double) * (size_t)nl);
}

void ref_elem_laplacian(void* h, int el, double c, double* K) {
    RefMesh* r = (RefMesh*)h;
    Field f("f", r->mesh, 1, r->bcs);
    Laplacian op("lap", &f, &f, c);
    op.AssembleElement(el, K);
}

void ref_elem_mass(void* h, int el, double c, double* M) {
    RefMesh* r = (RefMesh*)h;
    Field f("f", r->mesh, 1, r->bcs);
    MassMatrix op("mass", &f, &f, c);
    op.AssembleElement(el, M);
}

void ref_elem_field_rhs(void* h, int el, double c, const double* field, double* rhs) {
    RefMesh* r = (RefMesh*)h;
    Field f("f", r->mesh, 1, NULL);
    load_field(&f, field);
    FieldRHS op("rhs", r->mesh, c, &f);
    op.AssembleElement(el, rhs);
}

void ref_field_periodic_update(void* h, double* vals, int nDofs) {
    RefMesh* r = (RefMesh*)h;
    Field f("f", r->mesh, nDofs, r->bcs);
    load_field(&f, vals);
    f.PeriodicUpdate();
    store_field(&f, vals);
}

// Advector::FirstOrder through Advector::Advect; the departure points come from
// Advector::IntegrateRK2 and Mesh::GetEl called node by node.
void ref_advect_first(void* h, const double* q, int nDofs, const double* vel, double dt,
                      double* qSL, int* dep_el, double* dep_xy) {
    RefMesh* r = (RefMesh*)h;
    Quiet quiet;
    Field fq("q", r->mesh, nDofs, r->bcs);
    Field fv("vel", r->mesh, 2, NULL);
    load_field(&fq, q);
    load_field(&fv, vel);
    Advector adv(&fq, &fv);
    adv.Advect(dt);
    store_field(adv.fieldSL, qSL);
    if (dep_el || dep_xy) {
        for (int v = 0; v < r->mesh->nVertsTotal; v++) {
            double c[3];
            adv.IntegrateRK2(dt, &fv, r->mesh->verts[v], c);
            if (dep_el) r->mesh->GetEl(c, &dep_el[v]);
            if (dep_xy) {
                dep_xy[2 * (size_t)v] = c[0];
                dep_xy[2 * (size_t)v + 1] = c[1];
            }
        }
    }
}

void ref_advect_second(void* h, const double* q, const double* qPrev, int nDofs, const double* vel,
                       const double* velPrev, double dt, double* qSL) {
    RefMesh* r = (RefMesh*)h;
    Quiet quiet;
    Field fq("q", r->mesh, nDofs, r->bcs);
    Field fqp("qp", r->mesh, nDofs, r->bcs);
    Field fv("vel", r->mesh, 2, NULL);
    Field fvp("velp", r->mesh, 2, NULL);
    load_field(&fq, q);
    load_field(&fqp, qPrev);
    load_field(&fv, vel);
    load_field(&fvp, velPrev);
    Advector adv(&fq, &fv, &fqp, &fvp);
    adv.Advect(dt);
    store_field(adv.fieldSL, qSL);
}

// QuasiGeostrophicEqn::GenVel (src/qg, not buildable without PETSc) is
// Field::InterpDerivsGlobal at every vertex followed by u = -psi_y, v = psi_x
// (QuasiGeostrophicEqn.cpp:213-217); the derivative call is the reference's.
void ref_genvel(void* h, const double* psi, double* vel) {
    RefMesh* r = (RefMesh*)h;
    Field f("psi", r->mesh, 1, r->bcs);
    load_field(&f, psi);
    double g[2];
    double* dp[1] = {g};
    for (int v = 0; v < r->mesh->nVertsTotal; v++) {
        f.InterpDerivsGlobal(r->mesh->verts[v], dp);
        vel[2 * (size_t)v] = -1.0 * dp[0][1];
        vel[2 * (size_t)v + 1] = +1.0 * dp[0][0];
    }
}

}  // extern "C"
