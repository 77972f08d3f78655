/*
 * se_oracle.h -- CPU oracle for the SE spectral-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms (davelee2804/SE, src/base + src/linsys + src/qg) used as the
 * checker for the HIP product path.  Only tests/, the smoke() entry and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Every function cites the reference file:line it restates.  Functions whose
 * arithmetic has no independent pin (golden vectors from the compiled
 * reference, closed forms, brute force) are marked "parity unpinned".
 *
 * Index conventions follow the reference exactly:
 *   field node  v = J*nVerts[0] + I          (Mesh::IndexToTopo, Mesh.cpp:154-157)
 *   element     e = ey*nEls[0] + ex           (Mesh::GetEl, Mesh.cpp:136)
 *   local node  l = jj*(N+1) + ii             (Mesh::ElNodes, Mesh.cpp:138-152)
 *   dCij[i*(N+1)+j] = l_j'(x_i)               (Legendre.cpp:86-95)
 */
#ifndef SE_ORACLE_H
#define SE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- a1: GLL basis (Jacobi.cpp:21-76, Element.cpp:54-106, Legendre.cpp:15-95) ---- */
int    orc_gll(int N, double* abcissa, double* weights, double* dCij, double* pn);
double orc_leg_eval(const double* c, int n, double x);
double orc_leg_eval_deriv(const double* c, int n, double x);
double orc_leg_eval_2nd(const double* c, int n, double x);

/* ---- a2/a3: mesh + numbering (Mesh.cpp:17-181, BCs.cpp:14-173) ---- */
typedef struct orc_mesh orc_mesh;
/* bc = {bottom, top, left, right}; status: 0 ok, <0 reference exit(1) path */
orc_mesh* orc_mesh_create(const int* nEls, int N, const double* min, const double* max,
                          const int* periodic, const int* bc, int* status);
void   orc_mesh_destroy(orc_mesh* m);
/* out = {nVerts0, nVerts1, nV, nElsTotal, n, size0, N} */
void   orc_mesh_sizes(const orc_mesh* m, int* out);
void   orc_mesh_verts(const orc_mesh* m, double* verts);
void   orc_mesh_maps(const orc_mesh* m, int* f2v, int* v2f);
void   orc_mesh_gll(const orc_mesh* m, double* abcissa, double* weights, double* dCij);
void   orc_el_nodes(const orc_mesh* m, int el, int* nodes);
void   orc_get_el(const orc_mesh* m, const double* coord, int* el);
void   orc_global_to_local(const orc_mesh* m, const double* g, int* el, double* l);
void   orc_shape_funcs_at_coord(const orc_mesh* m, const double* l, double* Ni);
void   orc_interp_global(const orc_mesh* m, const double* vals, int nDofs, const double* g, double* out);
double orc_det_jac(const orc_mesh* m, int el, int pt);
void   orc_shape_func_derivs(const orc_mesh* m, int el, int pt, double* gGNix, double* detJac);

/* ---- a4-a7: element operators (Operator.cpp:63-84,157-177; RHSOp.cpp:31-50) ---- */
/* detj_mode: 0 = literal (detJac := 1.0, Operator.cpp:72), 1 = physical */
void orc_elem_laplacian(const orc_mesh* m, int el, double c, int detj_mode, double* K);
void orc_elem_mass(const orc_mesh* m, int el, double c, double* M);
void orc_elem_field_rhs(const orc_mesh* m, int el, double c, const double* field, double* rhs);
/* SVV::AssembleElement (SVV.cpp:120-176): adds c (Lxx + Lyy) into A (nl x nl) */
void orc_elem_svv(const orc_mesh* m, int el, double c, int mN, double* A);

/* ---- a8-a10: assembled operator + PCG (Matrix.cpp:47-184, Vector.cpp:40-121,
 *      HelmholtzEqn.cpp:21-136, PoissonEqn.cpp:30-134) ---- */
typedef struct orc_helm orc_helm;
/* A = c_mass*M + c_lap*K~ ; c_mass == 0 reproduces PoissonEqn (Laplacian only). */
orc_helm* orc_helm_create(const orc_mesh* m, double c_mass, double c_lap, int detj_mode);
void orc_helm_destroy(orc_helm* h);
/* the SVV operator of HelmholtzEqn(..., dt, svvCutoff); before assembly */
void orc_helm_set_svv(orc_helm* h, double dt, int mN);
void orc_helm_assemble(orc_helm* h, const double* sol_field);
long orc_helm_nnz(const orc_helm* h);
void orc_helm_csr(const orc_helm* h, long* rowptr, int* cols, double* vals);
void orc_helm_lift(const orc_helm* h, double* b);
void orc_helm_inv_diag(const orc_helm* h, double* dinv);
void orc_helm_matvec(const orc_helm* h, const double* x, double* y);
void orc_helm_rhs(const orc_helm* h, const double* rhs_field, double* f);
/* status: 0 converged, 1 max iterations, 2 breakdown/indefinite, 3 non-finite */
int  orc_helm_pcg(const orc_helm* h, const double* f, double* x, double rtol, int maxit,
                  int null_space, int* iters, double* true_relres, double* hist);
int  orc_helm_solve(orc_helm* h, const double* rhs_field, double* sol_field, double rtol,
                    int maxit, int null_space, int* iters, double* true_relres);
void orc_update_field(const orc_mesh* m, const double* x, double* field);
void orc_update_vec(const orc_mesh* m, const double* field, double* x);
void orc_field_periodic_update(const orc_mesh* m, double* field, int nDofs);

/* matrix-free restatement of the same operator with ideal (dx/2, dy/2) geometry,
 * used as the checker where the CSR is too large */
void orc_helm_apply_ideal(const orc_mesh* m, double c_mass, double c_lap, int detj_mode,
                          const double* x, double* y);

/* ---- a11-a16: semi-Lagrangian step (Advector.cpp, Field.cpp:57-118, Utils.cpp:23-40) ---- */
void   orc_interp_bicubic(const orc_mesh* m, const double* vals, int nDofs, const double* coord, double* psi);
void   orc_advect_first_interp(const orc_mesh* m, const double* q, int nDofs, const double* vel, double dt,
                               double* qSL, int interp);
void   orc_advect_first(const orc_mesh* m, const double* q, int nDofs, const double* vel,
                        double dt, double* qSL, int* dep_el, double* dep_xy);
void   orc_advect_second(const orc_mesh* m, const double* q, const double* qPrev, int nDofs,
                         const double* vel, const double* velPrev, double dt, double* qSL);
void   orc_sl_points(const orc_mesh* m, int order, const double* vel, const double* velPrev, double dt,
                     double* pts);
void   orc_genvel(const orc_mesh* m, const double* psi, double* vel);
double orc_calc_timestep(const orc_mesh* m, const double* vel, double courant);

/* ---- callers: QG steps (QuasiGeostrophicEqn.cpp:83-337;
 *      TwoLayerQGEqn.cpp:174-462).  NULL previous state: the first-order
 *      step; otherwise the second-order step from (prev, current). ---- */
/* SolveQ + DisRHS / DissipationRHS: alpha's unknowns from the a x mass system
 * (Dirichlet nodes keep alpha's values); tau, omegaHat, tau_e nullable */
void orc_qg_solveq(const orc_mesh* m, double a, double dt, const double* qhat, double rF, double Ei, double nu,
                   const double* tau, const double* omegaHat, const double* tau_e, double* alpha);
int orc_qg_step(const orc_mesh* m, double F, double beta, double nu, double r, const double* topo, double dt,
                int detj_mode, double* psi, double* omega, const double* psi_prev, const double* omega_prev,
                double rtol, int maxit, int* iters);
int orc_qg2_step(const orc_mesh* m, double F1, double F2, double beta, double E2, double r, double nu,
                 const double* tau_e, double dt, int detj_mode, double* psi1, double* omega1, double* psi2,
                 double* omega2, const double* psi1p, const double* omega1p, const double* psi2p,
                 const double* omega2p, double rtol, int maxit, int* iters);

/* threads of the oracle's parallel loops (CSR matvec, vector updates,
 * fixed-block dot products: results do not depend on the count) */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
