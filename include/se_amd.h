/*
 * se_amd.h -- C ABI of the MI355X (gfx950) native spectral-element hot path.
 *
 * Drop-in boundary for davelee2804/SE's matrix-free Helmholtz/Poisson CG solve
 * and semi-Lagrangian (SL) advection step.  Every entry point states the
 * reference interface it replaces as  [ref: file:line].  Plain C types only:
 * pointers, sizes, int status codes.  A context owns one HIP stream and all
 * device memory it allocates; a context is used by one host thread at a time.
 * Drive one device from one host thread: FDM batch solves on two contexts of
 * one device from two threads at once ran slower than one batch, and in
 * three of four trials some large members diverged (cause not found).
 *
 * Pointer conventions
 *   *_dev entry points take device pointers that are already resident in HBM
 *   and enqueue work on the context stream; plain entry points take host
 *   pointers, stage through the context and return after completion.
 *   Field arrays are node-major: field node v = J*nVerts[0] + I holds
 *   nDofs consecutive doubles (the reference's Field::vals[v][dof]).
 *   Vectors of unknowns use the reference's numbering (BCs::fieldToVecMap).
 *
 * Status codes: 0 = success, negative = error (text in se_last_error()).
 */
#ifndef SE_AMD_H
#define SE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SE_OK = 0,
    SE_ERR_ARG = -1,         /* invalid argument */
    SE_ERR_HIP = -2,         /* HIP runtime failure */
    SE_ERR_UNSUPPORTED = -3, /* configuration the reference rejects or leaves undefined */
    SE_ERR_NOMEM = -4,       /* allocation failed / over a memory budget */
    SE_ERR_STATE = -5        /* call out of order */
};

/* Laplacian Jacobian convention.  LITERAL reproduces the reference, which
 * overwrites detJ with 1 in Laplacian::AssembleElement [ref:
 * src/base/Operator.cpp:72]; PHYSICAL keeps the true detJ. */
enum { SE_DETJ_LITERAL = 0, SE_DETJ_PHYSICAL = 1 };

/* Solve precision.  FP64: everything in FP64.  FP32_IR / BF16_IR: the
 * low-precision work (the inner Jacobi-PCG iteration, or the
 * fast-diagonalisation preconditioner's transforms) in FP32, resp. with BF16
 * operands (Jacobi: BF16 search direction, FP32 VALU arithmetic; FDM: BF16
 * inputs to the matrix cores with FP32 accumulation), inside an FP64 loop
 * whose residual, norms and update are FP64.  FP64_FP32C / FP64_BF16C (FDM
 * preconditioner only): FP64 CG whose preconditioner's first application --
 * the one that sets the residual's floor -- runs in FP64 and later ones,
 * corrections of the residual it left, in FP32, resp. with BF16 operands
 * (which, like BF16_IR, may stagnate on a near-singular system). */
enum { SE_PREC_FP64 = 0, SE_PREC_FP32_IR = 1, SE_PREC_BF16_IR = 2, SE_PREC_FP64_FP32C = 3, SE_PREC_FP64_BF16C = 4 };

/* Preconditioner.  JACOBI: the reference's (PETSc -pc_type jacobi; FP64) or
 * the inner Jacobi-PCG with FP64 reliable updates (FP32_IR / BF16_IR).
 * FDM: fast diagonalisation of the operator's separable form (dense 1-D
 * eigenvector transforms on the matrix cores, in the solve's precision)
 * inside FP64 CG -- a near-exact inverse on the reference's uniform meshes,
 * so the CG needs a handful of iterations.  DEFAULT: JACOBI for FP64, FDM for
 * FP32_IR / BF16_IR. */
enum { SE_PC_DEFAULT = 0, SE_PC_JACOBI = 1, SE_PC_FDM = 2 };

/* Solve status */
enum {
    SE_SOLVE_CONVERGED = 0, /* true ||b - A x||_2 <= rtol ||b||_2 */
    SE_SOLVE_MAXIT = 1,     /* iteration limit reached */
    SE_SOLVE_BREAKDOWN = 2, /* p.Ap <= 0: operator or preconditioner indefinite */
    SE_SOLVE_NONFINITE = 3, /* NaN / Inf in the data */
    SE_SOLVE_STAGNATED = 4  /* mixed precision only: the inner precision cannot
                               resolve this system (refinement stopped making
                               progress); solve in FP64 or FP32_IR instead */
};

typedef struct se_ctx se_ctx;
typedef struct se_mesh se_mesh;
typedef struct se_op se_op;

const char* se_last_error(void);
int se_version(void);

/* ---- context -------------------------------------------------------------- */
/* [ref: none -- the reference is a serial CPU code; PetscInitialize is the
 *  closest analogue, models/qg-channel/BarotropicChannel.cpp:74] */
int se_ctx_create(int device, se_ctx** out);
int se_ctx_destroy(se_ctx* ctx);
int se_ctx_synchronize(se_ctx* ctx);
void* se_ctx_stream(se_ctx* ctx); /* the context's hipStream_t */
/* Average time (us) of one plain device copy of `bytes` (src and dst of that
 * size, allocated and freed by the call), over `reps` launches.  No
 * counterpart in the reference: the yardstick for the HBM byte counters and
 * the attainable copy bandwidth in the measurements. */
int se_ctx_copy_probe(se_ctx* ctx, size_t bytes, int reps, double* us);

/* ---- mesh and global numbering --------------------------------------------
 * [ref: Mesh::Mesh src/base/Mesh.cpp:17-84 with elType "legendre";
 *       BCs::BCs (scalar) src/base/BCs.cpp:14-173]
 * dirichlet = {bottom, top, left, right}.  Builds vertex coordinates, the
 * field<->unknown maps and the exact per-element Jacobian tables on device. */
int se_mesh_create(se_ctx* ctx, const int nEls[2], int N, const double min[2], const double max[2],
                   const int periodic[2], const int dirichlet[4], se_mesh** out);
int se_mesh_destroy(se_mesh* mesh);

typedef struct {
    int N;
    int nEls[2];
    int nVerts[2];
    int nV;        /* field nodes incl. periodic duplicates (Mesh::nVertsTotal) */
    int nElsTotal;
    int n;         /* unknowns (vecSize) */
    int size0;     /* BCs::size[0] */
} se_mesh_info;
int se_mesh_get_info(const se_mesh* mesh, se_mesh_info* out);
/* [ref: Legendre::Legendre src/base/Legendre.cpp:15-95] abcissa/weights N+1, dCij (N+1)^2 */
int se_mesh_gll(const se_mesh* mesh, double* abcissa, double* weights, double* dCij);
/* [ref: BCs::fieldToVecMap / vecToFieldMap src/base/BCs.h:8-9, Mesh::ElNodes
 *  src/base/Mesh.cpp:138-152]  host outputs (nV, n, nElsTotal*(N+1)^2); any may be NULL */
int se_mesh_maps(const se_mesh* mesh, int32_t* fieldToVec, int32_t* vecToField, int32_t* elNodes);
/* [ref: Mesh::verts src/base/Mesh.cpp:63-73]  host output nV x 2 */
int se_mesh_verts(const se_mesh* mesh, double* verts);
/* device pointers to the resident maps / coordinates (owned by the mesh) */
const int32_t* se_mesh_dev_field_to_vec(const se_mesh* mesh);
const double* se_mesh_dev_verts(const se_mesh* mesh);

/* ---- field <-> unknown vector ---------------------------------------------
 * [ref: Vector::UpdateVec src/base/Vector.cpp:63-81] */
int se_field_to_vec_dev(se_mesh* mesh, const double* d_field, double* d_vec);
/* [ref: Vector::UpdateField src/base/Vector.cpp:40-61 + Field::PeriodicUpdate
 *  src/base/Field.cpp:232-252]  Dirichlet nodes of d_field are left untouched. */
int se_vec_to_field_dev(se_mesh* mesh, const double* d_vec, double* d_field);

/* ---- Helmholtz / Poisson operator -----------------------------------------
 * A = c_mass*M + c_lap*K~ on the unknowns, matrix-free.
 * [ref: HelmholtzEqn::HelmholtzEqn src/linsys/HelmholtzEqn.cpp:21-82 with
 *       c_mass = c, c_lap = nu*dt;  PoissonEqn::PoissonEqn
 *       src/linsys/PoissonEqn.cpp:30-80 with c_mass = 0, c_lap = c;
 *       element blocks Laplacian::AssembleElement Operator.cpp:63-84,
 *       MassMatrix::AssembleElement Operator.cpp:157-177] */
int se_helm_create(se_mesh* mesh, double c_mass, double c_lap, int detj_mode, se_op** out);
int se_helm_destroy(se_op* op);
/* adds the spectral vanishing viscosity term of HelmholtzEqn(..., dt,
 * svvCutoff): A += dt (Lxx + Lyy) per element, the modal filter of the modes
 * above `cutoff` (1 .. N-1) in both directions, Lyy built from the x terms as
 * the reference builds it [ref: HelmholtzEqn.cpp:71-76, SVV::AssembleElement
 * src/base/SVV.cpp:120-176].  Before set_dirichlet; once per operator.  Such
 * operators solve with FP64 Jacobi-PCG, or with SE_PC_FDM in any precision,
 * alone or batched (the FDM is that of the operator without the SVV term; the
 * CG applies the exact operator); the mixed-precision Jacobi-PCG refuses
 * them. */
int se_helm_set_svv(se_op* op, double dt, int cutoff);
/* y = A x   [ref: MatMult on the matrix built by Matrix::Assemble
 *            src/base/Matrix.cpp:47-133] */
int se_helm_apply(se_op* op, const double* x, double* y);
int se_helm_apply_dev(se_op* op, const double* d_x, double* d_y);
/* [ref: Matrix::InverseDiagonal src/base/Matrix.cpp:162-184]  host output n */
int se_helm_inv_diag(se_op* op, double* dinv);
/* Dirichlet lifting from the solution field's boundary values
 * [ref: Matrix::Assemble lifting src/base/Matrix.cpp:101-124, run by
 *  HelmholtzEqn::Assemble src/linsys/HelmholtzEqn.cpp:91-103]; host nV in */
int se_helm_set_dirichlet(se_op* op, const double* sol_field);
int se_helm_set_dirichlet_dev(se_op* op, const double* d_sol_field);
int se_helm_lift(se_op* op, double* b); /* host output n */
/* f = M*rhs_field + b  [ref: Vector::Assemble of FieldRHS src/base/Vector.cpp:89-121,
 *  RHSOp.cpp:31-50, then VecAXPY HelmholtzEqn.cpp:111-118] */
int se_helm_rhs(se_op* op, const double* rhs_field, double* f);
int se_helm_rhs_dev(se_op* op, const double* d_rhs_field, double* d_f);

typedef struct {
    double rtol;        /* relative true-residual target (default 1e-10) */
    int maxit;          /* iteration limit, inner iterations counted (default 100000) */
    int precision;      /* SE_PREC_* (default FP64) */
    int null_space;     /* 1: constant null space (periodic Poisson) */
    int check_every;    /* iterations per graph replay (default 32).  FP64: a host convergence
                           check after every replay.  Mixed precision (single and batched):
                           after every replay for the first 16 checks, then after every 2nd,
                           from the 32nd check on after every 4th */
    double inner_rtol;  /* mixed solves: reliable update once the recursive residual
                           fell by this factor since the last one (0 = 0.1) */
    int max_outer;      /* limit on FP64 reliable updates of a mixed solve (default 1000) */
    double* history;    /* optional host array, maxit+1 recursive relative residuals (FP64 only) */
    int preconditioner; /* SE_PC_* (default SE_PC_DEFAULT) */
} se_solve_opts;

typedef struct {
    int status;               /* SE_SOLVE_* */
    int iters;                /* CG iterations (inner iterations for IR) */
    int outer_iters;          /* IR outer steps (1 for FP64) */
    double true_rel_residual; /* FP64 ||b - A x|| / ||b|| at exit (an FDM solve that
                                 converges at its first iteration from x = 0: its
                                 update's b - alpha A p, for x = alpha p) */
    double seconds;           /* device time of the solve */
} se_solve_info;

void se_solve_opts_default(se_solve_opts* opts);
/* Inner operator of the mixed-precision solves (no counterpart in the
 * reference, which solves in FP64 only): y = sign*A~ x with A~ the separable
 * (Kronecker-sum) form of A without its round-off-sized Jacobian cross terms,
 * FP32 band tables; the input is rounded to FP32 (SE_PREC_FP32_IR) or BF16
 * (SE_PREC_BF16_IR), arithmetic and output are FP32.  sign is +-1 so that
 * sign*A has a positive diagonal.  Host arrays of n values. */
int se_helm_apply_inner(se_op* op, int precision, const float* x, float* y, int* sign);
/* Fast-diagonalisation preconditioner (no counterpart in the reference,
 * whose PETSc solve takes -pc_type jacobi): z ~= (sign*A)^-1 r through the
 * eigenvectors of the operator's separable form, the four transforms on the
 * matrix cores in `precision` (SE_PREC_FP64 / FP32_IR / BF16_IR); the null
 * mode of a singular operator is left out of z.  Host arrays of n values.
 * Uses the operator's solver work vectors. */
int se_helm_apply_fdm(se_op* op, int precision, const double* r, double* z);
/* Measurement hook (no counterpart in the reference): after an FDM solve of
 * op, times `reps` back-to-back launches of each of the four transform
 * passes (the first with its residual pack) and of the FP64 operator apply,
 * with HIP events on the context stream: us_pass[4], flops_pass[4] (2 M N K
 * per member, padded sizes), us_apply, bytes_apply (16 B per unknown, 4 B
 * per field node). */
int se_helm_fdm_probe(se_op* op, int reps, double* us_pass, double* flops_pass, double* us_apply,
                      double* bytes_apply);
/* Measurement hook (no counterpart in the reference): after a mixed-precision
 * solve of op, times `reps` back-to-back launches of the inner iteration's
 * direction+operator kernel and update kernel with HIP events on the context
 * stream; outputs microseconds and algorithmic bytes per launch.  The
 * operator's mixed solver state is left as the probe's. */
int se_helm_probe(se_op* op, int precision, int reps, double* us_dir, double* us_upd, double* bytes_dir,
                  double* bytes_upd);
/* Jacobi-PCG on A x = f, x0 = 0  [ref: KSPSolve with -ksp_type cg -pc_type
 *  jacobi, src/linsys/HelmholtzEqn.cpp:120-131] (sign-normalised: a
 *  negative-definite A is solved as -A x = -f) */
int se_helm_pcg_dev(se_op* op, const double* d_f, double* d_x, const se_solve_opts* opts, se_solve_info* info);
int se_helm_pcg(se_op* op, const double* f, double* x, const se_solve_opts* opts, se_solve_info* info);
/* full HelmholtzEqn::Solve: f = M rhs + b, solve, UpdateField into sol_field
 * [ref: HelmholtzEqn::Solve src/linsys/HelmholtzEqn.cpp:105-136,
 *       PoissonEqn::Solve src/linsys/PoissonEqn.cpp:102-134] */
int se_helm_solve(se_op* op, const double* rhs_field, double* sol_field, const se_solve_opts* opts,
                  se_solve_info* info);
int se_helm_solve_dev(se_op* op, const double* d_rhs_field, double* d_sol_field, const se_solve_opts* opts,
                      se_solve_info* info);

/* ---- batched solves ----------------------------------------------------------
 * Many independent solves at once: mixed precision (SE_PREC_FP32_IR /
 * SE_PREC_BF16_IR) with either preconditioner, or SE_PREC_FP64 with the FDM
 * preconditioner.  No counterpart in the reference, which solves one
 * HelmholtzEqn at a time (KSPSolve, src/linsys/HelmholtzEqn.cpp:120-131); each
 * member's result is bitwise the one se_helm_pcg_dev gives for it alone with
 * the same options.
 * Members are operators created on the same context (any size, order 2..8 and
 * boundary conditions); an operator may belong to one batch at a time. */
typedef struct se_batch se_batch;
int se_batch_create(se_ctx* ctx, int count, se_op* const* ops, se_batch** out);
int se_batch_destroy(se_batch* batch);
int se_batch_size(const se_batch* batch);
/* The launch shapes the autotuner chose on the first solve for the
 * direction and the update kernel: workgroups per CU, 0 for one workgroup
 * per (member, tile) item, -1 before the first solve.  No counterpart in the
 * reference. */
int se_batch_launch_shape(const se_batch* batch, int* dir_per_cu, int* update_per_cu);
/* d_f[i], d_x[i]: device vectors of member i; infos: count entries (nullable) */
int se_batch_pcg_dev(se_batch* batch, const double* const* d_f, double* const* d_x, const se_solve_opts* opts,
                     se_solve_info* infos);
int se_batch_pcg(se_batch* batch, const double* const* f, double* const* x, const se_solve_opts* opts,
                 se_solve_info* infos);
/* Measurement hook (no counterpart in the reference): after a solve of this
 * batch, times `reps` back-to-back launches of each group's direction+operator
 * kernel and update kernel over all members, with HIP events on the context
 * stream.  Per order group (at most max_groups): the order, microseconds per
 * launch and algorithmic bytes per launch of each kernel.  Returns the number
 * of groups (>= 0) or a negative SE_ERR_* code; the members' solver states are
 * left as the probe's (the next solve starts afresh). */
int se_batch_probe(se_batch* batch, int precision, int reps, int max_groups, int* orders, double* us_dir,
                   double* us_upd, double* bytes_dir, double* bytes_upd);
/* The FDM counterpart of se_batch_probe, over all members at once (as
 * se_helm_fdm_probe); needs an FDM solve of the batch first. */
int se_batch_fdm_probe(se_batch* batch, int reps, double* us_pass, double* flops_pass, double* us_apply,
                       double* bytes_apply);

/* ---- semi-Lagrangian advection ---------------------------------------------
 * [ref: Advector::Advect / FirstOrder src/base/Advector.cpp:108-120,181-201,
 *  IntegrateRK2 :136-148, PeriodicUpdate :122-134, Mesh::GetEl Mesh.cpp:106-137,
 *  Field::InterpGlobal Field.cpp:57-76, Legendre::ShapeFuncsAtCoord
 *  Legendre.cpp:145-169, bicubic Interpolate (N < 3) Advector.cpp:248-337]
 * q, qSL: nV x nDofs (nDofs 1..3); vel: nV x 2; dep_el (nullable): nV. */
int se_sl_advect(se_mesh* mesh, const double* q, int nDofs, const double* vel, double dt, double* qSL,
                 int32_t* dep_el);
int se_sl_advect_dev(se_mesh* mesh, const double* d_q, int nDofs, const double* d_vel, double dt,
                     double* d_qSL, int32_t* d_dep_el);
/* Interpolation of the advected field.  DEFAULT: the reference's choice,
 * bicubic Lagrange on the 4 x 4 node stencil when N < 3, GLL cardinal
 * functions otherwise (Advector's constructors set cubicInterp = N < 3,
 * Advector.cpp:33,60); GLL / BICUBIC: as a caller that sets cubicInterp
 * itself (QuasiGeostrophicSpectral.cpp:84).  The mesh keeps the mode for its
 * SL steps (se_sl_advect*), velocities are always GLL-interpolated. */
enum { SE_INTERP_DEFAULT = 0, SE_INTERP_GLL = 1, SE_INTERP_BICUBIC = 2 };
int se_mesh_set_sl_interp(se_mesh* mesh, int mode);
/* Values of a field (nV x nDofs, nDofs 1..3) at npts points (xy: npts x 2)
 * [ref: Field::InterpGlobal Field.cpp:57-76 (GLL), Advector::Interpolate
 *  Advector.cpp:278-337 (BICUBIC)]; out: npts x nDofs */
int se_interp_points(se_mesh* mesh, const double* vals, int nDofs, int npts, const double* xy, int mode,
                     double* out);
int se_interp_points_dev(se_mesh* mesh, const double* d_vals, int nDofs, int npts, const double* d_xy, int mode,
                         double* d_out);
/* [ref: Advector::SecondOrder + ExtrapolateVelocities Advector.cpp:203-246,339-349] */
int se_sl_advect2(se_mesh* mesh, const double* q, const double* qPrev, int nDofs, const double* vel,
                  const double* velPrev, double dt, double* qSL);
int se_sl_advect2_dev(se_mesh* mesh, const double* d_q, const double* d_qPrev, int nDofs, const double* d_vel,
                      const double* d_velPrev, double dt, double* d_qSL);
/* se_sl_advect2_dev without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The extrapolated velocities go to a fixed
 * per-GPU scratch array instead of a buffer the library allocates, so meshes
 * are limited to SE_SL_NOALLOC_MAX_NODES field nodes (nV; SE_ERR_ARG above
 * it).  The result equals se_sl_advect2_dev's bit for bit.  Every call on one
 * GPU shares that scratch: calls that overlap on one GPU (different streams
 * or threads) race. */
#define SE_SL_NOALLOC_MAX_NODES 524288
int se_sl_advect2_dev_noalloc(se_mesh* mesh, const double* d_q, const double* d_qPrev, int nDofs,
                              const double* d_vel, const double* d_velPrev, double dt, double* d_qSL);
/* u = -dpsi/dy, v = dpsi/dx at every vertex  [ref: QuasiGeostrophicEqn::GenVel
 *  src/qg/QuasiGeostrophicEqn.cpp:207-223 via Field::InterpDerivsGlobal
 *  Field.cpp:96-118, Mesh::ShapeFuncDerivsAtCoord Mesh.cpp:252-285] */
int se_genvel(se_mesh* mesh, const double* psi, double* vel);
int se_genvel_dev(se_mesh* mesh, const double* d_psi, double* d_vel);
/* [ref: CalcTimeStep src/base/Utils.cpp:23-40] */
int se_calc_timestep(se_mesh* mesh, const double* vel, double courant, double* dt);
int se_calc_timestep_dev(se_mesh* mesh, const double* d_vel, double courant, double* dt);

/* ---- quasi-geostrophic steps ---------------------------------------------------
 * Time steps of the QG models.  The first step after set_state is the
 * reference's first-order step, every later one its second-order step from the
 * state before the previous step (as the reference's drivers call Solve).
 * SolveQ's PV update runs on the diagonal GLL mass matrix: each unknown is its
 * assembled right-hand side over its mass (the two copies of a periodic node
 * are averaged), Dirichlet (wall) nodes keep their GenPV value, so the wall
 * vorticity is kept.  With dissipation set, the step first inverts the
 * advected PV (the reference's first inversion, or pair) for the omega^ and
 * tau^ its dissipation terms use; without, that inversion only feeds terms
 * multiplied by zero and is skipped.  The inversions run with the precision
 * of `opts` (FP64 or mixed). */

/* Two layers [ref: TwoLayerQGEqn::FirstOrder / SecondOrder
 * src/qg/TwoLayerQGEqn.cpp:174-333, GenVel :335-352]: GenVel and GenPV of both
 * layers, semi-Lagrangian advection of both PVs, baroclinic Helmholtz
 * (lap - (F1+F2)) tau = a1 - a2, barotropic Poisson
 * lap chi = a1 + a2 + (F1-F2) tau - 2 beta y, then psi/omega of both layers.
 * As in the reference, the first-order step updates layer 1's PV only
 * (SolveQ for layer 1, TwoLayerQGEqn.cpp:214). */
typedef struct se_qg2 se_qg2;
int se_qg2_create(se_mesh* mesh, double F1, double F2, double beta, int detj_mode, se_qg2** out);
int se_qg2_destroy(se_qg2* q);
/* host fields, nV each; the next step is first order */
int se_qg2_set_state(se_qg2* q, const double* psi1, const double* omega1, const double* psi2, const double* omega2);
int se_qg2_get_state(se_qg2* q, double* psi1, double* omega1, double* psi2, double* omega2);
/* device addresses of the state fields (nullable outputs) */
int se_qg2_state_dev(se_qg2* q, double** psi1, double** omega1, double** psi2, double** omega2);
/* CFL time step of both layers' velocities [ref: CalcTimeStep src/base/Utils.cpp:23-40] */
int se_qg2_timestep(se_qg2* q, double courant, double* dt);
/* The dissipation of TwoLayerQGEqn(psi1, omega1, psi2, omega2, F1, F2, beta, E2,
 * r, nu, tau_e, dt) [ref: TwoLayerQGEqn.cpp:17-87, DissipationRHS :419-462]:
 * layer 1 relaxes tau = psi1 - psi2 toward tau_e at rate r F1, layer 2 at
 * -r F2 with Ekman drag E2 on omega2, both layers take viscosity nu.  tau_e
 * is a host field (nV) or NULL (zero); all zero (the default) is the
 * undissipated step. */
int se_qg2_set_dissipation(se_qg2* q, double E2, double r, double nu, const double* tau_e);
/* infos: 2 entries (baroclinic, barotropic), nullable; with dissipation each
 * describes that operator's second solve, with the iterations of both */
int se_qg2_step(se_qg2* q, double dt, const se_solve_opts* opts, se_solve_info* infos);

/* One layer, the channel model [ref: QuasiGeostrophicEqn::Solve / FirstOrder /
 * SecondOrder src/qg/QuasiGeostrophicEqn.cpp:83-205, GenVel :207-223, GenPV
 * :225-232]: alpha = omega - F psi + beta y + topo advected, (lap - F) psi =
 * alpha - beta y - topo, omega = alpha + F psi - beta y - topo. */
typedef struct se_qg se_qg;
int se_qg_create(se_mesh* mesh, double F, double beta, int detj_mode, se_qg** out);
int se_qg_destroy(se_qg* q);
/* viscosity nu and linear drag r on omega [ref: the nu, r arguments of
 * QuasiGeostrophicEqn(phi, omega, F, beta, nu, r, dt, topo)
 * QuasiGeostrophicEqn.cpp:28-64, DisRHS :303-337]; zero (the default) is the
 * undissipated step */
int se_qg_set_dissipation(se_qg* q, double nu, double r);
/* topography [ref: the topo argument, QuasiGeostrophicEqn.cpp:28-64]: a host
 * field (nV), or NULL for none */
int se_qg_set_topography(se_qg* q, const double* topo);
int se_qg_set_state(se_qg* q, const double* psi, const double* omega);
int se_qg_get_state(se_qg* q, double* psi, double* omega);
int se_qg_timestep(se_qg* q, double courant, double* dt);
/* info (nullable): the last inversion, with the iterations of both when dissipative */
int se_qg_step(se_qg* q, double dt, const se_solve_opts* opts, se_solve_info* info);

/* ---- request server -----------------------------------------------------------
 * No counterpart in the reference (a model driver steps one TwoLayerQGEqn at a
 * time).  A server owns a context on `device` and one host thread; clients
 * submit two-layer QG step requests from any thread.  Requests are admitted in
 * arrival order while their device memory fits `memory_budget`; every tick
 * advances each admitted request by one se_qg2_step-equivalent step, the
 * inversions of all admitted requests running as batched solves (continuous
 * batching: requests join and leave between ticks).  A request's result is
 * bitwise what it gets when served alone. */
typedef struct se_server se_server;
typedef struct {
    double memory_budget; /* bytes of device memory for admitted requests (default 200e9) */
    int max_active;       /* limit on admitted requests, 0 = none */
    int precision;        /* SE_PREC_* of the inversions (default SE_PREC_FP32_IR) */
    double rtol;          /* default 1e-10 */
    int detj_mode;        /* SE_DETJ_* of the operators (default literal, as the reference) */
    int preconditioner;   /* SE_PC_* of the inversions (default: FDM for FP32_IR / BF16_IR; SE_PREC_FP64
                             with the default solves every inversion alone with the reference's Jacobi-PCG) */
} se_server_opts;
typedef struct {
    int nEls[2];
    int N;
    double min[2], max[2];
    int periodic[2];
    int bc[4]; /* bottom, top, left, right Dirichlet */
    double F1, F2, beta;
    int steps;  /* time steps to run */
    double dt;  /* > 0: this step; otherwise the CFL-0.5 step of the initial state */
    const double *psi1, *omega1, *psi2, *omega2; /* initial fields (nV each), copied at submission */
    double E2, r, nu;    /* dissipation as se_qg2_set_dissipation (zero: none) */
    const double* tau_e; /* relaxation target (nV) or NULL, copied at submission */
} se_qg2_request;
typedef struct {
    int64_t id;
    int status;                     /* SE_OK or an SE_ERR_* code */
    int steps_done;
    double t_submit, t_admit, t_done; /* seconds on the server's clock */
    int64_t inner_iters;            /* CG iterations over all its inversions */
    double unknowns;                /* sum of the unknowns of its inversions */
} se_request_result;
typedef struct {
    int64_t submitted, admitted, completed, steps, ticks;
    double device_bytes; /* admitted + pooled model memory (estimate) */
} se_server_stats;
void se_server_opts_default(se_server_opts* opts);
int se_server_create(int device, const se_server_opts* opts, se_server** out);
int se_server_submit(se_server* server, const se_qg2_request* req, int64_t* id);
/* blocks until request `id` is done; final fields into the (nullable) arrays */
int se_server_wait(se_server* server, int64_t id, se_request_result* res, double* psi1, double* omega1,
                   double* psi2, double* omega2);
int se_server_get_stats(se_server* server, se_server_stats* stats);
/* serves what was submitted, then stops */
int se_server_destroy(se_server* server);

#ifdef __cplusplus
}
#endif
#endif /* SE_AMD_H */
