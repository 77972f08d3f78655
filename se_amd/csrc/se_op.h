// se_op.h -- operator object shared by the FP64 and mixed-precision solvers.
#pragma once
#include "helm_kernels.h"
#include "se_common.h"

struct se_op;
namespace se {
struct FdmOp;                   // fast-diagonalisation data of an operator (fdm.hip)
void fdm_release(se_op* op);
}  // namespace se

struct se_op {
    se_mesh* mesh = nullptr;
    double c_mass = 0.0, c_lap = 0.0;
    int detj = 0;
    int tile = 2;
    int sign = 1;          // sign of the operator diagonal (solve runs on sign*A)
    bool definite = true;
    se::DevBuf dinv;       // inverse diagonal of A
    se::DevBuf dinv_s;     // inverse diagonal of sign*A
    se::DevBuf lift;       // Dirichlet lifting b
    se::DevBuf tabs;       // metric factor tables: element columns, then element rows
    std::vector<double> tabs_host;
    se::DevBuf qtabs;      // FP32 cross-term tables (and transposes): columns, then rows
    // separable inner operator (FP32 band tables) and mixed-precision work
    se::DevBuf sep, sep_state, sep_part;
    int sep_ready = 0, sep_sign = 0;
    se::DevBuf state, partials;
    se::DevBuf b, x, r, p, q, f;             // FP64 CG work vectors
    se::DevBuf sx, sy, sfield, sfield2;      // host-staging buffers
    // mixed-precision work (allocated on first use)
    se::DevBuf lo_r, lo_p, lo_q, lo_z, lo_x, lo_dinv, lo_geo, d64;
    void* h_state = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_iters = 0;
    hipGraphExec_t graph_exec_lo = nullptr;
    int graph_iters_lo = 0;
    // spectral vanishing viscosity term (svv.hip): dt, cutoff (0: none), the
    // 1-D factors Mf | Kf
    double svv_dt = 0.0;
    int svv_mN = 0;
    se::DevBuf svv_mk;
    // fast-diagonalisation preconditioner (fdm.hip), built on first use
    se::FdmOp* fdm = nullptr;

    void release_all() {
        se::fdm_release(this);
        se::DevBuf* bufs[] = {&sep, &sep_state, &sep_part, &dinv, &dinv_s, &lift, &tabs, &qtabs, &state, &partials, &b, &x, &r, &p, &q, &f,
                              &sx, &sy, &sfield, &sfield2, &lo_r, &lo_p, &lo_q, &lo_z, &lo_x, &lo_dinv,
                              &lo_geo, &d64, &svv_mk};
        for (auto* b : bufs) b->release();
        if (graph_exec_lo) (void)hipGraphExecDestroy(graph_exec_lo);
        graph_exec_lo = nullptr;
        if (h_state) (void)hipHostFree(h_state);
        h_state = nullptr;
    }
};

namespace se {
OpParams op_params(const se_op* op, double scale);
int op_apply_dev(se_op* op, const double* d_x, double* d_y, double scale, bool dot);
int op_ensure_work(se_op* op);
int pcg_fp64(se_op* op, const double* d_f, double* d_x, const se_solve_opts& o, se_solve_info& info);
int pcg_mixed(se_op* op, const double* d_f, double* d_x, const se_solve_opts& o, se_solve_info& info);
// FDM-preconditioned CG (fdm.hip); solve_pc resolves SE_PC_DEFAULT
int pcg_fdm(se_op* op, const double* d_f, double* d_x, const se_solve_opts& o, se_solve_info& info);
int solve_pc(const se_solve_opts& o);
struct FdmPlan;
int fdm_plan_create(se_ctx* c, int count, se_op* const* ops, FdmPlan** out);
int fdm_plan_set_members(FdmPlan* pl, int count, se_op* const* ops);
void fdm_plan_destroy(FdmPlan* pl);
int fdm_plan_solve(FdmPlan* pl, const double* const* d_f, double* const* d_x, const se_solve_opts& o,
                   se_solve_info* infos);
int fdm_plan_probe(FdmPlan* pl, int reps, double* us_pass, double* flops_pass, double* us_apply,
                   double* bytes_apply);
// SVV term (svv.hip): y += scale S x (gate: only while that CG awaits its
// true-residual check); np partials of x.y; lift -= S x_D
int svv_apply(se_op* op, const double* d_x, double* d_y, double scale, const CgState* gate = nullptr);
int svv_dot(se_op* op, const double* d_x, const double* d_y, double* partials, int np);
int svv_lift(se_op* op, const double* d_field, double* d_lift);
}  // namespace se

// mixed-precision solve setup (mixed.hip), used by the batched solver
namespace se {
int tcg_prepare(se_op* op);
int tcg_tiles(const se_mesh* m);
// separable form of the operator along one axis (FP64, before any sign or
// coefficient): band rows g [nU][2N+1], line sums s0, s1 of the metric
// factors and the mass factor mm [nU]
void assemble_line(const se_mesh* m, const se_op* op, int axis, std::vector<double>& g, std::vector<double>& s0,
                   std::vector<double>& s1, std::vector<double>& mm);
}  // namespace se
