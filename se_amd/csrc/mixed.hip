// mixed.hip -- mixed-precision Helmholtz solve.
//
// Jacobi-preconditioned CG whose iteration runs in FP32 (SE_PREC_FP32_IR) or
// with a BF16 search direction (SE_PREC_BF16_IR, FP32 arithmetic) on the
// separable form of the operator (sep_kernels.h, tcg_kernels.h), refined in
// FP64 by reliable updates: whenever the recursive residual has dropped by
// delta^2 (delta = inner_rtol, default 0.1) since the last update, or the
// recursion reports convergence, the FP32 correction is added to the FP64
// solution and the residual is recomputed as b - A x with the exact FP64
// operator (op_apply_dev).  The recurrence then continues with the same
// search direction, so the iteration count stays close to the FP64 solver's.
// Convergence is declared on the FP64 true residual only.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "se_common.h"
#include "se_op.h"
#include "tcg_host.h"
#include "tcg_kernels.h"

using namespace se;

namespace se {

// unknown index of a field line, -1 on a Dirichlet line (BCs.cpp:91-172)
int line_unknown(int I, int nv, int per, int lo, int hi) {
    if (per) return I == nv - 1 ? 0 : I;
    if (lo && I == 0) return -1;
    if (hi && I == nv - 1) return -1;
    return I - lo;
}

// Per direction: assembled band rows g (from the element operators
// D^T diag(P0) D of every element on the line), the line sums of P1 (scale of
// the other direction's sweep) and of the mass factor M.
void assemble_line(const se_mesh* m, const se_op* op, int axis, std::vector<double>& g, std::vector<double>& s0,
                   std::vector<double>& s1, std::vector<double>& mm) {
    const int N = m->N, n1 = N + 1, B = 2 * N + 1, CS = 4 * n1 + n1 * n1;
    const int nE = m->nEls[axis], nv = m->nVerts[axis], per = m->periodic[axis];
    const int lo = axis == 0 ? m->bc[2] : m->bc[0];
    const int hi = axis == 0 ? (per ? 0 : m->bc[3]) : (per ? 0 : m->bc[1]);
    const int nU = m->nU[axis];
    g.assign((size_t)nU * B, 0.0);
    s0.assign(nU, 0.0);
    s1.assign(nU, 0.0);
    mm.assign(nU, 0.0);
    const std::vector<double>& D = m->gll.D;
    const size_t base = axis == 0 ? 0 : (size_t)m->nEls[0] * CS;
    std::vector<double> G((size_t)n1 * n1);
    for (int e = 0; e < nE; e++) {
        const double* T = &op->tabs_host[base + (size_t)e * CS];
        // x lines: the x sweep uses P0 = X0 and the y sweep is scaled by X1;
        // y lines: the y sweep uses P1 = Y1 and the x sweep is scaled by Y0
        const double* Pop = axis == 0 ? T : T + n1;
        for (int rr = 0; rr < n1; rr++)
            for (int k = 0; k < n1; k++) {
                double s = 0.0;
                for (int q = 0; q < n1; q++) s += D[q * n1 + rr] * Pop[q] * D[q * n1 + k];
                G[rr * n1 + k] = s;
            }
        for (int rr = 0; rr < n1; rr++) {
            const int a = line_unknown(e * N + rr, nv, per, lo, hi);
            if (a < 0) continue;
            s0[a] += T[rr];
            s1[a] += T[n1 + rr];
            mm[a] += T[3 * n1 + rr];
            for (int k = 0; k < n1; k++) {
                const int b = line_unknown(e * N + k, nv, per, lo, hi);
                if (b < 0) continue;
                g[(size_t)a * B + (k - rr + N)] += G[rr * n1 + k];
            }
        }
    }
}

}  // namespace se

namespace {

constexpr int kVB = 1024;

int vblocks(int n) {
    int b = (n + kVB * 4 - 1) / (kVB * 4);
    return std::max(1, std::min(b, 256));
}

__global__ __launch_bounds__(kVB) void k_ir_sum(int n, const double* __restrict__ f, double sgn,
                                                double* __restrict__ part) {
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += sgn * f[i];
    double v[1] = {s};
    block_store_partials<1, kVB>(v, part);
}

__global__ void k_f32_to_bf16(int n, const float* __restrict__ a, __hip_bfloat16* __restrict__ b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = __float2bfloat16(a[i]);
}

// ---------------------------------------------------------------------------
// separable band tables (host assembly, FP32 upload)
// ---------------------------------------------------------------------------
struct SepLayout {
    size_t gx, gy, sx, sy, mx, my, total;  // float offsets
};

SepLayout sep_layout(const se_mesh* m) {
    const int B = 2 * m->N + 1;
    SepLayout L;
    L.gx = 0;
    L.gy = L.gx + (size_t)m->nU[0] * B;
    L.sx = L.gy + (size_t)m->nU[1] * B;
    L.sy = L.sx + m->nU[0];
    L.mx = L.sy + m->nU[1];
    L.my = L.mx + m->nU[0];
    L.total = L.my + m->nU[1];
    return L;
}

int sep_build(se_op* op) {
    if (op->sep_ready) return SE_OK;
    const se_mesh* m = op->mesh;
    const double sgn = (double)op->sign;
    std::vector<double> gx, X0, X1, MX, gy, Y0, Y1, MY;
    assemble_line(m, op, 0, gx, X0, X1, MX);
    assemble_line(m, op, 1, gy, Y0, Y1, MY);
    const SepLayout L = sep_layout(m);
    std::vector<float> t(L.total, 0.0f);
    // sign folded into the x-sweep band, the scales and the mass factor:
    // sign*A~ = (sy*sgn) Gx + (sx*sgn) Gy + (mx*sgn) my
    for (size_t i = 0; i < gx.size(); i++) t[L.gx + i] = (float)gx[i];
    for (size_t i = 0; i < gy.size(); i++) t[L.gy + i] = (float)gy[i];
    for (int I = 0; I < m->nU[0]; I++) {
        t[L.sx + I] = (float)(sgn * op->c_lap * X1[I]);
        t[L.mx + I] = (float)(sgn * op->c_mass * MX[I]);
    }
    for (int J = 0; J < m->nU[1]; J++) {
        t[L.sy + J] = (float)(sgn * op->c_lap * Y0[J]);
        t[L.my + J] = (float)MY[J];
    }
    int rc = op->sep.alloc(sizeof(float) * std::max<size_t>(L.total, 1));
    if (rc) return rc;
    SE_HIP(hipMemcpy(op->sep.ptr, t.data(), sizeof(float) * L.total, hipMemcpyHostToDevice));
    op->sep_ready = 1;
    return SE_OK;
}

SepParams sep_params(const se_op* op) {
    const se_mesh* m = op->mesh;
    const SepLayout L = sep_layout(m);
    const float* b = op->sep.as<float>();
    SepParams S;
    S.nU0 = m->nU[0];
    S.nU1 = m->nU[1];
    S.per0 = m->periodic[0];
    S.per1 = m->periodic[1];
    S.gx = b + L.gx;
    S.gy = b + L.gy;
    S.sx = b + L.sx;
    S.sy = b + L.sy;
    S.mx = b + L.mx;
    S.my = b + L.my;
    return S;
}

int sep_tiles(const se_mesh* m) {
    return ((m->nU[0] + SEP_TX - 1) / SEP_TX) * ((m->nU[1] + TCG_TY - 1) / TCG_TY);
}

#define SE_SEP_ORDERS(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8)

template <typename T, typename TQ, bool DOT>
int launch_sep(const se_op* op, const T* in, TQ* out, double* part, const CgState* st) {
    const se_mesh* m = op->mesh;
    SepParams S = sep_params(op);
    dim3 grid(sep_tiles(m)), block(SEP_NT);
    hipStream_t s = m->ctx->stream;
    switch (m->N) {
#define SE_CASE(NN) \
    case NN: hipLaunchKernelGGL((k_sep_apply<NN, T, TQ, DOT>), grid, block, 0, s, S, in, out, part, st); break;
        SE_SEP_ORDERS(SE_CASE)
#undef SE_CASE
        default: return fail(SE_ERR_UNSUPPORTED, "separable operator built for orders 2..8");
    }
    SE_HIP(hipGetLastError());
    return SE_OK;
}

// ---------------------------------------------------------------------------
// tile CG
// ---------------------------------------------------------------------------
size_t part_total(const se_mesh* m) { return 7 * (size_t)sep_tiles(m) + 256; }

}  // namespace

namespace se {
// tables and work buffers of the mixed-precision solve
int tcg_prepare(se_op* op) {
    SE_CHECK(op->svv_mN == 0, SE_ERR_UNSUPPORTED, "mixed-precision solve: operators with SVV solve in FP64");
    const se_mesh* m = op->mesh;
    const size_t nb = (size_t)std::max(m->n, 1);
    int rc = op_ensure_work(op);
    if (!rc) rc = sep_build(op);
    if (!rc) rc = op->sep_part.alloc(sizeof(double) * part_total(m));
    if (!rc) rc = op->sep_state.alloc(sizeof(CgState));
    // p sized for 4-byte elements whatever its type, so switching precisions
    // keeps the addresses a captured graph holds
    if (!rc) rc = op->lo_r.alloc(sizeof(float) * nb);
    if (!rc) rc = op->lo_p.alloc(sizeof(float) * nb);
    if (!rc) rc = op->lo_q.alloc(sizeof(float) * nb);
    if (!rc) rc = op->lo_x.alloc(sizeof(float) * nb);
    return rc;
}

TcgMember tcg_member(se_op* op, const double* d_f, double rtol, int null_space, int np_s) {
    TcgMember M;
    M.S = sep_params(op);
    M.z = op->lo_r.as<float>();
    M.q = op->lo_q.as<float>();
    M.xlo = op->lo_x.as<float>();
    M.p = op->lo_p.ptr;
    M.b = op->b.as<double>();
    M.x = op->x.as<double>();
    M.q64 = op->q.as<double>();
    M.f = d_f;
    M.part = op->sep_part.as<double>();
    M.st = op->sep_state.as<CgState>();
    M.tiles = tcg_tiles(op->mesh);
    M.np_s = np_s;
    M.null_space = null_space;
    M.n = op->mesh->n;
    M.sgn = (double)op->sign;
    M.rtol = rtol;
    return M;
}

int tcg_tiles(const se_mesh* m) {
    return ((m->nU[0] + TCG_TX - 1) / TCG_TX) * ((m->nU[1] + TCG_TY - 1) / TCG_TY);
}
}  // namespace se

namespace {

template <typename T, int MODE>
int launch_dir_apply(se_op* op, const TcgMember& M) {
    hipStream_t s = op->mesh->ctx->stream;
    switch (op->mesh->N) {
#define SE_CASE(NN)                                                                                                   \
    case NN: hipLaunchKernelGGL((k_tcg_dir_apply<NN, T, MODE>), dim3(M.tiles), dim3(TCG_NT), 0, s, M); break;
        SE_SEP_ORDERS(SE_CASE)
#undef SE_CASE
        default: return fail(SE_ERR_UNSUPPORTED, "separable operator built for orders 2..8");
    }
    return SE_OK;
}

template <typename T>
int launch_update(se_op* op, const TcgMember& M) {
    hipStream_t s = op->mesh->ctx->stream;
    switch (op->mesh->N) {
#define SE_CASE(NN) \
    case NN: hipLaunchKernelGGL((k_tcg_update<NN, T>), dim3(M.tiles), dim3(TCG_NT), 0, s, M); break;
        SE_SEP_ORDERS(SE_CASE)
#undef SE_CASE
        default: return fail(SE_ERR_UNSUPPORTED, "separable operator built for orders 2..8");
    }
    return SE_OK;
}

// RESID = false: start (b, x, x_lo, p, r from f); true: r = b - q64
template <typename T, bool RESID>
int launch_start(se_op* op, const TcgMember& M) {
    hipStream_t s = op->mesh->ctx->stream;
    switch (op->mesh->N) {
#define SE_CASE(NN) \
    case NN: hipLaunchKernelGGL((k_tcg_start<NN, T, RESID>), dim3(M.tiles), dim3(TCG_NT), 0, s, M); break;
        SE_SEP_ORDERS(SE_CASE)
#undef SE_CASE
        default: return fail(SE_ERR_UNSUPPORTED, "separable operator built for orders 2..8");
    }
    return SE_OK;
}

template <typename T>
int tcg_iteration(se_op* op, const TcgMember& M) {
    int rc = launch_update<T>(op, M);
    if (!rc) rc = launch_dir_apply<T, TCG_NORMAL>(op, M);
    return rc;
}

// k iterations, captured once per (k, T) into a graph and replayed
template <typename T>
int tcg_batch(se_op* op, const TcgMember& M, int k) {
    hipStream_t s = op->mesh->ctx->stream;
    static const bool use_graph = getenv("SE_NO_GRAPH") == nullptr;
    const int key = k * 2 + (sizeof(T) == 2 ? 1 : 0);
    if (use_graph) {
        if (!op->graph_exec_lo || op->graph_iters_lo != key) {
            if (op->graph_exec_lo) {
                (void)hipGraphExecDestroy(op->graph_exec_lo);
                op->graph_exec_lo = nullptr;
            }
            hipGraph_t g = nullptr;
            SE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = SE_OK;
            for (int i = 0; i < k && !rc; i++) rc = tcg_iteration<T>(op, M);
            hipError_t e = hipStreamEndCapture(s, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(SE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&op->graph_exec_lo, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                op->graph_exec_lo = nullptr;
                return fail(SE_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            }
            op->graph_iters_lo = key;
        }
        SE_HIP(hipGraphLaunch(op->graph_exec_lo, s));
        return SE_OK;
    }
    for (int i = 0; i < k; i++) {
        int rc = tcg_iteration<T>(op, M);
        if (rc) return rc;
    }
    return SE_OK;
}

// x += x_lo; r = b - A x (FP64); r_lo = r; st->tt = ||r||^2 (projected)
template <typename T>
int reliable_update(se_op* op, const TcgMember& M) {
    hipStream_t s = op->mesh->ctx->stream;
    const int n = op->mesh->n;
    hipLaunchKernelGGL(k_tcg_flush, dim3(std::max(1, std::min((n + 4095) / 4096, 2048))), dim3(TCG_NT), 0, s, M);
    int rc = op_apply_dev(op, op->x.as<double>(), op->q.as<double>(), (double)op->sign, false);
    if (!rc) rc = launch_start<T, true>(op, M);
    if (rc) return rc;
    hipLaunchKernelGGL(k_tcg_true, dim3(1), dim3(TCG_NT), 0, s, M);
    SE_HIP(hipGetLastError());
    return SE_OK;
}

template <typename T>
int tcg_solve(se_op* op, const double* d_f, double* d_x, const se_solve_opts& o, se_solve_info& info) {
    se_mesh* m = op->mesh;
    se_ctx* c = m->ctx;
    hipStream_t s = c->stream;
    const int n = m->n, vb = vblocks(n);
    SE_CHECK(op->definite, SE_ERR_UNSUPPORTED, "operator diagonal changes sign: not a definite system");
    int rc = tcg_prepare(op);
    if (rc) return rc;
    const int null = o.null_space ? 1 : 0;
    const TcgMember M = tcg_member(op, d_f, o.rtol, null, vb);
    TcgControl ctl;
    ctl.configure(o);
    const int k = std::max(1, o.check_every);
    if (!op->h_state) SE_HIP(hipHostMalloc(&op->h_state, sizeof(CgState), hipHostMallocDefault));
    CgState* hs = (CgState*)op->h_state;
    CgState* st = M.st;
    auto pull = [&]() -> int {
        SE_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        SE_HIP(hipStreamSynchronize(s));
        return SE_OK;
    };

    SE_HIP(hipEventRecord(c->ev0, s));
    // state uploads go through the pinned host copy
    *hs = CgState{};
    hs->maxit = ctl.maxit;
    hs->null_space = null;
    hs->n = n;
    hs->status = CG_RUNNING;
    SE_HIP(hipMemcpyAsync(st, hs, sizeof(CgState), hipMemcpyHostToDevice, s));
    if (null) hipLaunchKernelGGL(k_ir_sum, dim3(vb), dim3(kVB), 0, s, n, d_f, (double)op->sign, M.part + 7 * (size_t)M.tiles);
    rc = launch_start<T, false>(op, M);
    if (!rc) rc = launch_dir_apply<T, TCG_FIRST>(op, M);
    if (!rc) rc = pull();
    if (rc) return rc;
    SE_HIP(hipGetLastError());
    ctl.started(*hs);
    for (int round = 0; ctl.status == CG_RUNNING; round++) {
        const int reps = replays_for_round(round);
        for (int r = 0; r < reps && !rc; r++) rc = tcg_batch<T>(op, M, k);
        if (!rc) rc = pull();
        if (rc) return rc;
        if (!ctl.due(*hs, k * reps)) continue;
        ctl.before_update(*hs);
        rc = reliable_update<T>(op, M);
        if (!rc) rc = pull();
        if (rc) return rc;
        if (ctl.after_update(*hs) != CG_RUNNING) break;
        SE_HIP(hipMemcpyAsync(st, hs, sizeof(CgState), hipMemcpyHostToDevice, s));
        rc = ctl.reset_dir ? launch_dir_apply<T, TCG_RESET>(op, M) : launch_dir_apply<T, TCG_NORMAL>(op, M);
        if (rc) return rc;
    }
    if (d_x != op->x.as<double>())
        SE_HIP(hipMemcpyAsync(d_x, op->x.ptr, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    SE_HIP(hipEventRecord(c->ev1, s));
    SE_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    ctl.fill(info, hs->iter, ms * 1e-3);
    return SE_OK;
}

}  // namespace

extern "C" int se_helm_apply_inner(se_op* op, int precision, const float* x, float* y, int* sign) {
    SE_CHECK(op && x && y, SE_ERR_ARG, "se_helm_apply_inner: NULL argument");
    SE_CHECK(precision == SE_PREC_FP32_IR || precision == SE_PREC_BF16_IR, SE_ERR_ARG,
             "se_helm_apply_inner: precision must be SE_PREC_FP32_IR or SE_PREC_BF16_IR");
    se_mesh* m = op->mesh;
    SE_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    const int n = m->n;
    if (sign) *sign = op->sign;
    if (n == 0) return SE_OK;
    int rc = sep_build(op);
    if (rc) return rc;
    DevBuf dx, dy, db;
    rc = dx.alloc(sizeof(float) * n);
    if (!rc) rc = dy.alloc(sizeof(float) * n);
    if (!rc && precision == SE_PREC_BF16_IR) rc = db.alloc(sizeof(__hip_bfloat16) * n);
    if (rc) return rc;
    SE_HIP(hipMemcpyAsync(dx.ptr, x, sizeof(float) * n, hipMemcpyHostToDevice, s));
    if (precision == SE_PREC_FP32_IR) {
        rc = launch_sep<float, float, false>(op, dx.as<float>(), dy.as<float>(), nullptr, nullptr);
    } else {
        hipLaunchKernelGGL(k_f32_to_bf16, dim3(std::max(1, std::min((n + 255) / 256, 2048))), dim3(256), 0, s, n,
                           dx.as<float>(), db.as<__hip_bfloat16>());
        rc = launch_sep<__hip_bfloat16, float, false>(op, db.as<__hip_bfloat16>(), dy.as<float>(), nullptr, nullptr);
    }
    if (rc) return rc;
    SE_HIP(hipMemcpyAsync(y, dy.ptr, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    SE_HIP(hipStreamSynchronize(s));
    return SE_OK;
}

namespace {
// time `reps` back-to-back launches of the direction+operator kernel and of
// the update kernel, continuing the recurrence of the last solve
template <typename T>
int helm_probe(se_op* op, int reps, double* out) {
    hipStream_t s = op->mesh->ctx->stream;
    se_ctx* c = op->mesh->ctx;
    const TcgMember M = tcg_member(op, nullptr, 0.0, 0, 1);
    CgState* hs = (CgState*)op->h_state;
    SE_HIP(hipMemcpyAsync(hs, M.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    SE_HIP(hipStreamSynchronize(s));
    hs->stop = 0;
    hs->status = CG_RUNNING;
    hs->tol2 = 0.0;
    hs->iter = 0;
    hs->maxit = 1 << 30;
    hs->rho = hs->rho_prev = 1.0;
    hs->beta = 0.0;
    SE_HIP(hipMemcpyAsync(M.st, hs, sizeof(CgState), hipMemcpyHostToDevice, s));
    int rc = launch_dir_apply<T, TCG_NORMAL>(op, M);
    if (!rc) rc = launch_update<T>(op, M);
    float ms_d = 0.f, ms_u = 0.f;
    SE_HIP(hipEventRecord(c->ev0, s));
    for (int r = 0; r < reps && !rc; r++) rc = launch_dir_apply<T, TCG_NORMAL>(op, M);
    SE_HIP(hipEventRecord(c->ev1, s));
    SE_HIP(hipEventSynchronize(c->ev1));
    (void)hipEventElapsedTime(&ms_d, c->ev0, c->ev1);
    SE_HIP(hipEventRecord(c->ev0, s));
    for (int r = 0; r < reps && !rc; r++) rc = launch_update<T>(op, M);
    SE_HIP(hipEventRecord(c->ev1, s));
    SE_HIP(hipEventSynchronize(c->ev1));
    (void)hipEventElapsedTime(&ms_u, c->ev0, c->ev1);
    if (rc) return rc;
    const double n = op->mesh->n, sT = (double)sizeof(T);
    out[0] = 1e3 * ms_d / reps;
    out[1] = 1e3 * ms_u / reps;
    out[2] = n * (8.0 + sT);
    out[3] = n * (20.0 + 2.0 * sT);
    return SE_OK;
}
}  // namespace

extern "C" int se_helm_probe(se_op* op, int precision, int reps, double* us_dir, double* us_upd, double* bytes_dir,
                             double* bytes_upd) {
    SE_CHECK(op && us_dir && us_upd && bytes_dir && bytes_upd, SE_ERR_ARG, "se_helm_probe: NULL argument");
    SE_CHECK(reps > 0, SE_ERR_ARG, "se_helm_probe: reps must be positive");
    SE_CHECK(precision == SE_PREC_FP32_IR || precision == SE_PREC_BF16_IR, SE_ERR_ARG,
             "se_helm_probe: precision must be SE_PREC_FP32_IR or SE_PREC_BF16_IR");
    SE_CHECK(op->sep_ready && op->h_state && op->sep_state.ptr, SE_ERR_STATE,
             "se_helm_probe: run a mixed-precision solve of this operator first");
    SE_HIP(hipSetDevice(op->mesh->ctx->device));
    double out[4];
    int rc = precision == SE_PREC_FP32_IR ? helm_probe<float>(op, reps, out) : helm_probe<__hip_bfloat16>(op, reps, out);
    if (rc) return rc;
    *us_dir = out[0], *us_upd = out[1], *bytes_dir = out[2], *bytes_upd = out[3];
    return SE_OK;
}

namespace se {
int pcg_mixed(se_op* op, const double* d_f, double* d_x, const se_solve_opts& o, se_solve_info& info) {
    // the inner operator is the separable Sy Gx + Gy Sx + My Mx; the SVV term
    // (band x band) is not of that form
    SE_CHECK(op->svv_mN == 0, SE_ERR_UNSUPPORTED, "mixed-precision solve: operators with SVV solve in FP64");
    if (o.precision == SE_PREC_FP32_IR) return tcg_solve<float>(op, d_f, d_x, o, info);
    if (o.precision == SE_PREC_BF16_IR) return tcg_solve<__hip_bfloat16>(op, d_f, d_x, o, info);
    return fail(SE_ERR_ARG, "unknown precision");
}
}  // namespace se
