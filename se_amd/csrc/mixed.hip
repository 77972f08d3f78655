// mixed.hip -- mixed-precision Helmholtz solve: FP64 iterative refinement
// around an inner Jacobi-PCG whose vectors are FP32 (SE_PREC_FP32_IR) or BF16
// (SE_PREC_BF16_IR, FP32 arithmetic), with the separable operator of
// sep_kernels.h.  The outer residual r = b - A x, its norm, the correction
// x += d and every global inner product are FP64; the outer residual uses the
// exact operator (op_apply_dev), so the returned x meets the FP64 target
// against the reference operator.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "se_common.h"
#include "se_op.h"
#include "sep_kernels.h"

using namespace se;

namespace {

constexpr int kVB = 1024;

int vblocks(int n) {
    int b = (n + kVB * 4 - 1) / (kVB * 4);
    return std::max(1, std::min(b, 256));
}

// unknown index of a field line, -1 on a Dirichlet line (BCs.cpp:91-172)
int line_unknown(int I, int nv, int per, int lo, int hi) {
    if (per) return I == nv - 1 ? 0 : I;
    if (lo && I == 0) return -1;
    if (hi && I == nv - 1) return -1;
    return I - lo;
}

// ---------------------------------------------------------------------------
// outer (FP64) kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kVB) void k_ir_sum(int n, const double* __restrict__ f, double sgn,
                                                double* __restrict__ part) {
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += sgn * f[i];
    double v[1] = {s};
    block_store_partials<1, kVB>(v, part);
}

// b = sgn f - mean (null space); x = 0; r = b; partials (sum r, sum r^2)
__global__ __launch_bounds__(kVB) void k_ir_start(int n, double sgn, int null_space, const double* __restrict__ f,
                                                  const double* __restrict__ part_s, int np_s, double* __restrict__ b,
                                                  double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ part) {
    double mu = 0.0;
    if (null_space) {
        double t[1];
        reduce_partials<1, kVB>(part_s, np_s, t);
        mu = t[0] / n;
    }
    double s = 0.0, s2 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double bi = sgn * f[i] - mu;
        b[i] = bi;
        r[i] = bi;
        x[i] = 0.0;
        s += bi;
        s2 += bi * bi;
    }
    double v[2] = {s, s2};
    block_store_partials<2, kVB>(v, part);
}

// r = b - q; partials (sum r, sum r^2)
__global__ __launch_bounds__(kVB) void k_ir_resid(int n, const double* __restrict__ b, const double* __restrict__ q,
                                                  double* __restrict__ r, double* __restrict__ part) {
    double s = 0.0, s2 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double t = b[i] - q[i];
        r[i] = t;
        s += t;
        s2 += t * t;
    }
    double v[2] = {s, s2};
    block_store_partials<2, kVB>(v, part);
}

// single block: out[0] = ||r - mean||^2 (or ||r||^2), out[1] = mean
__global__ __launch_bounds__(kVB) void k_ir_fin(int n, int null_space, const double* __restrict__ part, int np,
                                                double* __restrict__ out) {
    double t[2];
    reduce_partials<2, kVB>(part, np, t);
    if (threadIdx.x == 0) {
        double mu = null_space ? t[0] / n : 0.0;
        out[0] = null_space ? t[1] - t[0] * mu : t[1];
        out[1] = mu;
    }
}

__global__ void k_ir_xadd(int n, double* __restrict__ x, const float* __restrict__ d, double scale) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[i] += scale * (double)d[i];
}

__global__ void k_to_f32(int n, const double* __restrict__ a, double s, float* __restrict__ b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (float)(s * a[i]);
}

// ---------------------------------------------------------------------------
// inner Jacobi-PCG: r, x in FP32, the search direction p (the operator's
// input, read with its halo) in T = float or bf16; FP32 arithmetic, FP64
// reductions; same recurrence and stop logic as the FP64 solver.  The residual
// recurrence r -= alpha A p stays in FP32 with p as stored, so r tracks the
// inner system's residual; rounding p to bf16 only perturbs conjugacy.
// ---------------------------------------------------------------------------
// four consecutive elements per thread: one 16-byte (float) or 8-byte (bf16)
// access per vector, scalar accesses on a partial last quad
__device__ __forceinline__ void load4(const float* __restrict__ p, int i, int n, float (&v)[4]) {
    if (i + 3 < n) {
        const float4 t = *reinterpret_cast<const float4*>(p + i);
        v[0] = t.x, v[1] = t.y, v[2] = t.z, v[3] = t.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = i + k < n ? p[i + k] : 0.0f;
    }
}
__device__ __forceinline__ void load4(const __hip_bfloat16* __restrict__ p, int i, int n, float (&v)[4]) {
    if (i + 3 < n) {
        const uint2 t = *reinterpret_cast<const uint2*>(p + i);
        v[0] = __uint_as_float(t.x << 16), v[1] = __uint_as_float(t.x & 0xffff0000u);
        v[2] = __uint_as_float(t.y << 16), v[3] = __uint_as_float(t.y & 0xffff0000u);
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = i + k < n ? __bfloat162float(p[i + k]) : 0.0f;
    }
}
__device__ __forceinline__ void store4(float* __restrict__ p, int i, int n, const float (&v)[4]) {
    if (i + 3 < n) {
        *reinterpret_cast<float4*>(p + i) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i + k < n) p[i + k] = v[k];
    }
}
__device__ __forceinline__ unsigned bf16_bits(float x) {
    return (unsigned)__bfloat16_as_ushort(__float2bfloat16(x));
}
__device__ __forceinline__ void store4(__hip_bfloat16* __restrict__ p, int i, int n, const float (&v)[4]) {
    if (i + 3 < n) {
        uint2 t;
        t.x = bf16_bits(v[0]) | (bf16_bits(v[1]) << 16);
        t.y = bf16_bits(v[2]) | (bf16_bits(v[3]) << 16);
        *reinterpret_cast<uint2*>(p + i) = t;
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i + k < n) p[i + k] = __float2bfloat16(v[k]);
    }
}
__global__ __launch_bounds__(kVB) void k_in_start(int n, const double* __restrict__ r64, const double* __restrict__ mu_scale,
                                                  const float* __restrict__ dinv, float* __restrict__ r, float* __restrict__ x,
                                                  double* __restrict__ part) {
    const double mu = mu_scale[1], scale = 1.0 / sqrt(mu_scale[0]);
    double rr = 0.0, rz = 0.0, sz = 0.0, sr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float rv = (float)((r64[i] - mu) * scale);
        r[i] = rv;
        x[i] = 0.0f;
        const float z = dinv[i] * rv;
        rr += (double)rv * rv;
        rz += (double)rv * z;
        sz += z;
        sr += rv;
    }
    double v[4] = {rr, rz, sz, sr};
    block_store_partials<4, kVB>(v, part);
}

enum { IN_FIRST = 0, IN_NORMAL = 1 };

template <typename T>
__global__ __launch_bounds__(kVB) void k_in_dir(int mode, int n, double rtol, const float* __restrict__ dinv,
                                                const float* __restrict__ r, T* __restrict__ p,
                                                const double* __restrict__ part_b, int np_b, CgState* __restrict__ st) {
    // the first quad's loads are issued before the stop flag and the partial
    // sums are read, so their latencies overlap
    const int i0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x), stride = 4 * gridDim.x * blockDim.x;
    float rv[4], dv[4], pv[4];
    if (i0 < n) {
        load4(r, i0, n, rv);
        load4(dinv, i0, n, dv);
        if (mode != IN_FIRST) load4(p, i0, n, pv);
    }
    if (mode == IN_NORMAL && __hip_atomic_load(&st->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    double t[4];
    reduce_partials<4, kVB>(part_b, np_b, t);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const double rr = t[0];
    const bool finite = (rr - rr == 0.0) && (t[1] - t[1] == 0.0);
    const double mu = st->null_space ? t[2] / n : 0.0;
    const double rho_new = t[1] - mu * t[3];
    double beta = 0.0;
    if (mode == IN_FIRST) {
        if (lead) {
            st->bb = rr;
            st->tol2 = rtol * rtol * rr;
            st->rr = rr;
            st->rho = rho_new;
            st->iter = 0;
            st->best = rr;
            st->since_best = 0;
            st->stop = (rr == 0.0 || !finite) ? 1 : 0;
            st->status = rr == 0.0 ? 0 : (finite ? CG_RUNNING : 3);
        }
        if (rr == 0.0 || !finite) return;
    } else {
        const bool conv = finite && rr <= st->tol2;
        if (lead) {
            const int it = st->iter + 1;
            st->iter = it;
            st->rr = rr;
            if (!finite) {
                st->stop = 1;
                st->status = 3;
            } else if (conv) {
                st->stop = 1;
                st->status = 0;
            } else {
                st->rho = rho_new;
                st->beta = rho_new / st->rho_prev;
                if (rr < 0.999 * st->best) {
                    st->best = rr;
                    st->since_best = 0;
                } else {
                    st->since_best += 1;
                }
                if (st->since_best >= st->stall) {
                    st->stop = 1;
                    st->status = 4;
                } else if (it >= st->maxit) {
                    st->stop = 1;
                    st->status = 1;
                }
            }
        }
        if (!finite || conv) return;
        beta = rho_new / st->rho_prev;
    }
    const float fmu = (float)mu, fb = (float)beta;
    for (int i = i0; i < n; i += stride) {
        if (i != i0) {
            load4(r, i, n, rv);
            load4(dinv, i, n, dv);
            if (mode != IN_FIRST) load4(p, i, n, pv);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float z = dv[k] * rv[k] - fmu;
            pv[k] = mode == IN_FIRST ? z : z + fb * pv[k];
        }
        store4(p, i, n, pv);
    }
}

template <typename T, typename TQ>
__global__ __launch_bounds__(kVB) void k_in_update(int n, const T* __restrict__ p, const TQ* __restrict__ q,
                                                   const float* __restrict__ dinv, float* __restrict__ x,
                                                   float* __restrict__ r, const double* __restrict__ part_a, int np_a,
                                                   double* __restrict__ part_b, CgState* __restrict__ st) {
    const int i0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x), stride = 4 * gridDim.x * blockDim.x;
    float pv[4], qv[4], dv[4], xv[4], rv[4];
    if (i0 < n) {
        load4(p, i0, n, pv);
        load4(q, i0, n, qv);
        load4(dinv, i0, n, dv);
        load4(x, i0, n, xv);
        load4(r, i0, n, rv);
    }
    if (__hip_atomic_load(&st->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    double t[1];
    reduce_partials<1, kVB>(part_a, np_a, t);
    const double pq = t[0];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (!(pq > 0.0)) {
        if (lead) {
            st->pq = pq;
            st->stop = 1;
            st->status = (pq - pq == 0.0) ? 2 : 3;
        }
        return;
    }
    const double rho = st->rho;
    const double alpha = rho / pq;
    if (lead) {
        st->pq = pq;
        st->alpha = alpha;
        st->rho_prev = rho;
    }
    const float fa = (float)alpha;
    double rr = 0.0, rz = 0.0, sz = 0.0, sr = 0.0;
    for (int i = i0; i < n; i += stride) {
        if (i != i0) {
            load4(p, i, n, pv);
            load4(q, i, n, qv);
            load4(dinv, i, n, dv);
            load4(x, i, n, xv);
            load4(r, i, n, rv);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            xv[k] += fa * pv[k];
            rv[k] -= fa * qv[k];
            const float z = dv[k] * rv[k];
            rr += (double)rv[k] * rv[k];
            rz += (double)rv[k] * z;
            sz += z;
            sr += rv[k];
        }
        store4(x, i, n, xv);
        store4(r, i, n, rv);
    }
    double v[4] = {rr, rz, sz, sr};
    block_store_partials<4, kVB>(v, part_b);
}

// ---------------------------------------------------------------------------
// host: separable band tables
// ---------------------------------------------------------------------------
struct SepLayout {
    size_t gx, gy, sx, sy, mx, my, dinv, total;  // float offsets
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
    L.dinv = L.my + m->nU[1];
    L.total = L.dinv + (size_t)std::max(m->n, 1);
    return L;
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

int sep_build(se_op* op) {
    if (op->sep_ready) return SE_OK;
    const se_mesh* m = op->mesh;
    const double sgn = (double)op->sign;
    std::vector<double> gx, X0, X1, MX, gy, Y0, Y1, MY;
    assemble_line(m, op, 0, gx, X0, X1, MX);
    assemble_line(m, op, 1, gy, Y0, Y1, MY);
    const SepLayout L = sep_layout(m);
    std::vector<float> t(L.total, 0.0f);
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
    int rc = op->sep.alloc(sizeof(float) * L.total);
    if (rc) return rc;
    SE_HIP(hipMemcpy(op->sep.ptr, t.data(), sizeof(float) * L.dinv, hipMemcpyHostToDevice));
    // Jacobi diagonal of the sign-normalised exact operator, in FP32
    int n = m->n;
    hipLaunchKernelGGL(k_to_f32, dim3(std::max(1, std::min((n + 255) / 256, 2048))), dim3(256), 0, m->ctx->stream, n,
                       op->dinv.as<double>(), sgn, op->sep.as<float>() + L.dinv);
    SE_HIP(hipGetLastError());
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
    return ((m->nU[0] + SEP_TX - 1) / SEP_TX) * ((m->nU[1] + SEP_TY - 1) / SEP_TY);
}

template <typename T, typename TQ, bool DOT>
int launch_sep(const se_op* op, const T* in, TQ* out, double* part, const CgState* st) {
    const se_mesh* m = op->mesh;
    SepParams S = sep_params(op);
    dim3 grid(sep_tiles(m)), block(256);
    hipStream_t s = m->ctx->stream;
    switch (m->N) {
#define SE_SEP_CASE(NN) \
    case NN: hipLaunchKernelGGL((k_sep_apply<NN, T, TQ, DOT>), grid, block, 0, s, S, in, out, part, st); break;
        SE_SEP_CASE(2)
        SE_SEP_CASE(3)
        SE_SEP_CASE(4)
        SE_SEP_CASE(5)
        SE_SEP_CASE(6)
        SE_SEP_CASE(7)
        SE_SEP_CASE(8)
#undef SE_SEP_CASE
        default: return fail(SE_ERR_UNSUPPORTED, "separable operator built for orders 2..8");
    }
    SE_HIP(hipGetLastError());
    return SE_OK;
}

// work areas inside op->sep_part (doubles): A sep tiles, B 4 x VB, R 2 x VB, S VB, O 2
struct MixLayout {
    size_t a, b, r, s, o, total;
};
MixLayout mix_layout(const se_mesh* m) {
    MixLayout L;
    L.a = 0;
    L.b = (size_t)sep_tiles(m);
    L.r = L.b + 4 * 256;
    L.s = L.r + 2 * 256;
    L.o = L.s + 256;
    L.total = L.o + 2;
    return L;
}

// inner vectors live in op->lo_*: r (float), p (T), q (TQ), x (float)
template <typename T, typename TQ>
int inner_iteration(se_op* op) {
    const se_mesh* m = op->mesh;
    const int n = m->n, vb = vblocks(n);
    const MixLayout L = mix_layout(m);
    double* part = op->sep_part.as<double>();
    CgState* st = op->sep_state.as<CgState>();
    const float* dinv = op->sep.as<float>() + sep_layout(m).dinv;
    hipStream_t s = m->ctx->stream;
    int rc = launch_sep<T, TQ, true>(op, op->lo_p.as<T>(), op->lo_q.as<TQ>(), part + L.a, st);
    if (rc) return rc;
    hipLaunchKernelGGL((k_in_update<T, TQ>), dim3(vb), dim3(kVB), 0, s, n, op->lo_p.as<T>(), op->lo_q.as<TQ>(), dinv,
                       op->lo_x.as<float>(), op->lo_r.as<float>(), part + L.a, sep_tiles(m), part + L.b, st);
    hipLaunchKernelGGL((k_in_dir<T>), dim3(vb), dim3(kVB), 0, s, (int)IN_NORMAL, n, 0.0, dinv, op->lo_r.as<float>(),
                       op->lo_p.as<T>(), part + L.b, vb, st);
    return SE_OK;
}

template <typename T, typename TQ>
int inner_batch(se_op* op, int k) {
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
            for (int i = 0; i < k && !rc; i++) rc = inner_iteration<T, TQ>(op);
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
        int rc = inner_iteration<T, TQ>(op);
        if (rc) return rc;
    }
    return SE_OK;
}

template <typename T, typename TQ>
int ir_solve(se_op* op, const double* d_f, double* d_x, const se_solve_opts& o, se_solve_info& info) {
    se_mesh* m = op->mesh;
    se_ctx* c = m->ctx;
    hipStream_t s = c->stream;
    const int n = m->n, vb = vblocks(n);
    SE_CHECK(op->definite, SE_ERR_UNSUPPORTED, "operator diagonal changes sign: not a definite system");
    int rc = op_ensure_work(op);
    if (!rc) rc = sep_build(op);
    const MixLayout L = mix_layout(m);
    const size_t nb = (size_t)std::max(n, 1);
    if (!rc) rc = op->sep_part.alloc(sizeof(double) * L.total);
    if (!rc) rc = op->sep_state.alloc(sizeof(CgState));
    // sized for 4-byte elements whatever T is, so switching precisions keeps
    // the addresses a captured inner graph holds
    if (!rc) rc = op->lo_r.alloc(sizeof(float) * nb);
    if (!rc) rc = op->lo_p.alloc(sizeof(float) * nb);
    if (!rc) rc = op->lo_q.alloc(sizeof(float) * nb);
    if (!rc) rc = op->lo_x.alloc(sizeof(float) * nb);
    if (rc) return rc;
    double* part = op->sep_part.as<double>();
    double* outer = part + L.o;
    CgState* st = op->sep_state.as<CgState>();
    const float* dinv = op->sep.as<float>() + sep_layout(m).dinv;
    const double sgn = (double)op->sign;
    const int null = o.null_space ? 1 : 0;
    const double inner_tol = o.inner_rtol > 0.0 ? o.inner_rtol : (sizeof(T) == 2 ? 1e-3 : 1e-6);

    SE_HIP(hipEventRecord(c->ev0, s));
    if (null) hipLaunchKernelGGL(k_ir_sum, dim3(vb), dim3(kVB), 0, s, n, d_f, sgn, part + L.s);
    hipLaunchKernelGGL(k_ir_start, dim3(vb), dim3(kVB), 0, s, n, sgn, null, d_f, part + L.s, vb, op->b.as<double>(),
                       op->x.as<double>(), op->r.as<double>(), part + L.r);
    hipLaunchKernelGGL(k_ir_fin, dim3(1), dim3(kVB), 0, s, n, null, part + L.r, vb, outer);
    SE_HIP(hipGetLastError());
    double ho[2];
    SE_HIP(hipMemcpyAsync(ho, outer, sizeof(ho), hipMemcpyDeviceToHost, s));
    SE_HIP(hipStreamSynchronize(s));
    const double bb = ho[0];
    double rr = bb;
    int status = 1, total_inner = 0, outer_steps = 0;
    if (!(bb - bb == 0.0)) status = 3;
    else if (bb == 0.0) status = 0;
    if (!op->h_state) SE_HIP(hipHostMalloc(&op->h_state, sizeof(CgState), hipHostMallocDefault));
    CgState* hs = (CgState*)op->h_state;
    const int k = std::max(1, o.check_every);
    while (status == 1) {
        if (rr <= o.rtol * o.rtol * bb) {
            status = 0;
            break;
        }
        if (outer_steps >= o.max_outer || total_inner >= o.maxit) break;
        outer_steps++;
        // inner solve of A d = r / ||r|| to a relative residual that is no
        // tighter than the remaining outer reduction needs
        const double need = o.rtol * std::sqrt(bb / rr);
        const double tol = std::max(inner_tol, 0.5 * need);
        CgState hinit{};
        hinit.maxit = std::max(1, o.maxit - total_inner);
        hinit.null_space = null;
        hinit.n = n;
        hinit.status = CG_RUNNING;
        hinit.stall = 256;
        SE_HIP(hipMemcpyAsync(st, &hinit, sizeof(CgState), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_in_start, dim3(vb), dim3(kVB), 0, s, n, op->r.as<double>(), outer, dinv,
                           op->lo_r.as<float>(), op->lo_x.as<float>(), part + L.b);
        hipLaunchKernelGGL((k_in_dir<T>), dim3(vb), dim3(kVB), 0, s, (int)IN_FIRST, n, tol, dinv, op->lo_r.as<float>(),
                           op->lo_p.as<T>(), part + L.b, vb, st);
        SE_HIP(hipGetLastError());
        for (;;) {
            rc = inner_batch<T, TQ>(op, k);
            if (rc) return rc;
            SE_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
            SE_HIP(hipStreamSynchronize(s));
            if (hs->stop) break;
        }
        total_inner += hs->iter;
        if (hs->status == 3) {
            status = 3;
            break;
        }
        // a stalled inner solve that ended above its starting residual holds
        // no usable correction
        const bool stalled = hs->status == 4;
        if (stalled && !(hs->rr < hs->bb)) {
            status = 4;
            break;
        }
        // x += ||r|| d ; r = b - A x (exact FP64 operator)
        hipLaunchKernelGGL(k_ir_xadd, dim3(std::max(1, std::min((n + 255) / 256, 2048))), dim3(256), 0, s, n,
                           op->x.as<double>(), op->lo_x.as<float>(), std::sqrt(rr));
        rc = op_apply_dev(op, op->x.as<double>(), op->q.as<double>(), sgn, false);
        if (rc) return rc;
        hipLaunchKernelGGL(k_ir_resid, dim3(vb), dim3(kVB), 0, s, n, op->b.as<double>(), op->q.as<double>(),
                           op->r.as<double>(), part + L.r);
        hipLaunchKernelGGL(k_ir_fin, dim3(1), dim3(kVB), 0, s, n, null, part + L.r, vb, outer);
        SE_HIP(hipGetLastError());
        SE_HIP(hipMemcpyAsync(ho, outer, sizeof(ho), hipMemcpyDeviceToHost, s));
        SE_HIP(hipStreamSynchronize(s));
        const double rr_new = ho[0];
        if (!(rr_new - rr_new == 0.0)) {
            status = 3;
            break;
        }
        const double gain = rr_new / rr;
        rr = rr_new;
        if (hs->status == 2 && gain >= 1.0) {  // inner breakdown without progress
            status = 2;
            break;
        }
        if (gain >= 1.0 || (stalled && gain > 0.25)) {
            status = 4;
            break;
        }
    }
    if (d_x != op->x.as<double>())
        SE_HIP(hipMemcpyAsync(d_x, op->x.ptr, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    SE_HIP(hipEventRecord(c->ev1, s));
    SE_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    info.status = status;
    info.iters = total_inner;
    info.outer_iters = outer_steps;
    info.true_rel_residual = bb > 0.0 ? std::sqrt(std::max(rr, 0.0) / bb) : 0.0;
    info.seconds = ms * 1e-3;
    return SE_OK;
}

}  // namespace

namespace {
__global__ void k_f32_to_bf16(int n, const float* __restrict__ a, __hip_bfloat16* __restrict__ b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = __float2bfloat16(a[i]);
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

namespace se {
int pcg_mixed(se_op* op, const double* d_f, double* d_x, const se_solve_opts& o, se_solve_info& info) {
    if (o.precision == SE_PREC_FP32_IR) return ir_solve<float, float>(op, d_f, d_x, o, info);
    if (o.precision == SE_PREC_BF16_IR) return ir_solve<__hip_bfloat16, float>(op, d_f, d_x, o, info);
    return fail(SE_ERR_ARG, "unknown precision");
}
}  // namespace se
