// tcg_kernels.h -- the inner conjugate-gradient iteration of the
// mixed-precision solve, on 64 x 32 tiles of the unknown grid, two kernels per
// iteration:
//
//   k_tcg_dir_apply : totals of the previous update -> beta, convergence;
//                     p' = z + beta p on the tile and its halo (z = D~^-1 r,
//                     D~ the diagonal of the separable operator, computed from
//                     the line tables), q = A~ p', partial p'.q
//   k_tcg_update    : totals of p'.q -> alpha; p = p' (recomputed with the same
//                     instructions), x_lo += alpha p, r -= alpha q, partials of
//                     r.r, r.z, sum z, sum r
//
// The direction p is read with its halo by k_tcg_dir_apply while only
// k_tcg_update (which has no halo) writes it, so no workgroup can see a
// neighbour's new values.  Vectors r, q, x_lo are FP32, p is FP32 or BF16;
// every reduction is FP64 and deterministic (consumer-side fixed-order sums).
// FP64 enters through reliable updates driven by the host: x += x_lo in FP64,
// r = b - A x with the exact FP64 operator (k_tcg_flush, op apply,
// k_tcg_resid, k_tcg_true), after which the recurrence continues with the
// same direction.
#pragma once
#include "sep_kernels.h"

struct se_op;

namespace se {

constexpr int TCG_NT = 256, TCG_TY = 32, TCG_RPT = TCG_TY / (TCG_NT / 64);
constexpr int TCG_PRE = 4;  // preloaded partials per thread and value

// direction kernel modes: FIRST starts a solve (p = z, sets the reference
// norms), NORMAL continues (p = z + beta p), RESET continues the solve with
// p = z (after a reliable update that moved the residual far from the
// recursion's, the old direction is no longer conjugate to anything useful)
enum { TCG_FIRST = 0, TCG_NORMAL = 1, TCG_RESET = 2 };

// Jacobi diagonal inverse and the new direction: shared by both kernels so
// that the halo values one workgroup computes equal the owner's stored ones.
// (v_rcp_f32: one instruction, within 1 ulp; the same instruction in every
// kernel, so every kernel sees the same preconditioner)
__device__ __forceinline__ float tcg_dinv(float gxN, float sxI, float mxI, float gyN, float syJ, float myJ) {
    return __builtin_amdgcn_rcpf(__fmaf_rn(syJ, gxN, __fmaf_rn(sxI, gyN, mxI * myJ)));
}
template <typename T>
__device__ __forceinline__ float tcg_pnew(float r, float dinv, float mu, float beta, float pold) {
    return to_f<T>(from_f<T>(__fmaf_rn(beta, pold, __fmaf_rn(r, dinv, -mu))));
}

// Partial sums of the producer kernel, preloaded into registers together with
// the consumer's data loads, then reduced in a fixed order (identical totals
// in every workgroup).
template <int K>
struct TcgPart {
    double v[K][TCG_PRE];
    __device__ __forceinline__ void load(const double* __restrict__ part, int np) {
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int c = 0; c < TCG_PRE; c++) {
                const int b = threadIdx.x + c * TCG_NT;
                v[k][c] = b < np ? part[(size_t)k * np + b] : 0.0;
            }
    }
    __device__ __forceinline__ void reduce(const double* __restrict__ part, int np, double (&tot)[K]) {
        constexpr int NW = TCG_NT / 64;
        __shared__ double s_red[K][NW];
        __shared__ double s_tot[K];
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < K; k++) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < TCG_PRE; c++) s += v[k][c];
            for (int b = threadIdx.x + TCG_PRE * TCG_NT; b < np; b += TCG_NT) s += part[(size_t)k * np + b];
            s = wave_sum(s);
            if (lane == 0) s_red[k][wid] = s;
        }
        __syncthreads();
        if (threadIdx.x < K) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < NW; w++) s += s_red[threadIdx.x][w];
            s_tot[threadIdx.x] = s;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; k++) tot[k] = s_tot[k];
    }
};

// Owned cells of a tile: lane j owns column I0 + j, wave g rows J0 + g*RPT + k.
struct TcgTile {
    int I0, J0, I, lane, wv;
    bool colok;
    __device__ __forceinline__ TcgTile(const SepParams& S, int tile) {
        const int tilesX = (S.nU0 + SEP_TX - 1) / SEP_TX;
        I0 = (tile % tilesX) * SEP_TX;
        J0 = (tile / tilesX) * TCG_TY;
        lane = threadIdx.x & 63;
        wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        I = I0 + lane;
        colok = I < S.nU0;
    }
    __device__ __forceinline__ int row(int k) const { return J0 + wv * TCG_RPT + k; }
};

// One solve: the separable tables, the vectors and the partial-sum areas
// (part: A p.q [tiles] | B r stats [4 tiles] | T FP64 residual stats
// [2 tiles] | S mean of f [256]).
struct TcgMember {
    SepParams S;
    float *r, *q, *xlo;
    void* p;  // T
    double *b, *x, *q64;
    const double* f;
    double* part;
    CgState* st;
    int tiles, np_s, null_space, n;
    double sgn, rtol;
    __device__ __forceinline__ double* part_a() const { return part; }
    __device__ __forceinline__ double* part_b() const { return part + tiles; }
    __device__ __forceinline__ double* part_t() const { return part + 5 * (size_t)tiles; }
    __device__ __forceinline__ double* part_s() const { return part + 7 * (size_t)tiles; }
};

// ---------------------------------------------------------------------------
// k_tcg_dir_apply
// ---------------------------------------------------------------------------
template <int N, typename T, int MODE>
__device__ __forceinline__ void tcg_dir_apply_tile(const TcgMember& M, int tile) {
    constexpr bool FIRST = MODE == TCG_FIRST, NOP = MODE != TCG_NORMAL;  // NOP: the old p is not used
    const SepParams& S = M.S;
    const float* __restrict__ r = M.r;
    const T* __restrict__ p = (const T*)M.p;
    float* __restrict__ q = M.q;
    const int np = M.tiles;
    const double* __restrict__ part_b = M.part_b();
    const double* __restrict__ part_t = M.part_t();
    double* __restrict__ part_a = M.part_a();
    CgState* __restrict__ st = M.st;
    const double rtol = M.rtol;
    constexpr int B = 2 * N + 1;
    constexpr int TY = TCG_TY, NT = TCG_NT, NW = NT / 64, RPT = TCG_RPT;
    constexpr int W = SEP_TX + 2 * N, H = TY + 2 * N, WP = W + 1;
    constexpr int RL = (H + NW - 1) / NW;
    constexpr int NG = (TY * B + NT - 1) / NT;
    __shared__ float s_u[H * WP];
    __shared__ float s_gy[TY * B];
    __shared__ float s_cg[W], s_cs[W], s_cm[W];  // per tile column: gx centre, sx, mx
    __shared__ float s_rg[H], s_rs[H], s_rm[H];  // per tile row: gy centre, sy, my

    const TcgTile t(S, tile);
    const int lane = t.lane, wv = t.wv, I0 = t.I0, J0 = t.J0;

    // ---- every global load first ----
    int Ic0 = I0 - N + lane, Ic1 = I0 - N + 64 + lane;
    bool c0 = true, c1 = lane < 2 * N;
    if (S.per0) {
        Ic0 = Ic0 < 0 ? Ic0 + S.nU0 : (Ic0 >= S.nU0 ? Ic0 - S.nU0 : Ic0);
        Ic1 = Ic1 >= S.nU0 ? Ic1 - S.nU0 : Ic1;
    }
    c0 = c0 && Ic0 >= 0 && Ic0 < S.nU0;
    c1 = c1 && Ic1 >= 0 && Ic1 < S.nU0;
    Ic0 = c0 ? Ic0 : 0;
    Ic1 = c1 ? Ic1 : 0;
    float r0[RL], r1[RL], p0[RL], p1[RL];
    bool rok[RL];
#pragma unroll
    for (int k = 0; k < RL; k++) {
        const int a = wv + NW * k;
        int J = J0 - N + a;
        if (S.per1) J = J < 0 ? J + S.nU1 : (J >= S.nU1 ? J - S.nU1 : J);
        rok[k] = a < H && J >= 0 && J < S.nU1;
        const int Jc = rok[k] ? J : 0;
        const size_t rowo = (size_t)Jc * S.nU0;
        r0[k] = r[rowo + Ic0];
        r1[k] = r[rowo + Ic1];
        if (!NOP) {
            p0[k] = to_f<T>(p[rowo + Ic0]);
            p1[k] = to_f<T>(p[rowo + Ic1]);
        }
    }
    // row tables of tile row `lane` (H <= 64 rows)
    float rgl, rsl, rml;
    {
        int J = J0 - N + lane;
        if (S.per1) J = J < 0 ? J + S.nU1 : (J >= S.nU1 ? J - S.nU1 : J);
        const int Jc = (lane < H && J >= 0 && J < S.nU1) ? J : 0;
        rgl = S.gy[(size_t)Jc * B + N];
        rsl = S.sy[Jc];
        rml = S.my[Jc];
    }
    const float cg0 = S.gx[(size_t)Ic0 * B + N], cs0 = S.sx[Ic0], cm0 = S.mx[Ic0];
    const float cg1 = S.gx[(size_t)Ic1 * B + N], cs1 = S.sx[Ic1], cm1 = S.mx[Ic1];
    float gyv[NG];
#pragma unroll
    for (int k = 0; k < NG; k++) {
        const int i = threadIdx.x + NT * k;
        const bool ok = i < TY * B && J0 + i / B < S.nU1;
        const float v = S.gy[ok ? (size_t)J0 * B + i : 0];
        gyv[k] = ok ? v : 0.0f;
    }
    const int Ic = t.colok ? t.I : 0;
    float gxr[B];
#pragma unroll
    for (int s = 0; s < B; s++) gxr[s] = S.gx[(size_t)Ic * B + s];
    const float sxI = S.sx[Ic], mxI = S.mx[Ic];
    TcgPart<4> pb;
    pb.load(part_b, np);
    TcgPart<2> pt;
    if (FIRST) pt.load(part_t, np);
    // the stop flag, read by one thread and shared through LDS, so that the
    // whole workgroup takes the same branch even if the flag changes meanwhile
    __shared__ int s_stop;
    if (threadIdx.x == 0) s_stop = FIRST ? 0 : __hip_atomic_load(&st->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int null_space = st->null_space, iter = st->iter, maxit = st->maxit;
    const double rho_prev = st->rho_prev, tol2_in = st->tol2;

    // ---- tables to LDS, totals, state ----
    if (wv == 0 && lane < H) {
        s_rg[lane] = rgl;
        s_rs[lane] = rsl;
        s_rm[lane] = rml;
    }
    s_cg[lane] = cg0, s_cs[lane] = cs0, s_cm[lane] = cm0;
    if (lane < 2 * N) s_cg[64 + lane] = cg1, s_cs[64 + lane] = cs1, s_cm[64 + lane] = cm1;
#pragma unroll
    for (int k = 0; k < NG; k++) {
        const int i = threadIdx.x + NT * k;
        if (i < TY * B) s_gy[i] = gyv[k];
    }
    double tb[4];
    pb.reduce(part_b, np, tb);  // (its barriers publish the tables and s_stop)
    if (s_stop) return;
    const double rr = tb[0];
    const bool finite = (rr - rr == 0.0) && (tb[1] - tb[1] == 0.0);
    const double mu = null_space ? tb[2] / st->n : 0.0;
    const double rho_new = tb[1] - mu * tb[3];
    const bool lead = tile == 0 && threadIdx.x == 0;  // the member's first tile writes its state
    double beta = 0.0;
    if (FIRST) {
        double tt[2];
        pt.reduce(part_t, np, tt);
        const double bb = null_space ? tt[0] - tt[1] * (tt[1] / st->n) : tt[0];
        const bool fin = finite && (bb - bb == 0.0);
        if (lead) {
            st->bb = bb;
            st->tol2 = rtol * rtol * bb;
            st->rr = rr;
            st->rho = rho_new;
            st->beta = 0.0;
            st->mu = mu;
            st->iter = 0;
            st->stop = (bb == 0.0 || !fin) ? 1 : 0;
            st->status = bb == 0.0 ? 0 : (fin ? CG_RUNNING : 3);
        }
        if (bb == 0.0 || !fin) return;
    } else {
        const bool conv = finite && rr <= tol2_in;
        const bool over = iter >= maxit;
        if (lead) {
            st->rr = rr;
            if (!finite) {
                st->stop = 1;
                st->status = 3;
            } else if (conv) {
                st->stop = 1;
                st->status = CG_PENDING;
            } else if (over) {
                st->stop = 1;
                st->status = 1;
            } else {
                st->rho = rho_new;
                st->beta = NOP ? 0.0 : rho_new / rho_prev;
                st->mu = mu;
            }
        }
        if (!finite || conv || over) return;
        beta = NOP ? 0.0 : rho_new / rho_prev;
    }
    const float fmu = (float)mu, fb = (float)beta;

    // ---- new direction on the tile and its halo ----
#pragma unroll
    for (int k = 0; k < RL; k++) {
        const int a = wv + NW * k;
        if (a < H) {
            const float d0 = tcg_dinv(s_cg[lane], s_cs[lane], s_cm[lane], s_rg[a], s_rs[a], s_rm[a]);
            const float v0 = tcg_pnew<T>(r0[k], d0, fmu, fb, NOP ? 0.0f : p0[k]);
            s_u[a * WP + lane] = (rok[k] && c0) ? v0 : 0.0f;
            if (lane < 2 * N) {
                const float d1 =
                    tcg_dinv(s_cg[64 + lane], s_cs[64 + lane], s_cm[64 + lane], s_rg[a], s_rs[a], s_rm[a]);
                const float v1 = tcg_pnew<T>(r1[k], d1, fmu, fb, NOP ? 0.0f : p1[k]);
                s_u[a * WP + 64 + lane] = (rok[k] && c1) ? v1 : 0.0f;
            }
        }
    }
    __syncthreads();

    // ---- q = A~ p' on the owned cells ----
    const int j = lane, g = wv;
    float ax[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const float* rowp = &s_u[(g * RPT + k + N) * WP + j];
        float acc = 0.0f;
#pragma unroll
        for (int s = 0; s < B; s++) acc += gxr[s] * rowp[s];
        ax[k] = acc;
    }
    float win[RPT + 2 * N];
#pragma unroll
    for (int tt = 0; tt < RPT + 2 * N; tt++) win[tt] = s_u[(g * RPT + tt) * WP + j + N];
    double dot = 0.0;
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int rr_ = g * RPT + k;
        const int J = J0 + rr_;
        const float* gy = &s_gy[rr_ * B];
        float acc = 0.0f;
#pragma unroll
        for (int s = 0; s < B; s++) acc += gy[s] * win[k + s];
        const float u = win[k + N];
        const float y = s_rs[rr_ + N] * ax[k] + sxI * acc + (mxI * s_rm[rr_ + N]) * u;
        if (t.colok && J < S.nU1) {
            q[(size_t)J * S.nU0 + t.I] = y;
            dot += (double)u * (double)y;
        }
    }
    double v[1] = {dot};
    block_store_partials_at<1, NT>(v, part_a, tile, np);
}

// ---------------------------------------------------------------------------
// k_tcg_update
// ---------------------------------------------------------------------------
template <int N, typename T>
__device__ __forceinline__ void tcg_update_tile(const TcgMember& M, int tile) {
    const SepParams& S = M.S;
    T* __restrict__ p = (T*)M.p;
    float* __restrict__ r = M.r;
    const float* __restrict__ q = M.q;
    float* __restrict__ x = M.xlo;
    const int np = M.tiles;
    const double* __restrict__ part_a = M.part_a();
    double* __restrict__ part_b = M.part_b();
    CgState* __restrict__ st = M.st;
    constexpr int B = 2 * N + 1, RPT = TCG_RPT;
    const TcgTile t(S, tile);
    const int Ic = t.colok ? t.I : 0;
    float pv[RPT], rv[RPT], qv[RPT], xv[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int J = t.row(k);
        const int Jc = J < S.nU1 ? J : 0;
        const size_t o = (size_t)Jc * S.nU0 + Ic;
        pv[k] = to_f<T>(p[o]);
        rv[k] = r[o];
        qv[k] = q[o];
        xv[k] = x[o];
    }
    // row tables of the wave's row `lane` (lanes < RPT), read back per row
    // with readlane
    float rgl, rsl, rml;
    {
        const int J = t.row(t.lane < RPT ? t.lane : 0);
        const int Jc = J < S.nU1 ? J : 0;
        rgl = S.gy[(size_t)Jc * B + N];
        rsl = S.sy[Jc];
        rml = S.my[Jc];
    }
    const float cg = S.gx[(size_t)Ic * B + N], cs = S.sx[Ic], cm = S.mx[Ic];
    TcgPart<1> pa;
    pa.load(part_a, np);
    __shared__ int s_stop;
    if (threadIdx.x == 0) s_stop = __hip_atomic_load(&st->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const double rho = st->rho, beta = st->beta, mu = st->mu;
    double ta[1];
    pa.reduce(part_a, np, ta);
    if (s_stop) return;
    const double pq = ta[0];
    const bool lead = tile == 0 && threadIdx.x == 0;  // the member's first tile writes its state
    if (!(pq > 0.0)) {
        if (lead) {
            st->pq = pq;
            st->stop = 1;
            st->status = (pq - pq == 0.0) ? 2 : 3;
        }
        return;
    }
    const double alpha = rho / pq;
    if (lead) {
        st->pq = pq;
        st->alpha = alpha;
        st->rho_prev = rho;
        st->iter = st->iter + 1;
    }
    const float fa = (float)alpha, fb = (float)beta, fmu = (float)mu;
    double srr = 0.0, srz = 0.0, ssz = 0.0, ssr = 0.0;
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int J = t.row(k);
        if (t.colok && J < S.nU1) {
            const size_t o = (size_t)J * S.nU0 + t.I;
            const float dinv = tcg_dinv(cg, cs, cm, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rgl), k)),
                                        __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rsl), k)),
                                        __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rml), k)));
            const float pn = tcg_pnew<T>(rv[k], dinv, fmu, fb, pv[k]);
            const float xn = __fmaf_rn(fa, pn, xv[k]);
            const float rn = __fmaf_rn(-fa, qv[k], rv[k]);
            const float z = dinv * rn;
            p[o] = from_f<T>(pn);
            x[o] = xn;
            r[o] = rn;
            srr += (double)rn * rn;
            srz += (double)rn * z;
            ssz += z;
            ssr += rn;
        }
    }
    double v[4] = {srr, srz, ssz, ssr};
    block_store_partials_at<4, TCG_NT>(v, part_b, tile, np);
}

// ---------------------------------------------------------------------------
// start and reliable update
// ---------------------------------------------------------------------------
// b = sgn f - mu0; x = 0; x_lo = 0; p = 0; r_lo = b; partials of r_lo (part_b)
// and of b in FP64 (part_t: sum b^2, sum b).  RESID: r = b - q64 instead, with
// x, x_lo, p left alone.
template <int N, typename T, bool RESID>
__device__ __forceinline__ void tcg_start_tile(const TcgMember& M, int tile) {
    const SepParams& S = M.S;
    const double* __restrict__ f = M.f;
    const double sgn = M.sgn;
    const double* __restrict__ part_s = M.part_s();
    const int np_s = M.np_s;
    double* __restrict__ b = M.b;
    const double* __restrict__ q64 = M.q64;
    double* __restrict__ x = M.x;
    float* __restrict__ xlo = M.xlo;
    T* __restrict__ p = (T*)M.p;
    float* __restrict__ r = M.r;
    double* __restrict__ part_b = M.part_b();
    double* __restrict__ part_t = M.part_t();
    const int null_space = M.null_space;
    constexpr int B = 2 * N + 1, RPT = TCG_RPT;
    const TcgTile t(S, tile);
    const int Ic = t.colok ? t.I : 0;
    const size_t n = (size_t)S.nU0 * S.nU1;
    double mu0 = 0.0;
    if (!RESID && null_space) {
        double tot[1];
        reduce_partials<1, TCG_NT>(part_s, np_s, tot);
        mu0 = tot[0] / (double)n;
    }
    const float cg = S.gx[(size_t)Ic * B + N], cs = S.sx[Ic], cm = S.mx[Ic];
    double srr = 0.0, srz = 0.0, ssz = 0.0, ssr = 0.0, sbb = 0.0, sb = 0.0;
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int J = t.row(k);
        if (t.colok && J < S.nU1) {
            const size_t o = (size_t)J * S.nU0 + t.I;
            double bi;
            if (RESID) {
                bi = b[o] - q64[o];
            } else {
                bi = sgn * f[o] - mu0;
                b[o] = bi;
                x[o] = 0.0;
                xlo[o] = 0.0f;
                p[o] = from_f<T>(0.0f);
            }
            const float rn = (float)bi;
            r[o] = rn;
            const float dinv = tcg_dinv(cg, cs, cm, S.gy[(size_t)J * B + N], S.sy[J], S.my[J]);
            const float z = dinv * rn;
            srr += (double)rn * rn;
            srz += (double)rn * z;
            ssz += z;
            ssr += rn;
            sbb += bi * bi;
            sb += bi;
        }
    }
    double v[4] = {srr, srz, ssz, ssr};
    block_store_partials_at<4, TCG_NT>(v, part_b, tile, M.tiles);
    double w[2] = {sbb, sb};
    block_store_partials_at<2, TCG_NT>(w, part_t, tile, M.tiles);
}

// x += x_lo (FP64); x_lo = 0, on elements [i0, i1) of a member
__device__ __forceinline__ void tcg_flush_range(const TcgMember& M, int i0, int i1, int stride) {
    for (int i = i0 + threadIdx.x; i < i1; i += stride) {
        M.x[i] += (double)M.xlo[i];
        M.xlo[i] = 0.0f;
    }
}

// st->tt = ||r - mean||^2 (or ||r||^2) from part_t (one workgroup)
__device__ __forceinline__ void tcg_true_member(const TcgMember& M) {
    double t[2];
    reduce_partials<2, TCG_NT>(M.part_t(), M.tiles, t);
    if (threadIdx.x == 0) M.st->tt = M.st->null_space ? t[0] - t[1] * (t[1] / M.st->n) : t[0];
}

// ---------------------------------------------------------------------------
// kernels: one member (descriptor as kernel argument, one tile per workgroup)
// and batches (descriptor array, (member, tile) work items, the grid loops
// over the items)
// ---------------------------------------------------------------------------
template <int N, typename T, int MODE>
__global__ __launch_bounds__(TCG_NT) void k_tcg_dir_apply(TcgMember M) {
    tcg_dir_apply_tile<N, T, MODE>(M, blockIdx.x);
}
template <int N, typename T>
__global__ __launch_bounds__(TCG_NT) void k_tcg_update(TcgMember M) {
    tcg_update_tile<N, T>(M, blockIdx.x);
}
template <int N, typename T, bool RESID>
__global__ __launch_bounds__(TCG_NT) void k_tcg_start(TcgMember M) {
    tcg_start_tile<N, T, RESID>(M, blockIdx.x);
}
static __global__ __launch_bounds__(TCG_NT) void k_tcg_flush(TcgMember M) {
    const int n = M.n;
    const int per = (n + gridDim.x - 1) / gridDim.x;
    const int i0 = blockIdx.x * per;
    tcg_flush_range(M, i0, min(n, i0 + per), TCG_NT);
}
static __global__ __launch_bounds__(TCG_NT) void k_tcg_true(TcgMember M) { tcg_true_member(M); }

template <int N, typename T, int MODE>
__global__ __launch_bounds__(TCG_NT) void k_tcg_dir_apply_batch(const TcgMember* __restrict__ Ms,
                                                                const int2* __restrict__ work, int nwork) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        __syncthreads();
        tcg_dir_apply_tile<N, T, MODE>(Ms[it.x], it.y);
    }
}
template <int N, typename T>
__global__ __launch_bounds__(TCG_NT) void k_tcg_update_batch(const TcgMember* __restrict__ Ms,
                                                             const int2* __restrict__ work, int nwork) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        __syncthreads();
        tcg_update_tile<N, T>(Ms[it.x], it.y);
    }
}
template <int N, typename T, bool RESID>
__global__ __launch_bounds__(TCG_NT) void k_tcg_start_batch(const TcgMember* __restrict__ Ms,
                                                            const int2* __restrict__ work, int nwork) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        __syncthreads();
        tcg_start_tile<N, T, RESID>(Ms[it.x], it.y);
    }
}
// work items (member, chunk of TCG_FLUSH elements)
constexpr int TCG_FLUSH = 4096;
static __global__ __launch_bounds__(TCG_NT) void k_tcg_flush_batch(const TcgMember* __restrict__ Ms,
                                                            const int2* __restrict__ work, int nwork) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        const TcgMember& M = Ms[it.x];
        const int i0 = it.y * TCG_FLUSH;
        tcg_flush_range(M, i0, min(M.n, i0 + TCG_FLUSH), TCG_NT);
    }
}
// work items: member indices
static __global__ __launch_bounds__(TCG_NT) void k_tcg_true_batch(const TcgMember* __restrict__ Ms,
                                                           const int* __restrict__ list, int count) {
    for (int w = blockIdx.x; w < count; w += gridDim.x) {
        __syncthreads();
        tcg_true_member(Ms[list[w]]);
    }
}

// descriptor of an operator's mixed-precision solve (host, mixed.hip)
TcgMember tcg_member(se_op* op, const double* d_f, double rtol, int null_space, int np_s);

}  // namespace se
