// tcg_kernels.h -- the inner conjugate-gradient iteration of the
// mixed-precision solve, on 64 x 32 tiles of the unknown grid, two kernels per
// iteration:
//
//   k_tcg_dir_apply : totals of the previous update -> beta, convergence;
//                     p' = z + beta p on the tile and its halo, q = A~ p',
//                     partial p'.q
//   k_tcg_update    : totals of p'.q -> alpha; p = p' (recomputed with the same
//                     instructions), x_lo += alpha p, z -= alpha D~^-1 q,
//                     partials of r.r, r.z, sum z, sum r with r = D~ z
//
// The recursion stores the preconditioned residual z = D~^-1 r (D~ the
// diagonal of the separable operator, rebuilt from the line tables where it
// is needed), so the halo work of k_tcg_dir_apply is one FMA per cell.  The
// direction p is read with its halo by k_tcg_dir_apply while only
// k_tcg_update (which has no halo) writes it, so no workgroup can see a
// neighbour's new values.  Vectors z, q, x_lo are FP32, p is FP32 or BF16;
// every reduction is FP64 and deterministic (consumer-side fixed-order sums).
// FP64 enters through reliable updates driven by the host: x += x_lo in FP64,
// r = b - A x with the exact FP64 operator (k_tcg_flush, op apply,
// k_tcg_resid, k_tcg_true), after which the recurrence continues with the
// same direction.
#pragma once
#include "sep_kernels.h"

struct se_op;

namespace se {

// Tiles of TCG_TX x TCG_TY unknowns, one workgroup of eight waves each: wave w
// takes the 64 columns of half (w & 1) and the TCG_RPT rows of group (w >> 1).
constexpr int TCG_TX = 128, TCG_TY = 32, TCG_NT = 512;
constexpr int TCG_RG = TCG_NT / 64 / (TCG_TX / 64);  // row groups
constexpr int TCG_RPT = TCG_TY / TCG_RG;
constexpr int TCG_PRE = 2;  // preloaded partials per thread and value

// direction kernel modes: FIRST starts a solve (p = z, sets the reference
// norms), NORMAL continues (p = z + beta p), RESET continues the solve with
// p = z (after a reliable update that moved the residual far from the
// recursion's, the old direction is no longer conjugate to anything useful)
enum { TCG_FIRST = 0, TCG_NORMAL = 1, TCG_RESET = 2 };

// Jacobi diagonal of the separable operator from its line factors, and its
// inverse (v_rcp_f32: one instruction, within 1 ulp; the same instructions in
// every kernel, so every kernel sees the same preconditioner).
__device__ __forceinline__ float tcg_diag(float gxN, float sxI, float mxI, float gyN, float syJ, float myJ) {
    return __fmaf_rn(syJ, gxN, __fmaf_rn(sxI, gyN, mxI * myJ));
}
__device__ __forceinline__ float tcg_dinv(float diag) { return __builtin_amdgcn_rcpf(diag); }
// The new direction p' = (z - mu) + beta p from the stored preconditioned
// residual z = D~^-1 r: shared by both kernels so that the halo values one
// workgroup computes equal the owner's stored ones.
template <typename T>
__device__ __forceinline__ float tcg_pnew(float z, float mu, float beta, float pold) {
    return to_f<T>(from_f<T>(__fmaf_rn(beta, pold, z - mu)));
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

// Owned cells of a tile: lane j of wave w owns column I0 + 64 (w & 1) + j and
// rows J0 + (w >> 1) RPT + k.  Tiles are numbered column-major (consecutive
// tiles are vertical neighbours), so that a run of consecutive tiles is a
// vertical strip.
struct TcgTile {
    int I0, J0, I, c, lane, wv, g;  // c: column within the tile, g: row group
    bool colok;
    __device__ __forceinline__ TcgTile(const SepParams& S, int tile) {
        const int tilesY = (S.nU1 + TCG_TY - 1) / TCG_TY;
        I0 = (tile / tilesY) * TCG_TX;
        J0 = (tile % tilesY) * TCG_TY;
        lane = threadIdx.x & 63;
        wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        g = wv >> 1;
        c = 64 * (wv & 1) + lane;
        I = I0 + c;
        colok = I < S.nU0;
    }
    __device__ __forceinline__ int row(int k) const { return J0 + g * TCG_RPT + k; }
};

// One solve: the separable tables, the vectors and the partial-sum areas
// (part: A p.q [tiles] | B r stats [4 tiles] | T FP64 residual stats
// [2 tiles] | S mean of f [256]).
struct TcgMember {
    SepParams S;
    float *z, *q, *xlo;  // z = D~^-1 r, the preconditioned residual of the recursion
    void* p;             // T
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

// Address spaces: the solver vectors are global memory; the band tables are
// read-only, so they are read through the constant space, where a
// wave-uniform address (a row's coefficients) becomes a scalar load and the
// value an SGPR operand, and a per-lane address a vector load.
template <typename T> using gp = __attribute__((address_space(1))) T*;
template <typename T> using cgp = const __attribute__((address_space(1))) T*;
template <typename T> using kp = const __attribute__((address_space(4))) T*;

// element loads/stores of the direction vector through those address spaces
template <typename T> __device__ __forceinline__ float ldf(cgp<T> p, size_t i);
template <> __device__ __forceinline__ float ldf<float>(cgp<float> p, size_t i) { return p[i]; }
template <> __device__ __forceinline__ float ldf<__hip_bfloat16>(cgp<__hip_bfloat16> p, size_t i) {
    return __uint_as_float((unsigned)((cgp<unsigned short>)p)[i] << 16);
}
template <typename T> __device__ __forceinline__ void stf(gp<T> p, size_t i, float v);
template <> __device__ __forceinline__ void stf<float>(gp<float> p, size_t i, float v) { p[i] = v; }
template <> __device__ __forceinline__ void stf<__hip_bfloat16>(gp<__hip_bfloat16> p, size_t i, float v) {
    ((gp<unsigned short>)p)[i] = __bfloat16_as_ushort(__float2bfloat16(v));
}

struct TcgTab {
    kp<float> gx, gy, sx, sy, mx, my;
    __device__ __forceinline__ TcgTab(const SepParams& S)
        : gx((kp<float>)S.gx), gy((kp<float>)S.gy), sx((kp<float>)S.sx), sy((kp<float>)S.sy),
          mx((kp<float>)S.mx), my((kp<float>)S.my) {}
};

// ---------------------------------------------------------------------------
// k_tcg_dir_apply
// ---------------------------------------------------------------------------
template <int N, typename T, int MODE>
struct TcgDir {
    static constexpr bool FIRST = MODE == TCG_FIRST, NOP = MODE != TCG_NORMAL;  // NOP: the old p is not used
    static constexpr int B = 2 * N + 1;
    static constexpr int TY = TCG_TY, TX = TCG_TX, NT = TCG_NT, RG = TCG_RG, RPT = TCG_RPT;
    static constexpr int W = TX + 2 * N, H = TY + 2 * N, WP = W + 1;
    static constexpr int RL = (H + RG - 1) / RG;          // main-column window rows per thread
    static constexpr int NX = (H * 2 * N + NT - 1) / NT;  // right-halo cells per thread

    // ---- the member's totals and state (once per member and workgroup run) --
    struct Setup {
        TcgPart<4> pb;
        TcgPart<2> pt;
        __device__ __forceinline__ void load(const TcgMember& M) {
            pb.load(M.part_b(), M.tiles);
            if (FIRST) pt.load(M.part_t(), M.tiles);
        }
        // totals of the update kernel's partials -> beta, mu; convergence and
        // limits; the member's first tile writes the state.  False: no work.
        __device__ __forceinline__ bool run(const TcgMember& M, int tile, float& fb, float& fmu) {
            CgState* __restrict__ st = M.st;
            const int np = M.tiles;
            // the stop flag, read by one thread and shared through LDS, so that the
            // whole workgroup takes the same branch even if the flag changes meanwhile
            __shared__ int s_stop;
            if (threadIdx.x == 0)
                s_stop = FIRST ? 0 : cg_stopped(st);
            const int null_space = st->null_space, iter = st->iter, maxit = st->maxit;
            const double rho_prev = st->rho_prev, tol2_in = st->tol2;
            double tbv[4];
            pb.reduce(M.part_b(), np, tbv);  // (its barriers publish s_stop)
            if (s_stop) return false;
            const double rr = tbv[0];
            const bool finite = (rr - rr == 0.0) && (tbv[1] - tbv[1] == 0.0);
            const double mu = null_space ? tbv[2] / st->n : 0.0;
            const double rho_new = tbv[1] - mu * tbv[3];
            const bool lead = tile == 0 && threadIdx.x == 0;  // the member's first tile writes its state
            double beta = 0.0;
            if (FIRST) {
                double tt[2];
                pt.reduce(M.part_t(), np, tt);
                const double bb = null_space ? tt[0] - tt[1] * (tt[1] / st->n) : tt[0];
                const bool fin = finite && (bb - bb == 0.0);
                if (lead) {
                    st->bb = bb;
                    st->tol2 = M.rtol * M.rtol * bb;
                    st->rr = rr;
                    st->rho = rho_new;
                    st->beta = 0.0;
                    st->mu = mu;
                    st->iter = 0;
                    st->stop = (bb == 0.0 || !fin) ? 1 : 0;
                    st->status = bb == 0.0 ? 0 : (fin ? CG_RUNNING : 3);
                }
                if (bb == 0.0 || !fin) return false;
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
                if (!finite || conv || over) return false;
                beta = NOP ? 0.0 : rho_new / rho_prev;
            }
            fmu = (float)mu;
            fb = (float)beta;
            return true;
        }
    };

    // ---- one tile ---------------------------------------------------------
    // Tiles are numbered column-major within a member, so a workgroup's run
    // of consecutive items is a vertical strip.  The window of a tile is its
    // 32 owned rows with N halo rows above and below (H rows of 64 + 2N
    // columns); the next tile of the same strip shares 2N of those rows, so
    // it copies them from the previous window and loads and builds only its
    // 32 new rows.  Halo cells outside the domain (non-periodic sides) load a
    // clamped, valid cell: the band tables hold zero weights for neighbours
    // outside the domain, so those finite values never reach q.
    static constexpr int SU = H * WP;  // floats of one window in LDS

    // inputs of the window rows a >= a0 (a0 = 0: every row; a0 = 2N: the
    // rows a strip's next tile adds), held in registers until built
    float zr[RL], pr[RL], zh[NX], ph[NX];
    int a0;

    __device__ __forceinline__ void load(const TcgMember& M, int tile, bool cont) {
        const SepParams& S = M.S;
        const TcgTab tb(S);
        cgp<float> z = (cgp<float>)M.z;
        cgp<T> p = (cgp<T>)M.p;
        const TcgTile t(S, tile);
        const int c = t.c, g = t.g, I0 = t.I0, J0 = t.J0;
        a0 = cont ? 2 * N : 0;
        // main columns: the thread of tile column c holds window column c of
        // rows a0 + g + RG*k
        int Ic0 = I0 - N + c;
        if (S.per0) Ic0 = Ic0 < 0 ? Ic0 + S.nU0 : (Ic0 >= S.nU0 ? Ic0 - S.nU0 : Ic0);
        Ic0 = min(max(Ic0, 0), S.nU0 - 1);
#pragma unroll
        for (int k = 0; k < RL; k++) {
            const int a = a0 + g + RG * k;
            if (a < H) {
                int J = J0 - N + a;
                if (S.per1) J = J < 0 ? J + S.nU1 : (J >= S.nU1 ? J - S.nU1 : J);
                J = min(max(J, 0), S.nU1 - 1);
                const size_t rowo = (size_t)J * S.nU0;
                zr[k] = z[rowo + Ic0];
                if (!NOP) pr[k] = ldf<T>(p, rowo + Ic0);
            }
        }
        // right-halo columns TX .. TX+2N-1 of those rows, spread over all threads
        const int nh = (H - a0) * 2 * N;
#pragma unroll
        for (int k = 0; k < NX; k++) {
            const int i = threadIdx.x + NT * k;
            if (i < nh) {
                const int a = a0 + i / (2 * N), cc = TX + i % (2 * N);
                int J = J0 - N + a, I = I0 - N + cc;
                if (S.per1) J = J < 0 ? J + S.nU1 : (J >= S.nU1 ? J - S.nU1 : J);
                if (S.per0) I = I >= S.nU0 ? I - S.nU0 : I;
                J = min(max(J, 0), S.nU1 - 1);
                I = min(I, S.nU0 - 1);
                const size_t o = (size_t)J * S.nU0 + I;
                zh[k] = z[o];
                ph[k] = NOP ? 0.0f : ldf<T>(p, o);
            }
        }
    }

    // the window in su: rows a < a0 from the previous window prev (its rows
    // TY .. TY+2N-1), rows a >= a0 as p' = z + beta p of the loaded inputs
    __device__ __forceinline__ void build(float* su, const float* prev, float fb, float fmu) const {
        const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int c = 64 * (wv & 1) + (threadIdx.x & 63), g = wv >> 1;
        if (a0 > 0) {
            for (int i = threadIdx.x; i < 2 * N * W; i += NT) {
                const int a = i / W, c = i - a * W;
                su[a * WP + c] = prev[(a + TY) * WP + c];
            }
        }
#pragma unroll
        for (int k = 0; k < RL; k++) {
            const int a = a0 + g + RG * k;
            if (a < H) su[a * WP + c] = tcg_pnew<T>(zr[k], fmu, fb, NOP ? 0.0f : pr[k]);
        }
        const int nh = (H - a0) * 2 * N;
#pragma unroll
        for (int k = 0; k < NX; k++) {
            const int i = threadIdx.x + NT * k;
            if (i < nh) su[(a0 + i / (2 * N)) * WP + TX + i % (2 * N)] = tcg_pnew<T>(zh[k], fmu, fb, ph[k]);
        }
    }

    // the owned column's band row and line factors (the same for every tile
    // of a strip)
    struct Col {
        float gxr[B];
        float sxI, mxI;
        __device__ __forceinline__ void load(const TcgMember& M, int tile) {
            const TcgTab tb(M.S);
            const TcgTile t(M.S, tile);
            const int Ic = t.colok ? t.I : 0;
#pragma unroll
            for (int s = 0; s < B; s++) gxr[s] = tb.gx[(size_t)Ic * B + s];
            sxI = tb.sx[Ic];
            mxI = tb.mx[Ic];
        }
    };

    // q = A~ p' on the owned cells of the built window su, partial p'.q.
    // Two rows per step (not unrolled further): they share one y-window of
    // 2N+2 values, and the rows' band coefficients and line factors are
    // wave-uniform scalar loads (SGPR operands), so LDS carries only the
    // window and the x-neighbours and only a step's operands are live.
    static __device__ __forceinline__ void stencil(const TcgMember& M, int tile, const float* su, const Col& cl) {
        static_assert(RPT % 2 == 0, "rows are taken in pairs");
        const SepParams& S = M.S;
        const TcgTab tb(S);
        gp<float> q = (gp<float>)M.q;
        const TcgTile t(S, tile);
        const int j = t.c, g = t.g, J0 = t.J0;
        double dot = 0.0;
#pragma unroll 1
        for (int k = 0; k < RPT; k += 2) {
            const int row = g * RPT + k;
            const int Ja = min(J0 + row, S.nU1 - 1), Jb = min(J0 + row + 1, S.nU1 - 1);
            const kp<float> ga = tb.gy + (size_t)Ja * B, gb = tb.gy + (size_t)Jb * B;
            const float* yc = &su[row * WP + j + N];    // column j+N, rows row .. row+2N+1
            const float* xa = &su[(row + N) * WP + j];  // row row+N, columns j .. j+2N
            const float* xb = xa + WP;
            float win[2 * N + 2];
#pragma unroll
            for (int s = 0; s < 2 * N + 2; s++) win[s] = yc[s * WP];
            float axa = 0.0f, axb = 0.0f, aya = 0.0f, ayb = 0.0f;
#pragma unroll
            for (int s = 0; s < B; s++) {
                axa += cl.gxr[s] * xa[s];
                axb += cl.gxr[s] * xb[s];
                aya += ga[s] * win[s];
                ayb += gb[s] * win[s + 1];
            }
            const float ua = win[N], ub = win[N + 1];
            const float ya = tb.sy[Ja] * axa + cl.sxI * aya + (cl.mxI * tb.my[Ja]) * ua;
            const float yb = tb.sy[Jb] * axb + cl.sxI * ayb + (cl.mxI * tb.my[Jb]) * ub;
            if (t.colok && J0 + row < S.nU1) {
                q[(size_t)(J0 + row) * S.nU0 + t.I] = ya;
                dot += (double)ua * (double)ya;
            }
            if (t.colok && J0 + row + 1 < S.nU1) {
                q[(size_t)(J0 + row + 1) * S.nU0 + t.I] = yb;
                dot += (double)ub * (double)yb;
            }
        }
        double v[1] = {dot};
        block_store_partials_at<1, TCG_NT>(v, M.part_a(), tile, M.tiles);
    }

    // the next item continues this tile's strip
    static __device__ __forceinline__ bool next_in_strip(const SepParams& S, int tile, int next) {
        const int tilesY = (S.nU1 + TY - 1) / TY;
        return next == tile + 1 && next % tilesY != 0;
    }
};

// one tile per workgroup (single-member kernel)
template <int N, typename T, int MODE>
__device__ __forceinline__ void tcg_dir_apply_tile(const TcgMember& M, int tile) {
    using D = TcgDir<N, T, MODE>;
    __shared__ float s_u[D::SU];
    D d;
    d.load(M, tile, false);
    typename D::Col cl;
    cl.load(M, tile);
    typename D::Setup su;
    su.load(M);
    float fb = 0.0f, fmu = 0.0f;
    if (!su.run(M, tile, fb, fmu)) return;
    d.build(s_u, s_u, fb, fmu);
    __syncthreads();
    D::stencil(M, tile, s_u, cl);
}

// ---------------------------------------------------------------------------
// k_tcg_update
// ---------------------------------------------------------------------------
template <int N, typename T>
struct TcgUpd {
    static constexpr int B = 2 * N + 1, RPT = TCG_RPT;
    struct Setup {
        TcgPart<1> pa;
        __device__ __forceinline__ void load(const TcgMember& M) { pa.load(M.part_a(), M.tiles); }
        // total of p'.q -> alpha; the member's first tile writes the state
        __device__ __forceinline__ bool run(const TcgMember& M, int tile, float& fa, float& fb, float& fmu) {
            CgState* __restrict__ st = M.st;
            __shared__ int s_stop;
            if (threadIdx.x == 0) s_stop = cg_stopped(st);
            const double rho = st->rho, beta = st->beta, mu = st->mu;
            double ta[1];
            pa.reduce(M.part_a(), M.tiles, ta);
            if (s_stop) return false;
            const double pq = ta[0];
            const bool lead = tile == 0 && threadIdx.x == 0;  // the member's first tile writes its state
            if (!(pq > 0.0)) {
                if (lead) {
                    st->pq = pq;
                    st->stop = 1;
                    st->status = (pq - pq == 0.0) ? 2 : 3;
                }
                return false;
            }
            const double alpha = rho / pq;
            if (lead) {
                st->pq = pq;
                st->alpha = alpha;
                st->rho_prev = rho;
                st->iter = st->iter + 1;
            }
            fa = (float)alpha;
            fb = (float)beta;
            fmu = (float)mu;
            return true;
        }
    };

    float pv[RPT], zv[RPT], qv[RPT], xv[RPT], rg[RPT], rs[RPT], rm[RPT];
    float cg, cs, cm;

    __device__ __forceinline__ void load(const TcgMember& M, int tile) {
        const SepParams& S = M.S;
        const TcgTab tb(S);
        const TcgTile t(S, tile);
        const int Ic = t.colok ? t.I : 0;
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const int J = t.row(k);
            const int Jc = J < S.nU1 ? J : 0;
            const size_t o = (size_t)Jc * S.nU0 + Ic;
            pv[k] = ldf<T>((cgp<T>)M.p, o);
            zv[k] = ((cgp<float>)M.z)[o];
            qv[k] = ((cgp<float>)M.q)[o];
            xv[k] = ((cgp<float>)M.xlo)[o];
            rg[k] = tb.gy[(size_t)Jc * B + N];  // wave-uniform: scalar loads
            rs[k] = tb.sy[Jc];
            rm[k] = tb.my[Jc];
        }
        cg = tb.gx[(size_t)Ic * B + N];
        cs = tb.sx[Ic];
        cm = tb.mx[Ic];
    }

    // p = p', x_lo += alpha p, z -= alpha D~^-1 q (r = D~ z); partials of r.r,
    // r.z, sum z, sum r
    __device__ __forceinline__ void compute(const TcgMember& M, int tile, float fa, float fb, float fmu) {
        const SepParams& S = M.S;
        gp<T> p = (gp<T>)M.p;
        gp<float> z = (gp<float>)M.z;
        gp<float> x = (gp<float>)M.xlo;
        const TcgTile t(S, tile);
        double srr = 0.0, srz = 0.0, ssz = 0.0, ssr = 0.0;
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const int J = t.row(k);
            if (t.colok && J < S.nU1) {
                const size_t o = (size_t)J * S.nU0 + t.I;
                const float diag = tcg_diag(cg, cs, cm, rg[k], rs[k], rm[k]);
                const float dinv = tcg_dinv(diag);
                const float pn = tcg_pnew<T>(zv[k], fmu, fb, pv[k]);
                const float xn = __fmaf_rn(fa, pn, xv[k]);
                const float zn = __fmaf_rn(-fa, dinv * qv[k], zv[k]);
                const float rn = diag * zn;
                stf<T>(p, o, pn);
                x[o] = xn;
                z[o] = zn;
                srr += (double)rn * rn;
                srz += (double)rn * zn;
                ssz += zn;
                ssr += rn;
            }
        }
        double v[4] = {srr, srz, ssz, ssr};
        block_store_partials_at<4, TCG_NT>(v, M.part_b(), tile, M.tiles);
    }
};

template <int N, typename T>
__device__ __forceinline__ void tcg_update_tile(const TcgMember& M, int tile) {
    using U = TcgUpd<N, T>;
    U u;
    u.load(M, tile);
    typename U::Setup su;
    su.load(M);
    float fa = 0.0f, fb = 0.0f, fmu = 0.0f;
    if (!su.run(M, tile, fa, fb, fmu)) return;
    u.compute(M, tile, fa, fb, fmu);
}

// ---------------------------------------------------------------------------
// start and reliable update
// ---------------------------------------------------------------------------
// b = sgn f - mu0; x = 0; x_lo = 0; p = 0; z = D~^-1 r_lo with r_lo = b in
// FP32; partials of r_lo, z (part_b) and of b in FP64 (part_t: sum b^2,
// sum b).  RESID: r_lo = b - q64 instead, with x, x_lo, p left alone.
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
    float* __restrict__ zs = M.z;
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
            const float z = tcg_dinv(tcg_diag(cg, cs, cm, S.gy[(size_t)J * B + N], S.sy[J], S.my[J])) * rn;
            zs[o] = z;
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

// Batched kernels: the item count is `nwork`, or *nwork_dev when that is
// given (a work list the host compacts between graph replays).
__device__ __forceinline__ int work_count(int nwork, const int* __restrict__ nwork_dev) {
    return nwork_dev ? __builtin_amdgcn_readfirstlane(*nwork_dev) : nwork;
}

// Each workgroup takes one contiguous run of the work list.  The list is
// ordered by member, then tile, so a run is a strip of neighbouring tiles of
// one or two members: their halo rows meet in one L2, and the member's totals
// are reduced once per run instead of once per tile.  Workgroups are
// dispatched round-robin over the 8 XCDs; the workgroups of one XCD take
// consecutive runs, so that neighbours one tile row apart share that XCD's L2
// too.
__device__ __forceinline__ void tcg_run(int nwork, int& w0, int& w1) {
    const int G = gridDim.x;
    const int r = (G & 7) ? (int)blockIdx.x : (int)(blockIdx.x & 7) * (G >> 3) + (int)(blockIdx.x >> 3);
    const int chunk = (nwork + G - 1) / G;
    w0 = min(nwork, r * chunk);
    w1 = min(nwork, w0 + chunk);
}

template <int N, typename T, int MODE>
__global__ __launch_bounds__(TCG_NT) void k_tcg_dir_apply_batch(const TcgMember* __restrict__ Ms,
                                                                const int2* __restrict__ work, int nwork,
                                                                const int* __restrict__ nwork_dev) {
    using D = TcgDir<N, T, MODE>;
    __shared__ float s_u[2][D::SU];  // windows of consecutive tiles (ping-pong)
    nwork = work_count(nwork, nwork_dev);
    int w0, w1;
    tcg_run(nwork, w0, w1);
    if (w0 >= w1) return;
    int cur = -1;
    bool go = false;
    float fb = 0.0f, fmu = 0.0f;
    D d;
    typename D::Col cl;
    int2 it = work[w0];
    d.load(Ms[it.x], it.y, false);
    for (int w = w0; w < w1; w++) {
        const TcgMember& M = Ms[it.x];
        const bool cont = d.a0 > 0;  // this tile continues the previous tile's strip
        if (!cont) cl.load(M, it.y);
        if (it.x != cur) {  // a new member: its totals
            typename D::Setup su;
            su.load(M);
            go = su.run(M, it.y, fb, fmu);
            cur = it.x;
        }
        float* sb = s_u[w & 1];
        if (go) d.build(sb, s_u[(w & 1) ^ 1], fb, fmu);
        __syncthreads();  // the window is built
        // the next item's inputs are loaded while this tile computes
        int2 itn = it;
        if (w + 1 < w1) {
            itn = work[w + 1];
            const bool contn = go && itn.x == it.x && D::next_in_strip(M.S, it.y, itn.y);
            d.load(Ms[itn.x], itn.y, contn);
        }
        if (go) D::stencil(M, it.y, sb, cl);  // (ends with a barrier)
        it = itn;
    }
}
template <int N, typename T>
__global__ __launch_bounds__(TCG_NT) void k_tcg_update_batch(const TcgMember* __restrict__ Ms,
                                                             const int2* __restrict__ work, int nwork,
                                                             const int* __restrict__ nwork_dev) {
    using U = TcgUpd<N, T>;
    nwork = work_count(nwork, nwork_dev);
    int w0, w1;
    tcg_run(nwork, w0, w1);
    int cur = -1;
    bool go = false;
    float fa = 0.0f, fb = 0.0f, fmu = 0.0f;
    for (int w = w0; w < w1; w++) {
        const int2 it = work[w];
        const TcgMember& M = Ms[it.x];
        __syncthreads();
        U u;
        u.load(M, it.y);
        if (it.x != cur) {
            typename U::Setup su;
            su.load(M);
            go = su.run(M, it.y, fa, fb, fmu);
            cur = it.x;
        }
        if (go) u.compute(M, it.y, fa, fb, fmu);
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
