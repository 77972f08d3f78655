// fdm.hip -- fast-diagonalisation (FDM) preconditioner on the matrix cores
// and the FDM-preconditioned conjugate-gradient solves.
//
// The operator in its separable form (mixed.hip, assemble_line) has the
// near-exact inverse (Vy (x) Vx) D^-1 (Vy (x) Vx)^T (fdm_host.h).  On the
// unknown grid R (rows J, columns I) applying it is four dense transforms:
//   pass 1  T1t = Vx^T R^T  (= (R Vx)^T)      x-transform
//   pass 2  T2  = (T1t Vy) ./ D               y-transform, scaled by 1/D
//   pass 3  T3t = Vy T2^T   (= (T2 Vy^T)^T)   y back-transform
//   pass 4  Z   = T3t Vx^T                    x back-transform
// each an "NT" GEMM C = A Bt^T whose operands both have the summed index
// contiguous, on 64 x 64 output tiles (four waves of 32 x 32) with MFMA:
// v_mfma_f64_16x16x4f64 (FP64), v_mfma_f32_32x32x2f32 (FP32) or
// v_mfma_f32_32x32x16_bf16 (BF16 operands, FP32 accumulation).  Matrices are
// padded to multiples of 64 with zeros; pass 1 reads the FP64 residual with
// its true extent and pass 4 writes the FP64 result with it, together with
// per-tile partial sums of r.z, sum z and sum r.
//
// The solver around it is FP64 CG (x0 = 0, true-residual stop, constant null
// space by Euclidean-mean projection, as the Jacobi-PCG of helm.hip) whose
// preconditioner is the FDM in the solve's precision (or, in the FP64_*C
// precisions, FP64 for the first application and FP32 / BF16 operands for
// the corrections after it).  It runs members of a batch in lockstep, one
// launch per phase over (member, item) work lists, one host pull of the
// members' states per round; a single solve is a batch of one, so batched
// results are the single ones bit for bit.  The exact FP64 operator apply is
// k_apply_lines_batch; every reduction is per item with consumer-side
// fixed-order sums.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "apply_kernels.h"
#include "apply_lines.h"
#include "fdm_host.h"
#include "se_common.h"
#include "se_op.h"

namespace se {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// global-memory views of the operands: loads and stores through them are
// global_* instructions.  Through generic pointers they would be flat_*, which
// share the LDS counter, so every wait for an LDS operand read would also wait
// for the next chunk's prefetch (SE_GLOBAL, apply_kernels.h).

enum { FDM_OUT_PLAIN = 0, FDM_OUT_SCALE = 1, FDM_OUT_FINAL = 2 };
enum { FDM_F64 = 0, FDM_F32 = 1, FDM_BF16 = 2 };  // element types of the transforms

// One GEMM of one member: C = A Bt^T.  A: M x K (lda), Bt: N x K (ldb), both
// row-major, element type T.  Pass 1's Bt is the residual packed to T and
// padded by k_fdm_pack from r (true extent arows x acols, leading dimension
// acols).
struct GemmTask {
    const void* A;
    const void* Bt;
    void* C;
    const double* r;                       // FINAL: residual (for the partials)
    double* part;                          // FINAL: [3][tiles] partials r.z, sum z, sum r
    const double *lx, *ly, *ex, *ey;       // SCALE: D = lx[row] + ly[col] + ex[row] ey[col]
    double dtol;                           // SCALE: D below this is a null mode (1/D := 0)
    int M, N, K, lda, ldb, ldc;
    int arows, acols;
    int crows, ccols;                      // valid extent (SCALE: modes; FINAL: the unknown grid)
    int mode, tilesN;
    // the member's CG state and r.r partials (null outside a CG solve): an
    // iteration's transforms skip a member that no longer needs z
    const double* pv;
    const CgState* st;
    int nch;
    // FINAL, launched for a CG's first or restarted direction: z goes straight
    // to the CG's p (its direction, p = z - mean z), not to z; for the first,
    // the residual of the partials is b (r = b is not stored)
    double* C2;
    const double* b;
};
// where pass 4 of a CG's transforms writes: z, or p (restart: from r; start: from b)
enum { FDM_TO_Z = 0, FDM_TO_P = 1, FDM_TO_P_FROM_B = 2 };

// Output tiles are 64 x 64, four waves of 32 x 32 each.
constexpr int FDM_BT = 64;

// K-chunk staging.  A chunk of one operand tile is 64 rows x 128 bytes in
// LDS, copied straight from global memory by global_load_lds_dwordx4 (no
// register staging).  Two stages alternate: the copy of the next chunk runs
// during the current chunk's MFMAs, and one barrier per chunk separates them.
// One such instruction fills 64 consecutive 16-byte slots (lane i: slot i);
// the lanes choose which global vectors land there, which places vector c of
// row r at slot r VR + (c ^ swz(r)).  An MFMA operand read takes one vector
// from each of 16 consecutive rows; the swizzle puts them in 16 different
// 16-byte bank groups (plain 128-byte rows would share banks).  Every padded
// dimension is a multiple of 64, so KC divides 64.  (Rows of 256 bytes, 64 KB
// per workgroup, measured slower: two workgroups per CU instead of four.)
template <typename T>
struct GemmCfg {
    static constexpr int RB = 128;                         // bytes per row and chunk
    static constexpr int VR = RB / 16;                     // 16-byte vectors per row
    static constexpr int KC = RB / (int)sizeof(T);         // K per chunk
    static constexpr int VEC = 16 / (int)sizeof(T);        // elements per vector
    static constexpr int OPB = FDM_BT * RB;                // bytes of one operand's chunk
    static constexpr int NI = FDM_BT * VR / 256;           // copies per thread and operand
    static constexpr int STAGE = 2 * OPB;                  // both operands
    static_assert(64 % KC == 0 && NI * 256 == FDM_BT * VR, "chunk shape");
};
template <typename T>
__device__ __forceinline__ int swz(int r) {
    static_assert(GemmCfg<T>::VR == 8, "two rows per 256 bytes of banks");
    return (r >> 1) & 7;
}

// MFMA block edge per element type and the accumulators of one wave
template <typename T> struct MfmaBlock { static constexpr int E = 32; };   // 32 x 32 (FP32 / BF16 operands)
template <> struct MfmaBlock<double> { static constexpr int E = 16; };     // 16 x 16 (FP64)
template <typename T>
struct Acc {
    static constexpr int NB = (FDM_BT / 2) / MfmaBlock<T>::E;  // blocks per wave and dimension
    typename std::conditional<std::is_same<T, double>::value, f64x4, f32x16>::type c[NB][NB];
};
// one 16-byte operand vector
template <typename T> struct Vec16;
template <> struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct Vec16<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Vec16<__bf16> { typedef bf16x8 type; };

template <typename T> __device__ __forceinline__ T to_t(double v);
template <> __device__ __forceinline__ double to_t<double>(double v) { return v; }
template <> __device__ __forceinline__ float to_t<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ __bf16 to_t<__bf16>(double v) { return (__bf16)(float)v; }

// the MFMAs of one operand vector per block pair
__device__ __forceinline__ void mfma_vec(const Vec16<double>::type& a, const Vec16<double>::type& b, f64x4& c) {
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], b[1], c, 0, 0, 0);
}
__device__ __forceinline__ void mfma_vec(const Vec16<float>::type& a, const Vec16<float>::type& b, f32x16& c) {
#pragma unroll
    for (int e = 0; e < 4; e++) c = __builtin_amdgcn_mfma_f32_32x32x2f32(a[e], b[e], c, 0, 0, 0);
}
__device__ __forceinline__ void mfma_vec(const bf16x8& a, const bf16x8& b, f32x16& c) {
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// one K chunk of the wave's 32 x 32 block.  The lanes split into G = 64 / E
// groups along K (FP64: quarters l >> 4; FP32 / BF16: halves l >> 5); group g
// takes the same VR / G of the chunk's vectors from both operands, so the sum
// covers the chunk once: FP32 / BF16 the run [g VR / G, (g + 1) VR / G), FP64
// the vectors g + G v (its quarters read one row each and, on runs, meet in
// the same banks: measured 4 conflict cycles per LDS cycle, none when
// interleaved).  The wave's rows wr 32 + i E + (l mod E) differ by multiples
// of 16, so they share one swizzle.  The operand reads
// run RA vectors ahead of the MFMAs that use them (the scheduler would
// otherwise issue each read just before its MFMAs and wait for it there).
template <typename T>
__device__ __forceinline__ void mma_chunk(const char* As, const char* Bs, int wr, int wc, int l, Acc<T>& acc) {
    using C = GemmCfg<T>;
    using V = typename Vec16<T>::type;
    constexpr int E = MfmaBlock<T>::E, G = 64 / E, NB = Acc<T>::NB, VPG = C::VR / G, RA = 2;
    const int lr = l % E, g = l / E, sw = swz<T>(lr);
    const char* a = As + (wr * 32 + lr) * C::RB;
    const char* b = Bs + (wc * 32 + lr) * C::RB;
    V av[VPG][NB], bv[VPG][NB];
    auto read = [&](int v) {
        const int c = std::is_same<T, double>::value ? g + G * v : g * VPG + v;
        const int cs = (c ^ sw) * 16;
#pragma unroll
        for (int i = 0; i < NB; i++) {
            av[v][i] = *(const V*)(a + i * E * C::RB + cs);
            bv[v][i] = *(const V*)(b + i * E * C::RB + cs);
        }
    };
#pragma unroll
    for (int v = 0; v < RA; v++) read(v);
#pragma unroll
    for (int v = 0; v < VPG; v++) {
        if (v + RA < VPG) read(v + RA);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) mfma_vec(av[v][i], bv[v][j], acc.c[i][j]);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// every accumulator element with its (row, column) in the 64 x 64 tile:
// 32 x 32 blocks hold rows (v & 3) + 4 (l >> 5) + 8 (v >> 2), column l & 31;
// 16 x 16 FP64 blocks rows (l >> 4) + 4 v, column l & 15 (both measured with
// a one-wave product of random operands)
template <typename T, typename F>
__device__ __forceinline__ void for_each_acc(const Acc<T>& acc, int wr, int wc, int l, F&& f) {
    constexpr int NB = Acc<T>::NB;
#pragma unroll
    for (int bi = 0; bi < NB; bi++)
#pragma unroll
        for (int bj = 0; bj < NB; bj++) {
            if constexpr (std::is_same<T, double>::value) {
#pragma unroll
                for (int v = 0; v < 4; v++)
                    f(wr * 32 + 16 * bi + (l >> 4) + 4 * v, wc * 32 + 16 * bj + (l & 15), acc.c[bi][bj][v]);
            } else {
#pragma unroll
                for (int v = 0; v < 16; v++)
                    f(wr * 32 + 32 * bi + (v & 3) + 4 * (l >> 5) + 8 * (v >> 2), wc * 32 + 32 * bj + (l & 31),
                      (double)acc.c[bi][bj][v]);
            }
        }
}

// The copies of one chunk of one operand tile: thread-invariant element
// offsets (row ld + column) of the NI vectors this lane copies, and the
// wave-uniform LDS slot runs they fill.
template <typename T>
struct ChunkCopy {
    using C = GemmCfg<T>;
    int off[C::NI];
    __device__ __forceinline__ void init(int ld, int wv, int l) {
#pragma unroll
        for (int j = 0; j < C::NI; j++) {
            const int s = (wv * C::NI + j) * 64 + l, r = s / C::VR;
            off[j] = r * ld + ((s % C::VR) ^ swz<T>(r)) * C::VEC;
        }
    }
    __device__ __forceinline__ void issue(const SE_GLOBAL T* src, int k0, char* S, int wv) const {
#pragma unroll
        for (int j = 0; j < C::NI; j++)
            __builtin_amdgcn_global_load_lds((const SE_GLOBAL void*)(src + off[j] + k0),
                                             (__attribute__((address_space(3))) void*)(S + (wv * C::NI + j) * 1024), 16,
                                             0, 0);
    }
};

// the copies this thread issued have landed in LDS
__device__ __forceinline__ void wait_copies() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

template <typename T>
__device__ __forceinline__ void gemm_tile(const GemmTask& t, int tile, char* S0, char* S1, int to_p) {
    using C = GemmCfg<T>;
    constexpr int NB = Acc<T>::NB;
    const int tid = threadIdx.x, l = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wv >> 1, wc = wv & 1;
    const int tm = tile / t.tilesN, tn = tile - tm * t.tilesN;
    const int m0 = tm * FDM_BT, n0 = tn * FDM_BT;
    Acc<T> acc;
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if constexpr (std::is_same<T, double>::value) acc.c[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
            else
                for (int v = 0; v < 16; v++) acc.c[i][j][v] = 0.0f;
        }
    const SE_GLOBAL T* A = (const SE_GLOBAL T*)t.A + (size_t)m0 * t.lda;
    const SE_GLOBAL T* B = (const SE_GLOBAL T*)t.Bt + (size_t)n0 * t.ldb;
    ChunkCopy<T> ca, cb;
    ca.init(t.lda, wv, l);
    cb.init(t.ldb, wv, l);
    const int nk = t.K / C::KC;
    __syncthreads();  // the previous item is done reading both stages
    ca.issue(A, 0, S0, wv);
    cb.issue(B, 0, S0 + C::OPB, wv);
    for (int kc = 0; kc < nk; kc += 2) {
        wait_copies();
        __syncthreads();  // chunk kc is in S0 for every wave; S1's last reads are done
        if (kc + 1 < nk) {
            ca.issue(A, (kc + 1) * C::KC, S1, wv);
            cb.issue(B, (kc + 1) * C::KC, S1 + C::OPB, wv);
        }
        mma_chunk<T>(S0, S0 + C::OPB, wr, wc, l, acc);
        if (kc + 1 >= nk) break;
        wait_copies();
        __syncthreads();  // chunk kc + 1 is in S1; S0's reads are done
        if (kc + 2 < nk) {
            ca.issue(A, (kc + 2) * C::KC, S0, wv);
            cb.issue(B, (kc + 2) * C::KC, S0 + C::OPB, wv);
        }
        mma_chunk<T>(S1, S1 + C::OPB, wr, wc, l, acc);
    }
    if (t.mode == FDM_OUT_PLAIN) {
        SE_GLOBAL T* Cm = (SE_GLOBAL T*)t.C;
        for_each_acc<T>(acc, wr, wc, l,
                        [&](int row, int col, double v) { Cm[(size_t)(m0 + row) * t.ldc + n0 + col] = to_t<T>(v); });
    } else if (t.mode == FDM_OUT_SCALE) {
        SE_GLOBAL T* Cm = (SE_GLOBAL T*)t.C;
        const SE_GLOBAL double* lx = (const SE_GLOBAL double*)t.lx;
        const SE_GLOBAL double* ly = (const SE_GLOBAL double*)t.ly;
        const SE_GLOBAL double* ex = (const SE_GLOBAL double*)t.ex;
        const SE_GLOBAL double* ey = (const SE_GLOBAL double*)t.ey;
        for_each_acc<T>(acc, wr, wc, l, [&](int row, int col, double v) {
            const int gr = m0 + row, gc = n0 + col;
            double s = 0.0;
            if (gr < t.crows && gc < t.ccols) {
                const double D = lx[gr] + ly[gc] + ex[gr] * ey[gc];
                s = D > t.dtol ? 1.0 / D : 0.0;
            }
            Cm[(size_t)gr * t.ldc + gc] = to_t<T>(v * s);
        });
    } else {  // FDM_OUT_FINAL: FP64 z on the unknown grid, partials of r.z, sum z, sum r
        SE_GLOBAL double* Z = (SE_GLOBAL double*)(to_p != FDM_TO_Z ? t.C2 : t.C);
        const SE_GLOBAL double* R = (const SE_GLOBAL double*)(to_p == FDM_TO_P_FROM_B ? t.b : t.r);
        double rz = 0.0, sz = 0.0, sr = 0.0;
        for_each_acc<T>(acc, wr, wc, l, [&](int row, int col, double v) {
            const int gr = m0 + row, gc = n0 + col;
            if (gr < t.crows && gc < t.ccols) {
                const size_t i = (size_t)gr * t.ldc + gc;
                Z[i] = v;
                const double ri = R[i];
                rz += ri * v;
                sz += v;
                sr += ri;
            }
        });
        const double pv[3] = {rz, sz, sr};
        block_store_partials_at<3, 256>(pv, t.part, tile, (t.M / FDM_BT) * t.tilesN);
    }
}

// Inside a CG iteration (skip != 0): a member whose recursive residual
// already meets its target, or whose iteration has stopped, needs no new
// z = FDM(r).  The direction kernel that follows decides the same from the
// same partials summed in the same order (and leaves z unused), so every
// workgroup of the member's transforms skips it without a host round trip.
__device__ __forceinline__ bool fdm_skip(const GemmTask& t) {
    if (!t.pv) return false;
    if (cg_stopped(t.st)) return true;
    double tr[1];
    reduce_partials<1, 256>(t.pv, t.nch, tr);
    const double rr = tr[0];
    return !(rr - rr == 0.0) || rr <= t.st->tol2;
}

// pass `pass` of the listed members' transforms; work items (member, tile),
// or tiles of member 0 when work is null.  Two stages of 16 KB each: four
// workgroups per CU (one wave of each per SIMD, at most 128 VGPRs).
template <typename T>
__global__ __launch_bounds__(256, 4) void k_fdm_gemm(const GemmTask* __restrict__ tasks, int pass,
                                                    const int2* __restrict__ work, int nwork, int skip, int to_p) {
    __shared__ __attribute__((aligned(1024))) char S0[GemmCfg<T>::STAGE];
    __shared__ __attribute__((aligned(1024))) char S1[GemmCfg<T>::STAGE];
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work ? work[w] : make_int2(0, w);
        if (it.x < 0) continue;  // a padding item of the XCD-ordered list
        const GemmTask t = tasks[it.x * 4 + pass];
        if (skip && fdm_skip(t)) continue;
        gemm_tile<T>(t, it.y, S0, S1, to_p);
    }
}

// pass 1's Bt operand: the FP64 residual r (nU1 x nU0) rounded to T into the
// padded nP1 x nP0 array, zeros outside; one 64 x 64 block per item
template <typename T>
__global__ __launch_bounds__(256) void k_fdm_pack(const GemmTask* __restrict__ tasks, const int2* __restrict__ work,
                                                  int nwork, int skip) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work ? work[w] : make_int2(0, w);
        if (it.x < 0) continue;
        const GemmTask& t = tasks[it.x * 4];
        if (skip && fdm_skip(t)) continue;
        const int tilesN = t.K / FDM_BT;  // pass 1: K = nP0 columns of the packed array
        const int tm = it.y / tilesN, tn = it.y - tm * tilesN;
        SE_GLOBAL T* P = (SE_GLOBAL T*)t.Bt;
        const SE_GLOBAL double* r = (const SE_GLOBAL double*)t.r;
        for (int e = threadIdx.x; e < FDM_BT * FDM_BT; e += 256) {
            const int gr = tm * FDM_BT + e / FDM_BT, gc = tn * FDM_BT + e % FDM_BT;
            const double v = (gr < t.arows && gc < t.acols) ? r[(size_t)gr * t.acols + gc] : 0.0;
            P[(size_t)gr * t.ldb + gc] = to_t<T>(v);
        }
    }
}

// ---------------------------------------------------------------------------
// CG vector kernels over (member, chunk) items
// ---------------------------------------------------------------------------
constexpr int FB_NT = 256;
constexpr int FB_CH = 4096;  // unknowns per chunk

struct FbMember {
    const double* f;
    double *b, *x, *r, *p, *q, *z;
    double* pv;  // [nch]      r.r after start / update / restart
    double* pz;  // [3][ntz]   FDM partials r.z, sum z, sum r
    double* pa;  // [nap]      p.q
    double* pt;  // [2][nap]   sum t, sum t^2 of the true residual
    double* ps;  // [nch]      sum of sgn f
    double* xo;  //            the caller's solution vector
    void* rp;    //            r in the transforms' type T, padded: pass 1's operand
    void* rpc;   //            the same for the corrections' transforms (after an update)
    CgState* st;
    int n, nch, ntz, null_space;
    int fused;  // p.q and the true residual from the apply's epilogue (no SVV term)
    int nap;    // partials of p.q and of the true residual: apply tiles (fused), else nch
    int nU0, nP0, rprec;  //   unknown-grid row length, padded row length, T (FDM_F64 / F32 / BF16)
    int rcprec;           //   T of the corrections' transforms
    double sgn, rtol;
};

// r[i] into the packed residual (row i / nU0, column i mod nU0), rounded to T
// as the pack kernel rounds it
__device__ __forceinline__ void put_packed_to(const FbMember& M, void* rp, int prec, int i, double v) {
    const int J = i / M.nU0, I = i - J * M.nU0;
    const size_t k = (size_t)J * M.nP0 + I;
    if (prec == FDM_F64) ((SE_GLOBAL double*)rp)[k] = v;
    else if (prec == FDM_F32) ((SE_GLOBAL float*)rp)[k] = to_t<float>(v);
    else ((SE_GLOBAL __bf16*)rp)[k] = to_t<__bf16>(v);
}
// for the first (or a restarted) direction's transforms
__device__ __forceinline__ void put_packed(const FbMember& M, int i, double v) { put_packed_to(M, M.rp, M.rprec, i, v); }
// for a correction's transforms (the residual of an update)
__device__ __forceinline__ void put_packed_c(const FbMember& M, int i, double v) {
    put_packed_to(M, M.rpc, M.rcprec, i, v);
}

__device__ __forceinline__ void chunk_range(const FbMember& M, int ch, int& i0, int& i1) {
    i0 = ch * FB_CH;
    i1 = min(M.n, i0 + FB_CH);
}

__global__ __launch_bounds__(FB_NT) void k_fb_sum(const FbMember* __restrict__ Ms, const int2* __restrict__ work,
                                                  int nwork) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        __syncthreads();  // the previous item's reduction is done with its LDS
        const int2 it = work[w];
        const FbMember& M = Ms[it.x];
        int i0, i1;
        chunk_range(M, it.y, i0, i1);
        double s = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += FB_NT) s += M.sgn * M.f[i];
        const double v[1] = {s};
        block_store_partials_at<1, FB_NT>(v, M.ps, it.y, M.nch);
    }
}

// b = sgn f - mean (null space); partials of b.b.  (r = b and x = 0 are
// implicit: the first update reads b for r and writes x = alpha p, and a
// member stopped before it gets x = 0 from the first direction kernel or the
// update)
__global__ __launch_bounds__(FB_NT) void k_fb_start(const FbMember* __restrict__ Ms, const int2* __restrict__ work,
                                                    int nwork) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        const FbMember& M = Ms[it.x];
        __syncthreads();
        double mu = 0.0;
        if (M.null_space) {
            double t[1];
            reduce_partials<1, FB_NT>(M.ps, M.nch, t);
            mu = t[0] / M.n;
        }
        int i0, i1;
        chunk_range(M, it.y, i0, i1);
        double rr = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += FB_NT) {
            const double bi = M.sgn * M.f[i] - mu;
            if (M.b != M.f) M.b[i] = bi;  // (else b is f, and bi = f[i])
            put_packed(M, i, bi);
            rr += bi * bi;
        }
        const double v[1] = {rr};
        block_store_partials_at<1, FB_NT>(v, M.pv, it.y, M.nch);
    }
}

enum { FB_FIRST = 0, FB_NORMAL = 1, FB_RESTART = 2 };

// totals of r.r and of the FDM partials -> convergence, beta;
// p = (z - mean z) + beta p.  The member's chunk 0 writes its state.
__global__ __launch_bounds__(FB_NT) void k_fb_dir(const FbMember* __restrict__ Ms, const int2* __restrict__ work,
                                                  int nwork, int mode) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        const FbMember& M = Ms[it.x];
        CgState* __restrict__ st = M.st;
        __syncthreads();
        __shared__ int s_stop;
        if (threadIdx.x == 0) s_stop = mode == FB_NORMAL ? cg_stopped(st) : 0;
        double tr[1], tz[3];
        reduce_partials<1, FB_NT>(M.pv, M.nch, tr);  // (its barriers publish s_stop)
        if (s_stop) continue;
        reduce_partials<3, FB_NT>(M.pz, M.ntz, tz);
        const double rr = tr[0];
        const bool finite = (rr - rr == 0.0) && (tz[0] - tz[0] == 0.0);
        const double mu = M.null_space ? tz[1] / M.n : 0.0;
        const double rho = tz[0] - mu * tz[2];
        const bool lead = it.y == 0 && threadIdx.x == 0;
        double beta = 0.0;
        bool go = true;
        if (mode == FB_FIRST) {
            if (lead) {
                st->bb = rr;
                st->tol2 = M.rtol * M.rtol * rr;
                st->rr = rr;
                st->rho = rho;
                st->mu = mu;
                st->iter = 0;
                st->stop = (rr == 0.0 || !finite) ? 1 : 0;
                st->status = rr == 0.0 ? 0 : (finite ? CG_RUNNING : 3);
            }
            go = rr != 0.0 && finite;
        } else if (mode == FB_NORMAL) {
            const bool conv = finite && rr <= st->tol2;
            if (lead) {
                const int itn = st->iter + 1;
                const bool over = itn >= st->maxit;
                st->iter = itn;
                st->rr = rr;
                if (!finite) {
                    st->stop = 1;
                    st->status = 3;
                } else if (conv && itn == 1) {
                    // from x = 0 the first iterate is x = alpha p (one rounding)
                    // and r = b - alpha (A p): r is already an evaluation of
                    // b - A x, with the rounding a separate one would carry
                    st->stop = 1;
                    st->status = 0;
                    st->tt = rr;
                    st->mu_t = 0.0;
                } else if (conv) {
                    st->stop = 1;
                    st->status = CG_PENDING;
                } else if (over) {
                    st->stop = 1;
                    st->status = 1;
                } else {
                    st->rho = rho;
                    st->mu = mu;
                    st->beta = rho / st->rho_prev;
                }
            }
            // (the iteration limit is the lead's decision alone: other workgroups
            // may already see its new iteration count)
            go = finite && !conv;
            beta = rho / st->rho_prev;
        } else {  // restart from the true residual: p = z - mean z
            if (lead) {
                st->rr = rr;
                st->rho = rho;
                st->mu = mu;
                st->beta = 0.0;
            }
        }
        int i0, i1;
        chunk_range(M, it.y, i0, i1);
        if (!go) {
            // stopped before any update (b = 0, or not finite): x = 0
            if (mode == FB_FIRST)
                for (int i = i0 + threadIdx.x; i < i1; i += FB_NT) M.x[i] = 0.0;
            continue;
        }
        if (mode == FB_NORMAL) {
            for (int i = i0 + threadIdx.x; i < i1; i += FB_NT) M.p[i] = (M.z[i] - mu) + beta * M.p[i];
        } else if (mu != 0.0) {  // the transforms wrote z into p
            for (int i = i0 + threadIdx.x; i < i1; i += FB_NT) M.p[i] -= mu;
        }
    }
}

// partials of p.q (members with an SVV term)
__global__ __launch_bounds__(FB_NT) void k_fb_dot(const FbMember* __restrict__ Ms, const int2* __restrict__ work,
                                                  int nwork) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        __syncthreads();  // the previous item's reduction is done with its LDS
        const int2 it = work[w];
        const FbMember& M = Ms[it.x];
        if (M.fused) continue;  // the apply formed its partials
        int i0, i1;
        chunk_range(M, it.y, i0, i1);
        double s = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += FB_NT) s += M.p[i] * M.q[i];
        const double v[1] = {s};
        block_store_partials_at<1, FB_NT>(v, M.pa, it.y, M.nch);
    }
}

// alpha = rho / p.q; x += alpha p, r -= alpha q; partials of r.r
__global__ __launch_bounds__(FB_NT) void k_fb_update(const FbMember* __restrict__ Ms, const int2* __restrict__ work,
                                                     int nwork) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        const FbMember& M = Ms[it.x];
        CgState* __restrict__ st = M.st;
        __syncthreads();
        __shared__ int s_stop;
        if (threadIdx.x == 0) s_stop = cg_stopped(st);
        double t[1];
        reduce_partials<1, FB_NT>(M.pa, M.nap, t);
        if (s_stop) continue;
        const double pq = t[0];
        const bool lead = it.y == 0 && threadIdx.x == 0;
        const bool first = st->iter == 0;  // x = 0 until this update
        int i0, i1;
        chunk_range(M, it.y, i0, i1);
        if (!(pq > 0.0)) {
            if (lead) {
                st->pq = pq;
                st->stop = 1;
                st->status = (pq - pq == 0.0) ? 2 : 3;
            }
            if (first)
                for (int i = i0 + threadIdx.x; i < i1; i += FB_NT) M.x[i] = 0.0;
            continue;
        }
        const double rho = st->rho;
        const double alpha = rho / pq;
        if (lead) {
            st->pq = pq;
            st->alpha = alpha;
            st->rho_prev = rho;
        }
        double rr = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += FB_NT) {
            M.x[i] = first ? alpha * M.p[i] : M.x[i] + alpha * M.p[i];
            const double ri = (first ? M.b[i] : M.r[i]) - alpha * M.q[i];
            M.r[i] = ri;
            put_packed_c(M, i, ri);
            rr += ri * ri;
        }
        const double v[1] = {rr};
        block_store_partials_at<1, FB_NT>(v, M.pv, it.y, M.nch);
    }
}

// true residual t = b - q (q = A x): r = t; partials of sum t, sum t^2
// (members with an SVV term)
__global__ __launch_bounds__(FB_NT) void k_fb_true(const FbMember* __restrict__ Ms, const int2* __restrict__ work,
                                                   int nwork, int gated) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        __syncthreads();  // the previous item's reduction is done with its LDS
        const int2 it = work[w];
        const FbMember& M = Ms[it.x];
        if (M.fused) continue;  // the apply formed r and its partials
        if (gated && M.st->status != CG_PENDING) continue;
        int i0, i1;
        chunk_range(M, it.y, i0, i1);
        double s = 0.0, s2 = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += FB_NT) {
            const double ti = M.b[i] - M.q[i];
            M.r[i] = ti;
            s += ti;
            s2 += ti * ti;
        }
        const double v[2] = {s, s2};
        block_store_partials_at<2, FB_NT>(v, M.pt, it.y, M.nch);
    }
}

// per member (one workgroup each): ||t - mean t||^2 into the state
__global__ __launch_bounds__(FB_NT) void k_fb_true_fin(const FbMember* __restrict__ Ms, const int* __restrict__ list,
                                                       int count, int gated) {
    for (int w = blockIdx.x; w < count; w += gridDim.x) {
        const FbMember& M = Ms[list[w]];
        if (gated && M.st->status != CG_PENDING) continue;
        __syncthreads();
        double t[2];
        reduce_partials<2, FB_NT>(M.pt, M.nap, t);
        if (threadIdx.x == 0) {
            const double mt = M.null_space ? t[0] / M.n : 0.0;
            M.st->mu_t = mt;
            M.st->tt = M.null_space ? t[1] - t[0] * mt : t[1];
        }
    }
}

// x -> the caller's solution vector, all members in one launch
__global__ __launch_bounds__(FB_NT) void k_fb_out(const FbMember* __restrict__ Ms, const int2* __restrict__ work,
                                                  int nwork) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int2 it = work[w];
        const FbMember& M = Ms[it.x];
        int i0, i1;
        chunk_range(M, it.y, i0, i1);
        for (int i = i0 + threadIdx.x; i < i1; i += FB_NT) M.xo[i] = M.x[i];
    }
}

// continue from the true residual: r -= mean (null space); partials of r.r
__global__ __launch_bounds__(FB_NT) void k_fb_restart(const FbMember* __restrict__ Ms, const int2* __restrict__ work,
                                                      int nwork) {
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        __syncthreads();  // the previous item's reduction is done with its LDS
        const int2 it = work[w];
        const FbMember& M = Ms[it.x];
        const double mt = M.null_space ? M.st->mu_t : 0.0;
        int i0, i1;
        chunk_range(M, it.y, i0, i1);
        double rr = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += FB_NT) {
            const double ri = M.r[i] - mt;
            M.r[i] = ri;
            put_packed(M, i, ri);
            rr += ri * ri;
        }
        const double v[1] = {rr};
        block_store_partials_at<1, FB_NT>(v, M.pv, it.y, M.nch);
    }
}

// ---------------------------------------------------------------------------
// host: per-operator FDM data
// ---------------------------------------------------------------------------
int pad64(int n) { return std::max(64, (n + 63) / 64 * 64); }

struct FdmPlan;
void fdm_plan_destroy(FdmPlan* pl);

// A 1-D eigenvector matrix on the device in one transform precision, padded
// to nP x nP, and its transpose.  It depends only on the line operator (the
// mesh line and c_lap, not c_mass), so every operator of the same line on one
// device shares one copy: an ensemble of 512 members has a dozen distinct
// ones.
struct VMats {
    DevBuf v, vT;
};

struct FdmOp {
    FdmPlan* self = nullptr;  // the one-member plan of this operator's own solves
    int nU0 = 0, nU1 = 0, nP0 = 0, nP1 = 0;
    double dtol = 0.0;
    DevBuf lx, ly, ex, ey;   // FP64 eigenvalues and mass factors, padded
    DevBuf z, pv, pz, pa, pt, ps;
    DevBuf st;               // CgState
    struct Prec {
        std::shared_ptr<VMats> mx, my;  // Vx, Vy (shared)
        DevBuf t1, t2;
        DevBuf rp;  // the residual in T, padded nP1 x nP0 (zeros outside nU1 x nU0): pass 1's Bt
        bool ready = false;
    } prec[3];
    std::shared_ptr<const se_fdm::Fdm1D> ex1, ey1;
};

void fdm_release(se_op* op) {
    if (op->fdm) {
        if (op->fdm->self) fdm_plan_destroy(op->fdm->self);
        delete op->fdm;
    }
    op->fdm = nullptr;
}

namespace {

// eigen-decompositions of (G, S) shared by operators of the same line
// (keyed by the exact bytes of G and S)
std::mutex g_eig_mu;
std::map<std::string, std::shared_ptr<se_fdm::Fdm1D>> g_eig_cache;

std::shared_ptr<const se_fdm::Fdm1D> eig_1d(int nU, int N, int per, const std::vector<double>& g,
                                         const std::vector<double>& s) {
    std::string key((const char*)g.data(), g.size() * sizeof(double));
    key.append((const char*)s.data(), s.size() * sizeof(double));
    key.append((const char*)&N, sizeof(int));
    key.append((const char*)&per, sizeof(int));
    {
        std::lock_guard<std::mutex> lk(g_eig_mu);
        auto it = g_eig_cache.find(key);
        if (it != g_eig_cache.end()) return it->second;
    }
    auto e = std::make_shared<se_fdm::Fdm1D>();
    std::vector<double> zero(nU, 0.0);
    if (!se_fdm::fdm_1d(nU, N, per, g, s, zero, *e)) return nullptr;
    std::lock_guard<std::mutex> lk(g_eig_mu);
    g_eig_cache[key] = e;
    return e;
}

uint16_t bf16_bits(double v) {
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

// device copy of a padded FP64 matrix (rows x cols) in element type `prec`,
// transposed when `trans`
int upload_matrix(DevBuf& d, const std::vector<double>& h, int rows, int cols, bool trans, int prec) {
    const size_t cnt = (size_t)rows * cols;
    std::vector<double> t(cnt);
    for (int i = 0; i < rows; i++)
        for (int j = 0; j < cols; j++) t[trans ? (size_t)j * rows + i : (size_t)i * cols + j] = h[(size_t)i * cols + j];
    if (prec == FDM_F64) {
        int rc = d.alloc(sizeof(double) * cnt);
        if (rc) return rc;
        SE_HIP(hipMemcpy(d.ptr, t.data(), sizeof(double) * cnt, hipMemcpyHostToDevice));
    } else if (prec == FDM_F32) {
        std::vector<float> f(cnt);
        for (size_t i = 0; i < cnt; i++) f[i] = (float)t[i];
        int rc = d.alloc(sizeof(float) * cnt);
        if (rc) return rc;
        SE_HIP(hipMemcpy(d.ptr, f.data(), sizeof(float) * cnt, hipMemcpyHostToDevice));
    } else {
        std::vector<uint16_t> f(cnt);
        for (size_t i = 0; i < cnt; i++) f[i] = bf16_bits(t[i]);
        int rc = d.alloc(sizeof(uint16_t) * cnt);
        if (rc) return rc;
        SE_HIP(hipMemcpy(d.ptr, f.data(), sizeof(uint16_t) * cnt, hipMemcpyHostToDevice));
    }
    return SE_OK;
}

size_t elem_size(int prec) { return prec == FDM_F64 ? 8 : (prec == FDM_F32 ? 4 : 2); }

std::mutex g_vm_mu;
std::map<std::tuple<const se_fdm::Fdm1D*, int, int>, std::weak_ptr<VMats>> g_vm_cache;

// the shared device copy of e's eigenvectors in precision prec on device dev
int vmats_of(const std::shared_ptr<const se_fdm::Fdm1D>& e, int nP, int prec, int dev, std::shared_ptr<VMats>& out) {
    std::lock_guard<std::mutex> lk(g_vm_mu);
    const auto key = std::make_tuple(e.get(), prec, dev);
    auto it = g_vm_cache.find(key);
    if (it != g_vm_cache.end() && (out = it->second.lock())) return SE_OK;
    for (auto i = g_vm_cache.begin(); i != g_vm_cache.end();) i = i->second.expired() ? g_vm_cache.erase(i) : std::next(i);
    const int nU = e->nU;
    std::vector<double> h((size_t)nP * nP, 0.0);
    for (int i = 0; i < nU; i++)
        for (int k = 0; k < nU; k++) h[(size_t)i * nP + k] = e->V[(size_t)i * nU + k];
    auto vm = std::make_shared<VMats>();
    int rc = upload_matrix(vm->v, h, nP, nP, false, prec);
    if (!rc) rc = upload_matrix(vm->vT, h, nP, nP, true, prec);
    if (rc) return rc;
    g_vm_cache[key] = vm;
    out = vm;
    return SE_OK;
}

}  // namespace

// FDM tables of op (FP64 part: eigenvectors, eigenvalues, mass factors and
// work vectors) and the transforms in precision `prec`
int fdm_prepare(se_op* op, int prec) {
    // An SVV term (c (Mf (x) Kf), dense within each element in y) is outside
    // the separable form: the FDM is then that of the operator without it,
    // still a close preconditioner (the SVV term lives on the high modes;
    // kappa(A0^-1 A) ~ 1-3 on the reference's velocity solves, against tens
    // to hundreds for Jacobi), while the CG applies the exact operator.
    SE_CHECK(op->definite, SE_ERR_UNSUPPORTED, "FDM preconditioner: the operator is not definite");
    SE_CHECK(op->c_lap != 0.0, SE_ERR_UNSUPPORTED,
             "FDM preconditioner: a mass-only operator is diagonal (use Jacobi, which is exact for it)");
    const se_mesh* m = op->mesh;
    int rc = op_ensure_work(op);
    if (rc) return rc;
    if (!op->fdm) {
        std::unique_ptr<FdmOp> F(new FdmOp);
        const int N = m->N;
        std::vector<double> gx, X0, X1, MX, gy, Y0, Y1, MY;
        assemble_line(m, op, 0, gx, X0, X1, MX);
        assemble_line(m, op, 1, gy, Y0, Y1, MY);
        const double sgn = (double)op->sign;
        // sign*A = Sy (x) Gx + Gy (x) Sx + My (x) Mx  (the separable tables of mixed.hip)
        std::vector<double> sx(m->nU[0]), mx(m->nU[0]), sy(m->nU[1]), my(m->nU[1]);
        for (int I = 0; I < m->nU[0]; I++) {
            sx[I] = sgn * op->c_lap * X1[I];
            mx[I] = sgn * op->c_mass * MX[I];
        }
        for (int J = 0; J < m->nU[1]; J++) {
            sy[J] = sgn * op->c_lap * Y0[J];
            my[J] = MY[J];
        }
        F->ex1 = eig_1d(m->nU[0], N, m->periodic[0], gx, sx);
        F->ey1 = eig_1d(m->nU[1], N, m->periodic[1], gy, sy);
        SE_CHECK(F->ex1 && F->ey1, SE_ERR_STATE, "FDM preconditioner: 1-D eigenproblem failed");
        F->nU0 = m->nU[0];
        F->nU1 = m->nU[1];
        F->nP0 = pad64(F->nU0);
        F->nP1 = pad64(F->nU1);
        const se_fdm::Fdm1D &Ex = *F->ex1, &Ey = *F->ey1;
        std::vector<double> lx(F->nP0, 0.0), ex(F->nP0, 0.0), ly(F->nP1, 0.0), ey(F->nP1, 0.0);
        double lmax = 0.0, emx = 0.0, emy = 0.0;
        for (int k = 0; k < F->nU0; k++) {
            lx[k] = Ex.lambda[k];
            double e = 0.0;
            for (int i = 0; i < F->nU0; i++) e += Ex.V[(size_t)i * F->nU0 + k] * Ex.V[(size_t)i * F->nU0 + k] * mx[i];
            ex[k] = e;
            lmax = std::max(lmax, std::fabs(lx[k]));
            emx = std::max(emx, std::fabs(e));
        }
        double lmy = 0.0;
        for (int k = 0; k < F->nU1; k++) {
            ly[k] = Ey.lambda[k];
            double e = 0.0;
            for (int j = 0; j < F->nU1; j++) e += Ey.V[(size_t)j * F->nU1 + k] * Ey.V[(size_t)j * F->nU1 + k] * my[j];
            ey[k] = e;
            lmy = std::max(lmy, std::fabs(ly[k]));
            emy = std::max(emy, std::fabs(e));
        }
        // a mode whose D is this small is the constant null mode of a
        // singular (periodic / Neumann Poisson) operator
        F->dtol = 1e-11 * (lmax + lmy + emx * emy);
        rc = F->lx.alloc(sizeof(double) * F->nP0);
        if (!rc) rc = F->ex.alloc(sizeof(double) * F->nP0);
        if (!rc) rc = F->ly.alloc(sizeof(double) * F->nP1);
        if (!rc) rc = F->ey.alloc(sizeof(double) * F->nP1);
        if (rc) return rc;
        SE_HIP(hipMemcpy(F->lx.ptr, lx.data(), sizeof(double) * F->nP0, hipMemcpyHostToDevice));
        SE_HIP(hipMemcpy(F->ex.ptr, ex.data(), sizeof(double) * F->nP0, hipMemcpyHostToDevice));
        SE_HIP(hipMemcpy(F->ly.ptr, ly.data(), sizeof(double) * F->nP1, hipMemcpyHostToDevice));
        SE_HIP(hipMemcpy(F->ey.ptr, ey.data(), sizeof(double) * F->nP1, hipMemcpyHostToDevice));
        const int n = std::max(m->n, 1);
        const int nch = (n + FB_CH - 1) / FB_CH;
        const int ntz = (F->nP1 / 64) * (F->nP0 / 64);
        rc = F->z.alloc(sizeof(double) * n);
        if (!rc) rc = F->pv.alloc(sizeof(double) * nch);
        // p.q and true-residual partials: per chunk, or per apply tile when
        // the applies form them
        const int nap = std::max(nch, line_tiles(m->N, m->nEls[0], m->nEls[1]));
        if (!rc) rc = F->pa.alloc(sizeof(double) * nap);
        if (!rc) rc = F->ps.alloc(sizeof(double) * nch);
        if (!rc) rc = F->pt.alloc(sizeof(double) * 2 * nap);
        if (!rc) rc = F->pz.alloc(sizeof(double) * 3 * ntz);
        if (!rc) rc = F->st.alloc(sizeof(CgState));
        if (rc) return rc;
        op->fdm = F.release();
    }
    FdmOp* F = op->fdm;
    FdmOp::Prec& P = F->prec[prec];
    if (!P.ready) {
        const int dev = m->ctx->device;
        rc = vmats_of(F->ex1, F->nP0, prec, dev, P.mx);
        if (!rc) rc = vmats_of(F->ey1, F->nP1, prec, dev, P.my);
        if (!rc) rc = P.t1.alloc(elem_size(prec) * (size_t)F->nP0 * F->nP1);
        if (!rc) rc = P.t2.alloc(elem_size(prec) * (size_t)F->nP0 * F->nP1);
        if (!rc) rc = P.rp.alloc(elem_size(prec) * (size_t)F->nP0 * F->nP1);
        if (rc) return rc;
        // the padding is never written again: the CG kernels and the pack
        // store the unknown grid's nU1 x nU0 values only
        SE_HIP(hipMemset(P.rp.ptr, 0, elem_size(prec) * (size_t)F->nP0 * F->nP1));
        P.ready = true;
    }
    return SE_OK;
}

// the four passes of op's transforms, z = FDM(r), in precision prec
void fdm_tasks(se_op* op, int prec, const double* r, double* z, double* part, GemmTask (&t)[4]) {
    FdmOp* F = op->fdm;
    FdmOp::Prec& P = F->prec[prec];
    const int nP0 = F->nP0, nP1 = F->nP1;
    for (int k = 0; k < 4; k++) {
        t[k] = GemmTask{};
        t[k].dtol = F->dtol;
    }
    // pass 1: T1t[k'][J] = sum_I Vx[I][k'] R[J][I]   (A = Vx^T, Bt = R packed into rp)
    t[0].A = P.mx->vT.ptr;
    t[0].lda = nP0;
    t[0].Bt = P.rp.ptr;
    t[0].ldb = nP0;
    t[0].r = r;
    t[0].arows = F->nU1;  // the packed residual's true extent
    t[0].acols = F->nU0;
    t[0].C = P.t1.ptr;
    t[0].ldc = nP1;
    t[0].M = nP0;
    t[0].N = nP1;
    t[0].K = nP0;
    t[0].mode = FDM_OUT_PLAIN;
    // pass 2: T2[k'][j'] = sum_J T1t[k'][J] Vy[J][j'] / D(k', j')   (Bt = Vy^T)
    t[1].A = P.t1.ptr;
    t[1].lda = nP1;
    t[1].Bt = P.my->vT.ptr;
    t[1].ldb = nP1;
    t[1].C = P.t2.ptr;
    t[1].ldc = nP1;
    t[1].M = nP0;
    t[1].N = nP1;
    t[1].K = nP1;
    t[1].mode = FDM_OUT_SCALE;
    t[1].lx = F->lx.as<double>();
    t[1].ex = F->ex.as<double>();
    t[1].ly = F->ly.as<double>();
    t[1].ey = F->ey.as<double>();
    t[1].crows = F->nU0;
    t[1].ccols = F->nU1;
    // pass 3: T3t[J][k'] = sum_j' Vy[J][j'] T2[k'][j']   (A = Vy, Bt = T2)
    t[2].A = P.my->v.ptr;
    t[2].lda = nP1;
    t[2].Bt = P.t2.ptr;
    t[2].ldb = nP1;
    t[2].C = P.t1.ptr;
    t[2].ldc = nP0;
    t[2].M = nP1;
    t[2].N = nP0;
    t[2].K = nP1;
    t[2].mode = FDM_OUT_PLAIN;
    // pass 4: Z[J][I] = sum_k' T3t[J][k'] Vx[I][k']   (Bt = Vx), FP64 on the unknown grid
    t[3].A = P.t1.ptr;
    t[3].lda = nP0;
    t[3].Bt = P.mx->v.ptr;
    t[3].ldb = nP0;
    t[3].C = z;
    t[3].ldc = F->nU0;
    t[3].r = r;
    t[3].part = part;
    t[3].M = nP1;
    t[3].N = nP0;
    t[3].K = nP0;
    t[3].mode = FDM_OUT_FINAL;
    t[3].crows = F->nU1;
    t[3].ccols = F->nU0;
    for (int k = 0; k < 4; k++) t[k].tilesN = t[k].N / FDM_BT;
}

// the kernels' shape assumptions, checked on the host before any launch
bool fdm_tasks_valid(const GemmTask (&t)[4], int prec) {
    const int kc = prec == FDM_F64 ? GemmCfg<double>::KC : (prec == FDM_F32 ? GemmCfg<float>::KC : GemmCfg<__bf16>::KC);
    for (int k = 0; k < 4; k++) {
        if (t[k].M <= 0 || t[k].N <= 0 || t[k].K <= 0) return false;
        if (t[k].M % FDM_BT || t[k].N % FDM_BT || t[k].K % 64 || t[k].K % kc) return false;
        if (t[k].lda < t[k].K || t[k].ldb < t[k].K || t[k].lda % 64 || t[k].ldb % 64) return false;
    }
    // the pack fills pass 1's Bt (N x K) with its 64 x 64 blocks
    return t[0].arows <= t[0].N && t[0].acols <= t[0].K && t[3].crows <= t[3].M && t[3].ccols <= t[3].N;
}

int fdm_tiles(const se_op* op) { return (op->fdm->nP0 / FDM_BT) * (op->fdm->nP1 / FDM_BT); }

template <typename T>
void launch_gemm_t(const GemmTask* tasks, int pass, const int2* work, int nwork, hipStream_t s, int skip, bool pack,
                   int to_p) {
    const dim3 grid(nwork), block(256);
    if (pass == 0 && pack) hipLaunchKernelGGL((k_fdm_pack<T>), grid, block, 0, s, tasks, work, nwork, skip);
    hipLaunchKernelGGL((k_fdm_gemm<T>), grid, block, 0, s, tasks, pass, work, nwork, skip, to_p);
}

// pass `pass` of the listed items; pass 1 first packs the residual when
// `pack` (the CG kernels store it packed as they write r, so a solve does not);
// to_p (FDM_TO_*): where the last pass writes
int launch_gemm(int prec, const GemmTask* tasks, int pass, const int2* work, int nwork, hipStream_t s, int skip,
                bool pack, int to_p = FDM_TO_Z) {
    if (nwork <= 0) return SE_OK;
    if (prec == FDM_F64) launch_gemm_t<double>(tasks, pass, work, nwork, s, skip, pack, to_p);
    else if (prec == FDM_F32) launch_gemm_t<float>(tasks, pass, work, nwork, s, skip, pack, to_p);
    else launch_gemm_t<__bf16>(tasks, pass, work, nwork, s, skip, pack, to_p);
    return SE_OK;
}

int fdm_prec_of(int precision) {
    if (precision == SE_PREC_FP64 || precision == SE_PREC_FP64_FP32C || precision == SE_PREC_FP64_BF16C)
        return FDM_F64;
    return precision == SE_PREC_FP32_IR ? FDM_F32 : FDM_BF16;
}
// the transforms' precision for the corrections (the applications after the
// first, whose input is a residual the first application left)
int fdm_cprec_of(int precision) {
    if (precision == SE_PREC_FP64_FP32C) return FDM_F32;
    if (precision == SE_PREC_FP64_BF16C) return FDM_BF16;
    return fdm_prec_of(precision);
}

// ---------------------------------------------------------------------------
// host: the lockstep FDM-preconditioned CG over a set of operators
// ---------------------------------------------------------------------------
struct FdmPlan {
    se_ctx* ctx = nullptr;
    std::vector<se_op*> ops;
    int prec = -1, cprec = -1;  // the transforms' precision: first application, corrections
    DevBuf d_members, d_tasks;  // FbMember[count], GemmTask[4 count]
    DevBuf d_tasks_c;           // GemmTask[4 count] of the corrections (cprec != prec)
    std::vector<FbMember> h_members;
    // FP64 apply descriptors per order (p -> q and x -> q)
    struct Order {
        int N;
        std::vector<int> members;
        DevBuf d_ap, d_ax;  // ApplyMember[]
    };
    std::vector<Order> orders;   // one per order 2..8 (members may be empty)
    // Work lists of one member subset: (member, chunk) items of the vector
    // kernels, (member, tile) items of the transforms, per order the apply's
    // (local member, tile) items, and the member list, packed in one device
    // buffer.  Two independent sets -- the running members' iteration with
    // its true-residual checks, and restarts and the end of a solve -- each keep the
    // lists of the last kListCache subsets they were asked for (successive
    // solves of a batch meet the same subsets: all members, then the same
    // stragglers), so a repeated subset costs neither host work nor an upload.
    // A new subset's lists go up in one copy from pinned memory (a DMA queued
    // on the stream; from pageable memory the runtime stages the copy after
    // waiting for the stream, and the GPU idles around it); the staging is
    // rewritten only after a stream synchronisation, and lists launched
    // together in one round come from different sets.
    struct Lists {
        std::vector<char> on;  // the member flags these lists were built for
        DevBuf d;
        const int2 *vwork = nullptr, *gwork = nullptr, *apply[7] = {};
        const int* list = nullptr;
        int n_vwork = 0, n_gwork = 0, n_apply[7] = {}, n_list = 0;
        unsigned long long used = 0;
    };
    static constexpr int kListCache = 8;
    struct ListSet {
        std::vector<std::unique_ptr<Lists>> cache;
        const Lists* cur = nullptr;  // the lists of the set's current subset
        void* h = nullptr;           // pinned staging of an upload
        size_t hcap = 0;
        unsigned long long clock = 0;
        void clear() {  // (no launch from these lists may be pending)
            cache.clear();
            cur = nullptr;
        }
        ~ListSet() {
            if (h) (void)hipHostFree(h);
        }
    };
    enum { SET_ITER = 0, SET_OTHER = 1 };
    ListSet sets[2];
    CgState* h_states = nullptr;  // pinned, h_cap entries
    int h_cap = 0;
    DevBuf d_states;
    // the device descriptors (members, transforms, applies) were built for
    // these members and this precision, null-space flag and tolerance; a
    // later solve with the same ones only patches f and x
    bool desc_ok = false;
    // a member has an SVV term: it is added after the batched apply, so that
    // member's p.q and true residual come from separate kernels (k_fb_dot,
    // k_fb_true); every other member's come from the applies' epilogues
    // (AP_DOT, AP_TRUE), a member's way the same alone or in any batch
    bool any_svv = false;
    // b is the caller's f itself where b = sgn f - mean is f (sgn = 1, no null
    // space): the start then stores no b, and a new f moves b with it (the
    // first pass-4 tasks and the A x descriptors name b)
    std::vector<char> b_alias;
    std::vector<GemmTask> h_tasks;  // the host copy of d_tasks
    int desc_null = -1;
    double desc_rtol = 0.0;

    ~FdmPlan() {
        if (h_states) (void)hipHostFree(h_states);
    }
};

namespace {

template <int MODE>
int launch_apply_batch_m(int N, const ApplyMember* A, const int2* work, int nw, int grid, hipStream_t s, int gated) {
    switch (N) {
#define SE_CASE(NN)                                                                                        \
    case NN:                                                                                               \
        hipLaunchKernelGGL((k_apply_lines_batch<NN, MODE>), dim3(grid), dim3(256), 0, s, A, work, nw, gated); \
        break;
        SE_CASE(2) SE_CASE(3) SE_CASE(4) SE_CASE(5) SE_CASE(6) SE_CASE(7) SE_CASE(8)
#undef SE_CASE
        default: return fail(SE_ERR_UNSUPPORTED, "FDM solve built for orders 2..8");
    }
    return SE_OK;
}

// (grid <= nw workgroups; a multiple of 8 below nw, so that a workgroup's
// items keep its XCD's place in the XCD-ordered list)
int launch_apply_batch(int N, int mode, const ApplyMember* A, const int2* work, int nw, int grid, hipStream_t s,
                       int gated) {
    if (mode == AP_DOT) return launch_apply_batch_m<AP_DOT>(N, A, work, nw, grid, s, gated);
    if (mode == AP_TRUE) return launch_apply_batch_m<AP_TRUE>(N, A, work, nw, grid, s, gated);
    return launch_apply_batch_m<AP_PLAIN>(N, A, work, nw, grid, s, gated);
}

// Work lists XCD by XCD.  The dispatcher hands workgroup b to XCD b mod 8
// (one workgroup per item in these launches), so the item at position 8 j + x
// runs on XCD x.  A member's tiles go to one XCD when it is small against
// the whole list -- its operand panels or halo lines then stay in that XCD's
// L2 -- and are cut into contiguous runs spread over the XCDs otherwise (a
// single large member keeps all eight busy).  Runs are placed by longest
// processing time on their cost; within an XCD, the runs with the costliest
// tiles come first (a short tail), and runs of the same class (sharing
// eigenvector matrices) and tile cost are adjacent.  Positions a shorter XCD list
// leaves hold (-1, -1) items, which the kernels skip.
constexpr int kXcds = 8;
void xcd_interleave(const std::vector<int>& ids, const std::vector<int>& tiles, const std::vector<double>& cost,
                    const std::vector<const void*>& cls, std::vector<int2>& out) {
    struct Run {
        int id, t0, t1;
        double cost;
        const void* cls;
    };
    long total = 0;
    for (int t : tiles) total += t;
    const int cap = (int)std::max<long>(1, (total + 2 * kXcds - 1) / (2 * kXcds));  // tiles per run at most
    std::vector<Run> runs;
    for (size_t u = 0; u < ids.size(); u++) {
        const int n = tiles[u], parts = (n + cap - 1) / cap;
        for (int k = 0; k < parts; k++) {
            const int t0 = (int)((long)n * k / parts), t1 = (int)((long)n * (k + 1) / parts);
            runs.push_back({ids[u], t0, t1, cost[u] * (t1 - t0) / std::max(n, 1), cls[u]});
        }
    }
    std::stable_sort(runs.begin(), runs.end(), [](const Run& a, const Run& b) { return a.cost > b.cost; });
    std::vector<std::vector<const Run*>> x(kXcds);
    std::vector<double> load(kXcds, 0.0);
    for (const Run& r : runs) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        x[k].push_back(&r);
        load[k] += r.cost;
    }
    std::vector<std::vector<int2>> items(kXcds);
    size_t len = 0;
    for (int k = 0; k < kXcds; k++) {
        // the costliest tiles first (the shortest tail), a class's runs together
        std::stable_sort(x[k].begin(), x[k].end(), [](const Run* a, const Run* b) {
            const double ca = a->cost / (a->t1 - a->t0), cb = b->cost / (b->t1 - b->t0);
            return ca != cb ? ca > cb : a->cls < b->cls;
        });
        for (const Run* r : x[k])
            for (int t = r->t0; t < r->t1; t++) items[k].push_back(make_int2(r->id, t));
        len = std::max(len, items[k].size());
    }
    out.clear();
    for (size_t j = 0; j < len; j++)
        for (int k = 0; k < kXcds; k++) out.push_back(j < items[k].size() ? items[k][j] : make_int2(-1, -1));
}

// q = sign A p of the members of list set `set` (whose lists plan_lists
// built), with the partials of p.q when fused; from x: q = sign A x, or when
// fused the true residual r = b - sign A x with its partials (gated: only
// the members whose CG awaits that check)
int plan_apply(FdmPlan* pl, int set, bool from_x, hipStream_t s, bool gated = false) {
    const FdmPlan::Lists& L = *pl->sets[set].cur;
    const int mode = from_x ? AP_TRUE : AP_DOT;
    for (int k = 0; k < (int)pl->orders.size(); k++) {
        const FdmPlan::Order& o = pl->orders[k];
        const int nw = L.n_apply[k];
        if (o.members.empty() || nw == 0) continue;
        const ApplyMember* A = from_x ? o.d_ax.as<ApplyMember>() : o.d_ap.as<ApplyMember>();
        // a gated check runs for few members, if any: one workgroup per resident
        // slot (four per CU) walks the list, instead of one per tile whose
        // dispatch alone costs ~20 us at the ensemble's 14 k tiles
        const int slots = std::max(8, 4 * pl->ctx->cu_count / 8 * 8);
        const int grid = gated ? std::min(nw, slots) : nw;
        int rc = launch_apply_batch(o.N, mode, A, L.apply[k], nw, grid, s, gated ? 1 : 0);
        if (rc) return rc;
    }
    // members with an SVV term: add it to the batched apply's result
    for (int i = 0; i < (int)pl->ops.size(); i++) {
        se_op* op = pl->ops[i];
        if (!L.on[i] || op->svv_mN == 0) continue;
        int rc = svv_apply(op, from_x ? pl->h_members[i].x : op->p.as<double>(), op->q.as<double>(),
                           (double)op->sign, gated ? pl->d_states.as<CgState>() + i : nullptr);
        if (rc) return rc;
    }
    return SE_OK;
}

// the transforms' list: members' FLOPs as cost, grouped by x eigenvectors
void xcd_order(const FdmPlan* pl, const std::vector<char>& on, std::vector<int2>& out) {
    std::vector<int> ids, tiles;
    std::vector<double> cost;
    std::vector<const void*> cls;
    for (int i = 0; i < (int)pl->ops.size(); i++) {
        if (!on[i]) continue;
        const FdmOp* F = pl->ops[i]->fdm;
        ids.push_back(i);
        tiles.push_back(pl->h_members[i].ntz);
        cost.push_back((double)F->nP0 * F->nP1 * (F->nP0 + F->nP1));
        cls.push_back(F->ex1.get());
    }
    xcd_interleave(ids, tiles, cost, cls, out);
}

// list set `set` switched to the members flagged in `on`: their lists from
// the set's cache, or built and uploaded (one copy) into it
int plan_lists(FdmPlan* pl, int set, const std::vector<char>& on, hipStream_t s) {
    FdmPlan::ListSet& S = pl->sets[set];
    S.clock++;
    for (auto& e : S.cache)
        if (e->on == on) {
            e->used = S.clock;
            S.cur = e.get();
            return SE_OK;
        }
    std::vector<int2> vwork, gwork, apply[7];
    std::vector<int> list;
    for (int i = 0; i < (int)pl->ops.size(); i++) {
        if (!on[i]) continue;
        list.push_back(i);
        for (int c = 0; c < pl->h_members[i].nch; c++) vwork.push_back(make_int2(i, c));
    }
    xcd_order(pl, on, gwork);
    for (int k = 0; k < (int)pl->orders.size(); k++) {
        const FdmPlan::Order& o = pl->orders[k];
        std::vector<int> ids, tiles;
        std::vector<double> cost;
        std::vector<const void*> cls;
        for (int j = 0; j < (int)o.members.size(); j++) {
            const int i = o.members[j];
            if (!on[i]) continue;
            const se_mesh* m = pl->ops[i]->mesh;
            ids.push_back(j);
            tiles.push_back(line_tiles(m->N, m->nEls[0], m->nEls[1]));
            cost.push_back((double)tiles.back());
            cls.push_back(nullptr);
        }
        xcd_interleave(ids, tiles, cost, cls, apply[k]);
    }
    // packed: vector items, transform items, per-order apply items (int2), member list (int)
    size_t n2 = vwork.size() + gwork.size();
    for (const auto& a : apply) n2 += a.size();
    const size_t bytes = sizeof(int2) * n2 + sizeof(int) * list.size();
    if (bytes > S.hcap) {
        if (S.h) {
            SE_HIP(hipStreamSynchronize(s));  // no upload from the old staging is pending
            (void)hipHostFree(S.h);
            S.h = nullptr;
            S.hcap = 0;
        }
        SE_HIP(hipHostMalloc(&S.h, bytes, hipHostMallocDefault));
        S.hcap = bytes;
    }
    if ((int)S.cache.size() >= FdmPlan::kListCache) {  // the least recently used subset makes room
        SE_HIP(hipStreamSynchronize(s));                // no launch reads its lists
        auto lru = std::min_element(S.cache.begin(), S.cache.end(),
                                    [](const std::unique_ptr<FdmPlan::Lists>& a,
                                       const std::unique_ptr<FdmPlan::Lists>& b) { return a->used < b->used; });
        S.cache.erase(lru);
    }
    std::unique_ptr<FdmPlan::Lists> L(new FdmPlan::Lists);
    int rc = L->d.alloc(std::max<size_t>(bytes, sizeof(int2)));
    if (rc) return rc;
    char* h = static_cast<char*>(S.h);
    const char* d = L->d.as<char>();
    size_t off = 0;
    auto put = [&](const void* src, size_t nb) -> const char* {
        if (nb) std::memcpy(h + off, src, nb);
        const char* at = d + off;
        off += nb;
        return at;
    };
    L->vwork = reinterpret_cast<const int2*>(put(vwork.data(), sizeof(int2) * vwork.size()));
    L->gwork = reinterpret_cast<const int2*>(put(gwork.data(), sizeof(int2) * gwork.size()));
    for (int k = 0; k < 7; k++) {
        L->apply[k] = reinterpret_cast<const int2*>(put(apply[k].data(), sizeof(int2) * apply[k].size()));
        L->n_apply[k] = (int)apply[k].size();
    }
    L->list = reinterpret_cast<const int*>(put(list.data(), sizeof(int) * list.size()));
    L->n_vwork = (int)vwork.size();
    L->n_gwork = (int)gwork.size();
    L->n_list = (int)list.size();
    L->on = on;
    L->used = S.clock;
    if (bytes) SE_HIP(hipMemcpyAsync(L->d.ptr, S.h, bytes, hipMemcpyHostToDevice, s));
    S.cur = L.get();
    S.cache.push_back(std::move(L));
    return SE_OK;
}

int grid_of(size_t items, const se_ctx* c) {
    return (int)std::max<size_t>(1, std::min<size_t>(items, (size_t)16 * c->cu_count));
}

// one pass of the transforms of the listed members
int plan_pass(FdmPlan* pl, int set, int pass, hipStream_t s, int skip = 0, int to_p = FDM_TO_Z, bool corr = false) {
    const FdmPlan::Lists& L = *pl->sets[set].cur;
    const bool c = corr && pl->cprec != pl->prec;
    return launch_gemm(c ? pl->cprec : pl->prec, (c ? pl->d_tasks_c : pl->d_tasks).as<GemmTask>(), pass, L.gwork,
                       L.n_gwork, s, skip, false, pass == 3 ? to_p : FDM_TO_Z);
}

// z = FDM(r) of the listed members: four passes

// (to_p: into the CG's p, for a first or restarted direction; corr: in the
// corrections' precision, for the residual of an update)
int plan_fdm(FdmPlan* pl, int set, hipStream_t s, int skip = 0, int to_p = FDM_TO_Z, bool corr = false) {
    for (int pass = 0; pass < 4; pass++) {
        int rc = plan_pass(pl, set, pass, s, skip, to_p, corr);
        if (rc) return rc;
    }
    return SE_OK;
}

}  // namespace

// A plan's members can be replaced (the request server re-members one plan
// for every inversion pass): buffers only grow, so a long-lived plan stops
// allocating once it has seen its largest set.
int fdm_plan_set_members(FdmPlan* pl, int count, se_op* const* ops) {
    SE_HIP(hipStreamSynchronize(pl->ctx->stream));  // nothing of the previous set is in flight
    pl->ops.assign(ops, ops + count);
    if (pl->orders.empty()) {
        pl->orders.resize(7);
        for (int k = 0; k < 7; k++) pl->orders[k].N = k + 2;
    }
    for (auto& o : pl->orders) o.members.clear();
    for (int i = 0; i < count; i++) {
        const int N = ops[i]->mesh->N;
        SE_CHECK(N >= 2 && N <= 8, SE_ERR_UNSUPPORTED, "FDM solve built for orders 2..8");
        pl->orders[N - 2].members.push_back(i);
    }
    if (count > pl->h_cap) {
        if (pl->h_states) (void)hipHostFree(pl->h_states);
        pl->h_states = nullptr;
        pl->h_cap = 0;
        SE_HIP(hipHostMalloc((void**)&pl->h_states, sizeof(CgState) * count, hipHostMallocDefault));
        pl->h_cap = count;
    }
    int rc = pl->d_members.alloc(sizeof(FbMember) * std::max(count, 1));
    if (!rc) rc = pl->d_tasks.alloc(sizeof(GemmTask) * 4 * std::max(count, 1));
    if (rc) return rc;
    pl->prec = -1;
    pl->cprec = -1;
    pl->desc_ok = false;
    for (auto& L : pl->sets) L.clear();
    return SE_OK;
}

int fdm_plan_create(se_ctx* c, int count, se_op* const* ops, FdmPlan** out) {
    std::unique_ptr<FdmPlan> pl(new FdmPlan);
    pl->ctx = c;
    int rc = fdm_plan_set_members(pl.get(), count, ops);
    if (rc) return rc;
    *out = pl.release();
    return SE_OK;
}

void fdm_plan_destroy(FdmPlan* pl) { delete pl; }

// the FP64 apply descriptors of every order group: A p (+ p.q partials) and
// A x (or, fused, the true residual), with each member's current iterate
int upload_apply_descriptors(FdmPlan* pl, hipStream_t s) {
    for (auto& od : pl->orders) {
        if (od.members.empty()) continue;
        std::vector<ApplyMember> ap, ax;
        for (int i : od.members) {
            se_op* op = pl->ops[i];
            const FbMember& M = pl->h_members[i];
            ApplyMember A;
            A.P = op_params(op, (double)op->sign);
            A.in = op->p.as<double>();
            A.out = op->q.as<double>();
            if (M.fused) {  // + partials of p.q
                A.part = M.pa;
                A.np = M.nap;
            }
            ap.push_back(A);
            A.in = M.x;
            A.gate = M.st;
            if (M.fused) {  // r = b - A x, partials of sum t, sum t^2
                A.out = op->r.as<double>();
                A.b = M.b;
                A.part = M.pt;
            }
            ax.push_back(A);
        }
        int rc = od.d_ap.alloc(sizeof(ApplyMember) * ap.size());
        if (!rc) rc = od.d_ax.alloc(sizeof(ApplyMember) * ax.size());
        if (rc) return rc;
        SE_HIP(hipMemcpyAsync(od.d_ap.ptr, ap.data(), sizeof(ApplyMember) * ap.size(), hipMemcpyHostToDevice, s));
        SE_HIP(hipMemcpyAsync(od.d_ax.ptr, ax.data(), sizeof(ApplyMember) * ax.size(), hipMemcpyHostToDevice, s));
        SE_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
    }
    return SE_OK;
}

// The iterate of member i: the caller's solution vector itself unless it is
// also the right-hand side (then the operator's own x, copied out at the end)
double* iterate_of(FdmPlan* pl, int i, const double* const* d_f, double* const* d_x) {
    return d_x[i] != d_f[i] ? d_x[i] : pl->ops[i]->x.as<double>();
}

int fdm_plan_solve(FdmPlan* pl, const double* const* d_f, double* const* d_x, const se_solve_opts& o,
                   se_solve_info* infos) {
    se_ctx* c = pl->ctx;
    hipStream_t s = c->stream;
    const int count = (int)pl->ops.size();
    const int prec = fdm_prec_of(o.precision), cprec = fdm_cprec_of(o.precision);
    const int null = o.null_space ? 1 : 0;
    for (int i = 0; i < count; i++) {
        int rc = fdm_prepare(pl->ops[i], prec);
        if (!rc && cprec != prec) rc = fdm_prepare(pl->ops[i], cprec);
        if (rc) return rc;
    }
    // descriptors: built once per members / precision / null flag / tolerance;
    // a later solve patches only the right-hand sides and the results
    int rc = SE_OK;
    const bool reuse = pl->desc_ok && pl->prec == prec && pl->cprec == cprec && pl->desc_null == null &&
                       pl->desc_rtol == o.rtol;
    if (reuse) {
        bool changed = false, moved = false, bmoved = false;
        for (int i = 0; i < count; i++) {
            FbMember& M = pl->h_members[i];
            if (M.f != d_f[i] || M.xo != d_x[i]) {
                if (pl->b_alias[i] && M.f != d_f[i]) {
                    M.b = const_cast<double*>(d_f[i]);  // read only: b is f
                    pl->h_tasks[4 * (size_t)i + 3].b = M.b;
                    bmoved = true;
                }
                M.f = d_f[i];
                M.xo = d_x[i];
                double* x = iterate_of(pl, i, d_f, d_x);
                moved = moved || x != M.x;
                M.x = x;
                changed = true;
            }
        }
        if (changed && count > 0)
            SE_HIP(hipMemcpyAsync(pl->d_members.ptr, pl->h_members.data(), sizeof(FbMember) * count,
                                  hipMemcpyHostToDevice, s));
        if (bmoved)
            SE_HIP(hipMemcpyAsync(pl->d_tasks.ptr, pl->h_tasks.data(), sizeof(GemmTask) * 4 * count,
                                  hipMemcpyHostToDevice, s));
        if (moved || bmoved) {  // the A x descriptors name the iterates and b
            int e = upload_apply_descriptors(pl, s);
            if (e) return e;
        }
    } else {
        {
            int rc = pl->d_states.alloc(sizeof(CgState) * std::max(count, 1));
            if (rc) return rc;
        }
        pl->h_members.assign(count, FbMember{});
        std::vector<GemmTask> tasks(4 * (size_t)count), tasks_c(cprec != prec ? 4 * (size_t)count : 0);
        if (cprec != prec) {
            int e = pl->d_tasks_c.alloc(sizeof(GemmTask) * 4 * std::max(count, 1));
            if (e) return e;
        }
        pl->any_svv = false;
        for (se_op* op : pl->ops) pl->any_svv = pl->any_svv || op->svv_mN != 0;
        pl->b_alias.assign(count, 0);
        for (int i = 0; i < count; i++) {
            se_op* op = pl->ops[i];
            FdmOp* F = op->fdm;
            FbMember& M = pl->h_members[i];
            M.f = d_f[i];
            pl->b_alias[i] = op->sign == 1 && !null;
            M.b = pl->b_alias[i] ? const_cast<double*>(d_f[i]) : op->b.as<double>();  // (b is read only when f)
            M.x = iterate_of(pl, i, d_f, d_x);
            M.r = op->r.as<double>();
            M.p = op->p.as<double>();
            M.q = op->q.as<double>();
            M.z = F->z.as<double>();
            M.pv = F->pv.as<double>();
            M.pz = F->pz.as<double>();
            M.pa = F->pa.as<double>();
            M.pt = F->pt.as<double>();
            M.ps = F->ps.as<double>();
            M.xo = d_x[i];
            M.rp = F->prec[prec].rp.ptr;
            M.rpc = F->prec[cprec].rp.ptr;
            M.nU0 = F->nU0;
            M.nP0 = F->nP0;
            M.rprec = prec;
            M.rcprec = cprec;
            M.n = op->mesh->n;
            SE_CHECK(M.n == F->nU0 * F->nU1, SE_ERR_STATE, "FDM: the unknowns are not the nU1 x nU0 grid");
            M.nch = (std::max(M.n, 1) + FB_CH - 1) / FB_CH;
            M.fused = op->svv_mN == 0;
            M.nap = M.fused ? line_tiles(op->mesh->N, op->mesh->nEls[0], op->mesh->nEls[1]) : M.nch;
            M.ntz = fdm_tiles(op);
            M.null_space = null;
            M.sgn = (double)op->sign;
            M.rtol = o.rtol;
            M.st = pl->d_states.as<CgState>() + i;
            GemmTask t4[4];
            fdm_tasks(op, prec, M.r, M.z, M.pz, t4);
            SE_CHECK(fdm_tasks_valid(t4, prec), SE_ERR_STATE, "FDM: transform shapes break the kernels' assumptions");
            t4[3].C2 = M.p;
            t4[3].b = M.b;
            for (int k = 0; k < 4; k++) {
                t4[k].pv = M.pv;
                t4[k].st = M.st;
                t4[k].nch = M.nch;
                tasks[4 * (size_t)i + k] = t4[k];
            }
            if (cprec != prec) {
                GemmTask tc[4];
                fdm_tasks(op, cprec, M.r, M.z, M.pz, tc);
                SE_CHECK(fdm_tasks_valid(tc, cprec), SE_ERR_STATE,
                         "FDM: transform shapes break the kernels' assumptions");
                for (int k = 0; k < 4; k++) {
                    tc[k].pv = M.pv;
                    tc[k].st = M.st;
                    tc[k].nch = M.nch;
                    tasks_c[4 * (size_t)i + k] = tc[k];
                }
            }
        }
        int rc = SE_OK;
        if (count > 0) {
            SE_HIP(hipMemcpyAsync(pl->d_members.ptr, pl->h_members.data(), sizeof(FbMember) * count, hipMemcpyHostToDevice,
                                  s));
            SE_HIP(hipMemcpyAsync(pl->d_tasks.ptr, tasks.data(), sizeof(GemmTask) * 4 * count, hipMemcpyHostToDevice, s));
            pl->h_tasks = tasks;
            if (cprec != prec)
                SE_HIP(hipMemcpyAsync(pl->d_tasks_c.ptr, tasks_c.data(), sizeof(GemmTask) * 4 * count,
                                      hipMemcpyHostToDevice, s));
        }
        pl->prec = prec;
        pl->cprec = cprec;
        // FP64 apply descriptors
        rc = upload_apply_descriptors(pl, s);
        if (rc) return rc;
        for (auto& L : pl->sets) L.clear();  // new descriptors: new lists
        pl->desc_ok = true;
        pl->desc_null = null;
        pl->desc_rtol = o.rtol;
    }
    for (int i = 0; i < count; i++) {
        CgState h{};
        h.maxit = std::max(o.maxit, 0);
        h.null_space = null;
        h.n = pl->ops[i]->mesh->n;
        h.status = CG_RUNNING;
        pl->h_states[i] = h;
    }
    const FbMember* Ms = pl->d_members.as<FbMember>();
    auto pull = [&]() -> int {
        SE_HIP(hipMemcpyAsync(pl->h_states, pl->d_states.ptr, sizeof(CgState) * count, hipMemcpyDeviceToHost, s));
        SE_HIP(hipStreamSynchronize(s));
        return SE_OK;
    };
    auto vlaunch = [&](int set, void (*k)(const FbMember*, const int2*, int)) {
        const FdmPlan::Lists& L = *pl->sets[set].cur;
        const int nw = L.n_vwork;
        if (nw > 0) hipLaunchKernelGGL(k, dim3(grid_of(nw, c)), dim3(FB_NT), 0, s, Ms, L.vwork, nw);
    };
    auto dlaunch = [&](int set, int mode) {
        const FdmPlan::Lists& L = *pl->sets[set].cur;
        const int nw = L.n_vwork;
        if (nw > 0)
            hipLaunchKernelGGL(k_fb_dir, dim3(grid_of(nw, c)), dim3(FB_NT), 0, s, Ms, L.vwork, nw, mode);
    };
    // the true residual of the flagged members: t = b - A x into r, ||t - mean t||^2 into the state
    // (gated: of the flagged members only those whose CG has just marked
    // itself as awaiting the check, decided on the device)
    auto true_residual = [&](int set, const std::vector<char>& flags, bool gated) -> int {
        int e = plan_lists(pl, set, flags, s);
        if (!e) e = plan_apply(pl, set, true, s, gated);
        if (e) return e;
        const FdmPlan::Lists& L = *pl->sets[set].cur;
        if (pl->any_svv && L.n_vwork > 0)
            hipLaunchKernelGGL(k_fb_true, dim3(grid_of(L.n_vwork, c)), dim3(FB_NT), 0, s, Ms, L.vwork, L.n_vwork,
                               gated ? 1 : 0);
        if (L.n_list > 0)
            hipLaunchKernelGGL(k_fb_true_fin, dim3(L.n_list), dim3(FB_NT), 0, s, Ms, L.list, L.n_list, gated ? 1 : 0);
        return SE_OK;
    };
    const int IT = FdmPlan::SET_ITER, OT = FdmPlan::SET_OTHER;

    SE_HIP(hipEventRecord(c->ev0, s));
    SE_HIP(hipMemcpyAsync(pl->d_states.ptr, pl->h_states, sizeof(CgState) * count, hipMemcpyHostToDevice, s));
    // start: b, r = b, x = 0, z = FDM(r), p = z.  No host round trip: a member
    // stopped here (zero or non-finite right-hand side) is skipped by every
    // kernel of the first iteration and classified after it.
    std::vector<char> on(count, 1);
    rc = plan_lists(pl, IT, on, s);
    if (rc) return rc;
    if (null) vlaunch(IT, k_fb_sum);
    vlaunch(IT, k_fb_start);
    rc = plan_fdm(pl, IT, s, 0, FDM_TO_P_FROM_B);
    if (rc) return rc;
    dlaunch(IT, FB_FIRST);
    std::vector<int> status(count, CG_RUNNING);
    // progress of each member's recursive residual: a member whose r.r has
    // not fallen 4x within FB_STALL iterations cannot be resolved by the
    // preconditioner's precision (a near-singular system with BF16
    // transforms) and stops with SE_SOLVE_STAGNATED
    constexpr int FB_STALL = 40;
    std::vector<double> best(count, HUGE_VAL);
    std::vector<int> best_it(count, 0);
    int guard = 0;
    // Rounds: the running members' next iteration, then the true-residual
    // check of those whose recursive residual has just met the target (gated
    // on the device, so it needs no round trip of its own), then one pull.
    for (;;) {
        bool any_on = false;
        for (int i = 0; i < count; i++) any_on = any_on || on[i];
        if (!any_on) break;
        // q = A p, alpha, x and r, z = FDM(r), beta and p
        rc = plan_lists(pl, IT, on, s);
        if (!rc) rc = plan_apply(pl, IT, false, s);
        if (rc) return rc;
        if (pl->any_svv) vlaunch(IT, k_fb_dot);
        vlaunch(IT, k_fb_update);
        rc = plan_fdm(pl, IT, s, 1, FDM_TO_Z, true);
        if (rc) return rc;
        dlaunch(IT, FB_NORMAL);
        rc = true_residual(IT, on, true);
        if (rc) return rc;
        SE_HIP(hipGetLastError());
        rc = pull();
        if (rc) return rc;
        // running members: continuing, stagnated, stopped, or checked in this
        // round (converged, out of iterations, or restarted from the true residual)
        std::vector<char> again(count, 0);
        bool restart = false;
        for (int i = 0; i < count; i++) {
            if (!on[i]) continue;
            CgState& h = pl->h_states[i];
            if (!h.stop) {
                if (h.rr < 0.25 * best[i]) {
                    best[i] = h.rr;
                    best_it[i] = h.iter;
                } else if (h.iter - best_it[i] >= FB_STALL) {
                    status[i] = SE_SOLVE_STAGNATED;
                    on[i] = 0;
                }
                continue;
            }
            on[i] = 0;
            if (h.status != CG_PENDING) {
                status[i] = h.status;
            } else if (h.tt <= h.tol2) {
                status[i] = 0;
            } else if (h.iter >= h.maxit || ++guard > 100 * count + 1000) {
                status[i] = 1;
            } else {
                again[i] = 1;
                restart = true;
                h.stop = 0;
                h.status = CG_RUNNING;
            }
        }
        if (!restart) continue;
        // continue from the true residual (now in r) with a fresh direction
        SE_HIP(hipMemcpyAsync(pl->d_states.ptr, pl->h_states, sizeof(CgState) * count, hipMemcpyHostToDevice, s));
        rc = plan_lists(pl, OT, again, s);
        if (rc) return rc;
        vlaunch(OT, k_fb_restart);
        rc = plan_fdm(pl, OT, s, 0, FDM_TO_P);
        if (rc) return rc;
        dlaunch(OT, FB_RESTART);
        SE_HIP(hipGetLastError());
        for (int i = 0; i < count; i++)
            if (again[i]) on[i] = 1;
    }
    // final true residuals of the members that did not converge
    std::vector<char> fin(count, 0);
    bool anyfin = false;
    for (int i = 0; i < count; i++)
        if (status[i] != 0 && pl->h_states[i].bb > 0.0 && pl->h_states[i].bb - pl->h_states[i].bb == 0.0) {
            fin[i] = 1;
            anyfin = true;
        }
    if (anyfin) {
        rc = true_residual(OT, fin, false);
        if (!rc) rc = pull();
        if (rc) return rc;
    }
    {
        std::vector<char> out(count, 0);
        bool any_out = false;
        for (int i = 0; i < count; i++) {
            out[i] = d_x[i] != pl->h_members[i].x;  // the iterate is not the caller's vector
            any_out = any_out || out[i];
        }
        if (any_out) {
            rc = plan_lists(pl, anyfin ? IT : OT, out, s);  // (a set whose uploads are complete)
            if (rc) return rc;
            vlaunch(anyfin ? IT : OT, k_fb_out);
            SE_HIP(hipGetLastError());
        }
    }
    SE_HIP(hipEventRecord(c->ev1, s));
    SE_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    for (int i = 0; i < count; i++) {
        if (!infos) break;
        const CgState& h = pl->h_states[i];
        se_solve_info& inf = infos[i];
        inf.status = status[i] == CG_RUNNING ? 1 : status[i];
        inf.iters = h.iter;
        inf.outer_iters = 1;
        const double tt = (status[i] == 0 || fin[i]) ? h.tt : h.rr;
        inf.true_rel_residual = h.bb > 0.0 ? std::sqrt(std::max(tt, 0.0) / h.bb) : 0.0;
        inf.seconds = ms * 1e-3;
    }
    return SE_OK;
}

// Measurement: after a solve of the plan, time `reps` back-to-back launches
// of each transform pass over all members (pass 1 with its residual pack)
// and of the FP64 operator apply, with HIP events on the context stream.
// Per pass: microseconds per launch and flops per launch (2 M N K summed over
// the members' padded sizes); the apply: microseconds and algorithmic bytes
// (16 B per unknown for x and y).
int fdm_plan_probe(FdmPlan* pl, int reps, double* us_pass, double* flops_pass, double* us_apply, double* bytes_apply) {
    SE_CHECK(pl->prec >= 0, SE_ERR_STATE, "FDM probe: run an FDM solve first");
    se_ctx* c = pl->ctx;
    hipStream_t s = c->stream;
    const int count = (int)pl->ops.size();
    std::vector<char> all(count, 1);
    int rc = plan_lists(pl, FdmPlan::SET_OTHER, all, s);
    if (rc) return rc;
    for (int pass = 0; pass < 4; pass++) {
        double fl = 0.0;
        for (se_op* op : pl->ops) {
            const double nP0 = op->fdm->nP0, nP1 = op->fdm->nP1;
            fl += 2.0 * nP0 * nP1 * ((pass == 0 || pass == 3) ? nP0 : nP1);
        }
        flops_pass[pass] = fl;
        rc = plan_pass(pl, FdmPlan::SET_OTHER, pass, s);
        if (rc) return rc;
        SE_HIP(hipEventRecord(c->ev0, s));
        for (int r = 0; r < reps && !rc; r++) rc = plan_pass(pl, FdmPlan::SET_OTHER, pass, s);
        SE_HIP(hipEventRecord(c->ev1, s));
        SE_HIP(hipEventSynchronize(c->ev1));
        if (rc) return rc;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        us_pass[pass] = 1e3 * ms / reps;
    }
    double by = 0.0;
    for (se_op* op : pl->ops) by += 16.0 * op->mesh->n;
    *bytes_apply = by;
    rc = plan_apply(pl, FdmPlan::SET_OTHER, false, s);
    if (rc) return rc;
    SE_HIP(hipEventRecord(c->ev0, s));
    for (int r = 0; r < reps && !rc; r++) rc = plan_apply(pl, FdmPlan::SET_OTHER, false, s);
    SE_HIP(hipEventRecord(c->ev1, s));
    SE_HIP(hipEventSynchronize(c->ev1));
    if (rc) return rc;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    *us_apply = 1e3 * ms / reps;
    SE_HIP(hipGetLastError());
    return SE_OK;
}

// solve of one operator: a plan of one member, kept with the operator
int pcg_fdm(se_op* op, const double* d_f, double* d_x, const se_solve_opts& o, se_solve_info& info) {
    SE_CHECK(op->definite, SE_ERR_UNSUPPORTED, "operator diagonal changes sign: not a definite system");
    int rc = fdm_prepare(op, fdm_prec_of(o.precision));
    if (!rc && fdm_cprec_of(o.precision) != fdm_prec_of(o.precision)) rc = fdm_prepare(op, fdm_cprec_of(o.precision));
    if (rc) return rc;
    if (!op->fdm->self) {
        rc = fdm_plan_create(op->mesh->ctx, 1, &op, &op->fdm->self);
        if (rc) return rc;
    }
    return fdm_plan_solve(op->fdm->self, &d_f, &d_x, o, &info);
}

// resolved preconditioner of a solve
int solve_pc(const se_solve_opts& o) {
    if (o.preconditioner == SE_PC_JACOBI || o.preconditioner == SE_PC_FDM) return o.preconditioner;
    return o.precision == SE_PREC_FP64 ? SE_PC_JACOBI : SE_PC_FDM;
}

}  // namespace se

// z = FDM(r) of sign*A (host arrays of n values), in `precision`
extern "C" int se_helm_apply_fdm(se_op* op, int precision, const double* r, double* z) {
    using namespace se;
    SE_CHECK(op && r && z, SE_ERR_ARG, "se_helm_apply_fdm: NULL argument");
    SE_CHECK(precision == SE_PREC_FP64 || precision == SE_PREC_FP32_IR || precision == SE_PREC_BF16_IR, SE_ERR_ARG,
             "se_helm_apply_fdm: bad precision");
    SE_HIP(hipSetDevice(op->mesh->ctx->device));
    const int prec = fdm_prec_of(precision);
    int rc = fdm_prepare(op, prec);
    if (rc) return rc;
    const int n = op->mesh->n;
    hipStream_t s = op->mesh->ctx->stream;
    SE_HIP(hipStreamSynchronize(s));
    if (n > 0) SE_HIP(hipMemcpy(op->r.ptr, r, sizeof(double) * n, hipMemcpyHostToDevice));
    GemmTask t[4];
    fdm_tasks(op, prec, op->r.as<double>(), op->fdm->z.as<double>(), op->fdm->pz.as<double>(), t);
    SE_CHECK(fdm_tasks_valid(t, prec), SE_ERR_STATE, "se_helm_apply_fdm: transform shapes break the kernels' assumptions");
    DevBuf d;
    rc = d.alloc(sizeof(t));
    if (rc) return rc;
    SE_HIP(hipMemcpy(d.ptr, t, sizeof(t), hipMemcpyHostToDevice));
    for (int pass = 0; pass < 4; pass++) {
        rc = launch_gemm(prec, d.as<GemmTask>(), pass, nullptr, fdm_tiles(op), s, 0, true);
        if (rc) return rc;
    }
    SE_HIP(hipGetLastError());
    SE_HIP(hipStreamSynchronize(s));
    if (n > 0) SE_HIP(hipMemcpy(z, op->fdm->z.ptr, sizeof(double) * n, hipMemcpyDeviceToHost));
    return SE_OK;
}

// measurement hook of a single operator's FDM solve (after one)
extern "C" int se_helm_fdm_probe(se_op* op, int reps, double* us_pass, double* flops_pass, double* us_apply,
                                 double* bytes_apply) {
    using namespace se;
    SE_CHECK(op && us_pass && flops_pass && us_apply && bytes_apply && reps > 0, SE_ERR_ARG,
             "se_helm_fdm_probe: bad argument");
    SE_CHECK(op->fdm && op->fdm->self, SE_ERR_STATE, "se_helm_fdm_probe: run an FDM solve of this operator first");
    SE_HIP(hipSetDevice(op->mesh->ctx->device));
    return fdm_plan_probe(op->fdm->self, reps, us_pass, flops_pass, us_apply, bytes_apply);
}
