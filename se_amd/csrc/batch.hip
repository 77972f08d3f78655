// batch.hip -- many independent mixed-precision Helmholtz solves at once.
//
// Members (operators of any size, order and boundary conditions, created on
// one context) are grouped by GLL order.  One CG iteration of a group is two
// launches over the (member, tile) work items of all its members
// (k_tcg_update_batch, k_tcg_dir_apply_batch); the k-iteration chains of the
// groups are captured into one graph on forked streams, so the groups run
// concurrently.  Each member keeps its own recurrence state, convergence test
// and FP64 reliable updates: at every check the host reads all states, and
// the members that are due get their update in batched launches (flush,
// FP64 operator apply, residual, norm) over just their work items.  A
// finished member stays stopped: its work items exit at the flag.
//
// Launch shape: the work-item kernels loop over their items, so the grid
// size is free; an autotuner times the candidate grid sizes on the first
// checks of the first solve (real iterations, nothing discarded) and keeps
// the fastest for the batch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "apply_lines.h"
#include "se_common.h"
#include "se_op.h"
#include "tcg_host.h"
#include "tcg_kernels.h"

using namespace se;

namespace {

int vb_of(int n) { return std::max(1, std::min((n + 4095) / 4096, 256)); }

__global__ __launch_bounds__(1024) void k_batch_sum(int n, const double* __restrict__ f, double sgn,
                                                    double* __restrict__ part) {
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += sgn * f[i];
    double v[1] = {s};
    block_store_partials<1, 1024>(v, part);
}

struct Group {
    int N = 0;
    std::vector<int> members;   // batch member indices
    std::vector<int2> tiles;    // all (local member, tile) items
    DevBuf d_members, d_tiles;  // TcgMember[], int2[]
    // the iteration chain's work list: the items of the members still
    // running, compacted by the host between replays (count on the device)
    DevBuf d_active, d_nactive;
    std::vector<int2> h_active;
    int* h_nactive = nullptr;  // pinned
    DevBuf d_amembers;          // ApplyMember[] (FP64 operator of each member)
    // lists of the members due for a reliable update (rebuilt per check);
    // the host copies live here until the stream has consumed them
    DevBuf d_ut, d_uf, d_ua, d_ul, d_rt;
    std::vector<int2> h_ut, h_uf, h_ua, h_rt;
    std::vector<int> h_ul;
    std::vector<TcgMember> h_members;
    std::vector<ApplyMember> h_amembers;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int grid_dir = 0, grid_upd = 0;  // work-item kernel grids of the iteration chain
};

}  // namespace

struct se_batch {
    se_ctx* ctx = nullptr;
    std::vector<se_op*> ops;
    // false when a member has an SVV term: the inner Jacobi-PCG's separable
    // operator cannot carry it, so such a batch solves with FDM only
    bool jacobi_ok = true;
    std::vector<Group> groups;
    DevBuf d_states;
    CgState* h_states = nullptr;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_key = -1;
    // grid factors (workgroups per CU, 0: one per item) chosen by the
    // autotuner for the direction and the update kernel, plus one (0: not yet)
    int tuned_dir = 0, tuned_upd = 0;
    double tune_best = 0.0;
    // FDM-preconditioned solves of the members (fdm.hip), built on first use
    FdmPlan* fdm = nullptr;
    // pinned staging of the host -> device list uploads: filled between two
    // synchronisations of the stream, so no byte is reused while a copy of
    // it can still be pending (pageable copies would stage and block)
    char* h_arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;

    void release() {
        if (fdm) fdm_plan_destroy(fdm);
        fdm = nullptr;
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        if (h_arena) (void)hipHostFree(h_arena);
        h_arena = nullptr;
        arena_cap = arena_used = 0;
        for (auto& g : groups) {
            if (g.h_nactive) (void)hipHostFree(g.h_nactive);
            g.h_nactive = nullptr;
            if (g.stream) (void)hipStreamDestroy(g.stream);
            if (g.done) (void)hipEventDestroy(g.done);
            g.stream = nullptr;
            g.done = nullptr;
        }
        if (fork) (void)hipEventDestroy(fork);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        fork = t0 = t1 = nullptr;
        if (h_states) (void)hipHostFree(h_states);
        h_states = nullptr;
    }
};

namespace {

// the device copy of a host list, staged through the batch's pinned arena
template <class V>
int upload(se_batch* b, DevBuf& d, const std::vector<V>& v, hipStream_t s) {
    int rc = d.alloc(sizeof(V) * std::max<size_t>(v.size(), 1));
    if (rc || v.empty()) return rc;
    const size_t nb = sizeof(V) * v.size();
    if (b->arena_used + nb > b->arena_cap) {
        // the staged copies complete before the arena is reused or replaced
        SE_HIP(hipStreamSynchronize(s));
        b->arena_used = 0;
        if (nb > b->arena_cap) {
            if (b->h_arena) (void)hipHostFree(b->h_arena);
            b->h_arena = nullptr;
            b->arena_cap = 0;
            const size_t cap = std::max<size_t>(nb, (size_t)4 << 20);
            SE_HIP(hipHostMalloc((void**)&b->h_arena, cap, hipHostMallocDefault));
            b->arena_cap = cap;
        }
    }
    char* src = b->h_arena + b->arena_used;
    std::memcpy(src, v.data(), nb);
    b->arena_used += (nb + 255) & ~(size_t)255;
    SE_HIP(hipMemcpyAsync(d.ptr, src, nb, hipMemcpyHostToDevice, s));
    return SE_OK;
}

#define SE_BATCH_ORDERS(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8)

// ---- launches of one group's work-item kernels ----
template <typename T>
int g_update(const Group& g, int nwork, const int2* work, int grid, hipStream_t s, const int* ndev = nullptr) {
    if (nwork == 0) return SE_OK;
    const TcgMember* Ms = g.d_members.as<TcgMember>();
    switch (g.N) {
#define SE_CASE(NN)                                                                                              \
    case NN:                                                                                                     \
        hipLaunchKernelGGL((k_tcg_update_batch<NN, T>), dim3(grid), dim3(TCG_NT), 0, s, Ms, work, nwork, ndev); \
        break;
        SE_BATCH_ORDERS(SE_CASE)
#undef SE_CASE
        default: return fail(SE_ERR_UNSUPPORTED, "batched solve built for orders 2..8");
    }
    return SE_OK;
}
template <typename T, int MODE>
int g_dir_apply(const Group& g, int nwork, const int2* work, int grid, hipStream_t s, const int* ndev = nullptr) {
    if (nwork == 0) return SE_OK;
    const TcgMember* Ms = g.d_members.as<TcgMember>();
    switch (g.N) {
#define SE_CASE(NN)                                                                                                   \
    case NN:                                                                                                          \
        hipLaunchKernelGGL((k_tcg_dir_apply_batch<NN, T, MODE>), dim3(grid), dim3(TCG_NT), 0, s, Ms, work, nwork,     \
                           ndev);                                                                                     \
        break;
        SE_BATCH_ORDERS(SE_CASE)
#undef SE_CASE
        default: return fail(SE_ERR_UNSUPPORTED, "batched solve built for orders 2..8");
    }
    return SE_OK;
}
template <typename T, bool RESID>
int g_start(const Group& g, int nwork, const int2* work, hipStream_t s) {
    if (nwork == 0) return SE_OK;
    const TcgMember* Ms = g.d_members.as<TcgMember>();
    switch (g.N) {
#define SE_CASE(NN) \
    case NN: hipLaunchKernelGGL((k_tcg_start_batch<NN, T, RESID>), dim3(nwork), dim3(TCG_NT), 0, s, Ms, work, nwork); break;
        SE_BATCH_ORDERS(SE_CASE)
#undef SE_CASE
        default: return fail(SE_ERR_UNSUPPORTED, "batched solve built for orders 2..8");
    }
    return SE_OK;
}
int g_apply64(const Group& g, const GLL& gl, int nwork, const int2* work, hipStream_t s) {
    if (nwork == 0) return SE_OK;
    const ApplyMember* As = g.d_amembers.as<ApplyMember>();
    switch (g.N) {
#define SE_CASE(NN)                                                                                                  \
    case NN:                                                                                                         \
        hipLaunchKernelGGL((k_apply_lines_batch<NN, AP_PLAIN>), dim3(nwork), dim3(256), 0, s, As, work, nwork, 0);                                                                             \
        break;
        SE_BATCH_ORDERS(SE_CASE)
#undef SE_CASE
        default: return fail(SE_ERR_UNSUPPORTED, "batched solve built for orders 2..8");
    }
    return SE_OK;
}

int apply_tiles_b(const se_mesh* m) { return line_tiles(m->N, m->nEls[0], m->nEls[1]); }

// one graph: the k-iteration chains of every group on their own streams
template <typename T>
int chains(se_batch* b, int k) {
    hipStream_t s = b->ctx->stream;
    const int key = k * 4 + (sizeof(T) == 2 ? 1 : 0) + 2 * (b->tuned_dir > 0 ? 1 : 0);
    if (!b->graph || b->graph_key != key) {
        if (b->graph) (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
        hipGraph_t gr = nullptr;
        SE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = SE_OK;
        hipError_t e = hipEventRecord(b->fork, s);
        for (auto& g : b->groups) {
            if (e == hipSuccess) e = hipStreamWaitEvent(g.stream, b->fork, 0);
            // full-size grid argument; the kernels read the active count
            const int nw = (int)g.tiles.size();
            const int2* w = g.d_active.as<int2>();
            const int* nd = g.d_nactive.as<int>();
            for (int i = 0; i < k && !rc && e == hipSuccess; i++) {
                rc = g_update<T>(g, nw, w, g.grid_upd, g.stream, nd);
                if (!rc) rc = g_dir_apply<T, TCG_NORMAL>(g, nw, w, g.grid_dir, g.stream, nd);
            }
            if (e == hipSuccess) e = hipEventRecord(g.done, g.stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(s, g.done, 0);
        }
        hipError_t e2 = hipStreamEndCapture(s, &gr);
        if (rc || e != hipSuccess || e2 != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            if (rc) return rc;
            return fail(SE_ERR_HIP, std::string("batch graph capture: ") +
                                        hipGetErrorString(e != hipSuccess ? e : e2));
        }
        e = hipGraphInstantiate(&b->graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) {
            b->graph = nullptr;
            return fail(SE_ERR_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(e));
        }
        b->graph_key = key;
    }
    SE_HIP(hipGraphLaunch(b->graph, s));
    return SE_OK;
}

// each group's iteration work list := the items of its members still running
int compact(se_batch* b, const std::vector<TcgControl>& ctl, hipStream_t s) {
    for (auto& g : b->groups) {
        g.h_active.clear();
        for (const int2& it : g.tiles)
            if (ctl[g.members[it.x]].status == CG_RUNNING) g.h_active.push_back(it);
        *g.h_nactive = (int)g.h_active.size();
        int rc = upload(b, g.d_active, g.h_active, s);
        if (rc) return rc;
        SE_HIP(hipMemcpyAsync(g.d_nactive.ptr, g.h_nactive, sizeof(int), hipMemcpyHostToDevice, s));
    }
    return SE_OK;
}

void set_grids(se_batch* b, int fdir, int fupd) {
    for (auto& g : b->groups) {
        const int nw = std::max(1, (int)g.tiles.size());
        g.grid_dir = fdir <= 0 ? nw : std::min(nw, fdir * b->ctx->cu_count);
        g.grid_upd = fupd <= 0 ? nw : std::min(nw, fupd * b->ctx->cu_count);
    }
}

template <typename T>
int batch_solve(se_batch* b, const double* const* d_f, double* const* d_x, const se_solve_opts& o,
                se_solve_info* infos) {
    se_ctx* c = b->ctx;
    hipStream_t s = c->stream;
    const int count = (int)b->ops.size();
    const int null = o.null_space ? 1 : 0;
    const int k = std::max(1, o.check_every);
    std::vector<TcgControl> ctl(count);
    for (auto& t : ctl) t.configure(o);
    b->arena_used = 0;  // the previous solve ended with a synchronisation

    // descriptors, work items, states
    for (auto& g : b->groups) {
        std::vector<TcgMember>& Ms = g.h_members;
        std::vector<ApplyMember>& As = g.h_amembers;
        Ms.clear();
        As.clear();
        for (int i : g.members) {
            se_op* op = b->ops[i];
            TcgMember M = tcg_member(op, d_f[i], o.rtol, null, vb_of(op->mesh->n));
            M.st = b->d_states.as<CgState>() + i;
            Ms.push_back(M);
            ApplyMember A;
            A.P = op_params(op, (double)op->sign);
            A.in = op->x.as<double>();
            A.out = op->q.as<double>();
            As.push_back(A);
        }
        int rc = upload(b, g.d_members, Ms, s);
        if (!rc) rc = upload(b, g.d_amembers, As, s);
        if (rc) return rc;
    }
    for (int i = 0; i < count; i++) {
        CgState h{};
        h.maxit = ctl[i].maxit;
        h.null_space = null;
        h.n = b->ops[i]->mesh->n;
        h.status = CG_RUNNING;
        b->h_states[i] = h;
    }
    auto push = [&]() -> int {
        SE_HIP(hipMemcpyAsync(b->d_states.ptr, b->h_states, sizeof(CgState) * count, hipMemcpyHostToDevice, s));
        return SE_OK;
    };
    auto pull = [&]() -> int {
        SE_HIP(hipMemcpyAsync(b->h_states, b->d_states.ptr, sizeof(CgState) * count, hipMemcpyDeviceToHost, s));
        SE_HIP(hipStreamSynchronize(s));
        b->arena_used = 0;  // every staged upload has completed
        return SE_OK;
    };

    SE_HIP(hipEventRecord(c->ev0, s));
    int rc = push();
    if (rc) return rc;
    if (null) {
        for (int i = 0; i < count; i++) {
            se_op* op = b->ops[i];
            const int n = op->mesh->n;
            hipLaunchKernelGGL(k_batch_sum, dim3(vb_of(n)), dim3(1024), 0, s, n, d_f[i], (double)op->sign,
                               op->sep_part.as<double>() + 7 * (size_t)tcg_tiles(op->mesh));
        }
    }
    for (auto& g : b->groups) {
        const int nw = (int)g.tiles.size();
        rc = g_start<T, false>(g, nw, g.d_tiles.as<int2>(), s);
        if (!rc) rc = g_dir_apply<T, TCG_FIRST>(g, nw, g.d_tiles.as<int2>(), nw, s);
        if (rc) return rc;
    }
    SE_HIP(hipGetLastError());
    rc = pull();
    if (rc) return rc;
    int running = 0;
    for (int i = 0; i < count; i++) {
        ctl[i].started(b->h_states[i]);
        running += ctl[i].status == CG_RUNNING;
    }
    rc = compact(b, ctl, s);
    if (rc) return rc;

    // shape key of this batch: its largest group
    const Group* big = &b->groups[0];
    for (auto& g : b->groups)
        if (g.tiles.size() > big->tiles.size()) big = &g;
    int lg = 0;
    while ((size_t(1) << lg) < big->tiles.size()) lg++;
    const std::pair<int, int> shape(big->N, lg);
    if (!b->tuned_dir) {
        auto it = c->tuned_grid.find(shape);
        if (it != c->tuned_grid.end()) {
            b->tuned_dir = it->second / 64;
            b->tuned_upd = it->second % 64;
        }
    }
    // autotuner candidates: grid = all items (one item per workgroup), or
    // 16, 8, 6, 4, 3, 2 workgroups per CU (each a contiguous run of items),
    // one factor for both kernels: the order groups run concurrently in the
    // graph, so the whole chain is timed (a kernel's best grid alone is not
    // its best beside the other groups).  Phase 2j captures candidate j and
    // runs it once, phase 2j+1 times a replay; every launch is followed by
    // the regular check, so the check cadence (and with it every result) is
    // that of the single solve.
    static const int cand[] = {0, 16, 8, 6, 4, 3, 2};
    constexpr int ncand = (int)(sizeof(cand) / sizeof(cand[0]));
    int phase = b->tuned_dir ? -1 : 0;
    if (phase < 0) set_grids(b, b->tuned_dir - 1, b->tuned_upd - 1);
    for (int round = 0; running > 0; round++) {
        const int ci = phase / 2;
        if (phase >= 0 && phase % 2 == 0) {
            set_grids(b, cand[ci], cand[ci]);
            b->graph_key = -1;
        }
        if (phase >= 0 && phase % 2 == 1) SE_HIP(hipEventRecord(b->t0, s));
        // tuning ends before the replay count grows (14 phases < 16 rounds)
        const int reps = replays_for_round(round);
        for (int r = 0; r < reps && !rc; r++) rc = chains<T>(b, k);
        if (rc) return rc;
        if (phase >= 0 && phase % 2 == 1) {
            SE_HIP(hipEventRecord(b->t1, s));
            SE_HIP(hipEventSynchronize(b->t1));
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, b->t0, b->t1);
            if (ci == 0 || ms < b->tune_best) {
                b->tune_best = ms;
                b->tuned_dir = b->tuned_upd = cand[ci] + 1;
            }
        }
        if (phase >= 0 && ++phase == 2 * ncand) {
            phase = -1;
            set_grids(b, b->tuned_dir - 1, b->tuned_upd - 1);
            b->graph_key = -1;
            c->tuned_grid[shape] = b->tuned_dir * 64 + b->tuned_upd;
        }
        rc = pull();
        if (rc) return rc;

        // members due for a reliable update
        std::vector<int> due;
        for (int i = 0; i < count; i++)
            if (ctl[i].status == CG_RUNNING && ctl[i].due(b->h_states[i], k * reps)) {
                ctl[i].before_update(b->h_states[i]);
                due.push_back(i);
            }
        if (due.empty()) continue;
        std::vector<char> is_due(count, 0);
        for (int i : due) is_due[i] = 1;
        for (auto& g : b->groups) {
            std::vector<int2>&ut = g.h_ut, &uf = g.h_uf, &ua = g.h_ua;
            std::vector<int>& ul = g.h_ul;
            ut.clear(), uf.clear(), ua.clear(), ul.clear();
            for (int j = 0; j < (int)g.members.size(); j++) {
                const int i = g.members[j];
                if (!is_due[i]) continue;
                const se_mesh* m = b->ops[i]->mesh;
                ul.push_back(j);
                for (int t = 0; t < tcg_tiles(m); t++) ut.push_back(make_int2(j, t));
                for (int t = 0; t < (m->n + TCG_FLUSH - 1) / TCG_FLUSH; t++) uf.push_back(make_int2(j, t));
                for (int t = 0; t < apply_tiles_b(m); t++) ua.push_back(make_int2(j, t));
            }
            if (ul.empty()) continue;
            rc = upload(b, g.d_ut, ut, s);
            if (!rc) rc = upload(b, g.d_uf, uf, s);
            if (!rc) rc = upload(b, g.d_ua, ua, s);
            if (!rc) rc = upload(b, g.d_ul, ul, s);
            if (rc) return rc;
            const TcgMember* Ms = g.d_members.as<TcgMember>();
            hipLaunchKernelGGL(k_tcg_flush_batch, dim3((int)uf.size()), dim3(TCG_NT), 0, s, Ms, g.d_uf.as<int2>(),
                               (int)uf.size());
            rc = g_apply64(g, b->ops[g.members[0]]->mesh->gll, (int)ua.size(), g.d_ua.as<int2>(), s);
            if (!rc) rc = g_start<T, true>(g, (int)ut.size(), g.d_ut.as<int2>(), s);
            if (rc) return rc;
            hipLaunchKernelGGL(k_tcg_true_batch, dim3((int)ul.size()), dim3(TCG_NT), 0, s, Ms, g.d_ul.as<int>(),
                               (int)ul.size());
        }
        SE_HIP(hipGetLastError());
        rc = pull();
        if (rc) return rc;
        // 1: continue the direction, 2: restart it
        std::vector<char> restart(count, 0);
        bool finished = false;
        for (int i : due) {
            CgState hh = b->h_states[i];
            if (ctl[i].after_update(hh) == CG_RUNNING) {
                restart[i] = ctl[i].reset_dir ? 2 : 1;
            } else {
                hh.stop = 1;  // finished: its work items leave the iteration list
                running--;
                finished = true;
            }
            b->h_states[i] = hh;
        }
        rc = push();
        if (!rc && finished) rc = compact(b, ctl, s);
        if (rc) return rc;
        for (auto& g : b->groups) {
            // items of continuing members first, then of restarting ones
            std::vector<int2>& rt = g.h_rt;
            rt.clear();
            int ncont = 0;
            for (int pass = 1; pass <= 2; pass++) {
                for (int j = 0; j < (int)g.members.size(); j++) {
                    const int i = g.members[j];
                    if (restart[i] != pass) continue;
                    for (int t = 0; t < tcg_tiles(b->ops[i]->mesh); t++) rt.push_back(make_int2(j, t));
                }
                if (pass == 1) ncont = (int)rt.size();
            }
            if (rt.empty()) continue;
            rc = upload(b, g.d_rt, rt, s);
            const int2* w = g.d_rt.as<int2>();
            const int nres = (int)rt.size() - ncont;
            if (!rc) rc = g_dir_apply<T, TCG_NORMAL>(g, ncont, w, ncont, s);
            if (!rc) rc = g_dir_apply<T, TCG_RESET>(g, nres, w + ncont, nres, s);
            if (rc) return rc;
        }
    }
    for (int i = 0; i < count; i++) {
        se_op* op = b->ops[i];
        if (d_x[i] != op->x.as<double>())
            SE_HIP(hipMemcpyAsync(d_x[i], op->x.ptr, sizeof(double) * op->mesh->n, hipMemcpyDeviceToDevice, s));
    }
    SE_HIP(hipEventRecord(c->ev1, s));
    SE_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    for (int i = 0; i < count; i++)
        if (infos) ctl[i].fill(infos[i], b->h_states[i].iter, ms * 1e-3);
    return SE_OK;
}

}  // namespace

extern "C" int se_batch_create(se_ctx* ctx, int count, se_op* const* ops, se_batch** out) {
    SE_CHECK(ctx && out && (count == 0 || ops), SE_ERR_ARG, "se_batch_create: NULL argument");
    SE_CHECK(count >= 0, SE_ERR_ARG, "se_batch_create: negative count");
    SE_HIP(hipSetDevice(ctx->device));
    for (int i = 0; i < count; i++) {
        SE_CHECK(ops[i] != nullptr, SE_ERR_ARG, "se_batch_create: NULL member");
        SE_CHECK(ops[i]->mesh->ctx == ctx, SE_ERR_ARG, "se_batch_create: member created on another context");
        SE_CHECK(ops[i]->definite, SE_ERR_UNSUPPORTED, "se_batch_create: member operator is not definite");
        SE_CHECK(ops[i]->mesh->N >= 2 && ops[i]->mesh->N <= 8, SE_ERR_UNSUPPORTED,
                 "se_batch_create: orders 2..8");
        for (int j = 0; j < i; j++)
            SE_CHECK(ops[j] != ops[i], SE_ERR_ARG, "se_batch_create: an operator appears twice");
    }
    se_batch* b = new se_batch;
    b->ctx = ctx;
    b->ops.assign(ops, ops + count);
    for (int i = 0; i < count; i++)
        if (ops[i]->svv_mN > 0) b->jacobi_ok = false;
    int rc = SE_OK;
    for (int i = 0; i < count && !rc && b->jacobi_ok; i++) rc = tcg_prepare(ops[i]);
    if (!rc) rc = b->d_states.alloc(sizeof(CgState) * std::max(count, 1));
    if (!rc && hipHostMalloc((void**)&b->h_states, sizeof(CgState) * std::max(count, 1), hipHostMallocDefault) !=
                   hipSuccess)
        rc = fail(SE_ERR_NOMEM, "se_batch_create: pinned state buffer");
    if (!rc && (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
                hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess))
        rc = fail(SE_ERR_HIP, "se_batch_create: events");
    for (int N = 2; N <= 8 && !rc && b->jacobi_ok; N++) {
        Group g;
        g.N = N;
        for (int i = 0; i < count; i++)
            if (ops[i]->mesh->N == N) g.members.push_back(i);
        if (g.members.empty()) continue;
        for (int j = 0; j < (int)g.members.size(); j++)
            for (int t = 0; t < tcg_tiles(ops[g.members[j]]->mesh); t++) g.tiles.push_back(make_int2(j, t));
        if (hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&g.done, hipEventDisableTiming) != hipSuccess)
            rc = fail(SE_ERR_HIP, "se_batch_create: group stream");
        if (!rc) rc = upload(b, g.d_tiles, g.tiles, ctx->stream);
        if (!rc) rc = g.d_active.alloc(sizeof(int2) * std::max<size_t>(g.tiles.size(), 1));
        if (!rc) rc = g.d_nactive.alloc(sizeof(int));
        if (!rc && hipHostMalloc((void**)&g.h_nactive, sizeof(int), hipHostMallocDefault) != hipSuccess)
            rc = fail(SE_ERR_NOMEM, "se_batch_create: pinned count");
        b->groups.push_back(std::move(g));
    }
    if (!rc) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(SE_ERR_HIP, std::string("se_batch_create: ") + hipGetErrorString(e));
    }
    if (rc) {
        b->release();
        delete b;
        return rc;
    }
    *out = b;
    return SE_OK;
}

extern "C" int se_batch_destroy(se_batch* b) {
    if (!b) return SE_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    b->release();
    delete b;
    return SE_OK;
}

extern "C" int se_batch_size(const se_batch* b) { return b ? (int)b->ops.size() : 0; }

extern "C" int se_batch_launch_shape(const se_batch* b, int* dir_per_cu, int* update_per_cu) {
    SE_CHECK(b && dir_per_cu && update_per_cu, SE_ERR_ARG, "se_batch_launch_shape: NULL argument");
    *dir_per_cu = b->tuned_dir - 1;
    *update_per_cu = b->tuned_upd - 1;
    return SE_OK;
}

namespace {
// time `reps` back-to-back launches of each work-item kernel of every group,
// continuing the CG recurrence of the last solve with every member running
template <typename T>
int batch_probe(se_batch* b, int reps, int max_groups, int* orders, double* us_dir, double* us_upd,
                double* bytes_dir, double* bytes_upd) {
    hipStream_t s = b->ctx->stream;
    const int count = (int)b->ops.size();
    for (auto& g : b->groups)
        SE_CHECK(!g.h_members.empty(), SE_ERR_STATE, "se_batch_probe: run a solve of this batch first");
    for (int i = 0; i < count; i++) {
        CgState& h = b->h_states[i];
        h.stop = 0;
        h.status = CG_RUNNING;
        h.tol2 = 0.0;
        h.iter = 0;
        h.maxit = 1 << 30;
        h.rho = h.rho_prev = 1.0;
        h.beta = 0.0;
    }
    SE_HIP(hipMemcpyAsync(b->d_states.ptr, b->h_states, sizeof(CgState) * count, hipMemcpyHostToDevice, s));
    int ng = 0;
    for (auto& g : b->groups) {
        if (ng >= max_groups) break;
        const int nw = (int)g.tiles.size();
        double n = 0.0;
        for (int i : g.members) n += b->ops[i]->mesh->n;
        float ms_d = 0.f, ms_u = 0.f;
        int rc = g_dir_apply<T, TCG_NORMAL>(g, nw, g.d_tiles.as<int2>(), g.grid_dir > 0 ? g.grid_dir : nw, s);
        if (!rc) rc = g_update<T>(g, nw, g.d_tiles.as<int2>(), g.grid_upd > 0 ? g.grid_upd : nw, s);
        if (rc) return rc;
        SE_HIP(hipEventRecord(b->t0, s));
        for (int r = 0; r < reps && !rc; r++)
            rc = g_dir_apply<T, TCG_NORMAL>(g, nw, g.d_tiles.as<int2>(), g.grid_dir > 0 ? g.grid_dir : nw, s);
        SE_HIP(hipEventRecord(b->t1, s));
        SE_HIP(hipEventSynchronize(b->t1));
        (void)hipEventElapsedTime(&ms_d, b->t0, b->t1);
        SE_HIP(hipEventRecord(b->t0, s));
        for (int r = 0; r < reps && !rc; r++)
            rc = g_update<T>(g, nw, g.d_tiles.as<int2>(), g.grid_upd > 0 ? g.grid_upd : nw, s);
        SE_HIP(hipEventRecord(b->t1, s));
        SE_HIP(hipEventSynchronize(b->t1));
        (void)hipEventElapsedTime(&ms_u, b->t0, b->t1);
        if (rc) return rc;
        const double sT = (double)sizeof(T);
        orders[ng] = g.N;
        us_dir[ng] = 1e3 * ms_d / reps;
        us_upd[ng] = 1e3 * ms_u / reps;
        bytes_dir[ng] = n * (8.0 + sT);         // r, p read; q written
        bytes_upd[ng] = n * (20.0 + 2.0 * sT);  // p, r, q, x read; p, x, r written
        ng++;
    }
    // the members' states are now those of the probe: the next solve starts afresh
    return ng;
}
}  // namespace

extern "C" int se_batch_probe(se_batch* b, int precision, int reps, int max_groups, int* orders, double* us_dir,
                              double* us_upd, double* bytes_dir, double* bytes_upd) {
    SE_CHECK(b && orders && us_dir && us_upd && bytes_dir && bytes_upd, SE_ERR_ARG, "se_batch_probe: NULL argument");
    SE_CHECK(reps > 0 && max_groups > 0, SE_ERR_ARG, "se_batch_probe: reps and max_groups must be positive");
    SE_CHECK(precision == SE_PREC_FP32_IR || precision == SE_PREC_BF16_IR, SE_ERR_ARG,
             "se_batch_probe: precision must be SE_PREC_FP32_IR or SE_PREC_BF16_IR");
    SE_HIP(hipSetDevice(b->ctx->device));
    if (precision == SE_PREC_FP32_IR) return batch_probe<float>(b, reps, max_groups, orders, us_dir, us_upd, bytes_dir, bytes_upd);
    return batch_probe<__hip_bfloat16>(b, reps, max_groups, orders, us_dir, us_upd, bytes_dir, bytes_upd);
}

extern "C" int se_batch_pcg_dev(se_batch* b, const double* const* d_f, double* const* d_x, const se_solve_opts* opts,
                                se_solve_info* infos) {
    SE_CHECK(b && (b->ops.empty() || (d_f && d_x)), SE_ERR_ARG, "se_batch_pcg_dev: NULL argument");
    se_solve_opts o;
    se_solve_opts_default(&o);
    if (opts) o = *opts;
    SE_CHECK(o.preconditioner >= SE_PC_DEFAULT && o.preconditioner <= SE_PC_FDM, SE_ERR_ARG,
             "se_batch_pcg_dev: bad preconditioner");
    const bool fdm = solve_pc(o) == SE_PC_FDM;
    SE_CHECK(o.precision == SE_PREC_FP32_IR || o.precision == SE_PREC_BF16_IR ||
                 (fdm && (o.precision == SE_PREC_FP64 || o.precision == SE_PREC_FP64_FP32C ||
                          o.precision == SE_PREC_FP64_BF16C)),
             SE_ERR_UNSUPPORTED,
             "se_batch_pcg_dev: batched Jacobi solves are mixed precision (SE_PREC_FP32_IR or SE_PREC_BF16_IR); "
             "FP64 batches take the FDM preconditioner");
    for (size_t i = 0; i < b->ops.size(); i++)
        SE_CHECK(d_f[i] && d_x[i], SE_ERR_ARG, "se_batch_pcg_dev: NULL member vector");
    SE_HIP(hipSetDevice(b->ctx->device));
    if (b->ops.empty()) return SE_OK;
    if (fdm) {
        if (!b->fdm) {
            int rc = fdm_plan_create(b->ctx, (int)b->ops.size(), b->ops.data(), &b->fdm);
            if (rc) return rc;
        }
        return fdm_plan_solve(b->fdm, d_f, d_x, o, infos);
    }
    SE_CHECK(b->jacobi_ok, SE_ERR_UNSUPPORTED,
             "se_batch_pcg_dev: a batch with SVV members solves with the FDM preconditioner");
    if (o.precision == SE_PREC_FP32_IR) return batch_solve<float>(b, d_f, d_x, o, infos);
    return batch_solve<__hip_bfloat16>(b, d_f, d_x, o, infos);
}

extern "C" int se_batch_pcg(se_batch* b, const double* const* f, double* const* x, const se_solve_opts* opts,
                            se_solve_info* infos) {
    SE_CHECK(b && (b->ops.empty() || (f && x)), SE_ERR_ARG, "se_batch_pcg: NULL argument");
    const int count = (int)b->ops.size();
    SE_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    std::vector<const double*> df(count);
    std::vector<double*> dx(count);
    for (int i = 0; i < count; i++) {
        se_op* op = b->ops[i];
        SE_CHECK(f[i] && x[i], SE_ERR_ARG, "se_batch_pcg: NULL member vector");
        int rc = op_ensure_work(op);
        if (rc) return rc;
        SE_HIP(hipMemcpyAsync(op->f.ptr, f[i], sizeof(double) * op->mesh->n, hipMemcpyHostToDevice, s));
        df[i] = op->f.as<double>();
        dx[i] = op->x.as<double>();
    }
    int rc = se_batch_pcg_dev(b, df.data(), dx.data(), opts, infos);
    if (rc) return rc;
    for (int i = 0; i < count; i++)
        SE_HIP(hipMemcpyAsync(x[i], b->ops[i]->x.ptr, sizeof(double) * b->ops[i]->mesh->n, hipMemcpyDeviceToHost, s));
    SE_HIP(hipStreamSynchronize(s));
    return SE_OK;
}

extern "C" int se_batch_fdm_probe(se_batch* b, int reps, double* us_pass, double* flops_pass, double* us_apply,
                                  double* bytes_apply) {
    SE_CHECK(b && us_pass && flops_pass && us_apply && bytes_apply && reps > 0, SE_ERR_ARG,
             "se_batch_fdm_probe: bad argument");
    SE_CHECK(b->fdm != nullptr, SE_ERR_STATE, "se_batch_fdm_probe: run an FDM solve of this batch first");
    SE_HIP(hipSetDevice(b->ctx->device));
    return fdm_plan_probe(b->fdm, reps, us_pass, flops_pass, us_apply, bytes_apply);
}
