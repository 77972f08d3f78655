// serve.hip -- request server for two-layer QG time-step requests.
//
// A server owns a context (stream) and one host thread.  Clients submit
// requests (mesh, parameters, number of steps, initial state) from any thread;
// the server thread admits them in arrival order while their device memory
// fits the budget, and runs continuous batching: every tick advances each
// admitted request by one step, with the per-request kernels (velocities, PV,
// semi-Lagrangian advection, right-hand sides, field updates) launched per
// request and the two inversions of all admitted requests run as two batched
// mixed-precision solves.  A request that finishes leaves at the end of the
// tick and one that arrived meanwhile joins at the start of the next, so the
// work of requests at different stages overlaps on the device.  Meshes
// (numbering maps, geometry) are cached and model instances (fields, operators,
// solver buffers) pooled, so both stay resident in HBM across requests.
//
// A request's results do not depend on what else is in the batch: the batched
// solver gives every member bitwise its single-solve result, and every other
// kernel of a step touches only that request's data.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "qg.h"
#include "se_common.h"
#include "se_op.h"

using namespace se;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    se_qg2_request spec{};
    std::vector<double> init;  // psi1, om1, psi2, om2 (nV each)
    std::vector<double> tau_e; // relaxation target (nV) or empty
    std::vector<double> out;   // final fields
    se_request_result res{};
    bool done = false;
    se_qg2* q = nullptr;
    double dt = 0.0;
    size_t bytes = 0;
    std::string error;  // why the request failed (status != SE_OK)
};

using MeshKey = std::tuple<int, int, int, double, double, double, double, int, int, int, int, int, int>;
using PoolKey = std::tuple<se_mesh*, double, double, double, int>;

MeshKey mesh_key(const se_qg2_request& r) {
    return MeshKey(r.nEls[0], r.nEls[1], r.N, r.min[0], r.min[1], r.max[0], r.max[1], r.periodic[0], r.periodic[1],
                   r.bc[0], r.bc[1], r.bc[2], r.bc[3]);
}

}  // namespace

struct se_server {
    se_server_opts opts{};
    se_ctx* ctx = nullptr;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::map<int64_t, std::shared_ptr<Request>> all;
    bool stopping = false;
    int64_t next_id = 0;
    Clock::time_point t0;
    std::string error;
    // server-thread state
    std::map<MeshKey, se_mesh*> meshes;
    std::multimap<PoolKey, se_qg2*> pool;
    size_t used = 0;  // bytes of admitted requests and pooled instances
    // the inversions' solvers, kept across ticks: while the admitted set does
    // not change, every tick's passes find their FDM plan (descriptors and
    // work lists on the device) or Jacobi batch (graphs and launch shapes
    // tuned) as the last tick left it.  A pass with a new member set takes
    // over the least recently used of kPlans FDM plans (re-membered: its
    // buffers are reused, nothing is allocated unless the set is larger than
    // any before), or a new Jacobi batch (at most 8)
    static constexpr int kPlans = 4;
    struct PlanSlot {
        std::vector<se_op*> ops;
        FdmPlan* plan = nullptr;
        unsigned long long used = 0;
    };
    PlanSlot plans[kPlans];
    unsigned long long plan_clock = 0;
    std::map<std::vector<se_op*>, se_batch*> batches;
    // Lanes: the requests' own phases (advection, right-hand sides, updates)
    // run on kLanes streams beside the context's stream, which runs the
    // batched inversions; a small mesh fills a few CUs, so requests overlap.
    // Requests of one mesh share its scratch (the SL velocity fields) and so
    // one lane; every inversion waits for all lanes and they for it (events).
    static constexpr int kLanes = 3;  // with the main stream: the process's four hardware queues
    hipStream_t lanes[kLanes] = {};
    hipEvent_t lane_ev[kLanes] = {}, main_ev = nullptr;
    se_server_stats stats{};

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    void run();
    int lanes_fork();
    int lanes_join();
    int batch_solve(const std::vector<se_op*>& ops, const se_solve_opts& o, std::vector<se_solve_info>& inf);
    void drop_solvers(se_qg2* q);
    int tick(std::vector<std::shared_ptr<Request>>& active);
    int admit(std::vector<std::shared_ptr<Request>>& active);
    void retire(const std::shared_ptr<Request>& r, int status);
};

namespace {

}  // namespace

// the requests' phases launch on a lane: the context's stream is that lane
// while this is in scope (the server thread is the context's only user)
struct LaneScope {
    se_ctx* c;
    hipStream_t saved;
    LaneScope(se_ctx* ctx, hipStream_t s) : c(ctx), saved(ctx->stream) { c->stream = s; }
    ~LaneScope() { c->stream = saved; }
};

// the lanes start after everything queued on the main stream so far
int se_server::lanes_fork() {
    SE_HIP(hipEventRecord(main_ev, ctx->stream));
    for (int l = 0; l < kLanes; l++) SE_HIP(hipStreamWaitEvent(lanes[l], main_ev, 0));
    return SE_OK;
}

// the main stream continues after everything queued on the lanes so far
int se_server::lanes_join() {
    for (int l = 0; l < kLanes; l++) {
        SE_HIP(hipEventRecord(lane_ev[l], lanes[l]));
        SE_HIP(hipStreamWaitEvent(ctx->stream, lane_ev[l], 0));
    }
    return SE_OK;
}

// one inversion pass of the admitted requests as one batched solve
int se_server::batch_solve(const std::vector<se_op*>& ops, const se_solve_opts& o, std::vector<se_solve_info>& inf) {
    inf.assign(ops.size(), se_solve_info{});
    if (ops.empty()) return SE_OK;
    std::vector<const double*> f(ops.size());
    std::vector<double*> x(ops.size());
    for (size_t i = 0; i < ops.size(); i++) {
        f[i] = ops[i]->f.as<double>();
        x[i] = ops[i]->x.as<double>();
    }
    if (solve_pc(o) == SE_PC_FDM) {
        PlanSlot* slot = nullptr;
        for (PlanSlot& p : plans)
            if (p.plan && p.ops == ops) slot = &p;
        if (!slot) {  // an empty slot, else the least recently used
            for (PlanSlot& p : plans) {
                if (!p.plan) {
                    slot = &p;
                    break;
                }
                if (!slot || p.used < slot->used) slot = &p;
            }
            int rc = slot->plan ? fdm_plan_set_members(slot->plan, (int)ops.size(), ops.data())
                                : fdm_plan_create(ctx, (int)ops.size(), ops.data(), &slot->plan);
            if (rc) {
                if (slot->plan) fdm_plan_destroy(slot->plan);
                *slot = PlanSlot{};
                return rc;
            }
            slot->ops = ops;
        }
        slot->used = ++plan_clock;
        return fdm_plan_solve(slot->plan, f.data(), x.data(), o, inf.data());
    }
    // Jacobi inner CG
    auto it = batches.find(ops);
    if (it == batches.end()) {
        if (batches.size() >= 8) {  // bound the cache: drop the oldest sets
            for (auto& kv : batches) se_batch_destroy(kv.second);
            batches.clear();
        }
        se_batch* b = nullptr;
        int rc = se_batch_create(ctx, (int)ops.size(), ops.data(), &b);
        if (rc) return rc;
        it = batches.emplace(ops, b).first;
    }
    return se_batch_pcg_dev(it->second, f.data(), x.data(), &o, inf.data());
}

int se_server::admit(std::vector<std::shared_ptr<Request>>& active) {
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (queue.empty()) return SE_OK;
            if (opts.max_active > 0 && (int)active.size() >= opts.max_active) return SE_OK;
            r = queue.front();
        }
        const se_qg2_request& s = r->spec;
        // mesh (cached)
        se_mesh* m = nullptr;
        auto mk = mesh_key(s);
        auto it = meshes.find(mk);
        if (it != meshes.end()) {
            m = it->second;
        } else {
            int rc = se_mesh_create(ctx, s.nEls, s.N, s.min, s.max, s.periodic, s.bc, &m);
            if (rc) {
                std::lock_guard<std::mutex> lk(mu);
                queue.pop_front();
                retire(r, rc);
                continue;
            }
            meshes[mk] = m;
        }
        const size_t need = qg2_bytes(m->n, m->nV);
        PoolKey pk(m, s.F1, s.F2, s.beta, opts.detj_mode);
        auto pit = pool.find(pk);
        const bool pooled = pit != pool.end();
        if (!pooled) {
            // free pooled instances of other shapes while the budget is short
            while (used + need > (size_t)opts.memory_budget && !pool.empty()) {
                auto victim = pool.begin();
                used -= qg2_bytes(victim->second->mesh->n, victim->second->mesh->nV);
                drop_solvers(victim->second);
                se_qg2_destroy(victim->second);
                pool.erase(victim);
            }
            if (used + need > (size_t)opts.memory_budget) {
                if (active.empty() && pool.empty()) {  // can never fit
                    std::lock_guard<std::mutex> lk(mu);
                    queue.pop_front();
                    retire(r, SE_ERR_NOMEM);
                    continue;
                }
                return SE_OK;  // wait for memory to free up
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            queue.pop_front();
        }
        se_qg2* q = nullptr;
        if (pooled) {
            q = pit->second;
            pool.erase(pit);
        } else {
            int rc = se_qg2_create(m, s.F1, s.F2, s.beta, opts.detj_mode, &q);
            if (rc) {
                std::lock_guard<std::mutex> lk(mu);
                retire(r, rc);
                continue;
            }
            used += need;
        }
        r->q = q;
        r->bytes = need;
        int rc = se_qg2_set_state(q, r->init.data(), r->init.data() + m->nV, r->init.data() + 2 * (size_t)m->nV,
                                  r->init.data() + 3 * (size_t)m->nV);
        // a pooled instance carries the previous request's dissipation: set it every time
        if (!rc) rc = se_qg2_set_dissipation(q, s.E2, s.r, s.nu, r->tau_e.empty() ? nullptr : r->tau_e.data());
        double dt = s.dt;
        if (!rc && !(dt > 0.0)) rc = se_qg2_timestep(q, 0.5, &dt);
        if (rc) {
            std::lock_guard<std::mutex> lk(mu);
            retire(r, rc);
            continue;
        }
        r->dt = dt;
        r->res.t_admit = now();
        active.push_back(r);
        stats.admitted++;
    }
}

// forget every solver that holds an operator of q (before q is destroyed)
void se_server::drop_solvers(se_qg2* q) {
    for (PlanSlot& p : plans) {
        bool uses = false;
        for (se_op* op : p.ops) uses = uses || op == qg2_pass_op(q, 0) || op == qg2_pass_op(q, 1);
        if (uses && p.plan) {
            fdm_plan_destroy(p.plan);
            p = PlanSlot{};
        }
    }
    for (auto it = batches.begin(); it != batches.end();) {
        bool uses = false;
        for (se_op* op : it->first) uses = uses || op == qg2_pass_op(q, 0) || op == qg2_pass_op(q, 1);
        if (uses) {
            se_batch_destroy(it->second);
            it = batches.erase(it);
        } else {
            ++it;
        }
    }
}

void se_server::retire(const std::shared_ptr<Request>& r, int status) {
    // server thread, with mu held
    r->res.status = status;
    if (status != SE_OK && r->error.empty()) r->error = se_last_error();
    r->res.t_done = now();
    if (r->q) {
        se_mesh* m = r->q->mesh;
        if (status == SE_OK) {
            r->out.resize(4 * (size_t)m->nV);
            (void)se_qg2_get_state(r->q, r->out.data(), r->out.data() + m->nV, r->out.data() + 2 * (size_t)m->nV,
                                   r->out.data() + 3 * (size_t)m->nV);
        }
        pool.emplace(PoolKey(m, r->q->F1, r->q->F2, r->q->beta, opts.detj_mode), r->q);
        r->q = nullptr;
    }
    r->done = true;
    stats.completed++;
    cv_done.notify_all();
}

int se_server::tick(std::vector<std::shared_ptr<Request>>& active) {
    se_solve_opts o;
    se_solve_opts_default(&o);
    o.precision = opts.precision;
    o.rtol = opts.rtol;
    o.preconditioner = opts.preconditioner;
    const bool one_by_one = opts.precision == SE_PREC_FP64 && solve_pc(o) == SE_PC_JACOBI;
    std::vector<int> failed(active.size(), 0);
    // lanes: distinct meshes round-robin, a mesh's requests together
    std::vector<int> lane(active.size());
    {
        std::map<se_mesh*, int> of_mesh;
        for (size_t i = 0; i < active.size(); i++) {
            auto it = of_mesh.emplace(active[i]->q->mesh, (int)of_mesh.size() % kLanes).first;
            lane[i] = it->second;
        }
    }
    int rc0 = lanes_fork();
    if (rc0) return rc0;
    for (size_t i = 0; i < active.size(); i++) {
        LaneScope on(ctx, lanes[lane[i]]);
        int rc = qg2_phase_advect(active[i]->q, active[i]->dt);
        if (!rc) rc = qg2_phase_rhs(active[i]->q, 0);
        if (rc) failed[i] = rc;
    }
    rc0 = lanes_join();
    if (rc0) return rc0;
    // inversion k of every request that has one: 2 per step, 4 for the
    // dissipative ones (their first pair runs beside the others' only pair)
    for (int k = 0; k < 4; k++) {
        std::vector<se_op*> ops;
        std::vector<size_t> idx;
        for (size_t i = 0; i < active.size(); i++)
            if (!failed[i] && k < qg2_passes(active[i]->q)) {
                ops.push_back(qg2_pass_op(active[i]->q, k));
                idx.push_back(i);
            }
        if (ops.empty()) break;
        std::vector<se_solve_info> inf;
        int rc = one_by_one ? SE_OK : batch_solve(ops, o, inf);
        if (one_by_one) {  // the plain path: one solve after another
            inf.assign(ops.size(), se_solve_info{});
            for (size_t j = 0; j < ops.size() && !rc; j++)
                rc = se_helm_pcg_dev(ops[j], ops[j]->f.as<double>(), ops[j]->x.as<double>(), &o, &inf[j]);
        }
        if (rc) return rc;
        rc = lanes_fork();
        if (rc) return rc;
        for (size_t j = 0; j < idx.size(); j++) {
            LaneScope on(ctx, lanes[lane[idx[j]]]);
            Request& r = *active[idx[j]];
            const int passes = qg2_passes(r.q);
            r.res.inner_iters += inf[j].iters;
            r.res.unknowns += ops[j]->mesh->n;
            if (inf[j].status != SE_SOLVE_CONVERGED) {
                failed[idx[j]] = SE_ERR_STATE;
                r.error = std::string(passes == 4 && k < 2 ? "first " : "") + ((k & 1) ? "barotropic" : "baroclinic") +
                          " inversion ended with status " + std::to_string(inf[j].status) + " after " +
                          std::to_string(inf[j].iters) + " iterations";
                continue;
            }
            int rc2 = qg2_phase_update(r.q, k);
            if (!rc2) rc2 = k + 1 < passes ? qg2_phase_rhs(r.q, k + 1) : qg2_phase_post(r.q);
            if (rc2) failed[idx[j]] = rc2;
        }
        rc = lanes_join();
        if (rc) return rc;
    }
    SE_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<std::shared_ptr<Request>> keep;
    std::lock_guard<std::mutex> lk(mu);
    for (size_t i = 0; i < active.size(); i++) {
        Request& r = *active[i];
        if (!failed[i]) {
            r.res.steps_done++;
            stats.steps++;
        }
        if (failed[i] || r.res.steps_done >= r.spec.steps) {
            // the instance goes to the pool and stays in the memory account
            retire(active[i], failed[i]);
        } else {
            keep.push_back(active[i]);
        }
    }
    stats.ticks++;
    active.swap(keep);
    return SE_OK;
}

void se_server::run() {
    (void)hipSetDevice(ctx->device);
    bool ok = hipEventCreateWithFlags(&main_ev, hipEventDisableTiming) == hipSuccess;
    for (int l = 0; l < kLanes && ok; l++)
        ok = hipStreamCreateWithFlags(&lanes[l], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&lane_ev[l], hipEventDisableTiming) == hipSuccess;
    if (!ok) {  // no request can run: refuse them all
        std::lock_guard<std::mutex> lk(mu);
        error = "request server: cannot create its streams";
        stopping = true;
        while (!queue.empty()) {
            retire(queue.front(), SE_ERR_HIP);
            queue.pop_front();
        }
    }
    std::vector<std::shared_ptr<Request>> active;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stopping || !queue.empty() || !active.empty(); });
            if (stopping && queue.empty() && active.empty()) break;
        }
        int rc = admit(active);
        if (!rc && !active.empty()) rc = tick(active);
        if (rc) {
            std::lock_guard<std::mutex> lk(mu);
            error = se_last_error();
            for (auto& r : active) retire(r, rc);
            active.clear();
            while (!queue.empty()) {
                retire(queue.front(), rc);
                queue.pop_front();
            }
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    for (int l = 0; l < kLanes; l++) {
        if (lanes[l]) (void)hipStreamSynchronize(lanes[l]);
        if (lanes[l]) (void)hipStreamDestroy(lanes[l]);
        if (lane_ev[l]) (void)hipEventDestroy(lane_ev[l]);
    }
    if (main_ev) (void)hipEventDestroy(main_ev);
}

extern "C" void se_server_opts_default(se_server_opts* o) {
    if (!o) return;
    o->memory_budget = 200e9;
    o->max_active = 0;
    o->precision = SE_PREC_FP32_IR;
    o->rtol = 1e-10;
    o->detj_mode = SE_DETJ_LITERAL;
    o->preconditioner = SE_PC_DEFAULT;
}

extern "C" int se_server_create(int device, const se_server_opts* opts, se_server** out) {
    SE_CHECK(out, SE_ERR_ARG, "se_server_create: NULL argument");
    se_server* sv = new se_server;
    se_server_opts_default(&sv->opts);
    if (opts) sv->opts = *opts;
    int rc = se_ctx_create(device, &sv->ctx);
    if (rc) {
        delete sv;
        return rc;
    }
    sv->t0 = Clock::now();
    sv->th = std::thread([sv] { sv->run(); });
    *out = sv;
    return SE_OK;
}

extern "C" int se_server_submit(se_server* sv, const se_qg2_request* req, int64_t* id) {
    SE_CHECK(sv && req && id, SE_ERR_ARG, "se_server_submit: NULL argument");
    SE_CHECK(req->psi1 && req->omega1 && req->psi2 && req->omega2, SE_ERR_ARG, "se_server_submit: NULL field");
    SE_CHECK(req->steps >= 1, SE_ERR_ARG, "se_server_submit: steps must be >= 1");
    SE_CHECK(req->N >= 2 && req->N <= 8 && req->nEls[0] > 0 && req->nEls[1] > 0, SE_ERR_ARG,
             "se_server_submit: bad mesh");
    const size_t nV = ((size_t)req->nEls[0] * req->N + 1) * ((size_t)req->nEls[1] * req->N + 1);
    auto r = std::make_shared<Request>();
    r->spec = *req;
    r->init.resize(4 * nV);
    std::memcpy(r->init.data(), req->psi1, sizeof(double) * nV);
    std::memcpy(r->init.data() + nV, req->omega1, sizeof(double) * nV);
    std::memcpy(r->init.data() + 2 * nV, req->psi2, sizeof(double) * nV);
    std::memcpy(r->init.data() + 3 * nV, req->omega2, sizeof(double) * nV);
    if (req->tau_e) r->tau_e.assign(req->tau_e, req->tau_e + nV);
    r->spec.psi1 = r->spec.omega1 = r->spec.psi2 = r->spec.omega2 = r->spec.tau_e = nullptr;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        SE_CHECK(!sv->stopping, SE_ERR_STATE, "se_server_submit: server is stopping");
        r->id = sv->next_id++;
        r->res.id = r->id;
        r->res.t_submit = sv->now();
        sv->queue.push_back(r);
        sv->all[r->id] = r;
        sv->stats.submitted++;
    }
    sv->cv_work.notify_one();
    *id = r->id;
    return SE_OK;
}

extern "C" int se_server_wait(se_server* sv, int64_t id, se_request_result* res, double* psi1, double* omega1,
                              double* psi2, double* omega2) {
    SE_CHECK(sv, SE_ERR_ARG, "se_server_wait: NULL argument");
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->all.find(id);
        SE_CHECK(it != sv->all.end(), SE_ERR_ARG, "se_server_wait: unknown request id");
        r = it->second;
        sv->cv_done.wait(lk, [&] { return r->done; });
        sv->all.erase(id);
    }
    // r is done: the server thread no longer touches it
    if (res) *res = r->res;
    const size_t nV = r->out.size() / 4;
    double* dst[4] = {psi1, omega1, psi2, omega2};
    for (int k = 0; k < 4; k++)
        if (dst[k] && nV) std::memcpy(dst[k], r->out.data() + k * nV, sizeof(double) * nV);
    if (r->res.status != SE_OK) return fail(r->res.status, "se_server_wait: request failed: " + r->error);
    return SE_OK;
}

extern "C" int se_server_get_stats(se_server* sv, se_server_stats* st) {
    SE_CHECK(sv && st, SE_ERR_ARG, "se_server_get_stats: NULL argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    *st = sv->stats;
    st->device_bytes = (double)sv->used;
    return SE_OK;
}

extern "C" int se_server_destroy(se_server* sv) {
    if (!sv) return SE_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stopping = true;
    }
    sv->cv_work.notify_all();
    if (sv->th.joinable()) sv->th.join();
    (void)hipSetDevice(sv->ctx->device);
    for (auto& p : sv->plans)
        if (p.plan) fdm_plan_destroy(p.plan);
    for (auto& kv : sv->batches) se_batch_destroy(kv.second);
    for (auto& p : sv->pool) se_qg2_destroy(p.second);
    for (auto& m : sv->meshes) se_mesh_destroy(m.second);
    se_ctx_destroy(sv->ctx);
    delete sv;
    return SE_OK;
}
