// tcg_host.h -- host-side control of one mixed-precision solve (shared by the
// single solve and the batched solver): when to run a reliable update, and
// what to do after it.
#pragma once
#include <algorithm>
#include <cmath>

#include "helm_kernels.h"
#include "se_amd.h"

namespace se {

// Graph replays per host check in round `round` of a solve (both the single
// solve and the batched solver count rounds from the solve's start, so their
// cadences, and with them their results, agree): one for the first 16 rounds
// (the batched solver's launch-shape tuning times single replays there), then
// two, then four -- a solve that needs many iterations is in its tail, where
// few members remain and a host check costs more than the iterations it
// would save.
inline int replays_for_round(int round) { return round < 16 ? 1 : (round < 32 ? 2 : 4); }

struct TcgControl {
    // settings
    double delta2 = 0.01;  // reliable update once rr <= delta^2 * rr_ref
    int maxit = 100000, max_outer = 50, window = 256;
    // progress
    double bb = 0.0, tt = 0.0, rr_ref = 0.0, best = 0.0, rr_rec = 0.0;
    int status = CG_RUNNING, updates = 0, since = 0, flat = 0, inner_status = CG_RUNNING;
    bool reset_dir = false;  // restart the direction after this update

    void configure(const se_solve_opts& o) {
        const double delta = o.inner_rtol > 0.0 ? std::min(o.inner_rtol, 0.5) : 0.1;
        delta2 = delta * delta;
        maxit = std::max(o.maxit, 0);
        max_outer = o.max_outer;
    }
    // after the start kernels (state of the first direction kernel)
    void started(const CgState& h) {
        bb = tt = rr_ref = best = h.bb;
        status = h.stop ? h.status : CG_RUNNING;  // 0 (b = 0) or 3 (non-finite)
        updates = since = flat = 0;
    }
    // after k more iterations: is a reliable update due?
    bool due(const CgState& h, int k) {
        since += k;
        return h.stop || h.rr <= delta2 * rr_ref || since >= window;
    }
    // remember why the recursion stopped, before the update runs
    void before_update(const CgState& h) {
        inner_status = h.stop ? h.status : CG_RUNNING;
        rr_rec = h.rr;
    }
    // after the update: final status, or CG_RUNNING with the state to restart
    // from written into `h` (stop cleared, tolerance possibly tightened)
    int after_update(CgState& h) {
        updates++;
        since = 0;
        tt = h.tt;
        if (!(tt - tt == 0.0)) return status = 3;
        if (tt <= h.tol2) return status = 0;
        if (inner_status == 2 || inner_status == 3) return status = inner_status;
        if (h.iter >= maxit || inner_status == 1 || updates >= max_outer) return status = 1;
        // The inner precision cannot resolve this system when the true
        // residual stops improving altogether over several updates.
        if (tt < best) {
            best = tt;
            flat = 0;
        } else if (++flat >= 6) {
            return status = 4;
        }
        // A residual that moved far from the recursion's (more than 2x in
        // norm) is no longer orthogonal to the old direction: restart the
        // direction from it.  Otherwise the recursion continues unchanged.
        reset_dir = tt > 4.0 * rr_rec;
        // the recursion claimed convergence but the true residual disagrees:
        // ask the recursion for more from now on
        if (inner_status == CG_PENDING) h.tol2 *= 0.25;
        rr_ref = tt;
        h.stop = 0;
        h.status = CG_RUNNING;
        return status = CG_RUNNING;
    }
    void fill(se_solve_info& info, int iters, double seconds) const {
        info.status = status == CG_RUNNING ? 1 : status;
        info.iters = iters;
        info.outer_iters = updates;
        info.true_rel_residual = bb > 0.0 ? std::sqrt(std::max(tt, 0.0) / bb) : 0.0;
        info.seconds = seconds;
    }
};

}  // namespace se
