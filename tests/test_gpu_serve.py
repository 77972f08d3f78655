"""Request server: continuous batching of two-layer QG step requests.

No counterpart in the reference (a model driver steps one TwoLayerQGEqn at a
time, src/qg/TwoLayerQGEqn.cpp:174-243).  The bar: a request's result is the
same served alone, in a mix of other requests arriving while it runs, under a
memory budget that serialises admission, and stepped directly with se_qg2_step
(bitwise: the batched solver reproduces the single solves exactly)."""
import time

import numpy as np
import pytest

import se_amd
from se_amd import workloads as W

pytestmark = pytest.mark.gpu


def make_requests(gpu_ctx, sizes, steps):
    base = W.qg2_stream(n_requests=len(sizes))
    out = []
    for r, m_els, k in zip(base, sizes, steps):
        r = dict(r)
        r.update(nEls=(m_els, m_els), steps=k)
        m = se_amd.Mesh(gpu_ctx, r["nEls"], r["N"], r["lo"], r["hi"], r["periodic"], r["bc"])
        X, Y = W.mesh_coords(r["nEls"], r["N"], r["lo"], r["hi"], m.gll()[0])
        r["init"] = W.qg2_initial(r, X, Y)
        r["mesh"] = m
        out.append(r)
    return out


def serve(reqs, order=None, gap=0.0, **kw):
    sv = se_amd.Server(**kw)
    ids = {}
    for j in (order if order is not None else range(len(reqs))):
        ids[j] = sv.submit(reqs[j], *reqs[j]["init"])
        if gap:
            time.sleep(gap)
    out = {j: sv.wait(i) for j, i in ids.items()}
    st = sv.stats()
    sv.close()
    return out, st


PRECISIONS = [se_amd.PREC_FP32_IR, se_amd.PREC_BF16_IR, se_amd.PREC_FP64]


@pytest.mark.parametrize("prec", PRECISIONS, ids=["fp32", "bf16", "fp64"])
def test_served_result_same_alone_and_in_mix(gpu_ctx, prec):
    """FP32 / BF16 inversions run batched, FP64 ones one after another (the
    plain path); in each case the served result is the directly stepped one."""
    reqs = make_requests(gpu_ctx, [8, 10, 12, 9, 11], [3, 6, 2, 4, 5])
    alone, _ = serve(reqs, order=[1], precision=prec)
    mix, st = serve(reqs, order=[0, 2, 1, 3, 4], gap=0.002, precision=prec)
    assert st["completed"] == 5 and st["steps"] == sum(r["steps"] for r in reqs)
    for j in range(5):
        res, f = mix[j]
        assert res["status"] == 0 and res["steps_done"] == reqs[j]["steps"]
        assert res["t_submit"] <= res["t_admit"] <= res["t_done"]
    res_a, fa = alone[1]
    res_m, fm = mix[1]
    for a, b in zip(fa, fm):
        assert np.array_equal(a, b)
    assert res_a["inner_iters"] == res_m["inner_iters"]
    # and the same as stepping the model directly (single mixed-precision solves)
    r = reqs[1]
    q = se_amd.QG2(r["mesh"], r["F1"], r["F2"], r["beta"], se_amd.DETJ_LITERAL)
    q.set_state(*r["init"])
    dt = q.timestep(0.5)
    for _ in range(r["steps"]):
        q.step(dt, rtol=1e-10, precision=prec)
    for a, b in zip(q.state(), fm):
        assert np.array_equal(a, b)
    q.close()


def test_memory_budget_serialises_admission(gpu_ctx):
    reqs = make_requests(gpu_ctx, [9, 9, 9], [2, 3, 2])
    alone = [serve(reqs, order=[j])[0][j] for j in range(3)]
    # one 9x9 request's estimate fits, two do not
    nV = (9 * 6 + 1) ** 2
    n = 9 * 6 * (9 * 6 - 1)
    one = 8 * 28 * nV + n * (2 * (8 * 10 + 4 * 4) + 8 * 5)
    out, st = serve(reqs, memory_budget=1.5 * one)
    t = sorted((out[j][0]["t_admit"], out[j][0]["t_done"]) for j in range(3))
    for (a0, d0), (a1, d1) in zip(t, t[1:]):
        assert a1 >= d0  # the next request is admitted only after the previous one left
    for j in range(3):
        for a, b in zip(alone[j][1], out[j][1]):
            assert np.array_equal(a, b)


def step_directly(r, diss):
    q = se_amd.QG2(r["mesh"], r["F1"], r["F2"], r["beta"], se_amd.DETJ_LITERAL)
    if diss:
        q.set_dissipation(**diss)
    q.set_state(*r["init"])
    dt = q.timestep(0.5)
    for _ in range(r["steps"]):
        q.step(dt, rtol=1e-10, precision=se_amd.PREC_FP32_IR)
    out = q.state()
    q.close()
    return out


@pytest.mark.parametrize("max_active", [0, 1], ids=["mixed", "one-at-a-time"])
def test_dissipative_requests_in_a_mix(gpu_ctx, max_active):
    """Requests with dissipation (four inversions per step, the first pair
    batched beside the other requests' only pair) and without, on the same
    mesh shape: each result is bitwise the direct se_qg2_step one.  One at a
    time, the plain request reuses the pooled instance the dissipative one
    left, which must not keep its dissipation."""
    reqs = make_requests(gpu_ctx, [8, 8, 10], [3, 2, 3])
    xy = np.stack(np.meshgrid(*W.mesh_coords(reqs[0]["nEls"], reqs[0]["N"], reqs[0]["lo"], reqs[0]["hi"],
                                              reqs[0]["mesh"].gll()[0]), indexing="xy"), -1).reshape(-1, 2)
    tau_e = 0.02 * np.sin(np.pi * xy[:, 1]) * np.cos(2 * np.pi * xy[:, 0])
    diss = [dict(E2=0.2, r=0.05, nu=2e-4, tau_e=tau_e), None, dict(E2=0.1, r=0.0, nu=1e-4)]
    for r, d in zip(reqs, diss):
        if d:
            r.update(d)
    out, st = serve(reqs, max_active=max_active)
    assert st["completed"] == 3
    for j, r in enumerate(reqs):
        res, f = out[j]
        assert res["status"] == 0 and res["steps_done"] == r["steps"]
        for a, b in zip(step_directly(r, diss[j]), f):
            assert np.array_equal(a, b)


def test_oversized_request_is_refused(gpu_ctx):
    reqs = make_requests(gpu_ctx, [8], [1])
    sv = se_amd.Server(memory_budget=1e5)
    rid = sv.submit(reqs[0], *reqs[0]["init"])
    with pytest.raises(se_amd.SEError):
        sv.wait(rid)
    sv.close()
