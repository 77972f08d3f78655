"""Two-layer QG step on the device against the oracle's restatement of
TwoLayerQGEqn::FirstOrder (src/qg/TwoLayerQGEqn.cpp:174-243).

The step composes pieces pinned elsewhere (GenVel bit-exact, SL advection
within its conditioning bound, inversions to 1e-10), so the fields agree to
the solver tolerance amplified by the inversion's conditioning."""
import numpy as np
import pytest

import se_amd
from oracle import oracle as O
from se_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def setup(gpu_ctx, m_els=8, N=6, seed=7):
    req = W.qg2_stream(n_requests=1)[0]
    req.update(nEls=(m_els, m_els), N=N, seed=seed)
    m = se_amd.Mesh(gpu_ctx, req["nEls"], N, req["lo"], req["hi"], req["periodic"], req["bc"])
    X, Y = W.mesh_coords(req["nEls"], N, req["lo"], req["hi"], m.gll()[0])
    st = W.qg2_initial(req, X, Y)
    om = O.OracleMesh(req["nEls"], N, req["lo"], req["hi"], req["periodic"], req["bc"])
    return req, m, om, st


@pytest.mark.parametrize("detj", [0, 1], ids=["literal", "physical"])
def test_qg2_steps_match_oracle(gpu_ctx, detj):
    req, m, om, st = setup(gpu_ctx)
    q = se_amd.QG2(m, req["F1"], req["F2"], req["beta"], detj)
    q.set_state(*st)
    dt = q.timestep(0.5)
    ref = list(st)
    for k in range(3):
        infos = q.step(dt, rtol=1e-10)
        assert all(i["status"] == 0 for i in infos)
        s, it, *ref = om.qg2_step(req["F1"], req["F2"], req["beta"], dt, *ref, detj_mode=detj, rtol=1e-10)
        assert s == 0
        got = q.state()
        errs = [rel(a, b) for a, b in zip(got, ref)]
        print("step", k, "rel errors psi1 om1 psi2 om2", ["%.2e" % e for e in errs])
        assert max(errs) <= 1e-9, errs
    q.close()


def test_qg2_mixed_precision_step(gpu_ctx):
    req, m, _, st = setup(gpu_ctx, m_els=12, seed=9)
    q64 = se_amd.QG2(m, req["F1"], req["F2"], req["beta"], 0)
    q32 = se_amd.QG2(m, req["F1"], req["F2"], req["beta"], 0)
    q64.set_state(*st)
    q32.set_state(*st)
    dt = q64.timestep(0.5)
    for k in range(2):
        q64.step(dt, rtol=1e-10)
        infos = q32.step(dt, rtol=1e-10, precision=se_amd.PREC_FP32_IR)
        assert all(i["status"] == 0 and i["true_rel_residual"] <= 1e-10 for i in infos)
    errs = [rel(a, b) for a, b in zip(q32.state(), q64.state())]
    assert max(errs) <= 1e-8, errs
