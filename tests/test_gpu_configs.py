"""The benchmark configurations at their own sizes on the GPU.

* The 512-member Helmholtz ensemble (nEls ~ U{8,16,32,64}^2, p ~ U{4,6,8},
  lambda log-uniform in [1, 1e3]) solved as one batch, with FP64, FP64 then
  FP32 (corrections), and FP32 preconditioner transforms: every member meets the
  1e-10 true-residual target, measured with the exact FP64 operator apply
  (pinned at 1e-12 against the oracle's assembled CSR, a restatement of
  Matrix::Assemble, src/base/Matrix.cpp:47-133), and the small members also
  directly against that CSR.
* Two-layer QG steps (TwoLayerQGEqn::FirstOrder / SecondOrder,
  src/qg/TwoLayerQGEqn.cpp:174-333) at the served stream's smallest and
  largest meshes, 16^2 and 96^2 elements at p = 6, against the oracle's step,
  and served requests of those sizes equal to direct stepping, bit for bit.
"""
import numpy as np
import pytest

import se_amd
from oracle import oracle as O
from se_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("prec", [se_amd.PREC_FP64, se_amd.PREC_FP64_FP32C, se_amd.PREC_FP32_IR],
                         ids=["fp64", "fp64_fp32c", "fp32_ir"])
def test_ensemble_512_full_batch(gpu_ctx, prec):
    cfgs = W.ensemble_512(count=512, detj_mode=1)
    ms, hs, fs = [], [], []
    for c in cfgs:
        m = se_amd.Mesh(gpu_ctx, c["nEls"], c["N"], c["lo"], c["hi"], c["periodic"], c["bc"])
        h = se_amd.Helm(m, c["c_mass"], c["c_lap"], c["detj_mode"])
        fs.append(h.rhs(W.make_rhs(c, m.gll()[0])))
        ms.append(m)
        hs.append(h)
    assert sum(m.n for m in ms) == 26821632
    b = se_amd.Batch(gpu_ctx, hs)
    xs, infos = b.pcg(fs, precision=prec, rtol=1e-10, preconditioner=se_amd.PC_FDM)
    worst = 0.0
    for h, f, x, inf in zip(hs, fs, xs, infos):
        assert inf["status"] == se_amd.SOLVE_CONVERGED, inf
        assert inf["true_rel_residual"] <= 1e-10, inf
        worst = max(worst, rel(h.apply(x), f))
    print("512 members: max ||f - A x|| / ||f|| = %.2e (exact FP64 apply); iterations %d..%d"
          % (worst, min(i["iters"] for i in infos), max(i["iters"] for i in infos)))
    assert worst <= 1.5e-10
    # the 8 x 8 members (and one 16 x 16) against the oracle's assembled CSR
    small = [i for i, c in enumerate(cfgs) if c["nEls"][0] == 8][:12]
    small += [i for i, c in enumerate(cfgs) if c["nEls"][0] == 16][:1]
    for i in small:
        c = cfgs[i]
        om = O.OracleMesh(c["nEls"], c["N"], c["lo"], c["hi"], c["periodic"], c["bc"])
        oh = O.Helm(om, c["c_mass"], c["c_lap"], c["detj_mode"])
        assert rel(oh.matvec(xs[i]), fs[i]) <= 1.5e-10, (i, c["nEls"], c["N"])
    b.close()
    for h, m in zip(hs, ms):
        h.close()
        m.close()


def c5_request(gpu_ctx, m_els, seed=7):
    req = W.qg2_stream(n_requests=1)[0]
    req.update(nEls=(m_els, m_els), N=6, seed=seed)
    m = se_amd.Mesh(gpu_ctx, req["nEls"], 6, req["lo"], req["hi"], req["periodic"], req["bc"])
    X, Y = W.mesh_coords(req["nEls"], 6, req["lo"], req["hi"], m.gll()[0])
    return req, m, W.qg2_initial(req, X, Y)


@pytest.mark.parametrize("m_els,nsteps", [(16, 2), (96, 1)], ids=["16x16", "96x96"])
def test_qg2_at_stream_sizes_matches_oracle(gpu_ctx, m_els, nsteps):
    """FP64 steps (the plain path) against the oracle's, first order then second."""
    req, m, st = c5_request(gpu_ctx, m_els)
    om = O.OracleMesh(req["nEls"], 6, req["lo"], req["hi"], req["periodic"], req["bc"])
    q = se_amd.QG2(m, req["F1"], req["F2"], req["beta"], se_amd.DETJ_LITERAL)
    q.set_state(*st)
    dt = q.timestep(0.5)
    ref, prev = list(st), None
    for k in range(nsteps):
        infos = q.step(dt, rtol=1e-10)
        assert all(i["status"] == 0 for i in infos), infos
        s, _, *nxt = om.qg2_step(req["F1"], req["F2"], req["beta"], dt, *ref, detj_mode=0, rtol=1e-10, prev=prev)
        assert s == 0
        prev, ref = ref, nxt
        errs = [rel(a, b) for a, b in zip(q.state(), ref)]
        print("%dx%d step %d: rel errors psi1 om1 psi2 om2 %s" % (m_els, m_els, k, ["%.2e" % e for e in errs]))
        assert max(errs) <= 1e-9, errs
    q.close()


def test_served_stream_sizes_equal_direct_steps(gpu_ctx):
    """Requests of 16^2 and 96^2 elements (and one between) served together with
    mixed-precision inversions give what stepping each model directly gives."""
    sizes, steps = [16, 96, 40], [3, 2, 2]
    reqs = []
    for j, (m_els, k) in enumerate(zip(sizes, steps)):
        req, m, st = c5_request(gpu_ctx, m_els, seed=11 + j)
        req.update(steps=k)
        reqs.append((req, m, st))
    sv = se_amd.Server(precision=se_amd.PREC_FP32_IR)
    ids = [sv.submit(r, *st) for r, _, st in reqs]
    out = [sv.wait(i) for i in ids]
    sv.close()
    for (req, m, st), (res, fields) in zip(reqs, out):
        assert res["status"] == 0 and res["steps_done"] == req["steps"]
        q = se_amd.QG2(m, req["F1"], req["F2"], req["beta"], se_amd.DETJ_LITERAL)
        q.set_state(*st)
        dt = q.timestep(0.5)
        for _ in range(req["steps"]):
            infos = q.step(dt, rtol=1e-10, precision=se_amd.PREC_FP32_IR)
            assert all(i["status"] == 0 and i["true_rel_residual"] <= 1e-10 for i in infos)
        for a, b in zip(q.state(), fields):
            assert np.array_equal(a, b), req["nEls"]
        q.close()
