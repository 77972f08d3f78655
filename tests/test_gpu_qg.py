"""QG steps on the device against the oracle's restatements: the two-layer
TwoLayerQGEqn::FirstOrder / SecondOrder (src/qg/TwoLayerQGEqn.cpp:174-333) and
the one-layer QuasiGeostrophicEqn::FirstOrder / SecondOrder
(src/qg/QuasiGeostrophicEqn.cpp:83-205).  A model's first step after its
state is set is first order, the later ones second order.

The steps compose pieces pinned elsewhere (GenVel bit-exact, SL advection
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
    ref, prev = list(st), None
    for k in range(3):
        infos = q.step(dt, rtol=1e-10)
        assert all(i["status"] == 0 for i in infos)
        s, it, *nxt = om.qg2_step(req["F1"], req["F2"], req["beta"], dt, *ref, detj_mode=detj, rtol=1e-10,
                                  prev=prev)
        prev, ref = ref, nxt
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


@pytest.mark.parametrize("detj,nEls,nsteps", [(0, 10, 3), (1, 10, 3), (0, 32, 2)],
                         ids=["literal", "physical", "literal-32x32"])
def test_qg_channel_steps_match_oracle(gpu_ctx, detj, nEls, nsteps):
    """The one-layer channel of the 32 x 32, p = 6 configuration (and a 10 x 10
    mesh for both Jacobian modes): beta = 1, F = 100 (Ld = 0.1 L), psi = 0 on
    the walls, Gaussian vortices, CFL 0.5; the wall vorticity is kept."""
    cfg = W.qg_channel(nEls=nEls, N=6)
    m = se_amd.Mesh(gpu_ctx, cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], cfg["periodic"], cfg["bc"])
    om = O.OracleMesh(cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], cfg["periodic"], cfg["bc"])
    X, Y = W.mesh_coords(cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], m.gll()[0])
    rng = W.Rng(cfg["seed"])
    wall = np.repeat(np.sin(np.pi * np.asarray(Y)), len(X))
    psi = 0.01 * W.gaussian_vortices(X, Y, cfg["vortices"], rng, cfg["lo"], cfg["hi"]) * wall
    omega = W.gaussian_vortices(X, Y, cfg["vortices"], rng, cfg["lo"], cfg["hi"])
    q = se_amd.QG(m, cfg["F"], cfg["beta"], detj)
    q.set_state(psi, omega)
    dt = q.timestep(cfg["courant"])
    ref, prev = [psi, omega], None
    for k in range(nsteps):
        info = q.step(dt, rtol=1e-10)
        assert info["status"] == 0
        s, it, *nxt = om.qg_step(cfg["F"], cfg["beta"], dt, *ref, detj_mode=detj, rtol=1e-10, prev=prev)
        prev, ref = ref, nxt
        assert s == 0
        got = q.state()
        errs = [rel(a, b) for a, b in zip(got, ref)]
        print("step", k, "rel errors psi omega", ["%.2e" % e for e in errs])
        assert max(errs) <= 1e-9, errs
    on_wall = (np.abs(om.verts()[:, 1]) < 1e-12) | (np.abs(om.verts()[:, 1] - 1.0) < 1e-12)
    assert np.max(np.abs(q.state()[1][on_wall] - omega[on_wall])) <= 1e-12 * np.max(np.abs(omega))
    q.close()


def smooth(xy, kx, ly, amp):
    return amp * np.sin(ly * np.pi * xy[:, 1]) * np.cos(kx * 2.0 * np.pi * xy[:, 0])


@pytest.mark.parametrize("detj", [0, 1], ids=["literal", "physical"])
def test_qg2_dissipative_steps_match_oracle(gpu_ctx, detj):
    """TwoLayerQGEqn with thermal relaxation toward tau_e, Ekman drag and
    viscosity (DissipationRHS, TwoLayerQGEqn.cpp:419-462): the first inversion
    pair, the dissipative SolveQ of both layers and the final pair."""
    req, m, om, st = setup(gpu_ctx)
    xy = om.verts()
    diss = dict(E2=0.2, r=0.05, nu=2e-4, tau_e=smooth(xy, 1, 1, 0.02))
    q = se_amd.QG2(m, req["F1"], req["F2"], req["beta"], detj)
    q.set_dissipation(**diss)
    q.set_state(*st)
    dt = q.timestep(0.5)
    ref, prev = list(st), None
    for k in range(3):
        infos = q.step(dt, rtol=1e-10)
        assert all(i["status"] == 0 for i in infos)
        s, it, *nxt = om.qg2_step(req["F1"], req["F2"], req["beta"], dt, *ref, detj_mode=detj, rtol=1e-10,
                                  prev=prev, **diss)
        assert s == 0
        _, _, *plain = om.qg2_step(req["F1"], req["F2"], req["beta"], dt, *ref, detj_mode=detj, rtol=1e-10,
                                   prev=prev)
        assert max(rel(a, b) for a, b in zip(nxt, plain)) > 1e-6  # the dissipation moves the state
        prev, ref = ref, nxt
        got = q.state()
        errs = [rel(a, b) for a, b in zip(got, ref)]
        print("step", k, "rel errors psi1 om1 psi2 om2", ["%.2e" % e for e in errs])
        assert max(errs) <= 1e-9, errs
    q.close()


@pytest.mark.parametrize("detj,nEls,nsteps", [(0, 10, 3), (1, 10, 3), (0, 32, 2)],
                         ids=["literal", "physical", "literal-32x32"])
def test_qg_dissipative_steps_match_oracle(gpu_ctx, detj, nEls, nsteps):
    """QuasiGeostrophicEqn with viscosity, linear drag and topography (DisRHS,
    QuasiGeostrophicEqn.cpp:303-337): he1 with zero lifting, the dissipative
    SolveQ, he2; also at the channel's full 32 x 32, p = 6 size."""
    cfg = W.qg_channel(nEls=nEls, N=6)
    m = se_amd.Mesh(gpu_ctx, cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], cfg["periodic"], cfg["bc"])
    om = O.OracleMesh(cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], cfg["periodic"], cfg["bc"])
    X, Y = W.mesh_coords(cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], m.gll()[0])
    rng = W.Rng(cfg["seed"])
    wall = np.repeat(np.sin(np.pi * np.asarray(Y)), len(X))
    psi = 0.01 * W.gaussian_vortices(X, Y, cfg["vortices"], rng, cfg["lo"], cfg["hi"]) * wall
    omega = W.gaussian_vortices(X, Y, cfg["vortices"], rng, cfg["lo"], cfg["hi"])
    xy = om.verts()
    topo = smooth(xy, 2, 2, 5.0)
    diss = dict(nu=1e-4, r=0.3)
    q = se_amd.QG(m, cfg["F"], cfg["beta"], detj)
    q.set_dissipation(**diss)
    q.set_topography(topo)
    q.set_state(psi, omega)
    dt = q.timestep(cfg["courant"])
    ref, prev = [psi, omega], None
    for k in range(nsteps):
        info = q.step(dt, rtol=1e-10)
        assert info["status"] == 0
        s, it, *nxt = om.qg_step(cfg["F"], cfg["beta"], dt, *ref, detj_mode=detj, rtol=1e-10, prev=prev,
                                 topo=topo, **diss)
        assert s == 0
        _, _, *plain = om.qg_step(cfg["F"], cfg["beta"], dt, *ref, detj_mode=detj, rtol=1e-10, prev=prev,
                                  topo=topo)
        assert max(rel(a, b) for a, b in zip(nxt, plain)) > 1e-6
        prev, ref = ref, nxt
        got = q.state()
        errs = [rel(a, b) for a, b in zip(got, ref)]
        print("step", k, "rel errors psi omega", ["%.2e" % e for e in errs])
        assert max(errs) <= 1e-9, errs
    q.close()


def test_set_state_right_after_step_is_ordered(gpu_ctx):
    """set_state (and set_dissipation / set_topography) called right after a
    step replaces the state that step leaves behind, even though the step's
    last kernels may still be queued: the result equals a fresh model's."""
    req, m, _, st = setup(gpu_ctx, m_els=12, seed=11)
    a = se_amd.QG2(m, req["F1"], req["F2"], req["beta"], 0)
    a.set_state(*st)
    dt = a.timestep(0.5)
    a.step(dt, rtol=1e-10)
    a.set_state(*st)
    a.step(dt, rtol=1e-10)
    b = se_amd.QG2(m, req["F1"], req["F2"], req["beta"], 0)
    b.set_state(*st)
    b.step(dt, rtol=1e-10)
    assert all(np.array_equal(x, y) for x, y in zip(a.state(), b.state()))
    a.close()
    b.close()

    cfg = W.qg_channel(nEls=10, N=6)
    mc = se_amd.Mesh(gpu_ctx, cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], cfg["periodic"], cfg["bc"])
    X, Y = W.mesh_coords(cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], mc.gll()[0])
    q0 = W.gaussian_vortices(X, Y, cfg["vortices"], W.Rng(cfg["seed"]), cfg["lo"], cfg["hi"])
    c = se_amd.QG(mc, cfg["F"], cfg["beta"], 0)
    c.set_state(np.zeros(mc.nV), q0)
    dtc = 1e-3
    c.step(dtc, rtol=1e-10)
    c.set_topography(np.zeros(mc.nV))
    c.set_state(np.zeros(mc.nV), q0)
    c.step(dtc, rtol=1e-10)
    d = se_amd.QG(mc, cfg["F"], cfg["beta"], 0)
    d.set_topography(np.zeros(mc.nV))
    d.set_state(np.zeros(mc.nV), q0)
    d.step(dtc, rtol=1e-10)
    assert all(np.array_equal(x, y) for x, y in zip(c.state(), d.state()))
    c.close()
    d.close()
