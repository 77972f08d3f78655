"""Pins the oracle's dissipative QG pieces: SolveQ with DisRHS /
DissipationRHS (orc_qg_solveq: TwoLayerQGEqn.cpp:396-462,
QuasiGeostrophicEqn.cpp:277-337) and the steps that use it.

SolveQ's matrix is the (a x) GLL mass matrix, diagonal, so each unknown is its
assembled right-hand side over its mass: the mass cancels from the pointwise
terms, the two copies of a periodic node are averaged (they add into one
unknown through fieldToVecMap), and the viscous term is the weak Laplacian,
which for a field with a continuous gradient is the assembled stiffness
operator (the oracle's Helmholtz operator, pinned on the reference's own
matrices elsewhere).  The full steps are checked on exact solutions: with
beta = 0 a single Fourier mode psi_i = p_i sin(l y) cos(k x) in every layer
has J(psi_i, q_i) = 0, so the step leaves it in place and the dissipation
scales its amplitudes by factors that follow from the continuous terms
(linear drag r omega, viscosity nu lap omega, relaxation of the interface
tau = psi1 - psi2 toward tau_e, Ekman drag E2 on layer 2).  Topography that is
a multiple of the mode keeps it steady only if it enters the PV and the
inversion with opposite signs."""
import math

import numpy as np
import pytest

from oracle import oracle as O

K_X, L_Y = 2.0 * math.pi, math.pi
K2 = K_X * K_X + L_Y * L_Y


def channel(nEls=(8, 8), N=6):
    return O.OracleMesh(nEls, N, (0.0, 0.0), (1.0, 1.0), (1, 0), (1, 1, 0, 0))


def mode(xy, amp=1e-3):
    return amp * np.sin(L_Y * xy[:, 1]) * np.cos(K_X * xy[:, 0])


def seam_pairs(om):
    """(left, right) field indices of the x-periodic copies."""
    nv0, nv1 = om.nVerts
    return [(j * nv0, j * nv0 + nv0 - 1) for j in range(nv1)]


@pytest.mark.parametrize("a", [1.0, 1.5])
def test_solveq_averages_periodic_copies(a):
    om = channel((4, 3), 4)
    rng = np.random.default_rng(1)
    qhat = rng.standard_normal(om.nV)
    alpha0 = rng.standard_normal(om.nV)
    out = om.qg_solveq(a, 0.1, qhat, alpha0)
    dirichlet = om.maps()[0] < 0
    # Dirichlet nodes keep alpha, except that UpdateField's PeriodicUpdate
    # copies every row's left node onto its right copy, walls included
    kept = alpha0.copy()
    for lft, rgt in seam_pairs(om):
        kept[rgt] = kept[lft]
    assert np.array_equal(out[dirichlet], kept[dirichlet])
    seam = set()
    for lft, rgt in seam_pairs(om):
        if dirichlet[lft]:
            continue
        seam.update((lft, rgt))
        want = 0.5 * (qhat[lft] + qhat[rgt]) / a
        assert abs(out[lft] - want) <= 1e-14 * (1 + abs(want))
        assert out[rgt] == out[lft]
    inner = np.array([v for v in range(om.nV) if not dirichlet[v] and v not in seam])
    assert np.max(np.abs(out[inner] - qhat[inner] / a)) <= 1e-14 * np.max(np.abs(qhat))


def test_solveq_doubly_periodic_corner_averages_four_copies():
    om = O.OracleMesh((3, 3), 3, (0.0, 0.0), (1.0, 1.0), (1, 1), (0, 0, 0, 0))
    rng = np.random.default_rng(2)
    qhat = rng.standard_normal(om.nV)
    out = om.qg_solveq(1.0, 0.1, qhat, np.zeros(om.nV))
    nv0, nv1 = om.nVerts
    corners = [0, nv0 - 1, (nv1 - 1) * nv0, nv1 * nv0 - 1]
    want = np.mean(qhat[corners])
    for c in corners:
        assert abs(out[c] - want) <= 1e-14


def test_solveq_pointwise_terms():
    """alpha = (qhat + dt rF (tau - tau_e) - dt Ei omega_hat) / a on nodes off
    the seam; a wrong sign or a swapped tau / tau_e fails."""
    om = channel((4, 4), 5)
    rng = np.random.default_rng(3)
    qhat, tau, omh, tau_e = (rng.standard_normal(om.nV) for _ in range(4))
    a, dt, rF, Ei = 1.5, 0.3, 0.7, 0.4
    out = om.qg_solveq(a, dt, qhat, np.zeros(om.nV), rF=rF, Ei=Ei, tau=tau, omega_hat=omh, tau_e=tau_e)
    seam = {v for p in seam_pairs(om) for v in p}
    f2v = om.maps()[0]
    inner = np.array([v for v in range(om.nV) if f2v[v] >= 0 and v not in seam])
    want = (qhat + dt * rF * (tau - tau_e) - dt * Ei * omh) / a
    assert np.max(np.abs(out[inner] - want[inner])) <= 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("N", [3, 6])
def test_solveq_viscosity_is_the_stiffness_operator(N):
    """For omega_hat a polynomial of degree <= N vanishing on the walls (its
    gradient is continuous, so the element GlobalToLocal picks does not
    matter), -dt nu sum_e int grad N_i . grad omega_hat is -dt nu (K omega)_i
    with K the assembled stiffness matrix (physical Jacobian)."""
    om = O.OracleMesh((3, 4), N, (0.0, 0.0), (1.5, 1.0), (0, 0), (1, 1, 1, 1))
    xy = om.verts()
    x, y = xy[:, 0], xy[:, 1]
    omh = x * (1.5 - x) * (0.3 + x) * y * (1.0 - y)
    qhat = np.cos(x) + y
    a, dt, nu = 1.5, 0.2, 0.05
    out = om.qg_solveq(a, dt, qhat, np.zeros(om.nV), nu=nu, omega_hat=omh)
    K = O.Helm(om, 0.0, 1.0, 1)
    Mass = O.Helm(om, 1.0, 0.0, 1)
    Kw = K.matvec(om.update_vec(omh))
    winv = Mass.inv_diag()
    got = om.update_vec(out)
    want = (om.update_vec(qhat) - dt * nu * Kw * winv) / a
    visc = dt * nu * np.abs(Kw * winv) / a
    assert np.max(visc) > 1e-3
    # the gradient at a point comes from ShapeFuncDerivsAtCoord's Legendre
    # formulas, the stiffness from the derivative matrix: they agree to rounding
    assert np.max(np.abs(got - want)) <= 1e-9 * np.max(visc)


def one_layer_factor(dt, F, r, nu):
    # alpha = q - dt r omega + dt nu lap omega, omega = -K^2 psi, q = -(K^2+F) psi
    return 1.0 - dt * (r * K2 + nu * K2 * K2) / (K2 + F)


@pytest.mark.parametrize("second", [False, True])
def test_qg_step_damps_a_steady_mode(second):
    om = channel()
    xy = om.verts()
    F, dt, r, nu = 100.0, 0.05, 0.8, 2e-3
    psi = mode(xy)
    omega = -K2 * psi
    prev = (psi, omega) if second else None
    st, _, p, o = om.qg_step(F, 0.0, dt, psi, omega, detj_mode=1, rtol=1e-13, prev=prev, nu=nu, r=r)
    assert st == 0
    fac = one_layer_factor(dt, F, r, nu)
    assert 0.01 < 1.0 - fac < 0.1
    incr = np.linalg.norm((1.0 - fac) * psi)
    assert np.linalg.norm(p - fac * psi) <= 1e-2 * incr
    assert np.linalg.norm(o - fac * omega) <= 1e-2 * incr * K2
    # each term on its own moves the mode by its share
    for rr, nn in ((r, 0.0), (0.0, nu)):
        _, _, p1, _ = om.qg_step(F, 0.0, dt, psi, omega, detj_mode=1, rtol=1e-13, prev=prev, nu=nn, r=rr)
        f1 = one_layer_factor(dt, F, rr, nn)
        assert np.linalg.norm(p1 - f1 * psi) <= 1e-2 * np.linalg.norm((1.0 - f1) * psi)


@pytest.mark.parametrize("second", [False, True])
def test_qg_topography_keeps_a_steady_mode(second):
    om = channel()
    xy = om.verts()
    F, dt = 100.0, 0.05
    psi = mode(xy)
    omega = -K2 * psi
    topo = 40.0 * psi  # a multiple of the mode
    prev = (psi, omega) if second else None
    st, _, p, o = om.qg_step(F, 0.0, dt, psi, omega, detj_mode=1, rtol=1e-13, prev=prev, topo=topo)
    assert st == 0
    assert np.linalg.norm(p - psi) <= 1e-6 * np.linalg.norm(psi)
    assert np.linalg.norm(o - omega) <= 1e-6 * np.linalg.norm(omega)
    # topography entering the inversion with the PV's sign would change the mode
    bad = np.linalg.solve(np.array([[-(K2 + F)]]), np.array([-(K2 + F) + 2 * 40.0]))[0]
    assert abs(bad - 1.0) > 0.1


F1, F2 = 50.0, 150.0
B = np.array([[-K2 - F1, F1], [F2, -K2 - F2]])


def two_layer_expect(p, dt, r, E2, nu, e, second):
    """amplitudes after one step from the steady mode p (and prev = current
    for the second-order step); e = amplitude of tau_e."""
    q = B @ p
    tau = p[0] - p[1]
    s = 1.5 if second else 1.0
    a1 = q[0] + dt * r * F1 * (tau - e / s) + dt * nu * K2 * K2 * p[0]
    if second:
        a2 = q[1] - dt * r * F2 * (tau - e / s) + dt * E2 * K2 * p[1] + dt * nu * K2 * K2 * p[1]
    else:
        a2 = q[1]  # the first-order step does not solve layer 2
    return np.linalg.solve(B, np.array([a1, a2]))


@pytest.mark.parametrize("second", [False, True])
@pytest.mark.parametrize("terms", ["r", "E2", "nu", "tau_e", "all"])
def test_qg2_step_dissipation_on_a_steady_mode(second, terms):
    om = channel()
    xy = om.verts()
    dt = 0.05
    r = 0.5 if terms in ("r", "tau_e", "all") else 0.0
    E2 = 0.6 if terms in ("E2", "all") else 0.0
    nu = 2e-3 if terms in ("nu", "all") else 0.0
    e = 0.7 if terms in ("tau_e", "all") else 0.0
    p = np.array([1.0, -0.4])
    s = mode(xy)
    cur = [p[0] * s, -K2 * p[0] * s, p[1] * s, -K2 * p[1] * s]
    prev = cur if second else None
    tau_e = e * s if e else None
    st, _, p1, o1, p2, o2 = om.qg2_step(F1, F2, 0.0, dt, *cur, detj_mode=1, rtol=1e-12, prev=prev,
                                        E2=E2, r=r, nu=nu, tau_e=tau_e)
    assert st == 0
    want = two_layer_expect(p, dt, r, E2, nu, e, second)
    incr = np.linalg.norm(want - p)
    if terms == "E2" and not second:
        assert incr <= 1e-14 * np.linalg.norm(p)  # Ekman drag acts on layer 2 only
        want_fields = np.concatenate([p[0] * s, p[1] * s])
        assert np.linalg.norm(np.concatenate([p1, p2]) - want_fields) <= 1e-6 * np.linalg.norm(want_fields)
        return
    assert incr > 1e-3 * np.linalg.norm(p)
    got = np.concatenate([p1, p2])
    want_fields = np.concatenate([want[0] * s, want[1] * s])
    scale = incr * np.linalg.norm(s)
    assert np.linalg.norm(got - want_fields) <= 1e-2 * scale
    # omega_i = -K^2 psi_i for the mode
    assert np.linalg.norm(np.concatenate([o1, o2]) + K2 * want_fields) <= 1e-2 * scale * K2
