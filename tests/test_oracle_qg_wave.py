"""Pins the oracle's QG steps on exact nonlinear solutions: the one-layer
channel step (orc_qg_step: QuasiGeostrophicEqn::FirstOrder / SecondOrder,
src/qg/QuasiGeostrophicEqn.cpp:83-205) and the two-layer step (orc_qg2_step:
TwoLayerQGEqn::FirstOrder / SecondOrder, src/qg/TwoLayerQGEqn.cpp:174-333).

In an x-periodic channel with psi = 0 on the walls,
    psi_i = A a_i sin(l y) cos(k x - w t),   omega_i = -K^2 psi_i,  K^2 = k^2 + l^2,
has J(psi_i, q_i) = 0 for every layer (q_i - beta y is proportional to the
same function), so it solves the full equations when w B a = beta k a with
B = [[-K^2 - F1, F1], [F2, -K^2 - F2]]: the barotropic mode a = (1, 1),
w = -beta k / K^2, and the baroclinic mode a = (F1, -F2),
w = -beta k / (K^2 + F1 + F2); one layer: w = -beta k / (K^2 + F).  A step of
dt must move the wave by w dt: to O(w dt) relative to the increment for the
first-order step, O((w dt)^2) for the second-order step (started from the
wave at -dt and 0).  The reference's two-layer first-order step advects only
layer 1's PV (SolveQ is called for layer 1 only, TwoLayerQGEqn.cpp:214), so
its result is the inversion of layer 1's moved PV against layer 2's PV at
t = 0, which for one Fourier mode is B^-1 applied to each phase.  A flipped
beta, swapped F1/F2 or a wrong barotropic beta term moves the wave another
way and fails, as the reversed wave does in the control.  Dirichlet (wall)
nodes keep their PV through SolveQ's UpdateField, so the wall vorticity is
the same after every step."""
import math

import numpy as np
import pytest

from oracle import oracle as O

K_X, L_Y = 2.0 * math.pi, math.pi
K2 = K_X * K_X + L_Y * L_Y
F1, F2, BETA, AMP = 50.0, 150.0, 10.0, 1e-3
B = np.array([[-K2 - F1, F1], [F2, -K2 - F2]])


def mesh():
    return O.OracleMesh((8, 8), 6, (0.0, 0.0), (1.0, 1.0), (1, 0), (1, 1, 0, 0))


def shape(xy, phase):
    return AMP * np.sin(L_Y * xy[:, 1]) * np.cos(K_X * xy[:, 0] - phase)


def wave(xy, a, w, t):
    """psi1, omega1, psi2, omega2 of the wave at time t."""
    s = shape(xy, w * t)
    return [a[0] * s, -K2 * a[0] * s, a[1] * s, -K2 * a[1] * s]


def modes():
    return {"barotropic": ((1.0, 1.0), -BETA * K_X / K2),
            "baroclinic": ((F1 / F2, -1.0), -BETA * K_X / (K2 + F1 + F2))}


def check(got, start, exact, reverse, tol):
    incr = np.linalg.norm(exact - start)
    assert incr > 0
    err = np.linalg.norm(got - exact) / incr
    assert err <= tol, err
    assert np.linalg.norm(got - reverse) >= 1.5 * incr


def psis(f):
    return np.concatenate([f[0], f[2]])


def omegas(f):
    return np.concatenate([f[1], f[3]])


@pytest.mark.parametrize("mode", ["barotropic", "baroclinic"])
def test_qg2_second_order_step_moves_rossby_wave(mode):
    om = mesh()
    xy = om.verts()
    a, w = modes()[mode]
    dt = 0.02 / abs(w)
    prev, cur = wave(xy, a, w, -dt), wave(xy, a, w, 0.0)
    st, iters, p1, o1, p2, o2 = om.qg2_step(F1, F2, BETA, dt, *cur, detj_mode=1, rtol=1e-12, prev=prev)
    assert st == 0 and iters > 0
    ex, rv = wave(xy, a, w, dt), wave(xy, a, -w, dt)
    check(np.concatenate([p1, p2]), psis(cur), psis(ex), psis(rv), 0.01)
    check(np.concatenate([o1, o2]), omegas(cur), omegas(ex), omegas(rv), 0.01)


@pytest.mark.parametrize("mode", ["barotropic", "baroclinic"])
def test_qg2_first_order_step_moves_layer_one(mode):
    om = mesh()
    xy = om.verts()
    a, w = modes()[mode]
    dt = 0.02 / abs(w)
    cur = wave(xy, a, w, 0.0)
    st, iters, p1, o1, p2, o2 = om.qg2_step(F1, F2, BETA, dt, *cur, detj_mode=1, rtol=1e-12)
    assert st == 0 and iters > 0
    c = B @ np.array(a)  # PV amplitudes of the two layers (without beta y)
    Binv = np.linalg.inv(B)

    def expect(phase1):
        u = Binv @ np.array([c[0], 0.0])  # layer 1's PV at phase1, layer 2's at phase 0
        v = Binv @ np.array([0.0, c[1]])
        s1, s0 = shape(xy, phase1), shape(xy, 0.0)
        return np.concatenate([u[0] * s1 + v[0] * s0, u[1] * s1 + v[1] * s0])

    check(np.concatenate([p1, p2]), psis(cur), expect(w * dt), expect(-w * dt), 0.05)
    # layer 2's PV is the one it started with
    y = xy[:, 1]
    q2_new = o2 + F2 * (p1 - p2) + BETA * y
    q2_old = cur[3] + F2 * (cur[0] - cur[2]) + BETA * y
    assert np.max(np.abs(q2_new - q2_old)) <= 1e-12 * np.max(np.abs(q2_old))


@pytest.mark.parametrize("order", [1, 2])
def test_qg_step_moves_rossby_wave(order):
    om = mesh()
    xy = om.verts()
    F = 100.0
    w = -BETA * K_X / (K2 + F)
    dt = 0.02 / abs(w)
    psi, omega, _, _ = wave(xy, (1.0, 1.0), w, 0.0)
    prev = wave(xy, (1.0, 1.0), w, -dt)[:2] if order == 2 else None
    st, iters, p, o = om.qg_step(F, BETA, dt, psi, omega, detj_mode=1, rtol=1e-12, prev=prev)
    assert st == 0 and iters > 0
    ex, rv = wave(xy, (1.0, 1.0), w, dt), wave(xy, (1.0, 1.0), -w, dt)
    tol = 0.05 if order == 1 else 0.01
    check(p, psi, ex[0], rv[0], tol)
    check(o, omega, ex[1], rv[1], tol)


def test_second_order_is_second_order():
    """One step's error is O(dt^2) for the first-order step and O(dt^3) for the
    second-order one: halving dt divides them by 4 and by 8."""
    om = mesh()
    xy = om.verts()
    F = 100.0
    w = -BETA * K_X / (K2 + F)
    errs = {1: [], 2: []}
    for dt in (0.08 / abs(w), 0.04 / abs(w)):
        psi, omega, _, _ = wave(xy, (1.0, 1.0), w, 0.0)
        ex = wave(xy, (1.0, 1.0), w, dt)[0]
        for order in (1, 2):
            prev = wave(xy, (1.0, 1.0), w, -dt)[:2] if order == 2 else None
            _, _, p, _ = om.qg_step(F, BETA, dt, psi, omega, detj_mode=1, rtol=1e-13, prev=prev)
            errs[order].append(np.linalg.norm(p - ex))
    r1 = errs[1][0] / errs[1][1]
    r2 = errs[2][0] / errs[2][1]
    assert 3.5 <= r1 <= 4.5, r1
    assert 7.0 <= r2 <= 9.0, r2


def test_wall_vorticity_is_kept():
    """psi = A y (1-y) cos(k x) has omega = -A (2 + k^2 y (1-y)) cos(k x) != 0 on
    the walls; SolveQ leaves the walls' PV, so their omega is unchanged."""
    om = mesh()
    xy = om.verts()
    x, y = xy[:, 0], xy[:, 1]
    psi = 0.05 * y * (1 - y) * np.cos(K_X * x)
    omega = -0.05 * (2.0 + K_X * K_X * y * (1 - y)) * np.cos(K_X * x)
    wall = (y < 1e-12) | (y > 1 - 1e-12)
    assert np.max(np.abs(omega[wall])) > 0.05
    dt = 0.01
    _, _, p, o = om.qg_step(100.0, BETA, dt, psi, omega, detj_mode=1)
    _, _, p2, o2 = om.qg_step(100.0, BETA, dt, p, o, detj_mode=1, prev=(psi, omega))
    for oo in (o, o2):
        assert np.max(np.abs(oo[wall] - omega[wall])) <= 1e-12
        assert np.max(np.abs(oo[~wall] - omega[~wall])) > 1e-6  # the interior moves
    st = [psi, omega, 0.5 * psi, 0.5 * omega]
    _, _, q1, w1, q2, w2 = om.qg2_step(F1, F2, BETA, dt, *st, detj_mode=1)
    _, _, _, v1, _, v2 = om.qg2_step(F1, F2, BETA, dt, q1, w1, q2, w2, detj_mode=1, prev=st)
    for oo, o0 in ((w1, st[1]), (w2, st[3]), (v1, st[1]), (v2, st[3])):
        assert np.max(np.abs(oo[wall] - o0[wall])) <= 1e-12
