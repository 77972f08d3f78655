"""Pins the oracle's QG steps on an exact nonlinear solution: the two-layer
step (orc_qg2_step, the first-order TwoLayerQGEqn::Solve of
src/qg/TwoLayerQGEqn.cpp:174-243) and the one-layer channel step
(orc_qg_step, QuasiGeostrophicEqn::FirstOrder, src/qg/QuasiGeostrophicEqn.cpp:117-158).

In an x-periodic channel with psi = 0 on the walls,
    psi_i = A a_i sin(l y) cos(k x - w t),   omega_i = -K^2 psi_i,  K^2 = k^2 + l^2,
has J(psi_i, q_i) = 0 for both layers (q_i - beta y is proportional to the
same function), so it solves the full two-layer equations when
    w (-K^2 a_i + F_i (a_j - a_i)) = beta k a_i :
the barotropic mode a = (1, 1), w = -beta k / K^2, and the baroclinic mode
a = (F1, -F2), w = -beta k / (K^2 + F1 + F2).  One step of dt must move the
wave by w dt; the step is first order in time, so its increment matches the
exact increment to O(w dt).  A flipped beta, swapped F1/F2 or a wrong
barotropic beta term moves the wave another way and fails, as the reversed
wave does in the control.  One layer: q = omega - F psi + beta y and
w = -beta k / (K^2 + F)."""
import math

import numpy as np
import pytest

from oracle import oracle as O

K_X, L_Y = 2.0 * math.pi, math.pi
F1, F2, BETA, AMP = 50.0, 150.0, 10.0, 1e-3


def wave_fields(xy, a, w, t):
    s = AMP * np.sin(L_Y * xy[:, 1]) * np.cos(K_X * xy[:, 0] - w * t)
    K2 = K_X * K_X + L_Y * L_Y
    return [a[0] * s, -K2 * a[0] * s, a[1] * s, -K2 * a[1] * s]


@pytest.mark.parametrize("mode", ["barotropic", "baroclinic"])
def test_qg2_step_moves_rossby_wave(mode):
    om = O.OracleMesh((8, 8), 6, (0.0, 0.0), (1.0, 1.0), (1, 0), (1, 1, 0, 0))
    xy = om.verts()
    K2 = K_X * K_X + L_Y * L_Y
    if mode == "barotropic":
        a, w = (1.0, 1.0), -BETA * K_X / K2
    else:
        a, w = (F1 / F2, -1.0), -BETA * K_X / (K2 + F1 + F2)
    dt = 0.02 / abs(w)
    psi1, om1, psi2, om2 = wave_fields(xy, a, w, 0.0)
    st, iters, p1, o1, p2, o2 = om.qg2_step(F1, F2, BETA, dt, psi1, om1, psi2, om2, detj_mode=1, rtol=1e-12)
    assert st == 0 and iters > 0
    got = np.concatenate([p1, p2])
    start = np.concatenate([psi1, psi2])
    e1, _, e2, _ = wave_fields(xy, a, w, dt)
    exact = np.concatenate([e1, e2])
    incr = np.linalg.norm(exact - start)
    assert incr > 0
    assert np.linalg.norm(got - exact) <= 0.05 * incr, np.linalg.norm(got - exact) / incr
    # control: the same wave moving the other way is far from the step
    r1, _, r2, _ = wave_fields(xy, a, -w, dt)
    assert np.linalg.norm(got - np.concatenate([r1, r2])) >= 1.5 * incr
    # the vorticity follows the same wave
    got_o = np.concatenate([o1, o2])
    _, f1, _, f2 = wave_fields(xy, a, w, dt)
    exact_o = np.concatenate([f1, f2])
    start_o = np.concatenate([om1, om2])
    assert np.linalg.norm(got_o - exact_o) <= 0.05 * np.linalg.norm(exact_o - start_o)


def test_qg_step_moves_rossby_wave():
    om = O.OracleMesh((8, 8), 6, (0.0, 0.0), (1.0, 1.0), (1, 0), (1, 1, 0, 0))
    xy = om.verts()
    K2 = K_X * K_X + L_Y * L_Y
    F = 100.0
    w = -BETA * K_X / (K2 + F)
    dt = 0.02 / abs(w)
    psi, omega, _, _ = wave_fields(xy, (1.0, 1.0), w, 0.0)
    st, iters, p, o = om.qg_step(F, BETA, dt, psi, omega, detj_mode=1, rtol=1e-12)
    assert st == 0 and iters > 0
    exact, exact_o, _, _ = wave_fields(xy, (1.0, 1.0), w, dt)
    incr = np.linalg.norm(exact - psi)
    assert np.linalg.norm(p - exact) <= 0.05 * incr, np.linalg.norm(p - exact) / incr
    back, _, _, _ = wave_fields(xy, (1.0, 1.0), -w, dt)
    assert np.linalg.norm(p - back) >= 1.5 * incr
    assert np.linalg.norm(o - exact_o) <= 0.05 * np.linalg.norm(exact_o - omega)
