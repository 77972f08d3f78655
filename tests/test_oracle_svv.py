"""Pins the oracle's spectral vanishing viscosity operator (orc_elem_svv:
SVV::AssembleElement, src/base/SVV.cpp:120-176), restated literally in 2-D
(modal basis, Gauss-Jordan inverses, dense products).

Independent check: the 2-D construction is separable.  With Phi the 1-D
matrix of the hierarchical modes at the GLL points (phi_0 = (1+x)/2,
phi_1 = (1-x)/2, phi_k = (1-x^2) P^(1,1)_{k-2}(x)/4, here from scipy's Jacobi
polynomials), q_k = sqrt(exp(-(N-k)^2/(mN-k)^2)) for k > mN (else 0),
F = Phi diag(q) Phi^-1, W = diag(w), D the GLL derivative matrix,
    element SVV = 2 c detJ (F^T W F) (x) ((F D)^T W (F D))   (y (x) x),
the factor 2 because the reference builds Lyy from the x terms (SVV.cpp:165).
Consequences checked as well: the operator is symmetric positive
semi-definite, vanishes on fields of x-degree <= mN + 1 (their x-derivative
has no mode above the cutoff) and on every field of y alone (the quirk)."""
import numpy as np
import pytest
from scipy.special import eval_jacobi

from oracle import oracle as O


def mesh(N, nEls=(3, 2), lo=(0.0, -1.0), hi=(1.5, 1.0)):
    return O.OracleMesh(nEls, N, lo, hi, (0, 0), (0, 0, 0, 0))


def modes(N, x):
    Phi = np.zeros((N + 1, N + 1))
    Phi[:, 0] = 0.5 * (1 + x)
    Phi[:, 1] = 0.5 * (1 - x)
    for k in range(2, N + 1):
        Phi[:, k] = 0.25 * (1 - x * x) * eval_jacobi(k - 2, 1, 1, x)
    return Phi


def separable(N, mN, x, w, D, c, detj):
    Phi = modes(N, x)
    q = np.array([np.sqrt(np.exp(-((N - k) ** 2) / ((mN - k) ** 2))) if k > mN else 0.0 for k in range(N + 1)])
    F = Phi @ np.diag(q) @ np.linalg.inv(Phi)
    Mf = F.T @ np.diag(w) @ F
    G = F @ D
    Kf = G.T @ np.diag(w) @ G
    return 2.0 * c * detj * np.kron(Mf, Kf)


@pytest.mark.parametrize("N,mN", [(4, 1), (6, 2), (6, 3), (8, 4)])
def test_element_svv_is_the_separable_filter(N, mN):
    om = mesh(N)
    x, w, D = om.gll()
    D = np.asarray(D).reshape(N + 1, N + 1)
    hx, hy = 1.5 / 3, 2.0 / 2
    c = 0.37
    A = om.elem_svv(1, c, mN)
    want = separable(N, mN, np.asarray(x), np.asarray(w), D, c, hx * hy / 4)
    assert np.max(np.abs(A - want)) <= 1e-10 * np.max(np.abs(want))
    assert np.max(np.abs(want)) > 0
    # Lxx + a true Lyy (the y-derivative terms) would be a different operator
    sym = 0.5 * want + 0.5 * want.reshape(N + 1, N + 1, N + 1, N + 1).transpose(1, 0, 3, 2).reshape(A.shape)
    assert np.max(np.abs(A - sym)) > 1e-3 * np.max(np.abs(want))


@pytest.mark.parametrize("N,mN", [(5, 2), (8, 3)])
def test_element_svv_null_space_and_definiteness(N, mN):
    om = mesh(N)
    x = np.asarray(om.gll()[0])
    A = om.elem_svv(0, 1.0, mN)
    assert np.max(np.abs(A - A.T)) <= 1e-12 * np.max(np.abs(A))
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert ev.min() >= -1e-10 * ev.max()
    X, Y = np.meshgrid(x, x)  # element nodes, x fastest
    X, Y = X.ravel(), Y.ravel()
    for k in range(mN + 2):  # x-degree <= mN + 1: killed
        u = X ** k * (1 + Y + Y ** N)
        assert np.max(np.abs(A @ u)) <= 1e-10 * np.max(np.abs(A)) * np.max(np.abs(u))
    for u in (Y ** N, np.cos(3 * Y)):  # functions of y alone: killed (Lyy from the x terms)
        assert np.max(np.abs(A @ u)) <= 1e-10 * np.max(np.abs(A)) * np.max(np.abs(u))
    u = X ** N * Y ** N  # the top mode in both directions is not
    assert np.max(np.abs(A @ u)) > 1e-3 * np.max(np.abs(A))


def test_assembled_svv_adds_the_element_blocks():
    """The assembled Helmholtz operator with SVV is the one without plus the
    assembled SVV blocks (no Dirichlet rows in this mesh)."""
    N, mN, dt = 4, 2, 0.05
    om = mesh(N)
    base = O.Helm(om, 1.0, 0.01, 0)
    full = O.Helm(om, 1.0, 0.01, 0, svv=(dt, mN))
    diff = (full.scipy() - base.scipy()).toarray()
    f2v = om.maps()[0]
    want = np.zeros_like(diff)
    for el in range(om.nElsTotal):
        nodes = om.el_nodes(el)
        A = om.elem_svv(el, dt, mN)
        for a, na in enumerate(nodes):
            for b, nb in enumerate(nodes):
                want[f2v[na], f2v[nb]] += A[a, b]
    assert np.max(np.abs(diff - want)) <= 1e-12 * np.max(np.abs(want))
