"""CPU oracle pins from mathematics: closed forms, invariants, direct solves and
manufactured solutions.  These cover what the golden vectors of the compiled
reference cannot (the PETSc-side assembly, lifting and the CG solve), and each
is chosen so that a dropped term, a wrong sign or a transposed operand fails."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from numpy.polynomial import legendre as npleg

from oracle import oracle as O

TWO_PI = 2.0 * math.pi


# ---- a1: GLL basis ------------------------------------------------------------
def test_gll_closed_form_n4():
    a, w, D, _ = O.gll(4)
    s = math.sqrt(3.0 / 7.0)
    assert np.allclose(a, [-1, -s, 0, s, 1], atol=1e-15, rtol=0)
    assert np.allclose(w, [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], atol=1e-15, rtol=0)


@pytest.mark.parametrize("N", range(2, 13))
def test_gll_roots_weights_derivative(N):
    a, w, D, _ = O.gll(N)
    inner = np.sort(npleg.Legendre.basis(N).deriv().roots().real)
    assert np.max(np.abs(a[1:-1] - inner)) < 1e-13
    assert abs(w.sum() - 2.0) < 1e-12  # monomial-form P_N evaluation loses a few digits at high N
    # D is exact on polynomials of degree <= N, and D * 1 = 0
    for deg in range(N + 1):
        assert np.max(np.abs(D @ a ** deg - (deg * a ** (deg - 1) if deg else 0 * a))) < 1e-11 * max(1, N * N)
    # quadrature exact to degree 2N-1
    for deg in range(2 * N):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(w @ a ** deg - exact) < 1e-12


# ---- a2/a3: numbering -------------------------------------------------------------
def test_maps_doubly_periodic_closed_form():
    m = O.OracleMesh((8, 8), 4, (0, 0), (TWO_PI, TWO_PI), (1, 1))
    f2v, v2f = m.maps()
    nv0, nv1 = m.nVerts
    J, I = np.divmod(np.arange(m.nV), nv0)
    assert np.array_equal(f2v, (J % (nv1 - 1)) * (nv0 - 1) + (I % (nv0 - 1)))
    assert m.n == 1024
    assert np.array_equal(f2v[v2f], np.arange(m.n))


def test_maps_channel_closed_form():
    m = O.OracleMesh((32, 32), 6, (0, 0), (1, 1), (1, 0), (1, 1, 0, 0))
    f2v, v2f = m.maps()
    nv0, nv1 = m.nVerts
    J, I = np.divmod(np.arange(m.nV), nv0)
    exp = np.where((J == 0) | (J == nv1 - 1), -1, (J - 1) * (nv0 - 1) + (I % (nv0 - 1)))
    assert np.array_equal(f2v, exp)
    assert m.n == 36672 and m.nV == 37249


def test_maps_full_dirichlet_and_every_unknown_hit():
    m = O.OracleMesh((4, 3), 5, (0, 0), (1, 1), (0, 0), (1, 1, 1, 1))
    f2v, v2f = m.maps()
    hit = np.zeros(m.n, dtype=int)
    for e in range(m.nElsTotal):
        g = f2v[m.el_nodes(e)]
        hit[g[g >= 0]] += 1
    assert hit.min() >= 1
    assert m.n == (m.nVerts[0] - 2) * (m.nVerts[1] - 2)


def test_unsupported_bc_combinations():
    with pytest.raises(ValueError):
        O.OracleMesh((4, 4), 4, (0, 0), (1, 1), (0, 1))       # y-periodic only
    with pytest.raises(ValueError):
        O.OracleMesh((4, 4), 4, (0, 0), (1, 1), (1, 0), (0, 0, 1, 0))  # periodic + wall in x


# ---- a4-a6: element operators --------------------------------------------------------
def tensor_laplacian(m, c, detj_mode, el):
    """c[(2/dx)^2 W (x) G + (2/dy)^2 G (x) W] with G = D^T W D (node iy*(N+1)+ix)."""
    a, w, D = m.gll()
    G = D.T @ np.diag(w) @ D
    dx = (m.hi[0] - m.lo[0]) / m.nEls[0]
    dy = (m.hi[1] - m.lo[1]) / m.nEls[1]
    K = (2 / dx) ** 2 * np.kron(np.diag(w), G) + (2 / dy) ** 2 * np.kron(G, np.diag(w))
    if detj_mode:
        K *= dx * dy / 4
    return c * K


@pytest.mark.parametrize("mode", [0, 1])
def test_element_laplacian_tensor_form(mode):
    m = O.OracleMesh((3, 2), 6, (-1.0, 0.5), (2.0, 1.7), (0, 0), (0, 0, 0, 0))
    for e in range(m.nElsTotal):
        K = m.elem_laplacian(e, 1.7, mode)
        T = tensor_laplacian(m, 1.7, mode, e)
        assert np.max(np.abs(K - T)) < 1e-12 * np.max(np.abs(T))
        assert np.max(np.abs(K - K.T)) < 1e-12 * np.max(np.abs(K))
        assert np.max(np.abs(K.sum(axis=1))) < 1e-11 * np.max(np.abs(K))


def test_mass_is_gll_diagonal():
    m = O.OracleMesh((3, 2), 5, (0, 0), (3.0, 1.0), (0, 0))
    a, w, D = m.gll()
    dx, dy = 1.0, 0.5
    for e in range(m.nElsTotal):
        M = m.elem_mass(e, 2.0)
        # detJ from vertex coordinates carries ~1e-14 round-off (Mesh.cpp:183-198)
        assert np.allclose(np.diag(M), 2.0 * np.kron(w, w) * dx * dy / 4, rtol=1e-12, atol=0)


# ---- a8: assembly and lifting -------------------------------------------------------
def dense_assembly(m, h_cm, h_cl, mode):
    """Full (field-node) matrix summed in element order with the element blocks."""
    A = np.zeros((m.nV, m.nV))
    for e in range(m.nElsTotal):
        n = m.el_nodes(e)
        Ae = m.elem_laplacian(e, h_cl, mode)
        if h_cm != 0.0:
            Ae = m.elem_mass(e, h_cm) + Ae  # mass is applied first by the reference
        A[np.ix_(n, n)] += Ae
    return A


@pytest.mark.parametrize("args", [
    ((4, 3), 4, (0, 0), (1, 1), (0, 0), (1, 1, 1, 1), 2.0, 1.0, 0),
    ((5, 4), 3, (0, 0), (2, 1), (1, 0), (1, 1, 0, 0), -5.0, -1.0, 1),
    ((4, 4), 5, (0, 0), (TWO_PI, TWO_PI), (1, 1), (0, 0, 0, 0), 0.0, 1.0, 0),
])
def test_csr_and_lifting_against_dense(args):
    nEls, N, lo, hi, per, bc, cm, cl, mode = args
    m = O.OracleMesh(nEls, N, lo, hi, per, bc)
    f2v, v2f = m.maps()
    V = m.verts()
    ud = np.sin(3 * V[:, 0]) + V[:, 1] ** 2   # Dirichlet data on every node
    h = O.Helm(m, cm, cl, mode, sol_field=ud)
    A = h.scipy().toarray()
    Af = dense_assembly(m, cm, cl, mode)
    # restriction with periodic aliasing: R^T Af R where R maps unknowns to nodes
    R = np.zeros((m.nV, m.n))
    for v in range(m.nV):
        if f2v[v] >= 0:
            R[v, f2v[v]] = 1.0
    assert np.max(np.abs(A - R.T @ Af @ R)) < 1e-12 * np.max(np.abs(A))
    dmask = f2v < 0
    b_exp = -(R.T @ Af[:, dmask] @ ud[dmask])
    b = h.lift()
    assert np.max(np.abs(b - b_exp)) <= 1e-12 * max(1.0, np.max(np.abs(b_exp)))
    assert np.allclose(h.inv_diag(), 1.0 / np.diag(A), rtol=1e-15, atol=0)
    x = np.random.default_rng(0).standard_normal(m.n)
    assert np.max(np.abs(h.matvec(x) - A @ x)) < 1e-12 * np.max(np.abs(A @ x))


# ---- a10: Jacobi-PCG --------------------------------------------------------------------
def test_pcg_matches_direct_solve_helmholtz_channel():
    m = O.OracleMesh((8, 8), 6, (0, 0), (1, 1), (1, 0), (1, 1, 0, 0))
    h = O.Helm(m, -100.0, -1.0, 1)
    f = np.random.default_rng(1).standard_normal(m.n)
    x, info = h.pcg(f, rtol=1e-13, maxit=5000)
    assert info["status"] == 0 and info["relres"] <= 1e-13
    xd = spla.spsolve(h.scipy().tocsc(), f)
    assert np.max(np.abs(x - xd)) < 1e-9 * np.max(np.abs(xd))


def test_pcg_null_space_periodic_poisson():
    m = O.OracleMesh((6, 6), 4, (0, 0), (TWO_PI, TWO_PI), (1, 1))
    h = O.Helm(m, 0.0, -1.0, 1)
    f = np.random.default_rng(2).standard_normal(m.n)
    x, info = h.pcg(f, rtol=1e-12, null_space=1, maxit=5000)
    assert info["status"] == 0
    assert abs(x.mean()) < 1e-12 * np.max(np.abs(x))
    A = h.scipy().toarray()
    fp = f - f.mean()
    assert np.linalg.norm(fp - A @ x) <= 1.01e-12 * np.linalg.norm(fp)


def test_pcg_status_maxit_and_indefinite():
    m = O.OracleMesh((4, 4), 4, (0, 0), (1, 1), (0, 0), (1, 1, 1, 1))
    h = O.Helm(m, 1.0, 1.0, 1)
    f = np.ones(m.n)
    _, info = h.pcg(f, rtol=1e-14, maxit=2)
    assert info["status"] == 1 and info["iters"] == 2
    # a mass shift between the extreme diagonal ratios makes the diagonal change
    # sign: no sign normalisation makes that system definite
    kd = 1.0 / O.Helm(m, 0.0, 1.0, 1).inv_diag()
    md = 1.0 / O.Helm(m, 1.0, 0.0, 1).inv_diag()
    c = -float(np.median(kd / md))
    hi = O.Helm(m, c, 1.0, 1)
    d = 1.0 / hi.inv_diag()
    assert d.min() < 0 < d.max()
    _, info = hi.pcg(f, rtol=1e-10, maxit=50)
    assert info["status"] == 2


@pytest.mark.parametrize("N,tol", [(4, 2e-3), (6, 2e-5), (8, 2e-7)])
def test_manufactured_periodic_helmholtz_physical(N, tol):
    """(lambda M + K) psi = M f with psi = cos(2x + 3y): f = (lambda + 13) psi."""
    lam = 5.0
    m = O.OracleMesh((6, 6), N, (0, 0), (TWO_PI, TWO_PI), (1, 1))
    V = m.verts()
    psi = np.cos(2 * V[:, 0] + 3 * V[:, 1])
    h = O.Helm(m, lam, 1.0, 1)
    sol, info = h.solve((lam + 13.0) * psi, np.zeros(m.nV), rtol=1e-12)
    assert info["status"] == 0
    assert np.max(np.abs(sol - psi)) < tol


def test_reference_poisson_case_physical_and_literal():
    """models/poisson-equation/PoissonEquation.cpp:27-28,51-90 with PoissonEqn(c=-1):
    psi = cos(pi x) cos(pi y) on [-1,1]^2 with Dirichlet walls.  The physical
    operator reproduces it spectrally; the literal one (detJ := 1) does not."""
    m = O.OracleMesh((4, 4), 8, (-1, -1), (1, 1), (0, 0), (1, 1, 1, 1))
    V = m.verts()
    ana = np.cos(np.pi * V[:, 0]) * np.cos(np.pi * V[:, 1])
    omega = -2 * np.pi ** 2 * ana
    err = {}
    for mode in (0, 1):
        sol0 = np.where(m.maps()[0] < 0, ana, 0.0)
        h = O.Helm(m, 0.0, -1.0, mode, sol_field=sol0)
        sol, info = h.solve(omega, sol0, rtol=1e-12)
        assert info["status"] == 0
        err[mode] = np.max(np.abs(sol - ana))
    assert err[1] < 1e-7
    assert err[0] > 1e-2


def test_literal_equals_scaled_physical_on_periodic_poisson():
    """Literal K~ = K/detJ on a uniform mesh: psi_literal = detJ psi_physical."""
    m = O.OracleMesh((6, 6), 5, (0, 0), (TWO_PI, TWO_PI), (1, 1))
    V = m.verts()
    f = np.sin(V[:, 0]) * np.cos(2 * V[:, 1]) + 0.3 * np.cos(3 * V[:, 0])
    detJ = (TWO_PI / 6) ** 2 / 4
    sols = []
    for mode in (0, 1):
        h = O.Helm(m, 0.0, 1.0, mode)
        s, info = h.solve(f, np.zeros(m.nV), rtol=1e-13, null_space=1)
        assert info["status"] == 0
        sols.append(s)
    assert np.max(np.abs(sols[0] - detJ * sols[1])) < 1e-9 * np.max(np.abs(sols[0]))


def test_apply_ideal_matches_csr_on_small_mesh():
    m = O.OracleMesh((4, 3), 5, (0, 0), (1, 1), (1, 0), (1, 1, 0, 0))
    for mode in (0, 1):
        h = O.Helm(m, 3.0, -1.5, mode)
        x = np.random.default_rng(3).standard_normal(m.n)
        y1, y2 = h.matvec(x), m.apply_ideal(3.0, -1.5, mode, x)
        assert np.linalg.norm(y1 - y2) < 1e-13 * np.linalg.norm(y1)


# ---- a11-a16: semi-Lagrangian step ----------------------------------------------------
def test_constant_velocity_translation_closed_form():
    m = O.OracleMesh((8, 6), 5, (0, 0), (2.0, 1.5), (1, 1))
    V = m.verts()
    U = np.array([0.137, -0.071])
    dt = 0.9
    vel = np.tile(U, m.nV)
    q = np.sin(np.pi * V[:, 0]) * np.cos(TWO_PI / 1.5 * V[:, 1])
    qsl, dep, xy = m.advect_first(q, 1, vel, dt)
    exp = V - dt * U
    exp[:, 0] = np.where(exp[:, 0] < 0, exp[:, 0] + 2.0, exp[:, 0])
    exp[:, 1] = np.where(exp[:, 1] > 1.5, exp[:, 1] - 1.5, exp[:, 1])
    assert np.max(np.abs(xy - exp)) < 1e-13
    ex = np.floor(exp[:, 0] / 0.25).astype(int).clip(0, 7)
    ey = np.floor(exp[:, 1] / 0.25).astype(int).clip(0, 5)
    far = (np.abs(exp[:, 0] / 0.25 - np.rint(exp[:, 0] / 0.25)) > 1e-9) & \
          (np.abs(exp[:, 1] / 0.25 - np.rint(exp[:, 1] / 0.25)) > 1e-9)
    assert np.array_equal(dep[far], (ey * 8 + ex)[far])
    # smooth field translated: spectral accuracy of the GLL interpolation
    qe = np.sin(np.pi * exp[:, 0]) * np.cos(TWO_PI / 1.5 * exp[:, 1])
    assert np.max(np.abs(qsl - qe)) < 1e-3


def test_interpolation_reproduces_polynomials():
    N = 6
    m = O.OracleMesh((3, 2), N, (0, 0), (3.0, 2.0), (0, 0))
    V = m.verts()
    f = lambda x, y: (x - 0.3) ** 3 * (y + 0.1) ** 2 - 2.0 * x * y  # degree <= N in each element
    vals = f(V[:, 0], V[:, 1])
    pts = np.random.default_rng(4).uniform([0, 0], [3, 2], size=(50, 2))
    got = np.array([m.interp_global(vals, 1, p)[0] for p in pts])
    assert np.max(np.abs(got - f(pts[:, 0], pts[:, 1]))) < 1e-11


def test_genvel_exact_for_polynomials():
    m = O.OracleMesh((3, 3), 5, (0, 0), (1.5, 1.0), (0, 0))
    V = m.verts()
    psi = V[:, 0] ** 3 * V[:, 1] + V[:, 1] ** 2
    vel = m.genvel(psi)
    u = -(V[:, 0] ** 3 + 2 * V[:, 1])
    v = 3 * V[:, 0] ** 2 * V[:, 1]
    assert np.max(np.abs(vel[:, 0] - u)) < 1e-10
    assert np.max(np.abs(vel[:, 1] - v)) < 1e-10


def test_calc_timestep_closed_form():
    m = O.OracleMesh((4, 5), 4, (0, 0), (2.0, 1.0), (0, 0))
    vel = np.zeros((m.nV, 2))
    vel[7] = [3.0, 4.0]
    assert m.calc_timestep(vel.ravel(), 0.5) == 0.5 * 0.2 / 5.0
