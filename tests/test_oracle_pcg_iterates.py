"""The oracle's Jacobi-PCG iterates, not only its converged solution.

orc_helm_pcg restates the CG + Jacobi of the reference's KSPSolve
(src/linsys/HelmholtzEqn.cpp:120-131; x0 = 0, PETSc -ksp_type cg -pc_type
jacobi).  Its k-th iterate (maxit = k) is checked two independent ways:
  * against scipy.sparse.linalg.cg with the same diagonal preconditioner and
    x0 = 0, the textbook recurrence in another implementation;
  * against what the mathematics fixes for any correct PCG: x_k minimises the
    A-norm error over the Krylov space K_k(M^-1 A, M^-1 b), computed here by a
    dense Galerkin solve on an orthonormal basis of that space.
A wrong alpha, a wrong beta or a dropped preconditioner moves x_k off both
(checked by planting such a mistake in a numpy copy of the recurrence, which
fails the optimality check).  The oracle's OpenMP threads must not change a bit.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import oracle as O

CASES = [
    # label, nEls, N, periodic, bc, c_mass, c_lap, detj
    ("chan_helm", (6, 5), 5, (1, 0), (1, 1, 0, 0), -100.0, -1.0, 1),
    ("dir_poisson", (5, 5), 4, (0, 0), (1, 1, 1, 1), 0.0, -1.0, 0),
    ("per_helm", (4, 6), 6, (1, 1), (0, 0, 0, 0), 3.0, 1.0, 1),
]


def system(case):
    lab, nEls, N, per, bc, cm, cl, detj = case
    m = O.OracleMesh(nEls, N, (0, 0), (1, 1), per, bc)
    h = O.Helm(m, cm, cl, detj)
    A = h.scipy().toarray()
    s = 1.0 if A[0, 0] > 0 else -1.0  # the solver's sign normalisation
    f = np.random.default_rng(31).standard_normal(m.n)
    return m, h, s * A, s * f, f


def krylov_optimal(A, Minv, b, k):
    """argmin over x in K_k(Minv A, Minv b) of ||x - A^-1 b||_A."""
    v = Minv * b
    basis = []
    for _ in range(k):
        for q in basis:  # modified Gram-Schmidt, twice
            v = v - (q @ v) * q
        for q in basis:
            v = v - (q @ v) * q
        v = v / np.linalg.norm(v)
        basis.append(v)
        v = Minv * (A @ v)
    Q = np.stack(basis, 1)
    y = np.linalg.solve(Q.T @ A @ Q, Q.T @ b)
    return Q @ y


def textbook_pcg(A, Minv, b, k, beta_mistake=False):
    x = np.zeros_like(b)
    r = b.copy()
    z = Minv * r
    p = z.copy()
    rz = r @ z
    for _ in range(k):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        z = Minv * r
        rz_new = r @ z
        beta = (rz_new / rz) if not beta_mistake else (rz_new / rz) ** 0.5
        p = z + beta * p
        rz = rz_new
    return x


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("k", [1, 3, 8])
def test_oracle_iterates_match_scipy_cg_and_krylov_optimum(case, k):
    m, h, A, b, f = system(case)
    Minv = 1.0 / np.diag(A)
    x_orc, info = h.pcg(f, rtol=1e-300, maxit=k)
    assert info["iters"] == k and info["status"] == 1
    x_sp, flag = spla.cg(A, b, x0=np.zeros(m.n), rtol=1e-300, atol=0.0, maxiter=k,
                         M=spla.LinearOperator(A.shape, matvec=lambda v: Minv * v))
    assert flag == k
    x_opt = krylov_optimal(A, Minv, b, k)
    scale = np.linalg.norm(x_opt)
    assert np.linalg.norm(x_orc - x_sp) <= 1e-11 * scale
    assert np.linalg.norm(x_orc - x_opt) <= 1e-9 * scale


def test_a_planted_beta_mistake_fails_the_optimality_check():
    m, h, A, b, f = system(CASES[0])
    Minv = 1.0 / np.diag(A)
    x_opt = krylov_optimal(A, Minv, b, 8)
    good = textbook_pcg(A, Minv, b, 8)
    bad = textbook_pcg(A, Minv, b, 8, beta_mistake=True)
    scale = np.linalg.norm(x_opt)
    assert np.linalg.norm(good - x_opt) <= 1e-9 * scale
    assert np.linalg.norm(bad - x_opt) > 1e-4 * scale


def test_oracle_results_do_not_depend_on_its_thread_count():
    m = O.OracleMesh((24, 20), 6, (0, 0), (1, 1), (1, 0), (1, 1, 0, 0))
    before = O.get_threads()
    try:
        out = []
        for th in (1, 4):
            O.set_threads(th)
            h = O.Helm(m, -100.0, -1.0, 1)
            f = np.random.default_rng(3).standard_normal(m.n)
            x, info = h.pcg(f, rtol=1e-10)
            out.append((h.matvec(f), x, info["iters"]))
            h.close()
        assert np.array_equal(out[0][0], out[1][0])
        assert np.array_equal(out[0][1], out[1][1]) and out[0][2] == out[1][2]
    finally:
        O.set_threads(before)
