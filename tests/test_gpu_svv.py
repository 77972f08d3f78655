"""HelmholtzEqn with spectral vanishing viscosity on the device (se_helm_set_svv:
HelmholtzEqn.cpp:71-76, SVV::AssembleElement src/base/SVV.cpp:120-176)
against the oracle's assembled operator, whose element blocks restate the
reference's 2-D construction literally (pinned in tests/test_oracle_svv.py).

The device applies the separable 1-D factors with an owner-computes gather;
apply, inverse diagonal and Dirichlet lifting agree with the CSR to 1e-12
relative, solves to the solver tolerance: FP64 Jacobi-PCG, and CG preconditioned
by the fast diagonalisation of the operator without its SVV term (any
transform precision, alone or batched).  A vector (2-dof) field is two
such operators, one per dof with that dof's Dirichlet sides: the reference's
2-dof system (BCs.cpp:176-322) is block diagonal, dof 0's unknowns first."""
import numpy as np
import pytest

import se_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [((6, 5), 4, (1, 0), (1, 1, 0, 0), 2), ((5, 5), 6, (1, 1), (0, 0, 0, 0), 3),
         ((4, 6), 8, (0, 0), (1, 1, 1, 1), 4), ((7, 4), 5, (1, 0), (0, 1, 0, 0), 1)]


def build(gpu_ctx, nEls, N, per, bc, mN, c=1.0, nu_dt=0.01, dt=0.05, detj=0, seed=0):
    lo, hi = (0.0, 0.0), (2.0, 1.0)
    m = se_amd.Mesh(gpu_ctx, nEls, N, lo, hi, per, bc)
    om = O.OracleMesh(nEls, N, lo, hi, per, bc)
    rng = np.random.default_rng(seed)
    field = rng.standard_normal(om.nV)
    oh = O.Helm(om, c, nu_dt, detj, sol_field=field, svv=(dt, mN))
    h = se_amd.Helm(m, c, nu_dt, detj)
    h.set_svv(dt, mN)
    h.set_dirichlet(field)
    return m, om, oh, h, field, rng


@pytest.mark.parametrize("case", CASES, ids=lambda c: "N%d-mN%d-per%d%d" % (c[1], c[4], *c[2]))
def test_svv_operator_matches_oracle(gpu_ctx, case):
    m, om, oh, h, field, rng = build(gpu_ctx, *case)
    x = rng.standard_normal(om.n)
    want = oh.matvec(x)
    got = h.apply(x)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    dw, dg = oh.inv_diag(), h.inv_diag()
    assert np.max(np.abs(dg - dw) / np.abs(dw)) <= 1e-12
    lw, lg = oh.lift(), h.lift()
    if np.max(np.abs(lw)) > 0:
        assert np.max(np.abs(lg - lw)) <= 1e-12 * np.max(np.abs(lw))
    # the SVV term is there: without it the operator differs
    plain = se_amd.Helm(m, 1.0, 0.01, 0)
    assert np.max(np.abs(plain.apply(x) - want)) > 1e-6 * np.max(np.abs(want))


@pytest.mark.parametrize("case", CASES[:3], ids=lambda c: "N%d-mN%d" % (c[1], c[4]))
def test_svv_solve_matches_oracle(gpu_ctx, case):
    m, om, oh, h, field, rng = build(gpu_ctx, *case, detj=1)
    rhs = rng.standard_normal(om.nV)
    want, oinfo = oh.solve(rhs, field, rtol=1e-12)
    assert oinfo["status"] == 0
    got, info = h.solve(rhs, field, rtol=1e-12)
    assert info["status"] == 0 and info["true_rel_residual"] <= 1e-12
    assert np.max(np.abs(got - want)) <= 1e-9 * np.max(np.abs(want))


# FDM of the operator without its SVV term preconditions the exact operator
# (kappa(A0^-1 A) 1.02-2.45 on these cases, measured with the oracle's CSRs)
FDM_PRECS = [(se_amd.PREC_FP64, 30), (se_amd.PREC_FP32_IR, 30), (se_amd.PREC_BF16_IR, 150),
             (se_amd.PREC_FP64_FP32C, 30), (se_amd.PREC_FP64_BF16C, 150)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "N%d-mN%d-per%d%d" % (c[1], c[4], *c[2]))
def test_svv_solves_with_fdm_in_every_precision(gpu_ctx, case):
    m, om, oh, h, field, rng = build(gpu_ctx, *case)
    f = rng.standard_normal(om.n)
    for prec, itmax in FDM_PRECS:
        x, info = h.pcg(f, rtol=1e-10, precision=prec, preconditioner=se_amd.PC_FDM)
        print("prec %d: %d iterations, true rel residual %.2e" % (prec, info["iters"], info["true_rel_residual"]))
        assert info["status"] == se_amd.SOLVE_CONVERGED, info
        assert info["iters"] <= itmax, info
        assert np.linalg.norm(f - oh.matvec(x)) <= 1.5e-10 * np.linalg.norm(f)


def test_svv_members_batch_with_fdm(gpu_ctx):
    """SVV and plain members in one FDM batch: each result bitwise its single solve."""
    hs, ohs, fs = [], [], []
    for k, case in enumerate(CASES):
        m, om, oh, h, field, rng = build(gpu_ctx, *case, seed=10 + k)
        hs.append(h)
        ohs.append(oh)
        fs.append(rng.standard_normal(om.n))
    m = se_amd.Mesh(gpu_ctx, (6, 6), 5, (0.0, 0.0), (2.0, 1.0), (1, 0), (1, 1, 0, 0))
    om = O.OracleMesh((6, 6), 5, (0.0, 0.0), (2.0, 1.0), (1, 0), (1, 1, 0, 0))
    hs.append(se_amd.Helm(m, 3.0, 0.02, 1))
    ohs.append(O.Helm(om, 3.0, 0.02, 1))
    fs.append(np.random.default_rng(3).standard_normal(om.n))
    b = se_amd.Batch(gpu_ctx, hs)
    for prec in (se_amd.PREC_FP64, se_amd.PREC_FP32_IR, se_amd.PREC_FP64_BF16C):
        xs, infos = b.pcg(fs, precision=prec, rtol=1e-10, preconditioner=se_amd.PC_FDM)
        for h, oh, f, xb, ib in zip(hs, ohs, fs, xs, infos):
            assert ib["status"] == se_amd.SOLVE_CONVERGED, ib
            assert np.linalg.norm(f - oh.matvec(xb)) <= 1.5e-10 * np.linalg.norm(f)
            x1, i1 = h.pcg(f, rtol=1e-10, precision=prec, preconditioner=se_amd.PC_FDM)
            assert i1["iters"] == ib["iters"] and np.array_equal(x1, xb)
    # the inner Jacobi-PCG's separable operator cannot carry the SVV term
    with pytest.raises(se_amd.SEError):
        b.pcg(fs, precision=se_amd.PREC_FP32_IR, rtol=1e-10, preconditioner=se_amd.PC_JACOBI)
    b.close()


def test_svv_refusals(gpu_ctx):
    m, om, oh, h, field, rng = build(gpu_ctx, *CASES[0])
    with pytest.raises(se_amd.SEError):  # mixed-precision Jacobi-PCG
        h.pcg(rng.standard_normal(om.n), precision=se_amd.PREC_FP32_IR, preconditioner=se_amd.PC_JACOBI)
    with pytest.raises(se_amd.SEError):
        h.set_svv(0.05, 2)  # once per operator
    with pytest.raises(se_amd.SEError):
        se_amd.Helm(m, 1.0, 0.01, 0).set_svv(0.05, m.N)  # cutoff beyond the modes


def test_vector_helmholtz_is_two_dof_operators(gpu_ctx):
    """A velocity-like 2-dof field with different Dirichlet sides per dof
    (u: walls top and bottom; v: walls on all four sides) in a non-periodic
    box: each dof solves with its own operator; the stacked unknowns follow
    the reference's 2-dof numbering (dof 0's unknowns, then dof 1's)."""
    nEls, N, mN, dt = (5, 4), 6, 3, 0.02
    lo, hi = (0.0, 0.0), (2.0, 1.0)
    rng = np.random.default_rng(4)
    sides = [(1, 1, 0, 0), (1, 1, 1, 1)]
    sol, refs, sizes = [], [], []
    for d, bc in enumerate(sides):
        m = se_amd.Mesh(gpu_ctx, nEls, N, lo, hi, (0, 0), bc)
        om = O.OracleMesh(nEls, N, lo, hi, (0, 0), bc)
        field = rng.standard_normal(om.nV)
        rhs = rng.standard_normal(om.nV)
        h = se_amd.Helm(m, 1.0 / dt, 1e-3 * dt, 0)
        h.set_svv(dt, mN)
        h.set_dirichlet(field)  # HelmholtzEqn::Assemble
        x, info = h.solve(rhs, field, rtol=1e-12)
        assert info["status"] == 0
        oh = O.Helm(om, 1.0 / dt, 1e-3 * dt, 0, sol_field=field, svv=(dt, mN))
        want, oinfo = oh.solve(rhs, field, rtol=1e-12)
        assert oinfo["status"] == 0
        sol.append(om.update_vec(x))
        refs.append(om.update_vec(want))
        sizes.append(om.n)
    assert sizes[0] > sizes[1]  # the dofs have different unknowns
    got, want = np.concatenate(sol), np.concatenate(refs)
    assert np.max(np.abs(got - want)) <= 1e-9 * np.max(np.abs(want))
