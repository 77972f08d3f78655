"""Mixed-precision solves: FP64 iterative refinement around an FP32 / BF16
inner Jacobi-PCG on the separable inner operator.

The reference solves in FP64 only (KSPSolve, src/linsys/HelmholtzEqn.cpp:120-131),
so the bar is the FP64 one: the refined solution meets the 1e-10 relative
residual target against the oracle's assembled operator (a restatement of
Matrix::Assemble, src/base/Matrix.cpp:47-133), and agrees with the FP64 GPU
solve to what the conditioning allows.  The inner operator itself is checked
against the assembled operator at FP32 / BF16 resolution.
"""
import numpy as np
import pytest

import se_amd
from oracle import oracle as O
from se_amd import workloads as W

from test_gpu_parity import OP_CASES, rel

pytestmark = pytest.mark.gpu


def bf16_round(x):
    """Round float32 values to bfloat16 (nearest, ties to even), back as float32."""
    b = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


@pytest.fixture(scope="module", params=OP_CASES, ids=[c[0] for c in OP_CASES])
def case(request, gpu_ctx):
    lab, nEls, N, lo, hi, per, bc, cm, cl, detj = request.param
    om = O.OracleMesh(nEls, N, lo, hi, per, bc)
    oh = O.Helm(om, cm, cl, detj)
    m = se_amd.Mesh(gpu_ctx, nEls, N, lo, hi, per, bc)
    h = se_amd.Helm(m, cm, cl, detj)
    null = cm == 0.0 and per == (1, 1)
    yield dict(om=om, oh=oh, m=m, h=h, null=null, N=N, label=lab)
    h.close()
    m.close()


@pytest.mark.parametrize("prec", [se_amd.PREC_FP32_IR, se_amd.PREC_BF16_IR], ids=["fp32", "bf16"])
def test_inner_operator_matches_assembled(case, prec):
    """sign*A~ x agrees with sign*A x (FP64 CSR) to FP32 accumulation error,
    componentwise against |A||x|; the dropped Jacobian cross terms are ~1e-12."""
    oh, h, n = case["oh"], case["h"], case["m"].n
    x = np.random.default_rng(31).standard_normal(n).astype(np.float32)
    y, sgn = h.apply_inner(x, prec)
    xin = bf16_round(x) if prec == se_amd.PREC_BF16_IR else x
    xin = xin.astype(np.float64)
    y0 = sgn * oh.matvec(xin)
    absAx = abs(oh.scipy()) @ np.abs(xin)
    err = np.abs(y.astype(np.float64) - y0) / absAx
    assert np.max(err) <= 3e-6, np.max(err)
    assert sgn * oh.inv_diag()[0] > 0


# Near-singular system: periodic literal-mode Laplacian (detJ := 1 scales the
# stiffness by ~1/|J|) with c_mass = 1, so the constant mode's eigenvalue is
# ~1e-5 of the spectrum and dominates the solution; a BF16 search direction
# (8-bit mantissa) cannot resolve the other modes beside it.  The refinement
# must say so (SOLVE_STAGNATED) instead of running on.
BF16_UNRESOLVABLE = {"odd_13x11_p4"}


@pytest.mark.parametrize("prec", [se_amd.PREC_FP32_IR, se_amd.PREC_BF16_IR], ids=["fp32", "bf16"])
def test_refined_solve_meets_fp64_target(case, prec):
    oh, h, n, null = case["oh"], case["h"], case["m"].n, case["null"]
    f = np.random.default_rng(17).standard_normal(n)
    x, info = h.pcg(f, rtol=1e-10, null_space=null, precision=prec, preconditioner=se_amd.PC_JACOBI)
    if prec == se_amd.PREC_BF16_IR and case["label"] in BF16_UNRESOLVABLE:
        assert info["status"] == se_amd.SOLVE_STAGNATED, info
        # declared within 32 reliable-update windows of 256 iterations, far
        # below the 1000-update limit it would otherwise run to
        assert info["iters"] <= 32 * 256, info
        return
    assert info["status"] == se_amd.SOLVE_CONVERGED, info
    assert info["true_rel_residual"] <= 1e-10
    assert info["outer_iters"] >= 1
    fp = f - f.mean() if null else f
    r = fp - oh.matvec(x)
    if null:
        r = r - r.mean()
    assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(fp)
    x64, i64 = h.pcg(f, rtol=1e-10, null_space=null, preconditioner=se_amd.PC_JACOBI)
    if null:
        x, x64 = x - x.mean(), x64 - x64.mean()
    if case["N"] <= 6:
        assert rel(x, x64) <= 1e-7, rel(x, x64)


def test_refined_solve_deterministic(case):
    h, n, null = case["h"], case["m"].n, case["null"]
    f = np.random.default_rng(3).standard_normal(n)
    x1, i1 = h.pcg(f, rtol=1e-10, null_space=null, precision=se_amd.PREC_FP32_IR, preconditioner=se_amd.PC_JACOBI)
    x2, i2 = h.pcg(f, rtol=1e-10, null_space=null, precision=se_amd.PREC_FP32_IR, preconditioner=se_amd.PC_JACOBI)
    assert np.array_equal(x1, x2) and i1["iters"] == i2["iters"]


def test_refined_status_codes(gpu_ctx):
    m = se_amd.Mesh(gpu_ctx, (4, 4), 4, (0, 0), (1, 1), (0, 0), (1, 1, 1, 1))
    h = se_amd.Helm(m, 1.0, 1.0, 1)
    for prec in (se_amd.PREC_FP32_IR, se_amd.PREC_BF16_IR):
        _, info = h.pcg(np.zeros(m.n), precision=prec, preconditioner=se_amd.PC_JACOBI)
        assert info["status"] == se_amd.SOLVE_CONVERGED and info["iters"] == 0
        f = np.ones(m.n)
        f[2] = np.inf
        _, info = h.pcg(f, precision=prec, preconditioner=se_amd.PC_JACOBI)
        assert info["status"] == se_amd.SOLVE_NONFINITE
        # an outer-step limit that cannot reach the target reports maxit
        _, info = h.pcg(np.ones(m.n), rtol=1e-14, precision=prec, max_outer=1, inner_rtol=1e-2, preconditioner=se_amd.PC_JACOBI)
        assert info["status"] == se_amd.SOLVE_MAXIT and info["outer_iters"] == 1


def test_refined_full_size(gpu_ctx):
    """helmholtz_128 (128x128 elements, p=8, doubly periodic, physical
    Jacobian): the refined FP32 and BF16 solves reach the FP64 target with the
    exact operator and agree with the FP64 solve."""
    cfg = W.helmholtz_128()
    m = se_amd.Mesh(gpu_ctx, cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], cfg["periodic"], cfg["bc"])
    h = se_amd.Helm(m, cfg["c_mass"], cfg["c_lap"], cfg["detj_mode"])
    f = np.random.default_rng(29).standard_normal(m.n)
    x64, i64 = h.pcg(f, rtol=1e-10, preconditioner=se_amd.PC_JACOBI)
    assert i64["status"] == 0
    for prec in (se_amd.PREC_FP32_IR, se_amd.PREC_BF16_IR):
        x, info = h.pcg(f, rtol=1e-10, precision=prec, preconditioner=se_amd.PC_JACOBI)
        print("prec", prec, info, "fp64 iters", i64["iters"])
        assert info["status"] == 0 and info["true_rel_residual"] <= 1e-10
        r = f - h.apply(x)
        assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(f)
        assert rel(x, x64) <= 1e-6
