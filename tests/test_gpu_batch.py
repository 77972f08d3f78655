"""Batched mixed-precision solves.

The reference solves one HelmholtzEqn at a time (KSPSolve,
src/linsys/HelmholtzEqn.cpp:120-131); the batched solver must give every
member exactly what the single mixed-precision solve gives it alone (same
iterations, bitwise the same solution), whatever else is in the batch, and
each solution must meet the FP64 target against the oracle's assembled
operator (Matrix::Assemble, src/base/Matrix.cpp:47-133).
"""
import numpy as np
import pytest

import se_amd
from oracle import oracle as O
from se_amd import workloads as W

from test_gpu_parity import OP_CASES, TWO_PI

pytestmark = pytest.mark.gpu

DEFINITE = [c for c in OP_CASES if not (c[7] == 0.0 and c[5] == (1, 1))]
NULLSPACE = [c for c in OP_CASES if c[7] == 0.0 and c[5] == (1, 1)]


def build(gpu_ctx, cases):
    out = []
    for lab, nEls, N, lo, hi, per, bc, cm, cl, detj in cases:
        m = se_amd.Mesh(gpu_ctx, nEls, N, lo, hi, per, bc)
        h = se_amd.Helm(m, cm, cl, detj)
        om = O.OracleMesh(nEls, N, lo, hi, per, bc)
        out.append(dict(lab=lab, m=m, h=h, oh=O.Helm(om, cm, cl, detj)))
    return out


@pytest.mark.parametrize("prec", [se_amd.PREC_FP32_IR, se_amd.PREC_BF16_IR], ids=["fp32", "bf16"])
def test_batch_equals_single_solves(gpu_ctx, prec):
    mem = build(gpu_ctx, DEFINITE)
    if prec == se_amd.PREC_BF16_IR:   # the near-singular case stagnates in BF16 (see test_gpu_mixed)
        mem = [x for x in mem if x["lab"] != "odd_13x11_p4"]
    rng = np.random.default_rng(41)
    fs = [rng.standard_normal(x["m"].n) for x in mem]
    single = [x["h"].pcg(f, rtol=1e-10, precision=prec, preconditioner=se_amd.PC_JACOBI) for x, f in zip(mem, fs)]
    b = se_amd.Batch(gpu_ctx, [x["h"] for x in mem])
    xs, infos = b.pcg(fs, precision=prec, rtol=1e-10, preconditioner=se_amd.PC_JACOBI)
    for x, f, (x1, i1), xb, ib in zip(mem, fs, single, xs, infos):
        assert ib["status"] == se_amd.SOLVE_CONVERGED, (x["lab"], ib)
        assert ib["iters"] == i1["iters"] and ib["outer_iters"] == i1["outer_iters"], (x["lab"], ib, i1)
        assert np.array_equal(xb, x1), x["lab"]
        r = f - x["oh"].matvec(xb)
        assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(f), x["lab"]
    # a second solve with the same batch (tuned launch shape, cached graph)
    xs2, infos2 = b.pcg(fs, precision=prec, rtol=1e-10, preconditioner=se_amd.PC_JACOBI)
    for xb, xb2 in zip(xs, xs2):
        assert np.array_equal(xb, xb2)
    b.close()


def test_batch_null_space_members(gpu_ctx):
    mem = build(gpu_ctx, NULLSPACE)
    rng = np.random.default_rng(43)
    fs = [rng.standard_normal(x["m"].n) for x in mem]
    b = se_amd.Batch(gpu_ctx, [x["h"] for x in mem])
    xs, infos = b.pcg(fs, rtol=1e-10, null_space=True, preconditioner=se_amd.PC_JACOBI)
    for x, f, xb, ib in zip(mem, fs, xs, infos):
        x1, i1 = x["h"].pcg(f, rtol=1e-10, null_space=True, precision=se_amd.PREC_FP32_IR, preconditioner=se_amd.PC_JACOBI)
        assert ib["status"] == 0 and np.array_equal(xb, x1)
        fp = f - f.mean()
        r = fp - x["oh"].matvec(xb)
        r -= r.mean()
        assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(fp)
    b.close()


def test_batch_same_alone_and_in_any_mix(gpu_ctx):
    """A member's result does not depend on its batch companions or position."""
    mem = build(gpu_ctx, DEFINITE[:6])
    rng = np.random.default_rng(47)
    fs = [rng.standard_normal(x["m"].n) for x in mem]
    b1 = se_amd.Batch(gpu_ctx, [x["h"] for x in mem])
    xs1, _ = b1.pcg(fs, rtol=1e-10, preconditioner=se_amd.PC_JACOBI)
    b1.close()
    order = [4, 1, 5, 0]
    b2 = se_amd.Batch(gpu_ctx, [mem[j]["h"] for j in order])
    xs2, _ = b2.pcg([fs[j] for j in order], rtol=1e-10, preconditioner=se_amd.PC_JACOBI)
    b2.close()
    for k, j in enumerate(order):
        assert np.array_equal(xs2[k], xs1[j])


def test_batch_ensemble_members(gpu_ctx):
    """Members drawn like the 512-member ensemble (nEls in {8,16,32,64}^2,
    p in {4,6,8}, lambda log-uniform in [1, 1e3], doubly periodic, physical
    Jacobian): batch = single, and the FP64 target is met with the exact
    operator."""
    cfgs = W.ensemble_512(count=24)
    hs, fs = [], []
    for c in cfgs:
        m = se_amd.Mesh(gpu_ctx, c["nEls"], c["N"], c["lo"], c["hi"], c["periodic"], c["bc"])
        hs.append(se_amd.Helm(m, c["c_mass"], c["c_lap"], c["detj_mode"]))
        fs.append(np.random.default_rng(c["seed"]).standard_normal(m.n))
    b = se_amd.Batch(gpu_ctx, hs)
    xs, infos = b.pcg(fs, rtol=1e-10, preconditioner=se_amd.PC_JACOBI)
    for h, f, xb, ib in zip(hs, fs, xs, infos):
        assert ib["status"] == 0
        r = f - h.apply(xb)
        assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(f)
    for h, f, xb, ib in list(zip(hs, fs, xs, infos))[:6]:
        x1, i1 = h.pcg(f, rtol=1e-10, precision=se_amd.PREC_FP32_IR, preconditioner=se_amd.PC_JACOBI)
        assert np.array_equal(xb, x1) and i1["iters"] == ib["iters"]
    b.close()


def test_batch_argument_errors(gpu_ctx):
    m = se_amd.Mesh(gpu_ctx, (4, 4), 4, (0, 0), (1, 1), (0, 0), (1, 1, 1, 1))
    h = se_amd.Helm(m, 1.0, 1.0, 1)
    with pytest.raises(se_amd.SEError):
        se_amd.Batch(gpu_ctx, [h, h])
    b = se_amd.Batch(gpu_ctx, [h])
    with pytest.raises(se_amd.SEError):
        b.pcg([np.ones(m.n)], precision=se_amd.PREC_FP64, preconditioner=se_amd.PC_JACOBI)
    b.close()
