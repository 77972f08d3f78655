"""CPU oracle vs the golden vectors written by the compiled reference
(tests/golden/make_golden.py).  Every comparison is bit-exact."""
import numpy as np
import pytest

from golden_io import load_gll, load_mesh_case, mesh_args, mesh_cases
from oracle import oracle as O

CASES = mesh_cases()


@pytest.mark.parametrize("N", range(2, 13))
def test_gll_bit_exact(N):
    g = load_gll()
    a, w, D, _ = O.gll(N)
    assert np.array_equal(a, g["gll_%d_x" % N])
    assert np.array_equal(w, g["gll_%d_w" % N])
    assert np.array_equal(D, g["gll_%d_D" % N])


@pytest.fixture(scope="module", params=CASES)
def case(request):
    d = load_mesh_case(request.param)
    m = O.OracleMesh(*mesh_args(d))
    yield d, m
    m.close()


def test_sizes_verts_maps(case):
    d, m = case
    s = d["sizes"]
    assert (m.nVerts[0], m.nVerts[1], m.nV, m.nElsTotal, m.n, m.size0) == tuple(int(v) for v in s)
    assert np.array_equal(m.verts(), d["verts"])
    f2v, v2f = m.maps()
    assert np.array_equal(f2v, d["f2v"])
    assert np.array_equal(v2f, d["v2f"])


def test_element_operators(case):
    d, m = case
    for k, e in enumerate(d["els"]):
        assert np.array_equal(m.elem_laplacian(int(e), 1.3, 0), d["lap"][k])
        M = m.elem_mass(int(e), 0.7)
        assert np.array_equal(np.diag(M), d["mass_diag"][k])
        assert np.count_nonzero(M - np.diag(np.diag(M))) == 0
        assert np.array_equal(m.elem_field_rhs(int(e), 1.0, d["field"]), d["frhs"][k])
        assert np.array_equal(np.array([m.det_jac(int(e), p) for p in range(m.nl)]), d["detj"][k])


def test_genvel(case):
    d, m = case
    assert np.array_equal(m.genvel(d["psi"]), d["genvel"])


def test_advect_first(case):
    d, m = case
    q, dep, xy = m.advect_first(d["q"], 1, d["vel"], float(d["dt"]))
    assert np.array_equal(dep, d["adv1_dep"])
    assert np.array_equal(xy, d["adv1_xy"])
    assert np.array_equal(q, d["adv1_q"])
    v2 = m.advect_first(d["vel"], 2, d["vel"], float(d["dt"]), want_dep=False)[0]
    assert np.array_equal(v2, d["adv1_vel"])


def test_advect_second(case):
    d, m = case
    q = m.advect_second(d["q"], d["qprev"], 1, d["vel"], d["velprev"], float(d["dt"]))
    assert np.array_equal(q, d["adv2_q"])


def test_interp_and_local_coords(case):
    d, m = case
    vals = np.array([m.interp_global(d["field"], 1, p)[0] for p in d["pts"]])
    assert np.array_equal(vals, d["interp"])
    for p, e, l in zip(d["pts"], d["g2l_el"], d["g2l_l"]):
        e2, l2 = m.global_to_local(p)
        assert e2 == e and np.array_equal(l2, l)
