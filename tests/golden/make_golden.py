"""Write golden vectors from the compiled reference element layer.

Runs only in a container that has the reference tree: oracle/build_ref.sh
compiles the reference's own src/base classes (Legendre, Mesh, BCs, Field,
Operator, RHSOp, Advector) into oracle/_ref/libseref.so, and this script calls
them through oracle/ref_shim.cpp on small meshes with seeded inputs.  The
resulting .npz files are plain data (inputs and reference outputs); the tests
compare both the CPU oracle and the HIP path against them.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from se_amd.workloads import Rng  # noqa: E402

TWO_PI = 2.0 * np.pi

# (name, nEls, N, lo, hi, periodic, bc(bottom, top, left, right))
CASES = [
    ("per_8x8_p4", (8, 8), 4, (0.0, 0.0), (TWO_PI, TWO_PI), (1, 1), (0, 0, 0, 0)),
    ("chan_6x5_p6", (6, 5), 6, (0.0, 0.0), (1.0, 1.0), (1, 0), (1, 1, 0, 0)),
    ("dir_4x4_p8", (4, 4), 8, (-1.0, -1.0), (1.0, 1.0), (0, 0), (1, 1, 1, 1)),
    ("xper_top_5x3_p5", (5, 3), 5, (0.0, -2.0), (3.0, 2.0), (1, 0), (0, 1, 0, 0)),
    ("xper_3x4_p3", (3, 4), 3, (0.0, 0.0), (1.0, 2.0), (1, 0), (0, 0, 0, 0)),
    ("mixed_4x3_p7", (4, 3), 7, (-1.0, 0.0), (2.0, 1.0), (0, 0), (1, 0, 1, 0)),
    ("per_5x5_p2", (5, 5), 2, (0.0, 0.0), (1.0, 1.0), (1, 1), (0, 0, 0, 0)),
]


def inputs(m, lo, hi, seed):
    rng = Rng(seed)
    V = m.verts()
    L = np.array(hi) - np.array(lo)
    sx = TWO_PI / L[0]
    sy = TWO_PI / L[1]
    psi = np.sin(sx * V[:, 0]) * np.cos(0.5 * sy * V[:, 1]) + 0.25 * rng.normal(m.nV)
    q = rng.normal(m.nV)
    qprev = rng.normal(m.nV)
    umax = 0.4 * min(L[0] / m.nEls[0], L[1] / m.nEls[1])
    vel = np.stack([umax * (1.0 + 0.5 * np.cos(sy * V[:, 1])), umax * 0.7 * np.sin(sx * V[:, 0])], 1).ravel()
    velprev = vel * 0.9 + 0.05 * umax * rng.normal(2 * m.nV)
    pts = np.array(lo) + L * rng.uniform(2 * 40).reshape(40, 2)
    field = rng.normal(m.nV)
    return dict(psi=psi, q=q, qprev=qprev, vel=vel, velprev=velprev, pts=pts, field=field)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libseref.so is missing: run oracle/build_ref.sh first")
    out = {}
    for N in range(2, 13):
        r = O.RefMesh((2, 2), N, (0.0, 0.0), (1.0, 1.0), (0, 0), (0, 0, 0, 0))
        a, w, D = r.gll()
        out["gll_%d_x" % N], out["gll_%d_w" % N], out["gll_%d_D" % N] = a, w, D
        r.close()
    np.savez_compressed(os.path.join(HERE, "gll.npz"), **out)

    for ci, (name, nEls, N, lo, hi, per, bc) in enumerate(CASES):
        r = O.RefMesh(nEls, N, lo, hi, per, bc)
        inp = inputs(r, lo, hi, 7000 + ci)
        d = dict(nEls=np.array(nEls), N=np.array(N), lo=np.array(lo), hi=np.array(hi), periodic=np.array(per),
                 bc=np.array(bc), sizes=np.array([r.nVerts[0], r.nVerts[1], r.nV, r.nElsTotal, r.n, r.size0]))
        d.update(inp)
        d["verts"] = r.verts()
        d["f2v"], d["v2f"] = r.maps()
        els = sorted({0, r.nElsTotal // 2, r.nElsTotal - 1})
        if N >= 8:
            els = [r.nElsTotal // 2]
        d["els"] = np.array(els)
        d["lap"] = np.stack([r.elem_laplacian(e, 1.3) for e in els])
        d["mass_diag"] = np.stack([np.diag(r.elem_mass(e, 0.7)) for e in els])
        d["frhs"] = np.stack([r.elem_field_rhs(e, 1.0, inp["field"]) for e in els])
        d["detj"] = np.array([[r.det_jac(e, p) for p in range(r.nl)] for e in els])
        d["genvel"] = r.genvel(inp["psi"])
        dt = 0.9
        d["dt"] = np.array(dt)
        qsl, dep, xy = r.advect_first(inp["q"], 1, inp["vel"], dt)
        d["adv1_q"], d["adv1_dep"], d["adv1_xy"] = qsl, dep, xy
        d["adv1_vel"] = r.advect_first(inp["vel"], 2, inp["vel"], dt, want_dep=False)[0]
        d["adv2_q"] = r.advect_second(inp["q"], inp["qprev"], 1, inp["vel"], inp["velprev"], dt)
        d["interp"] = np.array([r.interp_global(inp["field"], 1, p)[0] for p in inp["pts"]])
        d["g2l_el"] = np.array([r.global_to_local(p)[0] for p in inp["pts"]])
        d["g2l_l"] = np.array([r.global_to_local(p)[1] for p in inp["pts"]])
        np.savez_compressed(os.path.join(HERE, "mesh_%s.npz" % name), **d)
        r.close()
        print("wrote", name)


if __name__ == "__main__":
    main()
