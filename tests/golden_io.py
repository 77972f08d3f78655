"""Loading helpers for the committed golden vectors (tests/golden/*.npz)."""
import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def mesh_cases():
    return sorted(os.path.basename(p)[5:-4] for p in glob.glob(os.path.join(GOLDEN, "mesh_*.npz")))


def load_mesh_case(name):
    with np.load(os.path.join(GOLDEN, "mesh_%s.npz" % name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def load_gll():
    with np.load(os.path.join(GOLDEN, "gll.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def mesh_args(d):
    return (tuple(int(v) for v in d["nEls"]), int(d["N"]), tuple(float(v) for v in d["lo"]),
            tuple(float(v) for v in d["hi"]), tuple(int(v) for v in d["periodic"]),
            tuple(int(v) for v in d["bc"]))
