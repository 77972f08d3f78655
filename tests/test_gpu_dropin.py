"""Drop-in HelmholtzEqn / PoissonEqn / Advector driven through the reference's
own Mesh, BCs and Field objects (tests/dropin/dropin_check.cpp), checked
against the oracle.  The program is built by tests/dropin/build.sh where the
reference tree is present (it links the reference's compiled base classes);
elsewhere the prebuilt binary travels with the tree.  A missing binary is a
failure, not a skip: this is the one check of the boundary through the
reference's own classes."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "dropin_check")


def test_dropin_classes_against_oracle():
    assert os.path.exists(BIN), ("%s is missing: build it with __graft_entry__.build() where the reference "
                                 "tree is present (tests/dropin/build.sh)" % BIN)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "dropin ok" in r.stdout
