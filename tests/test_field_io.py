"""Field files in the reference's HDF5 layout (libse_amd_io.so, host code):
Field::Save / Field::Read, src/base/Field.cpp:256-368.

The layout is pinned against the reference's own Field class: io_check (built
from the reference tree by tests/dropin/build.sh into oracle/_ref) has the
reference write a file that se_field_read reads and read a file that
se_field_save wrote, value for value, and checks the attributes.  Here:
the exports, a round trip and the shape check."""
import os
import re
import subprocess

import numpy as np
import pytest

import se_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IO_CHECK = os.path.join(ROOT, "oracle", "_ref", "io_check")

needs_io = pytest.mark.skipif(not os.path.exists(se_amd.IO_LIB_PATH), reason="libse_amd_io.so not built (no HDF5)")


@needs_io
def test_io_library_exports_every_declared_symbol():
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "se_amd_io.h")).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(se_[a-z0-9_]+)\s*\(", text)))
    assert len(names) == 4
    L = se_amd.io_lib()
    assert all(hasattr(L, n) for n in names)


@needs_io
def test_round_trip_and_attributes(tmp_path):
    nVerts = (17, 9)
    nV = nVerts[0] * nVerts[1]
    vals = np.random.default_rng(3).standard_normal((nV, 3))
    base = str(tmp_path / "state")
    se_amd.field_save(base, 42, vals, nVerts)
    assert os.path.exists(base + ".00042.h5")
    assert se_amd.field_info(base, 42) == (nV, 3, 2, (16, 8))
    back = se_amd.field_read(base, 42, nV, 3)
    assert np.array_equal(back, vals)
    with pytest.raises(se_amd.SEError):
        se_amd.field_read(base, 42, nV, 2)  # wrong shape
    with pytest.raises(se_amd.SEError):
        se_amd.field_read(base, 43, nV, 3)  # no such file


@needs_io
@pytest.mark.skipif(not os.path.exists(IO_CHECK), reason="reference build (oracle/_ref/io_check) absent")
def test_interop_with_reference_field(tmp_path):
    r = subprocess.run([IO_CHECK], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "io ok" in r.stdout
