"""Field and mesh files in the reference's HDF5 layout and their XDMF
descriptions (libse_amd_io.so, host code): Field::Save / Field::Read,
src/base/Field.cpp:256-368; Mesh::Save, src/base/Mesh.cpp:366-492; WriteXDMF*,
src/base/Utils.cpp:53-220.

The field and mesh layouts are pinned against the reference's own classes:
io_check (built from the reference tree by tests/dropin/build.sh into
oracle/_ref) has the reference write a field file that se_field_read reads and
read a file that se_field_save wrote, value for value, and compares the mesh
file of Mesh::Save with se_mesh_save's.  The XDMF writers live in the
reference's Utils.cpp, which includes PETSc and is not built here: their
output is checked for what the format fixes (well-formed XML whose every HDF
reference resolves to a dataset of the stated shape, whose hyperslabs select
the mesh coordinates and field components, with the step's time) -- parity
with the reference's exact text is unpinned.  Here also: the exports, a round
trip and the shape check."""
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
    assert len(names) == 10
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


def hyperslab(item, arrays):
    """Apply an XDMF HyperSlab DataItem (start, stride, count rows) to the array
    its HDF DataItem names; arrays maps "file.h5:/path" to numpy arrays."""
    sel, src = item.findall("DataItem")
    start, stride, count = np.array(sel.text.split(), dtype=int).reshape(3, 2)
    a = arrays[src.text.strip()]
    assert src.get("Dimensions") == "%d %d" % a.shape
    assert item.get("Dimensions") == "%d %d" % tuple(count)
    return a[start[0]:start[0] + count[0] * stride[0]:stride[0], start[1]:start[1] + count[1] * stride[1]:stride[1]]


@needs_io
def test_mesh_file_round_trip(tmp_path):
    nVerts = (13, 7)
    X, Y = np.linspace(-1.0, 2.0, nVerts[0]), np.linspace(0.0, 1.0, nVerts[1]) ** 2
    verts = np.stack(np.meshgrid(X, Y), -1).reshape(-1, 2)
    base = str(tmp_path / "grid")
    se_amd.mesh_save(base, nVerts, (-1.0, 0.0), (2.0, 1.0), verts)
    nv, lo, hi, v, conn = se_amd.mesh_read(base)
    assert nv == nVerts and np.array_equal(v, verts)
    assert np.array_equal(lo, [-1.0, 0.0]) and np.array_equal(hi, [2.0, 1.0])
    # every quad is a counter-clockwise unit cell of the vertex grid
    i, j = conn[:, 0] % nVerts[0], conn[:, 0] // nVerts[0]
    assert conn.shape == (12 * 6, 4) and np.all(i < nVerts[0] - 1) and np.all(j < nVerts[1] - 1)
    assert np.array_equal(conn[:, 1], conn[:, 0] + 1)
    assert np.array_equal(conn[:, 2], conn[:, 0] + 1 + nVerts[0])
    assert np.array_equal(conn[:, 3], conn[:, 0] + nVerts[0])
    area = 0.5 * sum((verts[conn[:, k], 0] * verts[conn[:, (k + 1) % 4], 1] -
                      verts[conn[:, (k + 1) % 4], 0] * verts[conn[:, k], 1]) for k in range(4))
    assert np.all(area > 0) and abs(area.sum() - 3.0) < 1e-12


@needs_io
def test_xdmf_description_resolves(tmp_path, monkeypatch):
    import xml.etree.ElementTree as ET
    monkeypatch.chdir(tmp_path)
    nVerts = (9, 5)
    nV = nVerts[0] * nVerts[1]
    verts = np.stack(np.meshgrid(np.linspace(0, 2, nVerts[0]), np.linspace(0, 1, nVerts[1])), -1).reshape(-1, 2)
    rng = np.random.default_rng(5)
    psi, vel, extra = rng.standard_normal(nV), rng.standard_normal((nV, 2)), rng.standard_normal((nV, 3))
    se_amd.mesh_save("mesh", nVerts, (0.0, 0.0), (2.0, 1.0), verts)
    for step, time in ((0, 0.0), (3, 0.25)):
        se_amd.xdmf_header(step)
        se_amd.xdmf_write("mesh", nVerts, [("psi", psi), ("vel", vel), ("extra", extra)], step, time, 0.01)
        se_amd.xdmf_footer(step)
    se_amd.xdmf_temporal(3, 3)
    assert open("time.00003.ts").read() == "time step:\t3\ttime:\t0.25\tdt:\t0.01\n"
    nv, lo, hi, v, conn = se_amd.mesh_read("mesh")
    arrays = {"mesh.h5:/vertices": v, "mesh.h5:/connectivity": conn,
              "psi.00003.h5:/data": se_amd.field_read("psi", 3, nV, 1),
              "vel.00003.h5:/data": se_amd.field_read("vel", 3, nV, 2)}
    assert np.array_equal(se_amd.field_read("extra", 3, nV, 3), extra)  # written, not described
    root = ET.parse("XDMF.00003.xmf").getroot()
    grids = root.findall("Domain/Grid")
    assert len(grids) == 1 and grids[0].get("Name") == "FEM_Grid_mesh"
    g = grids[0]
    assert float(g.find("Time").get("Value")) == 0.25
    topo = g.find("Topology")
    assert topo.get("Type") == "Quadrilateral" and int(topo.get("NumberOfElements")) == conn.shape[0]
    item = topo.find("DataItem")
    assert item.text.strip() == "mesh.h5:/connectivity" and item.get("Dimensions") == "%d 4" % conn.shape[0]
    geo = g.find("Geometry/DataItem")
    assert geo.get("Function") == "JOIN($0, $1, 0*$1)" and geo.get("Dimensions") == "%d 3" % nV
    xs, ys = geo.findall("DataItem")
    assert np.array_equal(hyperslab(xs, arrays)[:, 0], verts[:, 0])
    assert np.array_equal(hyperslab(ys, arrays)[:, 0], verts[:, 1])
    attrs = {a.get("Name"): a for a in g.findall("Attribute")}
    assert sorted(attrs) == ["psi", "vel"]
    assert attrs["psi"].get("Type") == "Scalar" and attrs["psi"].get("Center") == "Node"
    assert np.array_equal(hyperslab(attrs["psi"].find("DataItem"), arrays)[:, 0], psi)
    vx, vy = attrs["vel"].find("DataItem").findall("DataItem")
    assert attrs["vel"].get("Type") == "Vector"
    assert np.array_equal(hyperslab(vx, arrays)[:, 0], vel[:, 0])
    assert np.array_equal(hyperslab(vy, arrays)[:, 0], vel[:, 1])
    # the temporal collection includes steps 0 and 3, and the index files include it
    XI = "{http://www.w3.org/2001/XInclude}include"
    coll = ET.parse("XDMF.FilesField.xdmf").getroot().find("Grid")
    assert coll.get("CollectionType") == "Temporal"
    assert [e.get("href") for e in coll.findall(XI)] == ["XDMF.00000.xmf", "XDMF.00003.xmf"]
    for fn in ("XDMF.temporalAll.xmf", "XDMF.temporalFields.xmf"):
        inc = ET.parse(fn).getroot().findall("Domain/" + XI)
        assert [e.get("href") for e in inc] == ["XDMF.FilesField.xdmf"]
