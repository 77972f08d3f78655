// io_check.cpp -- field files written by the reference's own Field::Save are
// read by se_field_read, and files written by se_field_save are read by the
// reference's Field::Read (src/base/Field.cpp:256-368), value for value; the
// attributes and the dataset shape are those Field::Save writes.  The mesh
// file of se_mesh_save holds what the reference's Mesh::Save
// (src/base/Mesh.cpp:366-492) writes for the same mesh: attributes, /min,
// /max, /vertices and /connectivity, value for value.  Links the reference's
// src/base classes compiled in place (oracle/_ref/libsebase.so) and
// libse_amd_io.so.  Run in a scratch directory: writes four .h5 files.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "QuadPoint.h"
#include "LinAlg.h"
#include "Jacobi.h"
#include "Element.h"
#include "Mesh.h"
#include "BCs.h"
#include "Field.h"
#include "se_amd_io.h"
#include <hdf5.h>

static int bad(const char* what) {
    std::printf("io_check: FAILED %s (%s)\n", what, se_io_last_error());
    return 1;
}

int main() {
    int nEls[2] = {5, 3};
    double lo[2] = {0.0, -1.0}, hi[2] = {2.0, 1.5};
    bool per[2] = {true, false};
    const int N = 4, nDofs = 2;
    Mesh* mesh = new Mesh("mesh", nEls, "legendre", N, lo, hi, per);
    const int nV = mesh->nVertsTotal;
    // 1. the reference writes, we read
    Field* ref = new Field("ref_field", mesh, nDofs, NULL);
    for (int v = 0; v < nV; v++)
        for (int d = 0; d < nDofs; d++) ref->vals[v][d] = std::sin(0.37 * v + 1.3 * d) * std::pow(10.0, (v % 7) - 3);
    ref->Save(7);
    int shape0 = 0, shape1 = 0, dims = 0, res[2] = {0, 0};
    if (se_field_info("ref_field", 7, &shape0, &shape1, &dims, res)) return bad("se_field_info of the reference's file");
    if (shape0 != nV || shape1 != nDofs || dims != 2 || res[0] != mesh->nVerts[0] - 1 || res[1] != mesh->nVerts[1] - 1)
        return bad("attributes / shape of the reference's file");
    std::vector<double> got((size_t)nV * nDofs);
    if (se_field_read("ref_field", 7, got.data(), nV, nDofs)) return bad("se_field_read of the reference's file");
    for (int v = 0; v < nV; v++)
        for (int d = 0; d < nDofs; d++)
            if (std::memcmp(&got[(size_t)v * nDofs + d], &ref->vals[v][d], sizeof(double)) != 0)
                return bad("values of the reference's file");
    // 2. we write, the reference reads
    std::vector<double> mine((size_t)nV * nDofs);
    for (size_t i = 0; i < mine.size(); i++) mine[i] = std::cos(0.11 * i) / (1.0 + i);
    const int nVerts[2] = {mesh->nVerts[0], mesh->nVerts[1]};
    if (se_field_save("mine", 12, mine.data(), nV, nDofs, nVerts)) return bad("se_field_save");
    Field* back = new Field("mine", mesh, nDofs, NULL);
    back->Read(12);
    for (int v = 0; v < nV; v++)
        for (int d = 0; d < nDofs; d++)
            if (std::memcmp(&mine[(size_t)v * nDofs + d], &back->vals[v][d], sizeof(double)) != 0)
                return bad("values read back by the reference");
    // 3. a wrong shape is refused
    std::vector<double> small((size_t)(nV - 1) * nDofs);
    if (se_field_read("mine", 12, small.data(), nV - 1, nDofs) == SE_IO_OK) return bad("shape check");
    // 4. mesh files: the reference's Mesh::Save and se_mesh_save of the same mesh
    mesh->Save();
    std::vector<double> verts((size_t)nV * 2);
    for (int v = 0; v < nV; v++) verts[2 * (size_t)v] = mesh->verts[v][0], verts[2 * (size_t)v + 1] = mesh->verts[v][1];
    if (se_mesh_save("mine_mesh", nVerts, mesh->min, mesh->max, verts.data())) return bad("se_mesh_save");
    const int nq = (nVerts[0] - 1) * (nVerts[1] - 1);
    std::vector<double> va((size_t)nV * 2), vb((size_t)nV * 2);
    std::vector<int> ca((size_t)nq * 4), cb((size_t)nq * 4);
    int na[2], nb[2];
    double mina[2], maxa[2], minb[2], maxb[2];
    if (se_mesh_read("mesh", na, mina, maxa, va.data(), ca.data())) return bad("se_mesh_read of the reference's mesh");
    if (se_mesh_read("mine_mesh", nb, minb, maxb, vb.data(), cb.data())) return bad("se_mesh_read of ours");
    if (na[0] != nVerts[0] || na[1] != nVerts[1] || nb[0] != na[0] || nb[1] != na[1]) return bad("mesh resolution");
    if (std::memcmp(mina, minb, sizeof(mina)) || std::memcmp(maxa, maxb, sizeof(maxa))) return bad("mesh min / max");
    if (std::memcmp(va.data(), vb.data(), va.size() * sizeof(double))) return bad("mesh vertices");
    if (std::memcmp(ca.data(), cb.data(), ca.size() * sizeof(int))) return bad("mesh connectivity");
    for (const char* fn : {"mesh.h5", "mine_mesh.h5"}) {  // the dimensions attribute and its file type
        hid_t f = H5Fopen(fn, H5F_ACC_RDONLY, H5P_DEFAULT);
        hid_t a = H5Aopen_by_name(f, "/", "dimensions", H5P_DEFAULT, H5P_DEFAULT);
        hid_t t = H5Aget_type(a);
        int d = 0;
        const bool ok = a >= 0 && H5Aread(a, H5T_NATIVE_INT, &d) >= 0 && d == 2 && H5Tequal(t, H5T_STD_I32BE) > 0;
        H5Tclose(t);
        H5Aclose(a);
        H5Fclose(f);
        if (!ok) return bad("mesh dimensions attribute");
    }
    std::printf("io ok: %d nodes x %d dofs both ways; mesh file of %d quads\n", nV, nDofs, nq);
    delete back;
    delete ref;
    delete mesh;
    return 0;
}
