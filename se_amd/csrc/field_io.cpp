// field_io.cpp -- field and mesh files in the reference's HDF5 layout and the
// XDMF descriptions that index them (host code, libse_amd_io.so; HDF5 1.10 C
// API).
//
// Field::Save (src/base/Field.cpp:256-329) writes "<name>.<step:05u>.h5" with
// two integer attributes on the root group, "dimensions" (2) and
// "mesh resolution" (nVerts - 1 per direction), stored as big-endian 32-bit
// integers, and the dataset "/data" of native doubles, nVertsTotal x nDofs,
// node-major; Field::Read (:331-368) reads "/data" back.  Here the dataset
// is written and read in one call instead of one hyperslab per node: the
// file holds the same attributes, shape and values.
#include <hdf5.h>

#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "se_amd_io.h"

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
    g_err = msg;
    return SE_IO_ERR;
}

std::string file_name(const char* name, unsigned step) {
    char num[16];
    std::snprintf(num, sizeof(num), "%.5u", step);
    return std::string(name) + "." + num + ".h5";
}

// closes an HDF5 handle on scope exit
struct H {
    hid_t id;
    herr_t (*close)(hid_t);
    H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H() {
        if (id >= 0) close(id);
    }
    H(const H&) = delete;
    H& operator=(const H&) = delete;
};

int write_attr(hid_t group, const char* name, const int* vals, hsize_t n) {
    H space(H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (space.id < 0) return fail(std::string("H5Screate_simple for attribute ") + name);
    H attr(H5Acreate2(group, name, H5T_STD_I32BE, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (attr.id < 0) return fail(std::string("H5Acreate2 ") + name);
    if (H5Awrite(attr.id, H5T_NATIVE_INT, vals) < 0) return fail(std::string("H5Awrite ") + name);
    return SE_IO_OK;
}

int read_attr(hid_t group, const char* name, int* vals, hsize_t n) {
    if (H5Aexists(group, name) <= 0) return fail(std::string("attribute ") + name + " missing");
    H attr(H5Aopen(group, name, H5P_DEFAULT), H5Aclose);
    if (attr.id < 0) return fail(std::string("H5Aopen ") + name);
    H space(H5Aget_space(attr.id), H5Sclose);
    if (H5Sget_simple_extent_npoints(space.id) != (hssize_t)n) return fail(std::string("attribute ") + name + " size");
    if (H5Aread(attr.id, H5T_NATIVE_INT, vals) < 0) return fail(std::string("H5Aread ") + name);
    return SE_IO_OK;
}

struct Quiet {  // no HDF5 error-stack printing; errors come back as return codes
    Quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};

}  // namespace

extern "C" const char* se_io_last_error(void) { return g_err.c_str(); }

extern "C" int se_field_save(const char* name, unsigned step, const double* vals, int nVertsTotal, int nDofs,
                             const int nVerts[2]) {
    Quiet q;
    if (!name || !vals || !nVerts || nVertsTotal <= 0 || nDofs <= 0) return fail("se_field_save: bad argument");
    if ((long long)nVerts[0] * nVerts[1] != nVertsTotal) return fail("se_field_save: nVerts do not match nVertsTotal");
    const std::string fn = file_name(name, step);
    H file(H5Fcreate(fn.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail("se_field_save: cannot create " + fn);
    {
        H root(H5Gopen2(file.id, "/", H5P_DEFAULT), H5Gclose);
        if (root.id < 0) return fail("se_field_save: H5Gopen2 /");
        const int dims = 2;
        const int res[2] = {nVerts[0] - 1, nVerts[1] - 1};
        if (write_attr(root.id, "dimensions", &dims, 1) || write_attr(root.id, "mesh resolution", res, 2))
            return SE_IO_ERR;
    }
    hsize_t size[2] = {(hsize_t)nVertsTotal, (hsize_t)nDofs};
    H space(H5Screate_simple(2, size, nullptr), H5Sclose);
    if (space.id < 0) return fail("se_field_save: H5Screate_simple");
    H data(H5Dcreate2(file.id, "/data", H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
           H5Dclose);
    if (data.id < 0) return fail("se_field_save: H5Dcreate2 /data");
    if (H5Dwrite(data.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, vals) < 0)
        return fail("se_field_save: H5Dwrite /data");
    return SE_IO_OK;
}

extern "C" int se_field_info(const char* name, unsigned step, int* nVertsTotal, int* nDofs, int* dimensions,
                             int resolution[2]) {
    Quiet q;
    if (!name) return fail("se_field_info: bad argument");
    const std::string fn = file_name(name, step);
    H file(H5Fopen(fn.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail("se_field_info: cannot open " + fn);
    H data(H5Dopen2(file.id, "/data", H5P_DEFAULT), H5Dclose);
    if (data.id < 0) return fail("se_field_info: no /data in " + fn);
    H space(H5Dget_space(data.id), H5Sclose);
    hsize_t dims[2] = {0, 0};
    if (H5Sget_simple_extent_ndims(space.id) != 2 || H5Sget_simple_extent_dims(space.id, dims, nullptr) != 2)
        return fail("se_field_info: /data is not two-dimensional");
    if (nVertsTotal) *nVertsTotal = (int)dims[0];
    if (nDofs) *nDofs = (int)dims[1];
    H root(H5Gopen2(file.id, "/", H5P_DEFAULT), H5Gclose);
    if (dimensions && read_attr(root.id, "dimensions", dimensions, 1)) return SE_IO_ERR;
    if (resolution && read_attr(root.id, "mesh resolution", resolution, 2)) return SE_IO_ERR;
    return SE_IO_OK;
}

extern "C" int se_field_read(const char* name, unsigned step, double* vals, int nVertsTotal, int nDofs) {
    Quiet q;
    if (!name || !vals || nVertsTotal <= 0 || nDofs <= 0) return fail("se_field_read: bad argument");
    const std::string fn = file_name(name, step);
    H file(H5Fopen(fn.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail("se_field_read: cannot open " + fn);
    H data(H5Dopen2(file.id, "/data", H5P_DEFAULT), H5Dclose);
    if (data.id < 0) return fail("se_field_read: no /data in " + fn);
    H space(H5Dget_space(data.id), H5Sclose);
    hsize_t dims[2] = {0, 0};
    if (H5Sget_simple_extent_ndims(space.id) != 2 || H5Sget_simple_extent_dims(space.id, dims, nullptr) != 2 ||
        dims[0] != (hsize_t)nVertsTotal || dims[1] != (hsize_t)nDofs)
        return fail("se_field_read: /data of " + fn + " is not " + std::to_string(nVertsTotal) + " x " +
                    std::to_string(nDofs));
    if (H5Dread(data.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, vals) < 0)
        return fail("se_field_read: H5Dread /data");
    return SE_IO_OK;
}

// ---------------------------------------------------------------------------
// mesh files and XDMF descriptions
// ---------------------------------------------------------------------------
namespace {

int write_doubles(hid_t file, const char* path, const double* vals, int rank, const hsize_t* dims) {
    H space(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    if (space.id < 0) return fail(std::string("H5Screate_simple for ") + path);
    H data(H5Dcreate2(file, path, H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (data.id < 0) return fail(std::string("H5Dcreate2 ") + path);
    if (H5Dwrite(data.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, vals) < 0)
        return fail(std::string("H5Dwrite ") + path);
    return SE_IO_OK;
}

// reads dataset `path` of `count` elements of memory type `type` into out
int read_all(hid_t file, const char* path, hid_t type, void* out, hsize_t count) {
    H data(H5Dopen2(file, path, H5P_DEFAULT), H5Dclose);
    if (data.id < 0) return fail(std::string("no ") + path);
    H space(H5Dget_space(data.id), H5Sclose);
    if ((hsize_t)H5Sget_simple_extent_npoints(space.id) != count) return fail(std::string("size of ") + path);
    if (H5Dread(data.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) return fail(std::string("H5Dread ") + path);
    return SE_IO_OK;
}

std::string step_str(unsigned step) {
    char num[16];
    std::snprintf(num, sizeof(num), "%.5u", step);
    return num;
}

std::string xmf_name(unsigned step) { return "XDMF." + step_str(step) + ".xmf"; }

const char* const XML_HEAD =
    "<?xml version=\"1.0\" ?>\n"
    "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n"
    "<Xdmf xmlns:xi=\"http://www.w3.org/2001/XInclude\" Version=\"2.0\">\n"
    "\n"
    "<Domain>\n"
    "\n";
const char* const XML_FOOT = "</Domain>\n\n</Xdmf>\n\n";

// the node-centred HDF5 hyperslab of column `col` of an (nV x width) dataset
void slab(std::ofstream& f, const std::string& pad, int nV, int col, int width, const std::string& src,
          const char* name) {
    f << pad << "<DataItem ItemType=\"HyperSlab\" Dimensions=\"" << nV << " 1\"" << (name ? " Name=\"" : " ")
      << (name ? name : "") << (name ? "\">\n" : ">\n");
    f << pad << "   <DataItem Dimensions=\"3 2\" Format=\"XML\"> 0 " << col << " 1 1 " << nV << " 1 </DataItem>\n";
    f << pad << "   <DataItem Format=\"HDF\" NumberType=\"Float\" Precision=\"8\" Dimensions=\"" << nV << " "
      << width << "\">" << src << "</DataItem>\n";
    f << pad << "</DataItem>\n";
}

}  // namespace

extern "C" int se_mesh_save(const char* name, const int nVerts[2], const double min[2], const double max[2],
                            const double* verts) {
    Quiet q;
    if (!name || !nVerts || !min || !max || !verts || nVerts[0] < 2 || nVerts[1] < 2)
        return fail("se_mesh_save: bad argument");
    const std::string fn = std::string(name) + ".h5";
    H file(H5Fcreate(fn.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail("se_mesh_save: cannot create " + fn);
    const int nq[2] = {nVerts[0] - 1, nVerts[1] - 1};
    {
        H root(H5Gopen2(file.id, "/", H5P_DEFAULT), H5Gclose);
        if (root.id < 0) return fail("se_mesh_save: H5Gopen2 /");
        const int dims = 2;
        if (write_attr(root.id, "dimensions", &dims, 1) || write_attr(root.id, "mesh resolution", nq, 2))
            return SE_IO_ERR;
    }
    const hsize_t two = 2;
    if (write_doubles(file.id, "/min", min, 1, &two) || write_doubles(file.id, "/max", max, 1, &two))
        return SE_IO_ERR;
    const hsize_t nV = (hsize_t)nVerts[0] * nVerts[1];
    const hsize_t vd[2] = {nV, 2};
    if (write_doubles(file.id, "/vertices", verts, 2, vd)) return SE_IO_ERR;
    // linear quadrilaterals over the vertex grid, counter-clockwise from the
    // lower left vertex
    const size_t nquads = (size_t)nq[0] * nq[1];
    std::vector<int> conn(4 * nquads);
    for (size_t e = 0; e < nquads; e++) {
        const int i = (int)(e % nq[0]), j = (int)(e / nq[0]);
        const int v0 = j * nVerts[0] + i;
        conn[4 * e + 0] = v0;
        conn[4 * e + 1] = v0 + 1;
        conn[4 * e + 2] = v0 + 1 + nVerts[0];
        conn[4 * e + 3] = v0 + nVerts[0];
    }
    const hsize_t cd[2] = {(hsize_t)nquads, 4};
    H space(H5Screate_simple(2, cd, nullptr), H5Sclose);
    H data(H5Dcreate2(file.id, "/connectivity", H5T_NATIVE_INT, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
           H5Dclose);
    if (space.id < 0 || data.id < 0) return fail("se_mesh_save: H5Dcreate2 /connectivity");
    if (H5Dwrite(data.id, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, conn.data()) < 0)
        return fail("se_mesh_save: H5Dwrite /connectivity");
    return SE_IO_OK;
}

extern "C" int se_mesh_read(const char* name, int nVerts[2], double min[2], double max[2], double* verts,
                            int* conn) {
    Quiet q;
    if (!name || !nVerts) return fail("se_mesh_read: bad argument");
    const std::string fn = std::string(name) + ".h5";
    H file(H5Fopen(fn.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail("se_mesh_read: cannot open " + fn);
    int res[2];
    {
        H root(H5Gopen2(file.id, "/", H5P_DEFAULT), H5Gclose);
        if (read_attr(root.id, "mesh resolution", res, 2)) return SE_IO_ERR;
    }
    nVerts[0] = res[0] + 1;
    nVerts[1] = res[1] + 1;
    const hsize_t nV = (hsize_t)nVerts[0] * nVerts[1], nquads = (hsize_t)res[0] * res[1];
    if (min && read_all(file.id, "/min", H5T_NATIVE_DOUBLE, min, 2)) return SE_IO_ERR;
    if (max && read_all(file.id, "/max", H5T_NATIVE_DOUBLE, max, 2)) return SE_IO_ERR;
    if (verts && read_all(file.id, "/vertices", H5T_NATIVE_DOUBLE, verts, 2 * nV)) return SE_IO_ERR;
    if (conn && read_all(file.id, "/connectivity", H5T_NATIVE_INT, conn, 4 * nquads)) return SE_IO_ERR;
    return SE_IO_OK;
}

extern "C" int se_xdmf_header(unsigned step) {
    std::ofstream f(xmf_name(step));
    if (!f) return fail("se_xdmf_header: cannot create " + xmf_name(step));
    f << XML_HEAD;
    return f ? SE_IO_OK : fail("se_xdmf_header: write failed");
}

extern "C" int se_xdmf_footer(unsigned step) {
    std::ofstream f(xmf_name(step), std::ios::app);
    if (!f) return fail("se_xdmf_footer: cannot open " + xmf_name(step));
    f << XML_FOOT;
    return f ? SE_IO_OK : fail("se_xdmf_footer: write failed");
}

extern "C" int se_xdmf_write(const char* mesh_name, const int nVerts[2], const se_xdmf_field* fields, int nFields,
                             unsigned step, double time, double dt) {
    if (!mesh_name || !nVerts || (nFields > 0 && !fields) || nFields < 0) return fail("se_xdmf_write: bad argument");
    const std::string ts = step_str(step);
    {
        std::ofstream f("time." + ts + ".ts");
        if (!f) return fail("se_xdmf_write: cannot create time." + ts + ".ts");
        f << "time step:\t" << step << "\ttime:\t" << time << "\tdt:\t" << dt << "\n";
    }
    const int nV = nVerts[0] * nVerts[1];
    for (int k = 0; k < nFields; k++)
        if (se_field_save(fields[k].name, step, fields[k].vals, nV, fields[k].nDofs, nVerts)) return SE_IO_ERR;
    std::ofstream f(xmf_name(step), std::ios::app);
    if (!f) return fail("se_xdmf_write: cannot open " + xmf_name(step));
    const std::string mesh(mesh_name);
    const int nquads = (nVerts[0] - 1) * (nVerts[1] - 1);
    f << "   <Grid Name=\"FEM_Grid_" << mesh << "\">\n\n";
    f << "      <Time Value=\"" << time << "\" />\n\n";
    f << "         <Topology Type=\"Quadrilateral\" NumberOfElements=\"" << nquads << "\"> \n";
    f << "            <DataItem Format=\"HDF\" DataType=\"Int\"  Dimensions=\"" << nquads << " 4\">" << mesh
      << ".h5:/connectivity</DataItem>\n";
    f << "         </Topology>\n\n";
    f << "         <Geometry Type=\"XYZ\">\n";
    f << "            <DataItem ItemType=\"Function\"  Dimensions=\"" << nV << " 3\" Function=\"JOIN($0, $1, 0*$1)\">\n";
    slab(f, "               ", nV, 0, 2, mesh + ".h5:/vertices", "XCoords");
    slab(f, "               ", nV, 1, 2, mesh + ".h5:/vertices", "YCoords");
    f << "            </DataItem>\n";
    f << "         </Geometry>\n\n";
    for (int k = 0; k < nFields; k++) {
        const std::string src = std::string(fields[k].name) + "." + ts + ".h5:/data";
        if (fields[k].nDofs == 1) {
            f << "         <Attribute Type=\"Scalar\" Center=\"Node\" Name=\"" << fields[k].name << "\">\n";
            slab(f, "            ", nV, 0, 1, src, nullptr);
            f << "         </Attribute>\n\n";
        } else if (fields[k].nDofs == 2) {
            f << "         <Attribute Type=\"Vector\" Center=\"Node\" Name=\"" << fields[k].name << "\">\n";
            f << "            <DataItem ItemType=\"Function\"  Dimensions=\"" << nV
              << " 3\" Function=\"JOIN($0, $1, 0*$1)\">\n";
            slab(f, "               ", nV, 0, 2, src, "XValue");
            slab(f, "               ", nV, 1, 2, src, "YValue");
            f << "            </DataItem>\n";
            f << "         </Attribute>\n\n";
        } else {
            // the field file is written; XDMF describes scalars and 2-vectors only
            std::fprintf(stderr, "se_xdmf_write: field %s has %d dofs, not described\n", fields[k].name,
                         fields[k].nDofs);
        }
    }
    f << "   </Grid>\n\n";
    return f ? SE_IO_OK : fail("se_xdmf_write: write failed");
}

extern "C" int se_xdmf_temporal(int nTimeSteps, int dumpEvery) {
    if (dumpEvery < 1 || nTimeSteps < 0) return fail("se_xdmf_temporal: bad argument");
    const char* const include = "   <xi:include href=\"XDMF.FilesField.xdmf\" xpointer=\"xpointer(//Xdmf/Grid)\"/>\n\n";
    for (const char* fn : {"XDMF.temporalAll.xmf", "XDMF.temporalFields.xmf"}) {
        std::ofstream f(fn);
        if (!f) return fail(std::string("se_xdmf_temporal: cannot create ") + fn);
        f << XML_HEAD << include << XML_FOOT;
        if (!f) return fail(std::string("se_xdmf_temporal: write failed: ") + fn);
    }
    std::ofstream f("XDMF.FilesField.xdmf");
    if (!f) return fail("se_xdmf_temporal: cannot create XDMF.FilesField.xdmf");
    f << "<?xml version=\"1.0\" ?>\n";
    f << "<Xdmf xmlns:xi=\"http://www.w3.org/2001/XInclude\" Version=\"2.0\">\n";
    f << "<Grid GridType=\"Collection\" CollectionType=\"Temporal\" Name=\"FEM_Mesh_Fields\">\n";
    for (int s = 0; s <= nTimeSteps; s += dumpEvery)
        f << "    <xi:include href=\"" << xmf_name((unsigned)s) << "\" xpointer=\"xpointer(//Xdmf/Domain/Grid[1])\"/>\n";
    f << "</Grid>\n";
    f << "</Xdmf>\n";
    return f ? SE_IO_OK : fail("se_xdmf_temporal: write failed");
}
