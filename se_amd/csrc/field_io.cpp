// field_io.cpp -- field files in the reference's HDF5 layout (host code,
// libse_amd_io.so; HDF5 1.10 C API).
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
#include <string>

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
