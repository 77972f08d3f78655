/* se_amd_io.h -- field and mesh files in the reference's HDF5 layout and the
 * XDMF files that describe them (libse_amd_io.so, host code linked with HDF5;
 * separate from libse_amd.so so that the GPU library needs no HDF5).
 *
 * A field is nVertsTotal x nDofs doubles, node-major (Field::vals flattened);
 * the file is "<name>.<step:05u>.h5" (name may carry a directory), as the
 * reference names it.  Device fields are copied to the host first. */
#ifndef SE_AMD_IO_H
#define SE_AMD_IO_H

#ifdef __cplusplus
extern "C" {
#endif

enum { SE_IO_OK = 0, SE_IO_ERR = 1 };

const char* se_io_last_error(void);

/* [ref: Field::Save src/base/Field.cpp:256-329] root attributes "dimensions"
 * (= 2) and "mesh resolution" (= nVerts - 1 per direction) as big-endian
 * int32, dataset "/data" (native double, nVertsTotal x nDofs). */
int se_field_save(const char* name, unsigned step, const double* vals, int nVertsTotal, int nDofs,
                  const int nVerts[2]);

/* [ref: Field::Read src/base/Field.cpp:331-368] reads "/data"; its shape must
 * be nVertsTotal x nDofs. */
int se_field_read(const char* name, unsigned step, double* vals, int nVertsTotal, int nDofs);

/* the shape of "/data" and the two attributes of a field file (nullable
 * outputs); no counterpart in the reference, for checks and tools */
int se_field_info(const char* name, unsigned step, int* nVertsTotal, int* nDofs, int* dimensions,
                  int resolution[2]);

/* [ref: Mesh::Save src/base/Mesh.cpp:366-492] "<name>.h5": root attributes
 * "dimensions" (2) and "mesh resolution" (nVerts - 1), datasets "/min" and
 * "/max" (2 doubles each), "/vertices" (nVertsTotal x 2 doubles, verts as
 * Mesh::verts) and "/connectivity" ((nVerts0-1)(nVerts1-1) x 4 native int: the
 * linear quadrilaterals of the vertex grid, counter-clockwise from the lower
 * left vertex). */
int se_mesh_save(const char* name, const int nVerts[2], const double min[2], const double max[2],
                 const double* verts);
/* what se_mesh_save / Mesh::Save wrote (nullable outputs: verts nVertsTotal x 2,
 * conn nQuads x 4); no counterpart in the reference, for checks and restart
 * tools */
int se_mesh_read(const char* name, int nVerts[2], double min[2], double max[2], double* verts, int* conn);

/* XDMF descriptions, in the working directory as the reference writes them
 * [ref: WriteXDMFHeader / WriteXDMFFooter src/base/Utils.cpp:53-89]: the
 * header and footer of "XDMF.<step:05u>.xmf" */
int se_xdmf_header(unsigned step);
int se_xdmf_footer(unsigned step);
/* [ref: WriteXDMF src/base/Utils.cpp:91-170]: "time.<step:05u>.ts", each
 * field's file (as se_field_save), and the step's Grid (mesh "<mesh_name>.h5",
 * scalar and 2-vector node attributes) appended to XDMF.<step:05u>.xmf */
typedef struct {
    const char* name;   /* the field's file name stem */
    int nDofs;          /* 1 or 2 are described; others are written only */
    const double* vals; /* nVertsTotal x nDofs, node-major */
} se_xdmf_field;
int se_xdmf_write(const char* mesh_name, const int nVerts[2], const se_xdmf_field* fields, int nFields,
                  unsigned step, double time, double dt);
/* [ref: WriteXDMFTemporal src/base/Utils.cpp:172-220]: XDMF.temporalAll.xmf,
 * XDMF.temporalFields.xmf and the temporal collection XDMF.FilesField.xdmf of
 * steps 0, dumpEvery, ... <= nTimeSteps */
int se_xdmf_temporal(int nTimeSteps, int dumpEvery);

#ifdef __cplusplus
}
#endif
#endif
