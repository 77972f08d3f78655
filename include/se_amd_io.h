/* se_amd_io.h -- field files in the reference's HDF5 layout (libse_amd_io.so,
 * host code linked with HDF5; separate from libse_amd.so so that the GPU
 * library needs no HDF5).
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

#ifdef __cplusplus
}
#endif
#endif
