/* hermespy_amd drop record/replay store: a small C ABI over HDF5 (libhdf5 1.10).
 *
 * Replaces the HDF5 backend of the reference's serialization
 * (hermespy/core/factory.py:838-1240, HDFSerializationProcess /
 * HDFDeserializationProcess, driven by Scenario.record / replay / stop,
 * hermespy/core/scenario.py:368-566).  The conventions kept from it: root
 * attributes "version" (UTF-8 string) and "timestamp" (float64), an optional
 * campaign group, an "arrays" group, scalars as attributes (float64 / int64 /
 * UTF-8 strings), complex numbers as the compound {"r", "i"} that h5py reads as
 * complex.  Drops are stored column-wise in datasets that grow along their first
 * axis (one row per drop) rather than as one group per drop, so that whole
 * launches of drops are written with one call per column.
 *
 * Built as libhermespy_amd_h5.so (host code only).  Every function returns 0,
 * -1 for an invalid argument or -4 for a file / HDF5 failure;
 * hpx_h5_last_error() gives the message of the calling thread's last failure.
 */
#ifndef HERMESPY_AMD_H5_H
#define HERMESPY_AMD_H5_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types of attributes, datasets and buffers */
enum {
  HPX_H5_U8 = 0,
  HPX_H5_I8 = 1,
  HPX_H5_I32 = 2,
  HPX_H5_U32 = 3,
  HPX_H5_I64 = 4,
  HPX_H5_F32 = 5,
  HPX_H5_F64 = 6,
  HPX_H5_C64 = 7,  /* compound {r: float32, i: float32} */
  HPX_H5_C128 = 8, /* compound {r: float64, i: float64} */
  HPX_H5_STR = 9   /* variable-length UTF-8 string (attributes only) */
};

/* open modes */
enum { HPX_H5_READ = 0, HPX_H5_CREATE = 1, HPX_H5_TRUNCATE = 2, HPX_H5_APPEND = 3 };

#define HPX_H5_MAX_RANK 8

int hpx_h5_last_error(char* buffer, size_t size);

/* Opens (READ, APPEND) or creates (CREATE fails if the file exists, TRUNCATE
 * replaces it) a file; *handle receives the file handle. */
int hpx_h5_open(const char* path, int mode, int64_t* handle);
int hpx_h5_close(int64_t handle);

/* Creates a group and any missing parent groups (no error if it exists). */
int hpx_h5_create_group(int64_t handle, const char* path);

/* 1 if the link exists, 0 if not (parents included), negative on error. */
int hpx_h5_exists(int64_t handle, const char* path);

/* Attribute `name` of the object at `object` ("/" for the root): count == 1
 * writes a scalar, count > 1 a one-dimensional array; HPX_H5_STR writes the
 * NUL-terminated string `data`.  Replaces an existing attribute of that name
 * (factory.py:878-896). */
int hpx_h5_set_attr(int64_t handle, const char* object, const char* name, int dtype, int64_t count, const void* data);

/* Reads an attribute converted to dtype into data (capacity elements, or bytes
 * for HPX_H5_STR, which is NUL-terminated); *count receives the element count
 * (the string length for HPX_H5_STR).  Returns HPX_ERR_INVALID if it does not fit. */
int hpx_h5_get_attr(int64_t handle, const char* object, const char* name, int dtype, int64_t capacity, void* data,
                    int64_t* count);

/* Appends `rows` rows of shape row_shape[0 .. rank-1] (rank 0: scalars) to the
 * dataset at `path`, creating it (and its groups) on first use as a chunked
 * dataset of element type dtype whose first axis grows. */
int hpx_h5_append(int64_t handle, const char* path, int dtype, int rank, const int64_t* row_shape, int64_t rows,
                  const void* data);

/* Element type, rank and shape (first axis = rows) of the dataset at `path`. */
int hpx_h5_shape(int64_t handle, const char* path, int* dtype, int* rank, int64_t* shape);

/* Reads rows [row0, row0 + rows) of the dataset at `path` converted to dtype. */
int hpx_h5_read(int64_t handle, const char* path, int64_t row0, int64_t rows, int dtype, void* data);

/* Member names of the group at `path` in name order, each followed by '\n',
 * NUL-terminated; *size receives the bytes needed (including the NUL). */
int hpx_h5_list(int64_t handle, const char* path, char* names, int64_t capacity, int64_t* size);

#ifdef __cplusplus
}
#endif

#endif
