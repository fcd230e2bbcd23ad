/*
 * sopht_amd_h5.h — C ABI of the checkpoint/restart file layer (host only, no GPU).
 *
 * SophT saves and restores simulations through h5py (sopht/utils/io.py:205-427): one HDF5
 * file per save, groups /Eulerian/{Scalar,Vector,Parameters} and /Lagrangian/<grid>/...,
 * a root attribute "time". This library exposes the handful of HDF5 operations that layout
 * needs (groups, contiguous datasets, attributes, a listing of every object path) over the
 * HDF5 1.10 C library, so that sopht_amd.utils.IO writes and reads the same files as the
 * reference's IO class. It is a separate shared object (libsopht_amd_h5.so) so the GPU
 * library does not depend on libhdf5.
 *
 * Conventions
 *   - file: an HDF5 file id returned by sopht_h5_file_create / sopht_h5_file_open.
 *   - dtype codes: SOPHT_H5_F32, SOPHT_H5_F64, SOPHT_H5_I32, SOPHT_H5_I64. Datasets and
 *     attributes are stored little-endian in the type given (what h5py stores for a numpy
 *     array of that dtype); reads convert to the requested memory type.
 *   - rank 0 means a scalar dataspace (h5py's scalar attribute).
 *   - Return value: 0 on success, a negative SOPHT_H5_ERR_* code otherwise;
 *     sopht_h5_last_error() describes the last failure of the calling thread.
 */
#ifndef SOPHT_AMD_H5_H
#define SOPHT_AMD_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SOPHT_H5_F32 = 0, SOPHT_H5_F64 = 1, SOPHT_H5_I32 = 2, SOPHT_H5_I64 = 3 };
enum { SOPHT_H5_MAX_RANK = 8 };

enum {
  SOPHT_H5_OK = 0,
  SOPHT_H5_ERR_INVALID = -1,   /* bad argument (ValueError)                      */
  SOPHT_H5_ERR_NOT_FOUND = -2, /* file or object missing (FileNotFoundError/KeyError) */
  SOPHT_H5_ERR_HDF5 = -3       /* the HDF5 library failed (OSError)              */
};

const char* sopht_h5_last_error(void);
/* HDF5 library version as major*10000 + minor*100 + release. */
int sopht_h5_library_version(void);

/* h5py.File(name, "w") — io.py:218 (truncates an existing file). */
int sopht_h5_file_create(const char* path, int64_t* file);
/* h5py.File(name, "r") — io.py:342. */
int sopht_h5_file_open(const char* path, int64_t* file);
/* leaving the `with h5py.File(...)` block — io.py:218, 342. */
int sopht_h5_file_close(int64_t file);

/* Group.create_group — io.py:224-226, 263, 275-282. `path` is absolute ("/Eulerian"). */
int sopht_h5_group_create(int64_t file, const char* path);

/* Group.create_dataset(name, data=array) — io.py:240-252, 278-280, 289-296. A contiguous
 * dataset of `rank` dims `dims` (C order) holding `data`. */
int sopht_h5_dataset_write(int64_t file, const char* path, int dtype, int rank, const int64_t* dims,
                           const void* data);
/* Dataset.shape and the stored dtype code — io.py:354-365, 390-416. */
int sopht_h5_dataset_shape(int64_t file, const char* path, int* rank, int64_t* dims, int* dtype);
/* Dataset[...] into `out` (count elements, converted to `dtype`) — io.py:356-366, 390-416. */
int sopht_h5_dataset_read(int64_t file, const char* path, int dtype, int64_t count, void* out);

/* obj.attrs[name] = array — io.py:221, 263-266. Replaces an existing attribute. */
int sopht_h5_attr_write(int64_t file, const char* obj, const char* name, int dtype, int rank,
                        const int64_t* dims, const void* data);
/* shape and stored dtype code of obj.attrs[name] — io.py:345, 370-385. */
int sopht_h5_attr_shape(int64_t file, const char* obj, const char* name, int* rank, int64_t* dims,
                        int* dtype);
/* obj.attrs[name] into `out` (count elements, converted to `dtype`) — io.py:345, 370-385. */
int sopht_h5_attr_read(int64_t file, const char* obj, const char* name, int dtype, int64_t count,
                       void* out);

/* File.visit(list.append) — io.py:343-344: every group and dataset path below the root,
 * relative ("Eulerian/Scalar/f"), one per line, in increasing name order. Writes at most
 * cap bytes (NUL-terminated) and sets *needed to the full length + 1. */
int sopht_h5_list(int64_t file, char* buf, int64_t cap, int64_t* needed);

#ifdef __cplusplus
}
#endif

#endif /* SOPHT_AMD_H5_H */
