/*
 * Checkpoint/restart file layer over the HDF5 1.10 C library (host only).
 *
 * The reference reads and writes its snapshot files with h5py (sopht/utils/io.py:205-427).
 * h5py's defaults for what that code does are: contiguous datasets without filters, in the
 * numpy array's own little-endian type; attributes with a simple (or, for Python floats, a
 * scalar) dataspace; groups without tracked creation order. The calls below make those same
 * objects, so a file written here is laid out as the reference's and the reverse.
 */
#include "../../include/sopht_amd_h5.h"

#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static _Thread_local char g_err[512];

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

const char* sopht_h5_last_error(void) { return g_err; }

int sopht_h5_library_version(void) {
  unsigned maj = 0, min = 0, rel = 0;
  if (H5get_libversion(&maj, &min, &rel) < 0) return 0;
  return (int)(maj * 10000 + min * 100 + rel);
}

static hid_t mem_type(int dtype) {
  switch (dtype) {
    case SOPHT_H5_F32: return H5T_NATIVE_FLOAT;
    case SOPHT_H5_F64: return H5T_NATIVE_DOUBLE;
    case SOPHT_H5_I32: return H5T_NATIVE_INT32;
    case SOPHT_H5_I64: return H5T_NATIVE_INT64;
    default: return -1;
  }
}

static hid_t file_type(int dtype) {
  switch (dtype) {
    case SOPHT_H5_F32: return H5T_IEEE_F32LE;
    case SOPHT_H5_F64: return H5T_IEEE_F64LE;
    case SOPHT_H5_I32: return H5T_STD_I32LE;
    case SOPHT_H5_I64: return H5T_STD_I64LE;
    default: return -1;
  }
}

/* The dtype code of a stored type, or -1 for a type this layer does not map. */
static int code_of(hid_t type) {
  H5T_class_t cls = H5Tget_class(type);
  size_t size = H5Tget_size(type);
  if (cls == H5T_FLOAT) return size == 4 ? SOPHT_H5_F32 : size == 8 ? SOPHT_H5_F64 : -1;
  if (cls == H5T_INTEGER && H5Tget_sign(type) == H5T_SGN_2) {
    return size == 4 ? SOPHT_H5_I32 : size == 8 ? SOPHT_H5_I64 : -1;
  }
  return -1;
}

/* Whether every prefix of an absolute path names a link (H5Lexists needs the parents):
   1 yes, 0 no, -1 out of host memory. Paths of any length (the copy is on the heap). */
static int path_exists(hid_t file, const char* path) {
  size_t n = strlen(path);
  if (n == 0) return 0;
  if (strcmp(path, "/") == 0) return 1;
  char* buf = (char*)malloc(n + 1);
  if (!buf) return -1;
  memcpy(buf, path, n + 1);
  int found = 1;
  for (size_t i = 1; i <= n && found; ++i) {
    if (buf[i] == '/' || buf[i] == '\0') {
      char c = buf[i];
      buf[i] = '\0';
      htri_t e = H5Lexists(file, buf, H5P_DEFAULT);
      buf[i] = c;
      if (e <= 0) found = 0;
    }
  }
  free(buf);
  return found;
}

/* SOPHT_H5_OK when the object at path exists, else the recorded error. */
static int require_path(hid_t file, const char* path) {
  int e = path_exists(file, path);
  if (e < 0) return fail(SOPHT_H5_ERR_HDF5, "out of host memory looking up '%s'", path);
  if (e == 0) return fail(SOPHT_H5_ERR_NOT_FOUND, "no object '%s'", path);
  return SOPHT_H5_OK;
}

static hid_t make_space(int rank, const int64_t* dims) {
  if (rank == 0) return H5Screate(H5S_SCALAR);
  hsize_t d[SOPHT_H5_MAX_RANK];
  for (int i = 0; i < rank; ++i) d[i] = (hsize_t)dims[i];
  return H5Screate_simple(rank, d, NULL);
}

static int check_shape_args(int rank, const int64_t* dims) {
  if (rank < 0 || rank > SOPHT_H5_MAX_RANK || (rank > 0 && !dims)) return 0;
  for (int i = 0; i < rank; ++i)
    if (dims[i] < 0) return 0;
  return 1;
}

/* Rank, dims and dtype code of a dataspace + type pair. */
static int describe(hid_t space, hid_t type, int* rank, int64_t* dims, int* dtype) {
  int r = H5Sget_simple_extent_ndims(space);
  if (r < 0 || r > SOPHT_H5_MAX_RANK) return fail(SOPHT_H5_ERR_HDF5, "unsupported dataspace rank %d", r);
  hsize_t d[SOPHT_H5_MAX_RANK];
  if (r > 0 && H5Sget_simple_extent_dims(space, d, NULL) < 0) return fail(SOPHT_H5_ERR_HDF5, "H5Sget_simple_extent_dims failed");
  *rank = r;
  for (int i = 0; i < r; ++i) dims[i] = (int64_t)d[i];
  *dtype = code_of(type);
  return SOPHT_H5_OK;
}

int sopht_h5_file_create(const char* path, int64_t* file) {
  quiet();
  if (!path || !file) return fail(SOPHT_H5_ERR_INVALID, "null argument");
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return fail(SOPHT_H5_ERR_HDF5, "unable to create HDF5 file '%s'", path);
  *file = (int64_t)f;
  return SOPHT_H5_OK;
}

int sopht_h5_file_open(const char* path, int64_t* file) {
  quiet();
  if (!path || !file) return fail(SOPHT_H5_ERR_INVALID, "null argument");
  if (access(path, R_OK) != 0) return fail(SOPHT_H5_ERR_NOT_FOUND, "unable to open file '%s'", path);
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return fail(SOPHT_H5_ERR_HDF5, "'%s' is not a readable HDF5 file", path);
  *file = (int64_t)f;
  return SOPHT_H5_OK;
}

int sopht_h5_file_close(int64_t file) {
  quiet();
  if (H5Fclose((hid_t)file) < 0) return fail(SOPHT_H5_ERR_HDF5, "H5Fclose failed");
  return SOPHT_H5_OK;
}

int sopht_h5_group_create(int64_t file, const char* path) {
  quiet();
  if (!path || path[0] != '/') return fail(SOPHT_H5_ERR_INVALID, "group path must be absolute");
  hid_t g = H5Gcreate2((hid_t)file, path, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) return fail(SOPHT_H5_ERR_HDF5, "unable to create group '%s'", path);
  H5Gclose(g);
  return SOPHT_H5_OK;
}

int sopht_h5_dataset_write(int64_t file, const char* path, int dtype, int rank, const int64_t* dims,
                           const void* data) {
  quiet();
  if (!path || path[0] != '/') return fail(SOPHT_H5_ERR_INVALID, "dataset path must be absolute");
  if (mem_type(dtype) < 0) return fail(SOPHT_H5_ERR_INVALID, "unsupported dtype code %d", dtype);
  if (!check_shape_args(rank, dims)) return fail(SOPHT_H5_ERR_INVALID, "bad dataset shape");
  hid_t space = make_space(rank, dims);
  if (space < 0) return fail(SOPHT_H5_ERR_HDF5, "unable to make a dataspace for '%s'", path);
  hid_t d = H5Dcreate2((hid_t)file, path, file_type(dtype), space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  H5Sclose(space);
  if (d < 0) return fail(SOPHT_H5_ERR_HDF5, "unable to create dataset '%s'", path);
  int64_t count = 1;
  for (int i = 0; i < rank; ++i) count *= dims[i];
  herr_t rc = 0;
  if (count > 0) {
    if (!data) {
      H5Dclose(d);
      return fail(SOPHT_H5_ERR_INVALID, "null data for dataset '%s'", path);
    }
    rc = H5Dwrite(d, mem_type(dtype), H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  }
  H5Dclose(d);
  if (rc < 0) return fail(SOPHT_H5_ERR_HDF5, "unable to write dataset '%s'", path);
  return SOPHT_H5_OK;
}

int sopht_h5_dataset_shape(int64_t file, const char* path, int* rank, int64_t* dims, int* dtype) {
  quiet();
  if (!path || !rank || !dims || !dtype) return fail(SOPHT_H5_ERR_INVALID, "null argument");
  int found = require_path((hid_t)file, path);
  if (found != SOPHT_H5_OK) return found;
  hid_t d = H5Dopen2((hid_t)file, path, H5P_DEFAULT);
  if (d < 0) return fail(SOPHT_H5_ERR_HDF5, "'%s' is not a dataset", path);
  hid_t space = H5Dget_space(d), type = H5Dget_type(d);
  int rc = describe(space, type, rank, dims, dtype);
  H5Tclose(type);
  H5Sclose(space);
  H5Dclose(d);
  return rc;
}

int sopht_h5_dataset_read(int64_t file, const char* path, int dtype, int64_t count, void* out) {
  quiet();
  if (!path || (count > 0 && !out)) return fail(SOPHT_H5_ERR_INVALID, "null argument");
  if (mem_type(dtype) < 0) return fail(SOPHT_H5_ERR_INVALID, "unsupported dtype code %d", dtype);
  int found = require_path((hid_t)file, path);
  if (found != SOPHT_H5_OK) return found;
  hid_t d = H5Dopen2((hid_t)file, path, H5P_DEFAULT);
  if (d < 0) return fail(SOPHT_H5_ERR_HDF5, "'%s' is not a dataset", path);
  hid_t space = H5Dget_space(d);
  hssize_t n = H5Sget_simple_extent_npoints(space);
  H5Sclose(space);
  if (n != (hssize_t)count) {
    H5Dclose(d);
    return fail(SOPHT_H5_ERR_INVALID, "dataset '%s' holds %lld elements, buffer %lld", path, (long long)n,
                (long long)count);
  }
  herr_t rc = count > 0 ? H5Dread(d, mem_type(dtype), H5S_ALL, H5S_ALL, H5P_DEFAULT, out) : 0;
  H5Dclose(d);
  if (rc < 0) return fail(SOPHT_H5_ERR_HDF5, "unable to read dataset '%s'", path);
  return SOPHT_H5_OK;
}

int sopht_h5_attr_write(int64_t file, const char* obj, const char* name, int dtype, int rank,
                        const int64_t* dims, const void* data) {
  quiet();
  if (!obj || !name || !data) return fail(SOPHT_H5_ERR_INVALID, "null argument");
  if (mem_type(dtype) < 0) return fail(SOPHT_H5_ERR_INVALID, "unsupported dtype code %d", dtype);
  if (!check_shape_args(rank, dims)) return fail(SOPHT_H5_ERR_INVALID, "bad attribute shape");
  int found = require_path((hid_t)file, obj);
  if (found != SOPHT_H5_OK) return found;
  hid_t f = (hid_t)file;
  if (H5Aexists_by_name(f, obj, name, H5P_DEFAULT) > 0 && H5Adelete_by_name(f, obj, name, H5P_DEFAULT) < 0)
    return fail(SOPHT_H5_ERR_HDF5, "unable to replace attribute '%s' of '%s'", name, obj);
  hid_t space = make_space(rank, dims);
  if (space < 0) return fail(SOPHT_H5_ERR_HDF5, "unable to make a dataspace for attribute '%s'", name);
  hid_t a = H5Acreate_by_name(f, obj, name, file_type(dtype), space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  H5Sclose(space);
  if (a < 0) return fail(SOPHT_H5_ERR_HDF5, "unable to create attribute '%s' of '%s'", name, obj);
  herr_t rc = H5Awrite(a, mem_type(dtype), data);
  H5Aclose(a);
  if (rc < 0) return fail(SOPHT_H5_ERR_HDF5, "unable to write attribute '%s' of '%s'", name, obj);
  return SOPHT_H5_OK;
}

/* Opens obj.attrs[name]; on failure returns -1 with *code set and the error recorded. */
static hid_t open_attr(hid_t f, const char* obj, const char* name, int* code) {
  int found = require_path(f, obj);
  if (found != SOPHT_H5_OK) return *code = found, -1;
  if (H5Aexists_by_name(f, obj, name, H5P_DEFAULT) <= 0)
    return *code = fail(SOPHT_H5_ERR_NOT_FOUND, "no attribute '%s' on '%s'", name, obj), -1;
  hid_t a = H5Aopen_by_name(f, obj, name, H5P_DEFAULT, H5P_DEFAULT);
  if (a < 0) *code = fail(SOPHT_H5_ERR_HDF5, "unable to open attribute '%s' of '%s'", name, obj);
  return a;
}

int sopht_h5_attr_shape(int64_t file, const char* obj, const char* name, int* rank, int64_t* dims,
                        int* dtype) {
  quiet();
  if (!obj || !name || !rank || !dims || !dtype) return fail(SOPHT_H5_ERR_INVALID, "null argument");
  int code = SOPHT_H5_OK;
  hid_t a = open_attr((hid_t)file, obj, name, &code);
  if (a < 0) return code;
  hid_t space = H5Aget_space(a), type = H5Aget_type(a);
  int rc = describe(space, type, rank, dims, dtype);
  H5Tclose(type);
  H5Sclose(space);
  H5Aclose(a);
  return rc;
}

int sopht_h5_attr_read(int64_t file, const char* obj, const char* name, int dtype, int64_t count,
                       void* out) {
  quiet();
  if (!obj || !name || !out) return fail(SOPHT_H5_ERR_INVALID, "null argument");
  if (mem_type(dtype) < 0) return fail(SOPHT_H5_ERR_INVALID, "unsupported dtype code %d", dtype);
  int code = SOPHT_H5_OK;
  hid_t a = open_attr((hid_t)file, obj, name, &code);
  if (a < 0) return code;
  hid_t space = H5Aget_space(a);
  hssize_t n = H5Sget_simple_extent_npoints(space);
  H5Sclose(space);
  if (n != (hssize_t)count) {
    H5Aclose(a);
    return fail(SOPHT_H5_ERR_INVALID, "attribute '%s' holds %lld elements, buffer %lld", name, (long long)n,
                (long long)count);
  }
  herr_t rc = H5Aread(a, mem_type(dtype), out);
  H5Aclose(a);
  if (rc < 0) return fail(SOPHT_H5_ERR_HDF5, "unable to read attribute '%s' of '%s'", name, obj);
  return SOPHT_H5_OK;
}

struct listing {
  char* buf;
  int64_t cap;
  int64_t len; /* bytes the full listing needs, without the terminating NUL */
  int full;    /* a name did not fit: nothing more is written */
};

static herr_t list_one(hid_t group, const char* name, const H5L_info_t* info, void* op) {
  (void)group;
  (void)info;
  struct listing* L = (struct listing*)op;
  int64_t n = (int64_t)strlen(name);
  if (L->buf && !L->full && L->len + n + 1 < L->cap) {
    memcpy(L->buf + L->len, name, (size_t)n);
    L->buf[L->len + n] = '\n';
  } else {
    L->full = 1;
  }
  L->len += n + 1;
  return 0;
}

int sopht_h5_list(int64_t file, char* buf, int64_t cap, int64_t* needed) {
  quiet();
  if (!needed || cap < 0 || (cap > 0 && !buf)) return fail(SOPHT_H5_ERR_INVALID, "bad listing buffer");
  struct listing L = {buf, cap, 0, 0};
  if (H5Lvisit((hid_t)file, H5_INDEX_NAME, H5_ITER_INC, list_one, &L) < 0)
    return fail(SOPHT_H5_ERR_HDF5, "H5Lvisit failed");
  *needed = L.len + 1;
  if (buf && cap > 0) buf[L.full ? 0 : L.len] = '\0';
  return SOPHT_H5_OK;
}
