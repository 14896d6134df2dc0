/* HDF5 file primitives of fortnet_amd (declared and mapped to the reference's HDF5 calls in
 * include/fnet_h5.h).  Plain C over the HDF5 C and high-level libraries; built with gcc into
 * libfnet_h5.so, separate from the device library. */
#include "fnet_h5.h"

#include <hdf5.h>
#include <hdf5_hl.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_err[512];

static int fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return -1;
}

/* HDF5 prints its error stack by default; the host layer reports errors itself. */
static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

const char* fnet_h5_last_error(void) { return g_err; }

long long fnet_h5_create(const char* path) {
  quiet();
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return fail("cannot create HDF5 file '%s'", path);
  return (long long)f;
}

long long fnet_h5_open(const char* path, int rdwr) {
  quiet();
  hid_t f = H5Fopen(path, rdwr ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return fail("cannot open HDF5 file '%s'", path);
  return (long long)f;
}

int fnet_h5_close(long long file) {
  quiet();
  if (H5Fclose((hid_t)file) < 0) return fail("cannot close HDF5 file");
  return 0;
}

int fnet_h5_group_create(long long file, const char* path) {
  quiet();
  hid_t g = H5Gcreate2((hid_t)file, path, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) return fail("cannot create group '%s'", path);
  H5Gclose(g);
  return 0;
}

int fnet_h5_exists(long long file, const char* path) {
  quiet();
  if (strcmp(path, ".") == 0 || strcmp(path, "/") == 0) return 1;
  htri_t r = H5LTpath_valid((hid_t)file, path, 0);
  if (r < 0) return fail("cannot look up '%s'", path);
  return r > 0 ? 1 : 0;
}

int fnet_h5_delete(long long file, const char* path) {
  quiet();
  if (H5Ldelete((hid_t)file, path, H5P_DEFAULT) < 0) return fail("cannot delete '%s'", path);
  return 0;
}

typedef struct {
  char* buf;
  long long cap, used;
  int count;
} ListState;

static herr_t list_one(hid_t g, const char* name, const H5L_info_t* info, void* data) {
  (void)g;
  (void)info;
  ListState* st = (ListState*)data;
  const long long len = (long long)strlen(name) + 1;
  if (st->buf && st->used + len <= st->cap) memcpy(st->buf + st->used, name, (size_t)len);
  st->used += len;
  st->count += 1;
  return 0;
}

int fnet_h5_list(long long file, const char* path, char* buf, long long cap) {
  quiet();
  hid_t g = H5Gopen2((hid_t)file, path, H5P_DEFAULT);
  if (g < 0) return fail("cannot open group '%s'", path);
  ListState st = {buf, cap, 0, 0};
  hsize_t idx = 0;
  herr_t r = H5Literate(g, H5_INDEX_NAME, H5_ITER_INC, &idx, list_one, &st);
  H5Gclose(g);
  if (r < 0) return fail("cannot list group '%s'", path);
  if (st.used > cap) return (int)(-st.used - 1);
  return st.count;
}

int fnet_h5_attr_exists(long long file, const char* obj, const char* name) {
  quiet();
  htri_t r = H5Aexists_by_name((hid_t)file, obj, name, H5P_DEFAULT);
  if (r < 0) return fail("cannot look up attribute '%s' of '%s'", name, obj);
  return r > 0 ? 1 : 0;
}

static int class_code(H5T_class_t c) {
  switch (c) {
    case H5T_INTEGER: return FNET_H5_INT;
    case H5T_FLOAT: return FNET_H5_FLOAT;
    case H5T_STRING: return FNET_H5_STRING;
    default: return FNET_H5_OTHER;
  }
}

int fnet_h5_attr_info(long long file, const char* obj, const char* name, long long* n, int* cls,
                      long long* size) {
  quiet();
  hid_t a = H5Aopen_by_name((hid_t)file, obj, name, H5P_DEFAULT, H5P_DEFAULT);
  if (a < 0) return fail("no attribute '%s' on '%s'", name, obj);
  hid_t sp = H5Aget_space(a), ty = H5Aget_type(a);
  int rc = 0;
  if (sp < 0 || ty < 0) {
    rc = fail("cannot inspect attribute '%s' of '%s'", name, obj);
  } else {
    *n = (long long)H5Sget_simple_extent_npoints(sp);
    *cls = class_code(H5Tget_class(ty));
    *size = H5Tis_variable_str(ty) > 0 ? -1 : (long long)H5Tget_size(ty);
  }
  if (ty >= 0) H5Tclose(ty);
  if (sp >= 0) H5Sclose(sp);
  H5Aclose(a);
  return rc;
}

int fnet_h5_set_attr_int(long long file, const char* obj, const char* name, const int* v,
                         long long n) {
  quiet();
  if (n < 1) return fail("attribute '%s' needs at least one value", name);
  if (H5LTset_attribute_int((hid_t)file, obj, name, v, (size_t)n) < 0)
    return fail("cannot write attribute '%s' of '%s'", name, obj);
  return 0;
}

int fnet_h5_set_attr_double(long long file, const char* obj, const char* name, const double* v,
                            long long n) {
  quiet();
  if (n < 1) return fail("attribute '%s' needs at least one value", name);
  if (H5LTset_attribute_double((hid_t)file, obj, name, v, (size_t)n) < 0)
    return fail("cannot write attribute '%s' of '%s'", name, obj);
  return 0;
}

int fnet_h5_set_attr_string(long long file, const char* obj, const char* name, const char* s) {
  quiet();
  if (H5LTset_attribute_string((hid_t)file, obj, name, s) < 0)
    return fail("cannot write attribute '%s' of '%s'", name, obj);
  return 0;
}

static int read_attr(long long file, const char* obj, const char* name, hid_t memtype, void* v,
                     long long n) {
  quiet();
  hid_t a = H5Aopen_by_name((hid_t)file, obj, name, H5P_DEFAULT, H5P_DEFAULT);
  if (a < 0) return fail("no attribute '%s' on '%s'", name, obj);
  hid_t sp = H5Aget_space(a);
  int rc = 0;
  const long long have = sp >= 0 ? (long long)H5Sget_simple_extent_npoints(sp) : -1;
  if (have != n)
    rc = fail("attribute '%s' of '%s' has %lld values, %lld expected", name, obj, have, n);
  else if (H5Aread(a, memtype, v) < 0)
    rc = fail("cannot read attribute '%s' of '%s'", name, obj);
  if (sp >= 0) H5Sclose(sp);
  H5Aclose(a);
  return rc;
}

int fnet_h5_get_attr_int(long long file, const char* obj, const char* name, int* v, long long n) {
  return read_attr(file, obj, name, H5T_NATIVE_INT, v, n);
}

int fnet_h5_get_attr_double(long long file, const char* obj, const char* name, double* v,
                            long long n) {
  return read_attr(file, obj, name, H5T_NATIVE_DOUBLE, v, n);
}

int fnet_h5_get_attr_string(long long file, const char* obj, const char* name, char* buf,
                            long long cap) {
  quiet();
  if (cap < 1) return fail("no room for attribute '%s'", name);
  buf[0] = 0;
  hid_t a = H5Aopen_by_name((hid_t)file, obj, name, H5P_DEFAULT, H5P_DEFAULT);
  if (a < 0) return fail("no attribute '%s' on '%s'", name, obj);
  hid_t ty = H5Aget_type(a), sp = H5Aget_space(a);
  int rc = 0;
  if (ty < 0 || sp < 0 || H5Tget_class(ty) != H5T_STRING) {
    rc = fail("attribute '%s' of '%s' is not a string", name, obj);
  } else if (H5Sget_simple_extent_npoints(sp) != 1) {
    rc = fail("attribute '%s' of '%s' is not a single string", name, obj);
  } else {
    hid_t mt = H5Tcopy(H5T_C_S1);
    if (H5Tis_variable_str(ty) > 0) {
      char* p = NULL;
      H5Tset_size(mt, H5T_VARIABLE);
      if (H5Aread(a, mt, &p) < 0 || !p) {
        rc = fail("cannot read attribute '%s' of '%s'", name, obj);
      } else {
        strncpy(buf, p, (size_t)(cap - 1));
        buf[cap - 1] = 0;
        H5free_memory(p);
      }
    } else {
      const size_t len = H5Tget_size(ty);
      char* tmp = (char*)calloc(len + 1, 1);
      H5Tset_size(mt, len + 1);
      H5Tset_strpad(mt, H5T_STR_NULLTERM);
      if (!tmp || H5Aread(a, mt, tmp) < 0) {
        rc = fail("cannot read attribute '%s' of '%s'", name, obj);
      } else {
        strncpy(buf, tmp, (size_t)(cap - 1));
        buf[cap - 1] = 0;
      }
      free(tmp);
    }
    H5Tclose(mt);
  }
  if (sp >= 0) H5Sclose(sp);
  if (ty >= 0) H5Tclose(ty);
  H5Aclose(a);
  return rc;
}

static int make_dataset(long long file, const char* path, int rank, const long long* dims,
                        hid_t type, const void* data) {
  quiet();
  if (rank < 1 || rank > FNET_H5_MAX_RANK) return fail("dataset '%s': bad rank %d", path, rank);
  hsize_t d[FNET_H5_MAX_RANK];
  for (int r = 0; r < rank; ++r) {
    if (dims[r] < 0) return fail("dataset '%s': negative dimension", path);
    d[r] = (hsize_t)dims[r];
  }
  if (H5LTmake_dataset((hid_t)file, path, rank, d, type, data) < 0)
    return fail("cannot write dataset '%s'", path);
  return 0;
}

int fnet_h5_make_dataset_int(long long file, const char* path, int rank, const long long* dims,
                             const int* data) {
  return make_dataset(file, path, rank, dims, H5T_NATIVE_INT, data);
}

int fnet_h5_make_dataset_double(long long file, const char* path, int rank,
                                const long long* dims, const double* data) {
  return make_dataset(file, path, rank, dims, H5T_NATIVE_DOUBLE, data);
}

int fnet_h5_dataset_info(long long file, const char* path, int* rank, long long* dims, int* cls) {
  quiet();
  int nd = 0;
  if (H5LTget_dataset_ndims((hid_t)file, path, &nd) < 0) return fail("no dataset '%s'", path);
  if (nd > FNET_H5_MAX_RANK) return fail("dataset '%s': rank %d too large", path, nd);
  hsize_t d[FNET_H5_MAX_RANK];
  H5T_class_t c;
  size_t sz;
  if (H5LTget_dataset_info((hid_t)file, path, d, &c, &sz) < 0)
    return fail("cannot inspect dataset '%s'", path);
  *rank = nd;
  for (int r = 0; r < nd; ++r) dims[r] = (long long)d[r];
  *cls = class_code(c);
  return 0;
}

static int read_dataset(long long file, const char* path, hid_t memtype, void* data,
                        long long n) {
  quiet();
  int rank, cls;
  long long dims[FNET_H5_MAX_RANK];
  if (fnet_h5_dataset_info(file, path, &rank, dims, &cls) < 0) return -1;
  long long have = 1;
  for (int r = 0; r < rank; ++r) have *= dims[r];
  if (have != n) return fail("dataset '%s' has %lld values, %lld expected", path, have, n);
  if (cls != FNET_H5_INT && cls != FNET_H5_FLOAT)
    return fail("dataset '%s' is not numerical", path);
  if (H5LTread_dataset((hid_t)file, path, memtype, data) < 0)
    return fail("cannot read dataset '%s'", path);
  return 0;
}

int fnet_h5_read_dataset_int(long long file, const char* path, int* data, long long n) {
  return read_dataset(file, path, H5T_NATIVE_INT, data, n);
}

int fnet_h5_read_dataset_double(long long file, const char* path, double* data, long long n) {
  return read_dataset(file, path, H5T_NATIVE_DOUBLE, data, n);
}
