/* HDF5 I/O layer of fortnet_amd: the file primitives the reference's I/O modules are built on,
 * as a C-ABI over the HDF5 C and high-level ("lite") libraries.
 *
 * The reference reads and writes its three HDF5 files (the fnetdata dataset, the netstat and
 * fnetout) through the HDF5 Fortran high-level API, wrapped for 1-D/2-D arrays in
 * prog/fortnet/lib_hdf5fx/hdf5fx.F90:52-238:
 *   h5fcreate_f / h5fopen_f / h5fclose_f            -> fnet_h5_create / fnet_h5_open / fnet_h5_close
 *   h5gcreate_f                                     -> fnet_h5_group_create
 *   h5ltfind_dataset_f, h5ldelete_f                 -> fnet_h5_exists, fnet_h5_delete
 *   h5ltset_attribute_{int,double,string}_f         -> fnet_h5_set_attr_{int,double,string}
 *   h5ltget_attribute_{int,double,string}_f         -> fnet_h5_get_attr_{int,double,string}
 *   h5ltfxmake_dataset_{int,double}_f (hdf5fx.F90:147-238)
 *                                                   -> fnet_h5_make_dataset_{int,double}
 *   h5ltget_dataset_info_f + h5ltread_dataset_{int,double}_f (hdf5fx.F90:51-144)
 *                                                   -> fnet_h5_dataset_info,
 *                                                      fnet_h5_read_dataset_{int,double}
 * The schema (group and attribute names, Fortran array order) lives in the host layer
 * (fortnet_amd/h5io.py), which mirrors readHdfDataset (fnetdata.F90:994-1216), the netstat
 * writers and readers (netstat.F90:37-453, bpnn.F90:1062-1288, acsf.F90:1759-2135) and
 * writeFnetout (fnetout.F90:33-158).
 *
 * Conventions: a file is a handle (>= 0) returned by create/open; every object is named by its
 * path from the file root ("netstat/bpnn"); attributes are named by (object path, attribute
 * name), "." being the root.  Dimensions are C order (slowest first), as HDF5 stores them: a
 * Fortran array a(n1, n2) written by the reference has dims {n2, n1}.  Functions return 0 (or a
 * count) on success and a negative value on error; fnet_h5_last_error() then holds the message
 * (per thread).  Nothing here touches a GPU.
 */
#ifndef FNET_H5_H
#define FNET_H5_H

#ifdef __cplusplus
extern "C" {
#endif

/* Type classes reported by fnet_h5_attr_info / fnet_h5_dataset_info. */
#define FNET_H5_INT 0
#define FNET_H5_FLOAT 1
#define FNET_H5_STRING 2
#define FNET_H5_OTHER 3

/* Largest rank fnet_h5_dataset_info reports. */
#define FNET_H5_MAX_RANK 8

const char* fnet_h5_last_error(void);

/* h5fcreate_f(H5F_ACC_TRUNC_F) (netstat.F90:49, fnetout.F90:83). */
long long fnet_h5_create(const char* path);
/* h5fopen_f, read-only (rdwr = 0) or read-write (netstat.F90:337). */
long long fnet_h5_open(const char* path, int rdwr);
int fnet_h5_close(long long file);

/* h5gcreate_f (netstat.F90:52,343). */
int fnet_h5_group_create(long long file, const char* path);
/* h5ltfind_dataset_f: 1 if a link of this name exists in its parent group, 0 if not. */
int fnet_h5_exists(long long file, const char* path);
/* h5ldelete_f (bpnn.F90:1100-1107). */
int fnet_h5_delete(long long file, const char* path);
/* Names of the links of a group, in name order, each followed by a NUL, into buf.  Returns the
 * count; if cap is too small nothing is written and the return is -(bytes needed) - 1. */
int fnet_h5_list(long long file, const char* path, char* buf, long long cap);

/* Attributes. */
int fnet_h5_attr_exists(long long file, const char* obj, const char* name);
int fnet_h5_attr_info(long long file, const char* obj, const char* name, long long* n,
                      int* cls, long long* size);
int fnet_h5_set_attr_int(long long file, const char* obj, const char* name, const int* v,
                         long long n);
int fnet_h5_set_attr_double(long long file, const char* obj, const char* name, const double* v,
                            long long n);
int fnet_h5_set_attr_string(long long file, const char* obj, const char* name, const char* s);
/* Read n values (n must equal the attribute's element count). */
int fnet_h5_get_attr_int(long long file, const char* obj, const char* name, int* v, long long n);
int fnet_h5_get_attr_double(long long file, const char* obj, const char* name, double* v,
                            long long n);
/* Fixed-length or variable-length string, NUL-terminated into buf (truncated to cap - 1). */
int fnet_h5_get_attr_string(long long file, const char* obj, const char* name, char* buf,
                            long long cap);

/* Datasets: contiguous, native int32 / float64, as H5LTmake_dataset writes them. */
int fnet_h5_make_dataset_int(long long file, const char* path, int rank, const long long* dims,
                             const int* data);
int fnet_h5_make_dataset_double(long long file, const char* path, int rank,
                                const long long* dims, const double* data);
/* rank and dims (at most FNET_H5_MAX_RANK) and the type class of a dataset. */
int fnet_h5_dataset_info(long long file, const char* path, int* rank, long long* dims, int* cls);
/* Read n elements (n must equal the dataset's element count), converted to int32 / float64. */
int fnet_h5_read_dataset_int(long long file, const char* path, int* data, long long n);
int fnet_h5_read_dataset_double(long long file, const char* path, double* data, long long n);

#ifdef __cplusplus
}
#endif

#endif
