/* G4 files in the reference's HDF5 format (include/g4acc_io.h): plain C over the HDF5 C API, so
 * the library depends on libhdf5 and libc only.
 *
 * Format (HDF5Writer::execute for a complex func::function, include/dca/io/hdf5/hdf5_writer.hpp:
 * 244-262 and src/io/hdf5/hdf5_writer.cpp:153-193; element type hdf5_types.hpp:180-192):
 * compound { "r": F64, "i": F64 }, dataspace = dimensions slowest first, attributes
 * "domain-sizes" (uint64, first index fastest) and "name" (C string of strlen bytes). */
#include "../../include/g4acc_io.h"

#include <hdf5.h>

#include <stdio.h>
#include <string.h>

static _Thread_local char g_err[512];

static int fail(const char* what, const char* detail) {
  snprintf(g_err, sizeof g_err, "%s%s%s", what, detail ? ": " : "", detail ? detail : "");
  return -1;
}

const char* g4io_last_error(void) { return g_err; }

const char* g4io_channel_name(int channel) {
  /* FourPointType (include/dca/phys/four_point_type.hpp:27-33), toString
   * (src/phys/four_point_type.cpp:37-56). */
  switch (channel) {
    case 1: return "PARTICLE_HOLE_TRANSVERSE";
    case 2: return "PARTICLE_HOLE_MAGNETIC";
    case 3: return "PARTICLE_HOLE_CHARGE";
    case 4: return "PARTICLE_PARTICLE_UP_DOWN";
    default: return NULL;
  }
}

/* compound { "r": double, "i": double }, 16 bytes */
static hid_t complex_type(void) {
  hid_t t = H5Tcreate(H5T_COMPOUND, 2 * sizeof(double));
  if (t < 0)
    return t;
  if (H5Tinsert(t, "r", 0, H5T_NATIVE_DOUBLE) < 0 || H5Tinsert(t, "i", sizeof(double), H5T_NATIVE_DOUBLE) < 0) {
    H5Tclose(t);
    return -1;
  }
  return t;
}

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

static int write_attributes(hid_t ds, const char* name, const unsigned long long* dims, int ndims) {
  int rc = 0;
  if (H5Aexists(ds, "name") == 0) { /* an existing attribute is kept (hdf5_writer.cpp:180-182) */
    const size_t len = strlen(name);
    hid_t st = H5Tcopy(H5T_C_S1);
    hsize_t one = 1;
    hid_t sp = H5Screate_simple(1, &one, NULL);
    hid_t at = -1;
    if (st < 0 || sp < 0 || H5Tset_size(st, len > 0 ? len : 1) < 0 ||
        (at = H5Acreate2(ds, "name", st, sp, H5P_DEFAULT, H5P_DEFAULT)) < 0 || H5Awrite(at, st, name) < 0)
      rc = fail("writing attribute name", NULL);
    if (at >= 0)
      H5Aclose(at);
    if (sp >= 0)
      H5Sclose(sp);
    if (st >= 0)
      H5Tclose(st);
  }
  if (rc == 0 && H5Aexists(ds, "domain-sizes") == 0) {
    hsize_t n = (hsize_t)ndims;
    hsize_t d[32];
    hid_t sp = H5Screate_simple(1, &n, NULL);
    hid_t at = -1;
    for (int k = 0; k < ndims; ++k)
      d[k] = (hsize_t)dims[k];
    if (sp < 0 || (at = H5Acreate2(ds, "domain-sizes", H5T_NATIVE_HSIZE, sp, H5P_DEFAULT, H5P_DEFAULT)) < 0 ||
        H5Awrite(at, H5T_NATIVE_HSIZE, d) < 0)
      rc = fail("writing attribute domain-sizes", NULL);
    if (at >= 0)
      H5Aclose(at);
    if (sp >= 0)
      H5Sclose(sp);
  }
  return rc;
}

int g4io_write_function(const char* path, const char* group, const char* name, const double* values,
                        const unsigned long long* dims, int ndims, int truncate) {
  if (!path || !name || !values || !dims || ndims < 1 || ndims > 32)
    return fail("g4io_write_function", "invalid arguments");
  quiet();
  hid_t file = -1;
  if (!truncate) {
    FILE* f = fopen(path, "rb");
    if (f) {
      fclose(f);
      file = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
    }
  }
  if (file < 0)
    file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0)
    return fail("cannot open or create", path);
  int rc = 0;
  hid_t grp = -1, type = -1, space = -1, ds = -1;
  hsize_t rev[32];
  for (int k = 0; k < ndims; ++k)
    rev[k] = (hsize_t)dims[ndims - 1 - k]; /* slowest first (hdf5_writer.hpp:254-255) */
  if (group && *group) {
    grp = H5Lexists(file, group, H5P_DEFAULT) > 0 ? H5Gopen2(file, group, H5P_DEFAULT)
                                                  : H5Gcreate2(file, group, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (grp < 0)
      rc = fail("cannot open or create group", group);
  }
  const hid_t where = grp >= 0 ? grp : file;
  if (rc == 0 && ((type = complex_type()) < 0 || (space = H5Screate_simple(ndims, rev, NULL)) < 0))
    rc = fail("cannot build the complex type or dataspace", NULL);
  if (rc == 0) {
    if (H5Lexists(where, name, H5P_DEFAULT) > 0) {
      /* same shape: overwrite; else an error (hdf5_writer.cpp:155-167) */
      ds = H5Dopen2(where, name, H5P_DEFAULT);
      hid_t old = ds >= 0 ? H5Dget_space(ds) : -1;
      hsize_t od[32];
      const int on = old >= 0 ? H5Sget_simple_extent_ndims(old) : -1;
      if (on != ndims || H5Sget_simple_extent_dims(old, od, NULL) < 0 || memcmp(od, rev, sizeof(hsize_t) * ndims))
        rc = fail("Object size different than HDF5 dataset", name);
      if (old >= 0)
        H5Sclose(old);
    }
    else
      ds = H5Dcreate2(where, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (rc == 0 && ds < 0)
      rc = fail("cannot create dataset", name);
  }
  if (rc == 0 && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, values) < 0)
    rc = fail("cannot write dataset", name);
  if (rc == 0)
    rc = write_attributes(ds, name, dims, ndims);
  if (ds >= 0)
    H5Dclose(ds);
  if (space >= 0)
    H5Sclose(space);
  if (type >= 0)
    H5Tclose(type);
  if (grp >= 0)
    H5Gclose(grp);
  if (H5Fclose(file) < 0 && rc == 0)
    rc = fail("cannot close", path);
  return rc;
}

int g4io_read_function(const char* path, const char* dataset_path, double* values, size_t n_complex,
                       unsigned long long* dims, int max_dims, int* ndims) {
  if (!path || !dataset_path)
    return fail("g4io_read_function", "invalid arguments");
  quiet();
  hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0)
    return fail("cannot open", path);
  int rc = 0;
  hid_t ds = H5Dopen2(file, dataset_path, H5P_DEFAULT), space = -1, type = -1;
  hsize_t d[32];
  int n = -1;
  if (ds < 0)
    rc = fail("no dataset", dataset_path);
  if (rc == 0 && ((space = H5Dget_space(ds)) < 0 || (n = H5Sget_simple_extent_ndims(space)) < 1 || n > 32 ||
                  H5Sget_simple_extent_dims(space, d, NULL) < 0))
    rc = fail("cannot read the dataspace of", dataset_path);
  size_t total = 1;
  if (rc == 0) {
    if (ndims)
      *ndims = n;
    for (int k = 0; k < n; ++k) {
      total *= (size_t)d[k];
      if (dims && k < max_dims)
        dims[k] = (unsigned long long)d[n - 1 - k]; /* first index fastest */
    }
    if (dims && max_dims < n)
      rc = fail("dims capacity too small for", dataset_path);
  }
  if (rc == 0 && values) {
    hid_t ft = H5Dget_type(ds);
    /* the file type must be the {r, i} compound of two doubles */
    if (ft < 0 || H5Tget_class(ft) != H5T_COMPOUND || H5Tget_nmembers(ft) != 2 || H5Tget_member_index(ft, "r") < 0 ||
        H5Tget_member_index(ft, "i") < 0)
      rc = fail("not a complex {r, i} dataset", dataset_path);
    if (ft >= 0)
      H5Tclose(ft);
    if (rc == 0 && n_complex < total)
      rc = fail("output buffer too small for", dataset_path);
    if (rc == 0 && ((type = complex_type()) < 0 || H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, values) < 0))
      rc = fail("cannot read", dataset_path);
  }
  if (type >= 0)
    H5Tclose(type);
  if (space >= 0)
    H5Sclose(space);
  if (ds >= 0)
    H5Dclose(ds);
  H5Fclose(file);
  return rc;
}

int g4io_read_name(const char* path, const char* dataset_path, char* out, size_t capacity) {
  if (!path || !dataset_path || !out || capacity < 1)
    return fail("g4io_read_name", "invalid arguments");
  quiet();
  hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0)
    return fail("cannot open", path);
  int rc = 0;
  hid_t ds = H5Dopen2(file, dataset_path, H5P_DEFAULT), at = -1, ft = -1, mt = -1;
  if (ds < 0 || (at = H5Aopen(ds, "name", H5P_DEFAULT)) < 0 || (ft = H5Aget_type(at)) < 0)
    rc = fail("no name attribute on", dataset_path);
  if (rc == 0) {
    const size_t len = H5Tget_size(ft);
    if (len + 1 > capacity)
      rc = fail("name capacity too small for", dataset_path);
    else {
      mt = H5Tcopy(H5T_C_S1);
      H5Tset_size(mt, len + 1);
      memset(out, 0, capacity);
      if (H5Aread(at, mt, out) < 0)
        rc = fail("cannot read the name of", dataset_path);
    }
  }
  if (mt >= 0)
    H5Tclose(mt);
  if (ft >= 0)
    H5Tclose(ft);
  if (at >= 0)
    H5Aclose(at);
  if (ds >= 0)
    H5Dclose(ds);
  H5Fclose(file);
  return rc;
}
