/* Test infrastructure: direct HDF5 C-API access used to pin the MRI_Raw.h5
 * layout independently of libmrn_io.
 *
 * The tests write files element by element the way the reference's HDF5
 * wrapper does (hdf5_interface.cpp:82-135 for scalar attributes: STD_I32LE
 * ints, NATIVE_FLOAT floats; hdf5_interface.h:272-368 for arrays: Blitz
 * lengths reversed as the HDF5 dimensions, complex<float> as the compound
 * {real: float at 0, imag: float at 4}), and inspect what libmrn_io writes
 * (stored types, dimensions, raw bytes).  Nothing here is shared with the
 * library's own code. */
#include <hdf5.h>
#include <stdint.h>
#include <string.h>

static hid_t group_of(hid_t fid, const char *g) {
  H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
  if (H5Lexists(fid, g, H5P_DEFAULT) > 0) return H5Gopen2(fid, g, H5P_DEFAULT);
  return H5Gcreate2(fid, g, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
}

int64_t h5l_create(const char *path) { return H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT); }
int64_t h5l_open(const char *path) { return H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT); }
int h5l_close(int64_t fid) { return H5Fclose((hid_t)fid); }

static int attr(int64_t fid, const char *g, const char *name, hid_t ftype, hid_t mtype, const void *v) {
  hid_t gr = group_of((hid_t)fid, g);
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(gr, name, ftype, sp, H5P_DEFAULT, H5P_DEFAULT);
  int rc = (a < 0 || H5Awrite(a, mtype, v) < 0) ? -1 : 0;
  if (a >= 0) H5Aclose(a);
  H5Sclose(sp);
  H5Gclose(gr);
  return rc;
}

int h5l_attr_i32(int64_t fid, const char *g, const char *name, int32_t v) {
  return attr(fid, g, name, H5T_STD_I32LE, H5T_NATIVE_INT32, &v);
}
int h5l_attr_f32(int64_t fid, const char *g, const char *name, float v) {
  return attr(fid, g, name, H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, &v);
}

/* kind: 0 float32, 1 float64, 2 uint64, 3 complex<float> compound */
static hid_t type_of(int kind) {
  if (kind == 0) return H5Tcopy(H5T_NATIVE_FLOAT);
  if (kind == 1) return H5Tcopy(H5T_NATIVE_DOUBLE);
  if (kind == 2) return H5Tcopy(H5T_NATIVE_UINT64);
  hid_t t = H5Tcreate(H5T_COMPOUND, 8);
  H5Tinsert(t, "real", 0, H5T_NATIVE_FLOAT);
  H5Tinsert(t, "imag", 4, H5T_NATIVE_FLOAT);
  return t;
}

/* hdims: the HDF5 dimensions (slowest first) */
int h5l_dataset(int64_t fid, const char *g, const char *name, int kind, int rank, const uint64_t *hdims,
                const void *data) {
  hsize_t d[8];
  for (int i = 0; i < rank; i++) d[i] = (hsize_t)hdims[i];
  hid_t gr = group_of((hid_t)fid, g);
  hid_t sp = H5Screate_simple(rank, d, NULL);
  hid_t t = type_of(kind);
  hid_t ds = H5Dcreate2(gr, name, t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  int rc = (ds < 0 || H5Dwrite(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) ? -1 : 0;
  if (ds >= 0) H5Dclose(ds);
  H5Tclose(t);
  H5Sclose(sp);
  H5Gclose(gr);
  return rc;
}

/* Stored type of an attribute: class (H5T_class_t), size, byte order;
 * value read into v as 8 bytes of the stored type. */
int h5l_attr_info(int64_t fid, const char *g, const char *name, int *cls, int *size, int *order, void *v) {
  H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
  hid_t gr = H5Gopen2((hid_t)fid, g, H5P_DEFAULT);
  if (gr < 0) return -1;
  hid_t a = H5Aopen(gr, name, H5P_DEFAULT);
  if (a < 0) {
    H5Gclose(gr);
    return -1;
  }
  hid_t t = H5Aget_type(a);
  *cls = (int)H5Tget_class(t);
  *size = (int)H5Tget_size(t);
  *order = (int)H5Tget_order(t);
  memset(v, 0, 8);
  int rc = H5Aread(a, t, v) < 0 ? -1 : 0;
  H5Tclose(t);
  H5Aclose(a);
  H5Gclose(gr);
  return rc;
}

/* Dataset rank, HDF5 dimensions, type class and size; for a compound the
 * member count and, per member, class, offset and name (names joined by
 * ',' into `names`). */
int h5l_dataset_info(int64_t fid, const char *g, const char *name, int *rank, uint64_t *hdims, int *cls, int *size,
                     int *nmemb, int *mcls, int *moff, char *names, int names_cap) {
  H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
  hid_t gr = H5Gopen2((hid_t)fid, g, H5P_DEFAULT);
  if (gr < 0) return -1;
  hid_t ds = H5Dopen2(gr, name, H5P_DEFAULT);
  if (ds < 0) {
    H5Gclose(gr);
    return -1;
  }
  hid_t sp = H5Dget_space(ds);
  *rank = H5Sget_simple_extent_ndims(sp);
  hsize_t d[8];
  H5Sget_simple_extent_dims(sp, d, NULL);
  for (int i = 0; i < *rank; i++) hdims[i] = (uint64_t)d[i];
  hid_t t = H5Dget_type(ds);
  *cls = (int)H5Tget_class(t);
  *size = (int)H5Tget_size(t);
  *nmemb = 0;
  names[0] = 0;
  if (*cls == H5T_COMPOUND) {
    *nmemb = H5Tget_nmembers(t);
    for (int m = 0; m < *nmemb && m < 8; m++) {
      hid_t mt = H5Tget_member_type(t, m);
      mcls[m] = (int)H5Tget_class(mt);
      moff[m] = (int)H5Tget_member_offset(t, m);
      char *nm = H5Tget_member_name(t, m);
      if (strlen(names) + strlen(nm) + 2 < (size_t)names_cap) {
        if (m) strcat(names, ",");
        strcat(names, nm);
      }
      H5free_memory(nm);
      H5Tclose(mt);
    }
  }
  H5Tclose(t);
  H5Sclose(sp);
  H5Dclose(ds);
  H5Gclose(gr);
  return 0;
}

/* The dataset's bytes in its stored type. */
int h5l_dataset_raw(int64_t fid, const char *g, const char *name, void *buf) {
  hid_t gr = H5Gopen2((hid_t)fid, g, H5P_DEFAULT);
  if (gr < 0) return -1;
  hid_t ds = H5Dopen2(gr, name, H5P_DEFAULT);
  if (ds < 0) {
    H5Gclose(gr);
    return -1;
  }
  hid_t t = H5Dget_type(ds);
  int rc = H5Dread(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0 ? -1 : 0;
  H5Tclose(t);
  H5Dclose(ds);
  H5Gclose(gr);
  return rc;
}

int h5l_class_integer(void) { return H5T_INTEGER; }
int h5l_class_float(void) { return H5T_FLOAT; }
int h5l_class_compound(void) { return H5T_COMPOUND; }
int h5l_order_le(void) { return H5T_ORDER_LE; }
