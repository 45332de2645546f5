// MRI_Raw.h5 reader and writer (host code over the HDF5 C API).
//
// Layout of MRI_DATA::write_external_data / read_external_data
// (mri_data.cpp:571-902) through the reference's HDF5 wrapper
// (hdf5_interface.cpp): groups /Kdata and /Gating; int attributes written as
// STD_I32LE and float attributes as NATIVE_FLOAT (hdf5_interface.cpp:82-135);
// arrays with their Blitz lengths reversed as the HDF5 dimensions
// (hdf5_interface.h:249-368), complex<float> as the compound
// {real: float at 0, imag: float at 4} (hdf5_interface.h:272-292).
//
// Differences from the reference, each deliberate:
//  * the gating datasets are optional on reading (read_external_data reads
//    them unconditionally and fails on a file without them, although
//    write_external_data only writes them when there is ECG data);
//  * a missing or mis-sized array is an error status, where the reference
//    prints and resizes or rethrows;
//  * the reader accepts any stored type HDF5 converts to the requested one
//    (the reference reads with the file's own type into its array).
#include <hdf5.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/mrn_io.h"

namespace {

thread_local std::string g_err;

int raw_fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// owning HDF5 identifier
struct Hid {
  hid_t id = -1;
  herr_t (*closer)(hid_t) = nullptr;
  Hid() = default;
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), closer(c) {}
  Hid(const Hid &) = delete;
  Hid &operator=(const Hid &) = delete;
  Hid(Hid &&o) noexcept : id(o.id), closer(o.closer) { o.id = -1; }
  ~Hid() {
    if (id >= 0 && closer) closer(id);
  }
  bool ok() const { return id >= 0; }
};

// errors are reported through statuses; keep HDF5 from printing its stack
void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

hid_t complex_type() {
  hid_t t = H5Tcreate(H5T_COMPOUND, 2 * sizeof(float));
  H5Tinsert(t, "real", 0, H5T_NATIVE_FLOAT);
  H5Tinsert(t, "imag", sizeof(float), H5T_NATIVE_FLOAT);
  return t;
}

std::string enc_name(const char *base, int e) { return std::string(base) + "_E" + std::to_string(e); }
std::string coil_name(const char *base, int e, int c) {
  return std::string(base) + "_E" + std::to_string(e) + "_C" + std::to_string(c);
}

}  // namespace

struct mrn_raw {
  hid_t file = -1;
  bool write = false;
  mrn_raw_info info{};
  bool have_info = false;
};

namespace {

Hid open_group(mrn_raw *f, const char *name, bool create) {
  quiet();
  if (H5Lexists(f->file, name, H5P_DEFAULT) > 0) return Hid(H5Gopen2(f->file, name, H5P_DEFAULT), H5Gclose);
  if (!create) return Hid();
  return Hid(H5Gcreate2(f->file, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
}

bool exists(hid_t loc, const std::string &name) { return H5Lexists(loc, name.c_str(), H5P_DEFAULT) > 0; }

int put_int_attr(hid_t g, const char *name, int v) {
  if (H5Aexists(g, name) > 0) H5Adelete(g, name);
  Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
  Hid a(H5Acreate2(g, name, H5T_STD_I32LE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok() || H5Awrite(a.id, H5T_NATIVE_INT, &v) < 0)
    return raw_fail(MRN_EHIP, std::string("cannot write attribute ") + name);
  return MRN_OK;
}

int put_float_attr(hid_t g, const char *name, float v) {
  if (H5Aexists(g, name) > 0) H5Adelete(g, name);
  Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
  Hid a(H5Acreate2(g, name, H5T_NATIVE_FLOAT, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok() || H5Awrite(a.id, H5T_NATIVE_FLOAT, &v) < 0)
    return raw_fail(MRN_EHIP, std::string("cannot write attribute ") + name);
  return MRN_OK;
}

// required: absent is an error; otherwise absent leaves *v unchanged
int get_attr(hid_t g, const char *name, hid_t mem, void *v, bool required) {
  if (H5Aexists(g, name) <= 0)
    return required ? raw_fail(MRN_EINVAL, std::string("attribute /Kdata/") + name + " missing") : MRN_OK;
  Hid a(H5Aopen(g, name, H5P_DEFAULT), H5Aclose);
  if (!a.ok() || H5Aread(a.id, mem, v) < 0) return raw_fail(MRN_EINVAL, std::string("cannot read attribute ") + name);
  return MRN_OK;
}

// Blitz lengths of /group/name (reversed HDF5 dimensions), rank must be `rank`
int dataset_dims(hid_t g, const std::string &name, int rank, int64_t *dims) {
  if (!exists(g, name)) return raw_fail(MRN_EINVAL, "dataset " + name + " missing");
  Hid d(H5Dopen2(g, name.c_str(), H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return raw_fail(MRN_EINVAL, "cannot open dataset " + name);
  Hid sp(H5Dget_space(d.id), H5Sclose);
  const int r = H5Sget_simple_extent_ndims(sp.id);
  if (r != rank) return raw_fail(MRN_EINVAL, "dataset " + name + " has rank " + std::to_string(r));
  hsize_t hd[8];
  H5Sget_simple_extent_dims(sp.id, hd, nullptr);
  for (int i = 0; i < rank; i++) dims[i] = (int64_t)hd[rank - 1 - i];
  return MRN_OK;
}

int read_array(hid_t g, const std::string &name, hid_t mem, void *out, int64_t n) {
  if (!exists(g, name)) return raw_fail(MRN_EINVAL, "dataset " + name + " missing");
  Hid d(H5Dopen2(g, name.c_str(), H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return raw_fail(MRN_EINVAL, "cannot open dataset " + name);
  Hid sp(H5Dget_space(d.id), H5Sclose);
  const hssize_t np = H5Sget_simple_extent_npoints(sp.id);
  if (np != (hssize_t)n)
    return raw_fail(MRN_EINVAL, "dataset " + name + " has " + std::to_string((long long)np) + " points, expected " +
                                    std::to_string((long long)n));
  if (n > 0 && H5Dread(d.id, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
    return raw_fail(MRN_EINVAL, "cannot read dataset " + name + " as the requested type");
  return MRN_OK;
}

int write_array(hid_t g, const std::string &name, hid_t file_type, hid_t mem, const void *data, int rank,
                const int64_t *dims) {
  hsize_t hd[8];
  for (int i = 0; i < rank; i++) {
    if (dims[i] < 0) return raw_fail(MRN_EINVAL, "negative dimension for " + name);
    hd[i] = (hsize_t)dims[rank - 1 - i];
  }
  if (exists(g, name)) H5Ldelete(g, name.c_str(), H5P_DEFAULT);
  Hid sp(H5Screate_simple(rank, hd, nullptr), H5Sclose);
  Hid d(H5Dcreate2(g, name.c_str(), file_type, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return raw_fail(MRN_EHIP, "cannot create dataset " + name);
  int64_t n = 1;
  for (int i = 0; i < rank; i++) n *= dims[i];
  if (n > 0 && H5Dwrite(d.id, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return raw_fail(MRN_EHIP, "cannot write dataset " + name);
  return MRN_OK;
}

int check_open(mrn_raw *f, bool need_write) {
  if (!f || f->file < 0) return raw_fail(MRN_EINVAL, "file not open");
  if (need_write && !f->write) return raw_fail(MRN_ESTATE, "file opened for reading");
  return MRN_OK;
}

int coils_of(mrn_raw *f, int &C) {
  if (!f->have_info) {
    mrn_raw_info tmp;
    int rc = mrn_raw_get_info(f, &tmp);
    if (rc != MRN_OK) return rc;
  }
  C = f->info.num_coils;
  return MRN_OK;
}

static const char *kRotNames[12] = {"ix", "iy", "iz", "jx", "jy", "jz", "kx", "ky", "kz", "sx", "sy", "sz"};

}  // namespace

extern "C" {

int mrn_raw_open(const char *path, int write, mrn_raw **out) {
  if (!path || !out) return raw_fail(MRN_EINVAL, "null argument");
  *out = nullptr;
  quiet();
  hid_t fid = write ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                    : H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (fid < 0) return raw_fail(MRN_EINVAL, std::string("cannot ") + (write ? "create " : "open ") + path);
  mrn_raw *f = new mrn_raw();
  f->file = fid;
  f->write = write != 0;
  *out = f;
  return MRN_OK;
}

int mrn_raw_close(mrn_raw *f) {
  if (!f) return raw_fail(MRN_EINVAL, "null file");
  herr_t e = f->file >= 0 ? H5Fclose(f->file) : 0;
  delete f;
  return e < 0 ? raw_fail(MRN_EHIP, "H5Fclose failed") : MRN_OK;
}

int mrn_raw_get_info(mrn_raw *f, mrn_raw_info *info) {
  int rc = check_open(f, false);
  if (rc != MRN_OK) return rc;
  if (!info) return raw_fail(MRN_EINVAL, "null info");
  Hid g = open_group(f, "Kdata", false);
  if (!g.ok()) return raw_fail(MRN_EINVAL, "group /Kdata missing");
  mrn_raw_info r;
  memset(&r, 0, sizeof(r));
  // read_external_data requires these (mri_data.cpp:775-796)
  static const char *req[8] = {"Num_Encodings", "Num_Coils", "trajectory_typeX", "trajectory_typeY",
                               "trajectory_typeZ", "dft_neededX", "dft_neededY", "dft_neededZ"};
  int32_t *dst[8] = {&r.num_encodings, &r.num_coils, &r.trajectory_type[0], &r.trajectory_type[1],
                     &r.trajectory_type[2], &r.dft_needed[0], &r.dft_needed[1], &r.dft_needed[2]};
  for (int q = 0; q < 8; q++)
    if ((rc = get_attr(g.id, req[q], H5T_NATIVE_INT32, dst[q], true)) != MRN_OK) return rc;
  if ((rc = get_attr(g.id, "Num_Frames", H5T_NATIVE_INT32, &r.num_frames, false)) != MRN_OK) return rc;
  static const char *fovn[3] = {"fovx", "fovy", "fovz"};
  for (int a = 0; a < 3; a++)
    if ((rc = get_attr(g.id, fovn[a], H5T_NATIVE_FLOAT, &r.fov[a], false)) != MRN_OK) return rc;
  for (int q = 0; q < 12; q++)
    if ((rc = get_attr(g.id, kRotNames[q], H5T_NATIVE_FLOAT, &r.rot[q], false)) != MRN_OK) return rc;
  if (r.num_encodings < 0 || r.num_coils < 0) return raw_fail(MRN_EINVAL, "negative Num_Encodings or Num_Coils");
  f->info = r;
  f->have_info = true;
  *info = r;
  return MRN_OK;
}

int mrn_raw_put_info(mrn_raw *f, const mrn_raw_info *info) {
  int rc = check_open(f, true);
  if (rc != MRN_OK) return rc;
  if (!info || info->num_encodings < 0 || info->num_coils < 0) return raw_fail(MRN_EINVAL, "bad info");
  Hid g = open_group(f, "Kdata", true);
  if (!g.ok()) return raw_fail(MRN_EHIP, "cannot create group /Kdata");
  // the order of write_external_data (mri_data.cpp:576-605)
  if ((rc = put_int_attr(g.id, "Num_Encodings", info->num_encodings)) != MRN_OK) return rc;
  if ((rc = put_int_attr(g.id, "Num_Coils", info->num_coils)) != MRN_OK) return rc;
  if ((rc = put_int_attr(g.id, "Num_Frames", info->num_frames)) != MRN_OK) return rc;
  static const char *fovn[3] = {"fovx", "fovy", "fovz"};
  for (int a = 0; a < 3; a++)
    if ((rc = put_float_attr(g.id, fovn[a], info->fov[a])) != MRN_OK) return rc;
  for (int q = 0; q < 12; q++)
    if ((rc = put_float_attr(g.id, kRotNames[q], info->rot[q])) != MRN_OK) return rc;
  static const char *tt[3] = {"trajectory_typeX", "trajectory_typeY", "trajectory_typeZ"};
  static const char *dn[3] = {"dft_neededX", "dft_neededY", "dft_neededZ"};
  for (int a = 0; a < 3; a++)
    if ((rc = put_int_attr(g.id, tt[a], info->trajectory_type[a])) != MRN_OK) return rc;
  for (int a = 0; a < 3; a++)
    if ((rc = put_int_attr(g.id, dn[a], info->dft_needed[a] ? 1 : 0)) != MRN_OK) return rc;
  f->info = *info;
  f->have_info = true;
  return MRN_OK;
}

int mrn_raw_exists(mrn_raw *f, const char *group, const char *name) {
  int rc = check_open(f, false);
  if (rc != MRN_OK) return -rc;
  if (!group || !name) return -raw_fail(MRN_EINVAL, "null name");
  Hid g = open_group(f, group, false);
  if (!g.ok()) return 0;
  return exists(g.id, name) ? 1 : 0;
}

int mrn_raw_encoding_dims(mrn_raw *f, int e, int64_t dims[3]) {
  int rc = check_open(f, false);
  if (rc != MRN_OK) return rc;
  if (!dims || e < 0) return raw_fail(MRN_EINVAL, "bad argument");
  Hid g = open_group(f, "Kdata", false);
  if (!g.ok()) return raw_fail(MRN_EINVAL, "group /Kdata missing");
  return dataset_dims(g.id, enc_name("KX", e), 3, dims);
}

int mrn_raw_read_encoding(mrn_raw *f, int e, float *kx, float *ky, float *kz, float *kw, float *kt, void *kdata) {
  int rc = check_open(f, false);
  if (rc != MRN_OK) return rc;
  int64_t d[3];
  if ((rc = mrn_raw_encoding_dims(f, e, d)) != MRN_OK) return rc;
  const int64_t n = d[0] * d[1] * d[2];
  Hid g = open_group(f, "Kdata", false);
  const char *names[5] = {"KX", "KY", "KZ", "KW", "KT"};
  float *outs[5] = {kx, ky, kz, kw, kt};
  for (int q = 0; q < 5; q++)
    if (outs[q] && (rc = read_array(g.id, enc_name(names[q], e), H5T_NATIVE_FLOAT, outs[q], n)) != MRN_OK) return rc;
  if (kdata) {
    int C = 0;
    if ((rc = coils_of(f, C)) != MRN_OK) return rc;
    Hid ct(complex_type(), H5Tclose);
    for (int c = 0; c < C; c++)
      if ((rc = read_array(g.id, coil_name("KData", e, c), ct.id, (float *)kdata + 2 * n * c, n)) != MRN_OK)
        return rc;
  }
  return MRN_OK;
}

int mrn_raw_write_encoding(mrn_raw *f, int e, const int64_t dims[3], const float *kx, const float *ky,
                           const float *kz, const float *kw, const float *kt, const void *kdata) {
  int rc = check_open(f, true);
  if (rc != MRN_OK) return rc;
  if (!dims || e < 0 || !kx || !ky || !kz || !kw) return raw_fail(MRN_EINVAL, "bad argument");
  if (!f->have_info) return raw_fail(MRN_ESTATE, "mrn_raw_put_info first");
  Hid g = open_group(f, "Kdata", true);
  const int64_t n = dims[0] * dims[1] * dims[2];
  // KT, KX, KY, KZ, KW, then the coils (mri_data.cpp:607-690)
  const char *names[5] = {"KT", "KX", "KY", "KZ", "KW"};
  const float *ins[5] = {kt, kx, ky, kz, kw};
  for (int q = 0; q < 5; q++)
    if (ins[q] && (rc = write_array(g.id, enc_name(names[q], e), H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, ins[q], 3,
                                    dims)) != MRN_OK)
      return rc;
  if (kdata) {
    Hid ct(complex_type(), H5Tclose);
    for (int c = 0; c < f->info.num_coils; c++)
      if ((rc = write_array(g.id, coil_name("KData", e, c), ct.id, ct.id, (const float *)kdata + 2 * n * c, 3,
                            dims)) != MRN_OK)
        return rc;
  }
  return MRN_OK;
}

int mrn_raw_gating_dims(mrn_raw *f, int e, int64_t dims[2]) {
  int rc = check_open(f, false);
  if (rc != MRN_OK) return rc;
  if (!dims || e < 0) return raw_fail(MRN_EINVAL, "bad argument");
  Hid g = open_group(f, "Gating", false);
  if (!g.ok() || !exists(g.id, enc_name("ECG", e))) return raw_fail(MRN_ESTATE, "no gating data for this encoding");
  return dataset_dims(g.id, enc_name("ECG", e), 2, dims);
}

int mrn_raw_read_gating(mrn_raw *f, int e, double *ecg, double *resp, double *prep, double *time, uint64_t *clock,
                        void *k0) {
  int64_t d[2];
  int rc = mrn_raw_gating_dims(f, e, d);
  if (rc != MRN_OK) return rc;
  const int64_t n = d[0] * d[1];
  Hid g = open_group(f, "Gating", false);
  // the order of read_external_data (mri_data.cpp:851-886)
  const char *names[4] = {"ECG", "RESP", "PREP", "TIME"};
  double *outs[4] = {ecg, resp, prep, time};
  for (int q = 0; q < 4; q++)
    if (outs[q] && (rc = read_array(g.id, enc_name(names[q], e), H5T_NATIVE_DOUBLE, outs[q], n)) != MRN_OK)
      return rc;
  if (clock && (rc = read_array(g.id, enc_name("CLOCK", e), H5T_NATIVE_UINT64, clock, n)) != MRN_OK) return rc;
  if (k0) {
    int C = 0;
    if ((rc = coils_of(f, C)) != MRN_OK) return rc;
    Hid ct(complex_type(), H5Tclose);
    for (int c = 0; c < C; c++)
      if ((rc = read_array(g.id, coil_name("K0", e, c), ct.id, (float *)k0 + 2 * n * c, n)) != MRN_OK) return rc;
  }
  return MRN_OK;
}

int mrn_raw_write_gating(mrn_raw *f, int e, const int64_t dims[2], const double *ecg, const double *resp,
                         const double *prep, const double *time, const uint64_t *clock, const void *k0) {
  int rc = check_open(f, true);
  if (rc != MRN_OK) return rc;
  if (!dims || e < 0 || !ecg) return raw_fail(MRN_EINVAL, "bad argument (ECG is required: it marks gating)");
  if (!f->have_info) return raw_fail(MRN_ESTATE, "mrn_raw_put_info first");
  Hid g = open_group(f, "Gating", true);
  if (!g.ok()) return raw_fail(MRN_EHIP, "cannot create group /Gating");
  // ECG, CLOCK, RESP, PREP, TIME, K0 (mri_data.cpp:693-760)
  if ((rc = write_array(g.id, enc_name("ECG", e), H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, ecg, 2, dims)) != MRN_OK)
    return rc;
  if (clock &&
      (rc = write_array(g.id, enc_name("CLOCK", e), H5T_NATIVE_UINT64, H5T_NATIVE_UINT64, clock, 2, dims)) != MRN_OK)
    return rc;
  const char *names[3] = {"RESP", "PREP", "TIME"};
  const double *ins[3] = {resp, prep, time};
  for (int q = 0; q < 3; q++)
    if (ins[q] && (rc = write_array(g.id, enc_name(names[q], e), H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, ins[q], 2,
                                    dims)) != MRN_OK)
      return rc;
  if (k0) {
    const int64_t n = dims[0] * dims[1];
    Hid ct(complex_type(), H5Tclose);
    for (int c = 0; c < f->info.num_coils; c++)
      if ((rc = write_array(g.id, coil_name("K0", e, c), ct.id, ct.id, (const float *)k0 + 2 * n * c, 2, dims)) !=
          MRN_OK)
        return rc;
  }
  return MRN_OK;
}

int mrn_raw_noise_dims(mrn_raw *f, int64_t dims[2]) {
  int rc = check_open(f, false);
  if (rc != MRN_OK) return rc;
  if (!dims) return raw_fail(MRN_EINVAL, "bad argument");
  Hid g = open_group(f, "Kdata", false);
  if (!g.ok() || !exists(g.id, "Noise")) return raw_fail(MRN_ESTATE, "no noise samples");
  return dataset_dims(g.id, "Noise", 2, dims);
}

int mrn_raw_read_noise(mrn_raw *f, void *noise) {
  int64_t d[2];
  int rc = mrn_raw_noise_dims(f, d);
  if (rc != MRN_OK) return rc;
  if (!noise) return raw_fail(MRN_EINVAL, "null buffer");
  Hid g = open_group(f, "Kdata", false);
  Hid ct(complex_type(), H5Tclose);
  return read_array(g.id, "Noise", ct.id, noise, d[0] * d[1]);
}

int mrn_raw_write_noise(mrn_raw *f, const int64_t dims[2], const void *noise) {
  int rc = check_open(f, true);
  if (rc != MRN_OK) return rc;
  if (!dims || !noise) return raw_fail(MRN_EINVAL, "bad argument");
  Hid g = open_group(f, "Kdata", true);
  Hid ct(complex_type(), H5Tclose);
  return write_array(g.id, "Noise", ct.id, ct.id, noise, 2, dims);
}

int mrn_raw_params(const mrn_raw_info *info, const int32_t recon_N[3], mrn_params *p) {
  if (!info || !recon_N || !p) return raw_fail(MRN_EINVAL, "null argument");
  for (int a = 0; a < 3; a++) {
    if (recon_N[a] <= 0) return raw_fail(MRN_EINVAL, "recon size must be positive");
    p->N[a] = recon_N[a];
    p->grid_in[a] = info->trajectory_type[a] == 1 ? 1 : 0;  // MRI_DATA::NONCARTESIAN
    p->fft_in[a] = info->dft_needed[a] ? 1 : 0;
  }
  return MRN_OK;
}

const char *mrn_raw_error(void) { return g_err.c_str(); }

}  // extern "C"
