/*
 * mrn_io.h — C ABI of the MRI_Raw.h5 reader and writer (host code, HDF5).
 *
 * Replaces MRI_DATA::write_external_data / read_external_data
 * (src/mri_data.cpp:571-902) over the reference's HDF5 wrapper
 * (src/hdf5_interface.cpp:1-507, hdf5_interface.h:191-370), with the same
 * file layout so that files move between the two:
 *
 *   /Kdata  int32 (STD_I32LE) attributes Num_Encodings, Num_Coils, Num_Frames,
 *           trajectory_typeX/Y/Z (0 Cartesian, 1 non-Cartesian: MRI_DATA::TrajType),
 *           dft_neededX/Y/Z; float attributes fovx/y/z, ix iy iz jx jy jz
 *           kx ky kz sx sy sz (rotation);
 *           KX_E#, KY_E#, KZ_E#, KW_E#, KT_E#  float32, 3-D;
 *           KData_E#_C#  complex64 as the compound {float real; float imag}, 3-D;
 *           Noise  complex64, 2-D (optional)
 *   /Gating ECG_E#, RESP_E#, PREP_E#, TIME_E#  float64, 2-D; CLOCK_E#  uint64, 2-D;
 *           K0_E#_C#  complex64, 2-D (optional, written when gating is)
 *
 * Array shapes are given as the reference's Blitz lengths (dims[0] varies
 * fastest, column-major); HDF5 stores them reversed, in C order, so the
 * bytes are the same.  Data pointers are host memory.  Every call returns an
 * mrn_status (include/mrn.h) and never exits; mrn_raw_error() gives this
 * thread's last message.
 */
#ifndef MRN_IO_H
#define MRN_IO_H
#include <stdint.h>

#include "mrn.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mrn_raw_info {
  int32_t num_encodings;       /* Num_Encodings */
  int32_t num_coils;           /* Num_Coils */
  int32_t num_frames;          /* Num_Frames (written by the reference; 0 if absent) */
  int32_t trajectory_type[3];  /* trajectory_typeX/Y/Z: 0 CARTESIAN, 1 NONCARTESIAN */
  int32_t dft_needed[3];       /* dft_neededX/Y/Z */
  float fov[3];                /* fovx, fovy, fovz (0 if absent) */
  float rot[12];               /* ix iy iz jx jy jz kx ky kz sx sy sz (0 if absent) */
  int32_t reserved[8];
} mrn_raw_info;

typedef struct mrn_raw mrn_raw; /* an open MRI_Raw.h5 file */

/* Open an existing file for reading (write = 0) or create / truncate one
 * for writing (write = 1): HDF5::HDF5(name, "r" / "w"), hdf5_interface.cpp:6-17. */
int mrn_raw_open(const char *path, int write, mrn_raw **out);
int mrn_raw_close(mrn_raw *f);

/* The /Kdata attributes read_external_data reads (mri_data.cpp:769-796) plus
 * those write_external_data writes (571-605); absent optional attributes
 * read as 0. */
int mrn_raw_get_info(mrn_raw *f, mrn_raw_info *info);
int mrn_raw_put_info(mrn_raw *f, const mrn_raw_info *info);

/* 1 if /group/name exists, 0 if not (negative: error). */
int mrn_raw_exists(mrn_raw *f, const char *group, const char *name);

/* Blitz lengths (dims[0] fastest) of encoding e's KX_E# (mri_data.cpp:807-812). */
int mrn_raw_encoding_dims(mrn_raw *f, int e, int64_t dims[3]);
/* Encoding e: n = dims[0]*dims[1]*dims[2] samples per array; kdata [C][n]
 * complex64 (coil c = KData_E#_C#).  Any pointer may be NULL to skip that
 * array; a requested array that is absent or of another size is an error.
 * (mri_data.cpp:798-846) */
int mrn_raw_read_encoding(mrn_raw *f, int e, float *kx, float *ky, float *kz, float *kw, float *kt,
                          void *kdata);
/* Write encoding e (mri_data.cpp:607-690); kt and kdata may be NULL. */
int mrn_raw_write_encoding(mrn_raw *f, int e, const int64_t dims[3], const float *kx, const float *ky,
                           const float *kz, const float *kw, const float *kt, const void *kdata);

/* Gating of encoding e (mri_data.cpp:848-886, written at 692-760): dims
 * of ECG_E# (MRN_ESTATE when the file has no gating for e). */
int mrn_raw_gating_dims(mrn_raw *f, int e, int64_t dims[2]);
/* ecg, resp, prep, time [n] float64, clock [n] uint64, k0 [C][n] complex64
 * (K0_E#_C#), n = dims[0]*dims[1]; NULL skips an array. */
int mrn_raw_read_gating(mrn_raw *f, int e, double *ecg, double *resp, double *prep, double *time,
                        uint64_t *clock, void *k0);
int mrn_raw_write_gating(mrn_raw *f, int e, const int64_t dims[2], const double *ecg, const double *resp,
                         const double *prep, const double *time, const uint64_t *clock, const void *k0);

/* /Kdata/Noise, complex64 of Blitz lengths dims (mri_data.cpp:763-766, 888-893). */
int mrn_raw_noise_dims(mrn_raw *f, int64_t dims[2]);
int mrn_raw_read_noise(mrn_raw *f, void *noise);
int mrn_raw_write_noise(mrn_raw *f, const int64_t dims[2], const void *noise);

/* Fill the fields of p that gridFFT::precalc_gridding takes from MRI_DATA
 * (gridFFT.cpp:484-498): grid_in = trajectory_type == NONCARTESIAN, fft_in =
 * dft_needed, and N = recon_N (explicit: read_external_data never sets
 * recon_res).  Other fields of p are left as the caller set them
 * (mrn_params_default first). */
int mrn_raw_params(const mrn_raw_info *info, const int32_t recon_N[3], mrn_params *p);

const char *mrn_raw_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MRN_IO_H */
