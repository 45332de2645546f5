// mrn_gridfft.hpp — C++ facade with the reference's gridFFT call surface over the C ABI.
//
// mrn::gridFFT mirrors uwmri/mri_recon's class gridFFT (src/gridFFT.h:28-194):
// the same public knobs (overgrid, dwinX/Y/Z, grid_x/y/z, grid_scale_x/y/z,
// kernel_type, k_rad), read_commandline / precalc_gridding, and the per-coil
// forward / backward / backward_residual / forward_sos calls with and without a
// coil map, in the reference's argument order.  Arrays are HIP device pointers
// (complex<float> interleaved, x fastest) instead of Blitz arrays, so each call
// also takes the sample count M; the image size is the one given to
// precalc_gridding.  As in the reference, "forward" is A^H (k-space to image,
// X accumulates) and "backward" is A (image to k-space, data overwritten).
//
// Errors: the reference's methods return void and print or exit; these throw
// mrn::Error (status code and message) and never exit.
//
// Beside the operator: mrn::dcf_calc (RECON::dcf_calc), mrn::LowRankFrames
// (LOWRANKCOIL's update_threshold + thresh over frames) and mrn::GatedFrames
// (GATING's time scaling and RECON::pregate_data).
//
// mrn::SenseOperator is the batched form: all coils (and gated frames) of one
// acquisition resident on the device, binned gridding, and the SENSE-CG /
// power-method / FISTA loops of recon_lib.cpp (1523-2212) with FP64 scalars on
// the device and the iteration captured in a hipGraph.
#pragma once
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "mrn.h"

namespace mrn {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string &msg) : std::runtime_error(msg), status(status) {}
  int status;
};

// op = nullptr: the message of the calling thread's last op-less failure
inline void check(int s, const mrn_op *op, const char *what) {
  if (s == MRN_OK) return;
  std::string m = std::string(what) + ": " + mrn_status_string(s);
  const char *le = mrn_last_error(op);
  if (le && le[0]) m += std::string(" (") + le + ")";
  throw Error(s, m);
}

using cfloat = std::complex<float>;

// Trajectory flags precalc_gridding reads from MRI_DATA (gridFFT.cpp:484-498):
// trajectory_type NONCARTESIAN -> gridding along the axis, dft_needed -> FFT.
struct TrajectoryInfo {
  int noncartesian[3] = {1, 1, 1};  // x, y, z
  int dft_needed[3] = {1, 1, 1};
};

class gridFFT {
 public:
  // public knobs (gridFFT.h:30-105); set them before precalc_gridding,
  // except k_rad, which may change between calls as in recon_lib.cpp:2738-2786
  float overgrid;
  float dwinX, dwinY, dwinZ;
  float grid_x, grid_y, grid_z;
  float grid_scale_x, grid_scale_y, grid_scale_z;
  int kernel_type;
  float k_rad;
  int kernel_mode = MRN_KMODE_TABLE;
  // geometry after precalc_gridding (gridFFT.h:53-82)
  int Nx = 0, Ny = 0, Nz = 0;
  int Sx = 0, Sy = 0, Sz = 0;
  int og_sx = 0, og_sy = 0, og_sz = 0;
  float betaX = 0.f, betaY = 0.f, betaZ = 0.f;
  // HIP stream the calls are queued on (nullptr = the null stream)
  void *stream = nullptr;

  gridFFT() {
    mrn_params_default(&p_);
    from_params();
  }
  ~gridFFT() {
    if (op_) mrn_op_destroy(op_);
  }
  gridFFT(const gridFFT &) = delete;
  gridFFT &operator=(const gridFFT &) = delete;

  // gridFFT::read_commandline (gridFFT.cpp:187-245)
  void read_commandline(int numarg, const char **pstring) {
    to_params();
    check(mrn_parse_args(&p_, numarg, pstring), nullptr, "read_commandline");
    from_params();
  }

  // gridFFT::precalc_gridding (gridFFT.cpp:475-643): geometry, tables,
  // windows and the coil grid on the device
  void precalc_gridding(int nz, int ny, int nx, const TrajectoryInfo &t = TrajectoryInfo()) {
    to_params();
    p_.N[0] = nx;
    p_.N[1] = ny;
    p_.N[2] = nz;
    for (int a = 0; a < 3; a++) {
      p_.grid_in[a] = t.noncartesian[a];
      p_.fft_in[a] = t.dft_needed[a];
    }
    p_.ncoils = 1;
    p_.form = MRN_FORM_PLAIN;
    if (op_) mrn_op_destroy(op_);
    op_ = nullptr;
    check(mrn_op_create(&p_, &op_), nullptr, "precalc_gridding");
    int32_t S[3], og[3];
    float beta[3];
    check(mrn_get_geometry(op_, S, og, beta), op_, "precalc_gridding");
    Nx = nx, Ny = ny, Nz = nz;
    Sx = S[0], Sy = S[1], Sz = S[2];
    og_sx = og[0], og_sy = og[1], og_sz = og[2];
    betaX = beta[0], betaY = beta[1], betaZ = beta[2];
    applied_k_rad_ = k_rad;
  }

  // ---- with a coil map (gridFFT.h:118-141)
  // X += A^H (kw * taper * data) for one coil (gridFFT::forward, gridFFT.cpp:759-812)
  void forward(cfloat *X, const cfloat *smap, const cfloat *data, const float *kx, const float *ky,
               const float *kz, const float *kw, int64_t M) {
    sync_k_rad();
    check(mrn_coil_forward(op(), X, smap, data, M, kx, ky, kz, kw, stream), op_, "forward");
  }
  // data = A X for one coil (gridFFT::backward, gridFFT.cpp:824-844)
  void backward(const cfloat *X, const cfloat *smap, cfloat *data, const float *kx, const float *ky,
                const float *kz, const float *kw, int64_t M) {
    check(mrn_coil_backward(op(), X, smap, data, M, kx, ky, kz, kw, nullptr, stream), op_, "backward");
  }
  // data = A X - dataSub (gridFFT::backward_residual, gridFFT.cpp:856-876)
  void backward_residual(const cfloat *X, const cfloat *smap, cfloat *data, const float *kx,
                         const float *ky, const float *kz, const float *kw, int64_t M,
                         const cfloat *dataSub) {
    check(mrn_coil_backward(op(), X, smap, data, M, kx, ky, kz, kw, dataSub, stream), op_,
          "backward_residual");
  }

  // ---- without a coil map (gridFFT.h:144-156)
  void forward(cfloat *X, const cfloat *data, const float *kx, const float *ky, const float *kz,
               const float *kw, int64_t M) {
    forward(X, nullptr, data, kx, ky, kz, kw, M);
  }
  void backward(const cfloat *X, cfloat *data, const float *kx, const float *ky, const float *kz,
                const float *kw, int64_t M) {
    backward(X, nullptr, data, kx, ky, kz, kw, M);
  }
  // X += |w * img|^2 (gridFFT::forward_sos, gridFFT.cpp:722-748)
  void forward_sos(cfloat *X, const cfloat *data, const float *kx, const float *ky, const float *kz,
                   const float *kw, int64_t M) {
    sync_k_rad();
    check(mrn_coil_forward_sos(op(), X, data, M, kx, ky, kz, kw, stream), op_, "forward_sos");
  }

  // Kernel table / deapodisation window of one axis (grid_filterX/Y/Z, winx/y/z)
  std::vector<float> table(int axis) const { return host_vec(mrn_get_table, axis); }
  std::vector<float> window(int axis) const { return host_vec(mrn_get_window, axis); }

  void sync() { check(mrn_sync(op(), stream), op_, "sync"); }
  mrn_op *handle() const { return op_; }

 private:
  mrn_params p_{};
  mrn_op *op_ = nullptr;
  float applied_k_rad_ = 0.f;

  mrn_op *op() const {
    if (!op_) throw Error(MRN_ESTATE, "gridFFT: call precalc_gridding first");
    return op_;
  }
  void sync_k_rad() {
    if (k_rad != applied_k_rad_) {
      check(mrn_set_k_rad(op(), k_rad, stream), op_, "k_rad");
      applied_k_rad_ = k_rad;
    }
  }
  void from_params() {
    overgrid = p_.overgrid;
    dwinX = p_.dwin[0], dwinY = p_.dwin[1], dwinZ = p_.dwin[2];
    grid_x = p_.grid[0], grid_y = p_.grid[1], grid_z = p_.grid[2];
    grid_scale_x = p_.grid_scale[0], grid_scale_y = p_.grid_scale[1], grid_scale_z = p_.grid_scale[2];
    kernel_type = p_.kernel;
    k_rad = p_.k_rad;
  }
  void to_params() {
    p_.overgrid = overgrid;
    p_.dwin[0] = dwinX, p_.dwin[1] = dwinY, p_.dwin[2] = dwinZ;
    p_.grid[0] = grid_x, p_.grid[1] = grid_y, p_.grid[2] = grid_z;
    p_.grid_scale[0] = grid_scale_x, p_.grid_scale[1] = grid_scale_y, p_.grid_scale[2] = grid_scale_z;
    p_.kernel = kernel_type;
    p_.k_rad = k_rad;
    p_.kernel_mode = kernel_mode;
  }
  template <typename F>
  std::vector<float> host_vec(F fn, int axis) const {
    int n = 0;
    check(fn(op(), axis, nullptr, 0, &n), op_, "table");
    std::vector<float> v((size_t)n);
    check(fn(op(), axis, v.data(), n, &n), op_, "table");
    return v;
  }
};

// Batched operator over all coils and frames of one acquisition.
class SenseOperator {
 public:
  SenseOperator(const mrn_params &p) { check(mrn_op_create(&p, &op_), nullptr, "SenseOperator"); }
  ~SenseOperator() {
    if (op_) mrn_op_destroy(op_);
  }
  SenseOperator(const SenseOperator &) = delete;
  SenseOperator &operator=(const SenseOperator &) = delete;

  void *stream = nullptr;

  void set_traj(int frame, int64_t M, const float *kx, const float *ky, const float *kz,
                const float *kw) {
    check(mrn_set_traj(op_, frame, M, kx, ky, kz, kw, stream), op_, "set_traj");
  }
  void set_smaps(const cfloat *smaps) { check(mrn_set_smaps(op_, smaps, stream), op_, "set_smaps"); }
  void set_kdata(int frame, const cfloat *data) {
    check(mrn_set_kdata(op_, frame, data, stream), op_, "set_kdata");
  }
  // data[C][M] = A X (minus sub when given)
  void image_to_kspace(int frame, const cfloat *X, cfloat *data, const cfloat *sub = nullptr) {
    check(mrn_image_to_kspace(op_, frame, X, data, sub, stream), op_, "image_to_kspace");
  }
  // X (+)= sum_c A_c^H (kw taper data_c)
  void kspace_to_image(int frame, const cfloat *data, cfloat *X, bool accumulate = true) {
    check(mrn_kspace_to_image(op_, frame, data, X, accumulate ? 1 : 0, stream), op_, "kspace_to_image");
  }
  void normal(int frame, const cfloat *X, cfloat *out) {
    check(mrn_normal(op_, frame, X, out, stream), op_, "normal");
  }
  // SENSE-CG over all frames; returns the per-iteration trace (4 doubles each)
  std::vector<double> cg(cfloat *X, int iters, bool use_graph = true) {
    std::vector<double> tr(4 * (size_t)(iters > 0 ? iters : 1));
    check(mrn_cg(op_, X, iters, use_graph ? 1 : 0, tr.data(), stream), op_, "cg");
    tr.resize(4 * (size_t)(iters > 0 ? iters : 0));
    return tr;
  }
  double power_method(int iters = 30) {
    double step = 0.0;
    check(mrn_power_method(op_, iters, &step, stream), op_, "power_method");
    return step;
  }
  std::vector<double> fista(cfloat *X, int iters, double step, bool use_graph = true) {
    std::vector<double> r((size_t)(iters > 0 ? iters : 1));
    check(mrn_fista(op_, X, iters, step, use_graph ? 1 : 0, r.data(), stream), op_, "fista");
    r.resize((size_t)(iters > 0 ? iters : 0));
    return r;
  }
  void autotune(int frame = 0, const char *cache_path = nullptr) {
    check(mrn_autotune(op_, frame, cache_path, stream), op_, "autotune");
  }
  void sync() { check(mrn_sync(op_, stream), op_, "sync"); }
  mrn_op *handle() const { return op_; }

 private:
  mrn_op *op_ = nullptr;
};

// Pipe density compensation of one encoding (RECON::dcf_calc,
// recon_lib.cpp:854-917): kw (device, M) receives the weights; returns the
// per-iteration sums of |kw| the reference prints.
inline std::vector<double> dcf_calc(const mrn_dcf_params &p, int64_t M, const float *kx, const float *ky,
                                    const float *kz, float *kw, void *stream = nullptr) {
  std::vector<double> sums(p.iters > 0 ? p.iters : 1);
  check(mrn_dcf_calc(&p, M, kx, ky, kz, kw, sums.data(), stream), nullptr, "dcf_calc");
  sums.resize(p.iters > 0 ? p.iters : 0);
  return sums;
}

// LOWRANKCOIL for an image of Nt x Ne frames (clear.cpp:233-556): each call is
// update_threshold(X, dim, iter) followed by thresh(X, dim), as recon_lib does
// after a gradient step (recon_lib.cpp:2117-2126).  The call counter numbers
// the block-shift draws.
class LowRankFrames {
 public:
  mrn_llr_params p;
  void *stream = nullptr;
  LowRankFrames() { mrn_llr_params_default(&p); }
  LowRankFrames(int numarg, const char **pstring) {
    mrn_llr_params_default(&p);
    check(mrn_llr_parse_args(&p, numarg, pstring), nullptr, "LowRankFrames");
  }
  // returns smax (the median block singular value times the iteration factor)
  float threshold(cfloat *X, int Nt, int Ne, int Nx, int Ny, int Nz, int dim, int iter) {
    const int32_t N[3] = {Nx, Ny, Nz};
    float smax = 0.f;
    check(mrn_llr(&p, X, Nt, Ne, N, dim, iter, calls_++, &smax, stream), nullptr, "LowRankFrames::threshold");
    return smax;
  }

 private:
  int64_t calls_ = 0;
};

// GATING + RECON::pregate_data for one encoding: scale the per-readout gate
// times on the host, then select each frame's samples on the device.
class GatedFrames {
 public:
  mrn_gate_params p;
  void *stream = nullptr;
  double offset = 0.0, scale = 0.0, temporal_resolution = 0.0;
  GatedFrames() { mrn_gate_params_default(&p); }
  GatedFrames(int numarg, const char **pstring) {
    mrn_gate_params_default(&p);
    check(mrn_gate_parse_args(&p, numarg, pstring), nullptr, "GatedFrames");
  }
  // gating.cpp:880-1078, in place on host times (one per readout)
  void scale_times(std::vector<double> &times) {
    double info[3];
    check(mrn_gate_scale_times(&p, (int64_t)times.size(), times.data(), info), nullptr, "scale_times");
    offset = info[0];
    scale = info[1];
    temporal_resolution = info[2];
  }
  // samples per frame (outputs null) or the frame-major outputs (sized by the
  // sum of the counts of a first call); device pointers throughout
  std::vector<int64_t> pregate(int64_t xres, int64_t R, const double *gt, const float *resp, const float *kx,
                               const float *ky, const float *kz, const float *kw, int C, const cfloat *kdata,
                               float *kx_out = nullptr, float *ky_out = nullptr, float *kz_out = nullptr,
                               float *kw_out = nullptr, cfloat *kdata_out = nullptr) {
    std::vector<int64_t> counts(p.frames > 0 ? p.frames : 1);
    check(mrn_pregate(&p, xres, R, gt, resp, kx, ky, kz, kw, C, kdata, counts.data(), kx_out, ky_out, kz_out,
                      kw_out, kdata_out, stream),
          nullptr, "pregate");
    counts.resize(p.frames > 0 ? p.frames : 0);
    return counts;
  }
};

}  // namespace mrn
