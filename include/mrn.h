/*
 * mrn.h — C ABI of the MI355X-native NUFFT/SENSE encoding operator.
 *
 * Drop-in boundary for uwmri/mri_recon's gridFFT (src/gridFFT.h:28-194) and the
 * SENSE-CG / FISTA loops that call it (src/recon_lib.cpp:1523-2212).
 *
 * Conventions (those of the reference):
 *  - arrays are column-major with x fastest (Blitz ColumnMajorArray<3>): an
 *    image is complex64 [Nz][Ny][Nx] in C order; coil maps [C][Nz][Ny][Nx];
 *    k-space of one frame [C][M]; a trajectory is four float32 [M] arrays
 *    kx, ky, kz, kw in the reference's units (cycles/FOV, |k| <= N/2,
 *    src/mri_data.h:50-57).  complex64 = interleaved float re, im.
 *  - "image_to_kspace" is A (the reference's gridFFT::backward, a gather);
 *    "kspace_to_image" is A^H (the reference's gridFFT::forward, a scatter).
 *  - every pointer passed to an apply call is a HIP device pointer; the op owns
 *    its grids, bins, tables and workspaces, and apply calls allocate nothing.
 *  - every call returns an mrn_status and never exits the process.
 *  - calls are asynchronous on the given stream (NULL = the null stream);
 *    one op must not be used from two streams at once.  Solver calls with
 *    use_graph = 1 replay a captured hipGraph on the op's own stream, fenced
 *    with events against the caller's stream.
 */
#ifndef MRN_H
#define MRN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  MRN_OK = 0,
  MRN_EINVAL = 1,       /* bad argument (null pointer, size, frame index) */
  MRN_ENOMEM = 2,       /* device allocation failed */
  MRN_EHIP = 3,         /* a HIP runtime call failed */
  MRN_EUNSUPPORTED = 4, /* valid request this build does not implement */
  MRN_ESTATE = 5        /* call order (e.g. apply before set_traj) */
} mrn_status;

/* Kernel types, same values as the reference enum (src/gridFFT.h:19-22). */
enum { MRN_KERNEL_TRIANGLE = 0, MRN_KERNEL_KAISER = 1, MRN_KERNEL_SINC = 2, MRN_KERNEL_POLY = 3 };
/* Kernel evaluation: the reference's 600-point lookup table with linear
 * interpolation, or direct Kaiser-Bessel evaluation (smooth in k, used for
 * trajectory gradients). */
enum { MRN_KMODE_TABLE = 0, MRN_KMODE_EXACT_KB = 1 };
/* Execution form: batched (all coils resident, binned gridding, fused and
 * pruned FFT passes) or plain (one coil grid, the reference's per-coil
 * structure on the GPU). */
enum { MRN_FORM_BATCHED = 0, MRN_FORM_PLAIN = 1 };

typedef struct mrn_params {
  int32_t N[3];          /* image size Nx, Ny, Nz (precalc_gridding, gridFFT.cpp:475) */
  float overgrid;        /* -overgrid, default 1.5 (gridFFT.cpp:32) */
  float grid[3];         /* -grid_x/y/z; -1 = 16*ceil(overgrid*N/16)/N (503) */
  float dwin[3];         /* -dwinX/Y/Z kernel half width; -1 = kernel default */
  int32_t kernel;        /* MRN_KERNEL_* (-kaiser -triangle -sinc -poly_kernel) */
  int32_t grid_in[3];    /* 1 = non-Cartesian axis (trajectory_type, 484-486) */
  int32_t fft_in[3];     /* 1 = DFT along the axis (dft_needed, 488-498) */
  float grid_scale[3];   /* -grid_scale_x/y/z (343-345) */
  float k_rad;           /* A^H taper radius, default 9999 (51, 1096-1108) */
  int32_t ncoils;        /* coils C */
  int32_t nframes;       /* frames (gated bins / encodes), >= 1 */
  int32_t kernel_mode;   /* MRN_KMODE_* */
  int32_t form;          /* MRN_FORM_* */
  int32_t coil_chunk;    /* coils per grid cell block (1,2,4,8,16); 0 = auto */
  int32_t tile[3];       /* scatter tile edge (x,y,z): 0 (auto) or 4; other edges are MRN_EUNSUPPORTED */
  int32_t max_visits;    /* samples per scatter work item; 0 = auto */
  int32_t bin_edge;      /* sample sort bin edge in cells (1 or 2); 0 = auto */
  int32_t xcd_run;       /* scatter work items per XCD run (1 = hardware order); 0 = auto;
                            -1 = each XCD takes one contiguous range of the work list, of equal cost */
  int32_t xcd_run_gather;/* gather workgroups per XCD run; 0 = auto */
  int32_t gather_tiles;  /* dense tiles of the gather: 1 = stage the grid in LDS, 2 = box matrix-core
                          * gather (16-coil blocks; default 0 = direct gather) */
  int32_t scatter_valu;  /* 1 = vector-ALU scatter accumulate instead of the matrix cores (default 0) */
  int32_t fft_stockham;  /* 1 = LDS Stockham FFT passes instead of the four-step register form (default 0) */
  int32_t xcd_run_fft;   /* y and z FFT pass workgroups per XCD run (1 = hardware order); 0 = auto (8) */
  int32_t gather_coils;  /* coils per lane in the gather: 1, 2, 4 or 8; 0 = auto */
  int32_t fuse_frames;   /* multi-frame solver loops (batched form): 1 = one launch per stage over every
                          * frame (F grids resident), 0 = frame by frame, -1 (default) = fuse when the
                          * F grids take at most half the free device memory */
} mrn_params;

typedef struct mrn_op mrn_op;

/* Defaults of gridFFT::gridFFT() (gridFFT.cpp:28-60) for a 3-D
 * non-Cartesian acquisition. */
void mrn_params_default(mrn_params *p);

/* gridFFT::read_commandline (gridFFT.cpp:187-245): the same flags
 * (-overgrid -dwinX/Y/Z -grid_x/y/z -grid_scale_x/y/z -grid_in_x/y/z
 * -fft_in_x/y/z -kaiser -triangle -sinc -poly_kernel -fast_grid). */
int mrn_parse_args(mrn_params *p, int argc, const char **argv);

/* gridFFT::precalc_gridding + alloc_grid (gridFFT.cpp:66-75, 475-643):
 * geometry, kernel tables, deapodisation windows, device buffers. */
int mrn_op_create(const mrn_params *p, mrn_op **out);
int mrn_op_destroy(mrn_op *op);

/* Sx,Sy,Sz; og_sx,og_sy,og_sz (gridFFT.h:53-75); betaX/Y/Z (80-82). */
int mrn_get_geometry(const mrn_op *op, int32_t S[3], int32_t og[3], float beta[3]);
/* Host copies of the kernel table (grid_filterX/Y/Z, gridFFT.h:94-96) and
 * deapodisation window (winx/y/z, 85-87) of one axis; *len receives the
 * length; returns MRN_EINVAL when cap is too small. */
int mrn_get_table(const mrn_op *op, int axis, float *host_out, int cap, int *len);
int mrn_get_window(const mrn_op *op, int axis, float *host_out, int cap, int *len);

/* Trajectory of one frame (the kx, ky, kz, kw arguments of gridFFT::forward /
 * backward, gridFFT.h:118-141).  Sorts samples into tiles once; kept until
 * replaced.  Device pointers, M samples.  Waits for the device first; a
 * frame's arrays are reused when the new trajectory fits in them.  Stored
 * k-space data (mrn_set_kdata) is re-sorted for the new trajectory when M is
 * unchanged and dropped otherwise. */
int mrn_set_traj(mrn_op *op, int frame, int64_t M, const float *kx, const float *ky,
                 const float *kz, const float *kw, void *stream);
/* Coil sensitivities smaps[C][Nz][Ny][Nx] (the smap argument, gridFFT.h:119). */
int mrn_set_smaps(mrn_op *op, const void *smaps, void *stream);

/* A for every coil of one frame: data[C][M] = A_c X, overwriting data
 * (gridFFT::backward, gridFFT.cpp:824-844); with sub != NULL,
 * data = A_c X - sub (backward_residual, 856-876).  Samples with kw == 0 or
 * outside the grid give 0 (chop_grid_backward, 1196-1271). */
int mrn_image_to_kspace(mrn_op *op, int frame, const void *X, void *data, const void *sub,
                        void *stream);
/* A^H for every coil of one frame: X (+)= sum_c A_c^H (kw * taper * data_c)
 * (gridFFT::forward, gridFFT.cpp:791-812, summed over the caller's coil loop,
 * recon_lib.cpp:1565-1567).  accumulate = 1 reproduces the reference's +=. */
int mrn_kspace_to_image(mrn_op *op, int frame, const void *data, void *X, int accumulate,
                        void *stream);
/* out = sum_c A_c^H W A_c X for one frame (the E'E of recon_lib.cpp:1614-1622),
 * k-space kept on chip in the op's internal sample order. */
int mrn_normal(mrn_op *op, int frame, const void *X, void *out, void *stream);

/* One coil, one grid, trajectory per call: gridFFT::backward /
 * backward_residual / forward with a smap (gridFFT.h:118-141), or without
 * one when smap is NULL (gridFFT.h:144-156).  These run the plain per-coil
 * form (no binning); X accumulates in forward. */
int mrn_coil_backward(mrn_op *op, const void *X, const void *smap, void *data, int64_t M,
                      const float *kx, const float *ky, const float *kz, const float *kw,
                      const void *sub, void *stream);
int mrn_coil_forward(mrn_op *op, void *X, const void *smap, const void *data, int64_t M,
                     const float *kx, const float *ky, const float *kz, const float *kw,
                     void *stream);

/* X += |w * img|^2 for one coil without a coil map: gridFFT::forward_sos
 * (gridFFT.cpp:722-748, accumulate_sos 960-970), the sum-of-squares combine. */
int mrn_coil_forward_sos(mrn_op *op, void *X, const void *data, int64_t M, const float *kx,
                         const float *ky, const float *kz, const float *kw, void *stream);

/* Change the A^H taper radius (the public member gridFFT::k_rad,
 * gridFFT.h:102, which the reference's callers change between calls,
 * recon_lib.cpp:2738-2786); k_rad > 0, INFINITY disables the taper. */
int mrn_set_k_rad(mrn_op *op, float k_rad, void *stream);

/* Measured k-space of one frame data[C][M] for the solvers below. */
int mrn_set_kdata(mrn_op *op, int frame, const void *data, void *stream);

/* SENSE-CG (recon_lib.cpp:1523-1759) over all frames: X[F][Nz][Ny][Nx] is
 * zeroed and receives the image after `iters` iterations.  FP32 vectors, FP64
 * inner products and step sizes on the device.  use_graph = 1 captures one
 * iteration in a hipGraph and replays it.  trace (host, nullable) receives
 * 4 doubles per iteration: sum|R|^2, Re alpha, Im alpha, sum|R_new|^2. */
int mrn_cg(mrn_op *op, void *X, int iters, int use_graph, double *trace, void *stream);
/* The same loop in two parts: mrn_cg_init sets X = 0, R = P = sum E^H W d;
 * mrn_cg_iterate continues the iteration from the op's state for `iters`
 * more steps (the hipGraph of one step is captured once and cached until the
 * trajectory, maps or data change).  trace as in mrn_cg for these steps. */
int mrn_cg_init(mrn_op *op, void *X, void *stream);
int mrn_cg_iterate(mrn_op *op, void *X, int iters, int use_graph, double *trace, void *stream);

/* Power method for the FISTA step (recon_lib.cpp:1791-1915): iters
 * applications of E'E starting from X = -0.5-0.5i; *step = 1/max_eig. */
int mrn_power_method(mrn_op *op, int iters, double *step, void *stream);
/* FISTA without regulariser (recon_lib.cpp:1925-2105, fista_update
 * 2177-2212): X[F][Nz][Ny][Nx] starts at 0; resid (host, nullable) receives
 * sum|R|^2 per iteration. */
int mrn_fista(mrn_op *op, void *X, int iters, double step, int use_graph, double *resid,
              void *stream);

/* 3-D periodic wavelet of `frames` images X[frames][Nz][Ny][Nx] in place
 * (WAVELET3D::forward3D / backward3D, wavelet3D.cpp:563-631, with the
 * reference's default 4-tap Daubechies pair, 143-190): levels per axis
 * (x, y, z) are lowered until N / 2^L is an integer (99-115); shift[a] in
 * 0..3 is the axis' circular shift (the cycle-spin shift, 443-449). */
int mrn_wavelet(mrn_op *op, void *X, int frames, const int32_t levels[3], const int32_t shift[3],
                int inverse, void *stream);

/* The L1 wavelet proximal step RECON::L1_threshold (recon_lib.cpp:2416-2527)
 * with the fraction threshold (THRESHOLD, threshold.cpp:193-380, 766-800) on
 * X[frames][Nz][Ny][Nx]: per cycle spin, forward wavelet, the threshold taken
 * on the first spin over all frames (scaled by 1/spins), complex soft
 * threshold, backward wavelet.  shifts (host, [spins][3], nullable = drawn
 * from the op's seeded stream); *threshold (host, nullable) receives the
 * applied threshold. */
int mrn_l1_wavelet(mrn_op *op, void *X, int frames, double fraction, const int32_t levels[3], int spins,
                   const int32_t *shifts, double *threshold, void *stream);

/* Make mrn_fista apply the L1 wavelet step after every gradient step
 * (recon_lib.cpp:2145-2150); fraction 0 turns it off.  Cycle-spin shifts of
 * iteration i are entries 3*spins*i .. 3*spins*(i+1)-1 of the splitmix64
 * stream from `seed`, each (value >> 33) % 4 (the reference uses rand() % 4,
 * whose sequence depends on every other rand() call of the program). */
int mrn_set_fista_l1(mrn_op *op, double fraction, const int32_t levels[3], int spins, uint64_t seed);
/* FISTA / IST step rule (-iterative_step_type, recon_lib.cpp:346-360;
 * enum values of RECON::IterativeStepType, recon_lib.h:129-130).  MAXEIG:
 * the fixed step mrn_fista is given (1/max_eig from mrn_power_method).
 * CAUCHY: at iterations below cauchy_update_number (-cauchy_update_number,
 * default 30) the step is recomputed as complex<float>(sum|R|^2 / sum
 * conj(R) E'E R) with R the gradient of all frames, both sums complex
 * double (recon_lib.cpp:2013-2041, 2093-2100); later iterations keep the
 * last one; mrn_fista's step argument is used only before the first update. */
enum { MRN_STEP_CAUCHY = 0, MRN_STEP_MAXEIG = 1 };
int mrn_set_fista_step(mrn_op *op, int step_type, int cauchy_update_number);
/* Host copy of the complex step (re, im; 2 doubles per iteration) of the
 * first n iterations of the last mrn_fista with the Cauchy rule. */
int mrn_get_fista_steps(mrn_op *op, double *out, int n, void *stream);

/* Host copy of the threshold the last mrn_fista applied at each of its first
 * n iterations. */
int mrn_get_fista_thresholds(mrn_op *op, double *out, int n, void *stream);

/* Trajectory gradient: given X and the upstream gradient g[C][M] of the
 * k-space output y = A(X; k) of one frame (g = dL/dRe y + i dL/dIm y for a
 * real loss L), writes grad_k[M][3] (float32, reference k units)
 * = Re sum_c conj(g_cm) dy_cm/dk = dL/dk.  The oversampled grid of X is
 * recomputed, except straight after mrn_image_to_kspace of the same frame,
 * image pointer and stream with no other call on the op in between (and
 * mrn_get_grid never called): then the grid that call left is used, so X's
 * contents must not change in between.  No counterpart in the reference (its
 * trajectory is fixed). */
int mrn_traj_grad(mrn_op *op, int frame, const void *X, const void *g, float *grad_k,
                  void *stream);

/* Launch shapes of the batched form (mrn_get_tuning). */
typedef struct mrn_tuning {
  int32_t bin_edge;       /* sample bin edge in cells */
  int32_t max_visits;     /* samples per scatter work item */
  int32_t xcd_run;        /* scatter work items per XCD run */
  int32_t xcd_run_gather; /* gather workgroups per XCD run */
  int32_t gather_coils;   /* coils per lane in the gather */
  int32_t xcd_run_fft;    /* y and z FFT pass workgroups per XCD run */
  int32_t fft_stockham;   /* 1: LDS Stockham passes; 0: four-step passes with support pruning */
  int32_t scatter_valu;   /* 1: vector-ALU tile accumulate; 0: matrix cores */
  int32_t tile[3];        /* scatter tile edge (4, 4, 4) */
  int32_t from_cache;     /* 1: the last mrn_autotune applied a cached choice without timing */
  float normal_ms;        /* E'E time of the tuned frame with the choice (timed or cached) */
  int32_t gather_tiles;   /* 0: direct gather; 2: box matrix-core gather of dense tiles beside it */
  int32_t reserved[2];
} mrn_tuning;

/* Time the batched form's launch shapes on one frame's trajectory and keep
 * the fastest: scatter bin edge x samples per work item x accumulate form,
 * gather coils per lane x XCD run length, the FFT passes' XCD run length,
 * and the FFT pass form (four-step with pruning or Stockham).  The gather
 * form (gather_tiles) stays the op's own.  The choice
 * is looked up in, or else added to, the JSON file cache_path (NULL: no
 * cache), keyed by sizes, kernel, sample and coil counts and device; a hit
 * applies it without timing.  On failure the previous shapes are restored.
 * Every candidate computes the same operator up to FP32 summation order.
 * No counterpart in the reference. */
int mrn_autotune(mrn_op *op, int frame, const char *cache_path, void *stream);

/* The launch shapes in use.  No counterpart in the reference. */
int mrn_get_tuning(const mrn_op *op, mrn_tuning *out);

/* Average device time (ms, HIP events on `stream`) of each batched stage of
 * one frame over `reps` runs: [0] set_image+z FFT, [1] y FFT, [2] x FFT,
 * [3] gather (with the k-space weighting of E'E), [4] zeroing of the tiles
 * that several scatter work items share, [5] tile scatter, [6] x IFFT,
 * [7] y IFFT, [8] z IFFT+accumulate.  Each stage is one kernel per coil block
 * or one kernel over all blocks.  Diagnostic; no counterpart in the reference
 * (its -time_grid prints, gridFFT.cpp:727-747, time the same stages on the
 * host). */
#define MRN_NSTAGES 9
int mrn_stage_times(mrn_op *op, int frame, const void *X, int reps, float ms[MRN_NSTAGES],
                    void *stream);

/* Work lists of one frame's batched gridding (diagnostic; no counterpart
 * in the reference). */
typedef struct mrn_frame_stats {
  int64_t samples;         /* M */
  int64_t valid;           /* samples with kw != 0 inside the grid */
  int64_t tile_visits;     /* sample visits over the scatter's 4^3 tiles */
  int64_t scatter_atomics; /* float atomics of one scatter over all coils, upper bound
                              (split tiles' flushes and window-edge samples; zero sums skip theirs) */
  int32_t scatter_items;   /* scatter work items */
  int32_t split_items;     /* work items of tiles split over several items (flushed with atomics) */
  int32_t split_tiles;     /* tiles split over several work items */
  int32_t wide;            /* samples with one tap more than usual on some axis */
  int32_t gather_items;    /* tile-staged gather work items */
  int32_t reserved[7];
} mrn_frame_stats;
int mrn_get_frame_stats(mrn_op *op, int frame, mrn_frame_stats *out);

/* k-space support of one frame (host copies): xrng [Sz][Sy][2] receives,
 * per grid row (z, y), the first and last x cell that some sample's taps
 * reach (first > last: none; INT_MAX, INT_MIN: the row is outside plane z's
 * rows of support); yrng (nullable) [Sz][2] the first and last row y of
 * plane z with support.  The forward FFT passes store, and the inverse ones
 * load, only these cells; without pruning every cell is listed.
 * Diagnostic; no counterpart in the reference. */
int mrn_get_support(mrn_op *op, int frame, int32_t *xrng, int32_t *yrng);

/* The op's oversampled k-space grids (the reference's public member
 * gridFFT::k3d_grid, gridFFT.h:30, which RECON::kspace_residual zeroes and
 * fills through chop_grid_forward / chop_grid_backward, recon_lib.cpp:
 * 2244-2252): *grid receives the device pointer of complex64
 * [nch][Sz][Sy][Sx][CC], coil c at block c / CC, lane c % CC; dims =
 * {nch, CC, Sx, Sy, Sz}.  The plain form has one grid of one coil.  The
 * contents between calls are the last call's working state; callers may
 * read or write them between chop_grid calls. */
int mrn_get_grid(mrn_op *op, void **grid, int64_t dims[5]);
/* k3d_grid = 0 for every coil (recon_lib.cpp:2244). */
int mrn_grid_zero(mrn_op *op, void *stream);
/* gridFFT::chop_grid_forward (gridFFT.cpp:1009-1164) for every coil of
 * frame f: grid_c += the scatter of kw * taper * data_c (data [C][M]), added
 * to the grids' contents; no FFT.  Plain form: C = 1. */
int mrn_chop_grid_forward(mrn_op *op, int frame, const void *data, void *stream);
/* gridFFT::chop_grid_backward (gridFFT.cpp:1166-1322) for every coil of
 * frame f: data_c = the gather of grid_c at the samples (minus sub when not
 * NULL); kw == 0 and out-of-grid samples give 0; no FFT.  Plain form: C = 1. */
int mrn_chop_grid_backward(mrn_op *op, int frame, void *data, const void *sub, void *stream);

/* Wait for the work queued on `stream` (NULL = the null stream). */
int mrn_sync(mrn_op *op, void *stream);

/* ---- Pipe density compensation (next to the operator) ------------------
 * RECON::dcf_calc (recon_lib.cpp:854-917) with DCFgridFFT's polynomial
 * kernel (DCFgridFFT.cpp:56-440): kw = 1, then `iters` times
 * kw <- kw / (A^T A kw) (samples with |A^T A kw| < 1e-3 get 0), A the real
 * gridding onto a trunc(os*N + 16)-cell grid; finally kw = |kw * s^3| with
 * s = |k|/128*(acc - 1) + 1 (scale_kw, DCFgridFFT.cpp:291-312). */
enum { MRN_DCF_POLY = 0, MRN_DCF_POLY2 = 1 }; /* DCFgridFFT::POLY_KERNEL, POLY_KERNEL2 */
typedef struct mrn_dcf_params {
  int32_t N[3];    /* rcxres, rcyres, rczres (recon_lib.cpp:873-875) */
  int32_t iters;   /* -dcf_iter, default 20 (recon_lib.cpp:126, 417) */
  float dwin;      /* -dcf_dwin, default 2.1: kernel half width in cells */
  float scale;     /* -dcf_scale, default 1: k is multiplied by it for the DCF */
  float overgrid;  /* -dcf_overgrid, default 2.1 */
  float acc;       /* -dcf_acc, default 1: kernel widening with |k| */
  int32_t kernel;  /* MRN_DCF_POLY (what dcf_calc uses) or MRN_DCF_POLY2 */
  int32_t reserved[3];
} mrn_dcf_params;
void mrn_dcf_params_default(mrn_dcf_params *p);
/* -dcf_iter -dcf_dwin -dcf_scale -dcf_overgrid -dcf_acc (recon_lib.cpp:417-421);
 * other arguments are ignored. */
int mrn_dcf_parse_args(mrn_dcf_params *p, int argc, const char **argv);
/* Grid extents (Sx, Sy, Sz) = trunc(os*(Nz, Ny, Nx) + 16): the reference
 * declares the grid with the z size first (recon_lib.cpp:873). */
int mrn_dcf_grid_size(const mrn_dcf_params *p, int32_t S[3]);
/* One encoding's DCF (the body of recon_lib.cpp:879-913).  kx, ky, kz: device
 * [M] in the trajectory's units; kw: device [M], overwritten with the weights.
 * sums: optional host [iters], the printed sum of |kw| after each iteration.
 * The caller's trajectory is not modified (the reference scales it in place by
 * dcf_scale and back).  Needs 4*Sx*Sy*Sz bytes of device scratch for the call,
 * and returns when the weights are on the device.  Errors: mrn_last_error(NULL). */
int mrn_dcf_calc(const mrn_dcf_params *p, int64_t M, const float *kx, const float *ky, const float *kz,
                 float *kw, double *sums, void *stream);
/* Float atomics the grid scatters of the last mrn_dcf_calc on the current
 * device issued (all iterations).  Diagnostic; no counterpart in the
 * reference. */
int mrn_dcf_last_atomics(int64_t *count);

/* ---- Locally low-rank threshold over frames ------------------------------
 * LOWRANKCOIL::update_threshold + ::thresh (clear.cpp:233-556) as recon_lib
 * applies them after each gradient step (recon_lib.cpp:2117-2126), for an
 * image of Nt x Ne frames, X[(t*Ne + e)][Nz][Ny][Nx] complex64 on the device.
 * Blocks of b0 x b1 x b2 voxels (clamped to the volume; remainders untouched)
 * form (frames of the other index x voxels) x (frames of `dim`) matrices;
 * smax = median of the blocks' nonzero largest singular values times
 * 1 + amp*exp(-iter/t2); then block_iter passes, each with its own block
 * shift (periodic), soft-threshold every block's singular values by
 * alpha*smax/block_iter (alpha = alpha_time for dim 0, alpha_encode for 1;
 * 0 = update_threshold only, as the reference's early return).
 * Shifts: pass k of call c moves axis q by splitmix64(seed, 3*(c*block_iter
 * + k) + q) % b - b/2 (the reference's rand() % b - b/2 cannot be replayed). */
typedef struct mrn_llr_params {
  int32_t block[3];     /* -clear_nx/-clear_ny/-clear_nz, default 4 (clear.cpp:44-48) */
  int32_t block_iter;   /* -clear_block_iter, default 4 */
  float alpha_time;     /* -clear_alpha_time, default 0 */
  float alpha_encode;   /* -clear_alpha_encode, default 0 */
  float amp;            /* -clear_amp, default 4 */
  float t2;             /* -clear_t2, default 6 */
  uint64_t seed;        /* block shift stream */
  int32_t reserved[4];
} mrn_llr_params;
void mrn_llr_params_default(mrn_llr_params *p);
int mrn_llr_parse_args(mrn_llr_params *p, int argc, const char **argv);
/* One update_threshold + thresh call; N = {Nx, Ny, Nz}; at most 32 frames
 * along `dim` and about 4096 matrix entries per block.  smax (optional, host)
 * receives the threshold scale.  Returns when done; errors: mrn_last_error(NULL). */
int mrn_llr(const mrn_llr_params *p, void *X, int Nt, int Ne, const int32_t N[3], int dim, int iter,
            int64_t call, float *smax, void *stream);
/* mrn_llr without any device allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as mrn_llr; its scratch (8 bytes per block
 * plus about 2.5 KB) lives in a fixed 8 MiB device array, so calls with more
 * than about 1,048,000 blocks return MRN_EINVAL with the limit in
 * mrn_last_error(NULL).  Every call on one device shares that array: calls
 * that overlap on one device race. */
int mrn_llr_noalloc(const mrn_llr_params *p, void *X, int Nt, int Ne, const int32_t N[3], int dim, int iter,
                    int64_t call, float *smax, void *stream);
/* Time LLR (dim 0 over the op's frames) inside each FISTA iteration, after the
 * gradient step and before the L1 step; NULL or alpha_time = 0 turns it off.
 * The iteration index drives amp's decay and the shift stream's call index. */
int mrn_set_fista_llr(mrn_op *op, const mrn_llr_params *p);
/* smax of the last FISTA iteration's low-rank step. */
int mrn_get_fista_llr_smax(mrn_op *op, float *smax, void *stream);

/* ---- Gated frame selection --------------------------------------------
 * GATING::init_time_resolved's gate-time scaling (gating.cpp:880-1078) and
 * RECON::pregate_data (recon_lib.cpp:675-776) with GATING::weight_data /
 * hist_weight / tornado_weight (gating.cpp:1082-1200).  Enum values are the
 * reference's (gating.h:27-42). */
enum { MRN_VS_TORNADO = 0, MRN_VS_NONE = 1, MRN_VS_HIST = 2 };
enum { MRN_TORNADO_FLAT = 0, MRN_TORNADO_RADIAL = 1, MRN_TORNADO_VIPR = 2 };
enum { MRN_GATE_NONE = 0, MRN_GATE_RETRO_ECG = 1, MRN_GATE_ECG = 2, MRN_GATE_RESP = 3, MRN_GATE_TIME = 4,
       MRN_GATE_PREP = 5 };
typedef struct mrn_gate_params {
  int32_t gate_type;          /* -gating_type: sets how the time range is found */
  int32_t vs_type;            /* -viewshare_type tornado | none | hist (default none) */
  int32_t frames;             /* -rcframes */
  int32_t tornado_shape;      /* default VIPR (|k|^2 widening) */
  int32_t wdth_low, wdth_high;/* -vs_wdth_low/-vs_wdth_high, default 1 and 4 frames */
  float kmax;                 /* tornado |k| scale (the reference takes max |k| of encoding 0) */
  float min_quantile, max_quantile; /* -gate_min/max_quantile (respiratory range), 0 and 1 */
  int32_t ecg_outlier_reject; /* -ecg_out_reject / -wo_ecg_out_reject, default on */
  int32_t reset_dens;         /* 1: kw = 1 for kept samples (recon's reset_dens) */
  int32_t reserved[4];
} mrn_gate_params;
void mrn_gate_params_default(mrn_gate_params *p);
int mrn_gate_parse_args(mrn_gate_params *p, int argc, const char **argv);
/* Host, in place: per-readout gate times -> frame coordinates (min/max range;
 * [0, 2 x median) for retrospective ECG, outliers >= 1.3 x that dropped first;
 * respiratory quantiles (Hyndman-Fan 5) widened by half a frame), then floor
 * (no sharing) or rank bins (hist); info = {offset, scale, temporal resolution}. */
int mrn_gate_scale_times(const mrn_gate_params *p, int64_t R, double *times, double info[3]);
/* Device: R readouts of xres samples; gt [R] scaled gate times (double),
 * resp [R] respiratory weights or NULL; kx..kw [R*xres]; kdata [C][R*xres]
 * complex64.  counts (host, [frames]) receives the samples per frame.  With
 * kx_out NULL only the counts are computed; otherwise the outputs (sum of
 * counts entries; kdata_out C times that) receive every frame's samples in
 * order, frame after frame, k-space as [frame][C][count], kw scaled by
 * 1e6/count (or 1 with reset_dens).  A sample can be in several frames under
 * tornado sharing.  Returns when done; errors: mrn_last_error(NULL). */
int mrn_pregate(const mrn_gate_params *p, int64_t xres, int64_t R, const double *gt, const float *resp,
                const float *kx, const float *ky, const float *kz, const float *kw, int C, const void *kdata,
                int64_t *counts, float *kx_out, float *ky_out, float *kz_out, float *kw_out, void *kdata_out,
                void *stream);

const char *mrn_status_string(int status);
/* Message of the op's last failure ("" if none); for op = NULL, the message of
 * this thread's last failed call that takes no op (mrn_op_create, mrn_dcf_*,
 * mrn_llr, mrn_gate_*, mrn_pregate). */
const char *mrn_last_error(const mrn_op *op);

#ifdef __cplusplus
}
#endif
#endif /* MRN_H */
