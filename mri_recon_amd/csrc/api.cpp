// C-ABI entry points: trajectory/map/data setup, the operator A, A^H, E'E in
// the batched and plain forms, and the CG / power-method / FISTA drivers.
//
// Batched form, one frame (all coil blocks per launch):
//   A   : pre+z FFT (fused set_image) -> y FFT -> x FFT -> gather
//   A^H : tiled scatter -> x IFFT -> y IFFT -> z IFFT+accumulate (per block)
// with the FFT pruned to the lines and points that are non-zero or read.
// Plain form, one coil at a time (the reference's structure on the GPU,
// gridFFT.cpp:791-906): zero grid, set_image, per-axis FFT, per-sample gather;
// zero grid, per-sample atomic scatter, per-axis IFFT, accumulate.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>

#include "mrn_internal.hpp"

namespace mrn {
int recompute_taper(mrn_op *, FrameData &, hipStream_t);
int launch_gather_plain(mrn_op *, const float2 *, float2 *, int64_t, const float *, const float *,
                        const float *, const float *, const float2 *, hipStream_t);
int launch_scatter_plain(mrn_op *, float2 *, const float2 *, int64_t, const float *, const float *,
                         const float *, const float *, hipStream_t);
int build_bins(mrn_op *, FrameData &, int64_t, const float *, const float *, const float *,
               const float *, hipStream_t);
int op_zero(mrn_op *, float2 *, int64_t, hipStream_t);
int op_fill(mrn_op *, float2 *, int64_t, float2, hipStream_t);
int op_copy(mrn_op *, float2 *, const float2 *, int64_t, hipStream_t);
int op_smaps_in(mrn_op *, float2 *, const float2 *, hipStream_t);
int op_permute_in(mrn_op *, float2 *, const float2 *, const FrameData &, hipStream_t);
int op_permute_in_weighted(mrn_op *, float2 *, const float2 *, const FrameData &, hipStream_t);
int op_permute_out(mrn_op *, float2 *, const float2 *, const FrameData &, const float2 *, hipStream_t);
int op_set_image_plain(mrn_op *, float2 *, const float2 *, const float2 *, const PrePostArgs &, hipStream_t);
int op_accumulate_plain(mrn_op *, const float2 *, float2 *, const float2 *, const PrePostArgs &, hipStream_t,
                        int sos = 0);
int op_cg_alpha(mrn_op *, const float2 *, const float2 *, const float2 *, int64_t, double *, hipStream_t);
int op_cg_update(mrn_op *, float2 *, float2 *, float2 *, const float2 *, int64_t, double *, hipStream_t);
int op_power_step(mrn_op *, float2 *, const float2 *, int64_t, double *, hipStream_t);
int op_fista_momentum(mrn_op *, float2 *, float2 *, int64_t, hipStream_t);
int op_fista_finish(mrn_op *, float2 *, float2 *, int64_t, float, double *, hipStream_t, bool);
int op_cauchy_step(mrn_op *, const float2 *, const float2 *, int64_t, hipStream_t);
int op_set_step(mrn_op *, float, float, hipStream_t);
int traj_grad_impl(mrn_op *, int, const float2 *, const float2 *, float *, hipStream_t, bool);
int autotune_impl(mrn_op *, int, const char *, hipStream_t);

static inline bool batched(const mrn_op *op) { return op->p.form == MRN_FORM_BATCHED; }
static inline int64_t ncell(const mrn_op *op) {
  return (int64_t)op->g.S[0] * op->g.S[1] * op->g.S[2];
}
static inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }
static inline float fft_scale(int S) { return (float)(1. / sqrt((float)S)); }

static PrePostArgs prepost(const mrn_op *op, int ch, int overwrite, int CC) {
  PrePostArgs a;
  for (int q = 0; q < 3; q++) {
    a.N[q] = op->g.N[q];
    a.S[q] = op->g.S[q];
    a.og[q] = op->g.og[q];
  }
  a.CC = CC;
  a.lgCC = __builtin_ctz(CC);
  a.chunk_grid = ncell(op) * CC;
  a.chunk_smap = op->nvox * CC;
  a.ch = ch;
  a.overwrite = overwrite;
  a.scale = fft_scale(op->g.S[2]);
  a.xrun = op->xcd_run_fft;
  a.nf = 1;
  a.nchf = op->nch;
  a.xfs = 0;
  return a;
}

// ---------------------------------------------------- batched transforms
// Lines of the y and x passes of every coil block.  Forward passes load the
// crop range (the rest is zero) and store everything; inverse passes load
// everything and store the crop range (the rest is never read).
static LinesArgs ypass_args(const mrn_op *op, bool inverse, const FrameData *fr) {
  const Geom &g = op->g;
  const int CC = g.CC;
  LinesArgs y{};
  y.off0 = (int64_t)g.og[0] * CC;
  y.ystride = (int64_t)g.S[1] * g.S[0] * CC;
  y.zstride = ncell(op) * CC;
  y.pstride = (int64_t)g.S[0] * CC;
  y.rstride = 0;
  y.run = g.N[0] * CC;
  y.nlines = g.N[0] * CC;
  y.ld0 = inverse ? 0 : g.og[1];
  y.ld1 = inverse ? g.S[1] : g.og[1] + g.N[1];
  y.st0 = inverse ? g.og[1] : 0;
  y.st1 = inverse ? g.og[1] + g.N[1] : g.S[1];
  y.scale = fft_scale(g.S[1]);
  y.xrun = op->xcd_run_fft;  // neighbouring 128-byte runs of a row per XCD
  // the frame's rows of support in plane z: the forward pass stores, and the
  // inverse pass loads, only those ky
  y.rng = fr ? fr->yrng : nullptr;
  y.rng_zs = 1;
  y.rng_rs = 0;
  y.rng_store = inverse ? 0 : 1;
  y.rng_load = inverse ? 1 : 0;
  return y;
}

static LinesArgs xpass_args(const mrn_op *op, bool inverse, const FrameData *fr) {
  const Geom &g = op->g;
  const int CC = g.CC;
  LinesArgs x{};
  x.off0 = 0;
  x.ystride = (int64_t)g.S[1] * g.S[0] * CC;
  x.zstride = ncell(op) * CC;
  x.pstride = CC;
  x.rstride = (int64_t)g.S[0] * CC;
  x.run = CC;
  x.nlines = g.S[1] * CC;
  x.ld0 = inverse ? 0 : g.og[0];
  x.ld1 = inverse ? g.S[0] : g.og[0] + g.N[0];
  x.st0 = inverse ? g.og[0] : 0;
  x.st1 = inverse ? g.og[0] + g.N[0] : g.S[0];
  x.scale = fft_scale(g.S[0]);
  // the frame's k-space support: the forward pass stores, and the inverse
  // pass loads, only the x cells that some sample's taps reach
  x.rng = (fr && fft_x_pruned(op)) ? fr->xrng : nullptr;
  x.rng_zs = g.S[1];
  x.rng_rs = 1;
  x.rng_store = inverse ? 0 : 1;
  x.rng_load = inverse ? 1 : 0;
  return x;
}

// forward stage q of the batched transform: 0 set_image+z, 1 y, 2 x
static int fwd_stage(mrn_op *op, int q, const float2 *X, const FrameData *fr, hipStream_t st) {
  const Geom &g = op->g;
  if (q == 0) return fft_pre_z(op, op->grid, X, op->smaps, prepost(op, 0, 0, g.CC), st);
  const LinesArgs a = (q == 1) ? ypass_args(op, false, fr) : xpass_args(op, false, fr);
  const int S = g.S[q == 1 ? 1 : 0];
  return fft_lines(op, S, -1, op->grid, a, dim3(cdiv(a.nlines, FFT_LINES), g.S[2], op->nch),
                   op->tw_d[q == 1 ? 1 : 0], st);
}

// inverse stage q: 0 x, 1 y, 2 z+accumulate over the coil blocks in order
static int adj_stage(mrn_op *op, int q, float2 *X, bool overwrite, const FrameData *fr, hipStream_t st) {
  const Geom &g = op->g;
  if (q == 2) {
    for (int ch = 0; ch < op->nch; ch++)
      MRN_TRY(fft_post_z(op, op->grid, X, op->smaps, prepost(op, ch, overwrite && ch == 0, g.CC), st));
    return MRN_OK;
  }
  const LinesArgs a = (q == 1) ? ypass_args(op, true, fr) : xpass_args(op, true, fr);
  const int S = g.S[q == 1 ? 1 : 0];
  return fft_lines(op, S, 1, op->grid, a, dim3(cdiv(a.nlines, FFT_LINES), g.S[2], op->nch),
                   op->tw_d[q == 1 ? 1 : 0], st);
}

// image -> full k-space grid of every coil block
// fr: the frame whose samples read the grid (its k-space support), or null
static int fwd_grid(mrn_op *op, const float2 *X, const FrameData *fr, hipStream_t st) {
  for (int q = 0; q < 3; q++) MRN_TRY(fwd_stage(op, q, X, fr, st));
  return MRN_OK;
}

// ---------------------------------------------------- every frame at once
// A multi-frame op (batched form) keeps the frames' grids side by side in
// op->grid (op->fuse), so its solver loops run each stage once over every
// frame: the FFT passes take frame * nch + coil block as their z index, the
// gather and scatter read each frame's arrays from a table on the device
// (op->fdev).  fused_prepare builds the table, the frames' k-space work
// arrays and their supports at a fixed stride; it runs before any capture
// and again when the trajectories, data or settings change (op->version).
static bool frames_fused(const mrn_op *op) {
  return op->fuse && op->fdev && op->fdev_version == op->version && op->gather_tiles == 0;
}

static int fused_prepare(mrn_op *op, hipStream_t st) {
  if (!op->fuse || op->gather_tiles != 0) return MRN_OK;
  if (op->fdev && op->fdev_version == op->version) return MRN_OK;
  const Geom &g = op->g;
  const int F = op->F;
  for (int f = 0; f < F; f++)
    if (!op->frames[f].has_traj) return MRN_OK;  // not ready: frame by frame
  int64_t tot = 0;
  for (int f = 0; f < F; f++) tot += op->frames[f].Mv * g.CC * op->nch;
  if (tot > op->kwork_all_cap || !op->kwork_all) {
    hipFree(op->kwork_all);
    op->kwork_all = nullptr;
    op->kwork_all_cap = 0;
    MRN_HIP(op, hipMalloc(&op->kwork_all, sizeof(float2) * (tot > 0 ? tot : 1)));
    op->kwork_all_cap = tot;
  }
  if (!op->fdev) MRN_HIP(op, hipMalloc(&op->fdev, sizeof(FrameDev) * F));
  const int64_t nrow = (int64_t)g.S[1] * g.S[2];
  if (fft_x_pruned(op)) {
    if (!op->xrng_all) MRN_HIP(op, hipMalloc(&op->xrng_all, sizeof(int2) * nrow * F));
    for (int f = 0; f < F; f++)
      MRN_HIP(op, hipMemcpyAsync(op->xrng_all + f * nrow, op->frames[f].xrng, sizeof(int2) * nrow,
                                 hipMemcpyDeviceToDevice, st));
  }
  if (fft_xy_pruned(op)) {
    if (!op->yrng_all) MRN_HIP(op, hipMalloc(&op->yrng_all, sizeof(int2) * g.S[2] * F));
    for (int f = 0; f < F; f++)
      MRN_HIP(op, hipMemcpyAsync(op->yrng_all + (int64_t)f * g.S[2], op->frames[f].yrng, sizeof(int2) * g.S[2],
                                 hipMemcpyDeviceToDevice, st));
  }
  std::vector<FrameDev> h(F);
  int64_t off = 0;
  op->fmax_nsub = op->fmax_nheavy = op->fmax_nwide = op->fmax_xsplit = 0;
  op->fmax_Mv = 0;
  for (int f = 0; f < F; f++) {
    const FrameData &fr = op->frames[f];
    FrameDev &d = h[f];
    d.coords = fr.coords;
    d.taper = fr.taper;
    d.subs = fr.subs;
    d.vofs = fr.vofs;
    d.vlist = fr.vlist;
    d.heavy = fr.heavy;
    d.wide = fr.wide;
    d.xsplit = fr.xsplit;
    d.kwork = op->kwork_all + off;
    d.kdata = fr.kdata;
    d.Mv = fr.Mv;
    d.grid_off = (int64_t)f * op->nch * ncell(op) * g.CC;
    d.nsub = fr.nsub;
    d.nheavy = fr.nheavy;
    d.nwide = fr.nwide;
    d.pad = 0;
    off += fr.Mv * g.CC * op->nch;
    op->fmax_nsub = std::max(op->fmax_nsub, fr.nsub);
    op->fmax_nheavy = std::max(op->fmax_nheavy, fr.nheavy);
    op->fmax_nwide = std::max(op->fmax_nwide, fr.nwide);
    op->fmax_xsplit = std::max(op->fmax_xsplit, fr.xsplit_max);
    op->fmax_Mv = std::max(op->fmax_Mv, fr.Mv);
  }
  MRN_HIP(op, hipMemcpyAsync(op->fdev, h.data(), sizeof(FrameDev) * F, hipMemcpyHostToDevice, st));
  MRN_HIP(op, hipStreamSynchronize(st));  // h is freed on return
  op->fdev_version = op->version;
  return MRN_OK;
}

// image frames X (F x nvox) -> every frame's full k-space grid
static int fwd_grid_frames(mrn_op *op, const float2 *X, hipStream_t st) {
  const Geom &g = op->g;
  PrePostArgs a = prepost(op, 0, 0, g.CC);
  a.nf = op->F;
  a.xfs = op->nvox;
  MRN_TRY(fft_pre_z(op, op->grid, X, op->smaps, a, st));
  LinesArgs y = ypass_args(op, false, nullptr);
  y.rng = fft_xy_pruned(op) ? op->yrng_all : nullptr;
  y.rng_nz = op->nch;
  y.rng_fs = g.S[2];
  MRN_TRY(fft_lines(op, g.S[1], -1, op->grid, y, dim3(cdiv(y.nlines, FFT_LINES), g.S[2], op->nch * op->F),
                    op->tw_d[1], st));
  LinesArgs x = xpass_args(op, false, nullptr);
  x.rng = fft_x_pruned(op) ? op->xrng_all : nullptr;
  x.rng_nz = op->nch;
  x.rng_fs = (int64_t)g.S[1] * g.S[2];
  return fft_lines(op, g.S[0], -1, op->grid, x, dim3(cdiv(x.nlines, FFT_LINES), g.S[2], op->nch * op->F),
                   op->tw_d[0], st);
}

// every frame's grids -> image frames X (= or += sum over coils)
static int adj_grid_frames(mrn_op *op, float2 *X, bool overwrite, hipStream_t st) {
  const Geom &g = op->g;
  LinesArgs x = xpass_args(op, true, nullptr);
  x.rng = fft_x_pruned(op) ? op->xrng_all : nullptr;
  x.rng_nz = op->nch;
  x.rng_fs = (int64_t)g.S[1] * g.S[2];
  MRN_TRY(fft_lines(op, g.S[0], 1, op->grid, x, dim3(cdiv(x.nlines, FFT_LINES), g.S[2], op->nch * op->F),
                    op->tw_d[0], st));
  LinesArgs y = ypass_args(op, true, nullptr);
  y.rng = fft_xy_pruned(op) ? op->yrng_all : nullptr;
  y.rng_nz = op->nch;
  y.rng_fs = g.S[2];
  MRN_TRY(fft_lines(op, g.S[1], 1, op->grid, y, dim3(cdiv(y.nlines, FFT_LINES), g.S[2], op->nch * op->F),
                    op->tw_d[1], st));
  for (int ch = 0; ch < op->nch; ch++) {
    PrePostArgs a = prepost(op, ch, overwrite && ch == 0, g.CC);
    a.nf = op->F;
    a.xfs = op->nvox;
    MRN_TRY(fft_post_z(op, op->grid, X, op->smaps, a, st));
  }
  return MRN_OK;
}

int fwd_grid_pub(mrn_op *op, const float2 *X, const FrameData *fr, hipStream_t st) {
  return fwd_grid(op, X, fr, st);
}

// full grids of every coil block -> image (X = or += sum over coils)
// fr: the frame whose samples were scattered into the grid, or null
static int adj_grid(mrn_op *op, float2 *X, bool overwrite, const FrameData *fr, hipStream_t st) {
  for (int q = 0; q < 3; q++) MRN_TRY(adj_stage(op, q, X, overwrite, fr, st));
  return MRN_OK;
}

int adj_grid_pub(mrn_op *op, float2 *X, bool overwrite, const FrameData *fr, hipStream_t st) {
  return adj_grid(op, X, overwrite, fr, st);
}

// ---------------------------------------------------- plain transforms
static int plain_passes(mrn_op *op, int dir, hipStream_t st) {
  const Geom &g = op->g;
  float2 *G = op->cgrid;
  const int64_t plane = (int64_t)g.S[1] * g.S[0];
  auto zpass = [&](bool pruned) -> int {
    if (!g.fft_in[2]) return MRN_OK;
    LinesArgs a{};
    const int nx = pruned ? g.N[0] : g.S[0], ny = pruned ? g.N[1] : g.S[1];
    a.off0 = pruned ? (int64_t)g.og[1] * g.S[0] + g.og[0] : 0;
    a.ystride = g.S[0];
    a.zstride = 0;
    a.pstride = plane;
    a.run = nx;
    a.nlines = nx;
    a.ld0 = 0;
    a.ld1 = g.S[2];
    a.st0 = 0;
    a.st1 = g.S[2];
    a.scale = fft_scale(g.S[2]);
    return fft_lines(op, g.S[2], dir, G, a, dim3(cdiv(nx, FFT_LINES), ny, 1), op->tw_d[2], st);
  };
  auto ypass = [&](bool pruned) -> int {
    if (!g.fft_in[1]) return MRN_OK;
    LinesArgs a{};
    const int nx = pruned ? g.N[0] : g.S[0];
    a.off0 = pruned ? g.og[0] : 0;
    a.ystride = plane;
    a.pstride = g.S[0];
    a.run = nx;
    a.nlines = nx;
    a.ld0 = 0;
    a.ld1 = g.S[1];
    a.st0 = 0;
    a.st1 = g.S[1];
    a.scale = fft_scale(g.S[1]);
    return fft_lines(op, g.S[1], dir, G, a, dim3(cdiv(nx, FFT_LINES), g.S[2], 1), op->tw_d[1], st);
  };
  auto xpass = [&]() -> int {
    if (!g.fft_in[0]) return MRN_OK;
    LinesArgs a{};
    a.ystride = plane;
    a.pstride = 1;
    a.rstride = g.S[0];
    a.run = 1;
    a.nlines = g.S[1];
    a.ld0 = 0;
    a.ld1 = g.S[0];
    a.st0 = 0;
    a.st1 = g.S[0];
    a.scale = fft_scale(g.S[0]);
    return fft_lines(op, g.S[0], dir, G, a, dim3(cdiv(g.S[1], FFT_LINES), g.S[2], 1), op->tw_d[0], st);
  };
  if (dir < 0) {  // do_fft: z over crop lines, y over crop x, x over all (gridFFT.cpp:645-671)
    MRN_TRY(zpass(true));
    MRN_TRY(ypass(true));
    MRN_TRY(xpass());
  } else {  // do_ifft: x, y, z over all lines (gridFFT.cpp:673-711)
    MRN_TRY(xpass());
    MRN_TRY(ypass(false));
    MRN_TRY(zpass(false));
  }
  return MRN_OK;
}

static int plain_A_coil(mrn_op *op, const float2 *X, const float2 *smap, float2 *out, int64_t M,
                        const float *kx, const float *ky, const float *kz, const float *kw,
                        const float2 *sub, hipStream_t st) {
  MRN_TRY(op_zero(op, op->cgrid, ncell(op), st));
  MRN_TRY(op_set_image_plain(op, op->cgrid, X, smap, prepost(op, 0, 0, 1), st));
  MRN_TRY(plain_passes(op, -1, st));
  return launch_gather_plain(op, op->cgrid, out, M, kx, ky, kz, kw, sub, st);
}

static int plain_AH_coil(mrn_op *op, float2 *X, const float2 *smap, const float2 *data, int64_t M,
                         const float *kx, const float *ky, const float *kz, const float *kw,
                         hipStream_t st, int sos = 0) {
  MRN_TRY(op_zero(op, op->cgrid, ncell(op), st));
  MRN_TRY(launch_scatter_plain(op, op->cgrid, data, M, kx, ky, kz, kw, st));
  MRN_TRY(plain_passes(op, 1, st));
  return op_accumulate_plain(op, op->cgrid, X, smap, prepost(op, 0, 0, 1), st, sos);
}

static const float *traw(const FrameData &fr, int q) { return fr.traw + (int64_t)q * fr.M; }

// ---------------------------------------------------- frame-level engine
static int check_frame(mrn_op *op, int f, bool need_kdata) {
  if (f < 0 || f >= op->F) return fail(op, MRN_EINVAL, "frame index out of range");
  if (!op->frames[f].has_traj) return fail(op, MRN_ESTATE, "trajectory of this frame not set");
  if (need_kdata && !op->frames[f].has_kdata) return fail(op, MRN_ESTATE, "k-space data of this frame not set");
  return MRN_OK;
}

// out = E'E X of frame f (out overwritten)
static int normal_frame(mrn_op *op, int f, const float2 *X, float2 *out, hipStream_t st) {
  FrameData &fr = op->frames[f];
  if (batched(op)) {
    MRN_TRY(fwd_grid(op, X, &fr, st));
    MRN_TRY(launch_gather(op, op->grid, op->kwork, fr, nullptr, true, st));
    MRN_TRY(launch_scatter(op, op->grid, op->kwork, fr, st));
    return adj_grid(op, out, true, &fr, st);
  }
  MRN_TRY(op_zero(op, out, op->nvox, st));
  for (int c = 0; c < op->C; c++) {
    const float2 *sm = op->smaps + (int64_t)c * op->nvox;
    MRN_TRY(plain_A_coil(op, X, sm, op->kplain, fr.M, traw(fr, 0), traw(fr, 1), traw(fr, 2), traw(fr, 3), nullptr, st));
    MRN_TRY(plain_AH_coil(op, out, sm, op->kplain, fr.M, traw(fr, 0), traw(fr, 1), traw(fr, 2), traw(fr, 3), st));
  }
  return MRN_OK;
}

// R = E^H W (E X - d) of frame f (R overwritten)
static int grad_frame(mrn_op *op, int f, const float2 *X, float2 *R, hipStream_t st) {
  FrameData &fr = op->frames[f];
  if (batched(op)) {
    MRN_TRY(fwd_grid(op, X, &fr, st));
    MRN_TRY(launch_gather(op, op->grid, op->kwork, fr, fr.kdata, true, st));
    MRN_TRY(launch_scatter(op, op->grid, op->kwork, fr, st));
    return adj_grid(op, R, true, &fr, st);
  }
  MRN_TRY(op_zero(op, R, op->nvox, st));
  for (int c = 0; c < op->C; c++) {
    const float2 *sm = op->smaps + (int64_t)c * op->nvox;
    MRN_TRY(plain_A_coil(op, X, sm, op->kplain, fr.M, traw(fr, 0), traw(fr, 1), traw(fr, 2), traw(fr, 3),
                         fr.kraw + (int64_t)c * fr.M, st));
    MRN_TRY(plain_AH_coil(op, R, sm, op->kplain, fr.M, traw(fr, 0), traw(fr, 1), traw(fr, 2), traw(fr, 3), st));
  }
  return MRN_OK;
}

// out = E'E X over every frame (frames of nvox, side by side): one launch per
// stage over all frames when they are fused, else frame by frame
static int normal_frames(mrn_op *op, const float2 *X, float2 *out, hipStream_t st) {
  if (!frames_fused(op)) {
    for (int f = 0; f < op->F; f++) MRN_TRY(normal_frame(op, f, X + f * op->nvox, out + f * op->nvox, st));
    return MRN_OK;
  }
  MRN_TRY(fwd_grid_frames(op, X, st));
  MRN_TRY(launch_gather_frames(op, 0, st));
  MRN_TRY(launch_scatter_frames(op, st));
  return adj_grid_frames(op, out, true, st);
}

// R = E^H W (E X - d) over every frame, as normal_frames
static int grad_frames(mrn_op *op, const float2 *X, float2 *R, hipStream_t st) {
  if (!frames_fused(op)) {
    for (int f = 0; f < op->F; f++) MRN_TRY(grad_frame(op, f, X + f * op->nvox, R + f * op->nvox, st));
    return MRN_OK;
  }
  MRN_TRY(fwd_grid_frames(op, X, st));
  MRN_TRY(launch_gather_frames(op, 1, st));
  MRN_TRY(launch_scatter_frames(op, st));
  return adj_grid_frames(op, R, true, st);
}

// R = E^H W d of frame f (R overwritten)
static int adjoint_data_frame(mrn_op *op, int f, float2 *R, hipStream_t st) {
  FrameData &fr = op->frames[f];
  if (batched(op)) {
    MRN_TRY(launch_weight(op, op->kwork, fr.kdata, fr, st));
    MRN_TRY(launch_scatter(op, op->grid, op->kwork, fr, st));
    return adj_grid(op, R, true, &fr, st);
  }
  MRN_TRY(op_zero(op, R, op->nvox, st));
  for (int c = 0; c < op->C; c++)
    MRN_TRY(plain_AH_coil(op, R, op->smaps + (int64_t)c * op->nvox, fr.kraw + (int64_t)c * fr.M, fr.M,
                          traw(fr, 0), traw(fr, 1), traw(fr, 2), traw(fr, 3), st));
  return MRN_OK;
}

static int ensure_solver(mrn_op *op) {
  const int64_t n = (int64_t)op->F * op->nvox;
  if (!op->vR) {
    MRN_HIP(op, hipMalloc(&op->vR, sizeof(float2) * n));
    MRN_HIP(op, hipMalloc(&op->vP, sizeof(float2) * n));
    MRN_HIP(op, hipMalloc(&op->vL, sizeof(float2) * n));
    MRN_HIP(op, hipMalloc(&op->vX0, sizeof(float2) * n));
  }
  if (!op->dtrace) {
    MRN_HIP(op, hipMalloc(&op->dtrace, sizeof(double) * 4 * MRN_TRACE_CAP));
    op->dtrace_cap = MRN_TRACE_CAP;
  }
  if (!op->dstep_log) MRN_HIP(op, hipMalloc(&op->dstep_log, sizeof(double) * 2 * MRN_TRACE_CAP));
  return MRN_OK;
}

// Copy trace entries [first, first + count) of the ring (width doubles each).
static int copy_trace(mrn_op *op, double *host, int64_t first, int64_t count, int width, hipStream_t st) {
  for (int64_t j = 0; j < count;) {
    const int64_t pos = (first + j) % MRN_TRACE_CAP;
    const int64_t run = std::min<int64_t>(count - j, MRN_TRACE_CAP - pos);
    MRN_HIP(op, hipMemcpyAsync(host + j * width, op->dtrace + pos * width, sizeof(double) * width * run,
                               hipMemcpyDeviceToHost, st));
    j += run;
  }
  MRN_HIP(op, hipStreamSynchronize(st));
  return MRN_OK;
}

static void drop_cg_graph(mrn_op *op) {
  if (op->cg_exec) (void)hipGraphExecDestroy(op->cg_exec);
  if (op->cg_graph) (void)hipGraphDestroy(op->cg_graph);
  op->cg_exec = nullptr;
  op->cg_graph = nullptr;
  op->cg_X = nullptr;
}

// Run `body` iters times on the caller's stream, or capture one call of it
// in a hipGraph on the op's stream and replay it.
template <class Body>
static int run_iters(mrn_op *op, int iters, bool graph, hipStream_t st, Body body) {
  if (!graph) {
    for (int it = 0; it < iters; it++) MRN_TRY(body(st));
    return MRN_OK;
  }
  Events ev;
  MRN_HIP(op, ev.create(2, hipEventDisableTiming));
  MRN_HIP(op, hipEventRecord(ev[0], st));
  MRN_HIP(op, hipStreamWaitEvent(op->stream, ev[0], 0));
  GraphGuard gg;
  MRN_HIP(op, hipStreamBeginCapture(op->stream, hipStreamCaptureModeThreadLocal));
  const int rc = body(op->stream);
  const hipError_t ec = hipStreamEndCapture(op->stream, &gg.g);
  if (rc != MRN_OK) return rc;
  MRN_HIP(op, ec);
  MRN_HIP(op, hipGraphInstantiate(&gg.x, gg.g, nullptr, nullptr, 0));
  gg.launched = op->stream;
  for (int it = 0; it < iters; it++) {
    gg.was_launched = true;
    MRN_HIP(op, hipGraphLaunch(gg.x, op->stream));
  }
  MRN_HIP(op, hipEventRecord(ev[1], op->stream));
  MRN_HIP(op, hipStreamWaitEvent(st, ev[1], 0));
  MRN_HIP(op, hipStreamSynchronize(op->stream));
  return MRN_OK;
}

}  // namespace mrn

using namespace mrn;

extern "C" {

int mrn_set_traj(mrn_op *op, int f, int64_t M, const float *kx, const float *ky, const float *kz,
                 const float *kw, void *stream) {
  if (!op) return MRN_EINVAL;
  if (f < 0 || f >= op->F) return fail(op, MRN_EINVAL, "frame index out of range");
  if (M < 0 || M > INT32_MAX) return fail(op, MRN_EINVAL, "sample count out of range");
  if (M > 0 && (!kx || !ky || !kz || !kw)) return fail(op, MRN_EINVAL, "null trajectory pointer");
  hipStream_t st = pick(op, stream);
  FrameData &fr = op->frames[f];
  // the frame's arrays are reused when large enough: first let every launch
  // that may still read the old trajectory finish
  MRN_HIP(op, hipDeviceSynchronize());
  MRN_HIP(op, grow(fr.traw, fr.cap.traw, 4 * M));
  const float *src[4] = {kx, ky, kz, kw};
  for (int q = 0; q < 4; q++)
    if (M > 0) MRN_HIP(op, hipMemcpyAsync(fr.traw + q * M, src[q], sizeof(float) * M, hipMemcpyDeviceToDevice, st));
  fr.M = M;
  // stored data of another sample count no longer matches the trajectory
  // (kraw is [C][kraw_M]): the caller sets it again
  if (fr.has_kdata && fr.kraw_M != M) fr.has_kdata = false;
  if (batched(op)) {
    MRN_TRY(build_bins(op, fr, M, fr.traw, fr.traw + M, fr.traw + 2 * M, fr.traw + 3 * M, st));
    const int64_t need = fr.Mv * op->g.CC * op->nch;
    if (need > op->kwork_cap) {
      hipFree(op->kwork);
      op->kwork = nullptr;
      MRN_HIP(op, hipMalloc(&op->kwork, sizeof(float2) * (need > 0 ? need : 1)));
      op->kwork_cap = need;
    }
    if (fr.has_kdata) {  // re-sort stored data for the new trajectory
      MRN_HIP(op, grow(fr.kdata, fr.cap.kdata, need));
      MRN_TRY(op_permute_in(op, fr.kdata, fr.kraw, fr, st));
    }
  }
  if (M > op->kplain_cap) {
    hipFree(op->kplain);
    op->kplain = nullptr;
    MRN_HIP(op, hipMalloc(&op->kplain, sizeof(float2) * M));
    op->kplain_cap = M;
  }
  fr.has_traj = true;
  op->version++;
  MRN_HIP(op, hipStreamSynchronize(st));
  return MRN_OK;
}

int mrn_set_smaps(mrn_op *op, const void *smaps, void *stream) {
  if (!op || !smaps) return MRN_EINVAL;
  hipStream_t st = pick(op, stream);
  if (batched(op)) {
    MRN_TRY(op_smaps_in(op, op->smaps, (const float2 *)smaps, st));
  } else {
    MRN_HIP(op, hipMemcpyAsync(op->smaps, smaps, sizeof(float2) * op->nvox * op->C, hipMemcpyDeviceToDevice, st));
  }
  op->version++;
  return MRN_OK;
}

int mrn_set_kdata(mrn_op *op, int f, const void *data, void *stream) {
  if (!op || !data) return MRN_EINVAL;
  MRN_TRY(check_frame(op, f, false));
  hipStream_t st = pick(op, stream);
  FrameData &fr = op->frames[f];
  MRN_HIP(op, hipDeviceSynchronize());  // arrays reused in place: no launch may still read them
  MRN_HIP(op, grow(fr.kraw, fr.cap.kraw, (int64_t)op->C * fr.M));
  MRN_HIP(op, hipMemcpyAsync(fr.kraw, data, sizeof(float2) * op->C * fr.M, hipMemcpyDeviceToDevice, st));
  if (batched(op)) {
    const int64_t need = fr.Mv * op->g.CC * op->nch;
    MRN_HIP(op, grow(fr.kdata, fr.cap.kdata, need));
    MRN_TRY(op_permute_in(op, fr.kdata, fr.kraw, fr, st));
  }
  fr.kraw_M = fr.M;
  fr.has_kdata = true;
  op->version++;
  op->cg_ready = false;
  MRN_HIP(op, hipStreamSynchronize(st));
  return MRN_OK;
}

int mrn_image_to_kspace(mrn_op *op, int f, const void *X, void *data, const void *sub, void *stream) {
  if (!op || !X) return MRN_EINVAL;
  MRN_TRY(check_frame(op, f, false));
  if (!data && op->frames[f].M > 0) return fail(op, MRN_EINVAL, "null k-space pointer");
  hipStream_t st = pick(op, stream);
  FrameData &fr = op->frames[f];
  if (batched(op)) {
    op->fwd_f = -1;
    MRN_TRY(fwd_grid(op, (const float2 *)X, &fr, st));
    MRN_TRY(launch_gather(op, op->grid, op->kwork, fr, nullptr, false, st));
    MRN_TRY(op_permute_out(op, (float2 *)data, op->kwork, fr, (const float2 *)sub, st));
    // the grid now holds X's forward grid (mrn_traj_grad may reuse it)
    op->fwd_f = f;
    op->fwd_X = X;
    op->fwd_stream = stream;
    op->fwd_calls = op->calls;
    return MRN_OK;
  }
  for (int c = 0; c < op->C; c++)
    MRN_TRY(plain_A_coil(op, (const float2 *)X, op->smaps + (int64_t)c * op->nvox,
                         (float2 *)data + (int64_t)c * fr.M, fr.M, traw(fr, 0), traw(fr, 1), traw(fr, 2),
                         traw(fr, 3), sub ? (const float2 *)sub + (int64_t)c * fr.M : nullptr, st));
  return MRN_OK;
}

int mrn_kspace_to_image(mrn_op *op, int f, const void *data, void *X, int accumulate, void *stream) {
  if (!op || !X) return MRN_EINVAL;
  MRN_TRY(check_frame(op, f, false));
  if (!data && op->frames[f].M > 0) return fail(op, MRN_EINVAL, "null k-space pointer");
  hipStream_t st = pick(op, stream);
  FrameData &fr = op->frames[f];
  if (batched(op)) {
    MRN_TRY(op_permute_in_weighted(op, op->kwork, (const float2 *)data, fr, st));
    MRN_TRY(launch_scatter(op, op->grid, op->kwork, fr, st));
    return adj_grid(op, (float2 *)X, !accumulate, &fr, st);
  }
  if (!accumulate) MRN_TRY(op_zero(op, (float2 *)X, op->nvox, st));
  for (int c = 0; c < op->C; c++)
    MRN_TRY(plain_AH_coil(op, (float2 *)X, op->smaps + (int64_t)c * op->nvox,
                          (const float2 *)data + (int64_t)c * fr.M, fr.M, traw(fr, 0), traw(fr, 1),
                          traw(fr, 2), traw(fr, 3), st));
  return MRN_OK;
}

int mrn_normal(mrn_op *op, int f, const void *X, void *out, void *stream) {
  if (!op || !X || !out) return MRN_EINVAL;
  MRN_TRY(check_frame(op, f, false));
  return normal_frame(op, f, (const float2 *)X, (float2 *)out, pick(op, stream));
}

int mrn_coil_backward(mrn_op *op, const void *X, const void *smap, void *data, int64_t M,
                      const float *kx, const float *ky, const float *kz, const float *kw,
                      const void *sub, void *stream) {
  if (!op || !X || !data || (M > 0 && (!kx || !ky || !kz || !kw))) return MRN_EINVAL;
  return plain_A_coil(op, (const float2 *)X, (const float2 *)smap, (float2 *)data, M, kx, ky, kz, kw,
                      (const float2 *)sub, pick(op, stream));
}

int mrn_coil_forward(mrn_op *op, void *X, const void *smap, const void *data, int64_t M,
                     const float *kx, const float *ky, const float *kz, const float *kw, void *stream) {
  if (!op || !X || !data || (M > 0 && (!kx || !ky || !kz || !kw))) return MRN_EINVAL;
  return plain_AH_coil(op, (float2 *)X, (const float2 *)smap, (const float2 *)data, M, kx, ky, kz, kw,
                       pick(op, stream));
}

int mrn_set_k_rad(mrn_op *op, float k_rad, void *stream) {
  if (!op || !(k_rad > 0.0f)) return MRN_EINVAL;
  hipStream_t st = pick(op, stream);
  op->p.k_rad = k_rad;
  op->g.k_rad = k_rad;
  if (batched(op))
    for (auto &fr : op->frames)
      if (fr.has_traj) MRN_TRY(recompute_taper(op, fr, st));
  op->version++;
  return MRN_OK;
}

int mrn_coil_forward_sos(mrn_op *op, void *X, const void *data, int64_t M, const float *kx,
                         const float *ky, const float *kz, const float *kw, void *stream) {
  if (!op || !X || !data || (M > 0 && (!kx || !ky || !kz || !kw))) return MRN_EINVAL;
  return plain_AH_coil(op, (float2 *)X, nullptr, (const float2 *)data, M, kx, ky, kz, kw,
                       pick(op, stream), 1);
}

// one CG step (recon_lib.cpp:1578-1722) on the op's solver vectors
static int cg_step(mrn_op *op, float2 *X, bool with_trace, hipStream_t s) {
  const int64_t n = (int64_t)op->F * op->nvox;
  double *tr = with_trace ? op->dtrace : nullptr;
  MRN_TRY(normal_frames(op, op->vP, op->vL, s));
  MRN_TRY(op_cg_alpha(op, op->vR, op->vP, op->vL, n, tr, s));
  if (!batched(op)) {  // plain form: the reference's host-side scalar round trip
    double host_scal[16];
    MRN_HIP(op, hipMemcpyAsync(host_scal, op->dscal, sizeof(host_scal), hipMemcpyDeviceToHost, s));
    MRN_HIP(op, hipStreamSynchronize(s));
  }
  MRN_TRY(op_cg_update(op, X, op->vR, op->vP, op->vL, n, tr, s));
  if (!batched(op)) {
    double host_scal[16];
    MRN_HIP(op, hipMemcpyAsync(host_scal, op->dscal, sizeof(host_scal), hipMemcpyDeviceToHost, s));
    MRN_HIP(op, hipStreamSynchronize(s));
  }
  return MRN_OK;
}

int mrn_cg_init(mrn_op *op, void *Xv, void *stream) {
  if (!op || !Xv) return MRN_EINVAL;
  for (int f = 0; f < op->F; f++) MRN_TRY(check_frame(op, f, true));
  MRN_TRY(ensure_solver(op));
  hipStream_t st = pick(op, stream);
  MRN_TRY(fused_prepare(op, st));
  const int64_t n = (int64_t)op->F * op->nvox;
  MRN_HIP(op, hipMemsetAsync(op->dcounter, 0, sizeof(int), st));
  for (int f = 0; f < op->F; f++) MRN_TRY(adjoint_data_frame(op, f, op->vR + f * op->nvox, st));
  MRN_TRY(op_copy(op, op->vP, op->vR, n, st));
  MRN_TRY(op_zero(op, (float2 *)Xv, n, st));
  op->cg_ready = true;
  op->cg_count = 0;
  return MRN_OK;
}

int mrn_cg_iterate(mrn_op *op, void *Xv, int iters, int use_graph, double *trace, void *stream) {
  if (!op || !Xv || iters < 0) return MRN_EINVAL;
  if (!op->cg_ready) return fail(op, MRN_ESTATE, "mrn_cg_init not called");
  hipStream_t st = pick(op, stream);
  MRN_TRY(fused_prepare(op, st));  // before any capture
  float2 *X = (float2 *)Xv;
  const int64_t first = op->cg_count;
  if (use_graph && batched(op) && iters > 0) {
    if (!op->cg_exec || op->cg_X != Xv || op->cg_version != op->version) {
      drop_cg_graph(op);
      MRN_HIP(op, hipStreamBeginCapture(op->stream, hipStreamCaptureModeThreadLocal));
      int rc = cg_step(op, X, true, op->stream);
      hipError_t ec = hipStreamEndCapture(op->stream, &op->cg_graph);
      if (rc != MRN_OK || ec != hipSuccess) {
        drop_cg_graph(op);
        return rc != MRN_OK ? rc : fail(op, MRN_EHIP, hipGetErrorString(ec));
      }
      ec = hipGraphInstantiate(&op->cg_exec, op->cg_graph, nullptr, nullptr, 0);
      if (ec != hipSuccess) {
        drop_cg_graph(op);
        return fail(op, MRN_EHIP, hipGetErrorString(ec));
      }
      op->cg_X = Xv;
      op->cg_version = op->version;
    }
    Events ev;
    MRN_HIP(op, ev.create(2, hipEventDisableTiming));
    MRN_HIP(op, hipEventRecord(ev[0], st));
    MRN_HIP(op, hipStreamWaitEvent(op->stream, ev[0], 0));
    for (int it = 0; it < iters; it++) MRN_HIP(op, hipGraphLaunch(op->cg_exec, op->stream));
    MRN_HIP(op, hipEventRecord(ev[1], op->stream));
    MRN_HIP(op, hipStreamWaitEvent(st, ev[1], 0));
  } else {
    for (int it = 0; it < iters; it++) MRN_TRY(cg_step(op, X, true, st));
  }
  op->cg_count += iters;
  if (trace && iters > 0) MRN_TRY(copy_trace(op, trace, first, iters, 4, st));
  return MRN_OK;
}

int mrn_cg(mrn_op *op, void *Xv, int iters, int use_graph, double *trace, void *stream) {
  MRN_TRY(mrn_cg_init(op, Xv, stream));
  return mrn_cg_iterate(op, Xv, iters, use_graph, trace, stream);
}

int mrn_power_method(mrn_op *op, int iters, double *step, void *stream) {
  if (!op || iters <= 0) return MRN_EINVAL;
  for (int f = 0; f < op->F; f++) MRN_TRY(check_frame(op, f, false));
  MRN_TRY(ensure_solver(op));
  hipStream_t st = pick(op, stream);
  const int64_t n = (int64_t)op->F * op->nvox;
  MRN_HIP(op, hipMemsetAsync(op->dcounter, 0, sizeof(int), st));
  MRN_TRY(fused_prepare(op, st));
  MRN_TRY(op_fill(op, op->vX0, n, make_float2(-0.5f, -0.5f), st));
  for (int it = 0; it < iters; it++) {
    MRN_TRY(normal_frames(op, op->vX0, op->vR, st));
    MRN_TRY(op_power_step(op, op->vX0, op->vR, n, op->dtrace, st));
  }
  double sc[16];
  MRN_HIP(op, hipMemcpyAsync(sc, op->dscal, sizeof(double) * 16, hipMemcpyDeviceToHost, st));
  MRN_HIP(op, hipStreamSynchronize(st));
  if (step) *step = (double)(float)(1. / sc[4]);
  return MRN_OK;
}

int mrn_fista(mrn_op *op, void *Xv, int iters, double step, int use_graph, double *resid, void *stream) {
  if (!op || !Xv || iters < 0) return MRN_EINVAL;
  for (int f = 0; f < op->F; f++) MRN_TRY(check_frame(op, f, true));
  MRN_TRY(ensure_solver(op));
  hipStream_t st = pick(op, stream);
  float2 *X = (float2 *)Xv;
  const int64_t n = (int64_t)op->F * op->nvox;
  MRN_HIP(op, hipMemsetAsync(op->dcounter, 0, sizeof(int), st));
  MRN_TRY(op_zero(op, X, n, st));
  MRN_TRY(op_zero(op, op->vP, n, st));  // X_old
  const float stf = (float)step;
  const bool l1 = op->l1_fraction > 0.0f;
  const bool cauchy = op->step_type == MRN_STEP_CAUCHY;
  if (l1) MRN_TRY(ensure_l1_state(op));  // before any capture
  MRN_TRY(llr_fista_prepare(op));
  MRN_TRY(fused_prepare(op, st));
  if (cauchy) MRN_TRY(op_set_step(op, stf, 0.0f, st));  // the step until the first Cauchy update
  auto body = [&](hipStream_t s) -> int {
    MRN_TRY(op_fista_momentum(op, X, op->vP, n, s));
    MRN_TRY(grad_frames(op, X, op->vR, s));
    if (cauchy) {  // P = E'E R per frame, step = sum|R|^2 / sum conj(R) P (recon_lib.cpp:2013-2041)
      MRN_TRY(normal_frames(op, op->vR, op->vL, s));
      MRN_TRY(op_cauchy_step(op, op->vR, op->vL, n, s));
    }
    MRN_TRY(op_fista_finish(op, X, op->vR, n, stf, op->dtrace, s, cauchy));
    // locally low-rank threshold over the frames (recon_lib.cpp:2117-2121)
    if (op->llr_on) MRN_TRY(llr_fista_step(op, X, s));
    // L1 wavelet soft threshold after the gradient step (recon_lib.cpp:2145-2150)
    if (l1) MRN_TRY(l1_fista_step(op, X, s));
    return MRN_OK;
  };
  MRN_TRY(run_iters(op, iters, use_graph && batched(op), st, body));
  if (resid && iters > 0) MRN_TRY(copy_trace(op, resid, 0, iters, 1, st));
  return MRN_OK;
}

int mrn_wavelet(mrn_op *op, void *X, int frames, const int32_t levels[3], const int32_t shift[3], int inverse,
                void *stream) {
  if (!op || !X || !levels || !shift || frames < 1) return MRN_EINVAL;
  for (int a = 0; a < 3; a++)
    if (levels[a] < 0 || shift[a] < 0 || shift[a] > 3) return fail(op, MRN_EINVAL, "wavelet levels or shift");
  return wavelet_host(op, (float2 *)X, frames, levels, shift, inverse != 0, pick(op, stream));
}

int mrn_l1_wavelet(mrn_op *op, void *X, int frames, double fraction, const int32_t levels[3], int spins,
                   const int32_t *shifts, double *threshold, void *stream) {
  if (!op || !X || !levels || frames < 1) return MRN_EINVAL;
  if (!(fraction >= 0.0 && fraction <= 1.0) || spins < 1 || spins > MRN_MAX_SPINS)
    return fail(op, MRN_EINVAL, "fraction in [0,1] and 1 <= spins <= 64");
  for (int a = 0; a < 3; a++)
    if (levels[a] < 0) return fail(op, MRN_EINVAL, "wavelet levels");
  if (shifts)
    for (int q = 0; q < 3 * spins; q++)
      if (shifts[q] < 0 || shifts[q] > 3) return fail(op, MRN_EINVAL, "wavelet shifts are 0..3");
  return l1_wavelet_host(op, (float2 *)X, frames, (float)fraction, levels, spins, shifts, threshold,
                         pick(op, stream));
}

int mrn_set_fista_l1(mrn_op *op, double fraction, const int32_t levels[3], int spins, uint64_t seed) {
  if (!op || !levels) return MRN_EINVAL;
  if (!(fraction >= 0.0 && fraction <= 1.0) || spins < 1 || spins > MRN_MAX_SPINS)
    return fail(op, MRN_EINVAL, "fraction in [0,1] and 1 <= spins <= 64");
  for (int a = 0; a < 3; a++) {
    if (levels[a] < 0) return fail(op, MRN_EINVAL, "wavelet levels");
    op->l1_levels[a] = levels[a];
  }
  op->l1_fraction = (float)fraction;
  op->l1_spins = spins;
  op->l1_seed = seed;
  op->version++;
  return MRN_OK;
}

int mrn_set_fista_step(mrn_op *op, int step_type, int cauchy_update_number) {
  if (!op) return MRN_EINVAL;
  if (step_type != MRN_STEP_CAUCHY && step_type != MRN_STEP_MAXEIG)
    return fail(op, MRN_EINVAL, "step type is MRN_STEP_CAUCHY or MRN_STEP_MAXEIG");
  if (cauchy_update_number < 0) return fail(op, MRN_EINVAL, "cauchy_update_number < 0");
  op->step_type = step_type;
  op->cauchy_updates = cauchy_update_number;
  op->version++;
  return MRN_OK;
}

int mrn_get_fista_steps(mrn_op *op, double *out, int n, void *stream) {
  if (!op || !out || n < 0 || n > MRN_TRACE_CAP) return MRN_EINVAL;
  if (!op->dstep_log) return fail(op, MRN_ESTATE, "no FISTA run yet");
  hipStream_t st = pick(op, stream);
  MRN_HIP(op, hipMemcpyAsync(out, op->dstep_log, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, st));
  MRN_HIP(op, hipStreamSynchronize(st));
  return MRN_OK;
}

int mrn_get_fista_thresholds(mrn_op *op, double *out, int n, void *stream) {
  if (!op || !out || n < 0 || n > MRN_TRACE_CAP) return MRN_EINVAL;
  if (!op->dthr_log) return fail(op, MRN_ESTATE, "no L1 FISTA run yet");
  hipStream_t st = pick(op, stream);
  MRN_HIP(op, hipMemcpyAsync(out, op->dthr_log, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  MRN_HIP(op, hipStreamSynchronize(st));
  return MRN_OK;
}

int mrn_traj_grad(mrn_op *op, int f, const void *X, const void *g, float *grad_k, void *stream) {
  if (!op || !X || !g || !grad_k) return MRN_EINVAL;
  MRN_TRY(check_frame(op, f, false));
  if (!batched(op)) return fail(op, MRN_EUNSUPPORTED, "trajectory gradient needs the batched form");
  // straight after mrn_image_to_kspace of the same frame, image and stream the
  // grid already holds X's forward grid
  const bool grid_ready = !op->grid_exposed && op->fwd_f == f && op->fwd_X == X && op->fwd_stream == stream &&
                          op->fwd_calls == op->calls;
  return traj_grad_impl(op, f, (const float2 *)X, (const float2 *)g, grad_k, pick(op, stream), grid_ready);
}

int mrn_autotune(mrn_op *op, int f, const char *cache_path, void *stream) {
  if (!op) return MRN_EINVAL;
  MRN_TRY(check_frame(op, f, false));
  if (!batched(op)) return MRN_OK;
  MRN_TRY(ensure_solver(op));
  op->version++;
  return autotune_impl(op, f, cache_path, pick(op, stream));
}

int mrn_get_tuning(const mrn_op *op, mrn_tuning *out) {
  if (!op || !out) return MRN_EINVAL;
  memset(out, 0, sizeof(*out));
  out->bin_edge = op->t.B;
  out->max_visits = op->t.maxv;
  out->xcd_run = op->t.xrun;
  out->xcd_run_gather = op->t.xrun_gather;
  out->gather_coils = op->gather_coils > 0 ? std::min(op->gather_coils, op->g.CC) : std::min(4, op->g.CC);
  out->xcd_run_fft = op->xcd_run_fft;
  out->fft_stockham = op->fft_stockham ? 1 : 0;
  out->scatter_valu = (op->scatter_valu || (op->g.CC != 16 && op->g.CC != 8)) ? 1 : 0;
  for (int a = 0; a < 3; a++) out->tile[a] = op->t.T;
  out->from_cache = op->tune_from_cache;
  out->gather_tiles = op->gather_tiles;
  out->normal_ms = op->tune_ms;
  return MRN_OK;
}

int mrn_stage_times(mrn_op *op, int f, const void *Xv, int reps, float ms[MRN_NSTAGES], void *stream) {
  if (!op || !Xv || !ms || reps <= 0) return MRN_EINVAL;
  MRN_TRY(check_frame(op, f, false));
  if (!batched(op)) return fail(op, MRN_EUNSUPPORTED, "stage times cover the batched form");
  hipStream_t st = pick(op, stream);
  FrameData &fr = op->frames[f];
  const float2 *X = (const float2 *)Xv;
  MRN_TRY(ensure_solver(op));
  Events ev;
  MRN_HIP(op, ev.create(MRN_NSTAGES + 1));
  double acc[MRN_NSTAGES] = {0};
  for (int r = 0; r < reps; r++) {
    int e = 0;
    MRN_HIP(op, hipEventRecord(ev[e++], st));
    for (int q = 0; q < 3; q++) {
      MRN_TRY(fwd_stage(op, q, X, &fr, st));
      MRN_HIP(op, hipEventRecord(ev[e++], st));
    }
    MRN_TRY(launch_gather(op, op->grid, op->kwork, fr, nullptr, true, st));
    MRN_HIP(op, hipEventRecord(ev[e++], st));
    MRN_TRY(launch_scatter_zero(op, op->grid, fr, st));
    MRN_HIP(op, hipEventRecord(ev[e++], st));
    MRN_TRY(launch_scatter_tiles(op, op->grid, op->kwork, fr, st));
    MRN_HIP(op, hipEventRecord(ev[e++], st));
    for (int q = 0; q < 3; q++) {
      MRN_TRY(adj_stage(op, q, op->vX0, true, &fr, st));
      MRN_HIP(op, hipEventRecord(ev[e++], st));
    }
    MRN_HIP(op, hipEventSynchronize(ev[MRN_NSTAGES]));
    for (int q = 0; q < MRN_NSTAGES; q++) {
      float t = 0.f;
      MRN_HIP(op, hipEventElapsedTime(&t, ev[q], ev[q + 1]));
      acc[q] += t;
    }
  }
  for (int q = 0; q < MRN_NSTAGES; q++) ms[q] = (float)(acc[q] / reps);
  return MRN_OK;
}

int mrn_get_frame_stats(mrn_op *op, int f, mrn_frame_stats *out) {
  if (!op || !out) return MRN_EINVAL;
  MRN_TRY(check_frame(op, f, false));
  memset(out, 0, sizeof(*out));
  const FrameData &fr = op->frames[f];
  out->samples = fr.M;
  out->valid = fr.Mv;
  if (!batched(op)) return MRN_OK;
  out->scatter_items = fr.nsub;
  out->split_tiles = fr.nheavy;
  out->wide = fr.nwide;
  out->gather_items = fr.ngsub;
  if (fr.nsub > 0) {
    std::vector<int4> subs((size_t)fr.nsub);
    MRN_HIP(op, hipMemcpy(subs.data(), fr.subs, sizeof(int4) * fr.nsub, hipMemcpyDeviceToHost));
    for (const int4 &w : subs) {
      out->tile_visits += w.z - w.y;
      if (w.w) out->split_items++;
    }
  }
  // a split tile's items each add 64 cells x 2 CC reals; a wide sample adds at
  // most NT^3 taps x 2 CC reals; per coil block
  const int64_t per_cell = 2 * (int64_t)op->g.CC;
  const int64_t nt3 = (int64_t)op->t.NT * op->t.NT * op->t.NT;
  out->scatter_atomics = (out->split_items * 64 * per_cell + (int64_t)fr.nwide * nt3 * per_cell) * op->nch;
  return MRN_OK;
}

int mrn_get_support(mrn_op *op, int f, int32_t *xrng, int32_t *yrng) {
  if (!op || !xrng) return MRN_EINVAL;
  MRN_TRY(check_frame(op, f, false));
  const FrameData &fr = op->frames[f];
  const int64_t nrow = (int64_t)op->g.S[1] * op->g.S[2];
  if (!batched(op) || !fr.xrng) {  // every cell of every row
    for (int64_t r = 0; r < nrow; r++) {
      xrng[2 * r] = 0;
      xrng[2 * r + 1] = op->g.S[0] - 1;
    }
  } else {
    MRN_HIP(op, hipMemcpy(xrng, fr.xrng, sizeof(int2) * nrow, hipMemcpyDeviceToHost));
  }
  if (yrng) {
    if (fr.yrng && fft_xy_pruned(op)) {
      MRN_HIP(op, hipMemcpy(yrng, fr.yrng, sizeof(int2) * op->g.S[2], hipMemcpyDeviceToHost));
    } else {
      for (int z = 0; z < op->g.S[2]; z++) {
        yrng[2 * z] = 0;
        yrng[2 * z + 1] = op->g.S[1] - 1;
      }
    }
  }
  // the x passes prune only where the four-step form applies
  if (!fft_x_pruned(op))
    for (int64_t r = 0; r < nrow; r++) {
      xrng[2 * r] = 0;
      xrng[2 * r + 1] = op->g.S[0] - 1;
    }
  return MRN_OK;
}

int mrn_get_grid(mrn_op *op, void **grid, int64_t dims[5]) {
  if (!op || !grid || !dims) return MRN_EINVAL;
  const bool b = batched(op);
  *grid = b ? (void *)op->grid : (void *)op->cgrid;
  op->grid_exposed = true;  // the caller may write it: no grid is reused from now on
  dims[0] = b ? op->nch : 1;
  dims[1] = b ? op->g.CC : 1;
  for (int a = 0; a < 3; a++) dims[2 + a] = op->g.S[a];
  return MRN_OK;
}

int mrn_grid_zero(mrn_op *op, void *stream) {
  if (!op) return MRN_EINVAL;
  const int64_t n = batched(op) ? ncell(op) * op->g.CC * op->nch : ncell(op);
  return op_zero(op, batched(op) ? op->grid : op->cgrid, n, pick(op, stream));
}

int mrn_chop_grid_forward(mrn_op *op, int f, const void *data, void *stream) {
  if (!op) return MRN_EINVAL;
  MRN_TRY(check_frame(op, f, false));
  FrameData &fr = op->frames[f];
  if (!data && fr.M > 0) return fail(op, MRN_EINVAL, "null k-space pointer");
  hipStream_t st = pick(op, stream);
  if (!batched(op)) {
    if (op->C != 1) return fail(op, MRN_EUNSUPPORTED, "the plain form's grid holds one coil: chop_grid needs C = 1");
    return launch_scatter_plain(op, op->cgrid, (const float2 *)data, fr.M, traw(fr, 0), traw(fr, 1), traw(fr, 2),
                                traw(fr, 3), st);
  }
  MRN_TRY(op_permute_in_weighted(op, op->kwork, (const float2 *)data, fr, st));
  return launch_scatter(op, op->grid, op->kwork, fr, st, true);
}

int mrn_chop_grid_backward(mrn_op *op, int f, void *data, const void *sub, void *stream) {
  if (!op) return MRN_EINVAL;
  MRN_TRY(check_frame(op, f, false));
  FrameData &fr = op->frames[f];
  if (!data && fr.M > 0) return fail(op, MRN_EINVAL, "null k-space pointer");
  hipStream_t st = pick(op, stream);
  if (!batched(op)) {
    if (op->C != 1) return fail(op, MRN_EUNSUPPORTED, "the plain form's grid holds one coil: chop_grid needs C = 1");
    return launch_gather_plain(op, op->cgrid, (float2 *)data, fr.M, traw(fr, 0), traw(fr, 1), traw(fr, 2),
                               traw(fr, 3), (const float2 *)sub, st);
  }
  MRN_TRY(launch_gather(op, op->grid, op->kwork, fr, nullptr, false, st));
  return op_permute_out(op, (float2 *)data, op->kwork, fr, (const float2 *)sub, st);
}

int mrn_sync(mrn_op *op, void *stream) {
  if (!op) return MRN_EINVAL;
  MRN_HIP(op, hipStreamSynchronize(pick(op, stream)));
  return MRN_OK;
}

}  // extern "C"
