// Op construction: geometry, kernel tables, deapodisation windows, device
// buffers.  Host arithmetic keeps the reference's float/double types so the
// tables and windows come out as the reference computes them:
//   geometry            gridFFT.cpp:484-541
//   kernel tables       gridFFT.cpp:251-270, 276-473 (bessi0 1325-1352)
//   windows             gridFFT.cpp:550-631
//   flags               gridFFT.cpp:187-245
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "mrn_internal.hpp"

namespace mrn {

static const double kPi = 3.14159265359;  // the reference's PI constant
static const float kGridMod = 600.0f;     // table points per grid unit

int fail(mrn_op *op, int code, const char *msg) {
  if (op) op->err = msg ? msg : "";
  return code;
}

// Modified Bessel I0, Abramowitz & Stegun 9.8.1 / 9.8.2 (double polynomial,
// float result, as the reference evaluates it).
static float host_i0(float x) {
  float ax = fabsf(x);
  double y;
  if (ax < 3.75) {
    y = x / 3.75;
    y *= y;
    return (float)(1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
                   y * (0.360768e-1 + y * 0.45813e-2))))));
  }
  y = 3.75 / ax;
  return (float)((exp((double)ax) / sqrt((double)ax)) *
                 (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 +
                  y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 + y * (-0.1647633e-1 +
                  y * 0.392377e-2)))))))));
}

static float default_dwin(int kernel) {
  switch (kernel) {
    case MRN_KERNEL_TRIANGLE: return 1.0f;
    case MRN_KERNEL_POLY: return 2.0f;
    default: return 2.5f;
  }
}

// One axis' lookup table: length L+10, entries 0..L filled, scaled so the
// entries sum to 0.5 * 600.
static void make_table(mrn_op *op, int a) {
  const mrn_params &p = op->p;
  const float dw = op->g.dwin[a];
  const int L = (int)(dw * kGridMod);
  std::vector<float> f((size_t)L + 10, 0.0f);
  op->g.beta[a] = 0.0f;
  switch (p.kernel) {
    case MRN_KERNEL_TRIANGLE:
      for (int i = 0; i <= L; i++) f[i] = (float)(1.0 - (float)i / (float)L);
      break;
    case MRN_KERNEL_KAISER: {
      const float g = op->g.grid[a] * p.grid_scale[a];
      float beta = (float)(kPi * sqrtf((float)((dw * dw) / (g * g) * (g - 0.5) * (g - 0.5) - 0.8)));
      const float beta_min = (float)sqrt(pow(kPi * 2 * dw / g, 2.0) - 9.6752);
      if (beta_min > beta) beta = beta_min;
      op->g.beta[a] = beta;
      const float den = host_i0(beta);
      for (int i = 0; i <= L; i++) {
        const float pos = (float)i / (float)L;
        const float arg = sqrtf((float)(1.0 - pos * pos));
        f[i] = host_i0(beta * arg) / den;
      }
    } break;
    case MRN_KERNEL_SINC: {
      // Hamming term of the X table divides by the Z table length (gridFFT.cpp:414)
      const int Lz = (int)(op->g.dwin[2] * kGridMod);
      const int Lh = (a == 0) ? Lz : L;
      for (int i = 0; i <= L; i++) {
        const float pos = ((float)i * dw / p.grid_scale[a]) / (float)L;
        f[i] = (pos < 0.01f) ? 1.0f : (float)(sin(pos * kPi) / (pos * kPi));
        f[i] *= (float)(0.54 + 0.46 * cos(kPi * (float)i / (float)Lh));
      }
    } break;
    case MRN_KERNEL_POLY: {
      static const double c[6] = {1., 0.04522831, -3.36020304, 1.12417012, 2.82448025, -1.63447764};
      const int len = L + 10;
      for (int i = 0; i < len; i++) {
        const double x = (double)i / (double)len;
        double v = 0.0, xp = 1.0;
        for (int q = 0; q < 6; q++) {
          v += c[q] * xp;
          xp *= x;
        }
        f[i] = (float)v;
      }
    } break;
  }
  double s = 0.0;
  for (float v : f) s += v;
  const double sc = 0.5 * kGridMod / s;
  for (float &v : f) v = (float)(v * sc);
  op->g.knorm[a] = (float)sc;
  op->g.i0beta[a] = op->g.beta[a] > 0 ? host_i0(op->g.beta[a]) : 1.0f;
  op->g.tlen[a] = L + 10;
  op->table_h[a] = std::move(f);
}

// Deapodisation window of one axis (crop, fftshift sign and 1/S folded in).
static void make_window(mrn_op *op, int a) {
  const int S = op->g.S[a], og0 = op->g.og[a], og1 = op->g.og[a] + op->g.N[a];
  std::vector<float> w((size_t)S, 0.0f);
  const float dw = op->g.dwin[a];
  const std::vector<float> &f = op->table_h[a];
  if (op->g.fft_in[a] == 1 && op->g.grid_in[a] == 1) {
    for (int i = 0; i < S; i++) {
      const float ipos = (float)(i - (float)S / 2.0);
      float acc = 0.0f;
      for (int q = 0; q < dw * kGridMod; q++)
        acc = (float)(acc + 2 * cos(2 * kPi * ipos * q / kGridMod / (float)S) * f[q]);
      acc = kGridMod / acc;
      acc *= (float)(2 * (i % 2) - 1) / S;
      if (i < og0 || i > og1 - 1) acc *= 0.0f;
      w[i] = acc;
    }
    const float mx = *std::max_element(w.begin(), w.end());
    for (float &v : w) v /= mx;
  } else if (op->g.fft_in[a]) {
    for (int i = 0; i < S; i++) w[i] = (float)(2 * (i % 2) - 1) / S;
  } else {
    for (int i = 0; i < S; i++) w[i] = 1.0f;
  }
  op->win_h[a] = std::move(w);
}

static int pick_cc(int C, int req) {
  if (req > 0) return req;
  int cc = 16;
  while (cc > 1 && (C % cc) != 0) cc >>= 1;
  return cc;
}

}  // namespace mrn

using namespace mrn;

extern "C" {

void mrn_params_default(mrn_params *p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->overgrid = 1.5f;
  for (int a = 0; a < 3; a++) {
    p->grid[a] = -1.0f;
    p->dwin[a] = -1.0f;
    p->grid_in[a] = 1;
    p->fft_in[a] = 1;
    p->grid_scale[a] = 1.0f;
  }
  p->kernel = MRN_KERNEL_KAISER;
  p->k_rad = 9999.0f;
  p->ncoils = 1;
  p->nframes = 1;
  p->kernel_mode = MRN_KMODE_TABLE;
  p->form = MRN_FORM_BATCHED;
  p->fuse_frames = -1;
}

int mrn_parse_args(mrn_params *p, int argc, const char **argv) {
  if (!p || (argc > 0 && !argv)) return MRN_EINVAL;
  for (int i = 0; i < argc; i++) {
    const char *s = argv[i];
    auto fval = [&](float &dst) -> int {
      if (i + 1 >= argc) return MRN_EINVAL;
      dst = (float)atof(argv[++i]);
      return MRN_OK;
    };
    auto ival = [&](int32_t &dst) -> int {
      if (i + 1 >= argc) return MRN_EINVAL;
      dst = atoi(argv[++i]);
      return MRN_OK;
    };
    int rc = MRN_OK;
    if (!strcmp(s, "-overgrid")) rc = fval(p->overgrid);
    else if (!strcmp(s, "-dwinX")) rc = fval(p->dwin[0]);
    else if (!strcmp(s, "-dwinY")) rc = fval(p->dwin[1]);
    else if (!strcmp(s, "-dwinZ")) rc = fval(p->dwin[2]);
    else if (!strcmp(s, "-grid_x")) rc = fval(p->grid[0]);
    else if (!strcmp(s, "-grid_y")) rc = fval(p->grid[1]);
    else if (!strcmp(s, "-grid_z")) rc = fval(p->grid[2]);
    else if (!strcmp(s, "-grid_scale_x")) rc = fval(p->grid_scale[0]);
    else if (!strcmp(s, "-grid_scale_y")) rc = fval(p->grid_scale[1]);
    else if (!strcmp(s, "-grid_scale_z")) rc = fval(p->grid_scale[2]);
    else if (!strcmp(s, "-grid_in_x")) rc = ival(p->grid_in[0]);
    else if (!strcmp(s, "-grid_in_y")) rc = ival(p->grid_in[1]);
    else if (!strcmp(s, "-grid_in_z")) rc = ival(p->grid_in[2]);
    else if (!strcmp(s, "-fft_in_x")) rc = ival(p->fft_in[0]);
    else if (!strcmp(s, "-fft_in_y")) rc = ival(p->fft_in[1]);
    else if (!strcmp(s, "-fft_in_z")) rc = ival(p->fft_in[2]);
    else if (!strcmp(s, "-kaiser")) p->kernel = MRN_KERNEL_KAISER;
    else if (!strcmp(s, "-triangle")) p->kernel = MRN_KERNEL_TRIANGLE;
    else if (!strcmp(s, "-sinc")) p->kernel = MRN_KERNEL_SINC;
    else if (!strcmp(s, "-poly_kernel")) p->kernel = MRN_KERNEL_POLY;
    else if (!strcmp(s, "-fast_grid")) {  // special copy (gridFFT.cpp:237-243)
      p->overgrid = 1.0f;
      p->kernel = MRN_KERNEL_TRIANGLE;
      p->dwin[0] = p->dwin[1] = p->dwin[2] = 1.0f;
    }
    if (rc != MRN_OK) return rc;
  }
  return MRN_OK;
}

int mrn_op_create(const mrn_params *pin, mrn_op **out) {
  if (!pin || !out) return MRN_EINVAL;
  *out = nullptr;
  mrn::free_err_clear();
  const mrn_params &p = *pin;
  for (int a = 0; a < 3; a++)
    if (p.N[a] <= 0) return MRN_EINVAL;
  if (p.ncoils <= 0 || p.nframes <= 0) return MRN_EINVAL;
  if (p.kernel < 0 || p.kernel > 3) return MRN_EINVAL;
  if (p.form != MRN_FORM_BATCHED && p.form != MRN_FORM_PLAIN) return MRN_EINVAL;
  if (p.coil_chunk != 0 && (p.coil_chunk < 1 || p.coil_chunk > 16 || (p.coil_chunk & (p.coil_chunk - 1))))
    return MRN_EINVAL;
  mrn_op *op = new mrn_op();
  op->p = p;
  Geom &g = op->g;
  memset(&g, 0, sizeof(g));
  op->C = p.ncoils;
  op->F = p.nframes;
  for (int a = 0; a < 3; a++) {
    g.N[a] = p.N[a];
    g.grid_in[a] = p.grid_in[a] ? 1 : 0;
    g.fft_in[a] = p.fft_in[a] ? 1 : 0;
    float gr = p.grid[a];
    if (g.grid_in[a]) {
      if (gr == -1.0f) gr = (float)(16.0 * ceil((p.overgrid * (float)p.N[a]) / 16.0) / (float)p.N[a]);
    } else {
      gr = 1.0f;
    }
    if (!(gr >= 1.0f)) {
      delete op;
      return MRN_EINVAL;
    }
    g.grid[a] = gr;
    g.S[a] = (int)(gr * p.N[a]);
    g.og[a] = (int)((float)p.N[a] * (gr - 1) / 2.0);
    g.dwin[a] = (p.dwin[a] == -1.0f) ? default_dwin(p.kernel) : p.dwin[a];
    if (!(g.dwin[a] > 0.0f) || g.dwin[a] > 3.5f) {
      delete op;
      return mrn::free_fail(MRN_EUNSUPPORTED, "kernel half width must lie in (0, 3.5]");
    }
  }
  g.k_rad = p.k_rad;
  g.kmode = p.kernel_mode;
  for (int a = 0; a < 3; a++) make_table(op, a);
  for (int a = 0; a < 3; a++) make_window(op, a);
  // pack the tables, sharing identical ones
  std::vector<float> packed;
  for (int a = 0; a < 3; a++) {
    int found = -1;
    for (int b = 0; b < a; b++)
      if (op->table_h[b] == op->table_h[a]) found = g.toff[b];
    if (found >= 0) {
      g.toff[a] = found;
    } else {
      g.toff[a] = (int)packed.size();
      packed.insert(packed.end(), op->table_h[a].begin(), op->table_h[a].end());
    }
  }
  g.ntab = (int)packed.size();
  const bool batched = p.form == MRN_FORM_BATCHED;
  g.CC = batched ? pick_cc(op->C, p.coil_chunk) : 1;
  g.lgCC = 0;
  while ((1 << g.lgCC) < g.CC) g.lgCC++;
  op->nch = (op->C + g.CC - 1) / g.CC;
  op->nvox = (int64_t)g.N[0] * g.N[1] * g.N[2];
  // sample bins (B = 1: one cell per bin, so a tile's bin neighbourhood is
  // exact) and scatter tiles (T = 4 cells: one wavefront)
  int ntap = 1;
  for (int a = 0; a < 3; a++)
    if (g.grid_in[a]) ntap = std::max(ntap, (int)floorf(2 * g.dwin[a]) + 1);
  op->t.NT = ntap;
  op->t.NTb = taps_main(g);
  if (p.bin_edge != 0 && p.bin_edge != 1 && p.bin_edge != 2) {
    delete op;
    return MRN_EINVAL;
  }
  op->t.B = p.bin_edge > 0 ? p.bin_edge : 1;
  // scatter tiles are 4 x 4 x 4 cells: one wavefront of cells, and the two
  // 32-row halves of the matrix-core product
  for (int a = 0; a < 3; a++)
    if (p.tile[a] != 0 && p.tile[a] != 4) {
      delete op;
      return mrn::free_fail(MRN_EUNSUPPORTED, "scatter tile edge must be 0 (auto) or 4");
    }
  op->t.T = 4;
  for (int a = 0; a < 3; a++) {
    op->t.nb[a] = (g.S[a] + op->t.B - 1) / op->t.B;
    op->t.nt[a] = (g.S[a] + op->t.T - 1) / op->t.T;
  }
  op->t.nbins = op->t.nb[0] * op->t.nb[1] * op->t.nb[2];
  op->t.ntiles = op->t.nt[0] * op->t.nt[1] * op->t.nt[2];
  op->t.maxv = p.max_visits > 0 ? p.max_visits : 4096;
  op->t.xrun = p.xcd_run > 0 ? p.xcd_run : (p.xcd_run < 0 ? -1 : 1);
  op->t.xrun_gather = p.xcd_run_gather > 0 ? p.xcd_run_gather : 64;
  op->gather_tiles = p.gather_tiles == 2 ? 2 : p.gather_tiles != 0 ? 1 : 0;
  op->scatter_valu = p.scatter_valu != 0;
  op->fft_stockham = p.fft_stockham != 0;
  op->xcd_run_fft = p.xcd_run_fft > 0 ? p.xcd_run_fft : 8;
  if (p.gather_coils != 0 && p.gather_coils != 1 && p.gather_coils != 2 && p.gather_coils != 4 &&
      p.gather_coils != 8) {
    delete op;
    return mrn::free_fail(MRN_EINVAL, "gather_coils must be 0 (auto), 1, 2, 4 or 8");
  }
  op->gather_coils = p.gather_coils;
  if (4 * (g.ntab + 4) > 64 * 1024) {
    delete op;
    return mrn::free_fail(MRN_EUNSUPPORTED, "kernel tables larger than 64 KB of LDS");
  }
  if (batched) {
    for (int a = 0; a < 3; a++) {
      if (!g.fft_in[a] || !fft_size_supported(g.S[a])) {
        delete op;
        return mrn::free_fail(MRN_EUNSUPPORTED,
                              "the batched form needs a transform on every axis and grid sizes with "
                              "factors 2, 3 and 5; the plain form covers this geometry");
      }
    }
    if (op->t.nbins > (1 << 30)) {
      delete op;
      return mrn::free_fail(MRN_EUNSUPPORTED, "grid too large for the sample bins");
    }
  }
  op->frames.resize(op->F);
  hipError_t e = hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete op;
    return MRN_EHIP;
  }
  if (op->gather_tiles == 2) {  // default priority, like op->stream
    if (hipStreamCreateWithFlags(&op->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&op->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&op->ev_join, hipEventDisableTiming) != hipSuccess) {
      mrn_op_destroy(op);
      return MRN_EHIP;
    }
  }
  auto cleanup_fail = [&](int code) {
    mrn_op_destroy(op);
    return code;
  };
  auto amalloc = [&](void **ptr, size_t bytes) -> hipError_t { return hipMalloc(ptr, bytes > 0 ? bytes : 16); };
  if (amalloc((void **)&op->tables_d, sizeof(float) * g.ntab) != hipSuccess) return cleanup_fail(MRN_ENOMEM);
  hipMemcpy(op->tables_d, packed.data(), sizeof(float) * g.ntab, hipMemcpyHostToDevice);
  for (int a = 0; a < 3; a++) {
    if (amalloc((void **)&op->win_d[a], sizeof(float) * g.S[a]) != hipSuccess) return cleanup_fail(MRN_ENOMEM);
    hipMemcpy(op->win_d[a], op->win_h[a].data(), sizeof(float) * g.S[a], hipMemcpyHostToDevice);
    std::vector<float2> tw((size_t)g.S[a]);
    for (int q = 0; q < g.S[a]; q++) {
      const double ang = -2.0 * M_PI * (double)q / (double)g.S[a];
      tw[q] = make_float2((float)cos(ang), (float)sin(ang));
    }
    if (amalloc((void **)&op->tw_d[a], sizeof(float2) * g.S[a]) != hipSuccess) return cleanup_fail(MRN_ENOMEM);
    hipMemcpy(op->tw_d[a], tw.data(), sizeof(float2) * g.S[a], hipMemcpyHostToDevice);
  }
  const int64_t ncell = (int64_t)g.S[0] * g.S[1] * g.S[2];
  if (batched) {
    // frames fused into one launch per stage keep F grids side by side:
    // forced (fuse_frames > 0), off (0), or when they take at most half the
    // free device memory (-1)
    op->fuse = false;
    if (op->F > 1 && p.fuse_frames != 0) {
      const double bytes = (double)sizeof(float2) * ncell * g.CC * op->nch * op->F;
      size_t fm = 0, tm = 0;
      op->fuse = p.fuse_frames > 0 || (hipMemGetInfo(&fm, &tm) == hipSuccess && bytes <= 0.5 * (double)fm);
    }
    const int64_t ngrid = ncell * g.CC * op->nch * (op->fuse ? op->F : 1);
    if (amalloc((void **)&op->grid, sizeof(float2) * ngrid) != hipSuccess) return cleanup_fail(MRN_ENOMEM);
    // zeroed once: the forward x pass leaves the cells outside a frame's
    // k-space support as they are, and readers of whole tile regions weight
    // them by zero, so they must hold finite values
    if (hipMemset(op->grid, 0, sizeof(float2) * ngrid) != hipSuccess) return cleanup_fail(MRN_EHIP);
    if (amalloc((void **)&op->smaps, sizeof(float2) * op->nvox * g.CC * op->nch) != hipSuccess)
      return cleanup_fail(MRN_ENOMEM);
  } else {
    if (amalloc((void **)&op->smaps, sizeof(float2) * op->nvox * op->C) != hipSuccess)
      return cleanup_fail(MRN_ENOMEM);
  }
  if (amalloc((void **)&op->cgrid, sizeof(float2) * ncell) != hipSuccess) return cleanup_fail(MRN_ENOMEM);
  if (amalloc((void **)&op->dscal, sizeof(double) * 16) != hipSuccess) return cleanup_fail(MRN_ENOMEM);
  if (amalloc((void **)&op->dpart, sizeof(double) * 3 * 1024) != hipSuccess) return cleanup_fail(MRN_ENOMEM);
  if (amalloc((void **)&op->dcounter, sizeof(int) * 4) != hipSuccess) return cleanup_fail(MRN_ENOMEM);
  hipMemset(op->dscal, 0, sizeof(double) * 16);
  hipMemset(op->dcounter, 0, sizeof(int) * 4);
  if (hipDeviceSynchronize() != hipSuccess) return cleanup_fail(MRN_EHIP);
  *out = op;
  return MRN_OK;
}

int mrn_op_destroy(mrn_op *op) {
  if (!op) return MRN_EINVAL;
  if (op->stream) hipStreamSynchronize(op->stream);
  for (auto &fr : op->frames) {
    hipFree(fr.coords);
    hipFree(fr.taper);
    hipFree(fr.perm);
    hipFree(fr.inv);
    hipFree(fr.bin_start);
    hipFree(fr.subs);
    hipFree(fr.xsplit);
    hipFree(fr.heavy);
    hipFree(fr.vofs);
    hipFree(fr.vlist);
    hipFree(fr.gsubs);
    hipFree(fr.sparse);
    hipFree(fr.wide);
    hipFree(fr.xrng);
    hipFree(fr.yrng);
    hipFree(fr.kdata);
    hipFree(fr.kraw);
    hipFree(fr.traw);
  }
  op->bind.release();
  hipFree(op->fdev);
  hipFree(op->kwork_all);
  hipFree(op->xrng_all);
  hipFree(op->yrng_all);
  hipFree(op->tables_d);
  for (int a = 0; a < 3; a++) {
    hipFree(op->win_d[a]);
    hipFree(op->tw_d[a]);
  }
  hipFree(op->smaps);
  hipFree(op->grid);
  hipFree(op->cgrid);
  hipFree(op->kwork);
  hipFree(op->kplain);
  if (op->cg_exec) (void)hipGraphExecDestroy(op->cg_exec);
  if (op->cg_graph) (void)hipGraphDestroy(op->cg_graph);
  hipFree(op->vR);
  hipFree(op->vP);
  hipFree(op->vL);
  hipFree(op->vX0);
  hipFree(op->dpart);
  hipFree(op->dscal);
  hipFree(op->dtrace);
  hipFree(op->dstep_log);
  hipFree(op->dcounter);
  hipFree(op->dthr);
  hipFree(op->dsh);
  hipFree(op->dthr_log);
  hipFree(op->llr_sblocks);
  hipFree(op->llr_sel);
  hipFree(op->llr_scal);
  if (op->stream) hipStreamDestroy(op->stream);
  if (op->side) hipStreamDestroy(op->side);
  if (op->ev_fork) hipEventDestroy(op->ev_fork);
  if (op->ev_join) hipEventDestroy(op->ev_join);
  delete op;
  return MRN_OK;
}

int mrn_get_geometry(const mrn_op *op, int32_t S[3], int32_t og[3], float beta[3]) {
  if (!op) return MRN_EINVAL;
  for (int a = 0; a < 3; a++) {
    if (S) S[a] = op->g.S[a];
    if (og) og[a] = op->g.og[a];
    if (beta) beta[a] = op->g.beta[a];
  }
  return MRN_OK;
}

static int copy_out(const std::vector<float> &v, float *out, int cap, int *len) {
  if (len) *len = (int)v.size();
  if (!out) return MRN_OK;
  if (cap < (int)v.size()) return MRN_EINVAL;
  memcpy(out, v.data(), sizeof(float) * v.size());
  return MRN_OK;
}

int mrn_get_table(const mrn_op *op, int axis, float *out, int cap, int *len) {
  if (!op || axis < 0 || axis > 2) return MRN_EINVAL;
  return copy_out(op->table_h[axis], out, cap, len);
}

int mrn_get_window(const mrn_op *op, int axis, float *out, int cap, int *len) {
  if (!op || axis < 0 || axis > 2) return MRN_EINVAL;
  return copy_out(op->win_h[axis], out, cap, len);
}

const char *mrn_status_string(int s) {
  switch (s) {
    case MRN_OK: return "ok";
    case MRN_EINVAL: return "invalid argument";
    case MRN_ENOMEM: return "out of device memory";
    case MRN_EHIP: return "HIP runtime error";
    case MRN_EUNSUPPORTED: return "unsupported configuration";
    case MRN_ESTATE: return "call order: required state not set";
    default: return "unknown status";
  }
}

const char *mrn_last_error(const mrn_op *op) { return op ? op->err.c_str() : mrn::free_error(); }

}  // extern "C"
