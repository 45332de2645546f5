/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of uwmri/mri_recon's Pipe density compensation: the
 * iteration of RECON::dcf_calc over DCFgridFFT's real-valued polynomial-kernel
 * gridder.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, as the checker of mri_recon_amd's mrn_dcf_calc.
 *
 * Built into liboracle_f32.so (-DORC_REAL=float: the reference's arithmetic,
 * including its float/double mix, written as the same C expressions so that
 * the usual arithmetic conversions give the same roundings) and
 * liboracle_f64.so (-DORC_REAL=double: the same discretisation in FP64).
 *
 * Reference files followed (paths relative to the reference root):
 *   src/recon_lib.cpp:124-130     defaults (dcf_iter 20, dcf_dwin 2.1, dcf_scale 1,
 *                                 dcf_overgrid 2.1, dcf_acc 1)
 *   src/recon_lib.cpp:854-917     dcf_calc: grid extents, kw = 1, the ratio update
 *                                 with its 1e-3 cut, scale_kw, |kw|
 *   src/DCFgridFFT.cpp:56-90      polynomial tables (evaluated at i / len)
 *   src/DCFgridFFT.cpp:96-165     precalc_kernel: 600 points per cell, len, norm
 *   src/DCFgridFFT.cpp:171-289    forward (scatter; kw == 0 samples skipped)
 *   src/DCFgridFFT.cpp:291-312    scale_kw (kw *= scale^3)
 *   src/DCFgridFFT.cpp:314-440    backward (gather)
 *
 * The scatter runs over the samples in order (the reference adds with OpenMP
 * atomics in whatever order the threads reach them).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

typedef struct {
  int32_t N[3];    /* rcxres, rcyres, rczres */
  int32_t iters;   /* -dcf_iter */
  float dwin;      /* -dcf_dwin (same on every axis) */
  float scale;     /* -dcf_scale: k is multiplied by it for the DCF */
  float overgrid;  /* -dcf_overgrid */
  float acc;       /* -dcf_acc */
  int32_t kernel;  /* 0 = POLY_KERNEL, 1 = POLY_KERNEL2 */
} orc_dcf_params;

/* Table length: (int)(dwin * 600) entries plus 100 (DCFgridFFT.cpp:117-126). */
int orc_dcf_table_len(float dwin) { return (int)((float)dwin * (float)600.0f) + 100; }

/* Polynomial table, normalised so that it sums to 0.5 * 600
 * (DCFgridFFT.cpp:56-90, 160-164).  The polynomial's argument is i / len,
 * len the table length including its 100 spare entries, as in the reference. */
int orc_dcf_table(float dwin, int kernel, real *tab) {
  static const double P1[6] = {1., 0.04522831, -3.36020304, 1.12417012, 2.82448025, -1.63447764};
  static const double P2[6] = {1., 0.03056504, -3.01961845, 0.6679865, 2.77924058, -1.45923643};
  const double *c = kernel == 1 ? P2 : P1;
  const int len = orc_dcf_table_len(dwin);
  for (int i = 0; i < len; i++) {
    double x = (double)i / (double)len;
    double x2 = x * x;
    double x4 = x2 * x2;
    double x3 = x * x2;
    double x5 = x * x4;
    tab[i] = (real)(c[0] + c[1] * x + c[2] * x2 + c[3] * x3 + c[4] * x4 + c[5] * x5);
  }
  double sum = 0.0; /* the array library sums float arrays in double */
  for (int i = 0; i < len; i++) sum += tab[i];
  const double f = 0.5 * (double)600.0f / sum;
  for (int i = 0; i < len; i++) tab[i] = (real)(tab[i] * f);
  return len;
}

/* Grid extents of dcf_calc (recon_lib.cpp:873-875): the array is declared
 * (os*rczres + 16, os*rcyres + 16, os*rcxres + 16) with x fastest, so the
 * extent along kx is the one computed from rczres. */
void orc_dcf_grid_size(const orc_dcf_params *p, int32_t S[3]) {
  S[0] = (int)(p->overgrid * (float)p->N[2] + 16.0f);
  S[1] = (int)(p->overgrid * (float)p->N[1] + 16.0f);
  S[2] = (int)(p->overgrid * (float)p->N[0] + 16.0f);
}

/* Per-sample kernel scale (DCFgridFFT.cpp:198-199 forward and scale_kw, in
 * double; :345-346 backward, in float). */
static real scale_fwd(real kr, float acc) { return (real)(kr / 128.0 * (acc - 1) + 1); }
static real scale_bwd(real kr, float acc) { return (real)(kr / (real)128 * (real)(acc - 1) + (real)1); }

typedef struct {
  int S[3];
  real c[3];
  real dwin, gx, mod;
  const real *tab;
} dgeom;

static void dgeom_init(dgeom *g, const orc_dcf_params *p, const real *tab) {
  int32_t S[3];
  orc_dcf_grid_size(p, S);
  for (int a = 0; a < 3; a++) {
    g->S[a] = S[a];
    g->c[a] = (real)(S[a] / 2);
  }
  g->dwin = p->dwin;
  g->gx = p->overgrid;
  g->mod = (real)600.0f;
  g->tab = tab;
}

/* Tap range of one axis; 0 when the sample lies off the grid. */
static int drange(const dgeom *g, int a, real k, real scale, real *dk, int *s, int *e) {
  *dk = k * g->gx + g->c[a];
  *s = (int)ceil(*dk - scale * g->dwin);
  if (*s < 0) *s = 0;
  *e = (int)floor(*dk + scale * g->dwin);
  if (*e > g->S[a] - 1) *e = g->S[a] - 1;
  return !(*s >= g->S[a] || *e < 0);
}

/* Interpolated table value at distance |mod * (dk - l)| / scale, in the
 * reference's expression (table float, 1.0 double). */
static double dlerp(const dgeom *g, real dk, int l, real scale) {
  real del = (real)fabs((double)(g->mod * (dk - (real)l))) / scale;
  int i = (int)del;
  real d = del - (real)i;
  return g->tab[i] * (1.0 - d) + g->tab[i + 1] * d;
}

/* DCFgridFFT::forward (DCFgridFFT.cpp:171-289): image += sum_i w_i kw_i. */
void orc_dcf_forward(const orc_dcf_params *p, const real *tab, real *X, const real *kw, int64_t M,
                     const float *kx, const float *ky, const float *kz) {
  dgeom g;
  dgeom_init(&g, p, tab);
  for (int64_t i = 0; i < M; i++) {
    const real k[3] = {(real)kx[i] * (real)p->scale, (real)ky[i] * (real)p->scale, (real)kz[i] * (real)p->scale};
    const real kr = (real)sqrt((double)(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]));
    const real scale = scale_fwd(kr, p->acc);
    const real temp = kw[i];
    if (temp == 0.0 || !(scale > 0)) continue; /* scale <= 0: no taps (0 would divide by 0) */
    real dk[3];
    int s[3], e[3];
    if (!drange(&g, 0, k[0], scale, &dk[0], &s[0], &e[0])) continue;
    if (!drange(&g, 1, k[1], scale, &dk[1], &s[1], &e[1])) continue;
    if (!drange(&g, 2, k[2], scale, &dk[2], &s[2], &e[2])) continue;
    for (int lz = s[2]; lz <= e[2]; lz++) {
      real wtz = (real)dlerp(&g, dk[2], lz, scale);
      for (int ly = s[1]; ly <= e[1]; ly++) {
        real wty = (real)(wtz * dlerp(&g, dk[1], ly, scale));
        for (int lx = s[0]; lx <= e[0]; lx++) {
          real wtx = (real)(wty * dlerp(&g, dk[0], lx, scale));
          real temp2 = wtx * temp;
          X[((int64_t)lz * g.S[1] + ly) * g.S[0] + lx] += temp2;
        }
      }
    }
  }
}

/* DCFgridFFT::backward (DCFgridFFT.cpp:314-440): out_i = sum w_i image. */
void orc_dcf_backward(const orc_dcf_params *p, const real *tab, const real *X, real *out, int64_t M,
                      const float *kx, const float *ky, const float *kz) {
  dgeom g;
  dgeom_init(&g, p, tab);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < M; i++) {
    const real k[3] = {(real)kx[i] * (real)p->scale, (real)ky[i] * (real)p->scale, (real)kz[i] * (real)p->scale};
    const real kr = (real)sqrt((double)(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]));
    const real scale = scale_bwd(kr, p->acc);
    out[i] = 0;
    if (!(scale > 0)) continue;
    real dk[3];
    int s[3], e[3];
    if (!drange(&g, 0, k[0], scale, &dk[0], &s[0], &e[0])) continue;
    if (!drange(&g, 1, k[1], scale, &dk[1], &s[1], &e[1])) continue;
    if (!drange(&g, 2, k[2], scale, &dk[2], &s[2], &e[2])) continue;
    real temp = 0;
    for (int lz = s[2]; lz <= e[2]; lz++) {
      real wtz = (real)dlerp(&g, dk[2], lz, scale);
      for (int ly = s[1]; ly <= e[1]; ly++) {
        real wty = (real)(wtz * dlerp(&g, dk[1], ly, scale));
        for (int lx = s[0]; lx <= e[0]; lx++) {
          real wtx = (real)(wty * dlerp(&g, dk[0], lx, scale));
          temp += wtx * X[((int64_t)lz * g.S[1] + ly) * g.S[0] + lx];
        }
      }
    }
    out[i] = temp;
  }
}

/* RECON::dcf_calc for one encoding (recon_lib.cpp:879-913): kw = 1, then
 * `iters` times kw <- kw / (A^T A kw) with samples whose A^T A kw is below
 * 1e-3 in magnitude set to 0, then kw *= scale^3 and kw = |kw|.  sums[it]
 * (optional) is the reference's printed sum of |kw| after iteration it. */
int orc_dcf_calc(const orc_dcf_params *p, int64_t M, const float *kx, const float *ky, const float *kz,
                 real *kw, double *sums) {
  const int len = orc_dcf_table_len(p->dwin);
  real *tab = (real *)calloc((size_t)len, sizeof(real));
  int32_t S[3];
  orc_dcf_grid_size(p, S);
  const int64_t nS = (int64_t)S[0] * S[1] * S[2];
  real *X = (real *)malloc(sizeof(real) * (size_t)nS);
  real *kw2 = (real *)malloc(sizeof(real) * (size_t)(M > 0 ? M : 1));
  if (!tab || !X || !kw2) {
    free(tab);
    free(X);
    free(kw2);
    return -1;
  }
  orc_dcf_table(p->dwin, p->kernel, tab);
  for (int64_t i = 0; i < M; i++) kw[i] = 1;
  for (int it = 0; it < p->iters; it++) {
    memset(X, 0, sizeof(real) * (size_t)nS);
    orc_dcf_forward(p, tab, X, kw, M, kx, ky, kz);
    orc_dcf_backward(p, tab, X, kw2, M, kx, ky, kz);
    real kw_sum = 0;
    for (int64_t i = 0; i < M; i++) {
      if (fabs((double)kw2[i]) < (1e-3)) {
        kw[i] = 0;
      } else {
        kw[i] = kw[i] / kw2[i];
        kw_sum += (real)fabs((double)kw[i]);
      }
    }
    if (sums) sums[it] = kw_sum;
  }
  /* scale_kw (DCFgridFFT.cpp:291-312), then kw = |kw| */
  for (int64_t i = 0; i < M; i++) {
    const real k[3] = {(real)kx[i] * (real)p->scale, (real)ky[i] * (real)p->scale, (real)kz[i] * (real)p->scale};
    const real kr = (real)sqrt((double)(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]));
    const real scale = scale_fwd(kr, p->acc);
    kw[i] = (real)(kw[i] * pow((double)scale, 3.0));
    kw[i] = (real)fabs((double)kw[i]);
  }
  free(tab);
  free(X);
  free(kw2);
  return 0;
}
