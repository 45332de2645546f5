/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of uwmri/mri_recon's gridFFT operator (the NUFFT/SENSE
 * encoding operator) and of the SENSE-CG / FISTA loops that drive it.  It is
 * the checker for the HIP library under mri_recon_amd/: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing that is measured or shipped.
 *
 * Compiled twice: -DORC_REAL=float (the reference's FP32 arithmetic, keeping
 * its float/double mix where the reference has one) and -DORC_REAL=double (the
 * same discretisation evaluated in FP64: the "FP64 oracle").
 *
 * Reference files followed (paths relative to the reference root):
 *   src/gridFFT.cpp:28-60     defaults (overgrid 1.5, KB, k_rad 9999, grid_in)
 *   src/gridFFT.cpp:187-245   -fast_grid special copy (os 1, triangle, dwin 1)
 *   src/gridFFT.cpp:251-270   poly kernel table
 *   src/gridFFT.cpp:276-473   precalc_kernel (600 pts/unit table, beta, norm)
 *   src/gridFFT.cpp:475-643   precalc_gridding (geometry, deapod windows)
 *   src/gridFFT.cpp:645-711   do_fft (pruned z,y,x) / do_ifft (x,y,z)
 *   src/ArrayTemplates.cpp:276-427  fft3: per-axis 1-D DFT, x 1/sqrt(n)
 *   src/gridFFT.cpp:791-844   forward (A^H) / backward (A)
 *   src/gridFFT.cpp:856-876   backward_residual
 *   src/gridFFT.cpp:926-1002  accumulate / set_image
 *   src/gridFFT.cpp:1009-1164 chop_grid_forward (scatter, taper, atomics)
 *   src/gridFFT.cpp:1166-1322 chop_grid_backward (gather)
 *   src/gridFFT.cpp:1325-1352 bessi0 (Abramowitz & Stegun 9.8.1 / 9.8.2)
 *   src/recon_lib.cpp:1523-1759  SENSE-CG loop
 *   src/recon_lib.cpp:1791-1915  power method for the FISTA step size
 *   src/recon_lib.cpp:1925-2105  FISTA/IST gradient step
 *   src/recon_lib.cpp:2177-2212  fista_update momentum
 *   src/wavelet3D.cpp:99-190,563-791  3-D periodic wavelet (4-tap Daubechies pair)
 *   src/threshold.cpp:250-380,766-800 fraction threshold (bisection), soft threshold
 *   src/recon_lib.cpp:2416-2527  L1_threshold (cycle-spun wavelet soft threshold)
 *
 * Pinned by tests/test_oracle_*.py (closed forms, brute-force DFT/NDFT on tiny
 * shapes, adjointness, special-case reductions, dense solves).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#include <dlfcn.h>
#endif

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

#define ORC_PI 3.14159265359 /* the reference's PI (gridFFT.h:25) */

enum { K_TRIANGLE = 0, K_KAISER = 1, K_SINC = 2, K_POLY = 3 };

typedef struct {
  int32_t N[3];        /* Nx, Ny, Nz (image size) */
  float overgrid;      /* -overgrid, default 1.5 */
  float grid[3];       /* -grid_x/y/z, -1 = derive from overgrid */
  float dwin[3];       /* -dwinX/Y/Z half widths, -1 = kernel default */
  int32_t kernel;      /* K_* */
  int32_t grid_in[3];  /* 1 = non-Cartesian axis (KB interpolation) */
  int32_t fft_in[3];   /* 1 = DFT along axis */
  float grid_scale[3]; /* -grid_scale_x/y/z */
  float k_rad;         /* taper radius, default 9999 */
} orc_params;

typedef struct {
  real re, im;
} cpx;

static inline int imin_(int a, int b) { return a < b ? a : b; }
static inline int imax_(int a, int b) { return a > b ? a : b; }

typedef struct {
  orc_params p;
  int N[3], S[3], og[3], L[3], tlen[3];
  real grid[3], dwin[3], beta[3];
  real *table[3]; /* kernel lookup tables, length L+10 */
  real *win[3];   /* deapodisation windows, length S */
} orc_op;

/* ---------------------------------------------------------------- bessi0 */
/* Abramowitz & Stegun 9.8.1 (|x|<3.75) and 9.8.2; same evaluation types as
 * gridFFT.cpp:1325-1352 in the FP32 build (double polynomial, float result). */
static real orc_bessi0(real x) {
  real ax = fabs(x);
  double y;
  if (ax < 3.75) {
    y = (double)x / 3.75;
    y *= y;
    return (real)(1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 +
                  y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2))))));
  }
  y = 3.75 / (double)ax;
  return (real)((exp((double)ax) / sqrt((double)ax)) *
                (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 +
                 y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 +
                 y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2)))))))));
}

double orc_bessi0_eval(double x) { return (double)orc_bessi0((real)x); }

/* ---------------------------------------------------- kernel tables (a2) */
static void poly_table(real *out, int len) { /* gridFFT.cpp:251-270 */
  const double c0 = 1., c1 = 0.04522831, c2 = -3.36020304, c3 = 1.12417012,
               c4 = 2.82448025, c5 = -1.63447764;
  for (int i = 0; i < len; i++) {
    double x = (double)i / (double)len, x2 = x * x, x4 = x2 * x2;
    out[i] = (real)(c0 + c1 * x + c2 * x2 + c3 * x * x2 + c4 * x4 + c5 * x * x4);
  }
}

static void build_tables(orc_op *op) { /* gridFFT.cpp:276-473 */
  const real grid_mod = 600;
  for (int a = 0; a < 3; a++) {
    real dw = op->dwin[a];
    if (dw == -1) {
      dw = (op->p.kernel == K_TRIANGLE) ? 1.0 : (op->p.kernel == K_POLY) ? 2.0 : 2.5;
    }
    op->dwin[a] = dw;
  }
  for (int a = 0; a < 3; a++) {
    int L = (int)((float)op->dwin[a] * (float)grid_mod);
    op->L[a] = L;
    op->tlen[a] = L + 10;
    real *f = (real *)calloc((size_t)L + 10, sizeof(real));
    op->table[a] = f;
    op->beta[a] = 0;
    switch (op->p.kernel) {
      case K_TRIANGLE:
        for (int i = 0; i < L + 1; i++) {
          real pos = (real)i / (real)L;
          f[i] = (real)(1.0 - pos);
        }
        break;
      case K_KAISER: {
        real g = op->grid[a] * (real)op->p.grid_scale[a];
        real dw = op->dwin[a];
        /* sqrtf of the argument rounded to `real`, as the reference calls sqrtf */
        real beta = (real)(ORC_PI * (real)sqrt((real)((dw * dw) / (g * g) * (g - 0.5) * (g - 0.5) - 0.8)));
        real beta_min = (real)sqrt(pow(ORC_PI * 2 * dw / g, 2.0) - 9.6752);
        if (beta_min > beta) beta = beta_min;
        op->beta[a] = beta;
        real i0b = orc_bessi0(beta);
        for (int i = 0; i < L + 1; i++) {
          real pos = (real)i / (real)L;
          real arg = (real)sqrt((real)(1.0 - pos * pos));
          f[i] = orc_bessi0(beta * arg) / i0b;
        }
      } break;
      case K_SINC: {
        /* the X/Y/Z Hamming terms use grid_lengthZ for X (gridFFT.cpp:414) */
        int Lz = (int)((float)op->dwin[2] * (float)grid_mod);
        int Lh = (a == 0) ? Lz : L;
        for (int i = 0; i < L + 1; i++) {
          real pos = ((real)i * op->dwin[a] / (real)op->p.grid_scale[a]) / (real)L;
          if (pos < 0.01) {
            f[i] = 1.0;
          } else {
            f[i] = (real)(sin(pos * ORC_PI) / (pos * ORC_PI));
          }
          f[i] *= (real)(0.54 + 0.46 * cos(ORC_PI * (real)i / ((real)Lh)));
        }
      } break;
      case K_POLY:
        poly_table(f, L + 10);
        break;
    }
    double s = 0;
    for (int i = 0; i < L + 10; i++) s += f[i];
    double sc = 0.5 * grid_mod / s;
    for (int i = 0; i < L + 10; i++) f[i] = (real)(f[i] * sc);
  }
}

/* --------------------------------------------- geometry + windows (a1,a3) */
static void build_windows(orc_op *op) { /* gridFFT.cpp:550-631 */
  const real grid_mod = 600;
  for (int a = 0; a < 3; a++) {
    int S = op->S[a], ogs = op->og[a], oge = op->og[a] + op->N[a];
    real *w = (real *)calloc((size_t)S, sizeof(real));
    op->win[a] = w;
    if (op->p.fft_in[a] == 1 && op->p.grid_in[a] == 1) {
      for (int i = 0; i < S; i++) {
        real ipos = (real)(i - (real)S / 2.0);
        real acc = 0;
        for (int gp = 0; gp < op->dwin[a] * grid_mod; gp++) {
          double term = 2 * cos(2 * ORC_PI * ipos * gp / (real)grid_mod / (real)S) *
                        op->table[a][gp];
          acc = (real)(acc + term);
        }
        acc = (real)grid_mod / acc;
        real fact = (real)(2 * (i % 2) - 1);
        acc *= fact / S;
        acc *= (i < ogs) ? 0.0 : 1.0;
        acc *= (i > oge - 1) ? 0.0 : 1.0;
        w[i] = acc;
      }
      real mx = w[0];
      for (int i = 1; i < S; i++)
        if (w[i] > mx) mx = w[i];
      for (int i = 0; i < S; i++) w[i] /= mx;
    } else if (op->p.fft_in[a]) {
      for (int i = 0; i < S; i++) w[i] = (real)(2 * (i % 2) - 1) / S;
    } else {
      for (int i = 0; i < S; i++) w[i] = 1.0;
    }
  }
}

orc_op *orc_create(const orc_params *p) { /* gridFFT.cpp:475-643 */
  orc_op *op = (orc_op *)calloc(1, sizeof(orc_op));
  op->p = *p;
  /* The geometry is the reference's float arithmetic in both builds: grid_x
   * is a float member, Sx = (int)(grid_x * Nx), og = (int)((float)Nx *
   * (grid_x - 1) / 2.0).  In double, 36 * (64/36 - 1) / 2 falls just below
   * 14 and truncates to 13 -- another grid offset, not a rounding effect. */
  float gf[3];
  for (int a = 0; a < 3; a++) {
    op->N[a] = p->N[a];
    op->dwin[a] = p->dwin[a];
    float g = (float)p->grid[a];
    if (p->grid_in[a] == 1) {
      if (g == -1.0f) g = (float)(16.0 * ceil(((float)p->overgrid * (float)p->N[a]) / 16.0) / (float)p->N[a]);
    } else {
      g = 1.0f;
    }
    gf[a] = g;
    op->grid[a] = (real)g;
    op->S[a] = (int)(g * p->N[a]);
  }
  build_tables(op);
  for (int a = 0; a < 3; a++) op->og[a] = (int)((float)op->N[a] * (gf[a] - 1) / 2.0);
  build_windows(op);
  return op;
}

void orc_destroy(orc_op *op) {
  if (!op) return;
  for (int a = 0; a < 3; a++) {
    free(op->table[a]);
    free(op->win[a]);
  }
  free(op);
}

void orc_geometry(const orc_op *op, int32_t *S, int32_t *og, int32_t *L, double *grid,
                  double *dwin, double *beta) {
  for (int a = 0; a < 3; a++) {
    S[a] = op->S[a];
    og[a] = op->og[a];
    L[a] = op->L[a];
    grid[a] = op->grid[a];
    dwin[a] = op->dwin[a];
    beta[a] = op->beta[a];
  }
}

int orc_table(const orc_op *op, int axis, double *out) {
  for (int i = 0; i < op->tlen[axis]; i++) out[i] = op->table[axis][i];
  return op->tlen[axis];
}

int orc_window(const orc_op *op, int axis, double *out) {
  for (int i = 0; i < op->S[axis]; i++) out[i] = op->win[axis][i];
  return op->S[axis];
}

/* ------------------------------------------------------------ FFT (a5,a8) */
/* Mixed-radix Stockham autosort 1-D DFT with twiddles computed in double and
 * rounded to `real`.  Sign -1 = FFTW_FORWARD, +1 = FFTW_BACKWARD.  Stands in
 * for FFTW's 1-D plans in fft3 (ArrayTemplates.cpp:276-427). */
typedef struct {
  int n, nf, fac[32];
  cpx *tw; /* exp(sign*2*pi*i*q/n), q < n */
} fft_plan1;

static void plan1(fft_plan1 *pl, int n, int sign) {
  pl->n = n;
  pl->nf = 0;
  int m = n;
  static const int pref[] = {4, 2, 3, 5, 7};
  for (int k = 0; k < 5; k++)
    while (m % pref[k] == 0) {
      pl->fac[pl->nf++] = pref[k];
      m /= pref[k];
    }
  for (int f = 3; m > 1; f += 2) /* any remaining primes: generic radix */
    while (m % f == 0) {
      pl->fac[pl->nf++] = f;
      m /= f;
    }
  pl->tw = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  for (int q = 0; q < n; q++) {
    double ang = sign * 2.0 * M_PI * (double)q / (double)n;
    pl->tw[q].re = (real)cos(ang);
    pl->tw[q].im = (real)sin(ang);
  }
}

static void line_fft(const fft_plan1 *pl, cpx *a, cpx *b) {
  int n = pl->n, Ns = 1;
  cpx *src = a, *dst = b, v[64], o[64];
  for (int f = 0; f < pl->nf; f++) {
    int R = pl->fac[f], nR = n / R, tstep = n / (Ns * R);
    for (int j = 0; j < nR; j++) {
      int jm = j % Ns;
      for (int k = 0; k < R; k++) {
        cpx x = src[j + k * nR];
        cpx t = pl->tw[(jm * k * tstep) % n];
        v[k].re = x.re * t.re - x.im * t.im;
        v[k].im = x.re * t.im + x.im * t.re;
      }
      for (int q = 0; q < R; q++) {
        real sr = 0, si = 0;
        for (int k = 0; k < R; k++) {
          cpx t = pl->tw[((q * k) % R) * nR];
          sr += v[k].re * t.re - v[k].im * t.im;
          si += v[k].re * t.im + v[k].im * t.re;
        }
        o[q].re = sr;
        o[q].im = si;
      }
      int idxD = (j / Ns) * Ns * R + jm;
      for (int q = 0; q < R; q++) dst[idxD + q * Ns] = o[q];
    }
    cpx *t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  if (src != a) memcpy(a, src, sizeof(cpx) * (size_t)n);
}

/* Unitary 1-D DFT of `howmany` contiguous lines (test entry point). */
void orc_fft1d(int n, int sign, real *data, int howmany) {
  fft_plan1 pl;
  plan1(&pl, n, sign);
  real scale = (real)(1. / sqrt((real)n));
  cpx *tmp = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  for (int h = 0; h < howmany; h++) {
    cpx *ln = (cpx *)data + (size_t)h * n;
    line_fft(&pl, ln, tmp);
    for (int i = 0; i < n; i++) {
      ln[i].re *= scale;
      ln[i].im *= scale;
    }
  }
  free(tmp);
  free(pl.tw);
}

/* Optional line transforms through MKL's FFTW3 interface (libmkl_rt with its
 * sequential threading layer, one line per OpenMP thread at a time, as the
 * reference runs FFTW: ArrayTemplates.cpp:276-427).  Only the CPU timing
 * baseline turns it on; the checker keeps the textbook transform above. */
typedef void *(*mkl_plan_fn)(int, void *, void *, int, unsigned);
typedef void (*mkl_exec_fn)(const void *, void *, void *);
typedef void (*mkl_destroy_fn)(void *);
static struct {
  int on;
  mkl_plan_fn plan;
  mkl_exec_fn exec;
  mkl_destroy_fn destroy;
} g_mkl;

int orc_use_mkl_fft(int on) {
  if (!on) {
    g_mkl.on = 0;
    return 0;
  }
  if (!g_mkl.plan) {
    setenv("MKL_THREADING_LAYER", "SEQUENTIAL", 0);
    static const char *paths[] = {"libmkl_rt.so", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so"};
    void *h = NULL;
    for (int i = 0; i < 3 && !h; i++) h = dlopen(paths[i], RTLD_NOW | RTLD_LOCAL);
    if (!h) return 0;
    const int f = sizeof(real) == 4;
    g_mkl.plan = (mkl_plan_fn)dlsym(h, f ? "fftwf_plan_dft_1d" : "fftw_plan_dft_1d");
    g_mkl.exec = (mkl_exec_fn)dlsym(h, f ? "fftwf_execute_dft" : "fftw_execute_dft");
    g_mkl.destroy = (mkl_destroy_fn)dlsym(h, f ? "fftwf_destroy_plan" : "fftw_destroy_plan");
    if (!g_mkl.plan || !g_mkl.exec || !g_mkl.destroy) {
      g_mkl.plan = NULL;
      return 0;
    }
  }
  g_mkl.on = 1;
  return 1;
}

/* fft3 on a box [lo,hi) of a column-major grid (x fastest) along `axis`:
 * every line of the box is transformed over the full extent S[axis]. */
static void fft_axis(cpx *g, const int S[3], int axis, int sign, const int lo[3],
                     const int hi[3]) {
  int n = S[axis];
  fft_plan1 pl;
  plan1(&pl, n, sign);
  real scale = (real)(1. / sqrt((real)n));
  long st[3] = {1, S[0], (long)S[0] * S[1]};
  int a1 = (axis == 0) ? 1 : 0, a2 = (axis == 2) ? 1 : 2;
  long nl1 = hi[a1] - lo[a1], nl2 = hi[a2] - lo[a2];
  void *mp = NULL;
  const size_t lb = ((sizeof(cpx) * (size_t)n + 63) / 64) * 64;  /* 64-byte aligned line buffers */
  if (g_mkl.on) {
    cpx *pa = (cpx *)aligned_alloc(64, 2 * lb);
    mp = g_mkl.plan(n, pa, (char *)pa + lb, sign, 1u << 6 /* FFTW_ESTIMATE */);
    free(pa);
  }
#pragma omp parallel
  {
    cpx *buf = (cpx *)aligned_alloc(64, 2 * lb);
    cpx *out = (cpx *)((char *)buf + lb);
#pragma omp for schedule(static)
    for (long l = 0; l < nl1 * nl2; l++) {
      long i1 = lo[a1] + l % nl1, i2 = lo[a2] + l / nl1;
      cpx *base = g + i1 * st[a1] + i2 * st[a2];
      for (int p = 0; p < n; p++) buf[p] = base[p * st[axis]];
      const cpx *res = buf;
      if (mp) {
        g_mkl.exec(mp, buf, out);
        res = out;
      } else {
        line_fft(&pl, buf, out);
      }
      for (int p = 0; p < n; p++) {
        base[p * st[axis]].re = res[p].re * scale;
        base[p * st[axis]].im = res[p].im * scale;
      }
    }
    free(buf);
  }
  if (mp) g_mkl.destroy(mp);
  free(pl.tw);
}

/* do_fft (gridFFT.cpp:645-671): z over (N+1)^2 lines, y over (N+1)*S, x all.
 * Blitz Range(og, og_e) is inclusive; lines outside the crop are zero. */
static void do_fft(orc_op *op, cpx *g) {
  const int *S = op->S;
  if (op->p.fft_in[2]) {
    int lo[3] = {op->og[0], op->og[1], 0};
    int hi[3] = {imin_(op->og[0] + op->N[0] + 1, S[0]), imin_(op->og[1] + op->N[1] + 1, S[1]), S[2]};
    fft_axis(g, S, 2, -1, lo, hi);
  }
  if (op->p.fft_in[1]) {
    int lo[3] = {op->og[0], 0, 0};
    int hi[3] = {imin_(op->og[0] + op->N[0] + 1, S[0]), S[1], S[2]};
    fft_axis(g, S, 1, -1, lo, hi);
  }
  if (op->p.fft_in[0]) {
    int lo[3] = {0, 0, 0};
    int hi[3] = {S[0], S[1], S[2]};
    fft_axis(g, S, 0, -1, lo, hi);
  }
}

/* do_ifft (gridFFT.cpp:673-711): x, y, z over every line (unpruned). */
static void do_ifft(orc_op *op, cpx *g) {
  int lo[3] = {0, 0, 0};
  int hi[3] = {op->S[0], op->S[1], op->S[2]};
  if (op->p.fft_in[0]) fft_axis(g, op->S, 0, +1, lo, hi);
  if (op->p.fft_in[1]) fft_axis(g, op->S, 1, +1, lo, hi);
  if (op->p.fft_in[2]) fft_axis(g, op->S, 2, +1, lo, hi);
}

/* ------------------------------------------------ set_image / accumulate */
static inline cpx cmul(cpx a, cpx b) {
  cpx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
static inline cpx cmulc(cpx a, cpx b) { /* a * conj(b) */
  cpx r = {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
  return r;
}

/* gridFFT.cpp:829 (k3d_grid = 0) + set_image 972-1002 */
static void set_image(orc_op *op, cpx *g, const cpx *X, const cpx *smap) {
  const int *S = op->S, *N = op->N, *og = op->og;
  memset(g, 0, sizeof(cpx) * (size_t)S[0] * S[1] * S[2]);
#pragma omp parallel for
  for (int k = 0; k < N[2]; k++) {
    real wtz = op->win[2][k + og[2]];
    for (int j = 0; j < N[1]; j++) {
      real wty = wtz * op->win[1][j + og[1]];
      for (int i = 0; i < N[0]; i++) {
        real wt = wty * op->win[0][i + og[0]];
        long r = i + (long)N[0] * (j + (long)N[1] * k);
        cpx v = {wt * X[r].re, wt * X[r].im};
        if (smap) v = cmul(v, smap[r]);
        g[(i + og[0]) + (long)S[0] * ((j + og[1]) + (long)S[1] * (k + og[2]))] = v;
      }
    }
  }
}

/* gridFFT.cpp:926-958: X += wt * image * conj(smap) (accumulates) */
static void accumulate(orc_op *op, const cpx *g, cpx *X, const cpx *smap) {
  const int *S = op->S, *N = op->N, *og = op->og;
#pragma omp parallel for
  for (int k = 0; k < N[2]; k++) {
    real wtz = op->win[2][k + og[2]];
    for (int j = 0; j < N[1]; j++) {
      real wty = wtz * op->win[1][j + og[1]];
      for (int i = 0; i < N[0]; i++) {
        real wt = wty * op->win[0][i + og[0]];
        long r = i + (long)N[0] * (j + (long)N[1] * k);
        cpx im = g[(i + og[0]) + (long)S[0] * ((j + og[1]) + (long)S[1] * (k + og[2]))];
        cpx v = {wt * im.re, wt * im.im};
        if (smap) v = cmulc(v, smap[r]);
        X[r].re += v.re;
        X[r].im += v.im;
      }
    }
  }
}

/* gridFFT.cpp:912-924 then 960-970: image *= wt (deapp), X += image * conj(image) */
static void accumulate_sos(orc_op *op, const cpx *g, cpx *X) {
  const int *S = op->S, *N = op->N, *og = op->og;
#pragma omp parallel for
  for (int k = 0; k < N[2]; k++) {
    real wtz = op->win[2][k + og[2]];
    for (int j = 0; j < N[1]; j++) {
      real wty = wtz * op->win[1][j + og[1]];
      for (int i = 0; i < N[0]; i++) {
        real wt = wty * op->win[0][i + og[0]];
        long r = i + (long)N[0] * (j + (long)N[1] * k);
        cpx im = g[(i + og[0]) + (long)S[0] * ((j + og[1]) + (long)S[1] * (k + og[2]))];
        cpx v = {im.re * wt, im.im * wt};
        X[r].re += v.re * v.re + v.im * v.im;
      }
    }
  }
}

/* ----------------------------------------------------- tap weights (a6,a7) */
/* Per axis: first/last tap and the table-lerp factor of each tap, as in the
 * tap loops of chop_grid_backward/forward (gridFFT.cpp:1211-1271, 1276-1306).
 * lerp[] keeps the double-typed lerp value the reference multiplies in. */
#define MAXTAP 64
typedef struct {
  int s, e;
  double lerp[MAXTAP];
  real sgn[MAXTAP];
} taps1;

/* returns 0 if the sample is outside the grid on this axis */
static int axis_taps(const orc_op *op, int a, real k, taps1 *t, real *dk_out) {
  const real grid_mod = 600;
  int S = op->S[a];
  real c = (real)(S / 2);
  real dk = k * op->grid[a] + c;
  *dk_out = dk;
  if (op->p.grid_in[a]) {
    t->s = imax_((int)ceil(dk - op->dwin[a]), 0);
    t->e = imin_((int)floor(dk + op->dwin[a]), S - 1);
  } else {
    t->s = (int)(dk);
    t->e = t->s;
  }
  if (t->s >= S) return 0;
  if (t->e < 0) return 0;
  if (t->e - t->s + 1 > MAXTAP) return 0;
  for (int l = t->s; l <= t->e; l++) {
    real del = fabs(grid_mod * (dk - (real)l));
    real fr = del - (real)((int)del);
    int i0 = (int)del;
    const real *f = op->table[a];
    int n = op->tlen[a];
    double f0 = (i0 < n) ? f[i0] : 0.0, f1 = (i0 + 1 < n) ? f[i0 + 1] : 0.0;
    double v = f0 * (1.0 - (double)fr) + (double)((real)f1 * fr);
    /* the z weight is forced to 1 on a Cartesian z axis; x and y keep the
     * table value (gridFFT.cpp:1116-1118 has no x/y counterpart) */
    if (a == 2 && !op->p.grid_in[2]) v = 1.0;
    t->lerp[l - t->s] = v;
    t->sgn[l - t->s] = op->p.fft_in[a] ? (real)(2 * (l % 2) - 1) : (real)1;
  }
  return 1;
}

/* chop_grid_backward (gridFFT.cpp:1166-1322): gather, A's last stage. */
static void gather(const orc_op *op, const cpx *g, cpx *out, const float *kx, const float *ky,
                   const float *kz, const float *kw, long M, const cpx *sub) {
  const int *S = op->S;
#pragma omp parallel for schedule(dynamic, 1024)
  for (long m = 0; m < M; m++) {
    taps1 tx, ty, tz;
    real dk;
    cpx zero = {0, 0};
    if (kw[m] == 0.0f || !axis_taps(op, 0, kx[m], &tx, &dk) || !axis_taps(op, 1, ky[m], &ty, &dk) ||
        !axis_taps(op, 2, kz[m], &tz, &dk)) {
      out[m] = zero;
      continue;
    }
    real tr = 0, ti = 0;
    for (int lz = tz.s; lz <= tz.e; lz++) {
      real wtz = (real)tz.lerp[lz - tz.s];
      wtz *= tz.sgn[lz - tz.s];
      for (int ly = ty.s; ly <= ty.e; ly++) {
        real wty = (real)((double)wtz * ty.lerp[ly - ty.s]);
        wty *= ty.sgn[ly - ty.s];
        const cpx *row = g + (long)S[0] * (ly + (long)S[1] * lz);
        for (int lx = tx.s; lx <= tx.e; lx++) {
          real wtx = (real)((double)wty * tx.lerp[lx - tx.s]);
          wtx *= tx.sgn[lx - tx.s];
          tr += wtx * row[lx].re;
          ti += wtx * row[lx].im;
        }
      }
    }
    if (sub) {
      out[m].re = tr - sub[m].re;
      out[m].im = ti - sub[m].im;
    } else {
      out[m].re = tr;
      out[m].im = ti;
    }
  }
}

/* chop_grid_forward (gridFFT.cpp:1009-1164): scatter with kw and taper. */
static void scatter(const orc_op *op, cpx *g, const cpx *data, const float *kx, const float *ky,
                    const float *kz, const float *kw, long M) {
  const int *S = op->S;
  real k_rad = op->p.k_rad;
#pragma omp parallel for schedule(dynamic, 1024)
  for (long m = 0; m < M; m++) {
    real w = kw[m];
    if (w == 0) continue;
    cpx t = {data[m].re * w, data[m].im * w};
    if (t.re == 0 && t.im == 0) continue;
    taps1 tx, ty, tz;
    real dk;
    if (!axis_taps(op, 0, kx[m], &tx, &dk) || !axis_taps(op, 1, ky[m], &ty, &dk) ||
        !axis_taps(op, 2, kz[m], &tz, &dk))
      continue;
    real kr = 0;
    if (op->p.fft_in[0]) kr += (real)kx[m] * (real)kx[m];
    if (op->p.fft_in[1]) kr += (real)ky[m] * (real)ky[m];
    if (op->p.fft_in[2]) kr += (real)kz[m] * (real)kz[m];
    real taper = (real)exp(-kr / (2.0 * k_rad * k_rad));
    t.re *= taper;
    t.im *= taper;
    for (int lz = tz.s; lz <= tz.e; lz++) {
      real wtz = (real)tz.lerp[lz - tz.s];
      wtz *= tz.sgn[lz - tz.s];
      for (int ly = ty.s; ly <= ty.e; ly++) {
        real wty = (real)((double)wtz * ty.lerp[ly - ty.s]);
        wty *= ty.sgn[ly - ty.s];
        cpx *row = g + (long)S[0] * (ly + (long)S[1] * lz);
        for (int lx = tx.s; lx <= tx.e; lx++) {
          real wtx = (real)((double)wty * tx.lerp[lx - tx.s]);
          wtx *= tx.sgn[lx - tx.s];
          real RD = wtx * t.re, ID = wtx * t.im;
#pragma omp atomic
          row[lx].re += RD;
#pragma omp atomic
          row[lx].im += ID;
        }
      }
    }
  }
}

/* ----------------------------------------------------- per-coil operator */
static cpx *grid_alloc(const orc_op *op) {
  return (cpx *)malloc(sizeof(cpx) * (size_t)op->S[0] * op->S[1] * op->S[2]);
}

/* gridFFT::backward / backward_residual (gridFFT.cpp:824-876): A.
 * data[m] = (A_c X)[m]  (minus sub[m] when sub != NULL); overwrites data. */
void orc_backward(orc_op *op, const real *X, const real *smap, real *data, const float *kx,
                  const float *ky, const float *kz, const float *kw, long M, const real *sub) {
  cpx *g = grid_alloc(op);
  set_image(op, g, (const cpx *)X, (const cpx *)smap);
  do_fft(op, g);
  gather(op, g, (cpx *)data, kx, ky, kz, kw, M, (const cpx *)sub);
  free(g);
}

/* gridFFT::forward (gridFFT.cpp:791-812): X += A_c^H (kw * taper * data). */
void orc_forward(orc_op *op, real *X, const real *smap, const real *data, const float *kx,
                 const float *ky, const float *kz, const float *kw, long M) {
  cpx *g = grid_alloc(op);
  memset(g, 0, sizeof(cpx) * (size_t)op->S[0] * op->S[1] * op->S[2]);
  scatter(op, g, (const cpx *)data, kx, ky, kz, kw, M);
  do_ifft(op, g);
  accumulate(op, g, (cpx *)X, (const cpx *)smap);
  free(g);
}

/* gridFFT::forward_sos (gridFFT.cpp:722-748): X += |deapodised A^H (kw taper data)|^2. */
void orc_forward_sos(orc_op *op, real *X, const real *data, const float *kx, const float *ky,
                     const float *kz, const float *kw, long M) {
  cpx *g = grid_alloc(op);
  memset(g, 0, sizeof(cpx) * (size_t)op->S[0] * op->S[1] * op->S[2]);
  scatter(op, g, (const cpx *)data, kx, ky, kz, kw, M);
  do_ifft(op, g);
  accumulate_sos(op, g, (cpx *)X);
  free(g);
}

/* Intermediate stages for stage-level parity tests. */
void orc_stage_pre_fft(orc_op *op, const real *X, const real *smap, real *grid_out) {
  set_image(op, (cpx *)grid_out, (const cpx *)X, (const cpx *)smap);
  do_fft(op, (cpx *)grid_out);
}
void orc_stage_gather(orc_op *op, const real *grid, real *data, const float *kx, const float *ky,
                      const float *kz, const float *kw, long M, const real *sub) {
  gather(op, (const cpx *)grid, (cpx *)data, kx, ky, kz, kw, M, (const cpx *)sub);
}
void orc_stage_scatter(orc_op *op, real *grid_out, const real *data, const float *kx,
                       const float *ky, const float *kz, const float *kw, long M) {
  memset(grid_out, 0, sizeof(cpx) * (size_t)op->S[0] * op->S[1] * op->S[2]);
  scatter(op, (cpx *)grid_out, (const cpx *)data, kx, ky, kz, kw, M);
}
void orc_stage_ifft_post(orc_op *op, real *grid, real *X, const real *smap) {
  do_ifft(op, (cpx *)grid);
  accumulate(op, (const cpx *)grid, (cpx *)X, (const cpx *)smap);
}
void orc_fft3(orc_op *op, real *grid, int sign) {
  if (sign < 0)
    do_fft(op, (cpx *)grid);
  else
    do_ifft(op, (cpx *)grid);
}

/* ------------------------------------------------------- multi-coil loops */
static long nvox(const orc_op *op) { return (long)op->N[0] * op->N[1] * op->N[2]; }

/* data[c][m] = A_c X (- sub[c][m]) for every coil (smaps[c] per coil). */
void orc_image_to_kspace(orc_op *op, const real *X, const real *smaps, int C, real *data,
                         const float *kx, const float *ky, const float *kz, const float *kw,
                         long M, const real *sub) {
  long nv = nvox(op);
  for (int c = 0; c < C; c++)
    orc_backward(op, X, smaps ? smaps + 2 * nv * c : NULL, data + 2 * M * c, kx, ky, kz, kw, M,
                 sub ? sub + 2 * M * c : NULL);
}

/* X += sum_c A_c^H data[c] (coil loop of recon_lib.cpp:1565-1567). */
void orc_kspace_to_image(orc_op *op, real *X, const real *smaps, int C, const real *data,
                         const float *kx, const float *ky, const float *kz, const float *kw,
                         long M) {
  long nv = nvox(op);
  for (int c = 0; c < C; c++)
    orc_forward(op, X, smaps ? smaps + 2 * nv * c : NULL, data + 2 * M * c, kx, ky, kz, kw, M);
}

/* out = sum_c A_c^H W A_c P  (recon_lib.cpp:1614-1622); out is overwritten. */
void orc_normal(orc_op *op, const real *P, const real *smaps, int C, real *out, const float *kx,
                const float *ky, const float *kz, const float *kw, long M) {
  long nv = nvox(op);
  memset(out, 0, sizeof(cpx) * (size_t)nv);
  real *diff = (real *)malloc(sizeof(cpx) * (size_t)M);
  for (int c = 0; c < C; c++) {
    memset(diff, 0, sizeof(cpx) * (size_t)M);
    orc_backward(op, P, smaps ? smaps + 2 * nv * c : NULL, diff, kx, ky, kz, kw, M, NULL);
    orc_forward(op, out, smaps ? smaps + 2 * nv * c : NULL, diff, kx, ky, kz, kw, M);
  }
  free(diff);
}

/* ------------------------------------------------------------ SENSE-CG */
/* recon_lib.cpp:1523-1759 for one frame/encode, l2reg off.  X is zeroed
 * first.  trace (nullable) receives per iteration {sum_R0_R0, Re/Im alpha,
 * sum_R_R}: 4 doubles per iteration. */
void orc_cg(orc_op *op, real *Xo, const real *smaps, int C, const real *data, const float *kx,
            const float *ky, const float *kz, const float *kw, long M, int iters, double *trace) {
  long nv = nvox(op);
  int Nz = op->N[2];
  long plane = (long)op->N[0] * op->N[1];
  cpx *X = (cpx *)Xo;
  cpx *R = (cpx *)calloc((size_t)nv, sizeof(cpx));
  cpx *P = (cpx *)malloc(sizeof(cpx) * (size_t)nv);
  cpx *LHS = (cpx *)malloc(sizeof(cpx) * (size_t)nv);
  double *a_rr = (double *)malloc(sizeof(double) * Nz), *a_pl_r = (double *)malloc(sizeof(double) * Nz),
         *a_pl_i = (double *)malloc(sizeof(double) * Nz), *a_rn = (double *)malloc(sizeof(double) * Nz);
  memset(X, 0, sizeof(cpx) * (size_t)nv);
  orc_kspace_to_image(op, (real *)R, smaps, C, data, kx, ky, kz, kw, M);
  memcpy(P, R, sizeof(cpx) * (size_t)nv);
  for (int it = 0; it < iters; it++) {
    orc_normal(op, (const real *)P, smaps, C, (real *)LHS, kx, ky, kz, kw, M);
#pragma omp parallel for
    for (int k = 0; k < Nz; k++) {
      double rr = 0, pr = 0, pi = 0;
      for (long q = k * plane; q < (k + 1) * plane; q++) {
        real nr = R[q].re * R[q].re + R[q].im * R[q].im;
        rr += nr;
        real cr = P[q].re * LHS[q].re + P[q].im * LHS[q].im; /* conj(P)*LHS */
        real ci = P[q].re * LHS[q].im - P[q].im * LHS[q].re;
        pr += cr;
        pi += ci;
      }
      a_rr[k] = rr;
      a_pl_r[k] = pr;
      a_pl_i[k] = pi;
    }
    double s_rr = 0, s_pr = 0, s_pi = 0;
    for (int k = 0; k < Nz; k++) {
      s_rr += a_rr[k];
      s_pr += a_pl_r[k];
      s_pi += a_pl_i[k];
    }
    double den = s_pr * s_pr + s_pi * s_pi;
    double al_r = (s_rr * s_pr) / den, al_i = (-s_rr * s_pi) / den;
    real ar = (real)al_r, ai = (real)al_i;
#pragma omp parallel for
    for (int k = 0; k < Nz; k++) {
      double rn = 0;
      for (long q = k * plane; q < (k + 1) * plane; q++) {
        X[q].re += ar * P[q].re - ai * P[q].im;
        X[q].im += ar * P[q].im + ai * P[q].re;
        R[q].re -= ar * LHS[q].re - ai * LHS[q].im;
        R[q].im -= ar * LHS[q].im + ai * LHS[q].re;
        real nr = R[q].re * R[q].re + R[q].im * R[q].im;
        rn += nr;
      }
      a_rn[k] = rn;
    }
    double s_rn = 0;
    for (int k = 0; k < Nz; k++) s_rn += a_rn[k];
    real be = (real)(s_rn / s_rr);
#pragma omp parallel for
    for (long q = 0; q < nv; q++) {
      P[q].re = R[q].re + be * P[q].re;
      P[q].im = R[q].im + be * P[q].im;
    }
    if (trace) {
      trace[4 * it + 0] = s_rr;
      trace[4 * it + 1] = al_r;
      trace[4 * it + 2] = al_i;
      trace[4 * it + 3] = s_rn;
    }
  }
  free(R);
  free(P);
  free(LHS);
  free(a_rr);
  free(a_pl_r);
  free(a_pl_i);
  free(a_rn);
}

/* --------------------------------------------------------- FISTA pieces */
static double energy(const cpx *x, long n) { /* ArrayEnergy, ArrayTemplates.hpp:132-139 */
  double e = 0;
  for (long q = 0; q < n; q++) e += (double)(x[q].re * x[q].re + x[q].im * x[q].im);
  return e;
}

/* Power method, recon_lib.cpp:1803-1906.  The seed std::rand()/RAND_MAX - 0.5
 * is integer division, so X starts at -0.5-0.5i (up to the rare rand()==
 * RAND_MAX); that constant seed is used here.  Returns 1/max_eig, the step. */
double orc_power_method(orc_op *op, const real *smaps, int C, const float *kx, const float *ky,
                        const float *kz, const float *kw, long M, int iters, double *eig_trace) {
  long nv = nvox(op);
  cpx *X = (cpx *)malloc(sizeof(cpx) * (size_t)nv);
  cpx *R = (cpx *)malloc(sizeof(cpx) * (size_t)nv);
  for (long q = 0; q < nv; q++) {
    X[q].re = (real)-0.5;
    X[q].im = (real)-0.5;
  }
  double step = 0;
  for (int it = 0; it < iters; it++) {
    orc_normal(op, (const real *)X, smaps, C, (real *)R, kx, ky, kz, kw, M);
    double me = sqrt(energy(R, nv));
    real sc = (real)(1. / me);
    for (long q = 0; q < nv; q++) {
      X[q].re = sc * R[q].re;
      X[q].im = sc * R[q].im;
    }
    step = (double)(float)(1. / me);
    if (eig_trace) eig_trace[it] = me;
  }
  free(X);
  free(R);
  return step;
}

/* FISTA without a regulariser (recon_lib.cpp:1925-2105 + 2177-2212):
 * X starts at zero; per iteration fista_update, R = sum_c A_c^H W (A_c X - d_c),
 * X -= step * R.  resid (nullable) receives scale_RhR per iteration. */
void orc_fista(orc_op *op, real *Xo, const real *smaps, int C, const real *data, const float *kx,
               const float *ky, const float *kz, const float *kw, long M, int iters, double step,
               double *resid) {
  long nv = nvox(op);
  cpx *X = (cpx *)Xo;
  cpx *Xold = (cpx *)calloc((size_t)nv, sizeof(cpx));
  cpx *R = (cpx *)malloc(sizeof(cpx) * (size_t)nv);
  real *diff = (real *)malloc(sizeof(cpx) * (size_t)M);
  real st = (real)step;
  memset(X, 0, sizeof(cpx) * (size_t)nv);
  for (int it = 0; it < iters; it++) {
    memset(R, 0, sizeof(cpx) * (size_t)nv);
    real A = (real)(1.0 + (it - 1.0) / (it + 1.0));
    real B = (real)(-(it - 1.0) / (it + 1.0));
    for (long q = 0; q < nv; q++) {
      cpx x0 = Xold[q], x1 = X[q];
      X[q].re = A * x1.re + B * x0.re;
      X[q].im = A * x1.im + B * x0.im;
      Xold[q] = x1;
    }
    for (int c = 0; c < C; c++) {
      orc_backward(op, (const real *)X, smaps ? smaps + 2 * nv * c : NULL, diff, kx, ky, kz, kw,
                   M, data + 2 * M * c);
      orc_forward(op, (real *)R, smaps ? smaps + 2 * nv * c : NULL, diff, kx, ky, kz, kw, M);
    }
    if (resid) resid[it] = energy(R, nv);
    for (long q = 0; q < nv; q++) {
      real rr = R[q].re * st, ri = R[q].im * st;
      X[q].re -= rr;
      X[q].im -= ri;
    }
  }
  free(Xold);
  free(R);
  free(diff);
}

/* FISTA with the Cauchy step (recon_lib.cpp:2013-2041, 2093-2100): for
 * it < update_number, P = sum_c A_c^H W A_c R, step = complex<float>(RhR /
 * RhP), RhR = sum |R|^2, RhP = sum conj(R) P (complex double); then R *=
 * step (complex float), X -= R.  `step` is the value before the first
 * update; steps (nullable) receives re, im per iteration. */
void orc_fista_cauchy(orc_op *op, real *Xo, const real *smaps, int C, const real *data, const float *kx,
                      const float *ky, const float *kz, const float *kw, long M, int iters, double step,
                      int update_number, double *resid, double *steps) {
  long nv = nvox(op);
  cpx *X = (cpx *)Xo;
  cpx *Xold = (cpx *)calloc((size_t)nv, sizeof(cpx));
  cpx *R = (cpx *)malloc(sizeof(cpx) * (size_t)nv);
  cpx *P = (cpx *)malloc(sizeof(cpx) * (size_t)nv);
  real *diff = (real *)malloc(sizeof(cpx) * (size_t)M);
  float sr = (float)step, si = 0.0f;
  memset(X, 0, sizeof(cpx) * (size_t)nv);
  for (int it = 0; it < iters; it++) {
    memset(R, 0, sizeof(cpx) * (size_t)nv);
    real A = (real)(1.0 + (it - 1.0) / (it + 1.0));
    real B = (real)(-(it - 1.0) / (it + 1.0));
    for (long q = 0; q < nv; q++) {
      cpx x0 = Xold[q], x1 = X[q];
      X[q].re = A * x1.re + B * x0.re;
      X[q].im = A * x1.im + B * x0.im;
      Xold[q] = x1;
    }
    for (int c = 0; c < C; c++) {
      orc_backward(op, (const real *)X, smaps ? smaps + 2 * nv * c : NULL, diff, kx, ky, kz, kw, M,
                   data + 2 * M * c);
      orc_forward(op, (real *)R, smaps ? smaps + 2 * nv * c : NULL, diff, kx, ky, kz, kw, M);
    }
    double rr = energy(R, nv);
    if (it < update_number) {
      memset(P, 0, sizeof(cpx) * (size_t)nv);
      for (int c = 0; c < C; c++) {
        orc_backward(op, (const real *)R, smaps ? smaps + 2 * nv * c : NULL, diff, kx, ky, kz, kw, M, NULL);
        orc_forward(op, (real *)P, smaps ? smaps + 2 * nv * c : NULL, diff, kx, ky, kz, kw, M);
      }
      double pr = 0, pi = 0;
      for (long q = 0; q < nv; q++) {  /* P *= conj(R), summed */
        pr += (double)(P[q].re * R[q].re + P[q].im * R[q].im);
        pi += (double)(P[q].im * R[q].re - P[q].re * R[q].im);
      }
      const double den = pr * pr + pi * pi;
      sr = (float)(rr * pr / den);
      si = (float)(-rr * pi / den);
    }
    if (resid) resid[it] = rr;
    if (steps) {
      steps[2 * it] = sr;
      steps[2 * it + 1] = si;
    }
    for (long q = 0; q < nv; q++) {
      real r0 = R[q].re * (real)sr - R[q].im * (real)si, r1 = R[q].re * (real)si + R[q].im * (real)sr;
      X[q].re -= r0;
      X[q].im -= r1;
    }
  }
  free(Xold);
  free(R);
  free(P);
  free(diff);
}

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ====================================================== L1 wavelet proximal step
 * WAVELET3D (wavelet3D.cpp) with its default 4-tap Daubechies pair: lowpass
 * h = ((1+r3), (3+r3), (3-r3), (1-r3)) / (4 r2) in reversed order, highpass its
 * quadrature mirror, synthesis the time reverses (wavelet3D.cpp:143-190).
 * The coefficients are computed here from that closed form. */
#define ORC_WN 4
static void d4_filters(real lpf[4], real hpf[4], real slpf[4], real shpf[4]) {
  const double r3 = sqrt(3.0), n = 4.0 * sqrt(2.0);
  const double h0 = (1 + r3) / n, h1 = (3 + r3) / n, h2 = (3 - r3) / n, h3 = (1 - r3) / n;
  lpf[0] = (real)h3; lpf[1] = (real)h2; lpf[2] = (real)h1; lpf[3] = (real)h0;
  hpf[0] = (real)-h0; hpf[1] = (real)h1; hpf[2] = (real)-h2; hpf[3] = (real)h3;
  slpf[0] = (real)h0; slpf[1] = (real)h1; slpf[2] = (real)h2; slpf[3] = (real)h3;
  shpf[0] = (real)h3; shpf[1] = (real)-h2; shpf[2] = (real)h1; shpf[3] = (real)-h0;
}

/* levels per axis: the requested count, lowered until N / 2^L is an integer
 * (setup_wavelet_directions, wavelet3D.cpp:99-115) */
void orc_wavelet_levels(const int32_t N[3], const int32_t req[3], int32_t L[3]) {
  for (int a = 0; a < 3; a++) {
    int l = req[a];
    while (l > 0 && (N[a] % (1 << l)) != 0) l--;
    L[a] = l;
  }
}

/* index i modulo pts for any sign (the reference's (i + pts) % pts, which
 * only stays in range for lines at least as long as the filter) */
static int wmod(int i, int pts) { return ((i % pts) + pts) % pts; }

/* wave1D / iwave1D (wavelet3D.cpp:754-791), periodic, shifted by sh */
static void wave1d(const cpx *in, cpx *out, int pts, int sh, const real *lpf, const real *hpf) {
  for (int pos = 0; pos < pts / 2; pos++) {
    cpx lo = {0, 0}, hi = {0, 0};
    for (int fp = 0; fp < ORC_WN; fp++) {
      const int c = wmod(2 * pos - fp + sh, pts);
      lo.re += lpf[fp] * in[c].re;
      lo.im += lpf[fp] * in[c].im;
      hi.re += hpf[fp] * in[c].re;
      hi.im += hpf[fp] * in[c].im;
    }
    out[pos] = lo;
    out[pos + pts / 2] = hi;
  }
}
static void iwave1d(const cpx *in, cpx *out, int pts, int sh, const real *slpf, const real *shpf) {
  for (int p = 0; p < pts; p++) out[p].re = out[p].im = 0;
  for (int pos = 0; pos < pts / 2; pos++)
    for (int fp = 0; fp < ORC_WN; fp++) {
      const int c = wmod(2 * pos + fp - ORC_WN + 1 + sh, pts);
      out[c].re += in[pos].re * slpf[fp];
      out[c].im += in[pos].im * slpf[fp];
      out[c].re += in[pos + pts / 2].re * shpf[fp];
      out[c].im += in[pos + pts / 2].im * shpf[fp];
    }
}

/* lines along axis a of the sub-block [0,sx) x [0,sy) x [0,sz) of X[Nz][Ny][Nx] */
static void wave_axis(cpx *X, const int32_t N[3], int a, int sx, int sy, int sz, int sh, int inv) {
  real lpf[4], hpf[4], slpf[4], shpf[4];
  d4_filters(lpf, hpf, slpf, shpf);
  const int sz3[3] = {sx, sy, sz};
  const long st[3] = {1, N[0], (long)N[0] * N[1]};
  const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
  const int pts = sz3[a];
#pragma omp parallel for
  for (int v = 0; v < sz3[c]; v++) {
    cpx *in = (cpx *)malloc(sizeof(cpx) * (size_t)pts);
    cpx *out = (cpx *)malloc(sizeof(cpx) * (size_t)pts);
    for (int u = 0; u < sz3[b]; u++) {
      cpx *line = X + (long)u * st[b] + (long)v * st[c];
      for (int p = 0; p < pts; p++) in[p] = line[(long)p * st[a]];
      if (inv) iwave1d(in, out, pts, sh, slpf, shpf);
      else wave1d(in, out, pts, sh, lpf, hpf);
      for (int p = 0; p < pts; p++) line[(long)p * st[a]] = out[p];
    }
    free(in);
    free(out);
  }
}

/* forward3D / backward3D (wavelet3D.cpp:563-631), in place; inv = 0 forward */
void orc_wavelet3d(real *Xr, const int32_t N[3], const int32_t req_levels[3], const int32_t shift[3],
                   int inv) {
  cpx *X = (cpx *)Xr;
  int32_t L[3];
  orc_wavelet_levels(N, req_levels, L);
  int maxl = 0;
  for (int a = 0; a < 3; a++) maxl = L[a] > maxl ? L[a] : maxl;
  if (!inv) {
    int s[3] = {N[0], N[1], N[2]};
    for (int l = 0; l < maxl; l++) {
      if (l < L[0]) wave_axis(X, N, 0, s[0], s[1], s[2], shift[0], 0);
      if (l < L[1]) wave_axis(X, N, 1, s[0], s[1], s[2], shift[1], 0);
      if (l < L[2]) wave_axis(X, N, 2, s[0], s[1], s[2], shift[2], 0);
      for (int a = 0; a < 3; a++)
        if (l < L[a]) s[a] = N[a] >> (l + 1);
    }
  } else {
    int s[3];
    for (int a = 0; a < 3; a++) s[a] = N[a] >> L[a];
    for (int l = maxl - 1; l >= 0; l--) {
      for (int a = 0; a < 3; a++)
        if (l < L[a]) s[a] = N[a] >> l;
      if (l < L[2]) wave_axis(X, N, 2, s[0], s[1], s[2], shift[2], 1);
      if (l < L[1]) wave_axis(X, N, 1, s[0], s[1], s[2], shift[1], 1);
      if (l < L[0]) wave_axis(X, N, 0, s[0], s[1], s[2], shift[0], 1);
    }
  }
}

static real cabs_(cpx v) { return (real)sqrt((double)v.re * v.re + (double)v.im * v.im); }

/* THRESHOLD::get_threshold for complex-grouped coefficients
 * (threshold.cpp:250-380): bisection on the value for which
 * count(|c| < value) is within 0.1% of fraction * total, at most 64 steps,
 * starting at 0.5*(max-min) with the bracket [min, fraction*max]. */
double orc_fraction_threshold(const real *Xr, long n, double fraction_d) {
  const cpx *X = (const cpx *)Xr;
  const real fraction = (real)fraction_d;
  real mx = 0, mn = cabs_(X[0]);
  for (long q = 0; q < n; q++) {
    const real v = cabs_(X[q]);
    mx = v > mx ? v : mx;
    mn = v < mn ? v : mn;
  }
  const long total = n, target = (long)(fraction * (double)total);
  const long accuracy = (long)(0.001 * (double)total);
  if (target < 2) return 0.0;
  real value = (real)0.5 * (mx - mn), max_t = fraction * mx, min_t = mn;
  long found = total;
  int iter = 0;
  while (labs(found - target) > accuracy && iter < 64) {
    if (iter > 0) {
      if (found > target) max_t = value;
      else min_t = value;
      value = (real)0.5 * (max_t + min_t);
    }
    iter++;
    found = 0;
    for (long q = 0; q < n; q++)
      if (cabs_(X[q]) < value) found++;
  }
  return value;
}

/* THRESHOLD::thresholding, complex grouped (threshold.cpp:766-800) */
void orc_soft_threshold(real *Xr, long n, double value_d) {
  cpx *X = (cpx *)Xr;
  const real value = (real)value_d;
  for (long q = 0; q < n; q++) {
    const real cc = cabs_(X[q]);
    if (cc <= value) {
      X[q].re = X[q].im = 0;
    } else {
      const real f = (real)1 - value / cc;
      X[q].re *= f;
      X[q].im *= f;
    }
  }
}

/* RECON::L1_threshold with a wavelet transform, no phase rotation and no
 * temporal transform (recon_lib.cpp:2416-2527): per cycle spin, forward
 * wavelet of every frame, the threshold taken once (first spin) over all
 * frames with scale 1/spins, soft threshold, backward wavelet.
 * shifts[spin][3] are the per-axis wavelet shifts. Returns the threshold. */
double orc_l1_wavelet(real *X, int frames, const int32_t N[3], const int32_t levels[3], double fraction,
                      int spins, const int32_t *shifts) {
  const long nv = (long)N[0] * N[1] * N[2];
  real thr = 0;
  for (int sp = 0; sp < spins; sp++) {
    for (int f = 0; f < frames; f++) orc_wavelet3d(X + 2 * nv * f, N, levels, shifts + 3 * sp, 0);
    if (sp == 0) thr = ((real)1 / (real)spins) * (real)orc_fraction_threshold(X, nv * frames, fraction);
    orc_soft_threshold(X, nv * frames, thr);
    for (int f = 0; f < frames; f++) orc_wavelet3d(X + 2 * nv * f, N, levels, shifts + 3 * sp, 1);
  }
  return thr;
}

/* Cycle-spin shift stream shared with the HIP library: splitmix64 from `seed`,
 * shift = (next >> 33) % 4 per axis, three per spin (the reference draws
 * rand() % wN, wavelet3D.cpp:443-449). */
void orc_spin_shifts(uint64_t seed, long count, int32_t *out) {
  uint64_t st = seed;
  for (long q = 0; q < count; q++) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    out[q] = (int32_t)((z >> 33) % ORC_WN);
  }
}

/* FISTA with the L1 wavelet proximal step after each gradient step
 * (recon_lib.cpp:1925-2150): shifts of iteration `it` are the 3*spins values
 * starting at shifts + 3*spins*it. */
void orc_fista_l1(orc_op *op, real *Xo, const real *smaps, int C, const real *data, const float *kx,
                  const float *ky, const float *kz, const float *kw, long M, int iters, double step,
                  double fraction, const int32_t levels[3], int spins, const int32_t *shifts,
                  double *resid, double *thresh) {
  long nv = nvox(op);
  cpx *X = (cpx *)Xo;
  cpx *Xold = (cpx *)calloc((size_t)nv, sizeof(cpx));
  cpx *R = (cpx *)malloc(sizeof(cpx) * (size_t)nv);
  real *diff = (real *)malloc(sizeof(cpx) * (size_t)M);
  real st = (real)step;
  memset(X, 0, sizeof(cpx) * (size_t)nv);
  for (int it = 0; it < iters; it++) {
    memset(R, 0, sizeof(cpx) * (size_t)nv);
    real A = (real)(1.0 + (it - 1.0) / (it + 1.0));
    real B = (real)(-(it - 1.0) / (it + 1.0));
    for (long q = 0; q < nv; q++) {
      cpx x0 = Xold[q], x1 = X[q];
      X[q].re = A * x1.re + B * x0.re;
      X[q].im = A * x1.im + B * x0.im;
      Xold[q] = x1;
    }
    for (int c = 0; c < C; c++) {
      orc_backward(op, (const real *)X, smaps ? smaps + 2 * nv * c : NULL, diff, kx, ky, kz, kw,
                   M, data + 2 * M * c);
      orc_forward(op, (real *)R, smaps ? smaps + 2 * nv * c : NULL, diff, kx, ky, kz, kw, M);
    }
    if (resid) resid[it] = energy(R, nv);
    for (long q = 0; q < nv; q++) {
      real rr = R[q].re * st, ri = R[q].im * st;
      X[q].re -= rr;
      X[q].im -= ri;
    }
    const double t = orc_l1_wavelet((real *)X, 1, op->N, levels, fraction, spins, shifts + 3 * spins * it);
    if (thresh) thresh[it] = t;
  }
  free(Xold);
  free(R);
  free(diff);
}
