/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of uwmri/mri_recon's locally low-rank step over frames
 * (LOWRANKCOIL::update_threshold and ::thresh for an image indexed (t, e)),
 * the checker of mri_recon_amd's mrn_llr.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Reference files followed (paths relative to the reference root):
 *   src/clear.cpp:40-96     defaults and -clear_* flags (blocks 4x4x4,
 *                           block_iter 4, clear_amp 4, clear_t2 6)
 *   src/clear.cpp:233-350   update_threshold: unshifted blocks, largest
 *                           singular value of each, median of the nonzero ones
 *                           times 1 + amp*exp(-iter/t2)
 *   src/clear.cpp:356-556   thresh: block_iter passes, each with one random
 *                           block shift (periodic wrap), singular values
 *                           soft-thresholded by alpha*smax/block_iter
 *   src/recon_lib.cpp:2117-2126  the calls after each gradient step
 *
 * Block matrices are Np x N in double complex, Np = (frames of the other
 * index) * block voxels, N = frames of index `dim` (0 = t, 1 = e), as the
 * reference's Armadillo matrices.  Their SVD here is one-sided (Hestenes)
 * Jacobi in double: the reference calls LAPACK through Armadillo, which is
 * absent from this image; tests pin this routine against numpy's LAPACK SVD.
 * Random shifts: the reference draws rand() % b - b/2, which depends on every
 * other rand() call of its program; here shift q of pass `pass` of call
 * `call` is splitmix64(seed, 3 * (call * block_iter + pass) + q) % b - b/2.
 *
 * Image layout: X[(t * Ne + e) * nvox + (z * Ny + y) * Nx + x], complex
 * (re, im) in `real` (float in liboracle_f32, double in liboracle_f64).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;
typedef double complex dcx;

typedef struct {
  int32_t block[3];
  int32_t block_iter;
  float alpha_time;
  float alpha_encode;
  float amp;
  float t2;
  uint64_t seed;
} orc_llr_params;

static uint64_t splitmix(uint64_t seed, uint64_t index) {
  uint64_t z = seed + (index + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Block shifts of pass `pass` of thresh call `call` for the (clamped) block
 * sizes bs: 0 when block_iter == 1 (clear.cpp:412-416). */
void orc_llr_shift(const orc_llr_params *p, const int32_t bs[3], int64_t call, int pass, int32_t shift[3]) {
  for (int q = 0; q < 3; q++) {
    const int b = bs[q];
    shift[q] = p->block_iter == 1 ? 0
               : (int32_t)(splitmix(p->seed, 3ull * (uint64_t)(call * p->block_iter + pass) + q) % (uint64_t)b) - b / 2;
  }
}

/* One-sided Jacobi SVD of the Np x N column-major matrix A (overwritten with
 * A V, whose column norms are the singular values); V is N x N column-major.
 * Returns the sweeps used. */
int orc_jacobi_svd(dcx *A, int Np, int N, dcx *V) {
  for (int i = 0; i < N * N; i++) V[i] = 0;
  for (int i = 0; i < N; i++) V[i * N + i] = 1;
  int sweep;
  for (sweep = 0; sweep < 60; sweep++) {
    double off = 0, on = 0;
    for (int p = 0; p < N - 1; p++)
      for (int q = p + 1; q < N; q++) {
        dcx *ap = A + (int64_t)p * Np, *aq = A + (int64_t)q * Np;
        double al = 0, be = 0;
        dcx ga = 0;
        for (int i = 0; i < Np; i++) {
          al += creal(ap[i]) * creal(ap[i]) + cimag(ap[i]) * cimag(ap[i]);
          be += creal(aq[i]) * creal(aq[i]) + cimag(aq[i]) * cimag(aq[i]);
          ga += conj(ap[i]) * aq[i];
        }
        const double g = cabs(ga);
        on += al * be;
        off += g * g;
        if (g == 0.0 || g <= 1e-17 * sqrt(al * be)) continue;
        /* rotation J on columns p, q with (J^H G J)_pq = 0, G = A^H A */
        const double zeta = (be - al) / (2 * g);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
        const double c = 1 / sqrt(1 + t * t), s = c * t;
        const dcx e = ga / g;
        for (int i = 0; i < Np; i++) {
          const dcx x = ap[i], y = aq[i];
          ap[i] = c * x - s * conj(e) * y;
          aq[i] = s * e * x + c * y;
        }
        dcx *vp = V + (int64_t)p * N, *vq = V + (int64_t)q * N;
        for (int i = 0; i < N; i++) {
          const dcx x = vp[i], y = vq[i];
          vp[i] = c * x - s * conj(e) * y;
          vq[i] = s * e * x + c * y;
        }
      }
    if (off <= 1e-30 * (on > 0 ? on : 1)) break;
  }
  return sweep;
}

/* Singular values of an Np x N column-major complex double matrix (given as
 * interleaved re/im), descending. */
void orc_svd_values(const double *Ain, int Np, int N, double *s) {
  dcx *A = (dcx *)malloc(sizeof(dcx) * (size_t)Np * N);
  dcx *V = (dcx *)malloc(sizeof(dcx) * (size_t)N * N);
  for (int64_t i = 0; i < (int64_t)Np * N; i++) A[i] = Ain[2 * i] + I * Ain[2 * i + 1];
  orc_jacobi_svd(A, Np, N, V);
  for (int k = 0; k < N; k++) {
    double n2 = 0;
    for (int i = 0; i < Np; i++) n2 += creal(A[(int64_t)k * Np + i]) * creal(A[(int64_t)k * Np + i]) +
                                        cimag(A[(int64_t)k * Np + i]) * cimag(A[(int64_t)k * Np + i]);
    s[k] = sqrt(n2);
  }
  for (int a = 0; a < N; a++)  /* sort descending */
    for (int b = a + 1; b < N; b++)
      if (s[b] > s[a]) {
        double t = s[a];
        s[a] = s[b];
        s[b] = t;
      }
  free(A);
  free(V);
}

/* Soft threshold of the singular values of the Np x N block matrix A (in
 * place): A <- U diag(max(s - tau, 0)) V^H.  Returns the largest singular
 * value. */
static double svt(dcx *A, int Np, int N, double tau, dcx *V, double *s, int apply) {
  orc_jacobi_svd(A, Np, N, V);
  double smax = 0;
  for (int k = 0; k < N; k++) {
    double n2 = 0;
    for (int i = 0; i < Np; i++) {
      const dcx a = A[(int64_t)k * Np + i];
      n2 += creal(a) * creal(a) + cimag(a) * cimag(a);
    }
    s[k] = sqrt(n2);
    if (s[k] > smax) smax = s[k];
  }
  if (!apply) return smax;
  /* A V = U diag(s): scale column k by max(s - tau, 0) / s, then times V^H */
  dcx *W = (dcx *)malloc(sizeof(dcx) * (size_t)Np * N);
  for (int k = 0; k < N; k++) {
    const double f = s[k] > 0 ? fmax(s[k] - tau, 0.0) / s[k] : 0.0;
    for (int i = 0; i < Np; i++) W[(int64_t)k * Np + i] = A[(int64_t)k * Np + i] * f;
  }
  for (int j = 0; j < N; j++)
    for (int i = 0; i < Np; i++) {
      dcx acc = 0;
      for (int k = 0; k < N; k++) acc += W[(int64_t)k * Np + i] * conj(V[(int64_t)k * N + j]);
      A[(int64_t)j * Np + i] = acc;
    }
  free(W);
  return smax;
}

/* Gather (or scatter back) block (bi, bj, bk) with shift into the Np x N
 * matrix: column c = frame index along dim, row = (other frame index, z, y, x)
 * with x fastest (clear.cpp:417-461, 483-529). */
static void block_io(real *X, int Nt, int Ne, const int32_t Nv[3], const int32_t bs[3], int dim, int i0, int j0,
                     int k0, const int32_t sh[3], dcx *A, int Np, int back) {
  const int64_t nvox = (int64_t)Nv[0] * Nv[1] * Nv[2];
  const int N = dim == 0 ? Nt : Ne, Nother = dim == 0 ? Ne : Nt;
  for (int col = 0; col < N; col++) {
    int row = 0;
    for (int o = 0; o < Nother; o++) {
      const int t = dim == 0 ? col : o, e = dim == 0 ? o : col;
      real *F = X + 2 * ((int64_t)t * Ne + e) * nvox;
      for (int kk = k0 + sh[2]; kk < k0 + bs[2] + sh[2]; kk++)
        for (int jj = j0 + sh[1]; jj < j0 + bs[1] + sh[1]; jj++)
          for (int ii = i0 + sh[0]; ii < i0 + bs[0] + sh[0]; ii++) {
            const int px = (ii + Nv[0]) % Nv[0], py = (jj + Nv[1]) % Nv[1], pz = (kk + Nv[2]) % Nv[2];
            real *v = F + 2 * (((int64_t)pz * Nv[1] + py) * Nv[0] + px);
            if (back) {
              v[0] = (real)creal(A[(int64_t)col * Np + row]);
              v[1] = (real)cimag(A[(int64_t)col * Np + row]);
            } else {
              A[(int64_t)col * Np + row] = (double)v[0] + I * (double)v[1];
            }
            row++;
          }
    }
  }
  (void)Np;
}

static void clamp_blocks(const int32_t Nv[3], const int32_t b_in[3], int32_t bs[3], int nb[3]) {
  for (int a = 0; a < 3; a++) {
    bs[a] = b_in[a] > Nv[a] ? Nv[a] : b_in[a];
    nb[a] = Nv[a] / bs[a];
  }
}

/* Largest singular value of every unshifted block, in the reference's block
 * order (i outer, k inner: clear.cpp:269-281).  Returns the block count. */
int64_t orc_llr_block_smax(real *X, int Nt, int Ne, const int32_t Nv[3], const int32_t b_in[3], int dim,
                           double *out) {
  int32_t bs[3];
  int nb[3];
  clamp_blocks(Nv, b_in, bs, nb);
  const int N = dim == 0 ? Nt : Ne, Np = (dim == 0 ? Ne : Nt) * bs[0] * bs[1] * bs[2];
  const int64_t nblk = (int64_t)nb[0] * nb[1] * nb[2];
  const int32_t zero[3] = {0, 0, 0};
#pragma omp parallel
  {
    dcx *A = (dcx *)malloc(sizeof(dcx) * (size_t)Np * N);
    dcx *V = (dcx *)malloc(sizeof(dcx) * (size_t)N * N);
    double *s = (double *)malloc(sizeof(double) * N);
#pragma omp for schedule(dynamic, 16)
    for (int64_t b = 0; b < nblk; b++) {
      const int i = (int)(b / ((int64_t)nb[1] * nb[2])), j = (int)((b / nb[2]) % nb[1]), k = (int)(b % nb[2]);
      block_io(X, Nt, Ne, Nv, bs, dim, i * bs[0], j * bs[1], k * bs[2], zero, A, Np, 0);
      out[b] = svt(A, Np, N, 0.0, V, s, 0);
    }
    free(A);
    free(V);
    free(s);
  }
  return nblk;
}

static int cmp_d(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return x < y ? -1 : x > y;
}

/* Median of the nonzero entries (mean of the two middle ones for an even
 * count); 0 when there are none. */
double orc_median_nonzero(const double *v, int64_t n) {
  double *w = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  int64_t m = 0;
  for (int64_t i = 0; i < n; i++)
    if (v[i] != 0.0) w[m++] = v[i];
  double r = 0;
  if (m > 0) {
    qsort(w, (size_t)m, sizeof(double), cmp_d);
    r = (m & 1) ? w[m / 2] : w[m / 2 - 1] + (w[m / 2] - w[m / 2 - 1]) / 2;
  }
  free(w);
  return r;
}

/* update_threshold: smax = (float)(median * (float)(1 + amp * exp(-iter / t2))). */
float orc_llr_update_threshold(const orc_llr_params *p, real *X, int Nt, int Ne, const int32_t Nv[3], int dim,
                               int iter) {
  int32_t bs[3];
  int nb[3];
  clamp_blocks(Nv, p->block, bs, nb);
  const int64_t nblk = (int64_t)nb[0] * nb[1] * nb[2];
  double *sb = (double *)malloc(sizeof(double) * (size_t)(nblk > 0 ? nblk : 1));
  orc_llr_block_smax(X, Nt, Ne, Nv, p->block, dim, sb);
  const float amp_scale = (float)(1 + p->amp * expf(-((float)iter) / p->t2));
  const float smax = (float)(orc_median_nonzero(sb, nblk) * amp_scale);
  free(sb);
  return smax;
}

/* One thresh pass with the given shift and tau (the block loop of
 * clear.cpp:418-552). */
void orc_llr_pass(real *X, int Nt, int Ne, const int32_t Nv[3], const int32_t b_in[3], int dim,
                  const int32_t sh[3], double tau) {
  int32_t bs[3];
  int nb[3];
  clamp_blocks(Nv, b_in, bs, nb);
  const int N = dim == 0 ? Nt : Ne, Np = (dim == 0 ? Ne : Nt) * bs[0] * bs[1] * bs[2];
  const int64_t nblk = (int64_t)nb[0] * nb[1] * nb[2];
#pragma omp parallel
  {
    dcx *A = (dcx *)malloc(sizeof(dcx) * (size_t)Np * N);
    dcx *V = (dcx *)malloc(sizeof(dcx) * (size_t)N * N);
    double *s = (double *)malloc(sizeof(double) * N);
#pragma omp for schedule(dynamic, 16)
    for (int64_t b = 0; b < nblk; b++) {
      const int i = (int)(b / ((int64_t)nb[1] * nb[2])), j = (int)((b / nb[2]) % nb[1]), k = (int)(b % nb[2]);
      block_io(X, Nt, Ne, Nv, bs, dim, i * bs[0], j * bs[1], k * bs[2], sh, A, Np, 0);
      svt(A, Np, N, tau, V, s, 1);
      block_io(X, Nt, Ne, Nv, bs, dim, i * bs[0], j * bs[1], k * bs[2], sh, A, Np, 1);
    }
    free(A);
    free(V);
    free(s);
  }
}

/* update_threshold then thresh for `dim` (recon_lib.cpp:2117-2126); `call`
 * numbers the thresh calls for the shift stream.  Returns smax. */
float orc_llr(const orc_llr_params *p, real *X, int Nt, int Ne, const int32_t Nv[3], int dim, int iter,
              int64_t call) {
  const float alpha = dim == 0 ? p->alpha_time : p->alpha_encode;
  const float smax = orc_llr_update_threshold(p, X, Nt, Ne, Nv, dim, iter);
  double clear_alpha = alpha;
  if (clear_alpha == 0.0) return smax;
  clear_alpha *= smax;
  const double tau = clear_alpha / p->block_iter;
  int32_t bs[3];
  int nb[3];
  clamp_blocks(Nv, p->block, bs, nb);
  for (int pass = 0; pass < p->block_iter; pass++) {
    int32_t sh[3];
    orc_llr_shift(p, bs, call, pass, sh);
    orc_llr_pass(X, Nt, Ne, Nv, p->block, dim, sh, tau);
  }
  return smax;
}
