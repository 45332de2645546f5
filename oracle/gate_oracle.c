/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of uwmri/mri_recon's gated frame selection: the gate-time
 * scaling of GATING::init_time_resolved, the per-frame sample weights of
 * GATING::weight_data / hist_weight / tornado_weight, and RECON::pregate_data,
 * which keeps the samples of each frame and rescales their kw.  The checker of
 * mri_recon_amd's mrn_gate_scale_times / mrn_pregate; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Reference files followed (paths relative to the reference root):
 *   src/gating.h:27-47          enums (view sharing, tornado shape, gate type)
 *   src/gating.cpp:84-112       defaults (tornado widths 1 and 4, VIPR shape)
 *   src/gating.cpp:880-1078     init_time_resolved: time range (min/max;
 *                               2 x median for retrospective ECG with outlier
 *                               rejection; quantiles for respiratory), scaling
 *                               to frames, tornado frame centres and kmax,
 *                               floor for no sharing, rank bins for hist mode
 *   src/gating.cpp:1082-1200    weight_data, hist_weight, tornado_weight
 *   src/recon_lib.cpp:675-776   pregate_data: selection (weight > 0), frame-
 *                               major layout, kw * 1e6 / count (or 1)
 *
 * Gate times are one double per readout; a readout holds xres samples, the
 * trajectory and k-space are [readout][xres] (sample fastest).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { VS_TORNADO = 0, VS_NONE = 1, VS_HIST = 2 };
enum { T_FLAT = 0, T_RADIAL = 1, T_VIPR = 2 };
enum { G_NONE = 0, G_RETRO_ECG = 1, G_ECG = 2, G_RESP = 3, G_TIME = 4, G_PREP = 5 };

typedef struct {
  int32_t gate_type;
  int32_t vs_type;
  int32_t frames;
  int32_t tornado_shape;
  int32_t wdth_low, wdth_high;
  float kmax;
  float min_quantile, max_quantile;
  int32_t ecg_outlier_reject;
  int32_t reset_dens;
} orc_gate_params;

static int cmp_d(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return x < y ? -1 : x > y;
}

/* Median with the mean of the two middle values for an even count. */
static double median_of(double *v, int64_t n) {
  if (n <= 0) return 0.0;
  qsort(v, (size_t)n, sizeof(double), cmp_d);
  return (n & 1) ? v[n / 2] : v[n / 2 - 1] + (v[n / 2] - v[n / 2 - 1]) / 2;
}

/* Quantile by Hyndman & Fan's method 5 (the piecewise-linear rule Armadillo
 * documents for quantile()) on sorted v. */
double orc_quantile5(const double *sorted, int64_t n, double p) {
  if (n <= 0) return 0.0;
  const double h = (double)n * p + 0.5; /* 1-based position */
  if (h <= 1.0) return sorted[0];
  if (h >= (double)n) return sorted[n - 1];
  const int64_t lo = (int64_t)floor(h);
  return sorted[lo - 1] + (h - (double)lo) * (sorted[lo] - sorted[lo - 1]);
}

/* init_time_resolved after the gate signal is chosen: times (R, in place)
 * become frame coordinates; info = {offset, scale, temporal resolution}.
 * Returns 0, or -1 for an unknown setting. */
int orc_gate_scale(const orc_gate_params *p, int64_t R, double *t, double *info) {
  double mx = 0, mn = 0;
  double *tmp = (double *)malloc(sizeof(double) * (size_t)(R > 0 ? R : 1));
  memcpy(tmp, t, sizeof(double) * (size_t)R);
  if (p->gate_type == G_RETRO_ECG) {
    mn = 0;
    mx = 2.0 * median_of(tmp, R);
    if (p->ecg_outlier_reject) {
      memcpy(tmp, t, sizeof(double) * (size_t)R);
      int64_t m = 0;
      for (int64_t i = 0; i < R; i++)
        if (t[i] < 1.3 * mx) tmp[m++] = t[i];
      mx = 2.0 * median_of(tmp, m);
    }
  } else if (p->gate_type == G_RESP) {
    qsort(tmp, (size_t)R, sizeof(double), cmp_d);
    const double v0 = orc_quantile5(tmp, R, p->min_quantile), v1 = orc_quantile5(tmp, R, p->max_quantile);
    const double fw = (v1 - v0) / (p->frames - 1.0);
    mx = v1 + 0.5 * fw;
    mn = v0 - 0.5 * fw;
  } else {
    mx = -INFINITY;
    mn = INFINITY;
    for (int64_t i = 0; i < R; i++) {
      if (t[i] > mx) mx = t[i];
      if (t[i] < mn) mn = t[i];
    }
  }
  const double scale = (float)p->frames / (mx - mn) * (1 - 1e-9);
  const double offset = mn;
  for (int64_t i = 0; i < R; i++) {
    t[i] -= offset;
    t[i] *= scale;
  }
  if (info) {
    info[0] = offset;
    info[1] = scale;
    info[2] = (mx - mn) / ((float)p->frames);
  }
  if (p->vs_type == VS_NONE) {
    for (int64_t i = 0; i < R; i++) t[i] = floor(t[i]);
  } else if (p->vs_type == VS_HIST) {
    /* frame = rank * frames / R, rank = position in ascending order */
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(R > 0 ? R : 1));
    for (int64_t i = 0; i < R; i++) idx[i] = i;
    /* indices in ascending time, equal times in index order */
    for (int64_t w = 1; w < R; w *= 2) { /* bottom-up merge sort, stable */
      int64_t *buf = (int64_t *)malloc(sizeof(int64_t) * (size_t)R);
      for (int64_t lo = 0; lo < R; lo += 2 * w) {
        int64_t mid = lo + w < R ? lo + w : R, hi = lo + 2 * w < R ? lo + 2 * w : R, a = lo, b = mid, o = lo;
        while (a < mid && b < hi) buf[o++] = (t[idx[b]] < t[idx[a]]) ? idx[b++] : idx[a++];
        while (a < mid) buf[o++] = idx[a++];
        while (b < hi) buf[o++] = idx[b++];
      }
      memcpy(idx, buf, sizeof(int64_t) * (size_t)R);
      free(buf);
    }
    double *fr = (double *)malloc(sizeof(double) * (size_t)(R > 0 ? R : 1));
    for (int64_t r = 0; r < R; r++) fr[idx[r]] = (double)(int)((double)(r * p->frames) / (double)R);
    memcpy(t, fr, sizeof(double) * (size_t)R);
    free(fr);
    free(idx);
  } else if (p->vs_type != VS_TORNADO) {
    free(tmp);
    return -1;
  }
  free(tmp);
  return 0;
}

/* weight_data(..., NON_ITERATIVE, TIME_FRAME) of one sample for frame t: the
 * respiratory weight of its readout (if any) times the view-sharing weight. */
static float sample_weight(const orc_gate_params *p, double gt, const float *resp, int64_t j, float kx, float ky,
                           float kz, int t) {
  float w = 1.0f;
  if (resp) w *= resp[j];
  if (p->vs_type == VS_TORNADO) {
    const double t_diff = fabs(gt - (0.5 + (double)t));
    float kr = 0, kp = 0;
    if (p->tornado_shape == T_RADIAL) {
      kr = sqrtf(kx * kx + ky * ky);
      kp = 1.0f;
    } else if (p->tornado_shape == T_VIPR) {
      kr = sqrtf(kx * kx + ky * ky + kz * kz);
      kp = 2.0f;
    }
    const double wdth = 0.5 * ((p->wdth_high - p->wdth_low) * powf(kr / p->kmax, kp) + p->wdth_low);
    w = (float)(w * ((t_diff < wdth) ? (1. / wdth) : (0.0)));
  } else {
    w = (float)(w * ((floor(gt) == t) ? (1.0) : (0.0)));
  }
  return w;
}

/* pregate_data for one encoding: counts[t] samples per frame; when kx_out is
 * given, frame t's samples (weight > 0, in sample order) go to
 * [base_t, base_t + counts[t]) of kx/ky/kz/kw_out (base_t = sum of earlier
 * counts) and its k-space to kd_out + 2*C*base_t as [C][counts[t]]. */
void orc_pregate(const orc_gate_params *p, int64_t xres, int64_t R, const double *gt, const float *resp,
                 const float *kx, const float *ky, const float *kz, const float *kw, int C, const float *kd,
                 int64_t *counts, float *kx_out, float *ky_out, float *kz_out, float *kw_out, float *kd_out) {
  const int64_t M = xres * R;
  int64_t base = 0;
  for (int t = 0; t < p->frames; t++) {
    int64_t n = 0;
    for (int64_t i = 0; i < M; i++)
      if (sample_weight(p, gt[i / xres], resp, i / xres, kx[i], ky[i], kz[i], t) > 0) n++;
    counts[t] = n;
    if (kx_out) {
      const float kw_scale = (float)(1e6 / ((float)n));
      int64_t o = 0;
      for (int64_t i = 0; i < M; i++) {
        if (!(sample_weight(p, gt[i / xres], resp, i / xres, kx[i], ky[i], kz[i], t) > 0)) continue;
        kx_out[base + o] = kx[i];
        ky_out[base + o] = ky[i];
        kz_out[base + o] = kz[i];
        kw_out[base + o] = p->reset_dens ? 1.0f : kw_scale * kw[i];
        for (int c = 0; c < C; c++) {
          kd_out[2 * (C * base + (int64_t)c * n + o)] = kd[2 * ((int64_t)c * M + i)];
          kd_out[2 * (C * base + (int64_t)c * n + o) + 1] = kd[2 * ((int64_t)c * M + i) + 1];
        }
        o++;
      }
    }
    base += n;
  }
}
