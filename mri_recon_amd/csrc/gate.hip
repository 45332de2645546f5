// Gated frame selection: GATING::init_time_resolved's scaling of gate times
// (gating.cpp:880-1078, on the host: one value per readout), and on the
// device GATING::weight_data (gating.cpp:1082-1200) with RECON::pregate_data
// (recon_lib.cpp:675-776): for every frame t, keep the samples whose weight is
// positive, in sample order, frame after frame, with kw * 1e6 / count (or 1).
//
// Device side: a stream compaction, HBM-bound.  Pass 1 counts the selected
// samples of each (frame, 256-sample block); one exclusive scan over
// [frame][block] gives every block its output offset in frame-major order;
// pass 2 recomputes the selection and writes each kept sample at its offset
// plus its rank in the block (wavefront ballots), so the order is the input
// order.  Trajectory and k-space are read once per frame.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "mrn_internal.hpp"

namespace mrn {
namespace {

struct GateDev {
  int vs_type, shape, frames, reset;
  int wlo, whi;
  float kmax;
  long long xres, M, nb;  // samples per readout, samples, 256-sample blocks
};

// weight_data(NON_ITERATIVE, TIME_FRAME) > 0 for sample i in frame t
__device__ __forceinline__ bool selected(const GateDev &g, const double *__restrict__ gt,
                                         const float *__restrict__ resp, const float *__restrict__ kx,
                                         const float *__restrict__ ky, const float *__restrict__ kz, long long i,
                                         int t) {
  const long long j = i / g.xres;
  float w = 1.0f;
  if (resp) w *= resp[j];
  const double gtj = gt[j];
  if (g.vs_type == MRN_VS_TORNADO) {
    const double t_diff = fabs(gtj - (0.5 + (double)t));
    float kr = 0.0f, pw = 1.0f;
    if (g.shape == MRN_TORNADO_RADIAL) {
      kr = __fsqrt_rn(__fadd_rn(__fmul_rn(kx[i], kx[i]), __fmul_rn(ky[i], ky[i])));
      pw = __fdiv_rn(kr, g.kmax);  // (kr / kmax)^1
    } else if (g.shape == MRN_TORNADO_VIPR) {
      kr = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(kx[i], kx[i]), __fmul_rn(ky[i], ky[i])),
                                __fmul_rn(kz[i], kz[i])));
      const float q = __fdiv_rn(kr, g.kmax);
      pw = __fmul_rn(q, q);  // (kr / kmax)^2
    }
    const double wdth = 0.5 * (double)__fadd_rn(__fmul_rn((float)(g.whi - g.wlo), pw), (float)g.wlo);
    w = (float)((double)w * ((t_diff < wdth) ? (1.0 / wdth) : 0.0));
  } else {
    w = (floor(gtj) == (double)t) ? w : 0.0f;
  }
  return w > 0.0f;
}

__global__ __launch_bounds__(256) void k_gate_count(GateDev g, const double *__restrict__ gt,
                                                    const float *__restrict__ resp, const float *__restrict__ kx,
                                                    const float *__restrict__ ky, const float *__restrict__ kz,
                                                    int64_t *__restrict__ bcnt) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const int t = blockIdx.y;
  const bool s = i < g.M && selected(g, gt, resp, kx, ky, kz, i, t);
  const int n = __syncthreads_count(s);
  if (threadIdx.x == 0) bcnt[(long long)t * g.nb + blockIdx.x] = n;
}

// frame bases and counts from the scanned block offsets: base[t] = boff[t*nb]
__global__ void k_gate_frames(GateDev g, const int64_t *__restrict__ boff, int64_t *__restrict__ fb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= g.frames) fb[t] = boff[(long long)t * g.nb];
}

__global__ __launch_bounds__(256) void k_gate_scatter(GateDev g, const double *__restrict__ gt,
                                                      const float *__restrict__ resp, const float *__restrict__ kx,
                                                      const float *__restrict__ ky, const float *__restrict__ kz,
                                                      const float *__restrict__ kw, int C,
                                                      const float2 *__restrict__ kd, const int64_t *__restrict__ boff,
                                                      const int64_t *__restrict__ fb, float *__restrict__ ox,
                                                      float *__restrict__ oy, float *__restrict__ oz,
                                                      float *__restrict__ ow, float2 *__restrict__ okd) {
  __shared__ int wcount[4];
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const int t = blockIdx.y;
  const bool s = i < g.M && selected(g, gt, resp, kx, ky, kz, i, t);
  const unsigned long long m = __ballot(s);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wcount[wave] = __popcll(m);
  __syncthreads();
  if (!s) return;
  int before = __popcll(m & ((1ull << lane) - 1ull));
  for (int w = 0; w < wave; w++) before += wcount[w];
  const int64_t pos = boff[(long long)t * g.nb + blockIdx.x] + before;
  const int64_t base = fb[t], n = fb[t + 1] - fb[t], o = pos - base;
  ox[pos] = kx[i];
  oy[pos] = ky[i];
  oz[pos] = kz[i];
  const float scale = (float)(1e6 / (double)(float)n);
  ow[pos] = g.reset ? 1.0f : __fmul_rn(scale, kw[i]);
  for (int c = 0; c < C; c++) okd[base * C + (int64_t)c * n + o] = kd[(int64_t)c * g.M + i];
}

double median_sorted(const std::vector<double> &v) {
  const size_t n = v.size();
  if (n == 0) return 0.0;
  return (n & 1) ? v[n / 2] : v[n / 2 - 1] + (v[n / 2] - v[n / 2 - 1]) / 2;
}

// Hyndman & Fan method 5 on sorted values (Armadillo's documented quantile)
double quantile5(const std::vector<double> &v, double p) {
  const double n = (double)v.size();
  if (v.empty()) return 0.0;
  const double h = n * p + 0.5;
  if (h <= 1.0) return v.front();
  if (h >= n) return v.back();
  const size_t lo = (size_t)std::floor(h);
  return v[lo - 1] + (h - (double)lo) * (v[lo] - v[lo - 1]);
}

}  // namespace
}  // namespace mrn

using namespace mrn;

extern "C" {

void mrn_gate_params_default(mrn_gate_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->gate_type = MRN_GATE_NONE;
  p->vs_type = MRN_VS_NONE;
  p->frames = 1;
  p->tornado_shape = MRN_TORNADO_VIPR;
  p->wdth_low = 1;
  p->wdth_high = 4;
  p->kmax = 128.0f;
  p->min_quantile = 0.0f;
  p->max_quantile = 1.0f;
  p->ecg_outlier_reject = 1;
}

int mrn_gate_parse_args(mrn_gate_params *p, int argc, const char **argv) {
  if (!p || (argc > 0 && !argv)) return MRN_EINVAL;
  for (int i = 0; i < argc; i++) {
    const char *s = argv[i];
    const bool has = i + 1 < argc;
    auto word = [&](const char *w) { return has && !strcmp(argv[i + 1], w); };
    if (!strcmp(s, "-viewshare_type")) {
      if (word("tornado")) p->vs_type = MRN_VS_TORNADO;
      else if (word("hist")) p->vs_type = MRN_VS_HIST;
      else if (word("none")) p->vs_type = MRN_VS_NONE;
      else return MRN_EINVAL;
      i++;
    } else if (!strcmp(s, "-gating_type")) {
      if (word("retro_ecg")) p->gate_type = MRN_GATE_RETRO_ECG;
      else if (word("resp")) p->gate_type = MRN_GATE_RESP;
      else if (word("ecg")) p->gate_type = MRN_GATE_ECG;
      else if (word("time")) p->gate_type = MRN_GATE_TIME;
      else if (word("prep")) p->gate_type = MRN_GATE_PREP;
      else return MRN_EINVAL;
      i++;
    } else if (!strcmp(s, "-vs_wdth_low") || !strcmp(s, "-vs_wdth_high") || !strcmp(s, "-rcframes")) {
      if (!has) return MRN_EINVAL;
      const int v = atoi(argv[++i]);
      if (!strcmp(s, "-vs_wdth_low")) p->wdth_low = v;
      else if (!strcmp(s, "-vs_wdth_high")) p->wdth_high = v;
      else p->frames = v;
    } else if (!strcmp(s, "-gate_min_quantile") || !strcmp(s, "-gate_max_quantile")) {
      if (!has) return MRN_EINVAL;
      const float v = (float)atof(argv[++i]);
      if (!strcmp(s, "-gate_min_quantile")) p->min_quantile = v;
      else p->max_quantile = v;
    } else if (!strcmp(s, "-ecg_out_reject")) {
      p->ecg_outlier_reject = 1;
    } else if (!strcmp(s, "-wo_ecg_out_reject")) {
      p->ecg_outlier_reject = 0;
    }
  }
  return MRN_OK;
}

int mrn_gate_scale_times(const mrn_gate_params *p, int64_t R, double *t, double info[3]) {
  free_err_clear();
  if (!p || R < 1 || !t) return free_fail(MRN_EINVAL, "gate times");
  if (p->frames < 1) return free_fail(MRN_EINVAL, "frames < 1");
  if (p->vs_type != MRN_VS_TORNADO && p->vs_type != MRN_VS_NONE && p->vs_type != MRN_VS_HIST)
    return free_fail(MRN_EINVAL, "view sharing type");
  std::vector<double> v(t, t + R);
  double mx, mn;
  if (p->gate_type == MRN_GATE_RETRO_ECG) {
    std::sort(v.begin(), v.end());
    mn = 0.0;
    mx = 2.0 * median_sorted(v);
    if (p->ecg_outlier_reject) {
      std::vector<double> w;
      for (int64_t i = 0; i < R; i++)
        if (t[i] < 1.3 * mx) w.push_back(t[i]);
      std::sort(w.begin(), w.end());
      mx = 2.0 * median_sorted(w);
    }
  } else if (p->gate_type == MRN_GATE_RESP) {
    if (p->frames < 2) return free_fail(MRN_EINVAL, "respiratory gating needs 2 or more frames");
    std::sort(v.begin(), v.end());
    const double v0 = quantile5(v, p->min_quantile), v1 = quantile5(v, p->max_quantile);
    const double fw = (v1 - v0) / (p->frames - 1.0);
    mx = v1 + 0.5 * fw;
    mn = v0 - 0.5 * fw;
  } else {
    const auto mm = std::minmax_element(v.begin(), v.end());
    mn = *mm.first;
    mx = *mm.second;
  }
  if (!(mx > mn)) return free_fail(MRN_EINVAL, "gate times span no range");
  const double scale = (float)p->frames / (mx - mn) * (1 - 1e-9);
  const double offset = mn;
  for (int64_t i = 0; i < R; i++) t[i] = (t[i] - offset) * scale;
  if (info) {
    info[0] = offset;
    info[1] = scale;
    info[2] = (mx - mn) / ((float)p->frames);
  }
  if (p->vs_type == MRN_VS_NONE) {
    for (int64_t i = 0; i < R; i++) t[i] = std::floor(t[i]);
  } else if (p->vs_type == MRN_VS_HIST) {
    std::vector<int64_t> idx(R);
    for (int64_t i = 0; i < R; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return t[a] < t[b]; });
    for (int64_t r = 0; r < R; r++) v[idx[r]] = (double)(int)((double)(r * p->frames) / (double)R);
    std::copy(v.begin(), v.end(), t);
  }
  return MRN_OK;
}

int mrn_pregate(const mrn_gate_params *p, int64_t xres, int64_t R, const double *gt, const float *resp,
                const float *kx, const float *ky, const float *kz, const float *kw, int C, const void *kdata,
                int64_t *counts, float *kx_out, float *ky_out, float *kz_out, float *kw_out, void *kdata_out,
                void *stream) {
  free_err_clear();
  if (!p || !counts || xres < 1 || R < 1 || C < 0) return free_fail(MRN_EINVAL, "sizes");
  if (!gt || !kx || !ky || !kz || !kw || (C > 0 && !kdata)) return free_fail(MRN_EINVAL, "input arrays");
  if (p->frames < 1 || p->frames > 65535) return free_fail(MRN_EINVAL, "frames in 1..65535");
  if (p->vs_type == MRN_VS_TORNADO && !(p->kmax > 0.0f)) return free_fail(MRN_EINVAL, "tornado kmax");
  const bool fill = kx_out != nullptr;
  if (fill && (!ky_out || !kz_out || !kw_out || (C > 0 && !kdata_out)))
    return free_fail(MRN_EINVAL, "output arrays");
  GateDev g;
  g.vs_type = p->vs_type;
  g.shape = p->tornado_shape;
  g.frames = p->frames;
  g.reset = p->reset_dens != 0;
  g.wlo = p->wdth_low;
  g.whi = p->wdth_high;
  g.kmax = p->kmax;
  g.xres = xres;
  g.M = xres * R;
  g.nb = (g.M + 255) / 256;
  if (g.nb > 0x7fffffffll) return free_fail(MRN_EUNSUPPORTED, "too many samples");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nflat = (int64_t)g.frames * g.nb;
  int64_t *bcnt = nullptr;  // 64-bit: the scan accumulates in its input type
  int64_t *boff = nullptr, *fb = nullptr;
  void *tmp = nullptr;
  size_t tb = 0;
  int rc = MRN_OK;
  std::vector<int64_t> hfb(g.frames + 1);
  auto run = [&]() -> hipError_t {
    hipError_t e;
    if ((e = hipMalloc(&bcnt, sizeof(int64_t) * (nflat + 1))) != hipSuccess) return e;
    if ((e = hipMalloc(&boff, sizeof(int64_t) * (nflat + 1))) != hipSuccess) return e;
    if ((e = hipMalloc(&fb, sizeof(int64_t) * (g.frames + 1))) != hipSuccess) return e;
    if ((e = hipMemsetAsync(bcnt + nflat, 0, sizeof(int64_t), st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_gate_count, dim3((unsigned)g.nb, g.frames), dim3(256), 0, st, g, gt, resp, kx, ky, kz,
                       bcnt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, bcnt, boff, nflat + 1, st)) != hipSuccess) return e;
    if ((e = hipMalloc(&tmp, tb)) != hipSuccess) return e;
    if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, bcnt, boff, nflat + 1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_gate_frames, dim3((g.frames + 256) / 256), dim3(256), 0, st, g, boff, fb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(hfb.data(), fb, sizeof(int64_t) * (g.frames + 1), hipMemcpyDeviceToHost, st)) !=
        hipSuccess)
      return e;
    if (fill) {
      hipLaunchKernelGGL(k_gate_scatter, dim3((unsigned)g.nb, g.frames), dim3(256), 0, st, g, gt, resp, kx, ky, kz,
                         kw, C, (const float2 *)kdata, boff, fb, kx_out, ky_out, kz_out, kw_out, (float2 *)kdata_out);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipStreamSynchronize(st);
  };
  const hipError_t e = run();
  if (e != hipSuccess) {
    hipStreamSynchronize(st);
    rc = free_fail(MRN_EHIP, hipGetErrorString(e));
  } else {
    for (int t = 0; t < g.frames; t++) counts[t] = hfb[t + 1] - hfb[t];
  }
  hipFree(bcnt);
  hipFree(boff);
  hipFree(fb);
  hipFree(tmp);
  return rc;
}

}  // extern "C"
