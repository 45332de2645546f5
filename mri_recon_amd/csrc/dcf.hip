// Pipe density compensation on the device: RECON::dcf_calc
// (recon_lib.cpp:854-917) over DCFgridFFT's real-valued polynomial-kernel
// gridder (DCFgridFFT.cpp:96-440).
//
// The samples are first sorted by 8x8x8-cell brick of the grid and 4x4x4
// sub-brick (one device radix sort): a wavefront's samples then touch
// neighbouring cells, and the scatter's atomics and the gather's reads hit
// lines already in L2.  The iterations run on the sorted copy; the weights
// are put back in the caller's order at the end.
//
// The trajectory stays resident; one call runs every iteration on the stream:
//   kw = 1; repeat iters: X = 0; X += A^T kw (scatter); kw2 = A X (gather);
//   kw = |kw2| < 1e-3 ? 0 : kw / kw2;   then kw = |kw * scale^3|.
// The gather and the ratio update are one kernel; no per-iteration host sync.
//
// HBM layout: X is the real grid [Sz][Sy][Sx] (x fastest) with the reference's
// extents trunc(os*N + 16), the kx extent taken from Nz (recon_lib.cpp:873).
// The sample order is the caller's (radial spokes run outward, so neighbours
// in a wavefront touch neighbouring cells).
//
// Arithmetic follows the reference expression by expression: float where it
// is float and double where it promotes (the table interpolation and the
// forward scale), with the *_rn intrinsics so that no multiply-add is fused
// where the reference rounds twice.  The scatter adds with float atomics, as
// the reference's OpenMP atomics do, so the order of the additions into a
// cell differs from run to run; the gather sums in the reference's tap order.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include <hipcub/hipcub.hpp>

#include "mrn.h"

namespace mrn {

thread_local std::string g_free_err;  // message of the last failed call without an op

const char *free_error() { return g_free_err.c_str(); }
void free_err_clear() { g_free_err.clear(); }
int free_fail(int status, const char *msg) {
  g_free_err = msg ? msg : "";
  return status;
}

namespace {

constexpr int DCF_MOD = 600;          // table points per cell (DCFgridFFT.cpp:101-103)
constexpr int DCF_MAX_TABLE = 12288;  // table floats held in LDS (48 KB)
// sort brick edge in cells (5.12 M radial samples, 553^3, 20 iterations,
// plain brick order: 8^3 bricks 104.6 ms, 4^3 106.7, 2^3 119, 16^3 155; 8^3
// with a 4096-cell box 115; 8^3 with 4^3 sub-bricks 100.4 against 102.1)
constexpr int DCF_BRICK = 8;
static_assert(DCF_BRICK >= 2 && DCF_BRICK % 2 == 0, "the sort key splits each brick into 8 sub-bricks");

struct DcfGeom {
  int S[3];
  float c[3];   // grid centre Sx/2 (integer division), as float
  float gx;     // grid_x = overgrid
  float dwin;
  float scale;  // dcf_scale applied to k
  float acc;
  int len;
};

int dcf_fail(int status, const char *msg) {
  g_free_err = msg;
  return status;
}

#define DCF_HIP(call)                                                       \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      g_free_err = std::string(#call ": ") + hipGetErrorString(e_);         \
      return MRN_EHIP;                                                      \
    }                                                                       \
  } while (0)

// Polynomial table (DCFgridFFT.cpp:56-90, 160-164), on the host in double.
int dcf_table(float dwin, int kernel, float *tab) {
  static const double P1[6] = {1., 0.04522831, -3.36020304, 1.12417012, 2.82448025, -1.63447764};
  static const double P2[6] = {1., 0.03056504, -3.01961845, 0.6679865, 2.77924058, -1.45923643};
  const double *c = kernel == 1 ? P2 : P1;
  const int len = (int)(dwin * (float)DCF_MOD) + 100;
  if (!tab) return len;
  double sum = 0.0;
  for (int i = 0; i < len; i++) {
    const double x = (double)i / (double)len;
    const double x2 = x * x, x4 = x2 * x2, x3 = x * x2, x5 = x * x4;
    tab[i] = (float)(c[0] + c[1] * x + c[2] * x2 + c[3] * x3 + c[4] * x4 + c[5] * x5);
  }
  for (int i = 0; i < len; i++) sum += tab[i];  // float entries summed in double
  const double f = 0.5 * (double)DCF_MOD / sum;
  for (int i = 0; i < len; i++) tab[i] = (float)(tab[i] * f);
  return len;
}

void grid_size(const mrn_dcf_params *p, int S[3]) {
  S[0] = (int)(p->overgrid * (float)p->N[2] + 16.0f);
  S[1] = (int)(p->overgrid * (float)p->N[1] + 16.0f);
  S[2] = (int)(p->overgrid * (float)p->N[0] + 16.0f);
}

__device__ __forceinline__ float kr_of(const DcfGeom &g, float kx, float ky, float kz, float k[3]) {
  k[0] = __fmul_rn(kx, g.scale);
  k[1] = __fmul_rn(ky, g.scale);
  k[2] = __fmul_rn(kz, g.scale);
  return __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(k[0], k[0]), __fmul_rn(k[1], k[1])), __fmul_rn(k[2], k[2])));
}

// forward / scale_kw: kr / 128.0 * (acc - 1) + 1 in double
__device__ __forceinline__ float scale_fwd(const DcfGeom &g, float kr) {
  return (float)__dadd_rn(__dmul_rn(__ddiv_rn((double)kr, 128.0), (double)(g.acc - 1.0f)), 1.0);
}
// backward: kr / 128 * (acc - 1) + 1 in float
__device__ __forceinline__ float scale_bwd(const DcfGeom &g, float kr) {
  return __fadd_rn(__fmul_rn(__fdiv_rn(kr, 128.0f), __fsub_rn(g.acc, 1.0f)), 1.0f);
}

__device__ __forceinline__ bool axis(const DcfGeom &g, int a, float k, float scale, float &dk, int &s, int &e) {
  dk = __fadd_rn(__fmul_rn(k, g.gx), g.c[a]);
  const float r = __fmul_rn(scale, g.dwin);
  s = max((int)ceilf(__fsub_rn(dk, r)), 0);
  e = min((int)floorf(__fadd_rn(dk, r)), g.S[a] - 1);
  return !(s >= g.S[a] || e < 0);
}

// t[(int)del] * (1.0 - d) + t[(int)del + 1] * d, the first product in double
// and the second in float, as the reference's expression promotes
__device__ __forceinline__ double lerp(const float *tab, float dk, int l, float scale) {
  const float del = __fdiv_rn(fabsf(__fmul_rn((float)DCF_MOD, __fsub_rn(dk, (float)l))), scale);
  const int i = (int)del;
  const float d = __fsub_rn(del, (float)i);
  return __dadd_rn(__dmul_rn((double)tab[i], __dsub_rn(1.0, (double)d)), (double)__fmul_rn(tab[i + 1], d));
}

// Geometry of sample i for the forward gridder (DCFgridFFT.cpp:171-240): its
// kw, kernel scale, window centre dk and tap range [s, e] per axis.  False
// when the sample has no taps (i >= M, kw = 0, scale <= 0, or off the grid).
__device__ __forceinline__ bool fwd_taps(const DcfGeom &g, const float *__restrict__ kw,
                                         const float *__restrict__ kx, const float *__restrict__ ky,
                                         const float *__restrict__ kz, int64_t i, int64_t M, float &temp,
                                         float &scale, float dk[3], int s[3], int e[3]) {
  if (i >= M) return false;
  float k[3];
  scale = scale_fwd(g, kr_of(g, kx[i], ky[i], kz[i], k));
  temp = kw[i];
  bool live = temp != 0.0f && scale > 0.0f;  // a scale <= 0 (acc < 1, large |k|) has no taps
  live = live && axis(g, 0, k[0], scale, dk[0], s[0], e[0]);
  live = live && axis(g, 1, k[1], scale, dk[1], s[1], e[1]);
  live = live && axis(g, 2, k[2], scale, dk[2], s[2], e[2]);
  return live;
}

__device__ __forceinline__ void load_table(float *s_tab, const float *tab, int len) {
  for (int i = threadIdx.x; i < len; i += blockDim.x) s_tab[i] = tab[i];
  __syncthreads();
}

__global__ void k_dcf_fill(float *v, int64_t n, float value) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = value;
}

// DCFgridFFT::forward (DCFgridFFT.cpp:171-289): kernel weights with the
// reference's float/double promotions, grid += w * kw.  The samples run in
// DCF_BRICK^3-brick order, so in the dense centre of k-space a workgroup's
// 256 samples reach a box of a few thousand cells.  When the box (the
// bounding box of the taps of the workgroup's samples) holds at most DCF_BOX
// cells, the taps are summed there with FP64 LDS atomics (ds_add_f64; FP32
// LDS atomics are ~20x slower on gfx950) and each touched cell gets one
// float atomic in the grid, instead of one per tap.  Larger boxes (sparse
// regions) take the direct path, DCF_XL lanes per sample.  Launched with 256
// threads per workgroup.  The grid sums were already unordered (float
// atomics); a box only sums its part in double before rounding once.
// FP64 LDS box cells (64 KB).  20 iterations at 5.12 M radial samples, 553^3,
// 4^3 bricks: 105 ms; 4096 cells 113, 2048 120, no box 153 (the dense centre
// contends)
constexpr int DCF_BOX = 8192;
// Lanes per sample on the direct path.  The L2 serves float atomics per
// 64-byte line request: about 21 G requests/s on gfx950 when every lane of a
// wave has its own line, while lanes that share a line share its request
// (16 lanes on one line: 333 G atomics/s).  With one lane per x tap of a
// row, a row's 4-5 taps cost about one request instead of one each.
constexpr int DCF_XL = 8;

// float atomics the scatters of the last mrn_dcf_calc on this device issued
// to the grid (diagnostic: one 64-bit atomic per wavefront and launch)
static __device__ unsigned long long g_dcf_atomics;
__global__ void k_dcf_count_reset() { g_dcf_atomics = 0ull; }
__device__ __forceinline__ void count_atomics(unsigned long long n) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
  if ((threadIdx.x & 63) == 0 && n) atomicAdd(&g_dcf_atomics, n);
}
__global__ __launch_bounds__(256) void k_dcf_scatter_box(float *__restrict__ X, const float *__restrict__ kw,
                                                         const float *__restrict__ kx, const float *__restrict__ ky,
                                                         const float *__restrict__ kz, int64_t M, DcfGeom g,
                                                         const float *__restrict__ tab) {
  extern __shared__ double dsh[];
  double *s_acc = dsh;
  float *s_tab = reinterpret_cast<float *>(dsh + DCF_BOX);
  __shared__ int s_red[6][256 / 64];
  load_table(s_tab, tab, g.len);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float temp = 0.0f, scale = 0.0f, dk[3] = {0.f, 0.f, 0.f};
  int s[3] = {0, 0, 0}, e[3] = {-1, -1, -1};
  const bool live = fwd_taps(g, kw, kx, ky, kz, i, M, temp, scale, dk, s, e);
  // bounding box of the workgroup's taps
  int r6[6];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    r6[a] = live ? s[a] : INT_MAX;
    r6[3 + a] = live ? e[a] : INT_MIN;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      r6[a] = min(r6[a], __shfl_xor(r6[a], o));
      r6[3 + a] = max(r6[3 + a], __shfl_xor(r6[3 + a], o));
    }
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int a = 0; a < 6; a++) s_red[a][threadIdx.x >> 6] = r6[a];
  __syncthreads();
  int lo[3], ext[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    int mn = INT_MAX, mx = INT_MIN;
    for (int w = 0; w < 256 / 64; w++) {
      mn = min(mn, s_red[a][w]);
      mx = max(mx, s_red[3 + a][w]);
    }
    lo[a] = mn;
    ext[a] = mx >= mn ? mx - mn + 1 : 0;
  }
  const int64_t vol = (int64_t)ext[0] * ext[1] * ext[2];  // 0: no live sample
  if (vol == 0) return;                                     // (uniform)
  unsigned long long nat = 0;  // grid atomics of this thread
  if (vol > DCF_BOX) {  // (uniform) direct path: DCF_XL lanes per sample, one per x tap
    const int j = (int)(threadIdx.x % DCF_XL);
    const int64_t base = (int64_t)blockIdx.x * blockDim.x + threadIdx.x / DCF_XL;
    for (int r = 0; r < DCF_XL; r++) {  // the workgroup's samples, 256 / DCF_XL at a time
      float t2 = 0.0f, sc2 = 0.0f, dk2[3] = {0.f, 0.f, 0.f};
      int s2[3] = {0, 0, 0}, e2[3] = {-1, -1, -1};
      if (!fwd_taps(g, kw, kx, ky, kz, base + r * (256 / DCF_XL), M, t2, sc2, dk2, s2, e2)) continue;
      for (int lz = s2[2]; lz <= e2[2]; lz++) {
        const float wtz = (float)lerp(s_tab, dk2[2], lz, sc2);
        for (int ly = s2[1]; ly <= e2[1]; ly++) {
          const float wty = (float)__dmul_rn((double)wtz, lerp(s_tab, dk2[1], ly, sc2));
          float *row = X + ((int64_t)lz * g.S[1] + ly) * g.S[0];
          for (int lx = s2[0] + j; lx <= e2[0]; lx += DCF_XL) {
            const float wtx = (float)__dmul_rn((double)wty, lerp(s_tab, dk2[0], lx, sc2));
            atomicAdd(row + lx, __fmul_rn(wtx, t2));
            nat++;
          }
        }
      }
    }
    count_atomics(nat);
    return;
  }
  for (int q = threadIdx.x; q < vol; q += blockDim.x) s_acc[q] = 0.0;
  __syncthreads();
  if (live) {
    for (int lz = s[2]; lz <= e[2]; lz++) {
      const float wtz = (float)lerp(s_tab, dk[2], lz, scale);
      for (int ly = s[1]; ly <= e[1]; ly++) {
        const float wty = (float)__dmul_rn((double)wtz, lerp(s_tab, dk[1], ly, scale));
        double *brow = s_acc + ((lz - lo[2]) * ext[1] + (ly - lo[1])) * ext[0] - lo[0];
        for (int lx = s[0]; lx <= e[0]; lx++) {
          const float wtx = (float)__dmul_rn((double)wty, lerp(s_tab, dk[0], lx, scale));
          atomicAdd(brow + lx, (double)__fmul_rn(wtx, temp));
        }
      }
    }
  }
  __syncthreads();
  // one float atomic per touched cell; consecutive lanes take consecutive x
  for (int q = threadIdx.x; q < vol; q += blockDim.x) {
    const double v = s_acc[q];
    if (v == 0.0) continue;
    const int x = q % ext[0], yz = q / ext[0];
    const int y = yz % ext[1], z = yz / ext[1];
    atomicAdd(X + ((int64_t)(lo[2] + z) * g.S[1] + (lo[1] + y)) * g.S[0] + lo[0] + x, (float)v);
    nat++;
  }
  count_atomics(nat);
}

// DCFgridFFT::backward (DCFgridFFT.cpp:314-440) and the ratio update of
// dcf_calc (recon_lib.cpp:895-905), one thread per sample.  sum (optional)
// collects the iteration's sum of |kw| (the reference prints it).
__global__ __launch_bounds__(256) void k_dcf_gather_update(const float *__restrict__ X, float *__restrict__ kw,
                                                           const float *__restrict__ kx,
                                                           const float *__restrict__ ky,
                                                           const float *__restrict__ kz, int64_t M, DcfGeom g,
                                                           const float *__restrict__ tab, double *sum) {
  extern __shared__ float s_tab[];
  __shared__ double s_part[256 / 64];
  load_table(s_tab, tab, g.len);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double mine = 0.0;
  if (i < M) {
    float k[3];
    const float kr = kr_of(g, kx[i], ky[i], kz[i], k);
    const float scale = scale_bwd(g, kr);
    float temp = 0.0f;
    float dk[3];
    int s[3], e[3];
    if (scale > 0.0f && axis(g, 0, k[0], scale, dk[0], s[0], e[0]) && axis(g, 1, k[1], scale, dk[1], s[1], e[1]) &&
        axis(g, 2, k[2], scale, dk[2], s[2], e[2])) {
      for (int lz = s[2]; lz <= e[2]; lz++) {
        const float wtz = (float)lerp(s_tab, dk[2], lz, scale);
        for (int ly = s[1]; ly <= e[1]; ly++) {
          const float wty = (float)__dmul_rn((double)wtz, lerp(s_tab, dk[1], ly, scale));
          const float *row = X + ((int64_t)lz * g.S[1] + ly) * g.S[0];
          const int nx = e[0] - s[0] + 1;
          if (nx <= 8 && s[0] + 8 <= g.S[0]) {
            // the row's taps as two 16-byte loads (4-byte aligned; the 8
            // cells stay inside the row), summed in the same tap order
            float v[8];
            __builtin_memcpy(v, row + s[0], sizeof(v));
#pragma unroll
            for (int q = 0; q < 8; q++)
              if (q < nx) {
                const float wtx = (float)__dmul_rn((double)wty, lerp(s_tab, dk[0], s[0] + q, scale));
                temp = __fadd_rn(temp, __fmul_rn(wtx, v[q]));
              }
          } else {
            for (int lx = s[0]; lx <= e[0]; lx++) {
              const float wtx = (float)__dmul_rn((double)wty, lerp(s_tab, dk[0], lx, scale));
              temp = __fadd_rn(temp, __fmul_rn(wtx, row[lx]));
            }
          }
        }
      }
    }
    float w = kw[i];
    if ((double)fabsf(temp) < 1e-3) {
      w = 0.0f;
    } else {
      w = __fdiv_rn(w, temp);
      mine = (double)fabsf(w);
    }
    kw[i] = w;
  }
  if (sum) {
    for (int o = 32; o > 0; o >>= 1) mine += __shfl_down(mine, o);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += s_part[w];
      atomicAdd(sum, t);
    }
  }
}

// scale_kw (DCFgridFFT.cpp:291-312) and kw = |kw| (recon_lib.cpp:910)
__global__ void k_dcf_finish(float *__restrict__ kw, const float *__restrict__ kx, const float *__restrict__ ky,
                             const float *__restrict__ kz, int64_t M, DcfGeom g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  float k[3];
  const double s = (double)scale_fwd(g, kr_of(g, kx[i], ky[i], kz[i], k));
  kw[i] = fabsf((float)__dmul_rn((double)kw[i], __dmul_rn(__dmul_rn(s, s), s)));
}

// sort key: the DCF_BRICK^3-cell brick of the sample's nearest grid cell,
// then its (DCF_BRICK / 2)^3 sub-brick: the scatter's workgroups take whole
// bricks (their LDS boxes), the gather's wavefronts sub-bricks (their loads
// stay closer together)
__global__ void k_dcf_keys(const float *__restrict__ kx, const float *__restrict__ ky,
                           const float *__restrict__ kz, int64_t M, DcfGeom g, uint32_t *__restrict__ keys,
                           uint32_t *__restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  float k[3];
  kr_of(g, kx[i], ky[i], kz[i], k);
  uint32_t key = 0, sub = 0;
  for (int a = 2; a >= 0; a--) {
    const float dk = __fadd_rn(__fmul_rn(k[a], g.gx), g.c[a]);
    int c = (int)floorf(dk);
    c = c < 0 ? 0 : (c > g.S[a] - 1 ? g.S[a] - 1 : c);
    key = key * (uint32_t)((g.S[a] + DCF_BRICK - 1) / DCF_BRICK) + (uint32_t)(c / DCF_BRICK);
    sub = sub * 2u + (uint32_t)((c % DCF_BRICK) / (DCF_BRICK / 2));
  }
  keys[i] = key * 8u + sub;
  idx[i] = (uint32_t)i;
}

__global__ void k_dcf_gather3(const uint32_t *__restrict__ perm, const float *__restrict__ kx,
                              const float *__restrict__ ky, const float *__restrict__ kz, int64_t M,
                              float *__restrict__ sx, float *__restrict__ sy, float *__restrict__ sz) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const uint32_t i = perm[j];
  sx[j] = kx[i];
  sy[j] = ky[i];
  sz[j] = kz[i];
}

__global__ void k_dcf_unsort(const uint32_t *__restrict__ perm, const float *__restrict__ ws, int64_t M,
                             float *__restrict__ kw) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < M) kw[perm[j]] = ws[j];
}

int64_t nblocks(int64_t n, int b) { return (n + b - 1) / b; }

}  // namespace
}  // namespace mrn

using namespace mrn;

extern "C" {

void mrn_dcf_params_default(mrn_dcf_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->iters = 20;
  p->dwin = 2.1f;
  p->scale = 1.0f;
  p->overgrid = 2.1f;
  p->acc = 1.0f;
  p->kernel = MRN_DCF_POLY;
}

int mrn_dcf_parse_args(mrn_dcf_params *p, int argc, const char **argv) {
  if (!p || (argc > 0 && !argv)) return MRN_EINVAL;
  for (int i = 0; i < argc; i++) {
    const char *s = argv[i];
    const bool has = i + 1 < argc;
    if (!strcmp(s, "-dcf_iter")) {
      if (!has) return MRN_EINVAL;
      p->iters = atoi(argv[++i]);
    } else if (!strcmp(s, "-dcf_dwin") || !strcmp(s, "-dcf_scale") || !strcmp(s, "-dcf_overgrid") ||
               !strcmp(s, "-dcf_acc")) {
      if (!has) return MRN_EINVAL;
      const float v = (float)atof(argv[++i]);
      if (!strcmp(s, "-dcf_dwin")) p->dwin = v;
      else if (!strcmp(s, "-dcf_scale")) p->scale = v;
      else if (!strcmp(s, "-dcf_overgrid")) p->overgrid = v;
      else p->acc = v;
    }
  }
  return MRN_OK;
}

int mrn_dcf_grid_size(const mrn_dcf_params *p, int32_t S[3]) {
  if (!p || !S) return MRN_EINVAL;
  int s[3];
  grid_size(p, s);
  for (int a = 0; a < 3; a++) S[a] = s[a];
  return MRN_OK;
}

int mrn_dcf_last_atomics(int64_t *count) {
  if (!count) return MRN_EINVAL;
  unsigned long long v = 0;
  const hipError_t e = hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_dcf_atomics), sizeof(v), 0, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return dcf_fail(MRN_EHIP, hipGetErrorString(e));
  *count = (int64_t)v;
  return MRN_OK;
}

int mrn_dcf_calc(const mrn_dcf_params *p, int64_t M, const float *kx, const float *ky, const float *kz,
                 float *kw, double *sums, void *stream) {
  g_free_err.clear();
  if (!p) return dcf_fail(MRN_EINVAL, "null parameters");
  if (M < 0 || (M > 0 && (!kx || !ky || !kz || !kw))) return dcf_fail(MRN_EINVAL, "sample arrays");
  if (p->N[0] < 1 || p->N[1] < 1 || p->N[2] < 1) return dcf_fail(MRN_EINVAL, "image size");
  if (p->iters < 0) return dcf_fail(MRN_EINVAL, "dcf_iter < 0");
  if (!(p->dwin > 0.0f) || !(p->overgrid > 0.0f) || !std::isfinite(p->scale) || !std::isfinite(p->acc))
    return dcf_fail(MRN_EINVAL, "dcf_dwin and dcf_overgrid must be positive, dcf_scale and dcf_acc finite");
  if (p->kernel != MRN_DCF_POLY && p->kernel != MRN_DCF_POLY2) return dcf_fail(MRN_EINVAL, "dcf kernel");
  const int len = dcf_table(p->dwin, p->kernel, nullptr);
  if (len > DCF_MAX_TABLE) return dcf_fail(MRN_EUNSUPPORTED, "dcf_dwin too large for the kernel table");
  DcfGeom g;
  grid_size(p, g.S);
  for (int a = 0; a < 3; a++) {
    if (g.S[a] < 1) return dcf_fail(MRN_EINVAL, "grid size");
    g.c[a] = (float)(g.S[a] / 2);
  }
  g.gx = p->overgrid;
  g.dwin = p->dwin;
  g.scale = p->scale;
  g.acc = p->acc;
  g.len = len;
  if (M == 0) return MRN_OK;
  if (M > 0x7fffffffll) return dcf_fail(MRN_EUNSUPPORTED, "more than 2^31 - 1 samples");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nS = (int64_t)g.S[0] * g.S[1] * g.S[2];
  float *tab_h = new float[len];
  dcf_table(p->dwin, p->kernel, tab_h);
  float *X = nullptr, *tab = nullptr, *srt = nullptr;  // srt: sorted kx, ky, kz, kw
  uint32_t *keys = nullptr, *perm = nullptr;           // [2][M] each
  void *tmp = nullptr;
  size_t tb = 0;
  double *dsum = nullptr;
  int rc = MRN_OK;
  auto run = [&]() -> int {
    DCF_HIP(hipMalloc(&X, sizeof(float) * nS));
    DCF_HIP(hipMalloc(&tab, sizeof(float) * len));
    DCF_HIP(hipMalloc(&dsum, sizeof(double) * (p->iters > 0 ? p->iters : 1)));
    DCF_HIP(hipMemcpyAsync(tab, tab_h, sizeof(float) * len, hipMemcpyHostToDevice, st));
    const size_t lds = sizeof(float) * len;
    const size_t lds_box = sizeof(double) * DCF_BOX + lds;
    static const hipError_t box_attr = hipFuncSetAttribute((const void *)k_dcf_scatter_box,
                                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                                           (int)(sizeof(double) * DCF_BOX + sizeof(float) * DCF_MAX_TABLE));
    DCF_HIP(box_attr);
    const int fillb = 4096;
    const dim3 gs((unsigned)nblocks(M, 256));
    // brick order
    DCF_HIP(hipMalloc(&keys, sizeof(uint32_t) * 2 * M));
    DCF_HIP(hipMalloc(&perm, sizeof(uint32_t) * 2 * M));
    DCF_HIP(hipMalloc(&srt, sizeof(float) * 4 * M));
    hipLaunchKernelGGL(k_dcf_keys, gs, dim3(256), 0, st, kx, ky, kz, M, g, keys, perm);
    DCF_HIP(hipGetLastError());
    const uint64_t nkeys = 8ull * (uint64_t)((g.S[0] + DCF_BRICK - 1) / DCF_BRICK) * ((g.S[1] + DCF_BRICK - 1) / DCF_BRICK) *
                           ((g.S[2] + DCF_BRICK - 1) / DCF_BRICK);
    int bits = 1;
    while (bits < 32 && (1ull << bits) < nkeys) bits++;
    DCF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, keys + M, perm, perm + M, (int)M, 0, bits, st));
    DCF_HIP(hipMalloc(&tmp, tb));
    DCF_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, keys + M, perm, perm + M, (int)M, 0, bits, st));
    const uint32_t *order = perm + M;
    float *sx = srt, *sy = srt + M, *sz = srt + 2 * M, *sw = srt + 3 * M;
    hipLaunchKernelGGL(k_dcf_gather3, gs, dim3(256), 0, st, order, kx, ky, kz, M, sx, sy, sz);
    hipLaunchKernelGGL(k_dcf_fill, dim3(fillb), dim3(256), 0, st, sw, M, 1.0f);
    DCF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_dcf_count_reset, dim3(1), dim3(1), 0, st);
    if (p->iters > 0) {
      // zero the per-iteration sums (a double 0 is two float 0s)
      hipLaunchKernelGGL(k_dcf_fill, dim3(1), dim3(256), 0, st, reinterpret_cast<float *>(dsum),
                         (int64_t)2 * p->iters, 0.0f);
      DCF_HIP(hipGetLastError());
    }
    for (int it = 0; it < p->iters; it++) {
      hipLaunchKernelGGL(k_dcf_fill, dim3(fillb), dim3(256), 0, st, X, nS, 0.0f);
      hipLaunchKernelGGL(k_dcf_scatter_box, gs, dim3(256), lds_box, st, X, sw, sx, sy, sz, M, g, tab);
      hipLaunchKernelGGL(k_dcf_gather_update, gs, dim3(256), lds, st, X, sw, sx, sy, sz, M, g, tab,
                         sums ? dsum + it : nullptr);
      DCF_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_dcf_finish, gs, dim3(256), 0, st, sw, sx, sy, sz, M, g);
    hipLaunchKernelGGL(k_dcf_unsort, gs, dim3(256), 0, st, order, sw, M, kw);
    DCF_HIP(hipGetLastError());
    if (sums && p->iters > 0)
      DCF_HIP(hipMemcpyAsync(sums, dsum, sizeof(double) * p->iters, hipMemcpyDeviceToHost, st));
    DCF_HIP(hipStreamSynchronize(st));  // the scratch grid is freed below
    return MRN_OK;
  };
  rc = run();
  if (rc != MRN_OK) hipStreamSynchronize(st);
  hipFree(X);
  hipFree(tab);
  hipFree(dsum);
  hipFree(keys);
  hipFree(perm);
  hipFree(srt);
  hipFree(tmp);
  delete[] tab_h;
  return rc;
}

}  // extern "C"
