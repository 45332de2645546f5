// L1 wavelet proximal step of the FISTA loop, on the device.
//
// Follows RECON::L1_threshold (recon_lib.cpp:2416-2527) with its default
// spatial wavelet and fraction threshold, without phase rotation or temporal
// transforms: for each cycle spin, a 3-D periodic wavelet transform of every
// frame (WAVELET3D::forward3D, wavelet3D.cpp:563-594), the fraction threshold
// over all frames on the first spin, scaled by 1/spins
// (THRESHOLD::update_threshold / get_threshold, threshold.cpp:193-380), the
// complex soft threshold (thresholding, threshold.cpp:766-800), and the
// inverse transform (backward3D, wavelet3D.cpp:596-631).
//
// The wavelet is the 4-tap Daubechies pair of the reference's default
// (wavelet3D.cpp:143-190), computed from its closed form.  A line transform
// stages FL lines in LDS and writes the result back in place; the inverse is
// evaluated in gather form (each output sums its two contributing taps).
// The fraction threshold runs entirely on the device: a min/max pass, then 16
// rounds of (count below the bisection's next 15 candidate values, one-thread
// replay of up to 4 bisection steps), each round a no-op once the bracket has
// converged, so the whole step can be captured in a hipGraph.  |c| is evaluated as sqrt in double rounded to float, so the
// counts and the threshold do not depend on the order of operations.
#include <cmath>
#include <cstring>

#include "mrn_internal.hpp"

namespace mrn {

#define WL_LINES 16

struct D4 {
  float lpf[4], hpf[4], slpf[4], shpf[4];
};

static D4 d4_filters() {
  const double r3 = std::sqrt(3.0), n = 4.0 * std::sqrt(2.0);
  const double h0 = (1 + r3) / n, h1 = (3 + r3) / n, h2 = (3 - r3) / n, h3 = (1 - r3) / n;
  const double lp[4] = {h3, h2, h1, h0}, hp[4] = {-h0, h1, -h2, h3};
  const double sl[4] = {h0, h1, h2, h3}, shp[4] = {h3, -h2, h1, -h0};
  D4 f;
  for (int q = 0; q < 4; q++) {
    f.lpf[q] = (float)lp[q];
    f.hpf[q] = (float)hp[q];
    f.slpf[q] = (float)sl[q];
    f.shpf[q] = (float)shp[q];
  }
  return f;
}

__device__ __forceinline__ int wmod(int i, int n) {
  const int r = i % n;
  return r < 0 ? r + n : r;
}

// In-place one-level transform (INV = 0 analysis, 1 synthesis) of the lines
// along `axis` of the sub-block [0,s0) x [0,s1) x [0,s2) of frame blockIdx.y.
// The shift of the axis is read from sh[axis] (device memory; 0..3).  No
// integer division per element: the workgroup's 16 line bases are computed
// once, the x-axis walk steps by a precomputed quotient and remainder, and
// the periodic index wraps with one add or subtract for lines of 4 or more
// points (the indices reach [-3, pts + 2]).  Launched with 256 threads.
template <int INV>
__global__ __launch_bounds__(256) void k_wave_lines(float2 *X, int N0, int N1, int s0, int s1, int s2,
                                                    int axis, const int *__restrict__ sh, D4 f,
                                                    int64_t frame_stride) {
  extern __shared__ float2 wbuf[];  // [WL_LINES][pts]
  __shared__ int64_t s_base[WL_LINES];
  const int sz[3] = {s0, s1, s2};
  const int pts = sz[axis], h = pts >> 1;
  const int b = axis == 0 ? 1 : 0, c = axis == 2 ? 1 : 2;
  const int64_t st[3] = {1, N0, (int64_t)N0 * N1};
  const int64_t sa = st[axis];
  const int nlines = sz[b] * sz[c];
  const int l0 = blockIdx.x * WL_LINES;
  float2 *Xf = X + (int64_t)blockIdx.y * frame_stride;
  const int shift = sh[axis];
  const int tot = WL_LINES * pts;
  if (threadIdx.x < WL_LINES) {
    const int L = l0 + (int)threadIdx.x, u = L % sz[b], v = L / sz[b];
    s_base[threadIdx.x] = (int64_t)u * st[b] + (int64_t)v * st[c];
  }
  // element e = threadIdx.x + k * blockDim.x as (line, point): an x line is
  // contiguous (ln = e / pts, p = e % pts); along y and z the 16 lines of a
  // point are neighbouring x (or y) cells (p = e / 16, ln = e % 16)
  const int dq = (int)blockDim.x / pts, dr = (int)blockDim.x - dq * pts;
  const int q0 = (int)threadIdx.x / pts, r0 = (int)threadIdx.x - q0 * pts;
  auto walk = [&](auto &&fn) {
    if (axis == 0) {
      int ln = q0, p = r0;
      for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        fn(ln, p);
        p += dr;
        ln += dq;
        if (p >= pts) {
          p -= pts;
          ln++;
        }
      }
    } else {
      const int ln = (int)threadIdx.x % WL_LINES;
      for (int p = (int)threadIdx.x / WL_LINES; p < pts; p += (int)blockDim.x / WL_LINES) fn(ln, p);
    }
  };
  const bool near = pts >= 4;
  auto wrap = [&](int i) { return near ? (i < 0 ? i + pts : (i >= pts ? i - pts : i)) : wmod(i, pts); };
  __syncthreads();
  walk([&](int ln, int p) {
    if (l0 + ln < nlines) wbuf[ln * pts + p] = Xf[s_base[ln] + p * sa];
  });
  __syncthreads();
  walk([&](int ln, int p) {
    if (l0 + ln >= nlines) return;
    const float2 *in = wbuf + ln * pts;
    float2 o = make_float2(0.f, 0.f);
    if (!INV) {  // wave1D: out[pos] = sum lpf * in, out[pos + h] = sum hpf * in
      const bool lo = p < h;
      const int pos = lo ? p : p - h;
#pragma unroll
      for (int fp = 0; fp < 4; fp++) {
        const float2 v = in[wrap(2 * pos - fp + shift)];
        const float w = lo ? f.lpf[fp] : f.hpf[fp];
        o.x += w * v.x;
        o.y += w * v.y;
      }
    } else {  // iwave1D: in[pos] and in[pos + h] reach out[2 pos + fp - 3 + shift]
#pragma unroll
      for (int fp = 0; fp < 4; fp++) {
        const int t = p - fp + 3 - shift;
        if (t & 1) continue;
        const int pos = wrap(t) >> 1;
        const float2 a = in[pos], d = in[pos + h];
        o.x += a.x * f.slpf[fp];
        o.y += a.y * f.slpf[fp];
        o.x += d.x * f.shpf[fp];
        o.y += d.y * f.shpf[fp];
      }
    }
    Xf[s_base[ln] + p * sa] = o;
  });
}

// ------------------------------------------------------------ threshold
#define THR_DEPTH 4                         // bisection steps per counting pass
#define THR_NODES ((1 << THR_DEPTH) - 1)    // candidate values per pass
struct ThrState {
  unsigned int umin, umax;  // bit patterns of min / max |c| (non-negative floats)
  float value, max_t, min_t, thr;
  int iter, done;
  unsigned long long counter;
  long long found, target, accuracy, total;
  unsigned long long cnt[THR_NODES];  // count(|c| < candidate) of this pass
};

// The next THR_DEPTH values the bisection can query, breadth first: node k
// (value v, bracket [lo, hi]) has children 2k + 1 (count > target: max = v)
// and 2k + 2 (min = v), each at 0.5 * (max + min) as the sequential step
// computes it, so that a pass counts exactly the values the sequence visits.
__device__ __forceinline__ void thr_tree(const ThrState &s, float (&v)[THR_NODES]) {
  float lo[THR_NODES], hi[THR_NODES];
  v[0] = s.value;
  lo[0] = s.min_t;
  hi[0] = s.max_t;
#pragma unroll
  for (int k = 0; 2 * k + 2 < THR_NODES; k++) {
    lo[2 * k + 1] = lo[k];
    hi[2 * k + 1] = v[k];
    v[2 * k + 1] = 0.5f * (hi[2 * k + 1] + lo[2 * k + 1]);
    lo[2 * k + 2] = v[k];
    hi[2 * k + 2] = hi[k];
    v[2 * k + 2] = 0.5f * (hi[2 * k + 2] + lo[2 * k + 2]);
  }
}

__device__ __forceinline__ float cabs_rn(float2 v) {
  return (float)sqrt((double)v.x * v.x + (double)v.y * v.y);
}

__global__ void k_thr_reset(ThrState *s) {
  s->umin = 0x7f800000u;  // +inf
  s->umax = 0u;
  s->counter = 0ull;
}

__global__ __launch_bounds__(256) void k_thr_minmax(const float2 *__restrict__ X, int64_t n, ThrState *s) {
  float mn = INFINITY, mx = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float v = cabs_rn(X[i]);
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
  }
  for (int o = 32; o > 0; o >>= 1) {
    mn = fminf(mn, __shfl_xor(mn, o));
    mx = fmaxf(mx, __shfl_xor(mx, o));
  }
  // one atomic pair per workgroup: atomics on one address serialise in L2
  __shared__ float s_mn[256 / 64], s_mx[256 / 64];
  if ((threadIdx.x & 63) == 0) {
    s_mn[threadIdx.x >> 6] = mn;
    s_mx[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      mn = fminf(mn, s_mn[w]);
      mx = fmaxf(mx, s_mx[w]);
    }
    atomicMin(&s->umin, __float_as_uint(mn));
    atomicMax(&s->umax, __float_as_uint(mx));
  }
}

// get_threshold's set-up (threshold.cpp:250-336)
__global__ void k_thr_init(ThrState *s, long long total, float fraction) {
  const float mx = __uint_as_float(s->umax), mn = __uint_as_float(s->umin);
  s->total = total;
  s->target = (long long)(fraction * (double)total);
  s->accuracy = (long long)(0.001 * (double)total);
  s->value = 0.5f * (mx - mn);
  s->max_t = fraction * mx;
  s->min_t = mn;
  s->iter = 0;
  s->found = total;
  s->counter = 0ull;
  for (int k = 0; k < THR_NODES; k++) s->cnt[k] = 0ull;
  s->done = 0;
  if (s->target < 2) {
    s->value = 0.0f;
    s->done = 1;
  } else if (!(llabs(s->found - s->target) > s->accuracy)) {
    s->done = 1;
  }
}

// One counting pass: count(|c| < v) for every candidate of thr_tree.
__global__ __launch_bounds__(256) void k_thr_count(const float2 *__restrict__ X, int64_t n, ThrState *s) {
  __shared__ unsigned int s_cnt[THR_NODES];
  if (s->done) return;  // (uniform)
  float v[THR_NODES];
  thr_tree(*s, v);
  if (threadIdx.x < THR_NODES) s_cnt[threadIdx.x] = 0u;
  __syncthreads();
  unsigned int cnt[THR_NODES];
#pragma unroll
  for (int k = 0; k < THR_NODES; k++) cnt[k] = 0u;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float a = cabs_rn(X[i]);
#pragma unroll
    for (int k = 0; k < THR_NODES; k++) cnt[k] += a < v[k] ? 1u : 0u;
  }
#pragma unroll
  for (int k = 0; k < THR_NODES; k++) {
    unsigned int c = cnt[k];
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&s_cnt[k], c);
  }
  __syncthreads();
  if (threadIdx.x < THR_NODES && s_cnt[threadIdx.x])
    atomicAdd(&s->cnt[threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
}

// one bisection step after a count (threshold.cpp:337-378)
// Up to THR_DEPTH steps of threshold.cpp:250-380's bisection, from the
// counts of one pass: the same values, counts and stopping rule as one step
// per pass.
__global__ void k_thr_step(ThrState *s) {
  if (s->done) return;
  float v[THR_NODES];
  thr_tree(*s, v);
  int node = 0;
  for (int lvl = 0; lvl < THR_DEPTH; lvl++) {
    s->value = v[node];
    s->found = (long long)s->cnt[node];
    s->iter++;
    if (!(llabs(s->found - s->target) > s->accuracy && s->iter < 64)) {
      s->done = 1;
      break;
    }
    if (s->found > s->target) {
      s->max_t = v[node];
      node = 2 * node + 1;
    } else {
      s->min_t = v[node];
      node = 2 * node + 2;
    }
    if (lvl == THR_DEPTH - 1) s->value = 0.5f * (s->max_t + s->min_t);
  }
  for (int k = 0; k < THR_NODES; k++) s->cnt[k] = 0ull;
}

// global threshold = scale * value; optionally logged per FISTA iteration
__global__ void k_thr_final(ThrState *s, float scale, double *log, const int *counter, int cap) {
  s->thr = scale * s->value;
  if (log) {
    const int it = counter ? *counter - 1 : 0;
    if (it >= 0 && it < cap) log[it] = (double)s->thr;
  }
}

__global__ __launch_bounds__(256) void k_soft(float2 *X, int64_t n, const ThrState *__restrict__ s) {
  const float thr = s->thr;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = X[i];
    const float cc = cabs_rn(v);
    if (cc <= thr) {
      X[i] = make_float2(0.f, 0.f);
    } else {
      const float f = 1.0f - thr / cc;
      X[i] = make_float2(v.x * f, v.y * f);
    }
  }
}

// Cycle-spin shifts of FISTA iteration (*counter - 1): entries
// 3*spins*it .. 3*spins*(it+1) of the splitmix64 stream from `seed`, each
// (z >> 33) % 4 (the reference draws rand() % wN, wavelet3D.cpp:443-449).
__global__ void k_spin_shifts(int *sh, int count, unsigned long long seed, const int *counter) {
  const int q = threadIdx.x;
  if (q >= count) return;
  const long long it = counter ? (long long)(*counter - 1) : 0;
  unsigned long long z = seed + (unsigned long long)(it * count + q + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  sh[q] = (int)((z >> 33) % 4ull);
}

// ------------------------------------------------------------ host side
void wavelet_levels(const int N[3], const int req[3], int L[3]) {
  for (int a = 0; a < 3; a++) {
    int l = req[a] < 0 ? 0 : req[a];
    while (l > 0 && (N[a] % (1 << l)) != 0) l--;
    L[a] = l;
  }
}

int ensure_l1_state(mrn_op *op) {
  if (!op->dthr) MRN_HIP(op, hipMalloc(&op->dthr, sizeof(ThrState)));
  if (!op->dsh) MRN_HIP(op, hipMalloc(&op->dsh, sizeof(int) * 3 * MRN_MAX_SPINS));
  if (!op->dthr_log) MRN_HIP(op, hipMalloc(&op->dthr_log, sizeof(double) * MRN_TRACE_CAP));
  return MRN_OK;
}

static int wave_axis(mrn_op *op, float2 *X, int frames, const int s[3], int axis, const int *sh, bool inv,
                     hipStream_t st) {
  static const D4 f = d4_filters();
  const int *N = op->g.N;
  const int pts = s[axis];
  const int b = axis == 0 ? 1 : 0, c = axis == 2 ? 1 : 2;
  const int nlines = s[b] * s[c];
  dim3 g3((unsigned)((nlines + WL_LINES - 1) / WL_LINES), (unsigned)frames);
  const size_t shm = sizeof(float2) * WL_LINES * (size_t)pts;
  if (shm + sizeof(int64_t) * WL_LINES > 160 * 1024)  // + the kernel's line bases
    return fail(op, MRN_EUNSUPPORTED, "wavelet line too long");
  if (inv)
    hipLaunchKernelGGL(k_wave_lines<1>, g3, dim3(256), shm, st, X, N[0], N[1], s[0], s[1], s[2], axis, sh, f,
                       op->nvox);
  else
    hipLaunchKernelGGL(k_wave_lines<0>, g3, dim3(256), shm, st, X, N[0], N[1], s[0], s[1], s[2], axis, sh, f,
                       op->nvox);
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

// forward3D / backward3D of `frames` images, shifts sh[0..2] in device memory
int wavelet3d(mrn_op *op, float2 *X, int frames, const int req[3], const int *sh, bool inv, hipStream_t st) {
  const int *N = op->g.N;
  int L[3];
  wavelet_levels(N, req, L);
  int maxl = 0;
  for (int a = 0; a < 3; a++) maxl = L[a] > maxl ? L[a] : maxl;
  if (!inv) {
    int s[3] = {N[0], N[1], N[2]};
    for (int l = 0; l < maxl; l++) {
      for (int a = 0; a < 3; a++)
        if (l < L[a]) MRN_TRY(wave_axis(op, X, frames, s, a, sh, false, st));
      for (int a = 0; a < 3; a++)
        if (l < L[a]) s[a] = N[a] >> (l + 1);
    }
  } else {
    int s[3];
    for (int a = 0; a < 3; a++) s[a] = N[a] >> L[a];
    for (int l = maxl - 1; l >= 0; l--) {
      for (int a = 0; a < 3; a++)
        if (l < L[a]) s[a] = N[a] >> l;
      for (int a = 2; a >= 0; a--)
        if (l < L[a]) MRN_TRY(wave_axis(op, X, frames, s, a, sh, true, st));
    }
  }
  return MRN_OK;
}

static unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  return (unsigned)(b > 0 ? b : 1);
}

// fraction threshold of n coefficients into ThrState.thr (scaled), device only
static int fraction_threshold(mrn_op *op, const float2 *X, int64_t n, float fraction, float scale,
                              const int *log_counter, hipStream_t st) {
  ThrState *s = (ThrState *)op->dthr;
  hipLaunchKernelGGL(k_thr_reset, dim3(1), dim3(1), 0, st, s);
  hipLaunchKernelGGL(k_thr_minmax, dim3(grid_for(n) < 1024 ? grid_for(n) : 1024), dim3(256), 0, st, X, n, s);
  hipLaunchKernelGGL(k_thr_init, dim3(1), dim3(1), 0, st, s, (long long)n, fraction);
  for (int r = 0; r < (64 + THR_DEPTH - 1) / THR_DEPTH; r++) {
    hipLaunchKernelGGL(k_thr_count, dim3(grid_for(n)), dim3(256), 0, st, X, n, s);
    hipLaunchKernelGGL(k_thr_step, dim3(1), dim3(1), 0, st, s);
  }
  hipLaunchKernelGGL(k_thr_final, dim3(1), dim3(1), 0, st, s, scale, log_counter ? op->dthr_log : nullptr,
                     log_counter, MRN_TRACE_CAP);
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

// The L1 step over `frames` images; shifts of spin q at dsh[3q..3q+2].
int l1_wavelet_step(mrn_op *op, float2 *X, int frames, float fraction, const int levels[3], int spins,
                    const int *log_counter, hipStream_t st) {
  const int64_t n = (int64_t)frames * op->nvox;
  for (int q = 0; q < spins; q++) {
    const int *sh = op->dsh + 3 * q;
    MRN_TRY(wavelet3d(op, X, frames, levels, sh, false, st));
    if (q == 0) MRN_TRY(fraction_threshold(op, X, n, fraction, 1.0f / (float)spins, log_counter, st));
    hipLaunchKernelGGL(k_soft, dim3(grid_for(n)), dim3(256), 0, st, X, n, (const ThrState *)op->dthr);
    MRN_HIP(op, hipGetLastError());
    MRN_TRY(wavelet3d(op, X, frames, levels, sh, true, st));
  }
  return MRN_OK;
}

// FISTA's step: shifts for the current iteration from the device counter
int l1_fista_step(mrn_op *op, float2 *X, hipStream_t st) {
  MRN_TRY(ensure_l1_state(op));
  const int count = 3 * op->l1_spins;
  hipLaunchKernelGGL(k_spin_shifts, dim3(1), dim3(3 * MRN_MAX_SPINS), 0, st, op->dsh, count,
                     (unsigned long long)op->l1_seed, op->dcounter);
  MRN_HIP(op, hipGetLastError());
  return l1_wavelet_step(op, X, op->F, op->l1_fraction, op->l1_levels, op->l1_spins, op->dcounter, st);
}

int l1_wavelet_host(mrn_op *op, float2 *X, int frames, float fraction, const int levels[3], int spins,
                    const int32_t *shifts, double *thr, hipStream_t st) {
  MRN_TRY(ensure_l1_state(op));
  if (shifts) {
    MRN_HIP(op, hipMemcpyAsync(op->dsh, shifts, sizeof(int) * 3 * spins, hipMemcpyHostToDevice, st));
  } else {
    hipLaunchKernelGGL(k_spin_shifts, dim3(1), dim3(3 * MRN_MAX_SPINS), 0, st, op->dsh, 3 * spins,
                       (unsigned long long)op->l1_seed, nullptr);
  }
  MRN_TRY(l1_wavelet_step(op, X, frames, fraction, levels, spins, nullptr, st));
  if (thr) {
    ThrState h;
    MRN_HIP(op, hipMemcpyAsync(&h, op->dthr, sizeof(ThrState), hipMemcpyDeviceToHost, st));
    MRN_HIP(op, hipStreamSynchronize(st));
    *thr = (double)h.thr;
  }
  return MRN_OK;
}

int wavelet_host(mrn_op *op, float2 *X, int frames, const int levels[3], const int32_t shift[3], bool inv,
                 hipStream_t st) {
  MRN_TRY(ensure_l1_state(op));
  MRN_HIP(op, hipMemcpyAsync(op->dsh, shift, sizeof(int) * 3, hipMemcpyHostToDevice, st));
  return wavelet3d(op, X, frames, levels, op->dsh, inv, st);
}

}  // namespace mrn
