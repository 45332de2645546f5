// Elementwise and reduction kernels: layout changes at the boundary, the
// plain form's set_image / accumulate, and the CG / FISTA vector updates.
//
// Solver scalars never leave the device: inner products are reduced in FP64
// (FP32 products accumulated in double, as recon_lib.cpp:1671-1704 does with
// complex<double> per-slice sums), step sizes are computed by one-block
// kernels and read by the update kernels from device memory, so a whole
// iteration can be captured in a hipGraph.
#include "mrn_internal.hpp"

namespace mrn {

#define RED_THREADS 256
#define RED_BLOCKS 1024

// device scalar slots (double dscal[...])
enum { D_RR = 0, D_PR = 1, D_PI = 2, D_RN = 3, D_EIG = 4 };
// float slots after the doubles: float* f = (float*)(dscal + 8)
enum { F_AR = 0, F_AI = 1, F_BETA = 2, F_SCALE = 3, F_STEP = 4, F_SR = 5, F_SI = 6 };

__global__ void k_zero2(float2 *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_float2(0.f, 0.f);
}

__global__ void k_fill2(float2 *p, int64_t n, float2 v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_copy2(float2 *dst, const float2 *src, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// caller coil maps [C][z][y][x] -> coil blocks [nch][y][x][z][CC] (zero
// padded).  z inside y and x: a z pass reads a voxel column's maps as one
// contiguous run instead of one 8*CC-byte piece per z plane (and page).
__global__ void k_smaps_in(float2 *dst, const float2 *src, int64_t nvox, int Nx, int Ny, int Nz, int C,
                           int CC, int lgCC, int64_t total) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int cc = (int)(idx & (CC - 1));
    int64_t q = idx >> lgCC;
    const int z = (int)(q % Nz);
    q /= Nz;
    const int x = (int)(q % Nx);
    q /= Nx;
    const int y = (int)(q % Ny);
    const int ch = (int)(q / Ny);
    const int c = ch * CC + cc;
    dst[idx] = c < C ? src[(int64_t)c * nvox + ((int64_t)z * Ny + y) * Nx + x] : make_float2(0.f, 0.f);
  }
}

// caller k-space [C][M] -> internal [nch][Mv][CC] in tile order
__global__ void k_permute_in(float2 *kint, const float2 *data, const int32_t *perm, int64_t M,
                             int64_t Mv, int C, int CC, int lgCC, int64_t total) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int cc = (int)(idx & (CC - 1));
    const int64_t i = (idx >> lgCC) % Mv;
    const int ch = (int)((idx >> lgCC) / Mv);
    const int c = ch * CC + cc;
    kint[idx] = c < C ? data[(int64_t)c * M + perm[i]] : make_float2(0.f, 0.f);
  }
}

// internal [nch][Mv][CC] -> caller [C][M]; skipped samples give 0 and the
// residual is only taken on valid samples (chop_grid_backward 1196-1319)
__global__ void k_permute_out(float2 *data, const float2 *kint, const int32_t *inv,
                              const float2 *sub, int64_t M, int64_t Mv, int C, int CC, int lgCC) {
  const int64_t total = (int64_t)C * M;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(idx / M);
    const int64_t m = idx - (int64_t)c * M;
    const int i = inv[m];
    float2 v = make_float2(0.f, 0.f);
    if (i >= 0) {
      v = kint[((int64_t)(c >> lgCC) * Mv + i) * CC + (c & (CC - 1))];
      if (sub) {
        const float2 d = sub[idx];
        v.x -= d.x;
        v.y -= d.y;
      }
    }
    data[idx] = v;
  }
}

// The same two permutations with one lane per (caller sample m, coil block
// ch): the block's CC coils are read from (or written to) C rows at
// consecutive samples, so the caller side is coalesced, and the internal side
// moves one whole 8*CC-byte record per lane.  Every internal record belongs
// to exactly one valid caller sample (inv is a bijection onto [0, Mv)).
// W: the scatter's input weighting (d * kw) * taper of k_weight, on the way
template <int CC, bool W>
__global__ void k_permute_in_rec(float2 *__restrict__ kint, const float2 *__restrict__ data,
                                 const int32_t *__restrict__ inv, int64_t M, int64_t Mv, int C,
                                 const float4 *__restrict__ coords, const float *__restrict__ taper) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int ch = blockIdx.y;
  if (m >= M) return;
  const int i = inv[m];
  if (i < 0) return;
  float w = 1.f, tp = 1.f;
  if constexpr (W) {
    w = coords[i].w;
    tp = taper[i];
  }
  float2 v[CC];
#pragma unroll
  for (int cc = 0; cc < CC; cc++) {
    const int c = ch * CC + cc;
    v[cc] = c < C ? data[(int64_t)c * M + m] : make_float2(0.f, 0.f);
    if constexpr (W) v[cc] = make_float2((v[cc].x * w) * tp, (v[cc].y * w) * tp);
  }
  float2 *dst = kint + ((int64_t)ch * Mv + i) * CC;
  if constexpr (CC == 1) {
    dst[0] = v[0];
  } else {
#pragma unroll
    for (int cc = 0; cc < CC; cc += 2)
      *reinterpret_cast<float4 *>(dst + cc) = make_float4(v[cc].x, v[cc].y, v[cc + 1].x, v[cc + 1].y);
  }
}

template <int CC>
__global__ void k_permute_out_rec(float2 *__restrict__ data, const float2 *__restrict__ kint,
                                  const int32_t *__restrict__ inv, const float2 *__restrict__ sub, int64_t M,
                                  int64_t Mv, int C) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int ch = blockIdx.y;
  if (m >= M) return;
  const int i = inv[m];
  float2 v[CC];
  if (i >= 0) {
    const float2 *src = kint + ((int64_t)ch * Mv + i) * CC;
    if constexpr (CC == 1) {
      v[0] = src[0];
    } else {
#pragma unroll
      for (int cc = 0; cc < CC; cc += 2) {
        const float4 q = *reinterpret_cast<const float4 *>(src + cc);
        v[cc] = make_float2(q.x, q.y);
        v[cc + 1] = make_float2(q.z, q.w);
      }
    }
  } else {
#pragma unroll
    for (int cc = 0; cc < CC; cc++) v[cc] = make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int cc = 0; cc < CC; cc++) {
    const int c = ch * CC + cc;
    if (c < C) {
      float2 w = v[cc];
      if (sub && i >= 0) {
        const float2 d = sub[(int64_t)c * M + m];
        w.x -= d.x;
        w.y -= d.y;
      }
      data[(int64_t)c * M + m] = w;
    }
  }
}

// plain form: grid[og + r] = (w(r) * X(r)) * smap(r)  (grid pre-zeroed)
__global__ void k_set_image_plain(float2 *grid, const float2 *X, const float2 *smap,
                                  const float *wx, const float *wy, const float *wz,
                                  PrePostArgs a) {
  const int64_t n = (int64_t)a.N[0] * a.N[1] * a.N[2];
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(r % a.N[0]), j = (int)((r / a.N[0]) % a.N[1]), k = (int)(r / ((int64_t)a.N[0] * a.N[1]));
    const float wt = (wz[k + a.og[2]] * wy[j + a.og[1]]) * wx[i + a.og[0]];
    float2 v = make_float2(wt * X[r].x, wt * X[r].y);
    if (smap) v = cmul(v, smap[r]);
    grid[((int64_t)(k + a.og[2]) * a.S[1] + j + a.og[1]) * a.S[0] + i + a.og[0]] = v;
  }
}

// plain form: X(r) += (w(r) * img(og + r)) * conj(smap(r)) (gridFFT::accumulate,
// gridFFT.cpp:926-958), or with sos != 0 X(r) += |w(r) * img(og + r)|^2
// (deapp + accumulate_sos, gridFFT.cpp:912-924, 960-970)
__global__ void k_accumulate_plain(const float2 *grid, float2 *X, const float2 *smap,
                                   const float *wx, const float *wy, const float *wz,
                                   PrePostArgs a, int sos) {
  const int64_t n = (int64_t)a.N[0] * a.N[1] * a.N[2];
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(r % a.N[0]), j = (int)((r / a.N[0]) % a.N[1]), k = (int)(r / ((int64_t)a.N[0] * a.N[1]));
    const float wt = (wz[k + a.og[2]] * wy[j + a.og[1]]) * wx[i + a.og[0]];
    const float2 im = grid[((int64_t)(k + a.og[2]) * a.S[1] + j + a.og[1]) * a.S[0] + i + a.og[0]];
    float2 v = make_float2(wt * im.x, wt * im.y);
    if (sos) {
      v = make_float2(v.x * v.x + v.y * v.y, 0.0f);
    } else if (smap) {
      v = cmulc(v, smap[r]);
    }
    X[r] = make_float2(X[r].x + v.x, X[r].y + v.y);
  }
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// partials of sum|R|^2 and sum conj(P) * L (3 doubles per block)
__global__ __launch_bounds__(RED_THREADS) void k_dot_rr_pl(const float2 *R, const float2 *P,
                                                           const float2 *L, int64_t n,
                                                           double *part) {
  __shared__ double sh[RED_THREADS];
  double rr = 0, pr = 0, pi = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 r = R[i], p = P[i], l = L[i];
    rr += (double)(r.x * r.x + r.y * r.y);
    pr += (double)(p.x * l.x + p.y * l.y);
    pi += (double)(p.x * l.y - p.y * l.x);
  }
  rr = block_sum(rr, sh);
  pr = block_sum(pr, sh);
  pi = block_sum(pi, sh);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = rr;
    part[3 * blockIdx.x + 1] = pr;
    part[3 * blockIdx.x + 2] = pi;
  }
}

// alpha = sum|R|^2 / sum conj(P) L  (complex<double>, then complex<float>)
__global__ __launch_bounds__(RED_THREADS) void k_cg_alpha(const double *part, int nb,
                                                          double *dscal, double *trace,
                                                          const int *counter) {
  __shared__ double sh[RED_THREADS];
  double rr = 0, pr = 0, pi = 0;
  for (int b = threadIdx.x; b < nb; b += RED_THREADS) {
    rr += part[3 * b];
    pr += part[3 * b + 1];
    pi += part[3 * b + 2];
  }
  rr = block_sum(rr, sh);
  pr = block_sum(pr, sh);
  pi = block_sum(pi, sh);
  if (threadIdx.x == 0) {
    const double den = pr * pr + pi * pi;
    const double ar = rr * pr / den, ai = -rr * pi / den;
    dscal[D_RR] = rr;
    float *f = (float *)(dscal + 8);
    f[F_AR] = (float)ar;
    f[F_AI] = (float)ai;
    if (trace) {
      const int it = (*counter) % MRN_TRACE_CAP;
      trace[4 * it + 0] = rr;
      trace[4 * it + 1] = ar;
      trace[4 * it + 2] = ai;
    }
  }
}

// X += alpha P ; R -= alpha L ; partial sum|R|^2
__global__ __launch_bounds__(RED_THREADS) void k_cg_update(float2 *X, float2 *R, const float2 *P,
                                                           const float2 *L, int64_t n,
                                                           const double *dscal, double *part) {
  __shared__ double sh[RED_THREADS];
  const float *f = (const float *)(dscal + 8);
  const float ar = f[F_AR], ai = f[F_AI];
  double rn = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 p = P[i], l = L[i];
    float2 x = X[i], r = R[i];
    x.x += ar * p.x - ai * p.y;
    x.y += ar * p.y + ai * p.x;
    r.x -= ar * l.x - ai * l.y;
    r.y -= ar * l.y + ai * l.x;
    X[i] = x;
    R[i] = r;
    rn += (double)(r.x * r.x + r.y * r.y);
  }
  rn = block_sum(rn, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = rn;
}

// beta = sum|R_new|^2 / sum|R|^2 ; advances the iteration counter
__global__ __launch_bounds__(RED_THREADS) void k_cg_beta(const double *part, int nb, double *dscal,
                                                         double *trace, int *counter) {
  __shared__ double sh[RED_THREADS];
  double rn = 0;
  for (int b = threadIdx.x; b < nb; b += RED_THREADS) rn += part[b];
  rn = block_sum(rn, sh);
  if (threadIdx.x == 0) {
    dscal[D_RN] = rn;
    float *f = (float *)(dscal + 8);
    f[F_BETA] = (float)(rn / dscal[D_RR]);
    if (trace) trace[4 * ((*counter) % MRN_TRACE_CAP) + 3] = rn;
    *counter += 1;
  }
}

// P = R + beta P
__global__ void k_cg_p(float2 *P, const float2 *R, int64_t n, const double *dscal) {
  const float be = ((const float *)(dscal + 8))[F_BETA];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 r = R[i], p = P[i];
    P[i] = make_float2(r.x + be * p.x, r.y + be * p.y);
  }
}

// partial sum|V|^2
__global__ __launch_bounds__(RED_THREADS) void k_energy(const float2 *V, int64_t n, double *part) {
  __shared__ double sh[RED_THREADS];
  double e = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = V[i];
    e += (double)(v.x * v.x + v.y * v.y);
  }
  e = block_sum(e, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = e;
}

// power method: max_eig = sqrt(sum), scale = float(1/max_eig); optional trace
__global__ __launch_bounds__(RED_THREADS) void k_eig_finish(const double *part, int nb,
                                                            double *dscal, double *trace,
                                                            int *counter) {
  __shared__ double sh[RED_THREADS];
  double e = 0;
  for (int b = threadIdx.x; b < nb; b += RED_THREADS) e += part[b];
  e = block_sum(e, sh);
  if (threadIdx.x == 0) {
    const double me = sqrt(e);
    dscal[D_EIG] = me;
    ((float *)(dscal + 8))[F_SCALE] = (float)(1. / me);
    if (trace) trace[(*counter) % MRN_TRACE_CAP] = me;
    *counter += 1;
  }
}

// residual energy of a FISTA iteration into the trace; advances the counter
__global__ __launch_bounds__(RED_THREADS) void k_energy_finish(const double *part, int nb,
                                                               double *trace, int *counter) {
  __shared__ double sh[RED_THREADS];
  double e = 0;
  for (int b = threadIdx.x; b < nb; b += RED_THREADS) e += part[b];
  e = block_sum(e, sh);
  if (threadIdx.x == 0) {
    if (trace) trace[(*counter) % MRN_TRACE_CAP] = e;
    *counter += 1;
  }
}

// X = scale * R with the device scale (power method)
__global__ void k_scale_by(float2 *X, const float2 *R, int64_t n, const double *dscal) {
  const float s = ((const float *)(dscal + 8))[F_SCALE];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 r = R[i];
    X[i] = make_float2(s * r.x, s * r.y);
  }
}

// fista_update (recon_lib.cpp:2177-2212): X = A X + B X_old ; X_old = X
__global__ void k_fista_momentum(float2 *X, float2 *Xold, int64_t n, const int *counter) {
  const int it = *counter;
  const float A = (float)(1.0 + (it - 1.0) / (it + 1.0));
  const float B = (float)(-(it - 1.0) / (it + 1.0));
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 x0 = Xold[i], x1 = X[i];
    X[i] = make_float2(A * x1.x + B * x0.x, A * x1.y + B * x0.y);
    Xold[i] = x1;
  }
}

// R *= step ; X -= R  (recon_lib.cpp:2100-2105)
__global__ void k_fista_step(float2 *X, float2 *R, int64_t n, float step) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float2 r = R[i], x = X[i];
    r = make_float2(step * r.x, step * r.y);
    R[i] = r;
    X[i] = make_float2(x.x - r.x, x.y - r.y);
  }
}

// Cauchy step of the gradient iteration (recon_lib.cpp:2013-2041,
// 2093-2100): while the iteration is below update_number, step =
// complex<float>(sum|R|^2 / sum conj(R) E'E R) with both sums in complex
// double; afterwards the last step stays.  Logged per iteration.
__global__ __launch_bounds__(RED_THREADS) void k_cauchy_step(const double *part, int nb, double *dscal,
                                                             const int *counter, int update_number,
                                                             double *log) {
  __shared__ double sh[RED_THREADS];
  double rr = 0, pr = 0, pi = 0;
  for (int b = threadIdx.x; b < nb; b += RED_THREADS) {
    rr += part[3 * b];
    pr += part[3 * b + 1];
    pi += part[3 * b + 2];
  }
  rr = block_sum(rr, sh);
  pr = block_sum(pr, sh);
  pi = block_sum(pi, sh);
  if (threadIdx.x == 0) {
    const int it = *counter;
    float *f = (float *)(dscal + 8);
    if (it < update_number) {
      const double den = pr * pr + pi * pi;
      f[F_SR] = (float)(rr * pr / den);
      f[F_SI] = (float)(-rr * pi / den);
    }
    if (log) {
      log[2 * (it % MRN_TRACE_CAP)] = f[F_SR];
      log[2 * (it % MRN_TRACE_CAP) + 1] = f[F_SI];
    }
  }
}

// R *= step (complex<float>, from the device) ; X -= R
__global__ void k_fista_step_c(float2 *X, float2 *R, int64_t n, const double *dscal) {
  const float *f = (const float *)(dscal + 8);
  const float sr = f[F_SR], si = f[F_SI];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float2 r = R[i], x = X[i];
    r = make_float2(r.x * sr - r.y * si, r.x * si + r.y * sr);
    R[i] = r;
    X[i] = make_float2(x.x - r.x, x.y - r.y);
  }
}

__global__ void k_set_step(double *dscal, float sr, float si) {
  float *f = (float *)(dscal + 8);
  f[F_SR] = sr;
  f[F_SI] = si;
}

// ------------------------------------------------------------ launchers
static inline unsigned gs(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b > 0 ? b : 1);
}
static inline unsigned rb(int64_t n) {
  int64_t b = (n + RED_THREADS - 1) / RED_THREADS;
  if (b > RED_BLOCKS) b = RED_BLOCKS;
  return (unsigned)(b > 0 ? b : 1);
}

#define LAUNCH_CHECK(op) MRN_HIP(op, hipGetLastError())

int op_zero(mrn_op *op, float2 *p, int64_t n, hipStream_t st) {
  if (n <= 0) return MRN_OK;
  hipLaunchKernelGGL(k_zero2, dim3(gs(n)), dim3(256), 0, st, p, n);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
int op_fill(mrn_op *op, float2 *p, int64_t n, float2 v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill2, dim3(gs(n)), dim3(256), 0, st, p, n, v);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
int op_copy(mrn_op *op, float2 *d, const float2 *s, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_copy2, dim3(gs(n)), dim3(256), 0, st, d, s, n);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
int op_smaps_in(mrn_op *op, float2 *dst, const float2 *src, hipStream_t st) {
  const int64_t total = (int64_t)op->nch * op->nvox * op->g.CC;
  hipLaunchKernelGGL(k_smaps_in, dim3(gs(total)), dim3(256), 0, st, dst, src, op->nvox, op->g.N[0],
                     op->g.N[1], op->g.N[2], op->C, op->g.CC, op->g.lgCC, total);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
// The record kernels need 16-byte aligned records (CC >= 2) and one grid row
// per coil block; anything else takes the element-wise kernels.
static bool permute_by_record(const mrn_op *op, const void *kint, const FrameData &fr) {
  return fr.M > 0 && op->nch <= 65535 && (fr.M + 255) / 256 <= INT32_MAX &&
         (op->g.CC == 1 || ((uintptr_t)kint & 15) == 0);
}
static int permute_in_impl(mrn_op *op, float2 *kint, const float2 *data, const FrameData &fr, hipStream_t st,
                           bool weighted) {
  const int64_t total = (int64_t)op->nch * fr.Mv * op->g.CC;
  if (total == 0) return MRN_OK;
  if (permute_by_record(op, kint, fr)) {
    const dim3 grid((unsigned)((fr.M + 255) / 256), (unsigned)op->nch);
    switch (op->g.CC * 2 + (weighted ? 1 : 0)) {
#define MRN_PIN(n, w)                                                                                  \
  case n * 2 + w:                                                                                      \
    hipLaunchKernelGGL((k_permute_in_rec<n, w>), grid, dim3(256), 0, st, kint, data, fr.inv, fr.M, fr.Mv, \
                       op->C, fr.coords, fr.taper);                                                    \
    LAUNCH_CHECK(op);                                                                                  \
    return MRN_OK;
#define MRN_PIN2(n) MRN_PIN(n, 0) MRN_PIN(n, 1)
      MRN_PIN2(1) MRN_PIN2(2) MRN_PIN2(4) MRN_PIN2(8) MRN_PIN2(16)
#undef MRN_PIN2
#undef MRN_PIN
      default:
        break;
    }
  }
  hipLaunchKernelGGL(k_permute_in, dim3(gs(total)), dim3(256), 0, st, kint, data, fr.perm, fr.M,
                     fr.Mv, op->C, op->g.CC, op->g.lgCC, total);
  LAUNCH_CHECK(op);
  return weighted ? launch_weight(op, kint, kint, fr, st) : MRN_OK;
}
int op_permute_in(mrn_op *op, float2 *kint, const float2 *data, const FrameData &fr, hipStream_t st) {
  return permute_in_impl(op, kint, data, fr, st, false);
}
// the same with the scatter's input weighting kw * taper applied
int op_permute_in_weighted(mrn_op *op, float2 *kint, const float2 *data, const FrameData &fr, hipStream_t st) {
  return permute_in_impl(op, kint, data, fr, st, true);
}
int op_permute_out(mrn_op *op, float2 *data, const float2 *kint, const FrameData &fr,
                   const float2 *sub, hipStream_t st) {
  if (fr.M == 0) return MRN_OK;
  if (permute_by_record(op, kint, fr)) {
    const dim3 grid((unsigned)((fr.M + 255) / 256), (unsigned)op->nch);
    switch (op->g.CC) {
#define MRN_POUT(n)                                                                                    \
  case n:                                                                                              \
    hipLaunchKernelGGL(k_permute_out_rec<n>, grid, dim3(256), 0, st, data, kint, fr.inv, sub, fr.M, fr.Mv, \
                       op->C);                                                                         \
    LAUNCH_CHECK(op);                                                                                  \
    return MRN_OK;
      MRN_POUT(1) MRN_POUT(2) MRN_POUT(4) MRN_POUT(8) MRN_POUT(16)
#undef MRN_POUT
      default:
        break;
    }
  }
  hipLaunchKernelGGL(k_permute_out, dim3(gs((int64_t)op->C * fr.M)), dim3(256), 0, st, data, kint,
                     fr.inv, sub, fr.M, fr.Mv, op->C, op->g.CC, op->g.lgCC);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
int op_set_image_plain(mrn_op *op, float2 *grid, const float2 *X, const float2 *smap,
                       const PrePostArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_set_image_plain, dim3(gs(op->nvox)), dim3(256), 0, st, grid, X, smap,
                     op->win_d[0], op->win_d[1], op->win_d[2], a);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
int op_accumulate_plain(mrn_op *op, const float2 *grid, float2 *X, const float2 *smap,
                        const PrePostArgs &a, hipStream_t st, int sos) {
  hipLaunchKernelGGL(k_accumulate_plain, dim3(gs(op->nvox)), dim3(256), 0, st, grid, X, smap,
                     op->win_d[0], op->win_d[1], op->win_d[2], a, sos);
  LAUNCH_CHECK(op);
  return MRN_OK;
}

// CG pieces over n = F * nvox elements
int op_cg_alpha(mrn_op *op, const float2 *R, const float2 *P, const float2 *L, int64_t n,
                double *trace, hipStream_t st) {
  const unsigned nb = rb(n);
  hipLaunchKernelGGL(k_dot_rr_pl, dim3(nb), dim3(RED_THREADS), 0, st, R, P, L, n, op->dpart);
  hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(RED_THREADS), 0, st, op->dpart, (int)nb, op->dscal,
                     trace, op->dcounter);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
int op_cg_update(mrn_op *op, float2 *X, float2 *R, float2 *P, const float2 *L, int64_t n,
                 double *trace, hipStream_t st) {
  const unsigned nb = rb(n);
  hipLaunchKernelGGL(k_cg_update, dim3(nb), dim3(RED_THREADS), 0, st, X, R, P, L, n, op->dscal,
                     op->dpart);
  hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(RED_THREADS), 0, st, op->dpart, (int)nb, op->dscal,
                     trace, op->dcounter);
  hipLaunchKernelGGL(k_cg_p, dim3(gs(n)), dim3(256), 0, st, P, R, n, op->dscal);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
int op_power_step(mrn_op *op, float2 *X, const float2 *R, int64_t n, double *trace, hipStream_t st) {
  const unsigned nb = rb(n);
  hipLaunchKernelGGL(k_energy, dim3(nb), dim3(RED_THREADS), 0, st, R, n, op->dpart);
  hipLaunchKernelGGL(k_eig_finish, dim3(1), dim3(RED_THREADS), 0, st, op->dpart, (int)nb, op->dscal,
                     trace, op->dcounter);
  hipLaunchKernelGGL(k_scale_by, dim3(gs(n)), dim3(256), 0, st, X, R, n, op->dscal);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
int op_fista_momentum(mrn_op *op, float2 *X, float2 *Xold, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_fista_momentum, dim3(gs(n)), dim3(256), 0, st, X, Xold, n, op->dcounter);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
// device_step: the complex step in the device scalars (Cauchy), else `step`
int op_fista_finish(mrn_op *op, float2 *X, float2 *R, int64_t n, float step, double *trace,
                    hipStream_t st, bool device_step) {
  const unsigned nb = rb(n);
  hipLaunchKernelGGL(k_energy, dim3(nb), dim3(RED_THREADS), 0, st, R, n, op->dpart);
  hipLaunchKernelGGL(k_energy_finish, dim3(1), dim3(RED_THREADS), 0, st, op->dpart, (int)nb, trace,
                     op->dcounter);
  if (device_step)
    hipLaunchKernelGGL(k_fista_step_c, dim3(gs(n)), dim3(256), 0, st, X, R, n, op->dscal);
  else
    hipLaunchKernelGGL(k_fista_step, dim3(gs(n)), dim3(256), 0, st, X, R, n, step);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
// step from sum|R|^2 and sum conj(R) L, L = E'E R (before op_fista_finish)
int op_cauchy_step(mrn_op *op, const float2 *R, const float2 *L, int64_t n, hipStream_t st) {
  const unsigned nb = rb(n);
  hipLaunchKernelGGL(k_dot_rr_pl, dim3(nb), dim3(RED_THREADS), 0, st, R, R, L, n, op->dpart);
  hipLaunchKernelGGL(k_cauchy_step, dim3(1), dim3(RED_THREADS), 0, st, op->dpart, (int)nb, op->dscal,
                     op->dcounter, op->cauchy_updates, op->dstep_log);
  LAUNCH_CHECK(op);
  return MRN_OK;
}
int op_set_step(mrn_op *op, float sr, float si, hipStream_t st) {
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, st, op->dscal, sr, si);
  LAUNCH_CHECK(op);
  return MRN_OK;
}

}  // namespace mrn
