// Trajectory gradient of the operator and the launch-shape autotuner.
//
// Neither has a counterpart in the reference (its trajectory is an input and
// its loops have no tunables).  The gradient differentiates the gather of
// chop_grid_backward (gridFFT.cpp:1276-1313) with respect to the sample
// coordinate: d/dk_a of prod_b w_b(dk_b - l_b) = grid_a * w_a'(dk_a - l_a) *
// prod_{b != a} w_b.  In table mode w' is the slope of the table segment (the
// exact derivative of the piecewise-linear kernel); in exact-KB mode it is the
// analytic derivative of the Kaiser-Bessel function, using I1 from Abramowitz
// & Stegun 9.8.3 / 9.8.4.
#include <chrono>
#include <cmath>
#include <vector>

#include "mrn_internal.hpp"

namespace mrn {
int fwd_grid_pub(mrn_op *op, const float2 *X, const FrameData *fr, hipStream_t st);
int build_bins(mrn_op *, FrameData &, int64_t, const float *, const float *, const float *,
               const float *, hipStream_t);
int op_permute_in(mrn_op *, float2 *, const float2 *, const FrameData &, hipStream_t);
int op_fill(mrn_op *, float2 *, int64_t, float2, hipStream_t);

// I1(z)/z for z >= 0 (A&S 9.8.3 below 3.75, 9.8.4 above)
__device__ __forceinline__ float i1_over_z(float z) {
  if (z < 3.75f) {
    float t = z / 3.75f;
    t *= t;
    return 0.5f + t * (0.87890594f + t * (0.51498869f + t * (0.15084934f + t * (0.02658733f +
                  t * (0.00301532f + t * 0.00032411f)))));
  }
  float u = 3.75f / z;
  float i1 = (expf(z) * rsqrtf(z)) *
             (0.39894228f + u * (-0.03988024f + u * (-0.00362018f + u * (0.00163801f + u * (-0.01031555f +
              u * (0.02282967f + u * (-0.02895312f + u * (0.01787654f - u * 0.00420059f))))))));
  return i1 / z;
}

// d w(dk - l) / d dk, with the same fftshift sign as tap_weight.  A tap is
// used while |dk - l| <= dwin and the kernel drops to zero beyond, so at
// |dk - l| = dwin exactly the derivative is the one from inside the support
// (the limit along which the tap stays in the sum).
__device__ __forceinline__ float tap_dweight(const Geom &g, int a, const float *tab, float dk, int l) {
  if (a == 2 && !g.grid_in[2]) return 0.0f;
  const float u = dk - (float)l;
  const float su = u < 0.f ? -1.f : 1.f;
  float d;
  if (g.kmode == MRN_KMODE_EXACT_KB && g.grid_in[a]) {
    const float aw = g.dwin[a];
    const float x = fabsf(u);
    if (x > aw) return 0.0f;
    const float s = sqrtf(fmaxf(1.0f - (x / aw) * (x / aw), 0.0f));
    const float z = g.beta[a] * s;
    // dK/dx = knorm/I0(beta) * I1(beta s) * beta * ds/dx, ds/dx = -x/(aw^2 s)
    d = -g.knorm[a] / g.i0beta[a] * g.beta[a] * g.beta[a] * i1_over_z(z) * x / (aw * aw);
  } else {
    // slope of the table segment [i0, i0 + 1]; the last segment inside the
    // support is [L - 1, L] (the table holds L + 10 entries, zero beyond L)
    const float del = fabsf(600.0f * u);
    const int i0 = min((int)del, g.tlen[a] - 11);
    const float f0 = (i0 < g.tlen[a]) ? tab[i0] : 0.0f;
    const float f1 = (i0 + 1 < g.tlen[a]) ? tab[i0 + 1] : 0.0f;
    d = 600.0f * (f1 - f0);
  }
  d *= su;
  if (g.fft_in[a] && !(l & 1)) d = -d;
  return d;
}

// grad_k[m][a] += Re sum_c conj(g_cm) dY_cm/dk_a over the CC coils of a block.
// Thread = (sample, NC consecutive coils): one 8 NC-byte load per tap, as in
// the gather.  The kernel weights and their derivatives are evaluated once
// per axis and tap (3 NT of each, in registers), not once per tap of the 3-D
// stencil: in exact-KB mode each is an I0 or I1 evaluation with an
// exponential.  A sample with more than NT taps on an axis (exactly on the
// window edge) takes the loop path.
template <int NT, int NC>
__global__ __launch_bounds__(256, 3) void k_traj_grad(const float2 *__restrict__ grid,
                                                   const float2 *__restrict__ gin,
                                                   const float4 *__restrict__ coords,
                                                   const int32_t *__restrict__ perm, int64_t Mv,
                                                   Geom g, const float *__restrict__ tables,
                                                   int64_t chunk_grid, int64_t chunk_k,
                                                   float *grad_k) {
  extern __shared__ float tab_s[];
  for (int q = threadIdx.x; q < g.ntab; q += blockDim.x) tab_s[q] = tables[q];
  __syncthreads();
  constexpr int LGNC = NC == 8 ? 3 : NC == 4 ? 2 : NC == 2 ? 1 : 0;
  const int lgl = g.lgCC - LGNC;  // log2 of the lanes per sample
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = (int)(gid & ((1 << lgl) - 1)) * NC;
  const int64_t i = gid >> lgl;
  if (i >= Mv) return;
  const int ch = blockIdx.y;
  const float4 k = coords[i];
  const float dk[3] = {k.x, k.y, k.z};
  int s[3], e[3];
  bool wide = false;
  for (int a = 0; a < 3; a++) {
    axis_range(g, a, dk[a], s[a], e[a]);
    wide = wide || e[a] - s[a] + 1 > NT;
  }
  const float2 *G = grid + ch * chunk_grid + c;
  float dr[3][NC], di[3][NC];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int q = 0; q < NC; q++) dr[a][q] = di[a][q] = 0.f;
  auto tap = [&](const float2 *cell, float q0, float q1, float q2) {
    const CoilVec<NC> v = CoilVec<NC>::load(cell);
#pragma unroll
    for (int q = 0; q < NC; q++) {
      dr[0][q] += q0 * v.v[q].x;
      di[0][q] += q0 * v.v[q].y;
      dr[1][q] += q1 * v.v[q].x;
      di[1][q] += q1 * v.v[q].y;
      dr[2][q] += q2 * v.v[q].x;
      di[2][q] += q2 * v.v[q].y;
    }
  };
  if (!wide) {
    float w[3][NT], d[3][NT];
    int o[3][NT];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const bool in = s[a] + t <= e[a];
        const int l = in ? s[a] + t : s[a];
        w[a][t] = in ? tap_weight(g, a, tab_s + g.toff[a], dk[a], l) : 0.0f;
        d[a][t] = in ? tap_dweight(g, a, tab_s + g.toff[a], dk[a], l) : 0.0f;
        o[a][t] = l;
      }
    // rows of NT taps, software-pipelined as in the gather: row r + 1's loads
    // are issued before row r's products (two register buffers, held in that
    // order by the scheduling barriers).  Per row the x sums with the weights
    // and with their derivatives, then one update per axis:
    //   d/dkx += wy wz sum_x dwx G,  d/dky += dwy wz sum_x wx G,
    //   d/dkz += wy dwz sum_x wx G.
    CoilVec<NC> vb[2][NT];
    auto load_row = [&](int r, CoilVec<NC> *v) {
      const float2 *row = G + ((int64_t)o[2][r / NT] * g.S[1] + o[1][r % NT]) * g.S[0] * g.CC;
#pragma unroll
      for (int tx = 0; tx < NT; tx++) v[tx] = CoilVec<NC>::load(row + (int64_t)o[0][tx] * g.CC);
    };
    load_row(0, vb[0]);
#pragma unroll
    for (int r = 0; r < NT * NT; r++) {
      if (r + 1 < NT * NT) load_row(r + 1, vb[(r + 1) & 1]);
      __builtin_amdgcn_sched_barrier(0);
      const int tz = r / NT, ty = r % NT;
      float sdr[NC], sdi[NC], swr[NC], swi[NC];
#pragma unroll
      for (int q = 0; q < NC; q++) sdr[q] = sdi[q] = swr[q] = swi[q] = 0.f;
#pragma unroll
      for (int tx = 0; tx < NT; tx++)
#pragma unroll
        for (int q = 0; q < NC; q++) {
          const float2 v = vb[r & 1][tx].v[q];
          sdr[q] += d[0][tx] * v.x;
          sdi[q] += d[0][tx] * v.y;
          swr[q] += w[0][tx] * v.x;
          swi[q] += w[0][tx] * v.y;
        }
      const float wzy = w[1][ty] * w[2][tz], wzd = d[1][ty] * w[2][tz], dzy = w[1][ty] * d[2][tz];
#pragma unroll
      for (int q = 0; q < NC; q++) {
        dr[0][q] += wzy * sdr[q];
        di[0][q] += wzy * sdi[q];
        dr[1][q] += wzd * swr[q];
        di[1][q] += wzd * swi[q];
        dr[2][q] += dzy * swr[q];
        di[2][q] += dzy * swi[q];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    for (int lz = s[2]; lz <= e[2]; lz++) {
      const float wz = tap_weight(g, 2, tab_s + g.toff[2], dk[2], lz);
      const float dz = tap_dweight(g, 2, tab_s + g.toff[2], dk[2], lz);
      for (int ly = s[1]; ly <= e[1]; ly++) {
        const float wy = tap_weight(g, 1, tab_s + g.toff[1], dk[1], ly);
        const float dy = tap_dweight(g, 1, tab_s + g.toff[1], dk[1], ly);
        const float2 *row = G + ((int64_t)lz * g.S[1] + ly) * g.S[0] * g.CC;
        for (int lx = s[0]; lx <= e[0]; lx++) {
          const float wx = tap_weight(g, 0, tab_s + g.toff[0], dk[0], lx);
          const float dx = tap_dweight(g, 0, tab_s + g.toff[0], dk[0], lx);
          tap(row + (int64_t)lx * g.CC, dx * wy * wz, wx * dy * wz, wx * wy * dz);
        }
      }
    }
  }
  const int m = perm[i];
  for (int a = 0; a < 3; a++) {
    // Re(conj(g) * dY) * d(dk)/dk over this lane's coils
    float val = 0.f;
#pragma unroll
    for (int q = 0; q < NC; q++) {
      const float2 gv = gin[ch * chunk_k + i * g.CC + c + q];
      val += gv.x * dr[a][q] + gv.y * di[a][q];
    }
    val *= g.grid[a];
    // sum the lanes of this sample before one atomic
    for (int off = 1; off < (1 << lgl); off <<= 1) val += __shfl_xor(val, off);
    if (c == 0) atomicAdd(&grad_k[3 * m + a], val);
  }
}

int traj_grad_impl(mrn_op *op, int f, const float2 *X, const float2 *gup, float *grad_k,
                   hipStream_t st, bool grid_ready) {
  FrameData &fr = op->frames[f];
  MRN_HIP(op, hipMemsetAsync(grad_k, 0, sizeof(float) * 3 * fr.M, st));
  if (fr.Mv == 0) return MRN_OK;
  if (!grid_ready) MRN_TRY(fwd_grid_pub(op, X, &fr, st));
  MRN_TRY(op_permute_in(op, op->kwork, gup, fr, st));
  const Geom &g = op->g;
  dim3 g3((unsigned)((fr.Mv * g.CC + 255) / 256), op->nch);
  const int64_t cg = (int64_t)g.S[0] * g.S[1] * g.S[2] * g.CC, ck = fr.Mv * g.CC;
  // coils per lane: 4 (2 for 2-coil blocks, 1 for 1)
  const int nc = g.CC >= 4 ? 4 : g.CC;
  g3.x = (unsigned)((fr.Mv * (g.CC / nc) + 255) / 256);
  switch (taps_main(g) * 16 + nc) {
#define MRN_TG1(n, c)                                                                                        \
  case n * 16 + c:                                                                                          \
    hipLaunchKernelGGL((k_traj_grad<n, c>), g3, dim3(256), sizeof(float) * g.ntab, st, op->grid, op->kwork,  \
                       fr.coords, fr.perm, fr.Mv, g, op->tables_d, cg, ck, grad_k);                          \
    break;
#define MRN_TG(n) MRN_TG1(n, 1) MRN_TG1(n, 2) MRN_TG1(n, 4)
    MRN_TG(1) MRN_TG(2) MRN_TG(3) MRN_TG(4) MRN_TG(5) MRN_TG(6) MRN_TG(7) MRN_TG(8)
#undef MRN_TG
#undef MRN_TG1
    default:
      return fail(op, MRN_EUNSUPPORTED, "kernel width above 8 taps");
  }
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

}  // namespace mrn
