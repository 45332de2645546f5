// LDS-staged mixed-radix pencil FFT on the oversampled grid.
//
// Stands in for fft3 over FFTW 1-D plans (ArrayTemplates.cpp:276-427) as
// called by do_fft / do_ifft (gridFFT.cpp:645-711): unitary per-axis DFT,
// sign -1 forward (image -> k) and +1 inverse, scaled by 1/sqrt(S) per axis.
//
// A workgroup owns FFT_LINES (16) lines of one axis.  The lines are staged in
// LDS (one padded row per line), transformed by an in-place Stockham
// autosort (radix 8/4/2/3/5 butterflies in registers, twiddles from an LDS
// table computed in double on the host) and written back.  Loads skip the
// points known to be zero (pruning) and stores skip the points nobody reads.
// The first forward pass is fused with set_image (deapodisation x coil map x
// zero pad, gridFFT.cpp:972-986) and the last inverse pass with accumulate
// (deapodisation x conj(coil map), summed over the coils of a block,
// gridFFT.cpp:926-941).
#include <climits>

#include "mrn_internal.hpp"

namespace mrn {

__host__ __device__ constexpr int radix_at(int S, int i) {
  int m = S, r = 0;
  for (int k = 0; k <= i; k++) {
    if (m % 8 == 0) r = 8;
    else if (m % 4 == 0) r = 4;
    else if (m % 2 == 0) r = 2;
    else if (m % 3 == 0) r = 3;
    else if (m % 5 == 0) r = 5;
    else return 0;
    m /= r;
  }
  return r;
}

template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {  // (DIR * i) * a
  return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
__device__ __forceinline__ float2 add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 scl(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// Grid and coil-map traffic of the four-step passes: each element is read
// once and written once per pass, far more than the caches hold, so these
// loads and stores are nontemporal (streamed past the caches' keep policy).
typedef float nt_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_nt(const float2 *p) {
  const nt_f2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f2 *>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(float2 *p, float2 v) {
  nt_f2 q;
  q.x = v.x;
  q.y = v.y;
  __builtin_nontemporal_store(q, reinterpret_cast<nt_f2 *>(p));
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2 *x);

template <>
__device__ __forceinline__ void dft<2, -1>(float2 *x) {
  float2 a = x[0], b = x[1];
  x[0] = add(a, b);
  x[1] = sub(a, b);
}
template <>
__device__ __forceinline__ void dft<2, 1>(float2 *x) { dft<2, -1>(x); }

template <int DIR>
__device__ __forceinline__ void dft4(float2 &x0, float2 &x1, float2 &x2, float2 &x3) {
  float2 t0 = add(x0, x2), t1 = sub(x0, x2), t2 = add(x1, x3), t3 = mul_i<DIR>(sub(x1, x3));
  x0 = add(t0, t2);
  x2 = sub(t0, t2);
  x1 = add(t1, t3);
  x3 = sub(t1, t3);
}
template <>
__device__ __forceinline__ void dft<4, -1>(float2 *x) { dft4<-1>(x[0], x[1], x[2], x[3]); }
template <>
__device__ __forceinline__ void dft<4, 1>(float2 *x) { dft4<1>(x[0], x[1], x[2], x[3]); }

template <int DIR>
__device__ __forceinline__ void dft8(float2 *x) {
  const float r2 = 0.70710678118654752f;
  float2 e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
  float2 o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  // twiddles W8^k, W8 = exp(DIR * 2 pi i / 8)
  o1 = scl(add(o1, mul_i<DIR>(o1)), r2);
  o2 = mul_i<DIR>(o2);
  o3 = scl(sub(mul_i<DIR>(o3), o3), r2);
  x[0] = add(e0, o0);
  x[4] = sub(e0, o0);
  x[1] = add(e1, o1);
  x[5] = sub(e1, o1);
  x[2] = add(e2, o2);
  x[6] = sub(e2, o2);
  x[3] = add(e3, o3);
  x[7] = sub(e3, o3);
}
template <>
__device__ __forceinline__ void dft<8, -1>(float2 *x) { dft8<-1>(x); }
template <>
__device__ __forceinline__ void dft<8, 1>(float2 *x) { dft8<1>(x); }

template <int DIR>
__device__ __forceinline__ void dft3(float2 *x) {
  const float s = 0.86602540378443865f;
  float2 t = add(x[1], x[2]), u = mul_i<DIR>(sub(x[1], x[2]));
  float2 m = sub(x[0], scl(t, 0.5f));
  x[0] = add(x[0], t);
  u = scl(u, s);
  x[1] = add(m, u);
  x[2] = sub(m, u);
}
template <>
__device__ __forceinline__ void dft<3, -1>(float2 *x) { dft3<-1>(x); }
template <>
__device__ __forceinline__ void dft<3, 1>(float2 *x) { dft3<1>(x); }

template <int DIR>
__device__ __forceinline__ void dft5(float2 *x) {
  const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
  const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
  float2 a1 = add(x[1], x[4]), b1 = sub(x[1], x[4]);
  float2 a2 = add(x[2], x[3]), b2 = sub(x[2], x[3]);
  float2 x0 = x[0];
  float2 m1 = add(x0, add(scl(a1, c1), scl(a2, c2)));
  float2 m2 = add(x0, add(scl(a1, c2), scl(a2, c1)));
  float2 n1 = mul_i<DIR>(add(scl(b1, s1), scl(b2, s2)));
  float2 n2 = mul_i<DIR>(sub(scl(b1, s2), scl(b2, s1)));
  x[0] = add(x0, add(a1, a2));
  x[1] = add(m1, n1);
  x[4] = sub(m1, n1);
  x[2] = add(m2, n2);
  x[3] = sub(m2, n2);
}
template <>
__device__ __forceinline__ void dft<5, -1>(float2 *x) { dft5<-1>(x); }
template <>
__device__ __forceinline__ void dft<5, 1>(float2 *x) { dft5<1>(x); }

// One Stockham stage (radix R, NS = product of earlier radices) over all
// FFT_LINES lines in LDS, in place: every thread reads its butterflies into
// registers, the workgroup synchronises, then the results are written back.
template <int S, int R, int NS, int DIR>
__device__ __forceinline__ void fft_stage(float2 *buf, const float2 *tws) {
  constexpr int NBL = S / R;
  constexpr int NB = FFT_LINES * NBL;
  constexpr int PER = (NB + FFT_THREADS - 1) / FFT_THREADS;
  constexpr int TSTEP = S / (NS * R);
  float2 v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int b = tid + q * FFT_THREADS;
    if ((NB % FFT_THREADS) == 0 || b < NB) {
      const int line = b / NBL, j = b - line * NBL;
      const float2 *ln = buf + line * (S + FFT_PAD);
#pragma unroll
      for (int k = 0; k < R; k++) v[q][k] = ln[j + k * NBL];
      if (NS > 1) {
        const int jm = j % NS;
#pragma unroll
        for (int k = 1; k < R; k++) {
          float2 t = tws[jm * k * TSTEP];
          if (DIR > 0) t.y = -t.y;
          v[q][k] = cmul(v[q][k], t);
        }
      }
      dft<R, DIR>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int b = tid + q * FFT_THREADS;
    if ((NB % FFT_THREADS) == 0 || b < NB) {
      const int line = b / NBL, j = b - line * NBL;
      float2 *ln = buf + line * (S + FFT_PAD);
      const int idxD = (j / NS) * NS * R + (j % NS);
#pragma unroll
      for (int k = 0; k < R; k++) ln[idxD + k * NS] = v[q][k];
    }
  }
  __syncthreads();
}

template <int S, int DIR, int STAGE, int NS>
__device__ __forceinline__ void fft_stages(float2 *buf, const float2 *tws) {
  constexpr int R = radix_at(S, STAGE);
  if constexpr (R != 0) {
    fft_stage<S, R, NS, DIR>(buf, tws);
    fft_stages<S, DIR, STAGE + 1, NS * R>(buf, tws);
  }
}

template <int S>
__device__ __forceinline__ void load_twiddles(float2 *tws, const float2 *tw) {
  for (int q = threadIdx.x; q < S; q += FFT_THREADS) tws[q] = tw[q];
}

// Zero the points of every line outside [lo, hi).
template <int S>
__device__ __forceinline__ void zero_outside(float2 *buf, int lo, int hi) {
  const int nz = S - (hi - lo);
  if (nz <= 0) return;
  for (int idx = threadIdx.x; idx < FFT_LINES * nz; idx += FFT_THREADS) {
    const int l = idx / nz, q = idx - l * nz;
    const int p = q < lo ? q : q + (hi - lo);
    buf[l * (S + FFT_PAD) + p] = make_float2(0.f, 0.f);
  }
}

// Move points [p0, p0+np) of the workgroup's 16 lines between the grid and LDS
// (STORE = false: grid -> LDS, true: LDS -> grid, scaled).  Three layouts:
//  - run >= 16: the 16 lines are contiguous at every point (y/z passes, or an
//    x pass with 16 coils per cell): lane -> line, 128-byte rows;
//  - run < 16, a power of two dividing 16 (x pass, CC < 16): 16/run rows, each
//    a contiguous run*np block: shifts only;
//  - anything else (tiny plain-form grids): generic index arithmetic.
// In the first two a thread moves at most ceil(S/16) points; it issues all of
// its loads before the first LDS write (and all LDS reads before the first
// store), so that a workgroup keeps up to S*16*8 bytes in flight instead of
// one load per thread.
template <int S, bool STORE>
__device__ __forceinline__ void move_lines(float2 *__restrict__ grid, float2 *buf, const LinesArgs &a,
                                           int64_t base, int g0, int p0, int np) {
  constexpr int NJ = (S + 15) / 16;
  const int tid = threadIdx.x;
  if (a.run >= FFT_LINES) {
    const int l = tid & (FFT_LINES - 1);
    const int gl = g0 + l;
    if (gl >= a.nlines) return;
    const int r = gl / a.run, cc = gl - r * a.run;
    float2 *src = grid + base + (int64_t)r * a.rstride + cc;
    const int pb = p0 + (tid >> 4), pe = p0 + np;
    float2 v[NJ];
    if (STORE) {
#pragma unroll
      for (int j = 0; j < NJ; j++) {
        const int p = pb + 16 * j;
        if (p < pe) v[j] = scl(buf[l * (S + FFT_PAD) + p], a.scale);
      }
#pragma unroll
      for (int j = 0; j < NJ; j++) {
        const int p = pb + 16 * j;
        if (p < pe) src[(int64_t)p * a.pstride] = v[j];
      }
    } else {
#pragma unroll
      for (int j = 0; j < NJ; j++) {
        const int p = pb + 16 * j;
        if (p < pe) v[j] = src[(int64_t)p * a.pstride];
      }
#pragma unroll
      for (int j = 0; j < NJ; j++) {
        const int p = pb + 16 * j;
        if (p < pe) buf[l * (S + FFT_PAD) + p] = v[j];
      }
    }
  } else if ((a.run & (a.run - 1)) == 0 && a.pstride == a.run) {
    // element e = tid + 256 j of the 16 * np moved: row rr = e / (W np), then
    // (point, line-in-row) = (e % (W np)) / W, % W
    const int W = a.run;
    const int lgW = __builtin_ctz(W);
    const int rowlen = W * np;
    float2 v[NJ];
    int off[NJ];  // LDS index, -1 when not moved
    int64_t gidx[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
      const int e = tid + FFT_THREADS * j;
      const int rr = e / rowlen, q = e - rr * rowlen;
      const int c = q & (W - 1), p = p0 + (q >> lgW);
      const bool ok = e < FFT_LINES * np && g0 + rr * W < a.nlines;
      off[j] = ok ? (rr * W + c) * (S + FFT_PAD) + p : -1;
      gidx[j] = base + (int64_t)((g0 >> lgW) + rr) * a.rstride + (int64_t)p * W + c;
    }
    if (STORE) {
#pragma unroll
      for (int j = 0; j < NJ; j++)
        if (off[j] >= 0) v[j] = scl(buf[off[j]], a.scale);
#pragma unroll
      for (int j = 0; j < NJ; j++)
        if (off[j] >= 0) grid[gidx[j]] = v[j];
    } else {
#pragma unroll
      for (int j = 0; j < NJ; j++)
        if (off[j] >= 0) v[j] = grid[gidx[j]];
#pragma unroll
      for (int j = 0; j < NJ; j++)
        if (off[j] >= 0) buf[off[j]] = v[j];
    }
  } else {
    for (int idx = tid; idx < FFT_LINES * np; idx += FFT_THREADS) {
      const int l = idx / np, p = p0 + idx % np;
      const int gl = g0 + l;
      if (gl >= a.nlines) continue;
      const int r = gl / a.run, cc = gl - r * a.run;
      float2 *e = grid + base + (int64_t)r * a.rstride + cc + (int64_t)p * a.pstride;
      if (STORE)
        *e = scl(buf[l * (S + FFT_PAD) + p], a.scale);
      else
        buf[l * (S + FFT_PAD) + p] = *e;
    }
  }
}

// Grid -> grid pass over FFT_LINES lines (see LinesArgs).  Lines past
// nlines are never loaded (their LDS rows hold zeros or stale values, which
// only feed their own, unstored, transforms).
template <int S, int DIR>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_lines(float2 *__restrict__ grid, LinesArgs a,
                                                           const float2 *__restrict__ tw) {
  __shared__ float2 buf[FFT_LINES * (S + FFT_PAD)];
  __shared__ float2 tws[S];
  load_twiddles<S>(tws, tw);
  const int64_t base = a.off0 + (int64_t)blockIdx.y * a.ystride + (int64_t)blockIdx.z * a.zstride;
  const int g0 = blockIdx.x * FFT_LINES;
  zero_outside<S>(buf, a.ld0, a.ld1);
  move_lines<S, false>(grid, buf, a, base, g0, a.ld0, a.ld1 - a.ld0);
  __syncthreads();
  fft_stages<S, DIR, 0, 1>(buf, tws);
  move_lines<S, true>(grid, buf, a, base, g0, a.st0, a.st1 - a.st0);
}

// Fused set_image + forward z pass.  Workgroup: 16 lines (x, coil) of one
// image row y, coil block blockIdx.z.  Loads the crop z range straight from
// the image and the coil maps, writes the whole z line of the grid.
template <int S>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_pre_z(
    float2 *__restrict__ grid, const float2 *__restrict__ X, const float2 *__restrict__ smaps,
    const float *__restrict__ wx, const float *__restrict__ wy, const float *__restrict__ wz,
    PrePostArgs a, const float2 *__restrict__ tw) {
  __shared__ float2 buf[FFT_LINES * (S + FFT_PAD)];
  __shared__ float2 tws[S];
  load_twiddles<S>(tws, tw);
  // frame and coil block of this workgroup
  const int f = blockIdx.z / a.nchf, CC = a.CC, y = blockIdx.y, ch = blockIdx.z - f * a.nchf;
  X += f * a.xfs;
  grid += (int64_t)f * a.nchf * a.chunk_grid;
  // runs of consecutive workgroups (x columns) per XCD: neighbouring 128-byte
  // rows of the grid and of X share an L2
  const int bx = xcd_block(blockIdx.x, gridDim.x, a.xrun);
  const int nl = a.N[0] * CC, g0 = bx * FFT_LINES;
  const int ogz = a.og[2];
  zero_outside<S>(buf, ogz, ogz + a.N[2]);
  const float wyv = wy[a.og[1] + y];
  const float2 *sm = smaps + (int64_t)ch * a.chunk_smap;
  constexpr int NJ = (S + 15) / 16;
  // thread: line l = tid % 16, z points z0 + 16 j; all loads first
  const int l = threadIdx.x & (FFT_LINES - 1), z0 = threadIdx.x >> 4;
  const int gl = g0 + l;
  const bool lok = gl < nl;
  const int xi = gl >> a.lgCC;
  const float wxv = lok ? wx[a.og[0] + xi] : 0.f;
  float2 xv[NJ], sv[NJ];
  float wzv[NJ];
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    const int zi = z0 + 16 * j;
    if (lok && zi < a.N[2]) {
      const int64_t r = ((int64_t)zi * a.N[1] + y) * a.N[0];
      xv[j] = X[r + xi];
      sv[j] = sm[(((int64_t)y * a.N[0] + xi) * a.N[2] + zi) * CC + (gl & (CC - 1))];
      wzv[j] = wz[ogz + zi];
    }
  }
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    const int zi = z0 + 16 * j;
    if (zi < a.N[2]) {
      float2 v = make_float2(0.f, 0.f);
      if (lok) {
        const float wt = (wzv[j] * wyv) * wxv;
        v = cmul(make_float2(wt * xv[j].x, wt * xv[j].y), sv[j]);
      }
      buf[l * (S + FFT_PAD) + ogz + zi] = v;
    }
  }
  __syncthreads();
  fft_stages<S, -1, 0, 1>(buf, tws);
  if (lok) {
    float2 *gch = grid + (int64_t)ch * a.chunk_grid + ((int64_t)(a.og[1] + y) * a.S[0] + a.og[0]) * CC + gl;
    const int64_t pstride = (int64_t)a.S[1] * a.S[0] * CC;
    float2 v[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
      const int p = z0 + 16 * j;
      if (p < S) v[j] = scl(buf[l * (S + FFT_PAD) + p], a.scale);
    }
#pragma unroll
    for (int j = 0; j < NJ; j++) {
      const int p = z0 + 16 * j;
      if (p < S) gch[(int64_t)p * pstride] = v[j];
    }
  }
}

// Inverse z pass fused with accumulate: X(x,y,z) (+)= sum_c w * img_c * conj(S_c)
// over the CC coils of block a.ch, for the crop z range.
template <int S>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_post_z(
    const float2 *__restrict__ grid, float2 *__restrict__ X, const float2 *__restrict__ smaps,
    const float *__restrict__ wx, const float *__restrict__ wy, const float *__restrict__ wz,
    PrePostArgs a, const float2 *__restrict__ tw) {
  __shared__ float2 buf[FFT_LINES * (S + FFT_PAD)];
  __shared__ float2 tws[S];
  load_twiddles<S>(tws, tw);
  const int CC = a.CC, y = blockIdx.y, ch = a.ch;
  X += blockIdx.z * a.xfs;  // frame blockIdx.z
  grid += (int64_t)blockIdx.z * a.nchf * a.chunk_grid;
  const int bx = xcd_block(blockIdx.x, gridDim.x, a.xrun);  // as in k_fft_pre_z
  const int nl = a.N[0] * CC, g0 = bx * FFT_LINES;
  constexpr int NJ = (S + 15) / 16;
  const int nxl = FFT_LINES / CC;
  const int ogz = a.og[2];
  const float2 *sm = smaps + (int64_t)ch * a.chunk_smap;
  // Thread = (line l = x-in-group * CC + coil, z0 = tid / 16), output z = z0 + 16 j:
  // the 16 lanes of a z read one contiguous 128-byte run of coil maps (as the
  // forward z pass does), and the CC coil terms of a voxel are summed across
  // lanes.  Coil maps are loaded before the transform.
  const int l = threadIdx.x & (FFT_LINES - 1), z0 = threadIdx.x >> 4;
  const int c = l & (CC - 1), xl = l >> a.lgCC;
  const int xi = bx * nxl + xl;
  const bool xok = xi < a.N[0];
  float2 smv[NJ];
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    const int zi = z0 + 16 * j;
    smv[j] = make_float2(0.f, 0.f);
    if (xok && zi < a.N[2]) smv[j] = sm[(((int64_t)y * a.N[0] + xi) * a.N[2] + zi) * CC + c];
  }
  {
    const int p0 = z0;
    const int gl = g0 + l;
    const float2 *src = grid + (int64_t)ch * a.chunk_grid + ((int64_t)(a.og[1] + y) * a.S[0] + a.og[0]) * CC + gl;
    const int64_t pstride = (int64_t)a.S[1] * a.S[0] * CC;
    float2 v[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
      const int p = p0 + 16 * j;
      v[j] = make_float2(0.f, 0.f);
      if (gl < nl && p < S) v[j] = src[(int64_t)p * pstride];
    }
#pragma unroll
    for (int j = 0; j < NJ; j++) {
      const int p = p0 + 16 * j;
      if (p < S) buf[l * (S + FFT_PAD) + p] = v[j];
    }
  }
  __syncthreads();
  fft_stages<S, 1, 0, 1>(buf, tws);
  const float wyx = wy[a.og[1] + y] * (xok ? wx[a.og[0] + xi] : 0.f);
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    const int zi = z0 + 16 * j;
    const bool zok = zi < a.N[2];
    float2 t = make_float2(0.f, 0.f);
    if (xok && zok) {
      const float wt = wz[ogz + zi] * wyx;
      const float2 im = scl(buf[l * (S + FFT_PAD) + ogz + zi], a.scale);
      t = cmulc(make_float2(wt * im.x, wt * im.y), smv[j]);
    }
    for (int o = 1; o < CC; o <<= 1) {  // sum the CC coils (adjacent lanes)
      t.x += __shfl_xor(t.x, o);
      t.y += __shfl_xor(t.y, o);
    }
    if (xok && zok && c == 0) {
      const int64_t r = ((int64_t)zi * a.N[1] + y) * a.N[0] + xi;
      if (a.overwrite) {
        X[r] = t;
      } else {
        const float2 o = X[r];
        X[r] = make_float2(o.x + t.x, o.y + t.y);
      }
    }
  }
}

// ------------------------------------------------ four-step, register-resident
// For S = 16 P (64 <= S <= 512) a workgroup still owns 16 lines, but thread
// (l, t) = (tid % 16, tid / 16) keeps the P points p = t + 16 j of line l in
// registers, and
//   X[k1 + P k2] = sum_t W_16^(t k2) W_S^(t k1) sum_j x[t + 16 j] W_P^(j k1)
// (W_n = exp(DIR 2 pi i / n)).  Rows: the P-point DFT over j in registers and
// the twiddle W_S^(t k1); one LDS transpose; columns: thread (l, u) takes
// k1 = u, u + 16, ... < P and runs the 16-point DFT over t in registers,
// which gives the points k1 + P k2.  One LDS round trip per pass instead of
// one per radix stage.  Butterflies and twiddles are those of fft_stage.

// LDS row length: a multiple of 2 float2 (16-byte aligned columns), 2 mod 32
// so that the 16 lines of a column read start in distinct banks.
__host__ __device__ constexpr int fft4_ld(int S) { return S + (34 - S % 32) % 32; }
__host__ __device__ constexpr bool fft4_ok(int S) { return S % 16 == 0 && S >= 64 && S <= 512; }
// Line passes whose float2 transpose buffer would hold a CU to fewer than 3
// workgroups (S >= 400) move it as floats, real parts then imaginary parts:
// rows padded to 4 words mod 64 (16-byte aligned columns, the 16 lines'
// column reads starting in distinct banks).
__host__ __device__ constexpr bool fft4_split(int S) {
  return (FFT_LINES * fft4_ld(S) * 8 + S * 8) * 3 > 160 * 1024;
}
__host__ __device__ constexpr int fft4_ldf(int S) { return S + (68 - S % 64) % 64; }

// Stockham stages of a P-point DFT held in registers (natural order in and
// out); W_P^m = tws[m TSC], tws the S-point table.
template <int P, int DIR, int TSC, int STAGE, int NS>
__device__ __forceinline__ void reg_fft(float2 *v, const float2 *tws) {
  constexpr int R = radix_at(P, STAGE);
  if constexpr (R != 0) {
    constexpr int NBL = P / R, TSTEP = P / (NS * R);
    float2 o[P];
#pragma unroll
    for (int j = 0; j < NBL; j++) {
      float2 x[R];
#pragma unroll
      for (int k = 0; k < R; k++) x[k] = v[j + k * NBL];
      const int jm = j % NS;
#pragma unroll
      for (int k = 1; k < R; k++)
        if (jm * k != 0) {
          float2 tw = tws[jm * k * TSTEP * TSC];
          if (DIR > 0) tw.y = -tw.y;
          x[k] = cmul(x[k], tw);
        }
      dft<R, DIR>(x);
      const int d = (j / NS) * NS * R + jm;
#pragma unroll
      for (int k = 0; k < R; k++) o[d + k * NS] = x[k];
    }
#pragma unroll
    for (int q = 0; q < P; q++) v[q] = o[q];
    reg_fft<P, DIR, TSC, STAGE + 1, NS * R>(v, tws);
  }
}

// Rows: DFT over j, twiddle, write column entries row[k1 * 16 + t].
// Rows without the LDS write: DFT over j and twiddle, in place in v.
template <int S, int DIR>
__device__ __forceinline__ void fft4_rows_tw(float2 *v, int t, const float2 *tws) {
  constexpr int P = S / 16;
  reg_fft<P, DIR, 16, 0, 1>(v, tws);
#pragma unroll
  for (int k1 = 1; k1 < P; k1++) {
    float2 tw = tws[t * k1];
    if (DIR > 0) tw.y = -tw.y;
    v[k1] = cmul(v[k1], tw);
  }
}

template <int S, int DIR>
__device__ __forceinline__ void fft4_rows(float2 *v, int t, float2 *row, const float2 *tws) {
  constexpr int P = S / 16;
  reg_fft<P, DIR, 16, 0, 1>(v, tws);
#pragma unroll
  for (int k1 = 0; k1 < P; k1++) {
    float2 x = v[k1];
    if (k1 > 0) {
      float2 tw = tws[t * k1];
      if (DIR > 0) tw.y = -tw.y;
      x = cmul(x, tw);
    }
    row[k1 * 16 + t] = x;
  }
}

// Columns: the 16 entries of column k1 (16-byte aligned), DFT over t.
template <int S, int DIR>
__device__ __forceinline__ void fft4_cols(float2 *w, const float2 *col, const float2 *tws) {
  const float4 *c4 = reinterpret_cast<const float4 *>(col);
#pragma unroll
  for (int n = 0; n < 8; n++) {
    const float4 q = c4[n];
    w[2 * n] = make_float2(q.x, q.y);
    w[2 * n + 1] = make_float2(q.z, q.w);
  }
  reg_fft<16, DIR, S / 16, 0, 1>(w, tws);
}

// Sum over groups of `cc` adjacent lanes (1 < cc <= 16, a power of two) by
// DPP inside rows of 16 lanes; every lane of a group gets the same sum, added
// in the order of the xor butterflies 1, 2, 4, 8.  All lanes of a row must
// be active.
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float2 group_sum(float2 t, int cc) {
  if (cc > 1) t = make_float2(dpp_add<0xB1>(t.x), dpp_add<0xB1>(t.y));    // quad_perm [1,0,3,2]
  if (cc > 2) t = make_float2(dpp_add<0x4E>(t.x), dpp_add<0x4E>(t.y));    // quad_perm [2,3,0,1]
  if (cc > 4) t = make_float2(dpp_add<0x141>(t.x), dpp_add<0x141>(t.y));  // row_half_mirror
  if (cc > 8) t = make_float2(dpp_add<0x140>(t.x), dpp_add<0x140>(t.y));  // row_mirror
  return t;
}

// Grid -> grid pass, lines as in k_fft_lines (any LinesArgs layout).
template <int S, int DIR>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_lines4(float2 *__restrict__ grid, LinesArgs a,
                                                            const float2 *__restrict__ tw) {
  constexpr int P = S / 16;
  constexpr bool SPLIT = fft4_split(S);
  constexpr int LD = SPLIT ? fft4_ldf(S) : fft4_ld(S);  // floats or float2 per line
  __shared__ float4 buf4[SPLIT ? FFT_LINES * LD / 4 : FFT_LINES * LD / 2];
  __shared__ float2 tws[S];
  float2 *buf = reinterpret_cast<float2 *>(buf4);
  const int64_t base = a.off0 + (int64_t)blockIdx.y * a.ystride + (int64_t)blockIdx.z * a.zstride;
  const int l = threadIdx.x & (FFT_LINES - 1), t = threadIdx.x >> 4;
  const int gl = xcd_block(blockIdx.x, gridDim.x, a.xrun) * FFT_LINES + l;
  const bool ok = gl < a.nlines;
  const int r = ok ? gl / a.run : 0, cc = ok ? gl - r * a.run : 0;
  float2 *ln = grid + base + (int64_t)r * a.rstride + cc;
  int ld0 = a.ld0, ld1 = a.ld1, st0 = a.st0, st1 = a.st1;
  if (a.rng && ok) {  // k-space support of this line's row
    const int64_t rf = a.rng_fs ? (int64_t)(blockIdx.z / a.rng_nz) * a.rng_fs : 0;  // frame's support
    const int2 rr = a.rng[rf + (int64_t)blockIdx.y * a.rng_zs + (int64_t)r * a.rng_rs];
    if (a.rng_load) {
      ld0 = max(ld0, rr.x);
      ld1 = min(ld1, rr.y + 1);
      if (rr.x == INT_MAX) st1 = st0;  // no later pass reads this row
    }
    if (a.rng_store) {
      st0 = max(st0, rr.x);
      st1 = min(st1, rr.y + 1);
      if (rr.x > rr.y) ld1 = ld0;  // nothing to store
    }
  }
  float2 v[P];
#pragma unroll
  for (int j = 0; j < P; j++) {
    const int p = t + 16 * j;
    v[j] = make_float2(0.f, 0.f);
    if (ok && p >= ld0 && p < ld1) v[j] = ld_nt(ln + (int64_t)p * a.pstride);
  }
  load_twiddles<S>(tws, tw);
  __syncthreads();
  if constexpr (!SPLIT) {
    fft4_rows<S, DIR>(v, t, buf + l * LD, tws);
    __syncthreads();
    for (int k1 = t; k1 < P; k1 += 16) {
      float2 w[16];
      fft4_cols<S, DIR>(w, buf + l * LD + k1 * 16, tws);
      if (ok) {
#pragma unroll
        for (int k2 = 0; k2 < 16; k2++) {
          const int p = k1 + P * k2;
          if (p >= st0 && p < st1) st_nt(ln + (int64_t)p * a.pstride, scl(w[k2], a.scale));
        }
      }
    }
  } else {
    // real parts through LDS, then imaginary parts; thread (l, u) keeps the
    // real parts of its columns u and u + 16 between the two
    constexpr int NCOL = (P + 15) / 16;
    float *bf = reinterpret_cast<float *>(buf4) + l * LD;
    fft4_rows_tw<S, DIR>(v, t, tws);
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) bf[k1 * 16 + t] = v[k1].x;
    __syncthreads();
    float re[NCOL][16];
#pragma unroll
    for (int c = 0; c < NCOL; c++) {
      const int k1 = t + 16 * c;
      if (k1 < P) {
        const float4 *c4 = reinterpret_cast<const float4 *>(bf + k1 * 16);
#pragma unroll
        for (int n = 0; n < 4; n++) {
          const float4 q = c4[n];
          re[c][4 * n] = q.x;
          re[c][4 * n + 1] = q.y;
          re[c][4 * n + 2] = q.z;
          re[c][4 * n + 3] = q.w;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) bf[k1 * 16 + t] = v[k1].y;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NCOL; c++) {
      const int k1 = t + 16 * c;
      if (k1 < P) {
        float2 w[16];
        const float4 *c4 = reinterpret_cast<const float4 *>(bf + k1 * 16);
#pragma unroll
        for (int n = 0; n < 4; n++) {
          const float4 q = c4[n];
          w[4 * n] = make_float2(re[c][4 * n], q.x);
          w[4 * n + 1] = make_float2(re[c][4 * n + 1], q.y);
          w[4 * n + 2] = make_float2(re[c][4 * n + 2], q.z);
          w[4 * n + 3] = make_float2(re[c][4 * n + 3], q.w);
        }
        reg_fft<16, DIR, S / 16, 0, 1>(w, tws);
        if (ok) {
#pragma unroll
          for (int k2 = 0; k2 < 16; k2++) {
            const int p = k1 + P * k2;
            if (p >= st0 && p < st1) st_nt(ln + (int64_t)p * a.pstride, scl(w[k2], a.scale));
          }
        }
      }
    }
  }
}

// k_fft_pre_z on the four-step transform: same workgroups, same products.
template <int S>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_pre_z4(
    float2 *__restrict__ grid, const float2 *__restrict__ X, const float2 *__restrict__ smaps,
    const float *__restrict__ wx, const float *__restrict__ wy, const float *__restrict__ wz,
    PrePostArgs a, const float2 *__restrict__ tw) {
  constexpr int P = S / 16, LD = fft4_ld(S);
  __shared__ float4 buf4[FFT_LINES * LD / 2];
  __shared__ float2 tws[S];
  __shared__ float wzs[S];
  float2 *buf = reinterpret_cast<float2 *>(buf4);
  // frame and coil block of this workgroup
  const int f = blockIdx.z / a.nchf, CC = a.CC, y = blockIdx.y, ch = blockIdx.z - f * a.nchf;
  X += f * a.xfs;
  grid += (int64_t)f * a.nchf * a.chunk_grid;
  const int bx = xcd_block(blockIdx.x, gridDim.x, a.xrun);
  const int nl = a.N[0] * CC;
  const int ogz = a.og[2];
  const int l = threadIdx.x & (FFT_LINES - 1), t = threadIdx.x >> 4;
  const int gl = bx * FFT_LINES + l;
  const bool lok = gl < nl;
  const int xi = gl >> a.lgCC;
  const float wyv = wy[a.og[1] + y];
  const float wxv = lok ? wx[a.og[0] + xi] : 0.f;
  const float2 *sm = smaps + (int64_t)ch * a.chunk_smap;
  float2 xv[P], sv[P];
#pragma unroll
  for (int j = 0; j < P; j++) {
    const int zi = t + 16 * j - ogz;
    if (lok && zi >= 0 && zi < a.N[2]) {
      const int64_t r = ((int64_t)zi * a.N[1] + y) * a.N[0];
      xv[j] = X[r + xi];
      sv[j] = ld_nt(sm + (((int64_t)y * a.N[0] + xi) * a.N[2] + zi) * CC + (gl & (CC - 1)));
    }
  }
  // the z window in LDS with the twiddles, while the loads are in flight
  for (int q = threadIdx.x; q < S; q += FFT_THREADS) {
    tws[q] = tw[q];
    wzs[q] = wz[q];
  }
  __syncthreads();
  float2 v[P];
#pragma unroll
  for (int j = 0; j < P; j++) {
    const int zi = t + 16 * j - ogz;
    v[j] = make_float2(0.f, 0.f);
    if (lok && zi >= 0 && zi < a.N[2]) {
      const float wt = (wzs[ogz + zi] * wyv) * wxv;
      v[j] = cmul(make_float2(wt * xv[j].x, wt * xv[j].y), sv[j]);
    }
  }
  fft4_rows<S, -1>(v, t, buf + l * LD, tws);
  __syncthreads();
  float2 *gch = grid + (int64_t)ch * a.chunk_grid + ((int64_t)(a.og[1] + y) * a.S[0] + a.og[0]) * CC + gl;
  const int64_t pstride = (int64_t)a.S[1] * a.S[0] * CC;
  for (int k1 = t; k1 < P; k1 += 16) {
    float2 w[16];
    fft4_cols<S, -1>(w, buf + l * LD + k1 * 16, tws);
    if (lok) {
#pragma unroll
      for (int k2 = 0; k2 < 16; k2++) st_nt(gch + (int64_t)(k1 + P * k2) * pstride, scl(w[k2], a.scale));
    }
  }
}

// One column of k_fft_post_z4: the 16-point DFT, window x conj(coil map),
// the sum over the CC coils of a voxel (lanes of a row), X (+)= at the crop z.
template <int S>
__device__ __forceinline__ void post_col(float2 *__restrict__ X, const float2 *col, const float2 *tws,
                                         const float *wzs, const float2 *smv, const float2 *xo, int k1,
                                         float wyx, bool xok, int c, int64_t xr, int64_t zst,
                                         const PrePostArgs &a) {
  constexpr int P = S / 16;
  float2 w[16];
  fft4_cols<S, 1>(w, col, tws);
#pragma unroll
  for (int k2 = 0; k2 < 16; k2++) {
    const int z = k1 + P * k2, zi = z - a.og[2];
    const bool zok = zi >= 0 && zi < a.N[2];
    float2 tt = make_float2(0.f, 0.f);
    if (xok && zok) {
      const float wt = wzs[z] * wyx;
      const float2 im = scl(w[k2], a.scale);
      tt = cmulc(make_float2(wt * im.x, wt * im.y), smv[k2]);
    }
    tt = group_sum(tt, a.CC);
    if (xok && zok && c == 0) {
      const int64_t r = xr + (int64_t)zi * zst;
      X[r] = a.overwrite ? tt : make_float2(xo[k2].x + tt.x, xo[k2].y + tt.y);
    }
  }
}

// k_fft_post_z on the four-step transform: thread (l, u) ends with the z
// points k1 + P k2 of its line (k1 = u, then u + 16 where P > 16).  The coil
// maps of its first column (and X where it accumulates) load with the line,
// before the transform; a second column's load when it starts.  The z window
// sits in LDS beside the twiddles.
template <int S>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_post_z4(
    const float2 *__restrict__ grid, float2 *__restrict__ X, const float2 *__restrict__ smaps,
    const float *__restrict__ wx, const float *__restrict__ wy, const float *__restrict__ wz,
    PrePostArgs a, const float2 *__restrict__ tw) {
  constexpr int P = S / 16, LD = fft4_ld(S);
  __shared__ float4 buf4[FFT_LINES * LD / 2];
  __shared__ float2 tws[S];
  __shared__ float wzs[S];
  float2 *buf = reinterpret_cast<float2 *>(buf4);
  const int CC = a.CC, y = blockIdx.y, ch = a.ch;
  X += blockIdx.z * a.xfs;  // frame blockIdx.z
  grid += (int64_t)blockIdx.z * a.nchf * a.chunk_grid;
  const int bx = xcd_block(blockIdx.x, gridDim.x, a.xrun);
  const int nl = a.N[0] * CC, g0 = bx * FFT_LINES;
  const int nxl = FFT_LINES / CC;
  const int ogz = a.og[2];
  const int l = threadIdx.x & (FFT_LINES - 1), t = threadIdx.x >> 4;
  const int c = l & (CC - 1), xl = l >> a.lgCC;
  const int xi = bx * nxl + xl;
  const bool xok = xi < a.N[0];
  const int gl = g0 + l;
  // maps of voxel column (y, xi), coil c: [z] at stride CC
  const float2 *smc = smaps + (int64_t)ch * a.chunk_smap + (((int64_t)y * a.N[0] + xi) * a.N[2]) * CC + c;
  const int64_t xr = (int64_t)y * a.N[0] + xi, zst = (int64_t)a.N[1] * a.N[0];
  float2 smv[16], xo[16];
#pragma unroll
  for (int k2 = 0; k2 < 16; k2++) {
    const int zi = t + P * k2 - ogz;
    smv[k2] = xo[k2] = make_float2(0.f, 0.f);
    if (xok && t < P && zi >= 0 && zi < a.N[2]) {
      smv[k2] = ld_nt(smc + (int64_t)zi * CC);
      if (!a.overwrite && c == 0) xo[k2] = X[xr + (int64_t)zi * zst];
    }
  }
  const float2 *src = grid + (int64_t)ch * a.chunk_grid + ((int64_t)(a.og[1] + y) * a.S[0] + a.og[0]) * CC + gl;
  const int64_t pstride = (int64_t)a.S[1] * a.S[0] * CC;
  float2 v[P];
#pragma unroll
  for (int j = 0; j < P; j++) {
    v[j] = make_float2(0.f, 0.f);
    if (gl < nl) v[j] = ld_nt(src + (int64_t)(t + 16 * j) * pstride);
  }
  for (int q = threadIdx.x; q < S; q += FFT_THREADS) {
    tws[q] = tw[q];
    wzs[q] = wz[q];
  }
  __syncthreads();
  fft4_rows<S, 1>(v, t, buf + l * LD, tws);
  __syncthreads();
  const float wyx = wy[a.og[1] + y] * (xok ? wx[a.og[0] + xi] : 0.f);
  if (t < P)
    post_col<S>(X, buf + l * LD + t * 16, tws, wzs, smv, xo, t, wyx, xok, c, xr, zst, a);
  for (int k1 = t + 16; k1 < P; k1 += 16) {  // uniform over the 16 lanes of a row
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) {
      const int zi = k1 + P * k2 - ogz;
      smv[k2] = xo[k2] = make_float2(0.f, 0.f);
      if (xok && zi >= 0 && zi < a.N[2]) {
        smv[k2] = ld_nt(smc + (int64_t)zi * CC);
        if (!a.overwrite && c == 0) xo[k2] = X[xr + (int64_t)zi * zst];
      }
    }
    post_col<S>(X, buf + l * LD + k1 * 16, tws, wzs, smv, xo, k1, wyx, xok, c, xr, zst, a);
  }
}

// Direct O(S^2) DFT for sizes without a Stockham plan (plain form only).
__global__ __launch_bounds__(FFT_THREADS) void k_dft_lines(float2 *__restrict__ grid, LinesArgs a,
                                                           const float2 *__restrict__ tw, int S,
                                                           int dir) {
  extern __shared__ float2 dsm[];
  float2 *tws = dsm, *buf = dsm + S;
  for (int q = threadIdx.x; q < S; q += FFT_THREADS) tws[q] = tw[q];
  const int64_t base = a.off0 + (int64_t)blockIdx.y * a.ystride + (int64_t)blockIdx.z * a.zstride;
  const int g0 = blockIdx.x * FFT_LINES;
  for (int idx = threadIdx.x; idx < FFT_LINES * S; idx += FFT_THREADS) {
    const int l = idx / S, p = idx % S, gl = g0 + l;
    float2 v = make_float2(0.f, 0.f);
    if (gl < a.nlines && p >= a.ld0 && p < a.ld1) {
      const int r = gl / a.run, cc = gl - r * a.run;
      v = grid[base + (int64_t)r * a.rstride + cc + (int64_t)p * a.pstride];
    }
    buf[l * S + p] = v;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < FFT_LINES * (a.st1 - a.st0); idx += FFT_THREADS) {
    const int ns = a.st1 - a.st0;
    const int l = idx / ns, p = a.st0 + idx % ns, gl = g0 + l;
    if (gl >= a.nlines) continue;
    float sr = 0.f, si = 0.f;
    for (int q = 0; q < S; q++) {
      float2 t = tws[(int)(((int64_t)p * q) % S)];
      if (dir > 0) t.y = -t.y;
      const float2 x = buf[l * S + q];
      sr += x.x * t.x - x.y * t.y;
      si += x.x * t.y + x.y * t.x;
    }
    const int r = gl / a.run, cc = gl - r * a.run;
    grid[base + (int64_t)r * a.rstride + cc + (int64_t)p * a.pstride] =
        make_float2(sr * a.scale, si * a.scale);
  }
}

#define MRN_FFT_SIZES(X) \
  X(8) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(144) X(160) \
  X(192) X(240) X(256) X(288) X(320) X(384) X(400) X(432) X(480) X(512)

// The batched x passes run the four-step kernel with the frame's k-space
// support (forward: stores, inverse: loads limited to the support).
bool fft_x_pruned(const mrn_op *op) {
  return op->p.form != MRN_FORM_PLAIN && op->g.CC >= FFT_LINES / 2 && !op->fft_stockham && fft4_ok(op->g.S[0]) &&
         op->g.fft_in[0];
}

// ... and the y passes with the frame's rows of support per z plane
bool fft_xy_pruned(const mrn_op *op) { return fft_x_pruned(op) && fft4_ok(op->g.S[1]) && op->g.fft_in[1]; }

bool fft_size_supported(int S) {
  switch (S) {
#define MRN_CASE_TRUE(n) case n:
    MRN_FFT_SIZES(MRN_CASE_TRUE)
    return true;
#undef MRN_CASE_TRUE
    default:
      return false;
  }
}

int fft_lines(mrn_op *op, int S, int dir, float2 *grid, const LinesArgs &a, dim3 g3,
              const float2 *tw, hipStream_t st) {
  if (g3.x == 0 || g3.y == 0 || g3.z == 0) return MRN_OK;
  // four-step kernel where the 16 lines of a workgroup read each point as
  // 64-byte or longer runs (8 coils of two rows or 16 of one); the others
  // keep k_fft_lines' layouts, which coalesce short runs
  const bool four = a.run >= FFT_LINES / 2 && !op->fft_stockham;
  switch (S) {
#define MRN_CASE_LINES(n)                                                                    \
  case n:                                                                                    \
    if constexpr (fft4_ok(n)) {                                                              \
      if (four) {                                                                            \
        if (dir < 0)                                                                         \
          hipLaunchKernelGGL((k_fft_lines4<n, -1>), g3, dim3(FFT_THREADS), 0, st, grid, a, tw); \
        else                                                                                 \
          hipLaunchKernelGGL((k_fft_lines4<n, 1>), g3, dim3(FFT_THREADS), 0, st, grid, a, tw);  \
        break;                                                                               \
      }                                                                                      \
    }                                                                                        \
    if (dir < 0)                                                                             \
      hipLaunchKernelGGL((k_fft_lines<n, -1>), g3, dim3(FFT_THREADS), 0, st, grid, a, tw);   \
    else                                                                                     \
      hipLaunchKernelGGL((k_fft_lines<n, 1>), g3, dim3(FFT_THREADS), 0, st, grid, a, tw);    \
    break;
    MRN_FFT_SIZES(MRN_CASE_LINES)
#undef MRN_CASE_LINES
    default: {
      size_t sh = sizeof(float2) * (size_t)S * (FFT_LINES + 1);
      if (sh > 160 * 1024) return fail(op, MRN_EUNSUPPORTED, "FFT size too large for the DFT fallback");
      hipLaunchKernelGGL(k_dft_lines, g3, dim3(FFT_THREADS), sh, st, grid, a, tw, S, dir);
    }
  }
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

int fft_pre_z(mrn_op *op, float2 *grid, const float2 *X, const float2 *smaps,
              const PrePostArgs &a, hipStream_t st) {
  dim3 g3((a.N[0] * a.CC + FFT_LINES - 1) / FFT_LINES, a.N[1], op->nch * a.nf);
  switch (a.S[2]) {
#define MRN_CASE_PRE(n)                                                                    \
  case n:                                                                                  \
    if constexpr (fft4_ok(n)) {                                                            \
      if (!op->fft_stockham) {                                                             \
        hipLaunchKernelGGL((k_fft_pre_z4<n>), g3, dim3(FFT_THREADS), 0, st, grid, X, smaps,  \
                           op->win_d[0], op->win_d[1], op->win_d[2], a, op->tw_d[2]);       \
        break;                                                                             \
      }                                                                                    \
    }                                                                                      \
    hipLaunchKernelGGL((k_fft_pre_z<n>), g3, dim3(FFT_THREADS), 0, st, grid, X, smaps,     \
                       op->win_d[0], op->win_d[1], op->win_d[2], a, op->tw_d[2]);          \
    break;
    MRN_FFT_SIZES(MRN_CASE_PRE)
#undef MRN_CASE_PRE
    default:
      return fail(op, MRN_EUNSUPPORTED, "no fused FFT for this grid size");
  }
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

int fft_post_z(mrn_op *op, const float2 *grid, float2 *X, const float2 *smaps,
               const PrePostArgs &a, hipStream_t st) {
  dim3 g3((a.N[0] * a.CC + FFT_LINES - 1) / FFT_LINES, a.N[1], a.nf);
  switch (a.S[2]) {
#define MRN_CASE_POST(n)                                                                   \
  case n:                                                                                  \
    if constexpr (fft4_ok(n)) {                                                            \
      if (!op->fft_stockham) {                                                             \
        hipLaunchKernelGGL((k_fft_post_z4<n>), g3, dim3(FFT_THREADS), 0, st, grid, X, smaps, \
                           op->win_d[0], op->win_d[1], op->win_d[2], a, op->tw_d[2]);       \
        break;                                                                             \
      }                                                                                    \
    }                                                                                      \
    hipLaunchKernelGGL((k_fft_post_z<n>), g3, dim3(FFT_THREADS), 0, st, grid, X, smaps,    \
                       op->win_d[0], op->win_d[1], op->win_d[2], a, op->tw_d[2]);          \
    break;
    MRN_FFT_SIZES(MRN_CASE_POST)
#undef MRN_CASE_POST
    default:
      return fail(op, MRN_EUNSUPPORTED, "no fused FFT for this grid size");
  }
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

}  // namespace mrn
