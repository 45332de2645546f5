// Kaiser-Bessel interpolation between samples and the oversampled grid.
//
// Gather (A's last stage) follows chop_grid_backward (gridFFT.cpp:1166-1322);
// scatter (A^H's first stage) follows chop_grid_forward (gridFFT.cpp:1009-1164):
// the same tap ranges (ceil/floor of dk -/+ dwin, clipped to the grid, no
// wrap), table-lerp weights with the (-1)^l fftshift sign, kw/taper on the
// scatter side only, kw == 0 and out-of-grid samples skipped.
//
// Batched form: samples are sorted once per trajectory by the tile of their
// first tap.  The gather maps one lane to one (sample, coil) pair so that a
// tap reads CC contiguous coils (CC*8 bytes) of a grid cell.  The scatter is
// owner-computes: a workgroup owns one tile of the grid, accumulates every
// tap that lands in it in LDS (from the samples of the tile's own bin and its
// 7 lower neighbours) and writes the tile once with plain stores.  Tiles whose
// sample count exceeds one work item are split; only those are pre-zeroed and
// flushed with global atomics.
#include <hipcub/hipcub.hpp>

#include "mrn_internal.hpp"

namespace mrn {

__device__ __forceinline__ float grid_coord(const Geom &g, int a, float k) {
  // dk = k * grid + S/2 with the reference's rounding (no fused multiply-add)
  return __fadd_rn(__fmul_rn(k, g.grid[a]), (float)(g.S[a] / 2));
}

__device__ __forceinline__ void load_tables(float *tab_s, const float *tables, int ntab) {
  for (int q = threadIdx.x; q < ntab; q += blockDim.x) tab_s[q] = tables[q];
}

// ------------------------------------------------------------ binning
__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_traj_keys(const float *__restrict__ kx, const float *__restrict__ ky,
                            const float *__restrict__ kz, const float *__restrict__ kw, int64_t M,
                            Geom g, Tiles t, uint32_t *keys, int32_t *vals, int32_t *counts) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  uint32_t key = (uint32_t)t.ntiles;
  if (kw[m] != 0.0f) {
    const float k[3] = {kx[m], ky[m], kz[m]};
    bool ok = true;
    int b[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      int s, e;
      ok = ok && axis_range(g, a, grid_coord(g, a, k[a]), s, e);
      b[a] = ok ? s / t.T[a] : 0;
    }
    if (ok) key = (uint32_t)((b[2] * t.nt[1] + b[1]) * t.nt[0] + b[0]);
  }
  keys[m] = key;
  vals[m] = (int32_t)m;
  if (key < (uint32_t)t.ntiles) atomicAdd(&counts[key], 1);
}

__global__ void k_traj_finish(const float *__restrict__ kx, const float *__restrict__ ky,
                              const float *__restrict__ kz, const float *__restrict__ kw,
                              const int32_t *__restrict__ perm, int64_t Mv, Geom g,
                              float4 *coords, float *taper, int32_t *inv) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Mv) return;
  const int m = perm[i];
  const float k[3] = {kx[m], ky[m], kz[m]};
  coords[i] = make_float4(grid_coord(g, 0, k[0]), grid_coord(g, 1, k[1]), grid_coord(g, 2, k[2]), kw[m]);
  float kr = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; a++)
    if (g.fft_in[a]) kr += k[a] * k[a];
  // taper exp(-kr / (2 k_rad^2)) evaluated in double (gridFFT.cpp:1096-1108)
  taper[i] = (float)exp(-(double)kr / (2.0 * (double)g.k_rad * (double)g.k_rad));
  inv[m] = (int32_t)i;
}

__global__ void k_tile_nsub(const int32_t *__restrict__ bin_start, Tiles t, int32_t *nsub,
                            int32_t *visits) {
  int tile = blockIdx.x * blockDim.x + threadIdx.x;
  if (tile >= t.ntiles) return;
  const int tx = tile % t.nt[0], ty = (tile / t.nt[0]) % t.nt[1], tz = tile / (t.nt[0] * t.nt[1]);
  int v = 0;
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++) {
        const int bx = tx - dx, by = ty - dy, bz = tz - dz;
        if (bx < 0 || by < 0 || bz < 0) continue;
        const int b = (bz * t.nt[1] + by) * t.nt[0] + bx;
        v += bin_start[b + 1] - bin_start[b];
      }
  visits[tile] = v;
  nsub[tile] = v > t.maxv ? (v + t.maxv - 1) / t.maxv : 1;
}

__global__ void k_fill_subs(const int32_t *__restrict__ nsub, const int32_t *__restrict__ off,
                            const int32_t *__restrict__ visits, Tiles t, int4 *subs,
                            int32_t *heavy_flag) {
  int tile = blockIdx.x * blockDim.x + threadIdx.x;
  if (tile >= t.ntiles) return;
  const int n = nsub[tile], o = off[tile], v = visits[tile];
  for (int s = 0; s < n; s++) subs[o + s] = make_int4(tile, s * t.maxv, min((s + 1) * t.maxv, v), n > 1);
  heavy_flag[tile] = n > 1;
}

__global__ void k_compact(const int32_t *__restrict__ flag, const int32_t *__restrict__ off, int n,
                          int32_t *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) out[off[i]] = i;
}

// ------------------------------------------------------------ gather
// Taps per axis handled by the unrolled path; samples with more taps on an
// axis (a coordinate exactly an integer away from the window edge) take the
// loop path.
template <int NT>
__global__ __launch_bounds__(256) void k_gather(const float2 *__restrict__ grid,
                                                float2 *__restrict__ out,
                                                const float4 *__restrict__ coords,
                                                const float2 *__restrict__ sub, int64_t Mv, Geom g,
                                                const float *__restrict__ tables,
                                                int64_t chunk_grid, int64_t chunk_k) {
  extern __shared__ float tab_s[];
  load_tables(tab_s, tables, g.ntab);
  __syncthreads();
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = (int)(gid & (g.CC - 1));
  const int64_t i = gid >> g.lgCC;
  if (i >= Mv) return;
  const int ch = blockIdx.y;
  const float4 k = coords[i];
  const float dk[3] = {k.x, k.y, k.z};
  int s[3], n[3];
  bool wide = false;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    int e;
    axis_range(g, a, dk[a], s[a], e);
    n[a] = e - s[a] + 1;
    wide = wide || n[a] > NT;
  }
  const float2 *G = grid + ch * chunk_grid + c;
  float ar = 0.f, ai = 0.f;
  if (!wide) {
    float w[3][NT];
    int o[3][NT];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const bool in = t < n[a];
        w[a][t] = in ? tap_weight(g, a, tab_s + g.toff[a], dk[a], s[a] + t) : 0.0f;
        o[a][t] = in ? s[a] + t : s[a];
      }
#pragma unroll
    for (int tz = 0; tz < NT; tz++)
#pragma unroll
      for (int ty = 0; ty < NT; ty++) {
        const float wzy = w[2][tz] * w[1][ty];
        const float2 *row = G + ((int64_t)o[2][tz] * g.S[1] + o[1][ty]) * g.S[0] * g.CC;
#pragma unroll
        for (int tx = 0; tx < NT; tx++) {
          const float ww = wzy * w[0][tx];
          const float2 v = row[(int64_t)o[0][tx] * g.CC];
          ar += ww * v.x;
          ai += ww * v.y;
        }
      }
  } else {
    for (int lz = s[2]; lz < s[2] + n[2]; lz++) {
      const float wz = tap_weight(g, 2, tab_s + g.toff[2], dk[2], lz);
      for (int ly = s[1]; ly < s[1] + n[1]; ly++) {
        const float wzy = wz * tap_weight(g, 1, tab_s + g.toff[1], dk[1], ly);
        const float2 *row = G + ((int64_t)lz * g.S[1] + ly) * g.S[0] * g.CC;
        for (int lx = s[0]; lx < s[0] + n[0]; lx++) {
          const float ww = wzy * tap_weight(g, 0, tab_s + g.toff[0], dk[0], lx);
          const float2 v = row[(int64_t)lx * g.CC];
          ar += ww * v.x;
          ai += ww * v.y;
        }
      }
    }
  }
  const int64_t oi = ch * chunk_k + i * g.CC + c;
  if (sub) {
    const float2 d = sub[oi];
    ar -= d.x;
    ai -= d.y;
  }
  out[oi] = make_float2(ar, ai);
}

// ------------------------------------------------------------ scatter
__global__ void k_zero_tiles(float2 *grid, const int32_t *__restrict__ tiles, Geom g, Tiles t,
                             int64_t chunk_grid) {
  const int tile = tiles[blockIdx.x];
  const int ch = blockIdx.y;
  const int tx0 = (tile % t.nt[0]) * t.T[0], ty0 = ((tile / t.nt[0]) % t.nt[1]) * t.T[1],
            tz0 = (tile / (t.nt[0] * t.nt[1])) * t.T[2];
  const int ncell = t.T[0] * t.T[1] * t.T[2] * g.CC;
  for (int idx = threadIdx.x; idx < ncell; idx += blockDim.x) {
    const int cell = idx >> g.lgCC, c = idx & (g.CC - 1);
    const int x = tx0 + cell % t.T[0], y = ty0 + (cell / t.T[0]) % t.T[1], z = tz0 + cell / (t.T[0] * t.T[1]);
    if (x < g.S[0] && y < g.S[1] && z < g.S[2])
      grid[ch * chunk_grid + (((int64_t)z * g.S[1] + y) * g.S[0] + x) * g.CC + c] = make_float2(0.f, 0.f);
  }
}

template <int NT>
__global__ __launch_bounds__(256) void k_scatter_tiles(
    float2 *__restrict__ grid, const float2 *__restrict__ kin, const float4 *__restrict__ coords,
    const float *__restrict__ taper, const int32_t *__restrict__ bin_start,
    const int4 *__restrict__ subs, Geom g, Tiles t, const float *__restrict__ tables,
    int64_t chunk_grid, int64_t chunk_k, int CCs, int lgCCs) {
  extern __shared__ float smem[];
  float *tab_s = smem;
  float *acc = smem + ((g.ntab + 3) & ~3);
  const int4 sp = subs[blockIdx.x];
  const int tile = sp.x;
  const int nsc = g.CC >> lgCCs;
  const int ch = blockIdx.y / nsc, cb = (blockIdx.y % nsc) * CCs;
  const int tx = tile % t.nt[0], ty = (tile / t.nt[0]) % t.nt[1], tz = tile / (t.nt[0] * t.nt[1]);
  const int t0[3] = {tx * t.T[0], ty * t.T[1], tz * t.T[2]};
  const int nacc = t.T[0] * t.T[1] * t.T[2] * CCs;
  load_tables(tab_s, tables, g.ntab);
  for (int q = threadIdx.x; q < 2 * nacc; q += blockDim.x) acc[q] = 0.f;
  // the 8 source bins: this tile and its lower neighbours
  int bs[8], pre[9];
  pre[0] = 0;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int bx = tx - (q & 1), by = ty - ((q >> 1) & 1), bz = tz - (q >> 2);
    int cnt = 0, st = 0;
    if (bx >= 0 && by >= 0 && bz >= 0) {
      const int b = (bz * t.nt[1] + by) * t.nt[0] + bx;
      st = bin_start[b];
      cnt = bin_start[b + 1] - st;
    }
    bs[q] = st;
    pre[q + 1] = pre[q] + cnt;
  }
  __syncthreads();
  const int c = threadIdx.x & (CCs - 1);
  const int lane_s = threadIdx.x >> lgCCs, nls = blockDim.x >> lgCCs;
  for (int v = sp.y + lane_s; v < sp.z; v += nls) {
    int q = 0;
#pragma unroll
    for (int r = 1; r < 8; r++) q += (v >= pre[r]);
    const int64_t i = bs[q] + (v - pre[q]);
    const float4 k = coords[i];
    const float2 d = kin[ch * chunk_k + i * g.CC + cb + c];
    float vr = d.x * k.w, vi = d.y * k.w;
    if (vr == 0.f && vi == 0.f) continue;
    const float tp = taper[i];
    vr *= tp;
    vi *= tp;
    const float dk[3] = {k.x, k.y, k.z};
    int s[3], n[3], lo[3];
    bool wide = false;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      int e;
      axis_range(g, a, dk[a], s[a], e);
      n[a] = e - s[a] + 1;
      lo[a] = s[a] - t0[a];
      wide = wide || n[a] > NT;
    }
    if (!wide) {
      float w[3][NT];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int tt = 0; tt < NT; tt++) {
          const int l = lo[a] + tt;
          const bool in = tt < n[a] && l >= 0 && l < t.T[a];
          w[a][tt] = in ? tap_weight(g, a, tab_s + g.toff[a], dk[a], s[a] + tt) : 0.0f;
        }
#pragma unroll
      for (int zz = 0; zz < NT; zz++) {
        if (w[2][zz] == 0.f) continue;
#pragma unroll
        for (int yy = 0; yy < NT; yy++) {
          const float wzy = w[2][zz] * w[1][yy];
          if (wzy == 0.f) continue;
          const int rowc = ((lo[2] + zz) * t.T[1] + lo[1] + yy) * t.T[0] + lo[0];
#pragma unroll
          for (int xx = 0; xx < NT; xx++) {
            const float ww = wzy * w[0][xx];
            if (ww == 0.f) continue;
            float *a2 = acc + 2 * (((rowc + xx) << lgCCs) + c);
            atomicAdd(a2, ww * vr);
            atomicAdd(a2 + 1, ww * vi);
          }
        }
      }
    } else {
      for (int lz = max(s[2], t0[2]); lz < min(s[2] + n[2], t0[2] + t.T[2]); lz++) {
        const float wz = tap_weight(g, 2, tab_s + g.toff[2], dk[2], lz);
        for (int ly = max(s[1], t0[1]); ly < min(s[1] + n[1], t0[1] + t.T[1]); ly++) {
          const float wzy = wz * tap_weight(g, 1, tab_s + g.toff[1], dk[1], ly);
          for (int lx = max(s[0], t0[0]); lx < min(s[0] + n[0], t0[0] + t.T[0]); lx++) {
            const float ww = wzy * tap_weight(g, 0, tab_s + g.toff[0], dk[0], lx);
            const int cell = ((lz - t0[2]) * t.T[1] + (ly - t0[1])) * t.T[0] + (lx - t0[0]);
            float *a2 = acc + 2 * ((cell << lgCCs) + c);
            atomicAdd(a2, ww * vr);
            atomicAdd(a2 + 1, ww * vi);
          }
        }
      }
    }
  }
  __syncthreads();
  const bool heavy = sp.w != 0;
  float2 *gch = grid + ch * chunk_grid + cb;
  for (int idx = threadIdx.x; idx < nacc; idx += blockDim.x) {
    const int cell = idx >> lgCCs, cc = idx & (CCs - 1);
    const int x = t0[0] + cell % t.T[0], y = t0[1] + (cell / t.T[0]) % t.T[1],
              z = t0[2] + cell / (t.T[0] * t.T[1]);
    if (x >= g.S[0] || y >= g.S[1] || z >= g.S[2]) continue;
    const float2 val = make_float2(acc[2 * idx], acc[2 * idx + 1]);
    float2 *dst = gch + (((int64_t)z * g.S[1] + y) * g.S[0] + x) * g.CC + cc;
    if (heavy) {
      if (val.x != 0.f) atomicAdd(&dst->x, val.x);
      if (val.y != 0.f) atomicAdd(&dst->y, val.y);
    } else {
      *dst = val;
    }
  }
}

// ------------------------------------------------------------ plain form
// One coil grid [Sz][Sy][Sx], samples in the caller's order, one thread per
// sample, the reference's nested tap loops.
__global__ __launch_bounds__(256) void k_gather_plain(
    const float2 *__restrict__ grid, float2 *__restrict__ out, const float *__restrict__ kx,
    const float *__restrict__ ky, const float *__restrict__ kz, const float *__restrict__ kw,
    int64_t M, const float2 *__restrict__ sub, Geom g, const float *__restrict__ tables) {
  extern __shared__ float tab_s[];
  load_tables(tab_s, tables, g.ntab);
  __syncthreads();
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const float k[3] = {kx[m], ky[m], kz[m]};
  float dk[3];
  int s[3], e[3];
  bool ok = kw[m] != 0.0f;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    dk[a] = grid_coord(g, a, k[a]);
    ok = ok && axis_range(g, a, dk[a], s[a], e[a]);
  }
  if (!ok) {
    out[m] = make_float2(0.f, 0.f);
    return;
  }
  float ar = 0.f, ai = 0.f;
  for (int lz = s[2]; lz <= e[2]; lz++) {
    const float wz = tap_weight(g, 2, tab_s + g.toff[2], dk[2], lz);
    for (int ly = s[1]; ly <= e[1]; ly++) {
      const float wzy = wz * tap_weight(g, 1, tab_s + g.toff[1], dk[1], ly);
      const float2 *row = grid + ((int64_t)lz * g.S[1] + ly) * g.S[0];
      for (int lx = s[0]; lx <= e[0]; lx++) {
        const float ww = wzy * tap_weight(g, 0, tab_s + g.toff[0], dk[0], lx);
        const float2 v = row[lx];
        ar += ww * v.x;
        ai += ww * v.y;
      }
    }
  }
  if (sub) {
    ar -= sub[m].x;
    ai -= sub[m].y;
  }
  out[m] = make_float2(ar, ai);
}

__global__ __launch_bounds__(256) void k_scatter_plain(
    float2 *__restrict__ grid, const float2 *__restrict__ data, const float *__restrict__ kx,
    const float *__restrict__ ky, const float *__restrict__ kz, const float *__restrict__ kw,
    int64_t M, Geom g, const float *__restrict__ tables) {
  extern __shared__ float tab_s[];
  load_tables(tab_s, tables, g.ntab);
  __syncthreads();
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const float w = kw[m];
  if (w == 0.0f) return;
  const float2 d = data[m];
  float vr = d.x * w, vi = d.y * w;
  if (vr == 0.f && vi == 0.f) return;
  const float k[3] = {kx[m], ky[m], kz[m]};
  float dk[3];
  int s[3], e[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    dk[a] = grid_coord(g, a, k[a]);
    if (!axis_range(g, a, dk[a], s[a], e[a])) return;
  }
  float kr = 0.f;
#pragma unroll
  for (int a = 0; a < 3; a++)
    if (g.fft_in[a]) kr += k[a] * k[a];
  const float tp = (float)exp(-(double)kr / (2.0 * (double)g.k_rad * (double)g.k_rad));
  vr *= tp;
  vi *= tp;
  for (int lz = s[2]; lz <= e[2]; lz++) {
    const float wz = tap_weight(g, 2, tab_s + g.toff[2], dk[2], lz);
    for (int ly = s[1]; ly <= e[1]; ly++) {
      const float wzy = wz * tap_weight(g, 1, tab_s + g.toff[1], dk[1], ly);
      float2 *row = grid + ((int64_t)lz * g.S[1] + ly) * g.S[0];
      for (int lx = s[0]; lx <= e[0]; lx++) {
        const float ww = wzy * tap_weight(g, 0, tab_s + g.toff[0], dk[0], lx);
        atomicAdd(&row[lx].x, ww * vr);
        atomicAdd(&row[lx].y, ww * vi);
      }
    }
  }
}

// ------------------------------------------------------------ launchers
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

int taps_main(const Geom &g) {
  // unrolled tap count: 2*dwin when that is an integer (a sample exactly on
  // the window edge then takes the loop path), floor(2*dwin)+1 otherwise
  int nt = 1;
  for (int a = 0; a < 3; a++) {
    if (!g.grid_in[a]) continue;
    float w2 = 2.0f * g.dwin[a];
    int m = (w2 == floorf(w2)) ? (int)w2 : (int)floorf(w2) + 1;
    nt = m > nt ? m : nt;
  }
  return nt;
}

int launch_gather(mrn_op *op, const float2 *grid, float2 *out, const mrn::FrameData &fr,
                  const float2 *sub, hipStream_t st) {
  const Geom &g = op->g;
  if (fr.Mv == 0) return MRN_OK;
  dim3 g3(nblk(fr.Mv * g.CC, 256), op->nch);
  size_t sh = sizeof(float) * g.ntab;
  int64_t cg = (int64_t)g.S[0] * g.S[1] * g.S[2] * g.CC, ck = fr.Mv * g.CC;
  switch (taps_main(g)) {
#define MRN_G(n) \
  case n: hipLaunchKernelGGL((k_gather<n>), g3, dim3(256), sh, st, grid, out, fr.coords, sub, fr.Mv, g, op->tables_d, cg, ck); break;
    MRN_G(1) MRN_G(2) MRN_G(3) MRN_G(4) MRN_G(5) MRN_G(6) MRN_G(7) MRN_G(8)
#undef MRN_G
    default:
      return fail(op, MRN_EUNSUPPORTED, "kernel width above 8 taps");
  }
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

int launch_scatter(mrn_op *op, float2 *grid, const float2 *kin, const mrn::FrameData &fr,
                   hipStream_t st) {
  const Geom &g = op->g;
  const Tiles &t = op->t;
  int64_t cg = (int64_t)g.S[0] * g.S[1] * g.S[2] * g.CC, ck = fr.Mv * g.CC;
  if (fr.nheavy > 0) {
    hipLaunchKernelGGL(k_zero_tiles, dim3(fr.nheavy, op->nch), dim3(256), 0, st, grid, fr.heavy, g, t, cg);
    MRN_HIP(op, hipGetLastError());
  }
  int CCs = t.CCs, lg = 0;
  while ((1 << lg) < CCs) lg++;
  dim3 g3(fr.nsub, op->nch * (g.CC / CCs));
  size_t sh = sizeof(float) * (((g.ntab + 3) & ~3) + 2 * t.T[0] * t.T[1] * t.T[2] * CCs);
  switch (taps_main(g)) {
#define MRN_S(n)                                                                                   \
  case n:                                                                                          \
    hipLaunchKernelGGL((k_scatter_tiles<n>), g3, dim3(256), sh, st, grid, kin, fr.coords, fr.taper, \
                       fr.bin_start, fr.subs, g, t, op->tables_d, cg, ck, CCs, lg);                \
    break;
    MRN_S(1) MRN_S(2) MRN_S(3) MRN_S(4) MRN_S(5) MRN_S(6) MRN_S(7) MRN_S(8)
#undef MRN_S
    default:
      return fail(op, MRN_EUNSUPPORTED, "kernel width above 8 taps");
  }
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

int launch_gather_plain(mrn_op *op, const float2 *grid, float2 *out, int64_t M, const float *kx,
                        const float *ky, const float *kz, const float *kw, const float2 *sub,
                        hipStream_t st) {
  if (M == 0) return MRN_OK;
  Geom g = op->g;
  g.CC = 1;
  g.lgCC = 0;
  hipLaunchKernelGGL(k_gather_plain, dim3(nblk(M, 256)), dim3(256), sizeof(float) * g.ntab, st,
                     grid, out, kx, ky, kz, kw, M, sub, g, op->tables_d);
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

int launch_scatter_plain(mrn_op *op, float2 *grid, const float2 *data, int64_t M, const float *kx,
                         const float *ky, const float *kz, const float *kw, hipStream_t st) {
  if (M == 0) return MRN_OK;
  Geom g = op->g;
  g.CC = 1;
  g.lgCC = 0;
  hipLaunchKernelGGL(k_scatter_plain, dim3(nblk(M, 256)), dim3(256), sizeof(float) * g.ntab, st,
                     grid, data, kx, ky, kz, kw, M, g, op->tables_d);
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

// Sort one frame's samples into tiles and build the scatter work list.
int build_bins(mrn_op *op, mrn::FrameData &fr, int64_t M, const float *kx, const float *ky,
               const float *kz, const float *kw, hipStream_t st) {
  const Geom &g = op->g;
  const Tiles &t = op->t;
  const int nt = t.ntiles;
  uint32_t *keys = nullptr, *keys2 = nullptr;
  int32_t *vals = nullptr, *counts = nullptr, *nsub = nullptr, *visits = nullptr, *off = nullptr,
          *hflag = nullptr, *hoff = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0, b2 = 0;
  int rc = MRN_OK;
  auto cleanup = [&]() {
    hipFree(keys);
    hipFree(keys2);
    hipFree(vals);
    hipFree(counts);
    hipFree(nsub);
    hipFree(visits);
    hipFree(off);
    hipFree(hflag);
    hipFree(hoff);
    hipFree(tmp);
  };
#define MRN_BH(call)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      cleanup();                                                                             \
      return fail(op, e_ == hipErrorOutOfMemory ? MRN_ENOMEM : MRN_EHIP, hipGetErrorString(e_)); \
    }                                                                                        \
  } while (0)
  // release the previous trajectory of this frame
  hipFree(fr.coords);
  hipFree(fr.taper);
  hipFree(fr.perm);
  hipFree(fr.inv);
  hipFree(fr.bin_start);
  hipFree(fr.subs);
  hipFree(fr.heavy);
  fr.coords = nullptr;
  fr.taper = nullptr;
  fr.perm = nullptr;
  fr.inv = nullptr;
  fr.bin_start = nullptr;
  fr.subs = nullptr;
  fr.heavy = nullptr;
  fr.has_traj = false;
  fr.M = M;
  const int64_t Mm = M > 0 ? M : 1;
  MRN_BH(hipMalloc(&keys, sizeof(uint32_t) * Mm));
  MRN_BH(hipMalloc(&keys2, sizeof(uint32_t) * Mm));
  MRN_BH(hipMalloc(&vals, sizeof(int32_t) * Mm));
  MRN_BH(hipMalloc(&fr.perm, sizeof(int32_t) * Mm));
  MRN_BH(hipMalloc(&fr.inv, sizeof(int32_t) * Mm));
  MRN_BH(hipMalloc(&counts, sizeof(int32_t) * (nt + 1)));
  MRN_BH(hipMalloc(&fr.bin_start, sizeof(int32_t) * (nt + 1)));
  MRN_BH(hipMemsetAsync(counts, 0, sizeof(int32_t) * (nt + 1), st));
  if (M > 0) {
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(M, 256)), dim3(256), 0, st, fr.inv, -1, M);
    hipLaunchKernelGGL(k_traj_keys, dim3(nblk(M, 256)), dim3(256), 0, st, kx, ky, kz, kw, M, g, t,
                       keys, vals, counts);
    MRN_BH(hipGetLastError());
    int bits = 1;
    while ((1u << bits) <= (uint32_t)nt) bits++;
    MRN_BH(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys2, vals, fr.perm, (int)M, 0, bits, st));
    MRN_BH(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, counts, fr.bin_start, nt + 1, st));
    tmp_bytes = tmp_bytes > b2 ? tmp_bytes : b2;
    MRN_BH(hipMalloc(&tmp, tmp_bytes));
    MRN_BH(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys2, vals, fr.perm, (int)M, 0, bits, st));
  }
  size_t tb = tmp_bytes;
  if (!tmp) {
    MRN_BH(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, counts, fr.bin_start, nt + 1, st));
    MRN_BH(hipMalloc(&tmp, tb));
  }
  MRN_BH(hipcub::DeviceScan::ExclusiveSum(tmp, tb, counts, fr.bin_start, nt + 1, st));
  int32_t Mv32 = 0;
  MRN_BH(hipMemcpyAsync(&Mv32, fr.bin_start + nt, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MRN_BH(hipStreamSynchronize(st));
  fr.Mv = Mv32;
  MRN_BH(hipMalloc(&fr.coords, sizeof(float4) * (fr.Mv > 0 ? fr.Mv : 1)));
  MRN_BH(hipMalloc(&fr.taper, sizeof(float) * (fr.Mv > 0 ? fr.Mv : 1)));
  if (fr.Mv > 0)
    hipLaunchKernelGGL(k_traj_finish, dim3(nblk(fr.Mv, 256)), dim3(256), 0, st, kx, ky, kz, kw,
                       fr.perm, fr.Mv, g, fr.coords, fr.taper, fr.inv);
  MRN_BH(hipMalloc(&nsub, sizeof(int32_t) * (nt + 1)));
  MRN_BH(hipMalloc(&visits, sizeof(int32_t) * (nt + 1)));
  MRN_BH(hipMalloc(&off, sizeof(int32_t) * (nt + 1)));
  MRN_BH(hipMalloc(&hflag, sizeof(int32_t) * (nt + 1)));
  MRN_BH(hipMalloc(&hoff, sizeof(int32_t) * (nt + 1)));
  MRN_BH(hipMemsetAsync(nsub, 0, sizeof(int32_t) * (nt + 1), st));
  MRN_BH(hipMemsetAsync(hflag, 0, sizeof(int32_t) * (nt + 1), st));
  hipLaunchKernelGGL(k_tile_nsub, dim3(nblk(nt, 256)), dim3(256), 0, st, fr.bin_start, t, nsub, visits);
  MRN_BH(hipGetLastError());
  MRN_BH(hipcub::DeviceScan::ExclusiveSum(tmp, tb, nsub, off, nt + 1, st));
  int32_t nsub_total = 0;
  MRN_BH(hipMemcpyAsync(&nsub_total, off + nt, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MRN_BH(hipStreamSynchronize(st));
  fr.nsub = nsub_total;
  MRN_BH(hipMalloc(&fr.subs, sizeof(int4) * (nsub_total > 0 ? nsub_total : 1)));
  hipLaunchKernelGGL(k_fill_subs, dim3(nblk(nt, 256)), dim3(256), 0, st, nsub, off, visits, t,
                     fr.subs, hflag);
  MRN_BH(hipGetLastError());
  MRN_BH(hipcub::DeviceScan::ExclusiveSum(tmp, tb, hflag, hoff, nt + 1, st));
  int32_t nh = 0;
  MRN_BH(hipMemcpyAsync(&nh, hoff + nt, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MRN_BH(hipStreamSynchronize(st));
  fr.nheavy = nh;
  MRN_BH(hipMalloc(&fr.heavy, sizeof(int32_t) * (nh > 0 ? nh : 1)));
  if (nh > 0) {
    hipLaunchKernelGGL(k_compact, dim3(nblk(nt, 256)), dim3(256), 0, st, hflag, hoff, nt, fr.heavy);
    MRN_BH(hipGetLastError());
  }
  MRN_BH(hipStreamSynchronize(st));
#undef MRN_BH
  cleanup();
  fr.has_traj = true;
  return rc;
}

}  // namespace mrn
