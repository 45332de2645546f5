// Kaiser-Bessel interpolation between samples and the oversampled grid.
//
// Gather (A's last stage) follows chop_grid_backward (gridFFT.cpp:1166-1322);
// scatter (A^H's first stage) follows chop_grid_forward (gridFFT.cpp:1009-1164):
// the same tap ranges (ceil/floor of dk -/+ dwin, clipped to the grid, no
// wrap), table-lerp weights with the (-1)^l fftshift sign, kw/taper on the
// scatter side only, kw == 0 and out-of-grid samples skipped.
//
// Batched form: samples are sorted once per trajectory by the bin (B^3 cells)
// of their first tap.  The gather maps one lane to one (sample, coil) pair so
// that a tap reads CC contiguous coils (CC*8 bytes) of a grid cell.  The
// scatter is grid-driven: one wavefront owns a 4x4x4 tile of cells, one lane
// per cell with the CC coil sums in registers; the wavefront walks the samples
// of the bins that can reach the tile (wave-uniform loads) and every lane adds
// its own tap.  No atomics on this path (LDS float atomics run at about 3 cycles
// per lane on gfx950); a tile is written once with plain stores.  Tiles whose
// sample list exceeds one work item are split; only those are pre-zeroed and
// flushed with global atomics.
#include <hipcub/hipcub.hpp>

#include <climits>

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

// Sort key of a sample: the bin of its first tap (bins ordered z, y, x);
// skipped samples (kw == 0 or outside the grid) get the sentinel nbins.
__global__ void k_traj_keys(const float *__restrict__ kx, const float *__restrict__ ky,
                            const float *__restrict__ kz, const float *__restrict__ kw, int64_t M,
                            Geom g, Tiles t, uint32_t *keys, int32_t *vals, int32_t *counts) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = m < M;
  uint32_t key = (uint32_t)t.nbins;  // no bin (skipped, or past M)
  if (in && kw[m] != 0.0f) {
    const float k[3] = {kx[m], ky[m], kz[m]};
    bool ok = true;
    int b[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      int s, e;
      ok = ok && axis_range(g, a, grid_coord(g, a, k[a]), s, e);
      b[a] = ok ? s / t.B : 0;
    }
    if (ok) key = (uint32_t)((b[2] * t.nb[1] + b[1]) * t.nb[0] + b[0]);
  }
  if (in) {
    keys[m] = key;
    vals[m] = (int32_t)m;
  }
  // Consecutive samples of a trajectory mostly share a bin, and the centre
  // bins take a sample of every interleave: each run of equal keys among the
  // wavefront's lanes adds its length with one atomic from its first lane.
  // (Every lane reaches this point: blockDim.x is a multiple of 64.)
  const int lane = (int)(threadIdx.x & 63);
  const uint32_t prev = __shfl_up(key, 1);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long after = lane == 63 ? 0ull : heads >> (lane + 1);
  const int run = after ? __builtin_ctzll(after) + 1 : 64 - lane;
  if (head && key < (uint32_t)t.nbins) atomicAdd(&counts[key], run);
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

// taper of the valid samples for a new k_rad (same expression as k_traj_finish)
__global__ void k_taper(const float *__restrict__ traw, int64_t M, const int32_t *__restrict__ perm,
                        int64_t Mv, Geom g, float *taper) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Mv) return;
  const int64_t m = perm[i];
  float kr = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float k = traw[a * M + m];
    if (g.fft_in[a]) kr += k * k;
  }
  taper[i] = (float)exp(-(double)kr / (2.0 * (double)g.k_rad * (double)g.k_rad));
}

// Bins (per axis) whose samples can have a tap in tile `ti`: first taps in
// [ti*T - NTb + 1, ti*T + T - 1].  NTb is the usual tap count (2 dwin when
// that is an integer); a sample with more taps on some axis (its coordinate
// exactly on the window edge) can reach a tile outside this range and is
// scattered on its own by k_scatter_wide.
__host__ __device__ __forceinline__ void tile_bins(const Tiles &t, int a, int ti, int &b0, int &b1) {
  int lo = ti * t.T - t.NTb + 1;
  lo = lo < 0 ? 0 : lo;
  const int hi = ti * t.T + t.T - 1;
  b0 = lo / t.B;
  b1 = hi / t.B;
  if (b1 > t.nb[a] - 1) b1 = t.nb[a] - 1;
}

// xrng (null: every tile): a tile without samples gets no work item when no
// row's k-space support reaches its cells -- the inverse x pass never reads them
__global__ void k_tile_nsub(const int32_t *__restrict__ bin_start, Tiles t, int32_t *nsub,
                            int32_t *visits, const int2 *__restrict__ xrng, int S1, int S2) {
  int tile = blockIdx.x * blockDim.x + threadIdx.x;
  if (tile >= t.ntiles) return;
  const int tx = tile % t.nt[0], ty = (tile / t.nt[0]) % t.nt[1], tz = tile / (t.nt[0] * t.nt[1]);
  int bx0, bx1, by0, by1, bz0, bz1;
  tile_bins(t, 0, tx, bx0, bx1);
  tile_bins(t, 1, ty, by0, by1);
  tile_bins(t, 2, tz, bz0, bz1);
  int v = 0;
  for (int bz = bz0; bz <= bz1; bz++)
    for (int by = by0; by <= by1; by++) {
      const int b = (bz * t.nb[1] + by) * t.nb[0];
      v += bin_start[b + bx1 + 1] - bin_start[b + bx0];  // bins of a row are contiguous
    }
  visits[tile] = v;
  bool needed = v > 0 || !xrng;
  for (int z = tz * t.T; !needed && z < min(tz * t.T + t.T, S2); z++)
    for (int y = ty * t.T; !needed && y < min(ty * t.T + t.T, S1); y++) {
      const int2 rr = xrng[(int64_t)z * S1 + y];
      needed = rr.x <= tx * t.T + t.T - 1 && rr.y >= tx * t.T;
    }
  nsub[tile] = v > t.maxv ? (v + t.maxv - 1) / t.maxv : (needed ? 1 : 0);
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

// Visit counts widened for the offsets' scan (entry ntiles: 0).
__global__ void k_widen(const int32_t *__restrict__ v, int n, int64_t *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) out[i] = i < n ? v[i] : 0;
}

// Visit list: tile t's visits (the samples of its bin rows, rows in (bz, by)
// order, each row one contiguous run of internal indices) at
// vlist[vofs[t] ...], so that the scatter reads visit v of a tile with one
// load.  One wavefront per scatter work item (a tile's visits [sp.y, sp.z),
// at most maxv of them, so that a dense tile's list is written by several
// wavefronts), one lane per bin row: the lanes' run lengths are loaded
// together and scanned, then the wavefront writes the runs' parts inside its
// range one after the other, 64 consecutive entries per store.
__global__ void k_fill_visits(const int32_t *__restrict__ bin_start, Tiles t, const int64_t *__restrict__ vofs,
                              const int4 *__restrict__ subs, int nsub, int32_t *vlist) {
  const int item = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (item >= nsub) return;
  const int4 sp = subs[item];
  const int tile = sp.x;
  int64_t o = vofs[tile];
  const int vb = max(sp.y, 0), ve = min(sp.z, (int)(vofs[tile + 1] - o));
  if (ve <= vb) return;  // (uniform) nothing to write
  const int tx = tile % t.nt[0], ty = (tile / t.nt[0]) % t.nt[1], tz = tile / (t.nt[0] * t.nt[1]);
  int bx0, bx1, by0, by1, bz0, bz1;
  tile_bins(t, 0, tx, bx0, bx1);
  tile_bins(t, 1, ty, by0, by1);
  tile_bins(t, 2, tz, bz0, bz1);
  const int nry = by1 - by0 + 1, nrows = nry * (bz1 - bz0 + 1);
  int g0 = 0;  // the tile's visits before this group of rows
  for (int r0 = 0; r0 < nrows && g0 < ve; r0 += 64) {
    const int r = r0 + lane;
    int s0 = 0, n = 0;
    if (r < nrows) {
      const int b = ((bz0 + r / nry) * t.nb[1] + by0 + r % nry) * t.nb[0];
      s0 = bin_start[b + bx0];
      n = bin_start[b + bx1 + 1] - s0;
    }
    int x = n;  // inclusive prefix over the lanes
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    // row k's run: start sk, length nk, tile visits [ek, ek + nk); the part
    // inside [vb, ve) is written by the whole wavefront, lane q taking
    // entries q, q + 64, ... of it
    const int nr = min(64, nrows - r0);
    for (int k = 0; k < nr; k++) {
      const int nk = __shfl(n, k), sk = __shfl(s0, k), ek = g0 + __shfl(x, k) - nk;
      const int lo = max(ek, vb), hi = min(ek + nk, ve);
      for (int q = lo + lane; q < hi; q += 64) vlist[o + q] = sk + (q - ek);
    }
    g0 += __shfl(x, 63);
  }
}

// Gather work list: the samples whose first tap lies in tile t (its "home"
// samples: bins [t*T/B, (t*T+T-1)/B] per axis, one contiguous run per (bz, by)
// row), in items of at most `per` samples.  Tiles without samples get none.
__host__ __device__ __forceinline__ void home_bins(const Tiles &t, int a, int ti, int &b0, int &b1) {
  b0 = ti * t.T / t.B;
  b1 = (ti * t.T + t.T - 1) / t.B;
  if (b1 > t.nb[a] - 1) b1 = t.nb[a] - 1;
}

__global__ void k_gtile_count(const int32_t *__restrict__ bin_start, Tiles t, int per, int min_dense,
                              int32_t *nsub, int32_t *count, int32_t *nsparse) {
  int tile = blockIdx.x * blockDim.x + threadIdx.x;
  if (tile >= t.ntiles) return;
  const int tx = tile % t.nt[0], ty = (tile / t.nt[0]) % t.nt[1], tz = tile / (t.nt[0] * t.nt[1]);
  int bx0, bx1, by0, by1, bz0, bz1;
  home_bins(t, 0, tx, bx0, bx1);
  home_bins(t, 1, ty, by0, by1);
  home_bins(t, 2, tz, bz0, bz1);
  int v = 0;
  for (int bz = bz0; bz <= bz1; bz++)
    for (int by = by0; by <= by1; by++) {
      const int b = (bz * t.nb[1] + by) * t.nb[0];
      v += bin_start[b + bx1 + 1] - bin_start[b + bx0];
    }
  count[tile] = v;
  const bool dense = v >= min_dense;
  nsub[tile] = dense ? (v + per - 1) / per : 0;
  nsparse[tile] = dense ? 0 : v;
}

// internal indices of the home samples of the sparse tiles, tile by tile
__global__ void k_gtile_sparse(const int32_t *__restrict__ bin_start, Tiles t,
                               const int32_t *__restrict__ nsparse, const int32_t *__restrict__ soff,
                               int32_t *idx) {
  int tile = blockIdx.x * blockDim.x + threadIdx.x;
  if (tile >= t.ntiles || nsparse[tile] == 0) return;
  const int tx = tile % t.nt[0], ty = (tile / t.nt[0]) % t.nt[1], tz = tile / (t.nt[0] * t.nt[1]);
  int bx0, bx1, by0, by1, bz0, bz1;
  home_bins(t, 0, tx, bx0, bx1);
  home_bins(t, 1, ty, by0, by1);
  home_bins(t, 2, tz, bz0, bz1);
  int o = soff[tile];
  for (int bz = bz0; bz <= bz1; bz++)
    for (int by = by0; by <= by1; by++) {
      const int b = (bz * t.nb[1] + by) * t.nb[0];
      for (int q = bin_start[b + bx0]; q < bin_start[b + bx1 + 1]; q++) idx[o++] = q;
    }
}

__global__ void k_gtile_fill(const int32_t *__restrict__ nsub, const int32_t *__restrict__ off,
                             const int32_t *__restrict__ count, Tiles t, int per, int4 *gsubs) {
  int tile = blockIdx.x * blockDim.x + threadIdx.x;
  if (tile >= t.ntiles) return;
  const int n = nsub[tile], o = off[tile], v = count[tile];
  for (int q = 0; q < n; q++) gsubs[o + q] = make_int4(tile, q * per, min((q + 1) * per, v), 0);
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
template <int NT, int NC>
__global__ __launch_bounds__(256) void k_gather(const float2 *__restrict__ grid,
                                                float2 *__restrict__ out,
                                                const float4 *__restrict__ coords,
                                                const float2 *__restrict__ sub, int64_t Mv, Geom g,
                                                const float *__restrict__ tables,
                                                int64_t chunk_grid, int64_t chunk_k,
                                                const float *__restrict__ taper, int weighted,
                                                int xrun, const int32_t *__restrict__ list,
                                                int64_t nlist, const FrameDev *__restrict__ fdt, int fsub) {
  extern __shared__ float tab_s[];
  load_tables(tab_s, tables, g.ntab);
  __syncthreads();
  if (fdt) {  // every frame at once: frame blockIdx.z (sub = its data when fsub)
    const FrameDev &fd = fdt[blockIdx.z];
    grid += fd.grid_off;
    out = fd.kwork;
    coords = fd.coords;
    sub = fsub ? fd.kdata : nullptr;
    taper = fd.taper;
    Mv = fd.Mv;
    chunk_k = fd.Mv * g.CC;
  }
  const int64_t gid = (int64_t)xcd_block(blockIdx.x, gridDim.x, xrun) * blockDim.x + threadIdx.x;
  constexpr int LGNC = NC == 8 ? 3 : NC == 4 ? 2 : NC == 2 ? 1 : 0;
  const int lgl = g.lgCC - LGNC;  // log2 of the lanes per sample
  const int c = (int)(gid & ((1 << lgl) - 1)) * NC;
  int64_t i = gid >> lgl;  // sample (or position in `list` when given)
  if (list) {
    if (i >= nlist) return;
    i = list[i];
  } else if (i >= Mv) {
    return;
  }
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
  float ar[NC], ai[NC];
#pragma unroll
  for (int q = 0; q < NC; q++) ar[q] = ai[q] = 0.f;
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
    // rows of NT taps, software-pipelined: row r + 1's loads are issued
    // before row r's products (two register buffers), and the scheduling
    // barriers keep the compiler from trading them back for occupancy
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
      const float wzy = w[2][r / NT] * w[1][r % NT];
#pragma unroll
      for (int tx = 0; tx < NT; tx++) {
        const float ww = wzy * w[0][tx];
#pragma unroll
        for (int q = 0; q < NC; q++) {
          ar[q] += ww * vb[r & 1][tx].v[q].x;
          ai[q] += ww * vb[r & 1][tx].v[q].y;
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    for (int lz = s[2]; lz < s[2] + n[2]; lz++) {
      const float wz = tap_weight(g, 2, tab_s + g.toff[2], dk[2], lz);
      for (int ly = s[1]; ly < s[1] + n[1]; ly++) {
        const float wzy = wz * tap_weight(g, 1, tab_s + g.toff[1], dk[1], ly);
        const float2 *row = G + ((int64_t)lz * g.S[1] + ly) * g.S[0] * g.CC;
        for (int lx = s[0]; lx < s[0] + n[0]; lx++) {
          const float ww = wzy * tap_weight(g, 0, tab_s + g.toff[0], dk[0], lx);
          const CoilVec<NC> v = CoilVec<NC>::load(row + (int64_t)lx * g.CC);
#pragma unroll
          for (int q = 0; q < NC; q++) {
            ar[q] += ww * v.v[q].x;
            ai[q] += ww * v.v[q].y;
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NC; q++) {
    const int64_t oi = ch * chunk_k + i * g.CC + c + q;
    if (sub) {
      const float2 d = sub[oi];
      ar[q] -= d.x;
      ai[q] -= d.y;
    }
    if (weighted) {  // the scatter's input kw * taper * y (gridFFT.cpp:1040-1108)
      const float tp = taper[i];
      ar[q] = (ar[q] * k.w) * tp;
      ai[q] = (ai[q] * k.w) * tp;
    }
    out[oi] = make_float2(ar[q], ai[q]);
  }
}

// Whether the forward passes stored cell (x, y, z) for this frame: with its
// k-space support (xrng, pruned x passes) only the cells some tap reaches
// were stored, and the staged regions below hold zeros elsewhere instead of
// whatever an earlier call left in the grid.
__device__ __forceinline__ bool in_support(const int2 *__restrict__ xrng, const Geom &g, int x, int y, int z) {
  if (!xrng) return true;
  const int2 rr = xrng[(int64_t)z * g.S[1] + y];
  return x >= rr.x && x <= rr.y;
}

// Tile-staged gather.  Work item gsubs[blockIdx.x] = {tile, first, end} over
// the tile's home samples (first tap inside the 4x4x4 tile).  The workgroup
// loads the R^3 cells (R = 4 + NT - 1) that those samples' taps can reach, for
// coil block blockIdx.y, into LDS with coalesced 16-byte loads, then every
// home sample reads its NT^3 taps from LDS: lane = (sample, coil pair), so a
// wavefront serves 64 / (CCT/2) samples per pass.  A sample with more than NT
// taps on an axis (it sits exactly on the window edge) reads its taps from the
// grid instead.  Same arithmetic and order per coil as k_gather.
#define GT_PER 128   // home samples per work item
#define GT_MIN 64    // tiles with fewer home samples are gathered directly
template <int NT, int CCT>
__global__ __launch_bounds__(256) void k_gather_tile(const float2 *__restrict__ grid, float2 *__restrict__ out,
                                                     const float4 *__restrict__ coords,
                                                     const float2 *__restrict__ sub,
                                                     const int32_t *__restrict__ bin_start,
                                                     const int4 *__restrict__ gsubs, Geom g, Tiles t,
                                                     const float *__restrict__ tables, int64_t chunk_grid,
                                                     int64_t chunk_k, const float *__restrict__ taper,
                                                     int weighted, const int2 *__restrict__ xrng) {
  constexpr int R = NT + 3, HC = CCT / 2;  // region edge, float4 per cell
  constexpr int LPS = HC;                  // lanes per sample
  constexpr int SPP = 256 / LPS;           // samples per pass
  extern __shared__ float4 reg[];          // [R][R][R][HC]
  __shared__ int s_rs[32], s_rp[33];
  const int4 sp = gsubs[blockIdx.x];
  const int tile = sp.x;
  const int tix = tile % t.nt[0], tiy = (tile / t.nt[0]) % t.nt[1], tiz = tile / (t.nt[0] * t.nt[1]);
  const int t0[3] = {tix * 4, tiy * 4, tiz * 4};
  const int ch = blockIdx.y;
  const float4 *G4 = reinterpret_cast<const float4 *>(grid + ch * chunk_grid);
  // ---- stage the region: every load of a thread issued before its LDS writes
  constexpr int NE = R * R * R * HC, RN = (NE + 255) / 256;
  {
    float4 v[RN];
#pragma unroll
    for (int j = 0; j < RN; j++) {
      const int e = threadIdx.x + 256 * j;
      const int q = e % HC, cell = e / HC;
      const int rx = cell % R, ry = (cell / R) % R, rz = cell / (R * R);
      const int gx = t0[0] + rx, gy = t0[1] + ry, gz = t0[2] + rz;
      v[j] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (e < NE && gx < g.S[0] && gy < g.S[1] && gz < g.S[2] && in_support(xrng, g, gx, gy, gz))
        v[j] = G4[(((int64_t)gz * g.S[1] + gy) * g.S[0] + gx) * HC + q];
    }
#pragma unroll
    for (int j = 0; j < RN; j++) {
      const int e = threadIdx.x + 256 * j;
      if (e < NE) reg[e] = v[j];
    }
  }
  // ---- home rows of the tile and their prefix counts
  int bx0, bx1, by0, by1, bz0, bz1;
  home_bins(t, 0, tix, bx0, bx1);
  home_bins(t, 1, tiy, by0, by1);
  home_bins(t, 2, tiz, bz0, bz1);
  const int nry = by1 - by0 + 1, nrows = nry * (bz1 - bz0 + 1);
  if (threadIdx.x < 64) {
    const int r = threadIdx.x;
    int n = 0;
    if (r < nrows) {
      const int brow = ((bz0 + r / nry) * t.nb[1] + by0 + r % nry) * t.nb[0];
      const int s0 = bin_start[brow + bx0];
      n = bin_start[brow + bx1 + 1] - s0;
      s_rs[r] = s0;
    }
    int x = n;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o);
      if (r >= o) x += y;
    }
    if (r < nrows) s_rp[r + 1] = x;
    if (r == 0) s_rp[0] = 0;
  }
  __syncthreads();
  const int cq = threadIdx.x % LPS, c = cq * 2;
  for (int base = sp.y; base < sp.z; base += SPP) {
    const int v = base + threadIdx.x / LPS;
    if (v >= sp.z) continue;
    int lo = 0, hi = nrows - 1;  // last row with rp[row] <= v
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_rp[mid] <= v) lo = mid;
      else hi = mid - 1;
    }
    const int64_t i = (int64_t)s_rs[lo] + (v - s_rp[lo]);
    const float4 k = coords[i];
    const float dk[3] = {k.x, k.y, k.z};
    int s3[3], n3[3];
    bool local = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      int e;
      axis_range(g, a, dk[a], s3[a], e);
      n3[a] = e - s3[a] + 1;
      local = local && n3[a] <= NT && s3[a] - t0[a] >= 0 && s3[a] - t0[a] + n3[a] <= R;
    }
    float ar[2] = {0.f, 0.f}, ai[2] = {0.f, 0.f};
    if (local) {
      float w[3][NT];
      int o[3][NT];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int tt = 0; tt < NT; tt++) {
          const bool in = tt < n3[a];
          w[a][tt] = in ? tap_weight(g, a, tables + g.toff[a], dk[a], s3[a] + tt) : 0.0f;
          o[a][tt] = (s3[a] - t0[a]) + (in ? tt : 0);
        }
#pragma unroll
      for (int tz = 0; tz < NT; tz++)
#pragma unroll
        for (int ty = 0; ty < NT; ty++) {
          const float wzy = w[2][tz] * w[1][ty];
          const float4 *row = reg + ((o[2][tz] * R + o[1][ty]) * R) * HC + cq;
#pragma unroll
          for (int tx = 0; tx < NT; tx++) {
            const float ww = wzy * w[0][tx];
            const float4 q = row[o[0][tx] * HC];
            ar[0] += ww * q.x;
            ai[0] += ww * q.y;
            ar[1] += ww * q.z;
            ai[1] += ww * q.w;
          }
        }
    } else {  // taps outside the staged region: read the grid
      const float2 *Gc = grid + ch * chunk_grid + c;
      for (int lz = s3[2]; lz < s3[2] + n3[2]; lz++) {
        const float wz = tap_weight(g, 2, tables + g.toff[2], dk[2], lz);
        for (int ly = s3[1]; ly < s3[1] + n3[1]; ly++) {
          const float wzy = wz * tap_weight(g, 1, tables + g.toff[1], dk[1], ly);
          const float2 *row = Gc + ((int64_t)lz * g.S[1] + ly) * g.S[0] * CCT;
          for (int lx = s3[0]; lx < s3[0] + n3[0]; lx++) {
            const float ww = wzy * tap_weight(g, 0, tables + g.toff[0], dk[0], lx);
            const float4 q = *reinterpret_cast<const float4 *>(row + (int64_t)lx * CCT);
            ar[0] += ww * q.x;
            ai[0] += ww * q.y;
            ar[1] += ww * q.z;
            ai[1] += ww * q.w;
          }
        }
      }
    }
#pragma unroll
    for (int qq = 0; qq < 2; qq++) {
      const int64_t oi = ch * chunk_k + i * CCT + c + qq;
      if (sub) {
        const float2 d = sub[oi];
        ar[qq] -= d.x;
        ai[qq] -= d.y;
      }
      if (weighted) {
        const float tp = taper[i];
        ar[qq] = (ar[qq] * k.w) * tp;
        ai[qq] = (ai[qq] * k.w) * tp;
      }
      out[oi] = make_float2(ar[qq], ai[qq]);
    }
  }
}

// Make this wavefront's LDS writes visible to its other lanes (LDS executes a
// wavefront's instructions in order; the fence keeps the compiler from moving
// LDS accesses across this point and waits for the writes).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
typedef float mrn_v16f __attribute__((ext_vector_type(16)));
typedef float mrn_v4f __attribute__((ext_vector_type(4)));

// Tile-staged gather on the matrix cores (16-coil blocks).  A workgroup
// stages the cells that the home samples of work item gsubs[blockIdx.x] can
// reach (R = NT + 3 per axis, x padded to the even RX) in LDS, as
// k_gather_tile does; then each wavefront takes batches of 32 of those
// samples and forms
//   out[32 samples][32 reals] = W[32][cells] G[cells][32]
// with v_mfma_f32_32x32x2f32, W the samples' separable cell weights (zero
// outside their taps) and G the staged coil values.  Each staged value
// serves the 32 samples of a batch in one operand read, where the per-lane
// form reads every tap of every sample.  Step A: lane s < 32 locates sample
// s and writes its weights per axis and offset to LDS as [axis][offset][s]
// (consecutive lanes, consecutive banks), with its index, kw and taper.
// Step B walks the cells as (z, y) rows of x pairs: lane l takes x = 2p +
// l/32, so A = (wz wy) wx with wx fixed per lane for the whole batch (held
// in registers) and wz wy fixed per row, and B = G[cell][l%32].  A sample
// whose taps leave the region (NT + 1 taps on an axis, its coordinate
// exactly on the window edge) gets zero weights and is computed afterwards
// tap by tap from the grid.  The weight products and the epilogue
// (residual, kw * taper) are those of k_gather; the sum over cells runs in
// another order.
template <int NT>
__global__ __launch_bounds__(256) void k_gather_mm(const float2 *__restrict__ grid, float2 *__restrict__ out,
                                                   const float4 *__restrict__ coords,
                                                   const float2 *__restrict__ sub,
                                                   const int32_t *__restrict__ bin_start,
                                                   const int4 *__restrict__ gsubs, Geom g, Tiles t,
                                                   const float *__restrict__ tables, int64_t chunk_grid,
                                                   int64_t chunk_k, const float *__restrict__ taper,
                                                   int weighted, const int2 *__restrict__ xrng) {
  constexpr int CCT = 16, HC = 8, R = NT + 3, RX = (R + 1) & ~1, XP = RX / 2;
  extern __shared__ float4 reg[];  // [R][R][RX][HC]: cell c holds reals 32c .. 32c+31
  __shared__ int s_rs[32], s_rp[33];
  __shared__ float s_w[4][3][RX][32];  // [wave][axis][offset][sample]
  __shared__ int64_t s_idx[4][32];     // internal sample index, -1 if none
  __shared__ float s_kw[4][32], s_tp[4][32];
  const int4 sp = gsubs[blockIdx.x];
  const int tile = sp.x;
  const int tix = tile % t.nt[0], tiy = (tile / t.nt[0]) % t.nt[1], tiz = tile / (t.nt[0] * t.nt[1]);
  const int t0[3] = {tix * 4, tiy * 4, tiz * 4};
  const int ch = blockIdx.y;
  const float4 *G4 = reinterpret_cast<const float4 *>(grid + ch * chunk_grid);
  // ---- stage the region: every load of a thread issued before its LDS writes
  constexpr int NE = R * R * RX * HC, RN = (NE + 255) / 256;
  {
    float4 v[RN];
#pragma unroll
    for (int j = 0; j < RN; j++) {
      const int e = threadIdx.x + 256 * j;
      const int q = e % HC, cell = e / HC;
      const int rx = cell % RX, ry = (cell / RX) % R, rz = cell / (RX * R);
      const int gx = t0[0] + rx, gy = t0[1] + ry, gz = t0[2] + rz;
      v[j] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (e < NE && gx < g.S[0] && gy < g.S[1] && gz < g.S[2] && in_support(xrng, g, gx, gy, gz))
        v[j] = G4[(((int64_t)gz * g.S[1] + gy) * g.S[0] + gx) * HC + q];
    }
#pragma unroll
    for (int j = 0; j < RN; j++) {
      const int e = threadIdx.x + 256 * j;
      if (e < NE) reg[e] = v[j];
    }
  }
  // ---- home rows of the tile and their prefix counts
  int bx0, bx1, by0, by1, bz0, bz1;
  home_bins(t, 0, tix, bx0, bx1);
  home_bins(t, 1, tiy, by0, by1);
  home_bins(t, 2, tiz, bz0, bz1);
  const int nry = by1 - by0 + 1, nrows = nry * (bz1 - bz0 + 1);
  if (threadIdx.x < 64) {
    const int r = threadIdx.x;
    int n = 0;
    if (r < nrows) {
      const int brow = ((bz0 + r / nry) * t.nb[1] + by0 + r % nry) * t.nb[0];
      const int s0 = bin_start[brow + bx0];
      n = bin_start[brow + bx1 + 1] - s0;
      s_rs[r] = s0;
    }
    int x = n;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o);
      if (r >= o) x += y;
    }
    if (r < nrows) s_rp[r + 1] = x;
    if (r == 0) s_rp[0] = 0;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float *regf = reinterpret_cast<const float *>(reg);
  const int col = lane & 31, kh = lane >> 5;  // MFMA row / column index, k half
  for (int base = sp.y + 32 * wave; base < sp.z; base += 128) {
    // ---- A: samples of the batch, their weights per axis
    unsigned long long slow = 0;  // samples computed tap by tap
    if (lane < 32) {
      const int v = base + lane;
      int64_t i = -1;
      float kw = 0.f, tp = 0.f;
      bool local = false;
      float dk[3] = {0.f, 0.f, 0.f};
      int s3[3] = {0, 0, 0}, n3[3] = {0, 0, 0};
      if (v < sp.z) {
        int lo = 0, hi = nrows - 1;  // last row with rp[row] <= v
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_rp[mid] <= v) lo = mid;
          else hi = mid - 1;
        }
        i = (int64_t)s_rs[lo] + (v - s_rp[lo]);
        const float4 k = coords[i];
        kw = k.w;
        tp = weighted ? taper[i] : 1.0f;
        dk[0] = k.x;
        dk[1] = k.y;
        dk[2] = k.z;
        local = true;
#pragma unroll
        for (int a = 0; a < 3; a++) {
          int e;
          axis_range(g, a, dk[a], s3[a], e);
          n3[a] = e - s3[a] + 1;
          local = local && n3[a] <= NT && s3[a] - t0[a] >= 0 && s3[a] - t0[a] + n3[a] <= R;
        }
      }
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int r = 0; r < RX; r++) {
          const int tt = r - (s3[a] - t0[a]);
          const bool in = local && tt >= 0 && tt < n3[a];
          s_w[wave][a][r][lane] = in ? tap_weight(g, a, tables + g.toff[a], dk[a], t0[a] + r) : 0.0f;
        }
      s_idx[wave][lane] = i;
      s_kw[wave][lane] = kw;
      s_tp[wave][lane] = tp;
      slow = __ballot(i >= 0 && !local);
    }
    slow = __shfl(slow, 0);  // the wave-wide mask from lane 0
    wave_sync();
    // ---- B: out[32][32] = W[32][cells] G[cells][32]
    float wx[XP];
#pragma unroll
    for (int p = 0; p < XP; p++) wx[p] = s_w[wave][0][2 * p + kh][col];
    mrn_v16f acc0 = {}, acc1 = {};
    for (int cz = 0; cz < R; cz++) {
      const float wz = s_w[wave][2][cz][col];
      for (int cy = 0; cy < R; cy++) {
        const float wzy = wz * s_w[wave][1][cy][col];
        const float *brow = regf + ((cz * R + cy) * RX + kh) * 32 + col;
#pragma unroll
        for (int p = 0; p < XP; p++) {
          const float a = wzy * wx[p];
          const float b = brow[64 * p];
          if (p & 1) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc1, 0, 0, 0);
          else acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc0, 0, 0, 0);
        }
      }
    }
    // ---- epilogue: lane holds sample 8(r/4) + 4 kh + r%4, real col
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int smp = 8 * (r >> 2) + 4 * kh + (r & 3);
      const int64_t i = s_idx[wave][smp];
      if (i < 0 || ((slow >> smp) & 1ull)) continue;
      float val = acc0[r] + acc1[r];
      float *o = reinterpret_cast<float *>(out + ch * chunk_k + i * CCT) + col;
      if (sub) val -= reinterpret_cast<const float *>(sub + ch * chunk_k + i * CCT)[col];
      if (weighted) val = (val * s_kw[wave][smp]) * s_tp[wave][smp];
      *o = val;
    }
    // ---- samples whose taps leave the region: lane = coil, taps from the grid
    while (slow) {
      const int smp = __builtin_ctzll(slow);
      slow &= slow - 1;
      if (lane < CCT) {
        const int64_t i = s_idx[wave][smp];
        const float4 k = coords[i];
        const float dk[3] = {k.x, k.y, k.z};
        int s3[3], n3[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
          int e;
          axis_range(g, a, dk[a], s3[a], e);
          n3[a] = e - s3[a] + 1;
        }
        const float2 *Gc = grid + ch * chunk_grid + lane;
        float ar = 0.f, ai = 0.f;
        for (int lz = s3[2]; lz < s3[2] + n3[2]; lz++) {
          const float wz = tap_weight(g, 2, tables + g.toff[2], dk[2], lz);
          for (int ly = s3[1]; ly < s3[1] + n3[1]; ly++) {
            const float wzy = wz * tap_weight(g, 1, tables + g.toff[1], dk[1], ly);
            const float2 *row = Gc + ((int64_t)lz * g.S[1] + ly) * g.S[0] * CCT;
            for (int lx = s3[0]; lx < s3[0] + n3[0]; lx++) {
              const float ww = wzy * tap_weight(g, 0, tables + g.toff[0], dk[0], lx);
              const float2 q = row[(int64_t)lx * CCT];
              ar += ww * q.x;
              ai += ww * q.y;
            }
          }
        }
        const int64_t oi = ch * chunk_k + i * CCT + lane;
        if (sub) {
          const float2 d = sub[oi];
          ar -= d.x;
          ai -= d.y;
        }
        if (weighted) {
          const float tp = taper[i];
          ar = (ar * k.w) * tp;
          ai = (ai * k.w) * tp;
        }
        out[oi] = make_float2(ar, ai);
      }
    }
    wave_sync();
  }
}

// Matrix-core gather of the dense tiles without staging (16-coil blocks).
// Work item gsubs[blockIdx.x] = {tile, first, end} over the home samples of
// a tile (first tap inside the 4x4x4 tile); each wavefront takes batches of
// 32 of them and forms out[32 samples][32 reals] = W[32][cells] G[cells][32]
// with v_mfma_f32_32x32x2f32, as k_gather_mm does, but
//  * only over the box of region cells that some sample of the batch reaches
//    (consecutive home samples share bins: about 5 x 5 x 6 cells rather than
//    the (NT + 3)^3 region), and
//  * with one z plane of the box staged in LDS at a time (the 8 x cells of
//    each row of the box: 1 KB per row, one 16-byte load per lane, 8 KB per
//    wavefront), the next plane's loads in flight during this plane's
//    products.
// Each grid value read serves the 32 samples of the batch; the per-lane
// gather reads every tap of every sample through the texture path.  Cells
// outside the frame's k-space support (which the forward x pass did not
// store) are staged as 0.  Step A, the samples with taps outside the region
// and the epilogue are k_gather_mm's.
template <int NT>
__global__ __launch_bounds__(256) void k_gather_box(const float2 *__restrict__ grid, float2 *__restrict__ out,
                                                    const float4 *__restrict__ coords,
                                                    const float2 *__restrict__ sub,
                                                    const int32_t *__restrict__ bin_start,
                                                    const int4 *__restrict__ gsubs, int ngsub, Geom g, Tiles t,
                                                    const float *__restrict__ tables, int64_t chunk_grid,
                                                    int64_t chunk_k, const float *__restrict__ taper,
                                                    int weighted, const int2 *__restrict__ xrng) {
  constexpr int CCT = 16, R = NT + 3, RX = (R + 1) & ~1, XP = RX / 2;
  static_assert(RX <= 8, "region offsets are kept in 8-bit masks");
  __shared__ int s_rs[4][32], s_rp[4][33];
  __shared__ float s_w[4][3][RX][32];  // [wave][axis][offset][sample]
  __shared__ int64_t s_idx[4][32];     // internal sample index, -1 if none
  __shared__ float s_kw[4][32], s_tp[4][32];
  __shared__ float4 s_g[4][RX * 64];   // [wave][row of the plane][x cell][16 bytes]: one staged z plane
  extern __shared__ float tab_s[];     // kernel tables (g.ntab floats)
  // (before any wavefront can leave: the whole workgroup reaches the barrier)
  load_tables(tab_s, tables, g.ntab);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int sidx = xcd_block(blockIdx.x, gridDim.x, t.xrun) * 4 + wave;
  if (sidx >= ngsub) return;
  const int4 sp = gsubs[sidx];
  const int tile = sp.x;
  const int tix = tile % t.nt[0], tiy = (tile / t.nt[0]) % t.nt[1], tiz = tile / (t.nt[0] * t.nt[1]);
  const int t0[3] = {tix * 4, tiy * 4, tiz * 4};
  const int ch = blockIdx.y;
  // ---- home rows of the tile and their prefix counts (at most 16 rows)
  int bx0, bx1, by0, by1, bz0, bz1;
  home_bins(t, 0, tix, bx0, bx1);
  home_bins(t, 1, tiy, by0, by1);
  home_bins(t, 2, tiz, bz0, bz1);
  const int nry = by1 - by0 + 1, nrows = nry * (bz1 - bz0 + 1);
  {
    const int r = lane;
    int n = 0;
    if (r < nrows) {
      const int brow = ((bz0 + r / nry) * t.nb[1] + by0 + r % nry) * t.nb[0];
      const int s0 = bin_start[brow + bx0];
      n = bin_start[brow + bx1 + 1] - s0;
      s_rs[wave][r] = s0;
    }
    int x = n;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o);
      if (r >= o) x += y;
    }
    if (r < nrows) s_rp[wave][r + 1] = x;
    if (r == 0) s_rp[wave][0] = 0;
  }
  wave_sync();
  const int col = lane & 31, kh = lane >> 5;  // MFMA row / column index, k half
  for (int base = sp.y; base < sp.z; base += 32) {
    // ---- A: samples of the batch, their weights per axis and region
    // offset: lane (s, h) = (l % 32, l / 32) writes offsets 4h .. 4h + 3
    unsigned long long slow = 0;  // samples computed tap by tap
    unsigned box = 0;             // region offsets reached: bits 8 a + r
    {
      const int v = base + col;
      int64_t i = -1;
      float kw = 0.f, tp = 0.f;
      bool local = false;
      float dk[3] = {0.f, 0.f, 0.f};
      int s3[3] = {0, 0, 0}, n3[3] = {0, 0, 0};
      if (v < sp.z) {
        int lo = 0, hi = nrows - 1;  // last row with rp[row] <= v
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_rp[wave][mid] <= v) lo = mid;
          else hi = mid - 1;
        }
        i = (int64_t)s_rs[wave][lo] + (v - s_rp[wave][lo]);
        const float4 k = coords[i];
        kw = k.w;
        tp = weighted ? taper[i] : 1.0f;
        dk[0] = k.x;
        dk[1] = k.y;
        dk[2] = k.z;
        local = true;
#pragma unroll
        for (int a = 0; a < 3; a++) {
          int e;
          axis_range(g, a, dk[a], s3[a], e);
          n3[a] = e - s3[a] + 1;
          local = local && n3[a] <= NT && s3[a] - t0[a] >= 0 && s3[a] - t0[a] + n3[a] <= R;
        }
      }
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int u = 0; u < RX / 2; u++) {
          const int r = (RX / 2) * kh + u;
          const int tt = r - (s3[a] - t0[a]);
          const bool in = local && tt >= 0 && tt < n3[a];
          const float w = g.kmode == MRN_KMODE_EXACT_KB ? tap_weight(g, a, tab_s + g.toff[a], dk[a], t0[a] + r)
                                                        : tap_weight_clamped(g, a, tab_s + g.toff[a], dk[a], t0[a] + r);
          s_w[wave][a][r][col] = in ? w : 0.0f;
          box |= in ? 1u << (8 * a + r) : 0u;
        }
      if (kh == 0) {
        s_idx[wave][col] = i;
        s_kw[wave][col] = kw;
        s_tp[wave][col] = tp;
      }
      slow = __ballot(i >= 0 && !local) & 0xFFFFFFFFull;
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) box |= __shfl_xor(box, o);
    box = __builtin_amdgcn_readfirstlane(box);
    wave_sync();
    // ---- B: out[32][32] = W[32][box cells] G[box cells][32], one z plane
    // of the box at a time: the wavefront stages the plane's rows (all 8 x
    // cells of the region, 1 KB per row, one 16-byte load per lane) in LDS,
    // issues the next plane's loads, then runs the plane's products
    mrn_v16f acc0 = {}, acc1 = {};
    if (box & 0xFFu) {
      const unsigned ym = (box >> 8) & 0xFFu, zm = (box >> 16) & 0xFFu;
      const int zlo = __builtin_ctz(zm), zhi = 31 - __builtin_clz(zm);
      const int ylo = __builtin_ctz(ym), ny = 32 - __builtin_clz(ym) - ylo;
      float wx[XP];
#pragma unroll
      for (int p = 0; p < XP; p++) wx[p] = s_w[wave][0][2 * p + kh][col];
      // staging lane: cell x = t0 + (lane >> 3) of each row, 16 bytes q = lane & 7
      const int sx = t0[0] + (lane >> 3);
      const bool xin = (lane >> 3) < RX && sx < g.S[0];
      const float4 *G4 = reinterpret_cast<const float4 *>(grid + ch * chunk_grid) + (xin ? sx * 8 + (lane & 7) : 0);
      const int64_t rowst = (int64_t)g.S[0] * 8;  // float4 per grid row
      float4 st[RX];
      auto load_plane = [&](int cz) {
        const int gz = t0[2] + cz, gy0 = t0[1] + ylo;
        const float4 *P = G4 + ((int64_t)gz * g.S[1] + gy0) * rowst;
#pragma unroll
        for (int j = 0; j < RX; j++) {
          st[j] = make_float4(0.f, 0.f, 0.f, 0.f);
          if (j < ny) {
            int2 rr = make_int2(0, g.S[0] - 1);
            if (xrng) rr = xrng[(int64_t)gz * g.S[1] + gy0 + j];
            if (xin && sx >= rr.x && sx <= rr.y) st[j] = P[j * rowst];
          }
        }
      };
      auto store_plane = [&]() {
#pragma unroll
        for (int j = 0; j < RX; j++)
          if (j < ny) s_g[wave][j * 64 + lane] = st[j];
      };
      const float *pg = reinterpret_cast<const float *>(&s_g[wave][0]) + kh * 32 + col;
      load_plane(zlo);
      store_plane();
      wave_sync();
      for (int cz = zlo; cz <= zhi; cz++) {
        if (cz < zhi) load_plane(cz + 1);
        const float wz = s_w[wave][2][cz][col];
        for (int j = 0; j < ny; j++) {
          const float wzy = wz * s_w[wave][1][ylo + j][col];
#pragma unroll
          for (int p = 0; p < XP; p++) {
            const float a = wzy * wx[p];
            const float bv = pg[(j * 8 + 2 * p) * 32];
            if (p & 1) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bv, acc1, 0, 0, 0);
            else acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bv, acc0, 0, 0, 0);
          }
        }
        if (cz < zhi) {
          wave_sync();
          store_plane();
          wave_sync();
        }
      }
    }
    // ---- epilogue: lane holds sample 8(r/4) + 4 kh + r%4, real col
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int smp = 8 * (r >> 2) + 4 * kh + (r & 3);
      const int64_t i = s_idx[wave][smp];
      if (i < 0 || ((slow >> smp) & 1ull)) continue;
      float val = acc0[r] + acc1[r];
      float *o = reinterpret_cast<float *>(out + ch * chunk_k + i * CCT) + col;
      if (sub) val -= reinterpret_cast<const float *>(sub + ch * chunk_k + i * CCT)[col];
      if (weighted) val = (val * s_kw[wave][smp]) * s_tp[wave][smp];
      *o = val;
    }
    // ---- samples whose taps leave the region: lane = coil, taps from the grid
    while (slow) {
      const int smp = __builtin_ctzll(slow);
      slow &= slow - 1;
      if (lane < CCT) {
        const int64_t i = s_idx[wave][smp];
        const float4 k = coords[i];
        const float dk[3] = {k.x, k.y, k.z};
        int s3[3], n3[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
          int e;
          axis_range(g, a, dk[a], s3[a], e);
          n3[a] = e - s3[a] + 1;
        }
        const float2 *Gc = grid + ch * chunk_grid + lane;
        float ar = 0.f, ai = 0.f;
        for (int lz = s3[2]; lz < s3[2] + n3[2]; lz++) {
          const float wz = tap_weight(g, 2, tab_s + g.toff[2], dk[2], lz);
          for (int ly = s3[1]; ly < s3[1] + n3[1]; ly++) {
            const float wzy = wz * tap_weight(g, 1, tab_s + g.toff[1], dk[1], ly);
            const float2 *row = Gc + ((int64_t)lz * g.S[1] + ly) * g.S[0] * CCT;
            for (int lx = s3[0]; lx < s3[0] + n3[0]; lx++) {
              const float ww = wzy * tap_weight(g, 0, tab_s + g.toff[0], dk[0], lx);
              const float2 q = row[(int64_t)lx * CCT];
              ar += ww * q.x;
              ai += ww * q.y;
            }
          }
        }
        const int64_t oi = ch * chunk_k + i * CCT + lane;
        if (sub) {
          const float2 d = sub[oi];
          ar -= d.x;
          ai -= d.y;
        }
        if (weighted) {
          const float tp = taper[i];
          ar = (ar * k.w) * tp;
          ai = (ai * k.w) * tp;
        }
        out[oi] = make_float2(ar, ai);
      }
    }
    wave_sync();
  }
}

// LDS bytes of the matrix-core gather's region (dynamic part)
static size_t gather_mm_lds(int nt) {
  const size_t R = (size_t)nt + 3, RX = (R + 1) & ~(size_t)1;
  return R * R * RX * 16 * sizeof(float2);
}

// ------------------------------------------------------------ scatter
__global__ void k_zero_tiles(float2 *grid, const int32_t *__restrict__ tiles, Geom g, Tiles t,
                             int64_t chunk_grid, const FrameDev *__restrict__ fdt) {
  if (fdt) {  // every frame at once: frame blockIdx.z
    const FrameDev &fd = fdt[blockIdx.z];
    if ((int)blockIdx.x >= fd.nheavy) return;
    grid += fd.grid_off;
    tiles = fd.heavy;
  }
  const int tile = tiles[blockIdx.x];
  const int ch = blockIdx.y;
  const int tx0 = (tile % t.nt[0]) * t.T, ty0 = ((tile / t.nt[0]) % t.nt[1]) * t.T,
            tz0 = (tile / (t.nt[0] * t.nt[1])) * t.T;
  const int ncell = t.T * t.T * t.T * g.CC;
  for (int idx = threadIdx.x; idx < ncell; idx += blockDim.x) {
    const int cell = idx >> g.lgCC, c = idx & (g.CC - 1);
    const int x = tx0 + cell % t.T, y = ty0 + (cell / t.T) % t.T, z = tz0 + cell / (t.T * t.T);
    if (x < g.S[0] && y < g.S[1] && z < g.S[2])
      grid[ch * chunk_grid + (((int64_t)z * g.S[1] + y) * g.S[0] + x) * g.CC + c] = make_float2(0.f, 0.f);
  }
}

// Grid-driven scatter on the matrix cores (8- and 16-coil blocks).  A
// wavefront owns a 4x4x4 tile; its sums T[64 cells][2 CCT reals] are held as
// one 16 x 16 accumulator block per (z plane, 16 reals).  The wavefront
// walks the tile's visits in batches of SM_BATCH samples:
//  A. lane (s, h) = (l >> 1, l & 1) takes visit s of the batch (from the
//     tile's visit list) and writes its kernel weights of tile offsets 2h,
//     2h + 1 on each axis (0 outside the sample's taps) to LDS; a ballot per
//     z plane records which samples reach it.  The next batch's visits,
//     coordinates and coil values are loaded here, so that the loads fly
//     during step B.
//  B. per group of 4 samples and z plane that one of them reaches,
//     T[plane's 16 cells][16 reals] += W[16][4] D[4][16] by
//     v_mfma_f32_16x16x4f32: lane i + 16 k holds A = W[cell i][sample k] =
//     wz (wy wx) from LDS and B = D[sample k][real i], read straight from
//     k-space in this layout.
// One sample's taps cover 2.5 of a tile's 4 z planes on average (5 taps over
// two tiles), and the samples of a group are consecutive visits, mostly of
// one bin row, so most groups skip a plane or two.  Locating and weighting a
// sample takes two lanes, not a whole quad per tile offset, and not 64 lanes
// per cell.  Sums are FP32; only the order of the additions differs from a
// per-sample loop.  Each tile is written once with plain stores (atomics
// for a tile split over several work items).
// PAIR (8-coil blocks): a work item takes coil blocks 2 y and 2 y + 1 as its
// two 16-real blocks, so step A's weights serve both (the second is skipped
// when nch is odd and y is the last pair).
#define SM_BATCH 32
template <int CCT, int KM, bool PAIR>
__global__ __launch_bounds__(256, 4) void k_scatter_mm(
    float2 *__restrict__ grid, const float2 *__restrict__ kin, const float4 *__restrict__ coords,
    const int64_t *__restrict__ vofs, const int32_t *__restrict__ vlist, const int4 *__restrict__ subs, int nsub,
    Geom g, Tiles t, const float *__restrict__ tables, int64_t chunk_grid, int64_t chunk_k, int accum,
    const FrameDev *__restrict__ fdt, const int32_t *__restrict__ xs, int nch) {
  static_assert(CCT == 8 || CCT == 16, "matrix-core scatter needs 8 or 16 coils per cell");
  static_assert(!PAIR || CCT == 8, "coil blocks are paired only at 8 coils");
  constexpr int NB = PAIR ? 2 : CCT / 8;  // 16-real blocks of a cell's coil values (or the pair's blocks)
  constexpr int NG = SM_BATCH / 4;       // sample groups of a batch
  __shared__ __align__(16) float s_w[4][SM_BATCH * 12];  // [sample][axis][offset] weights
  extern __shared__ float tab_s[];         // kernel tables (g.ntab floats)
  // tables in LDS: a lookup waits for LDS, not for the L1/L2 round trip
  // (before any wavefront can leave: the whole workgroup reaches the barrier)
  load_tables(tab_s, tables, g.ntab);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (fdt) {  // every frame at once: frame blockIdx.z
    const FrameDev &fd = fdt[blockIdx.z];
    grid += fd.grid_off;
    kin = fd.kwork;
    coords = fd.coords;
    vofs = fd.vofs;
    vlist = fd.vlist;
    subs = fd.subs;
    nsub = fd.nsub;
    chunk_k = fd.Mv * g.CC;
    xs = fd.xsplit;
  }
  const int sidx = scatter_item(blockIdx.x, gridDim.x, wave, t.xrun, xs);
  if (sidx < 0 || sidx >= nsub) return;
  const int4 sp = subs[sidx];
  const int tile = sp.x;
  const int tix = tile % t.nt[0], tiy = (tile / t.nt[0]) % t.nt[1], tiz = tile / (t.nt[0] * t.nt[1]);
  const int t0[3] = {tix * 4, tiy * 4, tiz * 4};
  const int ch = PAIR ? 2 * (int)blockIdx.y : (int)blockIdx.y;
  const bool have1 = !PAIR || ch + 1 < nch;  // (uniform) the pair's second block exists
  const float *kch = reinterpret_cast<const float *>(kin + ch * chunk_k);
  const int64_t v0 = vofs[tile];
  const int32_t *vl = vlist + v0;
  const int vbeg = max(sp.y, 0), vend = min(sp.z, (int)(vofs[tile + 1] - v0));
  const int sl = lane >> 1, h = lane & 1;    // step A: sample slot, offset pair
  const int ci = lane & 15, kq = lane >> 4;  // step B: cell / real, sample of the group
  // internal index of visit v of this work item (the last one past the end:
  // a valid index, so that every lane loads)
  auto visit = [&](int v) -> int { return vl[min(v, vend - 1)]; };
  // coil values of a batch in operand layout: group q, block nb holds real
  // 16 nb + ci of sample 4 q + kq (whose index lane 2 (4 q + kq) holds in
  // idx).  Past the end idx repeats the last visit, whose weights there are
  // 0.  Byte offsets are 32-bit (the launcher checks that a coil block's
  // k-space stays below 4 GiB).
  const char *kb = reinterpret_cast<const char *>(kch);
  const char *kb1 = kb + chunk_k * (int64_t)sizeof(float2);  // PAIR: the second block's k-space
  const char *cb = reinterpret_cast<const char *>(coords);
  auto load_data = [&](int idx, float (&d)[NG][NB]) {
#pragma unroll
    for (int q = 0; q < NG; q++) {
      const uint32_t j = (uint32_t)__shfl(idx, 2 * (4 * q + kq));
#pragma unroll
      for (int nb = 0; nb < NB; nb++) {
        if (PAIR)
          d[q][nb] = (nb == 0 || have1) ? *reinterpret_cast<const float *>((nb ? kb1 : kb) + (j * 64u + 4u * ci))
                                        : 0.0f;
        else
          d[q][nb] = *reinterpret_cast<const float *>(kb + (j * (8u * CCT) + 4u * (16 * nb + ci)));
      }
    }
  };
  mrn_v4f acc[4][NB];
#pragma unroll
  for (int p = 0; p < 4; p++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[p][nb] = mrn_v4f{0.f, 0.f, 0.f, 0.f};
  // coordinates of this batch and the next; visit indices of the next two
  float4 kraw = make_float4(0.f, 0.f, 0.f, 0.f), kraw1 = kraw;
  int inext = 0, inext2 = 0;
  auto batch = [&](int base, const float (&cur)[NG][NB], float (&nxt)[NG][NB]) {
    // ---- A: weights of the batch
    const bool ok = base + sl < vend;
    const float dk[3] = {kraw.x, kraw.y, kraw.z};
    int lo[3], hi[3];  // the sample's taps as tile offsets [lo, hi] per axis
    bool any = ok;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      int s = 0, e = -1;
      const bool in = axis_range(g, a, dk[a], s, e);
      // the first NTb taps (the tiles' bin reach); k_scatter_wide adds the rest
      e = min(e, s + t.NTb - 1);
      lo[a] = max(s - t0[a], 0);
      hi[a] = min(e - t0[a], 3);
      any = any && in && lo[a] <= hi[a];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
      float w2[2];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int q = 2 * h + u;
        const float w = KM == 0 ? tap_weight_clamped(g, a, tab_s + g.toff[a], dk[a], t0[a] + q)
                                : tap_weight<KM>(g, a, tab_s + g.toff[a], dk[a], t0[a] + q);
        w2[u] = (any && q >= lo[a] && q <= hi[a]) ? w : 0.0f;
      }
      *reinterpret_cast<float2 *>(&s_w[wave][sl * 12 + 4 * a + 2 * h]) = make_float2(w2[0], w2[1]);
    }
    // bit 2 s of pm[p]: sample s reaches z plane p of the tile
    unsigned long long pm[4];
#pragma unroll
    for (int p = 0; p < 4; p++) pm[p] = __ballot(any && h == 0 && p >= lo[2] && p <= hi[2]);
    // ---- ahead: the next batch's coil values, the coordinates of the batch
    // after it (step A of the next batch starts with coordinates loaded one
    // batch earlier still), and the visit indices one batch further on
    const int nbase = base + SM_BATCH;
    if (nbase < vend) load_data(inext, nxt);
    float4 kraw2 = kraw1;
    if (nbase + SM_BATCH < vend) kraw2 = *reinterpret_cast<const float4 *>(cb + (uint32_t)inext2 * 16u);
    const int inext3 = visit(nbase + 2 * SM_BATCH + sl);
    kraw = kraw1;
    kraw1 = kraw2;
    inext = inext2;
    inext2 = inext3;
    wave_sync();
    // ---- B: accumulate, plane by plane
#pragma unroll
    for (int q = 0; q < NG; q++) {
      const float *w = &s_w[wave][(4 * q + kq) * 12];
      const float wyx = w[4 + (ci >> 2)] * w[ci & 3];
      const mrn_v4f wz = *reinterpret_cast<const mrn_v4f *>(w + 8);
#pragma unroll
      for (int p = 0; p < 4; p++) {
        if ((pm[p] >> (8 * q)) & 0xFFull) {
          const float a = wz[p] * wyx;
#pragma unroll
          for (int nb = 0; nb < NB; nb++)
            if (nb == 0 || have1)
              acc[p][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, cur[q][nb], acc[p][nb], 0, 0, 0);
        }
      }
    }
    wave_sync();
  };
  float dA[NG][NB], dB[NG][NB];
  if (vbeg < vend) {
    const int i0 = visit(vbeg + sl);
    inext = visit(vbeg + SM_BATCH + sl);
    inext2 = visit(vbeg + 2 * SM_BATCH + sl);
    kraw = *reinterpret_cast<const float4 *>(cb + (uint32_t)i0 * 16u);
    kraw1 = *reinterpret_cast<const float4 *>(cb + (uint32_t)inext * 16u);
    load_data(i0, dA);
  }
  for (int base = vbeg; base < vend; base += 2 * SM_BATCH) {
    batch(base, dA, dB);
    if (base + SM_BATCH >= vend) break;
    batch(base + SM_BATCH, dB, dA);
  }
  // accumulator block (p, nb), register r of lane l: cell 16 p + 4 (l / 16) + r,
  // real 16 nb + l % 16 of the cell's CCT coils (PAIR: real l % 16 of block
  // ch + nb)
#pragma unroll
  for (int p = 0; p < 4; p++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int cx = t0[0] + r, cy = t0[1] + kq, cz = t0[2] + p;
      if (cx >= g.S[0] || cy >= g.S[1] || cz >= g.S[2]) continue;
      float *dst = reinterpret_cast<float *>(grid + ch * chunk_grid +
                                             (((int64_t)cz * g.S[1] + cy) * g.S[0] + cx) * CCT) + ci;
#pragma unroll
      for (int nb = 0; nb < NB; nb++) {
        if (nb > 0 && !have1) continue;
        const float v = acc[p][nb][r];
        float *d = PAIR ? dst + nb * 2 * chunk_grid : dst + 16 * nb;
        if (sp.w) {
          if (v != 0.f) atomicAdd(d, v);
        } else if (accum) {  // the tile's only work item: a plain read-modify-write
          *d += v;
        } else {
          *d = v;
        }
      }
    }
}

// Grid-driven scatter on the vector ALUs (any coil block; the matrix-core
// form above takes 8- and 16-coil blocks unless told otherwise).  Lane l
// owns cell (l&3, (l>>2)&3, l>>4) of the wavefront's 4x4x4 tile and the CCT
// coil sums of coil block blockIdx.y; kin holds kw * taper * data (internal
// sample order).  The wavefront walks the visits in batches of SG_BATCH
// samples:
//  A. lane (s, q) = (l>>2, l&3) reads visit s of the batch from the tile's
//     visit list, loads its coordinates and, for each axis, the kernel weight
//     of tile offset q (0
//     outside the sample's taps).  The weights and the sample's CCT coil
//     values go to LDS; a ballot marks the samples with a tap in the tile.
//  B. for each marked sample, every lane forms its weight
//     (wz * wy) * wx from LDS and adds weight * data to its CCT sums.
// Per sample the wavefront thus evaluates 12 weights instead of 192, and a
// sample that misses the tile costs only its share of step A.  The sums
// accumulate in visit order, as a lane-per-sample walk would.
#define SG_BATCH 16
template <int CCT>
__global__ __launch_bounds__(256, 6) void k_scatter_gd(
    float2 *__restrict__ grid, const float2 *__restrict__ kin, const float4 *__restrict__ coords,
    const int64_t *__restrict__ vofs, const int32_t *__restrict__ vlist, const int4 *__restrict__ subs, int nsub,
    Geom g, Tiles t, const float *__restrict__ tables, int64_t chunk_grid, int64_t chunk_k, int accum,
    const FrameDev *__restrict__ fdt, const int32_t *__restrict__ xs) {
  __shared__ float s_w[4][SG_BATCH * 12];    // [sample][axis][offset] weights
  __shared__ float2 s_d[4][SG_BATCH * CCT];  // [sample][coil] data
  extern __shared__ float tab_s[];           // kernel tables (g.ntab floats)
  load_tables(tab_s, tables, g.ntab);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (fdt) {  // every frame at once: frame blockIdx.z
    const FrameDev &fd = fdt[blockIdx.z];
    grid += fd.grid_off;
    kin = fd.kwork;
    coords = fd.coords;
    vofs = fd.vofs;
    vlist = fd.vlist;
    subs = fd.subs;
    nsub = fd.nsub;
    chunk_k = fd.Mv * g.CC;
    xs = fd.xsplit;
  }
  const int sidx = scatter_item(blockIdx.x, gridDim.x, wave, t.xrun, xs);
  if (sidx < 0 || sidx >= nsub) return;
  const int4 sp = subs[sidx];
  const int tile = sp.x;
  const int tix = tile % t.nt[0], tiy = (tile / t.nt[0]) % t.nt[1], tiz = tile / (t.nt[0] * t.nt[1]);
  const int t0[3] = {tix * 4, tiy * 4, tiz * 4};
  const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
  const int ch = blockIdx.y;
  const float2 *kch = kin + ch * chunk_k;
  const int64_t v0 = vofs[tile];
  const int32_t *vl = vlist + v0;
  const int vbeg = max(sp.y, 0), vend = min(sp.z, (int)(vofs[tile + 1] - v0));
  float ar[CCT], ai[CCT];
#pragma unroll
  for (int c = 0; c < CCT; c++) ar[c] = ai[c] = 0.f;
  const int sl = lane >> 2, q = lane & 3;
  constexpr int CPL = CCT >= 4 ? CCT / 4 : 1;  // coil values a lane stages
  for (int base = vbeg; base < vend; base += SG_BATCH) {
    // ---- A: weights and data of the batch
    const int v = base + sl;
    const bool ok = v < vend;
    const int64_t i = vl[min(v, vend - 1)];
    float2 dv[CPL];
    if (q * CPL < CCT) {
#pragma unroll
      for (int c = 0; c < CPL; c++) dv[c] = kch[i * CCT + q * CPL + c];
    }
    const float4 k = ok ? coords[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float dk[3] = {k.x, k.y, k.z};
    unsigned long long hit = ~0ull;
    int sta[3], ena[3];
    bool oka[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      sta[a] = 0;
      ena[a] = -1;
      oka[a] = ok && axis_range(g, a, dk[a], sta[a], ena[a]);
      // the first NTb taps (the tiles' bin reach); k_scatter_wide adds the rest
      ena[a] = min(ena[a], sta[a] + t.NTb - 1);
    }
    const bool inall = oka[0] && oka[1] && oka[2];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int l = t0[a] + q;
      const bool in = inall && l >= sta[a] && l <= ena[a];
      s_w[wave][sl * 12 + a * 4 + q] = in ? tap_weight(g, a, tab_s + g.toff[a], dk[a], l) : 0.0f;
      unsigned long long m = __ballot(in);
      m |= m >> 1;
      m |= m >> 2;
      hit &= m;  // bit 4s: sample s has a tap in the tile on every axis so far
    }
    hit &= 0x1111111111111111ull;
    if (q * CPL < CCT && ok) {
#pragma unroll
      for (int c = 0; c < CPL; c++) s_d[wave][sl * CCT + q * CPL + c] = dv[c];
    }
    wave_sync();
    // ---- B: accumulate the samples that reach the tile
    while (hit) {
      const int s = __builtin_ctzll(hit) >> 2;
      hit &= hit - 1;
      const float *wp = &s_w[wave][s * 12];
      const float w = (wp[8 + lz] * wp[4 + ly]) * wp[lx];
      const float2 *dp = &s_d[wave][s * CCT];
#pragma unroll
      for (int c = 0; c < CCT; c++) {
        const float2 d = dp[c];
        ar[c] += w * d.x;
        ai[c] += w * d.y;
      }
    }
    wave_sync();
  }
  const int cx = t0[0] + lx, cy = t0[1] + ly, cz = t0[2] + lz;
  if (cx >= g.S[0] || cy >= g.S[1] || cz >= g.S[2]) return;
  float2 *dst = grid + ch * chunk_grid + (((int64_t)cz * g.S[1] + cy) * g.S[0] + cx) * CCT;
  if (sp.w) {
#pragma unroll
    for (int c = 0; c < CCT; c++) {
      if (ar[c] != 0.f) atomicAdd(&dst[c].x, ar[c]);
      if (ai[c] != 0.f) atomicAdd(&dst[c].y, ai[c]);
    }
  } else if (accum) {
#pragma unroll
    for (int c = 0; c < CCT; c++) {
      const float2 o = dst[c];
      dst[c] = make_float2(o.x + ar[c], o.y + ai[c]);
    }
  } else {
#pragma unroll
    for (int c = 0; c < CCT; c++) dst[c] = make_float2(ar[c], ai[c]);
  }
}

// The samples with more than NTb taps on some axis (fr.wide): the tile
// scatter adds their first NTb taps on every axis (the box its bins reach),
// and this kernel adds the rest -- the taps past that box, on the window's far
// edge -- with float atomics after the tiles are written.  Thread = (listed
// sample, coil) of coil block blockIdx.y.  Such samples sit exactly on a
// window edge, so the list is short (often empty), and each has one plane of
// edge taps per wide axis (25 of 150 taps for W = 5 on one axis).
__global__ void k_scatter_wide(float2 *grid, const float2 *__restrict__ kin, const float4 *__restrict__ coords,
                               const int32_t *__restrict__ wide, int nwide, Geom g, int NTb,
                               const float *__restrict__ tables, int64_t chunk_grid, int64_t chunk_k,
                               const FrameDev *__restrict__ fdt) {
  if (fdt) {  // every frame at once: frame blockIdx.z
    const FrameDev &fd = fdt[blockIdx.z];
    grid += fd.grid_off;
    kin = fd.kwork;
    coords = fd.coords;
    wide = fd.wide;
    nwide = fd.nwide;
    chunk_k = fd.Mv * g.CC;
  }
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nwide * g.CC) return;
  const int c = (int)(idx & (g.CC - 1));
  const int64_t i = wide[idx >> g.lgCC];
  const int ch = blockIdx.y;
  const float4 k = coords[i];
  const float dk[3] = {k.x, k.y, k.z};
  int s[3], e[3];
  for (int a = 0; a < 3; a++)
    if (!axis_range(g, a, dk[a], s[a], e[a])) return;
  const float2 v = kin[ch * chunk_k + i * g.CC + c];
  float2 *G = grid + ch * chunk_grid + c;
  for (int lz = s[2]; lz <= e[2]; lz++) {
    const float wz = tap_weight(g, 2, tables + g.toff[2], dk[2], lz);
    for (int ly = s[1]; ly <= e[1]; ly++) {
      const float wzy = wz * tap_weight(g, 1, tables + g.toff[1], dk[1], ly);
      float2 *row = G + ((int64_t)lz * g.S[1] + ly) * g.S[0] * g.CC;
      for (int lx = s[0]; lx <= e[0]; lx++) {
        // inside the first NTb taps on every axis: the tile scatter's
        if (lx < s[0] + NTb && ly < s[1] + NTb && lz < s[2] + NTb) continue;
        const float w = wzy * tap_weight(g, 0, tables + g.toff[0], dk[0], lx);
        atomicAdd(&row[(int64_t)lx * g.CC].x, w * v.x);
        atomicAdd(&row[(int64_t)lx * g.CC].y, w * v.y);
      }
    }
  }
}

// wide[i] = 1 when valid sample i has more than NTb taps on some axis
__global__ void k_wide_flag(const float4 *__restrict__ coords, int64_t Mv, Geom g, int NTb, int32_t *flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Mv) return;
  const float4 k = coords[i];
  const float dk[3] = {k.x, k.y, k.z};
  bool w = false;
  for (int a = 0; a < 3; a++) {
    int s, e;
    axis_range(g, a, dk[a], s, e);
    w = w || e - s + 1 > NTb;
  }
  flag[i] = w ? 1 : 0;
}

__global__ void k_compact64(const int32_t *__restrict__ flag, const int32_t *__restrict__ off, int64_t n,
                            int32_t *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) out[off[i]] = (int32_t)i;
}

// kout = (kin * kw) * taper for the valid samples of one frame, every block
__global__ void k_weight(float2 *kout, const float2 *kin, const float4 *__restrict__ coords,
                         const float *__restrict__ taper, int64_t Mv, int CC, int lgCC, int64_t total) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = (idx >> lgCC) % Mv;
    const float w = coords[i].w, tp = taper[i];
    const float2 d = kin[idx];
    kout[idx] = make_float2((d.x * w) * tp, (d.y * w) * tp);
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

// LDS bytes of the tile-staged gather (0 when it does not apply)
static size_t gather_tile_lds(const mrn_op *op, int nt) {
  const Geom &g = op->g;
  if ((g.CC != 8 && g.CC != 16) || nt < 3 || nt > 6 || op->t.T != 4 || (op->t.T % op->t.B) != 0) return 0;
  const size_t R = (size_t)nt + 3;
  const size_t bytes = R * R * R * (size_t)g.CC * sizeof(float2);
  return bytes <= 96 * 1024 ? bytes : 0;
}

template <int NT, int CCT>
static int launch_gather_tile(mrn_op *op, const float2 *grid, float2 *out, const mrn::FrameData &fr,
                              const float2 *sub, int wflag, size_t lds, hipStream_t st) {
  // once per instantiation (thread-safe static initialisation)
  static const hipError_t attr = hipFuncSetAttribute((const void *)k_gather_tile<NT, CCT>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
  MRN_HIP(op, attr);
  const Geom &g = op->g;
  int64_t cg = (int64_t)g.S[0] * g.S[1] * g.S[2] * g.CC, ck = fr.Mv * g.CC;
  hipLaunchKernelGGL((k_gather_tile<NT, CCT>), dim3(fr.ngsub, op->nch), dim3(256), lds, st, grid, out, fr.coords,
                     sub, fr.bin_start, fr.gsubs, g, op->t, op->tables_d, cg, ck, fr.taper, wflag,
                     fft_x_pruned(op) ? fr.xrng : nullptr);
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

template <int NT>
static int launch_gather_mm(mrn_op *op, const float2 *grid, float2 *out, const mrn::FrameData &fr,
                            const float2 *sub, int wflag, hipStream_t st) {
  static const hipError_t attr = hipFuncSetAttribute((const void *)k_gather_mm<NT>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
  MRN_HIP(op, attr);
  const Geom &g = op->g;
  int64_t cg = (int64_t)g.S[0] * g.S[1] * g.S[2] * g.CC, ck = fr.Mv * g.CC;
  const size_t lds = gather_mm_lds(NT);
  hipLaunchKernelGGL((k_gather_mm<NT>), dim3(fr.ngsub, op->nch), dim3(256), lds, st, grid, out, fr.coords, sub,
                     fr.bin_start, fr.gsubs, g, op->t, op->tables_d, cg, ck, fr.taper, wflag,
                     fft_x_pruned(op) ? fr.xrng : nullptr);
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

template <int NT>
static int launch_gather_box(mrn_op *op, const float2 *grid, float2 *out, const mrn::FrameData &fr,
                             const float2 *sub, int wflag, hipStream_t st) {
  const Geom &g = op->g;
  int64_t cg = (int64_t)g.S[0] * g.S[1] * g.S[2] * g.CC, ck = fr.Mv * g.CC;
  hipLaunchKernelGGL((k_gather_box<NT>), dim3((fr.ngsub + 3) / 4, op->nch), dim3(256), sizeof(float) * g.ntab, st,
                     grid, out, fr.coords, sub, fr.bin_start, fr.gsubs, fr.ngsub, g, op->t, op->tables_d, cg, ck,
                     fr.taper, wflag, fft_x_pruned(op) ? fr.xrng : nullptr);
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

// Direct gather over all samples, or over the samples listed in `list`.
// fdt: every frame of the op at once (op->fdev; out, sub from the table, sub
// when fsub), grid dimension z = frame.
static int launch_gather_direct(mrn_op *op, const float2 *grid, float2 *out, const mrn::FrameData &fr,
                                const float2 *sub, int wflag, const int32_t *list, int64_t nlist,
                                hipStream_t st, const FrameDev *fdt = nullptr, int fsub = 0) {
  const Geom &g = op->g;
  // coils per lane: 4 by default.  A wave then serves 16 samples (4 lanes
  // each) instead of 8, whose taps overlap more, so fewer of its loads miss
  // the L1 (16-coil blocks at 256^3, pipelined rows: 5.5 ms with 4 coils per
  // lane, 5.7 with 2, 7.5 with 8)
  int nc = op->gather_coils > 0 ? op->gather_coils : 4;
  if (nc > g.CC) nc = g.CC;
  const int64_t n = fdt ? op->fmax_Mv : list ? nlist : fr.Mv;
  if (n == 0) return MRN_OK;
  dim3 g3(nblk(n * (g.CC / nc), 256), op->nch, fdt ? op->F : 1);
  int64_t cg = (int64_t)g.S[0] * g.S[1] * g.S[2] * g.CC, ck = fr.Mv * g.CC;
  switch (taps_main(g) * 16 + nc) {
#define MRN_G1(n, c)                                                                                \
  case n * 16 + c:                                                                                  \
    hipLaunchKernelGGL((k_gather<n, c>), g3, dim3(256), sizeof(float) * g.ntab, st, grid, out,     \
                       fr.coords, sub, fr.Mv, g, op->tables_d, cg, ck, fr.taper, wflag,            \
                       op->t.xrun_gather, list, nlist, fdt, fsub);                                  \
    break;
#define MRN_G(n) MRN_G1(n, 1) MRN_G1(n, 2) MRN_G1(n, 4) MRN_G1(n, 8)
    MRN_G(1) MRN_G(2) MRN_G(3) MRN_G(4) MRN_G(5) MRN_G(6) MRN_G(7) MRN_G(8)
#undef MRN_G
#undef MRN_G1
    default:
      return fail(op, MRN_EUNSUPPORTED, "kernel width above 8 taps");
  }
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

// Gather: the tile-staged kernel over the dense tiles and the direct kernel
// over the samples of the sparse ones, or the direct kernel over everything.
int launch_gather(mrn_op *op, const float2 *grid, float2 *out, const mrn::FrameData &fr,
                  const float2 *sub, bool weighted, hipStream_t st) {
  const Geom &g = op->g;
  if (fr.Mv == 0) return MRN_OK;
  const int wf = weighted ? 1 : 0;
  const int nt = taps_main(g);
  const size_t lds = gather_tile_lds(op, nt);
  if (op->gather_tiles == 2 && g.CC == 16 && op->t.T == 4 && fr.ngsub > 0 && nt >= 3 && nt <= 5 && op->side) {
    // the box gather (matrix cores, LDS) of the dense tiles on the side
    // stream, beside the direct gather (texture path) of the sparse ones:
    // the two kernels are bound by different units.  Fork and join by events
    // (captured into a graph as two branches).
    MRN_HIP(op, hipEventRecord(op->ev_fork, st));
    MRN_HIP(op, hipStreamWaitEvent(op->side, op->ev_fork, 0));
    int rc = MRN_OK;
    switch (nt) {
      case 3: rc = launch_gather_box<3>(op, grid, out, fr, sub, wf, op->side); break;
      case 4: rc = launch_gather_box<4>(op, grid, out, fr, sub, wf, op->side); break;
      default: rc = launch_gather_box<5>(op, grid, out, fr, sub, wf, op->side); break;
    }
    if (rc == MRN_OK && fr.nsparse > 0)
      rc = launch_gather_direct(op, grid, out, fr, sub, wf, fr.sparse, fr.nsparse, st);
    // join even after an error, so that a capture in progress stays well formed
    MRN_HIP(op, hipEventRecord(op->ev_join, op->side));
    MRN_HIP(op, hipStreamWaitEvent(st, op->ev_join, 0));
    return rc;
  }
  if (lds && fr.ngsub > 0 && op->gather_tiles) {
    int rc = MRN_EUNSUPPORTED;
    switch (nt * 100 + g.CC) {
      case 308: rc = launch_gather_tile<3, 8>(op, grid, out, fr, sub, wf, lds, st); break;
      case 316: rc = launch_gather_mm<3>(op, grid, out, fr, sub, wf, st); break;
      case 408: rc = launch_gather_tile<4, 8>(op, grid, out, fr, sub, wf, lds, st); break;
      case 416: rc = launch_gather_mm<4>(op, grid, out, fr, sub, wf, st); break;
      case 508: rc = launch_gather_tile<5, 8>(op, grid, out, fr, sub, wf, lds, st); break;
      case 516: rc = launch_gather_mm<5>(op, grid, out, fr, sub, wf, st); break;
      case 608: rc = launch_gather_tile<6, 8>(op, grid, out, fr, sub, wf, lds, st); break;
      case 616: rc = launch_gather_tile<6, 16>(op, grid, out, fr, sub, wf, lds, st); break;  // region above 96 KB
      default: break;
    }
    if (rc != MRN_EUNSUPPORTED) {
      if (rc != MRN_OK) return rc;
      return fr.nsparse > 0 ? launch_gather_direct(op, grid, out, fr, sub, wf, fr.sparse, fr.nsparse, st)
                            : MRN_OK;
    }
  }
  return launch_gather_direct(op, grid, out, fr, sub, wf, nullptr, 0, st);
}

// Gather of every frame at once (the frames' grids side by side in op->grid,
// the results in their kwork arrays), the data subtracted when fsub.
int launch_gather_frames(mrn_op *op, int fsub, hipStream_t st) {
  return launch_gather_direct(op, op->grid, nullptr, op->frames[0], nullptr, 1, nullptr, 0, st, op->fdev, fsub);
}

// Zero the tiles that several work items flush with atomics (part of the scatter).
int launch_scatter_zero(mrn_op *op, float2 *grid, const mrn::FrameData &fr, hipStream_t st,
                        const FrameDev *fdt) {
  const Geom &g = op->g;
  const int nh = fdt ? op->fmax_nheavy : fr.nheavy;
  if (nh > 0) {
    const int64_t cg = (int64_t)g.S[0] * g.S[1] * g.S[2] * g.CC;
    hipLaunchKernelGGL(k_zero_tiles, dim3(nh, op->nch, fdt ? op->F : 1), dim3(256), 0, st, grid, fr.heavy, g,
                       op->t, cg, fdt);
    MRN_HIP(op, hipGetLastError());
  }
  return MRN_OK;
}

// The tile scatter proper; launch_scatter_zero must precede it unless
// accum (add into the grid's contents: tiles split over several work items
// add with atomics, the others read, add and store).
int launch_scatter_tiles(mrn_op *op, float2 *grid, const float2 *kin, const mrn::FrameData &fr,
                         hipStream_t st, bool accum, const FrameDev *fdt) {
  const int af = accum ? 1 : 0;
  const Geom &g = op->g;
  const Tiles &t = op->t;
  // fdt: every frame at once (grid dimension z = frame; the frames' largest
  // work lists size the grid, each frame's kernel reads its own)
  const int nsub = fdt ? op->fmax_nsub : fr.nsub, nwide = fdt ? op->fmax_nwide : fr.nwide;
  const int64_t Mv = fdt ? op->fmax_Mv : fr.Mv;
  int64_t cg = (int64_t)g.S[0] * g.S[1] * g.S[2] * g.CC, ck = fr.Mv * g.CC;
  // per-XCD ranges (xcd_run < 0): MRN_XCDS workgroup columns, each as long
  // as the longest range needs
  const int xmax = fdt ? op->fmax_xsplit : fr.xsplit_max;
  const bool slabs = t.xrun < 0 && (fdt || fr.xsplit);
  dim3 g3(slabs ? MRN_XCDS * ((xmax + 3) / 4) : (nsub + 3) / 4, op->nch, fdt ? op->F : 1);
  const size_t sh = sizeof(float) * g.ntab;  // kernel tables in LDS
  // no work item: a frame without valid samples whose x passes are limited to
  // its (empty) k-space support -- nothing reads the grid
  if (nsub == 0) return MRN_OK;
  const bool exact = g.kmode == MRN_KMODE_EXACT_KB;
  // the matrix-core form addresses a coil block's k-space and the coordinates
  // with 32-bit byte offsets
  const bool off32 = Mv * g.CC * 8 < (1ll << 32) && Mv * 16 < (1ll << 32);
  if ((g.CC == 16 || g.CC == 8) && !op->scatter_valu && off32) {
    // 8-coil blocks go in pairs (one work item, both blocks) when there are two or more
    const bool pair = g.CC == 8 && op->nch >= 2;
    dim3 gm = g3;
    if (pair) gm.y = (unsigned)((op->nch + 1) / 2);
#define MRN_MM(n, km, pr)                                                                                \
  hipLaunchKernelGGL((k_scatter_mm<n, km, pr>), gm, dim3(256), sh, st, grid, kin, fr.coords, fr.vofs,   \
                     fr.vlist, fr.subs, fr.nsub, g, t, op->tables_d, cg, ck, af, fdt, fr.xsplit, op->nch)
    if (g.CC == 16 && !exact) MRN_MM(16, 0, false);
    else if (g.CC == 16) MRN_MM(16, 1, false);
    else if (pair && !exact) MRN_MM(8, 0, true);
    else if (pair) MRN_MM(8, 1, true);
    else if (!exact) MRN_MM(8, 0, false);
    else MRN_MM(8, 1, false);
#undef MRN_MM
  } else {
    switch (g.CC) {
#define MRN_S(n)                                                                                 \
  case n:                                                                                        \
    hipLaunchKernelGGL((k_scatter_gd<n>), g3, dim3(256), sh, st, grid, kin, fr.coords, fr.vofs,  \
                       fr.vlist, fr.subs, fr.nsub, g, t, op->tables_d, cg, ck, af, fdt, fr.xsplit);   \
    break;
    MRN_S(1) MRN_S(2) MRN_S(4) MRN_S(8) MRN_S(16)
#undef MRN_S
      default:
        return fail(op, MRN_EUNSUPPORTED, "coil block size");
    }
  }
  MRN_HIP(op, hipGetLastError());
  if (nwide > 0) {  // after the tiles' plain stores
    hipLaunchKernelGGL(k_scatter_wide, dim3(nblk((int64_t)nwide * g.CC, 256), op->nch, fdt ? op->F : 1), dim3(256), 0,
                       st, grid, kin, fr.coords, fr.wide, fr.nwide, g, t.NTb, op->tables_d, cg, ck, fdt);
    MRN_HIP(op, hipGetLastError());
  }
  return MRN_OK;
}

int launch_scatter(mrn_op *op, float2 *grid, const float2 *kin, const mrn::FrameData &fr,
                   hipStream_t st, bool accum) {
  if (!accum) MRN_TRY(launch_scatter_zero(op, grid, fr, st));
  return launch_scatter_tiles(op, grid, kin, fr, st, accum);
}

// Scatter of every frame at once, from their kwork arrays into their grids.
int launch_scatter_frames(mrn_op *op, hipStream_t st) {
  MRN_TRY(launch_scatter_zero(op, op->grid, op->frames[0], st, op->fdev));
  return launch_scatter_tiles(op, op->grid, nullptr, op->frames[0], st, false, op->fdev);
}

int launch_weight(mrn_op *op, float2 *kout, const float2 *kin, const mrn::FrameData &fr,
                  hipStream_t st) {
  const int64_t total = (int64_t)op->nch * fr.Mv * op->g.CC;
  if (total == 0) return MRN_OK;
  int64_t b = (total + 255) / 256;
  if (b > 65536) b = 65536;
  hipLaunchKernelGGL(k_weight, dim3((unsigned)b), dim3(256), 0, st, kout, kin, fr.coords, fr.taper,
                     fr.Mv, op->g.CC, op->g.lgCC, total);
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

int recompute_taper(mrn_op *op, mrn::FrameData &fr, hipStream_t st) {
  if (fr.Mv == 0) return MRN_OK;
  hipLaunchKernelGGL(k_taper, dim3(nblk(fr.Mv, 256)), dim3(256), 0, st, fr.traw, fr.M, fr.perm, fr.Mv,
                     op->g, fr.taper);
  MRN_HIP(op, hipGetLastError());
  return MRN_OK;
}

// k-space support of a frame, per grid row (z, y): the x cells that any
// sample's taps reach.  A sample binned at b on an axis (b = first tap / B)
// reaches cells [b B, b B + B - 1 + W], W = floor(2 dwin) taps beyond the
// first on a gridded axis, 0 otherwise.  Pass 1: one wavefront per bin row
// (bz, by) finds its first and last nonempty bin in x.
__global__ void k_binrow_x(const int32_t *__restrict__ bin_start, Tiles t, int2 *brow) {
  const int row = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= t.nb[1] * t.nb[2]) return;  // whole wavefronts
  const int32_t *bs = bin_start + (int64_t)row * t.nb[0];
  int first = INT_MAX, last = -1;
  for (int b0 = 0; b0 < t.nb[0]; b0 += 64) {
    const int b = b0 + lane;
    const bool ne = b < t.nb[0] && bs[b + 1] > bs[b];
    const unsigned long long m = __ballot(ne);
    if (m) {
      if (first == INT_MAX) first = b0 + __builtin_ctzll(m);
      last = b0 + 63 - __builtin_clzll(m);
    }
  }
  if (lane == 0) brow[row] = make_int2(first, last);
}

// Pass 2: row (z, y) of cells takes the x extent of every bin row whose
// reach covers it; {1, 0} (empty) when none does.
__global__ void k_row_xrange(const int2 *__restrict__ brow, Tiles t, int S0, int S1, int S2, int W0, int W1,
                             int W2, int2 *xrng) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (int64_t)S1 * S2) return;
  const int y = (int)(r % S1), z = (int)(r / S1);
  const int B = t.B;
  // bins b with b B <= c <= b B + B - 1 + W: b in [ceil((c - B + 1 - W) / B), c / B]
  const int z0 = max(z - B + 1 - W2, 0), y0 = max(y - B + 1 - W1, 0);
  const int bz0 = (z0 + B - 1) / B, bz1 = min(z / B, t.nb[2] - 1);
  const int by0 = (y0 + B - 1) / B, by1 = min(y / B, t.nb[1] - 1);
  int first = INT_MAX, last = -1;
  for (int bz = bz0; bz <= bz1; bz++)
    for (int by = by0; by <= by1; by++) {
      const int2 v = brow[bz * t.nb[1] + by];
      first = min(first, v.x);
      last = max(last, v.y);
    }
  xrng[r] = last < 0 ? make_int2(1, 0) : make_int2(first * B, min(last * B + B - 1 + W0, S0 - 1));
}

// Pass 3 (y passes pruned too): plane z's rows with support span [y0, y1];
// the rows outside it are marked {INT_MAX, INT_MIN} (no y pass reads them,
// so the inverse x pass need not write them).
__global__ void k_plane_yrange(int2 *xrng, int S1, int S2, int2 *yrng) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= S2) return;
  int2 *row = xrng + (int64_t)z * S1;
  int y0 = S1, y1 = -1;
  for (int y = 0; y < S1; y++)
    if (row[y].x <= row[y].y) {
      y0 = min(y0, y);
      y1 = y;
    }
  yrng[z] = y1 < 0 ? make_int2(1, 0) : make_int2(y0, y1);
  for (int y = 0; y < S1; y++)
    if (y < y0 || y > y1) row[y] = make_int2(INT_MAX, INT_MIN);
}

// Cost of scatter work item i (entry n: 0, for the exclusive scan): its
// visits plus a fixed part for the tile's zeroing and epilogue stores.
__global__ void k_item_cost(const int4 *__restrict__ subs, int n, int64_t *cost) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cost[i] = max(subs[i].z - subs[i].y, 0) + 32;
  else if (i == n) cost[i] = 0;
}

// Boundaries of MRN_XCDS contiguous ranges of the work list of equal cost
// (pre: exclusive prefix of the costs, n + 1 entries): xs[x] is the first
// item whose prefix reaches x/MRN_XCDS of the total.  One thread per entry.
__global__ void k_xcd_split(const int64_t *__restrict__ pre, int n, int32_t *xs) {
  const int x = threadIdx.x;
  if (x > MRN_XCDS) return;
  if (x == 0 || x == MRN_XCDS) {
    xs[x] = x == 0 ? 0 : n;
    return;
  }
  const int64_t target = pre[n] * x / MRN_XCDS;
  int lo = 0, hi = n;  // first i in [0, n] with pre[i] >= target
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] >= target) hi = mid;
    else lo = mid + 1;
  }
  xs[x] = lo;
}

// Sort one frame's samples into tiles and build the scatter work list.  The
// frame's arrays and the bind's scratch only grow (FrameCaps, BindScratch), so
// binding a trajectory no larger than the last one allocates nothing; the
// device is synchronised first, as the hipFree of the old arrays did, so that
// no launch still reading the old trajectory sees the new one.
int build_bins(mrn_op *op, mrn::FrameData &fr, int64_t M, const float *kx, const float *ky,
               const float *kz, const float *kw, hipStream_t st) {
  const Geom &g = op->g;
  const Tiles &t = op->t;
  const int nt = t.ntiles, nbin = t.nbins;
  BindScratch &b = op->bind;
  FrameCaps &cp = fr.cap;
  size_t tmp_bytes = 0, b2 = 0;
#define MRN_BH(call)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      fr.has_traj = false;                                                                   \
      return fail(op, e_ == hipErrorOutOfMemory ? MRN_ENOMEM : MRN_EHIP, hipGetErrorString(e_)); \
    }                                                                                        \
  } while (0)
  MRN_BH(hipDeviceSynchronize());
  fr.has_traj = false;
  fr.nwide = 0;
  fr.ngsub = 0;
  fr.nsparse = 0;
  fr.nsub = 0;
  fr.nheavy = 0;
  fr.M = M;
  MRN_BH(grow(b.keys, b.c_keys, M));
  MRN_BH(grow(b.keys2, b.c_keys2, M));
  MRN_BH(grow(b.vals, b.c_vals, M));
  MRN_BH(grow(fr.perm, cp.perm, M));
  MRN_BH(grow(fr.inv, cp.inv, M));
  MRN_BH(grow(b.counts, b.c_counts, (int64_t)nbin + 1));
  MRN_BH(grow(fr.bin_start, cp.bin_start, (int64_t)nbin + 1));
  MRN_BH(hipMemsetAsync(b.counts, 0, sizeof(int32_t) * (nbin + 1), st));
  int bits = 1;
  while ((1u << bits) <= (uint32_t)nbin) bits++;
  {
    size_t b3 = 0;
    if (M > 0)
      MRN_BH(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, b.keys, b.keys2, b.vals, fr.perm, (int)M, 0,
                                                bits, st));
    MRN_BH(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, b.counts, fr.bin_start, nbin + 1, st));
    MRN_BH(hipcub::DeviceScan::ExclusiveSum(nullptr, b3, b.counts, fr.bin_start, nt + 1, st));
    b2 = b2 > b3 ? b2 : b3;
    tmp_bytes = tmp_bytes > b2 ? tmp_bytes : b2;
    MRN_BH(grow(b.tmp, b.c_tmp, (int64_t)tmp_bytes));
  }
  if (M > 0) {
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(M, 256)), dim3(256), 0, st, fr.inv, -1, M);
    hipLaunchKernelGGL(k_traj_keys, dim3(nblk(M, 256)), dim3(256), 0, st, kx, ky, kz, kw, M, g, t,
                       b.keys, b.vals, b.counts);
    MRN_BH(hipGetLastError());
    MRN_BH(hipcub::DeviceRadixSort::SortPairs(b.tmp, tmp_bytes, b.keys, b.keys2, b.vals, fr.perm, (int)M, 0, bits,
                                              st));
  }
  size_t tb = tmp_bytes;
  MRN_BH(hipcub::DeviceScan::ExclusiveSum(b.tmp, tb, b.counts, fr.bin_start, nbin + 1, st));
  int32_t Mv32 = 0;
  MRN_BH(hipMemcpyAsync(&Mv32, fr.bin_start + nbin, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MRN_BH(hipStreamSynchronize(st));
  fr.Mv = Mv32;
  MRN_BH(grow(fr.coords, cp.coords, fr.Mv));
  MRN_BH(grow(fr.taper, cp.taper, fr.Mv));
  if (fr.Mv > 0) {
    hipLaunchKernelGGL(k_traj_finish, dim3(nblk(fr.Mv, 256)), dim3(256), 0, st, kx, ky, kz, kw,
                       fr.perm, fr.Mv, g, fr.coords, fr.taper, fr.inv);
  }
  // samples with more taps than the tiles' bin reach assumes (only possible
  // when 2 dwin is an integer on some gridded axis)
  if (t.NTb < t.NT && fr.Mv > 0) {
    const int64_t n = fr.Mv + 1;
    MRN_BH(grow(b.wflag, b.c_wflag, n));
    MRN_BH(grow(b.woff, b.c_woff, n));
    MRN_BH(hipMemsetAsync(b.wflag, 0, sizeof(int32_t) * n, st));
    hipLaunchKernelGGL(k_wide_flag, dim3(nblk(fr.Mv, 256)), dim3(256), 0, st, fr.coords, fr.Mv, g, t.NTb, b.wflag);
    MRN_BH(hipGetLastError());
    size_t wb = 0;
    MRN_BH(hipcub::DeviceScan::ExclusiveSum(nullptr, wb, b.wflag, b.woff, (int)n, st));
    MRN_BH(grow(b.wtmp, b.c_wtmp, (int64_t)wb));
    MRN_BH(hipcub::DeviceScan::ExclusiveSum(b.wtmp, wb, b.wflag, b.woff, (int)n, st));
    int32_t nw = 0;
    MRN_BH(hipMemcpyAsync(&nw, b.woff + fr.Mv, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    MRN_BH(hipStreamSynchronize(st));
    if (nw > 0) {
      MRN_BH(grow(fr.wide, cp.wide, nw));
      hipLaunchKernelGGL(k_compact64, dim3(nblk(fr.Mv, 256)), dim3(256), 0, st, b.wflag, b.woff, fr.Mv, fr.wide);
      MRN_BH(hipGetLastError());
    }
    fr.nwide = nw;
  }
  // k-space support per grid row, for the x passes
  {
    const int nbr = t.nb[1] * t.nb[2];
    const int64_t nrow = (int64_t)g.S[1] * g.S[2];
    int W[3];
    for (int a = 0; a < 3; a++) W[a] = g.grid_in[a] ? (int)floorf(2.0f * g.dwin[a]) : 0;
    MRN_BH(grow(b.brow, b.c_brow, nbr));
    MRN_BH(grow(fr.xrng, cp.xrng, nrow));
    hipLaunchKernelGGL(k_binrow_x, dim3(nblk((int64_t)nbr * 64, 256)), dim3(256), 0, st, fr.bin_start, t, b.brow);
    hipLaunchKernelGGL(k_row_xrange, dim3(nblk(nrow, 256)), dim3(256), 0, st, b.brow, t, g.S[0], g.S[1], g.S[2],
                       W[0], W[1], W[2], fr.xrng);
    MRN_BH(hipGetLastError());
    if (fft_xy_pruned(op)) {
      MRN_BH(grow(fr.yrng, cp.yrng, g.S[2]));
      hipLaunchKernelGGL(k_plane_yrange, dim3(nblk(g.S[2], 64)), dim3(64), 0, st, fr.xrng, g.S[1], g.S[2], fr.yrng);
      MRN_BH(hipGetLastError());
    } else {  // a null yrng means "every row" to the y passes
      hipFree(fr.yrng);
      fr.yrng = nullptr;
      cp.yrng = 0;
    }
  }
  int32_t *nsub = nullptr, *visits = nullptr, *off = nullptr, *hflag = nullptr, *hoff = nullptr;
  MRN_BH(grow(b.nsub, b.c_nsub, (int64_t)nt + 1));
  MRN_BH(grow(b.visits, b.c_visits, (int64_t)nt + 1));
  MRN_BH(grow(b.off, b.c_off, (int64_t)nt + 1));
  MRN_BH(grow(b.hflag, b.c_hflag, (int64_t)nt + 1));
  MRN_BH(grow(b.hoff, b.c_hoff, (int64_t)nt + 1));
  nsub = b.nsub;
  visits = b.visits;
  off = b.off;
  hflag = b.hflag;
  hoff = b.hoff;
  MRN_BH(hipMemsetAsync(nsub, 0, sizeof(int32_t) * (nt + 1), st));
  MRN_BH(hipMemsetAsync(hflag, 0, sizeof(int32_t) * (nt + 1), st));
  hipLaunchKernelGGL(k_tile_nsub, dim3(nblk(nt, 256)), dim3(256), 0, st, fr.bin_start, t, nsub, visits,
                     fft_x_pruned(op) ? fr.xrng : nullptr, g.S[1], g.S[2]);
  MRN_BH(hipGetLastError());
  // visit list (offsets in 64 bits: about 8 visits per sample); it is
  // filled once the work items exist
  int64_t nvisits = 0;
  {
    MRN_BH(grow(fr.vofs, cp.vofs, (int64_t)nt + 1));
    MRN_BH(grow(b.vw, b.c_vw, (int64_t)nt + 1));
    hipLaunchKernelGGL(k_widen, dim3(nblk((int64_t)nt + 1, 256)), dim3(256), 0, st, visits, nt, b.vw);
    size_t vb = 0;
    MRN_BH(hipcub::DeviceScan::ExclusiveSum(nullptr, vb, b.vw, fr.vofs, nt + 1, st));
    MRN_BH(grow(b.vtmp, b.c_vtmp, (int64_t)vb));
    MRN_BH(hipcub::DeviceScan::ExclusiveSum(b.vtmp, vb, b.vw, fr.vofs, nt + 1, st));
    int64_t nv = 0;
    MRN_BH(hipMemcpyAsync(&nv, fr.vofs + nt, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    MRN_BH(hipStreamSynchronize(st));
    MRN_BH(grow(fr.vlist, cp.vlist, nv));
    nvisits = nv;
  }
  MRN_BH(hipcub::DeviceScan::ExclusiveSum(b.tmp, tb, nsub, off, nt + 1, st));
  int32_t nsub_total = 0;
  MRN_BH(hipMemcpyAsync(&nsub_total, off + nt, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MRN_BH(hipStreamSynchronize(st));
  MRN_BH(grow(fr.subs, cp.subs, nsub_total));
  fr.nsub = nsub_total;
  hipLaunchKernelGGL(k_fill_subs, dim3(nblk(nt, 256)), dim3(256), 0, st, nsub, off, visits, t,
                     fr.subs, hflag);
  MRN_BH(hipGetLastError());
  if (nvisits > 0 && nsub_total > 0) {  // the visit list, one wavefront per work item
    hipLaunchKernelGGL(k_fill_visits, dim3(nblk(nsub_total, 4)), dim3(256), 0, st, fr.bin_start, t, fr.vofs,
                       fr.subs, nsub_total, fr.vlist);
    MRN_BH(hipGetLastError());
  }
  // the work list cut into MRN_XCDS ranges of equal cost (xcd_run < 0)
  {
    const int64_t n = nsub_total;
    MRN_BH(grow(b.icost, b.c_icost, n + 1));
    MRN_BH(grow(b.ipre, b.c_ipre, n + 1));
    MRN_BH(grow(fr.xsplit, cp.xsplit, MRN_XCDS + 1));
    hipLaunchKernelGGL(k_item_cost, dim3(nblk(n + 1, 256)), dim3(256), 0, st, fr.subs, (int)n, b.icost);
    MRN_BH(hipGetLastError());
    size_t cb = 0;
    MRN_BH(hipcub::DeviceScan::ExclusiveSum(nullptr, cb, b.icost, b.ipre, (int)n + 1, st));
    MRN_BH(grow(b.vtmp, b.c_vtmp, (int64_t)cb));
    MRN_BH(hipcub::DeviceScan::ExclusiveSum(b.vtmp, cb, b.icost, b.ipre, (int)n + 1, st));
    hipLaunchKernelGGL(k_xcd_split, dim3(1), dim3(64), 0, st, b.ipre, (int)n, fr.xsplit);
    MRN_BH(hipGetLastError());
    int32_t xs[MRN_XCDS + 1];
    MRN_BH(hipMemcpyAsync(xs, fr.xsplit, sizeof(xs), hipMemcpyDeviceToHost, st));
    MRN_BH(hipStreamSynchronize(st));
    fr.xsplit_max = 0;
    for (int x = 0; x < MRN_XCDS; x++) fr.xsplit_max = std::max(fr.xsplit_max, xs[x + 1] - xs[x]);
  }
  MRN_BH(hipcub::DeviceScan::ExclusiveSum(b.tmp, tb, hflag, hoff, nt + 1, st));
  int32_t nh = 0;
  MRN_BH(hipMemcpyAsync(&nh, hoff + nt, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MRN_BH(hipStreamSynchronize(st));
  MRN_BH(grow(fr.heavy, cp.heavy, nh));
  fr.nheavy = nh;
  if (nh > 0) {
    hipLaunchKernelGGL(k_compact, dim3(nblk(nt, 256)), dim3(256), 0, st, hflag, hoff, nt, fr.heavy);
    MRN_BH(hipGetLastError());
  }
  // gather work list (home samples per tile), reusing the scratch arrays
  MRN_BH(hipStreamSynchronize(st));
  hipLaunchKernelGGL(k_gtile_count, dim3(nblk(nt, 256)), dim3(256), 0, st, fr.bin_start, t, GT_PER, GT_MIN,
                     nsub, visits, hflag);
  MRN_BH(hipGetLastError());
  MRN_BH(hipcub::DeviceScan::ExclusiveSum(b.tmp, tb, nsub, off, nt + 1, st));
  int32_t ng = 0;
  MRN_BH(hipMemcpyAsync(&ng, off + nt, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MRN_BH(hipStreamSynchronize(st));
  MRN_BH(grow(fr.gsubs, cp.gsubs, ng));
  fr.ngsub = ng;
  if (ng > 0) {
    hipLaunchKernelGGL(k_gtile_fill, dim3(nblk(nt, 256)), dim3(256), 0, st, nsub, off, visits, t, GT_PER, fr.gsubs);
    MRN_BH(hipGetLastError());
  }
  // samples of the sparse tiles, for the direct gather
  MRN_BH(hipcub::DeviceScan::ExclusiveSum(b.tmp, tb, hflag, hoff, nt + 1, st));
  int32_t nsp = 0;
  MRN_BH(hipMemcpyAsync(&nsp, hoff + nt, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MRN_BH(hipStreamSynchronize(st));
  MRN_BH(grow(fr.sparse, cp.sparse, nsp));
  fr.nsparse = nsp;
  if (nsp > 0) {
    hipLaunchKernelGGL(k_gtile_sparse, dim3(nblk(nt, 256)), dim3(256), 0, st, fr.bin_start, t, hflag, hoff,
                       fr.sparse);
    MRN_BH(hipGetLastError());
  }
  MRN_BH(hipStreamSynchronize(st));
#undef MRN_BH
  fr.has_traj = true;
  return MRN_OK;
}

}  // namespace mrn
