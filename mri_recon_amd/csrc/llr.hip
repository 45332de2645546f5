// Locally low-rank soft threshold over frames: LOWRANKCOIL::update_threshold
// and ::thresh (clear.cpp:233-556) for an image of Nt x Ne frames, as
// recon_lib calls them after each gradient step (recon_lib.cpp:2117-2126).
//
// Block b of a pass is the Np x N matrix A (N = frames along `dim`, Np = the
// other frames x block voxels).  One wavefront (or half of one) per block:
//   1-2. G = A^H A in FP64, A's rows streamed through LDS one chunk of LPB
//      rows at a time (periodic wrap, the pass's shift), so a block holds
//      N x (LPB + 1) values of A in LDS instead of all Np rows;
//   3. G = V diag(lambda) V^H by cyclic complex Jacobi with the round-robin
//      parallel order (N/2 disjoint rotations at a time, FP64, in LDS);
//   4. update_threshold: the block's largest singular value sqrt(max lambda);
//      thresh: A <- A V diag(f) V^H, f = max(s - tau, 0) / s, s = sqrt(lambda),
//      which is U diag(max(s - tau, 0)) V^H without forming U (A's chunks
//      read again from the image and written back in place).
// The threshold is the median of the nonzero block maxima (exact radix select
// on their bit patterns, all on the device) times 1 + amp*exp(-iter/t2), times
// alpha, over block_iter.  Everything after the launch parameters lives in
// device memory, so a FISTA iteration with this step can be captured in a
// hipGraph (no host round trip, no memset).
//
// Shifts: pass `pass` of call `call` shifts block axis q by
// splitmix64(seed, 3 * (call * block_iter + pass) + q) % b - b/2 (the
// reference draws rand() % b - b/2), 0 when block_iter == 1.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "mrn_internal.hpp"

namespace mrn {

namespace {

struct LlrGeom {
  int Nv[3];  // Nx, Ny, Nz
  int bs[3];  // block sizes, clamped to the volume
  int nb[3];  // blocks per axis
  int Nt, Ne, dim;
  int N, Np, bv;  // columns, rows, voxels per block
  int block_iter;
  unsigned long long seed;
  long long nvox;
  int lds_G, lds_V, lds_A, lds_R;  // byte offsets in one block's LDS
  int lds_block;                   // bytes of LDS per block (16-byte multiple)
  int lpb;                         // lanes per block: 64, or 32 (two blocks per wavefront)
  long long nblk;
};

constexpr int LLR_MAX_N = 32;
constexpr int LLR_LDS = 64 * 1024;  // the default dynamic LDS limit of a launch
constexpr int LLR_SWEEPS = 30;

__host__ __device__ inline unsigned long long splitmix(unsigned long long seed, unsigned long long index) {
  unsigned long long z = seed + (index + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// row-major N x N complex matrices in LDS, rows padded to N + 1 entries and
// the chunk's columns to LPB + 1, so that lanes walking a column (or reading
// several columns at one row) fall on different LDS banks
#define GM(r, c) G[(r) * LDG + (c)]
#define VM(r, c) V[(r) * LDG + (c)]

// Pair k of round r of the round-robin schedule over n + 1 players (n odd):
// (n, r) and ((r + k) mod n, (r - k) mod n).  Index n is the dummy when N is odd.
__device__ inline void rr_pair(int r, int k, int n, int &p, int &q) {
  if (k == 0) {
    p = n;
    q = r;
  } else {  // r, k < n
    p = r + k;
    if (p >= n) p -= n;
    q = r - k;
    if (q < 0) q += n;
  }
}

// LPB lanes per block: with LPB = 32 a wavefront works on two blocks, each
// half with its own LDS region (small N leaves most of 64 lanes idle in the
// rotation rounds otherwise).
template <bool THRESH, int LPB>
__global__ __launch_bounds__(64) void k_llr_block(float2 *__restrict__ X, LlrGeom g, int pass,
                                                  const int *__restrict__ counter, long long call,
                                                  const double *__restrict__ tau_dev, double *__restrict__ smax_out) {
  extern __shared__ __align__(16) unsigned char lds_all[];
  constexpr int BPW = 64 / LPB;
  const int half_id = threadIdx.x / LPB;
  unsigned char *lds = lds_all + half_id * g.lds_block;
  double2 *G = reinterpret_cast<double2 *>(lds + g.lds_G);
  double2 *V = reinterpret_cast<double2 *>(lds + g.lds_V);
  float2 *A = reinterpret_cast<float2 *>(lds + g.lds_A);
  double *R = reinterpret_cast<double *>(lds + g.lds_R);  // rotations (c, s, e.x, e.y), then f
  const int lane = threadIdx.x % LPB;
  const int N = g.N, Np = g.Np, LDG = N + 1;
  const long long bq = (long long)blockIdx.x * BPW + half_id;
  const bool live = bq < g.nblk;  // a padding half repeats the last block and writes nothing
  const long long b = live ? bq : g.nblk - 1;
  const int bi = (int)(b / ((long long)g.nb[1] * g.nb[2])), bj = (int)((b / g.nb[2]) % g.nb[1]),
            bk = (int)(b % g.nb[2]);
  int sh[3] = {0, 0, 0};
  if (THRESH && g.block_iter > 1) {
    const long long c = counter ? (long long)(*counter - 1) : call;
    for (int q = 0; q < 3; q++)
      sh[q] = (int)(splitmix(g.seed, 3ull * (unsigned long long)(c * g.block_iter + pass) + q) %
                    (unsigned long long)g.bs[q]) -
              g.bs[q] / 2;
  }
  const int o0 = bi * g.bs[0] + sh[0], o1 = bj * g.bs[1] + sh[1], o2 = bk * g.bs[2] + sh[2];
  // row r of A = (other frame index o, voxel v) with x fastest: its voxel
  // offset in a frame and o; column c adds the frame (t, e) offset
  auto row_of = [&](int row, int &o) -> long long {
    o = row / g.bv;
    const int v = row - o * g.bv;
    const int kk = v / (g.bs[0] * g.bs[1]), rem = v - kk * g.bs[0] * g.bs[1];
    const int jj = rem / g.bs[0], ii = rem - jj * g.bs[0];
    int px = o0 + ii, py = o1 + jj, pz = o2 + kk;  // within (-Nv, 2 Nv)
    px += px < 0 ? g.Nv[0] : (px >= g.Nv[0] ? -g.Nv[0] : 0);
    py += py < 0 ? g.Nv[1] : (py >= g.Nv[1] ? -g.Nv[1] : 0);
    pz += pz < 0 ? g.Nv[2] : (pz >= g.Nv[2] ? -g.Nv[2] : 0);
    return ((long long)pz * g.Nv[1] + py) * g.Nv[0] + px;
  };
  auto frame_of = [&](int col, int o) -> long long {
    return g.dim == 0 ? ((long long)col * g.Ne + o) * g.nvox : ((long long)o * g.Ne + col) * g.nvox;
  };
  // chunk of A: rows r0 .. r0 + LPB - 1, C[col][lane] (zero past Np)
  const int LDC = LPB + 1;
  auto load_chunk = [&](int r0) {
    const int row = r0 + lane;
    if (row < Np) {
      int o;
      const long long vo = row_of(row, o);
      for (int col = 0; col < N; col++) A[col * LDC + lane] = X[frame_of(col, o) + vo];
    } else {
      for (int col = 0; col < N; col++) A[col * LDC + lane] = make_float2(0.f, 0.f);
    }
  };
  // 1-2. Gram matrix: lanes own the pairs a <= c of the upper triangle and add
  // each chunk's sum to them, then mirror
  const int npair = N * (N + 1) / 2;
  for (int u = lane; u < N * N; u += LPB) {
    GM(u / N, u % N) = make_double2(0.0, 0.0);
    VM(u / N, u % N) = make_double2(u / N == u % N ? 1.0 : 0.0, 0.0);
  }
  for (int r0 = 0; r0 < Np; r0 += LPB) {
    load_chunk(r0);
    __syncthreads();
    const int nr = Np - r0 < LPB ? Np - r0 : LPB;
    for (int u = lane; u < npair; u += LPB) {
      int a = 0, rem = u;
      while (rem >= N - a) {
        rem -= N - a;
        a++;
      }
      const int c = a + rem;
      const float2 *Aa = A + a * LDC, *Ac = A + c * LDC;
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < nr; r++) {
        const double xr = Aa[r].x, xi = Aa[r].y, yr = Ac[r].x, yi = Ac[r].y;
        sr += xr * yr + xi * yi;  // conj(x) * y
        si += xr * yi - xi * yr;
      }
      const double2 gv = GM(a, c);
      GM(a, c) = make_double2(gv.x + sr, gv.y + si);
    }
    __syncthreads();  // the chunk is read before the next one replaces it
  }
  for (int u = lane; u < npair; u += LPB) {
    int a = 0, rem = u;
    while (rem >= N - a) {
      rem -= N - a;
      a++;
    }
    const int c = a + rem;
    if (c != a) {
      const double2 gv = GM(a, c);
      GM(c, a) = make_double2(gv.x, -gv.y);
    }
  }
  __syncthreads();
  // 3. Jacobi sweeps
  const int Npl = N + (N & 1), n = Npl - 1, half = Npl / 2;
  for (int sweep = 0; sweep < LLR_SWEEPS; sweep++) {
    double off = 0.0, dia = 0.0;
    for (int u = lane; u < N * N; u += LPB) {
      const double2 v = GM(u / N, u % N);
      const double m = v.x * v.x + v.y * v.y;
      if (u / N == u % N) dia += m;
      else off += m;
    }
    for (int s = LPB / 2; s > 0; s >>= 1) {
      off += __shfl_xor(off, s);
      dia += __shfl_xor(dia, s);
    }
    // both halves of a wavefront run the same number of sweeps (the rounds
    // hold barriers); each half stops rotating once its own block converged
    const bool done = off <= 1e-24 * dia || dia == 0.0;
    bool all_done = done;
    if (BPW == 2) all_done = done && __shfl_xor(done ? 1 : 0, 32) != 0;
    if (all_done) break;
    for (int r = 0; r < n; r++) {
      if (lane < half) {
        int p, q;
        rr_pair(r, lane, n, p, q);
        double c = 1.0, s = 0.0, ex = 1.0, ey = 0.0;
        if (p < N && q < N) {
          const double al = GM(p, p).x, be = GM(q, q).x;
          const double2 ga = GM(p, q);
          // Gram entries are sums of squared floats: no overflow in these squares;
          // z * z overflowing to inf gives t = 0, the right limit
          const double ag = sqrt(ga.x * ga.x + ga.y * ga.y);
          if (ag > 0.0 && !done) {
            const double z = (be - al) / (2.0 * ag);
            const double t = (z >= 0.0 ? 1.0 : -1.0) / (fabs(z) + sqrt(1.0 + z * z));
            c = 1.0 / sqrt(1.0 + t * t);
            s = c * t;
            ex = ga.x / ag;
            ey = ga.y / ag;
          }
        }
        R[4 * lane + 0] = c;
        R[4 * lane + 1] = s;
        R[4 * lane + 2] = ex;
        R[4 * lane + 3] = ey;
      }
      __syncthreads();
      // columns p, q of G and V: A <- A J
      for (int it = lane; it < half * N; it += LPB) {
        const int k = it / N, rr = it % N;
        int p, q;
        rr_pair(r, k, n, p, q);
        const double s = R[4 * k + 1];
        if (p >= N || q >= N || s == 0.0) continue;
        const double c = R[4 * k];
        const double2 e = make_double2(R[4 * k + 2], R[4 * k + 3]);
        const double2 se = make_double2(s * e.x, s * e.y), sec = make_double2(s * e.x, -s * e.y);
        double2 x = GM(rr, p), y = GM(rr, q);
        double2 m1 = cmul(sec, y), m2 = cmul(se, x);
        GM(rr, p) = make_double2(c * x.x - m1.x, c * x.y - m1.y);
        GM(rr, q) = make_double2(m2.x + c * y.x, m2.y + c * y.y);
        x = VM(rr, p);
        y = VM(rr, q);
        m1 = cmul(sec, y);
        m2 = cmul(se, x);
        VM(rr, p) = make_double2(c * x.x - m1.x, c * x.y - m1.y);
        VM(rr, q) = make_double2(m2.x + c * y.x, m2.y + c * y.y);
      }
      __syncthreads();
      // rows p, q of G: A <- J^H A, and the (p, q) pair set to 0
      for (int it = lane; it < half * N; it += LPB) {
        const int k = it / N, cc = it % N;
        int p, q;
        rr_pair(r, k, n, p, q);
        const double s = R[4 * k + 1];
        if (p >= N || q >= N || s == 0.0) continue;
        const double c = R[4 * k];
        const double2 e = make_double2(R[4 * k + 2], R[4 * k + 3]);
        const double2 se = make_double2(s * e.x, s * e.y), sec = make_double2(s * e.x, -s * e.y);
        const double2 x = GM(p, cc), y = GM(q, cc);
        const double2 m1 = cmul(se, y), m2 = cmul(sec, x);
        GM(p, cc) = cc == q ? make_double2(0.0, 0.0) : make_double2(c * x.x - m1.x, c * x.y - m1.y);
        GM(q, cc) = cc == p ? make_double2(0.0, 0.0) : make_double2(m2.x + c * y.x, m2.y + c * y.y);
      }
      __syncthreads();
    }
  }
  // 4. results
  if (!THRESH) {
    double m = 0.0;
    for (int a = lane; a < N; a += LPB) m = fmax(m, sqrt(fmax(GM(a, a).x, 0.0)));
    for (int s = LPB / 2; s > 0; s >>= 1) m = fmax(m, __shfl_xor(m, s));
    if (lane == 0 && live) smax_out[b] = m;
    return;
  }
  const double tau = *tau_dev;
  double *f = R;
  for (int a = lane; a < N; a += LPB) {
    const double sv = sqrt(fmax(GM(a, a).x, 0.0));
    f[a] = sv > 0.0 ? fmax(sv - tau, 0.0) / sv : 0.0;
  }
  __syncthreads();
  // P = V diag(f) V^H into G's space
  for (int u = lane; u < N * N; u += LPB) {
    const int a = u / N, c = u % N;
    double pr = 0.0, pi = 0.0;
    for (int k = 0; k < N; k++) {
      const double2 va = VM(a, k), vc = VM(c, k);
      const double w = f[k];
      pr += w * (va.x * vc.x + va.y * vc.y);  // va * conj(vc)
      pi += w * (va.y * vc.x - va.x * vc.y);
    }
    GM(a, c) = make_double2(pr, pi);
  }
  __syncthreads();
  // A P back to the image, chunk by chunk (a chunk's rows are read before
  // they are written; blocks of a pass share no voxel)
  for (int r0 = 0; r0 < Np; r0 += LPB) {
    load_chunk(r0);
    __syncthreads();
    const int row = r0 + lane;
    if (row < Np) {
      int o;
      const long long vo = row_of(row, o);
      for (int col = 0; col < N; col++) {
        double xr = 0.0, xi = 0.0;
        for (int a = 0; a < N; a++) {
          const float2 av = A[a * LDC + lane];
          const double2 pv = GM(a, col);
          xr += (double)av.x * pv.x - (double)av.y * pv.y;
          xi += (double)av.x * pv.y + (double)av.y * pv.x;
        }
        if (live) X[frame_of(col, o) + vo] = make_float2((float)xr, (float)xi);
      }
    }
    __syncthreads();
  }
}
#undef GM
#undef VM

// ---- exact median of the nonzero block maxima: radix select over the bit
// patterns of the (non-negative) doubles, 8 bits per pass, two ranks at once.
struct SelState {
  unsigned long long prefix[2], mask[2];
  long long rank[2];
  long long count;  // nonzero values
  unsigned int hist[2][256];
};

__global__ void k_sel_init(SelState *st) {
  const int t = threadIdx.x;
  if (t < 2) {
    st->prefix[t] = 0;
    st->mask[t] = 0;
    st->rank[t] = 0;
  }
  if (t == 0) st->count = 0;
  for (int i = t; i < 512; i += blockDim.x) (&st->hist[0][0])[i] = 0;
}

__global__ void k_sel_count(const double *__restrict__ v, long long n, SelState *st) {
  __shared__ unsigned int c;
  if (threadIdx.x == 0) c = 0;
  __syncthreads();
  unsigned int mine = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    mine += v[i] != 0.0;
  atomicAdd(&c, mine);
  __syncthreads();
  if (threadIdx.x == 0 && c) atomicAdd((unsigned long long *)&st->count, (unsigned long long)c);
}

__global__ void k_sel_ranks(SelState *st) {
  const long long K = st->count;
  st->rank[0] = K > 0 ? (K - 1) / 2 : 0;  // lower middle
  st->rank[1] = K / 2;                     // upper middle (the same one for odd K)
}

__global__ void k_sel_hist(const double *__restrict__ v, long long n, SelState *st, int shift) {
  __shared__ unsigned int h[2][256];
  for (int i = threadIdx.x; i < 512; i += blockDim.x) (&h[0][0])[i] = 0;
  __syncthreads();
  const unsigned long long p0 = st->prefix[0], p1 = st->prefix[1], m0 = st->mask[0], m1 = st->mask[1];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    if (v[i] == 0.0) continue;  // also -0.0, as the count
    const unsigned long long key = (unsigned long long)__double_as_longlong(v[i]);
    const int d = (int)((key >> shift) & 255ull);
    if ((key & m0) == p0) atomicAdd(&h[0][d], 1u);
    if ((key & m1) == p1) atomicAdd(&h[1][d], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 512; i += blockDim.x) {
    const unsigned int c = (&h[0][0])[i];
    if (c) atomicAdd(&(&st->hist[0][0])[i], c);
  }
}

__global__ void k_sel_step(SelState *st, int shift) {
  if (threadIdx.x < 2) {
    const int s = threadIdx.x;
    long long r = st->rank[s], acc = 0;
    int d = 0;
    for (; d < 255; d++) {
      const long long c = st->hist[s][d];
      if (r < acc + c) break;
      acc += c;
    }
    st->rank[s] = r - acc;
    st->prefix[s] |= (unsigned long long)d << shift;
    st->mask[s] |= 255ull << shift;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 512; i += blockDim.x) (&st->hist[0][0])[i] = 0;
}

// smax = (float)(median * (float)(1 + amp*exp(-iter/t2))) (clear.cpp:343-348);
// tau = alpha * smax / block_iter (clear.cpp:386-387, 458).
__global__ void k_llr_finish(const SelState *st, float amp, float t2, const int *counter, int iter_host, float alpha,
                             int block_iter, double *tau, float *smax_out) {
  const long long K = st->count;
  double med = 0.0;
  if (K > 0) {
    const double lo = __longlong_as_double((long long)st->prefix[0]);
    const double hi = __longlong_as_double((long long)st->prefix[1]);
    med = (K & 1) ? hi : lo + (hi - lo) / 2.0;
  }
  const int it = counter ? *counter - 1 : iter_host;
  const float amp_scale = 1.0f + amp * expf(-((float)it) / t2);
  const float smax = (float)(med * (double)amp_scale);
  *smax_out = smax;
  *tau = (double)alpha * (double)smax / (double)block_iter;
}

}  // namespace

// Geometry, limits and LDS layout; MRN_OK or an error status with `err`.
static int llr_geom(const mrn_llr_params *p, int Nt, int Ne, const int32_t N[3], int dim, LlrGeom &g,
                    const char *&err) {
  err = nullptr;
  if (Nt < 1 || Ne < 1 || dim < 0 || dim > 1) return (err = "frames or dim", MRN_EINVAL);
  for (int a = 0; a < 3; a++) {
    if (N[a] < 1 || p->block[a] < 1) return (err = "image size or block size", MRN_EINVAL);
    g.Nv[a] = N[a];
    g.bs[a] = p->block[a] > N[a] ? N[a] : p->block[a];
    g.nb[a] = N[a] / g.bs[a];
  }
  if (p->block_iter < 1) return (err = "clear_block_iter < 1", MRN_EINVAL);
  g.Nt = Nt;
  g.Ne = Ne;
  g.dim = dim;
  g.bv = g.bs[0] * g.bs[1] * g.bs[2];
  g.N = dim == 0 ? Nt : Ne;
  g.Np = (dim == 0 ? Ne : Nt) * g.bv;
  g.block_iter = p->block_iter;
  g.seed = p->seed;
  g.nvox = (long long)N[0] * N[1] * N[2];
  if (g.N > LLR_MAX_N) return (err = "more than 32 frames along the low-rank dimension", MRN_EUNSUPPORTED);
  const int Npl = g.N + (g.N & 1);
  g.lds_G = 0;
  g.lds_V = g.lds_G + 16 * g.N * (g.N + 1);
  g.lds_R = g.lds_V + 16 * g.N * (g.N + 1);
  g.lds_A = g.lds_R + 8 * (4 * (Npl / 2) > g.N ? 4 * (Npl / 2) : g.N);
  g.lds_A = (g.lds_A + 15) & ~15;
  // one block's LDS with a chunk of lpb rows of A
  auto block_bytes = [&](int lpb) { return (int)((g.lds_A + 8ll * (lpb + 1) * g.N + 15) & ~15ll); };
  g.lpb = (g.N <= 16 && 2ll * block_bytes(32) <= LLR_LDS) ? 32 : 64;
  g.lds_block = block_bytes(g.lpb);
  if ((long long)g.lds_block * (64 / g.lpb) > LLR_LDS)
    return (err = "low-rank block matrix too large for LDS", MRN_EUNSUPPORTED);
  g.nblk = (long long)g.nb[0] * g.nb[1] * g.nb[2];
  return MRN_OK;
}

static size_t llr_lds(const LlrGeom &g) { return (size_t)g.lds_block * (64 / g.lpb); }
static long long llr_blocks(const LlrGeom &g) { return (long long)g.nb[0] * g.nb[1] * g.nb[2]; }

// update_threshold + thresh on stream st.  Scratch: sblocks [nblocks] doubles,
// sel (SelState), tau (1 double), smax (1 float), all device.  counter: FISTA's
// device iteration counter (iteration = *counter - 1, also the call index of
// the shifts), or null to use iter / call.
static hipError_t llr_launch(const LlrGeom &g, float2 *X, float alpha, float amp, float t2, int iter, long long call,
                             const int *counter, double *sblocks, void *sel_v, double *tau, float *smax,
                             hipStream_t st) {
  SelState *sel = (SelState *)sel_v;
  const long long nblk = llr_blocks(g);
  const size_t lds = llr_lds(g);
  const unsigned nwg = (unsigned)((nblk + (64 / g.lpb) - 1) / (64 / g.lpb));
  auto blocks = [&](bool thresh, int pass) {
    if (g.lpb == 32) {
      if (thresh)
        hipLaunchKernelGGL((k_llr_block<true, 32>), dim3(nwg), dim3(64), lds, st, X, g, pass, counter, call, tau,
                           (double *)nullptr);
      else
        hipLaunchKernelGGL((k_llr_block<false, 32>), dim3(nwg), dim3(64), lds, st, X, g, 0, counter, call,
                           (const double *)nullptr, sblocks);
    } else {
      if (thresh)
        hipLaunchKernelGGL((k_llr_block<true, 64>), dim3(nwg), dim3(64), lds, st, X, g, pass, counter, call, tau,
                           (double *)nullptr);
      else
        hipLaunchKernelGGL((k_llr_block<false, 64>), dim3(nwg), dim3(64), lds, st, X, g, 0, counter, call,
                           (const double *)nullptr, sblocks);
    }
  };
  if (nblk > 0) blocks(false, 0);
  hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(256), 0, st, sel);
  const unsigned gb = (unsigned)((nblk + 255) / 256 < 1024 ? (nblk + 255) / 256 : 1024);
  if (nblk > 0) {
    hipLaunchKernelGGL(k_sel_count, dim3(gb), dim3(256), 0, st, sblocks, nblk, sel);
    hipLaunchKernelGGL(k_sel_ranks, dim3(1), dim3(1), 0, st, sel);
    for (int shift = 56; shift >= 0; shift -= 8) {
      hipLaunchKernelGGL(k_sel_hist, dim3(gb), dim3(256), 0, st, sblocks, nblk, sel, shift);
      hipLaunchKernelGGL(k_sel_step, dim3(1), dim3(256), 0, st, sel, shift);
    }
  }
  hipLaunchKernelGGL(k_llr_finish, dim3(1), dim3(1), 0, st, sel, amp, t2, counter, iter, alpha, g.block_iter, tau,
                     smax);
  if (alpha != 0.0f && nblk > 0)
    for (int pass = 0; pass < g.block_iter; pass++) blocks(true, pass);
  return hipGetLastError();
}

int llr_fista_prepare(mrn_op *op) {
  if (!op->llr_on) return MRN_OK;
  LlrGeom g;
  const char *err;
  const int32_t N[3] = {op->g.N[0], op->g.N[1], op->g.N[2]};
  const int rc = llr_geom(&op->llr, op->F, 1, N, 0, g, err);
  if (rc != MRN_OK) return fail(op, rc, err);
  const long long nblk = llr_blocks(g);
  if (op->llr_nblk < nblk) {
    hipFree(op->llr_sblocks);
    op->llr_sblocks = nullptr;
    MRN_HIP(op, hipMalloc(&op->llr_sblocks, sizeof(double) * (nblk > 0 ? nblk : 1)));
    op->llr_nblk = nblk;
  }
  if (!op->llr_sel) MRN_HIP(op, hipMalloc(&op->llr_sel, sizeof(SelState)));
  if (!op->llr_scal) MRN_HIP(op, hipMalloc(&op->llr_scal, sizeof(double) * 2));
  return MRN_OK;
}

// FISTA's step: time LLR on the op's frames, iteration from the device counter.
// llr_scal[0] holds tau and llr_scal[1] the latest smax (as a float).
int llr_fista_step(mrn_op *op, float2 *X, hipStream_t st) {
  LlrGeom g;
  const char *err;
  const int32_t N[3] = {op->g.N[0], op->g.N[1], op->g.N[2]};
  const int rc = llr_geom(&op->llr, op->F, 1, N, 0, g, err);
  if (rc != MRN_OK) return fail(op, rc, err);
  MRN_HIP(op, llr_launch(g, X, op->llr.alpha_time, op->llr.amp, op->llr.t2, 0, 0, op->dcounter, op->llr_sblocks,
                         op->llr_sel, op->llr_scal, reinterpret_cast<float *>(op->llr_scal + 1), st));
  return MRN_OK;
}

}  // namespace mrn

// ---- fixed scratch of mrn_llr_noalloc ------------------------------------
// The block maxima (8 bytes per block), the median selection state and the
// threshold scalars of one call live in this array instead of hipMalloc'd
// memory, for callers that cannot allocate.  Each device holds its own copy
// until the process ends; every call on one device shares it, so calls that
// overlap on one device race.  Every part is written by the call before it
// is read (the selection state is cleared by k_sel_init).
#define MRN_LLR_SCRATCH_BYTES (8u << 20)
static __device__ __align__(256) unsigned char scratch[MRN_LLR_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
size_t scratch_bytes(long long nblk) {
  return align256(sizeof(double) * (size_t)(nblk > 0 ? nblk : 1)) + align256(sizeof(mrn::SelState)) +
         align256(sizeof(double) + sizeof(float));
}
long long scratch_max_blocks() {
  return (long long)((MRN_LLR_SCRATCH_BYTES - align256(sizeof(mrn::SelState)) - 256) / sizeof(double));
}

// The device address of `scratch` on the current device, looked up once per
// device with that device current (host code cannot use the symbol's name
// as a device pointer).
hipError_t scratch_address(unsigned char **out, const char *&err) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) {
    err = "mrn_llr_noalloc: device index beyond its per-device table";
    return hipErrorInvalidDevice;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = (unsigned char *)p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}
}  // namespace

using namespace mrn;

extern "C" {

void mrn_llr_params_default(mrn_llr_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->block[0] = p->block[1] = p->block[2] = 4;
  p->block_iter = 4;
  p->amp = 4.0f;
  p->t2 = 6.0f;
}

int mrn_llr_parse_args(mrn_llr_params *p, int argc, const char **argv) {
  if (!p || (argc > 0 && !argv)) return MRN_EINVAL;
  for (int i = 0; i < argc; i++) {
    const char *s = argv[i];
    const bool has = i + 1 < argc;
    int32_t *iv = nullptr;
    float *fv = nullptr;
    if (!strcmp(s, "-clear_nx")) iv = &p->block[0];
    else if (!strcmp(s, "-clear_ny")) iv = &p->block[1];
    else if (!strcmp(s, "-clear_nz")) iv = &p->block[2];
    else if (!strcmp(s, "-clear_block_iter")) iv = &p->block_iter;
    else if (!strcmp(s, "-clear_alpha_time")) fv = &p->alpha_time;
    else if (!strcmp(s, "-clear_alpha_encode")) fv = &p->alpha_encode;
    else if (!strcmp(s, "-clear_amp")) fv = &p->amp;
    else if (!strcmp(s, "-clear_t2")) fv = &p->t2;
    if ((iv || fv) && !has) return MRN_EINVAL;
    if (iv) *iv = atoi(argv[++i]);
    if (fv) *fv = (float)atof(argv[++i]);
  }
  return MRN_OK;
}

int mrn_llr(const mrn_llr_params *p, void *X, int Nt, int Ne, const int32_t N[3], int dim, int iter,
            int64_t call, float *smax, void *stream) {
  free_err_clear();
  if (!p || !X || !N) return free_fail(MRN_EINVAL, "null argument");
  LlrGeom g;
  const char *err;
  const int rc = llr_geom(p, Nt, Ne, N, dim, g, err);
  if (rc != MRN_OK) return free_fail(rc, err);
  hipStream_t st = (hipStream_t)stream;
  const long long nblk = llr_blocks(g);
  double *sb = nullptr, *tau = nullptr;
  void *sel = nullptr;
  float *sm = nullptr;
  int out = MRN_OK;
  hipError_t e = hipMalloc(&sb, sizeof(double) * (nblk > 0 ? nblk : 1));
  if (e == hipSuccess) e = hipMalloc(&sel, sizeof(SelState));
  if (e == hipSuccess) e = hipMalloc(&tau, sizeof(double) + sizeof(float));
  if (e == hipSuccess) {
    sm = reinterpret_cast<float *>(tau + 1);
    const float alpha = dim == 0 ? p->alpha_time : p->alpha_encode;
    e = llr_launch(g, (float2 *)X, alpha, p->amp, p->t2, iter, (long long)call, nullptr, sb, sel, tau, sm, st);
  }
  float h = 0.0f;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, sm, sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    hipStreamSynchronize(st);
    out = free_fail(MRN_EHIP, hipGetErrorString(e));
  } else if (smax) {
    *smax = h;
  }
  hipFree(sb);
  hipFree(sel);
  hipFree(tau);
  return out;
}

int mrn_llr_noalloc(const mrn_llr_params *p, void *X, int Nt, int Ne, const int32_t N[3], int dim, int iter,
                    int64_t call, float *smax, void *stream) {
  free_err_clear();
  if (!p || !X || !N) return free_fail(MRN_EINVAL, "null argument");
  LlrGeom g;
  const char *err;
  const int rc = llr_geom(p, Nt, Ne, N, dim, g, err);
  if (rc != MRN_OK) return free_fail(rc, err);
  const long long nblk = llr_blocks(g);
  const size_t need = scratch_bytes(nblk);
  if (need > sizeof(scratch)) {
    char msg[200];
    snprintf(msg, sizeof msg,
             "mrn_llr_noalloc: %lld blocks need %zu bytes of scratch; its fixed scratch holds %zu bytes (%lld blocks)",
             nblk, need, sizeof(scratch), scratch_max_blocks());
    return free_fail(MRN_EINVAL, msg);
  }
  unsigned char *base = nullptr;
  const char *serr = nullptr;
  const hipError_t se = scratch_address(&base, serr);
  if (se != hipSuccess) return free_fail(MRN_EHIP, serr ? serr : hipGetErrorString(se));
  // layout: block maxima | selection state | tau, smax (256-byte aligned parts)
  double *sb = reinterpret_cast<double *>(base);
  void *sel = base + align256(sizeof(double) * (size_t)nblk);
  double *tau = reinterpret_cast<double *>(base + align256(sizeof(double) * (size_t)nblk) + align256(sizeof(SelState)));
  float *sm = reinterpret_cast<float *>(tau + 1);
  hipStream_t st = (hipStream_t)stream;
  const float alpha = dim == 0 ? p->alpha_time : p->alpha_encode;
  hipError_t e = llr_launch(g, (float2 *)X, alpha, p->amp, p->t2, iter, (long long)call, nullptr, sb, sel, tau, sm, st);
  float h = 0.0f;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, sm, sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(st);
    return free_fail(MRN_EHIP, hipGetErrorString(e));
  }
  if (smax) *smax = h;
  return MRN_OK;
}

int mrn_set_fista_llr(mrn_op *op, const mrn_llr_params *p) {
  if (!op) return MRN_EINVAL;
  if (!p || p->alpha_time == 0.0f) {
    op->llr_on = false;
    op->version++;
    return MRN_OK;
  }
  LlrGeom g;
  const char *err;
  const int32_t N[3] = {op->g.N[0], op->g.N[1], op->g.N[2]};
  const int rc = llr_geom(p, op->F, 1, N, 0, g, err);
  if (rc != MRN_OK) return fail(op, rc, err);
  op->llr = *p;
  op->llr_on = true;
  op->version++;
  return MRN_OK;
}

int mrn_get_fista_llr_smax(mrn_op *op, float *smax, void *stream) {
  if (!op || !smax) return MRN_EINVAL;
  if (!op->llr_scal) return fail(op, MRN_EINVAL, "no low-rank FISTA run yet");
  hipStream_t st = pick(op, stream);
  MRN_HIP(op, hipMemcpyAsync(smax, op->llr_scal + 1, sizeof(float), hipMemcpyDeviceToHost, st));
  MRN_HIP(op, hipStreamSynchronize(st));
  return MRN_OK;
}

}  // extern "C"
