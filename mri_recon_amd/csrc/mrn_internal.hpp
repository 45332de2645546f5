// Internal declarations shared by the HIP sources of libmrn.
//
// Data layout in HBM (all complex64 = float2):
//   image X            [Nz][Ny][Nx]                     (x fastest, like the reference)
//   coil maps (op)     [nch][Nz][Ny][Nx][CC]            coil-interleaved blocks of CC coils
//   grids (op)         [nch][Sz][Sy][Sx][CC]            one oversampled grid per coil block
//   k-space (op)       [nch][Mv][CC]                    valid samples in tile order
// CC coils share one grid cell (CC*8 bytes), so a gridding tap reads or adds
// one contiguous CC*8-byte run and the kernel weights are computed once per
// sample for all CC coils of the block.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mrn.h"

#define MRN_MAXT 8           // max taps per axis (dwin <= 3.5)
#define FFT_LINES 16         // 1-D lines per FFT workgroup
#define FFT_THREADS 256
#define FFT_PAD 1            // LDS line padding (float2) against bank conflicts
#define MRN_TRACE_CAP 65536  // solver trace ring (iterations)
#define MRN_MAX_SPINS 64     // wavelet cycle spins per L1 step

namespace mrn {

// ------------------------------------------------------------ geometry
struct Geom {
  int N[3], S[3], og[3];
  int CC, lgCC;
  int grid_in[3], fft_in[3];
  float grid[3];   // oversampling ratio per axis
  float dwin[3];
  int tlen[3];     // kernel table length (L + 10)
  int toff[3];     // table offset in the packed table buffer
  int ntab;        // packed table length (floats)
  int kmode;       // MRN_KMODE_*
  float beta[3];   // KB beta (exact mode)
  float knorm[3];  // table normalisation (exact mode)
  float i0beta[3]; // I0(beta) (exact mode)
  float k_rad;
};

// Sample bins and scatter tiles.  Samples are sorted by the B^3-cell bin of
// their first tap; scatter tile t covers cells [t*T, t*T+T) per axis (T = 4:
// one wavefront of cells) and reads the bins whose first taps lie in
// [t*T - NT + 1, t*T + T - 1].
struct Tiles {
  int B;        // bin edge (cells)
  int nb[3];    // bins per axis
  int nbins;
  int T;        // tile edge (cells)
  int nt[3];    // tiles per axis
  int ntiles;
  int NT;       // max taps per axis
  int NTb;      // taps per axis that set a tile's bin reach (samples with more: FrameData::wide)
  int maxv;     // samples per scatter work item
  int xrun;     // scatter work items per XCD run (1 = hardware order)
  int xrun_gather;  // gather workgroups per XCD run
};

// One frame's device arrays for a launch over every frame of the op at once
// (blockIdx.z = frame): the solver loops of a multi-frame op run one launch
// per stage, with the frames' grids side by side in op->grid.
struct FrameDev {
  const float4 *coords;
  const float *taper;
  const int4 *subs;
  const int64_t *vofs;
  const int32_t *vlist;
  const int32_t *heavy;
  const int32_t *wide;
  const int32_t *xsplit;   // [MRN_XCDS + 1] per-XCD ranges of the scatter work list
  float2 *kwork;           // this frame's [nch][Mv][CC] k-space work array
  const float2 *kdata;     // measured k-space, same layout
  int64_t Mv;
  int64_t grid_off;        // first element of this frame's coil blocks in the grid
  int nsub, nheavy, nwide, pad;
};

// Device arrays that only grow: binding a trajectory no larger than an
// earlier one of the same op reuses their memory, with no hipFree/hipMalloc
// (each of which waits for the whole device).  grow() keeps the contents only
// when the capacity suffices.
template <class T>
inline hipError_t grow(T *&p, size_t &cap, int64_t n) {
  const size_t want = n > 0 ? (size_t)n : 1;
  if (p && want <= cap) return hipSuccess;
  hipFree(p);
  p = nullptr;
  cap = 0;
  const hipError_t e = hipMalloc((void **)&p, sizeof(T) * want);
  if (e == hipSuccess) cap = want;
  return e;
}

// Capacities (elements) of a frame's grow-only arrays.
struct FrameCaps {
  size_t coords = 0, taper = 0, perm = 0, inv = 0, bin_start = 0, subs = 0, heavy = 0, vofs = 0, vlist = 0,
         gsubs = 0, wide = 0, sparse = 0, xrng = 0, yrng = 0, kdata = 0, kraw = 0, traw = 0, xsplit = 0;
};

// Scratch of a trajectory bind, kept by the op between binds (grow-only).
struct BindScratch {
  uint32_t *keys = nullptr, *keys2 = nullptr;
  int32_t *vals = nullptr, *counts = nullptr, *nsub = nullptr, *visits = nullptr, *off = nullptr,
          *hflag = nullptr, *hoff = nullptr, *wflag = nullptr, *woff = nullptr;
  int2 *brow = nullptr;
  int64_t *vw = nullptr, *icost = nullptr, *ipre = nullptr;
  uint8_t *tmp = nullptr, *wtmp = nullptr, *vtmp = nullptr;
  size_t c_keys = 0, c_keys2 = 0, c_vals = 0, c_counts = 0, c_nsub = 0, c_visits = 0, c_off = 0, c_hflag = 0,
         c_hoff = 0, c_wflag = 0, c_woff = 0, c_brow = 0, c_vw = 0, c_icost = 0, c_ipre = 0, c_tmp = 0, c_wtmp = 0,
         c_vtmp = 0;
  void release() {
    void *ps[] = {keys, keys2, vals, counts, nsub, visits, off, hflag, hoff, wflag, woff, brow, vw, icost, ipre,
                  tmp, wtmp, vtmp};
    for (void *q : ps) hipFree(q);
    *this = BindScratch{};
  }
};

struct FrameData {
  FrameCaps cap;
  int64_t M = 0;               // samples (caller order)
  int64_t Mv = 0;              // valid samples
  float4 *coords = nullptr;    // [Mv] dkx, dky, dkz, kw in tile order
  float *taper = nullptr;      // [Mv] exp(-|k|^2 / (2 k_rad^2))
  int32_t *perm = nullptr;     // [Mv] tile order -> caller index
  int32_t *inv = nullptr;      // [M] caller index -> tile order, -1 if skipped
  int32_t *bin_start = nullptr;// [nbins + 1] first sample of each bin
  int4 *subs = nullptr;        // [nsub] {tile, first visit, end visit, heavy}
  int32_t *xsplit = nullptr;   // [MRN_XCDS + 1] work-list ranges of equal cost, one per XCD
  int xsplit_max = 0;          // longest of those ranges (work items)
  int32_t *heavy = nullptr;    // [nheavy] tiles split over several work items
  int nsub = 0, nheavy = 0;
  int64_t *vofs = nullptr;     // [ntiles + 1] first entry of each tile's visits in vlist
  int32_t *vlist = nullptr;    // [vofs[ntiles]] samples of each tile's bin rows, in visit order
  int4 *gsubs = nullptr;       // [ngsub] gather work items {tile, first, end, 0} (home samples)
  int ngsub = 0;
  int32_t *wide = nullptr;     // [nwide] samples with more than Tiles::NTb taps on some axis
  int nwide = 0;
  int32_t *sparse = nullptr;   // [nsparse] samples of tiles below the dense limit (direct gather)
  int64_t nsparse = 0;
  int2 *xrng = nullptr;        // [Sz][Sy] x cells any tap reaches in grid row (z, y); x > y: none
                               // ({INT_MAX, INT_MIN}: row outside yrng[z], when yrng is set)
  int2 *yrng = nullptr;        // [Sz] rows y of plane z with support (y passes pruned), else null
  float2 *kdata = nullptr;     // measured k-space, internal layout [nch][Mv][CC]
  float2 *kraw = nullptr;      // measured k-space, caller layout [C][kraw_M] (plain form)
  int64_t kraw_M = -1;         // sample count kraw was set for
  float *traw = nullptr;       // trajectory kx|ky|kz|kw, caller order [4][M]
  bool has_traj = false, has_kdata = false;
};

}  // namespace mrn

struct mrn_op {
  mrn_params p;
  mrn::Geom g;
  mrn::Tiles t;
  int C = 0, F = 1, nch = 1;
  std::vector<float> table_h[3], win_h[3];
  float *tables_d = nullptr;   // packed tables
  float *win_d[3] = {nullptr, nullptr, nullptr};
  float2 *tw_d[3] = {nullptr, nullptr, nullptr};  // forward twiddles exp(-2 pi i q / S)
  float2 *smaps = nullptr;     // batched: [nch][Ny][Nx][Nz][CC]; plain: [C][Nz][Ny][Nx]
  float2 *grid = nullptr;      // [nch][Sz][Sy][Sx][CC]
  float2 *cgrid = nullptr;     // plain form: one coil grid [Sz][Sy][Sx]
  float2 *kwork = nullptr;     // [nch][Mv_max][CC]
  int64_t kwork_cap = 0;       // samples
  // frames fused into one launch per stage (batched form, F > 1): op->grid
  // holds F frames of nch coil blocks; fdev / kwork_all / supports are set up
  // by fused_prepare for op->version fdev_version
  bool fuse = false;
  mrn::FrameDev *fdev = nullptr;  // [F] on the device
  int fdev_version = -1;
  float2 *kwork_all = nullptr; // frames' k-space work arrays, frame after frame
  int64_t kwork_all_cap = 0;   // elements
  int2 *xrng_all = nullptr;    // [F][Sz][Sy] x support rows (k_row_xrange) of every frame
  int2 *yrng_all = nullptr;    // [F][Sz] y support of every frame, when the y passes are pruned
  int fmax_nsub = 0, fmax_nheavy = 0, fmax_nwide = 0, fmax_xsplit = 0;
  int64_t fmax_Mv = 0;
  float2 *kplain = nullptr;    // plain form: one coil of k-space [M_max]
  int64_t kplain_cap = 0;
  mrn::BindScratch bind;       // trajectory-bind scratch (grow-only)
  // The grid holds the forward grid of image fwd_X for frame fwd_f (from
  // mrn_image_to_kspace on fwd_stream) while op->calls == fwd_calls, i.e. no
  // other call has run on the op since; never once the grid's address was
  // handed out (mrn_get_grid), since the caller may write it.
  uint64_t calls = 0, fwd_calls = 0;
  int fwd_f = -1;
  const void *fwd_X = nullptr;
  void *fwd_stream = nullptr;
  bool grid_exposed = false;
  std::vector<mrn::FrameData> frames;
  hipStream_t stream = nullptr;
  // box matrix-core gather (gather_tiles = 2): the dense tiles run on `side`
  // beside the direct gather of the sparse ones (fork / join events)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // solver workspace
  float2 *vR = nullptr, *vP = nullptr, *vL = nullptr, *vX0 = nullptr;
  double *dpart = nullptr;     // reduction partials
  double *dscal = nullptr;     // device scalars (see solvers)
  double *dtrace = nullptr;    // per-iteration trace
  int dtrace_cap = 0;
  int *dcounter = nullptr;
  // cached hipGraph of one CG step (valid while version == cg_version)
  hipGraph_t cg_graph = nullptr;
  hipGraphExec_t cg_exec = nullptr;
  void *cg_X = nullptr;
  int version = 0, cg_version = -1;
  bool cg_ready = false;
  int64_t cg_count = 0;
  int64_t nvox = 0;
  std::string err;
  void *scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  // L1 wavelet proximal step of FISTA (fraction 0 = none)
  int gather_tiles = 0;        // dense tiles: 1 = LDS-staged gather, 2 = box matrix-core gather
  bool scatter_valu = false;   // VALU scatter accumulate (else matrix cores for 16-coil blocks)
  bool fft_stockham = false;   // LDS Stockham FFT passes (else four-step where the size allows)
  int xcd_run_fft = 8;         // y and z FFT pass workgroups per XCD run
  int gather_coils = 0;        // coils per lane in the direct gather (0 = auto)
  int step_type = MRN_STEP_MAXEIG;  // FISTA step: fixed (power method) or Cauchy
  int cauchy_updates = 30;     // iterations that recompute the Cauchy step
  double *dstep_log = nullptr; // Cauchy step per iteration (re, im)
  int tune_from_cache = 0;     // the last mrn_autotune applied a cached choice
  float tune_ms = 0.f;         // E'E of the tuned frame with the choice (ms)
  float l1_fraction = 0.0f;
  int l1_levels[3] = {4, 4, 4};
  int l1_spins = 4;
  uint64_t l1_seed = 0;
  void *dthr = nullptr;        // threshold search state
  int *dsh = nullptr;          // cycle-spin shifts [MRN_MAX_SPINS][3]
  double *dthr_log = nullptr;  // threshold per FISTA iteration
  // locally low-rank step of FISTA over the frames (off unless alpha_time > 0)
  bool llr_on = false;
  mrn_llr_params llr{};
  double *llr_sblocks = nullptr;  // largest singular value of each block
  long long llr_nblk = 0;
  void *llr_sel = nullptr;        // median selection state
  double *llr_scal = nullptr;     // tau, smax
};

namespace mrn {

int fail(mrn_op *op, int code, const char *msg);
#define MRN_HIP(op, call)                                                          \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      return mrn::fail(op, MRN_EHIP, (std::string(#call) + ": " + hipGetErrorString(e_)).c_str()); \
    }                                                                              \
  } while (0)
#define MRN_TRY(expr)          \
  do {                         \
    int s_ = (expr);           \
    if (s_ != MRN_OK) return s_; \
  } while (0)

// The stream of a call (NULL = null stream).  Every entry point that launches
// work picks its stream here, so op->calls counts those calls.
inline hipStream_t pick(mrn_op *op, void *s) {
  op->calls++;
  return (hipStream_t)s;
}

// Owning wrappers, released on every path out of a scope (errors included).
struct Events {
  std::vector<hipEvent_t> e;
  Events() = default;
  Events(const Events &) = delete;
  Events &operator=(const Events &) = delete;
  hipError_t create(int n, unsigned flags = hipEventDefault) {
    for (int i = 0; i < n; i++) {
      hipEvent_t x = nullptr;
      const hipError_t r = hipEventCreateWithFlags(&x, flags);
      if (r != hipSuccess) return r;
      e.push_back(x);
    }
    return hipSuccess;
  }
  hipEvent_t operator[](size_t i) const { return e[i]; }
  ~Events() {
    for (hipEvent_t x : e) (void)hipEventDestroy(x);
  }
};

// A captured graph and its executable; a launched executable is destroyed
// only after the stream it runs on has finished it.
struct GraphGuard {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  hipStream_t launched = nullptr;
  bool was_launched = false;
  GraphGuard() = default;
  GraphGuard(const GraphGuard &) = delete;
  GraphGuard &operator=(const GraphGuard &) = delete;
  ~GraphGuard() {
    if (was_launched) (void)hipStreamSynchronize(launched);
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
  }
};

// Runs f on scope exit unless released.
template <class F>
struct ScopeExit {
  F f;
  bool on = true;
  explicit ScopeExit(F fn) : f(fn) {}
  ScopeExit(const ScopeExit &) = delete;
  ~ScopeExit() {
    if (on) f();
  }
  void release() { on = false; }
};

// ------------------------------------------------------ device helpers
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// First/last tap on one axis (chop_grid_backward, gridFFT.cpp:1211-1271).
// Returns false when the sample lies outside the grid on this axis.
__device__ __forceinline__ bool axis_range(const Geom &g, int a, float dk, int &s, int &e) {
  if (g.grid_in[a]) {
    s = max((int)ceilf(dk - g.dwin[a]), 0);
    e = min((int)floorf(dk + g.dwin[a]), g.S[a] - 1);
  } else {
    s = (int)dk;
    e = s;
  }
  return !(s >= g.S[a] || e < 0);
}

// Modified Bessel I0 (Abramowitz & Stegun 9.8.1 / 9.8.2), float result.
__device__ __forceinline__ float dev_i0(float x) {
  float ax = fabsf(x);
  if (ax < 3.75f) {
    float y = x / 3.75f;
    y *= y;
    return 1.0f + y * (3.5156229f + y * (3.0899424f + y * (1.2067492f + y * (0.2659732f +
                  y * (0.360768e-1f + y * 0.45813e-2f)))));
  }
  float y = 3.75f / ax;
  return (expf(ax) * rsqrtf(ax)) *
         (0.39894228f + y * (0.1328592e-1f + y * (0.225319e-2f + y * (-0.157565e-2f +
          y * (0.916281e-2f + y * (-0.2057706e-1f + y * (0.2635537e-1f + y * (-0.1647633e-1f +
          y * 0.392377e-2f))))))));
}

// Kernel weight of tap l (with the fftshift sign), table or exact-KB mode.
// tab points at this axis' table (in LDS).  Mirrors the tap loops of
// gridFFT.cpp:1276-1306: the z weight is 1 on a Cartesian z axis.
// KM: -1 = the mode in g.kmode; 0 = table mode; 1 = exact KB (a kernel
// specialised for one mode passes it, and the other branch compiles away).
template <int KM = -1>
__device__ __forceinline__ float tap_weight(const Geom &g, int a, const float *tab, float dk,
                                            int l) {
  float w;
  const bool exact = KM < 0 ? g.kmode == MRN_KMODE_EXACT_KB : KM == 1;
  if (exact && g.grid_in[a]) {
    float x = fabsf(dk - (float)l) / g.dwin[a];
    w = (x <= 1.0f) ? g.knorm[a] * dev_i0(g.beta[a] * sqrtf(fmaxf(1.0f - x * x, 0.0f))) / g.i0beta[a]
                    : 0.0f;
  } else {
    float del = fabsf(600.0f * (dk - (float)l));
    int i0 = (int)del;
    float fr = del - (float)i0;
    float f0 = (i0 < g.tlen[a]) ? tab[i0] : 0.0f;
    float f1 = (i0 + 1 < g.tlen[a]) ? tab[i0 + 1] : 0.0f;
    w = f0 * (1.0f - fr) + f1 * fr;
  }
  if (a == 2 && !g.grid_in[2]) w = 1.0f;
  if (g.fft_in[a] && !(l & 1)) w = -w;
  return w;
}

// Table-mode weight of tap l for any l, with no guarded loads: the table
// index is clamped to the table's last pair.  Taps inside the window never
// reach the clamp (the table holds 600 dwin + 10 entries, zeros past
// 600 dwin) and get tap_weight's value; past the table both entries are
// zeros, so the weight is 0 there as in tap_weight.
__device__ __forceinline__ float tap_weight_clamped(const Geom &g, int a, const float *tab, float dk, int l) {
  const float del = fabsf(600.0f * (dk - (float)l));
  const int i0 = min((int)del, g.tlen[a] - 2);
  const float fr = del - (float)i0;
  float w = tab[i0] * (1.0f - fr) + tab[i0 + 1] * fr;
  if (a == 2 && !g.grid_in[2]) w = 1.0f;
  if (g.fft_in[a] && !(l & 1)) w = -w;
  return w;
}

// NC coils per lane (NC = 2: one 16-byte load per tap covers two coils, which
// halves the vector-memory instructions the texture path has to address).
template <int NC>
struct CoilVec;
template <>
struct CoilVec<1> {
  float2 v[1];
  __device__ __forceinline__ static CoilVec load(const float2 *p) {
    CoilVec r;
    r.v[0] = *p;
    return r;
  }
};
template <>
struct CoilVec<2> {
  float2 v[2];
  __device__ __forceinline__ static CoilVec load(const float2 *p) {
    const float4 q = *reinterpret_cast<const float4 *>(p);
    CoilVec r;
    r.v[0] = make_float2(q.x, q.y);
    r.v[1] = make_float2(q.z, q.w);
    return r;
  }
};
template <int NC>  // NC = 4, 8: NC / 2 16-byte loads of consecutive coils
struct CoilVec {
  float2 v[NC];
  __device__ __forceinline__ static CoilVec load(const float2 *p) {
    CoilVec r;
#pragma unroll
    for (int h = 0; h < NC / 2; h++) {
      const float4 q = reinterpret_cast<const float4 *>(p)[h];
      r.v[2 * h] = make_float2(q.x, q.y);
      r.v[2 * h + 1] = make_float2(q.z, q.w);
    }
    return r;
  }
};

// message of the last failed call without an op (dcf.hip)
const char *free_error();
void free_err_clear();
int free_fail(int status, const char *msg);
// locally low-rank FISTA step (llr.hip)
int llr_fista_prepare(mrn_op *op);
int llr_fista_step(mrn_op *op, float2 *X, hipStream_t st);

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Renumber them so that each XCD runs runs of `run` consecutive work items:
// neighbouring tiles and samples, which read the same grid cells and k-space
// samples, then share one L2, while the runs still interleave over the XCDs
// so that dense and sparse regions of k-space spread evenly.  run = 1 keeps
// the hardware order.  A bijection of [0, nb): full groups of 8*run blocks
// are permuted, the tail keeps its order.
#define MRN_XCDS 8
__device__ __forceinline__ int xcd_block(int b, int nb, int run) {
  const int G = MRN_XCDS * run;
  if (run <= 1 || b >= (nb / G) * G) return b;
  const int w = b % G;
  return (b - w) + (w % MRN_XCDS) * run + w / MRN_XCDS;
}

// Scatter work item of wavefront `wave` (4 per workgroup) of workgroup b.
// run < 0 with per-XCD ranges xs: the dispatcher deals workgroups to the XCDs
// in turn, so workgroup b runs on XCD b % MRN_XCDS and takes the next 4 items
// of that XCD's contiguous range of the work list (tiles in row order): a
// sample's neighbouring tiles then share one L2.  -1: no item.  Otherwise
// xcd_block's runs (run <= 1: hardware order).
__device__ __forceinline__ int scatter_item(int b, int nb, int wave, int run, const int32_t *xs) {
  if (run < 0 && xs) {
    const int x = b % MRN_XCDS;
    const int i = xs[x] + (b / MRN_XCDS) * 4 + wave;
    return i < xs[x + 1] ? i : -1;
  }
  return xcd_block(b, nb, run) * 4 + wave;
}

// ------------------------------------------------------ FFT pass args
struct LinesArgs {
  int64_t off0, ystride, zstride;  // base of the workgroup's lines
  int64_t pstride, rstride;        // point stride along the axis; row stride
  int run;                         // contiguous lines per row
  int nlines;                      // valid lines in the group dimension
  int ld0, ld1, st0, st1;          // loaded / stored point ranges
  float scale;
  int xrun;                        // workgroup groups per XCD run (xcd_block; 0/1 = hardware order)
  // x and y passes of the batched form: k-space support of the line's row,
  // rng[blockIdx.y * rng_zs + (line / run) * rng_rs] = {p0, p1}; points outside
  // are not stored (forward, rng_store) or are zero (inverse, rng_load).  An
  // empty range skips a forward line's loads; {INT_MAX, INT_MIN} also skips an
  // inverse line's stores (no later pass reads that row).
  const int2 *rng;
  int rng_zs, rng_rs, rng_store, rng_load;
  // launches over several frames (blockIdx.z = frame * rng_nz + coil block):
  // frame f's support at rng + f * rng_fs (one frame: rng_fs = 0)
  int rng_nz;
  int64_t rng_fs;
};

struct PrePostArgs {
  int N[3], S[3], og[3];
  int CC, lgCC;
  int64_t chunk_grid;   // grid elements per coil block
  int64_t chunk_smap;   // smap elements per coil block
  int ch;               // coil block (post: one launch per block)
  int overwrite;        // post: X = sum instead of X += sum
  float scale;
  int xrun;             // x-column workgroups per XCD run (xcd_block)
  // launches over several frames: pre blockIdx.z = frame * nchf + coil
  // block, post blockIdx.z = frame; the frame's image at X + frame * xfs and
  // its grids at coil blocks frame * nchf ... (one frame: nf = 1, xfs = 0)
  int nf, nchf;
  int64_t xfs;
};

// taps per axis of the unrolled gather and of the tiles' bin reach: 2 dwin
// when that is an integer (a sample exactly on the window edge has one more),
// floor(2 dwin) + 1 otherwise
int taps_main(const Geom &g);

// host-side launchers (defined in the .hip files)
int launch_gather(mrn_op *op, const float2 *grid, float2 *out, const FrameData &fr, const float2 *sub,
                  bool weighted, hipStream_t st);
int launch_weight(mrn_op *op, float2 *kout, const float2 *kin, const FrameData &fr, hipStream_t st);
// the scatter: overwrite the tiles it reaches (zeroing the split tiles first),
// or with accum add into the grid's contents
int launch_scatter(mrn_op *op, float2 *grid, const float2 *kin, const FrameData &fr, hipStream_t st,
                   bool accum = false);
int launch_scatter_zero(mrn_op *op, float2 *grid, const FrameData &fr, hipStream_t st,
                        const FrameDev *fdt = nullptr);
int launch_scatter_tiles(mrn_op *op, float2 *grid, const float2 *kin, const FrameData &fr, hipStream_t st,
                         bool accum = false, const FrameDev *fdt = nullptr);
// every frame at once (op->fdev set up): gather from / scatter into the
// frames' grids, k-space in their kwork arrays
int launch_gather_frames(mrn_op *op, int fsub, hipStream_t st);
int launch_scatter_frames(mrn_op *op, hipStream_t st);
int fft_lines(mrn_op *op, int S, int dir, float2 *grid, const LinesArgs &a, dim3 grid3,
              const float2 *tw, hipStream_t st);
bool fft_x_pruned(const mrn_op *op);
bool fft_xy_pruned(const mrn_op *op);
int fft_pre_z(mrn_op *op, float2 *grid, const float2 *X, const float2 *smaps,
              const PrePostArgs &a, hipStream_t st);
int fft_post_z(mrn_op *op, const float2 *grid, float2 *X, const float2 *smaps,
               const PrePostArgs &a, hipStream_t st);
bool fft_size_supported(int S);
int wavelet_host(mrn_op *op, float2 *X, int frames, const int levels[3], const int32_t shift[3], bool inv,
                 hipStream_t st);
int l1_wavelet_host(mrn_op *op, float2 *X, int frames, float fraction, const int levels[3], int spins,
                    const int32_t *shifts, double *thr, hipStream_t st);
int l1_fista_step(mrn_op *op, float2 *X, hipStream_t st);
int ensure_l1_state(mrn_op *op);

}  // namespace mrn
