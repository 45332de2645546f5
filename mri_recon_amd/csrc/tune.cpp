// Autotuner of the batched form's launch shapes (no counterpart in the
// reference, whose gridding has no launch shapes).
//
// Timed on the frame's own trajectory, one stage at a time, in this order:
//   1. scatter: sample bin edge B in {1, 2} x samples per work item
//      {1024, 4096, 16384} (the bins are rebuilt for each), and for 8- and
//      16-coil blocks the matrix-core or vector-ALU accumulate of the 4^3 tiles;
//   2. gather: coils per lane {1, 2, 4, 8} (<= coils per block) x XCD run
//      length {16, 64, 256} workgroups;
//   3. FFT: XCD run length of the y and z passes {1, 4, 8, 16};
//   4. FFT pass form: four-step register passes with k-space support
//      pruning, or LDS Stockham passes (bins rebuilt: pruning changes the
//      scatter's work list), timed over a whole E'E.
// Every candidate computes the same operator: the gather and FFT run
// candidates give the same bits, the others differ by FP32 summation order.
//
// The choice is cached in a JSON file keyed by the geometry, kernel, sample
// count, coil count and device; a cache hit applies it without timing.  On
// any failure the op's settings are restored and every frame is rebuilt for
// them, so the op stays usable.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cctype>
#include <cerrno>
#include <map>
#include <sstream>
#include <string>

#include "mrn_internal.hpp"

namespace mrn {
int build_bins(mrn_op *, FrameData &, int64_t, const float *, const float *, const float *, const float *,
               hipStream_t);
int op_permute_in(mrn_op *, float2 *, const float2 *, const FrameData &, hipStream_t);
int op_fill(mrn_op *, float2 *, int64_t, float2, hipStream_t);
int fwd_grid_pub(mrn_op *op, const float2 *X, const FrameData *fr, hipStream_t st);
int adj_grid_pub(mrn_op *op, float2 *X, bool overwrite, const FrameData *fr, hipStream_t st);

namespace {

struct Knobs {
  int B, maxv, xrun_gather, gather_coils, xcd_run_fft, fft_stockham, scatter_valu;
};

Knobs current(const mrn_op *op) {
  return Knobs{op->t.B, op->t.maxv, op->t.xrun_gather, op->gather_coils, op->xcd_run_fft,
               op->fft_stockham ? 1 : 0, op->scatter_valu ? 1 : 0};
}

void set_bins(mrn_op *op, int B) {
  op->t.B = B;
  for (int a = 0; a < 3; a++) op->t.nb[a] = (op->g.S[a] + B - 1) / B;
  op->t.nbins = op->t.nb[0] * op->t.nb[1] * op->t.nb[2];
}

void apply(mrn_op *op, const Knobs &k) {
  set_bins(op, k.B);
  op->t.maxv = k.maxv;
  op->t.xrun_gather = k.xrun_gather;
  op->gather_coils = k.gather_coils;
  op->xcd_run_fft = k.xcd_run_fft;
  op->fft_stockham = k.fft_stockham != 0;
  op->scatter_valu = k.scatter_valu != 0;
}

int rebuild(mrn_op *op, FrameData &fr, hipStream_t st) {
  MRN_TRY(build_bins(op, fr, fr.M, fr.traw, fr.traw + fr.M, fr.traw + 2 * fr.M, fr.traw + 3 * fr.M, st));
  if (fr.has_kdata) MRN_TRY(op_permute_in(op, fr.kdata, fr.kraw, fr, st));
  return MRN_OK;
}

int rebuild_all(mrn_op *op, hipStream_t st) {
  for (auto &fr : op->frames)
    if (fr.has_traj) MRN_TRY(rebuild(op, fr, st));
  MRN_HIP(op, hipStreamSynchronize(st));
  return MRN_OK;
}

// fault injection for the error-path test: MRN_TUNE_FAIL_AFTER=n fails the
// (n+1)-th timed candidate
int fail_after() {
  const char *s = getenv("MRN_TUNE_FAIL_AFTER");
  return s && *s ? atoi(s) : -1;
}

// ---------------------------------------------------------------- cache
std::string cache_key(const mrn_op *op, const FrameData &fr) {
  int dev = 0;
  hipDeviceProp_t prop;
  std::string dname = "unknown";
  int cus = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    dname = prop.gcnArchName;
    cus = prop.multiProcessorCount;
  }
  for (char &c : dname)
    if (c == '"' || c == '\\' || c == ':') c = '_';
  char buf[512];
  snprintf(buf, sizeof buf,
           "N=%dx%dx%d S=%dx%dx%d W=%.4gx%.4gx%.4g kernel=%d kmode=%d M=%lld C=%d CC=%d device=%s/%dcu",
           op->g.N[0], op->g.N[1], op->g.N[2], op->g.S[0], op->g.S[1], op->g.S[2], 2 * op->g.dwin[0],
           2 * op->g.dwin[1], 2 * op->g.dwin[2], op->p.kernel, op->g.kmode, (long long)fr.M, op->C, op->g.CC,
           dname.c_str(), cus);
  return buf;
}

typedef std::map<std::string, std::map<std::string, long long>> Cache;

// A JSON object of objects of integers: {"key": {"name": 1, ...}, ...}.
// Anything else makes the file unreadable as a cache (it is then rewritten).
struct Parser {
  const std::string &s;
  size_t i = 0;
  explicit Parser(const std::string &t) : s(t) {}
  void ws() {
    while (i < s.size() && isspace((unsigned char)s[i])) i++;
  }
  bool ch(char c) {
    ws();
    if (i < s.size() && s[i] == c) {
      i++;
      return true;
    }
    return false;
  }
  bool str(std::string &out) {
    ws();
    if (i >= s.size() || s[i] != '"') return false;
    out.clear();
    for (i++; i < s.size() && s[i] != '"'; i++) {
      if (s[i] == '\\' && i + 1 < s.size()) i++;
      out += s[i];
    }
    if (i >= s.size()) return false;
    i++;
    return true;
  }
  bool num(long long &v) {
    ws();
    const char *b = s.c_str() + i;
    char *e = nullptr;
    errno = 0;
    v = strtoll(b, &e, 10);
    if (e == b || errno) return false;
    i += (size_t)(e - b);
    return true;
  }
  bool cache(Cache &c) {
    if (!ch('{')) return false;
    if (ch('}')) return true;
    do {
      std::string key;
      if (!str(key) || !ch(':') || !ch('{')) return false;
      auto &m = c[key];
      if (!ch('}')) {
        do {
          std::string name;
          long long v;
          if (!str(name) || !ch(':') || !num(v)) return false;
          m[name] = v;
        } while (ch(','));
        if (!ch('}')) return false;
      }
    } while (ch(','));
    return ch('}');
  }
};

bool read_cache(const char *path, Cache &c) {
  FILE *f = fopen(path, "rb");
  if (!f) return false;
  std::string txt;
  char buf[4096];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) txt.append(buf, n);
  fclose(f);
  Parser p(txt);
  if (!p.cache(c)) {
    c.clear();
    return false;
  }
  return true;
}

void write_cache(const char *path, const Cache &c) {
  std::ostringstream o;
  o << "{\n";
  bool first = true;
  for (auto &kv : c) {
    o << (first ? "" : ",\n") << "  \"" << kv.first << "\": {";
    first = false;
    bool f2 = true;
    for (auto &nv : kv.second) {
      o << (f2 ? "" : ", ") << "\"" << nv.first << "\": " << nv.second;
      f2 = false;
    }
    o << "}";
  }
  o << "\n}\n";
  const std::string tmp = std::string(path) + ".tmp";
  FILE *f = fopen(tmp.c_str(), "wb");
  if (!f) return;  // the cache is an optimisation: no file, no cache
  const std::string txt = o.str();
  const bool ok = fwrite(txt.data(), 1, txt.size(), f) == txt.size();
  if (fclose(f) == 0 && ok) rename(tmp.c_str(), path);
  else remove(tmp.c_str());
}

std::map<std::string, long long> to_map(const Knobs &k, float ms) {
  return {{"bin_edge", k.B},          {"max_visits", k.maxv},         {"xcd_run_gather", k.xrun_gather},
          {"gather_coils", k.gather_coils}, {"xcd_run_fft", k.xcd_run_fft}, {"fft_stockham", k.fft_stockham},
          {"scatter_valu", k.scatter_valu}, {"normal_us", (long long)(ms * 1000.0f)}};
}

bool from_map(const std::map<std::string, long long> &m, Knobs &k, float &ms) {
  const char *names[8] = {"bin_edge",    "max_visits",   "xcd_run_gather", "gather_coils",
                          "xcd_run_fft", "fft_stockham", "scatter_valu",   "normal_us"};
  for (const char *n : names)
    if (!m.count(n)) return false;
  k.B = (int)m.at("bin_edge");
  k.maxv = (int)m.at("max_visits");
  k.xrun_gather = (int)m.at("xcd_run_gather");
  k.gather_coils = (int)m.at("gather_coils");
  k.xcd_run_fft = (int)m.at("xcd_run_fft");
  k.fft_stockham = (int)m.at("fft_stockham");
  k.scatter_valu = (int)m.at("scatter_valu");
  ms = (float)m.at("normal_us") / 1000.0f;
  // accept only values the op accepts
  return (k.B == 1 || k.B == 2) && k.maxv > 0 && k.xrun_gather > 0 && k.xcd_run_fft > 0 &&
         (k.gather_coils == 1 || k.gather_coils == 2 || k.gather_coils == 4 || k.gather_coils == 8);
}

}  // namespace

int autotune_impl(mrn_op *op, int f, const char *cache_path, hipStream_t st) {
  FrameData &fr = op->frames[f];
  const Knobs saved = current(op);
  op->tune_from_cache = 0;
  // restore the entry settings and rebuild every frame for them on any failure
  bool done = false;
  ScopeExit restore([&]() {
    if (done) return;
    std::string e = op->err;
    apply(op, saved);
    (void)rebuild_all(op, st);
    op->err = e;
  });
  const std::string key = cache_key(op, fr);
  Cache cache;
  if (cache_path && *cache_path) {
    read_cache(cache_path, cache);
    auto it = cache.find(key);
    Knobs k;
    float ms = 0.f;
    if (it != cache.end() && from_map(it->second, k, ms)) {
      if (k.scatter_valu && op->g.CC != 16 && op->g.CC != 8) k.scatter_valu = 0;
      if (k.gather_coils > op->g.CC) k.gather_coils = op->g.CC;
      apply(op, k);
      MRN_TRY(rebuild_all(op, st));
      op->tune_ms = ms;
      op->tune_from_cache = 1;
      done = true;
      return MRN_OK;
    }
  }
  Events ev;
  MRN_HIP(op, ev.create(2));
  int budget = fail_after();
  int ran = 0;
  // time `reps` calls of fn after one warm-up call
  auto timed = [&](int reps, auto fn, float &ms) -> int {
    if (budget >= 0 && ran++ >= budget) return fail(op, MRN_EHIP, "injected autotuner failure");
    MRN_TRY(fn());
    MRN_HIP(op, hipEventRecord(ev[0], st));
    for (int r = 0; r < reps; r++) MRN_TRY(fn());
    MRN_HIP(op, hipEventRecord(ev[1], st));
    MRN_HIP(op, hipEventSynchronize(ev[1]));
    MRN_HIP(op, hipEventElapsedTime(&ms, ev[0], ev[1]));
    ms /= reps;
    return MRN_OK;
  };
  const int64_t nk = fr.Mv * op->g.CC * op->nch;
  float2 *Xs = op->vX0;  // image scratch (ensure_solver has run)
  MRN_TRY(op_fill(op, Xs, op->nvox, make_float2(1.f, 0.f), st));
  Knobs best = saved;
  // 1. scatter
  float best_ms = 1e30f;
  for (int B : {1, 2})
    for (int maxv : {1024, 4096, 16384})
      for (int valu = 0; valu <= ((op->g.CC == 16 || op->g.CC == 8) ? 1 : 0); valu++) {
        Knobs k = best;
        k.B = B;
        k.maxv = maxv;
        k.scatter_valu = valu;
        apply(op, k);
        MRN_TRY(rebuild(op, fr, st));
        MRN_TRY(op_fill(op, op->kwork, nk, make_float2(1.f, 0.f), st));
        float ms = 0.f;
        MRN_TRY(timed(3, [&]() { return launch_scatter(op, op->grid, op->kwork, fr, st); }, ms));
        if (ms < best_ms) {
          best_ms = ms;
          best = k;
        }
      }
  apply(op, best);
  MRN_TRY(rebuild(op, fr, st));
  // 2. gather (the forward passes leave a grid for it to read)
  MRN_TRY(fwd_grid_pub(op, Xs, &fr, st));
  best_ms = 1e30f;
  Knobs gbest = best;
  for (int nc : {1, 2, 4, 8}) {
    if (nc > op->g.CC) continue;
    for (int run : {16, 64, 256}) {
      Knobs k = best;
      k.gather_coils = nc;
      k.xrun_gather = run;
      apply(op, k);
      float ms = 0.f;
      MRN_TRY(timed(3, [&]() { return launch_gather(op, op->grid, op->kwork, fr, nullptr, true, st); }, ms));
      if (ms < best_ms) {
        best_ms = ms;
        gbest = k;
      }
    }
  }
  best = gbest;
  apply(op, best);
  // 3. FFT run length of the y and z passes
  best_ms = 1e30f;
  Knobs fbest = best;
  for (int run : {1, 4, 8, 16}) {
    Knobs k = best;
    k.xcd_run_fft = run;
    apply(op, k);
    float ms = 0.f;
    MRN_TRY(timed(2, [&]() -> int {
      MRN_TRY(fwd_grid_pub(op, Xs, &fr, st));
      return adj_grid_pub(op, op->vR, true, &fr, st);
    }, ms));
    if (ms < best_ms) {
      best_ms = ms;
      fbest = k;
    }
  }
  best = fbest;
  // 4. FFT pass form, over a whole E'E
  auto normal = [&]() -> int {
    MRN_TRY(fwd_grid_pub(op, Xs, &fr, st));
    MRN_TRY(launch_gather(op, op->grid, op->kwork, fr, nullptr, true, st));
    MRN_TRY(launch_scatter(op, op->grid, op->kwork, fr, st));
    return adj_grid_pub(op, op->vR, true, &fr, st);
  };
  best_ms = 1e30f;
  Knobs nbest = best;
  for (int stock = 0; stock <= 1; stock++) {
    Knobs k = best;
    k.fft_stockham = stock;
    apply(op, k);
    MRN_TRY(rebuild(op, fr, st));
    float ms = 0.f;
    MRN_TRY(timed(2, normal, ms));
    if (ms < best_ms) {
      best_ms = ms;
      nbest = k;
    }
  }
  best = nbest;
  apply(op, best);
  MRN_TRY(rebuild_all(op, st));
  op->tune_ms = best_ms;
  done = true;
  if (cache_path && *cache_path) {
    cache[key] = to_map(best, best_ms);
    write_cache(cache_path, cache);
  }
  return MRN_OK;
}

}  // namespace mrn
