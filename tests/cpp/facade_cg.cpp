// SENSE-CG written the way recon_lib.cpp:1523-1759 writes it (a per-coil loop of
// gridding.backward / gridding.forward calls inside the iteration, FP64 inner
// products on the host), on the C++ gridFFT facade (include/mrn_gridfft.hpp),
// compared with the batched, graph-captured mrn::SenseOperator::cg on the same
// inputs.  Prints one line "facade_cg ... rel_err=<e>" and exits 0 when the two
// images agree to 1e-4 (relative L2).
//
//   facade_cg [N] [coils] [spokes] [points] [iters]
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mrn_gridfft.hpp"

using cfloat = std::complex<float>;
using cdouble = std::complex<double>;

#define HIPCHECK(x)                                                             \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));             \
      std::exit(2);                                                             \
    }                                                                           \
  } while (0)

template <typename T>
struct Dev {
  T *p = nullptr;
  size_t n = 0;
  explicit Dev(size_t n_) : n(n_) { HIPCHECK(hipMalloc((void **)&p, sizeof(T) * (n ? n : 1))); }
  ~Dev() { (void)hipFree(p); }
  void put(const std::vector<T> &h) { HIPCHECK(hipMemcpy(p, h.data(), sizeof(T) * n, hipMemcpyHostToDevice)); }
  void get(std::vector<T> &h) const {
    h.resize(n);
    HIPCHECK(hipMemcpy(h.data(), p, sizeof(T) * n, hipMemcpyDeviceToHost));
  }
};

// xorshift64* uniform in [0,1) and Box-Muller normals: a self-contained seeded source
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 2685821657736338717ULL + 1) {}
  double uni() {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return (double)((s * 2685821657736338717ULL) >> 11) * (1.0 / 9007199254740992.0);
  }
  double normal() {
    double u = uni(), v = uni();
    if (u < 1e-300) u = 1e-300;
    return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
  }
};

int main(int argc, char **argv) {
  const int N = argc > 1 ? std::atoi(argv[1]) : 32;
  const int C = argc > 2 ? std::atoi(argv[2]) : 4;
  const int NSP = argc > 3 ? std::atoi(argv[3]) : 600;
  const int NPT = argc > 4 ? std::atoi(argv[4]) : 32;
  const int ITERS = argc > 5 ? std::atoi(argv[5]) : 5;
  const int64_t M = (int64_t)NSP * NPT;
  const size_t nvox = (size_t)N * N * N;

  // golden-means kooshball, centre-out, kw = |k|^2 (reference units, |k| <= N/2)
  std::vector<float> kx(M), ky(M), kz(M), kw(M);
  for (int n = 0; n < NSP; n++) {
    const double gz = std::fmod(0.465571231876768 * n, 1.0), gp = std::fmod(0.682327803828019 * n, 1.0);
    const double dz = 2.0 * gz - 1.0, phi = 6.283185307179586 * gp, rxy = std::sqrt(std::fmax(1.0 - dz * dz, 0.0));
    for (int p = 0; p < NPT; p++) {
      const double r = 0.5 * N * p / (NPT - 1);
      const int64_t m = (int64_t)n * NPT + p;
      kx[m] = (float)(r * rxy * std::cos(phi));
      ky[m] = (float)(r * rxy * std::sin(phi));
      kz[m] = (float)(r * dz);
      kw[m] = kx[m] * kx[m] + ky[m] * ky[m] + kz[m] * kz[m];
    }
  }
  // image, and smooth coil maps normalised to sum_c |S_c|^2 = 1
  Rng rng(1234);
  std::vector<cfloat> X0(nvox), smaps(nvox * C);
  for (auto &v : X0) v = cfloat((float)(rng.normal() * 0.7071), (float)(rng.normal() * 0.7071));
  std::vector<double> sos(nvox, 0.0);
  for (int c = 0; c < C; c++) {
    const double a = 6.283185307179586 * c / C, cx = 0.6 * N * std::cos(a), cy = 0.6 * N * std::sin(a);
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          const double x = i - N / 2, y = j - N / 2, z = k - N / 2;
          const double d2 = ((x - cx) * (x - cx) + (y - cy) * (y - cy) + z * z) / (double)(N * N);
          const double ph = 0.3 * c + 0.05 * (x + 2 * y);
          const size_t r = ((size_t)k * N + j) * N + i;
          smaps[c * nvox + r] = cfloat((float)(std::exp(-d2) * std::cos(ph)), (float)(std::exp(-d2) * std::sin(ph)));
          sos[r] += std::norm(std::complex<double>(smaps[c * nvox + r]));
        }
  }
  for (int c = 0; c < C; c++)
    for (size_t r = 0; r < nvox; r++) smaps[c * nvox + r] /= (float)std::sqrt(sos[r]);

  Dev<float> dkx(M), dky(M), dkz(M), dkw(M);
  dkx.put(kx), dky.put(ky), dkz.put(kz), dkw.put(kw);
  Dev<cfloat> dsm(nvox * C), dX0(nvox), dd((size_t)M * C);
  dsm.put(smaps), dX0.put(X0);

  try {
    // ---- the reference's per-coil structure on the facade
    mrn::gridFFT gridding;
    gridding.overgrid = 1.5f;
    const char *args[] = {"-dwinX", "2.0", "-dwinY", "2.0", "-dwinZ", "2.0"};
    gridding.read_commandline(6, args);
    gridding.precalc_gridding(N, N, N);
    // simulate: d_c = A_c X0 (cf. commandline_recon.cxx:80-125)
    for (int c = 0; c < C; c++)
      gridding.backward(dX0.p, dsm.p + c * nvox, dd.p + (size_t)c * M, dkx.p, dky.p, dkz.p, dkw.p, M);

    Dev<cfloat> dR(nvox), dP(nvox), dL(nvox), ddiff(M);
    std::vector<cfloat> X(nvox, 0.f), R, P, L;
    HIPCHECK(hipMemset(dR.p, 0, sizeof(cfloat) * nvox));
    for (int c = 0; c < C; c++)  // R = E^H W d (recon_lib.cpp:1565-1567)
      gridding.forward(dR.p, dsm.p + c * nvox, dd.p + (size_t)c * M, dkx.p, dky.p, dkz.p, dkw.p, M);
    dR.get(R);
    P = R;
    for (int it = 0; it < ITERS; it++) {
      dP.put(P);
      HIPCHECK(hipMemset(dL.p, 0, sizeof(cfloat) * nvox));
      for (int c = 0; c < C; c++) {  // E'E P, one coil at a time (1614-1622)
        gridding.backward(dP.p, dsm.p + c * nvox, ddiff.p, dkx.p, dky.p, dkz.p, dkw.p, M);
        gridding.forward(dL.p, dsm.p + c * nvox, ddiff.p, dkx.p, dky.p, dkz.p, dkw.p, M);
      }
      dL.get(L);
      cdouble rr = 0.0, pl = 0.0, rn = 0.0;  // FP64 sums (1659-1687)
      for (size_t r = 0; r < nvox; r++) {
        rr += std::norm(std::complex<double>(R[r]));
        pl += std::conj(cdouble(P[r])) * cdouble(L[r]);
      }
      const cdouble a = rr / pl;
      const cfloat af((float)a.real(), (float)a.imag());
      for (size_t r = 0; r < nvox; r++) {
        X[r] += af * P[r];
        R[r] -= af * L[r];
        rn += std::norm(std::complex<double>(R[r]));
      }
      const cdouble b = rn / rr;
      const cfloat bf((float)b.real(), (float)b.imag());
      for (size_t r = 0; r < nvox; r++) P[r] = R[r] + bf * P[r];
    }

    // ---- batched, graph-captured form on the same inputs
    mrn_params p;
    mrn_params_default(&p);
    p.N[0] = p.N[1] = p.N[2] = N;
    p.overgrid = 1.5f;
    p.dwin[0] = p.dwin[1] = p.dwin[2] = 2.0f;
    p.ncoils = C;
    mrn::SenseOperator op(p);
    op.set_traj(0, M, dkx.p, dky.p, dkz.p, dkw.p);
    op.set_smaps(dsm.p);
    op.set_kdata(0, dd.p);
    Dev<cfloat> dXb(nvox);
    op.cg(dXb.p, ITERS, true);
    op.sync();
    std::vector<cfloat> Xb;
    dXb.get(Xb);

    double num = 0.0, den = 0.0;
    for (size_t r = 0; r < nvox; r++) {
      num += std::norm(std::complex<double>(Xb[r]) - std::complex<double>(X[r]));
      den += std::norm(std::complex<double>(X[r]));
    }
    const double err = std::sqrt(num / den);
    std::printf("facade_cg N=%d coils=%d M=%lld iters=%d S=%d beta=%.4f rel_err=%.3e\n", N, C,
                (long long)M, ITERS, gridding.Sx, gridding.betaX, err);
    return err < 1e-4 ? 0 : 1;
  } catch (const mrn::Error &e) {
    std::fprintf(stderr, "facade_cg: %s\n", e.what());
    return 3;
  }
}
