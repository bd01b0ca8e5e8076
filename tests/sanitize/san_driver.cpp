// Sanitizer driver for the host-only C++ of libxrtk and for the oracle.
//
// Built twice by the Makefile beside it, with g++/gcc only (no device code):
// once under AddressSanitizer + UndefinedBehaviorSanitizer, once under
// ThreadSanitizer.  It runs every host-only entry of the library
// (xrtk_prepare_wave, xrtk_beam_to_rays, the two CPU solves the auto entry
// points fall back to) and every oracle function on small inputs, including
// the edge cases of the device tests (no rays, no pixels, one pixel, a
// pixel on a ray, partial 64-pixel blocks, gather indices out of range), and
// checks each library result against the oracle:
//   prepare_wave, beam_to_rays       bit-identical
//   cpu_diffraction_f64              bit-identical to oracle_diffraction_f64
//                                    (same operation order and Kahan sums)
//   cpu_fallback_f32                 bit-identical to oracle_fallback_f32
// A sanitizer report aborts the run (halt_on_error); a mismatch exits 1.
// With --plant-overflow or --plant-race it runs only that planted fault.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/xrtk.h"
#include "../../xrt-rs_amd/csrc/host_copy.h"

namespace xrtk {
int cpu_fallback_f32(const XrtkGpuRay* rays, size_t nr, const XrtkGpuPixel* px, size_t np,
                     double* out4);
int cpu_diffraction_f64(const XrtkRaysF64* r, size_t nr, const double* px, const double* py,
                        const double* pz, size_t np, double* out4);
}  // namespace xrtk

extern "C" {
void oracle_diffraction_f64(const double* const cols[12], const double* path_up, size_t nr,
                            const double* px, const double* py, const double* pz, size_t np,
                            size_t pix_stride, int threads, double* out4);
void oracle_diffraction_f64_sincos(const double* const cols[12], const double* path_up,
                                   size_t nr, const double* px, const double* py,
                                   const double* pz, size_t np, size_t pix_stride, int threads,
                                   double* out4);
void oracle_diffraction_ld(const double* const cols[12], const double* path_up, size_t nr,
                           const double* px, const double* py, const double* pz, size_t np,
                           size_t pix_stride, int threads, double* out4);
void oracle_fallback_f32(const float* rays, size_t nr, const float* pix, size_t np, int threads,
                         double* out4);
void oracle_reduced_f32(const float* rays, size_t nr, const float* pix, size_t np, double* out4);
void oracle_prepare_wave(const double center[3], double dx, double dz, size_t nx, size_t nz,
                         double* px, double* py, double* pz);
void oracle_fresnel_grid(double lo, double hi, size_t n, double* v);
void oracle_fresnel_f64(const double* rx, const double* rz, const double* es_re,
                        const double* es_im, size_t nr, double energy, double dist,
                        const double* xp, size_t nx, const double* zp, size_t nz,
                        size_t pix_stride, int threads, double* out2);
double oracle_undulator_gamma(double electron_energy_gev);
double oracle_undulator_e1(double gamma, double kx, double ky, double period_mm);
void oracle_undulator_i_map(double gamma, double gamma2, double beam_current, double kx,
                            double ky, double period_mm, size_t n_periods, double phase_deg,
                            int n_steps, const double* energies, const double* thetas,
                            const double* psis, size_t n, int threads, double* intensity,
                            double* amp_s, double* amp_p);
void oracle_undulator_wgsl_f32(double gamma_d, double beam_current, double kx_d, double ky_d,
                               double period_mm, size_t n_periods, double phase_deg,
                               int n_steps, const float* obs4, size_t n, double* out5);
}

namespace {

int g_checks = 0, g_fail = 0;

void check(bool ok, const char* what) {
  ++g_checks;
  if (!ok) {
    ++g_fail;
    std::fprintf(stderr, "FAIL: %s\n", what);
  }
}

bool same_bits(const double* a, const double* b, size_t n) {
  return n == 0 || std::memcmp(a, b, n * sizeof(double)) == 0;
}

// Deterministic uniform numbers in [0, 1) (64-bit LCG, top 53 bits).
struct Lcg {
  uint64_t s;
  double next() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
  double sym(double h) { return (2.0 * next() - 1.0) * h; }
};

// Rays of a 0.2 mm aperture 10 m from a point source, 9-11 keV, per-ray
// normals tilted by up to 0.1 rad, complex Es and Ep.
struct Rays {
  std::vector<double> c[12];
  std::vector<double> path_up;
  size_t n = 0;
  explicit Rays(size_t nr, uint64_t seed) : n(nr) {
    Lcg g{seed};
    for (auto& v : c) v.resize(nr);
    path_up.assign(nr, 0.0);
    for (size_t j = 0; j < nr; ++j) {
      const double x = g.sym(0.1), z = g.sym(0.1), y = 0.0;
      double nx = g.sym(0.1), nz = g.sym(0.1), ny = 1.0;
      const double nn = std::sqrt(nx * nx + ny * ny + nz * nz);
      nx /= nn; ny /= nn; nz /= nn;
      const double d = std::sqrt(x * x + (y + 10000.0) * (y + 10000.0) + z * z);
      c[0][j] = x; c[1][j] = y; c[2][j] = z;
      c[3][j] = nx; c[4][j] = ny; c[5][j] = nz;
      c[6][j] = (nx * x + ny * (y + 10000.0) + nz * z) / d;
      c[7][j] = g.sym(1.0); c[8][j] = g.sym(1.0);
      c[9][j] = g.sym(0.5); c[10][j] = g.sym(0.5);
      c[11][j] = 9000.0 + 2000.0 * g.next();
    }
  }
  XrtkRaysF64 view() const {
    XrtkRaysF64 r;
    r.x = c[0].data(); r.y = c[1].data(); r.z = c[2].data();
    r.nx = c[3].data(); r.ny = c[4].data(); r.nz = c[5].data();
    r.nl = c[6].data();
    r.es_re = c[7].data(); r.es_im = c[8].data();
    r.ep_re = c[9].data(); r.ep_im = c[10].data();
    r.energy = c[11].data();
    return r;
  }
  void cols(const double* out[12]) const {
    for (int q = 0; q < 12; ++q) out[q] = c[q].data();
  }
};

struct Screen {
  std::vector<double> x, y, z;
  explicit Screen(size_t np, uint64_t seed) : x(np), y(np), z(np) {
    Lcg g{seed};
    for (size_t p = 0; p < np; ++p) {
      x[p] = g.sym(0.05);
      y[p] = 1000.0;
      z[p] = g.sym(0.05);
    }
  }
};

void test_prepare_wave() {
  const uint32_t shapes[][2] = {{1, 1}, {1, 7}, {5, 1}, {33, 17}, {64, 64}};
  for (const auto& s : shapes) {
    XrtkWaveParams w;
    w.center[0] = 0.25; w.center[1] = 1000.0; w.center[2] = -0.5;
    w.dx = 0.03; w.dz = 0.07; w.nx = s[0]; w.nz = s[1];
    const size_t n = (size_t)s[0] * s[1];
    std::vector<double> a(3 * n), b(3 * n);
    check(xrtk_prepare_wave(&w, a.data(), a.data() + n, a.data() + 2 * n) == XRTK_OK,
          "prepare_wave status");
    oracle_prepare_wave(w.center, w.dx, w.dz, s[0], s[1], b.data(), b.data() + n,
                        b.data() + 2 * n);
    check(same_bits(a.data(), b.data(), 3 * n), "prepare_wave vs oracle");
  }
  XrtkWaveParams w{};
  w.nx = 0; w.nz = 5;
  check(xrtk_prepare_wave(&w, nullptr, nullptr, nullptr) == XRTK_OK, "prepare_wave empty");
  w.nx = 2;
  check(xrtk_prepare_wave(&w, nullptr, nullptr, nullptr) == XRTK_EINVAL,
        "prepare_wave null outputs");
  check(xrtk_prepare_wave(nullptr, nullptr, nullptr, nullptr) == XRTK_EINVAL,
        "prepare_wave null params");
}

void test_beam_to_rays() {
  const size_t nb = 257;
  Lcg g{7};
  std::vector<double> b[7], es[4], normals(3 * nb);
  for (auto& v : b) { v.resize(nb); for (auto& x : v) x = g.sym(1.0); }
  for (auto& v : es) { v.resize(nb); for (auto& x : v) x = g.sym(1.0); }
  for (auto& x : normals) x = g.sym(1.0);
  std::vector<size_t> good;
  for (size_t i = 0; i < nb; i += 3) good.push_back(nb - 1 - i);  // descending, every 3rd
  const size_t ng = good.size();
  for (int amps = 0; amps < 2; ++amps) {
    std::vector<double> out(12 * ng, -7.0);
    double* cols[12];
    for (int q = 0; q < 12; ++q) cols[q] = out.data() + q * ng;
    const int st = xrtk_beam_to_rays(
        b[0].data(), b[1].data(), b[2].data(), b[3].data(), b[4].data(), b[5].data(),
        b[6].data(), amps ? es[0].data() : nullptr, amps ? es[1].data() : nullptr,
        amps ? es[2].data() : nullptr, amps ? es[3].data() : nullptr, nb, normals.data(),
        good.data(), ng, cols);
    check(st == XRTK_OK, "beam_to_rays status");
    bool ok = true;
    for (size_t k = 0; k < ng; ++k) {
      const size_t i = good[k];
      const double* n = normals.data() + 3 * i;
      const double nl = n[0] * b[3][i] + n[1] * b[4][i] + n[2] * b[5][i];
      const double want[12] = {b[0][i], b[1][i], b[2][i], n[0], n[1], n[2], nl,
                               amps ? es[0][i] : 1.0, amps ? es[1][i] : 0.0,
                               amps ? es[2][i] : 0.0, amps ? es[3][i] : 0.0, b[6][i]};
      for (int q = 0; q < 12; ++q) ok = ok && std::memcmp(&cols[q][k], &want[q], 8) == 0;
    }
    check(ok, "beam_to_rays gather");
  }
  // an index past the beam: EINVAL before anything is written
  std::vector<size_t> bad = {0, nb};
  std::vector<double> out(24, -7.0);
  double* cols[12];
  for (int q = 0; q < 12; ++q) cols[q] = out.data() + 2 * q;
  const int st = xrtk_beam_to_rays(b[0].data(), b[1].data(), b[2].data(), b[3].data(),
                                   b[4].data(), b[5].data(), b[6].data(), nullptr, nullptr,
                                   nullptr, nullptr, nb, normals.data(), bad.data(), 2, cols);
  bool untouched = true;
  for (double v : out) untouched = untouched && v == -7.0;
  check(st == XRTK_EINVAL && untouched, "beam_to_rays index out of range");
  check(xrtk_beam_to_rays(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                          nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr, 0,
                          cols) == XRTK_OK,
        "beam_to_rays empty");
  check(xrtk_beam_to_rays(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                          nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr, 0,
                          nullptr) == XRTK_EINVAL,
        "beam_to_rays null outputs");
}

void test_diffraction_f64() {
  const size_t nrs[] = {0, 1, 37, 600};
  const size_t nps[] = {0, 1, 63, 64, 65, 300};
  for (size_t nr : nrs) {
    Rays r(nr, 11 + nr);
    const XrtkRaysF64 v = r.view();
    const double* cols[12];
    r.cols(cols);
    for (size_t np : nps) {
      Screen s(np, 5 + np);
      if (nr && np) {  // one pixel exactly on a ray: the pair is skipped
        s.x[0] = r.c[0][0]; s.y[0] = r.c[1][0]; s.z[0] = r.c[2][0];
      }
      std::vector<double> a(4 * np + 4, 3.0), b(4 * np + 4, 3.0), c(4 * np + 4, 3.0);
      check(xrtk::cpu_diffraction_f64(&v, nr, s.x.data(), s.y.data(), s.z.data(), np,
                                      a.data()) == XRTK_OK,
            "cpu_diffraction_f64 status");
      oracle_diffraction_f64(cols, r.path_up.data(), nr, s.x.data(), s.y.data(), s.z.data(),
                             np, 1, 4, b.data());
      oracle_diffraction_f64_sincos(cols, r.path_up.data(), nr, s.x.data(), s.y.data(),
                                    s.z.data(), np, 1, 3, c.data());
      char what[96];
      std::snprintf(what, sizeof what, "cpu_diffraction_f64 vs oracle nr=%zu np=%zu", nr, np);
      check(same_bits(a.data(), b.data(), 4 * np + 4), what);
      check(same_bits(b.data(), c.data(), 4 * np + 4), "oracle sin_cos vs sincos");
      if (np >= 4 && nr >= 37) {
        // long double against f64 on every 3rd pixel: phases ~ 5e8 rad here,
        // so FP64 phase rounding bounds the gap
        std::vector<double> ld(4 * ((np + 2) / 3));
        oracle_diffraction_ld(cols, r.path_up.data(), nr, s.x.data(), s.y.data(), s.z.data(),
                              np, 3, 2, ld.data());
        double num = 0, den = 0;
        for (size_t q = 0; q < (np + 2) / 3; ++q)
          for (int e = 0; e < 4; ++e) {
            const double d = ld[4 * q + e] - b[4 * 3 * q + e];
            num += d * d;
            den += b[4 * 3 * q + e] * b[4 * 3 * q + e];
          }
        check(std::sqrt(num / den) < 1e-6, "oracle ld vs f64");
      }
    }
  }
}

void test_fallback_f32() {
  const size_t nrs[] = {0, 1, 129};
  const size_t nps[] = {0, 1, 65, 200};
  for (size_t nr : nrs) {
    Rays r(nr, 3 + nr);
    std::vector<XrtkGpuRay> rays(nr);
    std::vector<float> rf(12 * nr);
    for (size_t j = 0; j < nr; ++j) {
      float* f = rf.data() + 12 * j;
      const int order[12] = {0, 1, 2, 3, 4, 5, 6, 11, 7, 8, 9, 10};  // GpuRay field order
      for (int q = 0; q < 12; ++q) f[q] = (float)r.c[order[q]][j];
      std::memcpy(&rays[j], f, sizeof(XrtkGpuRay));
    }
    for (size_t np : nps) {
      Screen s(np, 9 + np);
      std::vector<XrtkGpuPixel> px(np);
      std::vector<float> pf(4 * np);
      for (size_t p = 0; p < np; ++p) {
        pf[4 * p] = (float)s.x[p]; pf[4 * p + 1] = (float)s.y[p];
        pf[4 * p + 2] = (float)s.z[p]; pf[4 * p + 3] = 0.f;
        std::memcpy(&px[p], &pf[4 * p], sizeof(XrtkGpuPixel));
      }
      if (nr && np) std::memcpy(&px[0], &rf[0], 12), std::memcpy(&pf[0], &rf[0], 12);
      std::vector<double> a(4 * np + 4, 3.0), b(4 * np + 4, 3.0), c(4 * np + 4, 3.0);
      check(xrtk::cpu_fallback_f32(rays.data(), nr, px.data(), np, a.data()) == XRTK_OK,
            "cpu_fallback_f32 status");
      oracle_fallback_f32(rf.data(), nr, pf.data(), np, 4, b.data());
      check(same_bits(a.data(), b.data(), 4 * np + 4), "cpu_fallback_f32 vs oracle");
      oracle_reduced_f32(rf.data(), nr, pf.data(), np, c.data());
      // finite for any rays; for none the reference's centroid is 0 / 0 and
      // its base phase makes every output NaN (the library gives zeros)
      bool ok = true;
      for (size_t i = 0; i < 4 * np; ++i) ok = ok && (nr ? std::isfinite(c[i]) : std::isnan(c[i]));
      check(ok, "oracle_reduced_f32 finite (NaN without rays)");
    }
  }
}

void test_fresnel_and_undulator() {
  const size_t nr = 300, nx = 17, nz = 9;
  Lcg g{21};
  std::vector<double> rx(nr), rz(nr), er(nr), ei(nr), xp(nx), zp(nz);
  for (size_t j = 0; j < nr; ++j) {
    rx[j] = g.sym(0.1); rz[j] = g.sym(0.1); er[j] = g.sym(1.0); ei[j] = g.sym(1.0);
  }
  oracle_fresnel_grid(-0.05, 0.05, nx, xp.data());
  oracle_fresnel_grid(-0.05, 0.05, nz, zp.data());
  std::vector<double> full(2 * nx * nz), part(2 * ((nx * nz + 4) / 5));
  oracle_fresnel_f64(rx.data(), rz.data(), er.data(), ei.data(), nr, 10000.0, 1000.0, xp.data(),
                     nx, zp.data(), nz, 1, 4, full.data());
  oracle_fresnel_f64(rx.data(), rz.data(), er.data(), ei.data(), nr, 10000.0, 1000.0, xp.data(),
                     nx, zp.data(), nz, 5, 2, part.data());
  bool ok = true;
  for (size_t q = 0; q < part.size() / 2; ++q)
    ok = ok && std::memcmp(&part[2 * q], &full[2 * 5 * q], 16) == 0;
  check(ok, "oracle_fresnel_f64 strided vs full");

  const double gamma = oracle_undulator_gamma(3.0);
  const double e1 = oracle_undulator_e1(gamma, 0.0, 1.2, 20.0);
  const size_t n = 77;
  std::vector<double> en(n), th(n), ps(n), in(n), as(2 * n), ap(2 * n), o5(5 * n);
  std::vector<float> obs(4 * n);
  for (size_t i = 0; i < n; ++i) {
    en[i] = e1 * (0.9 + 0.2 * g.next());
    th[i] = g.sym(2e-5);
    ps[i] = g.sym(2e-5);
    obs[4 * i] = (float)en[i]; obs[4 * i + 1] = (float)th[i];
    obs[4 * i + 2] = (float)ps[i]; obs[4 * i + 3] = 0.f;
  }
  oracle_undulator_i_map(gamma, gamma * gamma, 0.5, 0.0, 1.2, 20.0, 100, 0.0, 40, en.data(),
                         th.data(), ps.data(), n, 4, in.data(), as.data(), ap.data());
  oracle_undulator_wgsl_f32(gamma, 0.5, 0.0, 1.2, 20.0, 100, 0.0, 40, obs.data(), n, o5.data());
  bool fin = std::isfinite(e1) && e1 > 0;
  for (size_t i = 0; i < n; ++i) fin = fin && std::isfinite(in[i]) && in[i] >= 0;
  check(fin, "oracle undulator finite");
}

// Planted faults, run only when asked: they show that the build's
// instrumentation is live (the sanitizer must stop the run at them).
int planted(const char* what) {
  if (std::strcmp(what, "--plant-overflow") == 0) {
    std::vector<double> v(5, 1.0);
    volatile double* p = v.data();
    return (int)p[v.size()];  // one past the end
  }
  if (std::strcmp(what, "--plant-race") == 0) {
    static int counter = 0;
    std::thread a([] { for (int i = 0; i < 100000; ++i) ++counter; });
    std::thread b([] { for (int i = 0; i < 100000; ++i) ++counter; });
    a.join();
    b.join();
    return counter == 0;
  }
  std::fprintf(stderr, "unknown option %s\n", what);
  return 2;
}

}  // namespace

// par_memcpy and gather_chunks (the page-locked staging of the solves) on
// ragged sizes around the thread and chunk cuts: every byte lands, chunks
// are handed on in order, whole, and only once every part is in; after a
// refused chunk no further one is handed on and the copy still completes.
void test_host_copies() {
  const size_t MiB = size_t(1) << 20;
  const size_t sizes[] = {0, 1, 63, 64, 65, 4 * MiB - 1, 4 * MiB, 4 * MiB + 7, 8 * MiB,
                          8 * MiB + 13, 21 * MiB + 5, 33 * MiB};
  std::vector<char> src(33 * MiB + 64), dst(33 * MiB + 64);
  for (size_t i = 0; i < src.size(); ++i) src[i] = (char)(i * 2654435761u >> 13);
  for (size_t n : sizes) {
    std::fill(dst.begin(), dst.end(), 0);
    xrtk::par_memcpy(dst.data(), src.data(), n);
    check(std::memcmp(dst.data(), src.data(), n) == 0, "par_memcpy bytes");
    check(n == dst.size() || dst[n] == 0, "par_memcpy bound");
    const size_t chunks[] = {8 * MiB, 3 * MiB + 1};
    for (size_t chunk : chunks) {
      std::fill(dst.begin(), dst.end(), 0);
      size_t next = 0;
      bool whole = true;
      const bool ok = xrtk::gather_chunks(dst.data(), src.data(), n, chunk,
                                          [&](size_t c0, size_t len) {
                                            whole = whole && c0 == next && len > 0 &&
                                                    std::memcmp(dst.data() + c0, src.data() + c0,
                                                                len) == 0;
                                            next = c0 + len;
                                            return true;
                                          });
      check(ok && whole && next == n, "gather_chunks order and contents");
      check(std::memcmp(dst.data(), src.data(), n) == 0, "gather_chunks bytes");
      check(n == dst.size() || dst[n] == 0, "gather_chunks bound");
      // refuse the second chunk
      std::fill(dst.begin(), dst.end(), 0);
      int calls = 0;
      const bool ok2 = xrtk::gather_chunks(dst.data(), src.data(), n, chunk,
                                           [&](size_t, size_t) { return ++calls < 2; });
      const size_t nchunks = n ? (n + chunk - 1) / chunk : 0;
      check(calls == (int)std::min<size_t>(nchunks, 2), "gather_chunks stops handing on");
      check(ok2 == (nchunks < 2), "gather_chunks reports the refusal");
      check(std::memcmp(dst.data(), src.data(), n) == 0, "gather_chunks completes the copy");
    }
  }
}

// scatter_chunks (results out of the page-locked block): segments with a
// skipped one between them, ragged sizes around the chunk cuts, every byte
// to its segment and nothing past it; a refused chunk stops the copy.
void test_scatter() {
  const size_t MiB = size_t(1) << 20;
  const size_t ns[] = {1, 3, 1000, 131071, 131072, 655361, 1048577};
  for (size_t n : ns) {
    const size_t pad = ((n + 1) / 2) * 2 - n;
    const size_t total = 8 * n + 8 * pad + 32 * n;
    std::vector<char> src(total);
    for (size_t i = 0; i < total; ++i) src[i] = (char)(i * 40503u >> 7);
    std::vector<char> a(8 * n + 64, 0), b(16 * n + 64, 0), c(16 * n + 64, 0);
    const xrtk::Segment segs[4] = {{a.data(), 8 * n}, {nullptr, 8 * pad}, {b.data(), 16 * n},
                                   {c.data(), 16 * n}};
    const size_t chunks[] = {4 * MiB, 1 * MiB + 3, 64};
    for (size_t chunk : chunks) {
      if (chunk < 4096 && n > 1000) continue;  // (many tiny chunks: small n only)
      std::fill(a.begin(), a.end(), 0);
      std::fill(b.begin(), b.end(), 0);
      std::fill(c.begin(), c.end(), 0);
      size_t calls = 0;
      const bool ok = xrtk::scatter_chunks(src.data(), segs, 4, chunk, [&](size_t j) {
        return j == calls++;
      });
      check(ok && calls == (total + chunk - 1) / chunk, "scatter_chunks waits for every chunk");
      check(std::memcmp(a.data(), src.data(), 8 * n) == 0, "scatter_chunks segment 0");
      check(std::memcmp(b.data(), src.data() + 8 * (n + pad), 16 * n) == 0,
            "scatter_chunks segment 2");
      check(std::memcmp(c.data(), src.data() + 8 * (n + pad) + 16 * n, 16 * n) == 0,
            "scatter_chunks segment 3");
      bool clean = true;
      for (size_t i = 0; i < 64; ++i)
        clean = clean && a[8 * n + i] == 0 && b[16 * n + i] == 0 && c[16 * n + i] == 0;
      check(clean, "scatter_chunks bounds");
    }
    // refuse the first chunk: nothing is copied
    std::fill(a.begin(), a.end(), 0);
    const bool ok2 = xrtk::scatter_chunks(src.data(), segs, 4, 4096, [](size_t) { return false; });
    bool none = !ok2;
    for (size_t i = 0; i < 8 * n; ++i) none = none && a[i] == 0;
    check(none, "scatter_chunks stops at a refused chunk");
  }
}

int main(int argc, char** argv) {
  if (argc > 1) return planted(argv[1]);
  test_host_copies();
  test_scatter();
  test_prepare_wave();
  test_beam_to_rays();
  test_diffraction_f64();
  test_fallback_f32();
  test_fresnel_and_undulator();
  std::printf("sanitizer driver: %d checks, %d failed\n", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
