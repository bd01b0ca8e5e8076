// CPU branches of the auto entry points, run only when no gfx950 device
// exists (the reference's own behaviour without a GPU):
//   cpu_fallback_f32     kirchhoff_auto: the f64 loop of src/gpu/fallback.rs
//                        :29-71 over f32-typed inputs (promoted to f64, skip
//                        below 1e-20 mm, no Kahan);
//   cpu_diffraction_f64  diffraction_integral: the f64 loop of
//                        src/waves/diffraction.rs:67-117 (skip below 1e-30 mm,
//                        complex Kahan sums, the reference's operation order).
// Both run parallel over pixels with std::thread, as rayon's par_iter does
// there.  Built without FP contraction, as the reference's Rust is.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/xrtk.h"

namespace xrtk {

int cpu_fallback_f32(const XrtkGpuRay* rays, size_t nr, const XrtkGpuPixel* px, size_t np,
                     double* out4) {
  if (np == 0) return XRTK_OK;
  const double PI = 3.14159265358979323846;
  std::atomic<size_t> next{0};
  auto worker = [&] {
    for (;;) {
      const size_t p0 = next.fetch_add(64);
      if (p0 >= np) break;
      const size_t p1 = std::min(np, p0 + 64);
      for (size_t p = p0; p < p1; ++p) {
        double es_r = 0, es_i = 0, ep_r = 0, ep_i = 0;
        for (size_t j = 0; j < nr; ++j) {
          const XrtkGpuRay& r = rays[j];
          const double dx = (double)px[p].x - (double)r.x;
          const double dy = (double)px[p].y - (double)r.y;
          const double dz = (double)px[p].z - (double)r.z;
          const double path = std::sqrt(dx * dx + dy * dy + dz * dz);
          if (path < 1e-20) continue;
          const double inv_path = 1.0 / path;
          const double ns = ((double)r.nx * dx + (double)r.ny * dy + (double)r.nz * dz) * inv_path;
          const double k = (double)r.energy / XRTK_CHBAR * 1e7;
          const double phase = k * path;
          const double sp = std::sin(phase), cp = std::cos(phase);
          const double obl = (double)r.nl + ns;
          const double amp = k / (4.0 * PI) * obl * inv_path;
          const double ur = -amp * sp, ui = amp * cp;
          es_r += (double)r.es_re * ur - (double)r.es_im * ui;
          es_i += (double)r.es_re * ui + (double)r.es_im * ur;
          ep_r += (double)r.ep_re * ur - (double)r.ep_im * ui;
          ep_i += (double)r.ep_re * ui + (double)r.ep_im * ur;
        }
        out4[4 * p + 0] = es_r; out4[4 * p + 1] = es_i;
        out4[4 * p + 2] = ep_r; out4[4 * p + 3] = ep_i;
      }
    }
  };
  unsigned nt = std::max(1u, std::thread::hardware_concurrency());
  nt = (unsigned)std::min<size_t>(nt, (np + 63) / 64);
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(worker);
  worker();
  for (auto& t : th) t.join();
  return XRTK_OK;
}

namespace {
// Runs body(p) for every pixel p on all hardware threads, 64 pixels at a time.
template <class Body>
void over_pixels(size_t np, Body body) {
  std::atomic<size_t> next{0};
  auto worker = [&] {
    for (;;) {
      const size_t p0 = next.fetch_add(64);
      if (p0 >= np) break;
      const size_t p1 = std::min(np, p0 + 64);
      for (size_t p = p0; p < p1; ++p) body(p);
    }
  };
  unsigned nt = std::max(1u, std::thread::hardware_concurrency());
  nt = (unsigned)std::min<size_t>(nt, (np + 63) / 64);
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(worker);
  worker();
  for (auto& t : th) t.join();
}
}  // namespace

int cpu_diffraction_f64(const XrtkRaysF64* r, size_t nr, const double* px, const double* py,
                        const double* pz, size_t np, double* out4) {
  if (np == 0) return XRTK_OK;
  const double PI = 3.14159265358979323846;
  over_pixels(np, [&](size_t p) {
    // sums and Kahan compensations of Es and Ep (re, im)
    double s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    for (size_t j = 0; j < nr; ++j) {
      const double dx = px[p] - r->x[j];
      const double dy = py[p] - r->y[j];
      const double dz = pz[p] - r->z[j];
      const double path = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (path < 1e-30) continue;                                    // diffraction.rs:81
      const double ns = (r->nx[j] * dx + r->ny[j] * dy + r->nz[j] * dz) / path;
      const double k = r->energy[j] / XRTK_CHBAR * 1e7;
      const double phase = k * path;
      const double sp = std::sin(phase), cp = std::cos(phase);
      const double obl = r->nl[j] + ns;
      const double amp = k / (4.0 * PI) * obl / path;
      // u = i * amp * e^{i phase}                                   (:98)
      const double ur = -(amp * sp), ui = amp * cp;
      const double t[4] = {r->es_re[j] * ur - r->es_im[j] * ui, r->es_re[j] * ui + r->es_im[j] * ur,
                           r->ep_re[j] * ur - r->ep_im[j] * ui, r->ep_re[j] * ui + r->ep_im[j] * ur};
      for (int q = 0; q < 4; ++q) {                                  // Kahan, :100-110
        const double y = t[q] - c[q];
        const double n = s[q] + y;
        c[q] = (n - s[q]) - y;
        s[q] = n;
      }
    }
    for (int q = 0; q < 4; ++q) out4[4 * p + q] = s[q];
  });
  return XRTK_OK;
}

}  // namespace xrtk
