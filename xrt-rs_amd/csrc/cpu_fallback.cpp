// CPU branch of kirchhoff_auto: the f64 loop of src/gpu/fallback.rs:29-71
// over f32-typed inputs (promoted to f64, skip below 1e-20 mm, no Kahan),
// parallel over pixels with std::thread as rayon's par_iter is there.  It runs
// only when no gfx950 device exists.
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

}  // namespace xrtk
