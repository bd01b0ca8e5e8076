// extern "C" entry points of libxrtk (declared in include/xrtk.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "runtime.h"

using namespace xrtk;

namespace xrtk {
int cpu_fallback_f32(const XrtkGpuRay* rays, size_t nr, const XrtkGpuPixel* px, size_t np,
                     double* out4);
}

namespace {

bool env_flag(const char* name) {
  const char* v = std::getenv(name);
  return v && *v && std::strcmp(v, "0") != 0;
}

int create_ctx(int device, xrtk_ctx** out) {
  *out = nullptr;
  if (env_flag("XRTK_FORCE_NO_DEVICE")) return XRTK_ENODEV;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    return XRTK_ENODEV;
  }
  if (device < 0 || device >= count) return XRTK_ENODEV;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return XRTK_ENODEV;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return XRTK_ENODEV;
  if (hipSetDevice(device) != hipSuccess) return XRTK_ENODEV;
  std::unique_ptr<xrtk_ctx> c(new (std::nothrow) xrtk_ctx());
  if (!c) return XRTK_ENOMEM;
  c->device = device;
  c->name = prop.name;
  c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return XRTK_ELAUNCH;
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return XRTK_ELAUNCH;
  *out = c.release();
  return XRTK_OK;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

std::atomic<int> g_auto_backend{-1};

void record_timing(xrtk_ctx* ctx) {
  float a = 0, b = 0;
  if (hipEventElapsedTime(&a, ctx->ev[1], ctx->ev[2]) != hipSuccess) a = 0;
  if (hipEventElapsedTime(&b, ctx->ev[0], ctx->ev[3]) != hipSuccess) b = 0;
  (void)hipGetLastError();
  ctx->last_pair_ms = a;
  ctx->last_solve_ms = b;
}

void record_plan(xrtk_ctx* ctx, const Plan& p, int backend) {
  XrtkPlan& q = ctx->last_plan;
  q.mode = (int32_t)p.mode;
  q.pixels_per_lane = p.cfg.P;
  q.waves_per_group = p.cfg.W;
  q.tile_rays = p.cfg.tile;
  q.split = p.cfg.split;
  q.ray_src = p.cfg.ray_src;
  q.persistent = p.cfg.persistent;
  q.n_items = (int32_t)p.n_items;
  q.checked = p.checked;
  q.backend = backend;
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Groups ray indices by energy value (monochromatic sub-sets).  Returns the
// groups in order of first appearance.
template <class GetE>
std::vector<std::vector<uint32_t>> group_by_energy(size_t nr, GetE get) {
  std::vector<std::vector<uint32_t>> groups;
  std::map<double, size_t> where;
  for (size_t i = 0; i < nr; ++i) {
    const double e = get(i);
    auto it = where.find(e);
    if (it == where.end()) {
      where.emplace(e, groups.size());
      groups.emplace_back();
      groups.back().push_back((uint32_t)i);
      if (groups.size() > 64) return groups;
    } else {
      groups[it->second].push_back((uint32_t)i);
    }
  }
  return groups;
}

int mode_from(const XrtkOpts* o) {
  if (!o) return MODE_F32R;
  return o->mode == XRTK_MODE_F64 ? MODE_F64 : MODE_F32R;
}

// Solves instances whose rays already sit in ctx->rays64 and pixels in
// ctx->pix64, then copies the result to out4.
int finish_solve(xrtk_ctx* ctx, uint32_t mode, const XrtkOpts* opts,
                 std::vector<InstSpec>& specs, size_t nr_total, size_t np, double* out4) {
  LaunchCfg cfg = choose_cfg(ctx, mode, nr_total, np, opts);
  Plan plan;
  const int split = opts && opts->split > 0 ? opts->split : 0;
  int rc = run_instances(ctx, mode, cfg, specs, ctx->work, ctx->meta, true, &plan, split, 0,
                         false);
  if (rc != XRTK_OK) return rc;
  hipError_t e = hipMemcpyAsync(out4, ctx->outbuf.p, sizeof(double) * 4 * np,
                                hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return map_err(e);
  record_timing(ctx);
  // report whether any instance took the near-singular path
  plan.checked = 0;
  record_plan(ctx, plan, 1);
  return XRTK_OK;
}

}  // namespace

extern "C" {

void xrtk_opts_default(XrtkOpts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->mode = XRTK_MODE_F32R;
  o->persistent = -1;
}

xrtk_ctx* xrtk_ctx_shared(void) {
  static std::once_flag once;
  static xrtk_ctx* shared = nullptr;
  std::call_once(once, [] {
    int dev = 0;
    if (const char* d = std::getenv("XRTK_DEVICE")) dev = std::atoi(d);
    xrtk_ctx* c = nullptr;
    if (create_ctx(dev, &c) == XRTK_OK) {
      c->shared = true;
      shared = c;
    }
  });
  return shared;
}

int xrtk_ctx_create(int device, xrtk_ctx** out) {
  if (!out) return XRTK_EINVAL;
  try {
    return create_ctx(device, out);
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

void xrtk_ctx_destroy(xrtk_ctx* ctx) {
  if (!ctx || ctx->shared) return;
  DeviceGuard g(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto* ge : ctx->graphs) {
    if (ge->exec) (void)hipGraphExecDestroy(ge->exec);
    ge->work.release(); ge->meta.release(); ge->pix.release(); ge->outd.release();
    delete ge;
  }
  ctx->stage.release(); ctx->rays64.release(); ctx->pix64.release();
  ctx->work.release(); ctx->meta.release(); ctx->outbuf.release();
  for (auto& kv : ctx->raysets) kv.second.rays.release();
  for (auto& e : ctx->ev) if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int xrtk_ctx_device_name(const xrtk_ctx* ctx, char* buf, size_t len) {
  if (!ctx || !buf || !len) return XRTK_EINVAL;
  std::snprintf(buf, len, "%s", ctx->name.c_str());
  return XRTK_OK;
}

int xrtk_kirchhoff_gpu_f32(xrtk_ctx* ctx, const XrtkGpuRay* rays, size_t n_rays,
                           const XrtkGpuPixel* pixels, size_t n_pixels, double* out4,
                           const XrtkOpts* opts) {
  if (!ctx) return XRTK_ENODEV;
  if ((n_rays && !rays) || (n_pixels && (!pixels || !out4))) return XRTK_EINVAL;
  if (n_rays > 0xffffffffull || n_pixels > 0x7fffffffull) return XRTK_EINVAL;
  if (n_pixels == 0) return XRTK_OK;
  try {
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    uint32_t mode = mode_from(opts);
    const uint32_t nr = (uint32_t)n_rays, np = (uint32_t)n_pixels;
    // monochromatic groups for the reduced kernel
    std::vector<std::vector<uint32_t>> groups;
    if (mode == MODE_F32R && nr) {
      groups = group_by_energy(nr, [&](size_t i) { return (double)rays[i].energy; });
      if (groups.size() > 64) mode = MODE_F64;  // broadband: plain FP64 path
    }
    std::vector<XrtkGpuRay> perm;
    const XrtkGpuRay* src = rays;
    if (mode == MODE_F32R && groups.size() > 1) {
      perm.reserve(nr);
      for (auto& gr : groups)
        for (uint32_t i : gr) perm.push_back(rays[i]);
      src = perm.data();
    }
    const size_t rb = align256(sizeof(XrtkGpuRay) * nr), pb = sizeof(XrtkGpuPixel) * np;
    hipError_t e = ctx->stage.ensure(rb + pb);
    if (e == hipSuccess) e = ctx->rays64.ensure(sizeof(RayRec64) * std::max<uint32_t>(nr, 1));
    if (e == hipSuccess) e = ctx->pix64.ensure(sizeof(PixF64) * np);
    if (e == hipSuccess) e = ctx->outbuf.ensure(sizeof(double) * 4 * np);
    if (e != hipSuccess) return XRTK_ENOMEM;
    char* S = ctx->stage.as<char>();
    if (nr) e = hipMemcpyAsync(S, src, sizeof(XrtkGpuRay) * nr, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(S + rb, pixels, pb, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && nr)
      e = launch_rays_from_f32aos(reinterpret_cast<const float*>(S), ctx->rays64.as<RayRec64>(),
                                  nr, ctx->stream);
    if (e == hipSuccess) e = launch_pix_from_f32aos(S + rb, ctx->pix64.as<PixF64>(), np, ctx->stream);
    if (e != hipSuccess) return map_err(e);
    std::vector<InstSpec> specs;
    const double skip = 1e-20;  // kirchhoff.wgsl:85, fallback.rs:48
    if (mode == MODE_F32R && nr) {
      uint32_t off = 0;
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        InstSpec s;
        s.rays = ctx->rays64.as<RayRec64>() + off;
        s.nr = (uint32_t)groups[gi].size();
        s.pix = ctx->pix64.as<PixF64>();
        s.np = np;
        s.energy = (double)src[off].energy;
        s.skip = skip;
        s.out = ctx->outbuf.as<double>();
        s.accumulate = gi > 0;
        specs.push_back(s);
        off += s.nr;
      }
    } else {
      InstSpec s;
      s.rays = ctx->rays64.as<RayRec64>();
      s.nr = nr;
      s.pix = ctx->pix64.as<PixF64>();
      s.np = np;
      s.energy = 0.0;  // per-ray energies
      s.skip = skip;
      s.out = ctx->outbuf.as<double>();
      specs.push_back(s);
    }
    return finish_solve(ctx, mode, opts, specs, nr, np, out4);
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

int xrtk_kirchhoff_auto_f32(const XrtkGpuRay* rays, size_t n_rays, const XrtkGpuPixel* pixels,
                            size_t n_pixels, double* out4) {
  if ((n_rays && !rays) || (n_pixels && (!pixels || !out4))) return XRTK_EINVAL;
  xrtk_ctx* ctx = xrtk_ctx_shared();
  if (ctx) {
    const int rc = xrtk_kirchhoff_gpu_f32(ctx, rays, n_rays, pixels, n_pixels, out4, nullptr);
    g_auto_backend.store(1);
    return rc;
  }
  if (env_flag("XRTK_REQUIRE_GPU")) return XRTK_ENODEV;
  g_auto_backend.store(0);
  return cpu_fallback_f32(rays, n_rays, pixels, n_pixels, out4);
}

int xrtk_auto_last_backend(void) { return g_auto_backend.load(); }

static int upload_cols(xrtk_ctx* ctx, const XrtkRaysF64* r, const double* path_up,
                       const std::vector<uint32_t>* perm, uint32_t nr, RayRec64* dst) {
  if (!nr) return XRTK_OK;
  const double* cols[13] = {r->x,     r->y,     r->z,     r->nx,    r->ny,
                            r->nz,    r->nl,    r->es_re, r->es_im, r->ep_re,
                            r->ep_im, r->energy, path_up};
  for (int c = 0; c < 12; ++c)
    if (!cols[c]) return XRTK_EINVAL;
  const size_t cb = sizeof(double) * nr;
  hipError_t e = ctx->stage.ensure(13 * cb);
  if (e != hipSuccess) return XRTK_ENOMEM;
  char* S = ctx->stage.as<char>();
  std::vector<double> tmp;
  for (int c = 0; c < 13; ++c) {
    double* d = reinterpret_cast<double*>(S + c * cb);
    if (!cols[c]) {
      e = launch_zero_f64(d, nr, ctx->stream);
    } else if (perm) {
      tmp.resize(nr);
      for (uint32_t i = 0; i < nr; ++i) tmp[i] = cols[c][(*perm)[i]];
      e = hipMemcpyAsync(d, tmp.data(), cb, hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // tmp is reused
    } else {
      e = hipMemcpyAsync(d, cols[c], cb, hipMemcpyHostToDevice, ctx->stream);
    }
    if (e != hipSuccess) return map_err(e);
  }
  e = launch_rays_from_cols(reinterpret_cast<const double*>(S), dst, nr, ctx->stream);
  return map_err(e);
}

static int diffraction_common(xrtk_ctx* ctx, const XrtkRaysF64* rays, size_t n_rays,
                              const double* px, const double* py, const double* pz,
                              const XrtkWaveParams* grid, size_t n_pixels, double* out4,
                              const XrtkOpts* opts) {
  if (!ctx) return XRTK_ENODEV;
  if ((n_rays && !rays) || (n_pixels && !out4)) return XRTK_EINVAL;
  if (!grid && n_pixels && (!px || !py || !pz)) return XRTK_EINVAL;
  if (n_rays > 0xffffffffull || n_pixels > 0x7fffffffull) return XRTK_EINVAL;
  if (n_pixels == 0) return XRTK_OK;
  try {
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    uint32_t mode = mode_from(opts);
    const uint32_t nr = (uint32_t)n_rays, np = (uint32_t)n_pixels;
    std::vector<std::vector<uint32_t>> groups;
    std::vector<uint32_t> perm;
    if (mode == MODE_F32R && nr) {
      if (!rays->energy) return XRTK_EINVAL;
      groups = group_by_energy(nr, [&](size_t i) { return rays->energy[i]; });
      if (groups.size() > 64) mode = MODE_F64;
      else if (groups.size() > 1)
        for (auto& gr : groups) perm.insert(perm.end(), gr.begin(), gr.end());
    }
    // stage: 13 ray columns, then 3 pixel columns (sized once: a reallocation
    // would free memory that queued kernels still read)
    const size_t rb = align256(13 * sizeof(double) * (size_t)nr);
    hipError_t e = ctx->stage.ensure(rb + 3 * sizeof(double) * np);
    if (e == hipSuccess) e = ctx->rays64.ensure(sizeof(RayRec64) * std::max<uint32_t>(nr, 1));
    if (e == hipSuccess) e = ctx->pix64.ensure(sizeof(PixF64) * np);
    if (e == hipSuccess) e = ctx->outbuf.ensure(sizeof(double) * 4 * np);
    if (e != hipSuccess) return XRTK_ENOMEM;
    int rc = upload_cols(ctx, rays, nullptr, perm.empty() ? nullptr : &perm, nr,
                         ctx->rays64.as<RayRec64>());
    if (rc != XRTK_OK) return rc;
    if (grid) {
      e = launch_pix_grid(grid->center, grid->dx, grid->dz, grid->nx, grid->nz,
                          ctx->pix64.as<PixF64>(), ctx->stream);
    } else {
      char* S = ctx->stage.as<char>() + rb;
      const size_t cb = sizeof(double) * np;
      e = hipMemcpyAsync(S, px, cb, hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess) e = hipMemcpyAsync(S + cb, py, cb, hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(S + 2 * cb, pz, cb, hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess)
        e = launch_pix_from_cols(reinterpret_cast<const double*>(S), ctx->pix64.as<PixF64>(), np,
                                 ctx->stream);
    }
    if (e != hipSuccess) return map_err(e);
    std::vector<InstSpec> specs;
    const double skip = 1e-30;  // diffraction.rs:81
    if (mode == MODE_F32R && nr) {
      uint32_t off = 0;
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        InstSpec s;
        s.rays = ctx->rays64.as<RayRec64>() + off;
        s.nr = (uint32_t)groups[gi].size();
        s.pix = ctx->pix64.as<PixF64>();
        s.np = np;
        s.energy = rays->energy[groups[gi][0]];
        s.skip = skip;
        s.out = ctx->outbuf.as<double>();
        s.accumulate = gi > 0;
        specs.push_back(s);
        off += s.nr;
      }
    } else {
      InstSpec s;
      s.rays = ctx->rays64.as<RayRec64>();
      s.nr = nr;
      s.pix = ctx->pix64.as<PixF64>();
      s.np = np;
      s.energy = 0.0;
      s.skip = skip;
      s.out = ctx->outbuf.as<double>();
      specs.push_back(s);
    }
    return finish_solve(ctx, mode, opts, specs, nr, np, out4);
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

int xrtk_diffraction_f64(xrtk_ctx* ctx, const XrtkRaysF64* rays, size_t n_rays,
                         const double* pix_x, const double* pix_y, const double* pix_z,
                         size_t n_pixels, double* out4, const XrtkOpts* opts) {
  return diffraction_common(ctx, rays, n_rays, pix_x, pix_y, pix_z, nullptr, n_pixels, out4,
                            opts);
}

int xrtk_diffraction_grid_f64(xrtk_ctx* ctx, const XrtkRaysF64* rays, size_t n_rays,
                              const XrtkWaveParams* screen, double* out4, const XrtkOpts* opts) {
  if (!screen) return XRTK_EINVAL;
  const size_t np = (size_t)screen->nx * screen->nz;
  return diffraction_common(ctx, rays, n_rays, nullptr, nullptr, nullptr, screen, np, out4,
                            opts);
}

int xrtk_prepare_wave(const XrtkWaveParams* w, double* px, double* py, double* pz) {
  if (!w) return XRTK_EINVAL;
  const size_t n = (size_t)w->nx * w->nz;
  if (n && (!px || !py || !pz)) return XRTK_EINVAL;
  // prepare_wave.rs:43-65, same operation order (no contraction: see Makefile)
  size_t k = 0;
  for (uint32_t ix = 0; ix < w->nx; ++ix) {
    const double fx = w->nx > 1 ? (double)ix / (double)(w->nx - 1) - 0.5 : 0.0;
    const double x = w->center[0] + fx * 2.0 * w->dx;
    for (uint32_t iz = 0; iz < w->nz; ++iz) {
      const double fz = w->nz > 1 ? (double)iz / (double)(w->nz - 1) - 0.5 : 0.0;
      px[k] = x;
      py[k] = w->center[1];
      pz[k] = w->center[2] + fz * 2.0 * w->dz;
      ++k;
    }
  }
  return XRTK_OK;
}

int xrtk_beam_to_rays(const double* bx, const double* by, const double* bz, const double* ba,
                      const double* bb, const double* bc, const double* be, const double* es_re,
                      const double* es_im, const double* ep_re, const double* ep_im,
                      size_t n_beam, const double* normals3, const size_t* good, size_t n_good,
                      double* const out_cols[12]) {
  if (!out_cols) return XRTK_EINVAL;
  if (n_good && (!bx || !by || !bz || !ba || !bb || !bc || !be || !normals3 || !good))
    return XRTK_EINVAL;
  for (int c = 0; c < 12; ++c)
    if (n_good && !out_cols[c]) return XRTK_EINVAL;
  const bool amps = es_re && es_im && ep_re && ep_im;
  for (size_t k = 0; k < n_good; ++k) {
    const size_t i = good[k];
    if (i >= n_beam) return XRTK_EINVAL;
    const double* n = normals3 + 3 * i;
    // prepare_wave.rs:86-87: nl = n . (a, b, c)
    const double nl = n[0] * ba[i] + n[1] * bb[i] + n[2] * bc[i];
    out_cols[0][k] = bx[i]; out_cols[1][k] = by[i]; out_cols[2][k] = bz[i];
    out_cols[3][k] = n[0]; out_cols[4][k] = n[1]; out_cols[5][k] = n[2];
    out_cols[6][k] = nl;
    out_cols[7][k] = amps ? es_re[i] : 1.0;  // prepare_wave.rs:89-95
    out_cols[8][k] = amps ? es_im[i] : 0.0;
    out_cols[9][k] = amps ? ep_re[i] : 0.0;
    out_cols[10][k] = amps ? ep_im[i] : 0.0;
    out_cols[11][k] = be[i];
  }
  return XRTK_OK;
}

// ---- resident ray sets --------------------------------------------------
int xrtk_rayset_upload(xrtk_ctx* ctx, const XrtkRaysF64* rays, const double* path_up,
                       size_t n_rays, xrtk_rayset* out) {
  if (!ctx) return XRTK_ENODEV;
  if (!out || (n_rays && !rays) || n_rays > 0xffffffffull) return XRTK_EINVAL;
  try {
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    RaySet rs;
    rs.nr = (uint32_t)n_rays;
    rs.has_path = path_up != nullptr;
    if (n_rays) {
      if (!rays->energy) return XRTK_EINVAL;
      rs.energy = rays->energy[0];
      for (size_t i = 1; i < n_rays; ++i)
        if (rays->energy[i] != rs.energy) { rs.mono = false; break; }
    }
    if (rs.rays.ensure(sizeof(RayRec64) * std::max<size_t>(n_rays, 1)) != hipSuccess)
      return XRTK_ENOMEM;
    int rc = upload_cols(ctx, rays, path_up, nullptr, rs.nr, rs.rays.as<RayRec64>());
    if (rc == XRTK_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = XRTK_ELAUNCH;
    if (rc != XRTK_OK) {
      rs.rays.release();
      return rc;
    }
    const uint64_t h = ctx->next_rs++;
    ctx->raysets.emplace(h, rs);
    *out = h;
    return XRTK_OK;
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

int xrtk_rayset_release(xrtk_ctx* ctx, xrtk_rayset rs) {
  if (!ctx) return XRTK_ENODEV;
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->raysets.find(rs);
  if (it == ctx->raysets.end()) return XRTK_EINVAL;
  DeviceGuard g(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  it->second.rays.release();
  ctx->raysets.erase(it);
  return XRTK_OK;
}

int xrtk_rayset_bytes(xrtk_ctx* ctx, xrtk_rayset rs, size_t* bytes) {
  if (!ctx) return XRTK_ENODEV;
  if (!bytes) return XRTK_EINVAL;
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->raysets.find(rs);
  if (it == ctx->raysets.end()) return XRTK_EINVAL;
  *bytes = it->second.rays.cap;
  return XRTK_OK;
}

int xrtk_last_timing(xrtk_ctx* ctx, double* pair_ms, double* solve_ms) {
  if (!ctx) return XRTK_ENODEV;
  if (pair_ms) *pair_ms = ctx->last_pair_ms;
  if (solve_ms) *solve_ms = ctx->last_solve_ms;
  return XRTK_OK;
}

int xrtk_last_plan(xrtk_ctx* ctx, XrtkPlan* plan) {
  if (!ctx) return XRTK_ENODEV;
  if (!plan) return XRTK_EINVAL;
  *plan = ctx->last_plan;
  return XRTK_OK;
}

double xrtk_now(void) { return now_s(); }

}  // extern "C"
