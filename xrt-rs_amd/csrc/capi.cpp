// extern "C" entry points of libxrtk (declared in include/xrtk.h): context,
// one-shot solves, host helpers, resident ray sets, timing and plan queries.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "runtime.h"

using namespace xrtk;

namespace xrtk {
int cpu_fallback_f32(const XrtkGpuRay* rays, size_t nr, const XrtkGpuPixel* px, size_t np,
                     double* out4);
int cpu_diffraction_f64(const XrtkRaysF64* r, size_t nr, const double* px, const double* py,
                        const double* pz, size_t np, double* out4);
void load_tune_cache(xrtk_ctx* ctx);
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
  c->name = prop.name[0] ? prop.name : "AMD gfx950";
  c->name += std::string(" (") + prop.gcnArchName + ", " +
             std::to_string(prop.multiProcessorCount) + " CUs)";
  c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return XRTK_ELAUNCH;
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return XRTK_ELAUNCH;
  load_tune_cache(c.get());
  c->scratch = noalloc_scratch(device, &c->scratch_bytes);  // device is current here
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

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Groups ray indices by energy value (monochromatic sub-sets) in order of
// first appearance; stops at 65 groups.
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

uint32_t mode_from(const XrtkOpts* o) {
  if (!o) return MODE_F32R;
  return o->mode == XRTK_MODE_F64 ? MODE_F64 : MODE_F32R;
}

// Per-axis half-extent bound of slots around their tile mean, for an
// explicit pixel list laid out in tiles of `tsz` consecutive slots (the tail
// padded with the last pixel).
template <class Get>
void explicit_tile_dq(size_t np, uint32_t tsz, Get get, double dq[3]) {
  dq[0] = dq[1] = dq[2] = 0.0;
  for (size_t t0 = 0; t0 < np; t0 += tsz) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t s = t0; s < t0 + tsz; ++s) {
      double v[3];
      get(std::min(s, np - 1), v);
      for (int i = 0; i < 3; ++i) { lo[i] = std::min(lo[i], v[i]); hi[i] = std::max(hi[i], v[i]); }
    }
    for (int i = 0; i < 3; ++i) dq[i] = std::max(dq[i], hi[i] - lo[i]);
  }
}

Box grid_box(const XrtkWaveParams& w) {
  Box b;
  b.add(w.center[0] - std::fabs(w.dx), w.center[1], w.center[2] - std::fabs(w.dz));
  b.add(w.center[0] + std::fabs(w.dx), w.center[1], w.center[2] + std::fabs(w.dz));
  return b;
}

void grid_dq(const XrtkWaveParams& w, int P, double dq[3]) {
  uint32_t th, tw;
  grid_tile(P, w.nx, w.nz, &th, &tw);
  const double px = w.nx > 1 ? 2.0 * std::fabs(w.dx) / (w.nx - 1) : 0.0;
  const double pz = w.nz > 1 ? 2.0 * std::fabs(w.dz) / (w.nz - 1) : 0.0;
  dq[0] = px * th;
  dq[1] = 0.0;
  dq[2] = pz * tw;
}

// A float32 pixel list on a grid: the grid (prepare_wave's parameters), the
// offsets of its rows and columns from it in pixels, and their largest size.
struct ListGrid {
  XrtkWaveParams w;
  std::vector<float> ex, ez;
  double emax = 0.0;
};

// Whether a float32 pixel list lies on an nx x nz grid in the prepare_wave
// order (src/waves/prepare_wave.rs:40-68: entry ix * nz + iz, x a function of
// ix alone, z of iz alone, y the same everywhere), nx, nz >= 2, as prepare_wave
// gives it once rounded to float32.  The grid is the least-squares fit of the
// rows' and the columns' positions; each row's and column's float32 position
// is taken as an offset from it (the rounding, or any spacing that departs
// from it by at most 1/100 of a pixel).
bool fit_list_grid(const XrtkGpuPixel* p, size_t np, ListGrid& g) {
  if (np < 4) return false;
  size_t nz = 1;
  while (nz < np && p[nz].x == p[0].x) ++nz;
  if (nz < 2 || np % nz) return false;
  const size_t nx = np / nz;
  if (nx < 2) return false;
  const float y = p[0].y;
  for (size_t ix = 0; ix < nx; ++ix) {
    const XrtkGpuPixel* r = p + ix * nz;
    for (size_t iz = 0; iz < nz; ++iz)
      if (r[iz].x != r[0].x || r[iz].z != p[iz].z || r[iz].y != y) return false;
  }
  // least-squares line through the rows' (columns') positions: its centre
  // and half span are the grid's
  auto fit = [](size_t n, auto at, double* centre, double* half) {
    const double im = 0.5 * (double)(n - 1);
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) m += at(i);
    m /= (double)n;
    double sxy = 0.0, sxx = 0.0;
    for (size_t i = 0; i < n; ++i) {
      const double d = (double)i - im;
      sxy += d * (at(i) - m);
      sxx += d * d;
    }
    *centre = m;
    *half = sxy / sxx * im;
  };
  fit(nx, [&](size_t i) { return (double)p[i * nz].x; }, &g.w.center[0], &g.w.dx);
  fit(nz, [&](size_t i) { return (double)p[i].z; }, &g.w.center[2], &g.w.dz);
  g.w.center[1] = y;
  g.w.nx = (uint32_t)nx;
  g.w.nz = (uint32_t)nz;
  if (g.w.dx == 0.0 || g.w.dz == 0.0) return false;
  InstSpec t;  // the grid exactly as the kernels form it
  spec_set_grid(t, g.w, 1, 1);
  g.emax = 0.0;
  g.ex.resize(nx);
  g.ez.resize(nz);
  for (size_t ix = 0; ix < nx; ++ix) {
    const double e = ((double)p[ix * nz].x - (t.gx0 + (double)ix * t.gsx)) / t.gsx;
    g.emax = std::max(g.emax, std::fabs(e));
    g.ex[ix] = (float)e;
  }
  for (size_t iz = 0; iz < nz; ++iz) {
    const double e = ((double)p[iz].z - (t.gz0 + (double)iz * t.gsz)) / t.gsz;
    g.emax = std::max(g.emax, std::fabs(e));
    g.ez[iz] = (float)e;
  }
  return g.emax <= 0.01;
}

void record_timing(xrtk_ctx* ctx) {
  float a = 0, b = 0;
  if (hipEventElapsedTime(&a, ctx->ev[1], ctx->ev[2]) != hipSuccess) a = 0;
  if (hipEventElapsedTime(&b, ctx->ev[0], ctx->ev[3]) != hipSuccess) b = 0;
  (void)hipGetLastError();
  ctx->last_pair_ms = a;
  ctx->last_solve_ms = b;
}

static_assert(XRTK_VAR_CHECK == INST_F_CHECK && XRTK_VAR_TILE == INST_F_TILE &&
                  XRTK_VAR_PATHUP == INST_F_PATHUP && XRTK_VAR_UNIFORM_N == INST_F_UNIFORM_N &&
                  XRTK_VAR_NOFRACT == INST_F_NOFRACT && XRTK_VAR_LIN2 == INST_F_LIN2 &&
                  XRTK_VAR_PAX == INST_F_PAX && XRTK_VAR_SEPAMP == INST_F_SEPAMP &&
                  XRTK_VAR_LINE == INST_F_LINE && XRTK_VAR_LINEDEV == INST_F_LINEDEV &&
                  XRTK_VAR_EPZERO == INST_F_EPZERO && XRTK_VAR_MFMA == INST_F_MFMA,
              "variant flags of the ABI match the device planner's");

void record_plan(xrtk_ctx* ctx, const Plan& p, int backend) {
  XrtkPlan& q = ctx->last_plan;
  const bool all64 = p.n_inst && p.n_f64 == p.n_inst;
  const LaunchCfg& c = all64 ? p.cfg64 : p.cfg32;
  q.mode = all64 ? XRTK_MODE_F64 : XRTK_MODE_F32R;
  q.pixels_per_lane = c.P;
  q.waves_per_group = c.W;
  q.tile_rays = c.tile;
  q.split = c.split;
  q.ray_src = c.ray_src;
  q.persistent = c.persistent;
  q.n_items = (int32_t)(p.n_items32 + p.n_items64 + p.n_items_mfma);
  q.n_fp64 = (int32_t)p.n_f64;
  q.backend = backend;
  q.inst_flags = 0;
  q.n_mfma = 0;
  for (uint32_t f : p.flags) q.n_mfma += (f & INST_F_MFMA) ? 1 : 0;
}

// Pixel source of a one-shot solve.
struct PixSrc {
  const XrtkGpuPixel* f32 = nullptr;        // GpuPixel array
  const double *x = nullptr, *y = nullptr, *z = nullptr;  // f64 columns
  const XrtkWaveParams* grid = nullptr;     // prepare_wave grid
  size_t np = 0;
  void get(size_t i, double v[3]) const {
    if (f32) { v[0] = f32[i].x; v[1] = f32[i].y; v[2] = f32[i].z; }
    else { v[0] = x[i]; v[1] = y[i]; v[2] = z[i]; }
  }
};

// Ray source of a one-shot solve: f32 GpuRay records or f64 columns.
struct RaySrc {
  const XrtkGpuRay* f32 = nullptr;
  const XrtkRaysF64* f64 = nullptr;
  size_t nr = 0;
  double energy(size_t i) const { return f32 ? (double)f32[i].energy : f64->energy[i]; }
  void pos(size_t i, double v[3]) const {
    if (f32) { v[0] = f32[i].x; v[1] = f32[i].y; v[2] = f32[i].z; }
    else { v[0] = f64->x[i]; v[1] = f64->y[i]; v[2] = f64->z[i]; }
  }
};

// Whether every ray has the first ray's energy, and (if so) the box of the
// ray positions: one pass over the records, split over a few host threads.
bool one_energy_box(const RaySrc& rs, Box& box) {
  const size_t n = rs.nr;
  if (!n) return false;
  const double e0 = rs.energy(0);
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t T = std::min<size_t>(std::min(8u, hw), std::max<size_t>(1, n / 65536));
  std::vector<Box> part(T);
  std::vector<char> same(T, 1);
  auto run = [&](size_t t) {
    const size_t i0 = n * t / T, i1 = n * (t + 1) / T;
    Box b;
    for (size_t i = i0; i < i1; ++i) {
      if (rs.energy(i) != e0) { same[t] = 0; return; }
      double v[3];
      rs.pos(i, v);
      b.add(v[0], v[1], v[2]);
    }
    part[t] = b;
  };
  std::vector<std::thread> th;
  for (size_t t = 1; t < T; ++t) th.emplace_back(run, t);
  run(0);
  for (auto& x : th) x.join();
  for (size_t t = 0; t < T; ++t) {
    if (!same[t]) return false;
    for (int k = 0; k < 3; ++k) {
      box.lo[k] = std::min(box.lo[k], part[t].lo[k]);
      box.hi[k] = std::max(box.hi[k], part[t].hi[k]);
    }
  }
  return true;
}

// Uploads 13 f64 columns (path_up may be null: zeros) into `stage` at
// offset 0, optionally permuted, and converts them into dst.  pin: a
// page-locked block of 13 columns the copies pass through (or null).
int upload_cols(xrtk_ctx* ctx, DBuf& stage, const XrtkRaysF64* r, const double* path_up,
                const std::vector<uint32_t>* perm, uint32_t nr, RayRec64* dst,
                char* pin = nullptr) {
  if (!nr) return XRTK_OK;
  const double* cols[13] = {r->x,     r->y,     r->z,     r->nx,    r->ny,
                            r->nz,    r->nl,    r->es_re, r->es_im, r->ep_re,
                            r->ep_im, r->energy, path_up};
  for (int c = 0; c < 12; ++c)
    if (!cols[c]) return XRTK_EINVAL;
  const size_t cb = sizeof(double) * nr;
  if (stage.cap < 13 * cb) return XRTK_EINVAL;  // caller sizes the stage
  char* S = stage.as<char>();
  std::vector<double> tmp;
  hipError_t e = hipSuccess;
  for (int c = 0; c < 13 && e == hipSuccess; ++c) {
    double* d = reinterpret_cast<double*>(S + c * cb);
    if (!cols[c]) {
      e = launch_zero_f64(d, nr, ctx->stream);
    } else if (perm) {
      double* t = pin ? reinterpret_cast<double*>(pin + c * cb) : nullptr;
      if (!t) {
        tmp.resize(nr);
        t = tmp.data();
      }
      for (uint32_t i = 0; i < nr; ++i) t[i] = cols[c][(*perm)[i]];
      e = hipMemcpyAsync(d, t, cb, hipMemcpyHostToDevice, ctx->stream);
      if (!pin && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // tmp is reused
    } else {
      e = copy_h2d(ctx->stream, pin ? pin + c * cb : nullptr, d, cols[c], cb);
    }
  }
  if (e == hipSuccess) e = launch_rays_from_cols(reinterpret_cast<const double*>(S), dst, nr,
                                                 ctx->stream);
  return map_err(e);
}

// The one-shot pipeline: group, route, upload, solve, copy back.
int solve_oneshot(xrtk_ctx* ctx, Workspace& ws, const RaySrc& rs, const PixSrc& ps, double skip,
                  double* out4, const XrtkOpts* opts) {
  Range whole("xrtk.solve");
  Phases ph;
  const uint32_t want = mode_from(opts);
  const uint32_t nr = (uint32_t)rs.nr, np = (uint32_t)ps.np;
  // Staging sizes (the pixel part bounded before the list is examined: a
  // float32 list on a grid adds one offset per row and per column).
  const size_t ray_stage = rs.f32 ? align256(sizeof(XrtkGpuRay) * nr)
                                  : align256(13 * sizeof(double) * (size_t)nr);
  const size_t pix_bound = ps.grid  ? 0
                           : ps.f32 ? align256(sizeof(XrtkGpuPixel) * np) +
                                          sizeof(float) * ((size_t)np + 1)
                                    : 3 * sizeof(double) * (size_t)np;
  const size_t ob = sizeof(double) * 4 * np;
  // the page-locked block (not for the allocation-free workspace; a failed
  // allocation falls back to copies from the caller's memory), with 256
  // bytes past the staging for the variant flags
  char* pin = nullptr;
  size_t pin_tail = 0;
  if (!ws.stage.arena) {
    pin_tail = align256(std::max(ray_stage + pix_bound, ob));
    if (ws.hpin.ensure(pin_tail + 256) == hipSuccess) pin = static_cast<char*>(ws.hpin.p);
  }
  // float32 ray records in the given order go up first, through the block,
  // while the host plans; a solve that needs them regrouped by energy sends
  // them again in that order.  Every return waits for the stream, so the
  // block is free when the next solve fills it.
  bool rays_sent = false;
  struct WaitOnExit {
    hipStream_t st;
    bool armed = false;
    ~WaitOnExit() {
      if (armed) (void)hipStreamSynchronize(st);
    }
  } wait_on_exit{ctx->stream};
  if (pin && rs.f32 && nr) {
    ph.next("xrtk.upload");
    hipError_t e = ws.stage.ensure(ray_stage + pix_bound + 256);
    if (e == hipSuccess) e = ws.rays64.ensure(sizeof(RayRec64) * nr);
    if (e != hipSuccess) return XRTK_ENOMEM;
    wait_on_exit.armed = true;
    e = copy_h2d(ctx->stream, pin, ws.stage.p, rs.f32, sizeof(XrtkGpuRay) * nr);
    if (e == hipSuccess)
      e = launch_rays_from_f32aos(ws.stage.as<float>(), ws.rays64.as<RayRec64>(), nr,
                                  ctx->stream);
    if (e != hipSuccess) return map_err(e);
    rays_sent = true;
  }
  ph.next("xrtk.plan");
  // A float32 list on a grid is laid out as that grid, so that it may run on
  // the matrix cores (XRTK_ENGINE_VALU keeps the list layout).
  ListGrid lg;
  const bool list_grid = ps.f32 && want == MODE_F32R && nr &&
                         !(opts && opts->engine == XRTK_ENGINE_VALU) &&
                         fit_list_grid(ps.f32, np, lg);
  const XrtkWaveParams* grid = list_grid ? &lg.w : ps.grid;
  // the F32R shape fixes the slot layout even when only the FP64 kernel runs
  const bool f32_pixels = ps.f32 != nullptr && !list_grid;
  LaunchCfg c32 =
      choose_cfg(ctx, MODE_F32R, nr, np, want == MODE_F32R ? opts : nullptr, f32_pixels);
  LaunchCfg c64 = cfg64_for(c32);
  if (want == MODE_F64) c64 = choose_cfg(ctx, MODE_F64, nr, np, opts);
  if (!pair_cfg_supported(MODE_F32R, c32) || !pair_cfg_supported(MODE_F64, c64))
    return XRTK_EINVAL;
  // monochromatic groups for the reduced kernel (one energy: one group in
  // the given order, its box from the same pass)
  std::vector<std::vector<uint32_t>> groups;
  bool broadband = false;
  Box mono_box;
  const bool mono = nr && want == MODE_F32R && one_energy_box(rs, mono_box);
  if (nr && !mono) {
    groups = group_by_energy(nr, [&](size_t i) { return rs.energy(i); });
    broadband = groups.size() > 64;
  }
  // pixel slots and the routing bound inputs
  Box pbox;
  double dq[3] = {0, 0, 0};
  uint32_t npad;
  // (a list on a grid: its own box, and tiles wider by its offsets)
  auto list_box = [&]() {
    for (size_t i = 0; i < np; ++i) {
      double v[3];
      ps.get(i, v);
      pbox.add(v[0], v[1], v[2]);
    }
  };
  auto tile_dq = [&](int P, double* d) {
    grid_dq(*grid, P, d);
    if (list_grid) {
      d[0] += 2.0 * lg.emax * 2.0 * std::fabs(grid->dx) / (grid->nx - 1);
      d[2] += 2.0 * lg.emax * 2.0 * std::fabs(grid->dz) / (grid->nz - 1);
    }
  };
  if (grid) {
    npad = grid_slots(grid->nx, grid->nz, c32, c64);
    if (list_grid) list_box(); else pbox = grid_box(*grid);
    tile_dq(c32.P, dq);
  } else {
    npad = padded_slots(np, c32, c64);
    for (size_t i = 0; i < np; ++i) {
      double v[3];
      ps.get(i, v);
      pbox.add(v[0], v[1], v[2]);
    }
    if (want == MODE_F32R && !broadband)
      explicit_tile_dq(np, 64u * c32.P, [&](size_t i, double v[3]) { ps.get(i, v); }, dq);
  }
  // instances: one per energy group (F32R), routed by the residual bound;
  // broadband or F64 requests: one FP64 instance with per-ray energies
  struct G { uint32_t off, n, mode; double energy; };
  std::vector<G> gs;
  std::vector<Box> boxes;  // each group's ray box
  std::vector<uint32_t> perm;
  const bool reduced = nr && want == MODE_F32R && !broadband;
  if (mono) {
    const double e = rs.energy(0);
    gs.push_back(G{0, nr, route_mode(MODE_F32R, mono_box, pbox, dq, e), e});
    boxes.push_back(mono_box);
  } else if (reduced) {
    uint32_t off = 0;
    for (auto& gr : groups) {
      Box rb;
      for (uint32_t i : gr) {
        double v[3];
        rs.pos(i, v);
        rb.add(v[0], v[1], v[2]);
      }
      const double e = rs.energy(gr[0]);
      gs.push_back(G{off, (uint32_t)gr.size(), route_mode(MODE_F32R, rb, pbox, dq, e), e});
      boxes.push_back(rb);
      off += (uint32_t)gr.size();
    }
    if (groups.size() > 1)
      for (auto& gr : groups) perm.insert(perm.end(), gr.begin(), gr.end());
  } else if (nr) {
    gs.push_back(G{0, nr, MODE_F64, 0.0});
  }
  // An untuned grid solve takes smaller tiles while that keeps more
  // groups off the FP64 kernel: the routing bound grows with the tile's
  // extent (near-field screens at P = 16 can exceed it where P = 4 does
  // not).
  if (reduced && grid && !cfg_shape_given(ctx, MODE_F32R, nr, np, opts)) {
    auto n64 = [&](int P, double* d) {
      tile_dq(P, d);
      int n = 0;
      for (size_t g = 0; g < gs.size(); ++g)
        n += route_mode(MODE_F32R, boxes[g], pbox, d, gs[g].energy) == MODE_F64;
      return n;
    };
    int cur = 0;
    for (const G& g : gs) cur += g.mode == MODE_F64;
    while (cur > 0 && c32.P > 4) {
      double d2[3];
      const int m = n64(c32.P / 2, d2);
      if (m >= cur) break;
      c32.P /= 2;
      cur = m;
      std::memcpy(dq, d2, sizeof(dq));
      for (size_t g = 0; g < gs.size(); ++g)
        gs[g].mode = route_mode(MODE_F32R, boxes[g], pbox, dq, gs[g].energy);
      c64 = cfg64_for(c32);
      npad = grid_slots(grid->nx, grid->nz, c32, c64);
    }
  }
  // ---- buffers
  ph.next("xrtk.upload");
  const size_t offs = list_grid ? sizeof(float) * (lg.ex.size() + lg.ez.size()) : 0;
  const size_t pix_stage = ps.grid  ? 0
                           : ps.f32 ? sizeof(XrtkGpuPixel) * np + offs
                                    : 3 * sizeof(double) * (size_t)np;
  // (no larger than what the rays' early upload reserved: no reallocation)
  hipError_t e = ws.stage.ensure(ray_stage + pix_stage + 256);
  if (e == hipSuccess) e = ws.rays64.ensure(sizeof(RayRec64) * std::max<uint32_t>(nr, 1));
  if (e == hipSuccess) e = ws.pix64.ensure(sizeof(PixF64) * npad);
  if (e == hipSuccess) e = ws.perm.ensure(sizeof(uint32_t) * npad);
  if (e == hipSuccess) e = ws.outbuf.ensure(sizeof(double) * 4 * np);
  if (e != hipSuccess) return XRTK_ENOMEM;
  char* S = ws.stage.as<char>();
  if (rays_sent && !perm.empty()) {
    // the block is read by the first upload until the stream has run it
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return map_err(e);
    rays_sent = false;
  }
  // ---- rays
  if (rays_sent) {
    // already on the device in the given order
  } else if (rs.f32 && nr) {
    std::vector<XrtkGpuRay> tmp;
    const XrtkGpuRay* src = rs.f32;
    if (!perm.empty()) {
      tmp.reserve(nr);
      for (uint32_t i : perm) tmp.push_back(rs.f32[i]);
      src = tmp.data();
    }
    e = copy_h2d(ctx->stream, pin, S, src, sizeof(XrtkGpuRay) * nr);
    if (e == hipSuccess)
      e = launch_rays_from_f32aos(reinterpret_cast<const float*>(S), ws.rays64.as<RayRec64>(),
                                  nr, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // tmp goes out of scope
    if (e != hipSuccess) return map_err(e);
  } else if (nr) {
    const int rc = upload_cols(ctx, ws.stage, rs.f64, nullptr, perm.empty() ? nullptr : &perm, nr,
                               ws.rays64.as<RayRec64>(), pin);
    if (rc != XRTK_OK) return rc;
  }
  // ---- pixels
  const uint32_t* d_perm = nullptr;
  const float *d_ex = nullptr, *d_ez = nullptr;
  if (list_grid) {
    uint32_t th, tw;
    grid_tile(c32.P, grid->nx, grid->nz, &th, &tw);
    char* L = S + ray_stage;
    float* X = reinterpret_cast<float*>(L + sizeof(XrtkGpuPixel) * np);
    d_ex = X;
    d_ez = X + lg.ex.size();
    e = copy_h2d(ctx->stream, pin ? pin + ray_stage : nullptr, L, ps.f32,
                 sizeof(XrtkGpuPixel) * np);
    if (e == hipSuccess)
      e = hipMemcpyAsync(X, lg.ex.data(), sizeof(float) * lg.ex.size(), hipMemcpyHostToDevice,
                         ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(X + lg.ex.size(), lg.ez.data(), sizeof(float) * lg.ez.size(),
                         hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = launch_pix_grid_list(L, grid->nx, grid->nz, th, tw, npad, ws.pix64.as<PixF64>(),
                               ws.perm.as<uint32_t>(), ctx->stream);
    d_perm = ws.perm.as<uint32_t>();
  } else if (ps.grid) {
    uint32_t th, tw;
    grid_tile(c32.P, ps.grid->nx, ps.grid->nz, &th, &tw);
    e = launch_pix_grid(ps.grid->center, ps.grid->dx, ps.grid->dz, ps.grid->nx, ps.grid->nz, th,
                        tw, npad, ws.pix64.as<PixF64>(), ws.perm.as<uint32_t>(), ctx->stream);
    d_perm = ws.perm.as<uint32_t>();
  } else if (ps.f32) {
    e = copy_h2d(ctx->stream, pin ? pin + ray_stage : nullptr, S + ray_stage, ps.f32,
                 sizeof(XrtkGpuPixel) * np);
    if (e == hipSuccess)
      e = launch_pix_from_f32aos(S + ray_stage, ws.pix64.as<PixF64>(), ws.perm.as<uint32_t>(),
                                 np, npad, (uint32_t)c32.P, ctx->stream);
    d_perm = ws.perm.as<uint32_t>();
  } else {
    double* P3 = reinterpret_cast<double*>(S + ray_stage);
    const size_t cb = sizeof(double) * np;
    char* pp = pin ? pin + ray_stage : nullptr;
    e = copy_h2d(ctx->stream, pp, P3, ps.x, cb);
    if (e == hipSuccess) e = copy_h2d(ctx->stream, pp ? pp + cb : nullptr, P3 + np, ps.y, cb);
    if (e == hipSuccess)
      e = copy_h2d(ctx->stream, pp ? pp + 2 * cb : nullptr, P3 + 2 * np, ps.z, cb);
    if (e == hipSuccess)
      e = launch_pix_from_cols(P3, ws.pix64.as<PixF64>(), ws.perm.as<uint32_t>(), np, npad,
                               (uint32_t)c32.P, ctx->stream);
    d_perm = ws.perm.as<uint32_t>();
  }
  if (e != hipSuccess) return map_err(e);
  // ---- instances
  std::vector<InstSpec> specs;
  for (size_t gi = 0; gi < gs.size(); ++gi) {
    InstSpec s;
    s.mode = gs[gi].mode;
    s.rays = ws.rays64.as<RayRec64>() + gs[gi].off;
    s.nr = gs[gi].n;
    s.pix = ws.pix64.as<PixF64>();
    s.perm = d_perm;
    s.np = np;
    s.npad = npad;
    s.energy = s.mode == MODE_F32R ? gs[gi].energy : 0.0;  // F64: per-ray energies
    s.skip = skip;
    s.out = ws.outbuf.as<double>();
    s.accumulate = gi > 0;
    if (grid && np) {
      uint32_t th, tw;
      grid_tile(c32.P, grid->nx, grid->nz, &th, &tw);
      spec_set_grid(s, *grid, th, tw);
      if (list_grid) {
        s.gex = d_ex;
        s.gez = d_ez;
        s.gemax = lg.emax;
      }
    }
    specs.push_back(s);
  }
  if (specs.empty()) {  // no rays: zeros
    InstSpec s;
    s.mode = want;
    s.pix = ws.pix64.as<PixF64>();
    s.np = np;
    s.npad = npad;
    s.out = ws.outbuf.as<double>();
    specs.push_back(s);
  }
  Plan plan;
  const int split = opts && opts->split > 0 ? opts->split : 0;
  uint32_t force = 0;
  if (const char* f = std::getenv("XRTK_FORCE_FLAGS")) force = (uint32_t)std::strtoul(f, nullptr, 0);
  if (opts && opts->engine == XRTK_ENGINE_VALU) force |= 0x20000u;
  int rc = run_instances(ctx, c32, c64, specs, ws.work, ws.meta, true, &plan, split, force,
                         false, true, nullptr, /*route_sync=*/true);
  if (rc != XRTK_OK) return rc;
  ph.next("xrtk.readback");
  // (the variant flags into the page-locked block's tail, then a large
  // result through the block, which the inputs no longer need once the
  // stream has run this far; copy_d2h waits for the stream)
  uint32_t* fl_pin = pin ? reinterpret_cast<uint32_t*>(pin + pin_tail) : nullptr;
  if (plan.scr0 && fl_pin)
    e = hipMemcpyAsync(fl_pin, &plan.scr0->flags, sizeof(uint32_t), hipMemcpyDeviceToHost,
                       ctx->stream);
  const Segment res{out4, ob};
  if (e == hipSuccess) e = copy_d2h(ctx->stream, pin, ws.outbuf.p, &res, 1);
  if (e != hipSuccess) return map_err(e);
  record_timing(ctx);
  record_plan(ctx, plan, 1);
  if (plan.scr0) {
    uint32_t fl = 0;
    if (fl_pin)
      ctx->last_plan.inst_flags = (int32_t)*fl_pin;
    else if (hipMemcpy(&fl, &plan.scr0->flags, sizeof(fl), hipMemcpyDeviceToHost) == hipSuccess)
      ctx->last_plan.inst_flags = (int32_t)fl;
  }
  return XRTK_OK;
}

}  // namespace

namespace xrtk {

int kirchhoff_f32_ws(xrtk_ctx* ctx, Workspace& ws, const XrtkGpuRay* rays, size_t n_rays,
                     const XrtkGpuPixel* pixels, size_t n_pixels, double* out4,
                     const XrtkOpts* opts) {
  if ((n_rays && !rays) || (n_pixels && (!pixels || !out4))) return XRTK_EINVAL;
  if (n_rays > 0xffffffffull || n_pixels > 0x7fff0000ull) return XRTK_EINVAL;
  if (n_pixels == 0) return XRTK_OK;
  try {
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    RaySrc rs;
    rs.f32 = rays;
    rs.nr = n_rays;
    PixSrc ps;
    ps.f32 = pixels;
    ps.np = n_pixels;
    return solve_oneshot(ctx, ws, rs, ps, 1e-20 /* kirchhoff.wgsl:85, fallback.rs:48 */, out4,
                         opts);
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

}  // namespace xrtk

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
    try {
      if (create_ctx(dev, &c) == XRTK_OK) {
        c->shared = true;
        shared = c;
      }
    } catch (...) {
      shared = nullptr;
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
    ge->work.release(); ge->meta.release(); ge->pix.release(); ge->perm.release();
    ge->outd.release();
    delete ge;
  }
  ctx->ws.release();
  fresnel_release(ctx);
  ctx->undulator.release();
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
  return kirchhoff_f32_ws(ctx, ctx->ws, rays, n_rays, pixels, n_pixels, out4, opts);
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
  try {
    return cpu_fallback_f32(rays, n_rays, pixels, n_pixels, out4);
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

int xrtk_auto_last_backend(void) { return g_auto_backend.load(); }

int xrtk_diffraction_f64(xrtk_ctx* ctx, const XrtkRaysF64* rays, size_t n_rays,
                         const double* pix_x, const double* pix_y, const double* pix_z,
                         size_t n_pixels, double* out4, const XrtkOpts* opts) {
  if (!ctx) return XRTK_ENODEV;
  if ((n_rays && !rays) || (n_pixels && (!out4 || !pix_x || !pix_y || !pix_z)))
    return XRTK_EINVAL;
  if (n_rays && (!rays->x || !rays->y || !rays->z || !rays->energy)) return XRTK_EINVAL;
  if (n_rays > 0xffffffffull || n_pixels > 0x7fff0000ull) return XRTK_EINVAL;
  if (n_pixels == 0) return XRTK_OK;
  try {
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    RaySrc rs;
    rs.f64 = rays;
    rs.nr = n_rays;
    PixSrc ps;
    ps.x = pix_x; ps.y = pix_y; ps.z = pix_z;
    ps.np = n_pixels;
    return solve_oneshot(ctx, ctx->ws, rs, ps, 1e-30 /* diffraction.rs:81 */, out4, opts);
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

int xrtk_diffraction_auto_f64(const XrtkRaysF64* rays, size_t n_rays, const double* pix_x,
                              const double* pix_y, const double* pix_z, size_t n_pixels,
                              double* out4) {
  if ((n_rays && !rays) || (n_pixels && (!out4 || !pix_x || !pix_y || !pix_z)))
    return XRTK_EINVAL;
  if (n_rays) {
    const double* cols[12] = {rays->x,     rays->y,     rays->z,     rays->nx,
                              rays->ny,    rays->nz,    rays->nl,    rays->es_re,
                              rays->es_im, rays->ep_re, rays->ep_im, rays->energy};
    for (const double* c : cols)
      if (!c) return XRTK_EINVAL;
  }
  xrtk_ctx* ctx = xrtk_ctx_shared();
  if (ctx) {
    XrtkOpts o;
    xrtk_opts_default(&o);
    o.mode = XRTK_MODE_F64;
    const int rc = xrtk_diffraction_f64(ctx, rays, n_rays, pix_x, pix_y, pix_z, n_pixels, out4,
                                        &o);
    g_auto_backend.store(1);
    return rc;
  }
  if (env_flag("XRTK_REQUIRE_GPU")) return XRTK_ENODEV;
  g_auto_backend.store(0);
  try {
    return cpu_diffraction_f64(rays, n_rays, pix_x, pix_y, pix_z, n_pixels, out4);
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

void xrtk_debug_fail_launches(uint32_t n) { g_fail_launches.store(n); }

int xrtk_diffraction_grid_f64(xrtk_ctx* ctx, const XrtkRaysF64* rays, size_t n_rays,
                              const XrtkWaveParams* screen, double* out4, const XrtkOpts* opts) {
  if (!ctx) return XRTK_ENODEV;
  if (!screen || (n_rays && !rays)) return XRTK_EINVAL;
  if (n_rays && (!rays->x || !rays->y || !rays->z || !rays->energy)) return XRTK_EINVAL;
  const size_t np = (size_t)screen->nx * screen->nz;
  if (np && !out4) return XRTK_EINVAL;
  if (n_rays > 0xffffffffull || np > 0x7fff0000ull) return XRTK_EINVAL;
  if (np == 0) return XRTK_OK;
  try {
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    RaySrc rs;
    rs.f64 = rays;
    rs.nr = n_rays;
    PixSrc ps;
    ps.grid = screen;
    ps.np = np;
    return solve_oneshot(ctx, ctx->ws, rs, ps, 1e-30, out4, opts);
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

// ---- resident ray sets --------------------------------------------------
int xrtk_rayset_upload(xrtk_ctx* ctx, const XrtkRaysF64* rays, const double* path_up,
                       size_t n_rays, xrtk_rayset* out) {
  if (!ctx) return XRTK_ENODEV;
  if (!out || (n_rays && !rays) || n_rays > 0xffffffffull) return XRTK_EINVAL;
  try {
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    Range range("xrtk.rayset.upload");
    RaySet rs;
    rs.nr = (uint32_t)n_rays;
    rs.has_path = path_up != nullptr;
    if (n_rays) {
      if (!rays->x || !rays->y || !rays->z || !rays->energy) return XRTK_EINVAL;
      rs.energy = rays->energy[0];
      for (size_t i = 0; i < n_rays; ++i) {
        if (rays->energy[i] != rs.energy) rs.mono = false;
        rs.box.add(rays->x[i], rays->y[i], rays->z[i]);
      }
    }
    if (ctx->ws.stage.ensure(13 * sizeof(double) * std::max<size_t>(n_rays, 1) + 256) != hipSuccess ||
        rs.rays.ensure(sizeof(RayRec64) * std::max<size_t>(n_rays, 1)) != hipSuccess)
      return XRTK_ENOMEM;
    // (through the context's page-locked block, when it can have one; the
    // synchronisation below frees it for the next call)
    char* pin = nullptr;
    if (ctx->ws.hpin.ensure(13 * sizeof(double) * std::max<size_t>(n_rays, 1)) == hipSuccess)
      pin = static_cast<char*>(ctx->ws.hpin.p);
    int rc = upload_cols(ctx, ctx->ws.stage, rays, path_up, nullptr, rs.nr, rs.rays.as<RayRec64>(),
                         pin);
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

int xrtk_rayset_from_beam(xrtk_ctx* ctx, const double* bx, const double* by, const double* bz,
                          const double* ba, const double* bb, const double* bc, const double* be,
                          const double* es_re, const double* es_im, const double* ep_re,
                          const double* ep_im, const double* path, size_t n_beam,
                          const double* normals3, const size_t* good, size_t n_good,
                          xrtk_rayset* out) {
  static_assert(sizeof(size_t) == sizeof(uint64_t), "good indices go up as uint64");
  if (!ctx) return XRTK_ENODEV;
  if (!out) return XRTK_EINVAL;
  if (n_good && (!bx || !by || !bz || !ba || !bb || !bc || !be || !normals3 || !good))
    return XRTK_EINVAL;
  if (n_beam > 0xffffffffull || n_good > 0xffffffffull) return XRTK_EINVAL;
  for (size_t k = 0; k < n_good; ++k)
    if (good[k] >= n_beam) return XRTK_EINVAL;
  const bool amps = es_re && es_im && ep_re && ep_im;
  try {
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    Range range("xrtk.rayset.from_beam");
    RaySet rs;
    rs.nr = (uint32_t)n_good;
    rs.has_path = path != nullptr;
    if (n_good) {
      rs.energy = be[good[0]];
      for (size_t k = 0; k < n_good; ++k) {
        const size_t i = good[k];
        if (be[i] != rs.energy) rs.mono = false;
        rs.box.add(bx[i], by[i], bz[i]);
      }
    }
    // the beam's columns, its normals and the good indices go up once; the
    // gather runs on the device
    std::vector<const double*> cols = {bx, by, bz, ba, bb, bc, be};
    if (amps) cols.insert(cols.end(), {es_re, es_im, ep_re, ep_im});
    if (path) cols.push_back(path);
    const size_t nb = std::max<size_t>(n_beam, 1), cb = sizeof(double) * nb;
    const size_t off_n = align256(cols.size() * cb), off_g = off_n + align256(3 * cb);
    if (ctx->ws.stage.ensure(off_g + sizeof(uint64_t) * std::max<size_t>(n_good, 1) + 256) !=
            hipSuccess ||
        rs.rays.ensure(sizeof(RayRec64) * std::max<size_t>(n_good, 1)) != hipSuccess) {
      rs.rays.release();
      return XRTK_ENOMEM;
    }
    char* S = ctx->ws.stage.as<char>();
    hipError_t e = hipSuccess;
    if (n_good) {
      for (size_t c = 0; c < cols.size() && e == hipSuccess; ++c)
        e = hipMemcpyAsync(S + c * cb, cols[c], sizeof(double) * n_beam, hipMemcpyHostToDevice,
                           ctx->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(S + off_n, normals3, 3 * sizeof(double) * n_beam,
                           hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(S + off_g, good, sizeof(uint64_t) * n_good, hipMemcpyHostToDevice,
                           ctx->stream);
      if (e == hipSuccess)
        e = launch_rays_from_beam(reinterpret_cast<const double*>(S), (uint32_t)n_beam,
                                  reinterpret_cast<const double*>(S + off_n),
                                  reinterpret_cast<const uint64_t*>(S + off_g), (uint32_t)n_good,
                                  amps, path != nullptr, rs.rays.as<RayRec64>(), ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // the caller's buffers are free after
    if (e != hipSuccess) {
      rs.rays.release();
      return map_err(e);
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
  if (it->second.users->load() > 0) return XRTK_EAGAIN;  // a server still holds it
  DeviceGuard g(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  // cached batch graphs may reference this ray set: drop them
  for (auto*& ge : ctx->graphs) {
    if (ge->exec) (void)hipGraphExecDestroy(ge->exec);
    ge->work.release(); ge->meta.release(); ge->pix.release(); ge->perm.release();
    ge->outd.release();
    delete ge;
  }
  ctx->graphs.clear();
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

int xrtk_host_alloc(size_t bytes, void** out) {
  if (!out) return XRTK_EINVAL;
  *out = nullptr;
  if (!bytes) return XRTK_EINVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return XRTK_ENODEV;
  }
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return XRTK_ENOMEM;
  }
  *out = p;
  return XRTK_OK;
}

void xrtk_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

}  // extern "C"
