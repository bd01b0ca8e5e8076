// Batched solves over resident ray sets (one persistent pair-kernel launch,
// replayed from a hipGraph while the batch repeats) and the autotuner.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "runtime.h"

using namespace xrtk;

namespace {

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

uint64_t bits(double v) {
  uint64_t b;
  std::memcpy(&b, &v, 8);
  return b;
}

size_t a256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

namespace xrtk {

// The launch configuration of batched solves: fixed per context (the
// default, or what xrtk_autotune_jobs recorded), so that a job's result does
// not depend on which other jobs share the launch.
LaunchCfg batch_cfg(const xrtk_ctx* ctx, uint32_t mode, const XrtkOpts* o) {
  LaunchCfg c;
  if (mode == MODE_F32R) { c.P = 16; c.W = 1; } else { c.P = 1; c.W = 4; }
  c.ray_src = 1;
  c.persistent = 1;
  c.tile = mode == MODE_F32R ? 8192 : 0;
  if (ctx) {
    auto it = ctx->tune.find(std::make_tuple((int)mode, kBatchKey, kBatchKey));
    if (it != ctx->tune.end()) c = it->second;
  }
  if (o) {
    if (o->pixels_per_lane > 0) c.P = o->pixels_per_lane;
    if (o->waves_per_group > 0) c.W = o->waves_per_group;
    if (o->ray_src > 0) c.ray_src = o->ray_src;
    if (o->persistent == 0) c.persistent = 0;
    if (o->tile_rays > 0 && mode == MODE_F32R) c.tile = (o->tile_rays + 63) / 64 * 64;
  }
  return c;
}

// Builds the instance specs of a batch: pixel slots and permutations in
// `pixbuf` / `permbuf`, outputs in `outbuf` (device).  Each job's F32R/FP64
// choice comes from the residual bound of its ray set and screen.
int build_job_specs(xrtk_ctx* ctx, uint32_t want, const LaunchCfg& c32, const LaunchCfg& c64,
                    const XrtkJob* jobs, size_t n, DBuf& pixbuf, DBuf& permbuf, DBuf& outbuf,
                    std::vector<InstSpec>& specs) {
  size_t slots = 0, outs = 0;
  std::vector<size_t> slot_off(n), out_off(n);
  std::vector<uint32_t> npad(n);
  for (size_t i = 0; i < n; ++i) {
    const XrtkJob& j = jobs[i];
    const size_t np = (size_t)j.screen.nx * j.screen.nz;
    if (np > 0x7fff0000ull) return XRTK_EINVAL;
    if (np && !j.out4) return XRTK_EINVAL;
    auto it = ctx->raysets.find(j.rays);
    if (it == ctx->raysets.end()) return XRTK_EINVAL;
    const RaySet& rs = it->second;
    if (want == MODE_F32R && !(j.energy > 0) && !rs.mono) return XRTK_EINVAL;
    npad[i] = np ? grid_slots(j.screen.nx, j.screen.nz, c32, c64) : 0;
    slot_off[i] = slots;
    slots += npad[i];
    out_off[i] = outs;
    outs += a256(np);
  }
  if (pixbuf.ensure(sizeof(PixF64) * std::max<size_t>(slots, 1)) != hipSuccess ||
      permbuf.ensure(sizeof(uint32_t) * std::max<size_t>(slots, 1)) != hipSuccess ||
      outbuf.ensure(sizeof(double) * 4 * std::max<size_t>(outs, 1)) != hipSuccess)
    return XRTK_ENOMEM;
  specs.clear();
  for (size_t i = 0; i < n; ++i) {
    const XrtkJob& j = jobs[i];
    const RaySet& rs = ctx->raysets.find(j.rays)->second;
    const double e = j.energy > 0 ? j.energy : rs.energy;
    InstSpec s;
    Box pb;
    pb.add(j.screen.center[0] - std::fabs(j.screen.dx), j.screen.center[1],
           j.screen.center[2] - std::fabs(j.screen.dz));
    pb.add(j.screen.center[0] + std::fabs(j.screen.dx), j.screen.center[1],
           j.screen.center[2] + std::fabs(j.screen.dz));
    uint32_t th, tw;
    grid_tile(c32.P, j.screen.nx, j.screen.nz, &th, &tw);
    const double dq[3] = {
        j.screen.nx > 1 ? 2.0 * std::fabs(j.screen.dx) / (j.screen.nx - 1) * th : 0.0, 0.0,
        j.screen.nz > 1 ? 2.0 * std::fabs(j.screen.dz) / (j.screen.nz - 1) * tw : 0.0};
    s.mode = route_mode(want, rs.box, pb, dq, e);
    s.rays = rs.rays.as<RayRec64>();
    s.nr = rs.nr;
    s.pix = pixbuf.as<PixF64>() + slot_off[i];
    s.perm = permbuf.as<uint32_t>() + slot_off[i];
    s.np = j.screen.nx * j.screen.nz;
    s.npad = npad[i];
    // F32R needs one energy; the FP64 kernel takes the override or the rays'
    s.energy = s.mode == MODE_F32R ? e : (j.energy > 0 ? j.energy : 0.0);
    s.skip = 1e-30;
    s.static_flags = rs.has_path ? INST_F_PATHUP : 0u;
    s.out = outbuf.as<double>() + 4 * out_off[i];
    if (s.np) spec_set_grid(s, j.screen, th, tw);
    specs.push_back(s);
  }
  return XRTK_OK;
}

int enqueue_job_grids(xrtk_ctx* ctx, const LaunchCfg& c32, const XrtkJob* jobs,
                      const std::vector<InstSpec>& specs) {
  for (size_t i = 0; i < specs.size(); ++i) {
    const XrtkWaveParams& w = jobs[i].screen;
    uint32_t th, tw;
    grid_tile(c32.P, w.nx, w.nz, &th, &tw);
    hipError_t e = launch_pix_grid(w.center, w.dx, w.dz, w.nx, w.nz, th, tw, specs[i].npad,
                                   const_cast<PixF64*>(specs[i].pix),
                                   const_cast<uint32_t*>(specs[i].perm), ctx->stream);
    if (e != hipSuccess) return map_err(e);
  }
  return XRTK_OK;
}

}  // namespace xrtk

extern "C" {

int xrtk_solve_jobs(xrtk_ctx* ctx, const XrtkJob* jobs, size_t n_jobs, const XrtkOpts* opts) {
  if (!ctx) return XRTK_ENODEV;
  if (n_jobs && !jobs) return XRTK_EINVAL;
  if (!n_jobs) return XRTK_OK;
  try {
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    const uint32_t want = opts && opts->mode == XRTK_MODE_F64 ? MODE_F64 : MODE_F32R;
    const LaunchCfg c32 = batch_cfg(ctx, MODE_F32R, opts);
    const LaunchCfg c64 = batch_cfg(ctx, MODE_F64, want == MODE_F64 ? opts : nullptr);
    if (!pair_cfg_supported(MODE_F32R, c32) || !pair_cfg_supported(MODE_F64, c64))
      return XRTK_EINVAL;
    const char* ng = std::getenv("XRTK_NO_GRAPH");
    const bool use_graph = !(ng && *ng == '1');
    // XRTK_ENGINE_VALU keeps every instance off the matrix cores
    const uint32_t force = opts && opts->engine == XRTK_ENGINE_VALU ? 0x20000u : 0u;
    // graph key: everything the captured nodes bake in
    std::vector<uint64_t> key;
    key.push_back(n_jobs);
    key.push_back(want);
    key.push_back(force);
    for (const LaunchCfg* c : {&c32, &c64})
      key.push_back(((uint64_t)c->P << 32) | ((uint64_t)c->W << 16) |
                    ((uint64_t)c->ray_src << 8) | (uint64_t)c->persistent | ((uint64_t)c->tile << 40));
    for (size_t i = 0; i < n_jobs; ++i) {
      const XrtkJob& j = jobs[i];
      key.push_back(j.rays);
      key.push_back(bits(j.energy));
      key.push_back(bits(j.screen.center[0]));
      key.push_back(bits(j.screen.center[1]));
      key.push_back(bits(j.screen.center[2]));
      key.push_back(bits(j.screen.dx));
      key.push_back(bits(j.screen.dz));
      key.push_back(((uint64_t)j.screen.nx << 32) | j.screen.nz);
    }
    xrtk_ctx::GraphEntry* ge = nullptr;
    for (auto* g2 : ctx->graphs)
      if (g2->key == key) { ge = g2; break; }
    hipError_t e;
    Phases ph;
    // results -> host, on the stream behind the kernels
    auto enqueue_readback = [&]() -> hipError_t {
      for (size_t i = 0; i < n_jobs; ++i) {
        const size_t np = ge->outs[i].second;
        if (!np) continue;
        const hipError_t ce = hipMemcpyAsync(jobs[i].out4, ge->outs[i].first,
                                             sizeof(double) * 4 * np, hipMemcpyDeviceToHost,
                                             ctx->stream);
        if (ce != hipSuccess) return ce;
      }
      return hipSuccess;
    };
    bool read_back = false;
    if (!ge) {
      ph.next("xrtk.jobs.build");  // layout, grids and the first (eager) run
      auto* ne = new xrtk_ctx::GraphEntry();
      ne->key = key;
      auto drop = [&](xrtk_ctx::GraphEntry* x) {
        (void)hipStreamSynchronize(ctx->stream);
        if (x->exec) (void)hipGraphExecDestroy(x->exec);
        x->work.release(); x->meta.release(); x->pix.release(); x->perm.release();
        x->outd.release();
        delete x;
      };
      std::vector<InstSpec> specs;
      int rc = build_job_specs(ctx, want, c32, c64, jobs, n_jobs, ne->pix, ne->perm, ne->outd,
                               specs);
      if (rc != XRTK_OK) { drop(ne); return rc; }
      for (size_t i = 0; i < n_jobs; ++i) ne->outs.emplace_back(specs[i].out, (size_t)specs[i].np);
      ne->specs = specs;
      ne->jobs.assign(jobs, jobs + n_jobs);
      // first run: eager, with the metadata upload; then capture the kernels
      rc = enqueue_job_grids(ctx, c32, jobs, specs);
      if (rc == XRTK_OK)
        rc = run_instances(ctx, c32, c64, specs, ne->work, ne->meta, true, &ne->plan, 0, force,
                           true, true, nullptr, /*route_sync=*/true);
      if (rc == XRTK_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = XRTK_ELAUNCH;
      if (rc != XRTK_OK) { drop(ne); return rc; }
      if (use_graph) {
        ph.next("xrtk.jobs.capture");
        hipGraph_t graph = nullptr;
        e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
          int rc2 = enqueue_job_grids(ctx, c32, jobs, specs);
          if (rc2 == XRTK_OK)
            // (timed: the event records around the pair kernels are captured
            // too, so every replay times its own pair kernels)
            rc2 = run_instances(ctx, c32, c64, specs, ne->work, ne->meta, true, nullptr, 0,
                                force, true, /*upload=*/false,
                                ne->plan.routed ? &ne->plan.flags : nullptr);
          e = hipStreamEndCapture(ctx->stream, &graph);
          if (rc2 == XRTK_OK && e == hipSuccess && graph) {
            e = hipGraphInstantiate(&ne->exec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) ne->exec = nullptr;
          }
          if (graph) (void)hipGraphDestroy(graph);
        }
        (void)hipGetLastError();
      }
      // keep a bounded cache
      if (ctx->graphs.size() >= 16) {
        drop(ctx->graphs.front());
        ctx->graphs.erase(ctx->graphs.begin());
      }
      ctx->graphs.push_back(ne);
      ge = ne;
      float a = 0, b = 0;
      (void)hipEventElapsedTime(&a, ctx->ev[1], ctx->ev[2]);
      (void)hipEventElapsedTime(&b, ctx->ev[0], ctx->ev[3]);
      ctx->last_pair_ms = a;
      ctx->last_solve_ms = b;
    } else if (!ge->exec) {
      // no graph: the same kernels again, eagerly (the metadata the first
      // run uploaded is unchanged)
      ph.next("xrtk.jobs.replay");
      int rc = enqueue_job_grids(ctx, c32, ge->jobs.data(), ge->specs);
      if (rc == XRTK_OK)
        rc = run_instances(ctx, c32, c64, ge->specs, ge->work, ge->meta, true, nullptr, 0,
                           force, true, /*upload=*/false,
                           ge->plan.routed ? &ge->plan.flags : nullptr);
      if (rc != XRTK_OK) return rc;
      e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) return map_err(e);
      float a = 0, b = 0;
      (void)hipEventElapsedTime(&a, ctx->ev[1], ctx->ev[2]);
      (void)hipEventElapsedTime(&b, ctx->ev[0], ctx->ev[3]);
      ctx->last_pair_ms = a;
      ctx->last_solve_ms = b;
    } else {
      ph.next("xrtk.jobs.replay");
      (void)hipEventRecord(ctx->ev[0], ctx->stream);
      e = hipGraphLaunch(ge->exec, ctx->stream);
      if (e != hipSuccess) return map_err(e);
      (void)hipEventRecord(ctx->ev[3], ctx->stream);
      // the copies queue behind the graph: one wait for both
      e = enqueue_readback();
      if (e != hipSuccess) return map_err(e);
      read_back = true;
      e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) return map_err(e);
      float a = 0, b = 0;
      (void)hipEventElapsedTime(&b, ctx->ev[0], ctx->ev[3]);
      ctx->last_solve_ms = b;
      // the pair kernels' span, from the event records inside the graph
      ctx->last_pair_ms =
          hipEventElapsedTime(&a, ctx->ev[1], ctx->ev[2]) == hipSuccess && a > 0.f ? a : b;
      (void)hipGetLastError();
    }
    {
      XrtkPlan& q = ctx->last_plan;
      const Plan& p = ge->plan;
      const bool all64 = p.n_inst && p.n_f64 == p.n_inst;
      const LaunchCfg& c = all64 ? p.cfg64 : p.cfg32;
      q.mode = all64 ? XRTK_MODE_F64 : XRTK_MODE_F32R;
      q.pixels_per_lane = c.P; q.waves_per_group = c.W; q.tile_rays = c.tile; q.split = 0;
      q.ray_src = c.ray_src; q.persistent = c.persistent;
      q.n_items = (int32_t)(p.n_items32 + p.n_items64 + p.n_items_mfma);
      q.n_fp64 = (int32_t)p.n_f64;
      q.backend = 1;
      q.inst_flags = 0;
      q.n_mfma = 0;
      for (uint32_t f : p.flags) q.n_mfma += (f & INST_F_MFMA) ? 1 : 0;
    }
    // results -> host
    ph.next("xrtk.readback");
    if (!read_back) {
      e = enqueue_readback();
      if (e != hipSuccess) return map_err(e);
      e = hipStreamSynchronize(ctx->stream);
    } else {
      e = hipSuccess;
    }
    if (e == hipSuccess) {
      // the variant flags the device planner chose for the first job (known
      // on the host when the items were routed)
      if (ge->plan.routed && !ge->plan.flags.empty() && ge->plan.flags[0]) {
        ctx->last_plan.inst_flags = (int32_t)ge->plan.flags[0];
      } else if (ge->plan.scr0) {
        uint32_t fl = 0;
        if (hipMemcpy(&fl, &ge->plan.scr0->flags, sizeof(fl), hipMemcpyDeviceToHost) ==
            hipSuccess)
          ctx->last_plan.inst_flags = (int32_t)fl;
      }
    }
    return map_err(e);
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------
namespace {

int ilog2u(uint64_t v) {
  int l = 0;
  while (v > 1) { v >>= 1; ++l; }
  return l;
}

// Synthetic instance of the bucket: the flat-aperture point-source geometry
// (rays on y = 0 within +-0.1 mm, screen 1000 mm downstream, 10 keV).
void synth(size_t nr, size_t np, std::vector<double>& cols, std::vector<double>& pix) {
  cols.assign(13 * nr, 0.0);
  uint64_t s = 0x9e3779b97f4a7c15ull;
  auto u = [&] {
    s += 0x9e3779b97f4a7c15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (size_t i = 0; i < nr; ++i) {
    cols[0 * nr + i] = (u() - 0.5) * 0.2;
    cols[2 * nr + i] = (u() - 0.5) * 0.2;
    cols[4 * nr + i] = 1.0;
    cols[6 * nr + i] = 1.0;
    cols[7 * nr + i] = 1.0;
    cols[11 * nr + i] = 10000.0;
  }
  const size_t side = (size_t)std::max<double>(1.0, std::floor(std::sqrt((double)np)));
  pix.assign(3 * np, 0.0);
  for (size_t i = 0; i < np; ++i) {
    const size_t ix = i / side, iz = i % side;
    pix[i] = ((double)ix / (double)side - 0.5) * 0.04;
    pix[np + i] = 1000.0;
    pix[2 * np + i] = ((double)iz / (double)side - 0.5) * 0.04;
  }
}

void save_cache(xrtk_ctx* ctx) {
  const char* path = std::getenv("XRTK_TUNE_CACHE");
  if (!path || !*path) return;
  std::ofstream f(path);
  if (!f) return;
  f << "[\n";
  bool first = true;
  for (auto& kv : ctx->tune) {
    if (!first) f << ",\n";
    first = false;
    f << "  {\"mode\": " << std::get<0>(kv.first) << ", \"log2_rays\": " << std::get<1>(kv.first)
      << ", \"log2_pixels\": " << std::get<2>(kv.first) << ", \"P\": " << kv.second.P
      << ", \"W\": " << kv.second.W << ", \"ray_src\": " << kv.second.ray_src
      << ", \"persistent\": " << kv.second.persistent << ", \"split\": " << kv.second.split
      << ", \"tile\": " << kv.second.tile << "}";
  }
  f << "\n]\n";
}

}  // namespace

namespace xrtk {

// Loads the JSON cache written by save_cache (one object per line).
void load_tune_cache(xrtk_ctx* ctx) {
  const char* path = std::getenv("XRTK_TUNE_CACHE");
  if (!path || !*path) return;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    int mode, lr, lp, P, W, src, per, split, tile;
    if (std::sscanf(line.c_str(),
                    " {\"mode\": %d, \"log2_rays\": %d, \"log2_pixels\": %d, \"P\": %d, \"W\": %d, "
                    "\"ray_src\": %d, \"persistent\": %d, \"split\": %d, \"tile\": %d",
                    &mode, &lr, &lp, &P, &W, &src, &per, &split, &tile) == 9) {
      LaunchCfg c;
      c.P = P; c.W = W; c.ray_src = src; c.persistent = per; c.split = split; c.tile = tile;
      if (pair_cfg_supported((uint32_t)mode, c)) ctx->tune[std::make_tuple(mode, lr, lp)] = c;
    }
  }
}

}  // namespace xrtk

extern "C" int xrtk_autotune(xrtk_ctx* ctx, size_t n_rays, size_t n_pixels, int mode_in,
                             XrtkOpts* best, double* best_ms) {
  if (!ctx) return XRTK_ENODEV;
  if (!n_rays || !n_pixels || n_rays > 0xffffffffull || n_pixels > 0x7fffffffull)
    return XRTK_EINVAL;
  const uint32_t mode = mode_in == XRTK_MODE_F64 ? MODE_F64 : MODE_F32R;
  try {
    std::vector<double> cols, pix;
    // Time on the full pixel count and a ray sample: per-pair cost does not
    // depend on the ray count once chunks are long.
    const size_t nr = std::min<size_t>(n_rays, mode == MODE_F32R ? 131072 : 16384);
    synth(nr, n_pixels, cols, pix);
    XrtkRaysF64 r;
    r.x = &cols[0]; r.y = &cols[nr]; r.z = &cols[2 * nr];
    r.nx = &cols[3 * nr]; r.ny = &cols[4 * nr]; r.nz = &cols[5 * nr];
    r.nl = &cols[6 * nr]; r.es_re = &cols[7 * nr]; r.es_im = &cols[8 * nr];
    r.ep_re = &cols[9 * nr]; r.ep_im = &cols[10 * nr]; r.energy = &cols[11 * nr];
    std::vector<double> out(4 * n_pixels);
    std::vector<LaunchCfg> cands;
    if (mode == MODE_F32R) {
      for (int src : {1})
        for (int P : {2, 4, 8, 16})
          for (int W : {1, 4})
            for (int per : {0, 1}) {
              LaunchCfg c; c.P = P; c.W = W; c.ray_src = src; c.persistent = per; c.tile = 8192;
              if (pair_cfg_supported(mode, c)) cands.push_back(c);
            }
    } else {
      for (int P : {1, 2})
        for (int W : {1, 4})
          for (int per : {0, 1}) {
            LaunchCfg c; c.P = P; c.W = W; c.ray_src = 1; c.persistent = per; c.tile = 0;
            cands.push_back(c);
          }
    }
    double bt = 1e300;
    LaunchCfg bc = cands.front();
    for (const LaunchCfg& c : cands) {
      XrtkOpts o;
      xrtk_opts_default(&o);
      o.mode = (int)mode;
      o.pixels_per_lane = c.P; o.waves_per_group = c.W; o.ray_src = c.ray_src;
      o.persistent = c.persistent; o.tile_rays = c.tile;
      double t = 1e300;
      for (int rep = 0; rep < 3; ++rep) {
        const int rc = xrtk_diffraction_f64(ctx, &r, nr, &pix[0], &pix[n_pixels],
                                            &pix[2 * n_pixels], n_pixels, out.data(), &o);
        if (rc != XRTK_OK) { t = 1e300; break; }
        if (rep) t = std::min(t, ctx->last_pair_ms);
      }
      if (t < bt) { bt = t; bc = c; }
    }
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      ctx->tune[std::make_tuple((int)mode, ilog2u(n_rays), ilog2u(n_pixels))] = bc;
      save_cache(ctx);
    }
    if (best) {
      xrtk_opts_default(best);
      best->mode = (int)mode;
      best->pixels_per_lane = bc.P; best->waves_per_group = bc.W; best->ray_src = bc.ray_src;
      best->persistent = bc.persistent; best->tile_rays = bc.tile;
    }
    if (best_ms) *best_ms = bt * (double)n_rays / (double)nr;
    return XRTK_OK;
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

// Batch autotuner: times each candidate launch shape on the caller's batch
// and records the fastest as the context's batched-solve configuration
// (xrtk_solve_jobs, and servers created afterwards).  Each instance is cut to
// its first R rays for the timing, R = kTuneRays doubled until the cut batch
// holds kTunePairs pairs: the pixel slots, tiles and launch shape are those of
// the real batch, and the per-ray cost does not depend on the ray count once
// chunks are long and the launch fills the device.
namespace {
constexpr uint32_t kTuneRays = 65536;
// Pairs one timed launch should hold: about two rounds of full-size items
// (1024 slots x 8192 rays) over the 8192 resident one-wave groups, so that
// every candidate is timed on a full device.
constexpr double kTunePairs = 1.2e11;
struct ScratchBufs {
  DBuf pix, perm, outd, work, meta;
  ~ScratchBufs() { pix.release(); perm.release(); outd.release(); work.release(); meta.release(); }
};
}  // namespace

extern "C" int xrtk_autotune_jobs(xrtk_ctx* ctx, const XrtkJob* jobs, size_t n_jobs, int mode_in,
                                  XrtkOpts* best, double* best_ms) {
  if (!ctx) return XRTK_ENODEV;
  if (!n_jobs || !jobs) return XRTK_EINVAL;
  const uint32_t mode = mode_in == XRTK_MODE_F64 ? MODE_F64 : MODE_F32R;
  try {
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    std::vector<LaunchCfg> cands;
    for (int P : mode == MODE_F32R ? std::vector<int>{2, 4, 8, 16} : std::vector<int>{1, 2})
      for (int W : {1, 4})
        for (int per : {0, 1}) {
          LaunchCfg c;
          c.P = P; c.W = W; c.ray_src = 1; c.persistent = per;
          c.tile = mode == MODE_F32R ? 8192 : 0;
          if (pair_cfg_supported(mode, c)) cands.push_back(c);
        }
    if (cands.empty()) return XRTK_EINVAL;
    const LaunchCfg other = batch_cfg(ctx, mode == MODE_F32R ? MODE_F64 : MODE_F32R, nullptr);
    ScratchBufs b;
    double bt = 1e300;
    LaunchCfg bc = cands.front();
    uint32_t cut = 0;  // rays per instance in the timed launches
    for (const LaunchCfg& c : cands) {
      const LaunchCfg& c32 = mode == MODE_F32R ? c : other;
      const LaunchCfg& c64 = mode == MODE_F32R ? other : c;
      std::vector<InstSpec> specs;
      int rc = build_job_specs(ctx, mode, c32, c64, jobs, n_jobs, b.pix, b.perm, b.outd, specs);
      if (rc != XRTK_OK) return rc;
      if (!cut) {
        // kTuneRays, doubled while the cut batch holds fewer than kTunePairs
        uint32_t max_nr = 0;
        for (const InstSpec& s : specs) max_nr = std::max(max_nr, s.nr);
        auto pairs_at = [&](uint64_t R) {
          double p = 0;
          for (const InstSpec& s : specs) p += (double)std::min<uint64_t>(s.nr, R) * s.np;
          return p;
        };
        cut = kTuneRays;
        while (cut < max_nr && pairs_at(cut) < kTunePairs) cut = cut > (1u << 30) ? max_nr : 2 * cut;
      }
      for (InstSpec& s : specs) s.nr = std::min<uint32_t>(s.nr, cut);
      double t = 1e300;
      for (int rep = 0; rep < 3; ++rep) {
        rc = enqueue_job_grids(ctx, c32, jobs, specs);
        if (rc == XRTK_OK)
          rc = run_instances(ctx, c32, c64, specs, b.work, b.meta, true, nullptr, 0, 0, true,
                             true, nullptr, /*route_sync=*/true);
        if (rc == XRTK_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = XRTK_ELAUNCH;
        if (rc != XRTK_OK) return rc;
        float ms = 0.f;
        if (rep && hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]) == hipSuccess)
          t = std::min(t, (double)ms);
      }
      if (t < bt) { bt = t; bc = c; }
    }
    ctx->tune[std::make_tuple((int)mode, kBatchKey, kBatchKey)] = bc;
    save_cache(ctx);
    if (best) {
      xrtk_opts_default(best);
      best->mode = (int)mode;
      best->pixels_per_lane = bc.P; best->waves_per_group = bc.W; best->ray_src = bc.ray_src;
      best->persistent = bc.persistent; best->tile_rays = bc.tile;
    }
    if (best_ms) *best_ms = bt;
    return XRTK_OK;
  } catch (...) {
    return XRTK_ENOMEM;
  }
}
