// Host runtime: device context, HBM workspace, planner and the solve pipeline.
//
// Replaces the wgpu host side of physwkim/xrt-rs (src/gpu/kirchhoff.rs
// :85-298 and src/gpu/context.rs:20-67) with a HIP design: the context is
// created once per process (or explicitly), owns one stream and a grow-only
// workspace, and every solve is centroid -> bounds/plan -> FP32 records ->
// pair kernel -> FP64 merge on that stream, with no per-call compilation.
#include "runtime.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace xrtk {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int map_err(hipError_t e) {
  if (e == hipSuccess) return XRTK_OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return XRTK_ENOMEM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return XRTK_ENODEV;
  return XRTK_ELAUNCH;
}

hipError_t DBuf::ensure(size_t bytes) {
  if (bytes <= cap && p) return hipSuccess;
  release();
  size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
  want = (want + 4095) & ~size_t(4095);
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    cap = 0;
    e = hipMalloc(&p, bytes ? bytes : 256);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return hipErrorOutOfMemory;
    }
    cap = bytes ? bytes : 256;
    return hipSuccess;
  }
  cap = want;
  return hipSuccess;
}

void DBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

// ---------------------------------------------------------------------------
// planning
// ---------------------------------------------------------------------------
static int ilog2(uint64_t v) {
  int l = 0;
  while (v > 1) { v >>= 1; ++l; }
  return l;
}

static LaunchCfg default_cfg(uint32_t mode, uint64_t nr, uint64_t np) {
  LaunchCfg c;
  if (mode == MODE_F32R) {
    c.P = 4; c.W = 4; c.ray_src = 1; c.persistent = 0; c.tile = 1024;
    if (np < 64ull * 4 * 4 * 64) { c.P = 2; c.W = 4; }
  } else {
    c.P = 1; c.W = 4; c.ray_src = 1; c.persistent = 0; c.tile = 1024;
  }
  (void)nr;
  return c;
}

LaunchCfg choose_cfg(xrtk_ctx* ctx, uint32_t mode, uint64_t nr, uint64_t np, const XrtkOpts* o) {
  LaunchCfg c = default_cfg(mode, nr, np);
  auto it = ctx->tune.find(std::make_tuple((int)mode, ilog2(nr), ilog2(np)));
  if (it != ctx->tune.end()) c = it->second;
  if (o) {
    if (o->pixels_per_lane > 0) c.P = o->pixels_per_lane;
    if (o->waves_per_group > 0) c.W = o->waves_per_group;
    if (o->tile_rays > 0) c.tile = o->tile_rays;
    if (o->ray_src > 0) c.ray_src = o->ray_src;
    if (o->persistent >= 0 && o->persistent <= 1 && (o->persistent == 1 || o->split != 0 ||
                                                      o->pixels_per_lane != 0))
      c.persistent = o->persistent;
    if (o->split > 0) c.split = o->split;
  }
  if (c.tile < 64) c.tile = 64;
  c.tile = (c.tile + 63) / 64 * 64;
  return c;
}

// Work items per launch that fill the device: every CU holds ~8 waves of
// the pair kernel per SIMD; aim for 2 rounds so the tail stays short.
static uint64_t target_items(const xrtk_ctx* ctx, const LaunchCfg& cfg) {
  const uint64_t waves = (uint64_t)ctx->n_cu * 4 * 8;
  return std::max<uint64_t>(1, 2 * waves / cfg.W);
}

uint32_t chunk_len_for(const xrtk_ctx* ctx, uint32_t mode, const LaunchCfg& cfg, uint32_t nr,
                       uint32_t np, int split_override) {
  (void)mode;
  if (nr == 0) return cfg.tile;
  const uint64_t B = 64ull * cfg.W * cfg.P;
  const uint64_t pblocks = std::max<uint64_t>(1, (np + B - 1) / B);
  uint64_t split;
  if (split_override > 0) {
    split = (uint64_t)split_override;
  } else {
    split = (target_items(ctx, cfg) + pblocks - 1) / pblocks;
    const uint64_t min_chunk = 4096;  // keep per-item work >> prologue/epilogue
    split = std::min<uint64_t>(split, std::max<uint64_t>(1, nr / min_chunk));
  }
  split = std::max<uint64_t>(1, std::min<uint64_t>(split, nr));
  uint64_t cl = (nr + split - 1) / split;
  cl = (cl + cfg.tile - 1) / cfg.tile * cfg.tile;  // chunk edges on tile edges
  return (uint32_t)std::min<uint64_t>(cl, 0xffffffffu);
}

static void kr_dd(double k, double* hi, double* lo) {
  const long double two_pi = 6.283185307179586476925286766559005768L;
  const long double kr = (long double)k / two_pi;
  *hi = (double)kr;
  *lo = (double)(kr - (long double)*hi);
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

int run_instances(xrtk_ctx* ctx, uint32_t mode, const LaunchCfg& cfg,
                  const std::vector<InstSpec>& specs, DBuf& work, DBuf& meta, bool timed,
                  Plan* plan_out, int split_override, uint32_t force_flags,
                  bool persistent_sorted, bool upload) {
  hipStream_t st = ctx->stream;
  if (!pair_cfg_supported(mode, cfg)) return XRTK_EINVAL;
  const size_t n = specs.size();
  // ---- layout of the per-instance buffers in `work`
  struct Lay {
    size_t rays32, pix32, R64, partial, scr, part;
    uint32_t chunk_len, n_chunks;
  };
  std::vector<Lay> lay(n);
  size_t off = 0;
  for (size_t i = 0; i < n; ++i) {
    const InstSpec& s = specs[i];
    Lay& L = lay[i];
    L.chunk_len = chunk_len_for(ctx, mode, cfg, s.nr, s.np, split_override);
    L.n_chunks = s.nr ? (s.nr + L.chunk_len - 1) / L.chunk_len : 0;
    L.rays32 = off; off = align_up(off + (mode == MODE_F32R ? sizeof(RayRec32) * s.nr : 0), 256);
    L.pix32 = off; off = align_up(off + (mode == MODE_F32R ? sizeof(PixRec32) * s.np : 0), 256);
    L.R64 = off; off = align_up(off + sizeof(double) * s.np, 256);
    L.partial = off; off = align_up(off + sizeof(double) * 4 * (size_t)s.np * L.n_chunks, 256);
    L.scr = off; off = align_up(off + sizeof(InstScratch), 256);
    L.part = off; off = align_up(off + sizeof(double) * 4 * kPartBlocks, 256);
  }
  hipError_t e = work.ensure(off);
  if (e != hipSuccess) return XRTK_ENOMEM;
  char* W = work.as<char>();

  // ---- instance table and work items
  std::vector<KInst> insts(n);
  std::vector<WorkItem> items;
  const uint32_t B = 64u * cfg.W * cfg.P;
  uint32_t max_np = 0;
  std::vector<uint32_t> merge_list;
  for (size_t i = 0; i < n; ++i) {
    const InstSpec& s = specs[i];
    const Lay& L = lay[i];
    KInst& I = insts[i];
    std::memset(&I, 0, sizeof(I));
    I.rays64 = s.rays;
    I.rays32 = reinterpret_cast<const RayRec32*>(W + L.rays32);
    I.pix64 = s.pix;
    I.pix32 = reinterpret_cast<const PixRec32*>(W + L.pix32);
    I.R64 = reinterpret_cast<const double*>(W + L.R64);
    I.partial = reinterpret_cast<double*>(W + L.partial);
    I.out = s.out;
    I.scr = reinterpret_cast<InstScratch*>(W + L.scr);
    I.nr = s.nr; I.np = s.np; I.n_chunks = L.n_chunks; I.chunk_len = L.chunk_len;
    I.mode = mode; I.tile = cfg.tile; I.accumulate = s.accumulate;
    I.energy = s.energy;
    I.k = s.energy / kCHBAR_H * 1e7;
    kr_dd(I.k, &I.kr_hi, &I.kr_lo);
    I.skip = s.skip;
    max_np = std::max(max_np, s.np);
    for (uint32_t c = 0; c < L.n_chunks; ++c)
      for (uint32_t p0 = 0; p0 < s.np; p0 += B) items.push_back(WorkItem{(uint32_t)i, p0, c, 0});
  }
  // Batched launches: longest items first so the queue drains evenly.
  if (persistent_sorted && n > 1) {
    std::stable_sort(items.begin(), items.end(), [&](const WorkItem& a, const WorkItem& b) {
      const KInst& A = insts[a.inst];
      const KInst& Bq = insts[b.inst];
      const uint64_t la = std::min<uint64_t>(A.chunk_len, A.nr - (uint64_t)a.chunk * A.chunk_len);
      const uint64_t lb = std::min<uint64_t>(Bq.chunk_len, Bq.nr - (uint64_t)b.chunk * Bq.chunk_len);
      return la > lb;
    });
  }
  // merge lists: instances writing their own output merge in one launch;
  // instances that accumulate into an earlier output follow one by one.
  std::vector<uint32_t> first, rest;
  for (size_t i = 0; i < n; ++i)
    if (specs[i].nr) (specs[i].accumulate ? rest : first).push_back((uint32_t)i);
  std::vector<uint32_t> lists(first);
  lists.insert(lists.end(), rest.begin(), rest.end());
  uint32_t first_np = 0;
  for (uint32_t i : first) first_np = std::max(first_np, specs[i].np);

  const size_t meta_bytes = align_up(sizeof(KInst) * n, 256) +
                            align_up(sizeof(WorkItem) * std::max<size_t>(1, items.size()), 256) +
                            align_up(sizeof(uint32_t) * (lists.size() + 64), 256);
  e = meta.ensure(meta_bytes);
  if (e != hipSuccess) return XRTK_ENOMEM;
  char* M = meta.as<char>();
  KInst* d_insts = reinterpret_cast<KInst*>(M);
  WorkItem* d_items = reinterpret_cast<WorkItem*>(M + align_up(sizeof(KInst) * n, 256));
  uint32_t* d_lists = reinterpret_cast<uint32_t*>(
      M + align_up(sizeof(KInst) * n, 256) +
      align_up(sizeof(WorkItem) * std::max<size_t>(1, items.size()), 256));
  uint32_t* d_queue = d_lists + lists.size() + 32;
  if (upload) e = hipMemcpyAsync(d_insts, insts.data(), sizeof(KInst) * n, hipMemcpyHostToDevice, st);
  if (upload && e == hipSuccess && !items.empty())
    e = hipMemcpyAsync(d_items, items.data(), sizeof(WorkItem) * items.size(),
                       hipMemcpyHostToDevice, st);
  if (upload && e == hipSuccess && !lists.empty())
    e = hipMemcpyAsync(d_lists, lists.data(), sizeof(uint32_t) * lists.size(),
                       hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return map_err(e);

  if (timed) (void)hipEventRecord(ctx->ev[0], st);
  // ---- prep of every instance
  for (size_t i = 0; i < n; ++i) {
    const InstSpec& s = specs[i];
    const Lay& L = lay[i];
    if (s.nr == 0) {
      // No rays: the field is zero.  (The reference's GPU path divides by
      // the ray count in its centroid, kirchhoff.rs:97-104, and returns NaN.)
      if (s.np && !s.accumulate) {
        e = launch_zero_f64(s.out, (size_t)s.np * 4, st);
        if (e != hipSuccess) return map_err(e);
      }
      continue;
    }
    e = launch_prep(s.rays, s.nr, s.pix, s.np, reinterpret_cast<InstScratch*>(W + L.scr),
                    reinterpret_cast<double*>(W + L.part), mode, insts[i].kr_hi, insts[i].kr_lo,
                    s.skip, /*qfold_rev=*/8.0, force_flags, s.static_flags,
                    reinterpret_cast<RayRec32*>(W + L.rays32),
                    reinterpret_cast<PixRec32*>(W + L.pix32), reinterpret_cast<double*>(W + L.R64),
                    st);
    if (e != hipSuccess) return map_err(e);
  }
  // ---- pair kernel
  const uint32_t n_items = (uint32_t)items.size();
  uint32_t grid = n_items;
  if (cfg.persistent) {
    const uint32_t resident = (uint32_t)ctx->n_cu * (mode == MODE_F32R ? 32u : 16u) / cfg.W;
    grid = std::max<uint32_t>(1, std::min<uint32_t>(n_items, resident));
  }
  const uint32_t remap = (!cfg.persistent && n == 1) ? 1u : 0u;
  if (timed) (void)hipEventRecord(ctx->ev[1], st);
  e = launch_pair(mode, cfg, d_insts, d_items, n_items, d_queue, grid, remap, st);
  if (e != hipSuccess) return map_err(e);
  if (timed) (void)hipEventRecord(ctx->ev[2], st);
  // ---- merge
  if (!first.empty()) {
    e = launch_merge(d_insts, d_lists, (uint32_t)first.size(), first_np, st);
    if (e != hipSuccess) return map_err(e);
  }
  for (size_t j = 0; j < rest.size(); ++j) {
    e = launch_merge(d_insts, d_lists + first.size() + j, 1, specs[rest[j]].np, st);
    if (e != hipSuccess) return map_err(e);
  }
  if (timed) (void)hipEventRecord(ctx->ev[3], st);
  if (plan_out) {
    plan_out->mode = mode;
    plan_out->cfg = cfg;
    plan_out->n_items = n_items;
    plan_out->grid = grid;
    plan_out->remap = remap;
  }
  (void)max_np;
  (void)B;
  return XRTK_OK;
}

}  // namespace xrtk
