// Host runtime: device context, HBM workspace, planner and the solve pipeline.
//
// Replaces the wgpu host side of physwkim/xrt-rs (src/gpu/kirchhoff.rs
// :85-298 and src/gpu/context.rs:20-67) with a HIP design: the context is
// created once per process (or explicitly), owns one stream and a grow-only
// workspace, and every solve is centroid -> bounds/plan -> FP32 records ->
// pair kernels -> FP64 merge on that stream, with no per-call compilation.
#include "runtime.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <thread>

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

hipError_t copy_h2d(hipStream_t st, char* pin, void* dst, const void* src, size_t bytes) {
  if (!pin || bytes < (size_t(4) << 20))
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
  // chunks of 8 MiB: each chunk's DMA is queued as soon as it is in the
  // block, while the host threads copy the next
  hipError_t e = hipSuccess;
  gather_chunks(pin, static_cast<const char*>(src), bytes, size_t(8) << 20,
                [&](size_t c0, size_t n) {
                  e = hipMemcpyAsync(static_cast<char*>(dst) + c0, pin + c0, n,
                                     hipMemcpyHostToDevice, st);
                  return e == hipSuccess;
                });
  return e;
}

hipError_t copy_d2h(hipStream_t st, char* pin, const void* src, const Segment* segs,
                    size_t n_segs) {
  size_t bytes = 0;
  for (size_t k = 0; k < n_segs; ++k) bytes += segs[k].len;
  hipError_t e = hipSuccess;
  if (!pin || bytes < (size_t(4) << 20)) {
    size_t off = 0;
    for (size_t k = 0; k < n_segs && e == hipSuccess; off += segs[k].len, ++k)
      if (segs[k].dst && segs[k].len)
        e = hipMemcpyAsync(segs[k].dst, static_cast<const char*>(src) + off, segs[k].len,
                           hipMemcpyDeviceToHost, st);
    return e == hipSuccess ? hipStreamSynchronize(st) : e;
  }
  // chunks of 4 MiB into the block, each with an event; the host copies a
  // chunk out as soon as its event has fired, while later ones cross PCIe
  const size_t kChunk = size_t(4) << 20;
  const size_t C = (bytes + kChunk - 1) / kChunk;
  std::vector<hipEvent_t> ev(C, nullptr);
  for (size_t c = 0; c < C && e == hipSuccess; ++c)
    e = hipEventCreateWithFlags(&ev[c], hipEventDisableTiming);
  for (size_t c = 0; c < C && e == hipSuccess; ++c) {
    const size_t c0 = c * kChunk;
    e = hipMemcpyAsync(pin + c0, static_cast<const char*>(src) + c0,
                       std::min(kChunk, bytes - c0), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(ev[c], st);
  }
  if (e == hipSuccess)
    scatter_chunks(pin, segs, n_segs, kChunk, [&](size_t c) {
      e = hipEventSynchronize(ev[c]);
      return e == hipSuccess;
    });
  // (every queued copy has run before the block is handed back)
  const hipError_t es = hipStreamSynchronize(st);
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  return e != hipSuccess ? e : es;
}

hipError_t HBuf::ensure(size_t bytes) {
  if (bytes <= cap && p) return hipSuccess;
  release();
  size_t want = std::max<size_t>(bytes + bytes / 4, 1u << 20);
  want = (want + 4095) & ~size_t(4095);
  hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    return e;
  }
  cap = want;
  return hipSuccess;
}

void HBuf::release() {
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
}

hipError_t DBuf::ensure(size_t bytes) {
  if (bytes <= cap && p) return hipSuccess;
  if (arena) {
    if (p) return hipErrorOutOfMemory;  // a carved block does not grow
    const size_t off = (arena->off + 255) & ~size_t(255);
    if (off > arena->cap || bytes > arena->cap - off) return hipErrorOutOfMemory;
    p = arena->base + off;
    cap = bytes;
    arena->off = off + bytes;
    return hipSuccess;
  }
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
  if (p && !arena) (void)hipFree(p);
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

static LaunchCfg default_cfg(uint32_t mode, uint64_t nr, uint64_t np, bool f32_pixels) {
  LaunchCfg c;
  if (mode == MODE_F32R) {
    // one wave per group; 16 slots per lane: on 1 M x 512^2 the slot-line
    // form takes 57 ms against 64 ms at 8 slots, and the per-slot forms run
    // 16 slots as two passes of 8 at 8 slots' speed (107 ms).  float32
    // pixels sit off the slot line by their rounding: the line with per-slot
    // offsets takes 85.4 ms at 8 slots, 86.9 ms at 16 (two passes)
    c.P = f32_pixels ? 8 : 16; c.W = 1;
    if (np < 64ull * 4 * 4 * 64) { c.P = 2; c.W = 4; }
  } else {
    c.P = 1; c.W = 4;
  }
  c.ray_src = 1; c.persistent = 1; c.tile = mode == MODE_F32R ? 8192 : 0;
  (void)nr;
  return c;
}

LaunchCfg cfg64_for(const LaunchCfg& c32) {
  LaunchCfg c;
  c.P = 1; c.W = 4; c.ray_src = 1; c.persistent = c32.persistent; c.tile = 0;
  c.split = c32.split;
  return c;
}

bool cfg_shape_given(const xrtk_ctx* ctx, uint32_t mode, uint64_t nr, uint64_t np,
                     const XrtkOpts* o) {
  if (o && o->pixels_per_lane > 0) return true;
  return ctx->tune.count(std::make_tuple((int)mode, ilog2(nr), ilog2(np))) != 0;
}

LaunchCfg choose_cfg(xrtk_ctx* ctx, uint32_t mode, uint64_t nr, uint64_t np, const XrtkOpts* o,
                     bool f32_pixels) {
  LaunchCfg c = default_cfg(mode, nr, np, f32_pixels);
  auto it = ctx->tune.find(std::make_tuple((int)mode, ilog2(nr), ilog2(np)));
  if (it != ctx->tune.end()) c = it->second;
  if (o) {
    if (o->pixels_per_lane > 0) c.P = o->pixels_per_lane;
    if (o->waves_per_group > 0) c.W = o->waves_per_group;
    if (o->tile_rays > 0) c.tile = o->tile_rays;
    if (o->ray_src > 0) c.ray_src = o->ray_src;
    if (o->persistent == 0 || o->persistent == 1) c.persistent = o->persistent;
    if (o->split > 0) c.split = o->split;
  }
  if (mode == MODE_F32R) {
    if (c.tile < 64) c.tile = 64;
    c.tile = (c.tile + 63) / 64 * 64;
  } else {
    c.tile = 0;  // the FP64 kernel keeps Kahan sums over any chunk length
  }
  return c;
}

// Work items per launch that fill the device: ~8 resident waves per SIMD;
// aim for 2 rounds so the tail stays short.
static uint64_t target_items(const xrtk_ctx* ctx, const LaunchCfg& cfg) {
  const uint64_t waves = (uint64_t)ctx->n_cu * 4 * 8;
  return std::max<uint64_t>(1, 2 * waves / cfg.W);
}

uint32_t chunk_len_for(const xrtk_ctx* ctx, const LaunchCfg& cfg, uint32_t nr, uint32_t npad,
                       int split_override) {
  if (nr == 0) return cfg.tile;
  const uint64_t B = 64ull * cfg.W * cfg.P;
  const uint64_t pblocks = std::max<uint64_t>(1, (npad + B - 1) / B);
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
  cl = (cl + 63) / 64 * 64;  // whole LDS ray blocks
  if (cfg.tile > 0) cl = std::min<uint64_t>(cl, (uint64_t)cfg.tile);
  return (uint32_t)std::min<uint64_t>(cl, 0xffffffffu);
}

static uint32_t slot_block(const LaunchCfg& c32, const LaunchCfg& c64) {
  const uint32_t b32 = 64u * c32.P * c32.W, b64 = 64u * c64.P * c64.W;
  return std::max(b32, b64);  // both powers of two: the larger is a multiple
}

uint32_t padded_slots(uint32_t np, const LaunchCfg& c32, const LaunchCfg& c64) {
  const uint32_t b = slot_block(c32, c64);
  return (np + b - 1) / b * b;
}

void grid_tile(int P, uint32_t nx, uint32_t nz, uint32_t* th, uint32_t* tw) {
  const uint32_t n = 64u * P;
  const uint32_t w0 = P >= 8 ? 32u : 16u, h0 = n / w0;
  *tw = w0;
  *th = h0;
  const uint32_t w1 = 2 * w0, h1 = h0 / 2;
  if (w1 <= 64u && h1 >= 1u) {
    auto area = [&](uint64_t h, uint64_t w) {
      return ((nx + h - 1) / h * h) * ((nz + w - 1) / w * w);
    };
    if (area(h1, w1) < area(h0, w0)) {
      *tw = w1;
      *th = h1;
    }
  }
}

uint32_t grid_slots(uint32_t nx, uint32_t nz, const LaunchCfg& c32, const LaunchCfg& c64) {
  uint32_t th, tw;
  grid_tile(c32.P, nx, nz, &th, &tw);
  const uint64_t ntiles = (uint64_t)((nx + th - 1) / th) * ((nz + tw - 1) / tw);
  const uint64_t s = ntiles * th * tw;
  const uint32_t b = slot_block(c32, c64);
  return (uint32_t)((s + b - 1) / b * b);
}

double residual_bound_rev(const Box& rays, const Box& pix, const double dq[3], double energy) {
  if (rays.empty() || pix.empty()) return 0.0;
  double w2 = 0, gap2 = 0, lin = 0;
  for (int i = 0; i < 3; ++i) {
    const double w = rays.hi[i] - rays.lo[i];
    w2 += w * w;
    const double g = std::max({0.0, pix.lo[i] - rays.hi[i], rays.lo[i] - pix.hi[i]});
    gap2 += g * g;
    lin += w * dq[i];
  }
  const double w = std::sqrt(w2), den = 2.0 * std::sqrt(gap2) - w;
  if (!(std::sqrt(gap2) > w * (1.0 + 1e-3) + 1e-12) || den <= 0.0)
    return std::numeric_limits<double>::infinity();
  const double kr = energy / kCHBAR_H * 1e7 / (2.0 * M_PI);
  return kr * (2.0 * lin + w2 * w / den) / den;
}

uint32_t route_mode(uint32_t mode, const Box& rays, const Box& pix, const double dq[3],
                    double energy) {
  if (mode != MODE_F32R) return mode;
  return residual_bound_rev(rays, pix, dq, energy) > kMaxResidualRev ? MODE_F64 : MODE_F32R;
}

static void kr_dd(double k, double* hi, double* lo) {
  const long double two_pi = 6.283185307179586476925286766559005768L;
  const long double kr = (long double)k / two_pi;
  *hi = (double)kr;
  *lo = (double)(kr - (long double)*hi);
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

void spec_set_grid(InstSpec& s, const XrtkWaveParams& w, uint32_t th, uint32_t tw) {
  s.gnx = w.nx; s.gnz = w.nz; s.gth = th; s.gtw = tw;
  s.gx0 = w.nx > 1 ? w.center[0] - w.dx : w.center[0];
  s.gz0 = w.nz > 1 ? w.center[2] - w.dz : w.center[2];
  s.gsx = w.nx > 1 ? 2.0 * w.dx / (double)(w.nx - 1) : 0.0;
  s.gsz = w.nz > 1 ? 2.0 * w.dz / (double)(w.nz - 1) : 0.0;
  s.gy = w.center[1];
}

// Whether an instance may run on the matrix cores: a reduced-kernel instance
// with rays on a grid screen, unless force flag 0x20000 (XRTK_ENGINE_VALU)
// forbids it.  The planner kernel has the last word.
bool mfma_allowed(const InstSpec& s, uint32_t force_flags) {
  return !(force_flags & 0x20000u) && s.mode == MODE_F32R && s.nr > 0 && s.gnx > 0 &&
         s.gnz > 0 && s.gth > 0 && s.gtw > 0;
}

int run_instances(xrtk_ctx* ctx, const LaunchCfg& cfg32, const LaunchCfg& cfg64,
                  const std::vector<InstSpec>& specs, DBuf& work, DBuf& meta, bool timed,
                  Plan* plan_out, int split_override, uint32_t force_flags, bool sorted,
                  bool upload, const std::vector<uint32_t>* known_flags, bool route_sync) {
  Phases ph;
  ph.next("xrtk.layout");  // work items and instance metadata
  hipStream_t st = ctx->stream;
  if (!pair_cfg_supported(MODE_F32R, cfg32) || !pair_cfg_supported(MODE_F64, cfg64))
    return XRTK_EINVAL;
  const size_t n = specs.size();
  if (known_flags && known_flags->size() != n) return XRTK_EINVAL;
  const uint32_t sblock = slot_block(cfg32, cfg64);
  // ---- layout of the per-instance buffers in `work`
  struct Lay {
    size_t rays32, pix32, tiles, R64, partial, scr, part;
    uint32_t chunk_len, n_chunks;
  };
  std::vector<Lay> lay(n);
  size_t off = 0;
  for (size_t i = 0; i < n; ++i) {
    const InstSpec& s = specs[i];
    if (s.npad % sblock) return XRTK_EINVAL;
    const bool f32 = s.mode == MODE_F32R;
    const LaunchCfg& cfg = f32 ? cfg32 : cfg64;
    Lay& L = lay[i];
    L.chunk_len = chunk_len_for(ctx, cfg, s.nr, s.npad, split_override);
    L.n_chunks = s.nr ? (s.nr + L.chunk_len - 1) / L.chunk_len : 0;
    L.rays32 = off; off = align_up(off + (f32 ? sizeof(RayRec32) * s.nr : 0), 256);
    L.pix32 = off; off = align_up(off + (f32 ? sizeof(PixRec32) * s.npad : 0), 256);
    L.tiles = off;
    off = align_up(off + (f32 ? tile_table_bytes(s.npad, cfg32.P) : 0), 256);
    L.R64 = off; off = align_up(off + (f32 ? 2 * sizeof(double) * s.npad : 0), 256);
    L.partial = off;
    off = align_up(off + (f32 ? sizeof(float) : sizeof(double)) * 4 * (size_t)s.npad * L.n_chunks,
                   256);
    L.scr = off; off = align_up(off + sizeof(InstScratch), 256);
    L.part = off; off = align_up(off + sizeof(double) * 4 * kPartBlocks, 256);
  }
  hipError_t e = work.ensure(off);
  if (e != hipSuccess) return XRTK_ENOMEM;
  char* W = work.as<char>();

  // ---- instance table
  std::vector<KInst> insts(n);
  uint32_t n_f64 = 0;
  size_t cap32 = 0, cap64 = 0, capmf = 0;  // items of each kernel if every route were taken
  std::vector<uint8_t> mf_ok(n, 0);
  for (size_t i = 0; i < n; ++i) {
    const InstSpec& s = specs[i];
    const Lay& L = lay[i];
    const bool f32 = s.mode == MODE_F32R;
    KInst& I = insts[i];
    std::memset(&I, 0, sizeof(I));
    I.rays64 = s.rays;
    I.rays32 = reinterpret_cast<const RayRec32*>(W + L.rays32);
    I.pix64 = s.pix;
    I.pix32 = reinterpret_cast<const PixRec32*>(W + L.pix32);
    I.tiles = reinterpret_cast<const TileRec*>(W + L.tiles);
    I.perm = s.perm;
    I.R64 = reinterpret_cast<const double*>(W + L.R64);
    I.partial = W + L.partial;
    I.out = s.out;
    I.scr = reinterpret_cast<InstScratch*>(W + L.scr);
    I.nr = s.nr; I.np = s.np; I.npad = s.npad; I.n_chunks = L.n_chunks;
    I.chunk_len = L.chunk_len; I.mode = s.mode; I.tile = cfg32.tile;
    I.accumulate = s.accumulate;
    I.tile_slots = 64u * cfg32.P;
    I.energy = s.energy;
    I.k = s.energy / kCHBAR_H * 1e7;
    kr_dd(I.k, &I.kr_hi, &I.kr_lo);
    I.skip = s.skip;
    I.gx0 = s.gx0; I.gy = s.gy; I.gz0 = s.gz0; I.gsx = s.gsx; I.gsz = s.gsz;
    I.gnx = s.gnx; I.gnz = s.gnz; I.gth = s.gth; I.gtw = s.gtw;
    I.gex = s.gex; I.gez = s.gez; I.gemax = s.gemax;
    if (!f32) ++n_f64;
    const uint32_t B = f32 ? 64u * cfg32.W * cfg32.P : 64u * cfg64.W * cfg64.P;
    (f32 ? cap32 : cap64) += (size_t)L.n_chunks * (s.npad / B);
    mf_ok[i] = mfma_allowed(s, force_flags) ? 1u : 0u;
    if (mf_ok[i]) capmf += (size_t)L.n_chunks * ((s.gnx + 63u) / 64u) * ((s.gnz + 63u) / 64u);
  }
  // merge lists: instances writing their own output merge in one launch;
  // instances that accumulate into an earlier output follow one by one.
  std::vector<uint32_t> first, rest;
  for (size_t i = 0; i < n; ++i)
    if (specs[i].nr) (specs[i].accumulate ? rest : first).push_back((uint32_t)i);
  std::vector<uint32_t> lists(first);
  lists.insert(lists.end(), rest.begin(), rest.end());
  uint32_t first_np = 0;
  for (uint32_t i : first) first_np = std::max(first_np, specs[i].npad);

  // metadata: a fixed layout with room for the items of every route
  const size_t a_ins = align_up(sizeof(KInst) * std::max<size_t>(n, 1), 256);
  const size_t a_i32 = align_up(sizeof(WorkItem) * std::max<size_t>(1, cap32), 256);
  const size_t a_i64 = align_up(sizeof(WorkItem) * std::max<size_t>(1, cap64), 256);
  const size_t a_imf = align_up(sizeof(WorkItem) * std::max<size_t>(1, capmf), 256);
  const size_t a_lst = align_up(sizeof(uint32_t) * (lists.size() + 1), 256);
  e = meta.ensure(a_ins + a_i32 + a_i64 + 2 * a_imf + a_lst + 256);
  if (e != hipSuccess) return XRTK_ENOMEM;
  char* M = meta.as<char>();
  KInst* d_insts = reinterpret_cast<KInst*>(M);
  WorkItem* d_items32 = reinterpret_cast<WorkItem*>(M + a_ins);
  WorkItem* d_items64 = reinterpret_cast<WorkItem*>(M + a_ins + a_i32);
  // the MFMA kernels' items: the Es pass, then the Ep pass
  WorkItem* d_items_mf[2] = {reinterpret_cast<WorkItem*>(M + a_ins + a_i32 + a_i64),
                             reinterpret_cast<WorkItem*>(M + a_ins + a_i32 + a_i64 + a_imf)};
  uint32_t* d_lists = reinterpret_cast<uint32_t*>(M + a_ins + a_i32 + a_i64 + 2 * a_imf);
  uint32_t* d_queue =
      reinterpret_cast<uint32_t*>(M + a_ins + a_i32 + a_i64 + 2 * a_imf + a_lst);
  if (upload) {
    if (n) e = hipMemcpyAsync(d_insts, insts.data(), sizeof(KInst) * n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !lists.empty())
      e = hipMemcpyAsync(d_lists, lists.data(), sizeof(uint32_t) * lists.size(),
                         hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return map_err(e);
  }

  ph.next("xrtk.prep");
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
    e = launch_prep(s.rays, s.nr, s.pix, s.npad, s.perm, s.np,
                    reinterpret_cast<InstScratch*>(W + L.scr),
                    reinterpret_cast<double*>(W + L.part), s.mode, insts[i].kr_hi,
                    insts[i].kr_lo, s.skip, force_flags,
                    s.static_flags | (mf_ok[i] ? INST_F_MFMA_OK : 0u), cfg32.P, s.gsx, s.gsz,
                    s.gemax,
                    reinterpret_cast<RayRec32*>(W + L.rays32),
                    reinterpret_cast<PixRec32*>(W + L.pix32),
                    reinterpret_cast<TileRec*>(W + L.tiles), reinterpret_cast<double*>(W + L.R64),
                    st);
    if (e != hipSuccess) return map_err(e);
  }
  // ---- routing: the planner's flags of the instances that may run on the
  // matrix cores (given, or read back after the prep), so that each kernel
  // gets only the items it runs; without them both kernels get the items
  // and skip what the other takes
  std::vector<uint32_t> flags(n, 0u);
  bool have_flags = false;
  if (known_flags) {
    flags = *known_flags;
    have_flags = true;
  } else if (route_sync && capmf > 0) {
    ph.next("xrtk.route");
    e = hipStreamSynchronize(st);
    for (size_t i = 0; i < n && e == hipSuccess; ++i)
      if (mf_ok[i] && specs[i].nr)
        e = hipMemcpy(&flags[i], &reinterpret_cast<InstScratch*>(W + lay[i].scr)->flags,
                      sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return map_err(e);
    have_flags = true;
  }
  // ---- work items
  std::vector<WorkItem> items32, items64, items_mf[2];
  for (size_t i = 0; i < n; ++i) {
    const InstSpec& s = specs[i];
    const Lay& L = lay[i];
    const bool f32 = s.mode == MODE_F32R;
    const bool mf = mf_ok[i] && have_flags && (flags[i] & INST_F_MFMA);
    const bool valu_items = !mf;
    const bool mfma_items = mf_ok[i] && (!have_flags || mf);
    const uint32_t B = f32 ? 64u * cfg32.W * cfg32.P : 64u * cfg64.W * cfg64.P;
    std::vector<WorkItem>& items = f32 ? items32 : items64;
    if (valu_items)
      for (uint32_t c = 0; c < L.n_chunks; ++c)
        for (uint32_t p0 = 0; p0 < s.npad; p0 += B)
          items.push_back(WorkItem{(uint32_t)i, p0, c, 0});
    if (mfma_items) {
      // the Es pass, and the Ep pass unless every ray has Ep = 0 (without
      // routing both: an Ep item of such an instance returns at once)
      const uint32_t nb = ((s.gnx + 63u) / 64u) * ((s.gnz + 63u) / 64u);
      const uint32_t passes = (have_flags && (flags[i] & INST_F_EPZERO)) ? 1u : 2u;
      for (uint32_t pass = 0; pass < passes; ++pass)
        for (uint32_t c = 0; c < L.n_chunks; ++c)
          for (uint32_t b = 0; b < nb; ++b)
            items_mf[pass].push_back(WorkItem{(uint32_t)i, b, c, pass});
    }
  }
  // Batched launches: longest items first so the queue drains evenly.
  if (sorted && n > 1) {
    auto by_len = [&](const WorkItem& a, const WorkItem& b) {
      const KInst& A = insts[a.inst];
      const KInst& Bq = insts[b.inst];
      const uint64_t la = std::min<uint64_t>(A.chunk_len, A.nr - (uint64_t)a.chunk * A.chunk_len);
      const uint64_t lb = std::min<uint64_t>(Bq.chunk_len, Bq.nr - (uint64_t)b.chunk * Bq.chunk_len);
      return la > lb;
    };
    std::stable_sort(items32.begin(), items32.end(), by_len);
    std::stable_sort(items64.begin(), items64.end(), by_len);
    for (auto& v : items_mf) std::stable_sort(v.begin(), v.end(), by_len);
  }
  if (upload) {
    if (!items32.empty())
      e = hipMemcpyAsync(d_items32, items32.data(), sizeof(WorkItem) * items32.size(),
                         hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !items64.empty())
      e = hipMemcpyAsync(d_items64, items64.data(), sizeof(WorkItem) * items64.size(),
                         hipMemcpyHostToDevice, st);
    for (int k = 0; k < 2 && e == hipSuccess; ++k)
      if (!items_mf[k].empty())
        e = hipMemcpyAsync(d_items_mf[k], items_mf[k].data(), sizeof(WorkItem) * items_mf[k].size(),
                           hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return map_err(e);
  }
  // ---- pair kernels
  auto grid_of = [&](const LaunchCfg& cfg, uint32_t mode, size_t n_items) {
    uint32_t g = (uint32_t)n_items;
    if (cfg.persistent) {
      const uint32_t resident = (uint32_t)ctx->n_cu * (mode == MODE_F32R ? 32u : 16u) / cfg.W;
      g = std::max<uint32_t>(1, std::min<uint32_t>(g, resident));
    }
    return g;
  };
  ph.next("xrtk.pair");
  // (external records: a captured graph records them at each replay)
  if (timed) (void)hipEventRecordWithFlags(ctx->ev[1], st, hipEventRecordExternal);
  if (!items32.empty()) {
    const uint32_t remap = (!cfg32.persistent && n == 1) ? 1u : 0u;
    e = launch_pair(MODE_F32R, cfg32, d_insts, d_items32, (uint32_t)items32.size(), d_queue,
                    grid_of(cfg32, MODE_F32R, items32.size()), remap, st);
    if (e != hipSuccess) return map_err(e);
  }
  bool offsets = false;  // float32 pixel lists on a grid
  for (size_t i = 0; i < n; ++i) offsets = offsets || (mf_ok[i] && specs[i].gex);
  for (int k = 0; k < 2; ++k) {
    if (items_mf[k].empty()) continue;
    // one wave per group, up to two groups per SIMD resident
    const uint32_t g = std::min<uint32_t>((uint32_t)items_mf[k].size(), (uint32_t)ctx->n_cu * 8u);
    e = launch_mfma(d_insts, d_items_mf[k], (uint32_t)items_mf[k].size(), d_queue + 32 + k, g,
                    k == 1, offsets, st);
    if (e != hipSuccess) return map_err(e);
  }
  if (!items64.empty()) {
    const uint32_t remap = (!cfg64.persistent && n == 1) ? 1u : 0u;
    e = launch_pair(MODE_F64, cfg64, d_insts, d_items64, (uint32_t)items64.size(), d_queue + 16,
                    grid_of(cfg64, MODE_F64, items64.size()), remap, st);
    if (e != hipSuccess) return map_err(e);
  }
  if (timed) (void)hipEventRecordWithFlags(ctx->ev[2], st, hipEventRecordExternal);
  // ---- merge
  ph.next("xrtk.merge");
  if (!first.empty()) {
    e = launch_merge(d_insts, d_lists, (uint32_t)first.size(), first_np, st);
    if (e != hipSuccess) return map_err(e);
  }
  for (size_t j = 0; j < rest.size(); ++j) {
    e = launch_merge(d_insts, d_lists + first.size() + j, 1, specs[rest[j]].npad, st);
    if (e != hipSuccess) return map_err(e);
  }
  if (timed) (void)hipEventRecord(ctx->ev[3], st);
  if (plan_out) {
    plan_out->cfg32 = cfg32;
    plan_out->cfg64 = cfg64;
    plan_out->n_items32 = (uint32_t)items32.size();
    plan_out->n_items64 = (uint32_t)items64.size();
    plan_out->n_items_mfma = (uint32_t)(items_mf[0].size() + items_mf[1].size());
    plan_out->n_f64 = n_f64;
    plan_out->n_inst = (uint32_t)n;
    plan_out->scr0 = n ? reinterpret_cast<const InstScratch*>(W + lay[0].scr) : nullptr;
    plan_out->routed = have_flags;
    plan_out->flags = flags;
  }
  return XRTK_OK;
}

}  // namespace xrtk
