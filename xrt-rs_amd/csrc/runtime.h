// Host runtime of the Kirchhoff engine: context, HBM workspace, planner and
// the solve pipeline shared by the one-shot, batched and serving entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include "host_copy.h"

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/xrtk.h"
#include "internal.h"

namespace xrtk {

// ---- device launchers (device.hip) ----
hipError_t launch_rays_from_f32aos(const float* src, RayRec64* dst, uint32_t n, hipStream_t st);
hipError_t launch_rays_from_cols(const double* cols, RayRec64* dst, uint32_t n, hipStream_t st);
// cols: x y z a b c e [es_re es_im ep_re ep_im] [path] of n_beam beam rays;
// writes the n_good rays good[0..n_good) (beam_to_diffraction_rays).
hipError_t launch_rays_from_beam(const double* cols, uint32_t n_beam, const double* normals,
                                 const uint64_t* good, uint32_t n_good, bool has_amps,
                                 bool has_path, RayRec64* dst, hipStream_t st);
// Explicit pixel lists: slots in tiles of 64 P, each lane holding P
// consecutive list entries; perm[slot] = entry or ~0u (padding).
hipError_t launch_pix_from_f32aos(const void* src, PixF64* dst, uint32_t* perm, uint32_t n,
                                  uint32_t npad, uint32_t P, hipStream_t st);
hipError_t launch_pix_from_cols(const double* cols, PixF64* dst, uint32_t* perm, uint32_t n,
                                uint32_t npad, uint32_t P, hipStream_t st);
hipError_t launch_pix_grid(const double c[3], double dx, double dz, uint32_t nx, uint32_t nz,
                           uint32_t th, uint32_t tw, uint32_t npad, PixF64* dst, uint32_t* perm,
                           hipStream_t st);
// k_pix_grid's slots for a float32 pixel list (XrtkGpuPixel) on an nx x nz
// grid in the prepare_wave order, each slot holding the list's own point.
hipError_t launch_pix_grid_list(const void* list, uint32_t nx, uint32_t nz, uint32_t th,
                                uint32_t tw, uint32_t npad, PixF64* dst, uint32_t* perm,
                                hipStream_t st);
// perm: slot -> pixel index (~0u for padding), or null when slots [np, npad)
// are the padding.
hipError_t launch_prep(const RayRec64* rays, uint32_t nr, const PixF64* pix, uint32_t npad,
                       const uint32_t* perm, uint32_t np, InstScratch* scr, double* part,
                       uint32_t mode, double kr, double kr_lo, double skip, uint32_t force_flags,
                       uint32_t static_flags, int P, double gsx, double gsz, double gemax,
                       RayRec32* rays32,
                       PixRec32* pix32, TileRec* tiles, double* R64, hipStream_t st);
hipError_t launch_pair(uint32_t mode, const LaunchCfg& cfg, const KInst* insts,
                       const WorkItem* items, uint32_t n_items, uint32_t* queue, uint32_t grid,
                       uint32_t remap, hipStream_t st);
bool pair_cfg_supported(uint32_t mode, const LaunchCfg& c);
// Block items (inst, block, chunk) of grid instances on the matrix cores,
// persistent groups of one wave on queue[0]; items of instances the planner
// did not give the MFMA flag are skipped.  ep_pass: the kernel of the Ep
// sums (items with pad = 1), else that of the Es sums (pad = 0).  offsets:
// the kernel that applies the grid offsets of float32 pixel lists (KInst
// gex, gez; instances without them run on it as well).
hipError_t launch_mfma(const KInst* insts, const WorkItem* items, uint32_t n_items,
                       uint32_t* queue, uint32_t grid, bool ep_pass, bool offsets,
                       hipStream_t st);
// Pair-kernel launches still to be failed on purpose (xrtk_debug_fail_launches):
// launch_pair takes one and returns hipErrorLaunchFailure before enqueuing
// anything.
extern std::atomic<uint32_t> g_fail_launches;
hipError_t launch_merge(const KInst* insts, const uint32_t* list, uint32_t n_list,
                        uint32_t max_npad, hipStream_t st);
hipError_t launch_zero_f64(double* p, size_t n, hipStream_t st);

constexpr double kCHBAR_H = 1973.2697177417986;  // src/core/consts.rs:74
constexpr int kPartBlocks = 240;                 // centroid partial blocks (device.hip)
// Reduced-phase bound above which an instance runs on the FP64 kernel: the
// FP32 residual phase would carry more than ~1e-4 rad of rounding per pair.
constexpr double kMaxResidualRev = 512.0;

// A fixed block of device memory handed out front to back (256-byte
// aligned); nothing is ever allocated or freed through it.
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
  size_t off = 0;
};

// Grow-only device buffer.  With `arena` set it is carved once from that
// arena instead (ensure fails with hipErrorOutOfMemory when the arena is
// short, release does nothing).
struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
  Arena* arena = nullptr;
  hipError_t ensure(size_t bytes);
  void release();
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// Grow-only page-locked host buffer (hipHostMalloc).
struct HBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes);
  void release();
};

// Host memory to device memory on `st`.  With `pin` (a page-locked block,
// or null) a large copy is gathered into the block by a few host threads
// in 8 MiB chunks, each chunk's DMA queued as soon as it is in, about twice
// the rate of a copy from pageable memory; the block must stay as it is
// until the stream has run the copy.
hipError_t copy_h2d(hipStream_t st, char* pin, void* dst, const void* src, size_t bytes);
// Device memory to host memory on `st`, synchronously: the bytes of `src`
// (the segments back to back) to the segments' destinations.  With `pin` (a
// page-locked block at least as large, or null) a large copy lands in the
// block in 4 MiB chunks and a few host threads copy each chunk out as soon
// as it has arrived.
hipError_t copy_d2h(hipStream_t st, char* pin, const void* src, const Segment* segs,
                    size_t n_segs);

// Device buffers of one-shot solves: staged inputs, resident-form rays and
// slots, per-instance work (FP32 records, tiles, partial sums), launch
// metadata and the merged output; and the page-locked host block the inputs
// pass through on their way to `stage` (unused by an arena workspace, which
// allocates nothing).
struct Workspace {
  DBuf stage, rays64, pix64, perm, work, meta, outbuf;
  HBuf hpin;
  void release() {
    for (DBuf* b : {&stage, &rays64, &pix64, &perm, &work, &meta, &outbuf}) b->release();
    hpin.release();
  }
  void use_arena(Arena* a) {
    for (DBuf* b : {&stage, &rays64, &pix64, &perm, &work, &meta, &outbuf}) b->arena = a;
  }
};

// Axis-aligned bounding box.
struct Box {
  double lo[3] = {1e300, 1e300, 1e300};
  double hi[3] = {-1e300, -1e300, -1e300};
  void add(double x, double y, double z) {
    const double v[3] = {x, y, z};
    for (int i = 0; i < 3; ++i) {
      lo[i] = v[i] < lo[i] ? v[i] : lo[i];
      hi[i] = v[i] > hi[i] ? v[i] : hi[i];
    }
  }
  bool empty() const { return lo[0] > hi[0]; }
};

// A resident ray set.
struct RaySet {
  DBuf rays;          // RayRec64[nr]
  uint32_t nr = 0;
  bool has_path = false;
  bool mono = true;   // all energies equal
  double energy = 0;  // the common energy when mono
  Box box;            // positions
  std::shared_ptr<std::atomic<int>> users = std::make_shared<std::atomic<int>>(0);
};

// Description of one instance for a launch (all pointers are device memory).
struct InstSpec {
  uint32_t mode = MODE_F32R;
  const RayRec64* rays = nullptr;
  uint32_t nr = 0;
  const PixF64* pix = nullptr;   // npad slots
  const uint32_t* perm = nullptr;  // slot -> output index, or null (identity below np)
  uint32_t np = 0, npad = 0;
  double energy = 0;       // F32R: the instance energy; F64: > 0 overrides
  double skip = 1e-30;     // skip pairs with path below this [mm]
  uint32_t static_flags = 0;
  double* out = nullptr;   // np * 4 doubles
  uint32_t accumulate = 0; // add into out (energy groups after the first)
  // a prepare_wave grid screen laid out in tiles of gth x gtw pixels (gnx > 0):
  // such an instance may run on the matrix cores
  double gx0 = 0, gy = 0, gz0 = 0, gsx = 0, gsz = 0;
  uint32_t gnx = 0, gnz = 0, gth = 0, gtw = 0;
  // a float32 pixel list on that grid: device offsets of its rows (gnx) and
  // columns (gnz) from it, in pixels, at most gemax in magnitude
  const float* gex = nullptr;
  const float* gez = nullptr;
  double gemax = 0;
};
// Records the grid of a prepare_wave screen (src/waves/prepare_wave.rs:40-68:
// x = cx + (ix/(nx-1) - 1/2) 2 dx, likewise z; a dimension of 1 sits at the
// centre) in an instance spec.
void spec_set_grid(InstSpec& s, const XrtkWaveParams& w, uint32_t th, uint32_t tw);
// Whether an instance may run on the matrix cores: a reduced-kernel instance
// with rays on a grid screen, unless force flag 0x20000 (XRTK_ENGINE_VALU)
// forbids it.  The planner kernel decides from the instance's bounds.
bool mfma_allowed(const InstSpec& s, uint32_t force_flags);

struct Plan {
  LaunchCfg cfg32, cfg64;
  uint32_t n_items32 = 0, n_items64 = 0, n_items_mfma = 0;  // (MFMA: both passes)
  uint32_t n_f64 = 0;     // instances on the FP64 kernel
  uint32_t n_inst = 0;
  const InstScratch* scr0 = nullptr;  // device scratch of the first instance
  // routing: whether the items followed the planner's flags, and those flags
  // (per instance; 0 where no instance could run on the matrix cores)
  bool routed = false;
  std::vector<uint32_t> flags;
};

}  // namespace xrtk

struct xrtk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::mutex mu;
  std::string name;
  int n_cu = 256;
  bool shared = false;
  // one-shot workspace
  xrtk::Workspace ws;
  // this device's fixed scratch block (xrtk_kirchhoff_gpu_f32_noalloc)
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // resident ray sets
  std::map<uint64_t, xrtk::RaySet> raysets;
  uint64_t next_rs = 1;
  // autotune table: (mode, log2 nr, log2 np) -> cfg
  std::map<std::tuple<int, int, int>, xrtk::LaunchCfg> tune;
  // last solve
  XrtkPlan last_plan{};
  double last_pair_ms = 0, last_solve_ms = 0;
  // graph cache of xrtk_solve_jobs
  struct GraphEntry {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    xrtk::DBuf work, meta, pix, perm, outd;
    std::vector<std::pair<double*, size_t>> outs;  // device out and np of each job
    xrtk::Plan plan;
    // what an eager replay re-enqueues when there is no graph (XRTK_NO_GRAPH,
    // or a capture that failed): the instances and each job's screen
    std::vector<xrtk::InstSpec> specs;
    std::vector<XrtkJob> jobs;
  };
  std::vector<GraphEntry*> graphs;
  // Fresnel propagation workspace
  xrtk::DBuf fresnel;
  // undulator maps
  xrtk::DBuf undulator;
};

namespace xrtk {

// Chooses the launch configuration for a solve of nr rays on np pixels
// (f32_pixels: a float32 pixel list, which sits off the slot line by its
// rounding).
LaunchCfg choose_cfg(xrtk_ctx* ctx, uint32_t mode, uint64_t nr, uint64_t np, const XrtkOpts* o,
                     bool f32_pixels = false);
// Whether the pixels per lane of such a solve come from the caller's options
// or the context's tuning table (not the default).
bool cfg_shape_given(const xrtk_ctx* ctx, uint32_t mode, uint64_t nr, uint64_t np,
                     const XrtkOpts* o);
// Default FP64 configuration paired with a given F32R configuration.
LaunchCfg cfg64_for(const LaunchCfg& c32);
// Tune-table key (both log2 fields) of the batched-solve configuration.
constexpr int kBatchKey = -1;
// Launch configuration of batched solves (xrtk_solve_jobs and the server):
// the context's tuned batch entry, or the default; `o` overrides fields.
LaunchCfg batch_cfg(const xrtk_ctx* ctx, uint32_t mode, const XrtkOpts* o);
// Ray chunk length of one instance (a multiple of cfg.tile), intrinsic to the
// instance shape and the launch configuration.
uint32_t chunk_len_for(const xrtk_ctx* ctx, const LaunchCfg& cfg, uint32_t nr, uint32_t npad,
                       int split_override);
// Slot count of np pixels: a multiple of both launches' slot blocks.
uint32_t padded_slots(uint32_t np, const LaunchCfg& c32, const LaunchCfg& c64);
// Tile shape (th rows of ix, tw columns of iz) of 64 * P slots over an
// nx x nz prepare_wave grid.  tw divides 64, so a lane's P slots lie in one
// column, 64 / tw rows apart: 16 (P < 8) or 32 columns, or twice as many
// (half as tall, slots half as far apart) when that pads the grid less.
void grid_tile(int P, uint32_t nx, uint32_t nz, uint32_t* th, uint32_t* tw);
// Slots of a tiled nx x nz grid.
uint32_t grid_slots(uint32_t nx, uint32_t nz, const LaunchCfg& c32, const LaunchCfg& c64);

// Host-side routing: the bound of the tile-referenced residual phase for rays
// in `rays` and pixels in `pix`, with in-tile half-extent `dq` per axis.
// Returns +inf when a ray may lie on a pixel.
double residual_bound_rev(const Box& rays, const Box& pix, const double dq[3], double energy);
// Mode of an instance requested as `mode`: F32R unless the bound says FP64.
uint32_t route_mode(uint32_t mode, const Box& rays, const Box& pix, const double dq[3],
                    double energy);

// Runs prep, pair and merge for `specs` on ctx->stream (no host copies and
// no host sync, except the routing read below).  work/meta are grown as
// needed.  When `timed`, HIP events bracket the pair kernels and the whole
// device pipeline.  With upload == false the metadata copies are skipped
// (made by an earlier call with the same specs, flags and buffers), so the
// call only enqueues kernels and can be captured into a hipGraph.
// Routing between the VALU and MFMA pair kernels: with `known_flags` (the
// planner flags an earlier call with the same specs returned in
// Plan::flags), or with `route_sync` (the stream is synchronised after the
// prep kernels and the flags read back), each kernel gets only its own
// items; otherwise both get the items of a grid instance and the device
// flag decides.
int run_instances(xrtk_ctx* ctx, const LaunchCfg& cfg32, const LaunchCfg& cfg64,
                  const std::vector<InstSpec>& specs, DBuf& work, DBuf& meta, bool timed,
                  Plan* plan_out, int split_override, uint32_t force_flags, bool sorted,
                  bool upload = true, const std::vector<uint32_t>* known_flags = nullptr,
                  bool route_sync = false);

int map_err(hipError_t e);
// kirchhoff_gpu's one-shot solve with its device buffers in `ws` (the
// context's own workspace, or one over a fixed arena); takes the context
// lock.
int kirchhoff_f32_ws(xrtk_ctx* ctx, Workspace& ws, const XrtkGpuRay* rays, size_t n_rays,
                     const XrtkGpuPixel* pixels, size_t n_pixels, double* out4,
                     const XrtkOpts* opts);
// The device address and size of device `device`'s fixed scratch block,
// looked up once per device (noalloc.hip); nullptr when the lookup failed.
void* noalloc_scratch(int device, size_t* bytes);
double now_s();
// Releases the Fresnel workspace of a context.
void fresnel_release(xrtk_ctx* ctx);

// A roctx range over one host phase of a solve ("xrtk.upload",
// "xrtk.pair", ...), shown by rocprofv3 --marker-trace beside the kernel
// trace.  Ranges mark host time: a range around an enqueue covers the
// launch, the range around a readback includes the wait for the kernels
// queued before it.  Without a profiler attached the push and pop return at
// once.
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};
// Consecutive phases of one function: next() closes the open range and
// opens the next one; the last closes with the object.
struct Phases {
  bool open = false;
  void next(const char* name) {
    if (open) roctxRangePop();
    roctxRangePushA(name);
    open = true;
  }
  ~Phases() {
    if (open) roctxRangePop();
  }
};

}  // namespace xrtk
