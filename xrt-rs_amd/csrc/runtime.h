// Host runtime of the Kirchhoff engine: context, HBM workspace, planner and
// the solve pipeline shared by the one-shot, batched and serving entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
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
hipError_t launch_pix_from_f32aos(const void* src, PixF64* dst, uint32_t n, hipStream_t st);
hipError_t launch_pix_from_cols(const double* cols, PixF64* dst, uint32_t n, hipStream_t st);
hipError_t launch_pix_grid(const double c[3], double dx, double dz, uint32_t nx, uint32_t nz,
                           PixF64* dst, hipStream_t st);
hipError_t launch_prep(const RayRec64* rays, uint32_t nr, const PixF64* pix, uint32_t np,
                       InstScratch* scr, double* part, uint32_t mode, double kr_hi, double kr_lo,
                       double skip, double qfold_rev, uint32_t force_flags, uint32_t static_flags,
                       RayRec32* rays32, PixRec32* pix32, double* R64, hipStream_t st);
hipError_t launch_pair(uint32_t mode, const LaunchCfg& cfg, const KInst* insts,
                       const WorkItem* items, uint32_t n_items, uint32_t* queue, uint32_t grid,
                       uint32_t remap, hipStream_t st);
bool pair_cfg_supported(uint32_t mode, const LaunchCfg& c);
hipError_t launch_merge(const KInst* insts, const uint32_t* list, uint32_t n_list,
                        uint32_t max_np, hipStream_t st);
hipError_t launch_zero_f64(double* p, size_t n, hipStream_t st);

constexpr double kCHBAR_H = 1973.2697177417986;  // src/core/consts.rs:74
constexpr int kPartBlocks = 240;                 // centroid partial blocks (device.hip)

// Grow-only device buffer.
struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes);
  void release();
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// A resident ray set.
struct RaySet {
  DBuf rays;          // RayRec64[nr]
  uint32_t nr = 0;
  bool has_path = false;
  bool mono = true;   // all energies equal
  double energy = 0;  // the common energy when mono
};

// Description of one instance for a launch (all pointers are device memory).
struct InstSpec {
  const RayRec64* rays = nullptr;
  uint32_t nr = 0;
  const PixF64* pix = nullptr;
  uint32_t np = 0;
  double energy = 0;       // F32R: the instance energy; F64: > 0 overrides
  double skip = 1e-30;     // skip pairs with path below this [mm]
  uint32_t static_flags = 0;
  double* out = nullptr;   // np * 4 doubles
  uint32_t accumulate = 0; // add into out (energy groups after the first)
};

struct Plan {
  uint32_t mode = MODE_F32R;
  LaunchCfg cfg;
  uint32_t n_items = 0;
  uint32_t grid = 0;
  uint32_t remap = 0;
  int checked = 0;
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
  xrtk::DBuf stage, rays64, pix64, work, meta, outbuf;
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
    xrtk::DBuf work, meta, pix, outd;
    std::vector<std::pair<double*, size_t>> outs;  // device out and np of each job
  };
  std::vector<GraphEntry*> graphs;
};

namespace xrtk {

// Chooses the launch configuration for a solve of nr rays on np pixels.
LaunchCfg choose_cfg(xrtk_ctx* ctx, uint32_t mode, uint64_t nr, uint64_t np, const XrtkOpts* o);
// Ray chunk length of one instance (a multiple of cfg.tile), intrinsic to the
// instance shape and the launch configuration.
uint32_t chunk_len_for(const xrtk_ctx* ctx, uint32_t mode, const LaunchCfg& cfg, uint32_t nr,
                       uint32_t np, int split_override);

// Runs prep, pair and merge for `specs` on ctx->stream (no host sync, no
// host copies).  work/meta are grown as needed.  When `timed`, HIP events
// bracket the pair kernel and the whole device pipeline.
// With upload == false the metadata copies are skipped (they were made by an
// earlier call with the same specs and buffers), so the call only enqueues
// kernels and can be captured into a hipGraph.
int run_instances(xrtk_ctx* ctx, uint32_t mode, const LaunchCfg& cfg,
                  const std::vector<InstSpec>& specs, DBuf& work, DBuf& meta, bool timed,
                  Plan* plan_out, int split_override, uint32_t force_flags, bool persistent_sorted,
                  bool upload = true);

int map_err(hipError_t e);
double now_s();

}  // namespace xrtk
