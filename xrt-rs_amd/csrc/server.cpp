// Continuous-batching solver.
//
// A scheduler thread owns one HIP stream.  Requests (a resident ray set at
// one energy on one prepare_wave screen) queue on submit.  Each scheduler
// iteration
//   1. admits queued requests while the HBM budget and the active-request
//      cap allow, the one with the fewest pairs first (oldest first under
//      XRTK_SCHED_ARRIVAL), so that under overload small requests pass large
//      queued ones; their buffers come from a caching pool, their grid and
//      prep kernels are enqueued at once,
//   2. fills one "quantum" launch of the persistent pair kernel with work
//      items of the active requests, shortest remaining work first (the
//      request with the fewest pairs left fills the budget first; or oldest
//      first, XRTK_SCHED_ARRIVAL), and then one item of every active request
//      not yet in the launch, so a request that arrives while others run
//      joins at the next quantum, small requests are not stuck behind large
//      ones and every request advances in every quantum.  The quantum is a
//      whole number of rounds of full-size items over the resident
//      workgroups, and its items run longest first, so the workgroups end
//      the launch together,
//      A request on a grid screen learns after its prep whether the device
//      planner put it on the matrix cores (an async copy of its flags); its
//      items are the MFMA kernel's 64 x 64-pixel blocks then, and it joins
//      the quanta once the flags have landed,
//   3. enqueues, for each request whose last items went into that launch,
//      its FP64 merge and an async copy of the result into pinned memory,
//   4. retires requests whose copy has completed (result handed to the
//      caller's buffer, timestamps recorded, waiters woken); it also does so,
//      and admits, while it polls for the older launch to end.
// At most two quanta are in flight, so the host prepares the next launch
// while the device runs the current one.  A request's plan (chunking, slot
// layout, kernel shape) depends only on the request, so its result does not
// depend on which other requests share its launches.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "runtime.h"

using namespace xrtk;

namespace xrtk {
int build_job_specs(xrtk_ctx* ctx, uint32_t want, const LaunchCfg& c32, const LaunchCfg& c64,
                    const XrtkJob* jobs, size_t n, DBuf& pixbuf, DBuf& permbuf, DBuf& outbuf,
                    std::vector<InstSpec>& specs);
}

namespace {

constexpr size_t kAlign = 256;
size_t au(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Caching pool for device blocks or pinned host blocks.  Sizes round up to
// one of 8 steps per power of two (<= 12.5% slack); a cached block up to 2x
// the request is reused.  `allocated` counts every block held, cached or not.
struct Pool {
  bool pinned = false;
  std::multimap<size_t, void*> free_;
  size_t allocated = 0;
  static size_t bucket(size_t b) {
    if (b <= ((size_t)1 << 20)) return (size_t)1 << 20;
    size_t p = 1;
    while (p < b) p <<= 1;
    const size_t step = p / 16;
    return (b + step - 1) / step * step;
  }
  void drop(std::multimap<size_t, void*>::iterator it) {
    if (pinned) (void)hipHostFree(it->second);
    else (void)hipFree(it->second);
    allocated -= it->first;
    free_.erase(it);
  }
  // nullptr when the block cannot be had within `limit` bytes in total
  void* get(size_t bytes, size_t limit, size_t* got) {
    const size_t b = bucket(bytes);
    auto it = free_.lower_bound(b);
    if (it != free_.end() && it->first <= 2 * b) {
      void* p = it->second;
      *got = it->first;
      free_.erase(it);
      return p;
    }
    while (!free_.empty() && allocated + b > limit) drop(std::prev(free_.end()));
    if (allocated + b > limit) return nullptr;
    void* p = nullptr;
    hipError_t e = pinned ? hipHostMalloc(&p, b, hipHostMallocDefault) : hipMalloc(&p, b);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    allocated += b;
    *got = b;
    return p;
  }
  void put(void* p, size_t b) { free_.emplace(b, p); }
  void release_all() {
    while (!free_.empty()) drop(free_.begin());
  }
};

enum ReqState { QUEUED, ACTIVE, MERGING, DONE, FAILED };

struct Req {
  uint64_t id = 0;
  XrtkJob job{};
  double t[4] = {0, 0, 0, 0};
  int state = QUEUED;
  int rc = XRTK_OK;
  // plan
  InstSpec spec;
  LaunchCfg cfg;
  uint32_t chunk_len = 0, n_chunks = 0, pblocks = 0;
  uint64_t n_items = 0, next_item = 0;
  double item_pairs = 0;  // pairs of one full item
  size_t bytes = 0;
  // device block and its carve-up
  void* dblk = nullptr;
  size_t dblk_size = 0;
  void* hblk = nullptr;  // pinned result staging
  size_t hblk_size = 0;
  KInst kinst{};
  KInst* d_kinst = nullptr;
  uint32_t* d_list = nullptr;
  hipEvent_t done_ev = nullptr;
  // routing: a request that may run on the matrix cores waits for its
  // planner flags (copied into hblk, recorded by prep_ev) before its items
  // are fixed
  hipEvent_t prep_ev = nullptr;
  bool routed = true;
  uint32_t mblocks = 0;  // 64 x 64-pixel blocks of the screen (MFMA items)
  uint32_t mpasses = 1;  // MFMA sums: Es, and Ep unless every ray has Ep = 0
  bool mfma = false;
  std::shared_ptr<std::atomic<int>> users;  // pending requests on the ray set
};

struct Launch {
  hipEvent_t start = nullptr, end = nullptr;
  bool in_flight = false;
  size_t n_req = 0;
  double pairs = 0;
  void* meta = nullptr;  // device: KInst table + items + queue
  void* hmeta = nullptr; // pinned host staging of the same bytes
  size_t meta_size = 0;
  std::vector<KInst> h_insts;
  std::vector<WorkItem> h_items32, h_items64, h_items_mf[2];
};

}  // namespace

// Nominal pair rates of one resident wave of the pair kernels on gfx950
// (32 waves per CU for F32R, 16 for FP64); they size the server's work items.
static constexpr double kWavePairsF32 = 2.9e8, kWavePairsF64 = 3.2e7;
// Most ray chunks (partial sums per slot) the item cap may give a request.
static constexpr uint64_t kMaxChunks = 512;

struct xrtk_server {
  xrtk_ctx* ctx = nullptr;
  XrtkServerCfg cfg{};
  LaunchCfg c32, c64;
  // Longest ray chunk of a work item (0: the configuration's own), so that
  // one item takes at most half a quantum on one wave.  It depends only on
  // the configuration and quantum_ms, never on the other requests.
  uint32_t cap32 = 0, cap64 = 0;
  hipStream_t stream = nullptr;
  // grid and prep of requests that may run on the matrix cores: kept off the
  // quantum stream so their planner flags land while a quantum runs (the
  // quanta that take their items are enqueued after the flags are seen)
  hipStream_t pstream = nullptr;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  uint64_t next_id = 1;
  // submitted, not admitted, in admission order: key (planned pairs, id)
  // under shortest-remaining-first, (0, id) in arrival order
  std::map<std::pair<double, uint64_t>, Req*> queue;
  std::vector<Req*> active;             // admitted, items left or merging
  std::map<uint64_t, Req*> all;         // every request not yet collected
  Pool dpool, hpool;
  size_t hbm_in_use = 0, hbm_peak = 0;
  XrtkServerStats st{};
  Launch launches[2];
  int cur = 0;
  double rate_pairs_per_s = 1.2e12;     // refined from measured quanta

  // ---- scheduler helpers (called with `mu` held unless stated) ----
  bool plan_request(Req* r) {
    std::vector<InstSpec> specs;
    // build the spec's pixel/perm/out layout sizes only (no allocation)
    const XrtkJob& j = r->job;
    const size_t np64 = (size_t)j.screen.nx * j.screen.nz;
    if (np64 > 0x7fff0000ull) return false;  // the slot and merge indices are 32-bit
    const uint32_t np = (uint32_t)np64;
    auto it = ctx->raysets.find(j.rays);
    if (it == ctx->raysets.end()) return false;
    const RaySet& rs = it->second;
    const double e = j.energy > 0 ? j.energy : rs.energy;
    if (cfg.mode == XRTK_MODE_F32R && !(e > 0)) return false;
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
    InstSpec& s = r->spec;
    s.mode = route_mode(cfg.mode == XRTK_MODE_F64 ? MODE_F64 : MODE_F32R, rs.box, pb, dq, e);
    s.rays = rs.rays.as<RayRec64>();
    s.nr = rs.nr;
    s.np = np;
    s.npad = np ? grid_slots(j.screen.nx, j.screen.nz, c32, c64) : 0;
    s.energy = s.mode == MODE_F32R ? e : (j.energy > 0 ? j.energy : 0.0);
    s.skip = 1e-30;
    s.static_flags = rs.has_path ? INST_F_PATHUP : 0u;
    r->cfg = s.mode == MODE_F32R ? c32 : c64;
    r->chunk_len = chunk_len_for(ctx, r->cfg, s.nr, s.npad, 0);
    const uint32_t cap = s.mode == MODE_F32R ? cap32 : cap64;
    if (cap && r->chunk_len > cap) {
      // short items, but at most kMaxChunks partial sums per slot (a large
      // request then runs longer items rather than holding GBs of partials)
      const uint64_t floor_len = ((uint64_t)s.nr + kMaxChunks - 1) / kMaxChunks;
      const uint32_t min_len = (uint32_t)((floor_len + 63) / 64 * 64);
      r->chunk_len = std::min(r->chunk_len, std::max(cap, min_len));
    }
    r->n_chunks = s.nr ? (s.nr + r->chunk_len - 1) / r->chunk_len : 0;
    const uint32_t B = 64u * r->cfg.W * r->cfg.P;
    r->pblocks = s.npad / B;
    r->n_items = (uint64_t)r->n_chunks * r->pblocks;
    r->item_pairs = (double)B * r->chunk_len;
    if (np) spec_set_grid(s, j.screen, th, tw);
    r->mblocks = ((j.screen.nx + 63u) / 64u) * ((j.screen.nz + 63u) / 64u);
    r->routed = !mfma_allowed(s, 0u);
    // device block: pix64, perm, rays32, pix32, tiles, R64, partial, scr,
    // part, out, KInst, list
    const bool f32 = s.mode == MODE_F32R;
    size_t b = 0;
    b += au(sizeof(PixF64) * s.npad) + au(sizeof(uint32_t) * s.npad);
    b += au(f32 ? sizeof(RayRec32) * s.nr : 0) + au(f32 ? sizeof(PixRec32) * s.npad : 0);
    b += au(f32 ? tile_table_bytes(s.npad, c32.P) : 0);
    b += au(f32 ? 2 * sizeof(double) * s.npad : 0);
    b += au((f32 ? 16 : 32) * (size_t)s.npad * r->n_chunks);
    b += au(sizeof(InstScratch)) + au(sizeof(double) * 4 * kPartBlocks);
    b += au(sizeof(double) * 4 * (size_t)np) + au(sizeof(KInst)) + au(64);
    r->bytes = b;
    return true;
  }

  // Carves the block, fills the KInst, enqueues grid + prep.
  int start_request(Req* r) {
    Range range("xrtk.server.start");  // slots, upload and prep of one request
    hipStream_t st = r->routed ? stream : pstream;
    char* p = static_cast<char*>(r->dblk);
    InstSpec& s = r->spec;
    const bool f32 = s.mode == MODE_F32R;
    auto take = [&](size_t bytes) { char* q = p; p += au(bytes); return q; };
    PixF64* pix = reinterpret_cast<PixF64*>(take(sizeof(PixF64) * s.npad));
    uint32_t* perm = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * s.npad));
    RayRec32* rays32 = reinterpret_cast<RayRec32*>(take(f32 ? sizeof(RayRec32) * s.nr : 0));
    PixRec32* pix32 = reinterpret_cast<PixRec32*>(take(f32 ? sizeof(PixRec32) * s.npad : 0));
    TileRec* tiles = reinterpret_cast<TileRec*>(take(f32 ? tile_table_bytes(s.npad, c32.P) : 0));
    double* R64 = reinterpret_cast<double*>(take(f32 ? 2 * sizeof(double) * s.npad : 0));
    void* partial = take((f32 ? 16 : 32) * (size_t)s.npad * r->n_chunks);
    InstScratch* scr = reinterpret_cast<InstScratch*>(take(sizeof(InstScratch)));
    double* part = reinterpret_cast<double*>(take(sizeof(double) * 4 * kPartBlocks));
    double* out = reinterpret_cast<double*>(take(sizeof(double) * 4 * (size_t)s.np));
    r->d_kinst = reinterpret_cast<KInst*>(take(sizeof(KInst)));
    r->d_list = reinterpret_cast<uint32_t*>(take(64));
    s.pix = pix;
    s.perm = perm;
    s.out = out;
    KInst& I = r->kinst;
    std::memset(&I, 0, sizeof(I));
    I.rays64 = s.rays; I.rays32 = rays32; I.pix64 = pix; I.pix32 = pix32; I.tiles = tiles;
    I.perm = perm; I.R64 = R64; I.partial = partial; I.out = out; I.scr = scr;
    I.nr = s.nr; I.np = s.np; I.npad = s.npad; I.n_chunks = r->n_chunks;
    I.chunk_len = r->chunk_len; I.mode = s.mode; I.tile = c32.tile; I.accumulate = 0;
    I.tile_slots = 64u * c32.P;
    I.energy = s.energy;
    I.k = s.energy / kCHBAR_H * 1e7;
    I.gx0 = s.gx0; I.gy = s.gy; I.gz0 = s.gz0; I.gsx = s.gsx; I.gsz = s.gsz;
    I.gnx = s.gnx; I.gnz = s.gnz; I.gth = s.gth; I.gtw = s.gtw;
    {
      const long double two_pi = 6.283185307179586476925286766559005768L;
      const long double kr = (long double)I.k / two_pi;
      I.kr_hi = (double)kr;
      I.kr_lo = (double)(kr - (long double)I.kr_hi);
    }
    I.skip = s.skip;
    const XrtkWaveParams& w = r->job.screen;
    uint32_t th, tw;
    grid_tile(c32.P, w.nx, w.nz, &th, &tw);
    // KInst and the one-entry merge list go up through the pinned block (the
    // result lands there later, after these copies in stream order)
    char* hs = static_cast<char*>(r->hblk);
    std::memcpy(hs, &I, sizeof(KInst));
    std::memset(hs + sizeof(KInst), 0, 64);
    hipError_t e = hipMemcpyAsync(r->d_kinst, hs, sizeof(KInst), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(r->d_list, hs + sizeof(KInst), sizeof(uint32_t), hipMemcpyHostToDevice,
                         st);
    if (e == hipSuccess && s.npad)
      e = launch_pix_grid(w.center, w.dx, w.dz, w.nx, w.nz, th, tw, s.npad, pix, perm, st);
    if (e == hipSuccess && s.nr)
      e = launch_prep(s.rays, s.nr, pix, s.npad, perm, s.np, scr, part, s.mode, I.kr_hi, I.kr_lo,
                      s.skip, 0, s.static_flags | (r->routed ? 0u : INST_F_MFMA_OK), c32.P,
                      s.gsx, s.gsz, 0.0, rays32, pix32, tiles, R64, st);
    if (e == hipSuccess && !r->routed) {
      // the planner's flags, for the routing (hblk takes the result later)
      e = hipMemcpyAsync(r->hblk, &scr->flags, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipEventRecord(r->prep_ev, st);
    }
    if (e == hipSuccess && !s.nr && s.np) e = launch_zero_f64(out, (size_t)s.np * 4, st);
    return map_err(e);
  }

  void admit() {
    while (!queue.empty() && (int)active.size() < cfg.max_active) {
      Req* r = queue.begin()->second;
      if (hbm_in_use + Pool::bucket(r->bytes) > cfg.hbm_budget_bytes) break;  // wait
      size_t got = 0;
      void* d = dpool.get(r->bytes, cfg.hbm_budget_bytes, &got);
      if (!d) break;  // retry when something retires
      const size_t hb = std::max<size_t>(sizeof(double) * 4 * r->spec.np, 512);
      size_t hgot = 0;
      void* h = hpool.get(hb, (size_t)1 << 62, &hgot);
      if (!h) {
        dpool.put(d, got);
        break;
      }
      queue.erase(queue.begin());
      r->dblk = d; r->dblk_size = got; r->hblk = h; r->hblk_size = hgot;
      hbm_in_use += got;
      hbm_peak = std::max(hbm_peak, hbm_in_use);
      r->t[1] = now_s();
      r->state = ACTIVE;
      r->rc = start_request(r);
      if (r->rc != XRTK_OK) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamSynchronize(pstream);
        r->state = FAILED;
        finish(r);
        continue;
      }
      st.admitted++;
      active.push_back(r);
      if (r->n_items == 0) enqueue_merge(r);
    }
  }

  // A request whose launch failed (rc set) only records its event: its
  // partial sums were never all computed, so nothing is merged or copied.
  void enqueue_merge(Req* r) {
    Range range("xrtk.server.merge");  // merge and readback of one request
    r->state = MERGING;
    hipError_t e = hipSuccess;
    if (r->rc == XRTK_OK && r->spec.nr)
      e = launch_merge(r->d_kinst, r->d_list, 1, r->spec.npad, stream);
    if (e == hipSuccess && r->rc == XRTK_OK && r->spec.np)
      e = hipMemcpyAsync(r->hblk, r->spec.out, sizeof(double) * 4 * r->spec.np,
                         hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipEventRecord(r->done_ev, stream);
    if (e != hipSuccess) r->rc = map_err(e);
  }

  // Resident workgroups of one persistent pair-kernel launch.
  uint32_t grid_res(const LaunchCfg& c, uint32_t mode) const {
    return std::max<uint32_t>(1, (uint32_t)ctx->n_cu * (mode == MODE_F32R ? 32u : 16u) / c.W);
  }

  // Fixes the items of a request once its planner flags have landed: the
  // MFMA kernel's blocks when the planner put it on the matrix cores, else
  // the VALU kernel's tiles as planned.
  void route(Req* r) {
    const hipError_t q = hipEventQuery(r->prep_ev);
    if (q == hipErrorNotReady) return;
    r->routed = true;
    if (q != hipSuccess) {
      // the prep or the copy failed: the request fails without items
      (void)hipGetLastError();
      r->rc = XRTK_ELAUNCH;
      r->next_item = r->n_items;
      enqueue_merge(r);
      return;
    }
    uint32_t fl = 0;
    std::memcpy(&fl, r->hblk, sizeof(fl));
    if (fl & INST_F_MFMA) {
      // items: (pass, chunk, block); an item forms one of the two sums, so
      // it counts its pairs once per pass
      r->mfma = true;
      r->mpasses = (fl & INST_F_EPZERO) ? 1u : 2u;
      r->n_items = (uint64_t)r->mpasses * r->n_chunks * r->mblocks;
      r->item_pairs = 64.0 * 64.0 * r->chunk_len / r->mpasses;
    }
  }

  // Fills and enqueues one quantum.  Returns false when there was no work.
  bool launch_quantum() {
    Range range("xrtk.server.quantum");  // work items and pair launches of one quantum
    Launch& L = launches[cur];  // the caller has retired it
    double budget = std::max(1.0, cfg.quantum_ms) * 1e-3 * rate_pairs_per_s;
    if (c32.tile > 0) {
      // whole rounds of full-size items over the persistent workgroups, so
      // that they end the launch together
      const double rays = cap32 ? std::min<double>(cap32, c32.tile) : c32.tile;
      const double round_pairs = (double)grid_res(c32, MODE_F32R) * 64.0 * c32.W * c32.P * rays;
      budget = std::max(1.0, std::round(budget / round_pairs)) * round_pairs;
    }
    double pairs = 0;
    L.h_insts.clear();
    L.h_items32.clear();
    L.h_items64.clear();
    std::vector<Req*> finishing;
    // requests with work left, fewest remaining pairs first
    std::vector<Req*> order;
    for (Req* r : active) {
      if (r->state == ACTIVE && !r->routed) route(r);
      if (r->state == ACTIVE && r->routed && r->next_item < r->n_items) order.push_back(r);
    }
    auto left = [](const Req* r) { return (double)(r->n_items - r->next_item) * r->item_pairs; };
    if (cfg.policy == XRTK_SCHED_SHORTEST)
      std::stable_sort(order.begin(), order.end(),
                       [&](const Req* a, const Req* b) { return left(a) < left(b); });
    // where each request stood, to give its items back if the launch fails
    struct Undo { Req* r; uint64_t next_item; double t2; };
    std::vector<Undo> undo;
    // items with their pair counts, ordered longest first below
    std::vector<std::pair<double, WorkItem>> it32, it64, itmf[2];
    // take up to `max_items` items of r (0: until `limit` pairs are in)
    auto take = [&](Req* r, uint64_t max_items, double limit) {
      undo.push_back(Undo{r, r->next_item, r->t[2]});
      const uint32_t li = (uint32_t)L.h_insts.size();
      L.h_insts.push_back(r->kinst);
      const double B = r->mfma ? 64.0 * 64.0 / r->mpasses : 64.0 * r->cfg.W * r->cfg.P;
      const uint32_t nblk = r->mfma ? r->mblocks : r->pblocks;
      const uint64_t per_pass = (uint64_t)r->n_chunks * nblk;
      auto& items = r->spec.mode == MODE_F32R ? it32 : it64;
      uint64_t taken = 0;
      while (r->next_item < r->n_items &&
             (max_items ? taken < max_items : (pairs < limit || taken == 0))) {
        const uint64_t it = r->next_item++;
        const uint32_t pass = (uint32_t)(it / per_pass);
        const uint64_t rem = it % per_pass;
        const uint32_t chunk = (uint32_t)(rem / nblk);
        const uint32_t pb = (uint32_t)(rem % nblk);
        const double len =
            std::min<double>(r->chunk_len, (double)r->spec.nr - (double)chunk * r->chunk_len);
        // VALU items start at a slot, MFMA items name a block (one list per sum)
        if (r->mfma) itmf[pass].emplace_back(B * len, WorkItem{li, pb, chunk, pass});
        else items.emplace_back(B * len, WorkItem{li, pb * (uint32_t)B, chunk, 0});
        pairs += B * len;
        ++taken;
      }
      if (r->t[2] == 0) r->t[2] = now_s();
      if (r->next_item >= r->n_items) finishing.push_back(r);
    };
    // reserve[k]: one item of each of order[k..], the progress share
    std::vector<double> reserve(order.size() + 1, 0.0);
    for (size_t k = order.size(); k-- > 0;)
      reserve[k] = reserve[k + 1] + std::min(left(order[k]), order[k]->item_pairs);
    size_t filled = 0;
    for (; filled < order.size() && (pairs + reserve[filled] < budget || filled == 0); ++filled)
      take(order[filled], 0, budget - reserve[filled + 1]);
    // progress for the rest: one item each
    for (size_t k = filled; k < order.size(); ++k) take(order[k], 1, 0.0);
    if (L.h_insts.empty()) return false;
    // The workgroups of the persistent kernel pull items in list order:
    // longest first leaves the short items to even out the end of the
    // launch.  Each item writes its own partial, so the order does not
    // change any result.
    auto lpt = [](std::vector<std::pair<double, WorkItem>>& v, std::vector<WorkItem>& out) {
      std::stable_sort(v.begin(), v.end(),
                       [](const auto& a, const auto& b) { return a.first > b.first; });
      out.reserve(v.size());
      for (const auto& x : v) out.push_back(x.second);
    };
    lpt(it32, L.h_items32);
    lpt(it64, L.h_items64);
    for (int k = 0; k < 2; ++k) {
      L.h_items_mf[k].clear();
      lpt(itmf[k], L.h_items_mf[k]);
    }
    // metadata
    const size_t a_ins = au(sizeof(KInst) * L.h_insts.size());
    const size_t a_32 = au(sizeof(WorkItem) * std::max<size_t>(1, L.h_items32.size()));
    const size_t a_64 = au(sizeof(WorkItem) * std::max<size_t>(1, L.h_items64.size()));
    const size_t a_mf0 = au(sizeof(WorkItem) * std::max<size_t>(1, L.h_items_mf[0].size()));
    const size_t a_mf1 = au(sizeof(WorkItem) * std::max<size_t>(1, L.h_items_mf[1].size()));
    const size_t a_mf = a_mf0 + a_mf1;
    const size_t need = a_ins + a_32 + a_64 + a_mf + 256;
    if (need > L.meta_size) {
      if (L.meta) (void)hipFree(L.meta);
      if (L.hmeta) (void)hipHostFree(L.hmeta);
      L.meta = L.hmeta = nullptr;
      L.meta_size = std::max(need * 2, (size_t)1 << 20);
      if (hipMalloc(&L.meta, L.meta_size) != hipSuccess ||
          hipHostMalloc(&L.hmeta, L.meta_size, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        L.meta_size = 0;
        // give the items back; the next iteration retries
        for (const Undo& u : undo) {
          u.r->next_item = u.next_item;
          u.r->t[2] = u.t2;
        }
        return false;
      }
    }
    char* M = static_cast<char*>(L.meta);
    char* H = static_cast<char*>(L.hmeta);
    std::memcpy(H, L.h_insts.data(), sizeof(KInst) * L.h_insts.size());
    std::memcpy(H + a_ins, L.h_items32.data(), sizeof(WorkItem) * L.h_items32.size());
    std::memcpy(H + a_ins + a_32, L.h_items64.data(), sizeof(WorkItem) * L.h_items64.size());
    std::memcpy(H + a_ins + a_32 + a_64, L.h_items_mf[0].data(),
                sizeof(WorkItem) * L.h_items_mf[0].size());
    std::memcpy(H + a_ins + a_32 + a_64 + a_mf0, L.h_items_mf[1].data(),
                sizeof(WorkItem) * L.h_items_mf[1].size());
    KInst* d_ins = reinterpret_cast<KInst*>(M);
    WorkItem* d_32 = reinterpret_cast<WorkItem*>(M + a_ins);
    WorkItem* d_64 = reinterpret_cast<WorkItem*>(M + a_ins + a_32);
    WorkItem* d_mf[2] = {reinterpret_cast<WorkItem*>(M + a_ins + a_32 + a_64),
                         reinterpret_cast<WorkItem*>(M + a_ins + a_32 + a_64 + a_mf0)};
    uint32_t* d_q = reinterpret_cast<uint32_t*>(M + a_ins + a_32 + a_64 + a_mf);
    hipError_t e =
        hipMemcpyAsync(M, H, a_ins + a_32 + a_64 + a_mf, hipMemcpyHostToDevice, stream);
    (void)hipEventRecord(L.start, stream);
    // persistent groups pull items from the queue; otherwise one group per item
    auto grid_of = [&](const LaunchCfg& c, uint32_t mode, size_t n) {
      if (!c.persistent) return (uint32_t)n;
      return std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)n, grid_res(c, mode)));
    };
    if (e == hipSuccess && !L.h_items32.empty())
      e = launch_pair(MODE_F32R, c32, d_ins, d_32, (uint32_t)L.h_items32.size(), d_q,
                      grid_of(c32, MODE_F32R, L.h_items32.size()), 0, stream);
    for (int k = 0; k < 2; ++k)
      if (e == hipSuccess && !L.h_items_mf[k].empty())
        e = launch_mfma(d_ins, d_mf[k], (uint32_t)L.h_items_mf[k].size(), d_q + 32 + k,
                        std::min<uint32_t>((uint32_t)L.h_items_mf[k].size(),
                                           (uint32_t)ctx->n_cu * 8u),
                        k == 1, false, stream);
    if (e == hipSuccess && !L.h_items64.empty())
      e = launch_pair(MODE_F64, c64, d_ins, d_64, (uint32_t)L.h_items64.size(), d_q + 16,
                      grid_of(c64, MODE_F64, L.h_items64.size()), 0, stream);
    (void)hipEventRecord(L.end, stream);
    L.in_flight = true;
    L.n_req = L.h_insts.size();
    L.pairs = pairs;
    st.launches++;
    st.max_concurrent = std::max<uint64_t>(st.max_concurrent, L.n_req);
    if (e != hipSuccess) {
      // Every request with items in this launch loses them: each fails now
      // (its remaining items are dropped), rather than later merging partial
      // sums that were never computed.
      L.pairs = 0;
      for (const Undo& u : undo) {
        Req* r = u.r;
        r->rc = map_err(e);
        if (r->next_item < r->n_items) {
          r->next_item = r->n_items;
          finishing.push_back(r);
        }
      }
    }
    for (Req* r : finishing) enqueue_merge(r);
    cur ^= 1;
    return true;
  }

  void retire_launch(Launch& L) {
    if (!L.in_flight) return;
    float ms = 0;
    if (hipEventElapsedTime(&ms, L.start, L.end) == hipSuccess && ms > 0.05f) {
      st.busy_ms += ms;
      const double r = L.pairs / (ms * 1e-3);
      rate_pairs_per_s = 0.8 * rate_pairs_per_s + 0.2 * r;
    } else {
      (void)hipGetLastError();
    }
    st.pairs_done += L.pairs;
    L.in_flight = false;
  }

  // Retires merged requests whose copy has landed.
  void reap() {
    for (size_t i = 0; i < active.size();) {
      Req* r = active[i];
      if (r->state == MERGING) {
        const hipError_t q = hipEventQuery(r->done_ev);
        if (q == hipErrorNotReady) { ++i; continue; }
        if (q != hipSuccess && r->rc == XRTK_OK) r->rc = XRTK_ELAUNCH;
        (void)hipGetLastError();
        if (r->rc == XRTK_OK && r->spec.np)
          std::memcpy(r->job.out4, r->hblk, sizeof(double) * 4 * r->spec.np);
        r->state = r->rc == XRTK_OK ? DONE : FAILED;
        active.erase(active.begin() + i);
        finish(r);
        continue;
      }
      ++i;
    }
  }

  void finish(Req* r) {
    r->t[3] = now_s();
    if (r->dblk) {
      dpool.put(r->dblk, r->dblk_size);
      hbm_in_use -= r->dblk_size;
      r->dblk = nullptr;
    }
    if (r->hblk) {
      hpool.put(r->hblk, r->hblk_size);
      r->hblk = nullptr;
    }
    st.completed++;
    cv_done.notify_all();
  }

  void loop() {
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      admit();
      Launch& L = launches[cur];
      if (L.in_flight) {  // keep at most two quanta in flight
        // wait for the older launch, handing back merged requests and
        // admitting new ones meanwhile (a merge is done well before the
        // launch after it)
        while (hipEventQuery(L.end) == hipErrorNotReady) {
          reap();
          admit();
          lk.unlock();
          std::this_thread::sleep_for(std::chrono::microseconds(50));
          lk.lock();
        }
        retire_launch(L);
      }
      const bool launched = launch_quantum();
      reap();
      if (stop && queue.empty() && active.empty()) break;
      if (!launched) {
        bool merging = false;
        for (Req* r : active) merging |= r->state == MERGING || !r->routed;
        if (merging || !queue.empty()) {
          lk.unlock();
          std::this_thread::sleep_for(std::chrono::microseconds(100));
          lk.lock();
        } else {
          cv_work.wait_for(lk, std::chrono::milliseconds(20));
        }
      }
    }
    for (Launch& L2 : launches)
      if (L2.in_flight) {
        lk.unlock();
        (void)hipEventSynchronize(L2.end);
        lk.lock();
        retire_launch(L2);
      }
  }
};

extern "C" {

int xrtk_server_create(xrtk_ctx* ctx, const XrtkServerCfg* cfg, xrtk_server** out) {
  if (!out) return XRTK_EINVAL;
  *out = nullptr;
  if (!ctx) return XRTK_ENODEV;
  if (!cfg || cfg->hbm_budget_bytes == 0 || cfg->max_active <= 0) return XRTK_EINVAL;
  if (cfg->policy != XRTK_SCHED_SHORTEST && cfg->policy != XRTK_SCHED_ARRIVAL) return XRTK_EINVAL;
  try {
    std::unique_ptr<xrtk_server> s(new xrtk_server());
    s->ctx = ctx;
    s->cfg = *cfg;
    if (s->cfg.quantum_ms <= 0) s->cfg.quantum_ms = 10.0;
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      // the options shape the F32R kernel; the FP64 kernel takes them where
      // it has that shape, else the context's batch configuration
      s->c32 = batch_cfg(ctx, MODE_F32R, cfg->opts);
      s->c64 = batch_cfg(ctx, MODE_F64, cfg->opts);
      if (!pair_cfg_supported(MODE_F32R, s->c32)) return XRTK_EINVAL;
      if (!pair_cfg_supported(MODE_F64, s->c64)) s->c64 = batch_cfg(ctx, MODE_F64, nullptr);
    }
    auto cap = [&](const LaunchCfg& c, double wave_rate) -> uint32_t {
      const double rays = 0.5e-3 * s->cfg.quantum_ms * wave_rate / (64.0 * c.P);
      const uint32_t r = std::max<uint32_t>(512u, (uint32_t)std::min(rays, 1e9) / 64u * 64u);
      return c.tile > 0 && r >= (uint32_t)c.tile ? 0u : r;
    };
    s->cap32 = cap(s->c32, kWavePairsF32);
    s->cap64 = cap(s->c64, kWavePairsF64);
    s->hpool.pinned = true;
    int dev = -1;
    (void)hipGetDevice(&dev);
    if (hipSetDevice(ctx->device) != hipSuccess) return XRTK_ENODEV;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
      return XRTK_ELAUNCH;
    if (hipStreamCreateWithFlags(&s->pstream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipStreamDestroy(s->stream);
      return XRTK_ELAUNCH;
    }
    for (Launch& L : s->launches) {
      if (hipEventCreate(&L.start) != hipSuccess || hipEventCreate(&L.end) != hipSuccess)
        return XRTK_ELAUNCH;
    }
    if (dev >= 0) (void)hipSetDevice(dev);
    xrtk_server* raw = s.get();
    s->th = std::thread([raw] { raw->loop(); });
    *out = s.release();
    return XRTK_OK;
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

int xrtk_server_submit(xrtk_server* s, const XrtkJob* job, xrtk_ticket* ticket) {
  if (!s || !job || !ticket) return XRTK_EINVAL;
  try {
    std::unique_ptr<Req> r(new Req());
    r->job = *job;
    r->t[0] = now_s();
    {
      // The request holds its ray set from the lookup on (taken under the
      // context lock, which xrtk_rayset_release also takes), so the set
      // cannot be freed between this lookup and the admission.
      std::lock_guard<std::mutex> lk(s->ctx->mu);
      auto it = s->ctx->raysets.find(job->rays);
      if (it == s->ctx->raysets.end()) return XRTK_EINVAL;
      if (!s->plan_request(r.get())) return XRTK_EINVAL;
      r->users = it->second.users;
      r->users->fetch_add(1);
    }
    // gives the hold back on every early return below
    struct Hold {
      Req* r;
      ~Hold() { if (r) r->users->fetch_sub(1); }
    } hold{r.get()};
    if (r->spec.np && !job->out4) return XRTK_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    s->st.submitted++;
    if (s->stop) return XRTK_ESHUTDOWN;
    if (Pool::bucket(r->bytes) > s->cfg.hbm_budget_bytes) {
      s->st.rejected++;
      return XRTK_ENOMEM;
    }
    if (hipEventCreateWithFlags(&r->done_ev, hipEventDisableTiming) != hipSuccess)
      return XRTK_ELAUNCH;
    if (hipEventCreateWithFlags(&r->prep_ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipEventDestroy(r->done_ev);
      r->done_ev = nullptr;
      return XRTK_ELAUNCH;
    }
    hold.r = nullptr;  // kept by the request until it is collected
    r->id = s->next_id++;
    *ticket = r->id;
    Req* raw = r.release();
    s->all[raw->id] = raw;
    const double work =
        s->cfg.policy == XRTK_SCHED_SHORTEST ? (double)raw->n_items * raw->item_pairs : 0.0;
    s->queue.emplace(std::make_pair(work, raw->id), raw);
    s->cv_work.notify_all();
    return XRTK_OK;
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

int xrtk_server_poll(xrtk_server* s, xrtk_ticket t) {
  if (!s) return XRTK_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(t);
  if (it == s->all.end()) return XRTK_EINVAL;
  const Req* r = it->second;
  if (r->state == DONE) return XRTK_OK;
  if (r->state == FAILED) return r->rc ? r->rc : XRTK_ELAUNCH;
  return XRTK_EAGAIN;
}

int xrtk_server_wait(xrtk_server* s, xrtk_ticket t, double timeout_ms, double t_out[4]) {
  if (!s) return XRTK_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(t);
  if (it == s->all.end()) return XRTK_EINVAL;
  Req* r = it->second;
  auto done = [&] { return r->state == DONE || r->state == FAILED; };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, done);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms),
                                  done)) {
    return XRTK_ETIMEDOUT;
  }
  if (t_out)
    for (int k = 0; k < 4; ++k) t_out[k] = r->t[k];
  const int rc = r->state == DONE ? XRTK_OK : (r->rc ? r->rc : XRTK_ELAUNCH);
  if (r->users) r->users->fetch_sub(1);  // the ray set is held until collected
  (void)hipEventDestroy(r->done_ev);
  (void)hipEventDestroy(r->prep_ev);
  s->all.erase(it);
  delete r;
  return rc;
}

int xrtk_server_stats(xrtk_server* s, XrtkServerStats* st) {
  if (!s || !st) return XRTK_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->st;
  st->hbm_in_use = s->hbm_in_use;
  st->hbm_peak = s->hbm_peak;
  return XRTK_OK;
}

void xrtk_server_destroy(xrtk_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->stream);
  (void)hipStreamSynchronize(s->pstream);
  for (auto& kv : s->all) {
    if (kv.second->users) kv.second->users->fetch_sub(1);
    if (kv.second->done_ev) (void)hipEventDestroy(kv.second->done_ev);
    if (kv.second->prep_ev) (void)hipEventDestroy(kv.second->prep_ev);
    delete kv.second;
  }
  s->dpool.release_all();
  s->hpool.release_all();
  for (Launch& L : s->launches) {
    if (L.meta) (void)hipFree(L.meta);
    if (L.hmeta) (void)hipHostFree(L.hmeta);
    if (L.start) (void)hipEventDestroy(L.start);
    if (L.end) (void)hipEventDestroy(L.end);
  }
  (void)hipStreamDestroy(s->stream);
  (void)hipStreamDestroy(s->pstream);
  delete s;
}

}  // extern "C"
