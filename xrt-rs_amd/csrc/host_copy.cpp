// Host memory copies split over a few threads: the page-locked staging of
// the solves' inputs and results.  Plain C++, no HIP.
#include "host_copy.h"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace xrtk {

namespace {

size_t n_threads(size_t bytes) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return std::min<size_t>(std::min(8u, hw), std::max<size_t>(1, bytes >> 22));
}

// [a, b) of part t of T over n bytes (64-byte aligned cuts)
void part_range(size_t n, size_t t, size_t T, size_t& a, size_t& b) {
  a = (n * t / T) & ~size_t(63);
  b = t + 1 == T ? n : (n * (t + 1) / T) & ~size_t(63);
}

}  // namespace

void par_memcpy(void* dst, const void* src, size_t n) {
  const size_t T = n_threads(n);
  auto part = [&](size_t t) {
    size_t a, b;
    part_range(n, t, T, a, b);
    std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a);
  };
  // (a thread that cannot be started leaves its part, and the later ones,
  // to the calling thread)
  std::vector<std::thread> th;
  size_t started = 1;
  try {
    th.reserve(T);
    for (; started < T; ++started) th.emplace_back(part, started);
  } catch (...) {
  }
  part(0);
  for (size_t t = started; t < T; ++t) part(t);
  for (auto& x : th) x.join();
}

bool gather_chunks(char* dst, const char* src, size_t bytes, size_t chunk,
                   const std::function<bool(size_t, size_t)>& on_chunk) {
  if (!bytes) return true;
  chunk = std::max<size_t>(chunk, 64);
  const size_t C = (bytes + chunk - 1) / chunk;
  const size_t T = n_threads(bytes);
  std::unique_ptr<std::atomic<uint32_t>[]> done(new std::atomic<uint32_t>[C]);
  for (size_t c = 0; c < C; ++c) done[c].store(0, std::memory_order_relaxed);
  // every thread copies its share of chunk c; the calling thread hands
  // chunk c on once all shares are in
  auto share = [&](size_t t, size_t c) {
    const size_t c0 = c * chunk, n = std::min(chunk, bytes - c0);
    size_t a, b;
    part_range(n, t, T, a, b);
    std::memcpy(dst + c0 + a, src + c0 + a, b - a);
    done[c].fetch_add(1, std::memory_order_release);
  };
  // (the calling thread also takes the shares of threads that could not
  // be started)
  std::vector<std::thread> th;
  size_t started = 1;
  try {
    th.reserve(T);
    for (; started < T; ++started)
      th.emplace_back([&, t = started]() {
        for (size_t c = 0; c < C; ++c) share(t, c);
      });
  } catch (...) {
  }
  bool ok = true;
  for (size_t c = 0; c < C; ++c) {
    share(0, c);
    for (size_t t = started; t < T; ++t) share(t, c);
    while (done[c].load(std::memory_order_acquire) < T) std::this_thread::yield();
    if (ok) ok = on_chunk(c * chunk, std::min(chunk, bytes - c * chunk));
  }
  for (auto& x : th) x.join();
  return ok;
}

bool scatter_chunks(const char* src, const Segment* segs, size_t n_segs, size_t chunk,
                    const std::function<bool(size_t)>& wait_chunk) {
  size_t bytes = 0;
  for (size_t k = 0; k < n_segs; ++k) bytes += segs[k].len;
  if (!bytes) return true;
  chunk = std::max<size_t>(chunk, 64);
  const size_t C = (bytes + chunk - 1) / chunk;
  const size_t T = n_threads(bytes);
  // bytes [a, b) of the whole to their segments
  auto copy_range = [&](size_t a, size_t b) {
    size_t off = 0;
    for (size_t k = 0; k < n_segs && off < b; off += segs[k].len, ++k) {
      const size_t lo = std::max(a, off), hi = std::min(b, off + segs[k].len);
      if (lo < hi && segs[k].dst)
        std::memcpy(static_cast<char*>(segs[k].dst) + (lo - off), src + lo, hi - lo);
    }
  };
  auto share = [&](size_t t, size_t c) {
    const size_t c0 = c * chunk, n = std::min(chunk, bytes - c0);
    size_t a, b;
    part_range(n, t, T, a, b);
    copy_range(c0 + a, c0 + b);
  };
  // the workers sleep until the calling thread has released a chunk (the
  // chunks arrive over PCIe; spinning would take cores from the caller's
  // own threads meanwhile)
  std::mutex mu;
  std::condition_variable cv;
  size_t released = 0;  // chunks [0, released) may be copied
  bool stop = false;
  std::vector<std::thread> th;
  size_t started = 1;
  try {
    th.reserve(T);
    for (; started < T; ++started)
      th.emplace_back([&, t = started]() {
        for (size_t c = 0; c < C; ++c) {
          {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return stop || released > c; });
            if (released <= c) return;
          }
          share(t, c);
        }
      });
  } catch (...) {
  }
  bool ok = true;
  for (size_t c = 0; c < C; ++c) {
    ok = wait_chunk(c);
    {
      std::lock_guard<std::mutex> lk(mu);
      if (ok) released = c + 1; else stop = true;
    }
    cv.notify_all();
    if (!ok) break;
    share(0, c);
    for (size_t t = started; t < T; ++t) share(t, c);
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    stop = true;
  }
  cv.notify_all();
  for (auto& x : th) x.join();
  return ok;
}

}  // namespace xrtk
