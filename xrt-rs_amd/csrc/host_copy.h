// Host memory copies split over a few threads (no device dependency, so the
// same source builds into the sanitizer checks of the CPU code).
#pragma once

#include <cstddef>
#include <functional>

namespace xrtk {

// memcpy split over up to 8 host threads (at least 4 MiB each).
void par_memcpy(void* dst, const void* src, size_t n);

// Copies `bytes` from `src` to `dst` in chunks of `chunk` bytes, each chunk
// split over up to 8 host threads (at least 4 MiB of the whole per thread),
// and calls `on_chunk(offset, length)` on the calling thread, in order, as
// soon as every part of that chunk has landed.  After `on_chunk` returns
// false it is not called again; the copy still completes.  Returns whether
// every call returned true.
bool gather_chunks(char* dst, const char* src, size_t bytes, size_t chunk,
                   const std::function<bool(size_t, size_t)>& on_chunk);

// One destination of scatter_chunks: `len` bytes to `dst` (null: skipped).
struct Segment {
  void* dst;
  size_t len;
};

// Copies the bytes of `src` (the segments back to back) to the segments'
// destinations in chunks of `chunk` bytes of the whole: chunk j is copied,
// split over up to 8 host threads, once `wait_chunk(j)` has returned true on
// the calling thread, so later chunks may still be arriving in `src` while
// the earlier ones are copied out.  After a false return nothing more is
// copied.  Returns whether every call returned true.
bool scatter_chunks(const char* src, const Segment* segs, size_t n_segs, size_t chunk,
                    const std::function<bool(size_t)>& wait_chunk);

}  // namespace xrtk
