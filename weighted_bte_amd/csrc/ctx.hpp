// Per-(device, stream) execution context of the C ABI: a device workspace and pinned host
// staging, so that an entry point queues its work without allocating or waiting once the
// workspace has grown to the sizes the caller uses (the first calls of a shape, or
// bte_stream_reserve, do the growing).
//
//   * Device scratch is bump-allocated from one grow-only buffer per (device, stream) and given
//     back when the call returns; a later call (or a nested call made by this one) may reuse the
//     same bytes because its kernels run after these on the same stream.  A request that does
//     not fit is served by a stream-ordered allocation freed at the end of the call, and the
//     workspace grows to the high-water mark at the start of the next top-level call on the
//     stream (stream-ordered free and allocation, no host wait).
//   * Host data the kernels need (index arrays, flags, host-side results) goes through pinned
//     staging: one of four rotating buffers per stream, each reused only after the stream has
//     passed the call that last used it (an event; the host waits only when four calls are
//     still queued on the stream).
//   * While the stream is being captured into a graph, scratch and staging come from
//     allocations owned by the captured work (made in relaxed capture mode, kept until the
//     process ends), so graph replays never share bytes with eager calls.
#pragma once
#include <vector>

#include "common.hpp"

namespace bte {

struct StreamCtx;

class CallScope {
 public:
  explicit CallScope(hipStream_t s);
  ~CallScope();
  CallScope(const CallScope&) = delete;
  CallScope& operator=(const CallScope&) = delete;

  hipStream_t stream() const { return s_; }
  bool capturing() const { return capturing_; }

  // Device scratch, 256-byte aligned, valid for the work this call queues.
  int dev_bytes(void** p, size_t bytes);
  template <class T> int dev(T** p, size_t count) { return dev_bytes((void**)p, (count ? count : 1) * sizeof(T)); }
  // Pinned host memory, valid until the stream has run the work this call queues.
  int host_bytes(void** p, size_t bytes);
  template <class T> int host(T** p, size_t count) { return host_bytes((void**)p, (count ? count : 1) * sizeof(T)); }
  // New device scratch holding a copy of `src` (host memory the caller may reuse on return).
  template <class T> int upload(T** dst, const T* src, size_t count) {
    return upload_bytes((void**)dst, src, count * sizeof(T));
  }
  int upload_bytes(void** dst, const void* src, size_t bytes);

 private:
  StreamCtx* ctx_ = nullptr;
  hipStream_t s_;
  bool outer_ = false, capturing_ = false;
  size_t dev_mark_ = 0;
  std::vector<void*> overflow_;
};

// Grow the workspace of `stream` to at least dev_bytes of device scratch and host_bytes of
// staging per rotating buffer (synchronises the stream once).
int stream_reserve(hipStream_t s, size_t dev_bytes, size_t host_bytes);

}  // namespace bte
