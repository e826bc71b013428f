// Per-(device, stream) workspace and pinned staging of the C ABI (see ctx.hpp).
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>

#include "ctx.hpp"

namespace bte {

static constexpr int HOST_SLOTS = 4;
static constexpr size_t HOST_BLOCK_MIN = (size_t)1 << 20;

struct HostSlot {
  std::vector<std::pair<char*, size_t>> blocks;
  size_t blk = 0, top = 0;
  hipEvent_t ev = nullptr;
  bool pending = false;
};

struct StreamCtx {
  std::recursive_mutex mu;
  int depth = 0;
  char* dbase = nullptr;
  size_t dcap = 0, dtop = 0, dhigh = 0;
  HostSlot slots[HOST_SLOTS];
  int cur = 0, next = 0;
  std::vector<void*> graph_dev, graph_host;  // owned by captured work (kept to process end)
};

static std::mutex g_ctx_mu;
static std::map<std::pair<int, hipStream_t>, StreamCtx*> g_ctx;

static StreamCtx* ctx_for(hipStream_t s) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  StreamCtx*& c = g_ctx[std::make_pair(dev, s)];
  if (!c) c = new StreamCtx();
  return c;
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Runs an allocation that stream capture would otherwise refuse, with this thread's capture mode
// relaxed for the duration.
template <class Fn> static hipError_t relaxed(Fn fn) {
  hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
  (void)hipThreadExchangeStreamCaptureMode(&mode);
  const hipError_t e = fn();
  (void)hipThreadExchangeStreamCaptureMode(&mode);
  return e;
}

CallScope::CallScope(hipStream_t s) : s_(s) {
  ctx_ = ctx_for(s);
  ctx_->mu.lock();
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (s && hipStreamIsCapturing(s, &st) == hipSuccess) capturing_ = st == hipStreamCaptureStatusActive;
  outer_ = ctx_->depth++ == 0;
  if (outer_ && !capturing_) {
    StreamCtx& c = *ctx_;
    if (c.dhigh > c.dcap) {  // grow to the high-water mark of earlier calls (stream ordered)
      if (c.dbase) (void)hipFreeAsync(c.dbase, s);
      c.dbase = nullptr;
      c.dcap = 0;
      const size_t want = align_up(c.dhigh, (size_t)1 << 20);
      if (hipMallocAsync((void**)&c.dbase, want, s) == hipSuccess) c.dcap = want;
      else c.dbase = nullptr;
    }
    c.cur = c.next;
    c.next = (c.next + 1) % HOST_SLOTS;
    HostSlot& h = c.slots[c.cur];
    if (h.pending) {
      (void)hipEventSynchronize(h.ev);  // that call's copies out of this buffer are done
      h.pending = false;
    }
    h.blk = 0;
    h.top = 0;
  }
  dev_mark_ = ctx_->dtop;
}

CallScope::~CallScope() {
  for (void* p : overflow_) (void)hipFreeAsync(p, s_);
  StreamCtx& c = *ctx_;
  c.dtop = dev_mark_;
  c.depth--;
  if (outer_ && !capturing_) {
    HostSlot& h = c.slots[c.cur];
    if (!h.ev) (void)hipEventCreateWithFlags(&h.ev, hipEventDisableTiming);
    if (h.ev && hipEventRecord(h.ev, s_) == hipSuccess) h.pending = true;
  }
  c.mu.unlock();
}

int CallScope::dev_bytes(void** p, size_t bytes) {
  StreamCtx& c = *ctx_;
  const size_t b = align_up(bytes ? bytes : 1, 256);
  if (capturing_) {
    const hipError_t e = relaxed([&] { return hipMalloc(p, b); });
    if (e != hipSuccess) return set_error(BTE_ENOMEM, "scratch for captured work (%zu bytes): %s", b, hipGetErrorString(e));
    c.graph_dev.push_back(*p);
    return BTE_OK;
  }
  if (c.dtop + b <= c.dcap) {
    *p = c.dbase + c.dtop;
    c.dtop += b;
    c.dhigh = std::max(c.dhigh, c.dtop);
    return BTE_OK;
  }
  c.dhigh = std::max(c.dhigh, c.dtop + b);
  const hipError_t e = hipMallocAsync(p, b, s_);
  if (e != hipSuccess) return set_error(BTE_ENOMEM, "scratch (%zu bytes): %s", b, hipGetErrorString(e));
  overflow_.push_back(*p);
  return BTE_OK;
}

int CallScope::host_bytes(void** p, size_t bytes) {
  StreamCtx& c = *ctx_;
  const size_t b = align_up(bytes ? bytes : 1, 64);
  if (capturing_) {
    const hipError_t e = relaxed([&] { return hipHostMalloc(p, b, hipHostMallocDefault); });
    if (e != hipSuccess) return set_error(BTE_ENOMEM, "staging for captured work (%zu bytes): %s", b, hipGetErrorString(e));
    c.graph_host.push_back(*p);
    return BTE_OK;
  }
  HostSlot& h = c.slots[c.cur];
  while (h.blk < h.blocks.size() && h.top + b > h.blocks[h.blk].second) {
    ++h.blk;
    h.top = 0;
  }
  if (h.blk == h.blocks.size()) {
    const size_t cap = std::max({b, HOST_BLOCK_MIN, h.blocks.empty() ? 0 : 2 * h.blocks.back().second});
    char* blk = nullptr;
    const hipError_t e = hipHostMalloc((void**)&blk, cap, hipHostMallocDefault);
    if (e != hipSuccess) return set_error(BTE_ENOMEM, "staging (%zu bytes): %s", cap, hipGetErrorString(e));
    h.blocks.push_back({blk, cap});
    h.top = 0;
  }
  *p = h.blocks[h.blk].first + h.top;
  h.top += b;
  return BTE_OK;
}

int CallScope::upload_bytes(void** dst, const void* src, size_t bytes) {
  void* stage = nullptr;
  BTE_TRY_RC(host_bytes(&stage, bytes));
  BTE_TRY_RC(dev_bytes(dst, bytes));
  if (bytes) {
    memcpy(stage, src, bytes);
    BTE_HIP(hipMemcpyAsync(*dst, stage, bytes, hipMemcpyHostToDevice, s_));
  }
  return BTE_OK;
}

int stream_reserve(hipStream_t s, size_t dev_bytes, size_t host_bytes) {
  StreamCtx* c = ctx_for(s);
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (c->depth) return set_error(BTE_EINVAL, "stream_reserve: called inside a library call");
  BTE_HIP(hipStreamSynchronize(s));
  if (dev_bytes > c->dcap) {
    if (c->dbase) BTE_HIP(hipFreeAsync(c->dbase, s));
    c->dbase = nullptr;
    c->dcap = 0;
    const size_t want = align_up(dev_bytes, (size_t)1 << 20);
    BTE_HIP(hipMallocAsync((void**)&c->dbase, want, s));
    BTE_HIP(hipStreamSynchronize(s));
    c->dcap = want;
  }
  for (HostSlot& h : c->slots) h.pending = false;  // the stream has drained
  c->dhigh = std::max(c->dhigh, dev_bytes);
  for (HostSlot& h : c->slots) {
    size_t have = 0;
    for (auto& b : h.blocks) have = std::max(have, b.second);
    if (have < host_bytes) {
      char* blk = nullptr;
      BTE_HIP(hipHostMalloc((void**)&blk, host_bytes, hipHostMallocDefault));
      h.blocks.insert(h.blocks.begin(), {blk, host_bytes});
    }
  }
  return BTE_OK;
}

}  // namespace bte

extern "C" int bte_stream_reserve(void* stream, size_t device_bytes, size_t host_bytes) {
  return bte::stream_reserve((hipStream_t)stream, device_bytes, host_bytes);
}
