// runtime.hip -- device context, thread-local errors, pooled device memory and
// optional per-kernel event timing for librsvdpack.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <unordered_map>

#include "rsvd_internal.h"

namespace rsvd {

static thread_local char g_err[1024] = {0};

int set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  if (getenv("RSVD_VERBOSE")) fprintf(stderr, "[rsvdpack] %s\n", g_err);
  return -1;
}
void clear_error() { g_err[0] = 0; }
const char* last_error() { return g_err; }

static Context g_ctx;
static std::mutex g_init_mu;

// Device memory of the allocation-free entry point: a file-scope array the
// code object carries, so it exists on the device from the library's load
// until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns (kept in Context::scratch), never through the
// symbol itself.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

int ctx_get(Context** out) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_ctx.device < 0) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
      return set_error("no HIP device available (%s)",
                       e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    int dev = 0;
    RSVD_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    RSVD_HIP(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      return set_error("librsvdpack is built for gfx950 (MI355X); device %d is %s", dev,
                       prop.gcnArchName);
    RSVD_HIP(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    RSVD_HIP(hipEventCreateWithFlags(&g_ctx.entry_ev, hipEventDisableTiming));
    RSVD_HIP(hipHostMalloc(&g_ctx.pinned, 64, hipHostMallocDefault));
    void* sp = nullptr;
    RSVD_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    if (!sp) return set_error("hipGetSymbolAddress(scratch) returned NULL");
    g_ctx.scratch = static_cast<char*>(sp);
    g_ctx.num_cus = prop.multiProcessorCount;
    g_ctx.device = dev;
  }
  RSVD_HIP(hipSetDevice(g_ctx.device));
  *out = &g_ctx;
  return 0;
}

int order_after_null_stream(Context* c) {
  RSVD_HIP(hipEventRecord(c->entry_ev, nullptr));
  RSVD_HIP(hipStreamWaitEvent(c->stream, c->entry_ev, 0));
  return 0;
}

int read_small(Context* c, void* dst, const void* src, size_t bytes) {
  if (bytes > 64) return set_error("read_small: %zu bytes", bytes);
  RSVD_HIP(hipMemcpyAsync(c->pinned, src, bytes, hipMemcpyDeviceToHost, c->stream));
  RSVD_HIP(hipStreamSynchronize(c->stream));
  memcpy(dst, c->pinned, bytes);
  return 0;
}

// ---------------------------------------------------------------- pool
// Buffers are rounded up to 2 MiB granules above 2 MiB and to powers of two
// below, and recycled best-fit.  All frees are deferred to the stream order
// (the library only ever uses its own stream), so reuse is safe.
static size_t round_size(size_t b) {
  if (b <= 256) return 256;
  if (b < (2u << 20)) {
    size_t r = 256;
    while (r < b) r <<= 1;
    return r;
  }
  const size_t g = 2u << 20;
  return (b + g - 1) / g * g;
}

// Live allocations (pointer -> rounded size); the free list holds buffers
// that can be handed out again.
static std::unordered_map<void*, size_t> g_live;

// First fit in the scratch array, 256-byte aligned; no device call.
static int scratch_alloc(Context* c, void** p, size_t bytes) {
  const size_t want = (bytes + 255) / 256 * 256;
  size_t at = 0;
  size_t pos = 0;
  for (; pos < c->scratch_live.size(); ++pos) {
    const auto& r = c->scratch_live[pos];
    if (r.first >= at + want) break;
    at = r.first + r.second;
  }
  if (at + want > kScratchBytes)
    return set_error("allocation-free call: its device buffers need more than the %zu MiB "
                     "scratch (%zu bytes requested with %zu in use)",
                     kScratchBytes >> 20, bytes, at);
  c->scratch_live.insert(c->scratch_live.begin() + pos, {at, want});
  *p = c->scratch + at;
  return 0;
}

static bool scratch_free(Context* c, void* p) {
  if (!c->scratch || p < (void*)c->scratch || p >= (void*)(c->scratch + kScratchBytes))
    return false;
  const size_t off = (size_t)((char*)p - c->scratch);
  for (size_t i = 0; i < c->scratch_live.size(); ++i)
    if (c->scratch_live[i].first == off) {
      c->scratch_live.erase(c->scratch_live.begin() + i);
      break;
    }
  return true;
}

int dmalloc(void** p, size_t bytes) {
  Context* c = &g_ctx;
  if (c->scratch_mode) return scratch_alloc(c, p, bytes);
  size_t want = round_size(bytes);
  int best = -1;
  for (int i = 0; i < (int)c->free_blocks.size(); ++i) {
    size_t b = c->free_blocks[i].bytes;
    if (b >= want && b <= want + want / 4 + (2u << 20)) {
      if (best < 0 || b < c->free_blocks[best].bytes) best = i;
    }
  }
  if (best >= 0) {
    *p = c->free_blocks[best].p;
    g_live[*p] = c->free_blocks[best].bytes;
    c->free_blocks.erase(c->free_blocks.begin() + best);
    return 0;
  }
  hipError_t e = hipMalloc(p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    pool_trim();
    e = hipMalloc(p, want);
    if (e != hipSuccess)
      return set_error("hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
  }
  g_live[*p] = want;
  c->pooled_bytes += want;
  return 0;
}

void dfree(void* p) {
  if (!p) return;
  Context* c = &g_ctx;
  if (scratch_free(c, p)) return;
  auto it = g_live.find(p);
  if (it == g_live.end()) return;
  c->free_blocks.push_back({p, it->second});
  g_live.erase(it);
}

void pool_trim() {
  Context* c = &g_ctx;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& b : c->free_blocks) {
    (void)hipFree(b.p);
    c->pooled_bytes -= b.bytes;
  }
  c->free_blocks.clear();
}

// ---------------------------------------------------------------- kstats
KTimer::KTimer(Context* ctx, const char* name) : c(ctx) {
  if (!c->kstats) return;
  if (!c->kstats_only.empty()) {
    bool hit = false;
    for (const auto& s : c->kstats_only) hit = hit || s == name;
    if (!hit) return;
  }
  int i = 0;
  for (; i < (int)c->stats.size(); ++i)
    if (c->stats[i].name == name) break;
  if (i == (int)c->stats.size()) c->stats.push_back(KernelStat{name, 0.0, 0});
  idx = i;
  auto get_ev = [&]() {
    hipEvent_t e = nullptr;
    if (!c->event_pool.empty()) {
      e = c->event_pool.back();
      c->event_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    return e;
  };
  a = get_ev();
  b = get_ev();
  if (a) (void)hipEventRecord(a, c->stream);
}

KTimer::~KTimer() {
  if (idx < 0 || !a || !b) return;
  (void)hipEventRecord(b, c->stream);
  c->pending.push_back({idx, a, b});
}

}  // namespace rsvd
