// Library context: device, stream, error text and growable device workspaces.
#include <cstdarg>
#include <cstring>
#include <string>
#include <vector>

#include "ddla_internal.hpp"

namespace ddla {

namespace {
thread_local char g_err[1024] = "";
Context g_ctx;
}  // namespace

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_err; }

Context& ctx() { return g_ctx; }

int init_device(int device) {
  if (g_ctx.ready) {
    if (device == g_ctx.device) return 0;
    set_error("library already initialised on device %d", g_ctx.device);
    return DDLA_ERR_INVALID_ARG;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    set_error("no HIP device visible");
    return DDLA_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= count) {
    set_error("device %d out of range (%d visible)", device, count);
    return DDLA_ERR_INVALID_ARG;
  }
  DDLA_HIP_CHECK(hipSetDevice(device));
  hipDeviceProp_t prop;
  DDLA_HIP_CHECK(hipGetDeviceProperties(&prop, device));
  g_ctx.device = device;
  g_ctx.num_cus = prop.multiProcessorCount;
  DDLA_HIP_CHECK(hipStreamCreateWithFlags(&g_ctx.own_stream, hipStreamNonBlocking));
  g_ctx.stream = g_ctx.own_stream;
  // the auxiliary stream and its events exist from set-up on, so no entry point creates them
  DDLA_HIP_CHECK(hipStreamCreateWithFlags(&g_ctx.aux_stream, hipStreamNonBlocking));
  for (hipEvent_t& e : g_ctx.ev) DDLA_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  g_ctx.ready = true;
  // the no-allocation entry points' static scratch on this device
  DDLA_RETURN_IF(noalloc_setup(device, g_ctx.own_stream));
  return 0;
}

int aux_stream(hipStream_t* s, hipEvent_t** ev) {
  if (!g_ctx.aux_stream) {
    DDLA_HIP_CHECK(hipStreamCreateWithFlags(&g_ctx.aux_stream, hipStreamNonBlocking));
    for (hipEvent_t& e : g_ctx.ev) DDLA_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  *s = g_ctx.aux_stream;
  *ev = g_ctx.ev;
  return 0;
}

int pinned_host(void** p) {
  if (!g_ctx.pinned) DDLA_HIP_CHECK(hipHostMalloc(&g_ctx.pinned, PINNED_BYTES, hipHostMallocDefault));
  *p = g_ctx.pinned;
  return 0;
}

int ensure_init() {
  if (g_ctx.ready) {
    DDLA_HIP_CHECK(hipSetDevice(g_ctx.device));
    return 0;
  }
  return init_device(0);
}

hipStream_t stream() { return g_ctx.stream; }

int workspace(int slot, size_t bytes, void** out) {
  if (slot < 0 || slot >= WS_COUNT) {
    set_error("bad workspace slot %d", slot);
    return DDLA_ERR_INVALID_ARG;
  }
  if (bytes == 0) bytes = 256;
  if (g_ctx.ws_bytes[slot] < bytes) {
    if (g_ctx.ws[slot]) {
      // earlier queued work (on either library stream) may still use the old buffer
      DDLA_HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
      if (g_ctx.aux_stream) DDLA_HIP_CHECK(hipStreamSynchronize(g_ctx.aux_stream));
      DDLA_HIP_CHECK(hipFree(g_ctx.ws[slot]));
      g_ctx.ws[slot] = nullptr;
      g_ctx.ws_bytes[slot] = 0;
    }
    const size_t want = bytes + bytes / 4;  // headroom against repeated growth
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
      set_error("hipMalloc of %zu bytes for workspace %d failed", want, slot);
      return DDLA_ERR_ALLOC;
    }
    g_ctx.ws[slot] = p;
    g_ctx.ws_bytes[slot] = want;
    g_ctx.ws_cleared[slot] = 0;
  }
  *out = g_ctx.ws[slot];
  return 0;
}

int workspace_cleared(int slot, size_t bytes, void** out, hipStream_t s, bool* fresh) {
  DDLA_RETURN_IF(workspace(slot, bytes, out));
  if (bytes == 0) bytes = 256;
  const bool clear = g_ctx.ws_cleared[slot] < bytes;
  if (clear) {
    DDLA_HIP_CHECK(hipMemsetAsync(*out, 0, bytes, s));
    g_ctx.ws_cleared[slot] = bytes;
    g_ctx.ws_tag[slot] = 0;
  }
  if (fresh) *fresh = clear;
  return 0;
}

int next_tag(int slot, size_t bytes, unsigned step, hipStream_t s, unsigned* tag) {
  void* p = nullptr;
  DDLA_RETURN_IF(workspace_cleared(slot, bytes, &p, s));
  unsigned& t = g_ctx.ws_tag[slot];
  if (t > 0xFFFFFFFFu - step - 1u) {
    // the sequence would wrap: words with small tags may still be in the slot
    DDLA_HIP_CHECK(hipMemsetAsync(p, 0, g_ctx.ws_cleared[slot], s));
    t = 0;
  }
  *tag = t + 1u;
  t += step;
  return 0;
}

int check_abort(unsigned* word, hipStream_t s, const char* what) {
  if (!word) return 0;
  unsigned h = 0;
  DDLA_HIP_CHECK(hipMemcpyAsync(&h, word, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  DDLA_HIP_CHECK(hipStreamSynchronize(s));
  if (h == 0) return 0;
  DDLA_RETURN_IF(launch_fill_u32(word, 1, 0u, s));
  DDLA_HIP_CHECK(hipStreamSynchronize(s));
  set_error("%s: a workgroup waited more than %u poll rounds for another one and the launch "
            "gave up (workgroups not co-resident: other work on the device?)", what,
            g_ctx.wait_limit);
  return DDLA_ERR_HIP;
}

// ---------------------------------------------------------------- kernel timing ---------
namespace {
struct ProfEntry {
  std::string name;
  int64_t count = 0;
  double ms = 0.0, flops = 0.0, bytes = 0.0;
};
struct PendingEvent {
  int entry;
  hipEvent_t a, b;
  double flops, bytes;
};
bool g_prof_on = false;
std::string g_prof_only;  // empty: every kernel
std::vector<ProfEntry> g_prof;
std::vector<PendingEvent> g_pending;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t take_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

void resolve_pending() {
  if (g_pending.empty()) return;
  hipStreamSynchronize(g_ctx.stream);
  for (auto& p : g_pending) {
    float ms = 0.f;
    hipEventSynchronize(p.b);
    hipEventElapsedTime(&ms, p.a, p.b);
    ProfEntry& e = g_prof[p.entry];
    e.count += 1;
    e.ms += ms;
    e.flops += p.flops;
    e.bytes += p.bytes;
    g_event_pool.push_back(p.a);
    g_event_pool.push_back(p.b);
  }
  g_pending.clear();
}
}  // namespace

bool profiling() { return g_prof_on; }

ProfScope::ProfScope(const char* name, double flops, double bytes, hipStream_t s) : s_(s) {
  if (!g_prof_on || (!g_prof_only.empty() && g_prof_only != name)) return;
  int idx = -1;
  for (size_t i = 0; i < g_prof.size(); ++i)
    if (g_prof[i].name == name) idx = (int)i;
  if (idx < 0) {
    g_prof.push_back(ProfEntry{name});
    idx = (int)g_prof.size() - 1;
  }
  entry_ = idx;
  flops_ = flops;
  bytes_ = bytes;
  a_ = take_event();
  b_ = take_event();
  hipEventRecord(static_cast<hipEvent_t>(a_), s_);
}

ProfScope::~ProfScope() {
  if (entry_ < 0) return;
  hipEventRecord(static_cast<hipEvent_t>(b_), s_);
  g_pending.push_back(PendingEvent{entry_, static_cast<hipEvent_t>(a_),
                                   static_cast<hipEvent_t>(b_), flops_, bytes_});
  if (g_pending.size() > 4096) resolve_pending();
}

void profile_enable(bool on) {
  resolve_pending();
  g_prof_on = on;
}

void profile_only(const char* name) {
  resolve_pending();
  g_prof_only = name ? name : "";
}

void profile_reset() {
  resolve_pending();
  g_prof.clear();
}

int profile_read(int idx, char* name, size_t len, int64_t* count, double* ms, double* flops,
                 double* bytes) {
  resolve_pending();
  if (idx < 0) return (int)g_prof.size();
  if (idx >= (int)g_prof.size()) return DDLA_ERR_INVALID_ARG;
  const ProfEntry& e = g_prof[idx];
  if (name && len) {
    std::strncpy(name, e.name.c_str(), len - 1);
    name[len - 1] = '\0';
  }
  if (count) *count = e.count;
  if (ms) *ms = e.ms;
  if (flops) *flops = e.flops;
  if (bytes) *bytes = e.bytes;
  return 0;
}

int finalize() {
  if (!g_ctx.ready) return 0;
  resolve_pending();
  for (hipEvent_t e : g_event_pool) hipEventDestroy(e);
  g_event_pool.clear();
  DDLA_HIP_CHECK(hipDeviceSynchronize());
  for (int i = 0; i < WS_COUNT; ++i) {
    if (g_ctx.ws[i]) hipFree(g_ctx.ws[i]);
    g_ctx.ws[i] = nullptr;
    g_ctx.ws_bytes[i] = 0;
  }
  if (g_ctx.own_stream) hipStreamDestroy(g_ctx.own_stream);
  if (g_ctx.aux_stream) hipStreamDestroy(g_ctx.aux_stream);
  for (hipEvent_t e : g_ctx.ev)
    if (e) hipEventDestroy(e);
  g_ctx = Context{};
  return 0;
}

}  // namespace ddla
