// The no-allocation variant of the FP64 LU (ddla_dgetrf_noalloc_dev).
//
// ddla_dgetrf_dev takes its scratch -- the panel kernel's exchange area, the info word and
// the per-panel row lists -- from the library's growable workspaces, which call hipMalloc the
// first time (and whenever a larger order needs more).  Callers that must not allocate (code
// under graph capture, tight serving loops, harnesses that forbid hipMalloc) use this entry
// point instead: the same kernels and the same driver, with the scratch in a file-scope static
// device array whose size is fixed at build time (orders up to NOALLOC_MAX_N).
//
// The array's name is a host-side stand-in in host code; its device address is looked up once
// per device (hipGetSymbolAddress, with that device current) when the library sets up, and
// only that address is handed to kernels.  Every call on a device shares that device's array:
// calls that overlap on one device race (the library context is single-threaded anyway).
#include "ddla_internal.hpp"

namespace ddla {
namespace {

constexpr int64_t NOALLOC_MAX_N = 32768;
constexpr int64_t NOALLOC_MAX_PANELS = NOALLOC_MAX_N / 128;
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
// panel exchange area (keys, candidate rows, the panel's row list and the abort word: 1,116,672
// bytes, checked against panel_area_bytes() at set-up), then the info word, the per-panel
// counts and lists
constexpr size_t AREA_BYTES = 1116672;
constexpr size_t INFO_OFF = align256(AREA_BYTES);
constexpr size_t COUNTS_OFF = INFO_OFF + 256;
constexpr size_t LISTS_OFF = align256(COUNTS_OFF + NOALLOC_MAX_PANELS * sizeof(int));
constexpr size_t SCRATCH_BYTES = align256(LISTS_OFF + NOALLOC_MAX_PANELS * 4 * 128 * sizeof(int));

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
void* g_scratch[MAX_DEVICES] = {};    // device address of `scratch` on each device
unsigned g_tag[MAX_DEVICES] = {};     // last exchange tag used in that device's area

}  // namespace

int noalloc_setup(int device, hipStream_t s) {
  if (device < 0 || device >= MAX_DEVICES) {
    set_error("device %d beyond the %d devices the no-allocation scratch tracks", device,
              MAX_DEVICES);
    return DDLA_ERR_UNSUPPORTED;
  }
  if (g_scratch[device]) return 0;
  if (panel_area_bytes() > AREA_BYTES) {
    set_error("panel exchange area of %zu bytes exceeds the static %zu", panel_area_bytes(),
              AREA_BYTES);
    return DDLA_ERR_UNSUPPORTED;
  }
  void* p = nullptr;
  DDLA_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  // no exchange word may hold a live tag before the first panel
  DDLA_RETURN_IF(launch_fill_u32(p, SCRATCH_BYTES / 4, 0u, s));
  DDLA_HIP_CHECK(hipStreamSynchronize(s));
  g_scratch[device] = p;
  g_tag[device] = 0;
  return 0;
}

}  // namespace ddla

using namespace ddla;

extern "C" {

int ddla_dgetrf_noalloc_dev(ddla_dim_t n, double* a, ddla_inc_t lda, int32_t* dpiv) {
  DDLA_RETURN_IF(ensure_init());
  DDLA_ARG_CHECK(n >= 0 && lda >= n, "bad getrf arguments");
  if (n > NOALLOC_MAX_N) {
    set_error("n = %lld exceeds the no-allocation LU's static scratch (n <= %lld); use "
              "ddla_dgetrf_dev", (long long)n, (long long)NOALLOC_MAX_N);
    return DDLA_ERR_INVALID_ARG;
  }
  if (n == 0) return 0;
  DDLA_ARG_CHECK(a && dpiv, "null matrix or pivot pointer");
  const int dev = ctx().device;
  char* base = static_cast<char*>(g_scratch[dev]);
  if (!base) {
    set_error("no-allocation scratch not set up on device %d", dev);
    return DDLA_ERR_UNSUPPORTED;
  }
  hipStream_t s = stream();
  // each panel takes a fresh range of nb + 2 tags; restart from a cleared area before the
  // counter would wrap (words with small tags from long ago could match again)
  const unsigned need = static_cast<unsigned>((n + 127) / 128) * 130u;
  if (g_tag[dev] > 0xFFFFFFFFu - need - 1u) {
    DDLA_RETURN_IF(launch_fill_u32(base, AREA_BYTES / 4, 0u, s));
    g_tag[dev] = 0;
  }
  LuScratch sc;
  sc.panel_area = base;
  sc.dinfo = reinterpret_cast<int32_t*>(base + INFO_OFF);
  sc.counts = reinterpret_cast<int*>(base + COUNTS_OFF);
  sc.lists = reinterpret_cast<int*>(base + LISTS_OFF);
  sc.tag_counter = &g_tag[dev];
  int32_t info = 0;
  DDLA_RETURN_IF(dgetrf_internal(n, a, lda, dpiv, &info, &sc));
  return info;
}

}  // extern "C"
