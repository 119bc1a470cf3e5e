// Library lifecycle, error reporting and memory entry points.
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "octhip_internal.h"

namespace octhip {

namespace {
thread_local char g_last_error[1024] = "";
bool g_init = false;
int g_device = -1;
double* g_zero_row = nullptr;
}  // namespace

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

bool initialised() { return g_init; }
int current_device() { return g_device; }
const double* zero_row() { return g_zero_row; }

namespace {
// allocations left before injected failures start (octhip_inject_alloc_failure); < 0: off
std::atomic<long> g_alloc_budget{-1};
}  // namespace

int dalloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return OCTHIP_SUCCESS;
  for (long b = g_alloc_budget.load(); b >= 0; b = g_alloc_budget.load()) {
    if (b == 0) return fail(OCTHIP_ERR_ALLOC, "hipMalloc(%zu bytes): injected allocation failure", bytes);
    if (g_alloc_budget.compare_exchange_weak(b, b - 1)) break;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(OCTHIP_ERR_ALLOC, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
  }
  return OCTHIP_SUCCESS;
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

// ------------------------------------------------------------- profiling --
namespace {
struct ProfRec {
  const char* region;
  hipEvent_t start, stop;
};
bool g_prof = false;
bool g_capturing = false;
char g_only[64] = "";           // record only this region when set
std::vector<ProfRec> g_recs;
}  // namespace

void set_capturing(bool on) { g_capturing = on; }

ProfScope::ProfScope(const char* region, hipStream_t s) : s_(s) {
  roctxRangePushA(region);
  if (!g_prof || g_capturing) return;
  if (g_only[0] && std::strcmp(g_only, region) != 0) return;
  ProfRec r{region, nullptr, nullptr};
  if (hipEventCreate(&r.start) != hipSuccess || hipEventCreate(&r.stop) != hipSuccess) return;
  if (hipEventRecord(r.start, s) != hipSuccess) return;
  slot_ = (int)g_recs.size();
  g_recs.push_back(r);
}

ProfScope::~ProfScope() {
  if (slot_ >= 0) (void)hipEventRecord(g_recs[slot_].stop, s_);
  roctxRangePop();
}

ApiRange::ApiRange(const char* name) { roctxRangePushA(name); }
ApiRange::~ApiRange() { roctxRangePop(); }

static void prof_clear() {
  for (auto& r : g_recs) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  g_recs.clear();
}

}  // namespace octhip

using namespace octhip;

extern "C" int octhip_profiling_start(void) {
  OCTHIP_REQUIRE_INIT();
  prof_clear();
  g_prof = true;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_profiling_select(const char* region) {
  if (!region || !region[0]) {
    g_only[0] = '\0';
    return OCTHIP_SUCCESS;
  }
  OCTHIP_REQUIRE(std::strlen(region) < sizeof(g_only), "region name too long");
  std::strcpy(g_only, region);
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_profiling_stop(void) {
  g_prof = false;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_profiling_get(const char* region, double* total_ms, int* count) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(region && total_ms && count, "null argument");
  double tot = 0.0;
  int n = 0;
  for (auto& r : g_recs) {
    if (std::strcmp(r.region, region) != 0) continue;
    OCTHIP_HIP(hipEventSynchronize(r.stop));
    float ms = 0.f;
    OCTHIP_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
    tot += ms;
    ++n;
  }
  *total_ms = tot;
  *count = n;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_init(int device) {
  int ndev = 0;
  OCTHIP_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fail(OCTHIP_ERR_INVALID, "device %d out of range (%d devices)", device, ndev);
  if (g_init) {
    if (device != g_device)
      return fail(OCTHIP_ERR_STATE, "already initialised on device %d", g_device);
    return OCTHIP_SUCCESS;
  }
  OCTHIP_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  OCTHIP_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(OCTHIP_ERR_UNSUPPORTED, "device %d is %s; this library is built for gfx950", device,
                prop.gcnArchName);
  size_t bytes = sizeof(double) << MAX_LDF;
  int rc = dalloc((void**)&g_zero_row, bytes);
  if (rc) return rc;
  OCTHIP_HIP(hipMemset(g_zero_row, 0, bytes));
  OCTHIP_HIP(hipDeviceSynchronize());
  g_device = device;
  g_init = true;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_finalize(void) {
  if (!g_init) return OCTHIP_SUCCESS;
  OCTHIP_HIP(hipDeviceSynchronize());
  dfree(g_zero_row);
  g_zero_row = nullptr;
  g_init = false;
  g_device = -1;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_get_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return OCTHIP_ERR_INVALID;
  std::strncpy(buf, g_last_error, len - 1);
  buf[len - 1] = '\0';
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_version(int* major, int* minor) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_malloc(void** ptr, size_t bytes) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(ptr != nullptr, "octhip_malloc: null output pointer");
  return dalloc(ptr, bytes);
}

extern "C" int octhip_free(void* ptr) {
  OCTHIP_REQUIRE_INIT();
  if (ptr) OCTHIP_HIP(hipFree(ptr));
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_stream_synchronize(void* stream) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_HIP(hipStreamSynchronize((hipStream_t)stream));
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_inject_alloc_failure(long after) {
  octhip::g_alloc_budget.store(after < 0 ? -1 : after);
  return OCTHIP_SUCCESS;
}
