// fh_runtime.cpp -- error state, device selection and the device memory pool.
#include <map>
#include <mutex>

#include "fh_api_common.hpp"

namespace fh {

static thread_local std::string g_last_error;

void set_last_error(const std::string& s) { g_last_error = s; }

// Size classes: round up to 256 B below 1 MiB, to 64 KiB above.  Freed blocks
// are cached per (device, class).  The pool never returns memory to the
// driver before fh_pool_trim / process end, so steady-state loops do not call
// hipMalloc.
namespace {
std::mutex g_pool_mu;
std::multimap<std::pair<int, size_t>, void*> g_free;
std::map<void*, std::pair<int, size_t>> g_live;

size_t size_class(size_t b) {
  if (b <= (1u << 20)) return (b + 255) & ~size_t(255);
  return (b + 65535) & ~size_t(65535);
}
}  // namespace

// FH_POOL_POISON=1 (a checking mode for tests): every block handed out is
// first filled with 0xFF bytes (NaN as FP64/FP32, -1 as int32) on the null
// stream, so a kernel that reads memory it did not write gives a visibly
// different result.  Code that captures graphs must not run in this mode.
static void poison_block(void* p, size_t c) {
  const char* e = getenv("FH_POOL_POISON");
  if (e && e[0] == '1') FH_HIP(hipMemsetAsync(p, 0xFF, c, nullptr));
}

void* dev_alloc(size_t bytes) {
  int dev = 0;
  FH_HIP(hipGetDevice(&dev));
  size_t c = size_class(bytes);
  {
    std::unique_lock<std::mutex> lk(g_pool_mu);
    auto it = g_free.find({dev, c});
    if (it != g_free.end()) {
      void* p = it->second;
      g_free.erase(it);
      g_live[p] = {dev, c};
      lk.unlock();
      poison_block(p, c);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, c);
  if (e != hipSuccess) {
    pool_trim();
    FH_HIP(hipMalloc(&p, c));
  }
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_live[p] = {dev, c};
  }
  poison_block(p, c);
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  auto it = g_live.find(p);
  if (it == g_live.end()) return;
  g_free.insert({it->second, p});
  g_live.erase(it);
}

void pool_trim() {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  // blocks may still be referenced by queued work: wait before releasing them
  (void)hipDeviceSynchronize();
  for (auto& kv : g_free) (void)hipFree(kv.second);
  g_free.clear();
}

}  // namespace fh

extern "C" {

const char* fh_last_error(void) { return fh::g_last_error.c_str(); }

const char* fh_version(void) { return "faust_hip 0.1 (gfx950)"; }

int fh_dev_count(int* count) {
  FH_API_BEGIN
  FH_HIP(hipGetDeviceCount(count));
  FH_API_END
}

int fh_set_dev(int dev) {
  FH_API_BEGIN
  FH_HIP(hipSetDevice(dev));
  FH_API_END
}

int fh_autotune_clear(void) {
  FH_API_BEGIN
  fh::autotune_clear();
  FH_API_END
}

int fh_cur_dev(int* dev) {
  FH_API_BEGIN
  FH_HIP(hipGetDevice(dev));
  FH_API_END
}

int fh_stream_sync(void* stream) {
  FH_API_BEGIN
  FH_HIP(hipStreamSynchronize((hipStream_t)stream));
  FH_API_END
}

}  // extern "C"
