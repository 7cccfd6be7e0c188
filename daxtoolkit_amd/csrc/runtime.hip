// Runtime part of the C-ABI: errors, memory (the ArrayManagerExecution
// backend), synchronisation and the event timer.
#include <stdarg.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "common.h"

namespace daxhip {

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return DAX_HIP_OK;
  // clear the sticky error so later calls start clean
  // (as ArrayManagerExecutionCuda.h:46-84 does for out-of-memory)
  (void)hipGetLastError();
  set_error("%s: %s (%d)", what, hipGetErrorString(e), (int)e);
  return e == hipErrorOutOfMemory ? DAX_HIP_OOM : DAX_HIP_EXEC_ERROR;
}

namespace {
std::mutex g_pool_mu;
constexpr int kMaxDevices = 64;
hipMemPool_t g_pool[kMaxDevices] = {};

// A grow-only block per (device, stream) handed out as a stack: the
// temporaries of one call nest (their holders are scoped), so they are
// released in reverse order.  Reuse needs no synchronisation: every user of
// a block runs on its stream, in order.  Allocations that do not fit come
// from the pool, and the block grows to the high-water mark the next time it
// is idle.  The mark counts the spilled temporaries live at the time too, so
// one growth covers a call's whole peak.  Blocks above kArenaMax are not
// kept.
constexpr size_t kArenaMax = (size_t)256 << 20;
struct Arena {
  char* base = nullptr;
  size_t cap = 0, top = 0, want = 0;
  struct Live {
    size_t off, size;
    bool freed;
  };
  std::vector<Live> live;
  std::vector<std::pair<void*, size_t>> spills;  // live temporaries taken from the pool
  size_t spilled = 0;                            // their bytes
};
std::map<std::pair<int, uintptr_t>, Arena> g_arenas;  // guarded by g_pool_mu

hipMemPool_t pool_of(int dev, int* rc) {  // g_pool_mu held
  *rc = DAX_HIP_OK;
  if (!g_pool[dev]) {
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t np = nullptr;
    hipError_t e = hipMemPoolCreate(&np, &props);
    if (e != hipSuccess) {
      *rc = check_hip(e, "hipMemPoolCreate");
      return nullptr;
    }
    uint64_t keep = UINT64_MAX;  // never return memory to the device at synchronisation
    e = hipMemPoolSetAttribute(np, hipMemPoolAttrReleaseThreshold, &keep);
    if (e != hipSuccess) {
      (void)hipMemPoolDestroy(np);
      *rc = check_hip(e, "hipMemPoolSetAttribute");
      return nullptr;
    }
    g_pool[dev] = np;
  }
  return g_pool[dev];
}
}  // namespace

int temp_alloc(void** p, size_t bytes, hipStream_t s) {
  *p = nullptr;
  int dev = 0;
  DAXHIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("temp_alloc: device %d beyond the pool table", dev);
    return DAX_HIP_BAD_VALUE;
  }
  bytes = align_up(bytes ? bytes : 1, 256);
  hipMemPool_t pool;
  {
    std::lock_guard<std::mutex> lock(g_pool_mu);
    int rc = DAX_HIP_OK;
    pool = pool_of(dev, &rc);
    if (rc) return rc;
    Arena& a = g_arenas[{dev, (uintptr_t)s}];
    if (a.top == 0 && a.want > a.cap) {  // idle: grow to the high-water mark
      if (a.base) (void)hipFreeAsync(a.base, s);
      a.base = nullptr;
      a.cap = 0;
      void* nb = nullptr;
      if (hipMallocFromPoolAsync(&nb, a.want, pool, s) == hipSuccess) {
        a.base = (char*)nb;
        a.cap = a.want;
      } else {
        (void)hipGetLastError();
      }
    }
    if (bytes <= kArenaMax) {
      if (a.top + bytes <= a.cap) {
        *p = a.base + a.top;
        a.live.push_back({a.top, bytes, false});
        a.top += bytes;
        return DAX_HIP_OK;
      }
      a.want = std::max(a.want, std::min(kArenaMax, a.top + a.spilled + bytes));
    }
    hipError_t e = hipMallocFromPoolAsync(p, bytes, pool, s);
    if (e != hipSuccess) {
      *p = nullptr;
      return check_hip(e, "hipMallocFromPoolAsync (library temporary)");
    }
    a.spills.push_back({*p, bytes});
    a.spilled += bytes;
  }
  return DAX_HIP_OK;
}

void temp_free(void* p, hipStream_t s) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lock(g_pool_mu);
    for (auto& kv : g_arenas) {
      Arena& a = kv.second;
      if (kv.first.second != (uintptr_t)s) continue;
      bool spill = false;  // a temporary taken from the pool: returned below
      for (size_t i = 0; i < a.spills.size() && !spill; ++i)
        if (a.spills[i].first == p) {
          a.spilled -= a.spills[i].second;
          a.spills.erase(a.spills.begin() + (std::ptrdiff_t)i);
          spill = true;
        }
      if (spill) break;
      if (!a.base || (char*)p < a.base || (char*)p >= a.base + a.cap) continue;
      const size_t off = (size_t)((char*)p - a.base);
      for (auto it = a.live.rbegin(); it != a.live.rend(); ++it)
        if (it->off == off && !it->freed) {
          it->freed = true;
          break;
        }
      while (!a.live.empty() && a.live.back().freed) {
        a.top = a.live.back().off;
        a.live.pop_back();
      }
      return;
    }
  }
  (void)hipFreeAsync(p, s);
}

}  // namespace daxhip

using namespace daxhip;

namespace {
struct Timer {
  hipEvent_t start, stop;
};
}  // namespace

extern "C" {

const char* dax_hip_last_error(void) { return g_last_error; }

const char* dax_hip_version(void) { return "daxhip 0.1 gfx950"; }

int dax_hip_device_count(int* h_count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *h_count = 0;
    return check_hip(e, "hipGetDeviceCount");
  }
  *h_count = n;
  return DAX_HIP_OK;
}

// Execution arrays come from the device's memory pool (kept, never
// trimmed at synchronisation) in the order of the null stream, which the
// adapter's calls run on: a free returns the block once the work queued
// before it has finished, and a later allocation of a similar size reuses it
// without mapping new pages (a fresh 512 MB hipMalloc/hipFree pair costs
// milliseconds).  Out of memory, the pool's idle blocks are released and the
// allocation is tried once more.
int dax_hip_malloc(void** d_ptr, size_t bytes) {
  if (!d_ptr) {
    set_error("dax_hip_malloc: NULL output pointer");
    return DAX_HIP_BAD_VALUE;
  }
  *d_ptr = nullptr;
  if (bytes == 0) return DAX_HIP_OK;
  int dev = 0;
  DAXHIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("dax_hip_malloc: device %d beyond the pool table", dev);
    return DAX_HIP_BAD_VALUE;
  }
  hipMemPool_t pool;
  {
    std::lock_guard<std::mutex> lock(g_pool_mu);
    int rc = DAX_HIP_OK;
    pool = pool_of(dev, &rc);
    if (rc) return rc;
  }
  hipError_t e = hipMallocFromPoolAsync(d_ptr, bytes, pool, nullptr);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    (void)hipStreamSynchronize(nullptr);
    (void)hipMemPoolTrimTo(pool, 0);
    e = hipMallocFromPoolAsync(d_ptr, bytes, pool, nullptr);
  }
  if (e != hipSuccess) {
    *d_ptr = nullptr;
    int rc = check_hip(e, "hipMallocFromPoolAsync");
    if (rc == DAX_HIP_OOM)
      set_error("Failed to allocate execution array with %zu bytes.", bytes);
    return rc;
  }
  return DAX_HIP_OK;
}

int dax_hip_free(void* d_ptr) {
  if (!d_ptr) return DAX_HIP_OK;
  DAXHIP_CHECK(hipFreeAsync(d_ptr, nullptr));
  return DAX_HIP_OK;
}

int dax_hip_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes, void* stream) {
  if (bytes == 0) return DAX_HIP_OK;
  DAXHIP_CHECK(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return DAX_HIP_OK;
}

int dax_hip_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes, void* stream) {
  if (bytes == 0) return DAX_HIP_OK;
  DAXHIP_CHECK(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  DAXHIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return DAX_HIP_OK;
}

int dax_hip_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, void* stream) {
  if (bytes == 0) return DAX_HIP_OK;
  DAXHIP_CHECK(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return DAX_HIP_OK;
}

int dax_hip_memset(void* d_dst, int value, size_t bytes, void* stream) {
  if (bytes == 0) return DAX_HIP_OK;
  DAXHIP_CHECK(hipMemsetAsync(d_dst, value, bytes, (hipStream_t)stream));
  return DAX_HIP_OK;
}

int dax_hip_synchronize(void* stream) {
  DAXHIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return DAX_HIP_OK;
}

int dax_hip_timer_create(void** h_timer) {
  Timer* t = new Timer();
  hipError_t e = hipEventCreate(&t->start);
  if (e == hipSuccess) e = hipEventCreate(&t->stop);
  if (e != hipSuccess) {
    delete t;
    *h_timer = nullptr;
    return check_hip(e, "hipEventCreate");
  }
  *h_timer = t;
  return DAX_HIP_OK;
}

int dax_hip_timer_destroy(void* timer) {
  Timer* t = (Timer*)timer;
  if (!t) return DAX_HIP_OK;
  (void)hipEventDestroy(t->start);
  (void)hipEventDestroy(t->stop);
  delete t;
  return DAX_HIP_OK;
}

int dax_hip_timer_reset(void* timer, void* stream) {
  Timer* t = (Timer*)timer;
  DAXHIP_CHECK(hipEventRecord(t->start, (hipStream_t)stream));
  return DAX_HIP_OK;
}

int dax_hip_timer_elapsed(void* timer, void* stream, double* h_seconds) {
  Timer* t = (Timer*)timer;
  DAXHIP_CHECK(hipEventRecord(t->stop, (hipStream_t)stream));
  DAXHIP_CHECK(hipEventSynchronize(t->stop));
  float ms = 0.0f;
  DAXHIP_CHECK(hipEventElapsedTime(&ms, t->start, t->stop));
  *h_seconds = (double)ms / 1000.0;
  return DAX_HIP_OK;
}

}  // extern "C"
