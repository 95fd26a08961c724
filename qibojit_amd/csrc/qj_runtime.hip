// Runtime part of the C-ABI: error text, device memory, copies, per-device
// scratch and the matrix staging ring.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "qj_common.h"

namespace qj {

static thread_local char g_err[1024] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return QJ_OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return set_error(QJ_EMEMORY, "%s: %s", what, hipGetErrorString(e));
  return set_error(QJ_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int check_launch(const char* what) {
  return hip_status(hipGetLastError(), what);
}

// ---------------------------------------------------------------- scratch --
static constexpr int kMaxDevices = 64;
static std::mutex g_mutex;
static DeviceScratch* g_scratch[kMaxDevices] = {nullptr};

int scratch_for_current_device(DeviceScratch** out) {
  int dev = 0;
  int rc = hip_status(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(QJ_EHIP, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    DeviceScratch* s = new DeviceScratch();
    s->partials_cap = 1 << 20;  // 8 MiB of FP64 partials
    rc = hip_status(hipMalloc(&s->partials, s->partials_cap * sizeof(double)),
                    "scratch partials");
    if (rc) { delete s; return rc; }
    rc = hip_status(hipMalloc(&s->scalars, 64 * sizeof(double)), "scratch scalars");
    if (rc) { hipFree(s->partials); delete s; return rc; }
    s->big = nullptr;
    s->big_cap = 0;
    g_scratch[dev] = s;
  }
  *out = g_scratch[dev];
  return QJ_OK;
}

int scratch_reserve_big(DeviceScratch* sc, size_t bytes, hipStream_t stream) {
  if (bytes <= sc->big_cap) return QJ_OK;
  int rc = hip_status(hipStreamSynchronize(stream), "scratch grow sync");
  if (rc) return rc;
  if (sc->big) hipFree(sc->big);
  sc->big = nullptr;
  sc->big_cap = 0;
  rc = hip_status(hipMalloc(&sc->big, bytes), "scratch work buffer");
  if (rc) return rc;
  sc->big_cap = bytes;
  return QJ_OK;
}

// ------------------------------------------------------------ matrix ring --
// Pinned host slots + device slots; a slot is reused only after the event
// recorded behind its consuming kernel has completed, so uploads never need a
// per-gate host synchronisation.
static constexpr int kSlots = 64;
static constexpr size_t kSlotBytes = 256 * 1024;  // a 7-qubit complex128 matrix

struct Ring {
  char* host = nullptr;
  char* dev = nullptr;
  hipEvent_t ev[kSlots];
  bool pending[kSlots];
  int next = 0;
  // Oversized matrices get one-off buffers, freed once their event is done.
  void* big_dev[kSlots];
  void* big_host[kSlots];
};
static Ring* g_ring[kMaxDevices] = {nullptr};

static int ring_for_device(int dev, Ring** out) {
  if (!g_ring[dev]) {
    Ring* r = new Ring();
    int rc = hip_status(hipHostMalloc((void**)&r->host, kSlots * kSlotBytes, hipHostMallocDefault),
                        "pinned staging ring");
    if (rc) { delete r; return rc; }
    rc = hip_status(hipMalloc((void**)&r->dev, kSlots * kSlotBytes), "device staging ring");
    if (rc) { hipHostFree(r->host); delete r; return rc; }
    for (int i = 0; i < kSlots; ++i) {
      rc = hip_status(hipEventCreateWithFlags(&r->ev[i], hipEventDisableTiming), "ring event");
      if (rc) return rc;
      r->pending[i] = false;
      r->big_dev[i] = nullptr;
      r->big_host[i] = nullptr;
    }
    g_ring[dev] = r;
  }
  *out = g_ring[dev];
  return QJ_OK;
}

int ring_upload(const void* src, size_t bytes, hipStream_t stream, void** dev_out,
                int* ticket) {
  int dev = 0;
  int rc = hip_status(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(g_mutex);
  Ring* r = nullptr;
  rc = ring_for_device(dev, &r);
  if (rc) return rc;
  const int slot = r->next;
  r->next = (r->next + 1) % kSlots;
  if (r->pending[slot]) {
    rc = hip_status(hipEventSynchronize(r->ev[slot]), "ring slot wait");
    if (rc) return rc;
    r->pending[slot] = false;
    if (r->big_dev[slot]) { hipFree(r->big_dev[slot]); r->big_dev[slot] = nullptr; }
    if (r->big_host[slot]) { hipHostFree(r->big_host[slot]); r->big_host[slot] = nullptr; }
  }
  char* hsrc;
  char* ddst;
  if (bytes <= kSlotBytes) {
    hsrc = r->host + slot * kSlotBytes;
    ddst = r->dev + slot * kSlotBytes;
  } else {
    rc = hip_status(hipHostMalloc((void**)&hsrc, bytes, hipHostMallocDefault), "pinned matrix buffer");
    if (rc) return rc;
    rc = hip_status(hipMalloc((void**)&ddst, bytes), "device matrix buffer");
    if (rc) { hipHostFree(hsrc); return rc; }
    r->big_host[slot] = hsrc;
    r->big_dev[slot] = ddst;
  }
  std::memcpy(hsrc, src, bytes);
  rc = hip_status(hipMemcpyAsync(ddst, hsrc, bytes, hipMemcpyHostToDevice, stream),
                  "matrix upload");
  if (rc) return rc;
  *dev_out = ddst;
  *ticket = dev * kSlots + slot;
  return QJ_OK;
}

void ring_release(int ticket, hipStream_t stream) {
  const int dev = ticket / kSlots, slot = ticket % kSlots;
  std::lock_guard<std::mutex> lock(g_mutex);
  Ring* r = g_ring[dev];
  if (!r) return;
  if (hipEventRecord(r->ev[slot], stream) == hipSuccess) r->pending[slot] = true;
}

}  // namespace qj

using namespace qj;

extern "C" {

const char* qj_last_error(void) { return g_err; }

int qj_abi_version(void) { return 1; }

int qj_device_count(int* count) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  (void)hipGetLastError();
  *count = n;
  return QJ_OK;
}

int qj_set_device(int device) { return hip_status(hipSetDevice(device), "hipSetDevice"); }

int qj_malloc(void** ptr, uint64_t nbytes) {
  if (!ptr) return set_error(QJ_EVALUE, "qj_malloc: null output pointer");
  *ptr = nullptr;
  if (nbytes == 0) return QJ_OK;
  return hip_status(hipMalloc(ptr, nbytes), "hipMalloc");
}

int qj_free(void* ptr) {
  if (!ptr) return QJ_OK;
  return hip_status(hipFree(ptr), "hipFree");
}

int qj_memcpy_htod(void* dst, const void* src, uint64_t nbytes, void* stream) {
  if (!nbytes) return QJ_OK;
  return hip_status(hipMemcpyAsync(dst, src, nbytes, hipMemcpyHostToDevice, (hipStream_t)stream),
                    "hipMemcpyAsync H2D");
}

int qj_memcpy_dtoh(void* dst, const void* src, uint64_t nbytes, void* stream) {
  if (!nbytes) return QJ_OK;
  int rc = hip_status(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
                      "hipMemcpyAsync D2H");
  if (rc) return rc;
  return hip_status(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

int qj_memcpy_dtod(void* dst, const void* src, uint64_t nbytes, void* stream) {
  if (!nbytes) return QJ_OK;
  return hip_status(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToDevice, (hipStream_t)stream),
                    "hipMemcpyAsync D2D");
}

int qj_stream_synchronize(void* stream) {
  return hip_status(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

}  // extern "C"
