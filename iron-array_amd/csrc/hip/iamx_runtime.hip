// iamx_runtime.hip -- device selection, memory, workspaces, profiling.
//
// One process drives one GPU (device 0 of what HIP exposes).  All kernels run
// on a single non-blocking stream created at default priority; synchronous
// copies are stream-ordered copies followed by a stream sync.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "iamx_internal.h"

namespace iamx {

hipStream_t g_stream = nullptr;
hipStream_t g_cur = nullptr;
uint64_t g_launch_seq = 0;
static hipStream_t g_side = nullptr;  // second compute stream, default priority
int g_num_cus = 0;
static int g_init_state = 0;  // 0 not tried, 1 ok, <0 failed
static char g_err[512] = "";
static std::mutex g_mu;
static int *g_status = nullptr;
static int *g_status_host = nullptr, *g_status_hdev = nullptr;  // mapped host copy of the status word

// lane 0 copies the status word into mapped host memory (a vector store)
__global__ void k_status_out(const int *__restrict__ status, int *out) {
  if (threadIdx.x == 0) out[0] = __atomic_load_n(status, __ATOMIC_RELAXED);
}

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_hip(hipError_t e, const char *what) {
  return set_error(IAMX_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int ensure_init() {
  if (g_init_state == 1) return IAMX_OK;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_init_state == 1) return IAMX_OK;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) {
    g_init_state = -1;
    return set_error(IAMX_ENODEV, "no HIP device available (%s)",
                     e != hipSuccess ? hipGetErrorString(e) : "0 devices");
  }
  if ((e = hipSetDevice(0)) != hipSuccess) return check_hip(e, "hipSetDevice");
  hipDeviceProp_t p;
  if ((e = hipGetDeviceProperties(&p, 0)) != hipSuccess) return check_hip(e, "hipGetDeviceProperties");
  g_num_cus = p.multiProcessorCount;
  if ((e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking)) != hipSuccess)
    return check_hip(e, "hipStreamCreateWithFlags");
  if ((e = hipStreamCreateWithFlags(&g_side, hipStreamNonBlocking)) != hipSuccess)
    return check_hip(e, "hipStreamCreateWithFlags");
  g_cur = g_stream;
  if ((e = hipMalloc(&g_status, 256)) != hipSuccess) return check_hip(e, "hipMalloc(status)");
  if ((e = hipMemset(g_status, 0, 256)) != hipSuccess) return check_hip(e, "hipMemset(status)");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return check_hip(e, "hipDeviceSynchronize");
  // this device's copy of the no-allocation scratch
  const int rc = noalloc_setup();
  if (rc) return rc;
  g_init_state = 1;
  return IAMX_OK;
}

bool initialized() { return g_init_state == 1; }

int *status_word() { return g_status; }

// ---- profiling -------------------------------------------------------------------
struct ProfEntry {
  std::string name;
  int64_t launches = 0;
  double total_ms = 0.0;
};
struct Pending {
  int entry;
  hipEvent_t a, b;
};
static bool g_prof_on = false;
static std::vector<ProfEntry> g_prof;
static std::vector<Pending> g_pending;
static std::vector<hipEvent_t> g_event_pool;
static hipEvent_t g_first = nullptr, g_last = nullptr;
static double g_span_ms = 0.0;

static hipEvent_t get_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

static bool g_capturing = false;
bool capturing() { return g_capturing; }

ProfScope::ProfScope(const char *name) {
  if (!g_prof_on || g_capturing) return;
  int idx = -1;
  for (size_t i = 0; i < g_prof.size(); i++)
    if (g_prof[i].name == name) idx = (int)i;
  if (idx < 0) {
    g_prof.push_back(ProfEntry{name});
    idx = (int)g_prof.size() - 1;
  }
  Pending p{idx, get_event(), get_event()};
  if (!p.a || !p.b) return;
  (void)hipEventRecord(p.a, g_cur);
  if (!g_first) g_first = p.a;
  g_pending.push_back(p);
  slot = (int)g_pending.size() - 1;
}

ProfScope::~ProfScope() {
  if (slot < 0) return;
  (void)hipEventRecord(g_pending[slot].b, g_cur);
  g_last = g_pending[slot].b;
}

static void prof_collect() {
  if (g_pending.empty()) return;
  (void)hipStreamSynchronize(g_stream);
  if (g_side) (void)hipStreamSynchronize(g_side);
  for (auto &p : g_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      g_prof[p.entry].launches++;
      g_prof[p.entry].total_ms += ms;
    }
  }
  if (g_first && g_last) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g_first, g_last) == hipSuccess) g_span_ms += ms;
  }
  for (auto &p : g_pending) {
    g_event_pool.push_back(p.a);
    g_event_pool.push_back(p.b);
  }
  g_pending.clear();
  g_first = g_last = nullptr;
}

// ---- workspaces ---------------------------------------------------------------------
struct Workspace {
  void *ptr = nullptr;
  size_t bytes = 0;
};
static Workspace g_ws[IAMX_WORKSPACES];

}  // namespace iamx

using namespace iamx;

extern "C" {

IAMX_API int iamx_init(void) { return ensure_init(); }

IAMX_API const char *iamx_last_error(void) { return g_err; }

IAMX_API int iamx_device_info(char *name, int namelen, int *cus, char *arch, int archlen) {
  int rc = ensure_init();
  if (rc) return rc;
  hipDeviceProp_t p;
  IAMX_HIP_CHECK(hipGetDeviceProperties(&p, 0));
  if (name && namelen > 0) snprintf(name, (size_t)namelen, "%s", p.name);
  if (arch && archlen > 0) snprintf(arch, (size_t)archlen, "%s", p.gcnArchName);
  if (cus) *cus = p.multiProcessorCount;
  return IAMX_OK;
}

IAMX_API void *iamx_malloc(size_t bytes) {
  if (ensure_init()) return nullptr;
  void *p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
    set_error(IAMX_ENOMEM, "hipMalloc(%zu) failed", bytes);
    return nullptr;
  }
  return p;
}

IAMX_API void iamx_free(void *p) {
  if (p && g_init_state == 1) (void)hipFree(p);
}

IAMX_API int iamx_h2d(void *dst, const void *src, size_t bytes) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!bytes) return IAMX_OK;
  IAMX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
  IAMX_HIP_CHECK(hipStreamSynchronize(g_stream));
  return IAMX_OK;
}

IAMX_API int iamx_d2h(void *dst, const void *src, size_t bytes) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!bytes) return IAMX_OK;
  IAMX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
  IAMX_HIP_CHECK(hipStreamSynchronize(g_stream));
  return IAMX_OK;
}

// Persistent pinned staging buffer for device -> host downloads; grows
// (synchronising the stream first) and lives until iamx_shutdown.
static void *g_host_buf = nullptr;
static size_t g_host_bytes = 0;
// Downloads run on a copy stream of their own (default priority), each after
// an event that marks the work enqueued before it on the compute stream, so
// kernels enqueued after a download overlap its copy.
static hipEvent_t g_copy_ev[IAMX_COPY_EVENTS], g_ready_ev[IAMX_COPY_EVENTS];
static bool g_copy_ev_made = false;
static hipStream_t g_copy_stream = nullptr;

static hipStream_t g_up_stream_ref();
static void sync_streams() {
  (void)hipStreamSynchronize(g_stream);
  if (g_side) (void)hipStreamSynchronize(g_side);
  if (g_copy_stream) (void)hipStreamSynchronize(g_copy_stream);
  if (g_up_stream_ref()) (void)hipStreamSynchronize(g_up_stream_ref());
}

IAMX_API void *iamx_host_buffer(size_t bytes) {
  if (ensure_init()) return nullptr;
  if (g_host_buf && g_host_bytes >= bytes) return g_host_buf;
  if (g_host_buf) {
    sync_streams();
    (void)hipHostFree(g_host_buf);
    g_host_buf = nullptr;
    g_host_bytes = 0;
  }
  const size_t want = bytes + bytes / 8 + 4096;
  if (hipHostMalloc(&g_host_buf, want, hipHostMallocDefault) != hipSuccess) {
    g_host_buf = nullptr;
    set_error(IAMX_ENOMEM, "hipHostMalloc(%zu) failed", want);
    return nullptr;
  }
  g_host_bytes = want;
  return g_host_buf;
}

// Host memory the GPU writes directly: fine-grained (coherent) and mapped,
// so kernels store into it through its device pointer and the bytes are
// visible to the host once the stream is synchronised.  Grows like the
// staging buffer.
static void *g_direct_buf = nullptr, *g_direct_dev = nullptr;
static size_t g_direct_bytes = 0;

IAMX_API void *iamx_host_direct(size_t bytes, void **dev) {
  if (ensure_init()) return nullptr;
  if (!(g_direct_buf && g_direct_bytes >= bytes)) {
    if (g_direct_buf) {
      sync_streams();
      (void)hipHostFree(g_direct_buf);
      g_direct_buf = g_direct_dev = nullptr;
      g_direct_bytes = 0;
    }
    const size_t want = bytes + bytes / 8 + 4096;
    if (hipHostMalloc(&g_direct_buf, want, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
      g_direct_buf = nullptr;
      set_error(IAMX_ENOMEM, "hipHostMalloc(%zu, mapped) failed", want);
      return nullptr;
    }
    if (hipHostGetDevicePointer(&g_direct_dev, g_direct_buf, 0) != hipSuccess || !g_direct_dev) {
      (void)hipHostFree(g_direct_buf);
      g_direct_buf = g_direct_dev = nullptr;
      set_error(IAMX_ENOMEM, "hipHostGetDevicePointer failed");
      return nullptr;
    }
    g_direct_bytes = want;
  }
  *dev = g_direct_dev;
  return g_direct_buf;
}

IAMX_API int iamx_d2h_async(void *dst, const void *src, size_t bytes, int event) {
  int rc = ensure_init();
  if (rc) return rc;
  if (event < 0 || event >= IAMX_COPY_EVENTS) return set_error(IAMX_EINVAL, "iamx_d2h_async: event %d", event);
  if (capturing()) return set_error(IAMX_EINVAL, "iamx_d2h_async: not available under graph capture");
  if (!g_copy_ev_made) {
    for (int i = 0; i < IAMX_COPY_EVENTS; i++) {
      IAMX_HIP_CHECK(hipEventCreateWithFlags(&g_copy_ev[i], hipEventDisableTiming));
      IAMX_HIP_CHECK(hipEventCreateWithFlags(&g_ready_ev[i], hipEventDisableTiming));
    }
    IAMX_HIP_CHECK(hipStreamCreateWithFlags(&g_copy_stream, hipStreamNonBlocking));
    g_copy_ev_made = true;
  }
  IAMX_HIP_CHECK(hipEventRecord(g_ready_ev[event], g_cur));
  IAMX_HIP_CHECK(hipStreamWaitEvent(g_copy_stream, g_ready_ev[event], 0));
  if (bytes) IAMX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_copy_stream));
  IAMX_HIP_CHECK(hipEventRecord(g_copy_ev[event], g_copy_stream));
  return IAMX_OK;
}

IAMX_API int iamx_d2h_chain(void *dst, const void *src, size_t bytes, int event) {
  int rc = ensure_init();
  if (rc) return rc;
  if (event < 0 || event >= IAMX_COPY_EVENTS) return set_error(IAMX_EINVAL, "iamx_d2h_chain: event %d", event);
  if (!g_copy_stream) return set_error(IAMX_EINVAL, "iamx_d2h_chain: no download queued before");
  if (bytes) IAMX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_copy_stream));
  IAMX_HIP_CHECK(hipEventRecord(g_copy_ev[event], g_copy_stream));
  return IAMX_OK;
}

IAMX_API int iamx_event_sync(int event) {
  int rc = ensure_init();
  if (rc) return rc;
  if (event < 0 || event >= IAMX_COPY_EVENTS || !g_copy_ev_made)
    return set_error(IAMX_EINVAL, "iamx_event_sync: event %d not recorded", event);
  IAMX_HIP_CHECK(hipEventSynchronize(g_copy_ev[event]));
  return IAMX_OK;
}

IAMX_API int iamx_d2d(void *dst, const void *src, size_t bytes) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!bytes) return IAMX_OK;
  IAMX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream));
  return IAMX_OK;
}

IAMX_API int iamx_copy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!width || !height) return IAMX_OK;
  if (width > dpitch || width > spitch) return set_error(IAMX_EINVAL, "iamx_copy2d: width %zu exceeds a pitch", width);
  IAMX_HIP_CHECK(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToDevice, g_cur));
  return IAMX_OK;
}

IAMX_API int iamx_memset(void *dst, int value, size_t bytes) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!bytes) return IAMX_OK;
  IAMX_HIP_CHECK(hipMemsetAsync(dst, value, bytes, g_stream));
  return IAMX_OK;
}

IAMX_API int iamx_sync(void) {
  int rc = ensure_init();
  if (rc) return rc;
  IAMX_HIP_CHECK(hipStreamSynchronize(g_stream));
  if (g_side) IAMX_HIP_CHECK(hipStreamSynchronize(g_side));
  if (g_copy_stream) IAMX_HIP_CHECK(hipStreamSynchronize(g_copy_stream));
  if (g_up_stream_ref()) IAMX_HIP_CHECK(hipStreamSynchronize(g_up_stream_ref()));
  return IAMX_OK;
}

IAMX_API int iamx_use_stream(int side) {
  int rc = ensure_init();
  if (rc) return rc;
  if (side && capturing()) return set_error(IAMX_EINVAL, "iamx_use_stream: not available under graph capture");
  g_cur = side ? g_side : g_stream;
  return IAMX_OK;
}

IAMX_API int iamx_side_wait(void) {
  int rc = ensure_init();
  if (rc) return rc;
  if (capturing()) return set_error(IAMX_EINVAL, "iamx_side_wait: not available under graph capture");
  static hipEvent_t ev = nullptr;
  if (!ev) IAMX_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  IAMX_HIP_CHECK(hipEventRecord(ev, g_stream));
  IAMX_HIP_CHECK(hipStreamWaitEvent(g_side, ev, 0));
  return IAMX_OK;
}

// ---- streamed uploads ------------------------------------------------------------------------
// Operands too large to keep resident travel in panels through two stages:
// each stage is a pinned host buffer copied to a device buffer on an upload
// stream of its own (default priority).  The copy of a stage waits (on the
// device) for the library stream to release that stage, the library stream
// waits for the copy, and the host waits for a stage's previous copy before
// refilling its pinned buffer.  Only events, no device-side waits.
static void *g_up_host[2] = {nullptr, nullptr};
static size_t g_up_bytes[2] = {0, 0};
static hipEvent_t g_up_done[2], g_up_free[2];
static bool g_up_done_rec[2] = {false, false};
static hipStream_t g_up_stream = nullptr;

static int upload_setup() {
  if (g_up_stream) return IAMX_OK;
  for (int i = 0; i < 2; i++) {
    IAMX_HIP_CHECK(hipEventCreateWithFlags(&g_up_done[i], hipEventDisableTiming));
    IAMX_HIP_CHECK(hipEventCreateWithFlags(&g_up_free[i], hipEventDisableTiming));
  }
  IAMX_HIP_CHECK(hipStreamCreateWithFlags(&g_up_stream, hipStreamNonBlocking));
  return IAMX_OK;
}

IAMX_API void *iamx_upload_host(int stage, size_t bytes) {
  if (stage < 0 || stage > 1 || ensure_init() || upload_setup()) return nullptr;
  if (g_up_done_rec[stage] && hipEventSynchronize(g_up_done[stage]) != hipSuccess) {
    set_error(IAMX_EINVAL, "iamx_upload_host: waiting for stage %d", stage);
    return nullptr;
  }
  if (g_up_host[stage] && g_up_bytes[stage] >= bytes) return g_up_host[stage];
  if (g_up_host[stage]) {
    (void)hipHostFree(g_up_host[stage]);
    g_up_host[stage] = nullptr;
    g_up_bytes[stage] = 0;
  }
  const size_t want = bytes + bytes / 8 + 4096;
  if (hipHostMalloc(&g_up_host[stage], want, hipHostMallocDefault) != hipSuccess) {
    g_up_host[stage] = nullptr;
    set_error(IAMX_ENOMEM, "hipHostMalloc(%zu) failed", want);
    return nullptr;
  }
  g_up_bytes[stage] = want;
  return g_up_host[stage];
}

IAMX_API int iamx_upload(int stage, void *d_dst, size_t bytes) {
  int rc = ensure_init();
  if (rc) return rc;
  if (stage < 0 || stage > 1 || !g_up_host[stage] || bytes > g_up_bytes[stage])
    return set_error(IAMX_EINVAL, "iamx_upload: stage %d, %zu bytes", stage, bytes);
  if (capturing()) return set_error(IAMX_EINVAL, "iamx_upload: not available under graph capture");
  rc = upload_setup();
  if (rc) return rc;
  IAMX_HIP_CHECK(hipStreamWaitEvent(g_up_stream, g_up_free[stage], 0));
  if (bytes) IAMX_HIP_CHECK(hipMemcpyAsync(d_dst, g_up_host[stage], bytes, hipMemcpyHostToDevice, g_up_stream));
  IAMX_HIP_CHECK(hipEventRecord(g_up_done[stage], g_up_stream));
  g_up_done_rec[stage] = true;
  IAMX_HIP_CHECK(hipStreamWaitEvent(g_cur, g_up_done[stage], 0));
  return IAMX_OK;
}

static hipStream_t g_up_stream_ref() { return g_up_stream; }

IAMX_API int iamx_upload_release(int stage) {
  int rc = ensure_init();
  if (rc) return rc;
  if (stage < 0 || stage > 1) return set_error(IAMX_EINVAL, "iamx_upload_release: stage %d", stage);
  rc = upload_setup();
  if (rc) return rc;
  IAMX_HIP_CHECK(hipEventRecord(g_up_free[stage], g_cur));
  return IAMX_OK;
}

IAMX_API int iamx_mem_info(size_t *free_bytes, size_t *total_bytes) {
  int rc = ensure_init();
  if (rc) return rc;
  size_t f = 0, t = 0;
  IAMX_HIP_CHECK(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return IAMX_OK;
}

IAMX_API void *iamx_workspace(int slot, size_t bytes) {
  if (slot < 0 || slot >= IAMX_WORKSPACES || ensure_init()) return nullptr;
  Workspace &w = g_ws[slot];
  if (w.bytes >= bytes && w.ptr) return w.ptr;
  if (w.ptr) {
    sync_streams();
    (void)hipFree(w.ptr);
    w.ptr = nullptr;
    w.bytes = 0;
  }
  size_t want = bytes + bytes / 8 + 4096;
  if (hipMalloc(&w.ptr, want) != hipSuccess) {
    w.ptr = nullptr;
    set_error(IAMX_ENOMEM, "workspace %d: hipMalloc(%zu) failed", slot, want);
    return nullptr;
  }
  w.bytes = want;
  return w.ptr;
}

IAMX_API void iamx_shutdown(void) {
  if (g_init_state != 1) return;
  sync_streams();
  blt_shutdown();
  for (auto &w : g_ws) {
    if (w.ptr) (void)hipFree(w.ptr);
    w = Workspace{};
  }
  if (g_host_buf) (void)hipHostFree(g_host_buf);
  g_host_buf = nullptr;
  g_host_bytes = 0;
  if (g_direct_buf) (void)hipHostFree(g_direct_buf);
  g_direct_buf = g_direct_dev = nullptr;
  g_direct_bytes = 0;
  if (g_status_host) (void)hipHostFree(g_status_host);
  g_status_host = g_status_hdev = nullptr;
  for (int i = 0; i < 2; i++) {
    if (g_up_host[i]) (void)hipHostFree(g_up_host[i]);
    g_up_host[i] = nullptr;
    g_up_bytes[i] = 0;
  }
}

IAMX_API void iamx_prof_enable(int on) { g_prof_on = on != 0; }

IAMX_API void iamx_prof_reset(void) {
  prof_collect();
  g_prof.clear();
  g_span_ms = 0.0;
}

IAMX_API int iamx_prof_entry(int i, const char **name, int64_t *launches, double *total_ms) {
  prof_collect();
  if (i < 0 || i >= (int)g_prof.size()) return 0;
  if (name) *name = g_prof[i].name.c_str();
  if (launches) *launches = g_prof[i].launches;
  if (total_ms) *total_ms = g_prof[i].total_ms;
  return 1;
}

IAMX_API double iamx_prof_span_ms(void) {
  prof_collect();
  return g_span_ms;
}

IAMX_API int iamx_capture_begin(void) {
  int rc = ensure_init();
  if (rc) return rc;
  prof_collect();
  IAMX_HIP_CHECK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
  g_capturing = true;
  return IAMX_OK;
}

IAMX_API int iamx_capture_end(void **graph_exec) {
  int rc = ensure_init();
  if (rc) return rc;
  g_capturing = false;
  hipGraph_t g = nullptr;
  IAMX_HIP_CHECK(hipStreamEndCapture(g_stream, &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return check_hip(e, "hipGraphInstantiate");
  *graph_exec = (void *)ex;
  return IAMX_OK;
}

IAMX_API int iamx_graph_launch(void *graph_exec) {
  int rc = ensure_init();
  if (rc) return rc;
  ProfScope ps("graph_launch");
  ++g_launch_seq;
  IAMX_HIP_CHECK(hipGraphLaunch((hipGraphExec_t)graph_exec, g_stream));
  return IAMX_OK;
}

IAMX_API void iamx_graph_destroy(void *graph_exec) {
  if (graph_exec && g_init_state == 1) (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec);
}

// The status word reaches the host through a one-thread kernel storing it
// into mapped host memory, behind the work queued so far: one launch and one
// stream synchronisation, no staged copy.  A snapshot queued after the last
// launch (iamx_status_snapshot) is read without a new launch.
static uint64_t g_snap_seq = ~0ull;

IAMX_API int iamx_status_snapshot(void) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!g_status_host) {
    void *h = nullptr, *d = nullptr;
    IAMX_HIP_CHECK(hipHostMalloc(&h, 64, hipHostMallocMapped | hipHostMallocCoherent));
    if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess || !d) {
      (void)hipHostFree(h);
      return set_error(IAMX_ENOMEM, "hipHostGetDevicePointer(status) failed");
    }
    g_status_host = static_cast<int *>(h);
    g_status_hdev = static_cast<int *>(d);
  }
  *(volatile int *)g_status_host = 0;
  hipLaunchKernelGGL(k_status_out, dim3(1), dim3(64), 0, g_stream, (const int *)g_status, g_status_hdev);
  IAMX_HIP_CHECK(hipGetLastError());
  g_snap_seq = g_launch_seq;
  return IAMX_OK;
}

IAMX_API int iamx_decode_status(void) {
  int rc = ensure_init();
  if (rc) return rc;
  if (g_snap_seq != g_launch_seq && (rc = iamx_status_snapshot()) != 0) return rc;
  IAMX_HIP_CHECK(hipStreamSynchronize(g_stream));
  const int v = *(volatile int *)g_status_host;
  g_snap_seq = ~0ull;
  if (v) IAMX_HIP_CHECK(hipMemsetAsync(g_status, 0, sizeof(int), g_stream));
  return v;
}

}  // extern "C"
