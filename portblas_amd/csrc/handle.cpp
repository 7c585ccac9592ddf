// Handle, workspace pool, events, graph capture and error strings.
// Replaces blas::SB_Handle (include/sb_handle/portblas_handle.h:46-200) and
// Temp_Mem_Pool (include/sb_handle/temp_memory_pool.h:33-114): one handle owns
// one HIP stream, the device's compute-unit count and a scratch pool that only
// grows (so steady-state calls never allocate).
#include "pb_internal.hpp"

#include <cstring>

namespace {
thread_local std::string g_last_error;
}

namespace pb {

pb_status_t set_error(pb_status_t status, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

void clear_error() { g_last_error.clear(); }

pb_status_t enter(pb_handle_t h) {
  if (!h) return set_error(PB_STATUS_NOT_INITIALIZED, "null pb_handle_t");
  int cur = -1;
  PB_CHECK_HIP(hipGetDevice(&cur));
  if (cur != h->device) PB_CHECK_HIP(hipSetDevice(h->device));
  return PB_STATUS_SUCCESS;
}

pb_status_t workspace(pb_handle_t h, size_t bytes, void** ptr) {
  bytes = align_up(bytes < 256 ? 256 : bytes, 256);
  std::lock_guard<std::mutex> lock(h->mu);
  if (bytes <= h->workspace_bytes) {
    *ptr = h->workspace;
    return PB_STATUS_SUCCESS;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  PB_CHECK_HIP(hipStreamIsCapturing(h->stream, &cs));
  if (cs != hipStreamCaptureStatusNone)
    return set_error(PB_STATUS_ALLOC_FAILED,
                     "workspace of %zu bytes needed during graph capture but only "
                     "%zu reserved; call pb_handle_reserve_workspace first",
                     bytes, h->workspace_bytes);
  // Grow geometrically so a sequence of slightly larger calls does not
  // reallocate every time.
  size_t want = bytes;
  if (want < h->workspace_bytes + h->workspace_bytes / 2)
    want = align_up(h->workspace_bytes + h->workspace_bytes / 2, 256);
  if (h->ws) {
    // Work already queued on the stream may still use the old buffer.  Drop
    // the handle's reference only after it has finished; graphs captured
    // with the old buffer keep it alive until they are destroyed.
    PB_CHECK_HIP(hipStreamSynchronize(h->stream));
    h->ws.reset();
    h->workspace = nullptr;
    h->workspace_bytes = 0;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(PB_STATUS_ALLOC_FAILED, "hipMalloc(%zu) failed: %s", want,
                     hipGetErrorString(e));
  }
  h->ws = std::make_shared<pb_device_buffer>(p, want);
  h->workspace = p;
  h->workspace_bytes = want;
  *ptr = p;
  return PB_STATUS_SUCCESS;
}

unsigned int* counters(pb_handle_t h, int n) {
  (void)n;  // kernels on one stream run in order, so one arena prefix suffices
  return h->counters;
}

}  // namespace pb

// hipFree waits for the device's outstanding work, so a graph destroyed right
// after its last launch does not free memory that launch still uses.
pb_device_buffer::~pb_device_buffer() {
  if (ptr) (void)hipFree(ptr);
}

using namespace pb;

extern "C" {

int pb_abi_version(void) { return PB_ABI_VERSION; }

const char* pb_last_error(void) { return g_last_error.c_str(); }

const char* pb_status_string(pb_status_t s) {
  switch (s) {
    case PB_STATUS_SUCCESS: return "PB_STATUS_SUCCESS";
    case PB_STATUS_INVALID_VALUE: return "PB_STATUS_INVALID_VALUE";
    case PB_STATUS_NOT_SUPPORTED: return "PB_STATUS_NOT_SUPPORTED";
    case PB_STATUS_DEVICE_ERROR: return "PB_STATUS_DEVICE_ERROR";
    case PB_STATUS_ALLOC_FAILED: return "PB_STATUS_ALLOC_FAILED";
    case PB_STATUS_NOT_INITIALIZED: return "PB_STATUS_NOT_INITIALIZED";
  }
  return "PB_STATUS_UNKNOWN";
}

pb_status_t pb_handle_create(void* stream, pb_handle_t* out) {
  clear_error();
  if (!out) return set_error(PB_STATUS_INVALID_VALUE, "null output pointer");
  *out = nullptr;
  int dev = 0;
  PB_CHECK_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  PB_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(PB_STATUS_NOT_SUPPORTED,
                     "device %d is %s; this library is built for gfx950 only", dev,
                     prop.gcnArchName);
  auto* h = new pb_handle_s();
  h->stream = static_cast<hipStream_t>(stream);
  h->device = dev;
  h->num_cus = prop.multiProcessorCount;
  h->max_threads = prop.maxThreadsPerBlock;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&h->counters),
                           kCounterSlots * sizeof(unsigned int));
  if (e == hipSuccess)
    e = hipMemset(h->counters, 0, kCounterSlots * sizeof(unsigned int));
  if (e == hipSuccess) e = hipMalloc(&h->scalar_slot, 256);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&h->scale_guess), 256);
  if (e == hipSuccess) e = hipMemset(h->scale_guess, 0, 256);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&h->switch_event, hipEventDisableTiming);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    if (h->counters) (void)hipFree(h->counters);
    if (h->scalar_slot) (void)hipFree(h->scalar_slot);
    if (h->scale_guess) (void)hipFree(h->scale_guess);
    if (h->switch_event) (void)hipEventDestroy(h->switch_event);
    delete h;
    return set_error(PB_STATUS_DEVICE_ERROR, "counter arena setup failed: %s",
                     hipGetErrorString(e));
  }
  *out = h;
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_handle_destroy(pb_handle_t h) {
  clear_error();
  if (!h) return PB_STATUS_SUCCESS;
  PB_TRY(enter(h));
  hipError_t e = hipStreamSynchronize(h->stream);
  h->ws.reset();  // freed now unless a live graph still holds it
  if (h->counters) (void)hipFree(h->counters);
  if (h->scalar_slot) (void)hipFree(h->scalar_slot);
  if (h->scale_guess) (void)hipFree(h->scale_guess);
  if (h->switch_event) (void)hipEventDestroy(h->switch_event);
  delete h;
  if (e != hipSuccess)
    return set_error(PB_STATUS_DEVICE_ERROR, "stream sync at destroy: %s",
                     hipGetErrorString(e));
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_handle_set_stream(pb_handle_t h, void* stream) {
  clear_error();
  PB_TRY(enter(h));
  hipStream_t next = static_cast<hipStream_t>(stream);
  if (next != h->stream) {
    // The workspace and the counter arena are shared by every call: work
    // still queued on the old stream must finish before the new stream's
    // first kernel may touch them.
    PB_CHECK_HIP(hipEventRecord(h->switch_event, h->stream));
    PB_CHECK_HIP(hipStreamWaitEvent(next, h->switch_event, 0));
  }
  h->stream = next;
  return PB_STATUS_SUCCESS;
}

void* pb_handle_get_stream(pb_handle_t h) { return h ? (void*)h->stream : nullptr; }

int pb_handle_get_num_compute_units(pb_handle_t h) { return h ? h->num_cus : 0; }

pb_status_t pb_handle_wait(pb_handle_t h) {
  clear_error();
  PB_TRY(enter(h));
  PB_CHECK_HIP(hipStreamSynchronize(h->stream));
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_handle_reserve_workspace(pb_handle_t h, size_t bytes) {
  clear_error();
  PB_TRY(enter(h));
  void* p = nullptr;
  return workspace(h, bytes, &p);
}

size_t pb_handle_workspace_bytes(pb_handle_t h) { return h ? h->workspace_bytes : 0; }

pb_status_t pb_handle_wait_event(pb_handle_t h, void* ev) {
  clear_error();
  PB_TRY(enter(h));
  PB_REQUIRE(ev != nullptr, "null event");
  PB_CHECK_HIP(hipStreamWaitEvent(h->stream, static_cast<hipEvent_t>(ev), 0));
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_handle_record_event(pb_handle_t h, void* ev) {
  clear_error();
  PB_TRY(enter(h));
  PB_REQUIRE(ev != nullptr, "null event");
  PB_CHECK_HIP(hipEventRecord(static_cast<hipEvent_t>(ev), h->stream));
  return PB_STATUS_SUCCESS;
}

// -------------------------------------------------------------------------
// Graph capture of the handle's stream.
// -------------------------------------------------------------------------
struct pb_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // The workspace the captured kernels address; kept alive while the graph
  // lives, whatever the handle's workspace does in the meantime.
  std::shared_ptr<pb_device_buffer> ws;
};

pb_status_t pb_graph_capture_begin(pb_handle_t h) {
  clear_error();
  PB_TRY(enter(h));
  PB_REQUIRE(h->stream != nullptr,
             "graph capture needs a non-null stream on the handle");
  PB_CHECK_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  h->capturing = true;
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_graph_capture_end(pb_handle_t h, pb_graph_t* out) {
  clear_error();
  PB_TRY(enter(h));
  PB_REQUIRE(out != nullptr, "null output pointer");
  auto* g = new pb_graph_s();
  hipError_t e = hipStreamEndCapture(h->stream, &g->graph);
  h->capturing = false;
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return set_error(PB_STATUS_DEVICE_ERROR, "graph capture failed: %s",
                     hipGetErrorString(e));
  }
  {
    std::lock_guard<std::mutex> lock(h->mu);
    g->ws = h->ws;
  }
  *out = g;
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_graph_launch(pb_handle_t h, pb_graph_t g) {
  clear_error();
  PB_TRY(enter(h));
  PB_REQUIRE(g != nullptr, "null graph");
  PB_CHECK_HIP(hipGraphLaunch(g->exec, h->stream));
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_graph_destroy(pb_graph_t g) {
  clear_error();
  if (!g) return PB_STATUS_SUCCESS;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;  // drops the graph's hold on its workspace
  return PB_STATUS_SUCCESS;
}

}  // extern "C"
