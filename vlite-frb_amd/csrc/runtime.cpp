// Status strings, per-thread error text and thin HIP runtime wrappers.
#include <cstring>
#include <string>

#include "common.hpp"

namespace vfrb {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

vfrb_status hip_fail(hipError_t err, const char* what, const char* file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(err), file, line);
    if (err == hipErrorOutOfMemory) return VFRB_ERR_NOMEM;
    if (err == hipErrorNoDevice || err == hipErrorInvalidDevice) return VFRB_ERR_NODEVICE;
    return VFRB_ERR_HIP;
}

}  // namespace vfrb

using vfrb::as_stream;

extern "C" {

const char* vfrb_strerror(vfrb_status s) {
    switch (s) {
        case VFRB_OK: return "success";
        case VFRB_ERR_INVALID: return "invalid argument";
        case VFRB_ERR_UNSUPPORTED: return "unsupported shape";
        case VFRB_ERR_NOMEM: return "out of memory";
        case VFRB_ERR_HIP: return "HIP runtime error";
        case VFRB_ERR_NODEVICE: return "no usable GPU";
        case VFRB_ERR_IO: return "I/O error";
        case VFRB_ERR_FORMAT: return "VDIF format error";
        case VFRB_ERR_EOF: return "end of input";
        case VFRB_ERR_CAPACITY: return "capacity exceeded";
        default: return "unknown status";
    }
}

const char* vfrb_last_error(void) { return vfrb::g_last_error; }

const char* vfrb_version(void) { return "vfrb 1.0 (gfx950)"; }

vfrb_status vfrb_device_count(int32_t* count) {
    VFRB_REQUIRE(count, VFRB_ERR_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return VFRB_OK;
}

vfrb_status vfrb_set_device(int32_t device) {
    VFRB_HIP(hipSetDevice(device));
    return VFRB_OK;
}

vfrb_status vfrb_device_name(char* buf, int32_t cap) {
    VFRB_REQUIRE(buf && cap > 0, VFRB_ERR_INVALID, "bad buffer");
    int dev = 0;
    VFRB_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    VFRB_HIP(hipGetDeviceProperties(&prop, dev));
    snprintf(buf, cap, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return VFRB_OK;
}

vfrb_status vfrb_mem_info(size_t* free_bytes, size_t* total_bytes) {
    size_t f = 0, t = 0;
    VFRB_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return VFRB_OK;
}

vfrb_status vfrb_malloc(void** ptr, size_t bytes) {
    VFRB_REQUIRE(ptr, VFRB_ERR_INVALID, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return VFRB_OK;
    VFRB_HIP(hipMalloc(ptr, bytes));
    return VFRB_OK;
}

vfrb_status vfrb_free(void* ptr) {
    if (ptr) VFRB_HIP(hipFree(ptr));
    return VFRB_OK;
}

vfrb_status vfrb_host_alloc(void** ptr, size_t bytes) {
    VFRB_REQUIRE(ptr, VFRB_ERR_INVALID, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return VFRB_OK;
    VFRB_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return VFRB_OK;
}

vfrb_status vfrb_host_free(void* ptr) {
    if (ptr) VFRB_HIP(hipHostFree(ptr));
    return VFRB_OK;
}

static vfrb_status copy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, void* stream) {
    if (bytes == 0) return VFRB_OK;
    VFRB_REQUIRE(dst && src, VFRB_ERR_INVALID, "NULL pointer in copy");
    if (stream) {
        VFRB_HIP(hipMemcpyAsync(dst, src, bytes, kind, as_stream(stream)));
    } else {
        VFRB_HIP(hipMemcpy(dst, src, bytes, kind));
    }
    return VFRB_OK;
}

vfrb_status vfrb_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    return copy(dst, src, bytes, hipMemcpyHostToDevice, stream);
}
vfrb_status vfrb_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    return copy(dst, src, bytes, hipMemcpyDeviceToHost, stream);
}
vfrb_status vfrb_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    return copy(dst, src, bytes, hipMemcpyDeviceToDevice, stream);
}

vfrb_status vfrb_memset(void* dst, int32_t value, size_t bytes, void* stream) {
    if (bytes == 0) return VFRB_OK;
    VFRB_REQUIRE(dst, VFRB_ERR_INVALID, "NULL pointer in memset");
    VFRB_HIP(hipMemsetAsync(dst, value, bytes, as_stream(stream)));
    if (!stream) VFRB_HIP(hipStreamSynchronize(nullptr));
    return VFRB_OK;
}

vfrb_status vfrb_stream_create(void** stream) {
    VFRB_REQUIRE(stream, VFRB_ERR_INVALID, "stream is NULL");
    hipStream_t s;
    VFRB_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return VFRB_OK;
}

vfrb_status vfrb_stream_destroy(void* stream) {
    if (stream) VFRB_HIP(hipStreamDestroy(as_stream(stream)));
    return VFRB_OK;
}

vfrb_status vfrb_stream_synchronize(void* stream) {
    VFRB_HIP(hipStreamSynchronize(as_stream(stream)));
    return VFRB_OK;
}

vfrb_status vfrb_device_synchronize(void) {
    VFRB_HIP(hipDeviceSynchronize());
    return VFRB_OK;
}

vfrb_status vfrb_event_create(void** event) {
    VFRB_REQUIRE(event, VFRB_ERR_INVALID, "event is NULL");
    hipEvent_t e;
    VFRB_HIP(hipEventCreate(&e));
    *event = e;
    return VFRB_OK;
}

vfrb_status vfrb_event_destroy(void* event) {
    if (event) VFRB_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
    return VFRB_OK;
}

vfrb_status vfrb_event_record(void* event, void* stream) {
    VFRB_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(event), as_stream(stream)));
    return VFRB_OK;
}

vfrb_status vfrb_event_synchronize(void* event) {
    VFRB_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)));
    return VFRB_OK;
}

vfrb_status vfrb_stream_wait_event(void* stream, void* event) {
    VFRB_REQUIRE(event, VFRB_ERR_INVALID, "event is NULL");
    VFRB_HIP(hipStreamWaitEvent(vfrb::as_stream(stream), reinterpret_cast<hipEvent_t>(event), 0));
    return VFRB_OK;
}

struct vfrb_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

vfrb_status vfrb_graph_begin(void* stream) {
    VFRB_REQUIRE(stream, VFRB_ERR_INVALID, "capture needs a created stream, not the null stream");
    VFRB_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
    return VFRB_OK;
}

vfrb_status vfrb_graph_end(void* stream, vfrb_graph** graph) {
    VFRB_REQUIRE(stream && graph, VFRB_ERR_INVALID, "NULL argument");
    *graph = nullptr;
    hipGraph_t g = nullptr;
    VFRB_HIP(hipStreamEndCapture(as_stream(stream), &g));
    hipGraphExec_t x = nullptr;
    const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return vfrb::hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    *graph = new vfrb_graph{g, x};
    return VFRB_OK;
}

vfrb_status vfrb_graph_launch(vfrb_graph* graph, void* stream) {
    VFRB_REQUIRE(graph && graph->exec, VFRB_ERR_INVALID, "NULL graph");
    VFRB_HIP(hipGraphLaunch(graph->exec, as_stream(stream)));
    return VFRB_OK;
}

vfrb_status vfrb_graph_destroy(vfrb_graph* graph) {
    if (!graph) return VFRB_OK;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
    return VFRB_OK;
}

vfrb_status vfrb_event_elapsed_ms(float* ms, void* start, void* stop) {
    VFRB_REQUIRE(ms, VFRB_ERR_INVALID, "ms is NULL");
    VFRB_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                                 reinterpret_cast<hipEvent_t>(stop)));
    return VFRB_OK;
}

}  // extern "C"
