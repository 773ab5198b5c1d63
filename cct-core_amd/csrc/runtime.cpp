// Platform layer of libcogconv: errors, device, buffers, streams, events.
//
// Replaces the reference's OpenCL platform services for this path:
//   OpenCLPlatform      src/main/scala/cogx/platform/opencl/OpenCLPlatform.scala:38-244
//   OpenCLDevice        src/main/scala/cogx/platform/opencl/OpenCLDevice.scala:39-402
//   OpenCLBuffer        src/main/scala/cogx/platform/opencl/OpenCLBuffer.scala:53-523
//   OpenCLParallelCommandQueue  src/main/scala/cogx/platform/opencl/OpenCLParallelCommandQueue.scala:45-283
//   event status / timing       src/main/scala/cogx/platform/opencl/OpenCLDeviceKernel.scala:106-137
// Every entry point is noexcept in effect: errors become status codes plus a
// thread-local message, as the reference's JNI convention requires
// (src/main/java/cogjni/CogJNIWrapper.java:29-33).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "internal.h"

namespace cogconv {

static thread_local std::string g_last_error;

void clear_error() { g_last_error.clear(); }

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return COGCONV_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(COGCONV_ERR_OOM, "%s: %s", what, hipGetErrorString(e));
    return fail(COGCONV_ERR_RUNTIME, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

}  // namespace cogconv

using namespace cogconv;

extern "C" const char* cogconv_last_error(void) { return g_last_error.c_str(); }

extern "C" int cogconv_abi_version(void) { return COGCONV_ABI_VERSION; }

extern "C" int cogconv_init(int32_t device) {
    clear_error();
    int n = 0;
    int e = check_hip(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (e) return e;
    if (device < 0 || device >= n) return fail(COGCONV_ERR_INVALID_ARG, "device %d out of range (%d devices)", device, n);
    e = check_hip(hipSetDevice(device), "hipSetDevice");
    if (e) return e;
    e = check_hip(hipFree(nullptr), "context init");
    if (e) return e;
    return noalloc_setup(device);
}

extern "C" int cogconv_shutdown(void) {
    clear_error();
    return check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

extern "C" int cogconv_device_count(int32_t* count) {
    clear_error();
    if (!count) return fail(COGCONV_ERR_INVALID_ARG, "null count");
    int n = 0;
    hipError_t err = hipGetDeviceCount(&n);
    if (err == hipErrorNoDevice) n = 0;
    else if (err != hipSuccess) return check_hip(err, "hipGetDeviceCount");
    *count = n;
    return COGCONV_OK;
}

extern "C" int cogconv_device_props_get(int32_t device, cogconv_device_props* props) {
    clear_error();
    if (!props) return fail(COGCONV_ERR_INVALID_ARG, "null props");
    hipDeviceProp_t p;
    int e = check_hip(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
    if (e) return e;
    std::memset(props, 0, sizeof *props);
    std::snprintf(props->name, sizeof props->name, "%s", p.name);
    std::snprintf(props->arch, sizeof props->arch, "%s", p.gcnArchName);
    props->compute_units = p.multiProcessorCount;
    props->wavefront_size = p.warpSize;
    props->lds_bytes_per_block = (int32_t)p.sharedMemPerBlock;
    props->max_threads_per_block = p.maxThreadsPerBlock;
    props->clock_khz = p.clockRate;
    props->memory_clock_khz = p.memoryClockRate;
    props->memory_bus_width = p.memoryBusWidth;
    props->l2_bytes = p.l2CacheSize;
    props->hbm_bytes = (uint64_t)p.totalGlobalMem;
    return COGCONV_OK;
}

extern "C" int cogconv_device_synchronize(void) {
    clear_error();
    return check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

extern "C" int cogconv_buf_alloc(size_t bytes, void** dptr) {
    clear_error();
    if (!dptr) return fail(COGCONV_ERR_INVALID_ARG, "null pointer");
    *dptr = nullptr;
    if (bytes == 0) bytes = 256;
    return check_hip(hipMalloc(dptr, bytes), "hipMalloc");
}

extern "C" int cogconv_buf_free(void* dptr) {
    clear_error();
    if (!dptr) return COGCONV_OK;
    return check_hip(hipFree(dptr), "hipFree");
}

extern "C" int cogconv_h2d(void* dst, const void* src, size_t bytes) {
    clear_error();
    if ((!dst || !src) && bytes) return fail(COGCONV_ERR_INVALID_ARG, "null pointer");
    return check_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

extern "C" int cogconv_d2h(void* dst, const void* src, size_t bytes) {
    clear_error();
    if ((!dst || !src) && bytes) return fail(COGCONV_ERR_INVALID_ARG, "null pointer");
    return check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

extern "C" int cogconv_h2d_async(void* dst, const void* src, size_t bytes, cogconv_stream stream) {
    clear_error();
    if ((!dst || !src) && bytes) return fail(COGCONV_ERR_INVALID_ARG, "null pointer");
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, reinterpret_cast<hipStream_t>(stream)),
                     "hipMemcpyAsync H2D");
}

extern "C" int cogconv_d2h_async(void* dst, const void* src, size_t bytes, cogconv_stream stream) {
    clear_error();
    if ((!dst || !src) && bytes) return fail(COGCONV_ERR_INVALID_ARG, "null pointer");
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, reinterpret_cast<hipStream_t>(stream)),
                     "hipMemcpyAsync D2H");
}

extern "C" int cogconv_d2d_async(void* dst, const void* src, size_t bytes, cogconv_stream stream) {
    clear_error();
    if ((!dst || !src) && bytes) return fail(COGCONV_ERR_INVALID_ARG, "null pointer");
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, reinterpret_cast<hipStream_t>(stream)),
                     "hipMemcpyAsync D2D");
}

extern "C" int cogconv_fill_zero(void* dptr, size_t bytes, cogconv_stream stream) {
    clear_error();
    if (!dptr && bytes) return fail(COGCONV_ERR_INVALID_ARG, "null pointer");
    if (bytes == 0) return COGCONV_OK;
    return check_hip(launch_fill_zero(dptr, bytes, reinterpret_cast<hipStream_t>(stream)), "fill_zero");
}

extern "C" int cogconv_host_alloc(size_t bytes, void** hptr) {
    clear_error();
    if (!hptr) return fail(COGCONV_ERR_INVALID_ARG, "null pointer");
    *hptr = nullptr;
    return check_hip(hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
}

extern "C" int cogconv_host_free(void* hptr) {
    clear_error();
    if (!hptr) return COGCONV_OK;
    return check_hip(hipHostFree(hptr), "hipHostFree");
}

extern "C" int cogconv_stream_create(cogconv_stream* stream) {
    clear_error();
    if (!stream) return fail(COGCONV_ERR_INVALID_ARG, "null pointer");
    hipStream_t s = nullptr;
    // Default priority only.
    int e = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    *stream = reinterpret_cast<cogconv_stream>(s);
    return e;
}

extern "C" int cogconv_stream_destroy(cogconv_stream stream) {
    clear_error();
    if (!stream) return COGCONV_OK;
    return check_hip(hipStreamDestroy(reinterpret_cast<hipStream_t>(stream)), "hipStreamDestroy");
}

extern "C" int cogconv_stream_synchronize(cogconv_stream stream) {
    clear_error();
    return check_hip(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)), "hipStreamSynchronize");
}

extern "C" int cogconv_event_create(cogconv_event* event) {
    clear_error();
    if (!event) return fail(COGCONV_ERR_INVALID_ARG, "null pointer");
    hipEvent_t ev = nullptr;
    int e = check_hip(hipEventCreate(&ev), "hipEventCreate");
    *event = reinterpret_cast<cogconv_event>(ev);
    return e;
}

extern "C" int cogconv_event_destroy(cogconv_event event) {
    clear_error();
    if (!event) return COGCONV_OK;
    return check_hip(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)), "hipEventDestroy");
}

extern "C" int cogconv_event_record(cogconv_event event, cogconv_stream stream) {
    clear_error();
    if (!event) return fail(COGCONV_ERR_INVALID_ARG, "null event");
    return check_hip(hipEventRecord(reinterpret_cast<hipEvent_t>(event), reinterpret_cast<hipStream_t>(stream)),
                     "hipEventRecord");
}

extern "C" int cogconv_stream_wait_event(cogconv_stream stream, cogconv_event event) {
    clear_error();
    if (!event) return fail(COGCONV_ERR_INVALID_ARG, "null event");
    return check_hip(
        hipStreamWaitEvent(reinterpret_cast<hipStream_t>(stream), reinterpret_cast<hipEvent_t>(event), 0),
        "hipStreamWaitEvent");
}

extern "C" int cogconv_event_query(cogconv_event event) {
    clear_error();
    if (!event) return fail(COGCONV_ERR_INVALID_ARG, "null event");
    hipError_t e = hipEventQuery(reinterpret_cast<hipEvent_t>(event));
    if (e == hipErrorNotReady) return COGCONV_ERR_NOT_READY;
    return check_hip(e, "hipEventQuery");
}

extern "C" int cogconv_event_synchronize(cogconv_event event) {
    clear_error();
    if (!event) return fail(COGCONV_ERR_INVALID_ARG, "null event");
    return check_hip(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)), "hipEventSynchronize");
}

extern "C" int cogconv_event_elapsed_ms(cogconv_event start, cogconv_event end, float* ms) {
    clear_error();
    if (!start || !end || !ms) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    return check_hip(
        hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(end)),
        "hipEventElapsedTime");
}

extern "C" int cogconv_block_reduce_sum(const float* in, float* out, int64_t points, int32_t in_planes, int32_t factor,
                                        cogconv_stream stream) {
    clear_error();
    if (!in || !out) return fail(COGCONV_ERR_INVALID_ARG, "null buffer");
    if (points < 1 || in_planes < 1 || factor < 1)
        return fail(COGCONV_ERR_INVALID_ARG, "bad blockReduceSum sizes");
    if (in_planes % factor)
        return fail(COGCONV_ERR_ILLEGAL_OP, "tensor length %d is not a multiple of the reduction factor %d", in_planes,
                    factor);
    return check_hip(launch_block_reduce_sum(in, out, points, in_planes / factor, factor,
                                             reinterpret_cast<hipStream_t>(stream)),
                     "blockReduceSum launch");
}

extern "C" int cogconv_block_reduce_sum_epilogue(const float* in, float* out, const cogconv_field_type* in_type,
                                                 int32_t factor, const cogconv_epilogue* epi,
                                                 const float* const* operands, cogconv_stream stream) {
    clear_error();
    if (!in || !out || !in_type) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    const int64_t points = (int64_t)in_type->rows * in_type->cols;
    if (points < 1 || in_type->planes < 1 || factor < 1)
        return fail(COGCONV_ERR_INVALID_ARG, "bad blockReduceSum sizes");
    if (in_type->planes % factor)
        return fail(COGCONV_ERR_ILLEGAL_OP, "tensor length %d is not a multiple of the reduction factor %d",
                    in_type->planes, factor);
    cogconv_field_type t = *in_type;
    t.planes = in_type->planes / factor;
    t.tensor_order = t.planes == 1 ? 0 : 1;
    Epi e;
    int n = 0;
    if (int st = epi_build(t, epi, &e, &n)) return st;
    if (int st = epi_bind(&e, n, operands)) return st;
    return check_hip(launch_block_reduce_sum(in, out, points, t.planes, factor, reinterpret_cast<hipStream_t>(stream),
                                             e),
                     "blockReduceSum launch");
}
