// Device, context, memory and event entry points of libgip_hip.
#include <cstdio>
#include <cstring>

#include "gip_internal.h"

namespace gip {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return set_error(e == hipErrorOutOfMemory ? GIP_ENOMEM : GIP_EHIP,
                     std::string(what) + ": " + hipGetErrorString(e));
}

int reduce_grid(int n) {
    int g = (n + kBlock * 4 - 1) / (kBlock * 4);
    if (g < 1) g = 1;
    if (g > kMaxRedBlocks) g = kMaxRedBlocks;
    return g;
}

}  // namespace gip

using namespace gip;

struct gip_event {
    hipEvent_t ev;
    int device;
};

extern "C" {

const char *gip_last_error(void) { return g_last_error.c_str(); }

int gip_device_count(int *count) {
    if (!count) return set_error(GIP_EINVAL, "count is null");
    GIP_HIP_TRY(hipGetDeviceCount(count));
    return GIP_OK;
}

int gip_init(int device, gip_ctx **out) {
    if (!out) return set_error(GIP_EINVAL, "ctx out pointer is null");
    *out = nullptr;
    int n = 0;
    GIP_HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_error(GIP_EINVAL, "device index out of range");
    GIP_HIP_TRY(hipSetDevice(device));
    gip_ctx *c = new (std::nothrow) gip_ctx();
    if (!c) return set_error(GIP_ENOMEM, "context allocation failed");
    c->device = device;
    hipDeviceProp_t prop;
    GIP_HIP_TRY(hipGetDeviceProperties(&prop, device));
    c->num_cu = prop.multiProcessorCount;
    // default-priority, non-blocking stream
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->d_partials, sizeof(double) * kMaxRedBlocks * 4);
    if (e == hipSuccess) e = hipHostMalloc(&c->h_result, sizeof(double) * 4, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&c->h_result_dev, c->h_result, 0);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        int rc = hip_fail(e, "context setup");
        gip_finalize(c);
        return rc;
    }
    int rc = pcg_scratch_setup(device);  // the no-alloc solve's scratch on this device
    if (rc) {
        gip_finalize(c);
        return rc;
    }
    *out = c;
    return GIP_OK;
}

int gip_finalize(gip_ctx *c) {
    if (!c) return GIP_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->d_partials) hipFree(c->d_partials);
    if (c->h_result) hipHostFree(c->h_result);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return GIP_OK;
}

int gip_use_device(int device) {
    GIP_HIP_TRY(hipSetDevice(device));
    return GIP_OK;
}

int gip_sync(gip_ctx *c) {
    if (!c) return set_error(GIP_EINVAL, "null context");
    GIP_HIP_TRY(hipStreamSynchronize(c->stream));
    return GIP_OK;
}

void *gip_stream(gip_ctx *c) { return c ? (void *)c->stream : nullptr; }

int gip_malloc(gip_ctx *c, size_t bytes, void **dptr) {
    if (!c || !dptr) return set_error(GIP_EINVAL, "null argument");
    *dptr = nullptr;
    GIP_HIP_TRY(hipSetDevice(c->device));
    GIP_HIP_TRY(hipMalloc(dptr, bytes ? bytes : 1));
    return GIP_OK;
}

int gip_free(gip_ctx *c, void *dptr) {
    if (!c) return set_error(GIP_EINVAL, "null context");
    if (!dptr) return GIP_OK;
    GIP_HIP_TRY(hipStreamSynchronize(c->stream));
    GIP_HIP_TRY(hipFree(dptr));
    return GIP_OK;
}

int gip_memcpy_h2d(gip_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) return set_error(GIP_EINVAL, "null argument");
    if (!bytes) return GIP_OK;
    GIP_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    GIP_HIP_TRY(hipStreamSynchronize(c->stream));  // the host buffer may be reused at return
    return GIP_OK;
}

int gip_memcpy_d2h(gip_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) return set_error(GIP_EINVAL, "null argument");
    if (!bytes) return GIP_OK;
    GIP_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    GIP_HIP_TRY(hipStreamSynchronize(c->stream));
    return GIP_OK;
}

int gip_memcpy_d2d(gip_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) return set_error(GIP_EINVAL, "null argument");
    if (!bytes) return GIP_OK;
    GIP_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return GIP_OK;
}

int gip_event_create(gip_ctx *c, gip_event **ev) {
    if (!c || !ev) return set_error(GIP_EINVAL, "null argument");
    gip_event *e = new (std::nothrow) gip_event();
    if (!e) return set_error(GIP_ENOMEM, "event allocation failed");
    e->device = c->device;
    hipError_t r = hipEventCreate(&e->ev);
    if (r != hipSuccess) {
        delete e;
        return hip_fail(r, "hipEventCreate");
    }
    *ev = e;
    return GIP_OK;
}

int gip_event_record(gip_ctx *c, gip_event *ev) {
    if (!c || !ev) return set_error(GIP_EINVAL, "null argument");
    GIP_HIP_TRY(hipEventRecord(ev->ev, c->stream));
    return GIP_OK;
}

int gip_event_elapsed_ms(gip_event *a, gip_event *b, float *ms) {
    if (!a || !b || !ms) return set_error(GIP_EINVAL, "null argument");
    GIP_HIP_TRY(hipEventSynchronize(b->ev));
    GIP_HIP_TRY(hipEventElapsedTime(ms, a->ev, b->ev));
    return GIP_OK;
}

int gip_event_destroy(gip_event *ev) {
    if (!ev) return GIP_OK;
    hipEventDestroy(ev->ev);
    delete ev;
    return GIP_OK;
}

}  // extern "C"
