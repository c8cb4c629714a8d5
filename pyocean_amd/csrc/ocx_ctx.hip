// C ABI of the device context: a device, a stream on it, and the device memory and copies a host without HIP of its
// own needs to drive the library (declarations and contracts: include/ocx.h).
#include <hip/hip_runtime.h>

#include <new>

#include "../../include/ocx.h"
#include "ocx_host.hpp"

using namespace ocx_host;

struct ocx_ctx {
    int device;
    hipStream_t stream;
};

namespace {

int bind(ocx_ctx* c) {
    if (!c) return fail(OCX_E_ARG, "ctx is NULL");
    return hip_check(hipSetDevice(c->device), "hipSetDevice");
}

}  // namespace

extern "C" {

int ocx_create(int device, ocx_ctx** out) {
    if (!out) return fail(OCX_E_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    OCX_TRY(hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount"));
    if (device < 0 || device >= n) return fail(OCX_E_ARG, "device %d: %d device(s) visible", device, n);
    OCX_TRY(hip_check(hipSetDevice(device), "hipSetDevice"));
    hipStream_t s = nullptr;
    OCX_TRY(hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"));
    ocx_ctx* c = new (std::nothrow) ocx_ctx{device, s};
    if (!c) {
        (void)hipStreamDestroy(s);
        return fail(OCX_E_ARG, "out of host memory");
    }
    *out = c;
    return OCX_OK;
}

int ocx_destroy(ocx_ctx* ctx) {
    if (!ctx) return OCX_OK;
    OCX_TRY(bind(ctx));
    int rc = hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    const int rc2 = hip_check(hipStreamDestroy(ctx->stream), "hipStreamDestroy");
    delete ctx;
    return rc != OCX_OK ? rc : rc2;
}

void* ocx_ctx_stream(ocx_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int ocx_alloc(ocx_ctx* ctx, size_t bytes, void** dptr) {
    if (!dptr) return fail(OCX_E_ARG, "dptr is NULL");
    *dptr = nullptr;
    OCX_TRY(bind(ctx));
    if (bytes == 0) return OCX_OK;
    return hip_check(hipMalloc(dptr, bytes), "hipMalloc");
}

int ocx_free(ocx_ctx* ctx, void* dptr) {
    OCX_TRY(bind(ctx));
    if (!dptr) return OCX_OK;
    OCX_TRY(hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize"));
    return hip_check(hipFree(dptr), "hipFree");
}

int ocx_h2d(ocx_ctx* ctx, void* dst, const void* src, size_t bytes) {
    OCX_TRY(bind(ctx));
    if (bytes == 0) return OCX_OK;
    if (!dst || !src) return fail(OCX_E_ARG, "dst or src is NULL");
    OCX_TRY(hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy H2D"));
    return hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int ocx_d2h(ocx_ctx* ctx, void* dst, const void* src, size_t bytes) {
    OCX_TRY(bind(ctx));
    if (bytes == 0) return OCX_OK;
    if (!dst || !src) return fail(OCX_E_ARG, "dst or src is NULL");
    OCX_TRY(hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy D2H"));
    return hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int ocx_sync(ocx_ctx* ctx) {
    OCX_TRY(bind(ctx));
    return hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

}  // extern "C"
