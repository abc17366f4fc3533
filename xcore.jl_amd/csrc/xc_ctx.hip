// xc_ctx.hip -- context, errors, allocation and scratch for libxcore_hip.
#include <stdlib.h>
#include <string.h>

#include "xc_internal.h"

static thread_local char g_create_err[512] = "";

int xc_fail(xc_ctx* ctx, int code, const char* fmt, ...) {
    char* dst = ctx ? ctx->err : g_create_err;
    size_t n = ctx ? sizeof(ctx->err) : sizeof(g_create_err);
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(dst, n, fmt, ap);
    va_end(ap);
    return code;
}

int xc_finish(xc_ctx* ctx, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return xc_fail(ctx, XC_EHIP, "%s: launch failed: %s", what, hipGetErrorString(e));
    if (ctx->flags & XC_CTX_SYNC) {
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return xc_fail(ctx, XC_EHIP, "%s: %s", what, hipGetErrorString(e));
    }
    return XC_OK;
}

int xc_scratch(xc_ctx* ctx, int slot, size_t bytes, void** out) {
    if (slot < 0 || slot >= XC_MAX_SCRATCH) return xc_fail(ctx, XC_EINVAL_ARG, "scratch slot %d", slot);
    if (ctx->scratch_bytes[slot] < bytes) {
        if (ctx->scratch[slot]) {
            // the old buffer may still be read by queued work
            XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
            XC_TRY_HIP(ctx, hipFree(ctx->scratch[slot]));
            ctx->scratch[slot] = nullptr;
            ctx->scratch_bytes[slot] = 0;
        }
        size_t b = bytes < 256 ? 256 : bytes;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) return xc_fail(ctx, XC_ENOMEM, "scratch %zu bytes: %s", b, hipGetErrorString(e));
        ctx->scratch[slot] = p;
        ctx->scratch_bytes[slot] = b;
    }
    *out = ctx->scratch[slot];
    return XC_OK;
}

extern "C" {

int xc_create(xc_ctx** out, int device, unsigned flags) {
    if (!out) return XC_EINVAL_ARG;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || device < 0 || device >= n)
        return xc_fail(nullptr, XC_EHIP, "xc_create: device %d of %d (%s)", device, n, hipGetErrorString(e));
    e = hipSetDevice(device);
    if (e != hipSuccess) return xc_fail(nullptr, XC_EHIP, "hipSetDevice: %s", hipGetErrorString(e));
    xc_ctx* c = (xc_ctx*)calloc(1, sizeof(xc_ctx));
    if (!c) return XC_ENOMEM;
    c->device = device;
    c->flags = flags;
    // default-priority stream (no priority variants)
    e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        free(c);
        return xc_fail(nullptr, XC_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    c->stream = c->own_stream;
    // set-up of the no-allocation entry point on this device (the address of its scratch array)
    const int st = xc_noalloc_init_device();
    if (st != XC_OK) {
        hipStreamDestroy(c->own_stream);
        free(c);
        return st;
    }
    *out = c;
    return XC_OK;
}

int xc_destroy(xc_ctx* c) {
    if (!c) return XC_OK;
    hipStreamSynchronize(c->stream);
    for (int i = 0; i < XC_MAX_SCRATCH; ++i)
        if (c->scratch[i]) hipFree(c->scratch[i]);
    for (int i = 0; i < 8; ++i)
        if (c->ext[i] && c->ext_free[i]) c->ext_free[i](c, c->ext[i]);
    hipStreamDestroy(c->own_stream);
    free(c);
    return XC_OK;
}

int xc_set_stream(xc_ctx* c, void* s) {
    XC_CHECK_CTX(c);
    c->stream = (hipStream_t)s;  // NULL is the legacy default stream, not the context's own
    return XC_OK;
}

int xc_use_own_stream(xc_ctx* c) {
    XC_CHECK_CTX(c);
    c->stream = c->own_stream;
    return XC_OK;
}

void* xc_get_stream(xc_ctx* c) { return c ? (void*)c->stream : nullptr; }

int xc_set_flags(xc_ctx* c, unsigned flags) {
    XC_CHECK_CTX(c);
    c->flags = flags;
    return XC_OK;
}

int xc_sync(xc_ctx* c) {
    XC_CHECK_CTX(c);
    XC_TRY_HIP(c, hipStreamSynchronize(c->stream));
    return XC_OK;
}

const char* xc_last_error(const xc_ctx* c) { return c ? c->err : g_create_err; }

int xc_malloc(xc_ctx* c, size_t bytes, void** out) {
    XC_CHECK_CTX(c);
    if (!out) return XC_EINVAL_ARG;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 8);
    if (e != hipSuccess) return xc_fail(c, XC_ENOMEM, "xc_malloc(%zu): %s", bytes, hipGetErrorString(e));
    e = hipMemsetAsync(p, 0, bytes, c->stream);
    if (e != hipSuccess) return xc_fail(c, XC_EHIP, "xc_malloc memset: %s", hipGetErrorString(e));
    *out = p;
    return xc_finish(c, "xc_malloc");
}

int xc_free(xc_ctx* c, void* p) {
    XC_CHECK_CTX(c);
    if (!p) return XC_OK;
    XC_TRY_HIP(c, hipStreamSynchronize(c->stream));
    XC_TRY_HIP(c, hipFree(p));
    return XC_OK;
}

}  // extern "C"
