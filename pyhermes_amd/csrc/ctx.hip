// Library context, workspace management and error reporting.
#include <cstdarg>
#include <cstring>

#include "../../include/hermes_amd.h"
#include "hermes_common.h"

static thread_local char g_last_error[1024] = "";

void hermes_set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

void* hermes_workspace(hermes_ctx* ctx, int slot, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (ctx->ws_size[slot] >= bytes) return ctx->ws[slot];
    if (ctx->ws[slot]) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(ctx->ws[slot]);
        ctx->ws[slot] = nullptr;
        ctx->ws_size[slot] = 0;
    }
    size_t grow = bytes + bytes / 4;
    void* p = nullptr;
    if (hipMalloc(&p, grow) != hipSuccess) {
        hermes_set_error("hipMalloc of %zu bytes for workspace slot %d failed", grow, slot);
        return nullptr;
    }
    ctx->ws[slot] = p;
    ctx->ws_size[slot] = grow;
    return p;
}

void hermes_fft_cache_destroy(hermes_fft_cache* c);

extern "C" {

const char* hermes_last_error(void) { return g_last_error; }

int hermes_version(void) { return 100; }

int hermes_ctx_create(int device, hermes_ctx** out) {
    HERMES_REQUIRE(out != nullptr, "hermes_ctx_create: null output pointer");
    int n = 0;
    HERMES_CHECK_HIP(hipGetDeviceCount(&n));
    HERMES_REQUIRE(device >= 0 && device < n, "hermes_ctx_create: device %d out of range (%d devices)",
                   device, n);
    HERMES_CHECK_HIP(hipSetDevice(device));
    hermes_ctx* c = new hermes_ctx();
    c->device = device;
    c->stream = nullptr;
    *out = c;
    return 0;
}

int hermes_ctx_destroy(hermes_ctx* ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 8; ++i)
        if (ctx->ws[i]) (void)hipFree(ctx->ws[i]);
    hermes_fft_cache_destroy(ctx->fft);
    if (ctx->twiddles) (void)hipFree(ctx->twiddles);
    for (int i = 0; i < 2; ++i)
        if (ctx->gram_ev[i]) (void)hipEventDestroy(ctx->gram_ev[i]);
    delete ctx;
    return 0;
}

int hermes_ctx_set_stream(hermes_ctx* ctx, void* stream) {
    HERMES_REQUIRE(ctx != nullptr, "null context");
    ctx->stream = (hipStream_t)stream;
    return 0;
}

int hermes_sync(hermes_ctx* ctx) {
    HERMES_REQUIRE(ctx != nullptr, "null context");
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // extern "C"
