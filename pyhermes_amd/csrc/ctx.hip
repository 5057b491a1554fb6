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

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

hipError_t hermes_fill_u32(hipStream_t st, void* p, size_t n, uint32_t v) {
    if (n == 0) return hipSuccess;
    size_t nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    k_fill_u32<<<(unsigned)nb, 256, 0, st>>>((uint32_t*)p, n, v);
    return hipGetLastError();
}

void* hermes_workspace(hermes_ctx* ctx, int slot, size_t bytes) {
    if (slot < 0 || slot >= HERMES_WS_SLOTS) {
        hermes_set_error("workspace slot %d outside 0..%d", slot, HERMES_WS_SLOTS - 1);
        return nullptr;
    }
    if (bytes == 0) bytes = 256;
    if (ctx->ws_size[slot] >= bytes) return ctx->ws[slot];
    if (ctx->capturing) {
        hermes_set_error("workspace slot %d would have to grow to %zu bytes during stream capture; run the same "
                         "work once before capturing it",
                         slot, bytes);
        return nullptr;
    }
    if (ctx->ws[slot]) {
        // callers may have used the slot from another stream; growth is rare
        (void)hipDeviceSynchronize();
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

int hermes_read_word(hermes_ctx* ctx, const int* dev, int* out) {
    if (!ctx->host_word) HERMES_CHECK_HIP(hipHostMalloc((void**)&ctx->host_word, 64, hipHostMallocDefault));
    HERMES_CHECK_HIP(hipMemcpyAsync(ctx->host_word, dev, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    hipError_t e;
    while ((e = hipStreamQuery(ctx->stream)) == hipErrorNotReady) {
    }
    HERMES_CHECK_HIP(e);
    *out = *(volatile int*)ctx->host_word;
    return 0;
}

void hermes_fft_cache_destroy(hermes_fft_cache* c);

extern "C" {

const char* hermes_last_error(void) { return g_last_error; }

int hermes_version(void) { return 100; }

int hermes_release_workspaces(hermes_ctx* ctx) {
    HERMES_REQUIRE(ctx, "null context");
    HERMES_REQUIRE(!ctx->capturing, "hermes_release_workspaces: the stream is being captured");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    HERMES_CHECK_HIP(hipDeviceSynchronize());  // the slots may be in use on any stream
    for (int i = 0; i < HERMES_WS_SLOTS; ++i) {
        if (ctx->ws[i]) HERMES_CHECK_HIP(hipFree(ctx->ws[i]));
        ctx->ws[i] = nullptr;
        ctx->ws_size[i] = 0;
    }
    return 0;
}

int hermes_ctx_create(int device, hermes_ctx** out) {
    HERMES_REQUIRE(out != nullptr, "hermes_ctx_create: null output pointer");
    int n = 0;
    HERMES_CHECK_HIP(hipGetDeviceCount(&n));
    HERMES_REQUIRE(device >= 0 && device < n, "hermes_ctx_create: device %d out of range (%d devices)",
                   device, n);
    HERMES_CHECK_HIP(hipSetDevice(device));
    if (int rc = hermes_scratch_setup(device)) return rc;
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
    for (int i = 0; i < HERMES_WS_SLOTS; ++i)
        if (ctx->ws[i]) (void)hipFree(ctx->ws[i]);
    hermes_fft_cache_destroy(ctx->fft);
    if (ctx->twiddles) (void)hipFree(ctx->twiddles);
    if (ctx->dev_err) (void)hipFree(ctx->dev_err);
    if (ctx->host_word) (void)hipHostFree(ctx->host_word);
    for (hipEvent_t e : ctx->gram_ev) (void)hipEventDestroy(e);
    delete ctx;
    return 0;
}

int hermes_ctx_set_stream(hermes_ctx* ctx, void* stream) {
    HERMES_REQUIRE(ctx != nullptr, "null context");
    HERMES_REQUIRE(!ctx->capturing || (hipStream_t)stream == ctx->stream,
                   "hermes_ctx_set_stream: another stream is being captured");
    ctx->stream = (hipStream_t)stream;
    return 0;
}

int hermes_set_field_layout(hermes_ctx* ctx, int layout) {
    HERMES_REQUIRE(ctx != nullptr, "null context");
    HERMES_REQUIRE(layout == 0 || layout == 1, "hermes_set_field_layout: layout %d is neither 0 nor 1", layout);
    ctx->field_layout = layout;
    return 0;
}

int hermes_get_field_layout(hermes_ctx* ctx, int* layout_out) {
    HERMES_REQUIRE(ctx != nullptr && layout_out != nullptr, "hermes_get_field_layout: null argument");
    *layout_out = ctx->field_layout;
    return 0;
}

int hermes_capture_begin(hermes_ctx* ctx) {
    HERMES_REQUIRE(ctx != nullptr, "null context");
    HERMES_REQUIRE(!ctx->capturing, "hermes_capture_begin: already capturing");
    HERMES_REQUIRE(ctx->stream != nullptr,
                   "hermes_capture_begin: install an explicit stream first (the null stream cannot be captured)");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    if (!ctx->dev_err) {
        HERMES_CHECK_HIP(hipMalloc(&ctx->dev_err, 256));
        HERMES_CHECK_HIP(hermes_fill_u32(ctx->stream, ctx->dev_err, 64, 0u));
        HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    }
    HERMES_CHECK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    return 0;
}

int hermes_capture_end(hermes_ctx* ctx, void** exec_out) {
    HERMES_REQUIRE(ctx != nullptr && exec_out != nullptr, "hermes_capture_end: null argument");
    HERMES_REQUIRE(ctx->capturing, "hermes_capture_end: not capturing");
    *exec_out = nullptr;
    ctx->capturing = false;
    hipGraph_t graph = nullptr;
    HERMES_CHECK_HIP(hipStreamEndCapture(ctx->stream, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HERMES_CHECK_HIP(e);
    *exec_out = (void*)exec;
    return 0;
}

int hermes_graph_launch(hermes_ctx* ctx, void* exec) {
    HERMES_REQUIRE(ctx != nullptr && exec != nullptr, "hermes_graph_launch: null argument");
    HERMES_REQUIRE(!ctx->capturing, "hermes_graph_launch: the stream is being captured");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    HERMES_CHECK_HIP(hipGraphLaunch((hipGraphExec_t)exec, ctx->stream));
    return 0;
}

int hermes_graph_destroy(void* exec) {
    if (exec) HERMES_CHECK_HIP(hipGraphExecDestroy((hipGraphExec_t)exec));
    return 0;
}

int hermes_clear_errors(hermes_ctx* ctx) {
    HERMES_REQUIRE(ctx != nullptr, "null context");
    HERMES_REQUIRE(ctx->dev_err != nullptr, "hermes_clear_errors: no error word yet (capture once first)");
    HERMES_CHECK_HIP(hermes_fill_u32(ctx->stream, ctx->dev_err, 1, 0u));
    return 0;
}

int hermes_read_errors(hermes_ctx* ctx, int* out_host) {
    HERMES_REQUIRE(ctx != nullptr && out_host != nullptr, "hermes_read_errors: null argument");
    HERMES_REQUIRE(!ctx->capturing, "hermes_read_errors: the stream is being captured");
    *out_host = 0;
    if (!ctx->dev_err) return 0;
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    HERMES_CHECK_HIP(hipMemcpyAsync(out_host, ctx->dev_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int hermes_sync(hermes_ctx* ctx) {
    HERMES_REQUIRE(ctx != nullptr, "null context");
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // extern "C"
