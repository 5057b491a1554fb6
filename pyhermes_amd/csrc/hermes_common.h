// Shared definitions for the MI355X (gfx950) field-pipeline library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

struct hermes_fft_cache;

// Workspace slots (each slot is an independent grow-only buffer).
enum hermes_ws_slot {
    WS_PROJ_TILE = 0,   // projection: per-tile counters / offsets
    WS_PROJ_REC = 1,    // projection: bucketed particle records
    WS_REDUCE = 2,      // reductions: per-block partial sums
    WS_SPEC = 3,        // window / FFT staging
    WS_MISC = 4,
    WS_MC = 5,
    WS_MLP = 6,
    WS_CONTRACT = 7,
    WS_PROFILE = 8,     // radial profiles of the fused window kernel
};
constexpr int HERMES_WS_SLOTS = 9;

// One library context per device: stream, grow-only workspaces and FFT plans.
struct hermes_ctx {
    int device = 0;
    hipStream_t stream = nullptr;  // default stream unless the caller installs one
    // grow-only scratch buffers owned by the context
    void* ws[HERMES_WS_SLOTS] = {nullptr};
    size_t ws_size[HERMES_WS_SLOTS] = {0};
    hermes_fft_cache* fft = nullptr;
    float2* twiddles = nullptr;  // e^{+-2 pi i k / 2^j} tables for the LDS FFTs (first use)
    // autotuned strided-pass tile per (log2 L, inverse, axis: 0 = y lines, 1 = x lines); 0 = not tuned
    int fft_cols_variant[11][2][2] = {};
    // layout of the FP32 fields the Gram contractions read (hermes_set_field_layout): 0 = x-major, 1 = y-major
    int field_layout = 0;
    // optional HIP-event timing of the Gram contraction kernel (hermes_profile_gram)
    bool time_gram = false;
    std::vector<hipEvent_t> gram_ev;  // (start, end) per timed launch, summed when queried
    size_t gram_pending = 0;          // recorded pairs not yet summed
    double gram_ms = 0.0;
    int64_t gram_calls = 0;
    // stream capture (hermes_capture_begin/end): while set, nothing may allocate, copy to or
    // from the host, memset or synchronise; kernels report input errors into dev_err
    bool capturing = false;
    int* dev_err = nullptr;  // sticky device error word (see hermes_read_errors)
    int* host_word = nullptr;  // pinned host word for hermes_read_word (allocated on first use)
};


void hermes_set_error(const char* fmt, ...);
// Fill n 32-bit words with v on the stream (a kernel, so it can be captured into graphs).
hipError_t hermes_fill_u32(hipStream_t st, void* p, size_t n, uint32_t v);
void* hermes_workspace(hermes_ctx* ctx, int slot, size_t bytes);
// Resolves, on the current device, the device address of the allocation-free pair reduction's fixed
// scratch array (window.hip); called once per device by hermes_ctx_create.
int hermes_scratch_setup(int device);
// Reads one device int after the stream's preceding work, waiting by polling the stream (a blocking
// wait would add the host thread's wake-up latency to every eager call).  Not while capturing.
int hermes_read_word(hermes_ctx* ctx, const int* dev, int* out);
// Twiddle table e^{dir 2 pi i k / 2^log2L}, k < 2^log2L (device memory owned by ctx).
const float2* hermes_twiddles(hermes_ctx* ctx, int log2L, int dir);
// Inverse FFT along y and x of a batch of (L, L, L) complex64 grids already transformed along z.
int hermes_fft3_inverse_yx(hermes_ctx* ctx, int J, int batch, float2* data);

#define HERMES_CHECK_HIP(expr)                                                            \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            hermes_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                             __LINE__);                                                    \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

#define HERMES_REQUIRE(cond, ...)          \
    do {                                   \
        if (!(cond)) {                     \
            hermes_set_error(__VA_ARGS__); \
            return 2;                      \
        }                                  \
    } while (0)

#define HERMES_LAUNCH_CHECK()                                                          \
    do {                                                                               \
        hipError_t _e = hipGetLastError();                                             \
        if (_e != hipSuccess) {                                                        \
            hermes_set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                      \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

static inline unsigned int hermes_blocks(int64_t n, int threads, int64_t cap = 1 << 20) {
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned int)b;
}
