// In-place 3-D complex FFT (FP32) of a batch of (L, L, L) grids in three LDS-tiled passes.
//
// Pass 1 transforms contiguous z lines (rows of L complex), passes 2 and 3 the
// strided y and x lines; a workgroup of the strided passes stages a tile of TW
// adjacent z columns (TW * 8 contiguous bytes per line element) by L elements,
// so every global access is a run of TW complex values and every pass reads and
// writes each grid value once.  Used for the batched FP32 multipole fields in
// place of the rocFFT 3-D plan (which splits the strided axes into transposes).
#include <cmath>
#include <vector>

#include "../../include/hermes_amd.h"
#include "fft_lds.h"
#include "hermes_common.h"

using namespace hermes_fft;

// e^{dir 2 pi i k / 2^j} for j = 0..10, k < 2^j, computed in FP64 and rounded once.
const float2* hermes_twiddles(hermes_ctx* ctx, int log2L, int dir) {
    if (!ctx->twiddles && ctx->capturing) {
        hermes_set_error("FFT twiddle tables are built on first use; run an FFT once before capturing");
        return nullptr;
    }
    if (!ctx->twiddles) {
        std::vector<float2> h((size_t)11 * 2 * 1024);
        for (int j = 0; j <= 10; ++j)
            for (int d = 0; d < 2; ++d) {
                const int L = 1 << j;
                const double sgn = d ? 1.0 : -1.0;
                for (int k = 0; k < L; ++k) {
                    const double a = 2.0 * M_PI * (double)k / (double)L;
                    h[(size_t)(j * 2 + d) * 1024 + k] = make_float2((float)cos(a), (float)(sgn * sin(a)));
                }
            }
        float2* p = nullptr;
        if (hipMalloc(&p, h.size() * sizeof(float2)) != hipSuccess) {
            hermes_set_error("hipMalloc of the twiddle table failed");
            return nullptr;
        }
        if (hipMemcpy(p, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(p);
            hermes_set_error("upload of the twiddle table failed");
            return nullptr;
        }
        ctx->twiddles = p;
    }
    return ctx->twiddles + (size_t)(log2L * 2 + (dir > 0 ? 1 : 0)) * 1024;
}

namespace {

constexpr int FFT_THREADS = 256;

template <int LOG2L>
constexpr int rows_per_block() {
    return (2048 >> LOG2L) > 0 ? (2048 >> LOG2L) : 1;
}

// Pass over contiguous lines: rows_per_block() rows of L per workgroup.
template <int LOG2L, int DIR>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_rows(float2* __restrict__ data, int64_t nrows,
                                                          const float2* __restrict__ twg) {
    constexpr int L = 1 << LOG2L;
    constexpr int NL = rows_per_block<LOG2L>();
    using F = LdsFft<LOG2L, NL, FFT_THREADS, DIR, false>;
    __shared__ float2 lds[F::ELEMS];
    __shared__ float2 tw[L];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * NL;
    load_twiddles<LOG2L>(tw, twg, tid, FFT_THREADS);
    float2* base = data + r0 * L;
    for (int i = tid; i < NL * L; i += FFT_THREADS)
        if (r0 + i / L < nrows) lds[F::addr(i / L, i % L)] = base[i];
    __syncthreads();
    F::run(lds, tw, tid);
    for (int i = tid; i < NL * L; i += FFT_THREADS)
        if (r0 + i / L < nrows) base[i] = lds[F::addr(i / L, i % L)];
}

// Each element of a strided pass is read once and written once, and for grids of 256^3 and up no pass
// finds its input in the caches: the passes use streaming loads and stores (with them the two passes of a
// 512^3 inverse transform take 4.7% less time).
// Pass over strided lines.  Tile t = (outer, mid, zt): the lines at z = zt*TW .. zt*TW+TW-1 of one
// position `mid` on the other strided axis; element e of a line is read at in + e * es_in + mid * ms_in and
// written at out + e * es_out + mid * ms_out (plus the grid and z offsets; < 2^31 within a grid for
// L <= 1024).  in == out with equal geometries is the in-place pass (a workgroup reads its whole tile before
// it writes); otherwise the two grids do not overlap, and the write geometry may transpose the grid.
// LDS (dynamic) holds the L twiddles and TW * L complex values, element-major.
template <int LOG2L, int TW, int NT, int DIR>
__global__ __launch_bounds__(NT) void k_fft_cols(const float2* in, float2* out, const float2* __restrict__ twg,
                                                 int es_in, int64_t ms_in, int es_out, int64_t ms_out, int n_mid,
                                                 int64_t outer_stride, int ntile_z) {
    constexpr int L = 1 << LOG2L;
    constexpr bool STREAM = LOG2L >= 8;
    static_assert(NT % TW == 0, "a thread keeps one z line of the tile");
    using F = LdsFft<LOG2L, TW, NT, DIR, true>;
    extern __shared__ float2 smem[];
    float2* tw = smem;
    float2* lds = smem + L;
    const int tid = threadIdx.x;
    // Workgroups are dealt to the 8 XCDs round-robin; the remap hands each XCD a contiguous range of
    // tiles, so the ntile_z tiles of one line position (which touch the same pages: in the x pass one
    // page per element, 2 MB apart) are translated by one XCD's TLBs instead of eight.
    const int64_t nb = (int64_t)gridDim.x;
    int64_t b = (nb % 8 == 0) ? (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : (int64_t)blockIdx.x;
    const int zt = (int)(b % ntile_z);
    b /= ntile_z;
    const int mid = (int)(b % n_mid);
    const int64_t outer = b / n_mid;
    // this thread's z line is fixed (NT is a multiple of TW); its elements are NT / TW apart
    const int64_t zoff = outer * outer_stride + (zt * TW + tid % TW);
    const float2* src = in + zoff + (int64_t)mid * ms_in + (int64_t)(tid / TW) * es_in;
    float2* dst = out + zoff + (int64_t)mid * ms_out + (int64_t)(tid / TW) * es_out;
    const int64_t sstep = (int64_t)(NT / TW) * es_in, dstep = (int64_t)(NT / TW) * es_out;
    load_twiddles<LOG2L>(tw, twg, tid, NT);
    // every load of the thread is issued before the first LDS store (TW * L / NT of them), so a
    // workgroup keeps its whole tile in flight: the pass is bound by HBM latency otherwise
    constexpr int PER = (TW * L + NT - 1) / NT;
    constexpr bool EVEN = TW * L % NT == 0;  // otherwise (L = 8) threads beyond the tile idle
    float2 v[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (EVEN || tid + k * NT < TW * L) v[k] = STREAM ? hermes_fft::stream_load(src + k * sstep) : src[k * sstep];
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (EVEN || tid + k * NT < TW * L) lds[tid + k * NT] = v[k];
    __syncthreads();
    F::run(lds, tw, tid);
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (EVEN || tid + k * NT < TW * L) {
            if (STREAM)
                hermes_fft::stream_store(dst + k * dstep, lds[tid + k * NT]);
            else
                dst[k * dstep] = lds[tid + k * NT];
        }
}

template <int LOG2L>
constexpr int col_tile() {
    // TW * L = 8192 complex (64 KB) for L >= 128, all z lines for smaller grids
    return (8192 >> LOG2L) < (1 << LOG2L) ? (8192 >> LOG2L) : (1 << LOG2L);
}

// Column-pass tile shapes tried by the autotuner: (TW, threads).  Grids below 128^3 use variant 0 only.
// (TW, threads) per variant: base = col_tile(); 0: (base, 256), 1: (base, 512), 2: (base/2, 256),
// 3: (base/2, 512), 4: (2 base, 512), 5: (2 base, 1024).  Variants 4 and 5 stage 128 KB per workgroup.
constexpr int N_COL_VARIANTS = 6;

__host__ __device__ constexpr int col_tw_of(int base, int v) { return v >= 4 ? 2 * base : ((v & 2) ? base / 2 : base); }
__host__ __device__ constexpr int col_nt_of(int v) { return v >= 4 ? ((v & 1) ? 1024 : 512) : ((v & 1) ? 512 : 256); }

template <int LOG2L, int V>
struct ColVariant {
    static constexpr int TW = col_tw_of(col_tile<LOG2L>(), V);
    static constexpr int NT = col_nt_of(V);
};

// One strided pass of a batch of grids with tile variant V.  Axis 0: y lines of x-major grids (element
// (x, y, z) at (x L + y) L + z), in place; 1: their x lines, in place; 2: y lines from x-major in to y-major
// out (element (x, y, z) at (y L + x) L + z), out of place; 3: x lines of y-major grids, in place.
template <int LOG2L, int V, int DIR>
int launch_cols_variant(hermes_ctx* ctx, int batch, const float2* in, float2* out, const float2* twg, int axis) {
    constexpr int L = 1 << LOG2L;
    constexpr int TW = ColVariant<LOG2L, V>::TW, NT = ColVariant<LOG2L, V>::NT;
    const size_t shmem = sizeof(float2) * (size_t)(L + TW * L);
    auto kc = k_fft_cols<LOG2L, TW, NT, DIR>;
    HERMES_CHECK_HIP(hipFuncSetAttribute((const void*)kc, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    const int ntz = L / TW;
    const unsigned nblk = (unsigned)((int64_t)batch * L * ntz);
    const int64_t LL = (int64_t)L * L, grid = LL * L;
    switch (axis) {
        case 0:  // tiles (b * L + x, zt)
        case 3:  // tiles (b * L + y, zt) of y-major grids: the same strides
            kc<<<nblk, NT, shmem, ctx->stream>>>(in, out, twg, L, LL, L, LL, L, grid, ntz);
            break;
        case 1:  // tiles (b, y, zt)
            kc<<<nblk, NT, shmem, ctx->stream>>>(in, out, twg, (int)LL, L, (int)LL, L, L, grid, ntz);
            break;
        default:  // 2: tiles (b * L + x, zt), lines read along y, written along the y-major grid's y
            kc<<<nblk, NT, shmem, ctx->stream>>>(in, out, twg, L, LL, (int)LL, L, L, grid, ntz);
            break;
    }
    HERMES_LAUNCH_CHECK();
    return 0;
}

template <int LOG2L, int DIR>
int launch_cols(hermes_ctx* ctx, int variant, int batch, const float2* in, float2* out, const float2* twg, int axis) {
    if (LOG2L < 7) return launch_cols_variant<LOG2L, 0, DIR>(ctx, batch, in, out, twg, axis);
    switch (variant) {
        case 1: return launch_cols_variant<LOG2L, 1, DIR>(ctx, batch, in, out, twg, axis);
        case 2: return launch_cols_variant<LOG2L, 2, DIR>(ctx, batch, in, out, twg, axis);
        case 3: return launch_cols_variant<LOG2L, 3, DIR>(ctx, batch, in, out, twg, axis);
        case 4: return launch_cols_variant<LOG2L, 4, DIR>(ctx, batch, in, out, twg, axis);
        case 5: return launch_cols_variant<LOG2L, 5, DIR>(ctx, batch, in, out, twg, axis);
        default: return launch_cols_variant<LOG2L, 0, DIR>(ctx, batch, in, out, twg, axis);
    }
}

static float median3(const float t[3]) {
    return t[0] < t[1] ? (t[1] < t[2] ? t[1] : (t[0] < t[2] ? t[2] : t[0]))
                       : (t[0] < t[2] ? t[0] : (t[1] < t[2] ? t[2] : t[1]));
}

// Autotuner: on the first strided pass of a grid size and direction, time every variant on each axis
// of a scratch batch (median of three passes, HIP events on the context stream) and keep the
// fastest per axis in the context.  The scratch batch is freed afterwards.  Runs eagerly, so the first
// call of a size must not be inside stream capture.
template <int LOG2L, int DIR>
int tuned_variants(hermes_ctx* ctx, const float2* twg, int out[2]) {
    const int d = DIR > 0 ? 1 : 0;
    int* slot = ctx->fft_cols_variant[LOG2L][d];
    if (slot[0] > 0 && slot[1] > 0) {
        out[0] = slot[0] - 1;
        out[1] = slot[1] - 1;
        return 0;
    }
    if (LOG2L < 7) {
        slot[0] = slot[1] = 1;
        out[0] = out[1] = 0;
        return 0;
    }
    HERMES_REQUIRE(!ctx->capturing, "the FFT tile for L=%d is tuned on first use; run it once before capturing",
                   1 << LOG2L);
    constexpr int L = 1 << LOG2L;
    // a batch of about 4 GB (at most 16 grids), as the batched passes run: with one grid of 512^3 the
    // candidates' times differ by less than their tail effects; one grid if that much is not free
    const size_t grid_bytes = sizeof(float2) * (size_t)L * L * L;
    int nf = (int)(((size_t)1 << 32) / grid_bytes);
    nf = nf < 1 ? 1 : (nf > 16 ? 16 : nf);
    float2* scratch = nullptr;
    if (hipMalloc(&scratch, grid_bytes * nf) != hipSuccess) {
        (void)hipGetLastError();
        nf = 1;
        HERMES_CHECK_HIP(hipMalloc(&scratch, grid_bytes));
    }
    HERMES_CHECK_HIP(hipMemsetAsync(scratch, 0, grid_bytes * nf, ctx->stream));
    hipEvent_t e0, e1;
    HERMES_CHECK_HIP(hipEventCreate(&e0));
    HERMES_CHECK_HIP(hipEventCreate(&e1));
    int best[2] = {0, 0};
    float best_ms[2] = {1e30f, 1e30f};
    int rc = 0;
    for (int v = 0; v < N_COL_VARIANTS && rc == 0; ++v) {
        for (int axis = 0; axis < 2 && rc == 0; ++axis) {
            float t[3];
            rc = launch_cols<LOG2L, DIR>(ctx, v, nf, scratch, scratch, twg, axis);  // warm-up
            for (int r = 0; r < 3 && rc == 0; ++r) {
                if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = 1;
                if (rc == 0) rc = launch_cols<LOG2L, DIR>(ctx, v, nf, scratch, scratch, twg, axis);
                if (rc == 0 && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = 1;
                if (rc == 0 && hipEventSynchronize(e1) != hipSuccess) rc = 1;
                if (rc == 0 && hipEventElapsedTime(&t[r], e0, e1) != hipSuccess) rc = 1;
            }
            if (rc) break;
            const float med = median3(t);
            if (med < best_ms[axis]) {
                best_ms[axis] = med;
                best[axis] = v;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(scratch);
    if (rc) {
        if (rc == 1) hermes_set_error("FFT autotuning failed for L=%d", L);
        return rc;
    }
    for (int axis = 0; axis < 2; ++axis) {
        slot[axis] = best[axis] + 1;
        out[axis] = best[axis];
    }
    return 0;
}

// Strided passes (y lines, then x lines) of a batch of grids with the tuned tile shape of each axis
// (the x lines, 2 MB apart, gain from wider z tiles than the y lines).
template <int LOG2L, int DIR>
int run_yx(hermes_ctx* ctx, int batch, float2* data, const float2* twg) {
    int v[2] = {0, 0};
    int rc = tuned_variants<LOG2L, DIR>(ctx, twg, v);
    if (rc) return rc;
    rc = launch_cols<LOG2L, DIR>(ctx, v[0], batch, data, data, twg, 0);
    if (rc) return rc;
    return launch_cols<LOG2L, DIR>(ctx, v[1], batch, data, data, twg, 1);
}

// Strided passes from x-major grids `in` (transformed along z) to y-major grids `out`: the y lines are read
// along x-major rows 4 KB apart (L = 512) and written transposed, and the x lines of the result are then 4 KB
// apart as well; strided writes cost the HBM far less than strided reads.  Both passes take the y lines'
// tuned tile (the same read geometry).
template <int LOG2L, int DIR>
int run_yx_ymajor(hermes_ctx* ctx, int batch, const float2* in, float2* out, const float2* twg) {
    int v[2] = {0, 0};
    int rc = tuned_variants<LOG2L, DIR>(ctx, twg, v);
    if (rc) return rc;
    rc = launch_cols<LOG2L, DIR>(ctx, v[0], batch, in, out, twg, 2);
    if (rc) return rc;
    return launch_cols<LOG2L, DIR>(ctx, v[0], batch, out, out, twg, 3);
}

template <int LOG2L, int DIR>
int run_fft3(hermes_ctx* ctx, int batch, float2* data, bool z_pass) {
    constexpr int L = 1 << LOG2L;
    const float2* twg = hermes_twiddles(ctx, LOG2L, DIR);
    HERMES_REQUIRE(twg, "%s", hermes_last_error());
    if (z_pass) {
        const int64_t nrows = (int64_t)batch * L * L;
        constexpr int NL = rows_per_block<LOG2L>();
        k_fft_rows<LOG2L, DIR><<<(unsigned)((nrows + NL - 1) / NL), FFT_THREADS, 0, ctx->stream>>>(data, nrows, twg);
        HERMES_LAUNCH_CHECK();
    }
    return run_yx<LOG2L, DIR>(ctx, batch, data, twg);
}

template <int DIR>
int dispatch_fft3(hermes_ctx* ctx, int J, int batch, float2* data, bool z_pass) {
    switch (J) {
        case 3: return run_fft3<3, DIR>(ctx, batch, data, z_pass);
        case 4: return run_fft3<4, DIR>(ctx, batch, data, z_pass);
        case 5: return run_fft3<5, DIR>(ctx, batch, data, z_pass);
        case 6: return run_fft3<6, DIR>(ctx, batch, data, z_pass);
        case 7: return run_fft3<7, DIR>(ctx, batch, data, z_pass);
        case 8: return run_fft3<8, DIR>(ctx, batch, data, z_pass);
        case 9: return run_fft3<9, DIR>(ctx, batch, data, z_pass);
        case 10: return run_fft3<10, DIR>(ctx, batch, data, z_pass);
        default: break;
    }
    hermes_set_error("3-D FFT: J=%d outside 3..10", J);
    return 2;
}

// d = (float)(a * eps - b) as a real FP32 grid and as the real part of a complex one.
__global__ void k_delta_f32(const float* __restrict__ eps, int64_t n, double a, double b, float* __restrict__ d,
                            float2* __restrict__ dc) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = (float)(a * (double)eps[i] - b);
        if (d) d[i] = v;
        if (dc) dc[i] = make_float2(v, 0.0f);
    }
}

// (L, L, L/2+1) half spectrum of a real field from its full (L, L, L) transform.
__global__ void k_half_spectrum(const float2* __restrict__ full, int L, float2* __restrict__ half) {
    const int Lh = L / 2 + 1;
    const int64_t n = (int64_t)L * L * Lh;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / Lh;
        const int z = (int)(i - row * Lh);
        half[i] = full[row * L + z];
    }
}

}  // namespace

// Inverse strided passes only: completes a batch whose z lines were already transformed
// (the fused multipole window kernel does the z pass itself).
int hermes_fft3_inverse_yx(hermes_ctx* ctx, int J, int batch, float2* data) {
    return dispatch_fft3<1>(ctx, J, batch, data, false);
}

extern "C" {

int hermes_fft3_inverse_yx_f32(hermes_ctx* ctx, int J, int batch, void* data) {
    HERMES_REQUIRE(ctx && data, "hermes_fft3_inverse_yx_f32: null argument");
    HERMES_REQUIRE(J >= 3 && J <= 10, "hermes_fft3_inverse_yx_f32: J=%d outside 3..10", J);
    HERMES_REQUIRE(batch >= 1, "hermes_fft3_inverse_yx_f32: batch=%d", batch);
    return hermes_fft3_inverse_yx(ctx, J, batch, (float2*)data);
}

int hermes_fft3_inverse_yx_ymajor_f32(hermes_ctx* ctx, int J, int batch, const void* in, void* out) {
    HERMES_REQUIRE(ctx && in && out, "hermes_fft3_inverse_yx_ymajor_f32: null argument");
    HERMES_REQUIRE(J >= 3 && J <= 10, "hermes_fft3_inverse_yx_ymajor_f32: J=%d outside 3..10", J);
    HERMES_REQUIRE(batch >= 1, "hermes_fft3_inverse_yx_ymajor_f32: batch=%d", batch);
    const size_t bytes = sizeof(float2) * ((size_t)batch << (3 * J));
    const char *a = (const char*)in, *b = (const char*)out;
    HERMES_REQUIRE(a + bytes <= b || b + bytes <= a, "hermes_fft3_inverse_yx_ymajor_f32: in and out overlap");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const float2* twg = hermes_twiddles(ctx, J, 1);
    HERMES_REQUIRE(twg, "%s", hermes_last_error());
    const float2* src = (const float2*)in;
    float2* dst = (float2*)out;
    switch (J) {
        case 3: return run_yx_ymajor<3, 1>(ctx, batch, src, dst, twg);
        case 4: return run_yx_ymajor<4, 1>(ctx, batch, src, dst, twg);
        case 5: return run_yx_ymajor<5, 1>(ctx, batch, src, dst, twg);
        case 6: return run_yx_ymajor<6, 1>(ctx, batch, src, dst, twg);
        case 7: return run_yx_ymajor<7, 1>(ctx, batch, src, dst, twg);
        case 8: return run_yx_ymajor<8, 1>(ctx, batch, src, dst, twg);
        case 9: return run_yx_ymajor<9, 1>(ctx, batch, src, dst, twg);
        default: return run_yx_ymajor<10, 1>(ctx, batch, src, dst, twg);
    }
}

int hermes_fft3_tile_choice(hermes_ctx* ctx, int J, int inverse, int* tw_out, int* threads_out) {
    HERMES_REQUIRE(ctx && tw_out && threads_out, "hermes_fft3_tile_choice: null argument");
    HERMES_REQUIRE(J >= 3 && J <= 10, "hermes_fft3_tile_choice: J=%d outside 3..10", J);
    const int v = ctx->fft_cols_variant[J][inverse ? 1 : 0][0] - 1;
    if (v < 0) {
        *tw_out = 0;
        *threads_out = 0;
        return 0;
    }
    const int L = 1 << J;
    const int base = (8192 >> J) < L ? (8192 >> J) : L;
    *tw_out = col_tw_of(base, v);
    *threads_out = col_nt_of(v);
    return 0;
}

int hermes_fft3_set_tile(hermes_ctx* ctx, int J, int inverse, int tw, int threads) {
    HERMES_REQUIRE(ctx, "hermes_fft3_set_tile: null context");
    HERMES_REQUIRE(J >= 3 && J <= 10, "hermes_fft3_set_tile: J=%d outside 3..10", J);
    const int L = 1 << J;
    const int base = (8192 >> J) < L ? (8192 >> J) : L;
    int v = -1;
    for (int c = 0; c < (J < 7 ? 1 : N_COL_VARIANTS); ++c) {
        const int ctw = col_tw_of(base, c), cnt = col_nt_of(c);
        if (ctw == tw && cnt == threads) v = c;
    }
    HERMES_REQUIRE(v >= 0, "hermes_fft3_set_tile: (tw=%d, threads=%d) is not a candidate tile for J=%d", tw, threads,
                   J);
    for (int axis = 0; axis < 2; ++axis) ctx->fft_cols_variant[J][inverse ? 1 : 0][axis] = v + 1;
    return 0;
}

int hermes_delta_f32(hermes_ctx* ctx, int64_t n, const float* eps, double a, double b, float* d_out,
                     void* dc_out) {
    HERMES_REQUIRE(ctx && eps && (d_out || dc_out) && n >= 1, "hermes_delta_f32: invalid argument");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    k_delta_f32<<<hermes_blocks(n, 256, 8192), 256, 0, ctx->stream>>>(eps, n, a, b, d_out, (float2*)dc_out);
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_half_spectrum_f32(hermes_ctx* ctx, int J, const void* full, void* half) {
    HERMES_REQUIRE(ctx && full && half, "hermes_half_spectrum_f32: null argument");
    HERMES_REQUIRE(J >= 1 && J <= 10, "hermes_half_spectrum_f32: J=%d unsupported", J);
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int L = 1 << J;
    const int64_t n = (int64_t)L * L * (L / 2 + 1);
    k_half_spectrum<<<hermes_blocks(n, 256, 8192), 256, 0, ctx->stream>>>((const float2*)full, L, (float2*)half);
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_fft3_c2c_f32(hermes_ctx* ctx, int J, int batch, int inverse, void* data) {
    HERMES_REQUIRE(ctx && data, "hermes_fft3_c2c_f32: null argument");
    HERMES_REQUIRE(batch >= 1, "hermes_fft3_c2c_f32: batch=%d", batch);
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    return inverse ? dispatch_fft3<1>(ctx, J, batch, (float2*)data, true)
                   : dispatch_fft3<-1>(ctx, J, batch, (float2*)data, true);
}

}  // extern "C"
