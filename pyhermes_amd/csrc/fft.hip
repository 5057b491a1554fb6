// 3-D FFTs over (L, L, L) grids with rocFFT; plans cached per context.
// Replaces scipy.fft rfftn/irfftn/fftn/ifftn in convolution.py:284-295.
// Transforms are unnormalised (the inverse is not divided by N); callers fold
// 1/N into the spectral multiply.
#include <rocfft/rocfft.h>

#include <map>
#include <mutex>
#include <tuple>

#include "../../include/hermes_amd.h"
#include "hermes_common.h"

namespace {
std::once_flag g_rocfft_once;

struct Plan {
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    size_t work_bytes = 0;
    void* work = nullptr;
};

// key: (L, kind, double, batch, inplace, interleaved batch)
using PlanKey = std::tuple<int, int, int, int, int, int>;
}  // namespace

struct hermes_fft_cache {
    std::map<PlanKey, Plan> plans;
};

void hermes_fft_cache_destroy(hermes_fft_cache* c) {
    if (!c) return;
    for (auto& kv : c->plans) {
        if (kv.second.info) rocfft_execution_info_destroy(kv.second.info);
        if (kv.second.plan) rocfft_plan_destroy(kv.second.plan);
        if (kv.second.work) (void)hipFree(kv.second.work);
    }
    delete c;
}

namespace {
// kind: 0 = r2c forward, 1 = c2r inverse, 2 = c2c forward, 3 = c2c inverse
// ibatch != 0: complex transforms whose batch index is the fastest axis
// (element stride = batch, batch distance = 1), i.e. data laid out (L, L, L, batch).
int get_plan(hermes_ctx* ctx, int L, int kind, int dbl, int batch, int inplace, int ibatch, Plan** out) {
    std::call_once(g_rocfft_once, []() { rocfft_setup(); });
    if (!ctx->fft) ctx->fft = new hermes_fft_cache();
    PlanKey key{L, kind, dbl, batch, inplace, ibatch};
    auto it = ctx->fft->plans.find(key);
    if (it != ctx->fft->plans.end()) {
        *out = &it->second;
        return 0;
    }
    Plan p;
    size_t lengths[3] = {(size_t)L, (size_t)L, (size_t)L};
    rocfft_transform_type tt = kind == 0   ? rocfft_transform_type_real_forward
                               : kind == 1 ? rocfft_transform_type_real_inverse
                               : kind == 2 ? rocfft_transform_type_complex_forward
                                           : rocfft_transform_type_complex_inverse;
    rocfft_precision prec = dbl ? rocfft_precision_double : rocfft_precision_single;
    rocfft_result_placement pl = inplace ? rocfft_placement_inplace : rocfft_placement_notinplace;
    rocfft_plan_description desc = nullptr;
    if (ibatch) {
        rocfft_plan_description_create(&desc);
        size_t strides[3] = {(size_t)batch, (size_t)batch * L, (size_t)batch * L * L};
        rocfft_plan_description_set_data_layout(desc, rocfft_array_type_complex_interleaved,
                                                rocfft_array_type_complex_interleaved, nullptr, nullptr, 3, strides, 1,
                                                3, strides, 1);
    }
    rocfft_status s = rocfft_plan_create(&p.plan, pl, tt, prec, 3, lengths, (size_t)batch, desc);
    if (desc) rocfft_plan_description_destroy(desc);
    if (s != rocfft_status_success) {
        hermes_set_error("rocfft_plan_create failed (status %d) L=%d kind=%d dbl=%d batch=%d", (int)s, L, kind,
                         dbl, batch);
        return 3;
    }
    rocfft_plan_get_work_buffer_size(p.plan, &p.work_bytes);
    rocfft_execution_info_create(&p.info);
    if (p.work_bytes) {
        if (hipMalloc(&p.work, p.work_bytes) != hipSuccess) {
            hermes_set_error("hipMalloc of %zu bytes for the rocFFT work buffer failed", p.work_bytes);
            return 1;
        }
        rocfft_execution_info_set_work_buffer(p.info, p.work, p.work_bytes);
    }
    auto res = ctx->fft->plans.emplace(key, p);
    *out = &res.first->second;
    return 0;
}

int run(hermes_ctx* ctx, int J, int kind, int dbl, int batch, const void* in, void* out, int ibatch = 0) {
    HERMES_REQUIRE(ctx, "null context");
    HERMES_REQUIRE(J >= 1 && J <= 10, "FFT: J=%d unsupported", J);
    HERMES_REQUIRE(batch >= 1, "FFT: batch must be >= 1");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    int inplace = (in == out) ? 1 : 0;
    HERMES_REQUIRE(!(inplace && (kind == 0 || kind == 1)), "real<->complex FFTs must be out of place");
    Plan* p = nullptr;
    int rc = get_plan(ctx, 1 << J, kind, dbl, batch, inplace, ibatch, &p);
    if (rc) return rc;
    rocfft_execution_info_set_stream(p->info, ctx->stream);
    void* ib[1] = {const_cast<void*>(in)};
    void* ob[1] = {out};
    rocfft_status s = rocfft_execute(p->plan, ib, inplace ? nullptr : ob, p->info);
    if (s != rocfft_status_success) {
        hermes_set_error("rocfft_execute failed (status %d)", (int)s);
        return 3;
    }
    return 0;
}
}  // namespace

extern "C" {

int hermes_fft_r2c(hermes_ctx* ctx, int J, int batch, int dbl, const void* in_real, void* out_half) {
    return run(ctx, J, 0, dbl, batch, in_real, out_half);
}

int hermes_fft_c2r(hermes_ctx* ctx, int J, int batch, int dbl, void* in_half, void* out_real) {
    return run(ctx, J, 1, dbl, batch, in_half, out_real);
}

int hermes_fft_c2c(hermes_ctx* ctx, int J, int batch, int dbl, int inverse, const void* in, void* out) {
    return run(ctx, J, inverse ? 3 : 2, dbl, batch, in, out);
}

}  // extern "C"
