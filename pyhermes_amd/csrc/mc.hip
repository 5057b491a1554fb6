// Monte-Carlo 3PCF kernels.
//
// * hermes_interpolate: n(x) = sum_27 eps[(c - s) & mask] phi phi phi at scaled
//   positions (+ offset), FP64 (wavelet_grid.py:143-186).
// * hermes_mc_triplet_box: the box-random-centre estimator of
//   corr3pcf_kernels.py:100-140 for one (R1, R2, mu): sum over rotations and
//   centres of n1 n2 n3 / (n_rot N), rotation uniforms supplied by the caller,
//   triangle geometry of corr3pcf_kernels.py:21-63, deterministic FP64 sums.
// * hermes_mc_binned: binned triangle sampler over (r1, r2, cos theta) bins with
//   an optional importance proposal (tables built by the host from the learned
//   proposal), Philox4x32-10 counter-based random numbers, per-bin sums of
//   w f, (w f)^2 and counts accumulated with deterministic int64 fixed-point
//   atomics.  Optionally records per-sample features for proposal training.
#include "../../include/hermes_amd.h"
#include "hermes_common.h"

namespace {

constexpr double TWO_PI = 6.283185307179586476925286766559;

__device__ __forceinline__ double interp_at(const double* __restrict__ g, const double* __restrict__ phi, int L,
                                            int res, double sx, double sy, double sz) {
    const int64_t mask = L - 1;
    const int64_t xc = (int64_t)floor(sx), yc = (int64_t)floor(sy), zc = (int64_t)floor(sz);
    // s in (-2^-54, 0) rounds s - floor(s) to 1.0; keep the phi index inside the table
    const int64_t xf = min((int64_t)((sx - (double)xc) * res), (int64_t)res - 1);
    const int64_t yf = min((int64_t)((sy - (double)yc) * res), (int64_t)res - 1);
    const int64_t zf = min((int64_t)((sz - (double)zc) * res), (int64_t)res - 1);
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int64_t xg = (xc - a) & mask;
        const double px = phi[xf + a * res];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int64_t yg = (yc - b) & mask;
            const double py = phi[yf + b * res];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int64_t zg = (zc - c) & mask;
                acc += g[(xg * L + yg) * L + zg] * px * py * phi[zf + c * res];
            }
        }
    }
    return acc;
}

__global__ void k_interpolate(const double* __restrict__ g, const double* __restrict__ spos, int64_t n,
                              const double* __restrict__ phi, int L, int res, double dx, double dy, double dz,
                              double* __restrict__ out) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        out[p] = interp_at(g, phi, L, res, spos[3 * p] + dx, spos[3 * p + 1] + dy, spos[3 * p + 2] + dz);
}

// corr3pcf_kernels.py:21-63
__device__ __forceinline__ void triangle_offsets(double R1, double R2, double mu, double ph, double ct1, double al,
                                                 double o[6]) {
    const double st1 = sqrt(1.0 - ct1 * ct1);
    double sp, cp;
    sincos(ph, &sp, &cp);
    const double dx1 = st1 * cp, dy1 = st1 * sp, dz1 = ct1;
    o[0] = R1 * dx1;
    o[1] = R1 * dy1;
    o[2] = R1 * dz1;
    double ox1, oy1, oz1;
    if (fabs(dx1) > 1e-10 || fabs(dy1) > 1e-10) {
        ox1 = -dy1; oy1 = dx1; oz1 = 0.0;
    } else {
        ox1 = 0.0; oy1 = -dz1; oz1 = dy1;
    }
    const double nrm = sqrt(ox1 * ox1 + oy1 * oy1 + oz1 * oz1);
    ox1 /= nrm; oy1 /= nrm; oz1 /= nrm;
    const double ox2 = dy1 * oz1 - dz1 * oy1, oy2 = dz1 * ox1 - dx1 * oz1, oz2 = dx1 * oy1 - dy1 * ox1;
    double sa, ca;
    sincos(al, &sa, &ca);
    double st2 = 1.0 - mu * mu;
    if (st2 < 0.0) st2 = 0.0;
    const double st = sqrt(st2);
    o[3] = R2 * (mu * dx1 + st * (ca * ox1 + sa * ox2));
    o[4] = R2 * (mu * dy1 + st * (ca * oy1 + sa * oy2));
    o[5] = R2 * (mu * dz1 + st * (ca * oz1 + sa * oz2));
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

constexpr int MC_THREADS = 256;

// grid over (rotation, centre); per-block partial sums
__global__ __launch_bounds__(MC_THREADS) void k_triplet_box(double R1, double R2, double mu,
                                                            const double* __restrict__ spos,
                                                            const double* __restrict__ n1, int64_t n, int n_rot,
                                                            const double* __restrict__ abc,
                                                            const double* __restrict__ g2,
                                                            const double* __restrict__ g3,
                                                            const double* __restrict__ phi, int L, int res,
                                                            double* __restrict__ partial) {
    __shared__ double sh[MC_THREADS / 64];
    const int64_t total = n * (int64_t)n_rot;
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / n);
        const int64_t p = i - (int64_t)r * n;
        double o[6];
        triangle_offsets(R1, R2, mu, TWO_PI * abc[3 * r], 2.0 * abc[3 * r + 1] - 1.0, TWO_PI * abc[3 * r + 2], o);
        const double x = spos[3 * p], y = spos[3 * p + 1], z = spos[3 * p + 2];
        const double v2 = interp_at(g2, phi, L, res, x + o[0], y + o[1], z + o[2]);
        const double v3 = interp_at(g3, phi, L, res, x + o[3], y + o[4], z + o[5]);
        s += n1[p] * v2 * v3;
    }
    s = wave_sum(s);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < MC_THREADS / 64; ++w) t += sh[w];
        partial[blockIdx.x] = t;
    }
}

__global__ void k_sum_1(const double* __restrict__ partial, int nb, double* __restrict__ out) {
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < nb; ++i) t += partial[i];
        out[0] = t;
    }
}

// ---------------------------------------------------------------- Philox4x32-10
struct u4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u4 philox(uint64_t counter, uint64_t key) {
    uint32_t c0 = (uint32_t)counter, c1 = (uint32_t)(counter >> 32), c2 = 0x9E3779B9u, c3 = 0x85EBCA6Bu;
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0, h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
        const uint32_t n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
        c0 = n0; c1 = l1; c2 = n2; c3 = l0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {c0, c1, c2, c3};
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

struct BinnedArgs {
    int L, res, nr1, nr2, nmu;
    double rmin1, rmax1, rmin2, rmax2;  // grid units, uniform (linear) bins
    double scale_wf, scale_wf2;         // fixed-point scales for sum(w f) and sum((w f)^2)
    int64_t n_samples;
    uint64_t seed;
    int importance;
    double alpha;                        // defensive mixture weight of the uniform proposal
    int nbucket;
    int record;                          // write per-sample features for the first n_record samples
    int64_t n_record;
};

__device__ __forceinline__ int upper_bound_f(const float* __restrict__ cdf, int n, float u) {
    int lo = 0, hi = n;  // first index with cdf[i] > u
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    return lo < n ? lo : n - 1;
}

// Two uniforms in [0, 1) from stream `stream` of sample i.
__device__ __forceinline__ void uniforms2(uint64_t i, uint64_t seed, uint64_t stream, double& a, double& b) {
    const u4 r = philox(i, seed ^ (0x9E3779B97F4A7C15ULL * (stream + 1)));
    a = u53(r.x, r.y);
    b = u53(r.z, r.w);
}

__device__ __forceinline__ int cell_of(double cx, double cy, double cz, int L) {
    const int m = L - 1;
    return ((((int)floor(cx)) & m) * L + (((int)floor(cy)) & m)) * L + (((int)floor(cz)) & m);
}

// One sample per iteration: centre, bin (r1, r2, mu), orientation; f = n1 n2 n3.
// Plain: centre uniform in the box, bin uniform.  Importance (defensive mixture):
// with probability alpha as plain, otherwise bucket k ~ bucket_cdf, a uniform cell
// of that bucket, a uniform point in the cell, bin ~ bin_cdf[k].  The weight is
// p / q with p = 1/(V B): w = 1 / (alpha + (1 - alpha) * bucket_density[k] * bin_pdf[k][b]).
__global__ __launch_bounds__(MC_THREADS) void k_mc_binned(
    const double* __restrict__ g1, const double* __restrict__ g2, const double* __restrict__ g3,
    const double* __restrict__ phi, BinnedArgs A, const unsigned char* __restrict__ cell_bucket,
    const int* __restrict__ bucket_cells, const int64_t* __restrict__ bucket_off, const float* __restrict__ bucket_cdf,
    const double* __restrict__ bucket_density, const float* __restrict__ bin_cdf, const float* __restrict__ bin_pdf,
    unsigned long long* __restrict__ acc_wf, unsigned long long* __restrict__ acc_wf2,
    unsigned long long* __restrict__ acc_n, float* __restrict__ rec) {
    const int L = A.L;
    const int B = A.nr1 * A.nr2 * A.nmu;
    const double dr1 = (A.rmax1 - A.rmin1) / A.nr1, dr2 = (A.rmax2 - A.rmin2) / A.nr2, dmu = 2.0 / A.nmu;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < A.n_samples;
         i += (int64_t)gridDim.x * blockDim.x) {
        double ux, uy, uz, ubin, upick, umix, ubin2, w1, w2, w3, oph, oct, oal, spare;
        uniforms2((uint64_t)i, A.seed, 0, ux, uy);
        uniforms2((uint64_t)i, A.seed, 1, uz, ubin);
        uniforms2((uint64_t)i, A.seed, 2, upick, umix);
        uniforms2((uint64_t)i, A.seed, 3, ubin2, w1);
        uniforms2((uint64_t)i, A.seed, 4, w2, w3);
        uniforms2((uint64_t)i, A.seed, 5, oph, oct);
        uniforms2((uint64_t)i, A.seed, 6, oal, spare);
        double cx, cy, cz;
        int bin;
        if (!A.importance || umix < A.alpha) {
            cx = ux * L;
            cy = uy * L;
            cz = uz * L;
            bin = (int)(ubin * B);
            if (bin >= B) bin = B - 1;
        } else {
            const int k = upper_bound_f(bucket_cdf, A.nbucket, (float)ubin);
            const int64_t c0 = bucket_off[k], c1 = bucket_off[k + 1];
            int64_t pick = c0 + (int64_t)(upick * (double)(c1 - c0));
            if (pick >= c1) pick = c1 - 1;
            const int cell = bucket_cells[pick];
            const int zc = cell % L, yc = (cell / L) % L, xc = cell / (L * L);
            cx = xc + ux;
            cy = yc + uy;
            cz = zc + uz;
            bin = upper_bound_f(bin_cdf + (int64_t)k * B, B, (float)ubin2);
        }
        const int imu = bin % A.nmu, ir2 = (bin / A.nmu) % A.nr2, ir1 = bin / (A.nmu * A.nr2);
        const double r1 = A.rmin1 + (ir1 + w1) * dr1, r2 = A.rmin2 + (ir2 + w2) * dr2;
        const double mu = -1.0 + (imu + w3) * dmu;
        double o[6];
        triangle_offsets(r1, r2, mu, TWO_PI * oph, 2.0 * oct - 1.0, TWO_PI * oal, o);
        const double n1 = interp_at(g1, phi, L, A.res, cx, cy, cz);
        const double n2 = interp_at(g2, phi, L, A.res, cx + o[0], cy + o[1], cz + o[2]);
        const double n3 = interp_at(g3, phi, L, A.res, cx + o[3], cy + o[4], cz + o[5]);
        const double f = n1 * n2 * n3;
        const int cell = cell_of(cx, cy, cz, L);
        double w = 1.0;
        if (A.importance) {
            const int k = cell_bucket[cell];
            const double rel = bucket_density[k] * (double)bin_pdf[(int64_t)k * B + bin];
            w = 1.0 / (A.alpha + (1.0 - A.alpha) * rel);
        }
        const double wf = w * f;
        atomicAdd(&acc_wf[bin], (unsigned long long)__double2ll_rn(wf * A.scale_wf));
        atomicAdd(&acc_wf2[bin], (unsigned long long)__double2ll_rn(wf * wf * A.scale_wf2));
        atomicAdd(&acc_n[bin], 1ull);
        if (A.record && i < A.n_record) {
            float* rr = rec + 6 * i;
            rr[0] = (float)r1;
            rr[1] = (float)r2;
            rr[2] = (float)mu;
            rr[3] = (float)g1[cell];
            rr[4] = (float)f;
            rr[5] = (float)bin;
        }
        (void)spare;
    }
}

}  // namespace

extern "C" {

int hermes_interpolate(hermes_ctx* ctx, int J, const double* grid, const double* spos, int64_t n, const double* phi,
                       int phi_res, double dx, double dy, double dz, double* out) {
    HERMES_REQUIRE(ctx && grid && phi && (n == 0 || (spos && out)), "hermes_interpolate: null argument");
    HERMES_REQUIRE(J >= 2 && J <= 10, "J=%d unsupported", J);
    HERMES_REQUIRE(phi_res == 1024, "phi_res must be 1024");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    if (n == 0) return 0;
    k_interpolate<<<hermes_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(grid, spos, n, phi, 1 << J, phi_res, dx, dy,
                                                                         dz, out);
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_mc_triplet_box(hermes_ctx* ctx, int J, double R1, double R2, double mu, const double* spos,
                          const double* n1, int64_t n, int n_rot, const double* abc_host, const double* g2,
                          const double* g3, const double* phi, int phi_res, double* out_host) {
    HERMES_REQUIRE(ctx && spos && n1 && abc_host && g2 && g3 && phi && out_host, "hermes_mc_triplet_box: null arg");
    HERMES_REQUIRE(J >= 2 && J <= 10, "J=%d unsupported", J);
    HERMES_REQUIRE(n > 0 && n_rot > 0, "need centres and rotations");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int64_t total = n * (int64_t)n_rot;
    const int nb = (int)hermes_blocks(total, MC_THREADS, 8192);
    double* ws = (double*)hermes_workspace(ctx, WS_MC, sizeof(double) * ((size_t)nb + 8 + 3 * (size_t)n_rot));
    HERMES_REQUIRE(ws, "%s", hermes_last_error());
    double* d_abc = ws + nb + 8;
    HERMES_CHECK_HIP(hipMemcpyAsync(d_abc, abc_host, sizeof(double) * 3 * n_rot, hipMemcpyHostToDevice, ctx->stream));
    k_triplet_box<<<nb, MC_THREADS, 0, ctx->stream>>>(R1, R2, mu, spos, n1, n, n_rot, d_abc, g2, g3, phi, 1 << J,
                                                      phi_res, ws);
    HERMES_LAUNCH_CHECK();
    k_sum_1<<<1, 64, 0, ctx->stream>>>(ws, nb, ws + nb);
    HERMES_LAUNCH_CHECK();
    double s = 0.0;
    HERMES_CHECK_HIP(hipMemcpyAsync(&s, ws + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    *out_host = s / ((double)n_rot * (double)n);
    return 0;
}

int hermes_mc_binned(hermes_ctx* ctx, int J, const double* g1, const double* g2, const double* g3, const double* phi,
                     int phi_res, int nr1, int nr2, int nmu, double rmin1, double rmax1, double rmin2, double rmax2,
                     int64_t n_samples, uint64_t seed, double wf_bound, const hermes_mc_proposal* prop, float* record,
                     int64_t n_record, double* sum_wf_host, double* sum_wf2_host, int64_t* count_host) {
    HERMES_REQUIRE(ctx && g1 && g2 && g3 && phi && sum_wf_host && sum_wf2_host && count_host,
                   "hermes_mc_binned: null argument");
    HERMES_REQUIRE(J >= 2 && J <= 10, "J=%d unsupported", J);
    HERMES_REQUIRE(phi_res == 1024, "phi_res must be 1024");
    HERMES_REQUIRE(nr1 >= 1 && nr2 >= 1 && nmu >= 1 && (int64_t)nr1 * nr2 * nmu <= (1 << 20), "invalid binning");
    HERMES_REQUIRE(rmax1 > rmin1 && rmax2 > rmin2 && rmin1 >= 0.0 && rmin2 >= 0.0, "invalid radial ranges");
    HERMES_REQUIRE(n_samples >= 1 && wf_bound > 0.0, "n_samples and wf_bound must be positive");
    HERMES_REQUIRE(record == nullptr || n_record >= 0, "invalid record count");
    if (prop) {
        HERMES_REQUIRE(prop->cell_bucket && prop->bucket_cells && prop->bucket_off && prop->bucket_cdf &&
                           prop->bucket_density && prop->bin_cdf && prop->bin_pdf,
                       "hermes_mc_binned: incomplete proposal tables");
        HERMES_REQUIRE(prop->nbucket >= 1 && prop->nbucket <= 256, "nbucket must be in 1..256");
        HERMES_REQUIRE(prop->alpha > 0.0 && prop->alpha <= 1.0, "alpha must be in (0, 1]");
    }
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int B = nr1 * nr2 * nmu;
    BinnedArgs A;
    A.L = 1 << J;
    A.res = phi_res;
    A.nr1 = nr1;
    A.nr2 = nr2;
    A.nmu = nmu;
    A.rmin1 = rmin1;
    A.rmax1 = rmax1;
    A.rmin2 = rmin2;
    A.rmax2 = rmax2;
    A.n_samples = n_samples;
    A.seed = seed;
    A.importance = prop != nullptr;
    A.alpha = prop ? prop->alpha : 1.0;
    A.nbucket = prop ? prop->nbucket : 0;
    A.record = record != nullptr;
    A.n_record = record ? n_record : 0;
    // |w f| <= wf_bound for every sample; each bin sum stays below 2^62 after scaling
    A.scale_wf = ldexp(1.0, 62) / ((double)n_samples * wf_bound);
    A.scale_wf2 = ldexp(1.0, 62) / ((double)n_samples * wf_bound * wf_bound);
    unsigned long long* acc =
        (unsigned long long*)hermes_workspace(ctx, WS_MC, sizeof(unsigned long long) * 3 * (size_t)B);
    HERMES_REQUIRE(acc, "%s", hermes_last_error());
    HERMES_CHECK_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long) * 3 * (size_t)B, ctx->stream));
    k_mc_binned<<<hermes_blocks(n_samples, MC_THREADS, 16384), MC_THREADS, 0, ctx->stream>>>(
        g1, g2, g3, phi, A, prop ? prop->cell_bucket : nullptr, prop ? prop->bucket_cells : nullptr,
        prop ? prop->bucket_off : nullptr, prop ? prop->bucket_cdf : nullptr, prop ? prop->bucket_density : nullptr,
        prop ? prop->bin_cdf : nullptr, prop ? prop->bin_pdf : nullptr, acc, acc + B, acc + 2 * B, record);
    HERMES_LAUNCH_CHECK();
    std::vector<long long> h(3 * (size_t)B);
    HERMES_CHECK_HIP(hipMemcpyAsync(h.data(), acc, sizeof(long long) * 3 * (size_t)B, hipMemcpyDeviceToHost,
                                    ctx->stream));
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < B; ++b) {
        sum_wf_host[b] = (double)h[b] / A.scale_wf;
        sum_wf2_host[b] = (double)h[B + b] / A.scale_wf2;
        count_host[b] = h[2 * B + b];
    }
    return 0;
}

}  // extern "C"
