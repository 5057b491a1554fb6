// Small MLP for the Monte-Carlo importance proposal (no reference counterpart).
//
// s(x) = w3 . silu(W2 silu(W1 x + b1) + b2) + b3,  x in R^4, hidden width 64.
// Weights used by the forward pass are bf16 (rounded from FP32 master copies after
// every optimiser step); all arithmetic accumulates in FP32.
// Training loss per sample (y >= 0):  l = y exp(-s) + s   (minimised by exp(s) = E[y | x]),
// with dl/ds clipped to [-clip, 1].  Parameters live in one flat FP32 array:
//   W1 (64x4) | b1 (64) | W2 (64x64) | b2 (64) | w3 (64) | b3 (1)   = 4545 values.
#include <hip/hip_bf16.h>

#include "../../include/hermes_amd.h"
#include "hermes_common.h"

namespace {

constexpr int IN = 4, H = 64;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * IN, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H, NPARAM = OFF_B3 + 1;
constexpr int TILE = 32;  // samples per workgroup
constexpr int THREADS = 256;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float sg = 1.0f / (1.0f + __expf(-z));
    return sg * (1.0f + z * (1.0f - sg));
}

// Forward for a tile of samples; Z1/Z2 pre-activations and H1/H2 activations in LDS.
__device__ void tile_forward(const __hip_bfloat16* __restrict__ wb, const float* __restrict__ x, int n, int t0,
                             float (*Z1)[H], float (*H1)[H], float (*Z2)[H], float (*H2)[H], float* S,
                             float* wcache) {
    const int tid = threadIdx.x;
    for (int i = tid; i < NPARAM; i += blockDim.x) wcache[i] = __bfloat162float(wb[i]);
    __syncthreads();
    for (int e = tid; e < TILE * H; e += blockDim.x) {
        const int i = e / H, j = e % H;
        const int gi = t0 + i;
        float z = wcache[OFF_B1 + j];
        if (gi < n) {
#pragma unroll
            for (int k = 0; k < IN; ++k) z += wcache[OFF_W1 + j * IN + k] * x[(int64_t)gi * IN + k];
        }
        Z1[i][j] = z;
        H1[i][j] = silu(z);
    }
    __syncthreads();
    for (int e = tid; e < TILE * H; e += blockDim.x) {
        const int i = e / H, j = e % H;
        float z = wcache[OFF_B2 + j];
        for (int k = 0; k < H; ++k) z += wcache[OFF_W2 + j * H + k] * H1[i][k];
        Z2[i][j] = z;
        H2[i][j] = silu(z);
    }
    __syncthreads();
    if (tid < TILE) {
        float s = wcache[OFF_B3];
        for (int k = 0; k < H; ++k) s += wcache[OFF_W3 + k] * H2[tid][k];
        S[tid] = s;
    }
    __syncthreads();
}

__global__ __launch_bounds__(THREADS) void k_mlp_forward(const __hip_bfloat16* __restrict__ wb,
                                                         const float* __restrict__ x, int n, float* __restrict__ out) {
    __shared__ float Z1[TILE][H], H1[TILE][H], Z2[TILE][H], H2[TILE][H];
    __shared__ float S[TILE];
    __shared__ float wcache[NPARAM];
    const int t0 = blockIdx.x * TILE;
    tile_forward(wb, x, n, t0, Z1, H1, Z2, H2, S, wcache);
    if (threadIdx.x < TILE && t0 + (int)threadIdx.x < n) out[t0 + threadIdx.x] = S[threadIdx.x];
}

// Forward + backward of the loss for one tile; writes the tile's gradient (NPARAM)
// and loss sum to partial[blockIdx.x].
__global__ __launch_bounds__(THREADS) void k_mlp_grad(const __hip_bfloat16* __restrict__ wb,
                                                      const float* __restrict__ x, const float* __restrict__ y, int n,
                                                      float clip, float* __restrict__ partial) {
    __shared__ float Z1[TILE][H], H1[TILE][H], Z2[TILE][H], H2[TILE][H];
    __shared__ float S[TILE], G[TILE];
    __shared__ float wcache[NPARAM];
    const int tid = threadIdx.x;
    const int t0 = blockIdx.x * TILE;
    float* out = partial + (int64_t)blockIdx.x * (NPARAM + 1);
    tile_forward(wb, x, n, t0, Z1, H1, Z2, H2, S, wcache);
    // dl/ds, zero for padding samples
    if (tid < TILE) {
        const int gi = t0 + tid;
        float g = 0.0f, l = 0.0f;
        if (gi < n) {
            const float e = __expf(-S[tid]);
            const float ye = y[gi] * e;
            g = 1.0f - ye;
            if (g < -clip) g = -clip;
            l = ye + S[tid];
        }
        G[tid] = g;
        S[tid] = l;
    }
    __syncthreads();
    // gradients of layer 3: dw3[k] = sum_i G_i H2[i][k], db3 = sum_i G_i
    for (int k = tid; k < H; k += blockDim.x) {
        float a = 0.0f;
        for (int i = 0; i < TILE; ++i) a += G[i] * H2[i][k];
        out[OFF_W3 + k] = a;
    }
    if (tid == 0) {
        float a = 0.0f, ls = 0.0f;
        for (int i = 0; i < TILE; ++i) {
            a += G[i];
            ls += S[i];
        }
        out[OFF_B3] = a;
        out[NPARAM] = ls;
    }
    __syncthreads();
    // delta2 = G w3 * silu'(Z2)  (overwrite H2 after its use above)
    for (int e = tid; e < TILE * H; e += blockDim.x) {
        const int i = e / H, j = e % H;
        H2[i][j] = G[i] * wcache[OFF_W3 + j] * dsilu(Z2[i][j]);
    }
    __syncthreads();
    // dW2[j][k] = sum_i delta2[i][j] H1[i][k]; db2[j] = sum_i delta2[i][j]
    for (int e = tid; e < H * H; e += blockDim.x) {
        const int j = e / H, k = e % H;
        float a = 0.0f;
        for (int i = 0; i < TILE; ++i) a += H2[i][j] * H1[i][k];
        out[OFF_W2 + e] = a;
    }
    for (int j = tid; j < H; j += blockDim.x) {
        float a = 0.0f;
        for (int i = 0; i < TILE; ++i) a += H2[i][j];
        out[OFF_B2 + j] = a;
    }
    // delta1[i][k] = sum_j delta2[i][j] W2[j][k] * silu'(Z1[i][k])
    for (int e = tid; e < TILE * H; e += blockDim.x) {
        const int i = e / H, k = e % H;
        float a = 0.0f;
        for (int j = 0; j < H; ++j) a += H2[i][j] * wcache[OFF_W2 + j * H + k];
        Z2[i][k] = a * dsilu(Z1[i][k]);
    }
    __syncthreads();
    // dW1[j][k] = sum_i delta1[i][j] x[i][k]; db1[j] = sum_i delta1[i][j]
    for (int e = tid; e < H * IN; e += blockDim.x) {
        const int j = e / IN, k = e % IN;
        float a = 0.0f;
        for (int i = 0; i < TILE; ++i)
            if (t0 + i < n) a += Z2[i][j] * x[(int64_t)(t0 + i) * IN + k];
        out[OFF_W1 + e] = a;
    }
    for (int j = tid; j < H; j += blockDim.x) {
        float a = 0.0f;
        for (int i = 0; i < TILE; ++i) a += Z2[i][j];
        out[OFF_B1 + j] = a;
    }
}

// Tile partials are summed in fixed order in two stages: groups of RED_GROUP consecutive tiles
// (one thread per (group, parameter)), then the group sums; mean gradient and loss -> grad_out.
constexpr int RED_GROUP = 32;

__global__ void k_mlp_reduce_groups(const float* __restrict__ partial, int ntiles, double* __restrict__ gsum) {
    const int ngroups = (ntiles + RED_GROUP - 1) / RED_GROUP;
    const int64_t total = (int64_t)ngroups * (NPARAM + 1);
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int g = (int)(q / (NPARAM + 1)), p = (int)(q % (NPARAM + 1));
        const int t1 = min(ntiles, (g + 1) * RED_GROUP);
        double a = 0.0;
        for (int t = g * RED_GROUP; t < t1; ++t) a += partial[(int64_t)t * (NPARAM + 1) + p];
        gsum[q] = a;
    }
}

__global__ void k_mlp_reduce(const double* __restrict__ gsum, int ngroups, int n, float* __restrict__ grad_out) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p <= NPARAM; p += gridDim.x * blockDim.x) {
        double a = 0.0;
        for (int g = 0; g < ngroups; ++g) a += gsum[(int64_t)g * (NPARAM + 1) + p];
        grad_out[p] = (float)(a / n);
    }
}

// Adam step on the FP32 master weights, then refresh the bf16 copy.
__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, __hip_bfloat16* __restrict__ wb, float lr, float b1, float b2,
                           float eps, int step) {
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < NPARAM; p += gridDim.x * blockDim.x) {
        const float gp = g[p];
        const float mp = b1 * m[p] + (1.0f - b1) * gp;
        const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
        m[p] = mp;
        v[p] = vp;
        const float wn = w[p] - lr * (mp / c1) / (sqrtf(vp / c2) + eps);
        w[p] = wn;
        wb[p] = __float2bfloat16(wn);
    }
}

__global__ void k_to_bf16(const float* __restrict__ w, int n, __hip_bfloat16* __restrict__ wb) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) wb[p] = __float2bfloat16(w[p]);
}

}  // namespace

extern "C" {

int hermes_mlp_nparam(void) { return NPARAM; }

int hermes_mlp_to_bf16(hermes_ctx* ctx, const float* w, void* wb) {
    HERMES_REQUIRE(ctx && w && wb, "hermes_mlp_to_bf16: null argument");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    k_to_bf16<<<hermes_blocks(NPARAM, 256), 256, 0, ctx->stream>>>(w, NPARAM, (__hip_bfloat16*)wb);
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_mlp_forward(hermes_ctx* ctx, const void* wb, const float* x, int n, float* out) {
    HERMES_REQUIRE(ctx && wb && x && out && n >= 0, "hermes_mlp_forward: invalid argument");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    if (n == 0) return 0;
    k_mlp_forward<<<(n + TILE - 1) / TILE, THREADS, 0, ctx->stream>>>((const __hip_bfloat16*)wb, x, n, out);
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_mlp_grad(hermes_ctx* ctx, const void* wb, const float* x, const float* y, int n, float clip,
                    float* grad_out) {
    HERMES_REQUIRE(ctx && wb && x && y && grad_out && n >= 1 && n <= (1 << 18), "hermes_mlp_grad: invalid argument (batch must be 1..262144)");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int ntiles = (n + TILE - 1) / TILE;
    const int ngroups = (ntiles + RED_GROUP - 1) / RED_GROUP;
    const size_t part_bytes = ((sizeof(float) * (size_t)ntiles * (NPARAM + 1) + 255) / 256) * 256;
    char* ws = (char*)hermes_workspace(ctx, WS_MLP, part_bytes + sizeof(double) * (size_t)ngroups * (NPARAM + 1));
    HERMES_REQUIRE(ws, "%s", hermes_last_error());
    float* partial = (float*)ws;
    double* gsum = (double*)(ws + part_bytes);
    k_mlp_grad<<<ntiles, THREADS, 0, ctx->stream>>>((const __hip_bfloat16*)wb, x, y, n, clip, partial);
    HERMES_LAUNCH_CHECK();
    k_mlp_reduce_groups<<<hermes_blocks((int64_t)ngroups * (NPARAM + 1), 256), 256, 0, ctx->stream>>>(partial, ntiles,
                                                                                                     gsum);
    HERMES_LAUNCH_CHECK();
    k_mlp_reduce<<<hermes_blocks(NPARAM + 1, 256), 256, 0, ctx->stream>>>(gsum, ngroups, n, grad_out);
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_mlp_adam(hermes_ctx* ctx, float* w, float* m, float* v, const float* grad, void* wb, float lr, float beta1,
                    float beta2, float eps, int step) {
    HERMES_REQUIRE(ctx && w && m && v && grad && wb && step >= 1, "hermes_mlp_adam: invalid argument");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    k_mlp_adam<<<hermes_blocks(NPARAM, 256), 256, 0, ctx->stream>>>(w, m, v, grad, (__hip_bfloat16*)wb, lr, beta1,
                                                                     beta2, eps, step);
    HERMES_LAUNCH_CHECK();
    return 0;
}

// n_steps Adam steps over consecutive row ranges of (x, y): step s takes rows s * batch .. min((s + 1) * batch,
// n) - 1, step size lr_host[s] and Adam step counter step0 + s + 1; the same launches as hermes_mlp_grad then
// hermes_mlp_adam per step, without a host round trip between steps.
int hermes_mlp_train_steps(hermes_ctx* ctx, float* w, float* m, float* v, float* grad, void* wb, const float* x,
                           const float* y, int64_t n, int batch, float clip, const float* lr_host, int n_steps,
                           int step0, float beta1, float beta2, float eps) {
    HERMES_REQUIRE(ctx && w && m && v && grad && wb && x && y && lr_host && n >= 1 && batch >= 1 &&
                       batch <= (1 << 18) && n_steps >= 0 && step0 >= 0,
                   "hermes_mlp_train_steps: invalid argument");
    HERMES_REQUIRE((int64_t)n_steps * batch < n + batch, "hermes_mlp_train_steps: %d steps of %d rows exceed %lld rows",
                   n_steps, batch, (long long)n);
    for (int s = 0; s < n_steps; ++s) {
        const int64_t off = (int64_t)s * batch;
        const int nb = (int)(n - off < batch ? n - off : batch);
        if (int rc = hermes_mlp_grad(ctx, wb, x + off * IN, y + off, nb, clip, grad)) return rc;
        if (int rc = hermes_mlp_adam(ctx, w, m, v, grad, wb, lr_host[s], beta1, beta2, eps, step0 + s + 1)) return rc;
    }
    return 0;
}

}  // extern "C"
