// Spherical-harmonic 3PCF triple-product contraction.
//
// Plain FP64 path (faithful to compute_3d_result_gpu + reduce_complex_sum_kernel,
// corr3pcf_multipoles.py:33-58, :99-131):
//   S = sum_x d1[x] sum_{l1} F1[x-l1] sum_{l2} G3(l1,l2) F2[x-l2]   (complex128)
// fused with a deterministic FP64 reduction.
//
// Batched FP32 path (MFMA): with Gamma = Q diag(lam) Q^T,
//   sum_{l1,l2} G3(l1,l2) U[x-l1] V[x-l2] = sum_{abc} lam_abc U~_abc[x] V~_abc[x],
// U~_abc = separable 3-tap filter of U with Q's columns.  For 16 shells at once
//   G[s1][s2] = sum_x d1[x] sum_{abc,comp} lam_abc U~_{s1,abc,comp}[x] V~_{s2,abc,comp}[x]
// is a 16 x 16 Gram matrix accumulated with v_mfma_f32_16x16x4f32 (lane = shell
// x cell), flushed to FP64 after every x plane; per-workgroup FP64 partials are
// summed in fixed order.  comp runs over (re, im) so G = Re(U V) sums when V is
// given as (V.re, -V.im).
#include "../../include/hermes_amd.h"
#include "hermes_common.h"

namespace {

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------- plain FP64
constexpr int SUM_THREADS = 256;

__global__ __launch_bounds__(SUM_THREADS) void k_summand_f64(int L, const double* __restrict__ d1,
                                                             const double2* __restrict__ F1,
                                                             const double2* __restrict__ F2, const double* __restrict__ G,
                                                             double* __restrict__ partial) {
    __shared__ double sh[2][SUM_THREADS / 64];
    const int64_t N = (int64_t)L * L * L;
    double tr = 0.0, ti = 0.0;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < N; idx += (int64_t)gridDim.x * blockDim.x) {
        const int lz = (int)(idx % L), ly = (int)((idx / L) % L), lx = (int)(idx / ((int64_t)L * L));
        double s1r = 0.0, s1i = 0.0;
        for (int l1x = 0; l1x < 3; ++l1x) {
            const int i1x = (lx - l1x + L) % L;
            for (int l1y = 0; l1y < 3; ++l1y) {
                const int i1y = (ly - l1y + L) % L;
                for (int l1z = 0; l1z < 3; ++l1z) {
                    const int i1z = (lz - l1z + L) % L;
                    double s2r = 0.0, s2i = 0.0;
                    for (int l2x = 0; l2x < 3; ++l2x) {
                        const int i2x = (lx - l2x + L) % L;
                        double ryr = 0.0, ryi = 0.0;
                        for (int l2y = 0; l2y < 3; ++l2y) {
                            const int i2y = (ly - l2y + L) % L;
                            double rzr = 0.0, rzi = 0.0;
                            for (int l2z = 0; l2z < 3; ++l2z) {
                                const int i2z = (lz - l2z + L) % L;
                                const double2 f = F2[((int64_t)i2x * L + i2y) * L + i2z];
                                const double g = G[l1z * 3 + l2z];
                                rzr += g * f.x;
                                rzi += g * f.y;
                            }
                            const double g = G[l1y * 3 + l2y];
                            ryr += g * rzr;
                            ryi += g * rzi;
                        }
                        const double g = G[l1x * 3 + l2x];
                        s2r += g * ryr;
                        s2i += g * ryi;
                    }
                    const double2 a = F1[((int64_t)i1x * L + i1y) * L + i1z];
                    s1r += a.x * s2r - a.y * s2i;
                    s1i += a.x * s2i + a.y * s2r;
                }
            }
        }
        const double d = d1[idx];
        tr += d * s1r;
        ti += d * s1i;
    }
    tr = wave_sum_d(tr);
    ti = wave_sum_d(ti);
    if (threadIdx.x % 64 == 0) {
        sh[0][threadIdx.x / 64] = tr;
        sh[1][threadIdx.x / 64] = ti;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double t = 0.0;
        for (int w = 0; w < SUM_THREADS / 64; ++w) t += sh[threadIdx.x][w];
        partial[2 * blockIdx.x + threadIdx.x] = t;
    }
}

// Column sums of a row-major (nb x width) FP64 array, width a divisor of RED_THREADS.  Block b
// sums rows [b * rows_per_block, (b + 1) * rows_per_block) in a fixed order (thread phase p takes
// rows p, p + phases, ...; phases are then added in index order), so results are run-to-run identical.
constexpr int RED_THREADS = 256;

__global__ __launch_bounds__(RED_THREADS) void k_sum_rows(const double* __restrict__ in, int64_t nb, int width,
                                                          int64_t stride, int64_t rows_per_block,
                                                          double* __restrict__ out) {
    __shared__ double sh[RED_THREADS];
    const int phases = RED_THREADS / width;
    const int j = threadIdx.x % width, ph = threadIdx.x / width;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r1 = r0 + rows_per_block < nb ? r0 + rows_per_block : nb;
    double t = 0.0;
    if (ph < phases)
        for (int64_t r = r0 + ph; r < r1; r += phases) t += in[r * stride + j];
    sh[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x < width) {
        double s = 0.0;
        for (int p = 0; p < phases; ++p) s += sh[p * width + threadIdx.x];
        out[(int64_t)blockIdx.x * width + threadIdx.x] = s;
    }
}

constexpr int RED_MAX_BLOCKS = 1024;

// Two launches: <= RED_MAX_BLOCKS partial rows, then one row.  tmp holds RED_MAX_BLOCKS * width doubles.
// Column sums of nb rows (width <= RED_THREADS columns, rows stride apart) in a fixed order.
static void sum_rows(hipStream_t st, const double* in, int64_t nb, int width, double* tmp, double* out,
                     int64_t stride = 0) {
    if (stride == 0) stride = width;
    int64_t rpb = (nb + RED_MAX_BLOCKS - 1) / RED_MAX_BLOCKS;
    if (rpb < 16) rpb = 16;
    const int64_t nblk = (nb + rpb - 1) / rpb;
    k_sum_rows<<<(unsigned)nblk, RED_THREADS, 0, st>>>(in, nb, width, stride, rpb, tmp);
    k_sum_rows<<<1, RED_THREADS, 0, st>>>(tmp, nblk, width, width, nblk, out);
}

// ---------------------------------------------------------------- batched FP32 MFMA Gram
struct GramArgs {
    int L, S1, S2, cross, xseg;
    int xs, ys;       // element strides of x and y in the fields U, V (x-major: L^2, L; y-major: L, L^2)
    float vsign;      // sign applied to V's imaginary part in the cross case (-1: Re(U V))
    float Q[9];       // Q[tap * 3 + col]
    float lam[27];    // lam_a lam_b lam_c, index (a * 3 + b) * 3 + c
    float lamyz[9];   // lam_b lam_c, index b * 3 + c (auto kernel)
    float gx[9];      // Gamma[l1 * 3 + l2] = sum_a Q[l1][a] lam_a Q[l2][a] (auto kernel)
    float Qs[9];      // auto kernel: Q[tap][col] sqrt|lam_col|, columns sorted by ascending lam
};

constexpr int GRAM_WAVES = 4;

typedef float float4_t __attribute__((ext_vector_type(4)));

// Raw values of field f at plane x around cell (y, z): raw[ly * 3 + lz] = f[x, y-ly, z-lz].
__device__ __forceinline__ void load_plane(const float2* __restrict__ f, const GramArgs& g, int x, int y, int z,
                                           float2 raw[9]) {
    const int m = g.L - 1;
#pragma unroll
    for (int ly = 0; ly < 3; ++ly) {
        const float2* row = f + (int64_t)x * g.xs + (int64_t)((y - ly) & m) * g.ys;
#pragma unroll
        for (int lz = 0; lz < 3; ++lz) raw[ly * 3 + lz] = row[(z - lz) & m];
    }
}

// yz-filter: Y[b * 3 + c] = sum_{ly, lz} Q[ly][b] Q[lz][c] raw[ly * 3 + lz].
__device__ __forceinline__ void filter_plane(const float2 raw[9], const float* Q, float2 Y[9]) {
    float2 Z[3][3];  // Z[ly][c]
#pragma unroll
    for (int ly = 0; ly < 3; ++ly)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float q0 = Q[0 * 3 + c], q1 = Q[1 * 3 + c], q2 = Q[2 * 3 + c];
            const float2 u0 = raw[ly * 3 + 0], u1 = raw[ly * 3 + 1], u2 = raw[ly * 3 + 2];
            Z[ly][c] = make_float2(q0 * u0.x + q1 * u1.x + q2 * u2.x, q0 * u0.y + q1 * u1.y + q2 * u2.y);
        }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const float q0 = Q[0 * 3 + b], q1 = Q[1 * 3 + b], q2 = Q[2 * 3 + b];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            Y[b * 3 + c] = make_float2(q0 * Z[0][c].x + q1 * Z[1][c].x + q2 * Z[2][c].x,
                                       q0 * Z[0][c].y + q1 * Z[1][c].y + q2 * Z[2][c].y);
    }
}

// One x plane: x-filter of the three yz-filtered planes and 54 MFMA updates.
template <bool CROSS>
__device__ __forceinline__ void gram_plane(const float2 Yc[9], const float2 Y1[9], const float2 Y2[9],
                                           const float2 Vc[9], const float2 V1[9], const float2 V2[9],
                                           const float* Q, const float* lam, float dval, float vsign,
                                           float4_t& accr, float4_t& acci) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float q0 = Q[0 * 3 + a], q1 = Q[1 * 3 + a], q2 = Q[2 * 3 + a];
#pragma unroll
        for (int bc = 0; bc < 9; ++bc) {
            const float2 ut = make_float2(q0 * Yc[bc].x + q1 * Y1[bc].x + q2 * Y2[bc].x,
                                          q0 * Yc[bc].y + q1 * Y1[bc].y + q2 * Y2[bc].y);
            float2 vt;
            if (CROSS)
                vt = make_float2(q0 * Vc[bc].x + q1 * V1[bc].x + q2 * V2[bc].x,
                                 vsign * (q0 * Vc[bc].y + q1 * V1[bc].y + q2 * V2[bc].y));
            else
                vt = ut;
            const float dl = dval * lam[a * 9 + bc];
            accr = __builtin_amdgcn_mfma_f32_16x16x4f32(dl * ut.x, vt.x, accr, 0, 0, 0);
            acci = __builtin_amdgcn_mfma_f32_16x16x4f32(dl * ut.y, vt.y, acci, 0, 0, 0);
        }
    }
}

template <bool CROSS>
__global__ __launch_bounds__(64 * GRAM_WAVES) void k_gram_f32(const float* __restrict__ d1,
                                                              const float2* __restrict__ U,
                                                              const float2* __restrict__ V, GramArgs g,
                                                              double* __restrict__ partial) {
    __shared__ double red[GRAM_WAVES][256];
    const int L = g.L;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int s = lane % 16, k = lane / 16;
    const int nzg = L / 4, nyg = L / GRAM_WAVES;
    const int zg = blockIdx.x % nzg;
    const int yg = (blockIdx.x / nzg) % nyg;
    const int xs = blockIdx.x / (nzg * nyg);
    const int y = yg * GRAM_WAVES + wave;
    const int z = zg * 4 + k;
    const int x0 = xs * g.xseg, xend = x0 + g.xseg;
    const int64_t N = (int64_t)L * L * L;
    const bool ua = s < g.S1;
    const bool vb = s < g.S2;
    const float2* fu = U + (int64_t)(ua ? s : 0) * N;
    const float2* fv = (CROSS ? V : U) + (int64_t)(vb ? s : 0) * N;
    float Q[9], lam[27];
#pragma unroll
    for (int i = 0; i < 9; ++i) Q[i] = g.Q[i];
#pragma unroll
    for (int i = 0; i < 27; ++i) lam[i] = g.lam[i];
    const float2 zero = make_float2(0.f, 0.f);

    float2 rawU[9], rawV[9];
    float2 P0[9], P1[9], P2[9], W0[9], W1[9], W2[9];
    // warm-up: P1 = Y(x0-2), P2 = Y(x0-1)
#pragma unroll
    for (int i = 0; i < 9; ++i) { rawU[i] = zero; rawV[i] = zero; }
    if (ua) load_plane(fu, g, (x0 - 2) & (L - 1), y, z, rawU);
    if (CROSS && vb) load_plane(fv, g, (x0 - 2) & (L - 1), y, z, rawV);
    filter_plane(rawU, Q, P1);
    if (CROSS) filter_plane(rawV, Q, W1);
    if (ua) load_plane(fu, g, (x0 - 1) & (L - 1), y, z, rawU);
    if (CROSS && vb) load_plane(fv, g, (x0 - 1) & (L - 1), y, z, rawV);
    filter_plane(rawU, Q, P2);
    if (CROSS) filter_plane(rawV, Q, W2);
    if (ua) load_plane(fu, g, x0, y, z, rawU);
    if (CROSS && vb) load_plane(fv, g, x0, y, z, rawV);
    float dnext = d1[((int64_t)x0 * L + y) * L + z];  // d at plane x, loaded one step ahead

    double acc64[4] = {0.0, 0.0, 0.0, 0.0};
    // each step: filter the prefetched plane x, prefetch plane x+1, then the MFMA work of plane x
#define GRAM_STEP(CUR, PREV1, PREV2, WCUR, WPREV1, WPREV2)                                           \
    if (x < xend) {                                                                                  \
        filter_plane(rawU, Q, CUR);                                                                  \
        if (CROSS) filter_plane(rawV, Q, WCUR);                                                      \
        const float dval = dnext;                                                                    \
        if (x + 1 < xend) {                                                                          \
            if (ua) load_plane(fu, g, x + 1, y, z, rawU);                                            \
            if (CROSS && vb) load_plane(fv, g, x + 1, y, z, rawV);                                   \
            dnext = d1[((int64_t)(x + 1) * L + y) * L + z];                                          \
        }                                                                                            \
        float4_t accr = {0.f, 0.f, 0.f, 0.f}, acci = {0.f, 0.f, 0.f, 0.f};                           \
        gram_plane<CROSS>(CUR, PREV1, PREV2, WCUR, WPREV1, WPREV2, Q, lam, dval, g.vsign, accr, acci); \
        _Pragma("unroll") for (int j = 0; j < 4; ++j) acc64[j] += (double)accr[j] + (double)acci[j];  \
        ++x;                                                                                         \
    }
    for (int x = x0; x < xend;) {
        GRAM_STEP(P0, P2, P1, W0, W2, W1)
        GRAM_STEP(P1, P0, P2, W1, W0, W2)
        GRAM_STEP(P2, P1, P0, W2, W1, W0)
    }
#undef GRAM_STEP
    // lane holds G[4*(lane/16) + j][lane % 16]
#pragma unroll
    for (int j = 0; j < 4; ++j) red[wave][(4 * (lane / 16) + j) * 16 + (lane % 16)] = acc64[j];
    __syncthreads();
    for (int i = threadIdx.x; i < 256; i += blockDim.x) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < GRAM_WAVES; ++w) t += red[w][i];
        partial[(int64_t)blockIdx.x * 256 + i] = t;
    }
}

// Auto Gram with the z-direction Gamma stencil also moved onto d1 (the y stencil stays on the fields as
// the eigen-filter T_b(y) = sum_ly Qs[ly][b] F(y - ly), column b scaled by sqrt|lam_b|, sign s_b).  For the
// cells (p, r) of a column plane and delta = (dx, dz),
//   sum_{x, z} d(x, z) sum_{l1, l2} G[l1x][l2x] G[l1z][l2z] T^a(x - l1) T^b*(x - l2) = A[a][b] + A[b][a],
//   A = sum_{p, r} T^a(p, r) B^b(p, r),  B(p, r) = sum_{delta in H} w_delta C_delta(p, r) T*(p + dx, r + dz),
//   C_delta(p, r) = sum_{l1x} G[l1x][l1x - dx] E_dz(p + l1x, r),  E_dz(x', r) = sum_{l1z} G[l1z][l1z - dz] d(x', r + l1z),
// each sum over the taps that keep l1 - delta in 0 .. 2; H = {(0, 0) with weight 1/2, (0, 1), (0, 2),
// (1, -2 .. 2), (2, -2 .. 2)} (the pair delta, -delta gives the transposed term).  Per channel b one right
// operand B of 13 terms and one MFMA pair: 6 MFMAs per plane instead of 18, and no z filter.
//
// Each y-filtered plane q is read from LDS once (at z - 2 .. z + 2 of the lane's cell, step q - 2) and
// scattered into the right operands of the three planes it serves: B(q - 2) (dx = 2), B(q - 1) (dx = 1) and
// B(q) (dx = 0); B(q - 2) is then complete and meets its left operand T(q - 2, z), kept since it was read.
// C and E depend on the cell only; E of a plane is computed once and kept for three steps.
//
// LDS: y-filtered planes in 2 buffers (step p reads plane p+2 while plane p+3 is written), raw halo'd
// tiles in 2; the z halo is 2 on both sides.  Step p, after its one barrier: plane p+2 into the operands
// (and the MFMAs of plane p), y taps of plane p+3 from raw buffer (p+3)&1, store of staged plane p+4 into
// raw buffer (p+4)&1, load of plane p+5 into registers.  A workgroup of 8 waves owns 2 (y) x 16 (z) cells
// and 50 KB of LDS.
// Tile geometry with TY rows of 16 cells: 4 TY waves (4 along z with 4 cells each, times TY rows).  The raw
// tile keeps 16 shells per (row, z) position, the shell slot XOR-swizzled by z & 15 so that both the
// z-fastest staging stores and the shell-fastest y-tap reads spread over the banks.
template <int TY>
struct D2Geom {
    static constexpr int WAVES = 4 * TY, NROW = TY, TZ = 16;
    static constexpr int RY = TY + 2, RZ = TZ + 4;   // raw rows y0-2 .. y0+TY-1, z z0-2 .. z0+17
    static constexpr int RP = 16;
    static constexpr int TILE = RY * RZ * RP;        // float2 per raw buffer
    static constexpr int TB = TY * RZ * 16;          // float2 per channel of a T buffer
    static constexpr int TBUF = 3 * TB;              // float2 per T buffer: [b][row][z][shell]
    static constexpr int STAGE = RY * RZ * 16;       // elements staged per plane
    static constexpr int SPT = (STAGE + 64 * WAVES - 1) / (64 * WAVES);
    static constexpr int TPT = (TB + 64 * WAVES - 1) / (64 * WAVES);
    static_assert((2 * TILE + 2 * TBUF) * 8 >= WAVES * 256 * 8, "reduction reuses the tile buffers");
    // tiles, and the split variant's per-wave imaginary sums, in at most 80 KB: two workgroups per CU
    static_assert((2 * TILE + 2 * TBUF) * 8 + WAVES * 256 * 8 <= 80 * 1024, "LDS of two workgroups per CU");
};
// 2 rows per workgroup: 8 waves and at most 66 KB of LDS, so two workgroups and four waves per SIMD
using D2 = D2Geom<2>;

template <int NNEG, bool SPLIT = false>
__global__ __launch_bounds__(64 * D2::WAVES, 4) void k_gram_auto(const float* __restrict__ d1,
                                                           const float2* __restrict__ U, GramArgs g,
                                                           double* __restrict__ partial) {
    using GM = D2;
    __shared__ __align__(16) float2 smem[2 * GM::TILE + 2 * GM::TBUF];
    __shared__ double redi[SPLIT ? GM::WAVES * 256 : 1];  // imaginary sums per wave (SPLIT: registers are full)
    auto raw = [&](int q) -> float2* { return smem + (q & 1) * GM::TILE; };
    auto tb = [&](int q) -> float2* { return smem + 2 * GM::TILE + (q & 1) * GM::TBUF; };
    const int L = g.L, m = L - 1;
    const int tid = threadIdx.x;
    const int lane = tid % 64, wave = tid / 64;
    const int s = lane % 16, k = lane / 16;
    const int zw = wave % 4, yw = wave / 4;
    // Workgroups are dealt to the 8 XCDs round-robin; the remap gives each XCD a contiguous range of
    // logical blocks (x segment slowest, then columns 8 tiles wide walked along y), so co-resident
    // workgroups share halo rows in the XCD's L2 (columns 1, 2, 4, 16, 32 tiles wide measured 14%, 8%, 1%,
    // 1% and 4% slower).
    const int nzb = L / GM::TZ, nyb = L / GM::NROW;
    const int nb = nzb * nyb * (L / g.xseg);
    const int q = (nb % 8 == 0) ? (int)(blockIdx.x % 8) * (nb / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
    const int xs = q / (nzb * nyb);
    const int rr = q % (nzb * nyb);
    const int PZ = nzb < 8 ? nzb : 8;
    const int zb = (rr / (PZ * nyb)) * PZ + (rr % (PZ * nyb)) % PZ;
    const int yb = (rr % (PZ * nyb)) / PZ;
    const int z0 = zb * GM::TZ, y0 = yb * GM::NROW;
    const int x0 = xs * g.xseg, xend = x0 + g.xseg;
    const int64_t N = (int64_t)L * L * L;
    const int64_t LL = (int64_t)L * L;

    float Q[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Q[i] = g.Qs[i];
    const float G00 = g.gx[0], G11 = g.gx[4], G22 = g.gx[8], G10 = g.gx[3], G21 = g.gx[7], G20 = g.gx[6];
    const float H00 = 0.5f * G00, H11 = 0.5f * G11, H22 = 0.5f * G22;  // the weight 1/2 of delta = (0, 0)
    // this lane's cell (y0 + yw, z0 + cz); d1 offsets of z, z + 1, z + 2 within a plane
    const int cz = 4 * zw + k;
    const int yrow = (y0 + yw) * L, zc = z0 + cz;

    constexpr int SPT = GM::SPT;
    const float2* src[SPT];
    int dst[SPT];
    bool act[SPT];
#pragma unroll
    for (int i = 0; i < SPT; ++i) {
        const int e = tid + i * 64 * GM::WAVES;
        act[i] = e < GM::STAGE;
        const int ee = act[i] ? e : 0;
        const int sh = ee / (GM::RY * GM::RZ), ry = (ee / GM::RZ) % GM::RY, rz = ee % GM::RZ;
        src[i] = U + (int64_t)(sh < g.S1 ? sh : 0) * N + (int64_t)((y0 - 2 + ry) & m) * g.ys + ((z0 - 2 + rz) & m);
        dst[i] = (ry * GM::RZ + rz) * GM::RP + (sh ^ (rz & 15));
    }
    float2 stg[SPT];
    auto stage_load = [&](int plane) {
#pragma unroll
        for (int i = 0; i < SPT; ++i)
            if (act[i]) stg[i] = src[i][(int64_t)(plane & m) * g.xs];
    };
    auto stage_store = [&](int plane) {
#pragma unroll
        for (int i = 0; i < SPT; ++i)
            if (act[i]) raw(plane)[dst[i]] = stg[i];
    };
    // y taps: position u = (row * GM::RZ + z') * 16 + sh reads raw rows row + 2 - ly (tile row r holds
    // y0 - 2 + r) and writes T[b][row][z'][sh]
    int traw[GM::TPT];
    bool tact[GM::TPT];
#pragma unroll
    for (int i = 0; i < GM::TPT; ++i) {
        const int u = tid + i * 64 * GM::WAVES;
        tact[i] = u < GM::TB;
        const int uu = tact[i] ? u : 0;
        const int sh = uu % 16, zp = (uu / 16) % GM::RZ, row = uu / (16 * GM::RZ);
        traw[i] = ((row + 2) * GM::RZ + zp) * GM::RP + (sh ^ (zp & 15));
    }
    auto tphase = [&](int plane) {
        const float2* rw = raw(plane);
        float2* t = tb(plane);
#pragma unroll
        for (int i = 0; i < GM::TPT; ++i) {
            if (!tact[i]) continue;
            const float2 r0 = rw[traw[i]];
            const float2 r1 = rw[traw[i] - GM::RZ * GM::RP];
            const float2 r2 = rw[traw[i] - 2 * GM::RZ * GM::RP];
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const float q0 = Q[0 * 3 + b], q1 = Q[1 * 3 + b], q2 = Q[2 * 3 + b];
                t[b * GM::TB + tid + i * 64 * GM::WAVES] =
                    make_float2(q0 * r0.x + q1 * r1.x + q2 * r2.x, q0 * r0.y + q1 * r1.y + q2 * r2.y);
            }
        }
    };
    // E_dz of a plane from its d1 values at z, z + 1, z + 2 (index dz + 2); G symmetric
    auto efun = [&](const float dd[3], float E[5]) {
        E[0] = G20 * dd[0];
        E[1] = G10 * dd[0] + G21 * dd[1];
        E[2] = G00 * dd[0] + G11 * dd[1] + G22 * dd[2];
        E[3] = G10 * dd[1] + G21 * dd[2];
        E[4] = G20 * dd[2];
    };
    auto dload = [&](int plane, float dd[3]) {
        const float* dp = d1 + (int64_t)(plane & m) * LL;
        dd[0] = dp[yrow + (zc & m)];
        dd[1] = dp[yrow + ((zc + 1) & m)];
        dd[2] = dp[yrow + ((zc + 2) & m)];
    };
    const int tme = (yw * GM::RZ + cz + 2) * 16 + s;  // this lane's T index at dz = 0 (channel 0)

    double acc_slot[4] = {0.0, 0.0, 0.0, 0.0};  // this lane's G entries 4 (lane / 16) + j, column lane % 16
    double* islot = redi + (SPLIT ? wave * 256 + (4 * (lane / 16)) * 16 + (lane % 16) : 0);
    if (SPLIT) {
#pragma unroll
        for (int j = 0; j < 4; ++j) islot[j * 16] = 0.0;
    }

    // warm-up: T of plane x0, raw plane x0+1, staged plane x0+2; E of planes x0, x0+1; d of plane x0+2.  The
    // loop starts at p = x0 - 2, whose first two steps only fill the operands of planes x0 and x0+1.
    float E0[5], E1[5], E2[5], dn[3];
    float2 B0[3], B1[3], B2[3], A0[3], A1[3], A2[3];
    stage_load(x0);
    stage_store(x0);
    stage_load(x0 + 1);
    dload(x0, dn);
    efun(dn, E0);
    dload(x0 + 1, dn);
    efun(dn, E1);
    dload(x0 + 2, dn);
    __syncthreads();
    tphase(x0);
    stage_store(x0 + 1);
    stage_load(x0 + 2);

    float4_t accr = {0.f, 0.f, 0.f, 0.f}, acci = {0.f, 0.f, 0.f, 0.f};
    // step p: slots C* hold plane p (complete now), N* plane p+1, F* plane p+2 (started now); EA = E(p+2),
    // EB = E(p+3), EF = E(p+4) (computed now)
#define D2_STEP(FIN, MID, BC, BN, BF, AC, AN, AF, EA, EB, EF)                                         \
    if (p < xend) {                                                                                 \
        __syncthreads();                                                                            \
        efun(dn, EF);                                                                               \
        dload(p + 5, dn);                                                                           \
        float c2[5], c1[5];                                                                         \
        _Pragma("unroll") for (int j = 0; j < 5; ++j) {                                             \
            c2[j] = G20 * EA[j];                                                                    \
            c1[j] = G10 * EA[j] + G21 * EB[j];                                                      \
        }                                                                                           \
        const float c00 = H00 * EA[2] + H11 * EB[2] + H22 * EF[2];                                  \
        const float c01 = G00 * EA[3] + G11 * EB[3] + G22 * EF[3];                                  \
        const float c02 = G00 * EA[4] + G11 * EB[4] + G22 * EF[4];                                  \
        if (lead) {                                                                                 \
            tphase(p + 3);                                                                          \
            stage_store(p + 4);                                                                     \
            stage_load(p + 5);                                                                      \
        }                                                                                           \
        const float2* t = tb(p + 2) + tme;                                                          \
        _Pragma("unroll") for (int b = 0; b < 3; ++b) {                                             \
            const float sg = b < NNEG ? -1.0f : 1.0f;                                               \
            float2 u[5];                                                                            \
            _Pragma("unroll") for (int j = 0; j < 5; ++j) u[j] = t[b * GM::TB + (j - 2) * 16];       \
            /* plane p + 2 as dx = 2 of plane p: its right operand is complete */                   \
            if (FIN) {                                                                              \
                float bx = BC[b].x, by = BC[b].y;                                                   \
                _Pragma("unroll") for (int j = 0; j < 5; ++j) {                                     \
                    bx = (sg * c2[j]) * u[j].x + bx;                                                \
                    by = (sg * c2[j]) * u[j].y + by;                                                \
                }                                                                                   \
                accr = __builtin_amdgcn_mfma_f32_16x16x4f32(AC[b].x, bx, accr, 0, 0, 0);            \
                if (SPLIT)                                                                          \
                    acci = __builtin_amdgcn_mfma_f32_16x16x4f32(AC[b].y, by, acci, 0, 0, 0);        \
                else                                                                                \
                    accr = __builtin_amdgcn_mfma_f32_16x16x4f32(AC[b].y, by, accr, 0, 0, 0);        \
            }                                                                                       \
            /* dx = 1 of plane p + 1 */                                                             \
            if (MID) {                                                                              \
                _Pragma("unroll") for (int j = 0; j < 5; ++j) {                                     \
                    BN[b].x = (sg * c1[j]) * u[j].x + BN[b].x;                                      \
                    BN[b].y = (sg * c1[j]) * u[j].y + BN[b].y;                                      \
                }                                                                                   \
            }                                                                                       \
            /* dx = 0 of plane p + 2 (its first terms) and its left operand */                      \
            BF[b].x = (sg * c00) * u[2].x;                                                          \
            BF[b].y = (sg * c00) * u[2].y;                                                          \
            BF[b].x = (sg * c01) * u[3].x + BF[b].x;                                                \
            BF[b].y = (sg * c01) * u[3].y + BF[b].y;                                                \
            BF[b].x = (sg * c02) * u[4].x + BF[b].x;                                                \
            BF[b].y = (sg * c02) * u[4].y + BF[b].y;                                                \
            AF[b] = u[2];                                                                           \
        }                                                                                           \
        if (SPLIT && FIN) { /* the imaginary sums go to LDS every step (no registers to keep them) */ \
            _Pragma("unroll") for (int j = 0; j < 4; ++j) {                                         \
                islot[j * 16] += (double)acci[j];                                                   \
                acci[j] = 0.f;                                                                      \
            }                                                                                       \
        }                                                                                           \
        if (!lead) {                                                                                \
            tphase(p + 3);                                                                          \
            stage_store(p + 4);                                                                     \
            stage_load(p + 5);                                                                      \
        }                                                                                           \
        ++p;                                                                                        \
    }
    // Half of the waves (two per SIMD whichever way waves are dealt to SIMDs) do the y taps and staging of
    // a step before its operand work, the others after, so that every SIMD has LDS and arithmetic work from
    // different phases within a barrier interval (the y taps write a buffer that no wave reads this step).
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const bool lead = ((wv >> 2) ^ wv) & 1;
    int p = x0 - 2;
    // steps x0 - 2 and x0 - 1 only start the operands of planes x0 and x0 + 1
    D2_STEP(false, false, B0, B1, B2, A0, A1, A2, E0, E1, E2)
    D2_STEP(false, true, B1, B2, B0, A1, A2, A0, E1, E2, E0)
    while (p < xend) {
        D2_STEP(true, true, B2, B0, B1, A2, A0, A1, E2, E0, E1)
        D2_STEP(true, true, B0, B1, B2, A0, A1, A2, E0, E1, E2)
        D2_STEP(true, true, B1, B2, B0, A1, A2, A0, E1, E2, E0)
        // FP32 sums of at most six planes (144 products per entry) move into the FP64 slots
        if ((p - x0) % 6 == 0 || p >= xend) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc_slot[j] += (double)accr[j];
                accr[j] = 0.f;
            }
        }
    }
#undef D2_STEP
    // per-wave G blocks through LDS (the tile buffers are free after this barrier), summed in a fixed order
    constexpr int ROW = SPLIT ? 512 : 256;
    double* red = (double*)smem;  // [GM::WAVES][256]
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) red[wave * 256 + (4 * (lane / 16) + j) * 16 + (lane % 16)] = acc_slot[j];
    __syncthreads();
    for (int i = tid; i < 256; i += 64 * GM::WAVES) {
        double tsum = 0.0;
#pragma unroll
        for (int w = 0; w < GM::WAVES; ++w) tsum += red[w * 256 + i];
        partial[(int64_t)blockIdx.x * ROW + i] = tsum;
        if (SPLIT) {
            double usum = 0.0;
#pragma unroll
            for (int w = 0; w < GM::WAVES; ++w) usum += redi[w * 256 + i];
            partial[(int64_t)blockIdx.x * ROW + 256 + i] = usum;
        }
    }
}

// Lane-layout self check of the 16x16x4 FP32 MFMA: C = A B for A[m][k] = m + 16 k,
// B[k][n] = (k + 1) * (n - 7).
__global__ void k_mfma_selftest(float* out) {
    const int lane = threadIdx.x;
    const int m = lane % 16, kk = lane / 16;
    float a = (float)(m + 16 * kk), b = (float)((kk + 1) * (lane % 16 - 7));
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    for (int j = 0; j < 4; ++j) out[(4 * (lane / 16) + j) * 16 + (lane % 16)] = acc[j];
}

}  // namespace

extern "C" {

int hermes_multipole_summand_f64(hermes_ctx* ctx, int J, const double* d1, const void* F1, const void* F2,
                                 const double* gamma3x3_host, double* out_host) {
    HERMES_REQUIRE(ctx && d1 && F1 && F2 && gamma3x3_host && out_host, "hermes_multipole_summand_f64: null argument");
    HERMES_REQUIRE(J >= 2 && J <= 10, "J=%d unsupported", J);
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int L = 1 << J;
    const int64_t N = (int64_t)L * L * L;
    const int nb = (int)hermes_blocks(N, SUM_THREADS, 4096);
    double* ws = (double*)hermes_workspace(ctx, WS_CONTRACT,
                                           sizeof(double) * (2 * (size_t)nb + 2 + 9 + 2 * RED_MAX_BLOCKS));
    HERMES_REQUIRE(ws, "%s", hermes_last_error());
    double* dG = ws + 2 * nb + 2;
    HERMES_CHECK_HIP(hipMemcpyAsync(dG, gamma3x3_host, sizeof(double) * 9, hipMemcpyHostToDevice, ctx->stream));
    k_summand_f64<<<nb, SUM_THREADS, 0, ctx->stream>>>(L, d1, (const double2*)F1, (const double2*)F2, dG, ws);
    HERMES_LAUNCH_CHECK();
    sum_rows(ctx->stream, ws, nb, 2, dG + 9, ws + 2 * nb);
    HERMES_LAUNCH_CHECK();
    HERMES_CHECK_HIP(hipMemcpyAsync(out_host, ws + 2 * nb, sizeof(double) * 2, hipMemcpyDeviceToHost, ctx->stream));
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

// Waits for the recorded (start, end) event pairs of timed Gram launches and adds their
// durations to the totals (launches may sit on any stream).
static int gram_flush(hermes_ctx* ctx) {
    for (size_t i = 0; i < ctx->gram_pending; ++i) {
        float ms = 0.0f;
        HERMES_CHECK_HIP(hipEventSynchronize(ctx->gram_ev[2 * i + 1]));
        HERMES_CHECK_HIP(hipEventElapsedTime(&ms, ctx->gram_ev[2 * i], ctx->gram_ev[2 * i + 1]));
        ctx->gram_ms += ms;
        ctx->gram_calls += 1;
    }
    ctx->gram_pending = 0;
    return 0;
}

// Launches the Gram kernel and the fixed-order row reduction; the 16 x 16 raw sums land in
// out_dev (256 doubles; row i, column j at i * 16 + j).  No host synchronisation.
static int gram_launch(hermes_ctx* ctx, int J, const float* d1, const void* U, int n_u, const void* V, int n_v,
                       int v_conj, const double* Q3x3_host, const double* lam3_host, double* out_dev, GramArgs* gout,
                       bool split = false) {
    HERMES_REQUIRE(ctx && d1 && U && Q3x3_host && lam3_host && out_dev, "hermes_multipole_gram_f32: null argument");
    HERMES_REQUIRE(J >= 3 && J <= 10, "hermes_multipole_gram_f32: J=%d unsupported (needs L >= 8)", J);
    HERMES_REQUIRE(n_u >= 1 && n_u <= 16, "n_u=%d outside 1..16", n_u);
    HERMES_REQUIRE(V == nullptr || (n_v >= 1 && n_v <= 16), "n_v=%d outside 1..16", n_v);
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int L = 1 << J;
    GramArgs g;
    g.L = L;
    g.xs = ctx->field_layout ? L : L * L;  // x-major fields: (x L + y) L + z; y-major: (y L + x) L + z
    g.ys = ctx->field_layout ? L * L : L;
    g.S1 = n_u;
    g.cross = V != nullptr;
    g.S2 = g.cross ? n_v : n_u;
    g.vsign = v_conj ? 1.0f : -1.0f;  // Re(U V) needs -V.im, Re(U conj V) needs +V.im
    // planes per workgroup: the auto kernel's warm-up (two planes filtered ahead) and its final reduction
    // are paid once per segment, so it takes 512 planes (8192 workgroups at L = 512, 16 per workgroup slot;
    // 256- and 128-plane segments measured 2% and 3% slower)
    const int seg = g.cross ? 64 : 512;
    g.xseg = L < seg ? L : seg;
    for (int i = 0; i < 9; ++i) g.Q[i] = (float)Q3x3_host[i];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) g.lam[(a * 3 + b) * 3 + c] = (float)(lam3_host[a] * lam3_host[b] * lam3_host[c]);
    for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) g.lamyz[b * 3 + c] = (float)(lam3_host[b] * lam3_host[c]);
    for (int l1 = 0; l1 < 3; ++l1)
        for (int l2 = 0; l2 < 3; ++l2) {
            double t = 0.0;
            for (int a = 0; a < 3; ++a) t += Q3x3_host[l1 * 3 + a] * lam3_host[a] * Q3x3_host[l2 * 3 + a];
            g.gx[l1 * 3 + l2] = (float)t;
        }
    int order[3] = {0, 1, 2}, nneg = 0;  // eigenvalues ascending (insertion sort of three)
    for (int i = 1; i < 3; ++i)
        for (int j = i; j > 0 && lam3_host[order[j]] < lam3_host[order[j - 1]]; --j) {
            const int t = order[j];
            order[j] = order[j - 1];
            order[j - 1] = t;
        }
    for (int c = 0; c < 3; ++c) {
        const double lc = lam3_host[order[c]];
        if (lc < 0.0) ++nneg;
        for (int tap = 0; tap < 3; ++tap) g.Qs[tap * 3 + c] = (float)(Q3x3_host[tap * 3 + order[c]] * sqrt(fabs(lc)));
    }
    // auto kernel: 16 z x 4 y cells per workgroup (needs L >= 16); cross kernel: 4 z x 4 y
    HERMES_REQUIRE(g.cross || J >= 4, "hermes_multipole_gram_f32: J=%d, the auto Gram needs J >= 4", J);
    const int nblocks = g.cross ? (L / 4) * (L / GRAM_WAVES) * (L / g.xseg)
                                : (L / D2::TZ) * (L / D2::NROW) * (L / g.xseg);
    HERMES_REQUIRE(!split || !g.cross, "hermes_multipole_gram_f32: split sums are for the auto Gram");
    const int row = split ? 512 : 256;
    double* ws = (double*)hermes_workspace(
        ctx, WS_CONTRACT, sizeof(double) * ((size_t)nblocks * row + row + (size_t)RED_MAX_BLOCKS * row));
    HERMES_REQUIRE(ws, "%s", hermes_last_error());
    const bool timed = ctx->time_gram && !ctx->capturing;
    size_t slot = 0;
    if (timed) {
        if (ctx->gram_pending >= 4096 && gram_flush(ctx)) return 1;
        slot = 2 * ctx->gram_pending;
        while (ctx->gram_ev.size() < slot + 2) {
            hipEvent_t e = nullptr;
            HERMES_CHECK_HIP(hipEventCreate(&e));
            ctx->gram_ev.push_back(e);
        }
        HERMES_CHECK_HIP(hipEventRecord(ctx->gram_ev[slot], ctx->stream));
    }
    if (g.cross)
        k_gram_f32<true><<<nblocks, 64 * GRAM_WAVES, 0, ctx->stream>>>(d1, (const float2*)U, (const float2*)V, g, ws);
    else
        switch (nneg + (split ? 4 : 0)) {
            case 0: k_gram_auto<0><<<nblocks, 64 * D2::WAVES, 0, ctx->stream>>>(d1, (const float2*)U, g, ws); break;
            case 1: k_gram_auto<1><<<nblocks, 64 * D2::WAVES, 0, ctx->stream>>>(d1, (const float2*)U, g, ws); break;
            case 2: k_gram_auto<2><<<nblocks, 64 * D2::WAVES, 0, ctx->stream>>>(d1, (const float2*)U, g, ws); break;
            case 3: k_gram_auto<3><<<nblocks, 64 * D2::WAVES, 0, ctx->stream>>>(d1, (const float2*)U, g, ws); break;
            case 4: k_gram_auto<0, true><<<nblocks, 64 * D2::WAVES, 0, ctx->stream>>>(d1, (const float2*)U, g, ws); break;
            case 5: k_gram_auto<1, true><<<nblocks, 64 * D2::WAVES, 0, ctx->stream>>>(d1, (const float2*)U, g, ws); break;
            case 6: k_gram_auto<2, true><<<nblocks, 64 * D2::WAVES, 0, ctx->stream>>>(d1, (const float2*)U, g, ws); break;
            default: k_gram_auto<3, true><<<nblocks, 64 * D2::WAVES, 0, ctx->stream>>>(d1, (const float2*)U, g, ws); break;
        }
    HERMES_LAUNCH_CHECK();
    if (timed) {
        HERMES_CHECK_HIP(hipEventRecord(ctx->gram_ev[slot + 1], ctx->stream));
        ctx->gram_pending += 1;
    }
    double* tmp = ws + (size_t)nblocks * row + row;
    sum_rows(ctx->stream, ws, nblocks, 256, tmp, out_dev, row);
    if (split) sum_rows(ctx->stream, ws + 256, nblocks, 256, tmp, out_dev + 256, row);
    HERMES_LAUNCH_CHECK();
    if (gout) *gout = g;
    return 0;
}

int hermes_multipole_gram_f32(hermes_ctx* ctx, int J, const float* d1, const void* U, int n_u, const void* V, int n_v,
                              int v_conj, const double* Q3x3_host, const double* lam3_host, double* out_host) {
    HERMES_REQUIRE(ctx && out_host, "hermes_multipole_gram_f32: null argument");
    HERMES_REQUIRE(!ctx->capturing, "hermes_multipole_gram_f32 returns host data; capture hermes_multipole_gram_f32_dev");
    double* dst = (double*)hermes_workspace(ctx, WS_MISC, sizeof(double) * 256);
    HERMES_REQUIRE(dst, "%s", hermes_last_error());
    GramArgs g;
    if (int rc = gram_launch(ctx, J, d1, U, n_u, V, n_v, v_conj, Q3x3_host, lam3_host, dst, &g)) return rc;
    double h[256];
    HERMES_CHECK_HIP(hipMemcpyAsync(h, dst, sizeof(double) * 256, hipMemcpyDeviceToHost, ctx->stream));
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < g.S1; ++i)
        for (int j = 0; j < g.S2; ++j)
            out_host[i * g.S2 + j] = g.cross ? h[i * 16 + j] : h[i * 16 + j] + h[j * 16 + i];
    return 0;
}

int hermes_multipole_gram_f32_dev(hermes_ctx* ctx, int J, const float* d1, const void* U, int n_u, const void* V,
                                  int n_v, int v_conj, const double* Q3x3_host, const double* lam3_host,
                                  double* out_dev) {
    return gram_launch(ctx, J, d1, U, n_u, V, n_v, v_conj, Q3x3_host, lam3_host, out_dev, nullptr);
}

int hermes_multipole_gram_f32_dev_ri(hermes_ctx* ctx, int J, const float* d1, const void* U, int n_u,
                                     const double* Q3x3_host, const double* lam3_host, double* out_dev) {
    return gram_launch(ctx, J, d1, U, n_u, nullptr, 0, 0, Q3x3_host, lam3_host, out_dev, nullptr, true);
}

int hermes_profile_gram(hermes_ctx* ctx, int enable, double* total_ms_out, int64_t* calls_out) {
    HERMES_REQUIRE(ctx, "hermes_profile_gram: null context");
    HERMES_REQUIRE(!ctx->capturing, "hermes_profile_gram: the stream is being captured");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    if (gram_flush(ctx)) return 1;
    if (total_ms_out) *total_ms_out = ctx->gram_ms;
    if (calls_out) *calls_out = ctx->gram_calls;
    if (enable) {
        ctx->gram_ms = 0.0;
        ctx->gram_calls = 0;
    }
    ctx->time_gram = enable != 0;
    return 0;
}

int hermes_mfma_selftest(hermes_ctx* ctx, double* max_abs_err) {
    HERMES_REQUIRE(ctx && max_abs_err, "null argument");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    float* d = (float*)hermes_workspace(ctx, WS_MISC, sizeof(float) * 256);
    HERMES_REQUIRE(d, "%s", hermes_last_error());
    k_mfma_selftest<<<1, 64, 0, ctx->stream>>>(d);
    HERMES_LAUNCH_CHECK();
    float h[256];
    HERMES_CHECK_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    double err = 0.0;
    for (int m = 0; m < 16; ++m)
        for (int n = 0; n < 16; ++n) {
            double ref = 0.0;
            for (int k = 0; k < 4; ++k) ref += (double)(m + 16 * k) * (double)((k + 1) * (n - 7));
            double e = fabs(ref - h[m * 16 + n]);
            if (e > err) err = e;
        }
    *max_abs_err = err;
    return 0;
}

}  // extern "C"
