// Window stage: kernel builders, spectral multiply, fused pair reduction and the
// fused multipole window apply of the batched SH-3PCF path.
// P (|phi_hat|^2 samples, L+1 doubles) is a DEVICE pointer; window descriptors
// and shell radii are HOST arrays copied to the context workspace.
#include "../../include/hermes_amd.h"
#include "fft_lds.h"
#include "hermes_common.h"
#include "special.h"

#include <cmath>
#include <mutex>

namespace {

constexpr int RED_THREADS = 256;

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------- kernel builders
// convolution.py:14-38 + :263-281: K[x,y,z] = sum over the 8 mirrors {x, L-x} x ... of
// P[i]P[j]P[k] W(i/L, j/L, k/L), in the fold's summation order.
template <typename T>
__global__ void k_build_rfft(int L, const double* __restrict__ P, hermes_window_desc d, T* __restrict__ out) {
    const int Lh = L / 2 + 1;
    const int64_t n = (int64_t)L * L * Lh;
    const double inv_L = 1.0 / L;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int z = (int)(i % Lh);
        int y = (int)((i / Lh) % L);
        int x = (int)(i / ((int64_t)Lh * L));
        int xs[2] = {x, L - x}, ys[2] = {y, L - y}, zs[2] = {z, L - z};
        // fold order: (0,0,0) (1,0,0) (0,1,0) (0,0,1) (1,1,0) (1,0,1) (0,1,1) (1,1,1)
        const int ord[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
        double t = 0.0;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            int ii = xs[ord[a][0]], jj = ys[ord[a][1]], kk = zs[ord[a][2]];
            double wa = 0.0;
            // the mirror alias L - x sits at k = -(L - x) / L (the sign matters for a non-axis line of sight)
            const double sx = ord[a][0] ? -1.0 : 1.0, sy = ord[a][1] ? -1.0 : 1.0, sz = ord[a][2] ? -1.0 : 1.0;
            wa += P[ii] * P[jj] * P[kk] *
                  iso_window_f64(d.type, sx * ii * inv_L, sy * jj * inv_L, sz * kk * inv_L, d.a0, d.a1, d.axis, d.los);
            t = (a == 0) ? wa : t + wa;
        }
        if (d.type == HERMES_WIN_INVERSE_LAPLACIAN && i == 0) t = 0.0;
        out[i] = (T)t;
    }
}

// radial_multipole_windows.py:147-183 order: aliases (x-L | x) per axis, "0" first.
template <typename T>
__global__ void k_build_full(int L, const double* __restrict__ P, hermes_window_desc d, T* __restrict__ out) {
    const int64_t n = (int64_t)L * L * L;
    const double inv_L = 1.0 / L;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int z = (int)(i % L);
        int y = (int)((i / L) % L);
        int x = (int)(i / ((int64_t)L * L));
        double pi[2] = {P[L - x], P[x]}, pj[2] = {P[L - y], P[y]}, pk[2] = {P[L - z], P[z]};
        double ki[2] = {(x - L) * inv_L, x * inv_L}, kj[2] = {(y - L) * inv_L, y * inv_L},
               kk[2] = {(z - L) * inv_L, z * inv_L};
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            int u = (a >> 2) & 1, v = (a >> 1) & 1, w = a & 1;
            double p = pi[u] * pj[v] * pk[w];
            double2 wv = window_value_f64(d, ki[u], kj[v], kk[w]);
            if (a == 0) {
                re = p * wv.x;
                im = p * wv.y;
            } else {
                re += p * wv.x;
                im += p * wv.y;
            }
        }
        if (d.type == HERMES_WIN_INVERSE_LAPLACIAN && i == 0) re = im = 0.0;
        out[2 * i] = (T)re;
        out[2 * i + 1] = (T)im;
    }
}

// radial_multipole_windows.py:147-183 and :186-264: K[x,y,z] = sum over the 8 signed aliases of
// P P P U_l(|k|) Y_lm(k-hat), U_l from a radial profile (closed form or table).
template <typename T>
__global__ void k_build_full_profile(int L, const double* __restrict__ P, hermes_radial_profile prof, int l, int m,
                                     T* __restrict__ out) {
    const int64_t n = (int64_t)L * L * L;
    const double inv_L = 1.0 / L;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(i % L), y = (int)((i / L) % L), x = (int)(i / ((int64_t)L * L));
        const double pi[2] = {P[L - x], P[x]}, pj[2] = {P[L - y], P[y]}, pk[2] = {P[L - z], P[z]};
        const double ki[2] = {(x - L) * inv_L, x * inv_L}, kj[2] = {(y - L) * inv_L, y * inv_L},
                     kk[2] = {(z - L) * inv_L, z * inv_L};
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const int u = (a >> 2) & 1, v = (a >> 1) & 1, w = a & 1;
            const double p = pi[u] * pj[v] * pk[w];
            const double k = sqrt(ki[u] * ki[u] + kj[v] * kj[v] + kk[w] * kk[w]);
            const double rad = radial_profile_f64(prof, l, k);
            const double2 yv = ylm_f64(l, m, ki[u], kj[v], kk[w]);
            re += p * rad * yv.x;
            im += p * rad * yv.y;
        }
        out[2 * i] = (T)re;
        out[2 * i + 1] = (T)im;
    }
}

template <typename T>
__global__ void k_spec_mul(int64_t n, T* __restrict__ spec, const T* __restrict__ ker, int kc, T scale) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        T a = spec[2 * i], b = spec[2 * i + 1];
        if (kc) {
            T c = ker[2 * i] * scale, e = ker[2 * i + 1] * scale;
            spec[2 * i] = a * c - b * e;
            spec[2 * i + 1] = a * e + b * c;
        } else {
            T c = ker[i] * scale;
            spec[2 * i] = a * c;
            spec[2 * i + 1] = b * c;
        }
    }
}

// ---------------------------------------------------------------- pair reduction
// Window value at signed lattice alias (ax, ay, az), FP32 fast path for shells and
// shell multipoles; other window types use the FP64 formula.
__device__ __forceinline__ float window_value_f32(const hermes_window_desc& d, int ax, int ay, int az, int L,
                                                  float cnorm) {
    double n2 = (double)ax * ax + (double)ay * ay + (double)az * az;
    if (d.type == HERMES_WIN_SHELL) {
        double t = sqrt(n2) * d.a0 / L;
        return jl_f32(0, t);
    }
    if (d.type == HERMES_WIN_SHELL_MULTIPOLE) {
        double t = sqrt(n2) * d.a0 / L;
        float j = jl_f32(d.l, t);
        float2 y = ylm_f32(d.l, d.m, (float)ax, (float)ay, (float)az, cnorm);
        return j * y.x;  // m = 0 windows are real
    }
    double inv_L = 1.0 / L;
    // reached only by passes without non-axis line-of-sight windows
    return (float)iso_window_f64(d.type, ax * inv_L, ay * inv_L, az * inv_L, d.a0, d.a1, d.axis, nullptr);
}

// One thread per octant representative (x, y, z) in [0, L/2]^3; its kernel value is
// shared by the (up to) 4 modes (x|L-x, y|L-y, z) of the half spectrum.  GEN: the pass holds a
// line-of-sight window along a non-axis direction (descriptor axis 3), whose value differs between the 4
// modes; it is a separate instantiation so the other passes keep their register budget.
template <typename C, bool FP64, bool GEN>
__global__ __launch_bounds__(RED_THREADS) void k_pair_reduce(int L, const C* __restrict__ F1,
                                                             const C* __restrict__ F2, const double* __restrict__ P,
                                                             const hermes_window_desc* __restrict__ descs,
                                                             const float* __restrict__ cnorms, int nd,
                                                             double* __restrict__ partial) {
    extern __shared__ double wacc[];  // [waves][nd]
    const int Lh = L / 2 + 1;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int nwaves = blockDim.x / 64;
    for (int i = threadIdx.x; i < nwaves * nd; i += blockDim.x) wacc[i] = 0.0;
    __syncthreads();
    const int64_t nrep = (int64_t)Lh * Lh * Lh;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < nrep; base += stride) {
        int64_t r = base + threadIdx.x;
        bool active = r < nrep;
        double power = 0.0;
        double pw00 = 0.0, pw01 = 0.0, pw10 = 0.0, pw11 = 0.0;  // per mirror mode (x | L-x, y | L-y)
        int x = 0, y = 0, z = 0;
        if (active) {
            z = (int)(r % Lh);
            y = (int)((r / Lh) % Lh);
            x = (int)(r / ((int64_t)Lh * Lh));
            const int xm = (L - x) % L, ym = (L - y) % L;
            const int nxm = (xm != x) ? 2 : 1, nym = (ym != y) ? 2 : 1;
            const double wz = (z == 0 || z == L / 2) ? 1.0 : 2.0;
            for (int a = 0; a < nxm; ++a)
                for (int b = 0; b < nym; ++b) {
                    int xx = a ? xm : x, yy = b ? ym : y;
                    int64_t idx = ((int64_t)xx * L + yy) * Lh + z;
                    C f1 = F1[idx], f2 = F2[idx];
                    const double pab = wz * ((double)f1.x * (double)f2.x + (double)f1.y * (double)f2.y);
                    if (a == 0 && b == 0) pw00 = pab;
                    else if (a == 0) pw01 = pab;
                    else if (b == 0) pw10 = pab;
                    else pw11 = pab;
                    power += pab;
                }
        }
        for (int b = 0; b < nd; ++b) {
            double kb = 0.0;
            if (active) {
                const hermes_window_desc d = descs[b];
                // a window along a non-axis line of sight differs between the mirror modes: the aliases of
                // (L - x, y, z) are those of (x, y, z) with k_x negated, so each mode takes its own sum
                if (GEN && d.type >= HERMES_WIN_RING && d.axis == 3) {
                    const double inv_L = 1.0 / L;
                    for (int m = 0; m < 4; ++m) {
                        const double sx = (m & 2) ? -1.0 : 1.0, sy = (m & 1) ? -1.0 : 1.0;
                        const double pm = m == 0 ? pw00 : (m == 1 ? pw01 : (m == 2 ? pw10 : pw11));
                        if (pm == 0.0) continue;  // a mode that is its own mirror, or a zero product
                        double km = 0.0;
                        for (int a = 0; a < 8; ++a) {
                            int ax = (a & 4) ? x - L : x, ay = (a & 2) ? y - L : y, az = (a & 1) ? z - L : z;
                            int ix = ax < 0 ? -ax : ax, iy = ay < 0 ? -ay : ay, iz = az < 0 ? -az : az;
                            double p = P[ix] * P[iy] * P[iz];
                            const double w = window_value_f64(d, sx * ax * inv_L, sy * ay * inv_L, az * inv_L).x;
                            km += FP64 ? p * w : (double)((float)p * (float)w);
                        }
                        kb += pm * km;
                    }
                } else {
                    for (int a = 0; a < 8; ++a) {
                        int ax = (a & 4) ? x - L : x, ay = (a & 2) ? y - L : y, az = (a & 1) ? z - L : z;
                        int ix = ax < 0 ? -ax : ax, iy = ay < 0 ? -ay : ay, iz = az < 0 ? -az : az;
                        double p = P[ix] * P[iy] * P[iz];
                        if (FP64) {
                            double inv_L = 1.0 / L;
                            kb += p * window_value_f64(d, ax * inv_L, ay * inv_L, az * inv_L, GEN).x;
                        } else {
                            kb += (double)((float)p * window_value_f32(d, ax, ay, az, L, cnorms[b]));
                        }
                    }
                    kb *= power;
                }
                if (d.type == HERMES_WIN_INVERSE_LAPLACIAN && x == 0 && y == 0 && z == 0) kb = 0.0;
            }
            double v = wave_sum(kb);
            if (lane == 0) wacc[wave * nd + b] += v;
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nd; b += blockDim.x) {
        double s = 0.0;
        for (int w = 0; w < nwaves; ++w) s += wacc[w * nd + b];
        partial[(int64_t)blockIdx.x * nd + b] = s;
    }
}

// j_0 .. j_LMAX at q = 2 pi t (t in turns), FP32.  One hardware sin/cos and the
// upward recurrence when q >= LMAX + 1 (stable for every order there); otherwise
// each order on its own (jl_f32: recurrence where q >= l + 1, else the series).
template <int LMAX>
__device__ __forceinline__ void radial_jl(double t, float* j) {
    if (t == 0.0) {
        j[0] = 1.0f;
#pragma unroll
        for (int l = 1; l <= LMAX; ++l) j[l] = 0.0f;
        return;
    }
    const float q = (float)(2.0 * HERMES_PI * t);
    if (LMAX == 0 || q >= (float)LMAX + 1.0f) {
        const float fr = (float)(t - floor(t));
        const float sn = __builtin_amdgcn_sinf(fr), cs = __builtin_amdgcn_cosf(fr);
        const float iq = __builtin_amdgcn_rcpf(q);
        j[0] = sn * iq;
        if (LMAX >= 1) j[1] = (j[0] - cs) * iq;
#pragma unroll
        for (int n = 1; n < LMAX; ++n) j[n + 1] = (float)(2 * n + 1) * iq * j[n] - j[n - 1];
    } else {
#pragma unroll
        for (int l = 0; l <= LMAX; ++l) j[l] = jl_f32(l, t);
    }
}

// FP64 counterpart, j_0 .. j_LMAX at x = 2 pi t: one sincos and the upward
// recurrence of jl_f64 when x >= LMAX + 1 (jl_f64 takes the same branch for
// every order there); otherwise jl_f64 order by order.
template <int LMAX>
__device__ __forceinline__ void radial_jl(double t, double* j) {
    const double x = 2.0 * HERMES_PI * t;
    if (LMAX >= 1 && x >= (double)LMAX + 1.0) {
        double sx, cx;
        sincos(x, &sx, &cx);
        j[0] = sx / x;
        j[1] = sx / (x * x) - cx / x;
#pragma unroll
        for (int n = 1; n < LMAX; ++n) j[n + 1] = ((2.0 * n + 1.0) / x) * j[n] - j[n - 1];
    } else {
#pragma unroll
        for (int l = 0; l <= LMAX; ++l) j[l] = jl_f64(l, x);
    }
}

// Pair reduction for thin shells and m = 0 shell multipoles, grouped by radius.
// For octant representative (x, y, z) and its 8 aliases a, the window of radius R
// and order l is sum_a P(a) j_l(2 pi |a| R / L) P_l(a_z / |a|) (times the
// descriptor's constant: 1 for a shell, N_l0 for a multipole).  The alias geometry
// (P product, |a|, Legendre values) is computed once per representative and one
// sin/cos per (alias, radius) serves every order.  Channel c = (radius u, order l)
// accumulates sum_k K_c(k) Re F1 conj F2 in FP64; chan[u * (LMAX + 1) + l] is the
// channel of (u, l) or -1.
// T = float: FP32 windows (FP64 power and accumulation); T = double: all FP64.
//
// ROT (FP64 thin shells, equally spaced radii R_u = R0 + u dR): sin and cos of x = 2 pi |a| R_u / L
// come from those at R0 by rotating through the angle 2 pi |a| dR / L (both reduced from exact turn
// fractions once per alias), and 1 / x from 1 / (2 pi |a|) times 1 / R_u: four FMAs and two
// multiplies per (alias, radius) in place of an FP64 sincos and a division.  x < 1 takes radial_jl
// as before.  Each rotation adds about two ulps to |sin|, |cos| <= 1.
template <int LMAX, typename T, typename C, bool ROT = false>
__global__ __launch_bounds__(RED_THREADS) void k_pair_reduce_radial(int L, const C* __restrict__ F1,
                                                                    const C* __restrict__ F2,
                                                                    const double* __restrict__ P,
                                                                    const double* __restrict__ radii_over_L, int nu,
                                                                    double R0, double dR,
                                                                    const int* __restrict__ chan, int nch,
                                                                    double* __restrict__ partial) {
    static_assert(!ROT || (LMAX == 0 && sizeof(T) == 8), "rotated phases: FP64 thin shells");
    extern __shared__ double wacc[];  // [waves][nch]
    const int Lh = L / 2 + 1;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int nwaves = blockDim.x / 64;
    for (int i = threadIdx.x; i < nwaves * nch; i += blockDim.x) wacc[i] = 0.0;
    __syncthreads();
    const int64_t nrep = (int64_t)Lh * Lh * Lh;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < nrep; base += stride) {
        const int64_t r = base + threadIdx.x;
        const bool active = r < nrep;
        double power = 0.0;
        int x = 0, y = 0, z = 0;
        if (active) {
            z = (int)(r % Lh);
            y = (int)((r / Lh) % Lh);
            x = (int)(r / ((int64_t)Lh * Lh));
            const int xm = (L - x) % L, ym = (L - y) % L;
            const int nxm = (xm != x) ? 2 : 1, nym = (ym != y) ? 2 : 1;
            const double wz = (z == 0 || z == L / 2) ? 1.0 : 2.0;
            for (int a = 0; a < nxm; ++a)
                for (int b = 0; b < nym; ++b) {
                    const int xx = a ? xm : x, yy = b ? ym : y;
                    const int64_t idx = ((int64_t)xx * L + yy) * Lh + z;
                    const C f1 = F1[idx], f2 = F2[idx];
                    power += wz * ((double)f1.x * (double)f2.x + (double)f1.y * (double)f2.y);
                }
        }
        double sa[8];
        T wl[8][LMAX + 1];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const int ax = (a & 4) ? x - L : x, ay = (a & 2) ? y - L : y, az = (a & 1) ? z - L : z;
            const int ix = ax < 0 ? -ax : ax, iy = ay < 0 ? -ay : ay, iz = az < 0 ? -az : az;
            const T p = (T)(P[ix] * P[iy] * P[iz]);
            const int n2 = ax * ax + ay * ay + az * az;
            sa[a] = sqrt((double)n2);
            const T mu = n2 ? (T)((double)az / sa[a]) : (T)0;
            T pm = 1, pc = mu;  // P_0, P_1
            wl[a][0] = p;
            if (LMAX >= 1) wl[a][1] = n2 ? p * mu : (T)0;
#pragma unroll
            for (int n = 1; n < LMAX; ++n) {
                const T pn = ((T)(2 * n + 1) * mu * pc - (T)n * pm) / (T)(n + 1);
                pm = pc;
                pc = pn;
                wl[a][n + 1] = n2 ? p * pn : (T)0;
            }
        }
        // ROT: per alias sin, cos at the current radius, the rotation step and 1 / (2 pi |a|) (+inf for
        // a = 0, which then always takes radial_jl)
        double sn[8], cs[8], sd[8], cd[8], iqa[8];
        if (ROT) {
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const double t0 = sa[a] * R0, td = sa[a] * dR;
                sincospi(2.0 * (t0 - floor(t0)), &sn[a], &cs[a]);
                sincospi(2.0 * (td - floor(td)), &sd[a], &cd[a]);
                iqa[a] = sa[a] > 0.0 ? 1.0 / (2.0 * HERMES_PI * sa[a]) : __builtin_inf();
            }
        }
        for (int u = 0; u < nu; ++u) {
            const double R = radii_over_L[u];
            const double iR = ROT ? 1.0 / R : 0.0;
            T kb[LMAX + 1];
#pragma unroll
            for (int l = 0; l <= LMAX; ++l) kb[l] = 0;
            unsigned small = 0;  // ROT: aliases with x < 1, evaluated below
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                T j[LMAX + 1];
                if (ROT) {
                    if (u > 0) {
                        const double s2 = fma(sn[a], cd[a], cs[a] * sd[a]);
                        cs[a] = fma(cs[a], cd[a], -(sn[a] * sd[a]));
                        sn[a] = s2;
                    }
                    if (R >= iqa[a]) {  // x = R / iqa >= 1
                        kb[0] += sn[a] * (iqa[a] * iR) * wl[a][0];
                    } else {
                        small |= 1u << a;
                    }
                } else {
                    radial_jl<LMAX>(sa[a] * R, j);
#pragma unroll
                    for (int l = 0; l <= LMAX; ++l) kb[l] += j[l] * wl[a][l];
                }
            }
            // one copy of the small-argument path (it is rare: |a| R / L < 1 / (2 pi)), outside the
            // unrolled alias loop, so its registers do not add to the rotation state's
            for (; ROT && small; small &= small - 1) {
                const int a = __builtin_ctz(small);
                const int ax = (a & 4) ? x - L : x, ay = (a & 2) ? y - L : y, az = (a & 1) ? z - L : z;
                const int ix = ax < 0 ? -ax : ax, iy = ay < 0 ? -ay : ay, iz = az < 0 ? -az : az;
                const double t = sqrt((double)(ax * ax + ay * ay + az * az)) * R;
                kb[0] += (T)(jl_f64(0, 2.0 * HERMES_PI * t) * (P[ix] * P[iy] * P[iz]));
            }
#pragma unroll
            for (int l = 0; l <= LMAX; ++l) {
                const int c = chan[u * (LMAX + 1) + l];
                if (c >= 0) {
                    const double v = wave_sum(power * (double)kb[l]);
                    if (lane == 0) wacc[wave * nch + c] += v;
                }
            }
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nch; b += blockDim.x) {
        double s = 0.0;
        for (int w = 0; w < nwaves; ++w) s += wacc[w * nch + b];
        partial[(int64_t)blockIdx.x * nch + b] = s;
    }
}

// First stage of the fixed-order sum of block partials (nblocks rows of nch): groups of
// PART_GROUP consecutive rows, one thread per (group, column); the kernels below finish over
// the (nblocks / PART_GROUP) group rows.
constexpr int PART_GROUP = 64;

__global__ void k_sum_partial_groups(const double* __restrict__ partial, int nblocks, int nch,
                                     double* __restrict__ gsum) {
    const int ngroups = (nblocks + PART_GROUP - 1) / PART_GROUP;
    const int64_t total = (int64_t)ngroups * nch;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int g = (int)(q / nch), c = (int)(q % nch);
        const int r1 = min(nblocks, (g + 1) * PART_GROUP);
        double s = 0.0;
        for (int r = g * PART_GROUP; r < r1; ++r) s += partial[(int64_t)r * nch + c];
        gsum[q] = s;
    }
}

static inline int part_groups(int nblocks) { return (nblocks + PART_GROUP - 1) / PART_GROUP; }

// out[d] = factor[d] * (sum over rows of channel ch_of[d]), fixed order.
__global__ void k_plan_out(const double* __restrict__ partial, int nblocks, int nch, const int* __restrict__ ch_of,
                           const double* __restrict__ factor, int n_desc, double* __restrict__ out) {
    for (int d = blockIdx.x * blockDim.x + threadIdx.x; d < n_desc; d += gridDim.x * blockDim.x) {
        const int c = ch_of[d];
        double s = 0.0;
        for (int i = 0; i < nblocks; ++i) s += partial[(int64_t)i * nch + c];
        out[d] = factor[d] * s;
    }
}

__global__ void k_sum_partials(const double* __restrict__ partial, int nblocks, int nd, double* __restrict__ out) {
    // one thread per output; fixed order -> deterministic
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nd; b += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < nblocks; ++i) s += partial[(int64_t)i * nd + b];
        out[b] = s;
    }
}

// ---------------------------------------------------------------- multipole apply (FP32)
constexpr int MAX_SHELLS = 64;
constexpr int SHELL_CHUNK = 16;

struct ApplyArgs {
    int L, l, m, n_shell;
    float cnorm, scale;
    float sx, sy, sz;  // sign factors for x-, y- (1) and z-mirrors: (-1)^m, (-1)^(l+m)
    double radii_over_L[MAX_SHELLS];
};

// Several orders m of one l (the fused window + z-FFT kernel shares the radial factors among them).
constexpr int MAX_MP = 2;
struct MultiM {
    int n;                  // 1..MAX_MP
    int l[MAX_MP];          // degree of each order (one l, or l and l + 1 for the packed m = 0 pair)
    int m[MAX_MP];
    float cnorm[MAX_MP];    // N_l|m| (-1)^|m| (2|m|-1)!!
    float sx[MAX_MP], sz[MAX_MP];
};

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

__global__ __launch_bounds__(256) void k_multipole_apply(const float2* __restrict__ H, const float* __restrict__ Pf,
                                                         ApplyArgs A, float2* __restrict__ out) {
    const int L = A.L, Lh = L / 2 + 1;
    const int64_t nrep = (int64_t)Lh * Lh * Lh;
    const int64_t N = (int64_t)L * L * L;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nrep; r += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(r % Lh), y = (int)((r / Lh) % Lh), x = (int)(r / ((int64_t)Lh * Lh));
        const int xm = (L - x) % L, ym = (L - y) % L, zm = (L - z) % L;
        // spectrum values at the 8 mirror modes from 4 half-spectrum reads
        const float2 h00 = H[((int64_t)x * L + y) * Lh + z], h10 = H[((int64_t)xm * L + y) * Lh + z];
        const float2 h01 = H[((int64_t)x * L + ym) * Lh + z], h11 = H[((int64_t)xm * L + ym) * Lh + z];
        // alias geometry
        float pw[8];
        float2 yv[8];
        double nrm[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            int ax = (a & 4) ? x - L : x, ay = (a & 2) ? y - L : y, az = (a & 1) ? z - L : z;
            pw[a] = Pf[ax < 0 ? -ax : ax] * Pf[ay < 0 ? -ay : ay] * Pf[az < 0 ? -az : az];
            nrm[a] = sqrt((double)ax * ax + (double)ay * ay + (double)az * az);
            float2 yy = ylm_f32(A.l, A.m < 0 ? -A.m : A.m, (float)ax, (float)ay, (float)az, A.cnorm);
            if (A.m < 0) {  // Y_{l,-m} = (-1)^m conj(Y_{l,m})
                float sg = ((-A.m) & 1) ? -1.0f : 1.0f;
                yy = make_float2(sg * yy.x, -sg * yy.y);
            }
            yv[a] = yy;
        }
        for (int s0 = 0; s0 < A.n_shell; s0 += SHELL_CHUNK) {
            float2 K[SHELL_CHUNK];
#pragma unroll
            for (int s = 0; s < SHELL_CHUNK; ++s) K[s] = make_float2(0.0f, 0.0f);
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                if (pw[a] == 0.0f) continue;
#pragma unroll
                for (int s = 0; s < SHELL_CHUNK; ++s) {
                    if (s0 + s < A.n_shell) {
                        float j = jl_f32(A.l, nrm[a] * A.radii_over_L[s0 + s]) * pw[a];
                        K[s].x += j * yv[a].x;
                        K[s].y += j * yv[a].y;
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < SHELL_CHUNK; ++s) {
                if (s0 + s >= A.n_shell) break;
                float2* o = out + (int64_t)(s0 + s) * N;
                const float2 k = make_float2(K[s].x * A.scale, K[s].y * A.scale);
                const float2 kc = conjf2(k);
                // x-mirror: (-1)^m conj K; y-mirror: conj K; z-mirror: (-1)^(l+m) K
                o[((int64_t)x * L + y) * L + z] = cmulf(h00, k);
                if (xm != x) o[((int64_t)xm * L + y) * L + z] = cmulf(h10, make_float2(A.sx * kc.x, A.sx * kc.y));
                if (ym != y) o[((int64_t)x * L + ym) * L + z] = cmulf(h01, kc);
                if (xm != x && ym != y) o[((int64_t)xm * L + ym) * L + z] = cmulf(h11, make_float2(A.sx * k.x, A.sx * k.y));
                if (zm != z) {
                    const float szf = A.sz;
                    o[((int64_t)x * L + y) * L + zm] = cmulf(conjf2(h11), make_float2(szf * k.x, szf * k.y));
                    if (xm != x)
                        o[((int64_t)xm * L + y) * L + zm] =
                            cmulf(conjf2(h01), make_float2(szf * A.sx * kc.x, szf * A.sx * kc.y));
                    if (ym != y) o[((int64_t)x * L + ym) * L + zm] = cmulf(conjf2(h10), make_float2(szf * kc.x, szf * kc.y));
                    if (xm != x && ym != y)
                        o[((int64_t)xm * L + ym) * L + zm] =
                            cmulf(conjf2(h00), make_float2(szf * A.sx * k.x, szf * A.sx * k.y));
                }
            }
        }
    }
}

// Fused multipole window and inverse z-FFT (FP32).  Workgroup (x, y), 0 <= x, y <= L/2, evaluates
// K_{R_s,l,m} at its octant representatives (x, y, z <= L/2) exactly as k_multipole_apply does, forms
// the four mirror lines (x, y), (xm, y), (x, ym), (xm, ym) of H_full * K * scale in LDS with the same
// mirror rules, transforms them along z and writes each distinct line once.  Loops over shells.

// MP orders m of one l per launch: the radial factors j_l(2 pi |a| R_s) of the 8 aliases are
// evaluated once per (representative, shell) and shared by the MP windows; the weights
// P(a) Y_lm(a-hat) are kept per order (premultiplied).  The LDS z-FFT runs over 4 MP lines
// (line mi * 4 + row) and order mi's shell s lands at out + (mi * n_shell + s) * L^3.
template <int LOG2L>
constexpr int mpwin_threads() {
    // one z representative 0 .. L/2 - 1 per thread (at least a wave); for L >= 128 the Nyquist
    // plane z = L/2 is left over and its eight alias terms are computed by eight helper lanes
    return 64 * (((1 << LOG2L) / 2 + 63) / 64);
}

// PACK (MP = 2): the orders are (l, 0) and (l + 1, 0), whose windows are real with opposite parity
// under k -> -k, so the field of the even degree is real and that of the odd one imaginary: their
// windows are summed into one set of four lines (one field per shell, half the transforms), and the
// real and imaginary products of the Gram separate them again.
// PROF: the radial factor of shell s comes from the radial profile prof[s] (thin shell, table or closed-form
// monopole) instead of j_l(2 pi |a| R_s / L); prof is unused otherwise.
template <int LOG2L, int MP, bool PACK = false, bool PROF = false>
__global__ __launch_bounds__(mpwin_threads<LOG2L>()) void k_mpwin_zfft_multi(
    const float2* __restrict__ H, const float* __restrict__ Pf, ApplyArgs A, MultiM M,
    const float2* __restrict__ twg, float2* __restrict__ out, const hermes_radial_profile* __restrict__ prof) {
    static_assert(!PACK || MP == 2, "packing needs two orders");
    static_assert(!(PACK && PROF), "packed pairs take thin shells only");
    constexpr int L = 1 << LOG2L, Lh = L / 2 + 1;
    constexpr int NT = mpwin_threads<LOG2L>();
    constexpr bool NYQ = Lh > NT;  // z = L/2 by helper lanes 0..7
    constexpr int NLINE = PACK ? 4 : 4 * MP;  // lines transformed per shell
    using F = hermes_fft::LdsFft<LOG2L, NLINE, NT, 1, false>;
    __shared__ float2 lds[F::ELEMS];
    __shared__ float2 tw[L];
    __shared__ float2 nyq[8][MP];
    const int tid = threadIdx.x;
    const int x = blockIdx.x / Lh, y = blockIdx.x % Lh;
    const int xm = (L - x) & (L - 1), ym = (L - y) & (L - 1);
    const int64_t N = (int64_t)L * L * L;
    hermes_fft::load_twiddles<LOG2L>(tw, twg, tid, NT);
    // weights P(a) Y_lm(a-hat) of the 8 aliases per order (premultiplied) for this thread's z and, on the
    // helper lanes, alias tid of the Nyquist plane
    auto weights = [&](int z, int a, float2* w, double& nrm, bool& live) {
        const int ax = (a & 4) ? x - L : x, ay = (a & 2) ? y - L : y, az = (a & 1) ? z - L : z;
        const float pw = Pf[ax < 0 ? -ax : ax] * Pf[ay < 0 ? -ay : ay] * Pf[az < 0 ? -az : az];
        live = pw != 0.0f;
        nrm = sqrt((double)ax * ax + (double)ay * ay + (double)az * az);
#pragma unroll
        for (int mi = 0; mi < MP; ++mi) {
            const int m = M.m[mi], am = m < 0 ? -m : m;
            float2 yy = ylm_f32(M.l[mi], am, (float)ax, (float)ay, (float)az, M.cnorm[mi]);
            if (m < 0) {  // Y_{l,-m} = (-1)^m conj(Y_{l,m})
                const float sg = (am & 1) ? -1.0f : 1.0f;
                yy = make_float2(sg * yy.x, -sg * yy.y);
            }
            w[mi] = make_float2(pw * yy.x, pw * yy.y);
        }
    };
    const int z = tid;
    const bool zact = z < Lh;
    float2 wy[8][MP];
    double nrm[8];
    float2 h[4];
    unsigned live = 0;
    if (zact) {
        h[0] = H[((int64_t)x * L + y) * Lh + z];
        h[1] = H[((int64_t)xm * L + y) * Lh + z];
        h[2] = H[((int64_t)x * L + ym) * Lh + z];
        h[3] = H[((int64_t)xm * L + ym) * Lh + z];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            bool lv;
            weights(z, a, wy[a], nrm[a], lv);
            if (lv) live |= 1u << a;
        }
    }
    float2 wn[MP], hn[4];
    double nrm_n = 0.0;
    bool live_n = false;
    if (NYQ && tid < 8) weights(L / 2, tid, wn, nrm_n, live_n);
    if (NYQ && tid < MP) {
        hn[0] = H[((int64_t)x * L + y) * Lh + L / 2];
        hn[1] = H[((int64_t)xm * L + y) * Lh + L / 2];
        hn[2] = H[((int64_t)x * L + ym) * Lh + L / 2];
        hn[3] = H[((int64_t)xm * L + ym) * Lh + L / 2];
    }
    // the four mirror rows of representative z (and, for z != -z, the rows of -z) for order mi
    auto put = [&](int mi, int zz, const float2 hh[4], float2 Km) {
        const float2 k = make_float2(Km.x * A.scale, Km.y * A.scale);
        const float2 kc = conjf2(k);
        const float sx = M.sx[mi], sz = M.sz[mi];
        const int r0 = 4 * mi;
        lds[F::addr(r0 + 0, zz)] = cmulf(hh[0], k);
        lds[F::addr(r0 + 1, zz)] = cmulf(hh[1], make_float2(sx * kc.x, sx * kc.y));
        lds[F::addr(r0 + 2, zz)] = cmulf(hh[2], kc);
        lds[F::addr(r0 + 3, zz)] = cmulf(hh[3], make_float2(sx * k.x, sx * k.y));
        const int zm = (L - zz) & (L - 1);
        if (zm != zz) {
            lds[F::addr(r0 + 0, zm)] = cmulf(conjf2(hh[3]), make_float2(sz * k.x, sz * k.y));
            lds[F::addr(r0 + 1, zm)] = cmulf(conjf2(hh[2]), make_float2(sz * sx * kc.x, sz * sx * kc.y));
            lds[F::addr(r0 + 2, zm)] = cmulf(conjf2(hh[1]), make_float2(sz * kc.x, sz * kc.y));
            lds[F::addr(r0 + 3, zm)] = cmulf(conjf2(hh[0]), make_float2(sz * sx * k.x, sz * sx * k.y));
        }
    };
    // packed pair: the sum of the two orders' mirrored window values on each of the four lines (the
    // products with H distribute over the sum)
    auto put_pack = [&](int zz, const float2 hh[4], float2 K0, float2 K1) {
        const float2 k0 = make_float2(K0.x * A.scale, K0.y * A.scale), kc0 = conjf2(k0);
        const float2 k1 = make_float2(K1.x * A.scale, K1.y * A.scale), kc1 = conjf2(k1);
        const float sx0 = M.sx[0], sz0 = M.sz[0], sx1 = M.sx[1], sz1 = M.sz[1];
        auto mix = [](float a0, float2 u0, float a1, float2 u1) {
            return make_float2(a0 * u0.x + a1 * u1.x, a0 * u0.y + a1 * u1.y);
        };
        lds[F::addr(0, zz)] = cmulf(hh[0], mix(1.0f, k0, 1.0f, k1));
        lds[F::addr(1, zz)] = cmulf(hh[1], mix(sx0, kc0, sx1, kc1));
        lds[F::addr(2, zz)] = cmulf(hh[2], mix(1.0f, kc0, 1.0f, kc1));
        lds[F::addr(3, zz)] = cmulf(hh[3], mix(sx0, k0, sx1, k1));
        const int zm = (L - zz) & (L - 1);
        if (zm != zz) {
            lds[F::addr(0, zm)] = cmulf(conjf2(hh[3]), mix(sz0, k0, sz1, k1));
            lds[F::addr(1, zm)] = cmulf(conjf2(hh[2]), mix(sz0 * sx0, kc0, sz1 * sx1, kc1));
            lds[F::addr(2, zm)] = cmulf(conjf2(hh[1]), mix(sz0, kc0, sz1, kc1));
            lds[F::addr(3, zm)] = cmulf(conjf2(hh[0]), mix(sz0 * sx0, k0, sz1 * sx1, k1));
        }
    };
    // radial factor of an alias of norm nrm (grid points) for each order (one l, or l and l + 1 when packed)
    hermes_radial_profile pr{};
    auto radial = [&](double nrm_a, double rs, float* j) {
        if (PACK) {
            jl_pair_f32(M.l[0], nrm_a * rs, j[0], j[MP - 1]);
        } else {
            float jj;
            if (PROF)
                jj = pr.kind == HERMES_RADIAL_SHELL ? jl_f32(M.l[0], nrm_a * (pr.a0 / L))
                                                    : (float)radial_profile_f64(pr, M.l[0], nrm_a / L);
            else
                jj = jl_f32(M.l[0], nrm_a * rs);
#pragma unroll
            for (int mi = 0; mi < MP; ++mi) j[mi] = jj;
        }
    };
    for (int s = 0; s < A.n_shell; ++s) {
        const double rs = A.radii_over_L[s];
        if (PROF) pr = prof[s];
        if (zact) {
            float2 K[MP];
#pragma unroll
            for (int mi = 0; mi < MP; ++mi) K[mi] = make_float2(0.0f, 0.0f);
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                if (!((live >> a) & 1u)) continue;
                float j[MP];
                radial(nrm[a], rs, j);
#pragma unroll
                for (int mi = 0; mi < MP; ++mi) {
                    K[mi].x += j[mi] * wy[a][mi].x;
                    K[mi].y += j[mi] * wy[a][mi].y;
                }
            }
            if (PACK) {
                put_pack(z, h, K[0], K[MP - 1]);
            } else {
#pragma unroll
                for (int mi = 0; mi < MP; ++mi) put(mi, z, h, K[mi]);
            }
        }
        if (NYQ) {
            if (tid < 8) {
                float j[MP];
                if (live_n) {
                    radial(nrm_n, rs, j);
                } else {
#pragma unroll
                    for (int mi = 0; mi < MP; ++mi) j[mi] = 0.0f;
                }
#pragma unroll
                for (int mi = 0; mi < MP; ++mi) nyq[tid][mi] = make_float2(j[mi] * wn[mi].x, j[mi] * wn[mi].y);
            }
            hermes_fft::lds_barrier();
            if (tid < (PACK ? 1 : MP)) {
                float2 K[MP];
#pragma unroll
                for (int mi = 0; mi < MP; ++mi) K[mi] = make_float2(0.0f, 0.0f);
                for (int a = 0; a < 8; ++a) {
#pragma unroll
                    for (int mi = 0; mi < MP; ++mi) {
                        if (!PACK && mi != tid) continue;
                        K[mi].x += nyq[a][mi].x;
                        K[mi].y += nyq[a][mi].y;
                    }
                }
                if (PACK) {
                    put_pack(L / 2, hn, K[0], K[MP - 1]);
                } else {
#pragma unroll
                    for (int mi = 0; mi < MP; ++mi)
                        if (mi == tid) put(mi, L / 2, hn, K[mi]);
                }
            }
        }
        hermes_fft::lds_barrier();
        F::run(lds, tw, tid);
        // lines known at compile time: the skip of a mirror line that coincides with line 0 / 2 and the
        // line's base address are uniform, and a line's LDS reads are all issued before its stores
        constexpr int EPT = (L + NT - 1) / NT;  // elements per thread and line
#pragma unroll
        for (int line = 0; line < NLINE; ++line) {
            const int mi = line >> 2, row = line & 3;
            const bool dx = row & 1, dy = row & 2;
            if ((dx && xm == x) || (dy && ym == y)) continue;
            const int rx = dx ? xm : x, ry = dy ? ym : y;
            float2* __restrict__ o = out + (int64_t)(mi * A.n_shell + s) * N + ((int64_t)rx * L + ry) * L;
            float2 v[EPT];
#pragma unroll
            for (int q = 0; q < EPT; ++q)
                if (L % NT == 0 || tid + q * NT < L) v[q] = lds[F::addr(line, tid + q * NT)];
#pragma unroll
            for (int q = 0; q < EPT; ++q)
                if (L % NT == 0 || tid + q * NT < L) {
                    if (LOG2L >= 8)  // the fields are written once and read by the next pass from HBM
                        hermes_fft::stream_store(o + tid + q * NT, v[q]);
                    else
                        o[tid + q * NT] = v[q];
                }
        }
        hermes_fft::lds_barrier();
    }
}

__global__ void k_p_to_float(const double* __restrict__ P, int n, float* __restrict__ Pf) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) Pf[i] = (float)P[i];
}

// ---------------------------------------------------------------- FP64 sums
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_sum(const T* __restrict__ x, int64_t n, int ncomp,
                                                            double* __restrict__ partial) {
    __shared__ double sh[2][RED_THREADS / 64];
    double s0 = 0.0, s1 = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (ncomp == 1) {
            s0 += (double)x[i];
        } else {
            s0 += (double)x[2 * i];
            s1 += (double)x[2 * i + 1];
        }
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    if (threadIdx.x % 64 == 0) {
        sh[0][threadIdx.x / 64] = s0;
        sh[1][threadIdx.x / 64] = s1;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double t = 0.0;
        for (int w = 0; w < RED_THREADS / 64; ++w) t += sh[threadIdx.x][w];
        partial[2 * blockIdx.x + threadIdx.x] = t;
    }
}

__global__ void k_reduce_final(const double* __restrict__ partial, int nb, double* __restrict__ out) {
    if (threadIdx.x < 2) {
        double t = 0.0;
        for (int i = 0; i < nb; ++i) t += partial[2 * i + threadIdx.x];
        out[threadIdx.x] = t;
    }
}

int check_desc(const hermes_window_desc& d) {
    if (d.type < 0 || d.type > HERMES_WIN_CYLINDER) {
        hermes_set_error("unknown window type %d", d.type);
        return 2;
    }
    if (d.type >= HERMES_WIN_RING && (d.axis < 0 || d.axis > 3)) {
        hermes_set_error("line-of-sight window: axis %d is not 0, 1, 2 or 3", d.axis);
        return 2;
    }
    if (d.type >= HERMES_WIN_RING && d.axis == 3) {
        const double n2 = d.los[0] * d.los[0] + d.los[1] * d.los[1] + d.los[2] * d.los[2];
        if (!(std::fabs(n2 - 1.0) < 1e-12)) {
            hermes_set_error("line-of-sight window: los is not a unit vector (|los|^2 = %g)", n2);
            return 2;
        }
    }
    if (d.type == HERMES_WIN_SHELL_MULTIPOLE && (d.l < 0 || d.l > 40 || d.m > d.l || d.m < -d.l)) {
        hermes_set_error("invalid multipole (l=%d, m=%d)", d.l, d.m);
        return 2;
    }
    return 0;
}

double factorial_d(int n) {
    double o = 1.0;
    for (int i = 2; i <= n; ++i) o *= i;
    return o;
}

}  // namespace

// N_lm (-1)^m (2m-1)!! for the FP32 spherical harmonics (m >= 0).
double hermes_ylm_cnorm(int l, int m) {
    double n = sqrt((2.0 * l + 1.0) / (4.0 * HERMES_PI) * factorial_d(l - m) / factorial_d(l + m));
    double df = 1.0;
    for (int i = 1; i <= m; ++i) df *= (2.0 * i - 1.0);
    return ((m & 1) ? -1.0 : 1.0) * df * n;
}

extern "C" {

int hermes_window_build_rfft(hermes_ctx* ctx, int J, const double* P, const hermes_window_desc* desc, int dbl,
                             void* out) {
    HERMES_REQUIRE(ctx && desc && P && out, "hermes_window_build_rfft: null argument");
    HERMES_REQUIRE(J >= 1 && J <= 10, "J=%d unsupported", J);
    HERMES_REQUIRE(desc->type != HERMES_WIN_SHELL_MULTIPOLE, "multipole windows need the full complex layout");
    if (int rc = check_desc(*desc)) return rc;
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    int L = 1 << J;
    int64_t n = (int64_t)L * L * (L / 2 + 1);
    if (dbl) k_build_rfft<double><<<hermes_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(L, P, *desc, (double*)out);
    else k_build_rfft<float><<<hermes_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(L, P, *desc, (float*)out);
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_window_build_full(hermes_ctx* ctx, int J, const double* P, const hermes_window_desc* desc, int dbl,
                             void* out) {
    HERMES_REQUIRE(ctx && desc && P && out, "hermes_window_build_full: null argument");
    HERMES_REQUIRE(J >= 1 && J <= 10, "J=%d unsupported", J);
    if (int rc = check_desc(*desc)) return rc;
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    int L = 1 << J;
    int64_t n = (int64_t)L * L * L;
    if (dbl) k_build_full<double><<<hermes_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(L, P, *desc, (double*)out);
    else k_build_full<float><<<hermes_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(L, P, *desc, (float*)out);
    HERMES_LAUNCH_CHECK();
    return 0;
}

static int check_profile(const hermes_radial_profile& p, int l, const char* who) {
    HERMES_REQUIRE(p.kind >= HERMES_RADIAL_SHELL && p.kind <= HERMES_RADIAL_GAUSSIAN_SHELL, "%s: unknown radial kind %d",
                   who, p.kind);
    HERMES_REQUIRE(std::isfinite(p.a0) && std::isfinite(p.a1) && p.a0 >= 0.0, "%s: invalid profile lengths", who);
    HERMES_REQUIRE(p.kind <= HERMES_RADIAL_TABLE || l == 0,
                   "%s: closed-form profile kinds are monopoles (l = 0), got l = %d", who, l);
    if (p.kind == HERMES_RADIAL_TABLE)
        HERMES_REQUIRE(p.table && p.n_table >= 4 && std::isfinite(p.k_scale) && p.k_scale > 0.0,
                       "%s: a table profile needs a device table of >= 4 values and k_scale > 0", who);
    return 0;
}

int hermes_window_build_full_profile(hermes_ctx* ctx, int J, const double* P, const hermes_radial_profile* prof, int l,
                                     int m, int dbl, void* out) {
    HERMES_REQUIRE(ctx && P && prof && out, "hermes_window_build_full_profile: null argument");
    HERMES_REQUIRE(J >= 1 && J <= 10, "J=%d unsupported", J);
    HERMES_REQUIRE(l >= 0 && l <= 40 && m >= -l && m <= l, "invalid multipole (l=%d, m=%d)", l, m);
    if (int rc = check_profile(*prof, l, "hermes_window_build_full_profile")) return rc;
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int L = 1 << J;
    const int64_t n = (int64_t)L * L * L;
    if (dbl)
        k_build_full_profile<double><<<hermes_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(L, P, *prof, l, m,
                                                                                              (double*)out);
    else
        k_build_full_profile<float><<<hermes_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(L, P, *prof, l, m,
                                                                                             (float*)out);
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_spectrum_mul(hermes_ctx* ctx, int J, int half, int dbl, void* spec, const void* kernel, int kernel_complex,
                        double scale) {
    HERMES_REQUIRE(ctx && spec && kernel, "hermes_spectrum_mul: null argument");
    HERMES_REQUIRE(J >= 1 && J <= 10, "J=%d unsupported", J);
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    int L = 1 << J;
    int64_t n = (int64_t)L * L * (half ? (L / 2 + 1) : L);
    if (dbl)
        k_spec_mul<double><<<hermes_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(n, (double*)spec, (const double*)kernel,
                                                                                 kernel_complex, scale);
    else
        k_spec_mul<float><<<hermes_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(n, (float*)spec, (const float*)kernel,
                                                                                kernel_complex, (float)scale);
    HERMES_LAUNCH_CHECK();
    return 0;
}

// Radius-grouped channels of a descriptor list: unique (radius, order) pairs, each descriptor a
// channel times a constant.  False when some descriptor is not a thin shell or an m = 0 multipole
// of even order <= 10 (FP32) or <= 4 (FP64), or the channels do not fit the kernel's LDS.
struct RadialChannels {
    int LM = 0, nu = 0, nch = 0;
    int rot = 0;                 // FP64 thin shells at >= 3 equally spaced radii R0 + u dR: rotated phases
    double R0 = 0.0, dR = 0.0;
    std::vector<double> radii;   // R / L per radius
    std::vector<int> chan;       // nu x (LM + 1): channel or -1
    std::vector<int> ch_of;      // per descriptor
    std::vector<double> factor;  // per descriptor
};

static bool radial_channels(int J, int dbl, const hermes_window_desc* descs, int n_desc, RadialChannels& rc) {
    int lmax = 0;
    for (int b = 0; b < n_desc; ++b) {
        const hermes_window_desc& d = descs[b];
        if (d.type == HERMES_WIN_SHELL) continue;
        if (d.type != HERMES_WIN_SHELL_MULTIPOLE || d.m != 0 || d.l > (dbl ? 4 : 10)) return false;
        lmax = d.l > lmax ? d.l : lmax;
    }
    rc.LM = lmax == 0 ? 0 : (lmax <= 4 ? 4 : 10);
    const int L = 1 << J;
    rc.ch_of.resize(n_desc);
    rc.factor.resize(n_desc);
    for (int b = 0; b < n_desc; ++b) {
        const hermes_window_desc& d = descs[b];
        const double R = d.a0 / L;
        size_t u = 0;
        while (u < rc.radii.size() && rc.radii[u] != R) ++u;
        if (u == rc.radii.size()) {
            rc.radii.push_back(R);
            rc.chan.insert(rc.chan.end(), rc.LM + 1, -1);
        }
        const int l = d.type == HERMES_WIN_SHELL ? 0 : d.l;
        int& c = rc.chan[u * (rc.LM + 1) + l];
        if (c < 0) c = rc.nch++;
        rc.ch_of[b] = c;
        rc.factor[b] = d.type == HERMES_WIN_SHELL ? 1.0 : hermes_ylm_cnorm(d.l, 0);
    }
    rc.nu = (int)rc.radii.size();
    if (dbl && rc.LM == 0 && rc.nu >= 3) {
        const double r0 = rc.radii[0], d = (rc.radii[rc.nu - 1] - r0) / (rc.nu - 1);
        double rmax = 0.0;
        for (double r : rc.radii) rmax = std::max(rmax, std::fabs(r));
        bool uniform = d != 0.0;
        for (int u = 0; u < rc.nu && uniform; ++u) uniform = std::fabs(rc.radii[u] - (r0 + u * d)) <= 1e-12 * rmax;
        if (uniform) {
            rc.rot = 1;
            rc.R0 = r0;
            rc.dR = d;
        }
    }
    return sizeof(double) * (RED_THREADS / 64) * rc.nch <= 64 * 1024;
}

static hipError_t launch_radial(hipStream_t st, int J, int dbl, int LM, int nblocks, const void* F1, const void* F2,
                                const double* P, const double* d_radii, int nu, int rot, double R0, double dR,
                                const int* d_chan, int nch, double* partial) {
    const int L = 1 << J;
    const size_t shmem = sizeof(double) * (RED_THREADS / 64) * nch;
#define HERMES_RADIAL(LMV, T, C, ROT)                                                                      \
    k_pair_reduce_radial<LMV, T, C, ROT><<<nblocks, RED_THREADS, shmem, st>>>(L, (const C*)F1, (const C*)F2, P, \
                                                                             d_radii, nu, R0, dR, d_chan, nch, partial)
    if (dbl) {
        if (LM == 0 && rot) HERMES_RADIAL(0, double, double2, true);
        else if (LM == 0) HERMES_RADIAL(0, double, double2, false);
        else HERMES_RADIAL(4, double, double2, false);
    } else {
        if (LM == 0) HERMES_RADIAL(0, float, float2, false);
        else if (LM == 4) HERMES_RADIAL(4, float, float2, false);
        else HERMES_RADIAL(10, float, float2, false);
    }
#undef HERMES_RADIAL
    return hipGetLastError();
}

// Scratch of the pair reduction: the context's grow-only slot, or, for the allocation-free entry point, a
// fixed buffer of fixed_cap bytes (a request beyond it fails with status 2 and a message).
static char* pair_workspace(hermes_ctx* ctx, char* fixed_ws, size_t fixed_cap, size_t need) {
    if (!fixed_ws) return (char*)hermes_workspace(ctx, WS_REDUCE, need);
    if (need > fixed_cap) {
        hermes_set_error("hermes_pair_reduce_noalloc: this reduction needs %zu bytes of scratch, more than the "
                         "fixed %zu (split the windows over several calls)", need, fixed_cap);
        return nullptr;
    }
    return fixed_ws;
}

static int pair_blocks(int J) {
    const int Lh = (1 << J) / 2 + 1;
    return (int)hermes_blocks((int64_t)Lh * Lh * Lh, RED_THREADS, 2048);
}

// The radial path of hermes_pair_reduce.  Returns false (rc untouched) when it does not apply;
// true with rc = the status otherwise.
static bool pair_reduce_radial(hermes_ctx* ctx, int J, int dbl, const void* F1, const void* F2, const double* P,
                               const hermes_window_desc* descs, int n_desc, int nblocks, double* out_host,
                               char* fixed_ws, size_t fixed_cap, int& rc) {
    RadialChannels ch;
    if (!radial_channels(J, dbl, descs, n_desc, ch)) return false;
    rc = 0;
    const int nu = ch.nu, nch = ch.nch;
    const size_t off_chan = sizeof(double) * nu;
    const size_t off_part = ((off_chan + sizeof(int) * ch.chan.size() + 255) / 256) * 256;
    const int ng = part_groups(nblocks);
    const size_t need = off_part + sizeof(double) * ((size_t)nblocks * nch + (size_t)ng * nch + nch);
    char* ws = pair_workspace(ctx, fixed_ws, fixed_cap, need);
    if (!ws) {
        rc = fixed_ws ? 2 : 1;
        return true;
    }
    double* d_radii = (double*)ws;
    int* d_chan = (int*)(ws + off_chan);
    double* partial = (double*)(ws + off_part);
    double* gsum = partial + (size_t)nblocks * nch;
    double* d_out = gsum + (size_t)ng * nch;
    auto fail = [&](hipError_t e) {
        hermes_set_error("pair reduction: %s", hipGetErrorString(e));
        rc = 1;
        return true;
    };
    hipError_t e = hipMemcpyAsync(d_radii, ch.radii.data(), sizeof(double) * nu, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return fail(e);
    e = hipMemcpyAsync(d_chan, ch.chan.data(), sizeof(int) * ch.chan.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return fail(e);
    if ((e = launch_radial(ctx->stream, J, dbl, ch.LM, nblocks, F1, F2, P, d_radii, nu, ch.rot, ch.R0, ch.dR, d_chan,
                           nch, partial)) != hipSuccess)
        return fail(e);
    k_sum_partial_groups<<<hermes_blocks((int64_t)ng * nch, 256), 256, 0, ctx->stream>>>(partial, nblocks, nch, gsum);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    k_sum_partials<<<hermes_blocks(nch, 64), 64, 0, ctx->stream>>>(gsum, ng, nch, d_out);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    std::vector<double> chs(nch);
    e = hipMemcpyAsync(chs.data(), d_out, sizeof(double) * nch, hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return fail(e);
    for (int b = 0; b < n_desc; ++b) out_host[b] = ch.factor[b] * chs[ch.ch_of[b]];
    return true;
}

struct hermes_pair_plan {
    int J = 0, dbl = 0, LM = 0, nu = 0, nch = 0, n_desc = 0, nblocks = 0;
    int rot = 0;
    double R0 = 0.0, dR = 0.0;
    char* mem = nullptr;  // radii | chan | ch_of | factor | partial | group sums
    double* radii = nullptr;
    int* chan = nullptr;
    int* ch_of = nullptr;
    double* factor = nullptr;
    double* partial = nullptr;
    double* gsum = nullptr;
};

int hermes_pair_plan_destroy(hermes_pair_plan* plan) {
    if (!plan) return 0;
    if (plan->mem) (void)hipFree(plan->mem);
    delete plan;
    return 0;
}

int hermes_pair_plan_create(hermes_ctx* ctx, int J, int dbl, const hermes_window_desc* descs, int n_desc,
                            hermes_pair_plan** out) {
    HERMES_REQUIRE(ctx && descs && out, "hermes_pair_plan_create: null argument");
    HERMES_REQUIRE(!ctx->capturing, "hermes_pair_plan_create: the stream is being captured");
    HERMES_REQUIRE(J >= 1 && J <= 10, "J=%d unsupported", J);
    HERMES_REQUIRE(n_desc >= 1 && n_desc <= 4096, "hermes_pair_plan_create: n_desc=%d out of range", n_desc);
    *out = nullptr;
    for (int b = 0; b < n_desc; ++b) {
        if (int rc = check_desc(descs[b])) return rc;
        HERMES_REQUIRE(descs[b].type != HERMES_WIN_SHELL_MULTIPOLE || (descs[b].m == 0 && descs[b].l % 2 == 0),
                       "pair reduction supports only m = 0, even-l multipole windows");
    }
    RadialChannels ch;
    HERMES_REQUIRE(radial_channels(J, dbl, descs, n_desc, ch),
                   "hermes_pair_plan_create: plans take thin shells and even-l m = 0 shell multipoles "
                   "(l <= 10 in FP32, l <= 4 in FP64)");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    hermes_pair_plan* p = new hermes_pair_plan();
    p->J = J;
    p->dbl = dbl;
    p->LM = ch.LM;
    p->nu = ch.nu;
    p->nch = ch.nch;
    p->rot = ch.rot;
    p->R0 = ch.R0;
    p->dR = ch.dR;
    p->n_desc = n_desc;
    p->nblocks = pair_blocks(J);
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t o_chan = al(sizeof(double) * ch.nu), o_chof = o_chan + al(sizeof(int) * ch.chan.size()),
                 o_fac = o_chof + al(sizeof(int) * n_desc), o_part = o_fac + al(sizeof(double) * n_desc),
                 o_gs = o_part + al(sizeof(double) * (size_t)p->nblocks * ch.nch),
                 total = o_gs + sizeof(double) * (size_t)part_groups(p->nblocks) * ch.nch;
    if (hipMalloc(&p->mem, total) != hipSuccess) {
        delete p;
        hermes_set_error("hermes_pair_plan_create: hipMalloc of %zu bytes failed", total);
        return 1;
    }
    p->radii = (double*)p->mem;
    p->chan = (int*)(p->mem + o_chan);
    p->ch_of = (int*)(p->mem + o_chof);
    p->factor = (double*)(p->mem + o_fac);
    p->partial = (double*)(p->mem + o_part);
    p->gsum = (double*)(p->mem + o_gs);
    hipError_t e = hipMemcpy(p->radii, ch.radii.data(), sizeof(double) * ch.nu, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->chan, ch.chan.data(), sizeof(int) * ch.chan.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->ch_of, ch.ch_of.data(), sizeof(int) * n_desc, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->factor, ch.factor.data(), sizeof(double) * n_desc, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hermes_pair_plan_destroy(p);
        hermes_set_error("hermes_pair_plan_create: upload failed: %s", hipGetErrorString(e));
        return 1;
    }
    *out = p;
    return 0;
}

int hermes_pair_reduce_planned(hermes_ctx* ctx, const hermes_pair_plan* plan, const void* F1, const void* F2,
                               const double* P, double* out_dev) {
    HERMES_REQUIRE(ctx && plan && F1 && F2 && P && out_dev, "hermes_pair_reduce_planned: null argument");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    HERMES_CHECK_HIP(launch_radial(ctx->stream, plan->J, plan->dbl, plan->LM, plan->nblocks, F1, F2, P, plan->radii,
                                   plan->nu, plan->rot, plan->R0, plan->dR, plan->chan, plan->nch, plan->partial));
    const int ng = part_groups(plan->nblocks);
    k_sum_partial_groups<<<hermes_blocks((int64_t)ng * plan->nch, 256), 256, 0, ctx->stream>>>(
        plan->partial, plan->nblocks, plan->nch, plan->gsum);
    HERMES_LAUNCH_CHECK();
    k_plan_out<<<hermes_blocks(plan->n_desc, 64), 64, 0, ctx->stream>>>(plan->gsum, ng, plan->nch, plan->ch_of,
                                                                       plan->factor, plan->n_desc, out_dev);
    HERMES_LAUNCH_CHECK();
    return 0;
}

static int pair_reduce_impl(hermes_ctx* ctx, int J, int dbl, const void* F1, const void* F2, const double* P,
                            const hermes_window_desc* descs, int n_desc, double* out_host, char* fixed_ws,
                            size_t fixed_cap) {
    HERMES_REQUIRE(ctx && F1 && F2 && P && descs && out_host, "hermes_pair_reduce: null argument");
    HERMES_REQUIRE(J >= 1 && J <= 10, "J=%d unsupported", J);
    HERMES_REQUIRE(n_desc >= 1 && n_desc <= 4096, "hermes_pair_reduce: n_desc=%d out of range", n_desc);
    for (int b = 0; b < n_desc; ++b) {
        if (int rc = check_desc(descs[b])) return rc;
        HERMES_REQUIRE(descs[b].type != HERMES_WIN_SHELL_MULTIPOLE || descs[b].m == 0,
                       "pair reduction supports only m = 0 multipole windows");
        // the half-spectrum sum assumes W(-k) = W(k); an odd-l window is odd in k and its
        // pair product of real fields vanishes identically
        HERMES_REQUIRE(descs[b].type != HERMES_WIN_SHELL_MULTIPOLE || descs[b].l % 2 == 0,
                       "pair reduction supports only even-l multipole windows (odd l gives 0 for real fields)");
    }
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int L = 1 << J;
    const int Lh = L / 2 + 1;
    const int64_t nrep = (int64_t)Lh * Lh * Lh;
    const int nblocks = (int)hermes_blocks(nrep, RED_THREADS, 2048);
    {
        int rc = -1;
        if (pair_reduce_radial(ctx, J, dbl, F1, F2, P, descs, n_desc, nblocks, out_host, fixed_ws, fixed_cap, rc))
            return rc;
    }
    const int ng = part_groups(nblocks);
    size_t need = sizeof(hermes_window_desc) * n_desc + sizeof(float) * n_desc + 64 +
                  sizeof(double) * ((size_t)nblocks * n_desc + (size_t)ng * n_desc + n_desc);
    char* ws = pair_workspace(ctx, fixed_ws, fixed_cap, need);
    if (!ws) return fixed_ws ? 2 : 1;
    hermes_window_desc* d_desc = (hermes_window_desc*)ws;
    float* d_cn = (float*)(ws + sizeof(hermes_window_desc) * n_desc);
    double* partial = (double*)(ws + ((sizeof(hermes_window_desc) * n_desc + sizeof(float) * n_desc + 63) / 64) * 64);
    double* gsum = partial + (size_t)nblocks * n_desc;
    double* d_out = gsum + (size_t)ng * n_desc;
    std::vector<float> cn(n_desc);
    for (int b = 0; b < n_desc; ++b)
        cn[b] = descs[b].type == HERMES_WIN_SHELL_MULTIPOLE ? (float)hermes_ylm_cnorm(descs[b].l, 0) : 0.0f;
    HERMES_CHECK_HIP(hipMemcpyAsync(d_desc, descs, sizeof(hermes_window_desc) * n_desc, hipMemcpyHostToDevice, ctx->stream));
    HERMES_CHECK_HIP(hipMemcpyAsync(d_cn, cn.data(), sizeof(float) * n_desc, hipMemcpyHostToDevice, ctx->stream));
    size_t shmem = sizeof(double) * (RED_THREADS / 64) * n_desc;
    HERMES_REQUIRE(shmem <= 64 * 1024, "too many windows for one fused pass");
    bool gen = false;
    for (int b = 0; b < n_desc; ++b) gen = gen || (descs[b].type >= HERMES_WIN_RING && descs[b].axis == 3);
    if (dbl && gen)
        k_pair_reduce<double2, true, true><<<nblocks, RED_THREADS, shmem, ctx->stream>>>(
            L, (const double2*)F1, (const double2*)F2, P, d_desc, d_cn, n_desc, partial);
    else if (dbl)
        k_pair_reduce<double2, true, false><<<nblocks, RED_THREADS, shmem, ctx->stream>>>(
            L, (const double2*)F1, (const double2*)F2, P, d_desc, d_cn, n_desc, partial);
    else if (gen)
        k_pair_reduce<float2, false, true><<<nblocks, RED_THREADS, shmem, ctx->stream>>>(
            L, (const float2*)F1, (const float2*)F2, P, d_desc, d_cn, n_desc, partial);
    else
        k_pair_reduce<float2, false, false><<<nblocks, RED_THREADS, shmem, ctx->stream>>>(
            L, (const float2*)F1, (const float2*)F2, P, d_desc, d_cn, n_desc, partial);
    HERMES_LAUNCH_CHECK();
    k_sum_partial_groups<<<hermes_blocks((int64_t)ng * n_desc, 256), 256, 0, ctx->stream>>>(partial, nblocks, n_desc,
                                                                                            gsum);
    HERMES_LAUNCH_CHECK();
    k_sum_partials<<<hermes_blocks(n_desc, 64), 64, 0, ctx->stream>>>(gsum, ng, n_desc, d_out);
    HERMES_LAUNCH_CHECK();
    HERMES_CHECK_HIP(hipMemcpyAsync(out_host, d_out, sizeof(double) * n_desc, hipMemcpyDeviceToHost, ctx->stream));
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int hermes_pair_reduce(hermes_ctx* ctx, int J, int dbl, const void* F1, const void* F2, const double* P,
                       const hermes_window_desc* descs, int n_desc, double* out_host) {
    return pair_reduce_impl(ctx, J, dbl, F1, F2, P, descs, n_desc, out_host, nullptr, 0);
}

static int make_apply_args(hermes_ctx* ctx, int J, const double* P, const double* radii, int n_shell, int l,
                           int m, double scale, ApplyArgs* A, float** Pf_out) {
    HERMES_REQUIRE(n_shell >= 1 && n_shell <= MAX_SHELLS, "n_shell=%d outside 1..%d", n_shell, MAX_SHELLS);
    HERMES_REQUIRE(l >= 0 && l <= 40 && m >= -l && m <= l, "invalid multipole (l=%d, m=%d)", l, m);
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int L = 1 << J;
    float* Pf = (float*)hermes_workspace(ctx, WS_SPEC, sizeof(float) * (L + 1));
    HERMES_REQUIRE(Pf, "%s", hermes_last_error());
    k_p_to_float<<<hermes_blocks(L + 1, 256), 256, 0, ctx->stream>>>(P, L + 1, Pf);
    HERMES_LAUNCH_CHECK();
    A->L = L;
    A->l = l;
    A->m = m;
    A->n_shell = n_shell;
    const int am = m < 0 ? -m : m;
    A->cnorm = (float)hermes_ylm_cnorm(l, am);
    A->scale = (float)scale;
    A->sx = (am & 1) ? -1.0f : 1.0f;
    A->sy = 1.0f;
    A->sz = ((l + am) & 1) ? -1.0f : 1.0f;
    for (int s = 0; s < MAX_SHELLS; ++s) A->radii_over_L[s] = s < n_shell ? radii[s] / L : 0.0;
    *Pf_out = Pf;
    return 0;
}

int hermes_multipole_apply_f32(hermes_ctx* ctx, int J, const void* H, const double* P, const double* radii,
                               int n_shell, int l, int m, double scale, void* out) {
    HERMES_REQUIRE(ctx && H && P && radii && out, "hermes_multipole_apply_f32: null argument");
    HERMES_REQUIRE(J >= 2 && J <= 10, "J=%d unsupported", J);
    ApplyArgs A;
    float* Pf = nullptr;
    int rc = make_apply_args(ctx, J, P, radii, n_shell, l, m, scale, &A, &Pf);
    if (rc) return rc;
    const int L = 1 << J;
    const int64_t nrep = (int64_t)(L / 2 + 1) * (L / 2 + 1) * (L / 2 + 1);
    k_multipole_apply<<<hermes_blocks(nrep, 256, 1 << 20), 256, 0, ctx->stream>>>((const float2*)H, Pf, A,
                                                                                   (float2*)out);
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_multipole_window_ifft_f32(hermes_ctx* ctx, int J, const void* H, const double* P, const double* radii,
                                     int n_shell, int l, int m, double scale, void* out) {
    return hermes_multipole_window_ifft_multi_f32(ctx, J, H, P, radii, n_shell, l, &m, 1, scale, out);
}

// The window + z-pass launch of n_m orders of one l: thin shells at radii (prof_dev = nullptr), or per shell
// the device radial profiles prof_dev (radii unused).
static int mpwin_zpass(hermes_ctx* ctx, int J, const void* H, const double* P, const double* radii,
                       const hermes_radial_profile* prof_dev, int n_shell, int l, const int* ms_host, int n_m,
                       double scale, void* out);

int hermes_multipole_window_zpass_multi_f32(hermes_ctx* ctx, int J, const void* H, const double* P,
                                            const double* radii, int n_shell, int l, const int* ms_host, int n_m,
                                            double scale, void* out) {
    HERMES_REQUIRE(radii, "hermes_multipole_window_zpass_multi_f32: null argument");
    return mpwin_zpass(ctx, J, H, P, radii, nullptr, n_shell, l, ms_host, n_m, scale, out);
}

int hermes_multipole_window_ifft_profiles_f32(hermes_ctx* ctx, int J, const void* H, const double* P,
                                              const hermes_radial_profile* profs_host, int n_shell, int l,
                                              const int* ms_host, int n_m, double scale, void* out) {
    HERMES_REQUIRE(ctx && profs_host, "hermes_multipole_window_ifft_profiles_f32: null argument");
    HERMES_REQUIRE(!ctx->capturing, "hermes_multipole_window_ifft_profiles_f32: not capturable");
    HERMES_REQUIRE(n_shell >= 1 && n_shell <= MAX_SHELLS, "n_shell=%d outside 1..%d", n_shell, MAX_SHELLS);
    for (int s = 0; s < n_shell; ++s)
        if (int rc = check_profile(profs_host[s], l, "hermes_multipole_window_ifft_profiles_f32")) return rc;
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    auto* dev = (hermes_radial_profile*)hermes_workspace(ctx, WS_PROFILE, sizeof(hermes_radial_profile) * n_shell);
    HERMES_REQUIRE(dev, "%s", hermes_last_error());
    HERMES_CHECK_HIP(hipMemcpyAsync(dev, profs_host, sizeof(hermes_radial_profile) * n_shell, hipMemcpyHostToDevice,
                                    ctx->stream));
    const int rc = mpwin_zpass(ctx, J, H, P, nullptr, dev, n_shell, l, ms_host, n_m, scale, out);
    if (rc) return rc;
    // the pageable source of the copy must outlive it: wait before returning to the caller
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return hermes_fft3_inverse_yx(ctx, J, n_m * n_shell, (float2*)out);
}

static int mpwin_zpass(hermes_ctx* ctx, int J, const void* H, const double* P, const double* radii,
                       const hermes_radial_profile* prof_dev, int n_shell, int l, const int* ms_host, int n_m,
                       double scale, void* out) {
    HERMES_REQUIRE(ctx && H && P && ms_host && out, "hermes_multipole_window_zpass_multi_f32: null argument");
    HERMES_REQUIRE(J >= 3 && J <= 10, "hermes_multipole_window_zpass_multi_f32: J=%d outside 3..10", J);
    HERMES_REQUIRE(n_m >= 1 && n_m <= MAX_MP, "hermes_multipole_window_zpass_multi_f32: n_m=%d outside 1..%d", n_m,
                   MAX_MP);
    ApplyArgs A;
    float* Pf = nullptr;
    const double zeros[MAX_SHELLS] = {0.0};
    int rc = make_apply_args(ctx, J, P, radii ? radii : zeros, n_shell, l, ms_host[0], scale, &A, &Pf);
    if (rc) return rc;
    MultiM M{};
    M.n = n_m;
    for (int i = 0; i < n_m; ++i) {
        const int m = ms_host[i], am = m < 0 ? -m : m;
        HERMES_REQUIRE(m >= -l && m <= l, "invalid multipole (l=%d, m=%d)", l, m);
        M.l[i] = l;
        M.m[i] = m;
        M.cnorm[i] = (float)hermes_ylm_cnorm(l, am);
        M.sx[i] = (am & 1) ? -1.0f : 1.0f;
        M.sz[i] = ((l + am) & 1) ? -1.0f : 1.0f;
    }
    const float2* twg = hermes_twiddles(ctx, J, 1);
    HERMES_REQUIRE(twg, "%s", hermes_last_error());
    const unsigned nblk = (unsigned)(((1 << J) / 2 + 1) * ((1 << J) / 2 + 1));
    const float2* h = (const float2*)H;
    float2* o = (float2*)out;
#define HERMES_MPW(LG)                                                                                         \
    if (prof_dev && n_m == 1)                                                                                  \
        k_mpwin_zfft_multi<LG, 1, false, true><<<nblk, mpwin_threads<LG>(), 0, ctx->stream>>>(h, Pf, A, M, twg, o, \
                                                                                              prof_dev);       \
    else if (prof_dev)                                                                                         \
        k_mpwin_zfft_multi<LG, 2, false, true><<<nblk, mpwin_threads<LG>(), 0, ctx->stream>>>(h, Pf, A, M, twg, o, \
                                                                                              prof_dev);       \
    else if (n_m == 1)                                                                                         \
        k_mpwin_zfft_multi<LG, 1><<<nblk, mpwin_threads<LG>(), 0, ctx->stream>>>(h, Pf, A, M, twg, o, nullptr);   \
    else                                                                                                       \
        k_mpwin_zfft_multi<LG, 2><<<nblk, mpwin_threads<LG>(), 0, ctx->stream>>>(h, Pf, A, M, twg, o, nullptr);
    switch (J) {
        case 3: HERMES_MPW(3) break;
        case 4: HERMES_MPW(4) break;
        case 5: HERMES_MPW(5) break;
        case 6: HERMES_MPW(6) break;
        case 7: HERMES_MPW(7) break;
        case 8: HERMES_MPW(8) break;
        case 9: HERMES_MPW(9) break;
        default:
            // L = 1024: two orders per launch would not fit the registers; one launch per order
            for (int i = 0; i < n_m; ++i) {
                MultiM Mi{};
                Mi.n = 1;
                Mi.l[0] = M.l[i];
                Mi.m[0] = M.m[i];
                Mi.cnorm[0] = M.cnorm[i];
                Mi.sx[0] = M.sx[i];
                Mi.sz[0] = M.sz[i];
                if (prof_dev)
                    k_mpwin_zfft_multi<10, 1, false, true><<<nblk, mpwin_threads<10>(), 0, ctx->stream>>>(
                        h, Pf, A, Mi, twg, o + (size_t)i * n_shell * ((size_t)1 << 30), prof_dev);
                else
                    k_mpwin_zfft_multi<10, 1><<<nblk, mpwin_threads<10>(), 0, ctx->stream>>>(
                        h, Pf, A, Mi, twg, o + (size_t)i * n_shell * ((size_t)1 << 30), nullptr);
                HERMES_LAUNCH_CHECK();
            }
            break;
    }
#undef HERMES_MPW
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_multipole_window_zpass_pack0_f32(hermes_ctx* ctx, int J, const void* H, const double* P,
                                            const double* radii, int n_shell, int l, double scale, void* out) {
    HERMES_REQUIRE(ctx && H && P && radii && out, "hermes_multipole_window_zpass_pack0_f32: null argument");
    HERMES_REQUIRE(J >= 3 && J <= 9, "hermes_multipole_window_zpass_pack0_f32: J=%d outside 3..9", J);
    HERMES_REQUIRE(l >= 0 && l < 40, "hermes_multipole_window_zpass_pack0_f32: l=%d", l);
    ApplyArgs A;
    float* Pf = nullptr;
    int rc = make_apply_args(ctx, J, P, radii, n_shell, l, 0, scale, &A, &Pf);
    if (rc) return rc;
    MultiM M{};
    M.n = 2;
    for (int i = 0; i < 2; ++i) {
        M.l[i] = l + i;
        M.m[i] = 0;
        M.cnorm[i] = (float)hermes_ylm_cnorm(l + i, 0);
        M.sx[i] = 1.0f;
        M.sz[i] = ((l + i) & 1) ? -1.0f : 1.0f;
    }
    const float2* twg = hermes_twiddles(ctx, J, 1);
    HERMES_REQUIRE(twg, "%s", hermes_last_error());
    const unsigned nblk = (unsigned)(((1 << J) / 2 + 1) * ((1 << J) / 2 + 1));
    const float2* h = (const float2*)H;
    float2* o = (float2*)out;
#define HERMES_MPWP(LG) \
    k_mpwin_zfft_multi<LG, 2, true><<<nblk, mpwin_threads<LG>(), 0, ctx->stream>>>(h, Pf, A, M, twg, o, nullptr);
    switch (J) {
        case 3: HERMES_MPWP(3) break;
        case 4: HERMES_MPWP(4) break;
        case 5: HERMES_MPWP(5) break;
        case 6: HERMES_MPWP(6) break;
        case 7: HERMES_MPWP(7) break;
        case 8: HERMES_MPWP(8) break;
        default: HERMES_MPWP(9) break;
    }
#undef HERMES_MPWP
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_multipole_window_ifft_pack0_f32(hermes_ctx* ctx, int J, const void* H, const double* P,
                                           const double* radii, int n_shell, int l, double scale, void* out) {
    const int rc = hermes_multipole_window_zpass_pack0_f32(ctx, J, H, P, radii, n_shell, l, scale, out);
    if (rc) return rc;
    return hermes_fft3_inverse_yx(ctx, J, n_shell, (float2*)out);
}

int hermes_multipole_window_ifft_multi_f32(hermes_ctx* ctx, int J, const void* H, const double* P,
                                           const double* radii, int n_shell, int l, const int* ms_host, int n_m,
                                           double scale, void* out) {
    const int rc = hermes_multipole_window_zpass_multi_f32(ctx, J, H, P, radii, n_shell, l, ms_host, n_m, scale, out);
    if (rc) return rc;
    return hermes_fft3_inverse_yx(ctx, J, n_m * n_shell, (float2*)out);
}

int hermes_reduce_sum(hermes_ctx* ctx, const void* x, int64_t n, int ncomp, int dbl, double* out_host) {
    HERMES_REQUIRE(ctx && x && out_host, "hermes_reduce_sum: null argument");
    HERMES_REQUIRE(ncomp == 1 || ncomp == 2, "ncomp must be 1 or 2");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int nb = (int)hermes_blocks(n, RED_THREADS, 1024);
    double* ws = (double*)hermes_workspace(ctx, WS_MISC, sizeof(double) * (2 * nb + 2));
    HERMES_REQUIRE(ws, "%s", hermes_last_error());
    if (dbl) k_reduce_sum<double><<<nb, RED_THREADS, 0, ctx->stream>>>((const double*)x, n, ncomp, ws);
    else k_reduce_sum<float><<<nb, RED_THREADS, 0, ctx->stream>>>((const float*)x, n, ncomp, ws);
    HERMES_LAUNCH_CHECK();
    k_reduce_final<<<1, 64, 0, ctx->stream>>>(ws, nb, ws + 2 * nb);
    HERMES_LAUNCH_CHECK();
    double h[2];
    HERMES_CHECK_HIP(hipMemcpyAsync(h, ws + 2 * nb, sizeof(double) * 2, hipMemcpyDeviceToHost, ctx->stream));
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    out_host[0] = h[0];
    if (ncomp == 2) out_host[1] = h[1];
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free pair reduction
// hermes_pair_reduce_noalloc runs the kernels of hermes_pair_reduce with their scratch (window tables,
// per-block partial sums, group sums, the result row) in this fixed device array instead of the context's
// grow-only workspace, so it never allocates device memory.  Each device holds its own copy until the
// process ends; its address on a device comes from hipGetSymbolAddress, called once per device when the
// first context on it is created (hermes_ctx_create).  Host code must use that address, never the array's
// name.  Every call on a device shares the array, so calls that overlap on one device race; the kernels
// write every scratch word they read, so nothing needs clearing between calls.
constexpr size_t PAIR_SCRATCH_BYTES = size_t(8) << 20;
static __device__ __align__(256) char scratch[PAIR_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static char* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {nullptr};
static std::mutex g_scratch_mu;

int hermes_scratch_setup(int device) {
    HERMES_REQUIRE(device >= 0 && device < SCRATCH_MAX_DEVICES, "device %d beyond the scratch table", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        hermes_set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorString(e));
        return 1;
    }
    g_scratch_ptr[device] = (char*)ptr;
    return 0;
}

int hermes_pair_reduce_noalloc(hermes_ctx* ctx, int J, int dbl, const void* F1, const void* F2, const double* P,
                               const hermes_window_desc* descs, int n_desc, double* out_host) {
    HERMES_REQUIRE(ctx, "hermes_pair_reduce_noalloc: null context");
    HERMES_REQUIRE(ctx->device >= 0 && ctx->device < SCRATCH_MAX_DEVICES && g_scratch_ptr[ctx->device],
                   "hermes_pair_reduce_noalloc: no scratch address for device %d", ctx->device);
    return pair_reduce_impl(ctx, J, dbl, F1, F2, P, descs, n_desc, out_host, g_scratch_ptr[ctx->device],
                            PAIR_SCRATCH_BYTES);
}
