// Device special functions and window transfer functions.
//
// *_f64: faithful FP64 restatements of the reference algorithms
//   spherical_jn_numba      special_functions.py:49-128
//   assoc_legendre_numba    special_functions.py:273-304
//   spherical_harmonic_numba special_functions.py:307-334
//   window transfer kernels window_functions.py:29-129, :426-438
// *_f32: the FP32 fast path used by the batched kernels (no acos/atan2: Y_lm is
//   evaluated from ((kx + i ky)/|k|)^m and the Legendre polynomial part, j_l from
//   an exact revolution count and upward recurrence / power series).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/hermes_amd.h"

#define HERMES_PI 3.14159265358979323846

// ------------------------------------------------------------------ FP64 faithful
// Cylindrical Bessel J0(x) and J1(x), x >= 0, to FP64 rounding (the reference's jn_numba,
// special_functions.py:131-270, uses 8-digit rational approximations instead):
//   x <= 4       the power series sum_k (-x^2/4)^k / (k! (k+nu)!) (x/2)^nu (largest term 4, so little
//                cancellation);
//   4 < x < 25   Miller's backward recurrence J_{n-1} = (2n/x) J_n - J_{n+1} from n = x + 10 sqrt(x) + 30,
//                normalised by J0 + 2 (J2 + J4 + ...) = 1 (values stay below 1e60 on the way down);
//   x >= 25      Hankel's asymptotic expansion J_nu = sqrt(2/(pi x)) (P cos w - Q sin w),
//                w = x - (2 nu + 1) pi/4, with a_k = a_{k-1} (4 nu^2 - (2k-1)^2) / (8k) and
//                P = sum (-1)^k a_{2k} / x^{2k}, Q = sum (-1)^k a_{2k+1} / x^{2k+1}, summed while the
//                terms fall (their smallest is about e^{-2x}).
__host__ __device__ inline void bessel_j01_f64(double x, double& j0, double& j1) {
    if (x <= 4.0) {
        const double h = -0.25 * x * x;
        double t0 = 1.0, t1 = 0.5 * x, s0 = 1.0, s1 = t1;
        for (int k = 1; k <= 24; ++k) {
            t0 *= h / ((double)k * k);
            t1 *= h / ((double)k * (k + 1));
            s0 += t0;
            s1 += t1;
        }
        j0 = s0;
        j1 = s1;
        return;
    }
    if (x < 25.0) {
        const int n0 = 2 * (((int)(x + 10.0 * sqrt(x)) + 30) / 2);  // even start
        const double tox = 2.0 / x;
        double jp = 0.0, jc = 1e-30, even_sum = 0.0, a0 = 0.0, a1 = 0.0;
        for (int n = n0; n >= 1; --n) {
            const double jm = n * tox * jc - jp;  // J_{n-1}
            jp = jc;
            jc = jm;
            if (n == 1) {
                a0 = jc;
                a1 = jp;
            } else if (((n - 1) & 1) == 0) {
                even_sum += jc;  // J_{n-1} with n - 1 even and >= 2
            }
        }
        const double norm = a0 + 2.0 * even_sum;
        j0 = a0 / norm;
        j1 = a1 / norm;
        return;
    }
    double out[2];
    for (int nu = 0; nu < 2; ++nu) {
        const double mu4 = 4.0 * nu * nu;
        double a = 1.0, P = 1.0, Q = 0.0, prev = 1e300, xp = 1.0;
        for (int k = 1; k <= 60; ++k) {
            a *= (mu4 - (2.0 * k - 1.0) * (2.0 * k - 1.0)) / (8.0 * k);
            xp *= x;
            const double term = a / xp;
            if (fabs(term) >= prev || fabs(term) < 1e-17) break;
            prev = fabs(term);
            // k odd -> Q, k even -> P; sign (-1)^{floor(k/2)}
            const double sg = ((k / 2) & 1) ? -1.0 : 1.0;
            if (k & 1) Q += sg * term;
            else P += sg * term;
        }
        const double w = x - (2.0 * nu + 1.0) * 0.25 * HERMES_PI;
        out[nu] = sqrt(2.0 / (HERMES_PI * x)) * (P * cos(w) - Q * sin(w));
    }
    j0 = out[0];
    j1 = out[1];
}

__device__ inline double jl_f64(int l, double x) {
    if (l < 0) return __builtin_nan("");
    if (x == 0.0) return l == 0 ? 1.0 : 0.0;
    double ax = fabs(x);
    if (ax <= 5.0e-2) {
        if (l == 0) {
            double x2 = x * x;
            return 1.0 - x2 / 6.0 + x2 * x2 / 120.0;
        }
        double denom = 1.0;
        for (int n = 1; n <= l; ++n) denom *= 2.0 * n + 1.0;
        double x2 = x * x;
        return (pow(x, (double)l) / denom) *
               (1.0 - x2 / (2.0 * (2.0 * l + 3.0)) + x2 * x2 / (8.0 * (2.0 * l + 3.0) * (2.0 * l + 5.0)));
    }
    double sx, cx;
    sincos(x, &sx, &cx);
    double j0 = sx / x;
    if (l == 0) return j0;
    double j1 = sx / (x * x) - cx / x;
    if (l == 1) return j1;
    if (ax < l + 1.0) {
        int n_stop = l + 40;
        double jp1 = 0.0, jcur = 1.0, jt = 0.0, jd0 = 0.0, jd1 = 0.0;
        for (int n = n_stop; n > 0; --n) {
            double jm1 = ((2.0 * n + 1.0) / x) * jcur - jp1;
            if (n - 1 == l) jt = jm1;
            if (n - 1 == 1) jd1 = jm1;
            if (n - 1 == 0) jd0 = jm1;
            jp1 = jcur;
            jcur = jm1;
            if (fabs(jcur) > 1.0e200) {
                jp1 *= 1.0e-200; jcur *= 1.0e-200; jt *= 1.0e-200; jd0 *= 1.0e-200; jd1 *= 1.0e-200;
            } else if (fabs(jcur) < 1.0e-200 && jcur != 0.0) {
                jp1 *= 1.0e200; jcur *= 1.0e200; jt *= 1.0e200; jd0 *= 1.0e200; jd1 *= 1.0e200;
            }
        }
        if (fabs(jd0) >= fabs(jd1) || jd1 == 0.0) return jt * (j0 / jd0);
        return jt * (j1 / jd1);
    }
    double jm1 = j0, jcur = j1;
    for (int ell = 1; ell < l; ++ell) {
        double jn = ((2.0 * ell + 1.0) / x) * jcur - jm1;
        jm1 = jcur;
        jcur = jn;
    }
    return jcur;
}

__device__ inline double plm_f64(int l, int m, double x) {
    double pmm = 1.0;
    if (m > 0) {
        double t = 1.0 - x * x;
        double somx2 = sqrt(t > 0.0 ? t : 0.0);
        double fact = 1.0;
        for (int i = 0; i < m; ++i) {
            pmm *= -fact * somx2;
            fact += 2.0;
        }
    }
    if (l == m) return pmm;
    double pmmp1 = x * (2.0 * m + 1.0) * pmm;
    if (l == m + 1) return pmmp1;
    double a = pmm, b = pmmp1, pll = 0.0;
    for (int ell = m + 2; ell <= l; ++ell) {
        pll = ((2.0 * ell - 1.0) * x * b - (ell + m - 1.0) * a) / (ell - m);
        a = b;
        b = pll;
    }
    return pll;
}

__device__ inline double factorial_small_f64(int n) {
    double out = 1.0;
    for (int i = 2; i <= n; ++i) out *= i;
    return out;
}

__device__ inline double2 ylm_f64(int l, int m, double ki, double kj, double kk) {
    if (l < 0) return make_double2(__builtin_nan(""), 0.0);
    int am = m < 0 ? -m : m;
    if (am > l) return make_double2(0.0, 0.0);
    double k = sqrt(ki * ki + kj * kj + kk * kk);
    if (k == 0.0) {
        if (l == 0 && m == 0) return make_double2(1.0 / sqrt(4.0 * HERMES_PI), 0.0);
        return make_double2(0.0, 0.0);
    }
    double mu = kk / k;
    if (mu > 1.0) mu = 1.0;
    else if (mu < -1.0) mu = -1.0;
    double theta = acos(mu);
    double ph = atan2(kj, ki);
    double x = cos(theta);
    double P = plm_f64(l, am, x);
    double norm = sqrt((2.0 * l + 1.0) / (4.0 * HERMES_PI) * factorial_small_f64(l - am) / factorial_small_f64(l + am));
    double s = norm * P;
    double re = s * cos(am * ph), im = s * sin(am * ph);
    if (m >= 0) return make_double2(re, im);
    double sign = (am % 2 == 1) ? -1.0 : 1.0;
    return make_double2(sign * re, -sign * im);
}

__device__ inline double win_sphere_f64(double k, double R) {
    double q = 2 * HERMES_PI * k * R;
    if (q == 0.0) return 1.0;
    double s, c;
    sincos(q, &s, &c);
    return 3 * (s - q * c) / (q * q * q);
}

// Line-of-sight windows (window_functions.py:164-319).  Along a coordinate axis they are even in every
// component of k, so the octant folds apply to them as to the isotropic ones; along any other direction
// they are only even under k -> -k (the half-spectrum sums still apply, the octant sharing does not).
// axis 3: the unit vector n (k_par = k . n, k_perp from |k|^2 - k_par^2, window_functions.py:164-319).
__device__ inline double los_window_f64(int type, double ki, double kj, double kk, double R, double H, int axis,
                                        const double* n = nullptr) {
    double kpar, kperp2;
    if (n != nullptr && axis == 3) {
        kpar = ki * n[0] + kj * n[1] + kk * n[2];
        kperp2 = fmax(ki * ki + kj * kj + kk * kk - kpar * kpar, 0.0);
    } else {
        kpar = axis == 0 ? ki : (axis == 1 ? kj : kk);
        const double u = axis == 0 ? kj : ki, v = axis == 2 ? kj : kk;  // the two perpendicular components
        kperp2 = u * u + v * v;
    }
    const double qperp = 2.0 * HERMES_PI * sqrt(kperp2) * R;
    const double qpar = 2.0 * HERMES_PI * kpar * H;
    double j0, j1;
    bessel_j01_f64(qperp, j0, j1);
    const double perp = (type == HERMES_WIN_RING || type == HERMES_WIN_CYLSHELL)
                            ? j0
                            : (qperp == 0.0 ? 1.0 : 2.0 * j1 / qperp);
    const double par = (type == HERMES_WIN_RING || type == HERMES_WIN_DISK)
                           ? cos(qpar)
                           : (qpar == 0.0 ? 1.0 : sin(qpar) / qpar);
    return perp * par;
}

// Transfer function of one real, even window at k (cycles per grid cell): the isotropic ones and the
// line-of-sight ones.
__device__ inline double iso_window_f64(int type, double ki, double kj, double kk, double a0, double a1,
                                        int axis = 2, const double* los = nullptr) {
    if (type >= HERMES_WIN_RING && type <= HERMES_WIN_CYLINDER)
        return los_window_f64(type, ki, kj, kk, a0, a1, axis, los);
    switch (type) {
    case HERMES_WIN_SPHERE: return win_sphere_f64(sqrt(ki * ki + kj * kj + kk * kk), a0);
    case HERMES_WIN_GAUSSIAN: {
        double k = sqrt(ki * ki + kj * kj + kk * kk);
        double q = 2 * HERMES_PI * k * a0;
        return exp(-(q * q) / 2);
    }
    case HERMES_WIN_SHELL: {
        double k = sqrt(ki * ki + kj * kj + kk * kk);
        double q = 2 * HERMES_PI * k * a0;
        if (q == 0.0) return 1.0;
        return sin(q) / q;
    }
    case HERMES_WIN_THICK_SHELL: {
        double k = sqrt(ki * ki + kj * kj + kk * kk);
        double ri = a0 - 0.5 * a1;
        if (ri < 0.0) ri = 0.0;
        double ro = a0 + 0.5 * a1;
        double den = ro * ro * ro - ri * ri * ri;
        if (den == 0.0) return 1.0;
        return (ro * ro * ro * win_sphere_f64(k, ro) - ri * ri * ri * win_sphere_f64(k, ri)) / den;
    }
    case HERMES_WIN_GAUSSIAN_SHELL: {
        double k = sqrt(ki * ki + kj * kj + kk * kk);
        if (k == 0) return 1.0;
        double qs = 2 * HERMES_PI * k * a0;
        double qm = 2 * HERMES_PI * k * a1;
        double den = a0 * a0 + a1 * a1;
        if (den == 0.0) return 1.0;
        double sinc = (qs == 0.0) ? 1.0 : sin(qs) / qs;
        return (a1 * a1 * cos(qs) + a0 * a0 * sinc) / den * exp(-(qm * qm) / 2);
    }
    case HERMES_WIN_INVERSE_LAPLACIAN: {
        double k2 = ki * ki + kj * kj + kk * kk;
        if (k2 == 0.0) return 0.0;
        return -1.0 / (((2.0 * HERMES_PI) * (2.0 * HERMES_PI)) * k2);
    }
    }
    return __builtin_nan("");
}

// U_l(k) of a radial profile at k cycles per grid cell (hermes_radial_profile; FP64).
__device__ inline double radial_table_f64(const double* __restrict__ tab, int n, double x) {
    // x = k * k_scale, the fractional table index; radial_multipole_windows.py:186-210 lookup
    if (x <= 0.0) return tab[0];
    if (x >= (double)(n - 1)) return tab[n - 1];
    const int i = (int)x;
    const double t = x - i;
    if (i == 0 || i >= n - 2) return tab[i] * (1.0 - t) + tab[i + 1] * t;
    const double p0 = tab[i - 1], p1 = tab[i], p2 = tab[i + 1], p3 = tab[i + 2];
    return p1 + 0.5 * t * ((p2 - p0) + t * ((2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3) + t * (3.0 * (p1 - p2) + p3 - p0)));
}

__device__ inline double radial_profile_f64(const hermes_radial_profile& p, int l, double k) {
    switch (p.kind) {
    case HERMES_RADIAL_SHELL: return jl_f64(l, 2.0 * HERMES_PI * k * p.a0);
    case HERMES_RADIAL_TABLE: return radial_table_f64(p.table, p.n_table, k * p.k_scale);
    case HERMES_RADIAL_SPHERE: return iso_window_f64(HERMES_WIN_SPHERE, k, 0.0, 0.0, p.a0, p.a1);
    case HERMES_RADIAL_GAUSSIAN: return iso_window_f64(HERMES_WIN_GAUSSIAN, k, 0.0, 0.0, p.a0, p.a1);
    case HERMES_RADIAL_THICK_SHELL: return iso_window_f64(HERMES_WIN_THICK_SHELL, k, 0.0, 0.0, p.a0, p.a1);
    case HERMES_RADIAL_GAUSSIAN_SHELL: return iso_window_f64(HERMES_WIN_GAUSSIAN_SHELL, k, 0.0, 0.0, p.a0, p.a1);
    }
    return __builtin_nan("");
}

// Complex transfer function of a window at signed k (isotropic windows are real).  general = false
// compiles out the non-axis line-of-sight branch (for callers that never see axis-3 descriptors).
__device__ inline double2 window_value_f64(const hermes_window_desc& d, double ki, double kj, double kk,
                                           bool general = true) {
    if (d.type == HERMES_WIN_SHELL_MULTIPOLE) {
        double k = sqrt(ki * ki + kj * kj + kk * kk);
        double radial = jl_f64(d.l, 2.0 * HERMES_PI * k * d.a0);
        double2 y = ylm_f64(d.l, d.m, ki, kj, kk);
        return make_double2(radial * y.x, radial * y.y);
    }
    return make_double2(iso_window_f64(d.type, ki, kj, kk, d.a0, d.a1, d.axis, general ? d.los : nullptr), 0.0);
}

// ------------------------------------------------------------------ FP32 fast path
// j_l(2 pi t), FP32 fast path, from the exact revolution count t (FP64):
// l = 0 closed form; q >= l + 1 upward recurrence from j0, j1 (stable there,
// the same split as the reference); otherwise the power series
// j_l(q) = q^l/(2l+1)!! sum_k (-q^2/2)^k / (k! prod_{i=1..k}(2l+2i+1)).
__device__ inline float jl_f32(int l, double t) {
    if (t == 0.0) return l == 0 ? 1.0f : 0.0f;
    float q = (float)(2.0 * HERMES_PI * t);
    if (l == 0 || q >= (float)l + 1.0f) {
        // the argument is t turns: reduce in FP64, then the hardware sin/cos, which take turns
        // (sin(2 pi x); measured max abs error 1.1e-7 over a turn)
        const float fr = (float)(t - floor(t));
        const float s = __builtin_amdgcn_sinf(fr), c = __builtin_amdgcn_cosf(fr);
        const float iq = __builtin_amdgcn_rcpf(q);
        float j0 = s * iq;
        if (l == 0) return j0;
        float j1 = (j0 - c) * iq;
        float jm = j0, jc = j1;
        for (int n = 1; n < l; ++n) {
            float jn = (float)(2 * n + 1) * iq * jc - jm;
            jm = jc;
            jc = jn;
        }
        return jc;
    }
    float h = -0.5f * q * q;
    float ql = 1.0f, dfact = 1.0f;
    for (int n = 1; n <= l; ++n) {
        ql *= q;
        dfact *= (float)(2 * n + 1);
    }
    float term = 1.0f, sum = 1.0f;
    for (int k = 1; k < 80; ++k) {
        term *= h * __builtin_amdgcn_rcpf((float)(k * (2 * l + 2 * k + 1)));
        sum += term;
        if (fabsf(term) < 1e-9f * fabsf(sum)) break;
    }
    return ql / dfact * sum;
}

// j_l and j_{l+1} of 2 pi t together: the upward recurrence of jl_f32 run one step further where
// it serves both (q >= l + 2), otherwise each from jl_f32; the same values as two jl_f32 calls.
__device__ inline void jl_pair_f32(int l, double t, float& jl, float& jl1) {
    const float q = (float)(2.0 * HERMES_PI * t);
    if (t != 0.0 && q >= (float)l + 2.0f) {
        const float fr = (float)(t - floor(t));
        const float s = __builtin_amdgcn_sinf(fr), c = __builtin_amdgcn_cosf(fr);
        const float iq = __builtin_amdgcn_rcpf(q);
        float jm = s * iq;
        float jc = (jm - c) * iq;
        if (l == 0) {
            jl = jm;
            jl1 = jc;
            return;
        }
        for (int n = 1; n < l; ++n) {
            float jn = (float)(2 * n + 1) * iq * jc - jm;
            jm = jc;
            jc = jn;
        }
        jl = jc;
        jl1 = (float)(2 * l + 1) * iq * jc - jm;
        return;
    }
    jl = jl_f32(l, t);
    jl1 = jl_f32(l + 1, t);
}

// Y_l^m(a-hat) for integer lattice vector a (as floats), m >= 0; cnorm = N_lm (-1)^m (2m-1)!!.
__device__ inline float2 ylm_f32(int l, int m, float ax, float ay, float az, float cnorm) {
    float r2 = ax * ax + ay * ay + az * az;
    if (r2 == 0.0f) return make_float2((l == 0 && m == 0) ? 0.28209479177387814f : 0.0f, 0.0f);
    float ir = rsqrtf(r2);
    float x = az * ir;
    float ur = ax * ir, ui = ay * ir;  // sin(theta) e^{i phi}
    float pr = 1.0f, pi = 0.0f;
    for (int i = 0; i < m; ++i) {
        float t = pr * ur - pi * ui;
        pi = pr * ui + pi * ur;
        pr = t;
    }
    // polynomial part of P_l^m(x) / ((-1)^m (2m-1)!! (1-x^2)^{m/2})
    float pmm = 1.0f, poly;
    if (l == m) {
        poly = pmm;
    } else {
        float a = pmm, b = x * (float)(2 * m + 1);
        for (int ell = m + 2; ell <= l; ++ell) {
            float c = ((float)(2 * ell - 1) * x * b - (float)(ell + m - 1) * a) / (float)(ell - m);
            a = b;
            b = c;
        }
        poly = b;
    }
    float s = cnorm * poly;
    return make_float2(s * pr, s * pi);
}
