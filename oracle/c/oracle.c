/*
 * CPU oracle (C/OpenMP restatement of the PyHermes Numba kernels).
 * TEST INFRASTRUCTURE ONLY: used by tests/ as the checker and by bench.py's
 * cpu_baseline leg as the timed CPU baseline.  Never linked into the product.
 *
 * Each function names the reference file:line it restates (paths relative to
 * the PyHermes source tree).  Build: oracle/Makefile -> oracle/_build/liboracle.so
 * with -ffp-contract=off so products and sums round like the Numba kernels.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

static inline cplx c_make(double re, double im) { cplx c = {re, im}; return c; }
static inline cplx c_add(cplx a, cplx b) { return c_make(a.re + b.re, a.im + b.im); }
static inline cplx c_mul(cplx a, cplx b) {
    return c_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cplx c_scale(double s, cplx a) { return c_make(s * a.re, s * a.im); }

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------------------
 * Projection: wavelet_grid.py:17-70 (_project_scaling_grid_impl, full grid).
 * Serial loop over particles, 27 FP64 adds per particle.  Python negative
 * indexing wraps cells -1..-L; the oracle wraps every index modulo L so that
 * in-range inputs behave exactly like the reference (cell == L is undefined
 * behaviour there).  Optionally returns per-particle cell and fraction indices.
 * ------------------------------------------------------------------------- */
void oracle_project(const double *pos, const double *w, int64_t n,
                    const double *phi, int phi_len, int phi_res,
                    int J, double box, double *eps_out,
                    int32_t *cell_out, int32_t *frac_out) {
    const int64_t L = (int64_t)1 << J;
    const int support = phi_len / phi_res;
    const double sf = (double)L / box;
    memset(eps_out, 0, sizeof(double) * L * L * L);
    for (int64_t p = 0; p < n; ++p) {
        int32_t c[3], f[3];
        for (int d = 0; d < 3; ++d) {
            double s = pos[3 * p + d] * sf;
            c[d] = (int32_t)floor(s);
            f[d] = (int32_t)((s - (double)c[d]) * (double)phi_res);
        }
        if (cell_out) { cell_out[3 * p] = c[0]; cell_out[3 * p + 1] = c[1]; cell_out[3 * p + 2] = c[2]; }
        if (frac_out) { frac_out[3 * p] = f[0]; frac_out[3 * p + 1] = f[1]; frac_out[3 * p + 2] = f[2]; }
        double wt = w[p];
        for (int sx = 0; sx < support; ++sx) {
            double phx = wt * phi[f[0] + phi_res * sx];
            int64_t ix = (((int64_t)c[0] - sx) % L + L) % L;
            for (int sy = 0; sy < support; ++sy) {
                double phxy = phx * phi[f[1] + phi_res * sy];
                int64_t iy = (((int64_t)c[1] - sy) % L + L) % L;
                for (int sz = 0; sz < support; ++sz) {
                    int64_t iz = (((int64_t)c[2] - sz) % L + L) % L;
                    eps_out[(ix * L + iy) * L + iz] += phxy * phi[f[2] + phi_res * sz];
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * Isotropic window transfer functions: window_functions.py:29-129, :426-438.
 * code: 0 sphere(R) 1 gaussian(R) 2 shell(R) 3 thick_shell(R, dR)
 *       4 gaussian_shell(R_shell, R_smooth) 5 inverse_laplacian()
 * ------------------------------------------------------------------------- */
static double win_sphere(double ki, double kj, double kk, double R) {
    double k = sqrt(ki * ki + kj * kj + kk * kk);
    double q = 2 * M_PI * k * R;
    if (q == 0.0) return 1.0;
    return 3 * (sin(q) - q * cos(q)) / (q * q * q);
}

double oracle_window_eval(int code, double ki, double kj, double kk, double a0, double a1) {
    switch (code) {
    case 0: return win_sphere(ki, kj, kk, a0);
    case 1: {
        double k = sqrt(ki * ki + kj * kj + kk * kk);
        double q = 2 * M_PI * k * a0;
        return exp(-(q * q) / 2);
    }
    case 2: {
        double k = sqrt(ki * ki + kj * kj + kk * kk);
        double q = 2 * M_PI * k * a0;
        if (q == 0.0) return 1.0;
        return sin(q) / q;
    }
    case 3: {
        double ri = a0 - 0.5 * a1;
        if (ri < 0.0) ri = 0.0;
        double ro = a0 + 0.5 * a1;
        double den = ro * ro * ro - ri * ri * ri;
        if (den == 0.0) return 1.0;
        return (ro * ro * ro * win_sphere(ki, kj, kk, ro) - ri * ri * ri * win_sphere(ki, kj, kk, ri)) / den;
    }
    case 4: {
        double k = sqrt(ki * ki + kj * kj + kk * kk);
        if (k == 0) return 1.0;
        double qs = 2 * M_PI * k * a0;
        double qm = 2 * M_PI * k * a1;
        double den = a0 * a0 + a1 * a1;
        if (den == 0.0) return 1.0;
        double sinc = (qs == 0.0) ? 1.0 : sin(qs) / qs;
        return (a1 * a1 * cos(qs) + a0 * a0 * sinc) / den * exp(-(qm * qm) / 2);
    }
    case 5: {
        double k2 = ki * ki + kj * kj + kk * kk;
        if (k2 == 0.0) return 0.0;
        return -1.0 / (((2.0 * M_PI) * (2.0 * M_PI)) * k2);
    }
    }
    return NAN;
}

/* convolution.py:14-38 (bandwidth 1) then convolution.py:263-281 fold. */
void oracle_octant_rfft_kernel(int L, const double *P, int code, double a0, double a1,
                               double *out /* (L, L, L/2+1) */) {
    const int L1 = L + 1;
    double *wa = (double *)malloc(sizeof(double) * (size_t)L1 * L1 * L1);
    const double inv_L = 1.0 / L;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < L1; ++i)
        for (int j = 0; j < L1; ++j)
            for (int k = 0; k < L1; ++k) {
                double t = 0.0;
                t += P[i] * P[j] * P[k] * oracle_window_eval(code, i * inv_L, j * inv_L, k * inv_L, a0, a1);
                wa[((size_t)i * L1 + j) * L1 + k] = t;
            }
    const int Lh = L / 2 + 1;
#define WA(a, b, c) wa[((size_t)(a) * L1 + (b)) * L1 + (c)]
#pragma omp parallel for schedule(static)
    for (int x = 0; x < L; ++x)
        for (int y = 0; y < L; ++y)
            for (int z = 0; z < Lh; ++z)
                out[((size_t)x * L + y) * Lh + z] =
                    WA(x, y, z) + WA(L - x, y, z) + WA(x, L - y, z) + WA(x, y, L - z) +
                    WA(L - x, L - y, z) + WA(L - x, y, L - z) + WA(x, L - y, L - z) +
                    WA(L - x, L - y, L - z);
#undef WA
    free(wa);
}

/* convolution.py:73-157, bandwidth 1: the 8 explicit aliases, no octant symmetry. */
void oracle_full_rfft_kernel(int L, const double *P, int code, double a0, double a1, double *out) {
    const double inv_L = 1.0 / L;
    const int Lh = L / 2 + 1;
#pragma omp parallel for schedule(static)
    for (int x = 0; x < L; ++x) {
        int xm = L - x;
        for (int y = 0; y < L; ++y) {
            int ym = L - y;
            for (int z = 0; z < Lh; ++z) {
                int zm = L - z;
                double xp = x * inv_L, xn = -xm * inv_L;
                double yp = y * inv_L, yn = -ym * inv_L;
                double zp = z * inv_L, zn = -zm * inv_L;
                double t = 0.0;
                t += P[x] * P[y] * P[z] * oracle_window_eval(code, xp, yp, zp, a0, a1);
                t += P[xm] * P[y] * P[z] * oracle_window_eval(code, xn, yp, zp, a0, a1);
                t += P[x] * P[ym] * P[z] * oracle_window_eval(code, xp, yn, zp, a0, a1);
                t += P[x] * P[y] * P[zm] * oracle_window_eval(code, xp, yp, zn, a0, a1);
                t += P[xm] * P[ym] * P[z] * oracle_window_eval(code, xn, yn, zp, a0, a1);
                t += P[xm] * P[y] * P[zm] * oracle_window_eval(code, xn, yp, zn, a0, a1);
                t += P[x] * P[ym] * P[zm] * oracle_window_eval(code, xp, yn, zn, a0, a1);
                t += P[xm] * P[ym] * P[zm] * oracle_window_eval(code, xn, yn, zn, a0, a1);
                out[((size_t)x * L + y) * Lh + z] = t;
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * Special functions: special_functions.py:49-128 (spherical_jn_numba),
 * :273-304 (assoc_legendre_numba), :307-334 (spherical_harmonic_numba).
 * ------------------------------------------------------------------------- */
double oracle_spherical_jn(int l, double x) {
    if (l < 0) return NAN;
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
        return (pow(x, l) / denom) *
               (1.0 - x2 / (2.0 * (2.0 * l + 3.0)) + x2 * x2 / (8.0 * (2.0 * l + 3.0) * (2.0 * l + 5.0)));
    }
    double sx = sin(x), cx = cos(x);
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

double oracle_assoc_legendre(int l, int m, double x) {
    if (l < 0 || m < 0 || m > l) return NAN;
    double pmm = 1.0;
    if (m > 0) {
        double t = 1.0 - x * x;
        double somx2 = sqrt(t > 0.0 ? t : 0.0);
        double fact = 1.0;
        for (int i = 0; i < m; ++i) { pmm *= -fact * somx2; fact += 2.0; }
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

static double factorial_small(int n) {
    double out = 1.0;
    for (int i = 2; i <= n; ++i) out *= i;
    return out;
}

cplx oracle_sph_harm_c(int l, int m, double ki, double kj, double kk) {
    if (l < 0) return c_make(NAN, 0.0);
    if (abs(m) > l) return c_make(0.0, 0.0);
    double k = sqrt(ki * ki + kj * kj + kk * kk);
    if (k == 0.0) {
        if (l == 0 && m == 0) return c_make(1.0 / sqrt(4.0 * M_PI), 0.0);
        return c_make(0.0, 0.0);
    }
    double mu = kk / k;
    if (mu > 1.0) mu = 1.0; else if (mu < -1.0) mu = -1.0;
    double theta = acos(mu);
    double ph = atan2(kj, ki);
    double x = cos(theta);
    if (m >= 0) {
        double P = oracle_assoc_legendre(l, m, x);
        double norm = sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * factorial_small(l - m) / factorial_small(l + m));
        double s = norm * P;
        return c_make(s * cos(m * ph), s * sin(m * ph));
    }
    cplx yp = oracle_sph_harm_c(l, -m, ki, kj, kk);
    double sign = ((-m) % 2 == 1) ? -1.0 : 1.0;
    return c_make(sign * yp.re, -sign * yp.im);
}

void oracle_sph_harm(int l, int m, double ki, double kj, double kk, double *re, double *im) {
    cplx y = oracle_sph_harm_c(l, m, ki, kj, kk);
    *re = y.re;
    *im = y.im;
}

/* radial_multipole_windows.py:128-144, shell profile (code 0) only. */
static cplx radial_multipole_eval(double ki, double kj, double kk, double R, int l, int m) {
    if (l < 0) return c_make(NAN, 0.0);
    if (abs(m) > l) return c_make(0.0, 0.0);
    double k = sqrt(ki * ki + kj * kj + kk * kk);
    double radial = oracle_spherical_jn(l, 2.0 * M_PI * k * R);
    cplx ang = oracle_sph_harm_c(l, m, ki, kj, kk);
    return c_scale(radial, ang);
}

/* radial_multipole_windows.py:147-183 -> (L, L, L) complex128, interleaved re/im. */
void oracle_radial_multipole_window(int L, const double *P, double R, int l, int m, double *out) {
    const double inv_L = 1.0 / L;
#pragma omp parallel for schedule(static)
    for (int x = 0; x < L; ++x) {
        double pi0 = P[L - x], pi1 = P[x];
        double ki0 = (x - L) * inv_L, ki1 = x * inv_L;
        for (int y = 0; y < L; ++y) {
            double pj0 = P[L - y], pj1 = P[y];
            double kj0 = (y - L) * inv_L, kj1 = y * inv_L;
            for (int z = 0; z < L; ++z) {
                double pk0 = P[L - z], pk1 = P[z];
                double kk0 = (z - L) * inv_L, kk1 = z * inv_L;
                cplx t = c_scale(pi0 * pj0 * pk0, radial_multipole_eval(ki0, kj0, kk0, R, l, m));
                t = c_add(t, c_scale(pi0 * pj0 * pk1, radial_multipole_eval(ki0, kj0, kk1, R, l, m)));
                t = c_add(t, c_scale(pi0 * pj1 * pk0, radial_multipole_eval(ki0, kj1, kk0, R, l, m)));
                t = c_add(t, c_scale(pi0 * pj1 * pk1, radial_multipole_eval(ki0, kj1, kk1, R, l, m)));
                t = c_add(t, c_scale(pi1 * pj0 * pk0, radial_multipole_eval(ki1, kj0, kk0, R, l, m)));
                t = c_add(t, c_scale(pi1 * pj0 * pk1, radial_multipole_eval(ki1, kj0, kk1, R, l, m)));
                t = c_add(t, c_scale(pi1 * pj1 * pk0, radial_multipole_eval(ki1, kj1, kk0, R, l, m)));
                t = c_add(t, c_scale(pi1 * pj1 * pk1, radial_multipole_eval(ki1, kj1, kk1, R, l, m)));
                size_t o = ((size_t)x * L + y) * L + z;
                out[2 * o] = t.re;
                out[2 * o + 1] = t.im;
            }
        }
    }
}

/* radial_multipole_windows.py:17-126 closed-form radial profiles (code 0 shell j_l, 1 gaussian_shell,
 * 2 thick_shell, 3 sphere, 4 gaussian; the last four are monopoles) and :186-210 table lookup. */
static double profile_sinc(double q) { return q == 0.0 ? 1.0 : sin(q) / q; }
static double profile_sphere(double k, double R) {
    double q = 2.0 * M_PI * k * R;
    if (q == 0.0) return 1.0;
    return 3.0 * (sin(q) - q * cos(q)) / (q * q * q);
}
static double radial_profile_eval(double k, int code, double a0, double a1, int l) {
    if (code == 0) return oracle_spherical_jn(l, 2.0 * M_PI * k * a0);
    if (code == 1) {
        if (k == 0.0) return 1.0;
        double qs = 2.0 * M_PI * k * a0, qm = 2.0 * M_PI * k * a1, den = a0 * a0 + a1 * a1;
        if (den == 0.0) return 1.0;
        return (a1 * a1 * cos(qs) + a0 * a0 * profile_sinc(qs)) / den * exp(-0.5 * qm * qm);
    }
    if (code == 2) {
        double ri = a0 - 0.5 * a1, ro = a0 + 0.5 * a1;
        if (ri < 0.0) ri = 0.0;
        double den = ro * ro * ro - ri * ri * ri;
        if (den == 0.0) return 1.0;
        return (ro * ro * ro * profile_sphere(k, ro) - ri * ri * ri * profile_sphere(k, ri)) / den;
    }
    if (code == 3) return profile_sphere(k, a0);
    if (code == 4) {
        double q = 2.0 * M_PI * k * a0;
        return exp(-0.5 * q * q);
    }
    return NAN;
}
static double table_lookup(double k, const double *table, int n, double k_max) {
    if (k <= 0.0) return table[0];
    if (k >= k_max) return table[n - 1];
    double index = k * (n - 1) / k_max;
    int lo = (int)index;
    double f = index - lo;
    if (lo == 0 || lo >= n - 2) return table[lo] * (1.0 - f) + table[lo + 1] * f;
    double p0 = table[lo - 1], p1 = table[lo], p2 = table[lo + 1], p3 = table[lo + 2];
    double f2 = f * f, f3 = f2 * f;
    return 0.5 * (2.0 * p1 + (-p0 + p2) * f + (2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3) * f2 +
                  (-p0 + 3.0 * p1 - 3.0 * p2 + p3) * f3);
}

/* radial_multipole_windows.py:147-183 (closed-form profile, code as above) when table == NULL, else
 * :214-264 (U_l from the table over physical k in [0, k_max], k_phys = k_grid * g2p) -> (L, L, L) complex128. */
void oracle_profile_multipole_window(int L, const double *P, int code, double a0, double a1, const double *table,
                                     int n_table, double k_max, double g2p, int l, int m, double *out) {
    const double inv_L = 1.0 / L;
#pragma omp parallel for schedule(static)
    for (int x = 0; x < L; ++x) {
        double pis[2] = {P[L - x], P[x]}, kis[2] = {(x - L) * inv_L, x * inv_L};
        for (int y = 0; y < L; ++y) {
            double pjs[2] = {P[L - y], P[y]}, kjs[2] = {(y - L) * inv_L, y * inv_L};
            for (int z = 0; z < L; ++z) {
                double pks[2] = {P[L - z], P[z]}, kks[2] = {(z - L) * inv_L, z * inv_L};
                cplx t = c_make(0.0, 0.0);
                for (int a = 0; a < 8; ++a) {
                    int u = (a >> 2) & 1, v = (a >> 1) & 1, w = a & 1;
                    double ki = kis[u], kj = kjs[v], kk = kks[w];
                    double k = sqrt(ki * ki + kj * kj + kk * kk);
                    double rad = table ? table_lookup(k * g2p, table, n_table, k_max)
                                       : radial_profile_eval(k, code, a0, a1, l);
                    cplx ang = oracle_sph_harm_c(l, m, ki, kj, kk);
                    t = c_add(t, c_scale(pis[u] * pjs[v] * pks[w] * rad, ang));
                }
                size_t o = ((size_t)x * L + y) * L + z;
                out[2 * o] = t.re;
                out[2 * o + 1] = t.im;
            }
        }
    }
}

/* Pair product of a real field with the real, even window Re K_{l,0} of radial_multipole_window (even l),
 * from the field's rFFT half spectrum Fh (L, L, L/2+1) interleaved complex128: the full-grid sum
 * sum_k Re K(k) |F(k)|^2, with |F(-k)| = |F(k)| and K(-k) = K(k) folding it onto the half grid (planes
 * z = 0 and z = L/2 once, the others twice).  Each point's K is the 8-alias sum of the function above.
 * corr2pcf.py:366-372 in Parseval form; the caller divides by N^2. */
double oracle_pair_half_multipole(int L, const double *P, double R, int l, const double *Fh) {
    const double inv_L = 1.0 / L;
    const int Lh = L / 2 + 1;
    double total = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : total)
    for (int x = 0; x < L; ++x) {
        double pi0 = P[L - x], pi1 = P[x];
        double ki0 = (x - L) * inv_L, ki1 = x * inv_L;
        double row = 0.0;
        for (int y = 0; y < L; ++y) {
            double pj0 = P[L - y], pj1 = P[y];
            double kj0 = (y - L) * inv_L, kj1 = y * inv_L;
            for (int z = 0; z < Lh; ++z) {
                double pk0 = P[L - z], pk1 = P[z];
                double kk0 = (z - L) * inv_L, kk1 = z * inv_L;
                double t = pi0 * pj0 * pk0 * radial_multipole_eval(ki0, kj0, kk0, R, l, 0).re;
                t += pi0 * pj0 * pk1 * radial_multipole_eval(ki0, kj0, kk1, R, l, 0).re;
                t += pi0 * pj1 * pk0 * radial_multipole_eval(ki0, kj1, kk0, R, l, 0).re;
                t += pi0 * pj1 * pk1 * radial_multipole_eval(ki0, kj1, kk1, R, l, 0).re;
                t += pi1 * pj0 * pk0 * radial_multipole_eval(ki1, kj0, kk0, R, l, 0).re;
                t += pi1 * pj0 * pk1 * radial_multipole_eval(ki1, kj0, kk1, R, l, 0).re;
                t += pi1 * pj1 * pk0 * radial_multipole_eval(ki1, kj1, kk0, R, l, 0).re;
                t += pi1 * pj1 * pk1 * radial_multipole_eval(ki1, kj1, kk1, R, l, 0).re;
                size_t o = ((size_t)x * L + y) * Lh + z;
                double p2 = Fh[2 * o] * Fh[2 * o] + Fh[2 * o + 1] * Fh[2 * o + 1];
                double wz = (z == 0 || 2 * z == L) ? 1.0 : 2.0;
                row += wz * t * p2;
            }
        }
        total += row;
    }
    return total;
}

/* ---------------------------------------------------------------------------
 * corr3pcf_multipoles.py:134-170 (compute_3d_summand_cpu): returns the raw
 * complex sum (without the 4*pi/N factor).  F1, F2 interleaved complex128.
 * Per-thread partials are combined in thread order (deterministic for a
 * fixed thread count).
 * ------------------------------------------------------------------------- */
void oracle_summand(const double *data, const double *F1, const double *F2, const double *gamma,
                    int L, int support, double *out_re, double *out_im) {
    const int64_t N = (int64_t)L * L * L;
    const int64_t plane = (int64_t)L * L;
    int nt = oracle_num_threads();
    double *part = (double *)calloc(2 * (size_t)nt, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double tr = 0.0, ti = 0.0;
#pragma omp for schedule(static)
        for (int64_t idx = 0; idx < N; ++idx) {
            int64_t lx = idx / plane;
            int64_t rem = idx - lx * plane;
            int64_t ly = rem / L;
            int64_t lz = rem - ly * L;
            cplx s1 = c_make(0.0, 0.0);
            for (int l1x = 0; l1x < support; ++l1x) {
                int64_t i1x = ((lx - l1x) % L + L) % L;
                for (int l1y = 0; l1y < support; ++l1y) {
                    int64_t i1y = ((ly - l1y) % L + L) % L;
                    for (int l1z = 0; l1z < support; ++l1z) {
                        int64_t i1z = ((lz - l1z) % L + L) % L;
                        cplx s2 = c_make(0.0, 0.0);
                        for (int l2x = 0; l2x < support; ++l2x) {
                            int64_t i2x = ((lx - l2x) % L + L) % L;
                            double gx = gamma[l1x * support + l2x];
                            for (int l2y = 0; l2y < support; ++l2y) {
                                int64_t i2y = ((ly - l2y) % L + L) % L;
                                double gxy = gx * gamma[l1y * support + l2y];
                                for (int l2z = 0; l2z < support; ++l2z) {
                                    int64_t i2z = ((lz - l2z) % L + L) % L;
                                    double wgt = gxy * gamma[l1z * support + l2z];
                                    int64_t o = (i2x * L + i2y) * L + i2z;
                                    s2 = c_add(s2, c_make(wgt * F2[2 * o], wgt * F2[2 * o + 1]));
                                }
                            }
                        }
                        int64_t o1 = (i1x * L + i1y) * L + i1z;
                        s1 = c_add(s1, c_mul(c_make(F1[2 * o1], F1[2 * o1 + 1]), s2));
                    }
                }
            }
            cplx v = c_scale(data[idx], s1);
            tr += v.re;
            ti += v.im;
        }
        part[2 * tid] = tr;
        part[2 * tid + 1] = ti;
    }
    double sr = 0.0, si = 0.0;
    for (int t = 0; t < nt; ++t) { sr += part[2 * t]; si += part[2 * t + 1]; }
    free(part);
    *out_re = sr;
    *out_im = si;
}

/* ---------------------------------------------------------------------------
 * Monte-Carlo triplet estimator.
 * wavelet_grid.py:143-186 (interpolate_grid_at_pos_numba) and
 * corr3pcf_kernels.py:21-63 (generate_triangle_offsets), :100-140 (box-random
 * centres).  The per-rotation uniforms (a, b, c) are drawn by the caller.
 * ------------------------------------------------------------------------- */
/* One point.  A coordinate in (-2^-54, 0) rounds s - floor(s) to 1.0; the
 * fraction index is clamped to phi_res - 1 there (the reference indexes past
 * its table in that case). */
static double interp_one(const double *grid, const double *phi, int L, int phi_res, int support,
                         double sx, double sy, double sz) {
    const int64_t mask = L - 1;
    int64_t xc = (int64_t)floor(sx), yc = (int64_t)floor(sy), zc = (int64_t)floor(sz);
    int64_t xf = (int64_t)((sx - (double)xc) * phi_res);
    int64_t yf = (int64_t)((sy - (double)yc) * phi_res);
    int64_t zf = (int64_t)((sz - (double)zc) * phi_res);
    if (xf > phi_res - 1) xf = phi_res - 1;
    if (yf > phi_res - 1) yf = phi_res - 1;
    if (zf > phi_res - 1) zf = phi_res - 1;
    double acc = 0.0;
    for (int a = 0; a < support; ++a) {
        int64_t xg = (xc - a) & mask;
        double px = phi[xf + a * phi_res];
        for (int b = 0; b < support; ++b) {
            int64_t yg = (yc - b) & mask;
            double py = phi[yf + b * phi_res];
            for (int c = 0; c < support; ++c) {
                int64_t zg = (zc - c) & mask;
                double pz = phi[zf + c * phi_res];
                acc += grid[(xg * L + yg) * L + zg] * px * py * pz;
            }
        }
    }
    return acc;
}

void oracle_interpolate(double *out, const double *spos, int64_t n, const double *grid,
                        const double *phi, int L, int phi_res, int support,
                        double dx, double dy, double dz) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p)
        out[p] = interp_one(grid, phi, L, phi_res, support, spos[3 * p] + dx, spos[3 * p + 1] + dy,
                            spos[3 * p + 2] + dz);
}

void oracle_triangle_offsets(double R1, double R2, double mu, double phi, double ct1, double alpha,
                             double *o /* x2 y2 z2 x3 y3 z3 */) {
    double st1 = sqrt(1.0 - ct1 * ct1);
    double dx1 = st1 * cos(phi), dy1 = st1 * sin(phi), dz1 = ct1;
    o[0] = R1 * dx1; o[1] = R1 * dy1; o[2] = R1 * dz1;
    double ox1, oy1, oz1;
    if (fabs(dx1) > 1e-10 || fabs(dy1) > 1e-10) { ox1 = -dy1; oy1 = dx1; oz1 = 0.0; }
    else { ox1 = 0.0; oy1 = -dz1; oz1 = dy1; }
    double nrm = sqrt(ox1 * ox1 + oy1 * oy1 + oz1 * oz1);
    ox1 /= nrm; oy1 /= nrm; oz1 /= nrm;
    double ox2 = dy1 * oz1 - dz1 * oy1;
    double oy2 = dz1 * ox1 - dx1 * oz1;
    double oz2 = dx1 * oy1 - dy1 * ox1;
    double ca = cos(alpha), sa = sin(alpha);
    double st2 = 1.0 - mu * mu;
    if (st2 < 0.0) st2 = 0.0;
    double st = sqrt(st2);
    double dx2 = mu * dx1 + st * (ca * ox1 + sa * ox2);
    double dy2 = mu * dy1 + st * (ca * oy1 + sa * oy2);
    double dz2 = mu * dz1 + st * (ca * oz1 + sa * oz2);
    o[3] = R2 * dx2; o[4] = R2 * dy2; o[5] = R2 * dz2;
}

/* corr3pcf_kernels.py:100-140 with caller-supplied uniforms abc (n_rot x 3). */
double oracle_triplet_box_random(double R1, double R2, double mu, const double *centers, int64_t n,
                                 int n_rot, const double *abc, const double *e1, const double *e2,
                                 const double *e3, const double *phi, int L, int phi_res, int support) {
    double *n1 = (double *)malloc(sizeof(double) * n);
    double *n2 = (double *)malloc(sizeof(double) * n);
    double *n3 = (double *)malloc(sizeof(double) * n);
    oracle_interpolate(n1, centers, n, e1, phi, L, phi_res, support, 0.0, 0.0, 0.0);
    double total = 0.0;
    for (int r = 0; r < n_rot; ++r) {
        double o[6];
        oracle_triangle_offsets(R1, R2, mu, 2.0 * M_PI * abc[3 * r], 2.0 * abc[3 * r + 1] - 1.0,
                                2.0 * M_PI * abc[3 * r + 2], o);
        oracle_interpolate(n2, centers, n, e2, phi, L, phi_res, support, o[0], o[1], o[2]);
        oracle_interpolate(n3, centers, n, e3, phi, L, phi_res, support, o[3], o[4], o[5]);
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += n1[i] * n2[i] * n3[i];
        total += s;
    }
    free(n1); free(n2); free(n3);
    return total / ((double)n_rot * (double)n);
}

/* ---------------------------------------------------------------------------
 * Binned Monte-Carlo triangle sampler.  No reference counterpart as a whole:
 * it bins the reference's triangle products n1 n2 n3 (geometry
 * corr3pcf_kernels.py:21-63, interpolation wavelet_grid.py:143-186) over
 * (r1, r2, cos theta) bins, with uniforms from Philox4x32-10 (Salmon, Moraes,
 * Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11).
 * Sample i, stream s: counter (i_lo, i_hi, 0x9E3779B9, 0x85EBCA6B), key
 * seed ^ (0x9E3779B97F4A7C15 * (s + 1)); uniform = 53 bits from two words.
 * ------------------------------------------------------------------------- */
void oracle_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void mc_uniforms2(uint64_t i, uint64_t seed, uint64_t stream, double *a, double *b) {
    const uint64_t key = seed ^ (0x9E3779B97F4A7C15ULL * (stream + 1));
    const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0x9E3779B9u, 0x85EBCA6Bu};
    const uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    uint32_t r[4];
    oracle_philox4x32_10(ctr, k, r);
    *a = ((double)(r[0] >> 5) * 67108864.0 + (double)(r[1] >> 6)) / 9007199254740992.0;
    *b = ((double)(r[2] >> 5) * 67108864.0 + (double)(r[3] >> 6)) / 9007199254740992.0;
}

/* first index with cdf[i] > u, clamped to n - 1 */
static int mc_upper_bound(const float *cdf, int n, float u) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid; else lo = mid + 1;
    }
    return lo < n ? lo : n - 1;
}

static double mc_interp(const double *g, const double *phi, int L, int res, double sx, double sy, double sz) {
    return interp_one(g, phi, L, res, 3, sx, sy, sz);
}

/* Plain estimator when cell_bucket is NULL; otherwise the defensive-mixture
 * importance proposal on the sub-cell grid (side Ls = L sub): with probability
 * alpha a uniform draw, else bucket k ~ bucket_cdf, a uniform sub-cell of
 * bucket k, a uniform point p in it, bin ~ bin_cdf[k] and an anchor vertex a
 * (the apex when anchors = 1, else uniform over the three) placed at p.  With
 * R(k, b) = bucket_density[k] bin_pdf[k][b] and k_a the bucket of vertex a,
 * w = 1 / (alpha + (1 - alpha) mean_a R(k_a, bin)) (apex only when anchors = 1).
 * Outputs per-bin sum w f, sum (w f)^2 and counts (B = nr1 nr2 nmu,
 * bin = (ir1 nr2 + ir2) nmu + imu). */
void oracle_mc_binned(const double *g1, const double *g2, const double *g3, const double *phi, int L,
                      int phi_res, int nr1, int nr2, int nmu, double rmin1, double rmax1, double rmin2,
                      double rmax2, int64_t n, uint64_t seed, const uint8_t *cell_bucket,
                      const int32_t *bucket_cells, const int64_t *bucket_off, const float *bucket_cdf,
                      const double *bucket_density, const float *bin_cdf, const float *bin_pdf, int nbucket,
                      double alpha, int sub, int anchors, double *sum_wf, double *sum_wf2, int64_t *count) {
    const int B = nr1 * nr2 * nmu;
    const int64_t Ls = (int64_t)L * sub, lm = Ls - 1;
    const int imp = cell_bucket != NULL;
    const double dr1 = (rmax1 - rmin1) / nr1, dr2 = (rmax2 - rmin2) / nr2, dmu = 2.0 / nmu;
    memset(sum_wf, 0, sizeof(double) * B);
    memset(sum_wf2, 0, sizeof(double) * B);
    memset(count, 0, sizeof(int64_t) * B);
#pragma omp parallel
    {
        double *s1 = (double *)calloc(B, sizeof(double));
        double *s2 = (double *)calloc(B, sizeof(double));
        int64_t *cn = (int64_t *)calloc(B, sizeof(int64_t));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double ux, uy, uz, ubin, upick, umix, ubin2, w1, w2, w3, oph, oct, oal, uanc;
            mc_uniforms2((uint64_t)i, seed, 0, &ux, &uy);
            mc_uniforms2((uint64_t)i, seed, 1, &uz, &ubin);
            mc_uniforms2((uint64_t)i, seed, 2, &upick, &umix);
            mc_uniforms2((uint64_t)i, seed, 3, &ubin2, &w1);
            mc_uniforms2((uint64_t)i, seed, 4, &w2, &w3);
            mc_uniforms2((uint64_t)i, seed, 5, &oph, &oct);
            mc_uniforms2((uint64_t)i, seed, 6, &oal, &uanc);
            double px, py, pz;
            int bin, anchor = 0;
            if (!imp || umix < alpha) {
                px = ux * L; py = uy * L; pz = uz * L;
                bin = (int)(ubin * B);
                if (bin >= B) bin = B - 1;
            } else {
                const int k = mc_upper_bound(bucket_cdf, nbucket, (float)ubin);
                const int64_t c0 = bucket_off[k], c1 = bucket_off[k + 1];
                int64_t pick = c0 + (int64_t)(upick * (double)(c1 - c0));
                if (pick >= c1) pick = c1 - 1;
                const int64_t sc = bucket_cells[pick];
                px = (double)(sc / (Ls * Ls) + ux) / sub;
                py = (double)((sc / Ls) % Ls + uy) / sub;
                pz = (double)(sc % Ls + uz) / sub;
                bin = mc_upper_bound(bin_cdf + (int64_t)k * B, B, (float)ubin2);
                if (anchors == 3) { anchor = (int)(uanc * 3.0); if (anchor > 2) anchor = 2; }
            }
            const int imu = bin % nmu, ir2 = (bin / nmu) % nr2, ir1 = bin / (nmu * nr2);
            const double r1 = rmin1 + (ir1 + w1) * dr1, r2 = rmin2 + (ir2 + w2) * dr2;
            const double mu = -1.0 + (imu + w3) * dmu;
            double o[6];
            oracle_triangle_offsets(r1, r2, mu, 2.0 * M_PI * oph, 2.0 * oct - 1.0, 2.0 * M_PI * oal, o);
            double cx = px, cy = py, cz = pz;
            if (anchor == 1) { cx = px - o[0]; cy = py - o[1]; cz = pz - o[2]; }
            if (anchor == 2) { cx = px - o[3]; cy = py - o[4]; cz = pz - o[5]; }
            const double v[3][3] = {{cx, cy, cz}, {cx + o[0], cy + o[1], cz + o[2]},
                                    {cx + o[3], cy + o[4], cz + o[5]}};
            const double f = mc_interp(g1, phi, L, phi_res, v[0][0], v[0][1], v[0][2]) *
                             mc_interp(g2, phi, L, phi_res, v[1][0], v[1][1], v[1][2]) *
                             mc_interp(g3, phi, L, phi_res, v[2][0], v[2][1], v[2][2]);
            double w = 1.0;
            if (imp) {
                double rel = 0.0;
                const int na = anchors == 3 ? 3 : 1;
                for (int a = 0; a < na; ++a) {
                    const int64_t s3 = ((((int64_t)floor(v[a][0] * sub)) & lm) * Ls +
                                        (((int64_t)floor(v[a][1] * sub)) & lm)) * Ls +
                                       (((int64_t)floor(v[a][2] * sub)) & lm);
                    const int k = cell_bucket[s3];
                    rel += bucket_density[k] * (double)bin_pdf[(int64_t)k * B + bin];
                }
                if (na == 3) rel *= (1.0 / 3.0);
                w = 1.0 / (alpha + (1.0 - alpha) * rel);
            }
            s1[bin] += w * f;
            s2[bin] += (w * f) * (w * f);
            cn[bin] += 1;
        }
#pragma omp critical
        {
            for (int b = 0; b < B; ++b) { sum_wf[b] += s1[b]; sum_wf2[b] += s2[b]; count[b] += cn[b]; }
        }
        free(s1); free(s2); free(cn);
    }
}
