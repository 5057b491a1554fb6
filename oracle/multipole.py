"""Oracle: spherical-harmonic 3PCF multipoles.  Test infrastructure only."""

import numpy as np
from scipy.special import gammaln

from . import native
from .scaling import cal_gamma, fourier_power_spectrum
from .windows import convolve_complex, radial_multipole_window


def summand_c(delta1, F1, F2, gamma):
    """corr3pcf_multipoles.py:134-170 -> raw complex sum (no 4*pi/N)."""
    L = delta1.shape[0]
    d = native.f64(delta1)
    a = np.ascontiguousarray(F1, dtype=np.complex128)
    b = np.ascontiguousarray(F2, dtype=np.complex128)
    g = native.f64(gamma)
    re = native.ctypes.c_double()
    im = native.ctypes.c_double()
    native.lib().oracle_summand(native.ptr(d), native.ptr(a.view(np.float64)), native.ptr(b.view(np.float64)),
                                native.ptr(g), L, g.shape[0], native.ctypes.byref(re), native.ctypes.byref(im))
    return complex(re.value, im.value)


def summand_bruteforce(delta1, F1, F2, gamma):
    """Independent numpy form with np.roll shifts: sum_{l1,l2} G3 sum_x d[x] F1[x-l1] F2[x-l2]."""
    s = gamma.shape[0]
    total = 0.0 + 0.0j
    for l1 in np.ndindex(s, s, s):
        a = np.roll(F1, l1, axis=(0, 1, 2)) * delta1
        for l2 in np.ndindex(s, s, s):
            g = gamma[l1[0], l2[0]] * gamma[l1[1], l2[1]] * gamma[l1[2], l2[2]]
            total += g * np.sum(a * np.roll(F2, l2, axis=(0, 1, 2)))
    return total


def summand_pairs(delta1, F1s, F2s, gamma):
    """The raw sums of summand_c for every pair (F1s[i], F2s[j]) at once: (n1, n2) complex.

    Same sum, regrouped: with A_i the (27, N) stack of F1s[i] shifted by each tap l (F[x - l], np.roll)
    and B_j that of F2s[j], sum_x d1[x] sum_{l1,l2} G3[l1,l2] F1[x-l1] F2[x-l2] = sum G3 * (A_i D B_j^T)
    with D = diag(d1): one matrix product per call instead of 27 x 27 passes per pair."""
    s = gamma.shape[0]
    taps = list(np.ndindex(s, s, s))
    G3 = np.array([[gamma[a[0], b[0]] * gamma[a[1], b[1]] * gamma[a[2], b[2]] for b in taps] for a in taps])

    def stack(fields):
        out = np.empty((len(fields), len(taps), delta1.size), dtype=np.complex128)
        for i, f in enumerate(fields):
            for t, sh in enumerate(taps):
                out[i, t] = np.roll(f, sh, axis=(0, 1, 2)).ravel()
        return out

    A = stack(F1s)
    B = A if F2s is F1s else stack(F2s)
    n1, n2, T = A.shape[0], B.shape[0], len(taps)
    M = (A.reshape(n1 * T, -1) * delta1.ravel()) @ B.reshape(n2 * T, -1).T
    M = M.reshape(n1, T, n2, T)
    return np.einsum("itju,tu->ij", M, G3)


def m_summand(delta1, F1, F2, gamma):
    """corr3pcf_multipoles.py:393-413 CPU branch: 4*pi*raw/N."""
    return (4.0 * np.pi) * summand_c(delta1, F1, F2, gamma) / delta1.size


def combine_multipole_m_terms(m_values, l):
    """corr3pcf_multipoles.py:173-180."""
    coeff = complex(m_values[0])
    for m in range(1, l + 1):
        coeff += ((-1) ** m) * complex(m_values[m])
        coeff += ((-1) ** (-m)) * np.conj(complex(m_values[m]))
    coeff *= (-1) ** l
    return coeff.real


def multipole_fields(delta, P, R_grid, l, m, workers=1):
    """corr3pcf_multipoles.py:267-321 without caching: field @ radial_multipole window."""
    K = radial_multipole_window(delta.shape[0], P, R_grid, l, m)
    return convolve_complex(delta, K, workers)


def calc_ddd_multipole(d1, d2, d3, phi, R12_grid, R13_grid, l_min, l_max, workers=1,
                       phi_res=1024, return_m_values=False):
    """corr3pcf_multipoles.py:460-579 (CPU backend, no caching).  Radii in grid units."""
    L = d1.shape[0]
    P = fourier_power_spectrum(phi, 0, 1, L, phi_res)
    gamma = cal_gamma(phi, len(phi) // phi_res, phi_res)
    out = []
    mvals = {}
    for l in range(l_min, l_max + 1):
        mv = np.empty(l + 1, dtype=np.complex128)
        for m in range(l + 1):
            F1 = multipole_fields(d2, P, R12_grid, l, m, workers)
            F2 = multipole_fields(d3, P, R13_grid, l, -m, workers)
            mv[m] = m_summand(d1, F1, F2, gamma)
        mvals[l] = mv
        out.append(combine_multipole_m_terms(mv, l))
    out = np.asarray(out)
    if return_m_values:
        return out, mvals
    return out


def legendre_triple_coeff(l, lp, k):
    """special_functions.py:337-363."""
    s = l + lp + k
    if (k > l + lp) or (k < abs(l - lp)) or (s % 2 != 0):
        return 0.0
    g = s // 2
    log_val = (gammaln(s - 2 * l + 1) + gammaln(s - 2 * lp + 1) + gammaln(s - 2 * k + 1) - gammaln(s + 2)
               + 2 * (gammaln(g + 1) - gammaln(g - l + 1) - gammaln(g - lp + 1) - gammaln(g - k + 1)))
    return float((2 * k + 1) * np.exp(log_val))


def build_mixing_matrix(B, lmax):
    """special_functions.py:366-389."""
    B = np.asarray(B, dtype=np.float64)
    f = B / B[0]
    M = np.zeros((lmax + 1, lmax + 1))
    for l in range(lmax + 1):
        for lp in range(1, lmax + 1):
            if l == lp:
                M[0, l] += f[lp] / (2 * l + 1)
    for k in range(1, lmax + 1):
        for l in range(lmax + 1):
            for lp in range(1, lmax + 1):
                M[k, l] += legendre_triple_coeff(l, lp, k) * f[lp]
    M += np.eye(lmax + 1)
    M *= B[0]
    return M


def solve_multipoles_from_ratio(A, B, lmax):
    """special_functions.py:392-409 (without the warning print)."""
    M = build_mixing_matrix(np.asarray(B)[:lmax + 1], lmax)
    return np.linalg.solve(M, np.asarray(A, dtype=np.float64)[:lmax + 1]), M, np.linalg.cond(M)
