"""Host-side Legendre-multipole algebra used by the 3PCF multipole task.

Same definitions as special_functions.py:337-409: the Legendre triple-product
coefficient, the mixing matrix M with A = M C for A(mu) = B(mu) C(mu) truncated at
lmax, and the ratio solve C = M^-1 A.
"""

import numpy as np
from scipy.special import gammaln


def legendre_triple_coeff(l, lp, k):
    """(2k+1)/2 int P_lp P_l P_k dmu (Wigner 3j squared closed form)."""
    s = l + lp + k
    if k > l + lp or k < abs(l - lp) or s % 2:
        return 0.0
    g = s // 2
    v = (gammaln(s - 2 * l + 1) + gammaln(s - 2 * lp + 1) + gammaln(s - 2 * k + 1) - gammaln(s + 2)
         + 2 * (gammaln(g + 1) - gammaln(g - l + 1) - gammaln(g - lp + 1) - gammaln(g - k + 1)))
    return float((2 * k + 1) * np.exp(v))


def build_mixing_matrix(B, lmax):
    B = np.asarray(B, dtype=np.float64)
    if B[0] == 0:
        raise ValueError("B[0] is zero.")
    f = B / B[0]
    M = np.eye(lmax + 1)
    for l in range(1, lmax + 1):
        M[0, l] += f[l] / (2 * l + 1)
    for k in range(1, lmax + 1):
        for l in range(lmax + 1):
            for lp in range(1, lmax + 1):
                M[k, l] += legendre_triple_coeff(l, lp, k) * f[lp]
    return M * B[0]


def solve_multipoles_from_ratio(A, B, lmax, rcond_warning=1e12):
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    if A.shape[0] < lmax + 1 or B.shape[0] < lmax + 1:
        raise ValueError("A and B must have length at least lmax+1.")
    M = build_mixing_matrix(B[:lmax + 1], lmax)
    cond = np.linalg.cond(M)
    return np.linalg.solve(M, A[:lmax + 1]), M, cond
