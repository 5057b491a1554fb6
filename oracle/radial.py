"""Oracle: radial-profile multipole windows (non-thin edges).  Test infrastructure only.

Restates radial_profiles.py (the Hankel tables U_l(k) = 4 pi int r^2 w(r) j_l(2 pi k r) dr of the built-in
profiles by Gauss-Legendre quadrature, their node counts and supports, :28-36, :189-260, :297-415, :651-709)
and radial_multipole_windows.py:267-316 (which profile takes which builder), with the C restatement of the
builders (oracle/c/oracle.c: oracle_profile_multipole_window).  Pinned in tests/test_oracle.py by the
closed-form monopoles, scipy.integrate.quad at l >= 1 and the thin-shell limit of a thin thick shell.
"""

import math

import numpy as np
from numpy.polynomial.legendre import leggauss
from scipy.special import spherical_jn

from . import native
from .windows import convolve_complex

PROFILE_CODES = {"shell": 0, "gaussian_shell": 1, "thick_shell": 2, "sphere": 3, "gaussian": 4}


def _nodes_count(phase, minimum, per_cycle):
    return max(minimum, int(math.ceil(per_cycle * max(phase, 1.0))) + 1)


def _gl(a, b, n):
    x, w = leggauss(n)
    return 0.5 * (b - a) * (x + 1.0) + a, 0.5 * (b - a) * w


def _sphere_T(k, R):
    q = 2.0 * np.pi * np.asarray(k, dtype=np.float64) * R
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(q == 0.0, 1.0, 3.0 * (np.sin(q) - q * np.cos(q)) / np.where(q == 0.0, 1.0, q) ** 3)


def gaussian_shell_T(k, Rs, Rg):
    """window_functions.py:96-129."""
    k = np.asarray(k, dtype=np.float64)
    qs, qg = 2.0 * np.pi * k * Rs, 2.0 * np.pi * k * Rg
    with np.errstate(invalid="ignore", divide="ignore"):
        sinc = np.where(qs == 0.0, 1.0, np.sin(qs) / np.where(qs == 0.0, 1.0, qs))
    return np.where(k == 0.0, 1.0, (Rg * Rg * np.cos(qs) + Rs * Rs * sinc) / (Rs * Rs + Rg * Rg)
                    * np.exp(-0.5 * qg * qg))


def profile_nodes(radial_type, len_args, k_max):
    """(r, dr, w) with w normalised as the reference does for the built-in profiles."""
    t = radial_type
    if t == "sphere":
        R = float(len_args["R"])
        r, dr = _gl(0.0, R, _nodes_count(k_max * R, 1024, 32))
        return r, dr, np.full(r.shape, 3.0 / (4.0 * np.pi * R ** 3))
    if t == "thick_shell":
        R, dR = float(len_args["R"]), float(len_args["delta_R"])
        a, b = max(0.0, R - 0.5 * dR), R + 0.5 * dR
        r, dr = _gl(a, b, _nodes_count(k_max * (b - a), 1024, 32))
        return r, dr, np.full(r.shape, 3.0 / (4.0 * np.pi * (b ** 3 - a ** 3)))
    if t == "gaussian":
        R = float(len_args["R"])
        r, dr = _gl(0.0, 8.0 * R, _nodes_count(k_max * 8.0 * R, 1024, 32))
        w = (2.0 * np.pi * R * R) ** -1.5 * np.exp(-0.5 * (r / R) ** 2)
        return r, dr, w / (4.0 * np.pi * np.sum(dr * r * r * w))
    if t == "gaussian_shell":
        Rs, Rg = float(len_args["R_shell"]), float(len_args["R_smooth"])
        rmax, K = Rs + 8.0 * Rg, max(k_max, 8.0 / (2.0 * np.pi * Rg))
        r, dr = _gl(0.0, rmax, _nodes_count(k_max * rmax, 1024, 32))
        k, dk = _gl(0.0, K, _nodes_count(K * rmax, 1024, 32))
        w = np.array([np.sum(4.0 * np.pi * dk * k * k * gaussian_shell_T(k, Rs, Rg) * spherical_jn(0, 2.0 * np.pi * k * ri))
                      for ri in r])
        return r, dr, w / (4.0 * np.pi * np.sum(dr * r * r * w))
    raise ValueError(radial_type)


def hankel_table(radial_type, len_args, l, k_max):
    """radial_profiles.py:651-709 for the built-in profiles: U_l on n uniform k in [0, k_max]."""
    if radial_type == "gaussian_shell" and float(len_args["R_smooth"]) == 0.0:
        R = float(len_args["R_shell"])
        k = np.linspace(0.0, k_max, _nodes_count(k_max * R, 4097, 64))
        return spherical_jn(l, 2.0 * np.pi * k * R)
    r, dr, w = profile_nodes(radial_type, len_args, k_max)
    k = np.linspace(0.0, k_max, _nodes_count(k_max * r.max(), 4097, 64))
    out = np.empty(k.size)
    c = 4.0 * np.pi * dr * r * r * w
    for i in range(k.size):
        out[i] = np.dot(spherical_jn(l, 2.0 * np.pi * k[i] * r), c)
    return out


def profile_multipole_window(L, P, radial_type, len_args, l, m, box):
    """radial_multipole_windows.py:267-316: the (l, m) kernel (L, L, L) complex128 of one edge profile with
    physical len_args in a box of side ``box``."""
    t = radial_type
    g2p = L / box
    out = np.empty((L, L, L), dtype=np.complex128)
    P = native.f64(P)
    lib = native.lib()
    grid = {k: float(v) * g2p for k, v in len_args.items()}
    if t == "shell" or (l == 0 and t in PROFILE_CODES) or (t == "gaussian_shell" and len_args["R_smooth"] == 0.0):
        if t == "shell" or (t == "gaussian_shell" and l != 0):
            code, a0, a1 = 0, grid.get("R", grid.get("R_shell")), 0.0
        elif t == "gaussian_shell":
            code, a0, a1 = 1, grid["R_shell"], grid["R_smooth"]
        elif t == "thick_shell":
            code, a0, a1 = 2, grid["R"], grid["delta_R"]
        else:
            code, a0, a1 = PROFILE_CODES[t], grid["R"], 0.0
        lib.oracle_profile_multipole_window(L, native.ptr(P), code, a0, a1, None, 0, 0.0, 0.0, int(l), int(m),
                                            native.ptr(out.view(np.float64)))
        return out
    k_max = math.sqrt(3.0) * g2p
    tab = native.f64(hankel_table(t, len_args, l, k_max))
    lib.oracle_profile_multipole_window(L, native.ptr(P), 0, 0.0, 0.0, native.ptr(tab), tab.size, k_max, g2p,
                                        int(l), int(m), native.ptr(out.view(np.float64)))
    return out


def profile_fields(delta, P, radial_type, len_args, l, m, box, workers=1):
    """delta @ the (l, m) window of an edge profile (corr3pcf_multipoles.py:267-321)."""
    K = profile_multipole_window(delta.shape[0], P, radial_type, len_args, l, m, box)
    return convolve_complex(delta, K, workers)
