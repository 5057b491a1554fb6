"""Oracle: window kernels, special functions, FFT convolutions and pair products.

Test infrastructure only.
"""

import numpy as np
from scipy.fft import fftn, ifftn, irfftn, rfftn

from . import native

# Window type codes shared with oracle/c/oracle.c (window_functions.py:441-458 names).
WINDOW_CODES = {"sphere": 0, "gaussian": 1, "shell": 2, "thick_shell": 3, "gaussian_shell": 4,
                "inverse_laplacian": 5}


def window_args(win_type, rescaled_len_args):
    """Positional arguments of the built-in transfer functions (window_functions.py:29-129)."""
    if win_type in ("sphere", "gaussian", "shell"):
        return float(rescaled_len_args["R"]), 0.0
    if win_type == "thick_shell":
        return float(rescaled_len_args["R"]), float(rescaled_len_args["delta_R"])
    if win_type == "gaussian_shell":
        return float(rescaled_len_args["R_shell"]), float(rescaled_len_args["R_smooth"])
    if win_type == "inverse_laplacian":
        return 0.0, 0.0
    raise ValueError(win_type)


def octant_rfft_kernel(L, P, win_type, a0, a1=0.0):
    """convolution.py:14-38 + :263-281 (bandwidth 1) -> (L, L, L//2+1) float64."""
    out = np.empty((L, L, L // 2 + 1), dtype=np.float64)
    P = native.f64(P)
    native.lib().oracle_octant_rfft_kernel(L, native.ptr(P), WINDOW_CODES[win_type], a0, a1, native.ptr(out))
    if win_type == "inverse_laplacian":  # window.py:266-272 zero-mode convention
        out[0, 0, 0] = 0.0
    return out


def full_rfft_kernel(L, P, win_type, a0, a1=0.0):
    """convolution.py:73-157 (explicit 8 aliases, bandwidth 1) -> (L, L, L//2+1) float64."""
    out = np.empty((L, L, L // 2 + 1), dtype=np.float64)
    P = native.f64(P)
    native.lib().oracle_full_rfft_kernel(L, native.ptr(P), WINDOW_CODES[win_type], a0, a1, native.ptr(out))
    if win_type == "inverse_laplacian":
        out[0, 0, 0] = 0.0
    return out


def radial_multipole_window(L, P, R_grid, l, m):
    """radial_multipole_windows.py:147-183 (thin shell) -> (L, L, L) complex128."""
    out = np.empty((L, L, L), dtype=np.complex128)
    P = native.f64(P)
    native.lib().oracle_radial_multipole_window(L, native.ptr(P), float(R_grid), int(l), int(m),
                                                native.ptr(out.view(np.float64)))
    return out


def spherical_jn(l, x):
    """special_functions.py:49-128 (scalar)."""
    return native.lib().oracle_spherical_jn(int(l), float(x))


def sph_harm(l, m, ki, kj, kk):
    """special_functions.py:307-334 (scalar) -> complex."""
    re = native.ctypes.c_double()
    im = native.ctypes.c_double()
    native.lib().oracle_sph_harm(int(l), int(m), float(ki), float(kj), float(kk),
                                 native.ctypes.byref(re), native.ctypes.byref(im))
    return complex(re.value, im.value)


def assoc_legendre(l, m, x):
    """special_functions.py:273-304 (scalar)."""
    return native.lib().oracle_assoc_legendre(int(l), int(m), float(x))


def convolve_rfft(field, kernel, workers=1):
    """convolution.py:284-288."""
    sc = rfftn(field, workers=workers)
    sc *= kernel
    return irfftn(sc, s=field.shape, workers=workers)


def convolve_complex(field, kernel, workers=1):
    """convolution.py:291-295."""
    sc = fftn(field, workers=workers)
    sc *= kernel
    return ifftn(sc, workers=workers)


def pair_product(field1, field2, kernel, workers=1):
    """corr2pcf.py:366-372: sum_x (f1 @ W)(x) f2(x) / N."""
    conv = convolve_rfft(field1, kernel, workers)
    return float(np.einsum("ijk,ijk->", conv, field2, optimize=True) / conv.size)


def pair_product_parseval(field1, field2, kernel):
    """Independent Parseval form: (1/N^2) sum_k W(k) F1(k) conj(F2(k)) over the full spectrum."""
    L = field1.shape[0]
    F1 = np.fft.fftn(field1)
    F2 = np.fft.fftn(field2)
    # expand the rFFT kernel to the full grid by Hermitian symmetry of a real kernel
    Wf = np.empty((L, L, L))
    h = L // 2 + 1
    Wf[:, :, :h] = kernel
    ix = (-np.arange(L)) % L
    Wf[:, :, h:] = kernel[ix][:, ix][:, :, (L - np.arange(h, L)) % L]
    return float(np.real(np.sum(Wf * F1 * np.conj(F2))) / L ** 6)


def pair_product_realspace(field1, field2, win_real):
    """Brute force sum_x sum_y w(x - y) f1(y) f2(x) / N with a real-space window array."""
    L = field1.shape[0]
    conv = np.real(np.fft.ifftn(np.fft.fftn(win_real) * np.fft.fftn(field1)))
    return float(np.sum(conv * field2) / L ** 3)


# ---------------------------------------------------------------- line-of-sight windows
# window_functions.py:164-319: ring, disk, cylshell and cylinder windows about a line of sight n
# (normalised), with k_par = k . n, k_perp = sqrt(|k|^2 - k_par^2), q_perp = 2 pi k_perp R,
# q_par = 2 pi k_par H.  J0 and J1 are scipy.special's (the reference evaluates 8-digit rational
# approximations of them, special_functions.py:131-215).
LOS_WINDOW_TYPES = ("ring", "disk", "cylshell", "cylinder")


def los_window(win_type, ki, kj, kk, R, H, los=(0.0, 0.0, 1.0)):
    """Vectorised transfer function of one line-of-sight window (k in cycles per grid cell)."""
    from scipy.special import j0, j1
    n = np.asarray(los, dtype=np.float64)
    n = n / np.linalg.norm(n)
    ki, kj, kk = (np.asarray(a, dtype=np.float64) for a in (ki, kj, kk))
    kpar = ki * n[0] + kj * n[1] + kk * n[2]
    kperp = np.sqrt(np.maximum(ki * ki + kj * kj + kk * kk - kpar * kpar, 0.0))
    qp = 2.0 * np.pi * kperp * R
    qz = 2.0 * np.pi * kpar * H
    if win_type in ("ring", "cylshell"):
        perp = j0(qp)
    else:
        with np.errstate(invalid="ignore", divide="ignore"):
            perp = np.where(qp == 0.0, 1.0, 2.0 * j1(qp) / np.where(qp == 0.0, 1.0, qp))
    if win_type in ("ring", "disk"):
        par = np.cos(qz)
    else:
        with np.errstate(invalid="ignore", divide="ignore"):
            par = np.where(qz == 0.0, 1.0, np.sin(qz) / np.where(qz == 0.0, 1.0, qz))
    return perp * par


def los_rfft_kernel(L, P, win_type, R, H, los=(0.0, 0.0, 1.0)):
    """convolution.py:73-157 (bandwidth 1): K[x, y, z] = sum over the 8 signed aliases
    (x/L or -(L-x)/L per axis, the reference's summation order) of P P P W -> (L, L, L//2+1)."""
    P = np.asarray(P, dtype=np.float64)
    x = np.arange(L)[:, None, None]
    y = np.arange(L)[None, :, None]
    z = np.arange(L // 2 + 1)[None, None, :]
    pos = [(x, x / L), (y, y / L), (z, z / L)]
    neg = [(L - x, -(L - x) / L), (L - y, -(L - y) / L), (L - z, -(L - z) / L)]
    order = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1))
    out = np.zeros((L, L, L // 2 + 1))
    for sx, sy, sz in order:
        (ix, kx) = neg[0] if sx else pos[0]
        (iy, ky) = neg[1] if sy else pos[1]
        (iz, kz) = neg[2] if sz else pos[2]
        out = out + P[ix] * P[iy] * P[iz] * los_window(win_type, kx, ky, kz, R, H, los)
    return out


# ---------------------------------------------------------------- 2PCF multipoles xi_l(s)
# The reference has no xi_l task; the build defines it (corr2pcf.py:366-372 pair products with the shell
# multipole window of radial_multipole_windows.py:128-144, m = 0):
#   xi_l(s) = (-i)^l sqrt(4 pi (2l + 1)) <(delta @ W_{s,l,0}) delta> / rr,   W = j_l(2 pi |k| s) Y_l0(k-hat).
# Two independent routes below.  For even l they agree: the plane-wave expansion gives
# (2l + 1) int_0^1 dmu P_l(mu) J0(q_perp) cos(q_par) = i^l sqrt(4 pi (2l + 1)) j_l(q) Y_l0(k-hat) per alias,
# so the Gauss-Legendre mu-integral of the ring-window pair products xi(s, mu) (window_functions.py:164-192,
# corr2pcf.py:275-284 mapping R = s sqrt(1 - mu^2), H = s mu) pins the prefactor.

def xi_ell_window_route(delta, P, s_grid, ell, rr, workers=1):
    """xi_l(s) from the complex shell-multipole window: (-i)^l sqrt(4 pi (2l+1)) Re<(delta @ W_l0) delta>/rr
    (s in grid units; even l, where (-i)^l = (-1)^(l/2))."""
    if ell % 2:
        raise ValueError("xi_l of a real field is defined for even l")
    L = delta.shape[0]
    K = radial_multipole_window(L, P, float(s_grid), int(ell), 0)
    pp = float(np.real(np.sum(convolve_complex(delta, K, workers) * delta))) / delta.size
    return (-1.0) ** (ell // 2) * np.sqrt(4.0 * np.pi * (2 * ell + 1)) * pp / rr


def xi_ell_ring_route(delta, P, s_grid, ell, rr, n_mu=64, los=(0.0, 0.0, 1.0), workers=1):
    """xi_l(s) = (2l + 1) int_0^1 xi(s, mu) P_l(mu) dmu with xi(s, mu) from ring-window pair products at
    n_mu Gauss-Legendre nodes on [0, 1] (exact for the even integrand up to degree 4 n_mu - 1 in mu)."""
    from scipy.special import eval_legendre
    L = delta.shape[0]
    x, w = np.polynomial.legendre.leggauss(int(n_mu))
    mu, wmu = 0.5 * (x + 1.0), 0.5 * w
    s = float(s_grid)
    total = 0.0
    for m, wm in zip(mu, wmu):
        K = los_rfft_kernel(L, P, "ring", s * np.sqrt(1.0 - m * m), s * m, los)
        total += wm * eval_legendre(int(ell), m) * pair_product(delta, delta, K, workers) / rr
    return (2 * ell + 1) * total


def xi_ell_half_spectrum(delta, P, s_grid, ell, rr, workers=1):
    """xi_window_route from the rFFT half spectrum without building the window grid (large grids):
    (-i)^l sqrt(4 pi (2l+1)) sum_k Re K_l0(k) |F(k)|^2 / N^2 / rr for even l."""
    if ell % 2:
        raise ValueError("xi_l of a real field is defined for even l")
    Fh = np.ascontiguousarray(rfftn(np.asarray(delta, dtype=np.float64), workers=workers))
    P = native.f64(P)
    L = delta.shape[0]
    raw = native.lib().oracle_pair_half_multipole(L, native.ptr(P), float(s_grid), int(ell),
                                                   native.ptr(Fh.view(np.float64)))
    pp = raw / float(delta.size) ** 2
    return (-1.0) ** (ell // 2) * np.sqrt(4.0 * np.pi * (2 * ell + 1)) * pp / rr
