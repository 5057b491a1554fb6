"""Oracle: db2 scaling function samples, |phi_hat|^2 and the Gamma overlap matrix.

Test infrastructure only (see ``oracle/__init__.py``).
"""

import math

import numpy as np

# PyWavelets' db2 reconstruction low-pass filter (``pywt.Wavelet('db2').rec_lo``).
DB2_REC_LO = (
    0.48296291314469025,
    0.836516303737469,
    0.22414386804185735,
    -0.12940952255092145,
)


def _upsampling_convolution_full(inp, filt):
    """Zero-padded upsample-by-2 full convolution, pywt's ``rec_a`` step.

    out[2i+p] accumulates filt[2j+p]*inp[i-j] over ascending j, starting from 0.
    Output length 2N+F-2.
    """
    n = len(inp)
    f = len(filt)
    out = np.zeros(2 * n + f - 2, dtype=np.float64)
    for i in range(n + f // 2):
        lo = max(0, i - (n - 1))
        hi = min(i, f // 2 - 1)
        for j in range(lo, hi + 1):
            out[2 * i] += filt[2 * j] * inp[i - j]
            out[2 * i + 1] += filt[2 * j + 1] * inp[i - j]
    return out


def sample_scaling_function(wavelet_name="db2", level=10):
    """Follows wavelet_grid.py:10-14 (``wavefun(level)[0][:-1]``).

    ``wavefun`` for an orthogonal wavelet runs ``level`` cascade steps from the
    single coefficient sqrt(2)**level, keeps the central ``keep_length`` samples
    and pads one zero on the left and ``right_extent_length`` zeros on the right;
    for db2 at level 10 that is [0] + 3070 samples + [0, 0], and the reference
    drops the last sample.
    """
    if wavelet_name != "db2":
        raise ValueError("oracle supports only db2")
    filt = np.asarray(DB2_REC_LO, dtype=np.float64)
    coeffs = np.array([math.pow(math.sqrt(2.0), float(level))])
    for _ in range(level):
        coeffs = _upsampling_convolution_full(coeffs, filt)
    f_len = len(filt)
    p = 2 ** level
    output_length = (f_len - 1) * p + 1
    keep_length = 1
    for _ in range(level):
        keep_length = 2 * keep_length + (f_len - 2)
    if output_length - keep_length - 2 < 0:
        output_length = keep_length + 2
    right_extent = output_length - keep_length - 1
    if keep_length < len(coeffs):
        left = (len(coeffs) - keep_length) // 2
        coeffs = coeffs[left:left + keep_length]
    phi = np.concatenate(([0.0], coeffs, np.zeros(right_extent)))
    return phi[:-1]


def fourier_power_spectrum(values, k_min, k_max, k_count, phi_resolution):
    """Follows wavelet_grid.py:94-118 (dense DFT of the samples, squared)."""
    sample_count = values.shape[0]
    x_max = values.shape[0] / phi_resolution
    dx = (x_max - 0) / sample_count
    dk = (k_max - k_min) / k_count
    x_samples = np.arange(sample_count) * dx
    k_samples = np.arange(k_count + 1) * dk
    x_grid, k_grid = np.meshgrid(x_samples, k_samples)
    real_part = np.sum(values * dx * np.cos(-2 * np.pi * k_grid * x_grid), axis=1)
    imag_part = np.sum(values * dx * np.sin(-2 * np.pi * k_grid * x_grid), axis=1)
    return real_part ** 2 + imag_part ** 2


def cascade_power_closed_form(k, level=10):
    """|DTFT|^2 of the cascade samples at cycle frequency k (independent check).

    The cascade output after ``level`` steps has DTFT
    sqrt(2)^level * prod_j H(2^j w), H(w) = sum_n h_n e^{-iwn}, so with
    dx = 2^-level: P(k) = 2^-level * prod_{j<level} |H(2 pi k 2^j / 2^level)|^2.
    """
    h = np.asarray(DB2_REC_LO)
    n = np.arange(len(h))
    k = np.atleast_1d(np.asarray(k, dtype=np.float64))
    out = np.full(k.shape, 2.0 ** (-level))
    for j in range(level):
        w = 2 * np.pi * k * (2.0 ** j) / (2.0 ** level)
        hw = (h[None, :] * np.exp(-1j * w[:, None] * n[None, :])).sum(axis=1)
        out = out * np.abs(hw) ** 2
    return out


def cal_gamma(phi_array, phi_support, phi_resolution):
    """Follows corr3pcf_multipoles.py:22-30, including the np.roll wrap."""
    gamma = np.zeros((phi_support, phi_support))
    for l1 in range(phi_support):
        for l2 in range(phi_support):
            r1 = np.roll(phi_array, l1 * phi_resolution)
            r2 = np.roll(phi_array, l2 * phi_resolution)
            gamma[l1, l2] = np.sum(phi_array * r1 * r2) / phi_resolution
    return gamma
