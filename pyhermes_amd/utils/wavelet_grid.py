"""Scaling-function tables computed once on the host (tiny, O(3072 * L) work).

* ``sample_scaling_function`` replaces wavelet_grid.py:10-14 (PyWavelets'
  ``Wavelet(name).wavefun(level)[0][:-1]``) with an in-tree cascade, because
  PyWavelets is not a dependency of this package.
* ``fourier_power_spectrum`` follows wavelet_grid.py:94-118.
* ``cal_gamma`` follows corr3pcf_multipoles.py:22-30 (including its np.roll wrap).

The hot kernels that consume these tables run on the GPU (see ``pyhermes_amd.kernels``).
"""

import math

import numpy as np

# Reconstruction low-pass filters, as printed by PyWavelets for ``Wavelet(name).rec_lo``.
_REC_LO = {
    "db2": (0.48296291314469025, 0.836516303737469, 0.22414386804185735, -0.12940952255092145),
}


def _cascade_step(coeffs, filt):
    """One upsample-by-two full convolution with ``filt``.

    out[2i + p] = sum_j filt[2j + p] * coeffs[i - j], accumulated over ascending j.
    """
    n = coeffs.size
    half = filt.size // 2
    out = np.zeros(2 * n + filt.size - 2)
    idx = np.arange(n + half - 1)
    for j in range(half):
        src = idx - j
        valid = (src >= 0) & (src < n)
        term = np.where(valid, coeffs[np.clip(src, 0, n - 1)], 0.0)
        out[0::2] += filt[2 * j] * term
        out[1::2] += filt[2 * j + 1] * term
    return out


_PHI_CACHE = {}


def sample_scaling_function(wavelet_name="db2", level=10):
    """Scaling-function samples phi(i / 2^level), i = 0 .. (support * 2^level - 1).

    The cascade runs once per (wavelet, level) in a process; callers get their own copy.
    """
    key = (wavelet_name, int(level))
    if key not in _PHI_CACHE:
        _PHI_CACHE[key] = _cascade_samples(wavelet_name, int(level))
    return _PHI_CACHE[key].copy()


def _cascade_samples(wavelet_name, level):
    if wavelet_name not in _REC_LO:
        raise ValueError(f"Unsupported wavelet_mode '{wavelet_name}'. Supported: {sorted(_REC_LO)}.")
    filt = np.asarray(_REC_LO[wavelet_name], dtype=np.float64)
    coeffs = np.array([math.pow(math.sqrt(2.0), float(level))])
    for _ in range(level):
        coeffs = _cascade_step(coeffs, filt)
    keep = 1
    for _ in range(level):
        keep = 2 * keep + (filt.size - 2)
    out_len = max((filt.size - 1) * 2 ** level + 1, keep + 2)
    if keep < coeffs.size:
        left = (coeffs.size - keep) // 2
        coeffs = coeffs[left:left + keep]
    phi = np.concatenate(([0.0], coeffs, np.zeros(out_len - keep - 1)))
    return phi[:-1]


_POWER_CACHE = {}


def fourier_power_spectrum(values, k_min, k_max, k_count, phi_resolution):
    """|sum_i phi_i dx exp(-2 pi i k x_i)|^2 at k = k_min + j (k_max - k_min) / k_count, j = 0..k_count.

    A dense (k_count + 1) x len(phi) DFT on the host (about 30 ms at k_count = 512), so results are
    kept per (phi samples, k range, resolution); callers get their own copy.
    """
    values = np.ascontiguousarray(values, dtype=np.float64)
    key = (hash(values.tobytes()), values.shape[0], float(k_min), float(k_max), int(k_count), int(phi_resolution))
    hit = _POWER_CACHE.get(key)
    if hit is not None:
        return hit.copy()
    if len(_POWER_CACHE) > 64:
        _POWER_CACHE.clear()
    n = values.shape[0]
    dx = (n / phi_resolution) / n
    dk = (k_max - k_min) / k_count
    x = np.arange(n) * dx
    k = np.arange(k_count + 1) * dk
    xg, kg = np.meshgrid(x, k)
    re = np.sum(values * dx * np.cos(-2 * np.pi * kg * xg), axis=1)
    im = np.sum(values * dx * np.sin(-2 * np.pi * kg * xg), axis=1)
    out = re ** 2 + im ** 2
    _POWER_CACHE[key] = out
    return out.copy()


def cal_gamma(phi_array, phi_support, phi_resolution):
    """Triple-overlap weights Gamma[l1, l2] = sum_i phi_i phi[i - l1 R] phi[i - l2 R] / R (indices mod len)."""
    n = phi_array.shape[0]
    i = np.arange(n)
    gamma = np.zeros((phi_support, phi_support))
    for l1 in range(phi_support):
        for l2 in range(phi_support):
            gamma[l1, l2] = np.sum(phi_array * phi_array[(i - l1 * phi_resolution) % n]
                                   * phi_array[(i - l2 * phi_resolution) % n]) / phi_resolution
    return gamma
