"""Oracle: box-random Monte-Carlo triplet estimator.  Test infrastructure only."""

import numpy as np

from . import native


def rotation_uniforms(seed_base_rot, n_rot, mu_index):
    """corr3pcf_kernels.py:119-125: per rotation ``np.random.seed(s); rand(3)``.

    Numba's ``np.random.seed``/``rand`` is believed to reproduce numpy's legacy
    MT19937 ``RandomState(s).rand(3)``; that equality is parity unpinned (only
    an import of numba could confirm it).
    """
    out = np.empty((n_rot, 3))
    for irot in range(n_rot):
        s = seed_base_rot + irot
        if mu_index >= 0:
            s += mu_index * 1000003
        out[irot] = np.random.RandomState(s).rand(3)
    return out


def box_centers(n, L, seed):
    """sampling.py:6-10 via corr3pcf.py:1415-1416 (grid units)."""
    return np.random.default_rng(seed=seed).uniform(0.0, L, size=(n, 3))


def interpolate(grid, spos, phi, offset=(0.0, 0.0, 0.0), phi_res=1024):
    """wavelet_grid.py:143-186."""
    L = grid.shape[0]
    spos = native.f64(spos)
    g = native.f64(grid)
    phi = native.f64(phi)
    out = np.empty(spos.shape[0])
    native.lib().oracle_interpolate(native.ptr(out), native.ptr(spos), spos.shape[0], native.ptr(g),
                                    native.ptr(phi), L, phi_res, len(phi) // phi_res,
                                    float(offset[0]), float(offset[1]), float(offset[2]))
    return out


def interpolate_bruteforce(grid, spos, phi, offset=(0.0, 0.0, 0.0), phi_res=1024):
    """Independent numpy evaluation of n(x) = sum_27 eps[(c - s) & mask] phi phi phi."""
    L = grid.shape[0]
    sp = np.asarray(spos, dtype=np.float64) + np.asarray(offset)[None, :]
    cell = np.floor(sp).astype(np.int64)
    frac = ((sp - cell) * phi_res).astype(np.int64)
    support = len(phi) // phi_res
    out = np.zeros(sp.shape[0])
    for a in range(support):
        for b in range(support):
            for c in range(support):
                out += (grid[(cell[:, 0] - a) % L, (cell[:, 1] - b) % L, (cell[:, 2] - c) % L]
                        * phi[frac[:, 0] + a * phi_res] * phi[frac[:, 1] + b * phi_res]
                        * phi[frac[:, 2] + c * phi_res])
    return out


def triangle_offsets(R1, R2, mu, phi, ct1, alpha):
    """corr3pcf_kernels.py:21-63."""
    o = np.empty(6)
    native.lib().oracle_triangle_offsets(float(R1), float(R2), float(mu), float(phi), float(ct1),
                                         float(alpha), native.ptr(o))
    return o


def triplet_box_random(R1, R2, mu, centers, abc, e1, e2, e3, phi, phi_res=1024):
    """corr3pcf_kernels.py:100-140 with the rotation uniforms supplied."""
    c = native.f64(centers)
    abc = native.f64(abc)
    e1, e2, e3 = native.f64(e1), native.f64(e2), native.f64(e3)
    phi = native.f64(phi)
    return native.lib().oracle_triplet_box_random(
        float(R1), float(R2), float(mu), native.ptr(c), c.shape[0], abc.shape[0], native.ptr(abc),
        native.ptr(e1), native.ptr(e2), native.ptr(e3), native.ptr(phi), e1.shape[0], phi_res,
        len(phi) // phi_res)
