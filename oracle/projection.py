"""Oracle: catalogue -> scaling-function coefficients.  Test infrastructure only."""

import numpy as np

from . import native


def projection_weights(catalog_weight, field_value, weight_normalization="catalog"):
    """Follows sfc_projection.py:251-282 (FP64 sums and projection weights).

    Returns (projection_weight, sums) with sums = dict(catalog_weight_sum,
    catalog_weight_sq_sum, raw_field_weighted_sum, field_integral).
    """
    cw = np.asarray(catalog_weight, dtype=np.float64)
    fv = np.asarray(field_value, dtype=np.float64)
    cw_sum = float(np.sum(cw, dtype=np.float64))
    cw_sq = float(np.sum(cw ** 2, dtype=np.float64))
    raw = float(np.sum(cw * fv, dtype=np.float64))
    mode = {"unit": "field"}.get(weight_normalization, weight_normalization)
    if mode == "raw":
        den = 1.0
    elif mode == "catalog":
        den = cw_sum
    else:
        den = raw
    return cw * fv / den, {
        "catalog_weight_sum": cw_sum,
        "catalog_weight_sq_sum": cw_sq,
        "raw_field_weighted_sum": raw,
        "field_integral": raw / den,
    }


def cell_and_frac(pos, J, box, phi_res=1024):
    """Follows wavelet_grid.py:30-34: int32 floor cell and truncated fraction index."""
    sf = (1 << J) / box
    scaled = np.asarray(pos, dtype=np.float64)[:, :3] * sf
    cell = np.floor(scaled).astype(np.int32)
    frac = ((scaled - cell) * phi_res).astype(np.int32)
    return cell, frac


def project_c(pos, w, phi, J, box, phi_res=1024):
    """wavelet_grid.py:17-70 via the serial C restatement.  Returns (eps, cell, frac)."""
    pos = native.f64(pos)
    w = native.f64(w)
    phi = native.f64(phi)
    L = 1 << J
    n = pos.shape[0]
    eps = np.empty((L, L, L), dtype=np.float64)
    cell = np.empty((n, 3), dtype=np.int32)
    frac = np.empty((n, 3), dtype=np.int32)
    native.lib().oracle_project(native.ptr(pos), native.ptr(w), n, native.ptr(phi), len(phi), phi_res,
                                J, float(box), native.ptr(eps), native.ptr(cell, native.ctypes.c_int32),
                                native.ptr(frac, native.ctypes.c_int32))
    return eps, cell, frac


def project_bruteforce(pos, w, phi, J, box, phi_res=1024):
    """Independent numpy restatement with np.add.at and explicit modular wrap (tiny inputs)."""
    L = 1 << J
    support = len(phi) // phi_res
    cell, frac = cell_and_frac(pos, J, box, phi_res)
    eps = np.zeros((L, L, L))
    w = np.asarray(w, dtype=np.float64)
    for sx in range(support):
        for sy in range(support):
            for sz in range(support):
                val = (w * phi[frac[:, 0] + phi_res * sx] * phi[frac[:, 1] + phi_res * sy]
                       * phi[frac[:, 2] + phi_res * sz])
                np.add.at(eps, ((cell[:, 0] - sx) % L, (cell[:, 1] - sy) % L, (cell[:, 2] - sz) % L), val)
    return eps
