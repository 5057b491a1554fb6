"""Oracle: box-random Monte-Carlo triplet estimator.  Test infrastructure only."""

import ctypes

import numpy as np

from . import mt19937, native


def rotation_uniforms(seed_base_rot, n_rot, mu_index):
    """corr3pcf_kernels.py:119-125: per rotation ``np.random.seed(s); rand(3)`` inside the Numba kernel,
    restated with Numba's generator (MT19937 init_genrand + genrand_res53, oracle/mt19937.py), seed
    s = seed_base_rot + irot (+ mu_index * 1000003, corr3pcf.py:1462)."""
    out = np.empty((n_rot, 3))
    for irot in range(n_rot):
        s = seed_base_rot + irot
        if mu_index >= 0:
            s += mu_index * 1000003
        out[irot] = mt19937.uniforms(s, 3)
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


def philox4x32_10(counter, key):
    """Philox4x32-10 block (Salmon et al., SC'11): 4 counter words, 2 key words -> 4 words."""
    c = np.ascontiguousarray(counter, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.empty(4, dtype=np.uint32)
    native.lib().oracle_philox4x32_10(native.ptr(c, ctypes.c_uint32), native.ptr(k, ctypes.c_uint32),
                                      native.ptr(out, ctypes.c_uint32))
    return out


def binned_sampler(g1, g2, g3, phi, nbins, rrange1, rrange2, n, seed, tables=None, phi_res=1024):
    """Binned triangle sampler (plain, or importance with ``tables``).

    Restates the sampler over the reference's triangle geometry
    (corr3pcf_kernels.py:21-63) and interpolation (wavelet_grid.py:143-186):
    per bin sum(w f), sum((w f)^2) and the triangle count.  ``tables`` holds
    cell_bucket (u8), bucket_cells (i32), bucket_off (i64), bucket_cdf (f32),
    bucket_density (f64), bin_cdf (f32), bin_pdf (f32), nbucket, alpha and
    optionally sub (sub-cells per cell and axis) and anchors (1 or 3).
    Returns three (nr1, nr2, nmu) arrays.
    """
    L = g1.shape[0]
    nr1, nr2, nmu = (int(v) for v in nbins)
    B = nr1 * nr2 * nmu
    g = [native.f64(a) for a in (g1, g2, g3)]
    phi = native.f64(phi)
    s1, s2 = np.zeros(B), np.zeros(B)
    cnt = np.zeros(B, dtype=np.int64)
    if tables is None:
        tp = [None] * 7
        nbucket, alpha, sub, anchors = 0, 1.0, 1, 1
    else:
        arrs = [np.ascontiguousarray(tables[k], dtype=d) for k, d in
                (("cell_bucket", np.uint8), ("bucket_cells", np.int32), ("bucket_off", np.int64),
                 ("bucket_cdf", np.float32), ("bucket_density", np.float64), ("bin_cdf", np.float32),
                 ("bin_pdf", np.float32))]
        ctys = (ctypes.c_uint8, ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double, ctypes.c_float,
                ctypes.c_float)
        tp = [native.ptr(a, t) for a, t in zip(arrs, ctys)]
        nbucket, alpha = int(tables["nbucket"]), float(tables["alpha"])
        sub, anchors = int(tables.get("sub", 1)), int(tables.get("anchors", 1))
    native.lib().oracle_mc_binned(native.ptr(g[0]), native.ptr(g[1]), native.ptr(g[2]), native.ptr(phi), L, phi_res,
                                  nr1, nr2, nmu, float(rrange1[0]), float(rrange1[1]), float(rrange2[0]),
                                  float(rrange2[1]), int(n), int(seed) & 0xFFFFFFFFFFFFFFFF, *tp, nbucket, alpha,
                                  sub, anchors, native.ptr(s1), native.ptr(s2), native.ptr(cnt, ctypes.c_int64))
    shape = (nr1, nr2, nmu)
    return s1.reshape(shape), s2.reshape(shape), cnt.reshape(shape)


def bucketize(grid, phi, sub, edges, phi_res=1024):
    """Proposal buckets on the (L sub)^3 sub-cell grid: bucket = number of edges <= the
    interpolated field (wavelet_grid.py:143-186) at the sub-cell centre; the sub-cells
    listed bucket by bucket (stable).  Returns (cell_bucket, bucket_cells, bucket_off)."""
    L = grid.shape[0]
    Ls = L * int(sub)
    idx = np.arange(Ls ** 3, dtype=np.int64)
    xs, ys, zs = idx // (Ls * Ls), (idx // Ls) % Ls, idx % Ls
    centres = (np.stack([xs, ys, zs], axis=1) + 0.5) / sub
    v = interpolate(grid, centres, phi, phi_res=phi_res)
    cb = np.searchsorted(np.asarray(edges, dtype=np.float64), v, side="right").astype(np.uint8)
    cells = np.argsort(cb, kind="stable").astype(np.int32)
    counts = np.bincount(cb, minlength=len(edges) + 1)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return cb, cells, off, v


def triplet_particle_centers(R1, R2, mu, centers, weights, rho1, abc, e2, e3, phi, phi_res=1024):
    """corr3pcf_kernels.py:66-97 with the rotation uniforms supplied: the first vertex sits on each centre
    with its weight, rho1 sum_rot sum_i w_i n2(c_i + d12) n3(c_i + d13) / (n_rot sum_i w_i)."""
    c = np.asarray(centers, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)
    total = 0.0
    for a, b, g in np.asarray(abc, dtype=np.float64):
        o = triangle_offsets(R1, R2, mu, 2.0 * np.pi * a, 2.0 * b - 1.0, 2.0 * np.pi * g)
        n2 = interpolate(e2, c, phi, offset=(o[0], o[1], o[2]), phi_res=phi_res)
        n3 = interpolate(e3, c, phi, offset=(o[3], o[4], o[5]), phi_res=phi_res)
        total += float(np.sum(w * n2 * n3))
    return rho1 * total / (len(abc) * float(np.sum(w)))
