"""Seeded lognormal mock catalogues (synthetic inputs; the reference ships no generator).

Recipe: a Gaussian random field G on an Ng^3 grid with P(k) = A k^-1.5
(k = 2 pi |n| / box, P(0) = 0) drawn as white noise times sqrt(P) in Fourier
space; G is rescaled to var(G) = ln(1 + sigma_LN^2); rho = exp(G - var/2);
per-cell counts are multinomial(N, rho / sum rho) (exact N); positions are
(cell + U[0,1)^3) * box / Ng in float64.  Seeds: base + 1 (field), + 2 (counts),
+ 3 (in-cell offsets), + 4 (RSD / weights).
"""

import numpy as np


def gaussian_field(ng, box, seed, slope=-1.5, sigma_ln=1.0):
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((ng, ng, ng))
    wk = np.fft.rfftn(white)
    n = np.fft.fftfreq(ng) * ng
    nz = np.fft.rfftfreq(ng) * ng
    k = 2 * np.pi / box * np.sqrt(n[:, None, None] ** 2 + n[None, :, None] ** 2 + nz[None, None, :] ** 2)
    pk = np.zeros_like(k)
    pk[k > 0] = k[k > 0] ** slope
    g = np.fft.irfftn(wk * np.sqrt(pk), s=(ng, ng, ng), axes=(0, 1, 2))
    g -= g.mean()
    target = np.log(1.0 + sigma_ln ** 2)
    g *= np.sqrt(target / g.var())
    return g, target


def lognormal_mock(n_particles, ng, box, seed_base, sigma_ln=1.0, rsd_sigma_kms=None, weight_sigma=None):
    """Returns (pos (N,3) f64, weights (N,) f64 or None)."""
    g, var = gaussian_field(ng, box, seed_base + 1, sigma_ln=sigma_ln)
    rho = np.exp(g - var / 2.0).ravel()
    rho /= rho.sum()
    rng_c = np.random.default_rng(seed_base + 2)
    counts = rng_c.multinomial(int(n_particles), rho)
    cells = np.repeat(np.arange(ng ** 3, dtype=np.int64), counts)
    ix, iy, iz = np.unravel_index(cells, (ng, ng, ng))
    rng_u = np.random.default_rng(seed_base + 3)
    pos = (np.stack([ix, iy, iz], axis=1) + rng_u.random((cells.size, 3))) * (box / ng)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    weights = None
    rng_w = np.random.default_rng(seed_base + 4)
    if rsd_sigma_kms is not None:
        # plane-parallel redshift-space shift along z with H = 100 h km/s/Mpc, z = 0
        v = rng_w.normal(0.0, rsd_sigma_kms, cells.size)
        pos[:, 2] = np.mod(pos[:, 2] + v / 100.0, box)
        pos[pos[:, 2] >= box, 2] = 0.0
    if weight_sigma is not None:
        s2 = np.log(1.0 + weight_sigma ** 2)
        weights = np.exp(-s2 / 2.0 + np.sqrt(s2) * rng_w.standard_normal(cells.size))
    return pos, weights


def uniform_catalogue(n_particles, box, seed):
    return np.random.default_rng(seed).uniform(0.0, box, size=(int(n_particles), 3))
