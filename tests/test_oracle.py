"""Pins the CPU oracle with invariants, closed forms, library routines and brute force.

These run without a GPU.  A plausible mistake in the oracle (a dropped term, a
wrong sign or index, a transposed operand) fails at least one of them.
"""

import numpy as np
import pytest
import scipy.special as sps

from oracle import mc, multipole, projection, scaling, windows


# ---------------------------------------------------------------- phi / P / Gamma
def test_phi_invariants(phi):
    assert phi.shape == (3072,)
    assert phi[0] == 0.0 and phi[-1] == 0.0
    assert abs(phi.sum() - 1024.0) < 1e-11
    pu = phi.reshape(3, 1024).sum(axis=0)
    np.testing.assert_allclose(pu, 1.0, atol=1e-14)
    # db2 scaling function peak value (Daubechies 1988 tables give ~1.3660)
    assert abs(phi.max() - 1.3660) < 1e-3


def test_power_spectrum_matches_cascade_closed_form(phi):
    L = 64
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    closed = scaling.cascade_power_closed_form(np.arange(L + 1) / L)
    np.testing.assert_allclose(P, closed, rtol=1e-11, atol=1e-15)
    assert abs(P[0] - 1.0) < 1e-12
    assert P[L] < 1e-25


def test_gamma_bruteforce_and_partition(phi):
    g = scaling.cal_gamma(phi, 3, 1024)
    n = len(phi)
    i = np.arange(n)
    brute = np.zeros((3, 3))
    for l1 in range(3):
        for l2 in range(3):
            brute[l1, l2] = np.sum(phi * phi[(i - 1024 * l1) % n] * phi[(i - 1024 * l2) % n]) / 1024
    np.testing.assert_allclose(g, brute, rtol=0, atol=1e-15)
    np.testing.assert_allclose(g, g.T, atol=1e-16)
    assert abs(g.sum() - 1.0) < 1e-12


# ---------------------------------------------------------------- projection
def _positions(rng, n, box):
    pos = rng.uniform(0, box, size=(n, 3))
    pos[:5] = rng.uniform(0, box * 1e-3, size=(5, 3))  # exercise the negative-index wrap
    return pos


@pytest.mark.parametrize("J", [3, 4])
def test_projection_c_matches_bruteforce(phi, J):
    rng = np.random.default_rng(11 + J)
    box = 100.0
    pos = _positions(rng, 300, box)
    w = rng.lognormal(0, 0.3, 300)
    eps_c, cell, frac = projection.project_c(pos, w, phi, J, box)
    eps_b = projection.project_bruteforce(pos, w, phi, J, box)
    np.testing.assert_allclose(eps_c, eps_b, rtol=0, atol=1e-13 * np.abs(eps_b).max())
    assert abs(eps_c.sum() - w.sum()) < 1e-11 * w.sum()
    c2, f2 = projection.cell_and_frac(pos, J, box)
    np.testing.assert_array_equal(cell, c2)
    np.testing.assert_array_equal(frac, f2)


def test_projection_cell_frac_closed_form():
    # positions on exact binary fractions: with box == L, scale_factor == 1
    J = 4
    L = 1 << J
    rng = np.random.default_rng(3)
    c = rng.integers(0, L, size=(200, 3))
    f = rng.integers(0, 1024, size=(200, 3))
    pos = c + f / 1024.0 + 2.0 ** -20
    cell, frac = projection.cell_and_frac(pos, J, float(L))
    np.testing.assert_array_equal(cell, c)
    np.testing.assert_array_equal(frac, f)


def test_projection_whole_cell_shift_is_roll(phi):
    J, box = 4, 64.0
    L = 1 << J
    rng = np.random.default_rng(5)
    pos = rng.uniform(4 * box / L, box - 4 * box / L, size=(200, 3))
    w = rng.uniform(0.5, 1.5, 200)
    e0 = projection.project_bruteforce(pos, w, phi, J, box)
    e1 = projection.project_bruteforce(pos + np.array([1, 0, 2]) * (box / L), w, phi, J, box)
    np.testing.assert_allclose(e1, np.roll(e0, (1, 0, 2), axis=(0, 1, 2)), atol=1e-13)


def test_projection_weights_modes():
    cw = np.array([1.0, 2.0, 3.0])
    fv = np.array([2.0, 1.0, 0.5])
    pw, s = projection.projection_weights(cw, fv, "catalog")
    np.testing.assert_allclose(pw, cw * fv / 6.0)
    assert s["field_integral"] == pytest.approx(5.5 / 6.0)
    pw, s = projection.projection_weights(cw, fv, "unit")
    assert s["field_integral"] == pytest.approx(1.0)
    pw, s = projection.projection_weights(cw, fv, "raw")
    np.testing.assert_allclose(pw, cw * fv)


# ---------------------------------------------------------------- special functions
@pytest.mark.parametrize("l", list(range(0, 13)))
def test_spherical_jn_against_scipy(l):
    xs = np.concatenate([[1e-3, 0.02, 0.05, 0.06], np.linspace(0.1, l + 1.5, 17), [20.0, 77.7, 300.0, 650.0]])
    for x in xs:
        ours = windows.spherical_jn(l, x)
        ref = sps.spherical_jn(l, x)
        assert abs(ours - ref) <= 1e-9 * max(abs(ref), 1e-12) + 1e-15, (l, x, ours, ref)


def test_sph_harm_against_scipy():
    rng = np.random.default_rng(0)
    for _ in range(40):
        k = rng.normal(size=3)
        theta = np.arccos(k[2] / np.linalg.norm(k))
        ph = np.arctan2(k[1], k[0])
        for l in range(0, 11):
            for m in range(-l, l + 1):
                ours = windows.sph_harm(l, m, *k)
                ref = sps.sph_harm_y(l, m, theta, ph)
                assert abs(ours - ref) < 1e-12 * max(1.0, abs(ref)), (l, m)
    assert windows.sph_harm(0, 0, 0, 0, 0) == pytest.approx(1 / np.sqrt(4 * np.pi))
    assert windows.sph_harm(2, 1, 0, 0, 0) == 0


# ---------------------------------------------------------------- window kernels
@pytest.mark.parametrize("wt,args", [("sphere", (2.5, 0)), ("gaussian", (1.5, 0)), ("shell", (3.0, 0)),
                                      ("thick_shell", (3.0, 1.5)), ("gaussian_shell", (3.0, 1.0)),
                                      ("inverse_laplacian", (0, 0))])
def test_octant_fold_equals_explicit_alias_builder(phi, wt, args):
    L = 8
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    a = windows.octant_rfft_kernel(L, P, wt, *args)
    b = windows.full_rfft_kernel(L, P, wt, *args)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15)
    if wt != "inverse_laplacian":
        assert windows.native.lib().oracle_window_eval(windows.WINDOW_CODES[wt], 0, 0, 0, *args) == 1.0


def test_multipole_l0_equals_shell_kernel(phi):
    L = 8
    R = 2.3
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    K00 = windows.radial_multipole_window(L, P, R, 0, 0)
    Ks = windows.full_rfft_kernel(L, P, "shell", R)
    np.testing.assert_allclose(K00[:, :, :L // 2 + 1].real * np.sqrt(4 * np.pi), Ks, rtol=1e-12, atol=1e-15)
    assert np.abs(K00.imag).max() < 1e-15


def test_multipole_conjugation_symmetry(phi):
    L = 8
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    neg = (-np.arange(L)) % L
    for l, m in [(1, 1), (2, 1), (3, 2), (4, 0)]:
        Kp = windows.radial_multipole_window(L, P, 2.7, l, m)
        Km = windows.radial_multipole_window(L, P, 2.7, l, -m)
        Kp_neg = Kp[neg][:, neg][:, :, neg]
        np.testing.assert_allclose(Km, (-1) ** (l + m) * np.conj(Kp_neg), atol=1e-14)


# ---------------------------------------------------------------- convolution / pair product
def test_pair_product_parseval(phi):
    L = 8
    rng = np.random.default_rng(1)
    f1 = rng.normal(size=(L, L, L))
    f2 = rng.normal(size=(L, L, L))
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    K = windows.octant_rfft_kernel(L, P, "shell", 2.0)
    a = windows.pair_product(f1, f2, K)
    b = windows.pair_product_parseval(f1, f2, K)
    assert abs(a - b) < 1e-12 * max(1.0, abs(b))


def test_convolution_matches_realspace_circular(phi):
    L = 8
    rng = np.random.default_rng(2)
    f = rng.normal(size=(L, L, L))
    w = rng.normal(size=(L, L, L))
    K = np.fft.rfftn(w)
    conv = windows.convolve_rfft(f, K)
    brute = np.zeros_like(f)
    for s in np.ndindex(L, L, L):
        brute += w[s] * np.roll(f, s, axis=(0, 1, 2))
    np.testing.assert_allclose(conv, brute, atol=1e-10)


# ---------------------------------------------------------------- multipole contraction
def test_summand_matches_bruteforce(phi):
    L = 8
    rng = np.random.default_rng(4)
    d = rng.normal(size=(L, L, L))
    F1 = rng.normal(size=(L, L, L)) + 1j * rng.normal(size=(L, L, L))
    F2 = rng.normal(size=(L, L, L)) + 1j * rng.normal(size=(L, L, L))
    g = scaling.cal_gamma(phi, 3, 1024)
    a = multipole.summand_c(d, F1, F2, g)
    b = multipole.summand_bruteforce(d, F1, F2, g)
    assert abs(a - b) < 1e-11 * abs(b)


def test_summand_constant_closed_form(phi):
    L = 8
    g = scaling.cal_gamma(phi, 3, 1024)
    d = np.full((L, L, L), 0.7)
    F1 = np.full((L, L, L), 1.5 - 0.5j)
    F2 = np.full((L, L, L), -0.25 + 2.0j)
    v = multipole.m_summand(d, F1, F2, g)
    expect = 4 * np.pi * 0.7 * (1.5 - 0.5j) * (-0.25 + 2.0j)
    assert abs(v - expect) < 1e-12 * abs(expect)


def test_combine_m_terms_closed_form():
    vals = np.array([1.0 + 2.0j, 0.5 - 1.0j, 0.25 + 0.75j])
    got = multipole.combine_multipole_m_terms(vals, 2)
    expect = (vals[0] - 2 * vals[1].real + 2 * vals[2].real).real
    assert got == pytest.approx(expect)


def test_ratio_solve_against_legendre_product():
    from numpy.polynomial import legendre as leg
    rng = np.random.default_rng(9)
    lmax = 6
    B = np.concatenate([[2.0], rng.normal(size=lmax) * 0.3])
    C = rng.normal(size=lmax + 1)
    A = leg.legmul(B, C)[:lmax + 1]
    M = multipole.build_mixing_matrix(B, lmax)
    np.testing.assert_allclose(M @ C, A, rtol=1e-12, atol=1e-12)
    sol, _, _ = multipole.solve_multipoles_from_ratio(A, B, lmax)
    np.testing.assert_allclose(sol, C, rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------- Monte Carlo
def test_interpolation_matches_bruteforce(phi):
    L = 16
    rng = np.random.default_rng(6)
    grid = rng.normal(size=(L, L, L))
    sp = rng.uniform(0, L, size=(500, 3))
    a = mc.interpolate(grid, sp, phi, (1.3, -2.7, 0.4))
    b = mc.interpolate_bruteforce(grid, sp, phi, (1.3, -2.7, 0.4))
    np.testing.assert_allclose(a, b, atol=1e-13)
    const = mc.interpolate(np.full((L, L, L), 2.5), sp, phi)
    np.testing.assert_allclose(const, 2.5, atol=1e-12)


def test_triangle_offsets_geometry():
    rng = np.random.default_rng(8)
    for _ in range(50):
        R1, R2, mu = rng.uniform(1, 10), rng.uniform(1, 10), rng.uniform(-1, 1)
        o = mc.triangle_offsets(R1, R2, mu, rng.uniform(0, 2 * np.pi), rng.uniform(-1, 1),
                                rng.uniform(0, 2 * np.pi))
        a, b = o[:3], o[3:]
        assert np.linalg.norm(a) == pytest.approx(R1, rel=1e-12)
        assert np.linalg.norm(b) == pytest.approx(R2, rel=1e-12)
        assert a @ b / (R1 * R2) == pytest.approx(mu, abs=1e-12)


def test_triplet_estimator_constant_and_bruteforce(phi):
    L = 16
    rng = np.random.default_rng(10)
    centers = rng.uniform(0, L, size=(200, 3))
    abc = mc.rotation_uniforms(43, 5, 2)
    c = [np.full((L, L, L), v) for v in (1.5, 2.0, 0.5)]
    est = mc.triplet_box_random(3.0, 4.0, 0.3, centers, abc, *c, phi)
    assert est == pytest.approx(1.5, rel=1e-12)
    e = [rng.normal(size=(L, L, L)) for _ in range(3)]
    est = mc.triplet_box_random(3.0, 4.0, 0.3, centers, abc, *e, phi)
    n1 = mc.interpolate_bruteforce(e[0], centers, phi)
    tot = 0.0
    for r in range(abc.shape[0]):
        o = mc.triangle_offsets(3.0, 4.0, 0.3, 2 * np.pi * abc[r, 0], 2 * abc[r, 1] - 1, 2 * np.pi * abc[r, 2])
        tot += np.sum(n1 * mc.interpolate_bruteforce(e[1], centers, phi, o[:3])
                      * mc.interpolate_bruteforce(e[2], centers, phi, o[3:]))
    assert est == pytest.approx(tot / (abc.shape[0] * centers.shape[0]), rel=1e-11)


def test_mlp_gradient_finite_differences():
    """The oracle MLP gradient equals central differences of its mean loss (clip inactive)."""
    from oracle import mlp as omlp
    g = np.random.default_rng(11)
    n_par = omlp.H * omlp.IN + omlp.H + omlp.H * omlp.H + omlp.H + omlp.H + 1
    w = g.normal(0, 0.3, n_par)
    x = g.normal(size=(40, omlp.IN))
    y = g.exponential(1.0, 40)

    def loss(wv):
        s, _ = omlp.forward(wv, x)
        return float(np.mean(y * np.exp(-s) + s))

    gr, l0 = omlp.grad(w, x, y, clip=1e9)
    assert abs(l0 - loss(w)) < 1e-12
    idx = g.choice(n_par, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(n_par)
        e[i] = h
        fd = (loss(w + e) - loss(w - e)) / (2 * h)
        assert abs(fd - gr[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, gr[i])
    # bf16 rounding is round-to-nearest-even on the top 16 bits
    v = np.array([1.0, 1.00390625, 1.005859375, -3.1415927], dtype=np.float32)
    r = omlp.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125) and r[3] == np.float32(-3.140625)


# ---------------------------------------------------------------- binned MC sampler
_M64 = (1 << 64) - 1


def test_philox_known_answer_vectors():
    """Random123's published philox4x32-10 known-answer vectors."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert [int(v) for v in mc.philox4x32_10(ctr, key)] == want


def _u2(i, seed, stream):
    key = (seed ^ ((0x9E3779B97F4A7C15 * (stream + 1)) & _M64)) & _M64
    r = [int(v) for v in mc.philox4x32_10([i & 0xFFFFFFFF, i >> 32, 0x9E3779B9, 0x85EBCA6B],
                                          [key & 0xFFFFFFFF, key >> 32])]
    return (((r[0] >> 5) * 67108864 + (r[1] >> 6)) / 2.0 ** 53, ((r[2] >> 5) * 67108864 + (r[3] >> 6)) / 2.0 ** 53)


def test_binned_sampler_matches_per_sample_bruteforce(phi):
    """Each sample rebuilt from its Philox streams with brute-force interpolation."""
    L, n, seed = 16, 40, 987654321
    rng = np.random.default_rng(5)
    g = [rng.normal(size=(L, L, L)) for _ in range(3)]
    nr1, nr2, nmu, r1r, r2r = 3, 2, 4, (1.5, 5.0), (2.0, 6.5)
    s1, s2, cnt = mc.binned_sampler(*g, phi, (nr1, nr2, nmu), r1r, r2r, n, seed)
    B = nr1 * nr2 * nmu
    e1, e2, e3 = np.zeros(B), np.zeros(B), np.zeros(B, dtype=np.int64)
    for i in range(n):
        ux, uy = _u2(i, seed, 0)
        uz, ubin = _u2(i, seed, 1)
        w1 = _u2(i, seed, 3)[1]
        w2, w3 = _u2(i, seed, 4)
        oph, oct_ = _u2(i, seed, 5)
        oal = _u2(i, seed, 6)[0]
        b = min(int(ubin * B), B - 1)
        imu, ir2, ir1 = b % nmu, (b // nmu) % nr2, b // (nmu * nr2)
        r1 = r1r[0] + (ir1 + w1) * (r1r[1] - r1r[0]) / nr1
        r2 = r2r[0] + (ir2 + w2) * (r2r[1] - r2r[0]) / nr2
        mu = -1.0 + (imu + w3) * 2.0 / nmu
        o = mc.triangle_offsets(r1, r2, mu, 2 * np.pi * oph, 2 * oct_ - 1, 2 * np.pi * oal)
        assert abs(np.dot(o[:3], o[3:]) - r1 * r2 * mu) < 1e-9 * r1 * r2
        c = np.array([[ux * L, uy * L, uz * L]])
        f = (mc.interpolate_bruteforce(g[0], c, phi)[0] * mc.interpolate_bruteforce(g[1], c + o[:3], phi)[0]
             * mc.interpolate_bruteforce(g[2], c + o[3:], phi)[0])
        e1[b] += f
        e2[b] += f * f
        e3[b] += 1
    assert np.array_equal(cnt.ravel(), e3)
    np.testing.assert_allclose(s1.ravel(), e1, rtol=1e-12, atol=1e-12 * np.abs(e1).max())
    np.testing.assert_allclose(s2.ravel(), e2, rtol=1e-12)


def test_binned_sampler_uniform_bins_and_constant_field(phi):
    """Plain draws: bin counts uniform (chi-square); a constant field gives f = c^3 exactly."""
    L, n, c = 16, 200000, 1.3
    g = np.full((L, L, L), c)
    s1, s2, cnt = mc.binned_sampler(g, g, g, phi, (3, 4, 5), (1.0, 6.0), (1.0, 6.0), n, 17)
    B = cnt.size
    assert cnt.sum() == n
    chi2 = (((cnt - n / B) ** 2) / (n / B)).sum()
    assert chi2 < B + 6 * np.sqrt(2 * B), chi2
    np.testing.assert_allclose(s1, cnt * c ** 3, rtol=1e-12)


def _synthetic_tables(L, nbins, nbucket, alpha, seed, sub=1, anchors=1):
    """Importance tables from an arbitrary bucket assignment and proposal mass."""
    rng = np.random.default_rng(seed)
    B = int(np.prod(nbins))
    Ls = L * sub
    cell_bucket = rng.integers(0, nbucket, size=Ls ** 3).astype(np.uint8)
    order = np.argsort(cell_bucket, kind="stable").astype(np.int32)
    n_k = np.bincount(cell_bucket, minlength=nbucket).astype(np.float64)
    off = np.concatenate([[0], np.cumsum(n_k)]).astype(np.int64)
    Q = rng.uniform(0.05, 1.0, size=(nbucket, B)) * rng.uniform(0.1, 3.0, size=(nbucket, 1))
    Q /= Q.sum()
    P = Q.sum(axis=1)
    pdf = Q / P[:, None]
    cdf = np.cumsum(P).astype(np.float32)
    cdf[-1] = 2.0
    bcdf = np.cumsum(pdf, axis=1).astype(np.float32)
    bcdf[:, -1] = 2.0
    return {"cell_bucket": cell_bucket, "bucket_cells": order, "bucket_off": off, "bucket_cdf": cdf,
            "bucket_density": float(Ls) ** 3 * B * P / n_k, "bin_cdf": bcdf, "bin_pdf": pdf.astype(np.float32),
            "nbucket": nbucket, "alpha": alpha, "sub": sub, "anchors": anchors}


@pytest.mark.parametrize("sub,anchors", [(1, 1), (2, 3), (4, 3), (2, 1)])
def test_importance_weights_are_unbiased_on_constant_field(phi, sub, anchors):
    """E_q[w 1{bin b}] = 1/B for any proposal: per-bin (B/n) sum w f -> c^3 within the errors."""
    L, n, c = 16, 400000, 0.9
    nbins = (2, 3, 4)
    tab = _synthetic_tables(L, nbins, 6, 0.25, 3, sub, anchors)
    g = np.full((L, L, L), c)
    s1, s2, cnt = mc.binned_sampler(g, g, g, phi, nbins, (1.0, 6.0), (1.0, 6.0), n, 23, tables=tab)
    B = cnt.size
    est = B / n * s1.ravel()
    var = (B / n) ** 2 * (s2.ravel() - s1.ravel() ** 2 / n)
    z = (est - c ** 3) / np.sqrt(var)
    assert np.abs(z).max() < 5.0, z
    assert 0.5 < np.sqrt(np.mean(z ** 2)) < 1.6
    # the proposal really moved the counts away from uniform
    assert cnt.max() > 1.5 * cnt.min()


def test_anchored_proposal_is_unbiased_on_clustered_field(phi):
    """Three-anchor importance on sub-cells against plain draws on a lognormal field."""
    L, n = 16, 600000
    nbins, rr = (2, 2, 3), (1.5, 6.0)
    g = np.random.default_rng(31).lognormal(0.0, 0.8, size=(L, L, L)) - 1.0
    tab = _synthetic_tables(L, nbins, 8, 0.2, 4, sub=2, anchors=3)
    p1, p2, _ = mc.binned_sampler(g, g, g, phi, nbins, rr, rr, n, 41)
    i1, i2, _ = mc.binned_sampler(g, g, g, phi, nbins, rr, rr, n, 42, tables=tab)
    B = p1.size
    est = [B / n * a.ravel() for a in (p1, i1)]
    var = [(B / n) ** 2 * (b.ravel() - a.ravel() ** 2 / n) for a, b in ((p1, p2), (i1, i2))]
    z = (est[1] - est[0]) / np.sqrt(var[0] + var[1])
    assert np.abs(z).max() < 4.5, z
    assert np.sqrt(np.mean(z ** 2)) < 1.7


def test_bucketize_oracle_matches_definition(phi):
    """Buckets by edge count; the listing is a permutation sorted by (bucket, sub-cell)."""
    L, sub = 8, 2
    g = np.random.default_rng(8).normal(size=(L, L, L))
    edges = [-0.5, 0.0, 0.7]
    cb, cells, off, v = mc.bucketize(g, phi, sub, edges)
    Ls = L * sub
    idx = np.arange(Ls ** 3)
    c = (np.stack([idx // (Ls * Ls), (idx // Ls) % Ls, idx % Ls], axis=1) + 0.5) / sub
    np.testing.assert_allclose(v, mc.interpolate_bruteforce(g, c, phi), rtol=1e-12, atol=1e-13)
    assert np.array_equal(cb, (v[:, None] >= np.array(edges)[None, :]).sum(axis=1))
    assert np.array_equal(np.sort(cells), idx)
    key = cb[cells].astype(np.int64) * Ls ** 3 + cells
    assert np.all(np.diff(key) > 0)
    assert off[-1] == Ls ** 3 and np.array_equal(np.diff(off), np.bincount(cb, minlength=4))


def test_los_windows_obey_integral_identities():
    """Each line-of-sight window is pinned to another through an identity of the geometry:
    the mu-average of rings (R = s sqrt(1 - mu^2), H = s mu) is the spherical shell sin(q)/q; the
    area average of rings of radius r <= R is the disk; the average over heights |h| <= H of a ring
    (disk) is the cylinder shell (cylinder)."""
    rng = np.random.default_rng(8)
    k = rng.uniform(-0.45, 0.45, (40, 3))
    ki, kj, kk = k.T
    kn = np.linalg.norm(k, axis=1)
    mu, wmu = np.polynomial.legendre.leggauss(160)
    mu, wmu = 0.5 * (mu + 1.0), 0.5 * wmu  # [0, 1]
    for s in (1.7, 6.0):
        ring = sum(w * windows.los_window("ring", ki, kj, kk, s * np.sqrt(1 - m * m), s * m) for m, w in zip(mu, wmu))
        q = 2 * np.pi * kn * s
        np.testing.assert_allclose(ring, np.sin(q) / q, rtol=0, atol=1e-12)
    R, H = 3.1, 2.4
    t, wt = np.polynomial.legendre.leggauss(160)
    r, wr = 0.5 * R * (t + 1.0), 0.5 * R * wt
    disk = sum(w * 2 * rr / R ** 2 * windows.los_window("ring", ki, kj, kk, rr, H) for rr, w in zip(r, wr))
    np.testing.assert_allclose(windows.los_window("disk", ki, kj, kk, R, H), disk, rtol=0, atol=1e-12)
    h, wh = H * t, H * wt
    for avg, base in (("cylshell", "ring"), ("cylinder", "disk")):
        ref = sum(w / (2 * H) * windows.los_window(base, ki, kj, kk, R, hh) for hh, w in zip(h, wh))
        np.testing.assert_allclose(windows.los_window(avg, ki, kj, kk, R, H), ref, rtol=0, atol=1e-12)
    # the line of sight is a direction: along x it is the z-axis window with the axes exchanged
    for wt_ in windows.LOS_WINDOW_TYPES:
        np.testing.assert_allclose(windows.los_window(wt_, ki, kj, kk, R, H, los=(2.0, 0.0, 0.0)),
                                   windows.los_window(wt_, kk, kj, ki, R, H), rtol=0, atol=1e-15)
    assert windows.los_window("disk", 0.0, 0.0, 0.0, R, H) == 1.0
    assert windows.los_window("cylinder", 0.0, 0.0, 0.0, R, H) == 1.0


def test_los_kernel_mu_average_is_the_shell_kernel(phi):
    """At kernel level the Gauss-Legendre mu-average of the ring kernels (convolution.py:73-157 alias sums)
    equals the octant shell kernel of the C oracle (convolution.py:14-38 + :263-281)."""
    L = 8
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    s = 2.6
    mu, wmu = np.polynomial.legendre.leggauss(96)
    mu, wmu = 0.5 * (mu + 1.0), 0.5 * wmu
    avg = sum(w * windows.los_rfft_kernel(L, P, "ring", s * np.sqrt(1 - m * m), s * m) for m, w in zip(mu, wmu))
    shell = windows.octant_rfft_kernel(L, P, "shell", s)
    assert np.abs(avg - shell).max() <= 1e-11 * np.abs(shell).max()


def test_particle_centre_estimator_pins(phi):
    """corr3pcf_kernels.py:66-97 restated: with unit weights and rho1 = 1 it is the box-random estimator of a
    constant first field (the scaling function sums to one, so n1 = 1); a doubled weight equals a repeated
    centre; a common weight factor cancels."""
    L = 16
    rng = np.random.default_rng(31)
    e2, e3 = rng.normal(size=(L, L, L)), rng.normal(size=(L, L, L))
    cen = rng.uniform(0, L, (40, 3))
    abc = rng.uniform(size=(5, 3))
    R1, R2, mu = 2.5, 3.75, 0.3
    box = mc.triplet_box_random(R1, R2, mu, cen, abc, np.ones((L, L, L)), e2, e3, phi)
    part = mc.triplet_particle_centers(R1, R2, mu, cen, np.ones(40), 1.0, abc, e2, e3, phi)
    assert part == pytest.approx(box, rel=1e-12, abs=1e-15)
    w = rng.uniform(0.5, 2.0, 40)
    a = mc.triplet_particle_centers(R1, R2, mu, cen, np.r_[w[:-1], 2 * w[-1]], 0.7, abc, e2, e3, phi)
    b = mc.triplet_particle_centers(R1, R2, mu, np.r_[cen, cen[-1:]], np.r_[w, w[-1]], 0.7, abc, e2, e3, phi)
    assert a == pytest.approx(b, rel=1e-12)
    c = mc.triplet_particle_centers(R1, R2, mu, cen, 3.0 * w, 0.7, abc, e2, e3, phi)
    assert c == pytest.approx(mc.triplet_particle_centers(R1, R2, mu, cen, w, 0.7, abc, e2, e3, phi), rel=1e-12)


def _anisotropic_field(L, seed):
    rng = np.random.default_rng(seed)
    F = np.fft.fftn(rng.normal(size=(L, L, L)))
    k = np.fft.fftfreq(L)
    F *= np.exp(-20.0 * k[:, None, None] ** 2 - 5.0 * k[None, :, None] ** 2 - 2.0 * k[None, None, :] ** 2)
    d = np.real(np.fft.ifftn(F))
    return d - d.mean()


@pytest.mark.parametrize("ell", [0, 2, 4])
def test_xi_ell_prefactor_by_gauss_legendre_ring_route(phi, ell):
    """The build-defined xi_l(s) = (-i)^l sqrt(4 pi (2l+1)) <(delta @ j_l Y_l0) delta> / rr equals
    (2l+1) int_0^1 xi(s, mu) P_l(mu) dmu over the ring-window (s, mu) route (window_functions.py:164-192,
    corr2pcf.py:275-284) on an anisotropic field; a dropped (2l+1), a sign flip or the odd-l factor i^l
    in place of (-1)^(l/2) would each miss by far more than the tolerance."""
    L = 16
    d = _anisotropic_field(L, 5)
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    rr = 0.37
    for s in (2.5, 5.0, 6.75):
        a = windows.xi_ell_window_route(d, P, s, ell, rr)
        b = windows.xi_ell_ring_route(d, P, s, ell, rr)
        assert abs(a) > 1e-6
        assert a == pytest.approx(b, rel=1e-10)
        if ell:
            assert abs(a / (2 * ell + 1) - b) > 1e-3 * abs(b)
            assert abs(-a - b) > 1e-3 * abs(b)


def test_summand_pairs_regrouping_matches_summand(phi):
    """The matrix regrouping of the contraction for many field pairs equals the per-pair restatement of
    compute_3d_summand_cpu (corr3pcf_multipoles.py:134-170) for every pair, including swapped pairs."""
    L = 8
    rng = np.random.default_rng(41)
    d = rng.normal(size=(L, L, L))
    F1 = [rng.normal(size=(L, L, L)) + 1j * rng.normal(size=(L, L, L)) for _ in range(3)]
    F2 = [rng.normal(size=(L, L, L)) + 1j * rng.normal(size=(L, L, L)) for _ in range(2)]
    g = scaling.cal_gamma(phi, 3, 1024)
    got = multipole.summand_pairs(d, F1, F2, g)
    for i in range(3):
        for j in range(2):
            want = multipole.summand_c(d, F1[i], F2[j], g)
            assert abs(got[i, j] - want) <= 1e-12 * abs(want)
    same = multipole.summand_pairs(d, F1, F1, g)
    assert abs(same[2, 0] - multipole.summand_c(d, F1[2], F1[0], g)) <= 1e-12 * abs(same[2, 0])


@pytest.mark.parametrize("ell", [0, 2, 4])
def test_xi_ell_half_spectrum_sum_matches_window_route(phi, ell):
    """The half-spectrum Parseval sum of the multipole pair product (used at large grids) equals the
    convolution route with the full complex window (corr2pcf.py:366-372), even and odd L/2 planes included."""
    for L in (8, 16):
        d = _anisotropic_field(L, 7 + L)
        P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
        for s in (1.75, 0.3 * L):
            a = windows.xi_ell_window_route(d, P, s, ell, 0.8)
            b = windows.xi_ell_half_spectrum(d, P, s, ell, 0.8)
            assert b == pytest.approx(a, rel=1e-11, abs=1e-14)


# ---------------------------------------------------------------- radial-profile tables (non-thin edges)
RADIAL_CASES = [("sphere", {"R": 20.0}), ("thick_shell", {"R": 30.0, "delta_R": 10.0}), ("gaussian", {"R": 8.0}),
                ("gaussian_shell", {"R_shell": 30.0, "R_smooth": 5.0})]


@pytest.mark.parametrize("rtype,args", RADIAL_CASES)
def test_hankel_table_monopole_is_the_closed_form(rtype, args):
    """U_0 from the quadrature equals the closed-form isotropic window (window_functions.py:29-129) of the
    same profile; for the k-space Gaussian shell this also pins the inverse transform."""
    from oracle import radial
    k_max = np.sqrt(3.0) * 64 / 500.0
    tab = radial.hankel_table(rtype, args, 0, k_max)
    k = np.linspace(0.0, k_max, tab.size)
    if rtype == "gaussian_shell":
        want = radial.gaussian_shell_T(k, args["R_shell"], args["R_smooth"])
    elif rtype == "gaussian":
        want = np.exp(-0.5 * (2.0 * np.pi * k * args["R"]) ** 2)
    elif rtype == "sphere":
        want = radial._sphere_T(k, args["R"])
    else:
        a, b = args["R"] - 0.5 * args["delta_R"], args["R"] + 0.5 * args["delta_R"]
        want = (b ** 3 * radial._sphere_T(k, b) - a ** 3 * radial._sphere_T(k, a)) / (b ** 3 - a ** 3)
    assert np.abs(tab - want).max() <= 1e-8


def test_hankel_table_higher_orders_by_adaptive_quadrature():
    """U_2 and U_5 at sampled k against scipy.integrate.quad of 4 pi r^2 w(r) j_l(2 pi k r)."""
    from scipy.integrate import quad
    from oracle import radial
    k_max = np.sqrt(3.0) * 32 / 300.0
    for rtype, args, w, (lo, hi) in (
            ("sphere", {"R": 12.0}, lambda r: 3.0 / (4.0 * np.pi * 12.0 ** 3), (0.0, 12.0)),
            ("thick_shell", {"R": 20.0, "delta_R": 6.0}, lambda r: 3.0 / (4.0 * np.pi * (23.0 ** 3 - 17.0 ** 3)),
             (17.0, 23.0))):
        for l in (2, 5):
            tab = radial.hankel_table(rtype, args, l, k_max)
            k = np.linspace(0.0, k_max, tab.size)
            for j in (0, 37, tab.size // 3, tab.size - 1):
                ref = quad(lambda r: 4.0 * np.pi * r * r * w(r) * sps.spherical_jn(l, 2.0 * np.pi * k[j] * r), lo, hi,
                           limit=400, epsabs=1e-13, epsrel=1e-12)[0]
                assert abs(tab[j] - ref) <= 1e-9, (rtype, l, j)


def test_thin_thick_shell_tends_to_the_shell_transfer():
    """A thick shell of width 1e-4 R has U_l(k) = j_l(2 pi k R) up to O((2 pi k delta_R)^2)."""
    from oracle import radial
    k_max = np.sqrt(3.0) * 32 / 300.0
    R = 40.0
    tab = radial.hankel_table("thick_shell", {"R": R, "delta_R": 1e-4 * R}, 3, k_max)
    k = np.linspace(0.0, k_max, tab.size)
    assert np.abs(tab - sps.spherical_jn(3, 2.0 * np.pi * k * R)).max() <= 1e-6


@pytest.mark.parametrize("l", [0, 1, 4])
@pytest.mark.parametrize("rtype,args", RADIAL_CASES)
def test_product_hankel_tables_match_oracle(rtype, args, l):
    """The product's host tables (utils/radial_profiles.py) equal the oracle's restatement: node counts,
    supports, normalisation and the transform."""
    from oracle import radial
    from pyhermes_amd.utils import radial_profiles as rp
    k_max = np.sqrt(3.0) * 64 / 500.0
    k, tab = rp.build_radial_multipole_table(rtype, args, l, k_max)
    ref = radial.hankel_table(rtype, args, l, k_max)
    assert k == k_max and tab.shape == ref.shape
    assert np.abs(tab - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    d = rp.diagnose_radial_multipole_table(rtype, args, max(l, 1), k_max)
    assert d["passed"] and d["table_points"] == ref.size


def test_profile_window_table_path_matches_closed_form_monopole(phi):
    """The C oracle's tabulated builder (radial_multipole_windows.py:186-264) at l = 0 agrees with its
    closed-form builder (:147-183) to the table's interpolation error."""
    from oracle import radial
    L, box = 16, 200.0
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    g2p = L / box
    k_max = np.sqrt(3.0) * g2p
    for rtype, args in RADIAL_CASES:
        a = radial.profile_multipole_window(L, P, rtype, args, 0, 0, box)
        tab = native_f64(radial.hankel_table(rtype, args, 0, k_max))
        out = np.empty((L, L, L), dtype=np.complex128)
        from oracle import native
        native.lib().oracle_profile_multipole_window(L, native.ptr(native.f64(P)), 0, 0.0, 0.0, native.ptr(tab),
                                                     tab.size, k_max, g2p, 0, 0, native.ptr(out.view(np.float64)))
        assert np.abs(out - a).max() <= 1e-7 * np.abs(a).max(), rtype


def native_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def test_radial_profile_validation_messages():
    """The reference's request checks and messages (radial_profiles.py:115-186)."""
    from pyhermes_amd.utils import radial_profiles as rp
    with pytest.raises(ValueError, match="Unsupported radial profile 'cube'. Supported profiles: custom_kspace, "
                                         "custom_real, gaussian, gaussian_shell, shell, sphere, thick_shell."):
        rp.validate_radial_profile_request("cube", {})
    with pytest.raises(ValueError, match=r"radial profile 'sphere' requires len_args\['R'\]."):
        rp.validate_radial_profile_request("sphere", {})
    with pytest.raises(ValueError, match="requires delta_R > 0"):
        rp.validate_radial_profile_request("thick_shell", {"R": 1.0, "delta_R": 0.0})
    with pytest.raises(ValueError, match="non-negative R_shell and R_smooth"):
        rp.validate_radial_profile_request("gaussian_shell", {"R_shell": 1.0, "R_smooth": -1.0})
    with pytest.raises(ValueError, match=r"require other_args\['profile'\]"):
        rp.validate_radial_profile_request("custom_real", {}, None)
    with pytest.raises(ValueError, match="requires a real-space profile specification"):
        rp.validate_radial_profile_request("custom_real", {}, {"profile": rp.kspace_profile(np.cos, 5.0, 1.0)})
    prof = rp.real_space_profile(lambda r, R: (r <= R).astype(float), 10.0)
    rp.validate_radial_profile_request("custom_real", {"R": 10.0}, {"profile": prof})
    # a custom top-hat equals the built-in sphere's table
    k_max = np.sqrt(3.0) * 32 / 300.0
    _, a = rp.build_radial_multipole_table("custom_real", {"R": 10.0}, 2, k_max, {"profile": prof})
    _, b = rp.build_radial_multipole_table("sphere", {"R": 10.0}, 2, k_max)
    assert a.shape == b.shape and np.abs(a - b).max() <= 1e-10
