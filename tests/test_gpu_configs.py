"""GPU tests at the five benchmark configurations' own workloads.

Each test builds its configuration's seeded lognormal mock (pyhermes_amd.mocks; seeds 1000 k + {1..4} for
configuration k) at full size and runs it through the task API or the batched path, checking it against the
oracle (sampled where the full oracle would take minutes), against the FP64 path (FP32 error bar 1e-4) and,
for the batch, graph replays against eager runs bit for bit.
"""

import numpy as np
import pytest

from oracle import multipole as om
from oracle import native
from oracle import projection as oproj
from oracle import scaling
from oracle import windows as ow

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _project(pos, w, J, box, precision="fp64"):
    from pyhermes_amd.base import SFCProjection
    return SFCProjection({"SFCProjection": {"J": J, "box_size": box, "particle_pos": pos, "catalog_weight": w,
                                            "weight_normalization": "catalog", "precision": precision,
                                            "fout_path": ""}}).run(save_result=False)


def test_cfg1_2pcf_J7_100k_20_fp64_shells(phi):
    """Configuration 1: 100,000 lognormal-mock particles (128^3 GRF), box 1000 Mpc/h, unit weights, J = 7,
    isotropic 2PCF over 20 linear shell bins 5-150 Mpc/h, FP64 throughout -- projection, dd, delta_dd and xi
    against the full oracle (C projection, octant shell kernels, scipy convolutions), and the per-cell
    particle counts bit for bit."""
    import torch
    from pyhermes_amd import kernels
    from pyhermes_amd.mocks import lognormal_mock
    from pyhermes_amd.theory import Corr_2PCF
    box, J, n = 1000.0, 7, 100_000
    L = 1 << J
    _say("cfg1: mock 100k (128^3), projection J=7 FP64")
    pos, _ = lognormal_mock(n, 128, box, 1000)
    assert pos.shape == (n, 3)
    w = np.ones(n)
    field = _project(pos, w, J, box)
    res = Corr_2PCF({"Corr_2PCF": {"sfc_field": field, "random": "uniform",
                                   "sampling": {"s": {"min": 5.0, "max": 150.0, "n": 20}},
                                   "products": ["xi", "dd", "delta_dd", "rr"], "fout_path": ""}}).run(
        save_result=False)
    _say("cfg1: oracle")
    pw, sums = oproj.projection_weights(w, np.ones(n), "catalog")
    eps, cell, _ = oproj.project_c(pos, pw, phi, J, box)
    assert np.abs(field.epsilon - eps).max() <= 1e-12 * np.abs(eps).max()
    counts = kernels.project(kernels.as_device(pos, torch.float64), kernels.as_device(pw, torch.float64), J, box,
                             phi, want32=False, want64=False, want_counts=True)["counts"].cpu().numpy()
    want = np.zeros(L ** 3, dtype=np.int64)
    np.add.at(want, ((cell[:, 0] % L) * L + cell[:, 1] % L) * L + cell[:, 2] % L, 1)
    np.testing.assert_array_equal(counts.ravel(), want)
    rho = sums["field_integral"] / L ** 3
    delta = eps - rho
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    threads = native.num_threads()
    sc = (L / box) ** 6
    s = np.linspace(5.0, 150.0, 20)
    dd, ddd = [], []
    for sv in s:
        K = ow.octant_rfft_kernel(L, P, "shell", sv * L / box)
        dd.append(ow.pair_product(eps, eps, K, threads) * sc)
        ddd.append(ow.pair_product(delta, delta, K, threads) * sc)
    dd, ddd = np.array(dd), np.array(ddd)
    rr = rho * rho * sc
    np.testing.assert_allclose(res.dd, dd, rtol=1e-10)
    np.testing.assert_allclose(res.rr, rr, rtol=1e-13)
    xi_ref = ddd / rr
    err = np.abs(res.xi - xi_ref).max() / np.abs(xi_ref).max()
    _say(f"cfg1: xi max abs err / max |xi| = {err:.2e}")
    assert err <= 1e-10


def test_cfg2_xi_ell_J9_1M_rsd_30bins(phi):
    """Configuration 2: 1,000,000 lognormal-mock particles (256^3 GRF) with plane-parallel redshift-space
    shifts (sigma_v = 300 km/s along z) and lognormal weights (mean 1, sigma 0.3), J = 9, 30 shell bins over
    5-150 Mpc/h x l = 0, 2, 4 -- all 120 windows in one fused pass.  FP32 (FP32 coefficients, FP64
    normalisation and accumulators) against the FP64 path within 1e-4 of max |xi_l|; sampled bins of the
    FP64 path against the oracle (octant shell kernels for xi_0, the half-spectrum multipole sum for xi_2 and
    xi_4)."""
    from pyhermes_amd import kernels
    from pyhermes_amd.mocks import lognormal_mock
    from pyhermes_amd.theory import Corr_2PCF
    box, J, n, nbin = 1000.0, 9, 1_000_000, 30
    L = 1 << J
    ells = [0, 2, 4]
    assert nbin * (1 + len(ells)) <= kernels.PAIR_WINDOWS_PER_PASS  # one fused launch
    _say("cfg2: mock 1M (256^3), RSD + lognormal weights")
    pos, w = lognormal_mock(n, 256, box, 2000, rsd_sigma_kms=300.0, weight_sigma=0.3)
    assert pos.shape == (n, 3) and np.all((pos >= 0.0) & (pos < box))
    out = {}
    fields = {}
    for precision in ("fp32", "fp64"):
        _say(f"cfg2: projection J=9 and xi_l, {precision}")
        f = _project(pos, w, J, box, precision)
        fields[precision] = f
        out[precision] = Corr_2PCF({"Corr_2PCF": {"sfc_field": f, "random": "uniform",
                                                  "sampling": {"s": {"min": 5.0, "max": 150.0, "n": nbin}},
                                                  "products": ["xi"], "multipoles": ells, "fout_path": ""}}).run(
            save_result=False)
    for j, ell in enumerate(ells):
        a, b = out["fp32"].xi_l[j], out["fp64"].xi_l[j]
        err = np.abs(a - b).max() / np.abs(b).max()
        _say(f"cfg2: xi_{ell} FP32 vs FP64 max abs / max = {err:.2e}")
        assert err <= 1e-4, (ell, err)
    _say("cfg2: oracle on sampled bins")
    f64 = fields["fp64"]
    eps = f64.epsilon
    rho = f64.field_integral / L ** 3
    delta = eps - rho
    del fields
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    threads = native.num_threads()
    s = np.linspace(5.0, 150.0, nbin)
    got = out["fp64"]
    for b in (0, 29):
        K = ow.octant_rfft_kernel(L, P, "shell", s[b] * L / box)
        ref = ow.pair_product(delta, delta, K, threads) / rho ** 2
        del K
        assert got.xi[b] == pytest.approx(ref, rel=1e-9, abs=1e-12 * abs(got.xi).max()), b
        assert got.xi_l[0][b] == pytest.approx(ref, rel=1e-9, abs=1e-12 * abs(got.xi).max()), b
    for ell, b in ((2, 11), (4, 20)):
        ref = ow.xi_ell_half_spectrum(delta, P, s[b] * L / box, ell, rho * rho, threads)
        scale = np.abs(got.xi_l[ells.index(ell)]).max()
        _say(f"cfg2: xi_{ell} bin {b}: product {got.xi_l[ells.index(ell)][b]:.6e}, oracle {ref:.6e}")
        assert abs(got.xi_l[ells.index(ell)][b] - ref) <= 1e-9 * scale, (ell, b)


@pytest.fixture(scope="module")
def cfg3_fields():
    """Configuration 3's inputs: a 2,000,000-particle lognormal mock (256^3 GRF) and a 2,000,000-point uniform
    random catalogue in a 1000 Mpc/h box, both projected at J = 8 (FP64)."""
    from pyhermes_amd.mocks import lognormal_mock, uniform_catalogue
    box, J, n = 1000.0, 8, 2_000_000
    _say("cfg3: mock 2M (256^3) and 2M uniform randoms, projection J=8")
    pos, _ = lognormal_mock(n, 256, box, 3000)
    rpos = uniform_catalogue(n, box, 3005)
    field = _project(pos, np.ones(n), J, box)
    rfield = _project(rpos, np.ones(n), J, box)
    return field, rfield, box, J


def _cfg3_task(field, rfield, sampling, backend):
    from pyhermes_amd.theory import Corr_3PCF_Multipole
    return Corr_3PCF_Multipole({"Corr_3PCF_Multipole": {
        "sfc_field": field, "random": rfield, "l_min": 0, "l_max": 10,
        "products": ["zeta_l", "delta_ddd_l", "rrr_l"], "summation_backend": backend, "fout_path": "",
        "sampling": sampling,
        "binning_window12": {"type": "shell", "mapping": {"R": "r12"}},
        "binning_window13": {"type": "shell", "mapping": {"R": "r13"}}}}).run(save_result=False)


def test_cfg3_sh3pcf_J8_15_shells_lmax10_random_catalogue(cfg3_fields):
    """Configuration 3 through Corr_3PCF_Multipole: 15 x 15 shells over 10-120 Mpc/h, l <= 10, delta = D - R
    with the 2M-point random catalogue, rrr_l from the random field and zeta_l from the ratio solve, on the
    batched FP32 path; three sets -- the smallest-shell pair among them -- against the FP64 path (relative L2
    over l <= 1e-4 for zeta_l, delta_ddd_l and rrr_l)."""
    field, rfield, box, J = cfg3_fields
    shells = np.linspace(10.0, 120.0, 15)
    _say("cfg3: batched FP32 task, 225 samples")
    grid = _cfg3_task(field, rfield, {"mode": "grid", "r12": {"values": shells}, "r13": {"values": shells}}, "gpu")
    assert grid.zeta_l.shape == (225, 11)
    picks = [(0, 0), (0, 14), (5, 8)]
    r12 = [float(shells[i]) for i, _ in picks]
    r13 = [float(shells[j]) for _, j in picks]
    _say("cfg3: FP64 path on 3 sets")
    plain = _cfg3_task(field, rfield, {"mode": "paired", "r12": {"values": r12}, "r13": {"values": r13}},
                       "gpu_fp64")
    for k, (i, j) in enumerate(picks):
        row = i * 15 + j
        assert grid.sample_params["r12"][row] == r12[k] and grid.sample_params["r13"][row] == r13[k]
        for name in ("zeta_l", "delta_ddd_l", "rrr_l"):
            a, b = getattr(grid, name)[row], getattr(plain, name)[k]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            _say(f"cfg3: set ({r12[k]:.1f}, {r13[k]:.1f}) {name} FP32 vs FP64 rel L2 {rel:.2e}")
            assert rel <= 1e-4, (k, name, rel)
    # unique sets: ddd_l(r1, r2) = ddd_l(r2, r1) for the auto case
    dd = grid.delta_ddd_l.reshape(15, 15, 11)
    assert np.abs(dd - dd.transpose(1, 0, 2)).max() <= 1e-6 * np.abs(dd).max()


def test_cfg3_m_terms_against_oracle(cfg3_fields, phi):
    """Configuration 3's contraction per (l, m) against the oracle's m_summand (corr3pcf_multipoles.py:393-413
    on fields built as :267-321 do) for (l, m) = (0, 0), (5, 3), (10, 7) of the set (10, 33.6) Mpc/h on
    delta = D - R at J = 8: the FP64 path's complex v_m to 1e-9, the FP32 batched path's Re v_m to 1e-4 (plus
    1e-6 of the Cauchy-Schwarz bound of the sum)."""
    import torch
    from pyhermes_amd import kernels, multipole
    field, rfield, box, J = cfg3_fields
    L = 1 << J
    delta = field.epsilon - rfield.epsilon
    shells = np.linspace(10.0, 120.0, 15) * L / box
    i, j = 0, 3
    geo = multipole.MultipoleGeometry(phi, L)
    dd = kernels.as_device(delta, torch.float64)
    _say("cfg3 m-terms: FP64 path")
    _, mv64 = multipole.ddd_multipole_plain_f64(geo, dd, dd, dd, shells[i], shells[j], 0, 10, return_m_values=True)
    _say("cfg3 m-terms: FP32 batched path")
    d32 = kernels.as_device(delta, torch.float32)
    mv32 = {}
    multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, shells, shells, 0, 10, m_values=mv32)
    del dd, d32
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    gamma = scaling.cal_gamma(phi, 3, 1024)
    threads = native.num_threads()
    for l, m in ((0, 0), (5, 3), (10, 7)):
        _say(f"cfg3 m-terms: oracle (l, m) = ({l}, {m})")
        F1 = om.multipole_fields(delta, P, shells[i], l, m, threads)
        F2 = om.multipole_fields(delta, P, shells[j], l, -m, threads)
        ref = om.m_summand(delta, F1, F2, gamma)
        cs = 4.0 * np.pi / L ** 3 * np.sqrt(np.sum(delta ** 2) * np.sum(np.abs(F1) ** 2) * np.sum(np.abs(F2) ** 2))
        del F1, F2
        got64 = mv64[l][m]
        got32 = mv32[(l, m)][i, j]
        _say(f"cfg3 m-terms ({l}, {m}): oracle {ref:.6e}, FP64 {got64:.6e}, FP32 Re {got32:.6e}, CS {cs:.2e}")
        assert abs(got64 - ref) <= 1e-9 * abs(ref) + 1e-13 * cs
        assert abs(got32 - ref.real) <= 1e-4 * abs(ref.real) + 1e-6 * cs


def test_cfg4_mc_J8_1e8_triangles_learned_proposal():
    """Configuration 4: the J = 8 field of configuration 3's mock, 1e8 triangles in 15 x 15 (r1, r2) shells
    over 10-120 Mpc/h x 20 cos(theta) bins; the proposal MLP (3 layers, width 64, bf16 weights, FP32
    accumulation) is trained on 1e7 plain-MC triangle weights on the GPU.  At equal triangle count against
    uniform MC: median per-bin variance gain >= 3 (triangles the uniform sampler needs for the learned
    sampler's variance) and no bias (z-scores of the difference with rms in [0.8, 1.25])."""
    import torch
    from pyhermes_amd import kernels, mcsample
    from pyhermes_amd.mocks import lognormal_mock
    from pyhermes_amd.utils.wavelet_grid import sample_scaling_function
    box, J = 1000.0, 8
    L = 1 << J
    _say("cfg4: mock 2M (256^3), projection J=8")
    pos, _ = lognormal_mock(2_000_000, 256, box, 3000)
    phi = sample_scaling_function()
    proj = kernels.project(kernels.as_device(pos, torch.float64),
                           kernels.as_device(np.ones(pos.shape[0]), torch.float64), J, box, phi,
                           want32=False, want64=True)
    eps = proj["eps64"]
    delta = (eps / eps.mean() - 1.0).contiguous()
    del proj, eps
    phi_d = kernels.as_device(phi, torch.float64)
    s = L / box
    bins = mcsample.TriangleBins(15, 15, 20, 10.0 * s, 120.0 * s)
    _say("cfg4: proposal buckets, 1e7-record pilot, MLP training")
    prop = mcsample.LearnedProposal(bins, delta, phi_d, nbucket=128, alpha=0.02, sub=4, anchors=3, seed=1)
    rec = torch.empty((10_000_000, 8), dtype=torch.float32, device=delta.device)
    pilot = mcsample.run_mc(delta, delta, delta, phi_d, bins, 10_000_000, 11, proposal=prop.recording, record=rec)
    losses = prop.fit(rec, pilot, epochs=3)
    prop.build("uniform")
    del rec
    assert np.all(np.isfinite(losses))
    _say(f"cfg4: MLP epoch losses {[round(float(v), 4) for v in losses]}")
    _say("cfg4: 1e8 uniform and 1e8 learned triangles")
    plain = mcsample.run_mc(delta, delta, delta, phi_d, bins, 100_000_000, 12)
    imp = mcsample.run_mc(delta, delta, delta, phi_d, bins, 100_000_000, 13, proposal=prop)
    assert int(plain.counts.sum()) == 100_000_000
    cmp = mcsample.compare(plain, imp)
    ok = (plain.variance > 0) & (imp.variance > 0)
    gain = (plain.variance[ok] / imp.variance[ok]) * (imp.n / plain.n)
    med = float(np.median(gain))
    _say(f"cfg4: median gain {med:.2f} over {int(ok.sum())} bins, z_rms {cmp['z_rms']:.3f}, "
         f"z_mean {cmp['z_mean']:.3f}")
    assert ok.sum() >= 0.95 * bins.B
    assert med >= 3.0
    assert 0.8 <= cmp["z_rms"] <= 1.25


def test_cfg5_batch_64_uneven_catalogues_graph():
    """Configuration 5: 64 seeded lognormal sub-box catalogues, sizes log-uniform in 1e4..2e6 particles, box
    sides 250-1000 Mpc/h, J in {6, 7, 8}; each runs the isotropic 2PCF (20 bins) and the l <= 6 auto SH-3PCF
    (8 shells) in persistent kernels captured into one hipGraph.  Replays (after eager work and
    synchronisation) equal the eager run bit for bit; a J = 6 and a J = 8 catalogue are checked against the
    task API (Corr_2PCF xi, and the batched multipole path on the same FP32 field) to 1e-5."""
    import torch
    from pyhermes_amd import kernels, multipole
    from pyhermes_amd.batch import Catalogue, CatalogueBatch
    from pyhermes_amd.mocks import lognormal_mock
    from pyhermes_amd.theory import Corr_2PCF
    from pyhermes_amd.utils.wavelet_grid import sample_scaling_function
    rng = np.random.default_rng(5000)
    ncat = 64
    sizes = np.round(10.0 ** rng.uniform(4.0, np.log10(2e6), ncat)).astype(np.int64)
    boxes = rng.uniform(250.0, 1000.0, ncat)
    Js = rng.choice([6, 7, 8], ncat)
    _say(f"cfg5: {ncat} mocks, {int(sizes.sum())} particles, J counts {np.bincount(Js)[6:].tolist()}")
    cats = []
    for i in range(ncat):
        pos, w = lognormal_mock(int(sizes[i]), 64, float(boxes[i]), 5000 + 10 * i)
        cats.append(Catalogue(pos, w, float(boxes[i]), int(Js[i])))
    phi = sample_scaling_function()
    s_bins = np.linspace(10.0, 100.0, 20)
    shells = np.linspace(10.0, 80.0, 8)
    b = CatalogueBatch(cats, s_bins, shells, 6, phi)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        _say("cfg5: warm-up and eager batch")
        b.warmup()
        ref = b.run_eager()
        _say("cfg5: capture and replays")
        g = b.capture()
        for rep in range(2):
            b.pp_out.zero_()
            b.gram_out.zero_()
            torch.cuda.synchronize()
            b.replay(g)
            torch.cuda.synchronize()
            b.check_errors()
            got = b.results()
            for k, (r, q) in enumerate(zip(ref, got)):
                assert np.array_equal(r["xi"], q["xi"]), (rep, k)
                assert np.array_equal(r["zeta_l"], q["zeta_l"]), (rep, k)
        kernels.graph_destroy(g)
    for k in (int(np.nonzero(Js == 6)[0][0]), int(np.nonzero(Js == 8)[0][0])):
        c, r = b.cats[k], ref[k]
        _say(f"cfg5: catalogue {k} (n={c.n}, box={c.box:.0f}, J={c.J}) against the task API")
        f = _project(c.pos, c.weight, c.J, c.box, "fp32")
        xi = Corr_2PCF({"Corr_2PCF": {"sfc_field": f, "random": "uniform",
                                      "sampling": {"s": {"min": 10.0, "max": 100.0, "n": 20}},
                                      "products": ["xi"], "fout_path": ""}}).run(save_result=False).xi
        assert np.abs(r["xi"] - xi).max() <= 1e-5 * np.abs(xi).max()
        L = 1 << c.J
        rho = f.field_integral / L ** 3
        d = (f.device_epsilon() - np.float32(rho)).contiguous()
        geo = multipole.MultipoleGeometry(phi, L)
        radii = shells * L / c.box
        zeta = multipole.ddd_multipole_batched_f32(geo, d, d, d, radii, radii, 0, 6) / rho ** 3
        assert np.abs(r["zeta_l"] - zeta).max() <= 1e-5 * np.abs(zeta).max()
    b.close()
