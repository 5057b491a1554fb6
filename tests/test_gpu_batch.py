"""GPU: capture-safe entry points and uneven catalogue batches replayed from one hipGraph."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(phi, extra=()):
    from pyhermes_amd.batch import Catalogue, CatalogueBatch
    from pyhermes_amd.mocks import lognormal_mock
    cats = []
    for i, (n, box, J) in enumerate(((4000, 200.0, 4), (20000, 300.0, 5), (9000, 250.0, 5), (15000, 400.0, 4))
                                    + tuple(extra)):
        pos, w = lognormal_mock(n, 32, box, 700 + 10 * i, weight_sigma=0.3)
        cats.append(Catalogue(pos, w, box, J))
    return CatalogueBatch(cats, s_bins=np.linspace(20.0, 60.0, 4), shells=np.array([25.0, 40.0, 55.0]), lmax=2,
                          phi=phi)


def test_delta_and_half_spectrum(phi):
    import torch
    from pyhermes_amd import kernels
    L = 16
    rng = np.random.default_rng(3)
    eps = rng.normal(size=(L, L, L)).astype(np.float32)
    e = kernels.as_device(eps, torch.float32)
    d = torch.empty_like(e)
    dc = torch.empty((L, L, L), dtype=torch.complex64, device=e.device)
    kernels.delta_f32(e, 2.0, 0.25, d_out=d, dc_out=dc)
    ref = (2.0 * eps.astype(np.float64) - 0.25).astype(np.float32)
    assert np.array_equal(d.cpu().numpy(), ref)
    assert np.array_equal(dc.cpu().numpy().real, ref) and not dc.cpu().numpy().imag.any()
    kernels.fft3_c2c_f32(dc, inverse=False)
    H = torch.empty((L, L, L // 2 + 1), dtype=torch.complex64, device=e.device)
    kernels.half_spectrum_f32(dc, H)
    want = np.fft.rfftn(ref.astype(np.float64))
    assert np.abs(H.cpu().numpy() - want).max() <= 3e-6 * np.abs(want).max()


def test_pair_plan_and_gram_dev_match_host_entry_points(phi):
    import torch
    from pyhermes_amd import kernels
    from oracle import scaling
    J = 5
    L = 1 << J
    rng = np.random.default_rng(4)
    f = kernels.as_device(rng.normal(size=(L, L, L)), torch.float32)
    F = kernels.rfftn(f)
    P = kernels.as_device(scaling.fourier_power_spectrum(phi, 0, 1, L, 1024), torch.float64)
    descs = [kernels.window_desc("shell", R) for R in (1.5, 4.0, 9.0)]
    descs += [kernels.window_desc("shell_multipole", R, l=l) for l in (0, 2, 4) for R in (1.5, 9.0)]
    plan = kernels.PairPlan(J, descs)
    out = torch.zeros(len(descs), dtype=torch.float64, device=f.device)
    plan.reduce(F, F, P, out)
    assert np.array_equal(out.cpu().numpy(), kernels.pair_reduce(F, F, P, descs))
    plan.close()
    with pytest.raises(RuntimeError):
        kernels.PairPlan(J, [kernels.window_desc("sphere", 2.0)])
    gamma = scaling.cal_gamma(phi, 3, 1024)
    lam, Q = np.linalg.eigh(gamma)
    U = torch.complex(torch.randn((3, L, L, L), device=f.device), torch.randn((3, L, L, L), device=f.device))
    blk = torch.zeros(256, dtype=torch.float64, device=f.device)
    kernels.multipole_gram_f32_dev(f, U, Q, lam, blk)
    host = kernels.multipole_gram_f32(f, U, Q, lam)
    assert np.array_equal(kernels.gram_block_to_matrix(blk.cpu().numpy(), 3), host)


def test_batch_graph_replays_match_eager(phi):
    """Three replays, each after eager work and a synchronisation, reproduce the eager batch bit for bit."""
    import torch
    from pyhermes_amd import kernels
    print("batch: 4 catalogues, J 4-5, graph vs eager", flush=True)
    b = _batch(phi)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.warmup()
        ref = b.run_eager()
        g = b.capture()
        for rep in range(3):
            b.pp_out.zero_()
            b.gram_out.zero_()
            torch.cuda.synchronize()
            b.replay(g)
            torch.cuda.synchronize()
            b.check_errors()
            got = b.results()
            for r, q in zip(ref, got):
                assert np.array_equal(r["xi"], q["xi"]), rep
                assert np.array_equal(r["zeta_l"], q["zeta_l"]), rep
        kernels.graph_destroy(g)
    b.close()


def test_batch_matches_task_api(phi):
    """xi against Corr_2PCF (FP32 projection) and zeta_l against the batched FP32 multipole path."""
    import torch
    from pyhermes_amd import multipole
    from pyhermes_amd.base import SFCProjection
    from pyhermes_amd.theory import Corr_2PCF
    b = _batch(phi)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.warmup()
        res = b.run_eager()
    for c, r in zip(b.cats, res):
        print(f"batch vs API: n={c.n} box={c.box} J={c.J}", flush=True)
        f = SFCProjection({"SFCProjection": {"J": c.J, "box_size": c.box, "particle_pos": c.pos,
                                             "catalog_weight": c.weight, "weight_normalization": "catalog",
                                             "precision": "fp32", "fout_path": ""}}).run(save_result=False)
        xi = Corr_2PCF({"Corr_2PCF": {"sfc_field": f, "random": "uniform",
                                      "sampling": {"s": {"min": 20.0, "max": 60.0, "n": 4}},
                                      "products": ["xi"], "fout_path": ""}}).run(save_result=False).xi
        assert np.abs(r["xi"] - xi).max() <= 1e-5 * np.abs(xi).max()
        L = 1 << c.J
        rho = f.field_integral / L ** 3
        d = (f.device_epsilon() - np.float32(rho)).contiguous()
        geo = multipole.MultipoleGeometry(phi, L)
        radii = b.shells * L / c.box
        ddd = multipole.ddd_multipole_batched_f32(geo, d, d, d, radii, radii, 0, b.lmax)
        zeta = ddd / rho ** 3
        assert np.abs(r["zeta_l"] - zeta).max() <= 1e-5 * np.abs(zeta).max()
    b.close()


def test_batch_graph_reports_bad_positions(phi):
    """A non-finite position read during a replay sets the sticky error word; the eager call raises."""
    import torch
    from pyhermes_amd import kernels
    b = _batch(phi)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.warmup()
        g = b.capture()
        b.replay(g)
        torch.cuda.synchronize()
        b.check_errors()
        pos = b.dev[1]["pos"]
        keep = pos[5].clone()
        pos[5, 1] = float("nan")
        b.replay(g)
        torch.cuda.synchronize()
        with pytest.raises(ValueError):
            b.check_errors()
        with pytest.raises(RuntimeError):
            b.enqueue(1)
        pos[5] = keep
        kernels.clear_errors()
        b.replay(g)
        torch.cuda.synchronize()
        b.check_errors()
        kernels.graph_destroy(g)
    b.close()


def test_capture_refuses_to_allocate(phi):
    """Work that would grow a workspace fails cleanly inside a capture; the capture still ends."""
    import torch
    from pyhermes_amd import kernels
    b = _batch(phi, extra=((60000, 300.0, 5),))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        kernels.release_workspaces()  # whatever earlier tests grew them to
        for i in range(4):  # warm the levels with the small catalogues only
            b.enqueue(i)
        torch.cuda.synchronize()
        kernels.project_reserve(5, 20000)
        b._warm = True
        kernels.capture_begin()
        try:
            with pytest.raises(RuntimeError):
                b.enqueue(4)
        finally:
            g = kernels.capture_end()
        kernels.graph_destroy(g)
    b.close()
