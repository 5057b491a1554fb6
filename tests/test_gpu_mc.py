"""GPU parity: interpolation, box-random triplet estimator, binned sampler and the proposal MLP."""

import numpy as np
import pytest

from oracle import mc as omc
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def _grid(L, seed):
    rng = np.random.default_rng(seed)
    return rng.lognormal(0.0, 0.5, size=(L, L, L))


def test_interpolation_matches_oracle(phi):
    import torch
    from pyhermes_amd import kernels
    L = 32
    g = _grid(L, 1)
    sp = np.random.default_rng(2).uniform(0, L, size=(5000, 3))
    sp[:5] = [[0, 0, 0], [L - 1e-9, 0.5, 3], [1e-12, L - 1e-12, 7.5], [0.999999, 1.0, 2.0], [31.5, 31.5, 31.5]]
    got = kernels.interpolate(kernels.as_device(g, torch.float64), kernels.as_device(sp, torch.float64), phi,
                              (0.3, -5.2, 40.1)).cpu().numpy()
    ref = omc.interpolate(g, sp, phi, (0.3, -5.2, 40.1))
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-14 * np.abs(ref).max())


def test_triplet_box_matches_oracle(phi):
    import torch
    from pyhermes_amd import kernels
    L = 32
    g = [_grid(L, s) for s in (3, 4, 5)]
    centers = omc.box_centers(4000, L, 42 + 1000003)
    abc = omc.rotation_uniforms(43, 12, 3)
    ref = omc.triplet_box_random(3.5, 6.25, 0.3, centers, abc, *g, phi)
    gd = [kernels.as_device(a, torch.float64) for a in g]
    cd = kernels.as_device(centers, torch.float64)
    n1 = kernels.interpolate(gd[0], cd, phi)
    got = kernels.mc_triplet_box(3.5, 6.25, 0.3, cd, n1, abc, gd[1], gd[2], phi)
    assert abs(got - ref) <= 1e-12 * abs(ref)


def test_binned_sampler_constant_field_and_counts(phi):
    import torch
    from pyhermes_amd import kernels, mcsample
    L = 32
    c = 1.7
    g = kernels.as_device(np.full((L, L, L), c), torch.float64)
    bins = mcsample.TriangleBins(3, 4, 5, 2.0, 8.0)
    res = mcsample.run_mc(g, g, g, phi, bins, 200000, 7)
    assert res.counts.sum() == 200000
    np.testing.assert_allclose(res.sum_wf, res.counts * c ** 3, rtol=1e-9)
    again = mcsample.run_mc(g, g, g, phi, bins, 200000, 7)
    assert np.array_equal(again.sum_wf, res.sum_wf) and np.array_equal(again.counts, res.counts)


def test_mlp_gradient_matches_oracle():
    import torch
    from pyhermes_amd import kernels, mcsample
    mlp = mcsample.ProposalMLP(seed=3)
    rng = np.random.default_rng(4)
    n = 1000
    x = rng.uniform(-1, 1, size=(n, 4)).astype(np.float32)
    y = rng.exponential(1.0, size=n).astype(np.float32)
    xt, yt = kernels.as_device(x, torch.float32), kernels.as_device(y, torch.float32)
    kernels.mlp_grad(mlp.wb, xt, yt, 50.0, mlp.grad)
    got = mlp.grad.cpu().numpy().astype(np.float64)
    wb = omlp.bf16_round(mlp.w.cpu().numpy())
    np.testing.assert_array_equal(mlp.wb.float().cpu().numpy(), wb)
    ref, loss = omlp.grad(wb, x, y.astype(np.float64), 50.0)
    assert abs(got[-1] - loss) < 1e-5 * max(1.0, abs(loss))
    scale = np.abs(ref).max()
    assert np.abs(got[:-1] - ref).max() < 2e-5 * scale
    s_got = kernels.mlp_forward(mlp.wb, xt).cpu().numpy()
    s_ref, _ = omlp.forward(wb, x.astype(np.float64))
    np.testing.assert_allclose(s_got, s_ref, rtol=1e-5, atol=1e-5)


def test_learned_proposal_is_unbiased_and_reduces_variance(phi):
    import torch
    from pyhermes_amd import kernels, mcsample
    from pyhermes_amd.mocks import lognormal_mock
    from oracle import projection as oproj
    J, box = 6, 500.0
    L = 1 << J
    pos, _ = lognormal_mock(200000, 64, box, 5000)
    eps, _, _ = oproj.project_c(pos, np.full(pos.shape[0], 1.0), phi, J, box)
    g = kernels.as_device(eps, torch.float64)
    bins = mcsample.TriangleBins(4, 4, 6, 1.5, 9.0)
    rec = torch.zeros((400000, 6), dtype=torch.float32, device=g.device)
    pilot = mcsample.run_mc(g, g, g, phi, bins, 400000, 11, record=rec)
    prop = mcsample.LearnedProposal(bins, g, nbucket=16, alpha=0.2, seed=1)
    prop.fit(rec, pilot, epochs=4)
    prop.build()
    plain = mcsample.run_mc(g, g, g, phi, bins, 2_000_000, 12)
    imp = mcsample.run_mc(g, g, g, phi, bins, 2_000_000, 13, proposal=prop)
    cmp = mcsample.compare(plain, imp)
    print(cmp)
    assert 0.6 < cmp["z_rms"] < 1.6, cmp
    assert cmp["median_gain"] > 1.2, cmp
