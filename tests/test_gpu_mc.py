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


def _device_tables(tab):
    """hermes_mc_proposal over device copies of numpy proposal tables (kept alive on the struct)."""
    import torch
    from pyhermes_amd import _lib, kernels
    keep = {k: kernels.as_device(np.ascontiguousarray(tab[k]), dt) for k, dt in
            (("cell_bucket", torch.uint8), ("bucket_cells", torch.int32), ("bucket_off", torch.int64),
             ("bucket_cdf", torch.float32), ("bucket_density", torch.float64), ("bin_cdf", torch.float32),
             ("bin_pdf", torch.float32))}
    st = _lib.MCProposal()
    for k, t in keep.items():
        setattr(st, k, t.data_ptr())
    st.nbucket = int(tab["nbucket"])
    st.alpha = float(tab["alpha"])
    st.sub = int(tab.get("sub", 1))
    st.anchors = int(tab.get("anchors", 1))
    st._keep = keep
    return st


@pytest.mark.parametrize("importance,sub,anchors", [(False, 1, 1), (True, 1, 1), (True, 2, 3), (True, 4, 3),
                                                    (True, 2, 1)])
def test_binned_sampler_matches_oracle(phi, importance, sub, anchors):
    """Same Philox streams: bit-exact per-bin counts, sums to the fixed-point resolution."""
    import torch
    from pyhermes_amd import kernels
    from tests.test_oracle import _synthetic_tables
    L, n, seed = 32, 300000, 4242
    nbins, r1r, r2r = (3, 4, 5), (1.5, 9.0), (2.0, 11.0)
    g = [_grid(L, s) - 1.0 for s in (6, 7, 8)]
    tab = _synthetic_tables(L, nbins, 8, 0.2, 9, sub, anchors) if importance else None
    print(f"mc_binned vs oracle importance={importance} sub={sub} anchors={anchors}", flush=True)
    ref = omc.binned_sampler(*g, phi, nbins, r1r, r2r, n, seed, tables=tab)
    gd = [kernels.as_device(a, torch.float64) for a in g]
    prop = _device_tables(tab) if importance else None
    got = kernels.mc_binned(*gd, phi, nbins, r1r, r2r, n, seed, proposal=prop)
    assert np.array_equal(got[2], ref[2])
    for a, b in zip(got[:2], ref[:2]):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9 * np.abs(b).max())


def _edges_between_values(v, nbucket):
    """Quantile edges moved to midpoints between neighbouring sample values."""
    vs = np.sort(v)
    edges = []
    for q in np.linspace(0, 1, nbucket + 1)[1:-1]:
        j = min(max(int(q * vs.size), 1), vs.size - 1)
        while j < vs.size - 1 and vs[j] - vs[j - 1] < 1e-9:
            j += 1
        edges.append(0.5 * (vs[j - 1] + vs[j]))
    return np.array(edges)


@pytest.mark.parametrize("sub,nbucket", [(1, 7), (2, 7), (4, 7), (2, 1), (2, 256)])
def test_bucketize_matches_oracle(phi, sub, nbucket):
    """Bucket of every sub-cell and the stable per-bucket listing, bit for bit."""
    import torch
    from pyhermes_amd import kernels
    L = 16
    g = _grid(L, 21) - 1.0
    _, _, _, v = omc.bucketize(g, phi, sub, [])
    edges = _edges_between_values(v, nbucket)
    print(f"bucketize sub={sub} nbucket={nbucket}", flush=True)
    cb, cells, off, _ = omc.bucketize(g, phi, sub, edges)
    gcb, gcells, goff, counts = kernels.mc_bucketize(kernels.as_device(g, torch.float64), phi, sub, edges)
    assert np.array_equal(gcb.cpu().numpy(), cb)
    assert np.array_equal(gcells.cpu().numpy(), cells)
    assert np.array_equal(goff.cpu().numpy(), off)
    assert np.array_equal(counts, np.diff(off))


def test_binned_sampler_cap_paths_agree(phi):
    """Pilot cap, a small cap (overflow list + second pass) and a cap so small that the
    list overflows (rerun at the exact maximum) give the same sums and counts."""
    import torch
    from pyhermes_amd import kernels
    L, n, seed = 32, 1_200_000, 99
    nbins, rr = (2, 2, 3), (1.0, 7.0)
    g = kernels.as_device(_grid(L, 12) - 1.0, torch.float64)
    runs = []
    for cap in (0.0, 0.5, 1e-6):
        print(f"mc_binned cap={cap}", flush=True)
        runs.append(kernels.mc_binned(g, g, g, phi, nbins, rr, rr, n, seed, wf_cap=cap))
    for r in runs[1:]:
        assert np.array_equal(r[2], runs[0][2])
        for a, b in zip(r[:2], runs[0][:2]):
            np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * np.abs(b).max())
    assert runs[0][2].sum() == n


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


def test_mlp_train_steps_equal_per_step_calls():
    """The device-side training loop (hermes_mlp_train_steps) gives bit-identical weights, moments and bf16
    copy to the same steps issued one hermes_mlp_grad / hermes_mlp_adam pair at a time, with a short last
    batch."""
    import torch
    from pyhermes_amd import kernels, mcsample
    rng = np.random.default_rng(8)
    n, batch = 5000, 1024
    x = kernels.as_device(rng.uniform(-1, 1, size=(n, 4)).astype(np.float32), torch.float32)
    y = kernels.as_device(rng.exponential(1.0, size=n).astype(np.float32), torch.float32)
    steps = (n + batch - 1) // batch
    lrs = np.linspace(3e-3, 1e-3, steps).astype(np.float32)
    a, b = mcsample.ProposalMLP(seed=5), mcsample.ProposalMLP(seed=5)
    for s in range(steps):
        kernels.mlp_grad(a.wb, x[s * batch:(s + 1) * batch], y[s * batch:(s + 1) * batch], 50.0, a.grad)
        kernels.mlp_adam(a.w, a.m, a.v, a.grad, a.wb, float(lrs[s]), 7 + s + 1)
    kernels.mlp_train_steps(b.w, b.m, b.v, b.grad, b.wb, x, y, batch, 50.0, lrs, 7)
    for u, v in ((a.w, b.w), (a.m, b.m), (a.v, b.v), (a.wb, b.wb)):
        assert torch.equal(u, v)
    with pytest.raises(RuntimeError):
        kernels.mlp_train_steps(b.w, b.m, b.v, b.grad, b.wb, x, y, batch, 50.0, np.zeros(steps + 1), 0)


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
    rec = torch.zeros((400000, 8), dtype=torch.float32, device=g.device)
    prop = mcsample.LearnedProposal(bins, g, phi, nbucket=16, alpha=0.2, sub=2, anchors=3, seed=1)
    pilot = mcsample.run_mc(g, g, g, phi, bins, 400000, 11, proposal=prop.recording, record=rec)
    prop.fit(rec, pilot, epochs=4)
    prop.build()
    plain = mcsample.run_mc(g, g, g, phi, bins, 2_000_000, 12)
    imp = mcsample.run_mc(g, g, g, phi, bins, 2_000_000, 13, proposal=prop)
    cmp = mcsample.compare(plain, imp)
    print(cmp)
    assert 0.6 < cmp["z_rms"] < 1.6, cmp
    assert cmp["median_gain"] > 1.2, cmp
