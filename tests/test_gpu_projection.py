"""GPU parity: projection scatter and FFTs against the CPU oracle."""

import numpy as np
import pytest

from oracle import projection as oproj

pytestmark = pytest.mark.gpu


def _case(kind, n, J, box, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        pos = rng.uniform(0, box, size=(n, 3))
        pos[:7] = rng.uniform(0, box * 1e-4, size=(7, 3))  # negative-index wrap cells
        pos[7] = [0.0, 0.0, 0.0]
    else:
        from pyhermes_amd.mocks import lognormal_mock
        pos, _ = lognormal_mock(n, 32, box, seed)
    w = rng.lognormal(0.0, 0.3, pos.shape[0])
    return pos, w


@pytest.mark.parametrize("kind,n,J", [("uniform", 20000, 5), ("lognormal", 30000, 5), ("uniform", 5000, 3),
                                      ("lognormal", 200000, 7)])
def test_projection_matches_oracle(phi, kind, n, J):
    from pyhermes_amd import kernels
    box = 1000.0
    pos, w = _case(kind, n, J, box, 100 + J)
    print(f"projection case kind={kind} n={n} J={J}", flush=True)
    out = kernels.project(pos, w, J, box, phi, want32=True, want64=True, want_cells=True, want_counts=True)
    eps_o, cell_o, frac_o = oproj.project_c(pos, w, phi, J, box)
    np.testing.assert_array_equal(out["cell"].cpu().numpy(), cell_o)
    np.testing.assert_array_equal(out["frac"].cpu().numpy(), frac_o)
    L = 1 << J
    wc = np.mod(cell_o, L)
    counts_o = np.zeros((L, L, L), dtype=np.int64)
    np.add.at(counts_o, (wc[:, 0], wc[:, 1], wc[:, 2]), 1)
    np.testing.assert_array_equal(out["counts"].cpu().numpy(), counts_o)
    e64 = out["eps64"].cpu().numpy()
    scale = np.abs(eps_o).max()
    assert np.abs(e64 - eps_o).max() <= 1e-12 * scale
    rel64 = np.linalg.norm(e64 - eps_o) / np.linalg.norm(eps_o)
    assert rel64 < 1e-12
    e32 = out["eps32"].cpu().numpy().astype(np.float64)
    rel32 = np.linalg.norm(e32 - eps_o) / np.linalg.norm(eps_o)
    assert rel32 < 1e-6, rel32
    assert abs(e64.sum() - w.sum()) < 1e-10 * w.sum()


def test_projection_is_bit_deterministic(phi):
    from pyhermes_amd import kernels
    pos, w = _case("lognormal", 50000, 6, 1000.0, 7)
    a = kernels.project(pos, w, 6, 1000.0, phi, want32=True, want64=True)
    b = kernels.project(pos, w, 6, 1000.0, phi, want32=True, want64=True)
    assert bool((a["eps64"] == b["eps64"]).all()) and bool((a["eps32"] == b["eps32"]).all())


def test_projection_empty_and_errors(phi):
    from pyhermes_amd import kernels
    out = kernels.project(np.zeros((0, 3)), np.zeros(0), 4, 100.0, phi, want32=True, want64=True)
    assert float(out["eps64"].abs().max()) == 0.0
    bad = np.array([[np.nan, 1.0, 1.0]])
    with pytest.raises(RuntimeError, match="out-of-range"):
        kernels.project(bad, np.ones(1), 4, 100.0, phi)
    with pytest.raises(RuntimeError, match="J="):
        kernels.project(np.ones((1, 3)), np.ones(1), 12, 100.0, phi)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_fft_roundtrip_against_numpy(dtype):
    import torch
    from pyhermes_amd import kernels
    L = 16
    rng = np.random.default_rng(1)
    x = rng.normal(size=(L, L, L))
    td = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-12 if dtype == "f64" else 2e-5
    xt = kernels.as_device(x, td)
    spec = kernels.rfftn(xt)
    ref = np.fft.rfftn(x)
    got = spec.cpu().numpy()
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    back = kernels.irfftn_unnormalized(spec.clone(), L).cpu().numpy() / L ** 3
    assert np.abs(back - x).max() <= tol * 10
    c = x + 1j * rng.normal(size=(L, L, L))
    ct = kernels.as_device(c, torch.complex128 if dtype == "f64" else torch.complex64)
    f = kernels.fftn_c2c(ct)
    assert np.abs(f.cpu().numpy() - np.fft.fftn(c)).max() <= tol * np.abs(np.fft.fftn(c)).max()
    kernels.fftn_c2c(f, inverse=True, out=f)
    assert np.abs(f.cpu().numpy() / L ** 3 - c).max() <= tol * 10


@pytest.mark.parametrize("case", ["f32_only", "f64_only", "counts_only", "dense_tiles", "tiny_weights"])
def test_projection_output_modes_match_oracle(phi, case):
    """Each output combination on its own (FP32 only is the batched pipeline's mode), tiles holding more
    than one 512-record chunk, and weights small enough that the fixed-point scale exceeds 2^100."""
    from pyhermes_amd import kernels
    box, J = 1000.0, 6  # 64^3: the full 4 x 8 x 64 tiles
    rng = np.random.default_rng(55)
    if case == "dense_tiles":  # 6000 particles in a 4 x 8 x 31-cell box: tiles with > 512 records
        pos = rng.uniform(0.0, 1.0, size=(6000, 3)) * np.array([4.0, 8.0, 31.0]) * box / (1 << J) + 100.0
    else:
        pos = rng.uniform(0.0, box, size=(30000, 3))
    w = rng.lognormal(0.0, 0.3, pos.shape[0]) * (1e-35 if case == "tiny_weights" else 1.0)
    print(f"projection output mode {case}", flush=True)
    eps_o, cell_o, _ = oproj.project_c(pos, w, phi, J, box)
    want32, want64, counts = case != "f64_only" and case != "counts_only", case == "f64_only", case == "counts_only"
    out = kernels.project(pos, w, J, box, phi, want32=want32, want64=want64, want_counts=counts)
    if counts:
        L = 1 << J
        wc = np.mod(cell_o, L)
        counts_o = np.zeros((L, L, L), dtype=np.int64)
        np.add.at(counts_o, (wc[:, 0], wc[:, 1], wc[:, 2]), 1)
        np.testing.assert_array_equal(out["counts"].cpu().numpy(), counts_o)
        return
    if want64:
        e64 = out["eps64"].cpu().numpy()
        assert np.abs(e64 - eps_o).max() <= 1e-12 * np.abs(eps_o).max()
        return
    e32 = out["eps32"].cpu().numpy().astype(np.float64)
    rel32 = np.linalg.norm(e32 - eps_o) / np.linalg.norm(eps_o)
    assert rel32 < 1e-6, rel32
    both = kernels.project(pos, w, J, box, phi, want32=True, want64=True)["eps32"].cpu().numpy().astype(np.float64)
    # the FP32-only conversion and the FP64-then-FP32 route differ by at most one rounding step
    assert np.all(np.abs(e32 - both) <= 2.0 ** -22 * np.abs(both)), np.abs(e32 - both).max()
