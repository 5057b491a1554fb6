"""GPU parity at the headline's full size (J = 9, 512^3 cells, the 2,000,000-particle mock).

Exact oracle comparisons where the oracle finishes in seconds at this size (the projection, one shell's
pair product), and size-independent properties elsewhere: the FFT round trip and Parseval's identity,
and the FP32 batched multipoles against the FP64 GPU path on the same field.
"""

import numpy as np
import pytest

from oracle import projection as oproj
from oracle import scaling
from oracle import windows as ow

pytestmark = pytest.mark.gpu

J, L, BOX, NPART = 9, 512, 1000.0, 2_000_000


@pytest.fixture(scope="module")
def mock():
    from pyhermes_amd.mocks import lognormal_mock
    pos, _ = lognormal_mock(NPART, 256, BOX, 3000)
    w = np.random.default_rng(9).lognormal(0.0, 0.3, pos.shape[0])
    return pos, w


@pytest.fixture(scope="module")
def projected(mock, phi):
    from pyhermes_amd import kernels
    pos, w = mock
    print("full size: projection of 2,000,000 particles at J=9", flush=True)
    out = kernels.project(pos, w, J, BOX, phi, want32=True, want64=True, want_counts=True)
    return {k: v.cpu().numpy() for k, v in out.items() if v is not None}


def test_projection_full_size_matches_oracle(mock, phi, projected):
    pos, w = mock
    eps_o, cell_o, _ = oproj.project_c(pos, w, phi, J, BOX)
    e64 = projected["eps64"]
    assert np.abs(e64 - eps_o).max() <= 1e-12 * np.abs(eps_o).max()
    e32 = projected["eps32"].astype(np.float64)
    assert np.linalg.norm(e32 - eps_o) / np.linalg.norm(eps_o) < 1e-6
    assert int(projected["counts"].sum()) == NPART
    assert abs(e64.sum() - w.sum()) < 1e-9 * w.sum()  # the sampled scaling function sums to one


def test_projection_full_size_is_deterministic(mock, phi, projected):
    from pyhermes_amd import kernels
    pos, w = mock
    again = kernels.project(pos, w, J, BOX, phi, want32=True, want64=True)
    assert np.array_equal(again["eps64"].cpu().numpy(), projected["eps64"])
    assert np.array_equal(again["eps32"].cpu().numpy(), projected["eps32"])


def test_fft_full_size_roundtrip_and_parseval():
    import torch
    from pyhermes_amd import kernels
    print("full size: 512^3 complex64 FFT round trip", flush=True)
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.complex(torch.randn((L, L, L), device="cuda", generator=g),
                      torch.randn((L, L, L), device="cuda", generator=g))
    y = x.clone()
    kernels.fft3_c2c_f32(y, inverse=False)
    # Parseval: sum |X|^2 = N sum |x|^2 (FP64 sums of the FP32 values)
    ex = float((x.real.double() ** 2 + x.imag.double() ** 2).sum())
    ey = float((y.real.double() ** 2 + y.imag.double() ** 2).sum())
    assert abs(ey / (L ** 3 * ex) - 1.0) < 1e-5
    kernels.fft3_c2c_f32(y, inverse=True)
    err = float((y / L ** 3 - x).abs().max())
    assert err < 1e-5 * float(x.abs().max()), err


def test_pair_reduce_full_size_matches_oracle(projected, phi):
    import torch
    from pyhermes_amd import kernels
    print("full size: 2PCF pair reductions at J=9 against the oracle", flush=True)
    eps = projected["eps64"]
    delta = eps / eps.mean() - 1.0
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    Pd = kernels.as_device(P, torch.float64)
    radii = (5.0, 17.5, 40.0)
    descs = [kernels.window_desc("shell", R) for R in radii]
    F64 = kernels.rfftn(kernels.as_device(delta, torch.float64))
    got64 = kernels.pair_reduce(F64, F64, Pd, descs) / L ** 6
    F32 = kernels.rfftn(kernels.as_device(delta, torch.float32))
    got32 = kernels.pair_reduce(F32, F32, Pd, descs) / L ** 6
    want = ow.pair_product(delta, delta, ow.octant_rfft_kernel(L, P, "shell", radii[1], 0.0), workers=16)
    assert abs(got64[1] - want) <= 1e-10 * abs(want), (got64[1], want)
    assert np.all(np.abs(got32 - got64) <= 1e-5 * np.abs(got64).max())


def test_multipoles_full_size_fp32_matches_fp64(projected, phi):
    import torch
    from pyhermes_amd import multipole
    print("full size: J=9 SH-3PCF multipoles, FP32 batched vs FP64 plain", flush=True)
    eps = projected["eps64"]
    delta = eps - eps.mean()
    geo = multipole.MultipoleGeometry(phi, L)
    radii = np.array([10.0, 30.0, 60.0]) * L / BOX
    d32 = torch.as_tensor(delta, dtype=torch.float32, device="cuda")
    d64 = torch.as_tensor(delta, dtype=torch.float64, device="cuda")
    lmax = 2
    batched = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, lmax)
    plain = np.array(multipole.ddd_multipole_plain_f64(geo, d64, d64, d64, radii[0], radii[2], 0, lmax))
    rel = np.linalg.norm(batched[0, 2] - plain) / np.linalg.norm(plain)
    assert rel < 1e-4, rel
    # the auto Gram is symmetric in the shell pair
    assert np.allclose(batched[0, 2], batched[2, 0], rtol=1e-6, atol=1e-6 * np.abs(batched).max())
