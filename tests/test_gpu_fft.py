"""GPU parity: LDS-tiled 3-D FFT and the fused multipole window + inverse FFT."""

import numpy as np
import pytest

from oracle import scaling
from oracle import windows as ow

pytestmark = pytest.mark.gpu


def _rand_complex(shape, seed):
    g = np.random.default_rng(seed)
    return (g.normal(size=shape) + 1j * g.normal(size=shape)).astype(np.complex64)


@pytest.mark.parametrize("J", [3, 4, 5, 6, 7])
@pytest.mark.parametrize("inverse", [False, True])
def test_fft3_matches_numpy(J, inverse):
    import torch
    from pyhermes_amd import kernels
    L = 1 << J
    print(f"fft3 J={J} inverse={inverse}", flush=True)
    x = _rand_complex((2, L, L, L), 10 + J)
    d = kernels.as_device(x, torch.complex64)
    kernels.fft3_c2c_f32(d, inverse=inverse)
    got = d.cpu().numpy()
    xd = x.astype(np.complex128)
    ref = np.fft.ifftn(xd, axes=(1, 2, 3)) * L ** 3 if inverse else np.fft.fftn(xd, axes=(1, 2, 3))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err <= 3e-6, err


@pytest.mark.parametrize("J", [3, 4, 5, 6, 7])
@pytest.mark.parametrize("l,m", [(0, 0), (1, 1), (3, -2), (6, 6), (10, 3)])
def test_window_ifft_matches_unfused_path(phi, J, l, m):
    """The fused window + inverse FFT against the separate apply and rocFFT; J = 7 (L = 128) runs the
    branch where the Nyquist plane z = L/2 is evaluated by helper lanes."""
    import torch
    from pyhermes_amd import kernels
    L = 1 << J
    print(f"window_ifft J={J} l={l} m={m}", flush=True)
    g = np.random.default_rng(J * 100 + l * 10 + m + 50)
    field = g.normal(size=(L, L, L)).astype(np.float32)
    H = kernels.rfftn(kernels.as_device(field, torch.float32))
    P = kernels.as_device(scaling.fourier_power_spectrum(phi, 0, 1, L, 1024), torch.float64)
    radii = np.array([1.5, 2.25, 0.4 * L, 3.1])
    scale = 1.0 / L ** 3
    ref = kernels.multipole_apply_f32(H, P, radii, l, m, scale)
    kernels.fftn_c2c(ref, inverse=True, out=ref)
    got = kernels.multipole_window_ifft_f32(H, P, radii, l, m, scale)
    r, gnp = ref.cpu().numpy(), got.cpu().numpy()
    for s in range(radii.size):
        err = np.abs(gnp[s] - r[s]).max() / max(np.abs(r[s]).max(), 1e-30)
        assert err <= 3e-6, (s, err)


@pytest.mark.parametrize("J,l,m", [(4, 0, 0), (4, 2, 1), (4, 4, -3), (7, 0, 0), (7, 5, 2), (7, 10, -7)])
def test_window_ifft_matches_oracle(phi, J, l, m):
    """Against the oracle's window builder and scipy convolution; J = 7 covers the Nyquist helper lanes."""
    import torch
    from pyhermes_amd import kernels
    L = 1 << J
    g = np.random.default_rng(77 + l)
    field = g.normal(size=(L, L, L))
    H = kernels.rfftn(kernels.as_device(field, torch.float32))
    Pn = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    P = kernels.as_device(Pn, torch.float64)
    radii = np.array([2.0, 3.3, 0.3 * L])
    got = kernels.multipole_window_ifft_f32(H, P, radii, l, m, 1.0 / L ** 3).cpu().numpy()
    for s, R in enumerate(radii):
        ref = ow.convolve_complex(field, ow.radial_multipole_window(L, Pn, R, l, m))
        err = np.abs(got[s] - ref).max() / np.abs(ref).max()
        assert err <= 2e-5, (s, err)


def test_window_ifft_rejects_bad_shapes(phi):
    import torch
    from pyhermes_amd import kernels
    L = 16
    H = torch.zeros((L, L, L), dtype=torch.complex64, device="cuda")  # full, not half spectrum
    P = kernels.as_device(np.ones(L + 1), torch.float64)
    with pytest.raises(ValueError):
        kernels.multipole_window_ifft_f32(H, P, [1.0], 0, 0, 1.0)
    with pytest.raises(ValueError):
        kernels.fft3_c2c_f32(torch.zeros((4, 4, 8), dtype=torch.complex64, device="cuda"))


@pytest.mark.parametrize("J", [7, 8])
def test_every_strided_tile_matches_numpy(J):
    """Each candidate tile of the strided passes, pinned in turn, gives the same transform."""
    import torch
    from pyhermes_amd import kernels
    L = 1 << J
    base = min(8192 // L, L)
    x = _rand_complex((1, L, L, L), 300 + J)
    ref = np.fft.ifftn(x.astype(np.complex128), axes=(1, 2, 3)) * L ** 3
    scale = np.abs(ref).max()
    for tw, nt in ((base, 256), (base, 512), (base // 2, 256), (base // 2, 512), (2 * base, 512), (2 * base, 1024)):
        if True:
            print(f"strided tile J={J} tw={tw} threads={nt}", flush=True)
            kernels.fft3_set_tile(J, tw, nt, inverse=True)
            d = kernels.as_device(x, torch.complex64)
            kernels.fft3_c2c_f32(d, inverse=True)
            err = np.abs(d.cpu().numpy() - ref).max() / scale
            assert err <= 3e-6, (tw, nt, err)
    with pytest.raises(RuntimeError):
        kernels.fft3_set_tile(J, 3, 256)


def test_autotuner_records_a_candidate():
    import torch
    from pyhermes_amd import kernels
    J = 6
    d = torch.zeros((2, 1 << J, 1 << J, 1 << J), dtype=torch.complex64, device="cuda")
    kernels.fft3_c2c_f32(d, inverse=False)
    tw, nt = kernels.fft3_tile_choice(J, inverse=False)
    assert (tw, nt) == (1 << J, 256)


@pytest.mark.parametrize("J", [4, 6, 7])
def test_window_ifft_order_pairs_match_single_orders(phi, J):
    """Two orders of one l in one launch (shared radial factors) equal two single-order launches to FP32
    rounding (the two instantiations may contract multiply-adds differently)."""
    import torch
    from pyhermes_amd import kernels
    L = 1 << J
    g = np.random.default_rng(90 + J)
    H = kernels.rfftn(kernels.as_device(g.normal(size=(L, L, L)).astype(np.float32), torch.float32))
    P = kernels.as_device(scaling.fourier_power_spectrum(phi, 0, 1, L, 1024), torch.float64)
    radii = np.array([1.5, 0.3 * L, 3.25])
    S = radii.size
    for l, ms in ((3, [1, 2]), (4, [0, -3]), (10, [9, 10])):
        print(f"window pair J={J} l={l} m={ms}", flush=True)
        out = torch.empty((2 * S, L, L, L), dtype=torch.complex64, device=H.device)
        kernels.multipole_window_ifft_multi_f32(H, P, radii, l, ms, 1.0 / L ** 3, out)
        for i, m in enumerate(ms):
            one = kernels.multipole_window_ifft_f32(H, P, radii, l, m, 1.0 / L ** 3)
            err = float((out[i * S:(i + 1) * S] - one).abs().max() / one.abs().max())
            print(f"  m={m}: max rel diff {err:.2e}", flush=True)
            assert err <= 1e-6, (l, m, err)


def test_inverse_yx_streaming_passes_match_numpy():
    """From 256^3 up the strided passes use streaming loads and stores: the in-place y and x inverse passes at
    J=8 against numpy's unnormalised 2-D inverse transform over (x, y)."""
    import torch
    from pyhermes_amd import kernels
    J = 8
    L = 1 << J
    print(f"inverse yx passes vs numpy J={J}", flush=True)
    h = _rand_complex((1, L, L, L), 808)
    x = kernels.as_device(h, torch.complex64)
    kernels.fft3_inverse_yx_f32(x)
    want = np.fft.ifft2(h.astype(np.complex128), axes=(1, 2)) * (L * L)
    got = x.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


@pytest.mark.parametrize("J", [5, 7, 8])
def test_inverse_yx_into_ymajor_order(J):
    """The strided passes from x-major grids into y-major order give the in-place x-major transform,
    transposed in (x, y); overlapping grids are refused."""
    import torch
    from pyhermes_amd import kernels
    L = 1 << J
    print(f"inverse yx into y-major order J={J}", flush=True)
    x = kernels.as_device(_rand_complex((3, L, L, L), 500 + J), torch.complex64)
    ref = x.clone()
    kernels.fft3_inverse_yx_f32(ref)
    out = torch.empty_like(x)
    kernels.fft3_inverse_yx_ymajor_f32(x, out)
    want = ref.transpose(1, 2)  # [b, y, x, z]
    scale = float(ref.abs().max())
    assert float((out - want).abs().max()) <= 2e-6 * scale
    with pytest.raises(RuntimeError):
        kernels.fft3_inverse_yx_ymajor_f32(x[:2], x[1:])
