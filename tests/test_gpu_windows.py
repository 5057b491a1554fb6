"""GPU parity: window kernels, fused pair reduction and fused multipole apply vs the oracle."""

import numpy as np
import pytest

from oracle import scaling, windows as ow

pytestmark = pytest.mark.gpu

ISO = [("sphere", 2.5, 0.0), ("gaussian", 1.5, 0.0), ("shell", 3.0, 0.0), ("thick_shell", 3.0, 1.5),
       ("gaussian_shell", 3.0, 1.0), ("inverse_laplacian", 0.0, 0.0)]


def _P(phi, L):
    import torch
    from pyhermes_amd import kernels
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    return P, kernels.as_device(P, torch.float64)


@pytest.mark.parametrize("wt,a0,a1", ISO)
def test_build_rfft_matches_octant_oracle(phi, wt, a0, a1):
    from pyhermes_amd import kernels
    J = 4
    L = 1 << J
    P, Pd = _P(phi, L)
    got = kernels.window_build_rfft(J, Pd, kernels.window_desc(wt, a0, a1), dbl=True).cpu().numpy()
    ref = ow.octant_rfft_kernel(L, P, wt, a0, a1)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    got32 = kernels.window_build_rfft(J, Pd, kernels.window_desc(wt, a0, a1), dbl=False).cpu().numpy()
    assert np.abs(got32 - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("l,m", [(0, 0), (1, 1), (2, -1), (3, 2), (5, 0), (10, 7)])
def test_build_full_multipole_matches_oracle(phi, l, m):
    from pyhermes_amd import kernels
    J = 4
    L = 1 << J
    P, Pd = _P(phi, L)
    R = 3.7
    got = kernels.window_build_full(J, Pd, kernels.window_desc("shell_multipole", R, l=l, m=m)).cpu().numpy()
    ref = ow.radial_multipole_window(L, P, R, l, m)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("dbl", [True, False])
def test_pair_reduce_matches_convolution_oracle(phi, dbl):
    import torch
    from pyhermes_amd import kernels
    J = 5
    L = 1 << J
    P, Pd = _P(phi, L)
    rng = np.random.default_rng(3)
    f1 = rng.normal(size=(L, L, L))
    f2 = f1 + 0.3 * rng.normal(size=(L, L, L))
    td = torch.float64 if dbl else torch.float32
    F1 = kernels.rfftn(kernels.as_device(f1, td))
    F2 = kernels.rfftn(kernels.as_device(f2, td))
    descs = [kernels.window_desc("shell", R) for R in (1.5, 3.0, 5.25, 9.0)]
    descs += [kernels.window_desc("sphere", 2.0), kernels.window_desc("gaussian_shell", 3.0, 1.0)]
    descs += [kernels.window_desc("shell_multipole", R, l=l) for R in (2.0, 6.5) for l in (0, 2, 4)]
    got = kernels.pair_reduce(F1, F2, Pd, descs) / L ** 6
    tol = 1e-11 if dbl else 2e-5
    for d, g in zip(descs, got):
        if d.type == 6:
            K = ow.radial_multipole_window(L, P, d.a0, d.l, 0)
            conv = np.real(ow.convolve_complex(f1, K))
            ref = float(np.sum(conv * f2) / L ** 3)
        else:
            wt = [k for k, v in kernels.WIN_CODES.items() if v == d.type][0]
            ref = ow.pair_product(f1, f2, ow.octant_rfft_kernel(L, P, wt, d.a0, d.a1))
        scale = np.sqrt(np.mean(f1 ** 2) * np.mean(f2 ** 2))
        assert abs(g - ref) <= tol * scale, (d.type, d.a0, d.l, g, ref)


@pytest.mark.parametrize("case", ["shells", "ell4", "ell10"])
@pytest.mark.parametrize("J", [5, 6])
def test_pair_reduce_radial_path(phi, case, J):
    """Shell / m = 0 multipole lists take the radius-grouped FP32 kernel: it matches the
    convolution oracle, the generic FP32 kernel and the FP64 kernel."""
    import torch
    from pyhermes_amd import kernels
    L = 1 << J
    P, Pd = _P(phi, L)
    rng = np.random.default_rng(30 + J)
    f1 = rng.normal(size=(L, L, L))
    f2 = f1 + 0.5 * rng.normal(size=(L, L, L))
    radii = (0.7, 1.5, 3.0, 0.45 * L)
    if case == "shells":
        descs = [kernels.window_desc("shell", R) for R in radii]
    elif case == "ell4":
        descs = [kernels.window_desc("shell", R) for R in radii]
        descs += [kernels.window_desc("shell_multipole", R, l=l) for l in (0, 2, 4) for R in radii]
    else:
        descs = [kernels.window_desc("shell_multipole", R, l=l) for l in (2, 6, 8, 10) for R in radii[1:]]
    print(f"radial pair reduce J={J} case={case}", flush=True)
    F1 = kernels.rfftn(kernels.as_device(f1, torch.float32))
    F2 = kernels.rfftn(kernels.as_device(f2, torch.float32))
    got = kernels.pair_reduce(F1, F2, Pd, descs)
    generic = kernels.pair_reduce(F1, F2, Pd, descs + [kernels.window_desc("sphere", 2.0)])[:-1]
    G1, G2 = F1.to(torch.complex128), F2.to(torch.complex128)
    fp64 = kernels.pair_reduce(G1, G2, Pd, descs)
    fp64_generic = kernels.pair_reduce(G1, G2, Pd, descs + [kernels.window_desc("sphere", 2.0)])[:-1]
    scale = np.sqrt(np.mean(f1 ** 2) * np.mean(f2 ** 2)) * L ** 6
    assert np.abs(fp64 - fp64_generic).max() <= 1e-11 * scale
    assert np.abs(got - generic).max() <= 1e-5 * scale
    assert np.abs(got - fp64).max() <= 1e-5 * scale
    for d, g in zip(descs[::3], got[::3]):
        K = ow.radial_multipole_window(L, P, d.a0, d.l, 0) if d.type == 6 else ow.octant_rfft_kernel(
            L, P, "shell", d.a0, 0.0)
        ref = (float(np.sum(np.real(ow.convolve_complex(f1, K)) * f2) / L ** 3) if d.type == 6
               else ow.pair_product(f1, f2, K))
        assert abs(g / L ** 6 - ref) <= 2e-5 * scale / L ** 6, (d.type, d.a0, d.l)


@pytest.mark.parametrize("J", [5, 6])
def test_pair_reduce_uniform_shells_fp64(phi, J):
    """FP64 thin shells at equally spaced radii (linear binning) take the kernel with rotated phases:
    it matches the same kernel without rotation (one radius moved by 1e-9, which turns it off) and the
    convolution oracle to FP64 rounding."""
    import torch
    from pyhermes_amd import kernels
    L = 1 << J
    P, Pd = _P(phi, L)
    rng = np.random.default_rng(70 + J)
    f1 = rng.normal(size=(L, L, L))
    f2 = f1 + 0.5 * rng.normal(size=(L, L, L))
    radii = np.linspace(0.05, 0.45 * L, 20)
    bumped = radii.copy()
    bumped[7] *= 1.0 + 1e-9
    print(f"uniform FP64 shells J={J}", flush=True)
    F1 = kernels.rfftn(kernels.as_device(f1, torch.float64))
    F2 = kernels.rfftn(kernels.as_device(f2, torch.float64))
    got = kernels.pair_reduce(F1, F2, Pd, [kernels.window_desc("shell", R) for R in radii])
    plain = kernels.pair_reduce(F1, F2, Pd, [kernels.window_desc("shell", R) for R in bumped])
    scale = np.sqrt(np.mean(f1 ** 2) * np.mean(f2 ** 2)) * L ** 6
    keep = np.arange(20) != 7
    assert np.abs(got - plain)[keep].max() <= 1e-12 * scale
    assert abs(got[7] - plain[7]) <= 1e-6 * scale
    for i in (0, 1, 10, 19):
        ref = ow.pair_product(f1, f2, ow.octant_rfft_kernel(L, P, "shell", radii[i], 0.0))
        assert abs(got[i] / L ** 6 - ref) <= 1e-11 * scale / L ** 6, (i, radii[i])


# line of sight: a coordinate axis (0, 1, 2) or a direction vector (axis 3 descriptors)
LOS_CASES = [("ring", 1.5, 2.5, 2), ("disk", 3.0, 1.25, 2), ("cylshell", 2.25, 3.5, 0), ("cylinder", 6.0, 0.75, 1),
             ("ring", 40.0, 9.0, 2),  # reaches J0/J1 arguments of several hundred (asymptotic range)
             ("ring", 2.5, 3.0, (1.0, 1.0, 0.0)), ("disk", 1.75, 2.0, (1.0, -2.0, 2.0)),
             ("cylshell", 3.0, 1.5, (0.3, 0.0, -1.0)), ("cylinder", 2.0, 4.0, (-1.0, 2.0, 0.5))]


def _los_case(wt, R, H, spec):
    """(descriptor, line-of-sight vector) of one LOS case."""
    from pyhermes_amd import kernels
    if isinstance(spec, tuple):
        n = np.asarray(spec) / np.linalg.norm(spec)
        return kernels.window_desc(wt, R, H, axis=3, los=n), tuple(spec)
    return kernels.window_desc(wt, R, H, axis=spec), tuple(1.0 if i == spec else 0.0 for i in range(3))


@pytest.mark.parametrize("wt,R,H,axis", LOS_CASES)
def test_build_rfft_los_windows_match_oracle(phi, wt, R, H, axis):
    """Line-of-sight windows (window_functions.py:164-319) in the rFFT layout against the oracle's signed
    8-alias sums (convolution.py:73-157) with scipy's J0/J1, FP64 and FP32."""
    from pyhermes_amd import kernels
    J = 5
    L = 1 << J
    P, Pd = _P(phi, L)
    desc, los = _los_case(wt, R, H, axis)
    ref = ow.los_rfft_kernel(L, P, wt, R, H, los)
    got = kernels.window_build_rfft(J, Pd, desc, dbl=True).cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), np.abs(got - ref).max()
    got32 = kernels.window_build_rfft(J, Pd, desc, dbl=False).cpu().numpy()
    assert np.abs(got32 - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("dbl", [True, False])
def test_pair_reduce_los_windows_match_oracle(phi, dbl):
    """The fused pair pass over ring / disk / cylshell / cylinder windows (several axes) equals the oracle's
    convolution pair products."""
    import torch
    from pyhermes_amd import kernels
    J = 5
    L = 1 << J
    P, Pd = _P(phi, L)
    rng = np.random.default_rng(17)
    f1 = rng.normal(size=(L, L, L))
    f2 = f1 + 0.4 * rng.normal(size=(L, L, L))
    td = torch.float64 if dbl else torch.float32
    F1 = kernels.rfftn(kernels.as_device(f1, td))
    F2 = kernels.rfftn(kernels.as_device(f2, td))
    print(f"LOS pair reduce J={J} dbl={dbl}", flush=True)
    cases = [_los_case(*c) for c in LOS_CASES]
    got = kernels.pair_reduce(F1, F2, Pd, [d for d, _ in cases])
    got /= L ** 6
    scale = np.sqrt(np.mean(f1 ** 2) * np.mean(f2 ** 2))
    tol = 1e-11 if dbl else 2e-5
    for (wt, R, H, a), (_, los), g in zip(LOS_CASES, cases, got):
        ref = ow.pair_product(f1, f2, ow.los_rfft_kernel(L, P, wt, R, H, los))
        assert abs(g - ref) <= tol * scale, (wt, R, H, a, g, ref)


def test_rejects_bad_los_axis():
    import torch
    from pyhermes_amd import kernels
    L = 16
    Pd = kernels.as_device(np.ones(L + 1), torch.float64)
    bad = kernels.window_desc("ring", 1.0, 1.0)
    bad.axis = 4
    with pytest.raises(RuntimeError, match="axis"):
        kernels.window_build_rfft(4, Pd, bad)
    bad.axis = 3  # with a zero los vector
    with pytest.raises(RuntimeError, match="unit vector"):
        kernels.window_build_rfft(4, Pd, bad)


def test_pair_reduce_rejects_odd_multipoles():
    import torch
    from pyhermes_amd import kernels
    L = 16
    F = torch.zeros((L, L, L // 2 + 1), dtype=torch.complex64, device="cuda")
    Pd = kernels.as_device(np.ones(L + 1), torch.float64)
    with pytest.raises(RuntimeError):
        kernels.pair_reduce(F, F, Pd, [kernels.window_desc("shell_multipole", 2.0, l=3)])


@pytest.mark.parametrize("l,m", [(0, 0), (2, 1), (3, -2), (6, 6)])
def test_multipole_apply_matches_oracle(phi, l, m):
    import torch
    from pyhermes_amd import kernels
    J = 4
    L = 1 << J
    P, Pd = _P(phi, L)
    rng = np.random.default_rng(5)
    d = rng.normal(size=(L, L, L))
    H = kernels.rfftn(kernels.as_device(d, torch.float32))
    radii = [1.25, 2.5, 4.0, 6.75, 7.9]
    out = kernels.multipole_apply_f32(H, Pd, radii, l, m, 1.0 / L ** 3).cpu().numpy()
    Fd = np.fft.fftn(d)
    for s, R in enumerate(radii):
        ref = Fd * ow.radial_multipole_window(L, P, R, l, m) / L ** 3
        err = np.abs(out[s] - ref).max() / np.abs(ref).max()
        assert err < 2e-5, (s, R, err)


def test_reduce_sum(phi):
    import torch
    from pyhermes_amd import kernels
    rng = np.random.default_rng(0)
    x = rng.normal(size=(17, 33, 65))
    assert kernels.reduce_sum(kernels.as_device(x, torch.float64)) == pytest.approx(x.sum(), rel=1e-13, abs=1e-12)
    c = x + 1j * rng.normal(size=x.shape)
    s = kernels.reduce_sum(kernels.as_device(c, torch.complex64))
    assert abs(s - c.sum()) < 1e-4 * np.abs(c).sum()
