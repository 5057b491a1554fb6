"""GPU parity of the non-thin radial profiles of multipole windows (radial_multipole_windows.py:186-316,
radial_profiles.py): the device builders against the oracle's restatement, and the SH-3PCF task with
sphere / thick-shell / Gaussian / Gaussian-shell edges on the FP64 and FP32 paths."""

import numpy as np
import pytest

from oracle import multipole as om
from oracle import radial as orad
from oracle import scaling

pytestmark = pytest.mark.gpu

CASES = [("sphere", {"R": 30.0}), ("thick_shell", {"R": 45.0, "delta_R": 15.0}), ("gaussian", {"R": 12.0}),
         ("gaussian_shell", {"R_shell": 40.0, "R_smooth": 8.0}), ("gaussian_shell", {"R_shell": 40.0, "R_smooth": 0.0})]


@pytest.mark.parametrize("rtype,args", CASES)
@pytest.mark.parametrize("l,m", [(0, 0), (2, 1), (5, -3)])
def test_build_full_profile_matches_oracle(phi, rtype, args, l, m):
    """Closed-form monopoles, thin shells and Hankel-table windows (FP64 and FP32) against the C oracle."""
    import torch
    from pyhermes_amd import kernels
    from pyhermes_amd.edges import EdgeProfile
    J, box = 4, 400.0
    L = 1 << J
    print(f"profile window {rtype} {args} l={l} m={m}", flush=True)
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    Pd = kernels.as_device(P, torch.float64)
    edge = EdgeProfile(rtype, args, {}, L, box)
    ref = orad.profile_multipole_window(L, P, rtype, args, l, m, box)
    got = edge.window_full(J, Pd, l, m, dbl=True).cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    got32 = edge.window_full(J, Pd, l, m, dbl=False).cpu().numpy()
    assert np.abs(got32 - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("J", [4, 7])
def test_fused_profile_window_matches_unfused(phi, J):
    """The fused window + inverse FFT with per-shell profiles (mixed kinds in one launch, two orders sharing
    the radial factors) against the full-grid builder, the spectral product and rocFFT; J = 7 runs the
    Nyquist helper lanes."""
    import torch
    from pyhermes_amd import kernels
    from pyhermes_amd.edges import EdgeProfile
    L, box = 1 << J, 300.0
    print(f"fused profile window J={J}", flush=True)
    g = np.random.default_rng(60 + J)
    H = kernels.rfftn(kernels.as_device(g.normal(size=(L, L, L)).astype(np.float32), torch.float32))
    Pd = kernels.as_device(scaling.fourier_power_spectrum(phi, 0, 1, L, 1024), torch.float64)
    edges = [EdgeProfile("sphere", {"R": 25.0}, {}, L, box), EdgeProfile("shell", {"R": 40.0}, {}, L, box),
             EdgeProfile("thick_shell", {"R": 50.0, "delta_R": 20.0}, {}, L, box),
             EdgeProfile("gaussian_shell", {"R_shell": 30.0, "R_smooth": 6.0}, {}, L, box)]
    S = len(edges)
    Hc = torch.fft.fftn(torch.fft.irfftn(H.to(torch.complex128), s=(L, L, L)).to(torch.complex128))
    for l, ms in ((0, [0]), (3, [1, -2]), (6, [6])):
        out = torch.empty((len(ms) * S, L, L, L), dtype=torch.complex64, device=H.device)
        kernels.multipole_window_ifft_profiles_f32(H, Pd, [e.profile(l) for e in edges], l, ms, 1.0 / L ** 3, out)
        for i, m in enumerate(ms):
            for s, e in enumerate(edges):
                K = e.window_full(J, Pd, l, m, dbl=True)
                ref = torch.fft.ifftn(Hc * K) * 1.0
                got = out[i * S + s].to(torch.complex128)
                err = float((got - ref).abs().max() / ref.abs().max())
                assert err <= 3e-5, (l, m, s, err)


def test_window_func_radial_multipole_profiles(phi):
    """WindowFunc type radial_multipole with a non-thin radial_type (window.py:212-238) equals the oracle's
    kernel; a user real-space top-hat profile gives the sphere's kernel."""
    import torch
    from pyhermes_amd.io import WindowFunc
    from pyhermes_amd.utils.radial_profiles import real_space_profile
    J, box = 4, 400.0
    L = 1 << J
    info = {"J": J, "box_size": box, "phi_resolution": 1024, "wavelet_mode": "db2", "wavelet_level": 10}
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    for rtype, args in CASES[:3]:
        w = WindowFunc({"type": "radial_multipole", "len_args": dict(args),
                        "other_args": {"radial_type": rtype, "l": 3, "m": 2}}, info)
        ref = orad.profile_multipole_window(L, P, rtype, args, 3, 2, box)
        assert np.abs(w.as_array() - ref).max() <= 1e-12 * np.abs(ref).max()
    prof = real_space_profile(lambda r, R: (r <= R).astype(float), 30.0)
    w = WindowFunc({"type": "radial_multipole", "len_args": {"R": 30.0},
                    "other_args": {"radial_type": "custom_real", "l": 2, "m": 0,
                                   "profile_config": {"profile": prof}}}, info)
    ref = orad.profile_multipole_window(L, P, "sphere", {"R": 30.0}, 2, 0, box)
    assert np.abs(w.as_array() - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("backend,rtol", [("gpu_fp64", 1e-9), ("gpu", 1e-4)])
def test_multipole_task_with_profile_edges(phi, backend, rtol):
    """Corr_3PCF_Multipole with sphere edges for r12 and Gaussian-shell edges for r13 (grid sampling) against
    the oracle's calc_DDD_multipole restatement built on the profile windows (corr3pcf_multipoles.py:460-579);
    the table diagnostics are recorded in the result."""
    from pyhermes_amd.base import SFCProjection
    from pyhermes_amd.mocks import lognormal_mock
    from pyhermes_amd.theory import Corr_3PCF_Multipole
    print(f"multipole task with profile edges, {backend}", flush=True)
    box, J = 200.0, 4
    L = 1 << J
    pos, _ = lognormal_mock(20000, 16, box, 4300)
    field = SFCProjection({"SFCProjection": {"J": J, "box_size": box, "particle_pos": pos,
                                             "catalog_weight": np.ones(pos.shape[0]), "fout_path": ""}}).run(
        save_result=False)
    lmax = 3
    res = Corr_3PCF_Multipole({"Corr_3PCF_Multipole": {
        "sfc_field": field, "random": "uniform", "l_min": 0, "l_max": lmax, "products": ["delta_ddd_l"],
        "summation_backend": backend, "fout_path": "",
        "sampling": {"mode": "grid", "r12": {"values": [30.0, 45.0]}, "r13": {"values": [35.0]}},
        "binning_window12": {"type": "sphere", "mapping": {"R": "r12"}},
        "binning_window13": {"type": "gaussian_shell", "len_args": {"R_smooth": 6.0},
                             "mapping": {"R_shell": "r13"}}}}).run(save_result=False)
    eps = field.epsilon
    delta = eps - field.field_integral / L ** 3
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    gamma = scaling.cal_gamma(phi, 3, 1024)
    scale = (L / box) ** 9
    for k, r12 in enumerate((30.0, 45.0)):
        ref = []
        for l in range(lmax + 1):
            v = []
            for m in range(l + 1):
                F1 = orad.profile_fields(delta, P, "sphere", {"R": r12}, l, m, box)
                F2 = orad.profile_fields(delta, P, "gaussian_shell", {"R_shell": 35.0, "R_smooth": 6.0}, l, -m, box)
                v.append(om.m_summand(delta, F1, F2, gamma))
            ref.append(om.combine_multipole_m_terms(v, l))
        ref = np.array(ref) * scale
        err = np.linalg.norm(res.delta_ddd_l[k] - ref) / np.linalg.norm(ref)
        assert err <= rtol, (k, err)
    diags = res.corr3pcf_multipole_info["radial_profile_diagnostics_result"]
    assert len(diags) == 2 * lmax * 2 and all(d["passed"] for d in diags)
