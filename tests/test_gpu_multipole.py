"""GPU parity: SH-3PCF contraction (plain FP64 and batched FP32 MFMA) vs the oracle."""

import numpy as np
import pytest

from oracle import multipole as om
from oracle import scaling

pytestmark = pytest.mark.gpu


def test_mfma_lane_layout():
    from pyhermes_amd import kernels
    assert kernels.mfma_selftest() == 0.0


def _fields(L, n, seed, smooth=False):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        f = rng.normal(size=(L, L, L)) + 1j * rng.normal(size=(L, L, L))
        out.append(f)
    return out


def test_summand_f64_matches_oracle(phi):
    import torch
    from pyhermes_amd import kernels
    L = 16
    rng = np.random.default_rng(1)
    d = rng.normal(size=(L, L, L))
    F1, F2 = _fields(L, 2, 2)
    g = scaling.cal_gamma(phi, 3, 1024)
    got = kernels.multipole_summand_f64(kernels.as_device(d, torch.float64), kernels.as_device(F1, torch.complex128),
                                        kernels.as_device(F2, torch.complex128), g)
    ref = om.summand_c(d, F1, F2, g)
    assert abs(got - ref) < 1e-11 * abs(ref)


@pytest.mark.parametrize("cross,L", [(False, 16), (True, 16), (False, 32), (False, 64)])
def test_gram_f32_matches_oracle(phi, cross, L):
    import torch
    from pyhermes_amd import kernels
    print(f"gram cross={cross} L={L}", flush=True)
    rng = np.random.default_rng(7)
    d = rng.normal(size=(L, L, L))
    U = np.stack(_fields(L, 5, 3))
    V = np.stack(_fields(L, 4, 4)) if cross else None
    g = scaling.cal_gamma(phi, 3, 1024)
    lam, Q = np.linalg.eigh(g)
    Ut = kernels.as_device(U, torch.complex64)
    Vt = kernels.as_device(V, torch.complex64) if cross else None
    got = kernels.multipole_gram_f32(kernels.as_device(d, torch.float32), Ut, Q, lam, V=Vt)
    for i in range(U.shape[0]):
        for j in range(got.shape[1]):
            W = V[j] if cross else np.conj(U[j])
            ref = om.summand_c(d, U[i], W, g).real
            scale = np.sqrt(np.sum(d ** 2) * np.sum(np.abs(U[i]) ** 2) * np.sum(np.abs(W) ** 2) / L ** 3)
            assert abs(got[i, j] - ref) <= 2e-6 * scale, (i, j, got[i, j], ref)


@pytest.mark.parametrize("nneg", [0, 1, 2, 3])
def test_gram_auto_for_any_eigenvalue_signs(nneg):
    """The auto contraction for a symmetric Gamma with 0..3 negative eigenvalues, eigenpairs given unsorted."""
    import torch
    from pyhermes_amd import kernels
    L = 16
    rng = np.random.default_rng(40 + nneg)
    Qm, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    lam = np.array([0.9, 0.3, 0.05]) * np.where(np.arange(3) < nneg, -1.0, 1.0)
    perm = rng.permutation(3)
    lam, Qm = lam[perm], Qm[:, perm]
    g = Qm @ np.diag(lam) @ Qm.T
    d = rng.normal(size=(L, L, L))
    U = np.stack(_fields(L, 4, 50 + nneg))
    got = kernels.multipole_gram_f32(kernels.as_device(d, torch.float32), kernels.as_device(U, torch.complex64),
                                     Qm, lam)
    for i in range(U.shape[0]):
        for j in range(U.shape[0]):
            W = np.conj(U[j])
            ref = om.summand_c(d, U[i], W, g).real
            scale = np.sqrt(np.sum(d ** 2) * np.sum(np.abs(U[i]) ** 2) * np.sum(np.abs(W) ** 2) / L ** 3)
            assert abs(got[i, j] - ref) <= 2e-6 * scale, (nneg, i, j, got[i, j], ref)


@pytest.mark.parametrize("L,lmax", [(16, 3)])
def test_ddd_paths_match_oracle(phi, L, lmax):
    import torch
    from pyhermes_amd import kernels, multipole
    from pyhermes_amd.mocks import lognormal_mock
    from oracle import projection as oproj
    box = 200.0
    J = L.bit_length() - 1
    pos, _ = lognormal_mock(20000, 16, box, 4000)
    w = np.full(pos.shape[0], 1.0 / pos.shape[0])
    eps, _, _ = oproj.project_c(pos, w, phi, J, box)
    delta = eps - eps.mean()
    geo = multipole.MultipoleGeometry(phi, L)
    radii = np.array([2.0, 3.5, 5.0])
    ref = np.zeros((3, 3, lmax + 1))
    for i in range(3):
        for j in range(3):
            ref[i, j] = om.calc_ddd_multipole(delta, delta, delta, phi, radii[i], radii[j], 0, lmax)
    dd = kernels.as_device(delta, torch.float64)
    plain = np.array(multipole.ddd_multipole_plain_f64(geo, dd, dd, dd, radii[0], radii[2], 0, lmax))
    assert np.linalg.norm(plain - ref[0, 2]) <= 1e-10 * np.linalg.norm(ref[0, 2])
    d32 = kernels.as_device(delta, torch.float32)
    batched = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, lmax)
    for i in range(3):
        for j in range(3):
            rel = np.linalg.norm(batched[i, j] - ref[i, j]) / np.linalg.norm(ref[i, j])
            assert rel < 1e-4, (i, j, rel)
    # cross path (separate field 3) gives the same numbers
    d32b = d32.clone()
    cross = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32b, radii, radii, 0, lmax)
    assert np.linalg.norm(cross - batched) <= 1e-4 * np.linalg.norm(batched)
    # more than 8 shells take the unpacked auto contraction (one (l, m) per Gram): same numbers
    radii9 = np.concatenate([radii, [1.5, 2.5, 3.0, 4.0, 4.5, 6.0]])
    wide = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii9, radii9, 0, lmax)
    assert np.linalg.norm(wide[:3, :3] - batched) <= 1e-5 * np.linalg.norm(batched)


def test_schedules_and_split_window_agree(phi):
    """The split window entry points reproduce the fused one bit for bit, and the three stream
    schedules of the batched auto path give identical multipoles."""
    import torch
    from pyhermes_amd import kernels, multipole
    L = 32
    rng = np.random.default_rng(21)
    delta = rng.normal(size=(L, L, L))
    d32 = kernels.as_device(delta - delta.mean(), torch.float32)
    geo = multipole.MultipoleGeometry(phi, L)
    H = kernels.rfftn(d32)
    radii = np.array([2.0, 3.5, 5.0, 7.0])
    fused = torch.empty((2 * radii.size, L, L, L), dtype=torch.complex64, device=d32.device)
    kernels.multipole_window_ifft_multi_f32(H, geo.P, radii, 3, [1, 2], 1.0 / L ** 3, fused)
    split = torch.empty_like(fused)
    kernels.multipole_window_zpass_multi_f32(H, geo.P, radii, 3, [1, 2], 1.0 / L ** 3, split)
    kernels.fft3_inverse_yx_f32(split)
    assert torch.equal(fused, split)
    runs = {s: multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, 4, schedule=s)
            for s in ("serial", "split", "overlap")}
    torch.cuda.synchronize()
    assert np.array_equal(runs["serial"], runs["split"])
    assert np.array_equal(runs["serial"], runs["overlap"])


@pytest.mark.parametrize("L", [32, 128])
def test_packed_m0_window_and_split_gram(phi, L):
    """The packed (l, 0) + (l + 1, 0) window equals the sum of the two single windows; the split Gram's
    real and imaginary blocks are the Grams of the real and imaginary parts.  L = 128 runs the window
    kernel's Nyquist helper lanes."""
    import torch
    from pyhermes_amd import kernels, multipole
    rng = np.random.default_rng(23)
    delta = rng.normal(size=(L, L, L))
    d32 = kernels.as_device(delta - delta.mean(), torch.float32)
    geo = multipole.MultipoleGeometry(phi, L)
    H = kernels.rfftn(d32)
    radii = np.array([1.5, 3.0, 4.5, 6.0, 8.0])
    S = radii.size
    for l in (0, 3, 6):
        print(f"packed window L={L} l={l}", flush=True)
        packed = torch.empty((S, L, L, L), dtype=torch.complex64, device=d32.device)
        kernels.multipole_window_ifft_pack0_f32(H, geo.P, radii, l, 1.0 / L ** 3, packed)
        a = kernels.multipole_window_ifft_f32(H, geo.P, radii, l, 0, 1.0 / L ** 3)
        b = kernels.multipole_window_ifft_f32(H, geo.P, radii, l + 1, 0, 1.0 / L ** 3)
        want = a + b
        assert float((packed - want).abs().max()) <= 2e-6 * float(want.abs().max())
        even, odd = (a, b) if l % 2 == 0 else (b, a)
        # the even degree's field is real, the odd degree's imaginary
        assert float(even.imag.abs().max()) <= 1e-6 * float(even.abs().max())
        assert float(odd.real.abs().max()) <= 1e-6 * float(odd.abs().max())
        blk = torch.zeros(512, dtype=torch.float64, device=d32.device)
        kernels.multipole_gram_f32_dev_ri(d32, packed, geo.Q, geo.lam, blk)
        h = blk.cpu().numpy()
        re_only = torch.complex(packed.real.contiguous(), torch.zeros_like(packed.real))
        im_only = torch.complex(packed.imag.contiguous(), torch.zeros_like(packed.imag))
        g_re = kernels.multipole_gram_f32(d32, re_only, geo.Q, geo.lam)
        g_im = kernels.multipole_gram_f32(d32, im_only, geo.Q, geo.lam)
        # (two instantiations of the kernel, free to contract and order their FP32 arithmetic differently,
        # and the split one moves its imaginary sums into FP64 every plane, the plain one every six: the
        # blocks agree to FP32 rounding of the products and partial sums)
        assert np.allclose(kernels.gram_block_to_matrix(h[:256], S), g_re, rtol=1e-6, atol=1e-6 * np.abs(g_re).max())
        assert np.allclose(kernels.gram_block_to_matrix(h[256:], S), g_im, rtol=1e-6, atol=1e-6 * np.abs(g_im).max())


def test_packed_pipeline_matches_unpacked_and_fp64(phi):
    """With 10 shells the batched auto path packs the m = 0 terms; it agrees with the unpacked (split)
    schedule and with the plain FP64 path, and the packed overlap schedule (windows of one unit beside
    the Grams of the previous one, on two streams) gives the serial one's numbers bit for bit."""
    import torch
    from pyhermes_amd import kernels, multipole
    L = 32
    rng = np.random.default_rng(29)
    delta = rng.normal(size=(L, L, L))
    delta -= delta.mean()
    d32 = kernels.as_device(delta, torch.float32)
    d64 = kernels.as_device(delta, torch.float64)
    geo = multipole.MultipoleGeometry(phi, L)
    radii = np.linspace(1.5, 9.0, 10)
    lmax = 5
    assert multipole.packs_m0(L, radii.size)
    packed = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, lmax)
    unpacked = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, lmax, schedule="split")
    assert not multipole.packs_m0(L, radii.size, "split")
    assert np.linalg.norm(packed - unpacked) <= 1e-5 * np.linalg.norm(unpacked)
    assert multipole.packs_m0(L, radii.size, "overlap")
    overlapped = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, lmax, schedule="overlap")
    torch.cuda.synchronize()
    assert np.array_equal(packed, overlapped)
    plain = np.array(multipole.ddd_multipole_plain_f64(geo, d64, d64, d64, radii[2], radii[7], 0, lmax))
    assert np.linalg.norm(packed[2, 7] - plain) <= 1e-4 * np.linalg.norm(plain)


def test_headline_path_every_set_matches_oracle(phi):
    """The headline's exact code path -- auto case, 15 shells, m = 0 terms packed in pairs, l <= 10, FP32
    batched -- on every (r12, r13) set at L = 32 against the oracle's restatement of calc_DDD_multipole
    (corr3pcf_multipoles.py:460-579): relative L2 error over l <= 1e-4 per set.  The oracle's shell fields
    are built once per (shell, l, m) and its contraction regrouped over all shell pairs (summand_pairs,
    pinned against summand_c on the CPU); one set is also run through calc_ddd_multipole verbatim."""
    import torch
    from oracle import native
    from oracle import projection as oproj
    from pyhermes_amd import kernels, multipole
    from pyhermes_amd.mocks import lognormal_mock
    L, lmax, S = 32, 10, 15
    J = 5
    box = 250.0
    print(f"headline path L={L} S={S} l<={lmax}: product", flush=True)
    pos, _ = lognormal_mock(60000, 32, box, 4100)
    w = np.full(pos.shape[0], 1.0 / pos.shape[0])
    eps, _, _ = oproj.project_c(pos, w, phi, J, box)
    delta = eps / eps.mean() - 1.0
    radii = np.linspace(1.5, 12.0, S)
    geo = multipole.MultipoleGeometry(phi, L)
    assert multipole.packs_m0(L, S)
    d32 = kernels.as_device(delta, torch.float32)
    got = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, lmax)
    print("headline path: oracle fields", flush=True)
    threads = native.num_threads()
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    g = scaling.cal_gamma(phi, 3, 1024)
    fields = {}
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            fields[(l, m)] = [om.multipole_fields(delta, P, R, l, m, threads) for R in radii]
    print("headline path: oracle contraction", flush=True)
    ref = np.zeros((S, S, lmax + 1))
    for l in range(lmax + 1):
        v = [4.0 * np.pi * om.summand_pairs(delta, fields[(l, m)], fields[(l, -m)], g) / L ** 3
             for m in range(l + 1)]
        for i in range(S):
            for j in range(S):
                ref[i, j, l] = om.combine_multipole_m_terms([vm[i, j] for vm in v], l)
    one = om.calc_ddd_multipole(delta, delta, delta, phi, radii[0], radii[9], 0, lmax, workers=threads)
    assert np.linalg.norm(ref[0, 9] - one) <= 1e-10 * np.linalg.norm(one)
    worst = 0.0
    for i in range(S):
        for j in range(S):
            rel = np.linalg.norm(got[i, j] - ref[i, j]) / np.linalg.norm(ref[i, j])
            worst = max(worst, rel)
            assert rel <= 1e-4, (i, j, rel)
    print(f"headline path: worst relative L2 error over 225 ordered sets {worst:.2e}", flush=True)


def test_gram_reads_ymajor_fields(phi):
    """Auto, cross and split Gram sums over y-major fields (hermes_set_field_layout(1)) equal those over the
    same fields in x-major order, bit for bit; and the packed auto path in y-major order (the default)
    agrees with its in-place x-major order to FP32 rounding."""
    import torch
    from pyhermes_amd import kernels, multipole
    L = 64
    rng = np.random.default_rng(41)
    d = kernels.as_device(rng.normal(size=(L, L, L)), torch.float32)
    U = kernels.as_device(rng.normal(size=(5, L, L, L)) + 1j * rng.normal(size=(5, L, L, L)), torch.complex64)
    V = kernels.as_device(rng.normal(size=(4, L, L, L)) + 1j * rng.normal(size=(4, L, L, L)), torch.complex64)
    Ut, Vt = U.transpose(1, 2).contiguous(), V.transpose(1, 2).contiguous()
    lam, Q = np.linalg.eigh(scaling.cal_gamma(phi, 3, 1024))
    blk_x = torch.zeros(512, dtype=torch.float64, device=U.device)
    blk_y = torch.zeros_like(blk_x)
    auto_x = kernels.multipole_gram_f32(d, U, Q, lam)
    cross_x = kernels.multipole_gram_f32(d, U, Q, lam, V=V, v_conj=True)
    kernels.multipole_gram_f32_dev_ri(d, U, Q, lam, blk_x)
    with kernels.field_layout("ymajor"):
        auto_y = kernels.multipole_gram_f32(d, Ut, Q, lam)
        cross_y = kernels.multipole_gram_f32(d, Ut, Q, lam, V=Vt, v_conj=True)
        kernels.multipole_gram_f32_dev_ri(d, Ut, Q, lam, blk_y)
    assert kernels.get_field_layout() == "xmajor"
    assert np.array_equal(auto_x, auto_y) and np.array_equal(cross_x, cross_y)
    assert torch.equal(blk_x, blk_y)
    geo = multipole.MultipoleGeometry(phi, L)
    delta = rng.normal(size=(L, L, L))
    d32 = kernels.as_device(delta - delta.mean(), torch.float32)
    radii = np.linspace(1.5, 12.0, 10)
    assert multipole.packs_m0(L, radii.size)
    ym = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, 6)
    xm = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, 6, field_order="xmajor")
    assert np.linalg.norm(ym - xm) <= 1e-6 * np.linalg.norm(xm)
