"""Headline benchmark: spherical-harmonic 3PCF multipole sets per second at J=9, l <= 10.

Workload (synthetic): a 2,000,000-particle seeded lognormal mock (Gaussian field
with P(k) ~ k^-1.5 on a 256^3 grid, sigma_LN = 1), box 1000 Mpc/h, projected at
J=9 (512^3), delta = eps - rho; 15 thin radial shells R = linspace(10, 120, 15)
Mpc/h for both edges; multipoles l = 0..10 (66 m-terms per set).

A "set" is one row of delta_ddd_l (l = 0..10) for one (r12, r13) pair; the 15 shells
give 120 unique (r12 <= r13) sets (225 ordered).  One step = one pass of the
FP32 batched path over all pairs (one r2c FFT, then per (l, m) one fused
window + inverse-FFT call for all 15 shells and one MFMA Gram contraction).  value = unique sets / s.

Also measured in the same run: the plain FP64 GPU path on one set (the
per-set restatement of calc_DDD_multipole), the FP32-vs-FP64 relative L2 error
on that set, per-stage HIP-event timings, the roofline of the dominant kernel,
and the CPU oracle on a bounded sample (see cpu_baseline.sample).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_F32_PEAK_TFLOPS = 157.3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def gram_traffic_from_profile(workload):
    """HBM bytes per k_gram_auto launch from the committed rocprofv3 --pmc summary, if it covers this
    workload (FETCH_SIZE and WRITE_SIZE scaled by their reading for a 1 GiB copy); else None."""
    path = os.path.join(ROOT, "profiles", "k_gram_auto_traffic.txt")
    if not os.path.exists(path):
        return None
    fields = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                fields[k.strip()] = v.strip()
    if fields.get("workload") != workload or "traffic_bytes_per_launch" not in fields:
        return None
    return float(fields["traffic_bytes_per_launch"])


def cpu_baseline(J_sample, lm_terms, shells_grid_at_sample, box, seed):
    """Time the CPU oracle (C/OpenMP restatement + scipy.fft) on a bounded sample.

    Sample: the given (l, m) terms of one set at J_sample (mock projected at that J).
    Returns seconds per m-term at J_sample and the thread count used.
    """
    import numpy as np
    from scipy.fft import fftn, ifftn

    from oracle import multipole as om, native, projection as oproj, scaling
    from oracle.windows import radial_multipole_window
    from pyhermes_amd.mocks import lognormal_mock

    threads = native.num_threads()
    L = 1 << J_sample
    phi = scaling.sample_scaling_function()
    pos, _ = lognormal_mock(300000, 64, box, seed)
    w = np.full(pos.shape[0], 1.0 / pos.shape[0])
    eps, _, _ = oproj.project_c(pos, w, phi, J_sample, box)
    delta = eps - 1.0 / L ** 3
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    gamma = scaling.cal_gamma(phi, 3, 1024)
    R1, R2 = shells_grid_at_sample
    t0 = time.perf_counter()
    for (l, m) in lm_terms:
        K1 = radial_multipole_window(L, P, R1, l, m)
        F1 = ifftn(fftn(delta, workers=threads) * K1, workers=threads)
        K2 = radial_multipole_window(L, P, R2, l, -m)
        F2 = ifftn(fftn(delta, workers=threads) * K2, workers=threads)
        om.summand_c(delta, F1, F2, gamma)
    dt = time.perf_counter() - t0
    return dt / len(lm_terms), threads


def mc_cpu_baseline(delta, phi, bins, n, seed):
    """Time the C/OpenMP oracle sampler (plain draws) on n triangles; returns (triangles/s, threads)."""
    from oracle import mc as omc, native
    t0 = time.perf_counter()
    omc.binned_sampler(delta, delta, delta, phi, bins.shape, bins.r1, bins.r2, n, seed)
    return n / (time.perf_counter() - t0), native.num_threads()


def run_mc_workload(args):
    """Monte-Carlo 3PCF with the learned proposal against uniform MC at equal triangle count.

    The J=8 field of the 2,000,000-particle mock (delta = eps / mean - 1, FP64), triangles
    binned in 15 x 15 (r1, r2) shells over 10-120 Mpc/h x 20 cos(theta) bins.  The MLP is
    trained on the records of a plain run; then plain and learned runs of the same size
    are compared bin by bin.  value = median over bins of the per-triangle variance ratio
    (triangles the uniform sampler needs for the learned sampler's variance).
    """
    import numpy as np
    import torch

    from pyhermes_amd import kernels, mcsample
    from pyhermes_amd.mocks import lognormal_mock
    from pyhermes_amd.utils.wavelet_grid import sample_scaling_function

    assert torch.cuda.is_available(), "bench needs a HIP device"
    torch.cuda.set_device(0)
    box, J = 1000.0, args.mc_J
    L = 1 << J
    log(f"mock: {args.particles} particles, ng=256, box={box}; projection J={J}")
    pos, _ = lognormal_mock(args.particles, 256, box, 3000)
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
    n_train, n_run = int(args.mc_train), int(args.mc_triangles)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    mcsample.run_mc(delta, delta, delta, phi_d, bins, 1 << 20, 1)  # warm
    prop, t_buckets = timed(lambda: mcsample.LearnedProposal(bins, delta, phi_d, nbucket=args.mc_buckets,
                                                             alpha=args.mc_alpha, sub=args.mc_sub,
                                                             anchors=args.mc_anchors, seed=1))
    log(f"proposal buckets on the {L * args.mc_sub}^3 sub-cell grid: {t_buckets:.2f} s")
    rec = torch.empty((n_train, 8), dtype=torch.float32, device=delta.device)
    pilot, t_pilot = timed(lambda: mcsample.run_mc(delta, delta, delta, phi_d, bins, n_train, 11,
                                                   proposal=prop.recording, record=rec))
    log(f"plain pilot with records: {n_train:.2e} triangles in {t_pilot:.2f} s")
    losses, t_fit = timed(lambda: prop.fit(rec, pilot, epochs=args.mc_epochs, log=log))
    _, t_build = timed(lambda: prop.build(args.mc_allocation))
    del rec
    log(f"MLP trained in {t_fit:.2f} s (losses {[round(v, 4) for v in losses]}), tables in {t_build:.2f} s")
    plain, t_plain = timed(lambda: mcsample.run_mc(delta, delta, delta, phi_d, bins, n_run, 12))
    log(f"uniform MC: {n_run:.2e} triangles in {t_plain:.2f} s")
    imp, t_imp = timed(lambda: mcsample.run_mc(delta, delta, delta, phi_d, bins, n_run, 13, proposal=prop))
    log(f"learned MC: {n_run:.2e} triangles in {t_imp:.2f} s")
    cmp = mcsample.compare(plain, imp)
    ok = (plain.variance > 0) & (imp.variance > 0)
    gain = (plain.variance[ok] / imp.variance[ok]) * (imp.n / plain.n)
    log(f"compare: {cmp}")
    cpu = None
    if not args.skip_cpu:
        n_cpu = int(args.mc_cpu_triangles)
        log(f"CPU oracle sampler: {n_cpu:.1e} plain triangles")
        rate, threads = mc_cpu_baseline(delta.cpu().numpy(), phi, bins, n_cpu, 12)
        cpu = {"value": rate, "unit": "triangles/s", "cores": threads, "kind": "port",
               "sample": f"C/OpenMP oracle binned sampler, {n_cpu:.0e} uniform triangles on the same field"}
    line = {
        "metric": "MC 3PCF triangles saved by the learned proposal at equal variance (median over bins)",
        "value": float(np.median(gain)),
        "unit": "x fewer triangles",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64 sampling, bf16/fp32 MLP",
        "data": "synthetic lognormal mock (seeded)",
        "config": {"workload": f"mc3pcf_J{J}_{n_run:.0e}_triangles_15x15x20", "particles": args.particles,
                   "J": J, "bins": [15, 15, 20], "r_range_mpc": [10.0, 120.0], "train_records": n_train,
                   "triangles_per_run": n_run, "nbucket": args.mc_buckets, "alpha": args.mc_alpha,
                   "sub": args.mc_sub, "anchors": args.mc_anchors, "allocation": args.mc_allocation,
                   "epochs": args.mc_epochs},
        "gain": {"median": float(np.median(gain)), "mean": cmp["mean_gain"], "total": cmp["total_gain"],
                 "p10": float(np.percentile(gain, 10)), "p90": float(np.percentile(gain, 90)),
                 "frac_bins_ge_3": float(np.mean(gain >= 3.0)), "bins": int(ok.sum())},
        "bias": {"z_mean": cmp["z_mean"], "z_rms": cmp["z_rms"]},
        "triangles_per_s": {"uniform": n_run / t_plain, "learned": n_run / t_imp,
                            "pilot_with_records": n_train / t_pilot},
        "seconds": {"buckets": t_buckets, "pilot": t_pilot, "train": t_fit, "build": t_build, "uniform": t_plain,
                    "learned": t_imp},
        "mlp_losses": [float(v) for v in losses],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def _time_events(fn, reps):
    """Mean HIP-event milliseconds of fn() on torch's current stream (the library's stream)."""
    import torch
    e0, e1 = _events()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        out = fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps, out


def run_2pcf_workload(args):
    """Isotropic 2PCF, FP64 throughout: 100,000-particle lognormal mock (P(k) ~ k^-1.5 on 128^3, sigma_LN = 1),
    box 1000 Mpc/h, unit weights, projection at J=7, 20 linear shell bins over 5-150 Mpc/h.

    Times the drop-in API (SFCProjection.run + Corr_2PCF.run, host catalogue in, xi out) and the device-resident
    kernels, and runs the full CPU oracle on the same catalogue (the configuration is CPU-sized) for parity.
    """
    import numpy as np
    import torch

    from oracle import native, projection as oproj, scaling, windows as ow
    from pyhermes_amd import kernels
    from pyhermes_amd.base import SFCProjection
    from pyhermes_amd.mocks import lognormal_mock
    from pyhermes_amd.theory import Corr_2PCF
    from pyhermes_amd.utils.wavelet_grid import sample_scaling_function

    assert torch.cuda.is_available(), "bench needs a HIP device"
    torch.cuda.set_device(0)
    box, J, n = 1000.0, 7, 100_000
    L = 1 << J
    pos, _ = lognormal_mock(n, 128, box, 1000)
    w = np.ones(pos.shape[0])
    s_edges = np.linspace(5.0, 150.0, 20)

    def api():
        f = SFCProjection({"SFCProjection": {"J": J, "box_size": box, "particle_pos": pos, "catalog_weight": w,
                                             "weight_normalization": "catalog", "fout_path": ""}}).run(
            save_result=False)
        return Corr_2PCF({"Corr_2PCF": {"sfc_field": f, "random": "uniform",
                                        "sampling": {"s": {"min": 5.0, "max": 150.0, "n": 20}},
                                        "products": ["xi"], "fout_path": ""}}).run(save_result=False)

    api()
    torch.cuda.synchronize()
    reps = max(args.steps, 3)
    t0 = time.perf_counter()
    for _ in range(reps):
        res = api()
    torch.cuda.synchronize()
    t_api = (time.perf_counter() - t0) / reps
    log(f"API projection + 2PCF: {t_api * 1e3:.2f} ms")
    # device-resident kernels: projection, r2c FFT, fused 20-bin pair reduction
    phi = sample_scaling_function()
    pos_d = kernels.as_device(pos, torch.float64)
    pw = kernels.as_device(w / w.sum(), torch.float64)
    ms_proj, proj = _time_events(lambda: kernels.project(pos_d, pw, J, box, phi, want32=False, want64=True), reps)
    delta = (proj["eps64"] - 1.0 / L ** 3).contiguous()
    ms_fft, F = _time_events(lambda: kernels.rfftn(delta), reps)
    Pd = kernels.as_device(scaling.fourier_power_spectrum(phi, 0, 1, L, 1024), torch.float64)
    descs = [kernels.window_desc("shell", sv * L / box) for sv in s_edges]
    ms_pair, _ = _time_events(lambda: kernels.pair_reduce(F, F, Pd, descs), reps)
    log(f"kernels: projection {ms_proj:.3f} ms, rfft {ms_fft:.3f} ms, pair reduce {ms_pair:.3f} ms")
    # full CPU oracle
    threads = native.num_threads()
    t0 = time.perf_counter()
    pwn, _ = oproj.projection_weights(w, np.ones_like(w), "catalog")
    eps, _, _ = oproj.project_c(pos, pwn, phi, J, box)
    rho = eps.sum() / L ** 3
    d = eps - rho
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    xi_cpu = np.array([ow.pair_product(d, d, ow.octant_rfft_kernel(L, P, "shell", sv * L / box), threads)
                       for sv in s_edges])
    xi_cpu = xi_cpu / rho ** 2
    t_cpu = time.perf_counter() - t0
    err = float(np.max(np.abs(res.xi - xi_cpu)) / np.max(np.abs(xi_cpu)))
    log(f"CPU oracle: {t_cpu:.2f} s ({threads} threads); max |xi_gpu - xi_cpu| / max |xi| = {err:.2e}")
    line = {
        "metric": "isotropic 2PCF (20 shell bins) per catalogue, FP64, projection + correlation",
        "value": 1.0 / t_api, "unit": "catalogues/s", "n_gpus": 1, "higher_is_better": True,
        "ms_per_catalogue": t_api * 1e3, "dtype": "f64", "data": "synthetic lognormal mock (seeded)",
        "config": {"workload": "2pcf_J7_100k_20bins", "particles": n, "J": J, "bins": 20, "s_range_mpc": [5, 150]},
        "kernels_ms": {"projection": ms_proj, "rfft": ms_fft, "pair_reduce_20_bins": ms_pair},
        "parity": {"xi_max_abs_err_over_max_xi": err, "bar": 1e-5},
        "cpu_baseline": {"value": 1.0 / t_cpu, "unit": "catalogues/s", "cores": threads, "kind": "port",
                         "sample": "the whole configuration: C/OpenMP projection, octant shell kernels and "
                                   "scipy.fft pair products for all 20 bins"},
    }
    print(json.dumps(line))


def run_smu_workload(args):
    """Anisotropic 2PCF xi(s, mu): the 100,000-particle lognormal mock of the 2pcf workload (box 1000 Mpc/h,
    unit weights, J=7), 20 s bins over 5-150 Mpc/h x 10 mu bins over [0.05, 0.95] with ring binning windows
    (mapping smu_to_RH, line of sight z), FP64 throughout, 200 windows in one fused pass.

    Times the drop-in API (SFCProjection.run + Corr_2PCF.run) and the fused pass on the device, checks sampled
    entries against the oracle's convolution pair products and times the oracle on a bounded sample.
    """
    import numpy as np
    import torch

    from oracle import native, projection as oproj, scaling, windows as ow
    from pyhermes_amd.base import SFCProjection
    from pyhermes_amd.mocks import lognormal_mock
    from pyhermes_amd.theory import Corr_2PCF
    from pyhermes_amd.utils.wavelet_grid import sample_scaling_function

    assert torch.cuda.is_available(), "bench needs a HIP device"
    torch.cuda.set_device(0)
    box, J, n = 1000.0, 7, 100_000
    L = 1 << J
    pos, _ = lognormal_mock(n, 128, box, 1000)
    w = np.ones(pos.shape[0])
    s = np.linspace(5.0, 150.0, 20)
    mu = np.linspace(0.05, 0.95, 10)

    def api():
        f = SFCProjection({"SFCProjection": {"J": J, "box_size": box, "particle_pos": pos, "catalog_weight": w,
                                             "weight_normalization": "catalog", "fout_path": ""}}).run(
            save_result=False)
        return Corr_2PCF({"Corr_2PCF": {"sfc_field": f, "random": "uniform", "binning_window": "ring",
                                        "sampling": {"s": s, "mu": mu}, "products": ["xi"],
                                        "fout_path": ""}}).run(save_result=False), f

    api()
    torch.cuda.synchronize()
    reps = max(args.steps, 3)
    t0 = time.perf_counter()
    for _ in range(reps):
        res, field = api()
    torch.cuda.synchronize()
    t_api = (time.perf_counter() - t0) / reps
    log(f"API projection + xi(s, mu) (200 ring windows): {t_api * 1e3:.2f} ms")
    from pyhermes_amd import kernels
    from pyhermes_amd.theory.corr2pcf import binning_desc, map_smu_to_RH
    eps = field.epsilon
    rho = field.field_integral / L ** 3
    delta = eps - rho
    F = kernels.rfftn(kernels.as_device(delta, torch.float64))
    phi = sample_scaling_function()
    P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
    Pd = kernels.as_device(P, torch.float64)
    bw = {"type": "ring", "len_args": {"R": None, "H": None}, "los_args": {"nx": 0.0, "ny": 0.0, "nz": 1.0}}
    samples = [(sv, m) for sv in s for m in mu]
    descs = [binning_desc(map_smu_to_RH({"s": sv, "mu": m}, bw), L, box) for sv, m in samples]
    ms_pair, _ = _time_events(lambda: kernels.pair_reduce(F, F, Pd, descs), reps)
    log(f"fused pass: {ms_pair:.3f} ms for {len(descs)} windows")
    # parity on sampled entries, and the bounded CPU sample: oracle kernels + pair products for 4 windows
    threads = native.num_threads()
    pick = [(0, 0), (7, 3), (13, 9), (19, 5)]
    errs = []
    t0 = time.perf_counter()
    for i, j in pick:
        R, H = s[i] * np.sqrt(1 - mu[j] ** 2) * L / box, s[i] * mu[j] * L / box
        ref = ow.pair_product(delta, delta, ow.los_rfft_kernel(L, P, "ring", R, H), threads) / rho ** 2
        errs.append((abs(res.xi[i, j] - ref) / abs(ref), abs(res.xi[i, j] - ref) / np.max(np.abs(res.xi))))
    t_sample = time.perf_counter() - t0
    t_cpu = t_sample / len(pick) * len(descs)
    err_rel = float(max(e[0] for e in errs))
    err = float(max(e[1] for e in errs))
    log(f"oracle: {t_sample:.2f} s for {len(pick)} windows; max |xi_gpu - xi_cpu| / |xi_cpu| = {err_rel:.2e}, "
        f"/ max |xi| = {err:.2e}")
    line = {
        "metric": "anisotropic 2PCF xi(s, mu) (20 x 10 ring bins) per catalogue, FP64, projection + correlation",
        "value": 1.0 / t_api, "unit": "catalogues/s", "n_gpus": 1, "higher_is_better": True,
        "ms_per_catalogue": t_api * 1e3, "dtype": "f64", "data": "synthetic lognormal mock (seeded)",
        "config": {"workload": "smu_J7_100k_20x10_ring", "particles": n, "J": J, "s_bins": 20, "mu_bins": 10,
                   "s_range_mpc": [5, 150], "mu_range": [0.05, 0.95], "windows_per_pass": len(descs)},
        "kernels_ms": {"pair_reduce_200_ring_windows": ms_pair},
        "parity": {"xi_max_rel_err_sampled": err_rel, "xi_max_abs_err_over_max_xi_sampled": err,
                   "entries_checked": len(pick), "bar": 1e-5},
        "cpu_baseline": {"value": 1.0 / t_cpu, "unit": "catalogues/s", "cores": threads, "kind": "port",
                         "sample": f"oracle ring kernels (scipy J0) + scipy.fft pair products for {len(pick)} of "
                                   f"the {len(descs)} windows ({t_sample:.2f} s), scaled to all windows; "
                                   f"projection not included"},
    }
    print(json.dumps(line))


def run_xi_ell_workload(args):
    """Anisotropic 2PCF multipoles: 1,000,000-particle lognormal mock (256^3 recipe) with plane-parallel
    redshift-space shifts (sigma_v = 300 km/s along z) and lognormal weights (mean 1, sigma 0.3), box
    1000 Mpc/h, projection at J=9; 30 shell bins over 5-150 Mpc/h x l = 0, 2, 4 in one fused pair reduction.
    FP32 coefficients/spectra/windows with FP64 normalisation and accumulators, against the FP64 path.
    """
    import numpy as np
    import torch

    from oracle import native, projection as oproj, scaling, windows as ow
    from pyhermes_amd import kernels
    from pyhermes_amd.base import SFCProjection
    from pyhermes_amd.mocks import lognormal_mock
    from pyhermes_amd.theory import Corr_2PCF
    from pyhermes_amd.utils.wavelet_grid import sample_scaling_function

    assert torch.cuda.is_available(), "bench needs a HIP device"
    torch.cuda.set_device(0)
    box, J, n = 1000.0, 9, 1_000_000
    L = 1 << J
    pos, w = lognormal_mock(n, 256, box, 2000, rsd_sigma_kms=300.0, weight_sigma=0.3)
    ells = [0, 2, 4]
    nbin = 30

    def api(precision):
        f = SFCProjection({"SFCProjection": {"J": J, "box_size": box, "particle_pos": pos, "catalog_weight": w,
                                             "weight_normalization": "catalog", "precision": precision,
                                             "fout_path": ""}}).run(save_result=False)
        return Corr_2PCF({"Corr_2PCF": {"sfc_field": f, "random": "uniform",
                                        "sampling": {"s": {"min": 5.0, "max": 150.0, "n": nbin}},
                                        "products": ["xi"], "multipoles": ells, "fout_path": ""}}).run(
            save_result=False)

    api("fp32")
    torch.cuda.synchronize()
    reps = max(args.steps, 3)
    t0 = time.perf_counter()
    for _ in range(reps):
        r32 = api("fp32")
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    r64 = api("fp64")
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    errs = {}
    for j, l in enumerate(ells):
        a, b = r32.xi_l[j], r64.xi_l[j]
        errs[f"xi_{l}"] = {"max_abs_over_max": float(np.max(np.abs(a - b)) / np.max(np.abs(b))),
                           "rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b))}
    log(f"API FP32 {t32 * 1e3:.1f} ms, FP64 {t64 * 1e3:.1f} ms; errors {errs}")
    # device-resident kernels of the FP32 path
    phi = sample_scaling_function()
    pos_d = kernels.as_device(pos, torch.float64)
    pw = kernels.as_device(w / w.sum(), torch.float64)
    ms_proj, proj = _time_events(lambda: kernels.project(pos_d, pw, J, box, phi, want32=True, want64=False), reps)
    d32 = (proj["eps32"] - np.float32(1.0 / L ** 3)).contiguous()
    ms_fft, F = _time_events(lambda: kernels.rfftn(d32), reps)
    Pd = kernels.as_device(scaling.fourier_power_spectrum(phi, 0, 1, L, 1024), torch.float64)
    s = np.linspace(5.0, 150.0, nbin)
    descs = [kernels.window_desc("shell", sv * L / box) for sv in s]
    descs += [kernels.window_desc("shell_multipole", sv * L / box, l=l, m=0) for l in ells for sv in s]
    ms_pair, _ = _time_events(lambda: kernels.pair_reduce(F, F, Pd, descs), reps)
    half_bytes = L * L * (L // 2 + 1) * 8
    log(f"kernels: projection {ms_proj:.3f} ms, rfft {ms_fft:.3f} ms, pair reduce ({len(descs)} windows) "
        f"{ms_pair:.3f} ms")
    cpu = None
    if not args.skip_cpu:
        # bounded sample: the oracle's pair products for one radius (shell + l = 0, 2, 4) at J=8, scaled by
        # the grid volume (x8) and the window count (x30 radii)
        Js = 8
        Ls = 1 << Js
        threads = native.num_threads()
        rng = np.random.default_rng(0)
        idx = rng.choice(pos.shape[0], 250_000, replace=False)
        pwn, _ = oproj.projection_weights(w[idx], np.ones(idx.size), "catalog")
        eps, _, _ = oproj.project_c(pos[idx], pwn, phi, Js, box)
        d = eps - eps.sum() / Ls ** 3
        P = scaling.fourier_power_spectrum(phi, 0, 1, Ls, 1024)
        R = 60.0 * Ls / box
        t0 = time.perf_counter()
        ow.pair_product(d, d, ow.octant_rfft_kernel(Ls, P, "shell", R), threads)
        for l in ells:
            K = ow.radial_multipole_window(Ls, P, R, l, 0)
            np.sum(np.real(ow.convolve_complex(d, K, threads)) * d)
        t_sample = time.perf_counter() - t0
        t_full = t_sample * 8.0 * nbin
        cpu = {"value": 1.0 / t_full, "unit": "catalogues/s", "cores": threads, "kind": "port",
               "sample": f"C/OpenMP + scipy.fft oracle pair products for one radius (shell + l=0,2,4) at J={Js} "
                         f"({t_sample:.2f} s), scaled x8 (J=9 volume) and x{nbin} radii; projection not included"}
    line = {
        "metric": "2PCF multipoles xi_0,2,4 (30 bins) per catalogue, FP32 path vs FP64",
        "value": 1.0 / t32, "unit": "catalogues/s", "n_gpus": 1, "higher_is_better": True,
        "ms_per_catalogue": t32 * 1e3, "fp64_ms_per_catalogue": t64 * 1e3,
        "dtype": "f32 (f64 normalisation and accumulators)", "data": "synthetic lognormal mock (seeded), RSD + weights",
        "config": {"workload": "xi_ell_J9_1M_rsd_30bins_l024", "particles": n, "J": J, "bins": nbin,
                   "multipoles": ells, "windows_per_launch": len(descs)},
        "kernels_ms": {"projection_fp32": ms_proj, "rfft_fp32": ms_fft, "pair_reduce": ms_pair},
        "pair_reduce_hbm_frac": half_bytes / (ms_pair * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "fp32_vs_fp64": errs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def run_batch_workload(args):
    """Uneven batch: 64 seeded lognormal sub-box catalogues, sizes log-uniform in 1e4..2e6 particles, box
    sides 250-1000 Mpc/h, each with its own J in {6, 7, 8}; every catalogue runs the isotropic 2PCF (20 shell
    bins, 10-100 Mpc/h) and the auto SH-3PCF multipoles (8 shells, 10-80 Mpc/h, l <= 6).  The whole batch is
    captured once into a hipGraph and replayed; value = catalogues/s of the replay.  Per-catalogue latency is
    the replay time of a graph of that catalogue alone (p50/p99 over the 64).  Also reported: the same batch
    enqueued eagerly, the warm-up (FFT tables, tuned tiles, workspaces) and the capture time.
    """
    import numpy as np
    import torch

    from oracle import native, projection as oproj, scaling, windows as ow
    from pyhermes_amd import kernels
    from pyhermes_amd.batch import Catalogue, CatalogueBatch, timed_replays, wall
    from pyhermes_amd.mocks import lognormal_mock
    from pyhermes_amd.utils.wavelet_grid import sample_scaling_function

    assert torch.cuda.is_available(), "bench needs a HIP device"
    torch.cuda.set_device(0)
    rng = np.random.default_rng(5000)
    ncat = args.batch_catalogues
    sizes = np.round(10.0 ** rng.uniform(4.0, np.log10(2e6), ncat)).astype(np.int64)
    boxes = rng.uniform(250.0, 1000.0, ncat)
    Js = rng.choice([6, 7, 8], ncat)
    log(f"batch: {ncat} catalogues, {int(sizes.sum())} particles, J counts "
        f"{ {int(j): int((Js == j).sum()) for j in (6, 7, 8)} }")
    cats = []
    for i in range(ncat):
        pos, w = lognormal_mock(int(sizes[i]), 64, float(boxes[i]), 5000 + 10 * i)
        cats.append(Catalogue(pos, w, float(boxes[i]), int(Js[i])))
    phi = sample_scaling_function()
    s_bins = np.linspace(10.0, 100.0, 20)
    shells = np.linspace(10.0, 80.0, 8)
    lmax = 6
    b = CatalogueBatch(cats, s_bins, shells, lmax, phi)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        _, t_warm = wall(b.warmup)
        res_eager, t_eager = wall(b.run_eager)
        log(f"warm-up {t_warm:.2f} s; eager batch {t_eager * 1e3:.1f} ms")
        t0 = time.perf_counter()
        g_all = b.capture()
        t_capture = time.perf_counter() - t0
        b.replay(g_all)  # first replay (uploads the graph)
        torch.cuda.synchronize()
        reps = max(args.steps, 3)
        ms_all = timed_replays(b, [g_all] * reps, stream)
        b.check_errors()
        res_graph = b.results()
        same = all(np.array_equal(r["xi"], q["xi"]) and np.array_equal(r["zeta_l"], q["zeta_l"])
                   for r, q in zip(res_eager, res_graph))
        log(f"capture {t_capture * 1e3:.1f} ms; graph replay {np.mean(ms_all):.1f} ms; identical to eager: {same}")
        graphs = [b.capture([i]) for i in range(ncat)]
        timed_replays(b, graphs, stream)  # first replay of each
        lat = np.array(timed_replays(b, graphs, stream))
        eager_lat = []
        for i in range(ncat):
            _, t = wall(lambda: b.enqueue(i))
            eager_lat.append(t * 1e3)
        for g in graphs + [g_all]:
            kernels.graph_destroy(g)
    t_graph = float(np.mean(ms_all)) * 1e-3
    by_J = {int(j): {"catalogues": int((Js == j).sum()),
                     "p50_ms": float(np.percentile(lat[Js == j], 50)) if (Js == j).any() else None}
            for j in (6, 7, 8)}
    cpu = None
    if not args.skip_cpu:
        # bounded sample: the oracle's 20 shell-bin pair products and 3 (l, m) terms of one shell pair for one
        # J=6 catalogue, scaled to the batch: x 8^(J-6) volume per catalogue, 20 bins and 28 m-terms x 36 pairs
        i6 = int(np.nonzero(Js == 6)[0][0]) if (Js == 6).any() else 0
        c = cats[i6]
        L = 64
        threads = native.num_threads()
        pwn, _ = oproj.projection_weights(c.weight, np.ones(c.n), "catalog")
        eps, _, _ = oproj.project_c(c.pos, pwn, phi, 6, c.box)
        d = eps - eps.sum() / L ** 3
        P = scaling.fourier_power_spectrum(phi, 0, 1, L, 1024)
        gamma = scaling.cal_gamma(phi, 3, 1024)
        from oracle import multipole as om
        t0 = time.perf_counter()
        for sv in s_bins:
            ow.pair_product(d, d, ow.octant_rfft_kernel(L, P, "shell", sv * L / c.box), threads)
        t_2pcf = time.perf_counter() - t0
        t0 = time.perf_counter()
        terms = [(0, 0), (2, 1), (4, 3)]
        for (l, m) in terms:
            F1 = ow.convolve_complex(d, ow.radial_multipole_window(L, P, shells[2] * L / c.box, l, m), threads)
            F2 = ow.convolve_complex(d, ow.radial_multipole_window(L, P, shells[5] * L / c.box, l, -m), threads)
            om.summand_c(d, F1, F2, gamma)
        t_term = (time.perf_counter() - t0) / len(terms)
        n_pairs = len(shells) * (len(shells) + 1) // 2
        n_m = (lmax + 1) * (lmax + 2) // 2
        t_cpu = sum(8.0 ** (int(j) - 6) * (t_2pcf + t_term * n_m * n_pairs) for j in Js)
        cpu = {"value": ncat / t_cpu, "unit": "catalogues/s", "cores": threads, "kind": "port",
               "sample": f"C/OpenMP + scipy.fft oracle on one J=6 catalogue: 20 shell-bin pair products "
                         f"({t_2pcf:.2f} s) and 3 (l,m) terms of one shell pair ({t_term:.2f} s each), scaled by "
                         f"8^(J-6) per catalogue, {n_m} m-terms x {n_pairs} pairs"}
    line = {
        "metric": "uneven catalogue batch (2PCF 20 bins + SH-3PCF l<=6, 8 shells) catalogues/s from one hipGraph",
        "value": ncat / t_graph, "unit": "catalogues/s", "n_gpus": 1, "higher_is_better": True,
        "ms_per_batch": t_graph * 1e3, "dtype": "f32 (f64 normalisation and accumulators)",
        "data": "synthetic lognormal mocks (seeded)",
        "config": {"workload": f"batch_{ncat}_catalogues_J678_2pcf20_sh3pcf_l6_8shells",
                   "particles_total": int(sizes.sum()), "sizes_range": [int(sizes.min()), int(sizes.max())],
                   "J_counts": {str(j): int((Js == j).sum()) for j in (6, 7, 8)}},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max()), "by_J": by_J},
        "eager": {"catalogues_per_s": ncat / t_eager, "p50_ms": float(np.percentile(eager_lat, 50)),
                  "p99_ms": float(np.percentile(eager_lat, 99))},
        "graph_identical_to_eager": bool(same),
        "setup_s": {"warmup": t_warm, "capture_batch": t_capture},
        "cpu_baseline": cpu,
    }
    b.close()
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["sh3pcf", "mc", "2pcf", "smu", "xi_ell", "batch"], default="sh3pcf",
                    help="sh3pcf: the multipole sets/s headline; mc: learned vs uniform Monte-Carlo 3PCF; "
                         "2pcf: FP64 isotropic 2PCF at J=7; smu: FP64 xi(s, mu) with ring windows at J=7; "
                         "xi_ell: FP32 2PCF multipoles with RSD at J=9; "
                         "batch: 64 uneven catalogues from one hipGraph")
    ap.add_argument("--batch-catalogues", type=int, default=64)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--J", type=int, default=9)
    ap.add_argument("--lmax", type=int, default=10)
    ap.add_argument("--nshell", type=int, default=15)
    ap.add_argument("--particles", type=int, default=2_000_000)
    ap.add_argument("--schedule", choices=["serial", "split", "overlap"], default="serial",
                    help="stream schedule of the multipole pipeline")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-pass-probe", action="store_true",
                    help="skip timing the strided inverse passes against a device copy of the same bytes")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB once (HBM counter calibration)")
    ap.add_argument("--mc-J", type=int, default=8)
    ap.add_argument("--mc-triangles", type=float, default=1e8)
    ap.add_argument("--mc-train", type=float, default=1e7)
    ap.add_argument("--mc-epochs", type=int, default=3)
    ap.add_argument("--mc-buckets", type=int, default=128)
    ap.add_argument("--mc-alpha", type=float, default=0.02)
    ap.add_argument("--mc-sub", type=int, default=4)
    ap.add_argument("--mc-anchors", type=int, choices=[1, 3], default=3)
    ap.add_argument("--mc-allocation", choices=["uniform", "relative"], default="uniform")
    ap.add_argument("--mc-cpu-triangles", type=float, default=1e8)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload != "sh3pcf":
        {"mc": run_mc_workload, "2pcf": run_2pcf_workload, "smu": run_smu_workload, "xi_ell": run_xi_ell_workload,
         "batch": run_batch_workload}[args.workload](args)
        return

    import numpy as np
    import torch

    from pyhermes_amd import kernels, multipole
    from pyhermes_amd.mocks import lognormal_mock
    from pyhermes_amd.utils.wavelet_grid import sample_scaling_function

    assert torch.cuda.is_available(), "bench needs a HIP device"
    torch.cuda.set_device(0)
    box = 1000.0
    J = args.J
    L = 1 << J
    N = L ** 3
    radii_phys = np.linspace(10.0, 120.0, args.nshell)
    radii = radii_phys * L / box
    lmax = args.lmax
    n_mterms = (lmax + 1) * (lmax + 2) // 2
    n_unique = args.nshell * (args.nshell + 1) // 2
    n_ordered = args.nshell ** 2

    log(f"mock: {args.particles} particles, ng=256, box={box}")
    pos, _ = lognormal_mock(args.particles, 256, box, 3000)
    w = np.full(pos.shape[0], 1.0 / pos.shape[0])
    phi = sample_scaling_function()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    pos_d = kernels.as_device(pos, torch.float64)
    w_d = kernels.as_device(w, torch.float64)
    kernels.project(pos_d, w_d, J, box, phi, want32=True, want64=True)  # warm
    torch.cuda.synchronize()
    ev0.record()
    proj = kernels.project(pos_d, w_d, J, box, phi, want32=True, want64=False)
    ev1.record()
    torch.cuda.synchronize()
    proj_ms = ev0.elapsed_time(ev1)
    proj64 = kernels.project(pos_d, w_d, J, box, phi, want32=False, want64=True)
    rho = 1.0 / N
    d64 = proj64["eps64"] - rho
    d32 = proj["eps32"] - np.float32(rho)
    del proj, proj64
    log(f"projection J={J}: {proj_ms:.3f} ms (eps32 output)")

    geo = multipole.MultipoleGeometry(phi, L)
    log(f"FP32 batched: warmup {args.warmup}, steps {args.steps}")
    for _ in range(args.warmup):
        res32 = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, lmax,
                                                    schedule=args.schedule)
    torch.cuda.synchronize()
    timer = multipole.StageTimer()
    kernels.profile_gram(True)  # HIP events around the contraction kernel, on the library's stream
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res32 = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, lmax, timer=timer,
                                                    schedule=args.schedule)
    torch.cuda.synchronize()
    t_fp32 = (time.perf_counter() - t0) / args.steps
    gram_total_ms, gram_calls = kernels.profile_gram(False)
    stages = {k: round(v[0] / args.steps, 3) for k, v in timer.summary().items()}
    gram_ms = gram_total_ms / gram_calls
    log(f"FP32 step {t_fp32 * 1e3:.1f} ms; stages (ms/step) {stages}")

    out = {}
    i1, i2 = 4, 9
    if not args.skip_fp64:
        log("plain FP64 GPU path: one set")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res64 = multipole.ddd_multipole_plain_f64(geo, d64, d64, d64, radii[i1], radii[i2], 0, lmax,
                                                  progress=lambda l: log(f"  FP64 l={l} done"))
        torch.cuda.synchronize()
        t_fp64 = time.perf_counter() - t0
        rel = float(np.linalg.norm(res32[i1, i2] - res64) / np.linalg.norm(res64))
        out["fp64_plain"] = {"sets_per_s": 1.0 / t_fp64, "s_per_set": t_fp64, "set": [i1, i2]}
        out["fp32_vs_fp64_rel_l2"] = rel
        out["speedup_vs_fp64_plain"] = (n_unique / t_fp32) * t_fp64
        log(f"FP64 set {t_fp64:.2f} s; rel L2 err FP32 vs FP64 = {rel:.3e}")

    cpu = None
    if not args.skip_cpu:
        # terms across l (the j_l recurrence costs more at high l); the sample runs at J - 2 and is scaled by
        # N log N (the FFTs' growth; the window builds and the contraction grow only as N, so this scaling
        # does not flatter the CPU)
        Js = max(J - 2, 4)
        terms = [t for t in [(0, 0), (5, 3), (10, 7)] if t[0] <= lmax]
        scale_R = (1 << Js) / box
        log(f"CPU oracle sample: {terms} of one set at J={Js}")
        per_term, threads = cpu_baseline(Js, terms, (radii_phys[i1] * scale_R, radii_phys[i2] * scale_R), box, 3000)
        growth = (8.0 ** (J - Js)) * (3 * J) / (3 * Js)
        s_per_set = per_term * growth * n_mterms
        cpu = {"value": 1.0 / s_per_set, "unit": "sets/s", "cores": threads, "kind": "port",
               "sample": f"C/OpenMP oracle + scipy.fft, (l,m) in {terms} of one set at J={Js} "
                         f"({per_term:.2f} s per m-term), scaled by N log N (x{growth:.1f}) to J={J} and "
                         f"x{n_mterms} m-terms"}
        log(f"CPU: {per_term:.2f} s per m-term at J={Js} -> {1.0 / s_per_set:.3e} sets/s at J={J}")

    # roofline of the dominant kernel, the auto Gram contraction k_gram_auto.  Flop basis: the contraction
    # with the x stencil on d1 and the y and z stencils on the fields, an S x S Gram of 18 real rank-1
    # products per cell (9 yz eigen slots x re/im against a 3-term right operand).  The kernel reaches the
    # same sums with the z stencil on d1 as well: 6 MFMA products per cell against 13-term right operands
    # formed on the VALU, so the figure is the rate of the contraction, not of the instructions executed.
    S = args.nshell
    gram_flops = 2.0 * S * S * 18 * N
    achieved = gram_flops / (gram_ms * 1e-3) / 1e12
    workload = f"sh3pcf_multipoles_J{J}_lmax{lmax}_{S}shells"
    traffic = gram_traffic_from_profile(workload)
    summ = timer.summary()
    n_modes = L * L * (L // 2 + 1)
    # fused window + inverse FFT: per call (one or two orders m of one l) the half spectrum (complex64) is
    # read once and S complex64 fields per order written by the window/z pass, then read and written by
    # each of the y and x passes.  With the split schedule the window/z launches and the y/x passes are
    # timed apart (they run beside each other on two streams); the stage time is their sum.
    win_key = "window_z" if "window_z" in summ else "window_ifft"
    win_calls = summ[win_key][1] / args.steps
    win_ms = (summ[win_key][0] + (summ["yx_passes"][0] if "yx_passes" in summ else 0.0)) / args.steps
    n_fields = multipole.fields_per_set(0, lmax, L, S, args.schedule)  # m = 0 terms packed in pairs
    win_bytes = win_calls * n_modes * 8 + n_fields * (S * N * 8 + 2 * (2 * S * N * 8))
    # algorithmic bytes of the stage: the half spectrum read once per call and every output field written
    # once (the intermediate passes' traffic is the FFT's, not the window's)
    win_alg_bytes = win_calls * n_modes * 8 + n_fields * S * N * 8
    # the two strided inverse passes alone on one call's worth of fields (S of them, x-major in, y-major
    # out), against a device copy of the same bytes: both move every byte once per read and once per write
    passes = None
    if not args.skip_pass_probe:
        n_probe = max(1, min(S, (16 << 30) // (N * 8)))  # at most 16 GiB per buffer
        src = torch.zeros((n_probe, L, L, L), dtype=torch.complex64, device="cuda")
        dst = torch.empty_like(src)

        def med_ms(fn):
            ts = []
            for _ in range(4):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                fn()
                e1.record()
                torch.cuda.synchronize()
                ts.append(e0.elapsed_time(e1))
            return float(np.median(ts[1:]))
        t_pass = med_ms(lambda: kernels.fft3_inverse_yx_ymajor_f32(src, dst))
        t_copy = med_ms(lambda: dst.copy_(src))
        moved = 2 * n_probe * N * 8  # one read and one write of the fields
        passes = {"fields": n_probe, "ms_both_passes": t_pass, "tbs_per_pass": 2 * moved / (t_pass * 1e-3) / 1e12,
                  "device_copy_ms": t_copy, "device_copy_tbs": moved / (t_copy * 1e-3) / 1e12}
        del src, dst
        log(f"strided passes: {t_pass:.2f} ms for {n_probe} fields; device copy of one pass's bytes {t_copy:.2f} ms")
    if args.copy_probe_mib > 0:
        n_el = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n_el, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written")
        del src, dst
    line = {
        "metric": "J=9 3PCF(l<=10) sets/s, FP32+batched vs FP64 GPU at <=1e-4 rel err; %HBM peak",
        "value": n_unique / t_fp32,
        "unit": "sets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_fp32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 accumulation)",
        "data": "synthetic lognormal mock (seeded)",
        "config": {"workload": workload, "particles": args.particles, "schedule": args.schedule,
                   "J": J, "l_max": lmax, "n_shell": S, "unique_sets_per_step": n_unique,
                   "ordered_sets_per_step": n_ordered, "m_terms_per_set": n_mterms},
        "roofline": {"bound": "mfma", "kernel": "k_gram_auto", "achieved": achieved,
                     "peak": MFMA_F32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / MFMA_F32_PEAK_TFLOPS,
                     "traffic": traffic, "avg_launch_ms": gram_ms,
                     "flops_basis": "2 S^2 18 L^3 per launch (x-delta-form contraction)",
                     "algorithmic_bytes": S * N * 8 + N * 4},
        "cpu_baseline": cpu,
        "stages_ms_per_step": stages,
        "projection_ms": proj_ms,
        "projection_hbm_frac": ((args.particles * 32 + N * 4) / (proj_ms * 1e-3) / 1e9) / HBM_PEAK_GBS,
        "window_ifft": {"ms_per_step": win_ms, "calls_per_step": win_calls, "field_groups_per_step": n_fields,
                        "bytes_per_step": win_bytes,
                        "hbm_frac": win_bytes / (win_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "algorithmic_bytes_per_step": win_alg_bytes,
                        "hbm_frac_algorithmic": win_alg_bytes / (win_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "strided_passes": passes},
    }
    line.update(out)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
