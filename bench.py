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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["sh3pcf", "mc"], default="sh3pcf",
                    help="sh3pcf: the multipole sets/s headline; mc: learned vs uniform Monte-Carlo 3PCF")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--J", type=int, default=9)
    ap.add_argument("--lmax", type=int, default=10)
    ap.add_argument("--nshell", type=int, default=15)
    ap.add_argument("--particles", type=int, default=2_000_000)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB once (HBM counter calibration)")
    ap.add_argument("--mc-J", type=int, default=8)
    ap.add_argument("--mc-triangles", type=float, default=1e8)
    ap.add_argument("--mc-train", type=float, default=1e7)
    ap.add_argument("--mc-epochs", type=int, default=3)
    ap.add_argument("--mc-buckets", type=int, default=64)
    ap.add_argument("--mc-alpha", type=float, default=0.05)
    ap.add_argument("--mc-sub", type=int, default=4)
    ap.add_argument("--mc-anchors", type=int, choices=[1, 3], default=3)
    ap.add_argument("--mc-allocation", choices=["uniform", "relative"], default="uniform")
    ap.add_argument("--mc-cpu-triangles", type=float, default=1e8)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "mc":
        run_mc_workload(args)
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
        res32 = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, lmax)
    torch.cuda.synchronize()
    timer = multipole.StageTimer()
    kernels.profile_gram(True)  # HIP events around the contraction kernel, on the library's stream
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res32 = multipole.ddd_multipole_batched_f32(geo, d32, d32, d32, radii, radii, 0, lmax, timer=timer)
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
        Js = J - 1
        terms = [(0, 0), (1, 0), (1, 1)]
        scale_R = (1 << Js) / box
        log(f"CPU oracle sample: {terms} of one set at J={Js}")
        per_term, threads = cpu_baseline(Js, terms, (radii_phys[i1] * scale_R, radii_phys[i2] * scale_R), box, 3000)
        vol = 8.0 ** (J - Js)
        s_per_set = per_term * vol * n_mterms
        cpu = {"value": 1.0 / s_per_set, "unit": "sets/s", "cores": threads, "kind": "port",
               "sample": f"C/OpenMP oracle + scipy.fft, (l,m) in {terms} of one set at J={Js} "
                         f"({per_term:.2f} s per m-term), scaled x{vol:.0f} to J={J} and x{n_mterms} m-terms"}
        log(f"CPU: {per_term:.2f} s per m-term at J={Js} -> {1.0 / s_per_set:.3e} sets/s at J={J}")

    # roofline of the dominant kernel, the auto Gram contraction k_gram_auto: per launch it forms the
    # S x S Gram of 18 real products per cell (9 yz slots x re/im against the combined right operand)
    S = args.nshell
    gram_flops = 2.0 * S * S * 18 * N
    achieved = gram_flops / (gram_ms * 1e-3) / 1e12
    workload = f"sh3pcf_multipoles_J{J}_lmax{lmax}_{S}shells"
    traffic = gram_traffic_from_profile(workload)
    summ = timer.summary()
    n_modes = L * L * (L // 2 + 1)
    # fused window + inverse FFT of S shells: the half spectrum (complex64) read once and S complex64
    # fields written by the window/z pass, then read and written by each of the y and x passes
    win_ms = summ["window_ifft"][0] / summ["window_ifft"][1]
    win_bytes = n_modes * 8 + S * N * 8 + 2 * (2 * S * N * 8)
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
        "config": {"workload": workload, "particles": args.particles,
                   "J": J, "l_max": lmax, "n_shell": S, "unique_sets_per_step": n_unique,
                   "ordered_sets_per_step": n_ordered, "m_terms_per_set": n_mterms},
        "roofline": {"bound": "mfma", "kernel": "k_gram_auto", "achieved": achieved,
                     "peak": MFMA_F32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / MFMA_F32_PEAK_TFLOPS,
                     "traffic": traffic, "avg_launch_ms": gram_ms,
                     "algorithmic_bytes": S * N * 8 + N * 4},
        "cpu_baseline": cpu,
        "stages_ms_per_step": stages,
        "projection_ms": proj_ms,
        "projection_hbm_frac": ((args.particles * 32 + N * 4) / (proj_ms * 1e-3) / 1e9) / HBM_PEAK_GBS,
        "window_ifft": {"avg_call_ms": win_ms, "bytes": win_bytes,
                        "hbm_frac": win_bytes / (win_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    line.update(out)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
