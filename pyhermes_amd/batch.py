"""Uneven batches of catalogues through one captured hipGraph.

Each catalogue has its own particle count, box and level J.  Per catalogue the device work is
the drop-in path of the tasks: projection of the weighted particles (FP32 coefficients, FP64
weights and normalisation; sfc_projection.py:221-324), delta = eps - rho with the uniform
random density rho (_prepare_product_fields, corr3pcf_multipole.py:1363-1384), the FP32 3-D
transform and half spectrum, the fused pair reduction of every isotropic 2PCF bin
(pair_product_with_window, corr2pcf.py:366-372), and per (l, m) the fused multipole window +
inverse transform of every shell and the MFMA Gram of the auto SH-3PCF contraction
(compute_multipole_m_summand, corr3pcf_multipoles.py:393-457).  Host post-processing turns
the device sums into xi(s) and zeta_l(r1, r2) exactly as Corr_2PCF / Corr_3PCF_Multipole do.

Every call is stream-ordered with no host synchronisation, allocation or memset, so the whole
batch can be captured once (hermes_capture_begin/end) and replayed as one graph; buffers are
sized for the largest level and reused catalogue after catalogue.  One eager pass over one
catalogue of each level before capturing builds the FFT tables, tunes the FFT tiles and grows
the library's workspaces (``warmup``).  No reference counterpart as a batch: the reference runs
one catalogue per task.
"""

import time

import numpy as np
import torch

from . import _lib, kernels
from .multipole import _window_runs, combine_m_terms
from .utils.wavelet_grid import cal_gamma, fourier_power_spectrum, sample_scaling_function


class Catalogue:
    """Positions (n, 3) in box units, catalogue weights (n,), box side and level J."""

    def __init__(self, pos, weight, box, J):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.weight = (np.ones(self.pos.shape[0]) if weight is None
                       else np.ascontiguousarray(weight, dtype=np.float64))
        self.box = float(box)
        self.J = int(J)
        if self.pos.ndim != 2 or self.pos.shape[1] != 3 or self.weight.shape != (self.pos.shape[0],):
            raise ValueError("Catalogue: pos must be (n, 3) and weight (n,)")
        if not 4 <= self.J <= 10:
            raise ValueError(f"Catalogue: J={self.J} outside 4..10 (the auto Gram needs L >= 16)")

    @property
    def n(self):
        return self.pos.shape[0]


class CatalogueBatch:
    """Isotropic 2PCF (thin shells at ``s_bins``) and auto SH-3PCF multipoles (thin shells at
    ``shells``, l <= ``lmax``) for every catalogue; lengths in the catalogue's box units."""

    def __init__(self, catalogues, s_bins, shells, lmax, phi=None):
        if not catalogues:
            raise ValueError("CatalogueBatch needs at least one catalogue")
        self.cats = list(catalogues)
        self.s_bins = np.asarray(s_bins, dtype=np.float64)
        self.shells = np.asarray(shells, dtype=np.float64)
        self.lmax = int(lmax)
        if not 1 <= self.shells.size <= 16:
            raise ValueError("CatalogueBatch: 1..16 SH-3PCF shells")
        self.lm = [(l, m) for l in range(self.lmax + 1) for m in range(l + 1)]
        # the Gram's MFMA tile is 16 shells wide: pack 16 // S fields into one contraction and keep the
        # diagonal blocks (the cross-term blocks are not used).  The m = 0 terms of consecutive degrees
        # share one field ((l, 0) real part, (l + 1, 0) imaginary part; a Gram with the real and imaginary
        # products apart separates them); the other terms have a field each.
        self.pack = max(1, 16 // self.shells.size)
        pack_m0 = max(c.J for c in self.cats) <= 9  # the packed window kernel covers L <= 512
        pairs = [l for l in range(0, self.lmax, 2)] if pack_m0 else []  # (l, 0) with (l + 1, 0)
        single = [(l, m) for l, m in self.lm if m > 0 or not pack_m0 or (l % 2 == 0 and l == self.lmax)]
        self.groups = [("plain", single[i:i + self.pack]) for i in range(0, len(single), self.pack)]
        self.groups += [("pack0", pairs[i:i + self.pack]) for i in range(0, len(pairs), self.pack)]
        phi = sample_scaling_function() if phi is None else np.asarray(phi, dtype=np.float64)
        self.phi = phi
        self.phi_d = kernels.as_device(phi, torch.float64)
        self.phi_abs_max = float(np.abs(phi).max())
        gamma = cal_gamma(phi, phi.size // 1024, 1024)
        lam, Q = np.linalg.eigh(gamma)
        self.Q, self.lam = Q, lam
        self.P = {}
        for J in sorted({c.J for c in self.cats}):
            L = 1 << J
            self.P[J] = kernels.as_device(fourier_power_spectrum(phi, 0, 1, L, 1024), torch.float64)
        # per catalogue: device inputs, projection weights, rho, window tables
        self.dev = []
        for c in self.cats:
            L = 1 << c.J
            wsum = float(np.sum(c.weight, dtype=np.float64))
            if not np.isfinite(wsum) or wsum == 0.0:
                raise ValueError("CatalogueBatch: catalogue weights must have a finite, non-zero sum")
            pw = c.weight / wsum  # weight_normalization "catalog": field_integral = 1
            rho = float(np.sum(pw, dtype=np.float64)) / L ** 3
            descs = [kernels.window_desc("shell", float(s) * L / c.box) for s in self.s_bins]
            self.dev.append({
                "pos": kernels.as_device(c.pos, torch.float64),
                "w": kernels.as_device(pw, torch.float64),
                "rho": rho,
                "plan": kernels.PairPlan(c.J, descs),
                "radii": self.shells * L / c.box,
            })
        Jmax = max(c.J for c in self.cats)
        Lm = 1 << Jmax
        dev = self.phi_d.device
        S = self.shells.size
        self.eps = torch.empty(Lm ** 3, dtype=torch.float32, device=dev)
        self.d = torch.empty(Lm ** 3, dtype=torch.float32, device=dev)
        self.dc = torch.empty(Lm ** 3, dtype=torch.complex64, device=dev)
        self.H = torch.empty(Lm * Lm * (Lm // 2 + 1), dtype=torch.complex64, device=dev)
        self.U = torch.empty(self.pack * S * Lm ** 3, dtype=torch.complex64, device=dev)
        # the window z passes write x-major fields here; the strided passes move them into U in y-major order
        # (both passes then read lines L elements apart), which the Grams read
        self.stage = torch.empty_like(self.U)
        self.pp_out = torch.zeros((len(self.cats), self.s_bins.size), dtype=torch.float64, device=dev)
        self.gram_out = torch.zeros((len(self.cats), len(self.groups), 512), dtype=torch.float64, device=dev)
        self._warm = False

    # ------------------------------------------------------------ device work
    def enqueue(self, i):
        """All device work of catalogue i on the library stream (capturable)."""
        c, dv = self.cats[i], self.dev[i]
        J = c.J
        L = 1 << J
        N = L ** 3
        eps = self.eps[:N]
        _lib.call("hermes_project", _lib.context(), _lib.ptr(dv["pos"]), _lib.ptr(dv["w"]), c.n, J, c.box,
                  _lib.ptr(self.phi_d), self.phi.size, 1024, self.phi_abs_max, _lib.ptr(eps), None, None, None,
                  None)
        d = self.d[:N].view(L, L, L)
        dc = self.dc[:N].view(L, L, L)
        kernels.delta_f32(eps, 1.0, dv["rho"], d_out=d, dc_out=dc)
        kernels.fft3_c2c_f32(dc, inverse=False)
        H = self.H[:L * L * (L // 2 + 1)].view(L, L, L // 2 + 1)
        kernels.half_spectrum_f32(dc, H)
        dv["plan"].reduce(H, H, self.P[J], self.pp_out[i])
        S = self.shells.size
        for k, (kind, group) in enumerate(self.groups):
            n = len(group) * S
            if kind == "pack0":
                for pos, l in enumerate(group):
                    Sj = self.stage[pos * S * N:(pos + 1) * S * N].view(S, L, L, L)
                    kernels.multipole_window_zpass_pack0_f32(H, self.P[J], dv["radii"], l, 1.0 / N, Sj)
            else:
                pos = 0
                for l, ms in _window_runs(group):  # orders of one l share the radial factors
                    Sj = self.stage[pos * S * N:(pos + len(ms)) * S * N].view(len(ms) * S, L, L, L)
                    kernels.multipole_window_zpass_multi_f32(H, self.P[J], dv["radii"], l, ms, 1.0 / N, Sj)
                    pos += len(ms)
            U = self.U[:n * N].view(n, L, L, L)
            kernels.fft3_inverse_yx_ymajor_f32(self.stage[:n * N].view(n, L, L, L), U)
            with kernels.field_layout("ymajor"):
                if kind == "pack0":
                    kernels.multipole_gram_f32_dev_ri(d, U, self.Q, self.lam, self.gram_out[i, k])
                else:
                    kernels.multipole_gram_f32_dev(d, U, self.Q, self.lam, self.gram_out[i, k, :256])

    def warmup(self):
        """One eager catalogue per level (FFT tables, tuned tiles, workspaces) and the projection
        workspaces for the largest catalogue of each level."""
        seen = set()
        for i, c in enumerate(self.cats):
            if c.J not in seen:
                seen.add(c.J)
                self.enqueue(i)
        for J in seen:
            kernels.project_reserve(J, max(c.n for c in self.cats if c.J == J))
        torch.cuda.synchronize()
        self._warm = True

    def run_eager(self):
        """Enqueue every catalogue without a graph; returns the host results."""
        for i in range(len(self.cats)):
            self.enqueue(i)
        torch.cuda.synchronize()
        return self.results()

    def capture(self, indices=None):
        """Capture the work of the given catalogues (default: all) into one graph handle.
        Must run under a non-default torch stream (``torch.cuda.stream(s)``)."""
        if not self._warm:
            raise RuntimeError("CatalogueBatch.capture: call warmup() first")
        idx = range(len(self.cats)) if indices is None else indices
        kernels.capture_begin()
        try:
            kernels.clear_errors()
            for i in idx:
                self.enqueue(i)
        finally:
            graph = kernels.capture_end()
        return graph

    @staticmethod
    def replay(graph):
        kernels.graph_launch(graph)

    def check_errors(self):
        if kernels.read_errors():
            raise ValueError("CatalogueBatch: a catalogue holds non-finite or out-of-range positions")

    # ------------------------------------------------------------ host post-processing
    def results(self):
        """Per catalogue: xi (n_bins,) and zeta_l (n_shell, n_shell, lmax + 1), uniform randoms."""
        pp = self.pp_out.cpu().numpy()
        gr = self.gram_out.cpu().numpy()
        S = self.shells.size
        out = []
        for i, c in enumerate(self.cats):
            L = 1 << c.J
            N = L ** 3
            rho = self.dev[i]["rho"]
            xi = pp[i] / float(L) ** 6 / rho ** 2
            zeta = np.zeros((S, S, self.lmax + 1))
            v_re = {}
            for k, (kind, group) in enumerate(self.groups):
                if kind == "pack0":
                    re = kernels.gram_block_to_matrix(gr[i, k, :256], len(group) * S)
                    im = kernels.gram_block_to_matrix(gr[i, k, 256:], len(group) * S)
                    for j, l in enumerate(group):  # l even: real part; l + 1 odd: imaginary part, sign (-1)^(l+1)
                        v_re[(l, 0)] = 4.0 * np.pi * re[j * S:(j + 1) * S, j * S:(j + 1) * S] / N
                        v_re[(l + 1, 0)] = -4.0 * np.pi * im[j * S:(j + 1) * S, j * S:(j + 1) * S] / N
                    continue
                full = kernels.gram_block_to_matrix(gr[i, k, :256], len(group) * S)
                for j, (l, m) in enumerate(group):
                    G = ((-1.0) ** (l + m)) * full[j * S:(j + 1) * S, j * S:(j + 1) * S]
                    v_re[(l, m)] = 4.0 * np.pi * G / N
            for l in range(self.lmax + 1):
                zeta[:, :, l] = combine_m_terms([v_re[(l, m)] for m in range(l + 1)], l) / rho ** 3
            out.append({"xi": xi, "zeta_l": zeta})
        return out

    def close(self):
        for dv in self.dev:
            dv["plan"].close()


def timed_replays(batch, graphs, stream):
    """Milliseconds of each graph's replay (HIP events on ``stream``, one replay at a time)."""
    ms = []
    for g in graphs:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        batch.replay(g)
        e1.record(stream)
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    return ms


def wall(fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0
