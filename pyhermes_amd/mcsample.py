"""Binned Monte-Carlo 3PCF with a learned importance proposal.

The plain estimator samples triangles uniformly: centre uniform in the box,
(r1, r2, cos theta) uniform over nr1 x nr2 x nmu bins, orientation uniform, and
records f = n1 n2 n3.  Per bin b the estimate is zeta_b = (B / N) sum_{i in b} w_i f_i
with w = 1, its variance (B^2 / N^2) (sum (w f)^2 - (sum w f)^2 / N).

The learned proposal q(centre, bin) reweights the same estimator (w = p / q, so
it stays unbiased): a 3-layer MLP (4 -> 64 -> 64 -> 1, bf16 weights, FP32
accumulation, trained on the GPU by the HIP kernels in csrc/mlp.hip) maps
(r1, r2, cos theta, local density) to log E[f^2 | ...], fitted on the plain
estimator's own triangle weights.  The proposal is
q(c, b) ~ sqrt(E[f^2 | b, rho(c)]) / |zeta_b(pilot)| -- the allocation that
minimises the summed relative variance -- discretised on density buckets of the
centre cell and mixed with the uniform proposal (weight alpha) so that
w <= 1 / alpha.
"""

import ctypes

import numpy as np
import torch

from . import _lib, kernels


class TriangleBins:
    def __init__(self, nr1, nr2, nmu, rmin, rmax, rmin2=None, rmax2=None):
        self.nr1, self.nr2, self.nmu = int(nr1), int(nr2), int(nmu)
        self.r1 = (float(rmin), float(rmax))
        self.r2 = (float(rmin if rmin2 is None else rmin2), float(rmax if rmax2 is None else rmax2))

    @property
    def shape(self):
        return (self.nr1, self.nr2, self.nmu)

    @property
    def B(self):
        return self.nr1 * self.nr2 * self.nmu

    def centres(self):
        r1 = self.r1[0] + (np.arange(self.nr1) + 0.5) * (self.r1[1] - self.r1[0]) / self.nr1
        r2 = self.r2[0] + (np.arange(self.nr2) + 0.5) * (self.r2[1] - self.r2[0]) / self.nr2
        mu = -1.0 + (np.arange(self.nmu) + 0.5) * 2.0 / self.nmu
        g = np.meshgrid(r1, r2, mu, indexing="ij")
        return [a.ravel() for a in g]


class MCResult:
    def __init__(self, sum_wf, sum_wf2, counts, n_samples, B):
        self.sum_wf, self.sum_wf2, self.counts = sum_wf, sum_wf2, counts
        self.n = int(n_samples)
        self.estimate = B / self.n * sum_wf
        self.variance = (B / self.n) ** 2 * np.maximum(sum_wf2 - sum_wf ** 2 / self.n, 0.0)


def _phi_bound(phi, res=1024):
    """max over fractions of sum_s |phi[f + res s]|: |n(x)| <= max|eps| * bound^3."""
    return float(np.abs(np.asarray(phi)).reshape(-1, res).sum(axis=0).max())


def _wf_bound(grids, phi, weight_max=1.0):
    a = _phi_bound(phi) ** 3
    b = 1.0
    for g in grids:
        b *= float(g.abs().max()) * a
    return max(b * weight_max * 1.0000001, 1e-300)


def run_mc(g1, g2, g3, phi, bins, n_samples, seed, proposal=None, record=None):
    wmax = 1.0 / proposal.alpha if proposal is not None else 1.0
    bound = _wf_bound((g1, g2, g3), phi, wmax)
    s1, s2, cnt = kernels.mc_binned(g1, g2, g3, phi, bins.shape, bins.r1, bins.r2, n_samples, seed, bound,
                                    proposal=proposal.struct if proposal is not None else None, record=record)
    return MCResult(s1, s2, cnt, n_samples, bins.B)


class ProposalMLP:
    """FP32 master weights, Adam moments and the bf16 copy used by the HIP kernels."""

    def __init__(self, seed=0, device=None):
        n = kernels.mlp_nparam()
        g = np.random.default_rng(seed)
        w = np.zeros(n, dtype=np.float32)
        # W1 (64x4) | b1 | W2 (64x64) | b2 | w3 | b3
        w[0:256] = g.normal(0, 1.0 / np.sqrt(4), 256)
        w[320:4416] = g.normal(0, 1.0 / np.sqrt(64), 4096)
        w[4480:4544] = g.normal(0, 1.0 / np.sqrt(64), 64) * 0.1
        self.w = kernels.as_device(w, torch.float32)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.wb = torch.empty(n, dtype=torch.bfloat16, device=self.w.device)
        kernels.mlp_to_bf16(self.w, self.wb)
        self.grad = torch.empty(n + 1, dtype=torch.float32, device=self.w.device)
        self.step = 0

    def forward(self, x):
        return kernels.mlp_forward(self.wb, x)

    def train(self, x, y, epochs=3, batch=8192, lr=3e-3, clip=50.0, seed=0, log=None):
        n = int(x.shape[0])
        gen = torch.Generator(device=x.device).manual_seed(seed)
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(n, device=x.device, generator=gen)
            tot = 0.0
            for s0 in range(0, n, batch):
                idx = perm[s0:s0 + batch]
                xb, yb = x[idx].contiguous(), y[idx].contiguous()
                kernels.mlp_grad(self.wb, xb, yb, clip, self.grad)
                self.step += 1
                kernels.mlp_adam(self.w, self.m, self.v, self.grad, self.wb, lr, self.step)
            # epoch loss on a fixed subsample
            sub = min(n, 65536)
            kernels.mlp_grad(self.wb, x[:sub].contiguous(), y[:sub].contiguous(), clip, self.grad)
            tot = float(self.grad[-1])
            losses.append(tot)
            if log:
                log(f"proposal MLP epoch {ep + 1}/{epochs}: loss {tot:.4f}")
        return losses


class LearnedProposal:
    """Fits the proposal MLP on plain-MC records and builds the sampling tables."""

    def __init__(self, bins, grid, nbucket=32, alpha=0.1, seed=0):
        self.bins = bins
        self.grid = grid
        self.L = grid.shape[0]
        self.nbucket = int(nbucket)
        self.alpha = float(alpha)
        self.mlp = ProposalMLP(seed=seed)
        flat = grid.reshape(-1)
        self.rho_scale = float(flat.std()) or 1.0
        t = torch.asinh(flat / self.rho_scale)
        qs = torch.linspace(0, 1, self.nbucket + 1, device=grid.device, dtype=torch.float64)[1:-1]
        sample = t[torch.randint(0, t.numel(), (min(t.numel(), 1 << 22),), device=grid.device,
                                 generator=torch.Generator(device=grid.device).manual_seed(seed))]
        self.edges = torch.quantile(sample.float(), qs.float()).double()
        self.cell_bucket = torch.bucketize(t, self.edges).to(torch.uint8)
        order = torch.sort(self.cell_bucket.to(torch.int32), stable=True)
        self.bucket_cells = order.indices.to(torch.int32)
        counts = torch.bincount(self.cell_bucket.to(torch.int64), minlength=self.nbucket)
        self.n_k = counts.cpu().numpy().astype(np.float64)
        self.bucket_off = torch.cat([torch.zeros(1, dtype=torch.int64, device=grid.device),
                                     torch.cumsum(counts, 0)]).contiguous()
        # representative feature value of each bucket: mean of t over its cells
        tsum = torch.zeros(self.nbucket, dtype=torch.float64, device=grid.device)
        tsum.index_add_(0, self.cell_bucket.to(torch.int64), t)
        self.t_k = (tsum.cpu().numpy() / np.maximum(self.n_k, 1.0))
        self.t_scale = float(max(np.abs(self.t_k).max(), 1e-6))
        self.zeta_pilot = None

    def features(self, r1, r2, mu, t):
        b = self.bins
        x = torch.stack([
            (r1 - b.r1[0]) / (b.r1[1] - b.r1[0]) * 2 - 1,
            (r2 - b.r2[0]) / (b.r2[1] - b.r2[0]) * 2 - 1,
            mu,
            t / self.t_scale,
        ], dim=1)
        return x.to(torch.float32).contiguous()

    def fit(self, records, pilot, epochs=3, log=None):
        """records: (n, 6) float32 device tensor from a plain run; pilot: its MCResult."""
        r1, r2, mu, rho, f = (records[:, i].double() for i in range(5))
        t = torch.asinh(rho / self.rho_scale)
        x = self.features(r1, r2, mu, t)
        y = (f * f)
        y = (y / y.mean()).to(torch.float32).contiguous()
        self.zeta_pilot = np.abs(pilot.estimate).ravel()
        return self.mlp.train(x, y, epochs=epochs, log=log)

    def build(self):
        """Evaluate the MLP on (bucket, bin) centres and upload the proposal tables."""
        b = self.bins
        B = b.B
        r1c, r2c, muc = b.centres()
        dev = self.grid.device
        T = np.repeat(self.t_k, B)
        R1 = np.tile(r1c, self.nbucket)
        R2 = np.tile(r2c, self.nbucket)
        MU = np.tile(muc, self.nbucket)
        x = self.features(*(torch.from_numpy(a).to(dev) for a in (R1, R2, MU, T)))
        s = self.mlp.forward(x).double().cpu().numpy().reshape(self.nbucket, B)
        s = np.clip(s - s.max(), -60.0, 0.0)
        score = np.exp(0.5 * s)
        if self.zeta_pilot is not None:
            z = np.maximum(self.zeta_pilot, 1e-3 * self.zeta_pilot.max() + 1e-300)
            score = score / z[None, :]
        g_k = score.sum(axis=1) / B
        bin_pdf = score / score.sum(axis=1, keepdims=True)
        P = self.n_k * g_k
        P = P / P.sum()
        V = float(self.L) ** 3
        dens = np.where(self.n_k > 0, V * B * P / np.maximum(self.n_k, 1.0), 0.0)
        cdf = np.cumsum(P).astype(np.float32)
        last = int(np.nonzero(P > 0)[0].max())
        cdf[last:] = 2.0
        bcdf = np.cumsum(bin_pdf, axis=1).astype(np.float32)
        bcdf[:, -1] = 2.0
        self._tables = {
            "bucket_cdf": kernels.as_device(cdf, torch.float32),
            "bucket_density": kernels.as_device(dens, torch.float64),
            "bin_cdf": kernels.as_device(bcdf, torch.float32),
            "bin_pdf": kernels.as_device(bin_pdf.astype(np.float32), torch.float32),
        }
        st = _lib.MCProposal()
        st.cell_bucket = self.cell_bucket.data_ptr()
        st.bucket_cells = self.bucket_cells.data_ptr()
        st.bucket_off = self.bucket_off.data_ptr()
        st.bucket_cdf = self._tables["bucket_cdf"].data_ptr()
        st.bucket_density = self._tables["bucket_density"].data_ptr()
        st.bin_cdf = self._tables["bin_cdf"].data_ptr()
        st.bin_pdf = self._tables["bin_pdf"].data_ptr()
        st.nbucket = self.nbucket
        st.alpha = self.alpha
        self.struct = st
        return self


def compare(plain, imp):
    """Variance gain and bias z-scores of an importance run against a plain run."""
    ok = (plain.variance > 0) & (imp.variance > 0)
    gain = plain.variance[ok] / imp.variance[ok] * (imp.n / plain.n)  # per-triangle efficiency ratio
    z = (imp.estimate - plain.estimate)[ok] / np.sqrt(plain.variance[ok] + imp.variance[ok])
    return {"median_gain": float(np.median(gain)), "mean_gain": float(np.mean(gain)),
            "total_gain": float(plain.variance[ok].sum() / imp.variance[ok].sum() * imp.n / plain.n),
            "z_mean": float(z.mean()), "z_rms": float(np.sqrt(np.mean(z ** 2))), "bins": int(ok.sum())}
