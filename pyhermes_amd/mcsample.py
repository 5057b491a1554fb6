"""Binned Monte-Carlo 3PCF with a learned importance proposal.

The plain estimator samples triangles uniformly: centre uniform in the box,
(r1, r2, cos theta) uniform over nr1 x nr2 x nmu bins, orientation uniform, and
records f = n1 n2 n3.  Per bin b the estimate is zeta_b = (B / N) sum_{i in b} w_i f_i
with w = 1, its variance (B^2 / N^2) (sum (w f)^2 - (sum w f)^2 / N).

The learned proposal q(centre, bin) reweights the same estimator (w = p / q, so
it stays unbiased): a 3-layer MLP (4 -> 64 -> 64 -> 1, bf16 weights, FP32
accumulation, trained on the GPU by the HIP kernels in csrc/mlp.hip) maps
(r1, r2, cos theta, local density) to log E[f^2 | ...], fitted on the plain
estimator's own triangle weights.  Within a bin the proposal places centres
with density ~ sqrt(E[f^2 | b, rho(c)]) -- the variance-minimising choice given
rho(c) -- discretised on density buckets of the centre cell; across bins it
keeps the uniform share (default) or weights by 1 / |zeta_b(pilot)|.  It is
mixed with the uniform proposal (weight alpha) so that w <= 1 / alpha.
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


def run_mc(g1, g2, g3, phi, bins, n_samples, seed, proposal=None, record=None):
    """proposal: a built LearnedProposal, a hermes_mc_proposal struct, or None (plain draws)."""
    st = getattr(proposal, "struct", proposal)
    s1, s2, cnt = kernels.mc_binned(g1, g2, g3, phi, bins.shape, bins.r1, bins.r2, n_samples, seed,
                                    proposal=st, record=record)
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

    def train(self, x, y, epochs=3, batch=8192, lr=3e-3, lr_final=1.5e-4, clip=50.0, seed=0, log=None):
        """Adam on mini-batches; the step size decays from lr to lr_final on a cosine."""
        n = int(x.shape[0])
        gen = torch.Generator(device=x.device).manual_seed(seed)
        losses = []
        total = epochs * ((n + batch - 1) // batch)
        it = 0
        for ep in range(epochs):
            perm = torch.randperm(n, device=x.device, generator=gen)
            xp, yp = x[perm], y[perm]  # the epoch's order once; the batches are then contiguous row ranges
            steps = (n + batch - 1) // batch
            its = np.arange(it, it + steps, dtype=np.float64)
            rates = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * its / max(total - 1, 1)))
            kernels.mlp_train_steps(self.w, self.m, self.v, self.grad, self.wb, xp, yp, batch, clip, rates,
                                    self.step)
            self.step += steps
            it += steps
            # epoch loss on a fixed subsample
            sub = min(n, 65536)
            kernels.mlp_grad(self.wb, x[:sub].contiguous(), y[:sub].contiguous(), clip, self.grad)
            tot = float(self.grad[-1])
            losses.append(tot)
            if log:
                log(f"proposal MLP epoch {ep + 1}/{epochs}: loss {tot:.4f}")
        return losses


class LearnedProposal:
    """Fits the proposal MLP on plain-MC records and builds the sampling tables.

    The proposal lives on the sub-cell grid of side L * sub: the local density of a
    sub-cell is the interpolated field at its centre, and sub-cells are grouped into
    ``nbucket`` quantile buckets of it (hermes_mc_bucketize).  With ``anchors`` = 3
    a proposal draw places any of the three vertices (chosen uniformly) in a
    favoured sub-cell and weights by the balance heuristic over the three.
    """

    def __init__(self, bins, grid, phi, nbucket=64, alpha=0.05, sub=4, anchors=3, seed=0, n_quantile=1 << 22):
        self.bins = bins
        self.grid = grid
        self.L = int(grid.shape[0])
        self.nbucket = int(nbucket)
        self.alpha = float(alpha)
        self.sub = int(sub)
        self.anchors = int(anchors)
        self.phi = phi if isinstance(phi, torch.Tensor) else kernels.as_device(np.asarray(phi, np.float64),
                                                                               torch.float64)
        self.mlp = ProposalMLP(seed=seed)
        Ls = self.L * self.sub
        # bucket edges: quantiles (linear interpolation between order statistics) of the field at a random
        # sample of sub-cell centres, all on the device from the sorted sample
        dev = grid.device
        rng = np.random.default_rng(seed)
        idx = torch.from_numpy(rng.integers(0, Ls ** 3, size=min(int(n_quantile), Ls ** 3))).to(dev)
        centres = (torch.stack([idx // (Ls * Ls), (idx // Ls) % Ls, idx % Ls], dim=1).double() + 0.5) / self.sub
        vs = torch.sort(kernels.interpolate(grid, centres.contiguous(), self.phi)).values
        n = int(vs.numel())
        pos = torch.linspace(0.0, 1.0, self.nbucket + 1, dtype=torch.float64, device=dev)[1:-1] * (n - 1)
        lo = pos.floor().long()
        hi = torch.clamp(lo + 1, max=n - 1)
        edges_t = vs[lo] + (vs[hi] - vs[lo]) * (pos - lo.double())
        self.edges = edges_t.cpu().numpy()
        self.cell_bucket, self.bucket_cells, self.bucket_off, counts = kernels.mc_bucketize(
            grid, self.phi, self.sub, self.edges)
        self.n_k = counts.astype(np.float64)
        # representative feature of each bucket: mean of asinh(v / scale) over its sampled sub-cells; bucket k
        # (edges[k - 1] <= v < edges[k]) is a run of the sorted sample, summed by prefix sums
        self.rho_scale = float(vs.std(correction=0)) or 1.0
        csum = torch.cat([torch.zeros(1, dtype=torch.float64, device=dev),
                          torch.cumsum(torch.asinh(vs / self.rho_scale), 0)])
        bnd = torch.cat([torch.zeros(1, dtype=torch.long, device=dev), torch.searchsorted(vs, edges_t),
                         torch.full((1,), n, dtype=torch.long, device=dev)])
        cnt = (bnd[1:] - bnd[:-1]).double().cpu().numpy()
        t_k = ((csum[bnd[1:]] - csum[bnd[:-1]]).cpu().numpy()) / np.maximum(cnt, 1.0)
        te = np.arcsinh(np.asarray(self.edges) / self.rho_scale)
        for k in np.nonzero(cnt == 0)[0]:  # unsampled bucket: its edge interval
            t_k[k] = te[min(k, te.size - 1)] if te.size else 0.0
        self.t_k = t_k
        self.t_scale = float(max(np.abs(self.t_k).max(), 1e-6))
        self.zeta_pilot = None
        self._recording = self._struct(np.ones(self.nbucket) / self.nbucket,
                                       np.full((self.nbucket, bins.B), 1.0 / bins.B), 1.0)

    def features(self, r1, r2, mu, t):
        b = self.bins
        x = torch.stack([
            (r1 - b.r1[0]) / (b.r1[1] - b.r1[0]) * 2 - 1,
            (r2 - b.r2[0]) / (b.r2[1] - b.r2[0]) * 2 - 1,
            mu,
            t / self.t_scale,
        ], dim=1)
        return x.to(torch.float32).contiguous()

    @property
    def recording(self):
        """Proposal struct for plain draws (alpha = 1) that record every vertex's bucket."""
        return self._recording

    def fit(self, records, pilot, epochs=3, log=None, lr=3e-3, batch=16384):
        """records: (n, 8) float32 device tensor from a plain run with ``recording``;
        pilot: its MCResult.  One training pair per record and anchor vertex."""
        r1, r2, mu, f = (records[:, i].double() for i in range(4))
        tk = torch.from_numpy(self.t_k).to(records.device)
        xs = []
        for a in range(self.anchors):
            k = records[:, 5 + a].long()
            if bool((k < 0).any()):
                raise ValueError("records lack proposal buckets: record a plain run with .recording")
            xs.append(self.features(r1, r2, mu, tk[k]))
        x = torch.cat(xs).contiguous()
        y = (f * f).repeat(self.anchors)
        y = (y / y.mean()).to(torch.float32).contiguous()
        self.zeta_pilot = np.abs(pilot.estimate).ravel()
        return self.mlp.train(x, y, epochs=epochs, batch=batch, lr=lr, log=log)

    def joint(self, allocation="uniform"):
        """Proposal mass Q[k, b] of density bucket k and triangle bin b (sums to 1).

        Within a bin the centre density follows sqrt(E[f^2 | bin, bucket]) from the
        MLP.  ``allocation`` spreads the mass over bins: "uniform" keeps every
        bin's expected share at 1/B, as in the plain estimator, so each bin gains
        on its own; "relative" weights bin b by 1/|zeta_b(pilot)|, minimising the
        summed relative variance over bins.
        """
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
        mass = self.n_k[:, None] * np.exp(0.5 * s)
        if allocation == "uniform":
            Q = mass / np.maximum(mass.sum(axis=0, keepdims=True), 1e-300) / B
        elif allocation == "relative":
            if self.zeta_pilot is None:
                raise ValueError("relative allocation needs a pilot run (fit)")
            z = np.maximum(self.zeta_pilot, 1e-3 * self.zeta_pilot.max() + 1e-300)
            Q = mass / z[None, :]
            Q = Q / Q.sum()
        else:
            raise ValueError(f"unknown allocation {allocation!r}")
        return Q

    def build(self, allocation="uniform"):
        """Evaluate the MLP on (bucket, bin) centres and upload the proposal tables."""
        B = self.bins.B
        Q = self.joint(allocation)
        P = Q.sum(axis=1)
        bin_pdf = Q / np.maximum(P[:, None], 1e-300)
        bin_pdf[P <= 0] = 1.0 / B
        self.struct = self._struct(P, bin_pdf, self.alpha)
        return self

    def _struct(self, P, bin_pdf, alpha):
        V = float(self.L * self.sub) ** 3
        B = self.bins.B
        Pm = np.where(self.n_k > 0, P, 0.0)
        Pm = Pm / Pm.sum()
        dens = np.where(self.n_k > 0, V * B * Pm / np.maximum(self.n_k, 1.0), 0.0)
        cdf = np.cumsum(Pm).astype(np.float32)
        last = int(np.nonzero(Pm > 0)[0].max())
        cdf[last:] = 2.0
        bcdf = np.cumsum(bin_pdf, axis=1).astype(np.float32)
        bcdf[:, -1] = 2.0
        tables = {
            "bucket_cdf": kernels.as_device(cdf, torch.float32),
            "bucket_density": kernels.as_device(dens, torch.float64),
            "bin_cdf": kernels.as_device(bcdf, torch.float32),
            "bin_pdf": kernels.as_device(bin_pdf.astype(np.float32), torch.float32),
        }
        st = _lib.MCProposal()
        st.cell_bucket = self.cell_bucket.data_ptr()
        st.bucket_cells = self.bucket_cells.data_ptr()
        st.bucket_off = self.bucket_off.data_ptr()
        st.bucket_cdf = tables["bucket_cdf"].data_ptr()
        st.bucket_density = tables["bucket_density"].data_ptr()
        st.bin_cdf = tables["bin_cdf"].data_ptr()
        st.bin_pdf = tables["bin_pdf"].data_ptr()
        st.nbucket = self.nbucket
        st.sub = self.sub
        st.alpha = float(alpha)
        st.anchors = self.anchors
        st._tables = tables  # keeps the device arrays alive with the struct
        return st


def compare(plain, imp):
    """Variance gain and bias z-scores of an importance run against a plain run."""
    ok = (plain.variance > 0) & (imp.variance > 0)
    gain = plain.variance[ok] / imp.variance[ok] * (imp.n / plain.n)  # per-triangle efficiency ratio
    z = (imp.estimate - plain.estimate)[ok] / np.sqrt(plain.variance[ok] + imp.variance[ok])
    return {"median_gain": float(np.median(gain)), "mean_gain": float(np.mean(gain)),
            "total_gain": float(plain.variance[ok].sum() / imp.variance[ok].sum() * imp.n / plain.n),
            "z_mean": float(z.mean()), "z_rms": float(np.sqrt(np.mean(z ** 2))), "bins": int(ok.sum())}
