"""Tensor-level wrappers over the HIP C ABI (include/hermes_amd.h).

All tensors are torch tensors on the HIP device; outputs are allocated here
(caller-owned afterwards) unless passed in.  No CPU fallback exists.
"""

import numpy as np
import torch

from . import _lib
from ._lib import ptr


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("pyhermes_amd requires a HIP device")
    return torch.device("cuda", torch.cuda.current_device())


def as_device(a, dtype):
    """numpy / torch -> contiguous device tensor of ``dtype`` (no copy if already there)."""
    dev = _dev()
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dtype).contiguous()


# ---------------------------------------------------------------- projection
def project(pos, w, J, box, phi, want32=True, want64=False, want_cells=False, want_counts=False,
            phi_res=1024):
    """Scatter weighted particles onto the (L, L, L) scaling-coefficient grid.

    Replaces project_scaling_grid_numba (wavelet_grid.py:58-70).  Returns a dict
    with keys among eps32, eps64, cell, frac, counts.
    """
    ctx = _lib.context()
    pos_t = as_device(pos, torch.float64)
    w_t = as_device(w, torch.float64)
    phi_np = np.asarray(phi, dtype=np.float64)
    phi_t = as_device(phi_np, torch.float64)
    n = int(pos_t.shape[0])
    if pos_t.ndim != 2 or pos_t.shape[1] != 3 or w_t.shape != (n,):
        raise ValueError("project: pos must be (N, 3) and w must be (N,)")
    L = 1 << int(J)
    dev = pos_t.device
    out = {}
    e32 = torch.empty((L, L, L), dtype=torch.float32, device=dev) if want32 else None
    e64 = torch.empty((L, L, L), dtype=torch.float64, device=dev) if want64 else None
    cell = torch.empty((n, 3), dtype=torch.int32, device=dev) if want_cells else None
    frac = torch.empty((n, 3), dtype=torch.int32, device=dev) if want_cells else None
    cnt = torch.empty((L, L, L), dtype=torch.int32, device=dev) if want_counts else None
    _lib.call("hermes_project", ctx, ptr(pos_t), ptr(w_t), n, int(J), float(box), ptr(phi_t), phi_np.size,
              int(phi_res), float(np.abs(phi_np).max()), ptr(e32), ptr(e64), ptr(cell), ptr(frac), ptr(cnt))
    for k, v in (("eps32", e32), ("eps64", e64), ("cell", cell), ("frac", frac), ("counts", cnt)):
        if v is not None:
            out[k] = v
    return out


# ---------------------------------------------------------------- FFT
def _J_of(L):
    J = int(L).bit_length() - 1
    if (1 << J) != L:
        raise ValueError(f"grid size {L} is not a power of two")
    return J


def rfftn(x):
    """(.., L, L, L) real -> (.., L, L, L//2+1) complex, unnormalised (scipy rfftn convention)."""
    L = x.shape[-1]
    batch = int(np.prod(x.shape[:-3])) if x.ndim > 3 else 1
    dbl = 1 if x.dtype == torch.float64 else 0
    cdt = torch.complex128 if dbl else torch.complex64
    out = torch.empty(x.shape[:-1] + (L // 2 + 1,), dtype=cdt, device=x.device)
    _lib.call("hermes_fft_r2c", _lib.context(), _J_of(L), batch, dbl, ptr(x.contiguous()), ptr(out))
    return out


def irfftn_unnormalized(spec, L, out=None):
    """(.., L, L, L//2+1) complex -> (.., L, L, L) real, NOT divided by N.  ``spec`` is clobbered."""
    batch = int(np.prod(spec.shape[:-3])) if spec.ndim > 3 else 1
    dbl = 1 if spec.dtype == torch.complex128 else 0
    rdt = torch.float64 if dbl else torch.float32
    if out is None:
        out = torch.empty(spec.shape[:-1] + (L,), dtype=rdt, device=spec.device)
    _lib.call("hermes_fft_c2r", _lib.context(), _J_of(L), batch, dbl, ptr(spec), ptr(out))
    return out


def fftn_c2c(x, inverse=False, out=None):
    """Complex (.., L, L, L) FFT, unnormalised; in place when ``out is x``."""
    L = x.shape[-1]
    batch = int(np.prod(x.shape[:-3])) if x.ndim > 3 else 1
    dbl = 1 if x.dtype == torch.complex128 else 0
    if out is None:
        out = torch.empty_like(x)
    _lib.call("hermes_fft_c2c", _lib.context(), _J_of(L), batch, dbl, 1 if inverse else 0, ptr(x), ptr(out))
    return out
