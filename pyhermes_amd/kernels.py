"""Tensor-level wrappers over the HIP C ABI (include/hermes_amd.h).

All tensors are torch tensors on the HIP device; outputs are allocated here
(caller-owned afterwards) unless passed in.  No CPU fallback exists.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import ptr

_DP = ctypes.POINTER(ctypes.c_double)


def ctypes_byref(obj):
    return ctypes.byref(obj)


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


def fft3_c2c_f32(x, inverse=False):
    """In-place complex64 3-D FFT of (.., L, L, L) by the LDS-tiled passes (unnormalised)."""
    if x.dtype != torch.complex64 or not x.is_contiguous() or x.ndim < 3:
        raise ValueError("fft3_c2c_f32 needs a contiguous complex64 (.., L, L, L) tensor")
    L = x.shape[-1]
    if x.shape[-3:] != (L, L, L):
        raise ValueError(f"fft3_c2c_f32: grid shape {tuple(x.shape[-3:])} is not cubic")
    batch = int(np.prod(x.shape[:-3])) if x.ndim > 3 else 1
    _lib.call("hermes_fft3_c2c_f32", _lib.context(), _J_of(L), batch, 1 if inverse else 0, ptr(x))
    return x


def fft3_tile_choice(J, inverse=True):
    """(tile width, threads) the autotuner picked for the strided FFT passes, or (0, 0) if untuned."""
    tw, nt = ctypes.c_int(0), ctypes.c_int(0)
    _lib.call("hermes_fft3_tile_choice", _lib.context(), int(J), 1 if inverse else 0, ctypes.byref(tw),
              ctypes.byref(nt))
    return tw.value, nt.value


def fft3_set_tile(J, tw, threads, inverse=True):
    """Pin the strided-pass tile shape (one of the autotuner's candidates)."""
    _lib.call("hermes_fft3_set_tile", _lib.context(), int(J), 1 if inverse else 0, int(tw), int(threads))


# ---------------------------------------------------------------- windows
WIN_CODES = {"sphere": 0, "gaussian": 1, "shell": 2, "thick_shell": 3, "gaussian_shell": 4,
             "inverse_laplacian": 5, "shell_multipole": 6, "ring": 7, "disk": 8, "cylshell": 9, "cylinder": 10}
LOS_TYPES = ("ring", "disk", "cylshell", "cylinder")


def window_desc(wtype, a0=0.0, a1=0.0, l=0, m=0, axis=2, los=None):
    """A hermes_window_desc; lengths in grid units.  axis: line of sight (0 x, 1 y, 2 z, or 3 along the
    unit vector ``los``) of the ring, disk, cylshell and cylinder windows (a0 = R, a1 = H); ignored by the
    others."""
    if wtype in LOS_TYPES and int(axis) not in (0, 1, 2, 3):
        raise ValueError(f"line-of-sight axis must be 0, 1, 2 or 3, got {axis!r}")
    general = wtype in LOS_TYPES and int(axis) == 3
    if general:
        n = np.asarray(los, dtype=np.float64).reshape(3)
        if not np.all(np.isfinite(n)) or abs(float(n @ n) - 1.0) > 1e-12:
            raise ValueError(f"axis 3 needs a unit line-of-sight vector, got {los!r}")
    d = _lib.WindowDesc(WIN_CODES[wtype], int(l), int(m), int(axis) if wtype in LOS_TYPES else 0, float(a0),
                        float(a1))
    if general:
        d.los[0], d.los[1], d.los[2] = (float(v) for v in n)
    return d


def _np_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def window_build_rfft(J, P, desc, dbl=True):
    """(L, L, L//2+1) real kernel of one isotropic window (P: device tensor of L+1 doubles)."""
    L = 1 << J
    out = torch.empty((L, L, L // 2 + 1), dtype=torch.float64 if dbl else torch.float32, device=P.device)
    _lib.call("hermes_window_build_rfft", _lib.context(), int(J), ptr(P), ctypes_byref(desc), 1 if dbl else 0,
              ptr(out))
    return out


def window_build_full(J, P, desc, dbl=True):
    """(L, L, L) complex kernel of one window."""
    L = 1 << J
    out = torch.empty((L, L, L), dtype=torch.complex128 if dbl else torch.complex64, device=P.device)
    _lib.call("hermes_window_build_full", _lib.context(), int(J), ptr(P), ctypes_byref(desc), 1 if dbl else 0,
              ptr(out))
    return out


RADIAL_KINDS = {"shell": 0, "table": 1, "sphere": 2, "gaussian": 3, "thick_shell": 4, "gaussian_shell": 5}


def radial_profile(kind, a0=0.0, a1=0.0, table=None, k_scale=0.0):
    """A hermes_radial_profile: lengths a0, a1 in grid units; for kind 'table' a device float64 tensor of U_l
    samples with k_scale samples per cycle per grid cell (the tensor must outlive every use)."""
    if kind not in RADIAL_KINDS:
        raise ValueError(f"unknown radial profile kind {kind!r}")
    if kind == "table" and (table is None or table.dtype != torch.float64 or not table.is_contiguous()):
        raise ValueError("a table profile needs a contiguous float64 device tensor")
    return _lib.RadialProfile(RADIAL_KINDS[kind], 0 if table is None else int(table.numel()), float(a0), float(a1),
                              float(k_scale), None if table is None else int(table.data_ptr()))


def window_build_full_profile(J, P, prof, l, m, dbl=True):
    """(L, L, L) complex kernel U_l(|k|) Y_lm(k-hat) of one radial profile (alias sums as the shell builder)."""
    L = 1 << J
    out = torch.empty((L, L, L), dtype=torch.complex128 if dbl else torch.complex64, device=P.device)
    _lib.call("hermes_window_build_full_profile", _lib.context(), int(J), ptr(P), ctypes_byref(prof), int(l), int(m),
              1 if dbl else 0, ptr(out))
    return out


def multipole_window_ifft_profiles_f32(H, P, profiles, l, ms, scale, out):
    """multipole_window_ifft_multi_f32 with a radial profile per shell (``profiles``: RadialProfile list)."""
    S = len(profiles)
    ms = [int(m) for m in ms]
    L = _check_multipole_inputs(H, P, len(ms) * S, out)
    arr = (ctypes.c_int * len(ms))(*ms)
    profs = (_lib.RadialProfile * S)(*profiles)
    _lib.call("hermes_multipole_window_ifft_profiles_f32", _lib.context(), _J_of(L), ptr(H), ptr(P), profs, S,
              int(l), arr, len(ms), float(scale), ptr(out))
    return out


def spectrum_mul(spec, kernel, scale=1.0):
    """spec *= kernel * scale in place (half spectrum with a real kernel, or full with a complex one)."""
    L = spec.shape[0]
    half = 1 if spec.shape[-1] != L else 0
    dbl = 1 if spec.dtype == torch.complex128 else 0
    kc = 1 if kernel.is_complex() else 0
    _lib.call("hermes_spectrum_mul", _lib.context(), _J_of(L), half, dbl, ptr(spec), ptr(kernel), kc, float(scale))
    return spec


PAIR_WINDOWS_PER_PASS = 2048  # one fused pass keeps a sum per window and wave in 64 KB of LDS


def pair_reduce(F1, F2, P, descs):
    """sum_k K_b(k) Re(F1 conj F2) over the full grid, for every descriptor (host float64 array); longer
    descriptor lists run as several fused passes of at most PAIR_WINDOWS_PER_PASS windows."""
    L = F1.shape[0]
    descs = list(descs)
    out = np.zeros(len(descs), dtype=np.float64)
    dbl = 1 if F1.dtype == torch.complex128 else 0
    for b0 in range(0, len(descs), PAIR_WINDOWS_PER_PASS):
        part = descs[b0:b0 + PAIR_WINDOWS_PER_PASS]
        arr = (_lib.WindowDesc * len(part))(*part)
        chunk = np.zeros(len(part), dtype=np.float64)
        _lib.call("hermes_pair_reduce", _lib.context(), _J_of(L), dbl, ptr(F1), ptr(F2), ptr(P), arr, len(part),
                  chunk.ctypes.data_as(_DP))
        out[b0:b0 + len(part)] = chunk
    return out


def multipole_apply_f32(H, P, radii_grid, l, m, scale, out=None):
    """(S, L, L, L) complex64 spectra H_full * K_{R_s,l,m} * scale from the half spectrum H (complex64)."""
    r = _np_f64(radii_grid)
    S = r.size
    L = _check_multipole_inputs(H, P, S, out)
    if out is None:
        out = torch.empty((S, L, L, L), dtype=torch.complex64, device=H.device)
    _lib.call("hermes_multipole_apply_f32", _lib.context(), _J_of(L), ptr(H), ptr(P), r.ctypes.data_as(_DP), S,
              int(l), int(m), float(scale), ptr(out))
    return out


def _check_multipole_inputs(H, P, S, out):
    if H.dtype != torch.complex64 or not H.is_contiguous() or H.ndim != 3:
        raise ValueError("half spectrum H must be a contiguous complex64 (L, L, L/2+1) tensor")
    L = H.shape[0]
    if tuple(H.shape) != (L, L, L // 2 + 1):
        raise ValueError(f"half spectrum shape {tuple(H.shape)} != {(L, L, L // 2 + 1)}")
    if P.dtype != torch.float64 or P.numel() < L + 1:
        raise ValueError("P must hold L + 1 float64 values")
    if out is not None and (out.dtype != torch.complex64 or not out.is_contiguous()
                            or tuple(out.shape) != (S, L, L, L)):
        raise ValueError(f"out must be a contiguous complex64 {(S, L, L, L)} tensor")
    return L


def multipole_window_ifft_f32(H, P, radii_grid, l, m, scale, out=None):
    """(S, L, L, L) complex64 fields IFFT3(H_full * K_{R_s,l,m} * scale), window and inverse FFT fused."""
    r = _np_f64(radii_grid)
    S = r.size
    L = _check_multipole_inputs(H, P, S, out)
    if out is None:
        out = torch.empty((S, L, L, L), dtype=torch.complex64, device=H.device)
    _lib.call("hermes_multipole_window_ifft_f32", _lib.context(), _J_of(L), ptr(H), ptr(P), r.ctypes.data_as(_DP),
              S, int(l), int(m), float(scale), ptr(out))
    return out


def multipole_window_ifft_multi_f32(H, P, radii_grid, l, ms, scale, out):
    """(len(ms) * S, L, L, L) complex64: the window + inverse FFT of every shell for each order in ``ms``
    (1 or 2 orders of one l, sharing the radial factors); order i's shell s at index i * S + s."""
    r = _np_f64(radii_grid)
    S = r.size
    ms = [int(m) for m in ms]
    L = _check_multipole_inputs(H, P, len(ms) * S, out)
    arr = (ctypes.c_int * len(ms))(*ms)
    _lib.call("hermes_multipole_window_ifft_multi_f32", _lib.context(), _J_of(L), ptr(H), ptr(P),
              r.ctypes.data_as(_DP), S, int(l), arr, len(ms), float(scale), ptr(out))
    return out


def multipole_window_zpass_multi_f32(H, P, radii_grid, l, ms, scale, out):
    """First half of multipole_window_ifft_multi_f32: the windows and the z pass only (``out`` holds the
    fields transformed along z); fft3_inverse_yx_f32 completes them, possibly on another stream."""
    r = _np_f64(radii_grid)
    S = r.size
    ms = [int(m) for m in ms]
    L = _check_multipole_inputs(H, P, len(ms) * S, out)
    arr = (ctypes.c_int * len(ms))(*ms)
    _lib.call("hermes_multipole_window_zpass_multi_f32", _lib.context(), _J_of(L), ptr(H), ptr(P),
              r.ctypes.data_as(_DP), S, int(l), arr, len(ms), float(scale), ptr(out))
    return out


def fft3_inverse_yx_ymajor_f32(src, dst):
    """The y and x inverse passes of the x-major batch ``src`` (z lines already transformed) into ``dst`` in
    y-major order (dst[..., y, x, z] holds the transform at (x, y, z)); src is only read."""
    for t in (src, dst):
        if t.dtype != torch.complex64 or not t.is_contiguous() or t.ndim < 3:
            raise ValueError("fft3_inverse_yx_ymajor_f32 needs contiguous complex64 (.., L, L, L) tensors")
    if src.shape != dst.shape:
        raise ValueError(f"fft3_inverse_yx_ymajor_f32: shapes {tuple(src.shape)} and {tuple(dst.shape)} differ")
    L = src.shape[-1]
    batch = int(np.prod(src.shape[:-3])) if src.ndim > 3 else 1
    _lib.call("hermes_fft3_inverse_yx_ymajor_f32", _lib.context(), _J_of(L), batch, ptr(src), ptr(dst))
    return dst


FIELD_LAYOUTS = {"xmajor": 0, "ymajor": 1}


def get_field_layout():
    """Layout of the fields the Gram contractions read: "xmajor" (default) or "ymajor"."""
    v = ctypes.c_int(0)
    _lib.call("hermes_get_field_layout", _lib.context(), ctypes.byref(v))
    return {c: n for n, c in FIELD_LAYOUTS.items()}[v.value]


def set_field_layout(name):
    """Select the layout of the fields U, V read by the multipole_gram_f32* contractions."""
    if name not in FIELD_LAYOUTS:
        raise ValueError(f"field layout {name!r} is not one of {sorted(FIELD_LAYOUTS)}")
    _lib.call("hermes_set_field_layout", _lib.context(), FIELD_LAYOUTS[name])


class field_layout:
    """Context manager: the Gram contractions read fields in layout ``name`` inside the block."""

    def __init__(self, name):
        self.name = name
        self.prev = None

    def __enter__(self):
        self.prev = get_field_layout()
        set_field_layout(self.name)
        return self

    def __exit__(self, *exc):
        set_field_layout(self.prev)
        return False


def fft3_inverse_yx_f32(x):
    """The y and x inverse passes (unnormalised) of a contiguous complex64 (.., L, L, L) batch whose z
    lines are already transformed."""
    if x.dtype != torch.complex64 or not x.is_contiguous() or x.ndim < 3:
        raise ValueError("fft3_inverse_yx_f32 needs a contiguous complex64 (.., L, L, L) tensor")
    L = x.shape[-1]
    batch = int(np.prod(x.shape[:-3])) if x.ndim > 3 else 1
    _lib.call("hermes_fft3_inverse_yx_f32", _lib.context(), _J_of(L), batch, ptr(x))
    return x


def reduce_sum(x):
    """Deterministic FP64 sum of a real or complex tensor."""
    out = np.zeros(2, dtype=np.float64)
    comp = 2 if x.is_complex() else 1
    dbl = 1 if x.dtype in (torch.float64, torch.complex128) else 0
    n = x.numel()
    _lib.call("hermes_reduce_sum", _lib.context(), ptr(x.contiguous()), n, comp, dbl, out.ctypes.data_as(_DP))
    return complex(out[0], out[1]) if comp == 2 else float(out[0])


# ---------------------------------------------------------------- contraction
def multipole_summand_f64(d1, F1, F2, gamma):
    """Raw complex sum of the plain FP64 triple-product contraction (no 4 pi / N)."""
    L = d1.shape[0]
    g = _np_f64(gamma)
    out = np.zeros(2, dtype=np.float64)
    _lib.call("hermes_multipole_summand_f64", _lib.context(), _J_of(L), ptr(d1), ptr(F1), ptr(F2),
              g.ctypes.data_as(_DP), out.ctypes.data_as(_DP))
    return complex(out[0], out[1])


def multipole_gram_f32(d1, U, Q, lam, V=None, v_conj=False):
    """(n_u, n_v) FP64 Gram sums; see hermes_multipole_gram_f32."""
    L = d1.shape[0]
    n_u = U.shape[0]
    n_v = V.shape[0] if V is not None else n_u
    q = _np_f64(Q)
    lm = _np_f64(lam)
    out = np.zeros((n_u, n_v), dtype=np.float64)
    _lib.call("hermes_multipole_gram_f32", _lib.context(), _J_of(L), ptr(d1), ptr(U), n_u, ptr(V), n_v,
              1 if v_conj else 0, q.ctypes.data_as(_DP), lm.ctypes.data_as(_DP), out.ctypes.data_as(_DP))
    return out


def multipole_gram_f32_dev(d1, U, Q, lam, out_dev, V=None, v_conj=False):
    """Raw 16 x 16 Gram block into out_dev (256 doubles on the device, no host sync); see
    hermes_multipole_gram_f32_dev."""
    L = d1.shape[0]
    n_u = U.shape[0]
    n_v = V.shape[0] if V is not None else n_u
    if out_dev.dtype != torch.float64 or out_dev.numel() < 256 or not out_dev.is_contiguous():
        raise ValueError("out_dev must be a contiguous float64 tensor of at least 256 elements")
    q = _np_f64(Q)
    lm = _np_f64(lam)
    _lib.call("hermes_multipole_gram_f32_dev", _lib.context(), _J_of(L), ptr(d1), ptr(U), n_u, ptr(V), n_v,
              1 if v_conj else 0, q.ctypes.data_as(_DP), lm.ctypes.data_as(_DP), ptr(out_dev))
    return out_dev


def multipole_gram_f32_dev_ri(d1, U, Q, lam, out_dev):
    """Auto Gram with the real-part and imaginary-part products summed apart: out_dev (512 doubles on the
    device) holds the real block then the imaginary block; see hermes_multipole_gram_f32_dev_ri."""
    L = d1.shape[0]
    if out_dev.dtype != torch.float64 or out_dev.numel() < 512 or not out_dev.is_contiguous():
        raise ValueError("out_dev must be a contiguous float64 tensor of at least 512 elements")
    q = _np_f64(Q)
    lm = _np_f64(lam)
    _lib.call("hermes_multipole_gram_f32_dev_ri", _lib.context(), _J_of(L), ptr(d1), ptr(U), U.shape[0],
              q.ctypes.data_as(_DP), lm.ctypes.data_as(_DP), ptr(out_dev))
    return out_dev


def multipole_window_ifft_pack0_f32(H, P, radii_grid, l, scale, out):
    """(S, L, L, L) complex64: the (l, 0) and (l + 1, 0) windows + inverse FFT of every shell packed into
    one field per shell; the even degree's field is its real part, the odd degree's its imaginary part."""
    r = _np_f64(radii_grid)
    S = r.size
    L = _check_multipole_inputs(H, P, S, out)
    _lib.call("hermes_multipole_window_ifft_pack0_f32", _lib.context(), _J_of(L), ptr(H), ptr(P),
              r.ctypes.data_as(_DP), S, int(l), float(scale), ptr(out))
    return out


def multipole_window_zpass_pack0_f32(H, P, radii_grid, l, scale, out):
    """The z pass of multipole_window_ifft_pack0_f32 alone (out: S fields transformed along z only)."""
    r = _np_f64(radii_grid)
    S = r.size
    L = _check_multipole_inputs(H, P, S, out)
    _lib.call("hermes_multipole_window_zpass_pack0_f32", _lib.context(), _J_of(L), ptr(H), ptr(P),
              r.ctypes.data_as(_DP), S, int(l), float(scale), ptr(out))
    return out


def gram_block_to_matrix(block, n_u, n_v=None):
    """(n_u, n_v) Gram sums from a raw block of hermes_multipole_gram_f32_dev (auto when n_v is None)."""
    h = np.asarray(block, dtype=np.float64).reshape(16, 16)
    if n_v is None:
        return h[:n_u, :n_u] + h[:n_u, :n_u].T
    return h[:n_u, :n_v].copy()


def profile_gram(enable):
    """Start (True: reset) or stop timing the Gram kernel; returns (total_ms, launches) so far."""
    ms, n = ctypes.c_double(0.0), ctypes.c_int64(0)
    _lib.call("hermes_profile_gram", _lib.context(), 1 if enable else 0, ctypes.byref(ms), ctypes.byref(n))
    return ms.value, n.value


def mfma_selftest():
    out = np.zeros(1, dtype=np.float64)
    _lib.call("hermes_mfma_selftest", _lib.context(), out.ctypes.data_as(_DP))
    return float(out[0])


# ---------------------------------------------------------------- graphs and batch helpers
def capture_begin():
    """Start capturing the library stream (torch's current stream; must not be the null stream)."""
    _lib.call("hermes_capture_begin", _lib.context())


def capture_end():
    """End the capture; returns an instantiated graph handle for graph_launch / graph_destroy."""
    h = ctypes.c_void_p()
    _lib.call("hermes_capture_end", _lib.context(), ctypes.byref(h))
    return h


def graph_launch(exec_handle):
    _lib.call("hermes_graph_launch", _lib.context(), exec_handle)


def graph_destroy(exec_handle):
    if exec_handle:
        _lib.call("hermes_graph_destroy", exec_handle)


def clear_errors():
    _lib.call("hermes_clear_errors", _lib.context())


def read_errors():
    v = ctypes.c_int(0)
    _lib.call("hermes_read_errors", _lib.context(), ctypes.byref(v))
    return int(v.value)


def project_reserve(J, n_max):
    _lib.call("hermes_project_reserve", _lib.context(), int(J), int(n_max))


def release_workspaces():
    """Free the library's scratch workspaces on the current device (they grow again on demand)."""
    _lib.call("hermes_release_workspaces", _lib.context())


def delta_f32(eps32, a, b, d_out=None, dc_out=None):
    """d = a * eps - b (FP32) into d_out (real) and/or dc_out (complex64 real part)."""
    if d_out is None and dc_out is None:
        raise ValueError("delta_f32 needs an output")
    for t, dt in ((d_out, torch.float32), (dc_out, torch.complex64)):
        if t is not None and (t.dtype != dt or t.numel() != eps32.numel() or not t.is_contiguous()):
            raise ValueError("delta_f32: output shape/dtype mismatch")
    _lib.call("hermes_delta_f32", _lib.context(), int(eps32.numel()), ptr(eps32), float(a), float(b), ptr(d_out),
              ptr(dc_out))


def half_spectrum_f32(full, half):
    """half (L, L, L/2+1) <- full (L, L, L) complex64 transform of a real field."""
    L = full.shape[-1]
    if full.dtype != torch.complex64 or half.dtype != torch.complex64 or half.shape[-1] != L // 2 + 1:
        raise ValueError("half_spectrum_f32: expects complex64 (L,L,L) -> (L,L,L/2+1)")
    _lib.call("hermes_half_spectrum_f32", _lib.context(), _J_of(L), ptr(full), ptr(half))
    return half


class PairPlan:
    """A fused pair-reduction plan (hermes_pair_plan): window tables on the device, results on the device."""

    def __init__(self, J, descs, dbl=False):
        self.n = len(descs)
        arr = (_lib.WindowDesc * self.n)(*descs)
        h = ctypes.c_void_p()
        _lib.call("hermes_pair_plan_create", _lib.context(), int(J), 1 if dbl else 0, arr, self.n, ctypes.byref(h))
        self.handle = h
        self.J = int(J)

    def reduce(self, F1, F2, P, out_dev):
        if out_dev.dtype != torch.float64 or out_dev.numel() < self.n:
            raise ValueError("out_dev must hold n_desc float64 values")
        _lib.call("hermes_pair_reduce_planned", _lib.context(), self.handle, ptr(F1), ptr(F2), ptr(P), ptr(out_dev))
        return out_dev

    def close(self):
        if self.handle:
            _lib.load().hermes_pair_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- Monte Carlo
def interpolate(grid, spos, phi, offset=(0.0, 0.0, 0.0), out=None):
    """n(x) at scaled positions (+offset) from an FP64 coefficient grid (device tensors)."""
    L = grid.shape[0]
    phi_t = phi if isinstance(phi, torch.Tensor) else as_device(np.asarray(phi, dtype=np.float64), torch.float64)
    n = int(spos.shape[0])
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=grid.device)
    _lib.call("hermes_interpolate", _lib.context(), _J_of(L), ptr(grid), ptr(spos), n, ptr(phi_t), 1024,
              float(offset[0]), float(offset[1]), float(offset[2]), ptr(out))
    return out


def mc_triplet_box(R1, R2, mu, spos, n1, abc, grid2, grid3, phi):
    """Box-random-centre triplet estimate for one (R1, R2, mu) (grid units)."""
    L = grid2.shape[0]
    phi_t = phi if isinstance(phi, torch.Tensor) else as_device(np.asarray(phi, dtype=np.float64), torch.float64)
    abc = _np_f64(abc)
    out = np.zeros(1, dtype=np.float64)
    _lib.call("hermes_mc_triplet_box", _lib.context(), _J_of(L), float(R1), float(R2), float(mu), ptr(spos), ptr(n1),
              int(spos.shape[0]), int(abc.shape[0]), abc.ctypes.data_as(_DP), ptr(grid2), ptr(grid3), ptr(phi_t), 1024,
              out.ctypes.data_as(_DP))
    return float(out[0])


def mc_binned(g1, g2, g3, phi, nbins, rrange1, rrange2, n_samples, seed, wf_cap=0.0, proposal=None, record=None):
    """Binned triangle sampler; returns (sum_wf, sum_wf2, counts) as (nr1, nr2, nmu) host arrays.

    ``wf_cap`` sets the fixed-point scale of the sums (0: measured on a pilot pass).
    """
    L = g1.shape[0]
    phi_t = phi if isinstance(phi, torch.Tensor) else as_device(np.asarray(phi, dtype=np.float64), torch.float64)
    nr1, nr2, nmu = (int(v) for v in nbins)
    B = nr1 * nr2 * nmu
    swf = np.zeros(B)
    swf2 = np.zeros(B)
    cnt = np.zeros(B, dtype=np.int64)
    prop = ctypes.byref(proposal) if proposal is not None else None
    if record is not None and (record.dtype != torch.float32 or record.ndim != 2 or record.shape[1] != 8
                               or not record.is_contiguous()):
        raise ValueError("record must be a contiguous (n, 8) float32 tensor")
    rec_ptr = ptr(record) if record is not None else None
    n_rec = int(record.shape[0]) if record is not None else 0
    _lib.call("hermes_mc_binned", _lib.context(), _J_of(L), ptr(g1), ptr(g2), ptr(g3), ptr(phi_t), 1024, nr1, nr2,
              nmu, float(rrange1[0]), float(rrange1[1]), float(rrange2[0]), float(rrange2[1]), int(n_samples),
              int(seed) & 0xFFFFFFFFFFFFFFFF, float(wf_cap), prop, rec_ptr, n_rec, swf.ctypes.data_as(_DP),
              swf2.ctypes.data_as(_DP), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    shape = (nr1, nr2, nmu)
    return swf.reshape(shape), swf2.reshape(shape), cnt.reshape(shape)


def mc_bucketize(grid, phi, sub, edges):
    """Proposal buckets on the (L*sub)^3 sub-cell grid: (cell_bucket u8, bucket_cells i32, bucket_off i64)
    device tensors and the per-bucket sub-cell counts (host int64)."""
    L = grid.shape[0]
    Ls = L * int(sub)
    phi_t = phi if isinstance(phi, torch.Tensor) else as_device(np.asarray(phi, dtype=np.float64), torch.float64)
    e = np.ascontiguousarray(edges, dtype=np.float64)
    nb = e.size + 1
    cb = torch.empty(Ls ** 3, dtype=torch.uint8, device=grid.device)
    cells = torch.empty(Ls ** 3, dtype=torch.int32, device=grid.device)
    off = torch.empty(nb + 1, dtype=torch.int64, device=grid.device)
    counts = np.zeros(nb, dtype=np.int64)
    _lib.call("hermes_mc_bucketize", _lib.context(), _J_of(L), ptr(grid), ptr(phi_t), 1024, int(sub),
              e.ctypes.data_as(_DP), nb, ptr(cb), ptr(cells), ptr(off),
              counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return cb, cells, off, counts


# ---------------------------------------------------------------- proposal MLP
def mlp_nparam():
    return int(_lib.load().hermes_mlp_nparam())


def mlp_to_bf16(w, wb):
    _lib.call("hermes_mlp_to_bf16", _lib.context(), ptr(w), ptr(wb))


def mlp_forward(wb, x, out=None):
    n = int(x.shape[0])
    if out is None:
        out = torch.empty(n, dtype=torch.float32, device=x.device)
    _lib.call("hermes_mlp_forward", _lib.context(), ptr(wb), ptr(x), n, ptr(out))
    return out


def mlp_grad(wb, x, y, clip, grad_out):
    _lib.call("hermes_mlp_grad", _lib.context(), ptr(wb), ptr(x), ptr(y), int(x.shape[0]), float(clip), ptr(grad_out))
    return grad_out


def mlp_train_steps(w, m, v, grad, wb, x, y, batch, clip, lrs, step0, beta1=0.9, beta2=0.999, eps=1e-8):
    """len(lrs) Adam steps over consecutive row ranges of batch rows of the contiguous (n, 4) x and (n,) y."""
    for t in (x, y):
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError("mlp_train_steps needs contiguous float32 x and y")
    n = int(x.shape[0])
    if x.ndim != 2 or x.shape[1] != 4 or y.shape != (n,):
        raise ValueError("mlp_train_steps: x must be (n, 4) and y (n,)")
    lr = np.ascontiguousarray(lrs, dtype=np.float32)
    _lib.call("hermes_mlp_train_steps", _lib.context(), ptr(w), ptr(m), ptr(v), ptr(grad), ptr(wb), ptr(x), ptr(y), n,
              int(batch), float(clip), lr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), int(lr.size), int(step0),
              float(beta1), float(beta2), float(eps))


def mlp_adam(w, m, v, grad, wb, lr, step, beta1=0.9, beta2=0.999, eps=1e-8):
    _lib.call("hermes_mlp_adam", _lib.context(), ptr(w), ptr(m), ptr(v), ptr(grad), ptr(wb), float(lr), float(beta1),
              float(beta2), float(eps), int(step))
