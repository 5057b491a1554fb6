"""Device orchestration of the spherical-harmonic 3PCF multipoles.

Two paths compute ``ddd_l`` for shell pairs (r12, r13):

* ``ddd_multipole_plain_f64`` -- a per-set restatement of calc_DDD_multipole
  (corr3pcf_multipoles.py:460-579) on the GPU in FP64: for every (l, m >= 0) two
  window builds, two forward and two inverse complex FFTs and the FP64
  contraction, exactly as the reference sequences them.
* ``ddd_multipole_batched_f32`` -- all shell pairs at once: one real FFT per
  field, then per (l, m) one fused window + inverse-FFT call producing every
  shell's field (window, z pass, y and x passes) and one MFMA Gram contraction
  (FP32 products, FP64 accumulation).  When fields 2 and 3 and their shell lists are
  identical, F_{l,-m} = (-1)^(l+m) conj(F_{l,m}) removes the second set of FFTs.

Both return ddd_l following combine_multipole_m_terms (corr3pcf_multipoles.py:173-180):
  ddd_l = (-1)^l Re[v_0 + sum_{m>=1} (-1)^m (v_m + conj v_m)],
  v_m = 4 pi / N * sum_x d1 sum_{l1,l2} G3 F_{r12,l,m}[x-l1] F_{r13,l,-m}[x-l2].
"""

import numpy as np
import torch

from . import kernels
from .edges import EdgeProfile
from .utils.wavelet_grid import cal_gamma, fourier_power_spectrum


class MultipoleGeometry:
    """Host tables shared by both paths: |phi_hat|^2 (device), Gamma and its eigen-decomposition."""

    def __init__(self, phi, L, phi_resolution=1024, device=None):
        self.L = int(L)
        self.J = int(L).bit_length() - 1
        self.phi = np.asarray(phi, dtype=np.float64)
        self.P_host = fourier_power_spectrum(self.phi, 0, 1, self.L, phi_resolution)
        self.P = kernels.as_device(self.P_host, torch.float64)
        self.gamma = cal_gamma(self.phi, self.phi.size // phi_resolution, phi_resolution)
        lam, Q = np.linalg.eigh(self.gamma)
        self.lam = lam
        self.Q = Q  # Q[tap, eigenvector]


def combine_m_terms(v_re_by_m, l):
    """(-1)^l [Re v_0 + 2 sum_{m>=1} (-1)^m Re v_m] (only real parts enter)."""
    out = np.array(v_re_by_m[0], dtype=np.float64, copy=True)
    for m in range(1, l + 1):
        out = out + 2.0 * ((-1) ** m) * np.asarray(v_re_by_m[m])
    return ((-1) ** l) * out


# ---------------------------------------------------------------- plain FP64
def _multipole_field_f64(geo, delta_c, edge, l, m):
    """field @ radial_multipole window (complex_full_fft), FP64: fftn -> x K -> ifftn.  ``edge``: a thin-shell
    radius in grid units or an EdgeProfile."""
    if isinstance(edge, EdgeProfile):
        K = edge.window_full(geo.J, geo.P, l, m, dbl=True)
    else:
        K = kernels.window_build_full(geo.J, geo.P, kernels.window_desc("shell_multipole", edge, l=l, m=m), dbl=True)
    spec = kernels.fftn_c2c(delta_c)
    kernels.spectrum_mul(spec, K, 1.0 / geo.L ** 3)
    del K
    kernels.fftn_c2c(spec, inverse=True, out=spec)
    return spec


def ddd_multipole_plain_f64(geo, d1, d2, d3, R12_grid, R13_grid, l_min, l_max, return_m_values=False,
                            progress=None):
    """One (r12, r13) set, every (l, m) recomputed from scratch as the reference does.  d*: (L,L,L) float64."""
    N = geo.L ** 3
    d2c = torch.complex(d2, torch.zeros_like(d2))
    d3c = d2c if d3 is d2 else torch.complex(d3, torch.zeros_like(d3))
    out = []
    mvals = {}
    for l in range(l_min, l_max + 1):
        vm = np.empty(l + 1, dtype=np.complex128)
        for m in range(l + 1):
            F1 = _multipole_field_f64(geo, d2c, R12_grid, l, m)
            F2 = _multipole_field_f64(geo, d3c, R13_grid, l, -m)
            raw = kernels.multipole_summand_f64(d1, F1, F2, geo.gamma)
            vm[m] = 4.0 * np.pi * raw / N
            del F1, F2
        mvals[l] = vm
        out.append(combine_m_terms([v.real for v in vm], l))
        if progress is not None:
            progress(l)
    out = np.asarray(out, dtype=np.float64)
    return (out, mvals) if return_m_values else out


# ---------------------------------------------------------------- batched FP32
class StageTimer:
    """HIP-event timing of pipeline stages on torch's current stream (the stream the library uses)."""

    def __init__(self):
        self.events = {}

    def start(self, name):
        e0 = torch.cuda.Event(enable_timing=True)
        e0.record()
        return (name, e0)

    def stop(self, token):
        name, e0 = token
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record()
        self.events.setdefault(name, []).append((e0, e1))

    def summary(self):
        """{stage: (total_ms, count)} -- call after synchronising."""
        return {k: (sum(a.elapsed_time(b) for a, b in v), len(v)) for k, v in self.events.items()}


class _NoTimer:
    def start(self, name):
        return None

    def stop(self, token):
        pass


def _shell_blocks(n, size=16):
    return [(i, min(n, i + size)) for i in range(0, n, size)]


def _edges(values):
    """Thin-shell radii (grid units, float64 array) -- also when every EdgeProfile is a thin shell -- or the
    list of EdgeProfile for any other radial profile."""
    if isinstance(values, np.ndarray) or not any(isinstance(v, EdgeProfile) for v in values):
        return np.asarray(values, dtype=np.float64)
    if not all(isinstance(v, EdgeProfile) for v in values):
        raise TypeError("edges must be all radii or all EdgeProfile")
    radii = [v.thin_radius for v in values]
    return np.asarray(radii, dtype=np.float64) if all(r is not None for r in radii) else list(values)


def _window_runner(edges):
    """win(H, P, l, ms, scale, out): the fused window + inverse FFT of every edge for orders ms of one l."""
    if isinstance(edges, np.ndarray):
        return lambda H, P, l, ms, scale, out: kernels.multipole_window_ifft_multi_f32(H, P, edges, l, ms, scale, out)
    return lambda H, P, l, ms, scale, out: kernels.multipole_window_ifft_profiles_f32(
        H, P, [e.profile(l) for e in edges], l, ms, scale, out)


def _same_edges(a, b):
    if isinstance(a, np.ndarray) and isinstance(b, np.ndarray):
        return a.shape == b.shape and np.array_equal(a, b)
    return not isinstance(a, np.ndarray) and not isinstance(b, np.ndarray) and list(a) == list(b)


def ddd_multipole_batched_f32(geo, d1, d2, d3, radii12_grid, radii13_grid, l_min, l_max, timer=None,
                              schedule="serial", m_values=None, field_order="ymajor"):
    """All (r12, r13) shell pairs: returns (n12, n13, n_l) float64.  d*: (L,L,L) float32 device tensors.
    ``schedule`` picks the stream schedule of the auto case with at most 16 shells (see
    _ddd_auto_pipelined): "serial", "split" or "overlap".  A dict passed as ``m_values`` receives
    Re v_{l,m} (n12 x n13, m >= 0) per (l, m), the terms combine_multipole_m_terms sums.  ``field_order``: the
    packed auto path transforms its shell fields into y-major order ("ymajor", the default: the strided
    inverse-FFT passes then read along 4 KB strides only) or in place in x-major order ("xmajor")."""
    if field_order not in ("ymajor", "xmajor"):
        raise ValueError(f"field_order {field_order!r} is neither 'ymajor' nor 'xmajor'")
    L, J = geo.L, geo.J
    tm = timer if timer is not None else _NoTimer()
    N = L ** 3
    r12 = _edges(radii12_grid)
    r13 = _edges(radii13_grid)
    win12, win13 = _window_runner(r12), _window_runner(r13)
    n12, n13 = len(r12), len(r13)
    profiled = not isinstance(r12, np.ndarray) or not isinstance(r13, np.ndarray)
    auto = (d3 is d2) and _same_edges(r12, r13)
    t = tm.start("fft_r2c")
    H2 = kernels.rfftn(d2)
    H3 = H2 if d3 is d2 else kernels.rfftn(d3)
    tm.stop(t)
    n_l = l_max - l_min + 1
    result = np.zeros((n12, n13, n_l), dtype=np.float64)
    if auto and n12 <= 16:
        if not profiled and packs_m0(geo.L, n12, schedule):
            return _ddd_auto_packed(geo, d1, H2, r12, l_min, l_max, tm, m_values, overlap=schedule == "overlap",
                                    ymajor=field_order == "ymajor")
        return _ddd_auto_pipelined(geo, d1, H2, r12, l_min, l_max, tm, schedule="serial" if profiled else schedule,
                                   m_values=m_values, win=win12)
    blocks12 = _shell_blocks(n12)
    blocks13 = _shell_blocks(n13)
    U = torch.empty((n12, L, L, L), dtype=torch.complex64, device=d1.device)
    V = None if auto else torch.empty((n13, L, L, L), dtype=torch.complex64, device=d1.device)
    for li, l in enumerate(range(l_min, l_max + 1)):
        v_re = []
        for m in range(l + 1):
            t = tm.start("window_ifft")
            win12(H2, geo.P, l, [m], 1.0 / N, U)
            tm.stop(t)
            G = np.zeros((n12, n13))
            t = tm.start("gram")
            if auto:
                sign = (-1.0) ** (l + m)
                for a0, a1 in blocks12:
                    for b0, b1 in blocks13:
                        if b0 == a0:
                            g = kernels.multipole_gram_f32(d1, U[a0:a1], geo.Q, geo.lam)
                        else:
                            g = kernels.multipole_gram_f32(d1, U[a0:a1], geo.Q, geo.lam, V=U[b0:b1], v_conj=True)
                        G[a0:a1, b0:b1] = sign * g
            else:
                tm.stop(t)
                t = tm.start("window_ifft")
                win13(H3, geo.P, l, [-m], 1.0 / N, V)
                tm.stop(t)
                t = tm.start("gram")
                for a0, a1 in blocks12:
                    for b0, b1 in blocks13:
                        G[a0:a1, b0:b1] = kernels.multipole_gram_f32(d1, U[a0:a1], geo.Q, geo.lam, V=V[b0:b1])
            tm.stop(t)
            v_re.append(4.0 * np.pi * G / N)
            if m_values is not None:
                m_values[(l, m)] = v_re[-1]
        result[:, :, li] = combine_m_terms(v_re, l)
    return result


_SIDE = {}


def packs_m0(L, n_shell, schedule="serial"):
    """Whether the batched auto path packs the m = 0 terms of consecutive degrees in pairs (serial or
    overlap schedule, 9..16 shells so that each term has its own Gram, L <= 512)."""
    return schedule in ("serial", "overlap") and 8 < n_shell <= 16 and L <= 512


def fields_per_set(l_min, l_max, L, n_shell, schedule="serial"):
    """Shell-field groups (window + transform + Gram launches) the batched auto path computes."""
    terms = sum(l + 1 for l in range(l_min, l_max + 1))
    return terms - ((l_max - l_min + 1) // 2 if packs_m0(L, n_shell, schedule) else 0)


def _packed_units(l_min, l_max):
    """Work units of the packed auto path: the (l, 0) terms of consecutive degrees l_min, l_min + 1, ...
    in pairs (one field per shell), a last unpaired (l_max, 0), then the orders m >= 1 of each l two at
    a time (sharing the radial factors)."""
    units = []
    l = l_min
    while l + 1 <= l_max:
        units.append(("pack0", l, None))
        l += 2
    if l == l_max:
        units.append(("multi", l, [0]))
    for l in range(max(l_min, 1), l_max + 1):
        ms = list(range(1, l + 1))
        for i in range(0, len(ms), 2):
            units.append(("multi", l, ms[i:i + 2]))
    return units


def _ddd_auto_packed(geo, d1, H, r, l_min, l_max, tm, m_values=None, overlap=False, ymajor=True):
    """Auto case with 9..16 shells, the m = 0 terms packed in pairs: the windows of (l, 0) and (l + 1, 0)
    are real with opposite parity under k -> -k, so one transform yields the even degree's fields as real
    parts and the odd degree's as imaginary parts, and one Gram with the real and imaginary products summed
    apart separates them (for l <= 10: 61 field groups per set instead of 66).  Every other term is computed
    as in _ddd_auto_pipelined, one Gram per term.
    overlap: two field buffers; the windows and transforms of unit k + 1 run on the caller's stream while
    the Grams of unit k run on a second stream (the memory-bound strided FFT passes and the ALU-bound Gram
    share the CUs: a Gram workgroup takes at most 66 KB of LDS and half of a SIMD's registers).
    ymajor: the windows' z pass writes x-major fields into a staging buffer and the strided passes move them
    into y-major order in the buffer the Grams read (hermes_fft3_inverse_yx_ymajor_f32); otherwise the
    strided passes run in place."""
    L = geo.L
    N = L ** 3
    S = r.size
    units = _packed_units(l_min, l_max)
    bufs = [torch.empty((2 * S, L, L, L), dtype=torch.complex64, device=d1.device)
            for _ in range(2 if overlap and len(units) > 1 else 1)]
    stage = torch.empty((2 * S, L, L, L), dtype=torch.complex64, device=d1.device) if ymajor else None
    n_blocks = sum(1 if kind == "pack0" else len(ms) for kind, _, ms in units)
    blocks = torch.zeros((n_blocks, 512), dtype=torch.float64, device=d1.device)
    main = torch.cuda.current_stream(d1.device)
    side = _side_stream(d1.device) if overlap else main
    side.wait_stream(main)  # d1, H and the buffers are ready on the caller's stream
    gram_done = []
    owners = []
    for k, (kind, l, ms) in enumerate(units):
        U = bufs[k % len(bufs)]
        if overlap and k >= len(bufs):
            main.wait_event(gram_done[k - len(bufs)])  # the buffer's previous Grams are done
        t = tm.start("window_ifft")
        n = S if kind == "pack0" else len(ms) * S
        if ymajor:
            if kind == "pack0":
                kernels.multipole_window_zpass_pack0_f32(H, geo.P, r, l, 1.0 / N, stage[:S])
            else:
                kernels.multipole_window_zpass_multi_f32(H, geo.P, r, l, ms, 1.0 / N, stage[:n])
            kernels.fft3_inverse_yx_ymajor_f32(stage[:n], U[:n])
        elif kind == "pack0":
            kernels.multipole_window_ifft_pack0_f32(H, geo.P, r, l, 1.0 / N, U[:S])
        else:
            kernels.multipole_window_ifft_multi_f32(H, geo.P, r, l, ms, 1.0 / N, U[:n])
        tm.stop(t)
        if overlap:
            win_done = torch.cuda.Event()
            win_done.record(main)
        with torch.cuda.stream(side), kernels.field_layout("ymajor" if ymajor else "xmajor"):
            if overlap:
                side.wait_event(win_done)
            t = tm.start("gram")
            if kind == "pack0":
                kernels.multipole_gram_f32_dev_ri(d1, U[:S], geo.Q, geo.lam, blocks[len(owners)])
                owners.append(("pack0", l, None))
            else:
                for i, m in enumerate(ms):
                    kernels.multipole_gram_f32_dev(d1, U[i * S:(i + 1) * S], geo.Q, geo.lam,
                                                   blocks[len(owners), :256])
                    owners.append(("term", l, m))
            tm.stop(t)
            if overlap:
                ev = torch.cuda.Event()
                ev.record(side)
                gram_done.append(ev)
    main.wait_stream(side)
    host = blocks.cpu().numpy()
    v_re = {}
    for b, (kind, l, m) in enumerate(owners):
        if kind == "pack0":
            even, odd = (l, l + 1) if l % 2 == 0 else (l + 1, l)
            v_re[(even, 0)] = 4.0 * np.pi * kernels.gram_block_to_matrix(host[b, :256], S) / N
            v_re[(odd, 0)] = -4.0 * np.pi * kernels.gram_block_to_matrix(host[b, 256:], S) / N
        else:
            v_re[(l, m)] = 4.0 * np.pi * ((-1.0) ** (l + m)) * kernels.gram_block_to_matrix(host[b, :256], S) / N
    if m_values is not None:
        m_values.update(v_re)
    result = np.zeros((S, S, l_max - l_min + 1), dtype=np.float64)
    for li, l in enumerate(range(l_min, l_max + 1)):
        result[:, :, li] = combine_m_terms([v_re[(l, m)] for m in range(l + 1)], l)
    return result


def _side_stream(device):
    """A second stream (default priority) per device for the contraction of the pipelined path."""
    key = torch.device(device).index
    if key not in _SIDE:
        _SIDE[key] = torch.cuda.Stream(device=device)
    return _SIDE[key]


def _window_runs(group):
    """Split consecutive (l, m) terms into window launches: runs of one l, at most two orders each."""
    runs = []
    for l, m in group:
        if runs and runs[-1][0] == l and len(runs[-1][1]) < 2:
            runs[-1][1].append(m)
        else:
            runs.append((l, [m]))
    return runs


def _ddd_auto_pipelined(geo, d1, H, r, l_min, l_max, tm, schedule="serial", m_values=None, win=None):
    """Auto case with S <= 16 shells.  Each iteration takes G = max(16 // S, 2) consecutive (l, m) terms:
    window + inverse-FFT launches over runs of one l, two orders at a time (they share the radial factors),
    fill a buffer of G * S shell fields, and Grams over chunks of 16 // S terms follow (the Gram's MFMA tile is
    16 shells wide; only each term's diagonal S x S block is kept).  The Gram sums stay on the device until
    the end.  Schedules (two buffers, a second stream, events between them, except "serial"):
      "serial":  everything in order on the caller's stream (the default: neither other schedule gains on
                 the headline workload).
      "split":   the windows and z pass of iteration k + 1 run on the caller's stream while the y and x passes
                 and then the Grams of iteration k run on the second stream.  The VALU-bound window kernel
                 and the memory-bound passes do not share CUs well: the passes slow down by about the
                 window kernel's time (46.95 against 47.25 sets/s).
      "overlap": whole windows + transforms of iteration k + 1 beside the Grams of iteration k (both stages
                 are then ALU-bound; measured at ~3% over the serial schedule before later kernel changes)."""
    if schedule not in ("split", "serial", "overlap"):
        raise ValueError(f"unknown schedule {schedule!r}")
    overlap = schedule == "overlap"
    split = schedule == "split"
    L = geo.L
    N = L ** 3
    S = len(r)
    if win is None:
        win = _window_runner(r)
    if split and not isinstance(r, np.ndarray):
        raise ValueError("the split schedule takes thin-shell radii only")
    pack = 16 // S
    G = max(pack, 2)
    lm = [(l, m) for l in range(l_min, l_max + 1) for m in range(l + 1)]
    iters = [lm[i:i + G] for i in range(0, len(lm), G)]
    chunks = [[it[j:j + pack] for j in range(0, len(it), pack)] for it in iters]
    n_blocks = sum(len(c) for c in chunks)
    nbuf = min(2, len(iters)) if (overlap or split) else 1
    bufs = [torch.empty((G * S, L, L, L), dtype=torch.complex64, device=d1.device) for _ in range(nbuf)]
    blocks = torch.zeros((n_blocks, 256), dtype=torch.float64, device=d1.device)
    main = torch.cuda.current_stream(d1.device)
    side = _side_stream(d1.device) if (overlap or split) else main
    side.wait_stream(main)  # d1, H and the buffers are ready on the caller's stream
    gram_done = []
    owners = []
    for k, group in enumerate(iters):
        U = bufs[k % len(bufs)]
        if (overlap or split) and k >= len(bufs):
            main.wait_event(gram_done[k - len(bufs)])  # the buffer's previous Grams are done
        t = tm.start("window_z" if split else "window_ifft")
        pos = 0
        for l, ms in _window_runs(group):
            if split:
                kernels.multipole_window_zpass_multi_f32(H, geo.P, r, l, ms, 1.0 / N, U[pos * S:(pos + len(ms)) * S])
            else:
                win(H, geo.P, l, ms, 1.0 / N, U[pos * S:(pos + len(ms)) * S])
            pos += len(ms)
        tm.stop(t)
        n_fields = pos * S
        grams = []
        pos = 0
        for chunk in chunks[k]:
            grams.append((U[pos * S:(pos + len(chunk)) * S], len(owners)))
            owners.append(chunk)
            pos += len(chunk)
        if split:
            z_done = torch.cuda.Event()
            z_done.record(main)
            with torch.cuda.stream(side):
                side.wait_event(z_done)
                t = tm.start("yx_passes")
                kernels.fft3_inverse_yx_f32(U[:n_fields])
                tm.stop(t)
                t = tm.start("gram")
                for Uc, b in grams:
                    kernels.multipole_gram_f32_dev(d1, Uc, geo.Q, geo.lam, blocks[b])
                tm.stop(t)
                ev = torch.cuda.Event()
                ev.record(side)
                gram_done.append(ev)
            continue
        if not overlap:
            t = tm.start("gram")
            for Uc, b in grams:
                kernels.multipole_gram_f32_dev(d1, Uc, geo.Q, geo.lam, blocks[b])
            tm.stop(t)
            continue
        win_done = torch.cuda.Event()
        win_done.record(main)
        with torch.cuda.stream(side):
            side.wait_event(win_done)
            for Uc, b in grams:
                kernels.multipole_gram_f32_dev(d1, Uc, geo.Q, geo.lam, blocks[b])
            ev = torch.cuda.Event()
            ev.record(side)
            gram_done.append(ev)
    main.wait_stream(side)
    host = blocks.cpu().numpy()
    v_re = {}
    for b, chunk in enumerate(owners):
        full = kernels.gram_block_to_matrix(host[b], len(chunk) * S)
        for j, (l, m) in enumerate(chunk):
            v_re[(l, m)] = 4.0 * np.pi * ((-1.0) ** (l + m)) * full[j * S:(j + 1) * S, j * S:(j + 1) * S] / N
    if m_values is not None:
        m_values.update(v_re)
    result = np.zeros((S, S, l_max - l_min + 1), dtype=np.float64)
    for li, l in enumerate(range(l_min, l_max + 1)):
        result[:, :, li] = combine_m_terms([v_re[(l, m)] for m in range(l + 1)], l)
    return result
