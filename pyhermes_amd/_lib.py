"""ctypes binding of libhermes_amd.so (the HIP/gfx950 C ABI declared in include/hermes_amd.h).

The product has no CPU fallback: if the library is missing or no HIP device is
present, every compute entry point raises RuntimeError.
"""

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhermes_amd.so")

_lib = None
_ctx = {}
_lock = threading.Lock()

c_int, c_i64, c_double, c_void_p = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p

# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGNATURES = {
    "hermes_version": [],
    "hermes_last_error": [],
    "hermes_ctx_create": [c_int, ctypes.POINTER(c_void_p)],
    "hermes_ctx_destroy": [c_void_p],
    "hermes_ctx_set_stream": [c_void_p, c_void_p],
    "hermes_sync": [c_void_p],
    "hermes_release_workspaces": [c_void_p],
    "hermes_project": [c_void_p, c_void_p, c_void_p, c_i64, c_int, c_double, c_void_p, c_int, c_int, c_double,
                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "hermes_fft_r2c": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p],
    "hermes_fft_c2r": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p],
    "hermes_fft_c2c": [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p],
    "hermes_fft3_c2c_f32": [c_void_p, c_int, c_int, c_int, c_void_p],
    "hermes_fft3_tile_choice": [c_void_p, c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)],
    "hermes_fft3_set_tile": [c_void_p, c_int, c_int, c_int, c_int],
    "hermes_fft3_inverse_yx_ymajor_f32": [c_void_p, c_int, c_int, c_void_p, c_void_p],
    "hermes_set_field_layout": [c_void_p, c_int],
    "hermes_get_field_layout": [c_void_p, ctypes.POINTER(c_int)],
    "hermes_profile_gram": [c_void_p, c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_i64)],
    "hermes_capture_begin": [c_void_p],
    "hermes_capture_end": [c_void_p, ctypes.POINTER(c_void_p)],
    "hermes_graph_launch": [c_void_p, c_void_p],
    "hermes_graph_destroy": [c_void_p],
    "hermes_clear_errors": [c_void_p],
    "hermes_read_errors": [c_void_p, ctypes.POINTER(c_int)],
    "hermes_project_reserve": [c_void_p, c_int, c_i64],
    "hermes_delta_f32": [c_void_p, c_i64, c_void_p, c_double, c_double, c_void_p, c_void_p],
    "hermes_half_spectrum_f32": [c_void_p, c_int, c_void_p, c_void_p],
}


class RadialProfile(ctypes.Structure):
    """hermes_radial_profile (include/hermes_amd.h)."""
    _fields_ = [("kind", ctypes.c_int32), ("n_table", ctypes.c_int32), ("a0", ctypes.c_double),
                ("a1", ctypes.c_double), ("k_scale", ctypes.c_double), ("table", ctypes.c_void_p)]


class WindowDesc(ctypes.Structure):
    """hermes_window_desc (include/hermes_amd.h)."""
    _fields_ = [("type", ctypes.c_int32), ("l", ctypes.c_int32), ("m", ctypes.c_int32), ("axis", ctypes.c_int32),
                ("a0", ctypes.c_double), ("a1", ctypes.c_double), ("los", ctypes.c_double * 3)]


_dpp = ctypes.POINTER(ctypes.c_double)
_SIGNATURES.update({
    "hermes_window_build_rfft": [c_void_p, c_int, c_void_p, ctypes.POINTER(WindowDesc), c_int, c_void_p],
    "hermes_window_build_full": [c_void_p, c_int, c_void_p, ctypes.POINTER(WindowDesc), c_int, c_void_p],
    "hermes_spectrum_mul": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_double],
    "hermes_pair_reduce": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, ctypes.POINTER(WindowDesc), c_int,
                           _dpp],
    "hermes_window_build_full_profile": [c_void_p, c_int, c_void_p, ctypes.POINTER(RadialProfile), c_int, c_int,
                                         c_int, c_void_p],
    "hermes_multipole_window_ifft_profiles_f32": [c_void_p, c_int, c_void_p, c_void_p, ctypes.POINTER(RadialProfile),
                                                  c_int, c_int, ctypes.POINTER(c_int), c_int, c_double, c_void_p],
    "hermes_pair_reduce_noalloc": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, ctypes.POINTER(WindowDesc),
                                   c_int, _dpp],
    "hermes_multipole_apply_f32": [c_void_p, c_int, c_void_p, c_void_p, _dpp, c_int, c_int, c_int, c_double,
                                   c_void_p],
    "hermes_multipole_window_ifft_f32": [c_void_p, c_int, c_void_p, c_void_p, _dpp, c_int, c_int, c_int, c_double,
                                   c_void_p],
    "hermes_multipole_window_ifft_multi_f32": [c_void_p, c_int, c_void_p, c_void_p, _dpp, c_int, c_int,
                                               ctypes.POINTER(c_int), c_int, c_double, c_void_p],
    "hermes_multipole_window_zpass_multi_f32": [c_void_p, c_int, c_void_p, c_void_p, _dpp, c_int, c_int,
                                               ctypes.POINTER(c_int), c_int, c_double, c_void_p],
    "hermes_fft3_inverse_yx_f32": [c_void_p, c_int, c_int, c_void_p],
    "hermes_reduce_sum": [c_void_p, c_void_p, c_i64, c_int, c_int, _dpp],
    "hermes_multipole_summand_f64": [c_void_p, c_int, c_void_p, c_void_p, c_void_p, _dpp, _dpp],
    "hermes_multipole_gram_f32": [c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, _dpp, _dpp,
                                  _dpp],
    "hermes_multipole_gram_f32_dev": [c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, _dpp,
                                      _dpp, c_void_p],
    "hermes_multipole_gram_f32_dev_ri": [c_void_p, c_int, c_void_p, c_void_p, c_int, _dpp, _dpp, c_void_p],
    "hermes_multipole_window_ifft_pack0_f32": [c_void_p, c_int, c_void_p, c_void_p, _dpp, c_int, c_int, c_double,
                                               c_void_p],
    "hermes_multipole_window_zpass_pack0_f32": [c_void_p, c_int, c_void_p, c_void_p, _dpp, c_int, c_int, c_double,
                                                c_void_p],
    "hermes_pair_plan_create": [c_void_p, c_int, c_int, ctypes.POINTER(WindowDesc), c_int, ctypes.POINTER(c_void_p)],
    "hermes_pair_plan_destroy": [c_void_p],
    "hermes_pair_reduce_planned": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "hermes_mfma_selftest": [c_void_p, _dpp],
    "hermes_interpolate": [c_void_p, c_int, c_void_p, c_void_p, c_i64, c_void_p, c_int, c_double, c_double, c_double,
                           c_void_p],
    "hermes_mc_triplet_box": [c_void_p, c_int, c_double, c_double, c_double, c_void_p, c_void_p, c_i64, c_int, _dpp,
                              c_void_p, c_void_p, c_void_p, c_int, _dpp],
    "hermes_mc_binned": [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_double,
                         c_double, c_double, c_double, c_i64, ctypes.c_uint64, c_double, c_void_p, c_void_p, c_i64,
                         _dpp, _dpp, ctypes.POINTER(ctypes.c_int64)],
    "hermes_mc_bucketize": [c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, _dpp, c_int, c_void_p, c_void_p,
                            c_void_p, ctypes.POINTER(ctypes.c_int64)],
    "hermes_mlp_nparam": [],
    "hermes_mlp_to_bf16": [c_void_p, c_void_p, c_void_p],
    "hermes_mlp_forward": [c_void_p, c_void_p, c_void_p, c_int, c_void_p],
    "hermes_mlp_grad": [c_void_p, c_void_p, c_void_p, c_void_p, c_int, ctypes.c_float, c_void_p],
    "hermes_mlp_adam": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_float, ctypes.c_float,
                        ctypes.c_float, ctypes.c_float, c_int],
    "hermes_mlp_train_steps": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_i64,
                               c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_float), c_int, c_int, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float],
})


class MCProposal(ctypes.Structure):
    """hermes_mc_proposal (include/hermes_amd.h)."""
    _fields_ = [("cell_bucket", c_void_p), ("bucket_cells", c_void_p), ("bucket_off", c_void_p),
                ("bucket_cdf", c_void_p), ("bucket_density", c_void_p), ("bin_cdf", c_void_p),
                ("bin_pdf", c_void_p), ("nbucket", ctypes.c_int32), ("sub", ctypes.c_int32),
                ("alpha", c_double), ("anchors", ctypes.c_int32), ("reserved", ctypes.c_int32)]
_RESTYPES = {"hermes_last_error": ctypes.c_char_p}


def declared_symbols():
    return list(_SIGNATURES)


def load():
    """Load the shared library (raises RuntimeError if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"pyhermes_amd: HIP library not found at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (make -C pyhermes_amd/csrc).")
    lib = ctypes.CDLL(LIB_PATH)
    for name, argtypes in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = _RESTYPES.get(name, c_int)
    _lib = lib
    return lib


def last_error():
    return load().hermes_last_error().decode("utf-8", "replace")


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed (status {rc}): {last_error()}")


def call(name, *args):
    fn = getattr(load(), name)
    check(fn(*args), name)


def context(device=None):
    """Per-device library context bound to torch's current HIP stream."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("pyhermes_amd requires a HIP device (torch.cuda.is_available() is False)")
    if device is None:
        device = torch.cuda.current_device()
    with _lock:
        ctx = _ctx.get(device)
        if ctx is None:
            handle = c_void_p()
            call("hermes_ctx_create", int(device), ctypes.byref(handle))
            ctx = handle
            _ctx[device] = ctx
    stream = torch.cuda.current_stream(device).cuda_stream
    call("hermes_ctx_set_stream", ctx, c_void_p(stream))
    return ctx


def ptr(t):
    """Device pointer of a torch tensor (or None -> NULL)."""
    if t is None:
        return None
    return c_void_p(t.data_ptr())
