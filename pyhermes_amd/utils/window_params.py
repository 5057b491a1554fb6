"""Window-parameter normalisation (same rules as pyhermes.utils.window_params)."""

import copy

import numpy as np

ANISOTROPIC_AUTO_WINDOW_TYPES = {"ring", "disk", "cylinder", "cylshell"}
COMPLEX_RFFT_WINDOW_TYPES = {"directional_derivative"}
COMPLEX_FULL_FFT_WINDOW_TYPES = {"legendre_multipole", "radial_multipole"}
VALID_KERNEL_MODES = {"auto", "octant", "full_rfft", "complex_rfft", "complex_full_fft"}
LOS_ARG_KEYS = ("nx", "ny", "nz")
DEFAULT_LOS_ARGS = {"nx": 0.0, "ny": 0.0, "nz": 1.0}
LOS_AWARE_WINDOW_TYPES = ANISOTROPIC_AUTO_WINDOW_TYPES | COMPLEX_RFFT_WINDOW_TYPES
BUILTIN_BINNING_WINDOW_TYPES = {"shell", "thick_shell"} | ANISOTROPIC_AUTO_WINDOW_TYPES


def default_binning_window():
    return {"type": "shell", "len_args": {"R": None}, "los_args": {}, "other_args": {}, "mapping": "s_to_R",
            "kernel_mode": "octant"}


def normalize_len_args(len_args):
    if len_args is None:
        return {}
    if isinstance(len_args, dict):
        return copy.deepcopy(len_args)
    if isinstance(len_args, str):
        return {len_args: None}
    if isinstance(len_args, (list, tuple)):
        out = {}
        for item in len_args:
            if not isinstance(item, str):
                raise TypeError("len_args entries must be strings.")
            out[item] = None
        return out
    raise TypeError("len_args must be a dict, string, list, tuple, or None.")


def merge_len_arg_defaults(len_args, names):
    out = normalize_len_args(len_args)
    for name in names:
        out.setdefault(name, None)
    return out


def normalize_los_args(los_args, window_type=None):
    if los_args is None:
        los_args = {}
    if isinstance(los_args, (list, tuple, np.ndarray)):
        arr = np.asarray(los_args, dtype=np.float64)
        if arr.shape != (3,):
            raise ValueError("los_args array must contain exactly three values: [nx, ny, nz].")
        los_args = {k: float(v) for k, v in zip(LOS_ARG_KEYS, arr)}
    elif isinstance(los_args, dict):
        los_args = copy.deepcopy(los_args)
    else:
        raise TypeError("los_args must be a dict, length-3 array, or None.")
    if not los_args and window_type in LOS_AWARE_WINDOW_TYPES:
        los_args = copy.deepcopy(DEFAULT_LOS_ARGS)
    if los_args:
        if not all(k in los_args for k in LOS_ARG_KEYS):
            raise ValueError("los_args must define nx, ny, and nz together.")
        vec = np.array([los_args[k] for k in LOS_ARG_KEYS], dtype=np.float64)
        if np.linalg.norm(vec) == 0.0:
            raise ValueError("los_args vector must be non-zero.")
        los_args = {k: float(los_args[k]) for k in LOS_ARG_KEYS}
    return los_args


def default_kernel_mode(window_type, has_custom_func=False):
    if has_custom_func:
        return "full_rfft"
    if window_type in COMPLEX_FULL_FFT_WINDOW_TYPES:
        return "complex_full_fft"
    if window_type in COMPLEX_RFFT_WINDOW_TYPES:
        return "complex_rfft"
    if window_type in ANISOTROPIC_AUTO_WINDOW_TYPES:
        return "auto"
    return "octant"


def normalize_kernel_mode(kernel_mode):
    kernel_mode = str(kernel_mode).strip().lower()
    if kernel_mode not in VALID_KERNEL_MODES:
        raise ValueError(f"Unsupported kernel_mode '{kernel_mode}'. Supported values are {sorted(VALID_KERNEL_MODES)}.")
    return kernel_mode


def callable_metadata(value):
    target = getattr(value, "py_func", value)
    module = getattr(target, "__module__", type(value).__module__)
    name = getattr(target, "__qualname__", getattr(target, "__name__", type(value).__name__))
    return {"kind": "callable", "name": f"{module}.{name}"}


def serialize_window_params(value):
    if isinstance(value, dict):
        return {("custom_func" if k == "func" else k): (callable_metadata(v) if k == "func" else
                                                           serialize_window_params(v)) for k, v in value.items()}
    if callable(value):
        return callable_metadata(value)
    if isinstance(value, np.ndarray):
        return value.tolist()
    if isinstance(value, (list, tuple)):
        return [serialize_window_params(v) for v in value]
    if isinstance(value, np.generic):
        return value.item()
    return copy.deepcopy(value)


def apply_builtin_binning_window_defaults(binning_window):
    params = copy.deepcopy(binning_window)
    wtype = params.get("type")
    if not wtype:
        return params
    wtype = str(wtype).strip().lower()
    params["type"] = wtype
    if params.get("func") is not None:
        return params
    if wtype == "shell":
        params["len_args"] = merge_len_arg_defaults(params.get("len_args", {}), ("R",))
        params.setdefault("mapping", "s_to_R")
    elif wtype == "thick_shell":
        params["len_args"] = merge_len_arg_defaults(params.get("len_args", {}), ("R", "delta_R"))
        params.setdefault("mapping", "s_to_R")
    elif wtype in LOS_AWARE_WINDOW_TYPES:
        params["len_args"] = merge_len_arg_defaults(params.get("len_args", {}), ("R", "H"))
        params.setdefault("los_args", copy.deepcopy(DEFAULT_LOS_ARGS))
        params.setdefault("mapping", "smu_to_RH")
    params.setdefault("los_args", {})
    params.setdefault("other_args", {})
    return params


def binning_window_from_string(binning_window):
    wtype = binning_window.strip().lower()
    if wtype not in BUILTIN_BINNING_WINDOW_TYPES:
        raise ValueError(f"Unsupported binning_window string '{binning_window}'. Supported built-in strings are "
                         "'shell', 'thick_shell', 'ring', 'disk', 'cylinder', and 'cylshell'.")
    return apply_builtin_binning_window_defaults({"type": wtype})


def normalize_binning_window_template(binning_window):
    if binning_window is None:
        binning_window = default_binning_window()
    elif isinstance(binning_window, str):
        binning_window = binning_window_from_string(binning_window)
    elif not isinstance(binning_window, dict):
        raise TypeError(f"Unsupported binning_window input: expected dict, string, or None, got {type(binning_window)}.")
    out = apply_builtin_binning_window_defaults(binning_window)
    if not out.get("type"):
        out["type"] = "custom" if out.get("func") is not None else "shell"
    has_func = out.get("func") is not None
    if has_func and not callable(out["func"]):
        raise TypeError("Custom binning-window parameter 'func' must be callable.")
    out["len_args"] = normalize_len_args(out.get("len_args", {}))
    out["los_args"] = normalize_los_args(out.get("los_args", {}), None if has_func else out.get("type"))
    out.setdefault("other_args", {})
    mode = out.get("kernel_mode") or default_kernel_mode(out.get("type"), has_custom_func=has_func)
    out["kernel_mode"] = normalize_kernel_mode(mode)
    return out
