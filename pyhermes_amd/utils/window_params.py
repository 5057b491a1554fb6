"""Window-parameter schema.

A window is described by a dict with ``type`` (or a callable ``func``), ``len_args`` (named lengths),
``los_args`` (line of sight {nx, ny, nz}), ``other_args``, ``kernel_mode`` (the layout its kernel is built
in) and, for binning windows, ``mapping`` (how a sample fills the lengths).  The keys, defaults, error
types and messages are the reference's (pyhermes/utils/window_params.py), so user configurations load
unchanged.  Here every per-type property lives in one table, ``_RULES``; the coercers and the single
template normaliser read it.
"""

import copy
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np


@dataclass(frozen=True)
class _TypeRule:
    lengths: Tuple[str, ...] = ()   # len_args a binning template gets (as None) when it does not set them
    mapping: Optional[str] = None   # sampling -> lengths rule of a binning template
    layout: str = "octant"          # default kernel_mode
    sight_line: bool = False        # takes los_args (defaulting to the z axis)
    binning: bool = False           # may be named by a bare string as a built-in binning window


_CYL = _TypeRule(lengths=("R", "H"), mapping="smu_to_RH", layout="auto", sight_line=True, binning=True)
_RULES = {
    "shell": _TypeRule(lengths=("R",), mapping="s_to_R", binning=True),
    "thick_shell": _TypeRule(lengths=("R", "delta_R"), mapping="s_to_R", binning=True),
    "ring": _CYL, "disk": _CYL, "cylinder": _CYL, "cylshell": _CYL,
    "directional_derivative": _TypeRule(lengths=("R", "H"), mapping="smu_to_RH", layout="complex_rfft",
                                        sight_line=True),
    "legendre_multipole": _TypeRule(layout="complex_full_fft"),
    "radial_multipole": _TypeRule(layout="complex_full_fft"),
}
_PLAIN = _TypeRule()


def _rule(window_type):
    return _RULES.get(window_type, _PLAIN)


def _types_where(pred):
    return {t for t, r in _RULES.items() if pred(r)}


ANISOTROPIC_AUTO_WINDOW_TYPES = _types_where(lambda r: r.layout == "auto")
COMPLEX_RFFT_WINDOW_TYPES = _types_where(lambda r: r.layout == "complex_rfft")
COMPLEX_FULL_FFT_WINDOW_TYPES = _types_where(lambda r: r.layout == "complex_full_fft")
LOS_AWARE_WINDOW_TYPES = _types_where(lambda r: r.sight_line)
BUILTIN_BINNING_WINDOW_TYPES = _types_where(lambda r: r.binning)
VALID_KERNEL_MODES = ("auto", "complex_full_fft", "complex_rfft", "full_rfft", "octant")
CUSTOM_FUNC_KERNEL_MODE = "full_rfft"
LOS_ARG_KEYS = ("nx", "ny", "nz")
Z_SIGHT_LINE = (0.0, 0.0, 1.0)


def default_binning_window():
    """The binning window of a task that names none: a thin shell whose radius is the sample's s."""
    return {"type": "shell", "len_args": {"R": None}, "los_args": {}, "other_args": {}, "mapping": "s_to_R",
            "kernel_mode": "octant"}


# ---------------------------------------------------------------- coercers
def normalize_len_args(len_args):
    """len_args as a dict of name -> value: a dict is copied, a name or a list of names become unset (None)
    entries."""
    if len_args is None:
        return {}
    if isinstance(len_args, dict):
        return copy.deepcopy(len_args)
    names = [len_args] if isinstance(len_args, str) else len_args
    if not isinstance(names, (list, tuple)):
        raise TypeError("len_args must be a dict, string, list, tuple, or None.")
    if not all(isinstance(n, str) for n in names):
        raise TypeError("len_args entries must be strings.")
    return dict.fromkeys(names)


def normalize_los_args(los_args, window_type=None):
    """los_args as {nx, ny, nz} floats, or {} for a window without a line of sight.  A length-3 array is
    accepted; a line-of-sight window given none looks along z."""
    if los_args is None:
        los_args = {}
    if isinstance(los_args, (list, tuple, np.ndarray)):
        vec = np.asarray(los_args, dtype=np.float64)
        if vec.shape != (3,):
            raise ValueError("los_args array must contain exactly three values: [nx, ny, nz].")
        los_args = dict(zip(LOS_ARG_KEYS, vec.tolist()))
    elif not isinstance(los_args, dict):
        raise TypeError("los_args must be a dict, length-3 array, or None.")
    if not los_args:
        return dict(zip(LOS_ARG_KEYS, Z_SIGHT_LINE)) if window_type in LOS_AWARE_WINDOW_TYPES else {}
    if any(k not in los_args for k in LOS_ARG_KEYS):
        raise ValueError("los_args must define nx, ny, and nz together.")
    out = {k: float(los_args[k]) for k in LOS_ARG_KEYS}
    if not any(out.values()):
        raise ValueError("los_args vector must be non-zero.")
    return out


def default_kernel_mode(window_type, has_custom_func=False):
    return CUSTOM_FUNC_KERNEL_MODE if has_custom_func else _rule(window_type).layout


def normalize_kernel_mode(kernel_mode):
    mode = str(kernel_mode).strip().lower()
    if mode not in VALID_KERNEL_MODES:
        raise ValueError(f"Unsupported kernel_mode '{mode}'. Supported values are {list(VALID_KERNEL_MODES)}.")
    return mode


def resolve_kernel_mode(requested, window_type, has_custom_func=False):
    """The requested kernel_mode, or the type's default when none is requested; validated."""
    return normalize_kernel_mode(requested or default_kernel_mode(window_type, has_custom_func))


# ---------------------------------------------------------------- serialisation
def callable_metadata(value):
    """How a callable is recorded in saved parameters: its module-qualified name (a JIT wrapper's Python
    function when it has one)."""
    target = getattr(value, "py_func", value)
    module = getattr(target, "__module__", None) or getattr(value, "__module__", None) or type(value).__module__
    name = getattr(target, "__qualname__", None) or getattr(target, "__name__", None) or type(value).__name__
    return {"kind": "callable", "name": f"{module}.{name}"}


def serialize_window_params(value):
    """A JSON/pickle-safe copy of window parameters: callables become their metadata (a "func" entry is
    stored as "custom_func"), arrays and numpy scalars become Python values."""
    if isinstance(value, dict):
        out = {}
        for key, item in value.items():
            if key == "func":
                out["custom_func"] = callable_metadata(item)
            else:
                out[key] = serialize_window_params(item)
        return out
    if callable(value):
        return callable_metadata(value)
    if isinstance(value, np.ndarray):
        return value.tolist()
    if isinstance(value, (list, tuple)):
        return [serialize_window_params(v) for v in value]
    if isinstance(value, np.generic):
        return value.item()
    return copy.deepcopy(value)


# ---------------------------------------------------------------- binning-window templates
def _with_type_defaults(params):
    """A named built-in type without a func gets its rule's unset lengths, line of sight, other_args and
    mapping; anything else is returned with its type name cleaned."""
    wtype = params.get("type")
    if not wtype:
        return params
    wtype = str(wtype).strip().lower()
    params["type"] = wtype
    rule = _RULES.get(wtype)
    if params.get("func") is not None or rule is None or not (rule.binning or rule.sight_line):
        return params
    lengths = normalize_len_args(params.get("len_args", {}))
    for name in rule.lengths:
        lengths.setdefault(name, None)
    params["len_args"] = lengths
    params.setdefault("los_args", dict(zip(LOS_ARG_KEYS, Z_SIGHT_LINE)) if rule.sight_line else {})
    params.setdefault("other_args", {})
    if rule.mapping is not None:
        params.setdefault("mapping", rule.mapping)
    return params


def normalize_binning_window_template(binning_window):
    """The binning-window template of a 2PCF task: None (a thin shell), a built-in type name, or a dict
    with a type or a callable func; returns a dict with every schema key set."""
    if binning_window is None:
        params = default_binning_window()
    elif isinstance(binning_window, str):
        name = binning_window.strip().lower()
        if name not in BUILTIN_BINNING_WINDOW_TYPES:
            raise ValueError(f"Unsupported binning_window string '{binning_window}'. Supported built-in strings are "
                             "'shell', 'thick_shell', 'ring', 'disk', 'cylinder', and 'cylshell'.")
        params = {"type": name}
    elif isinstance(binning_window, dict):
        params = copy.deepcopy(binning_window)
    else:
        raise TypeError(f"Unsupported binning_window input: expected dict, string, or None, got {type(binning_window)}.")
    params = _with_type_defaults(params)
    custom = params.get("func") is not None
    if not params.get("type"):
        params["type"] = "custom" if custom else "shell"
    if custom and not callable(params["func"]):
        raise TypeError("Custom binning-window parameter 'func' must be callable.")
    params["len_args"] = normalize_len_args(params.get("len_args", {}))
    params["los_args"] = normalize_los_args(params.get("los_args", {}), None if custom else params["type"])
    params.setdefault("other_args", {})
    params["kernel_mode"] = resolve_kernel_mode(params.get("kernel_mode"), params["type"], custom)
    return params
