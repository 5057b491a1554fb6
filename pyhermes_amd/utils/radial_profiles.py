"""Radial profiles of multipole windows and their Hankel tables U_l(k).

An edge window of the SH-3PCF with a radial profile w(r) has the transfer function U_l(|k|) Y_lm(k-hat) with

    U_l(k) = 4 pi int_0^inf r^2 w(r) j_l(2 pi k r) dr            (k in cycles per unit length)

(radial_profiles.py / radial_multipole_windows.py of PyHermes).  Thin shells (w a delta at R) have the
closed form j_l(2 pi k R) and need no table; the other profiles are tabulated on a uniform k grid over
[0, k_max] by Gauss-Legendre quadrature in r, and the device kernels interpolate the table with a local
cubic.  Profiles given in k space (the Gaussian-smoothed shell, user k-space profiles) are first taken to
r space by the inverse transform w(r) = 4 pi int k^2 T(k) j_0(2 pi k r) dk on Gauss-Legendre nodes in k.

The node and table counts, the support and normalisation of each built-in profile, the user-profile
dictionaries and the validation messages are the reference's (radial_profiles.py:28-36 defaults,
:39-104 profile constructors, :115-186 checks, :189-260 built-in profiles), so the same request yields the
same table.  The quadrature itself is the textbook Gauss-Legendre rule (numpy.polynomial.legendre.leggauss)
and scipy.special.spherical_jn, the library routines the reference also calls.
"""

import math
from collections import OrderedDict

import numpy as np
from numpy.polynomial.legendre import leggauss
from scipy.special import spherical_jn

BUILTIN_RADIAL_PROFILE_TYPES = {"shell", "sphere", "thick_shell", "gaussian", "gaussian_shell"}
CUSTOM_RADIAL_PROFILE_TYPES = {"custom_real", "custom_kspace"}
SUPPORTED_RADIAL_PROFILE_TYPES = BUILTIN_RADIAL_PROFILE_TYPES | CUSTOM_RADIAL_PROFILE_TYPES

# minimum node counts and nodes per cycle of the largest phase k_max * extent
R_NODES_MIN, R_NODES_PER_CYCLE = 1024, 32
TABLE_MIN, TABLE_PER_CYCLE = 4097, 64
DIAGNOSTIC_PROBES = 33
_OVERRIDE_KEYS = ("r_min", "r_max", "inverse_k_max", "normalization", "allow_signed", "quadrature_points",
                  "inverse_points", "table_points")
_NUMERIC_OVERRIDES = ("quadrature_points", "inverse_points", "table_points")


# ---------------------------------------------------------------- user profile descriptions
def real_space_profile(func, r_max, *, r_min=0.0, normalization="unit_integral", allow_signed=False,
                       quadrature_points=None, table_points=None, args=None):
    """A user real-space profile: func(r, **len_args, **args) -> w(r) on [r_min, r_max]; by default
    normalised to 4 pi int r^2 w dr = 1."""
    if not callable(func):
        raise TypeError("real_space_profile requires a callable func(r, **len_args).")
    return {"kind": "real_callable", "func": func, "r_min": float(r_min), "r_max": float(r_max),
            "normalization": str(normalization), "allow_signed": bool(allow_signed),
            "quadrature_points": quadrature_points, "table_points": table_points,
            "args": {} if args is None else dict(args)}


def tabulated_real_space_profile(r, w, *, normalization="unit_integral", allow_signed=False,
                                 quadrature_points=None, table_points=None):
    """A user real-space profile given as samples w(r) (linearly interpolated between them)."""
    r = np.asarray(r, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    if r.ndim != 1 or w.ndim != 1 or r.size != w.size or r.size < 2:
        raise ValueError("tabulated_real_space_profile requires matching one-dimensional r and w arrays.")
    return {"kind": "real_tabulated", "r": r, "w": w, "normalization": str(normalization),
            "allow_signed": bool(allow_signed), "quadrature_points": quadrature_points,
            "table_points": table_points}


def kspace_profile(func, r_max, inverse_k_max, *, r_min=0.0, normalization="unit_integral", allow_signed=True,
                   inverse_points=None, quadrature_points=None, table_points=None, args=None):
    """A user isotropic k-space transfer T(k) = func(k, **len_args, **args), taken to r space on
    [r_min, r_max] by the inverse transform over [0, inverse_k_max]."""
    if not callable(func):
        raise TypeError("kspace_profile requires a callable func(k, **len_args).")
    return {"kind": "kspace_callable", "func": func, "r_min": float(r_min), "r_max": float(r_max),
            "inverse_k_max": float(inverse_k_max), "normalization": str(normalization),
            "allow_signed": bool(allow_signed), "inverse_points": inverse_points,
            "quadrature_points": quadrature_points, "table_points": table_points,
            "args": {} if args is None else dict(args)}


# ---------------------------------------------------------------- validation
def _length(len_args, name, radial_type):
    if name not in len_args:
        raise ValueError(f"radial profile '{radial_type}' requires len_args['{name}'].")
    value = float(len_args[name])
    if not np.isfinite(value):
        raise ValueError(f"radial profile '{radial_type}' requires finite len_args['{name}'].")
    return value


def _user_profile(profile_config):
    """The profile dict of a custom type: other_args['profile'] (or other_args itself), with the override
    keys of other_args applied on top."""
    if profile_config is None:
        raise ValueError("Custom radial profiles require other_args['profile'].")
    if not isinstance(profile_config, dict):
        raise TypeError("Custom radial profile configuration must be a dictionary.")
    base = profile_config.get("profile", profile_config)
    if not isinstance(base, dict):
        raise TypeError("other_args['profile'] must be a dictionary created by real_space_profile or kspace_profile.")
    merged = dict(base)
    merged.update({k: profile_config[k] for k in _OVERRIDE_KEYS if k in profile_config})
    return merged


def _check_support(profile):
    kind = str(profile.get("kind", "")).strip().lower()
    if kind == "real_tabulated":
        r = np.asarray(profile.get("r"), dtype=np.float64)
        w = np.asarray(profile.get("w"), dtype=np.float64)
        if r.ndim != 1 or w.ndim != 1 or r.size != w.size or r.size < 2:
            raise ValueError("real_tabulated profiles require matching one-dimensional r and w arrays.")
        if not (np.all(np.isfinite(r)) and np.all(np.isfinite(w))) or np.any(np.diff(r) <= 0.0):
            raise ValueError("real_tabulated profile radii must be finite and strictly increasing.")
        return
    if kind not in ("real_callable", "kspace_callable"):
        raise ValueError(f"Unsupported custom radial profile kind '{kind}'.")
    if not callable(profile.get("func")):
        raise TypeError(f"Profile kind '{kind}' requires a callable func.")
    if "r_min" not in profile or "r_max" not in profile:
        raise ValueError(f"Profile kind '{kind}' requires finite r_min and r_max.")
    lo, hi = float(profile["r_min"]), float(profile["r_max"])
    if not (np.isfinite(lo) and np.isfinite(hi)) or lo < 0.0 or hi <= lo:
        raise ValueError("Custom radial profile support must satisfy 0 <= r_min < r_max.")
    if kind == "kspace_callable":
        kmax = float(profile.get("inverse_k_max", np.nan))
        if not np.isfinite(kmax) or kmax <= 0.0:
            raise ValueError("kspace_callable profiles require a finite positive inverse_k_max.")


def validate_radial_profile_request(radial_type, len_args, profile_config=None):
    """Checks one profile request without building anything (radial_profiles.py:115-141 rules)."""
    radial_type = str(radial_type).strip().lower()
    if radial_type not in SUPPORTED_RADIAL_PROFILE_TYPES:
        raise ValueError(f"Unsupported radial profile '{radial_type}'. Supported profiles: "
                         f"{', '.join(sorted(SUPPORTED_RADIAL_PROFILE_TYPES))}.")
    if radial_type in ("shell", "sphere", "gaussian", "thick_shell"):
        if _length(len_args, "R", radial_type) < 0.0:
            raise ValueError(f"radial profile '{radial_type}' requires a non-negative R.")
    if radial_type == "thick_shell" and _length(len_args, "delta_R", radial_type) <= 0.0:
        raise ValueError("radial profile 'thick_shell' requires delta_R > 0.")
    if radial_type == "gaussian_shell":
        if min(_length(len_args, "R_shell", radial_type), _length(len_args, "R_smooth", radial_type)) < 0.0:
            raise ValueError("radial profile 'gaussian_shell' requires non-negative R_shell and R_smooth.")
    if radial_type in CUSTOM_RADIAL_PROFILE_TYPES:
        profile = _user_profile(profile_config)
        space = "real" if radial_type == "custom_real" else "kspace"
        if not str(profile.get("kind", "")).startswith(space):
            raise ValueError(f"radial profile '{radial_type}' requires a {space}-space profile specification.")
        _check_support(profile)


def is_thin_shell(radial_type, len_args):
    """Thin shells -- and Gaussian shells without smoothing -- have the closed-form j_l transfer."""
    t = str(radial_type).strip().lower()
    return t == "shell" or (t == "gaussian_shell" and float(len_args["R_smooth"]) == 0.0)


def has_analytic_monopole(radial_type):
    """Built-in profiles whose l = 0 transfer is evaluated in closed form (radial_multipole_windows.py:
    286-294) rather than from a table."""
    return str(radial_type).strip().lower() in BUILTIN_RADIAL_PROFILE_TYPES


# ---------------------------------------------------------------- quadrature
def _count(extent_cycles, minimum, per_cycle, configured, label):
    if configured is not None:
        n = int(configured)
        if n < 8:
            raise ValueError(f"{label} must be at least 8.")
        return n
    return max(int(minimum), int(math.ceil(per_cycle * max(float(extent_cycles), 1.0))) + 1)


_GL = {}


def _gauss_legendre(lo, hi, n):
    n = int(n)
    if n < 8:
        raise ValueError("Radial Hankel quadrature requires at least 8 points.")
    if n not in _GL:
        _GL[n] = leggauss(n)
    x, w = _GL[n]
    half = 0.5 * (hi - lo)
    return half * (x + 1.0) + lo, half * w


def _call_profile(func, x, len_args, extra):
    kwargs = {**dict(extra), **dict(len_args)}
    try:
        v = np.asarray(func(x, **kwargs), dtype=np.float64)
    except Exception as exc:
        raise ValueError("Custom radial profile callable failed for the requested sampled parameters.") from exc
    if v.ndim == 0:
        v = np.full(x.shape, float(v))
    if v.shape != x.shape or not np.all(np.isfinite(v)):
        raise ValueError("Custom radial profile must return finite values with the same shape as its input.")
    return v


def monopole_transfer(radial_type, k, len_args):
    """Closed-form l = 0 transfer of a built-in profile, k in cycles per unit length (the isotropic window
    functions of window_functions.py:29-129)."""
    t = str(radial_type).strip().lower()
    k = np.asarray(k, dtype=np.float64)

    def sphere(radius):
        q = 2.0 * np.pi * k * radius
        out = np.ones_like(q)
        nz = q != 0.0
        out[nz] = 3.0 * (np.sin(q[nz]) - q[nz] * np.cos(q[nz])) / q[nz] ** 3
        return out

    if t == "shell":
        return spherical_jn(0, 2.0 * np.pi * k * _length(len_args, "R", t))
    if t == "sphere":
        return sphere(_length(len_args, "R", t))
    if t == "gaussian":
        q = 2.0 * np.pi * k * _length(len_args, "R", t)
        return np.exp(-0.5 * q * q)
    if t == "thick_shell":
        R, dR = _length(len_args, "R", t), _length(len_args, "delta_R", t)
        lo, hi = max(0.0, R - 0.5 * dR), R + 0.5 * dR
        den = hi ** 3 - lo ** 3
        return np.ones_like(k) if den == 0.0 else (hi ** 3 * sphere(hi) - lo ** 3 * sphere(lo)) / den
    if t == "gaussian_shell":
        Rs, Rg = _length(len_args, "R_shell", t), _length(len_args, "R_smooth", t)
        qs, qg = 2.0 * np.pi * k * Rs, 2.0 * np.pi * k * Rg
        den = Rs * Rs + Rg * Rg
        if den == 0.0:
            return np.ones_like(k)
        sinc = np.ones_like(qs)
        nz = qs != 0.0
        sinc[nz] = np.sin(qs[nz]) / qs[nz]
        return (Rg * Rg * np.cos(qs) + Rs * Rs * sinc) / den * np.exp(-0.5 * qg * qg)
    raise ValueError(f"No built-in k-space monopole is available for '{radial_type}'.")


def _with_overrides(options, profile_config):
    out = dict(options)
    if profile_config is None:
        return out
    if not isinstance(profile_config, dict):
        raise TypeError("Radial profile configuration must be a dictionary.")
    out.update({k: profile_config[k] for k in _NUMERIC_OVERRIDES if k in profile_config})
    return out


def _builtin_real_space(radial_type, len_args):
    """Support, density and normalisation rule of the built-in real-space profiles (radial_profiles.py:
    189-226)."""
    t = radial_type
    R = _length(len_args, "R", t)
    if t == "sphere":
        if R <= 0.0:
            raise ValueError("sphere radial multipoles require R > 0.")
        return 0.0, R, lambda r: np.where(r <= R, 3.0 / (4.0 * np.pi * R ** 3), 0.0), "provided"
    if t == "thick_shell":
        dR = _length(len_args, "delta_R", t)
        if dR <= 0.0:
            raise ValueError("thick_shell radial multipoles require delta_R > 0.")
        lo, hi = max(0.0, R - 0.5 * dR), R + 0.5 * dR
        vol = 4.0 * np.pi * (hi ** 3 - lo ** 3) / 3.0
        return lo, hi, lambda r: np.where((r >= lo) & (r <= hi), 1.0 / vol, 0.0), "provided"
    if R <= 0.0:
        raise ValueError("gaussian radial multipoles require R > 0.")
    c = (2.0 * np.pi * R * R) ** (-1.5)
    return 0.0, 8.0 * R, lambda r: c * np.exp(-0.5 * (r / R) ** 2), "unit_integral"


class RadialNodes:
    """Gauss-Legendre nodes r_i, weights dr_i and normalised profile values w(r_i) of one physical profile,
    resolved for transfers up to k_max."""

    def __init__(self, radial_type, len_args, k_max, profile_config=None):
        t = str(radial_type).strip().lower()
        self.radial_type = t
        if t in ("sphere", "thick_shell", "gaussian"):
            lo, hi, density, norm = _builtin_real_space(t, len_args)
            opts = _with_overrides({"normalization": norm, "allow_signed": False}, profile_config)
            n = _count(k_max * (hi - lo), R_NODES_MIN, R_NODES_PER_CYCLE, opts.get("quadrature_points"),
                       "quadrature_points")
            r, dr = _gauss_legendre(lo, hi, n)
            w = density(r)
        elif t == "custom_real":
            opts = _user_profile(profile_config)
            _check_support(opts)
            kind = str(opts["kind"]).strip().lower()
            if kind == "real_tabulated":
                tr, tw = np.asarray(opts["r"], np.float64), np.asarray(opts["w"], np.float64)
                lo, hi = float(tr[0]), float(tr[-1])
                n = _count(k_max * (hi - lo), max(R_NODES_MIN, tr.size), R_NODES_PER_CYCLE,
                           opts.get("quadrature_points"), "quadrature_points")
                r, dr = _gauss_legendre(lo, hi, n)
                w = np.interp(r, tr, tw)
            elif kind == "real_callable":
                lo, hi = float(opts["r_min"]), float(opts["r_max"])
                n = _count(k_max * (hi - lo), R_NODES_MIN, R_NODES_PER_CYCLE, opts.get("quadrature_points"),
                           "quadrature_points")
                r, dr = _gauss_legendre(lo, hi, n)
                w = _call_profile(opts["func"], r, len_args, opts.get("args", {}))
            else:
                raise ValueError(f"Expected a real-space profile, got kind '{kind}'.")
        elif t in ("gaussian_shell", "custom_kspace"):
            r, dr, w, opts = self._from_kspace(t, len_args, k_max, profile_config)
        else:
            raise ValueError(f"Unsupported tabulated radial profile '{radial_type}'.")
        self.r, self.dr = r, dr
        self.options = opts
        self.w = self._normalised(np.asarray(w, dtype=np.float64), opts)

    @staticmethod
    def _from_kspace(t, len_args, k_max, profile_config):
        """w(r) = 4 pi int_0^K k^2 T(k) j_0(2 pi k r) dk (radial_profiles.py:359-415)."""
        if t == "gaussian_shell":
            Rs, Rg = _length(len_args, "R_shell", t), _length(len_args, "R_smooth", t)
            if Rg == 0.0:
                raise ValueError("gaussian_shell with R_smooth=0 must use the exact shell path.")
            opts = _with_overrides({"kind": "builtin_inverse", "r_min": 0.0, "r_max": Rs + 8.0 * Rg,
                                    "inverse_k_max": max(k_max, 8.0 / (2.0 * np.pi * Rg)),
                                    "normalization": "unit_integral", "allow_signed": True}, profile_config)

            def transfer(k):
                return monopole_transfer("gaussian_shell", k, len_args)
        else:
            opts = _user_profile(profile_config)
            _check_support(opts)
            if str(opts["kind"]).strip().lower() != "kspace_callable":
                raise ValueError("custom_kspace requires a profile created by kspace_profile.")

            def transfer(k):
                return _call_profile(opts["func"], k, len_args, opts.get("args", {}))
        lo, hi, K = float(opts["r_min"]), float(opts["r_max"]), float(opts["inverse_k_max"])
        n_r = _count(k_max * (hi - lo), R_NODES_MIN, R_NODES_PER_CYCLE, opts.get("quadrature_points"),
                     "quadrature_points")
        r, dr = _gauss_legendre(lo, hi, n_r)
        n_k = _count(K * hi, R_NODES_MIN, R_NODES_PER_CYCLE, opts.get("inverse_points"), "inverse_points")
        k, dk = _gauss_legendre(0.0, K, n_k)
        coef = 4.0 * np.pi * dk * k * k * transfer(k)
        w = np.concatenate([spherical_jn(0, 2.0 * np.pi * np.outer(r[i:i + 64], k)) @ coef
                            for i in range(0, r.size, 64)])
        return r, dr, w, opts

    def _normalised(self, w, opts):
        if not np.all(np.isfinite(w)):
            raise ValueError("Radial profile evaluation produced non-finite values.")
        if not bool(opts.get("allow_signed", False)) and np.any(w < 0.0):
            raise ValueError("Real-space radial profile has negative values; set allow_signed=True for compensated "
                             "filters.")
        rule = str(opts.get("normalization", "unit_integral")).strip().lower()
        if rule == "provided":
            return w
        if rule != "unit_integral":
            raise ValueError("Radial profile normalization must be 'unit_integral' or 'provided'.")
        vol = self.volume_integral(w)
        if not np.isfinite(vol) or np.isclose(vol, 0.0):
            raise ValueError("Radial profile has a non-finite or zero volume integral.")
        return w / vol

    def volume_integral(self, w=None):
        """4 pi int r^2 w dr."""
        return float(4.0 * np.pi * np.sum(self.dr * self.r ** 2 * (self.w if w is None else w)))

    def transfer(self, l, k):
        """U_l(k) = sum_i 4 pi dr_i r_i^2 w(r_i) j_l(2 pi k r_i) at the given k values."""
        k = np.asarray(k, dtype=np.float64)
        coef = 4.0 * np.pi * self.dr * self.r ** 2 * self.w
        return np.concatenate([spherical_jn(int(l), 2.0 * np.pi * np.outer(k[i:i + 64], self.r)) @ coef
                               for i in range(0, k.size, 64)]) if k.size else np.empty(0)


# ---------------------------------------------------------------- tables (cached)
def _key_of(value):
    if callable(value):
        return ("callable", id(value))
    if isinstance(value, np.ndarray):
        return ("array", value.shape, str(value.dtype), value.tobytes())
    if isinstance(value, dict):
        return tuple(sorted((str(k), _key_of(v)) for k, v in value.items()))
    if isinstance(value, (list, tuple)):
        return tuple(_key_of(v) for v in value)
    if isinstance(value, np.generic):
        return value.item()
    return value


class _LRU(OrderedDict):
    def __init__(self, limit):
        super().__init__()
        self.limit = limit

    def fetch(self, key, make):
        if key in self:
            self.move_to_end(key)
            return self[key]
        value = make()
        self[key] = value
        while len(self) > self.limit:
            self.popitem(last=False)
        return value


_NODES = _LRU(256)
_TABLES = _LRU(256)


def radial_nodes(radial_type, len_args, k_max, profile_config=None):
    key = (str(radial_type).strip().lower(), _key_of(dict(len_args)), _key_of(profile_config or {}), float(k_max))
    return _NODES.fetch(key, lambda: RadialNodes(radial_type, len_args, k_max, profile_config))


def build_radial_multipole_table(radial_type, len_args, l, k_max, profile_config=None):
    """(k_max, U_l at k_j = j k_max / (n - 1), j < n) for one physical profile; n grows with k_max times the
    profile's largest radius (radial_profiles.py:651-709).  Thin shells need no table."""
    t = str(radial_type).strip().lower()
    validate_radial_profile_request(t, len_args, profile_config)
    if t == "shell":
        raise ValueError("shell uses the analytic multipole kernel and does not require a radial table.")
    if not np.isfinite(k_max) or k_max <= 0.0:
        raise ValueError("k_max must be finite and positive when tabulating radial multipole windows.")
    key = (t, _key_of(dict(len_args)), _key_of(profile_config or {}), int(l), float(k_max))

    def make():
        if t == "gaussian_shell" and float(len_args["R_smooth"]) == 0.0:
            R = _length(len_args, "R_shell", t)
            n = _count(k_max * R, TABLE_MIN, TABLE_PER_CYCLE, None, "table_points")
            k = np.linspace(0.0, float(k_max), n)
            return float(k_max), np.ascontiguousarray(spherical_jn(int(l), 2.0 * np.pi * k * R))
        nodes = radial_nodes(t, len_args, k_max, profile_config)
        n = _count(k_max * float(np.max(nodes.r)), TABLE_MIN, TABLE_PER_CYCLE, nodes.options.get("table_points"),
                   "table_points")
        k = np.linspace(0.0, float(k_max), n)
        return float(k_max), np.ascontiguousarray(nodes.transfer(l, k), dtype=np.float64)

    return _TABLES.fetch(key, make)


def interpolate_table(k, table, k_max):
    """The kernels' table lookup: linear in the first and last interval, elsewhere the local cubic through
    the four nearest samples (Catmull-Rom); clamped to table[0] / table[-1] outside [0, k_max]."""
    k = np.clip(np.asarray(k, dtype=np.float64), 0.0, float(k_max))
    n = table.size
    x = k * (n - 1) / float(k_max)
    i = np.minimum(np.floor(x).astype(np.intp), n - 2)
    t = x - i
    out = table[i] * (1.0 - t) + table[i + 1] * t
    mid = (i > 0) & (i < n - 2)
    j, u = i[mid], t[mid]
    p0, p1, p2, p3 = table[j - 1], table[j], table[j + 1], table[j + 2]
    out[mid] = p1 + 0.5 * u * ((p2 - p0) + u * ((2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3)
                                                 + u * (3.0 * (p1 - p2) + p3 - p0)))
    out[x >= n - 1] = table[-1]
    return out


def diagnose_radial_multipole_table(radial_type, len_args, l, k_max, profile_config=None, *, tolerance=1.0e-5,
                                    probe_count=DIAGNOSTIC_PROBES):
    """Normalisation and convergence of one table (radial_profiles.py:559-648): the table at probe_count
    midpoints of [0, k_max] against the transform with doubled quadrature, U_l(0) against its exact value
    (the volume integral for l = 0, else 0), and for k-space profiles at l = 0 the round trip of the inverse
    and forward transforms against the given transfer."""
    t = str(radial_type).strip().lower()
    validate_radial_profile_request(t, len_args, profile_config)
    if is_thin_shell(t, len_args):
        return {"radial_type": t, "l": int(l), "path": "analytic_shell", "passed": True}
    if not np.isfinite(tolerance) or tolerance <= 0.0:
        raise ValueError("Radial-profile diagnostic tolerance must be finite and positive.")
    probe_count = int(probe_count)
    if probe_count < 3:
        raise ValueError("Radial-profile diagnostics require at least three probe points.")
    tk, table = build_radial_multipole_table(t, len_args, l, k_max, profile_config)
    nodes = radial_nodes(t, len_args, k_max, profile_config)
    finer = dict(profile_config or {})
    finer["quadrature_points"] = max(2 * nodes.r.size, 2 * R_NODES_MIN)
    if t in ("gaussian_shell", "custom_kspace"):
        finer["inverse_points"] = finer["quadrature_points"]
    fine = radial_nodes(t, len_args, k_max, finer)
    probes = float(k_max) * (np.arange(probe_count) + 0.5) / probe_count
    ref = fine.transfer(l, probes)
    conv = float(np.max(np.abs(interpolate_table(probes, table, tk) - ref)) / max(1.0, float(np.max(np.abs(ref)))))
    target = nodes.volume_integral() if int(l) == 0 else 0.0
    zero_err = abs(float(table[0]) - target)
    roundtrip = None
    if t in ("gaussian_shell", "custom_kspace") and int(l) == 0:
        given = (monopole_transfer(t, probes, len_args) if t == "gaussian_shell"
                 else _call_profile(_user_profile(profile_config)["func"], probes, len_args,
                                    _user_profile(profile_config).get("args", {})))
        back = nodes.transfer(0, probes)
        roundtrip = float(np.max(np.abs(back - given)) / max(1.0, float(np.max(np.abs(given)))))
    worst = max([conv, zero_err] + ([roundtrip] if roundtrip is not None else []))
    return {"radial_type": t, "l": int(l),
            "path": "inverse_hankel" if t in ("gaussian_shell", "custom_kspace") else "real_space_hankel",
            "table_points": int(table.size), "quadrature_points": int(nodes.r.size),
            "refined_quadrature_points": int(fine.r.size), "probe_count": probe_count,
            "zero_mode": float(table[0]), "zero_mode_target": float(target), "zero_mode_error": float(zero_err),
            "table_convergence_error": conv, "inverse_roundtrip_error": roundtrip, "tolerance": float(tolerance),
            "passed": bool(worst <= tolerance)}
