"""WindowFunc: Fourier window kernels built on the HIP device (drop-in for pyhermes.io.WindowFunc).

Built-in isotropic windows (sphere, gaussian, shell, thick_shell, gaussian_shell,
inverse_laplacian) and the line-of-sight windows (ring, disk, cylshell, cylinder,
window_functions.py:164-319) are built by hermes_window_build_rfft in the rFFT
layout (window.py:274-332); thin-shell ``radial_multipole`` and
``legendre_multipole`` windows by hermes_window_build_full in the complex
full-FFT layout (window.py:197-264); ``radial_multipole`` windows of the other radial
profiles (sphere, thick_shell, gaussian, gaussian_shell, custom real/k-space profiles)
by hermes_window_build_full_profile from a closed-form monopole or a host-built Hankel
table (radial_multipole_windows.py:267-316).  A user ``func`` is evaluated on the host
over the same alias sums and then used on the device.  Line-of-sight windows take any
line of sight; a non-axis one needs kernel_mode 'auto' or 'full_rfft' (the octant
fold assumes a window even in each k component).  The derivative windows of the
reference raise NotImplementedError.
"""

import copy
import inspect

import numpy as np
import torch

from .. import kernels
from ..utils import func_util
from ..utils.wavelet_grid import fourier_power_spectrum, sample_scaling_function
from ..utils.window_params import (ANISOTROPIC_AUTO_WINDOW_TYPES, COMPLEX_FULL_FFT_WINDOW_TYPES,
                                   COMPLEX_RFFT_WINDOW_TYPES, default_kernel_mode, normalize_kernel_mode,
                                   normalize_los_args, serialize_window_params)
from .sfc_field import SFCField

ISOTROPIC_TYPES = {"sphere", "gaussian", "shell", "thick_shell", "gaussian_shell", "inverse_laplacian"}
LOS_TYPES = set(kernels.LOS_TYPES)
KNOWN_UNSUPPORTED = {"cubic", "cw", "cws", "gdw", "directional_derivative", "laplacian"}
ZERO_MODE_ZERO_WINDOW_TYPES = {"inverse_laplacian"}


def _iso_desc(wtype, rescaled):
    if wtype in ("sphere", "gaussian", "shell"):
        return kernels.window_desc(wtype, float(rescaled["R"]))
    if wtype == "thick_shell":
        return kernels.window_desc(wtype, float(rescaled["R"]), float(rescaled["delta_R"]))
    if wtype == "gaussian_shell":
        return kernels.window_desc(wtype, float(rescaled["R_shell"]), float(rescaled["R_smooth"]))
    return kernels.window_desc("inverse_laplacian")


def los_axis(los_args):
    """(axis, unit vector) of a line of sight {nx, ny, nz}: axis 0 x, 1 y, 2 z when |n| / norm is within
    numpy.isclose's default tolerances of a coordinate axis (either sign; the windows are even in k) -- the
    reference's rule for an axis-aligned line of sight (window.py:150-170) --, else 3 with the normalised
    direction, as the reference divides k . n by |n| (convolution.py:73-157)."""
    vec = np.array([los_args["nx"], los_args["ny"], los_args["nz"]], dtype=np.float64)
    norm = float(np.linalg.norm(vec))
    if not np.isfinite(norm) or norm == 0.0:
        raise ValueError(f"line-of-sight vector must be finite and non-zero, got {tuple(vec)}")
    unit = np.abs(vec / norm)
    for axis in range(3):
        target = np.zeros(3)
        target[axis] = 1.0
        if np.all(np.isclose(unit, target)):
            return axis, None
    return 3, vec / norm


def los_desc(wtype, rescaled, los_args):
    """Descriptor of a ring / disk / cylshell / cylinder window (len_args R, H in grid units)."""
    try:
        R, H = float(rescaled["R"]), float(rescaled["H"])
    except (KeyError, TypeError, ValueError) as exc:
        raise ValueError(f"{wtype} requires finite len_args {{'R': ..., 'H': ...}}.") from exc
    if not (np.isfinite(R) and np.isfinite(H)) or R < 0.0:
        raise ValueError(f"{wtype} requires a finite non-negative R and a finite H.")
    axis, n = los_axis(los_args)
    return kernels.window_desc(wtype, R, H, axis=axis, los=n)


class WindowFunc(SFCField):
    def __init__(self, win_params, sfc_params, bandwidth=1, threads=1, phi_array=None, phi_fourier_power=None):
        super().__init__(threads=threads)
        missing = [k for k in self._REQUIRED_ARGV if k not in sfc_params]
        if missing:
            self.logger.error(f"WindowFunc missing required keys: {missing}")
            func_util.safe_exit(1)
        self.J = int(sfc_params["J"])
        self.L = 1 << self.J
        self.bandwidth = int(bandwidth)
        self.box_size = sfc_params["box_size"]
        self.phi_resolution = int(sfc_params["phi_resolution"])
        self.wavelet_mode = sfc_params["wavelet_mode"]
        self.wavelet_level = sfc_params["wavelet_level"]
        if self.bandwidth != 1:
            raise NotImplementedError("pyhermes_amd builds window kernels with bandwidth=1 only.")
        self.input_params = {**dict(win_params), "J": self.J, "box_size": self.box_size,
                             "phi_resolution": self.phi_resolution, "wavelet_mode": self.wavelet_mode,
                             "wavelet_level": self.wavelet_level, "bandwidth": self.bandwidth}
        self.phi_array = sample_scaling_function(self.wavelet_mode, self.wavelet_level) if phi_array is None \
            else phi_array
        if phi_fourier_power is None:
            phi_fourier_power = fourier_power_spectrum(self.phi_array, 0, 1, self.L, self.phi_resolution)
        if not isinstance(phi_fourier_power, np.ndarray):
            self.logger.error("WindowFunc requires phi_fourier_power to be a numpy.ndarray.")
            func_util.safe_exit(1)
        self.phi_fourier_power = phi_fourier_power
        self.window_params = dict(win_params)
        custom = win_params.get("func")
        if custom is not None and not callable(custom):
            raise TypeError("Custom window parameter 'func' must be callable.")
        self.has_custom_func = custom is not None
        if self.has_custom_func:
            self.type = win_params.get("type", None) or "custom"
            self.func = custom
        else:
            if "type" not in win_params or not win_params["type"]:
                self.logger.error("Window parameters need a 'type' or a 'func'.")
                func_util.safe_exit(1)
            self.type = win_params["type"]
            self.func = None
            if self.type in KNOWN_UNSUPPORTED:
                raise NotImplementedError(f"Window type '{self.type}' is outside the field-pipeline scope of "
                                          "pyhermes_amd.")
            if (self.type not in ISOTROPIC_TYPES and self.type not in LOS_TYPES
                    and self.type not in COMPLEX_FULL_FFT_WINDOW_TYPES):
                self.logger.error(f"Unsupported input window function type: {self.type}")
                func_util.safe_exit(1)
        mode = win_params.get("kernel_mode", None) or default_kernel_mode(self.type, self.has_custom_func)
        self.kernel_mode = normalize_kernel_mode(mode)
        self.input_params["kernel_mode"] = self.kernel_mode
        self.window_params["kernel_mode"] = self.kernel_mode
        self.len_args = copy.deepcopy(win_params.get("len_args", {})) or {}
        self.input_params["len_args"] = copy.deepcopy(self.len_args)
        self.window_params["len_args"] = copy.deepcopy(self.len_args)
        self.rescale_len_args = {k: v * self.L / self.box_size for k, v in self.len_args.items()
                                 if v is not None}
        self.los_args = normalize_los_args(win_params.get("los_args", {}), None if self.has_custom_func else self.type)
        self.input_params["los_args"] = self.los_args
        self.window_params["los_args"] = self.los_args
        self.other_args = win_params.get("other_args", {}) or {}
        self.window_args = dict(self.rescale_len_args)
        self.window_args.update(self.los_args)
        self.window_args.update(copy.deepcopy(self.other_args))
        self._validate_builtin_arguments()
        self.w_kernel = None
        self._dev_kernel = {}
        self.is_composite_window = False

    def _validate_builtin_arguments(self):
        """window.py:131-142; line-of-sight windows: R, H, a non-zero line of sight (octant mode only along
        an axis), checked up front."""
        if not self.has_custom_func and self.type in LOS_TYPES:
            if los_desc(self.type, self.rescale_len_args, self.los_args).axis == 3 and self.kernel_mode == "octant":
                raise NotImplementedError("pyhermes_amd builds line-of-sight windows along a non-axis "
                                          "direction with kernel_mode 'auto' / 'full_rfft' only")
            return
        if self.has_custom_func or self.type != "thick_shell":
            return
        try:
            radius = float(self.rescale_len_args["R"])
            width = float(self.rescale_len_args["delta_R"])
        except (KeyError, TypeError, ValueError) as exc:
            raise ValueError("thick_shell requires finite len_args {'R': ..., 'delta_R': ...}.") from exc
        if not np.isfinite(radius) or radius < 0.0:
            raise ValueError("thick_shell requires a finite non-negative R.")
        if not np.isfinite(width) or width <= 0.0:
            raise ValueError("thick_shell requires a finite positive delta_R.")

    # ------------------------------------------------------------ kernel construction
    def _radial_multipole_kernel(self, P, dbl):
        """window.py:212-238: radial_multipole windows of any radial profile -- closed form for thin shells and
        built-in monopoles, else the profile's Hankel table (profile_config = other_args['profile_config'])."""
        from ..edges import EdgeProfile
        try:
            radial_type = str(self.other_args["radial_type"])
            l, m = int(self.other_args["l"]), int(self.other_args["m"])
        except KeyError as exc:
            self.logger.error(f"radial_multipole missing required other_args key: {exc}")
            func_util.safe_exit(1)
        try:
            edge = EdgeProfile(radial_type, self.len_args, self.other_args.get("profile_config", {}), self.L,
                               self.box_size)
            return edge.window_full(self.J, P, l, m, dbl=dbl)
        except (TypeError, ValueError) as exc:
            self.logger.error(f"Failed to build radial_multipole window: {exc}")
            func_util.safe_exit(1)

    def _multipole_desc(self):
        oa = self.other_args
        if "R" not in self.rescale_len_args:
            self.logger.error(f"{self.type} requires len_args={{'R': ...}}.")
            func_util.safe_exit(1)
        return kernels.window_desc("shell_multipole", float(self.rescale_len_args["R"]), l=int(oa["l"]),
                                   m=int(oa["m"]))

    def _host_custom_kernel(self):
        """Evaluate a user transfer function over the 8-alias sums (convolution.py:14-38, :73-157)."""
        sig = inspect.signature(self.func)
        names = list(sig.parameters)[3:]
        args = []
        for n in names:
            if n in self.window_args:
                args.append(self.window_args[n])
            elif sig.parameters[n].default is not inspect.Parameter.empty:
                args.append(sig.parameters[n].default)
            else:
                self.logger.error(f"Missing keyword arguments: {[n]}")
                func_util.safe_exit(1)
        f = np.vectorize(lambda a, b, c: self.func(a, b, c, *args), otypes=[np.float64])
        L, P = self.L, self.phi_fourier_power
        Lh = L // 2 + 1
        x = np.arange(L)[:, None, None]
        y = np.arange(L)[None, :, None]
        z = np.arange(Lh)[None, None, :]
        if self.kernel_mode == "octant":
            i = np.arange(L + 1)
            wa = P[i][:, None, None] * P[i][None, :, None] * P[i][None, None, :] * f(
                i[:, None, None] / L, i[None, :, None] / L, i[None, None, :] / L)
            k = 0.0
            for xs in (x, L - x):
                for ys in (y, L - y):
                    for zs in (z, L - z):
                        k = k + wa[xs, ys, zs]
            return k
        k = 0.0
        for xs, sx in ((x, 1), (L - x, -1)):
            for ys, sy in ((y, 1), (L - y, -1)):
                for zs, sz in ((z, 1), (L - z, -1)):
                    k = k + P[xs] * P[ys] * P[zs] * f(sx * xs / L, sy * ys / L, sz * zs / L)
        return k

    def device_kernel(self, dbl=True):
        """The kernel as a device tensor (FP64 or FP32), built once per precision."""
        key = bool(dbl)
        if key in self._dev_kernel:
            return self._dev_kernel[key]
        if self.w_kernel is not None and not isinstance(self.w_kernel, torch.Tensor):
            host = np.asarray(self.w_kernel)
            t = kernels.as_device(host, torch.from_numpy(host[:0]).dtype)
        elif isinstance(self.w_kernel, torch.Tensor):
            t = self.w_kernel
        elif self.has_custom_func:
            if self.kernel_mode in ("complex_rfft", "complex_full_fft"):
                self.logger.error("Custom window functions support real rFFT kernel modes only in pyhermes_amd.")
                func_util.safe_exit(1)
            t = kernels.as_device(self._host_custom_kernel(), torch.float64)
        else:
            P = kernels.as_device(self.phi_fourier_power[:self.L + 1], torch.float64)
            if self.type in COMPLEX_FULL_FFT_WINDOW_TYPES:
                if self.kernel_mode not in ("auto", "complex_full_fft"):
                    self.logger.error(f"Window type '{self.type}' requires kernel_mode='complex_full_fft' or "
                                      "kernel_mode='auto'.")
                    func_util.safe_exit(1)
                if self.type == "radial_multipole":
                    t = self._radial_multipole_kernel(P, dbl)
                else:
                    t = kernels.window_build_full(self.J, P, self._multipole_desc(), dbl=dbl)
            elif self.type in LOS_TYPES:
                # even in every component of k for an axis line of sight: the rFFT kernel of the reference's
                # "auto" (full_rfft) mode and the octant fold sum the same alias values.  Along any other
                # direction only the signed-alias (full_rfft) sum is built.
                t = kernels.window_build_rfft(self.J, P, los_desc(self.type, self.rescale_len_args, self.los_args),
                                              dbl=dbl)
            else:
                t = kernels.window_build_rfft(self.J, P, _iso_desc(self.type, self.rescale_len_args), dbl=dbl)
        self._dev_kernel[key] = t
        return t

    def _apply_zero_mode_convention(self):
        if not self.has_custom_func and self.type in ZERO_MODE_ZERO_WINDOW_TYPES and self.w_kernel is not None:
            self.w_kernel[0, 0, 0] = 0.0

    def as_array(self):
        if self.w_kernel is None:
            self.w_kernel = self.device_kernel(dbl=True).cpu().numpy()
            self._apply_zero_mode_convention()
        elif isinstance(self.w_kernel, torch.Tensor):
            self.w_kernel = self.w_kernel.cpu().numpy()
        return self.w_kernel

    # ------------------------------------------------------------ window arithmetic (window.py:345-516)
    def _window_operation_descriptor(self):
        return serialize_window_params(getattr(self, "window_params", {}))

    def _check_window_compatibility(self, other):
        for key in ("L", "bandwidth", "box_size", "phi_resolution", "wavelet_mode", "wavelet_level"):
            if getattr(self, key) != getattr(other, key):
                self.logger.error(f"Cannot combine WindowFunc objects with different {key}: "
                                  f"{getattr(self, key)} vs {getattr(other, key)}.")
                func_util.safe_exit(1)

    def _spawn_composite(self, kernel, operation):
        new = self.__class__.__new__(self.__class__)
        new.__dict__.update({k: v for k, v in self.__dict__.items() if k not in ("w_kernel", "_dev_kernel")})
        new.type = "composite"
        new.func = None
        new.has_custom_func = False
        new.kernel_mode = "materialized"
        new.len_args, new.rescale_len_args, new.los_args = {}, {}, {}
        new.other_args = {"operation": copy.deepcopy(operation)}
        new.window_args = {}
        new.window_params = {"type": "composite", "len_args": {}, "los_args": {}, "other_args": new.other_args,
                             "kernel_mode": "materialized"}
        new.w_kernel = kernel
        new._dev_kernel = {}
        new.is_composite_window = True
        return new

    def _binary_window_op(self, other, name, fn):
        if not isinstance(other, WindowFunc):
            return NotImplemented
        self._check_window_compatibility(other)
        a, b = self.device_kernel(True), other.device_kernel(True)
        if a.shape != b.shape:
            self.logger.error(f"Cannot combine WindowFunc kernels with different shapes: {tuple(a.shape)} vs "
                              f"{tuple(b.shape)}.")
            func_util.safe_exit(1)
        op = {"op": name, "left": self._window_operation_descriptor(), "right": other._window_operation_descriptor()}
        return self._spawn_composite(fn(a, b), op)

    def _scalar_window_op(self, scalar, name, fn):
        if not (np.isscalar(scalar) and not isinstance(scalar, (str, bytes)) and np.isrealobj(scalar)):
            return NotImplemented
        op = {"op": name, "window": self._window_operation_descriptor(), "scalar": float(scalar)}
        return self._spawn_composite(fn(self.device_kernel(True), float(scalar)), op)

    def __add__(self, other):
        return self._binary_window_op(other, "add", lambda a, b: a + b)

    def __sub__(self, other):
        return self._binary_window_op(other, "sub", lambda a, b: a - b)

    def __neg__(self):
        return self._scalar_window_op(-1.0, "mul", lambda w, s: w * s)

    def __mul__(self, other):
        if isinstance(other, WindowFunc):
            return self._binary_window_op(other, "projected_kernel_product", lambda a, b: a * b)
        return self._scalar_window_op(other, "mul", lambda w, s: w * s)

    def __rmul__(self, other):
        return self.__mul__(other)

    def __truediv__(self, other):
        if np.isscalar(other) and float(other) == 0.0:
            self.logger.error("Cannot divide WindowFunc by zero.")
            func_util.safe_exit(1)
        return self._scalar_window_op(other, "div", lambda w, s: w / s)

    def copy(self, copy_kernel=True):
        if self.is_composite_window and not copy_kernel:
            self.logger.error("Composite WindowFunc cannot be copied with copy_kernel=False.")
            func_util.safe_exit(1)
        new = self.__class__.__new__(self.__class__)
        for key, value in self.__dict__.items():
            if key in ("comm", "logger", "func"):
                setattr(new, key, value)
            elif key == "w_kernel":
                setattr(new, key, (value.clone() if isinstance(value, torch.Tensor) else value.copy())
                        if (copy_kernel and value is not None) else None)
            elif key == "_dev_kernel":
                setattr(new, key, dict(value) if copy_kernel else {})
            elif isinstance(value, np.ndarray):
                setattr(new, key, value.copy())
            else:
                setattr(new, key, copy.deepcopy(value))
        return new
