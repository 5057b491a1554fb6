"""SFCField: scaling-function-coefficient field resident in HBM (drop-in for pyhermes.io.SFCField).

``epsilon`` keeps the reference's numpy semantics for users (a host copy is made
on access), while the coefficients stay on the HIP device for every operation:
``field @ window`` (sfc_field.py:93-160) runs rocFFT plus the HIP spectral
multiply, the derived field integral (sfc_field.py:312-324) is an FP64 device
reduction, and ``field_density_at_pos`` (sfc_field.py:487-530) is the HIP
interpolation kernel.
"""

import copy

import numpy as np
import torch

from .. import kernels
from ..utils import func_util
from .base import HermesData


# Normalisation vocabularies (sfc_field.py:21-47): accepted spellings -> canonical name, and the error for
# anything else.  A projection folds 'unit' into 'field' (unit weights are the field normalisation of the
# weights); tasks and field values keep 'unit' (rescale the field to unit integral).
_FOUR = ("raw", "catalog", "field", "unit")
_VOCABULARIES = {
    "projection": ({**{m: m for m in _FOUR}, "unit": "field"},
                   "weight_normalization must be one of 'raw', 'catalog', 'field', or 'unit'."),
    "task": ({m: m for m in _FOUR}, "task weight_normalization must be one of 'raw', 'catalog', 'field', or 'unit'."),
    "value": ({m: m for m in _FOUR}, "normalization must be None, 'raw', 'catalog', 'field', or 'unit'."),
}


def _canonical(vocabulary, value, none_ok=False):
    names, message = _VOCABULARIES[vocabulary]
    if value is None and none_ok:
        return None
    key = None if value is None else str(value).strip().lower()
    if key not in names:
        raise ValueError(message)
    return names[key]


def normalize_weight_normalization(weight_normalization):
    """Projection weight normalisation: 'raw', 'catalog' or 'field' ('unit' means 'field')."""
    return _canonical("projection", weight_normalization)


def normalize_task_weight_normalization(weight_normalization):
    """A task's weight normalisation: 'raw', 'catalog', 'field' or 'unit'."""
    return _canonical("task", weight_normalization)


def normalize_field_value_normalization(normalization):
    """The normalisation a field's values are wanted in: None (as they are) or one of the four names."""
    return _canonical("value", normalization, none_ok=True)


def _scalar(x):
    return np.isscalar(x) and not isinstance(x, (str, bytes))


class SFCField(HermesData):
    _REQUIRED_ARGV = ("J", "box_size", "phi_resolution", "wavelet_mode", "wavelet_level")

    def __init__(self, *args, threads=None, **kwargs):
        data_path = kwargs.pop("data_path", None)
        self.sfc_info = {}
        self._eps = None
        self._eps_host = None
        super().__init__(*args, threads=threads, **kwargs)
        if data_path:
            self.sfc_info["sfc_field_path"] = data_path
            self.load_sfc_field(data_path)

    # ------------------------------------------------------------ storage
    @property
    def epsilon(self):
        if self._eps is None:
            return None
        if isinstance(self._eps, np.ndarray):
            return self._eps
        if self._eps_host is None:
            self._eps_host = self._eps.detach().cpu().numpy()
        return self._eps_host

    @epsilon.setter
    def epsilon(self, value):
        self._eps = value
        self._eps_host = None

    def device_epsilon(self, dtype=None):
        """Coefficients as a device tensor (moved to HBM once and kept there)."""
        if self._eps is None:
            raise ValueError("epsilon is None.")
        if isinstance(self._eps, np.ndarray):
            host = self._eps
            self._eps = kernels.as_device(host, torch.from_numpy(host[:0]).dtype)
            self._eps_host = host
        t = self._eps
        if dtype is not None and t.dtype != dtype:
            t = t.to(dtype)
        return t

    def _set_device(self, t):
        self._eps = t
        self._eps_host = None

    # ------------------------------------------------------------ copies
    def _spawn_like(self):
        new = self.__class__(threads=self.threads)
        new.comm, new.rank, new.logger = self.comm, self.rank, self.logger
        new.sfc_info = copy.deepcopy(self.sfc_info)
        new.task_params = copy.deepcopy(self.task_params)
        new._eps = None
        new.saveflag = False
        return new

    def copy(self):
        new = self._spawn_like()
        if isinstance(self._eps, np.ndarray):
            new._eps = self._eps.copy()
        elif self._eps is not None:
            new._eps = self._eps.clone()
        new.format_sfc_params()
        return new

    # ------------------------------------------------------------ convolution
    def _conv(self, other):
        if not hasattr(other, "device_kernel") and not hasattr(other, "as_array"):
            self.logger.error(f"Convolution requires `other` to implement as_array(); got {type(other)}.")
            func_util.safe_exit(1)
        eps = self.device_epsilon()
        if eps.ndim != 3:
            self.logger.error(f"Convolution expects 3D arrays; got a.ndim={eps.ndim}.")
            func_util.safe_exit(1)
        dbl = eps.dtype in (torch.float64, torch.complex128)
        if hasattr(other, "device_kernel"):
            K = other.device_kernel(dbl=dbl)
        else:
            K = kernels.as_device(other.as_array(), torch.float64 if dbl else torch.float32)
        nx, ny, nz = eps.shape
        N = nx * ny * nz
        if tuple(K.shape) == (nx, ny, nz // 2 + 1) and not K.is_complex() and not eps.is_complex():
            spec = kernels.rfftn(eps)
            kernels.spectrum_mul(spec, K.to(torch.float64 if dbl else torch.float32), 1.0 / N)
            conv = kernels.irfftn_unnormalized(spec, nx)
        elif tuple(K.shape) == (nx, ny, nz):
            cdt = torch.complex128 if dbl else torch.complex64
            ec = eps.to(cdt) if eps.is_complex() else torch.complex(eps, torch.zeros_like(eps)).to(cdt)
            spec = kernels.fftn_c2c(ec)
            Kc = K.to(cdt) if K.is_complex() else torch.complex(K, torch.zeros_like(K)).to(cdt)
            kernels.spectrum_mul(spec, Kc, 1.0 / N)
            conv = kernels.fftn_c2c(spec, inverse=True, out=spec)
        else:
            self.logger.error(
                f"Convolution requires same shape; got a.shape={tuple(eps.shape)}, b.shape={tuple(K.shape)}. "
                f"Expected an rFFT kernel with shape {(nx, ny, nz // 2 + 1)} or a complex full-FFT kernel with "
                f"shape {(nx, ny, nz)}. In this API, the convolution window is expected on the right-hand side "
                "(signal @ window).")
            func_util.safe_exit(1)
        new = self._spawn_like()
        new._set_device(conv)
        windows = self.sfc_info.get("convolution_of", [])
        new.sfc_info.update({"convolution_of": windows + [getattr(other, "window_params", {})]})
        self._mark_derived(new)
        new.format_sfc_params()
        return new

    def __matmul__(self, other):
        if isinstance(other, SFCField):
            return self._conv(other)
        return NotImplemented

    def __rmatmul__(self, other):
        if isinstance(other, SFCField):
            return other._conv(self)
        return NotImplemented

    # ------------------------------------------------------------ arithmetic (device)
    def _binary(self, other, op, name):
        if isinstance(other, SFCField):
            if self._eps is None or other._eps is None:
                self.logger.error(f"Cannot perform {name}: epsilon is None.")
                func_util.safe_exit(1)
            a, b = self.device_epsilon(), other.device_epsilon()
            if a.shape != b.shape:
                self.logger.error(f"Shape mismatch in {name}: {tuple(a.shape)} vs {tuple(b.shape)}")
                func_util.safe_exit(1)
            res = op(a, b)
        else:
            if self._eps is None:
                self.logger.error(f"Cannot perform {name}: epsilon is None.")
                func_util.safe_exit(1)
            res = op(self.device_epsilon(), other)
        new = self._spawn_like()
        new._set_device(res)
        self._mark_derived(new)
        new.format_sfc_params()
        return new

    def __add__(self, other):
        if isinstance(other, SFCField) or _scalar(other):
            return self._binary(other, lambda a, b: a + b, "addition")
        return NotImplemented

    __radd__ = __add__

    def __sub__(self, other):
        if isinstance(other, SFCField) or _scalar(other):
            return self._binary(other, lambda a, b: a - b, "subtraction")
        return NotImplemented

    def __rsub__(self, other):
        if _scalar(other):
            return self._binary(other, lambda a, b: b - a, "subtraction")
        return NotImplemented

    def __mul__(self, other):
        if isinstance(other, SFCField) or _scalar(other):
            return self._binary(other, lambda a, b: a * b, "multiplication")
        return NotImplemented

    __rmul__ = __mul__

    def __truediv__(self, other):
        if _scalar(other) and np.isrealobj(other):
            if float(other) == 0.0:
                self.logger.error("Cannot divide SFCField by zero.")
                func_util.safe_exit(1)
            return self._binary(float(other), lambda a, b: a / b, "division")
        return NotImplemented

    def __neg__(self):
        return self._binary(-1.0, lambda a, b: a * b, "negation")

    # ------------------------------------------------------------ metadata
    def _derived_field_integral(self, field):
        if field._eps is None:
            return None
        total = kernels.reduce_sum(field.device_epsilon())
        if isinstance(total, complex):
            if not (np.isfinite(total.real) and np.isfinite(total.imag)):
                return None
            if abs(total.imag) <= 1e-12 * max(1.0, abs(total.real)):
                return float(total.real)
            return complex(total)
        return float(total) if np.isfinite(total) else None

    def _mark_derived(self, field):
        field.sfc_info.update({
            "catalog_weight_sum": None,
            "catalog_weight_sq_sum": None,
            "raw_field_weighted_sum": None,
            "field_integral": self._derived_field_integral(field),
            "weight_normalization": None,
            "particle_count": None,
            "particle_data_retrievable": False,
            "particle_data_path": "",
            "particle_data_format": "",
            "field_kind": "derived_field",
        })

    def format_sfc_params(self):
        missing = [k for k in self._REQUIRED_ARGV if k not in self.sfc_info]
        if missing:
            self.logger.error(f"SFCField missing required keys: {missing}")
            func_util.safe_exit(1)
        for key, value in self.sfc_info.items():
            setattr(self, key, value)

    # ------------------------------------------------------------ normalisation (sfc_field.py:340-423)
    def _normalization_denominator(self, weight_normalization):
        mode = normalize_weight_normalization(weight_normalization)
        if mode == "raw":
            return 1.0
        label = "catalog_weight_sum" if mode == "catalog" else "raw_field_weighted_sum"
        value = getattr(self, label, None)
        if value is None or not np.isfinite(value) or np.isclose(value, 0.0):
            raise ValueError(f"Cannot use weight_normalization='{mode}' without a finite, non-zero {label}.")
        return float(value)

    def _particle_file(self):
        """The particle file behind a catalogue field (sfc_field.py:425-445 conditions and messages)."""
        if getattr(self, "field_kind", None) != "catalog_field":
            raise ValueError("Only catalog fields can provide particle data. Derived fields do not retain a unique "
                             "particle catalogue.")
        if not bool(getattr(self, "particle_data_retrievable", True)):
            raise ValueError("This field no longer maps to one retrievable particle catalogue. Provide "
                             "particle_pos1 and particle_weight1 explicitly for particle-centred statistics.")
        path = getattr(self, "particle_data_path", "")
        if not path:
            raise ValueError("This field has no saved particle file (project with save_particle_data=True); "
                             "reading the input catalogue again is not provided by pyhermes_amd.")
        return path

    def get_particle_data(self):
        """The catalogue behind a catalogue field, from the particle file SFCProjection saved
        (save_particle_data=True): pos, catalog_weight, normalized_catalog_weight, field_value and
        projection_weight, the last in the field's own weight normalisation (sfc_field.py:425-453).  Reading
        the original input catalogue again (sfc_field.py:455-480) needs the reference's file readers and is
        not provided."""
        path = self._particle_file()
        den = self._normalization_denominator(getattr(self, "weight_normalization", "raw"))
        with np.load(path, allow_pickle=False) as archive:
            missing = {"pos", "catalog_weight", "field_value"} - set(archive.files)
            if missing:
                self.logger.error(f"Particle dataset '{path}' must contain "
                                  f"{sorted({'pos', 'catalog_weight', 'field_value'})} arrays.")
                func_util.safe_exit(1)
            arrays = {k: np.asarray(archive[k]) for k in ("pos", "catalog_weight", "field_value")}
        w = arrays["catalog_weight"].astype(np.float64)
        q = arrays["field_value"].astype(np.float64)
        return {"pos": arrays["pos"], "catalog_weight": w, "normalized_catalog_weight": w / float(self.catalog_weight_sum),
                "field_value": q, "projection_weight": w * q / den}

    def _field_integral_for_normalization(self, weight_normalization):
        raw = getattr(self, "raw_field_weighted_sum", None)
        if raw is None or not np.isfinite(raw):
            return None
        return float(raw / self._normalization_denominator(weight_normalization))

    def switch_weight_normalization(self, weight_normalization):
        mode = normalize_weight_normalization(weight_normalization)
        if getattr(self, "field_kind", None) != "catalog_field":
            raise ValueError("weight normalization can only be switched for catalog fields.")
        current = normalize_weight_normalization(getattr(self, "weight_normalization", "raw") or "raw")
        new = self.copy()
        if current == mode:
            return new
        factor = self._normalization_denominator(current) / self._normalization_denominator(mode)
        new._set_device(new.device_epsilon() * factor)
        new.sfc_info["weight_normalization"] = mode
        new.sfc_info["field_integral"] = self._field_integral_for_normalization(mode)
        new.format_sfc_params()
        return new

    def switch_normalization(self, weight_normalization):
        return self.with_normalization(weight_normalization)

    def to_unit_weight(self):
        if getattr(self, "field_kind", None) == "catalog_field":
            return self.switch_weight_normalization("field")
        total = getattr(self, "field_integral", None)
        if total is None:
            total = self._derived_field_integral(self)
        if total is None or not np.isfinite(total) or np.isclose(total, 0.0):
            raise ValueError("Cannot rescale a derived field with zero or non-finite epsilon sum.")
        new = self.copy()
        new._set_device(new.device_epsilon() / total)
        self._mark_derived(new)
        new.format_sfc_params()
        return new

    def with_normalization(self, normalization=None):
        """This field in another normalisation (sfc_field.py:397-405): None keeps it, 'unit' rescales to unit
        integral, 'raw' / 'catalog' / 'field' re-weight a catalogue field."""
        mode = normalize_field_value_normalization(normalization)
        if mode in (None, "unit"):
            return self if mode is None else self.to_unit_weight()
        if getattr(self, "field_kind", None) == "catalog_field":
            return self.switch_weight_normalization(mode)
        raise ValueError("normalization='raw', 'catalog', or 'field' can only be applied to catalog fields.")

    _VALUE_UNITS = ("auto", "grid", "physical")

    def _resolve_value_unit(self, value_unit, sample_field=None):
        """'grid' (per grid cell) or 'physical' (per box volume unit); 'auto' is physical for catalogue fields
        and grid for derived ones (sfc_field.py:407-414)."""
        unit = str(value_unit).strip().lower()
        if unit not in self._VALUE_UNITS:
            raise ValueError("value_unit must be 'auto', 'grid', or 'physical'.")
        if unit != "auto":
            return unit
        kind = getattr(self if sample_field is None else sample_field, "field_kind", None)
        return "physical" if kind == "catalog_field" else "grid"

    def field_mean_density(self, normalization=None, value_unit="auto"):
        """Mean density: the field integral over the volume, in grid cells or box units (sfc_field.py:416-423);
        None when the integral is unknown."""
        field = self.with_normalization(normalization)
        unit = self._resolve_value_unit(value_unit, sample_field=field)
        total = getattr(field, "field_integral", None)
        if total is None:
            return None
        return total / (float(field.V) if unit == "grid" else float(field.box_size) ** 3)

    # ------------------------------------------------------------ sampling at positions
    def field_density_at_pos(self, pos, epsilon=None, filter=None, normalization=None, value_unit="auto"):
        """n(x) at positions (box units) via the HIP interpolation kernel (wavelet_grid.py:143-186)."""
        if epsilon is not None and normalization is not None:
            raise ValueError("normalization cannot be applied when an explicit epsilon array is provided.")
        sample = self.with_normalization(normalization)
        value_unit = self._resolve_value_unit(value_unit, sample_field=sample)
        if epsilon is None:
            grid = (sample._conv(filter) if filter is not None else sample).device_epsilon(torch.float64)
        else:
            grid = kernels.as_device(epsilon, torch.float64)
        spos = kernels.as_device(np.asarray(pos, dtype=np.float64) * float(sample.scale_factor), torch.float64)
        nx = kernels.interpolate(grid, spos, sample.phi_array).cpu().numpy()
        if value_unit == "physical":
            return nx * float(sample.scale_factor) ** 3
        return nx

    def as_array(self):
        return self.epsilon

    def phi_fourier_power_for(self, L):
        """|phi_hat|^2 at k = j / L, j = 0..L (wavelet_grid.py:94-118), cached per L."""
        from ..utils.wavelet_grid import fourier_power_spectrum, sample_scaling_function
        cached = getattr(self, "_phi_power_cache", None)
        if cached is not None and cached[0] == L:
            return cached[1]
        phi = getattr(self, "phi_array", None)
        if phi is None:
            phi = sample_scaling_function(self.wavelet_mode, self.wavelet_level)
        P = fourier_power_spectrum(np.asarray(phi), 0, 1, int(L), int(self.phi_resolution))
        self._phi_power_cache = (L, P)
        return P

    # ------------------------------------------------------------ I/O (sfc_field.py:543-584)
    def _payload(self):
        if not self.sfc_info:
            self.logger.error('The dictionary "sfc_info" is empty.')
            func_util.safe_exit(1)
        return {"sfc_info": self.sfc_info, "epsilon": self.epsilon}

    def _restore(self, dataset):
        if "epsilon" not in dataset:
            self.logger.error("Failed to load the dataset. The file is missing the 'epsilon' key.")
            func_util.safe_exit(1)
        self.epsilon = dataset["epsilon"]
        if dataset.get("sfc_info"):
            self.sfc_info.update(dataset["sfc_info"])
        self.format_sfc_params()

    def load_sfc_field(self, f_in, single=True, trusted=False):
        return self.load(f_in, trusted=trusted)

    def save_sfc_field(self, f_out, single=True, overwrite=False):
        return self.save(f_out, overwrite=overwrite)
