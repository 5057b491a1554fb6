"""Corr_2PCF: pair products with shell binning windows on the HIP device.

Drop-in for pyhermes.theory.Corr_2PCF (corr2pcf.py:226-382, :597-731, :936-1264):
samplings 's' (mapping 's_to_R', isotropic binning windows), ('s', 'mu') ('smu_to_RH')
and ('rp', 'pi') ('rppi_to_RH') with the line-of-sight ring / disk / cylshell /
cylinder windows, or a callable mapping; results have the shape of the sampling
grid (np.ndindex order over the mapping's sampling names, corr2pcf.py:226-232).
Instead of one window build and two FFTs per sample (corr2pcf.py:366-372), each
field is transformed once and every sample is evaluated in one fused Parseval pass
(hermes_pair_reduce):
  pp_b = sum_x (f1 @ W_b)(x) f2(x) / N = (1/N^2) sum_k W_b(k) Re F1(k) conj F2(k).
Products dd, dr, rd, delta_dd, rr, xi follow the reference; uniform randoms are
the scalar density rho; values are scaled by (L / box)^6.

Extension (no reference counterpart): ``multipoles: [0, 2, 4]`` adds
xi_l(s) = (-i)^l sqrt(4 pi (2l + 1)) pp_l(s) / rr from j_l(2 pi k s) Y_l0 windows,
evaluated in the same fused pass (s sampling only).
"""

import copy
import time

import numpy as np
import torch

from .. import kernels
from ..io import Corr2PCFData, SFCField, WindowFunc, normalize_task_weight_normalization
from ..pipeline import TaskBase
from ..utils import func_util
from ..io.window import los_desc
from ..utils.window_params import LOS_AWARE_WINDOW_TYPES, default_binning_window, normalize_binning_window_template

PRODUCT_NAMES = ("dd", "dr", "rd", "delta_dd", "rr", "xi")
PRODUCT_INPUT_FLAGS = {"dd": (True, False), "dr": (True, True), "rd": (True, True), "delta_dd": (True, True),
                       "rr": (False, True), "xi": (True, True)}
NONNEGATIVE_SAMPLING_ARGS = ("s", "rp", "pi")


# Built-in binning-window mappings (corr2pcf.py:267-322): a sample fills the window's unset lengths.
def _fill(params, key, value):
    if params["len_args"].get(key) is None:
        params["len_args"][key] = value


def map_s_to_R(sample, bw):
    p = copy.deepcopy(bw)
    p.setdefault("len_args", {})
    _fill(p, "R", sample["s"])
    return p


def map_smu_to_RH(sample, bw):
    p = copy.deepcopy(bw)
    p.setdefault("len_args", {})
    s, mu = sample["s"], sample["mu"]
    _fill(p, "R", s * np.sqrt(max(0.0, 1.0 - mu * mu)))
    _fill(p, "H", s * mu)
    return p


def map_rppi_to_RH(sample, bw):
    p = copy.deepcopy(bw)
    p.setdefault("len_args", {})
    _fill(p, "R", sample["rp"])
    _fill(p, "H", sample["pi"])
    return p


BINNING_MAPPINGS = {"s_to_R": (("s",), map_s_to_R), "smu_to_RH": (("s", "mu"), map_smu_to_RH),
                    "rppi_to_RH": (("rp", "pi"), map_rppi_to_RH)}


def binning_desc(params, L, box):
    """Window descriptor of one sample's binning-window parameters (lengths in box units)."""
    wtype = str(params.get("type", "")).strip().lower()
    if params.get("func") is not None:
        raise NotImplementedError("custom binning-window functions are outside pyhermes_amd's fused pair pass.")
    la = {k: (None if v is None else float(v) * L / box) for k, v in (params.get("len_args") or {}).items()}

    def need(*keys):
        missing = [k for k in keys if la.get(k) is None]
        if missing:
            raise ValueError(f"binning window '{wtype}' needs len_args {missing} (set them or use a mapping).")
        return [la[k] for k in keys]

    if wtype in ("shell", "sphere", "gaussian"):
        return kernels.window_desc(wtype, *need("R"))
    if wtype == "thick_shell":
        return kernels.window_desc(wtype, *need("R", "delta_R"))
    if wtype == "gaussian_shell":
        return kernels.window_desc(wtype, *need("R_shell", "R_smooth"))
    if wtype in LOS_AWARE_WINDOW_TYPES and wtype in kernels.LOS_TYPES:
        R, H = need("R", "H")
        return los_desc(wtype, {"R": R, "H": H}, params.get("los_args") or {"nx": 0.0, "ny": 0.0, "nz": 1.0})
    raise NotImplementedError(f"binning window type '{wtype}' is not available in pyhermes_amd's pair pass.")


def normalize_products(products):
    if isinstance(products, str):
        products = [products]
    elif products is None:
        products = ["xi"]
    elif not isinstance(products, (list, tuple, set)):
        raise TypeError(f"Unsupported products input: expected string or array of strings, got {type(products)}.")
    out = []
    for item in products:
        if not isinstance(item, str):
            raise TypeError("Each product name must be a string.")
        name = item.strip().lower()
        if name not in PRODUCT_NAMES:
            raise ValueError(f"Unsupported product '{item}'. Allowed values are {sorted(PRODUCT_NAMES)}.")
        if name not in out:
            out.append(name)
    return out


def expand_products(products):
    out = list(products)
    if "xi" in out:
        for dep in ("delta_dd", "rr"):
            if dep not in out:
                out.append(dep)
    return out


def normalize_sampling_spec(name, spec):
    """corr2pcf.py:87-99, :210-223: {min, max, n} -> linspace, or a monotonic 1-D array."""
    if isinstance(spec, dict):
        n = int(spec["n"])
        if n <= 0:
            raise ValueError(f"'{name}' sampling n must be positive.")
        arr = np.linspace(float(spec["min"]), float(spec["max"]), n)
    else:
        arr = np.asarray(spec, dtype=np.float64)
        if arr.ndim != 1:
            raise TypeError(f"'{name}' must be a 1D array-like input.")
        if arr.size == 0:
            raise ValueError(f"'{name}' must contain at least one sampling point.")
        d = np.diff(arr)
        if np.any(d < 0.0) and np.any(d > 0.0):
            raise ValueError(f"'{name}' values must be monotonic.")
    if name in NONNEGATIVE_SAMPLING_ARGS and np.any(arr < 0):
        raise ValueError(f"'{name}' sampling values must be non-negative.")
    return np.ascontiguousarray(arr, dtype=np.float64)


def _is_density(x):
    return isinstance(x, (float, int, np.floating))


def pair_products_with_descs(a, b, descs, cache=None):
    """sum_x (a @ W_d)(x) b(x) / N for every window descriptor d in one fused pass.

    a, b: SFCField or scalar densities (corr2pcf.py:375-382 uniform shortcut
    rho_a rho_b).  cache holds per-field spectra across calls.
    """
    cache = {} if cache is None else cache
    if _is_density(a) or _is_density(b):
        va = a if _is_density(a) else a.field_mean_density(value_unit="grid")
        vb = b if _is_density(b) else b.field_mean_density(value_unit="grid")
        return np.full(len(descs), float(va) * float(vb))
    def spectrum(f):
        if id(f) not in cache:
            cache[id(f)] = (f, kernels.rfftn(f.device_epsilon()))  # keep f alive so its id stays unique
        return cache[id(f)][1]

    Fa = spectrum(a)
    Fb = Fa if b is a else spectrum(b)
    if Fb.dtype != Fa.dtype:
        Fb = Fb.to(Fa.dtype)
    L = int(a.device_epsilon().shape[0])
    if "P" not in cache:
        cache["P"] = kernels.as_device(a.phi_fourier_power_for(L), torch.float64)
    return kernels.pair_reduce(Fa, Fb, cache["P"], descs) / float(L) ** 6


def shell_pair_products(a, b, radii, L, box, cache=None):
    """Pair products with thin-shell windows of physical radii (the default 's_to_R' binning)."""
    descs = [kernels.window_desc("shell", float(r) * L / box) for r in np.atleast_1d(radii)]
    return pair_products_with_descs(a, b, descs, cache)


class Corr_2PCF(TaskBase):
    def __init__(self, param_task=None):
        if param_task is None:
            param_task = {"Corr_2PCF": {}}
        self.task_name = str(self.__class__.__name__)
        super().__init__(param_task=param_task)
        self.format_params()
        self._fields_prepared = False

    def format_params(self):
        tp = self.task_params
        self.sfc_field = tp.get("sfc_field", "")
        self.sfc_field1 = tp.get("sfc_field1", "") or self.sfc_field
        self.sfc_field2 = tp.get("sfc_field2", "") or self.sfc_field
        self.random = tp.get("random", None)
        self.random1 = tp.get("random1", None) or self.random
        self.random2 = tp.get("random2", None) or self.random
        self.weight_normalization = normalize_task_weight_normalization(tp.get("weight_normalization", "catalog"))
        window = tp.get("window", None)
        self.window = window if (window and window.get("type")) else None
        for i in (1, 2):
            wi = tp.get(f"window{i}", None)
            wi = wi if (wi and wi.get("type")) else None
            setattr(self, f"window{i}", wi or (dict(self.window) if self.window else None))
        self.binning_window = normalize_binning_window_template(tp.get("binning_window", default_binning_window()))
        self.sampling = copy.deepcopy(tp["sampling"])
        self.threads = int(tp["threads"])
        self.products = normalize_products(tp.get("products", "xi"))
        self.multipoles = tp.get("multipoles", None)
        self.memory_strategy = str(tp.get("memory_strategy", "speed")).strip().lower()
        if self.memory_strategy not in ("speed", "memory"):
            raise ValueError("Corr_2PCF memory_strategy must be either 'speed' or 'memory'.")
        self.fout_path = tp["fout_path"]

    # ------------------------------------------------------------ inputs
    def _resolve_sampling(self):
        """corr2pcf.py:453-473: the mapping names the sampling axes (a callable mapping takes the
        sampling dict's own keys)."""
        if not isinstance(self.sampling, dict) or not self.sampling:
            raise ValueError("Corr_2PCF requires a non-empty 'sampling' dictionary.")
        mapping = self.binning_window.get("mapping")
        if isinstance(mapping, str):
            if mapping not in BINNING_MAPPINGS:
                raise ValueError(f"Unsupported binning_window mapping '{mapping}'. Supported built-in mappings: "
                                 f"{sorted(BINNING_MAPPINGS)}.")
            required, self._mapper = BINNING_MAPPINGS[mapping]
        elif callable(mapping):
            required, self._mapper = tuple(self.sampling.keys()), mapping
        else:
            raise TypeError("binning_window mapping must be a string or callable.")
        missing = [n for n in required if n not in self.sampling]
        extra = [n for n in self.sampling if n not in required]
        if missing or extra:
            raise ValueError(f"sampling keys must match mapping '{mapping}': expected {list(required)}, "
                             f"missing={missing}, extra={extra}.")
        self.sampling_names = tuple(required)
        self.sampling_arrays = {n: normalize_sampling_spec(n, self.sampling[n]) for n in self.sampling_names}
        self.sampling_shape = tuple(self.sampling_arrays[n].size for n in self.sampling_names)
        if self.multipoles and self.sampling_names != ("s",):
            raise ValueError("Corr_2PCF multipoles need the 's' sampling (mapping 's_to_R').")

    def _load(self, value, label):
        if isinstance(value, SFCField):
            return value
        if isinstance(value, str) and value:
            return SFCField(data_path=value, threads=self.threads)
        self.logger.error(f"Missing input for {label}.")
        func_util.safe_exit(1)

    def _field_in_task_normalization(self, field):
        self._density_scale = float(field.scale_factor) ** 3
        if self.weight_normalization == "unit":
            return field.with_normalization("unit")
        if getattr(field, "field_kind", None) != "catalog_field":
            return field.copy()
        return field.switch_weight_normalization(self.weight_normalization)

    def _window_obj(self, win, field):
        if win is None:
            return None
        if isinstance(win, WindowFunc):
            return win
        return WindowFunc(win, field.sfc_info, threads=self.threads)

    def prepare_input_fields(self, sfc_field1=None, sfc_field2=None, random1=None, random2=None, window1=None,
                             window2=None, binning_window=None, sync_runtime=True):
        self.corr2pcf_data = Corr2PCFData(threads=self.threads)
        if binning_window is not None:
            self.binning_window = normalize_binning_window_template(binning_window)
        self._resolve_sampling()
        expanded = expand_products(self.products)
        needs_data = any(PRODUCT_INPUT_FLAGS[p][0] for p in expanded)
        needs_random = any(PRODUCT_INPUT_FLAGS[p][1] for p in expanded)
        inputs = {1: (sfc_field1 or self.sfc_field1, random1 if random1 is not None else self.random1,
                      window1 or self.window1),
                  2: (sfc_field2 or self.sfc_field2, random2 if random2 is not None else self.random2,
                      window2 or self.window2)}
        self.fields, self.randoms = {}, {}
        for i, (fsrc, rsrc, win) in inputs.items():
            base = self._field_in_task_normalization(self._load(fsrc, f"field vertex {i}"))
            w = self._window_obj(win, base)
            final = base @ w if w is not None else base
            final.format_sfc_params()
            self.fields[i] = final
            setattr(self.corr2pcf_data, f"sfc_info{i}", final.sfc_info)
            if needs_random:
                if rsrc in (None, ""):
                    self.logger.error(f"Missing input for random vertex {i}. Products {self.products} require "
                                      f"'random{i}' or shared 'random'.")
                    func_util.safe_exit(1)
                if rsrc == "uniform":
                    rho = base.field_mean_density(value_unit="grid")
                    if rho is None:
                        raise ValueError("random='uniform' requires a SFCField field with a defined field_integral.")
                    self.randoms[i] = float(rho)
                else:
                    rb = self._field_in_task_normalization(self._load(rsrc, f"random vertex {i}"))
                    rw = self._window_obj(win, rb)
                    self.randoms[i] = rb @ rw if rw is not None else rb
        if not needs_data:
            self.fields = {}
        self._fields_prepared = True

    # ------------------------------------------------------------ products
    def _descs(self, L, box):
        """One descriptor per sample, in np.ndindex order over the sampling axes, then the multipoles."""
        bw = {k: v for k, v in self.binning_window.items() if k != "mapping"}
        descs = []
        for idx in np.ndindex(self.sampling_shape):
            sample = {n: float(self.sampling_arrays[n][i]) for n, i in zip(self.sampling_names, idx)}
            params = self._mapper(sample, bw)
            if not isinstance(params, dict):
                raise TypeError("binning_window mapping must return a binning_window dictionary.")
            descs.append(binning_desc(params, L, box))
        ells = list(self.multipoles or [])
        s = self.sampling_arrays.get("s")
        for l in ells:
            if int(l) % 2:
                raise ValueError("Corr_2PCF multipoles supports even l only (xi_l of an auto-correlation is real).")
            for sv in s:
                descs.append(kernels.window_desc("shell_multipole", float(sv) * L / box, l=int(l), m=0))
        return descs, ells

    def _pair_products(self, a, b, descs, cache):
        return pair_products_with_descs(a, b, descs, cache)

    def run(self, save_result=True, overwrite=False):
        t0 = time.perf_counter()
        if not self._fields_prepared:
            self.prepare_input_fields()
        expanded = expand_products(self.products)
        ref = next(iter(self.fields.values())) if self.fields else next(
            v for v in self.randoms.values() if not _is_density(v))
        L, box = int(ref.L), float(ref.box_size)
        descs, ells = self._descs(L, box)
        ns = int(np.prod(self.sampling_shape))
        cache = {}
        delta = {}
        if "delta_dd" in expanded:
            delta[1] = self.fields[1] - self.randoms[1]
            same = self.fields[1] is self.fields[2] or (self.sfc_field1 is self.sfc_field2 and
                                                         self.randoms[1] == self.randoms[2])
            delta[2] = delta[1] if same else self.fields[2] - self.randoms[2]
        out = {}
        scale = float(ref.scale_factor) ** 6
        pairs = {"dd": ("fields", 1, "fields", 2), "dr": ("fields", 1, "randoms", 2),
                 "rd": ("randoms", 1, "fields", 2), "rr": ("randoms", 1, "randoms", 2)}
        for prod in ("dd", "dr", "rd", "rr"):
            if prod in expanded:
                s1, i1, s2, i2 = pairs[prod]
                a, b = getattr(self, s1)[i1], getattr(self, s2)[i2]
                out[prod] = self._pair_products(a, b, descs, cache) * scale
        if "delta_dd" in expanded:
            out["delta_dd"] = self._pair_products(delta[1], delta[2], descs, cache) * scale
        data = self.corr2pcf_data
        data.set_sampling(self.sampling_names, self.sampling_arrays)
        for p in PRODUCT_NAMES:
            v = out.get(p)
            setattr(data, p, None if v is None else v[:ns].reshape(self.sampling_shape).copy())
        if "xi" in expanded:
            data.xi = data.delta_dd / data.rr
        if ells:
            data.multipoles = np.asarray(ells)
            rr0 = out["rr"][:ns] if "rr" in out else None
            dd_l = out["delta_dd"][ns:].reshape(len(ells), ns)
            xi_l = np.empty((len(ells), ns))
            for j, l in enumerate(ells):
                xi_l[j] = ((-1) ** (l // 2)) * np.sqrt(4 * np.pi * (2 * l + 1)) * dd_l[j] / rr0
            data.xi_l = xi_l
        data.corr2pcf_info = {"sampling": {k: v.tolist() for k, v in self.sampling_arrays.items()},
                              "products": list(self.products), "multipoles": ells,
                              "binning_window": {k: v for k, v in self.binning_window.items() if k != "func"}}
        if save_result and self.fout_path:
            data.save_corr2pcf(self.fout_path, overwrite=overwrite)
        self.logger.info(f"The time for task: {time.perf_counter() - t0:.4f} sec")
        return data
