"""Corr_3PCF_Multipole: spherical-harmonic 3PCF multipoles on the HIP device.

Drop-in for pyhermes.theory.Corr_3PCF_Multipole (serial execution mode):
parameters, sampling ('grid' / 'paired', corr3pcf_multipole.py:229-322),
products ddd_l / rrr_l / delta_ddd_l / zeta_l (:944-952, :1363-1395, :1554-1589)
and result container follow the reference.  The per-sample loop of
calc_DDD_multipole is replaced by:

* summation_backend 'gpu' (default): the batched FP32 path -- every distinct
  r12 and r13 radius is evaluated at once (one fused window launch, one batched
  inverse FFT and one MFMA Gram contraction per (l, m)), and each sample picks
  its (r12, r13) entry;
* summation_backend 'gpu_fp64': the plain FP64 per-sample restatement
  (two window builds, four FFTs and the FP64 contraction per (l, m));
* summation_backend 'cpu' is accepted for compatibility and runs 'gpu_fp64'
  (this package has no CPU path).

Edge binning windows take the reference's radial profiles: thin shells (closed
form), the built-in sphere / thick_shell / gaussian / gaussian_shell profiles and
custom real- or k-space profiles (Hankel tables, utils/radial_profiles.py, with the
reference's table diagnostics recorded in corr3pcf_multipole_info).  The reference's
pair_mpi / sample_mpi modes distribute samples over MPI ranks; one process per
GPU runs them as 'serial'.
"""

import copy
import itertools
import time

import numpy as np
import torch

from .. import kernels, multipole
from ..io import Corr3PCFMultipoleData, SFCField, WindowFunc, normalize_task_weight_normalization
from ..pipeline import TaskBase
from ..utils import func_util
from ..edges import EdgeProfile
from ..utils.radial_profiles import diagnose_radial_multipole_table, is_thin_shell, validate_radial_profile_request
from ..utils.special_functions import solve_multipoles_from_ratio

ALLOWED = {"ddd_l", "rrr_l", "delta_ddd_l", "zeta_l"}
INPUT_FLAGS = {"ddd_l": (True, False), "rrr_l": (False, True), "delta_ddd_l": (True, True), "zeta_l": (True, True)}


def _is_density(x):
    return isinstance(x, (float, int, np.floating))


class Corr_3PCF_Multipole(TaskBase):
    def __init__(self, param_task=None):
        if param_task is None:
            param_task = {"Corr_3PCF_Multipole": {}}
        self.task_name = str(self.__class__.__name__)
        super().__init__(param_task=param_task)
        self.format_params()
        self._fields_prepared = False

    # ------------------------------------------------------------ parameters
    def format_params(self):
        tp = self.task_params
        for key in ("r12", "r13", "r1", "r2"):
            if key in tp:
                raise ValueError("Corr_3PCF_Multipole no longer accepts top-level radius keys "
                                 f"['{key}']. Define radii through sampling and map them into binning_window12/13.")
        self.sfc_field = tp.get("sfc_field", "")
        self.random = tp.get("random", None)
        for i in (1, 2, 3):
            setattr(self, f"sfc_field{i}", tp.get(f"sfc_field{i}", "") or self.sfc_field)
            r = tp.get(f"random{i}", None)
            setattr(self, f"random{i}", self.random if r in (None, "") else r)
        self.weight_normalization = normalize_task_weight_normalization(tp.get("weight_normalization", "catalog"))
        window = tp.get("window", None)
        self.window = window if (isinstance(window, dict) and window.get("type")) else None
        for i in (1, 2, 3):
            wi = tp.get(f"window{i}", None)
            wi = wi if (isinstance(wi, dict) and wi.get("type")) else None
            setattr(self, f"window{i}", wi or (dict(self.window) if self.window else None))
        self.binning_window12_template = self._edge_template(tp.get("binning_window12"), "binning_window12")
        self.binning_window13_template = self._edge_template(tp.get("binning_window13"), "binning_window13")
        self.sampling = self._normalize_sampling(tp.get("sampling"))
        self.samples = self._expand_sampling(self.sampling)
        names = list(self.samples[0])
        self.sample_params = {n: np.asarray([s[n] for s in self.samples]) for n in names}
        self.binning_window12 = [self._sample_window(self.binning_window12_template, s, "binning_window12")
                                 for s in self.samples]
        self.binning_window13 = [self._sample_window(self.binning_window13_template, s, "binning_window13")
                                 for s in self.samples]
        self.r12 = self.sample_params.get("r12")
        self.r13 = self.sample_params.get("r13")
        self.l_min, self.l_max = int(tp["l_min"]), int(tp["l_max"])
        if self.l_min < 0 or self.l_max < 0 or self.l_min > self.l_max:
            raise ValueError("Corr_3PCF_Multipole requires 0 <= l_min <= l_max.")
        for name, wins in (("binning_window12", self.binning_window12), ("binning_window13", self.binning_window13)):
            for k, w in enumerate(wins):
                try:
                    validate_radial_profile_request(w["type"], w["len_args"], w.get("other_args", {}))
                except (TypeError, ValueError, KeyError) as exc:
                    raise ValueError(f"Invalid {name} for sample {k}: {exc}") from exc
        backend = str(tp.get("summation_backend", "gpu")).strip().lower()
        if backend not in ("gpu", "gpu_fp64", "cpu"):
            raise ValueError("summation_backend must be either 'gpu', 'gpu_fp64' or 'cpu'.")
        if backend == "cpu":
            self.logger.warning("summation_backend='cpu': pyhermes_amd has no CPU path; using the FP64 GPU path.")
            backend = "gpu_fp64"
        self.summation_backend = backend
        self.execution_mode = str(tp["execution_mode"]).strip().lower()
        if self.execution_mode not in {"serial", "pair_mpi", "sample_mpi"}:
            raise ValueError("Corr_3PCF_Multipole execution_mode must be 'serial', 'pair_mpi', or 'sample_mpi'.")
        self.zeta_condition_warning = float(tp.get("zeta_condition_warning", 1e12))
        self.radial_profile_diagnostics_enabled = bool(tp.get("radial_profile_diagnostics", True))
        self.radial_profile_diagnostic_tolerance = float(tp.get("radial_profile_diagnostic_tolerance", 1.0e-5))
        if not np.isfinite(self.radial_profile_diagnostic_tolerance) or self.radial_profile_diagnostic_tolerance <= 0:
            raise ValueError("radial_profile_diagnostic_tolerance must be finite and positive.")
        self.radial_profile_diagnostic_probes = int(tp.get("radial_profile_diagnostic_probes", 33))
        if self.radial_profile_diagnostic_probes < 3:
            raise ValueError("radial_profile_diagnostic_probes must be at least 3.")
        self.radial_profile_diagnostics = []
        self.threads = int(tp["threads"])
        self.products = self._normalize_products(tp.get("products", "zeta_l"))
        self.fout_path = tp["fout_path"]
        self.rho = None

    def _edge_template(self, value, name):
        if not isinstance(value, dict):
            raise TypeError(f"{name} must be a dictionary.")
        if not value.get("type"):
            raise ValueError(f"{name} must define a non-empty 'type'.")
        p = copy.deepcopy(value)
        p["type"] = str(p["type"]).strip().lower()
        for k in ("len_args", "other_args", "mapping"):
            if p.get(k) is None:
                p[k] = {}
            if not isinstance(p[k], dict):
                raise TypeError(f"{name}.{k} must be a dictionary.")
        return p

    def _normalize_sampling(self, sampling):
        if not isinstance(sampling, dict):
            raise TypeError("Corr_3PCF_Multipole requires a 'sampling' dictionary.")
        s = copy.deepcopy(sampling)
        mode = str(s.pop("mode", "grid")).strip().lower()
        if mode not in {"grid", "paired"}:
            raise ValueError("sampling.mode must be 'grid' or 'paired'.")
        values = {}
        for name, spec in s.items():
            if isinstance(spec, dict):
                if "values" in spec:
                    arr = np.asarray(spec["values"], dtype=np.float64)
                elif {"min", "max", "n"}.issubset(spec):
                    arr = np.linspace(float(spec["min"]), float(spec["max"]), int(spec["n"]))
                elif {"start", "stop", "step"}.issubset(spec):
                    step = float(spec["step"])
                    if step == 0.0:
                        raise ValueError(f"sampling.{name}.step must be non-zero.")
                    arr = np.arange(float(spec["start"]), float(spec["stop"]) + 0.5 * step, step)
                else:
                    raise ValueError(f"sampling.{name} must use 'values', 'min/max/n', 'start/stop/step', or a scalar.")
            else:
                arr = np.atleast_1d(np.asarray(spec, dtype=np.float64))
            if arr.ndim != 1 or arr.size == 0:
                raise ValueError(f"sampling.{name} must expand to a non-empty 1D array.")
            values[name] = arr
        if not values:
            raise ValueError("sampling must contain at least one sampled parameter.")
        return {"mode": mode, "values": values}

    def _expand_sampling(self, sampling):
        vals = sampling["values"]
        names = list(vals)
        if sampling["mode"] == "grid":
            return [{n: float(v) for n, v in zip(names, combo)} for combo in itertools.product(*(vals[n] for n in names))]
        n = max(a.size for a in vals.values())
        if any(a.size not in (1, n) for a in vals.values()):
            raise ValueError("sampling.mode='paired' requires each parameter to have length 1 or the shared length.")
        return [{k: float(a[0] if a.size == 1 else a[i]) for k, a in vals.items()} for i in range(n)]

    def _sample_window(self, template, sample, name):
        p = copy.deepcopy(template)
        mapping = p.pop("mapping")
        for target, source in mapping.items():
            if isinstance(source, str) and source not in sample:
                raise KeyError(f"Sampling parameter '{source}' is not defined.")
            value = float(sample[source]) if isinstance(source, str) else float(source)
            section, key = ("len_args", target) if "." not in target else target.split(".", 1)
            if section not in ("len_args", "other_args"):
                raise ValueError(f"Unsupported binning-window mapping target '{target}'.")
            p.setdefault(section, {})[key] = value
        if not p["len_args"]:
            raise ValueError(f"{name} produced an empty len_args dictionary.")
        return p

    def _normalize_products(self, products):
        if isinstance(products, str):
            products = [products]
        elif products is None:
            products = ["zeta_l"]
        out = []
        for item in products:
            name = str(item).strip().lower()
            if name not in ALLOWED:
                raise ValueError(f"Unsupported product '{item}'. Allowed values are {sorted(ALLOWED)}.")
            if name not in out:
                out.append(name)
        return out

    def _expanded_products(self):
        out = list(self.products)
        if "zeta_l" in out:
            for dep in ("delta_ddd_l", "rrr_l"):
                if dep not in out:
                    out.append(dep)
        return out

    # ------------------------------------------------------------ inputs
    def _load(self, value, label):
        if isinstance(value, SFCField):
            return value
        if isinstance(value, str) and value:
            return SFCField(data_path=value, threads=self.threads)
        self.logger.error(f"Missing input for {label}.")
        func_util.safe_exit(1)

    def _normalized(self, field):
        self._density_scale = float(field.scale_factor) ** 3
        if self.weight_normalization == "unit":
            return field.with_normalization("unit")
        if getattr(field, "field_kind", None) != "catalog_field":
            return field.copy()
        return field.switch_weight_normalization(self.weight_normalization)

    def _apply_window(self, field, win):
        if win is None:
            return field
        w = win if isinstance(win, WindowFunc) else WindowFunc(win, field.sfc_info, threads=self.threads)
        return field @ w

    def prepare_input_fields(self, sfc_field1=None, sfc_field2=None, sfc_field3=None, random1=None, random2=None,
                             random3=None, window1=None, window2=None, window3=None):
        self.corr3pcf_multipole_data = Corr3PCFMultipoleData()
        if "zeta_l" in self._expanded_products() and self.l_min > 0:
            self.logger.error("zeta_l requires l_min=0 because the ratio solve needs multipoles from l=0.")
            func_util.safe_exit(1)
        exp = self._expanded_products()
        needs_data = any(INPUT_FLAGS[p][0] for p in exp)
        needs_random = any(INPUT_FLAGS[p][1] for p in exp)
        given = {1: sfc_field1, 2: sfc_field2, 3: sfc_field3}
        rgiven = {1: random1, 2: random2, 3: random3}
        wgiven = {1: window1, 2: window2, 3: window3}
        srcs = {i: given[i] if given[i] is not None else getattr(self, f"sfc_field{i}") for i in (1, 2, 3)}
        wins = {i: wgiven[i] if wgiven[i] is not None else getattr(self, f"window{i}") for i in (1, 2, 3)}
        rsrcs = {i: rgiven[i] if rgiven[i] is not None else getattr(self, f"random{i}") for i in (1, 2, 3)}
        cache = {}
        self.fields, self.randoms, self.rho_vertices = {}, {}, [None, None, None]
        ref = None
        for i in (1, 2, 3):
            src = srcs[i]
            if src in (None, "") and not needs_data:
                continue
            key = (id(src), repr(wins[i]))
            if key not in cache:
                base = self._normalized(self._load(src, f"field vertex {i}"))
                cache[key] = (base, self._apply_window(base, wins[i]))
            base, final = cache[key]
            if ref is None:
                ref = base
            self.fields[i] = final
            self.rho_vertices[i - 1] = base.field_mean_density(value_unit="grid")
            setattr(self.corr3pcf_multipole_data, f"sfc_info{i}", final.sfc_info)
        if ref is None:
            self.logger.error("At least one SFCField input is required to define the grid geometry and shared density.")
            func_util.safe_exit(1)
        self.reference_sfc = ref
        self.rho = ref.field_mean_density(value_unit="grid")
        if needs_random:
            rcache = {}
            for i in (1, 2, 3):
                r = rsrcs[i]
                if r in (None, ""):
                    self.logger.error(f"Missing input for random vertex {i}. Use 'uniform' for the analytic uniform "
                                      "random field.")
                    func_util.safe_exit(1)
                if r == "uniform":
                    rho_i = float(self.rho_vertices[i - 1] if self.rho_vertices[i - 1] is not None else self.rho)
                    self.randoms[i] = self._uniform_after_window(rho_i, wins[i])
                else:
                    key = (id(r), repr(wins[i]))
                    if key not in rcache:
                        rcache[key] = self._apply_window(self._normalized(self._load(r, f"random vertex {i}")), wins[i])
                    self.randoms[i] = rcache[key]
        self._fields_prepared = True

    def _uniform_after_window(self, rho, win):
        """corr3pcf_multipole.py:607-619: W(0) = 1 keeps the scalar shortcut."""
        if win is None:
            return float(rho)
        w = win if isinstance(win, WindowFunc) else WindowFunc(win, self.reference_sfc.sfc_info, threads=self.threads)
        k0 = complex(w.device_kernel(True).reshape(-1)[0].item())
        if abs(k0.imag) <= 1e-12 * max(1.0, abs(k0.real)) and np.isclose(k0.real, 1.0, rtol=1e-10, atol=1e-12):
            return float(rho)
        u = self.reference_sfc._spawn_like()
        L = int(self.reference_sfc.L)
        u._set_device(torch.full((L, L, L), float(rho), dtype=torch.float64, device=kernels._dev()))
        u.sfc_info["field_integral"] = float(rho) * L ** 3
        u.sfc_info["field_kind"] = "derived_field"
        u.format_sfc_params()
        return u @ w

    # ------------------------------------------------------------ products
    def _materialize(self, v):
        if _is_density(v):
            L = int(self.reference_sfc.L)
            return torch.full((L, L, L), float(v), dtype=torch.float64, device=kernels._dev())
        return v.device_epsilon(torch.float64)

    def _product_tensors(self, product):
        if product == "ddd_l":
            src = [self.fields[i] for i in (1, 2, 3)]
        elif product == "delta_ddd_l":
            src = [self.fields[i] - self.randoms[i] for i in (1, 2, 3)]
            if self.fields[2] is self.fields[3] and self.randoms[2] == self.randoms[3]:
                src[2] = src[1]
            if self.fields[1] is self.fields[2] and self.randoms[1] == self.randoms[2]:
                src[0] = src[1]
        else:
            src = [self.randoms[i] for i in (1, 2, 3)]
        tens, seen = [], {}
        for s in src:
            if id(s) not in seen:
                seen[id(s)] = self._materialize(s)
            tens.append(seen[id(s)])
        return tens

    def _edge_list(self, windows, L, box):
        """Per sample: the thin-shell radius in grid units, or an EdgeProfile (corr3pcf_multipoles.py:231-252
        turns each edge window into a radial_multipole window of its profile)."""
        out, cache = [], {}
        for w in windows:
            if w["type"] == "shell":
                out.append(float(w["len_args"]["R"]) * L / box)
                continue
            edge = EdgeProfile(w["type"], w["len_args"], w.get("other_args", {}), L, box)
            out.append(cache.setdefault(edge.key, edge))
        return out

    @staticmethod
    def _unique_edges(edges):
        """Distinct edges in a stable order and each sample's index into them."""
        keys, index, uniq = {}, [], []
        for e in edges:
            k = e.key if isinstance(e, EdgeProfile) else ("shell", float(e))
            if k not in keys:
                keys[k] = len(uniq)
                uniq.append(e)
            index.append(keys[k])
        if all(not isinstance(e, EdgeProfile) for e in uniq):
            order = np.argsort(np.asarray(uniq, dtype=np.float64), kind="stable")
            rank = np.empty_like(order)
            rank[order] = np.arange(order.size)
            return np.asarray(uniq, dtype=np.float64)[order], rank[np.asarray(index)]
        return uniq, np.asarray(index)

    def _compute(self, product):
        d1, d2, d3 = self._product_tensors(product)
        L = int(self.reference_sfc.L)
        box = float(self.reference_sfc.box_size)
        geo = self._geometry(L)
        e12 = self._edge_list(self.binning_window12, L, box)
        e13 = self._edge_list(self.binning_window13, L, box)
        n_l = self.l_max - self.l_min + 1
        rows = np.empty((len(self.samples), n_l))
        if self.summation_backend == "gpu":
            u12, i12 = self._unique_edges(e12)
            u13, i13 = self._unique_edges(e13)
            u12, u13 = self._profiles_or_radii(u12, L, box), self._profiles_or_radii(u13, L, box)
            f1 = d1.float()
            f2 = d2.float()
            f3 = f2 if d3 is d2 else d3.float()
            res = multipole.ddd_multipole_batched_f32(geo, f1, f2, f3, u12, u13, self.l_min, self.l_max)
            for k in range(len(self.samples)):
                rows[k] = res[i12[k], i13[k]]
        else:
            for k in range(len(self.samples)):
                rows[k] = multipole.ddd_multipole_plain_f64(geo, d1, d2, d3, e12[k], e13[k], self.l_min, self.l_max)
        return rows

    @staticmethod
    def _profiles_or_radii(uniq, L, box):
        """Radii stay a float64 array; a mixed list becomes all EdgeProfile (thin shells as 'shell')."""
        if isinstance(uniq, np.ndarray):
            return uniq
        return [e if isinstance(e, EdgeProfile) else EdgeProfile("shell", {"R": float(e) * box / L}, {}, L, box)
                for e in uniq]

    def _record_radial_profile_diagnostics(self):
        """corr3pcf_multipole.py:668-727: each tabulated edge table checked against a refined quadrature (and
        its k = 0 value); failures are logged as warnings and kept in the result's info."""
        self.radial_profile_diagnostics = []
        if not self.radial_profile_diagnostics_enabled:
            return
        L, box = int(self.reference_sfc.L), float(self.reference_sfc.box_size)
        k_max = float(np.sqrt(3.0) * L / box)
        for edge_name, windows in (("binning_window12", self.binning_window12),
                                   ("binning_window13", self.binning_window13)):
            for k, w in enumerate(windows):
                t = str(w["type"]).strip().lower()
                if is_thin_shell(t, w["len_args"]):
                    continue
                first = self.l_min if t.startswith("custom_") else max(1, self.l_min)
                for l in range(first, self.l_max + 1):
                    d = diagnose_radial_multipole_table(t, w["len_args"], l, k_max,
                                                        profile_config=w.get("other_args", {}),
                                                        tolerance=self.radial_profile_diagnostic_tolerance,
                                                        probe_count=self.radial_profile_diagnostic_probes)
                    d.update({"edge": edge_name, "sample_idx": int(k)})
                    self.radial_profile_diagnostics.append(d)
                    if not d["passed"]:
                        self.logger.warning(f"Radial profile diagnostic exceeded tolerance | {edge_name}[{k}], "
                                            f"type={t}, l={l}, tolerance={self.radial_profile_diagnostic_tolerance:.3e}."
                                            " Increase profile quadrature/table points or verify the profile support.")

    def _geometry(self, L):
        if getattr(self, "_geo", None) is None or self._geo.L != L:
            phi = getattr(self.reference_sfc, "phi_array", None)
            if phi is None:
                from ..utils.wavelet_grid import sample_scaling_function
                phi = sample_scaling_function(self.reference_sfc.wavelet_mode, self.reference_sfc.wavelet_level)
            self._geo = multipole.MultipoleGeometry(np.asarray(phi), L, int(self.reference_sfc.phi_resolution))
        return self._geo

    def _store(self, name, values):
        data = self.corr3pcf_multipole_data
        data.sample_params = copy.deepcopy(self.sample_params)
        data.binning_window12 = copy.deepcopy(self.binning_window12)
        data.binning_window13 = copy.deepcopy(self.binning_window13)
        data.r12, data.r13 = self.r12, self.r13
        data.l = np.arange(self.l_min, self.l_max + 1, dtype=np.int32)
        values = np.asarray(values, dtype=np.float64)
        if values.ndim == 1:
            values = values.reshape(1, -1)
        factor = getattr(self, "_density_scale", 1.0) ** 3
        setattr(data, name, values * factor if not np.isclose(factor, 1.0) else values)

    def run(self, save_result=True, overwrite=False):
        t0 = time.perf_counter()
        if not self._fields_prepared:
            self.prepare_input_fields()
        exp = self._expanded_products()
        self._record_radial_profile_diagnostics()
        for product in exp:
            if product == "zeta_l":
                continue
            if product == "rrr_l" and all(_is_density(self.randoms[i]) for i in (1, 2, 3)):
                l_arr = np.arange(self.l_min, self.l_max + 1)
                row = np.zeros(l_arr.size)
                if 0 in l_arr:
                    row[0] = float(self.randoms[1]) * float(self.randoms[2]) * float(self.randoms[3])
                self._store(product, np.repeat(row[None, :], len(self.samples), axis=0))
                continue
            self._store(product, self._compute(product))
        data = self.corr3pcf_multipole_data
        if "zeta_l" in exp:
            zs, conds = [], []
            for k in range(data.delta_ddd_l.shape[0]):
                z, _, c = solve_multipoles_from_ratio(data.delta_ddd_l[k], data.rrr_l[k], self.l_max)
                zs.append(z)
                conds.append(c)
            data.zeta_l = np.asarray(zs)
            data.zeta_condition = np.asarray(conds)
            bad = np.where(data.zeta_condition > self.zeta_condition_warning)[0]
            if bad.size:
                self.logger.warning(f"zeta_l mixing matrices exceed zeta_condition_warning for samples {bad.tolist()}.")
        data.corr3pcf_multipole_info = {"l_min": self.l_min, "l_max": self.l_max, "n_samples": len(self.samples),
                                        "summation_backend": self.summation_backend,
                                        "products": list(self.products), "rho_vertices": self.rho_vertices,
                                        "radial_profile_diagnostics_result": copy.deepcopy(
                                            self.radial_profile_diagnostics)}
        if save_result and self.fout_path:
            data.save_corr3pcf_multipole(self.fout_path, overwrite=overwrite)
        self.logger.info(f"The time for task: {time.perf_counter() - t0:.4f} sec")
        return data
