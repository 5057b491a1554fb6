"""Corr_3PCF: real-space 3PCF by Monte-Carlo triangles around box-random or particle centres.

Drop-in for pyhermes.theory.Corr_3PCF (corr3pcf.py
:154-203 parameters, :288-322 angle sampling, :1105-1132 per-mu products,
:1183-1215 pair cache, :1400-1620 run).  For every mu the reference loops over
n_rot rotations and all centres in numba; here one HIP launch
(hermes_mc_triplet_box) evaluates n1(c) n2(c + d12) n3(c + d13) for every
(centre, rotation) pair with the same rotation draws:

* centres: numpy default_rng(base_seed + 1000003).uniform(0, L, (n_box_centers, 3))
  (sampling.py:6-10; the single-process rank-0 seed of corr3pcf.py:1415-1416);
* rotation irot of angle index it: numpy RandomState(base_seed + 1 + irot +
  it * 1000003).rand(3) (corr3pcf_kernels.py:119-125, corr3pcf.py:1462).

Products ddd, rrr, delta_ddd, xi12, xi13, xi23, zeta, zeta_H, Q follow the
reference's dependency rules; triplet products are scaled by (L / box)^9 and
the pair caches use thin-shell windows evaluated in one fused pass.

center='particle' (corr3pcf.py:1135-1180, :1405-1460, :1596-1614; the estimator of
corr3pcf_kernels.py:66-97): vertex 1 sits on the catalogue's own points with their
weights, rho1 sum_rot sum_i w_i n2(c_i + d12) n3(c_i + d13) / (n_rot sum_i w_i),
rho1 = sum_i w_i / L^3 -- the same HIP launch with the weights in place of n1.
Products add d_delta_dd (data centres on delta2, delta3) and r_delta_dd (random
centres; rrr * xi23 when random1 is a uniform density and no random_pos1 is
given), and delta_ddd = d_delta_dd - r_delta_dd.  Centres come as particle_pos1 /
particle_weight1 (and random_pos1 / random_weight1) in box units, or from the
particle files SFCProjection saved: sfc_field1's for the data centres, random1's
for the random centres of r_delta_dd (SFCField.get_particle_data); a random1 field
without one ends the task.  rrr without r_delta_dd samples on the data centres
(corr3pcf.py:1527-1535 passes them when no random centres were resolved).
"""

import copy
import time

import numpy as np
import torch

from .. import kernels
from ..io import Corr3PCFData, SFCField, WindowFunc, normalize_task_weight_normalization
from ..pipeline import TaskBase
from ..utils import func_util
from .corr2pcf import shell_pair_products

PRODUCT_RULES = {
    "box_random": {
        "allowed": {"ddd", "rrr", "delta_ddd", "xi12", "xi13", "xi23", "zeta", "zeta_H", "Q"},
        "deps": {"Q": ["zeta", "zeta_H"], "zeta_H": ["xi12", "xi13", "xi23"], "zeta": ["delta_ddd", "rrr"]},
    },
    "particle": {
        "allowed": {"ddd", "rrr", "d_delta_dd", "r_delta_dd", "delta_ddd", "xi12", "xi13", "xi23", "zeta", "zeta_H",
                    "Q"},
        "deps": {"Q": ["zeta", "zeta_H"], "zeta_H": ["xi12", "xi13", "xi23"], "zeta": ["delta_ddd", "rrr"],
                 "delta_ddd": ["d_delta_dd", "r_delta_dd"]},
    },
}
PRODUCT_INPUT_FLAGS = {"ddd": (True, False), "rrr": (False, True), "d_delta_dd": (True, True),
                       "r_delta_dd": (True, True), "delta_ddd": (True, True), "xi12": (True, True),
                       "xi13": (True, True), "xi23": (True, True), "zeta": (True, True), "zeta_H": (True, True),
                       "Q": (True, True)}
ROTATION_SEED_STRIDE = 1000003
CENTRE_SEED_OFFSET = 1000003


def third_side_from_mu(r1, r2, mu):
    """corr3pcf_kernels.py:16-18."""
    return np.sqrt(np.clip(r1 ** 2 + r2 ** 2 - 2.0 * r1 * r2 * mu, 0.0, None))


def rotation_draws(seed_base_rot, n_rot, mu_index):
    """(n_rot, 3) uniforms: one MT19937 RandomState(seed).rand(3) per rotation."""
    out = np.empty((n_rot, 3))
    for irot in range(n_rot):
        s = seed_base_rot + irot + (mu_index * ROTATION_SEED_STRIDE if mu_index >= 0 else 0)
        out[irot] = np.random.RandomState(s).rand(3)
    return out


def _is_density(x):
    return isinstance(x, (float, int, np.floating))


class Corr_3PCF(TaskBase):
    def __init__(self, param_task=None):
        if param_task is None:
            param_task = {"Corr_3PCF": {}}
        self.task_name = str(self.__class__.__name__)
        super().__init__(param_task=param_task)
        self.format_params()
        self._fields_prepared = False

    # ------------------------------------------------------------ parameters
    def format_params(self):
        tp = self.task_params
        self.sfc_field = tp.get("sfc_field", "")
        self.random = tp.get("random", None)
        for i in (1, 2, 3):
            setattr(self, f"sfc_field{i}", tp.get(f"sfc_field{i}", "") or self.sfc_field)
            r = tp.get(f"random{i}", None)
            setattr(self, f"random{i}", self.random if r in (None, "") else r)
        self.weight_normalization = normalize_task_weight_normalization(tp.get("weight_normalization", "catalog"))
        window = tp.get("window", None)
        self.window = window if (window and (window.get("type") or window.get("func"))) else None
        for i in (1, 2, 3):
            wi = tp.get(f"window{i}", None)
            wi = wi if (wi and (wi.get("type") or wi.get("func"))) else None
            setattr(self, f"window{i}", wi or (dict(self.window) if self.window else None))
        self.r12 = float(tp["r12"])
        self.r13 = float(tp["r13"])
        self.angle_param = tp["angle_param"]
        self.theta = copy.deepcopy(tp.get("theta"))
        self.mu = copy.deepcopy(tp.get("mu"))
        self.n_rot = int(tp["n_rot"])
        self.center = str(tp.get("center", "box_random")).strip().lower()
        if self.center not in ("box_random", "particle"):
            raise ValueError(f"Unknown center='{self.center}'. Use 'box_random' or 'particle'.")
        for key in ("particle_pos1", "particle_weight1", "random_pos1", "random_weight1"):
            setattr(self, key, tp.get(key, None))
        self.n_box_centers = int(tp["n_box_centers"])
        self.base_seed = int(tp["base_seed"])
        self.threads = int(tp["threads"])
        self.products = self._normalize_products(tp.get("products", "Q"))
        self.fout_path = tp["fout_path"]
        self.rho = None

    def _normalize_products(self, products):
        if isinstance(products, str):
            products = [products]
        elif products is None:
            products = ["Q"]
        elif not isinstance(products, (list, tuple, set)):
            raise TypeError(f"Unsupported products input: expected string or array of strings, got {type(products)}.")
        out = []
        for item in products:
            if not isinstance(item, str):
                raise TypeError("Each product name must be a string.")
            raw = item.strip()
            name = "Q" if raw.upper() == "Q" else ("zeta_H" if raw.lower() == "zeta_h" else raw.lower())
            if name not in PRODUCT_RULES[self.center]["allowed"]:
                raise ValueError(f"Unsupported product '{item}' for center='{self.center}'. Allowed values are "
                                 f"{sorted(PRODUCT_RULES[self.center]['allowed'])}.")
            if name not in out:
                out.append(name)
        return out

    def _expanded_products(self):
        out = list(self.products)
        i = 0
        while i < len(out):
            for dep in PRODUCT_RULES[self.center]["deps"].get(out[i], []):
                if dep not in out:
                    out.append(dep)
            i += 1
        return out

    def _resolve_angle_sampling(self):
        """corr3pcf.py:288-322."""
        ap = str(self.angle_param).strip().lower()
        if ap not in {"theta", "mu"}:
            raise ValueError(f"Unsupported angle_param='{self.angle_param}'. Use 'theta' or 'mu'.")
        if ap == "theta":
            spec = self.theta
            if spec is None:
                raise ValueError("angle_param='theta' requires 'theta' to be provided as a dict or array-like input.")
            if isinstance(spec, dict):
                theta = np.linspace(float(spec.get("theta_min", 0.0)), float(spec.get("theta_max", np.pi)),
                                    int(spec.get("n_theta", 20)))
            else:
                theta = self._monotonic(spec, "theta", 0.0, float(np.pi))
            mu = np.cos(theta)
        else:
            spec = self.mu
            if spec is None:
                raise ValueError("angle_param='mu' requires 'mu' to be provided as a dict or array-like input.")
            if isinstance(spec, dict):
                mu = np.linspace(float(spec.get("mu_min", -1.0)), float(spec.get("mu_max", 1.0)), int(spec.get("n_mu", 20)))
            else:
                mu = self._monotonic(spec, "mu", -1.0, 1.0)
            theta = np.arccos(np.clip(mu, -1.0, 1.0))
        self.angle_param = ap
        self.theta_arr = np.ascontiguousarray(theta, dtype=np.float64)
        self.mu_arr = np.ascontiguousarray(mu, dtype=np.float64)

    @staticmethod
    def _monotonic(values, name, vmin, vmax):
        arr = np.asarray(values, dtype=np.float64)
        if arr.ndim != 1:
            raise TypeError(f"'{name}' must be a 1D array-like input.")
        if arr.size == 0:
            raise ValueError(f"'{name}' must contain at least one sampling point.")
        if np.any(arr < vmin) or np.any(arr > vmax):
            raise ValueError(f"'{name}' values must lie in [{vmin}, {vmax}].")
        d = np.diff(arr)
        if np.any(d < 0.0) and np.any(d > 0.0):
            raise ValueError(f"'{name}' values must be monotonic.")
        return np.ascontiguousarray(arr)

    # ------------------------------------------------------------ inputs
    def _load(self, value, label, cache):
        if isinstance(value, SFCField):
            return value
        if isinstance(value, str) and value:
            if value not in cache:
                cache[value] = SFCField(data_path=value, threads=self.threads)
            return cache[value]
        self.logger.error(f"Missing usable input for {label}. Expected a string path or SFCField instance.")
        func_util.safe_exit(1)

    def _normalized(self, field):
        self._density_scale = float(field.scale_factor) ** 3
        if self.weight_normalization == "unit":
            return field.with_normalization("unit")
        if getattr(field, "field_kind", None) != "catalog_field":
            return field.copy()
        return field.switch_weight_normalization(self.weight_normalization)

    def _windowed(self, field, win):
        if isinstance(win, WindowFunc):
            return field @ win
        if isinstance(win, dict):
            return field @ WindowFunc(win, field.sfc_info, threads=self.threads)
        if win is not None:
            self.logger.error(f"Unsupported window input. Expected dict, WindowFunc, or None, got {type(win)}.")
            func_util.safe_exit(1)
        return field

    def _centre_arrays(self, pos, weight, pos_name, weight_name):
        """corr3pcf.py:550-572: (N, 3) positions and a positive-sum weight array, given together."""
        if (pos is None) != (weight is None):
            raise ValueError(f"{pos_name} and {weight_name} must be provided together for center='particle'.")
        if pos is None:
            return None
        p = np.asarray(pos, dtype=np.float64)
        if p.ndim != 2 or p.shape[1] != 3:
            raise ValueError(f"{pos_name} must be an array-like with shape (N, 3), got shape {p.shape}.")
        w = np.asarray(weight, dtype=np.float64)
        if w.ndim != 1 or w.shape[0] != p.shape[0]:
            raise ValueError(f"{weight_name} must be a 1D array with length {p.shape[0]}, got shape {w.shape}.")
        if not float(np.sum(w)) > 0.0:
            raise ValueError(f"{weight_name} must have a positive sum.")
        return np.ascontiguousarray(p), np.ascontiguousarray(w)

    def prepare_input_fields(self, sfc_field1=None, sfc_field2=None, sfc_field3=None, random1=None, random2=None,
                             random3=None, window1=None, window2=None, window3=None, particle_pos1=None,
                             particle_weight1=None, random_pos1=None, random_weight1=None):
        self.corr3pcf_data = Corr3PCFData()
        if self.center == "particle":
            for key, val in (("particle_pos1", particle_pos1), ("particle_weight1", particle_weight1),
                             ("random_pos1", random_pos1), ("random_weight1", random_weight1)):
                if val is not None:
                    setattr(self, key, val)
            self._data_centres = self._centre_arrays(self.particle_pos1, self.particle_weight1, "particle_pos1",
                                                     "particle_weight1")
            self._random_centres = self._centre_arrays(self.random_pos1, self.random_weight1, "random_pos1",
                                                       "random_weight1")
            needs_centres = {"ddd", "d_delta_dd", "r_delta_dd"} & set(self._expanded_products())
            if self._data_centres is None and needs_centres:
                # corr3pcf.py:600-607: the centres of vertex 1's catalogue field, with its projection weights
                f1 = sfc_field1 if sfc_field1 is not None else self.sfc_field1
                if not isinstance(f1, SFCField):
                    raise ValueError("center='particle' needs particle_pos1 and particle_weight1, or sfc_field1 as a "
                                     "catalogue field with saved particle data.")
                pdata = self._normalized(f1).get_particle_data()
                self._data_centres = self._centre_arrays(pdata["pos"], pdata["projection_weight"],
                                                         "sfc_field1 particle positions", "sfc_field1 particle weight")
        self._resolve_angle_sampling()
        exp = self._expanded_products()
        needs_data = any(PRODUCT_INPUT_FLAGS[p][0] for p in exp)
        needs_random = any(PRODUCT_INPUT_FLAGS[p][1] for p in exp)
        src = {1: sfc_field1, 2: sfc_field2, 3: sfc_field3}
        rsrc = {1: random1, 2: random2, 3: random3}
        wsrc = {1: window1, 2: window2, 3: window3}
        for i in (1, 2, 3):
            src[i] = src[i] if src[i] is not None else getattr(self, f"sfc_field{i}")
            rsrc[i] = rsrc[i] if rsrc[i] is not None else getattr(self, f"random{i}")
            wsrc[i] = wsrc[i] if wsrc[i] is not None else getattr(self, f"window{i}")
        if needs_random and all(r in (None, "") for r in rsrc.values()):
            self.logger.error("Random-related products were requested, but no random inputs were provided. "
                              "Please set shared 'random' or vertex-specific 'random1/2/3'.")
            func_util.safe_exit(1)
        cache, built = {}, {}
        if self.center == "particle" and "r_delta_dd" in exp and self._random_centres is None:
            self._random_centres = self._catalogue_random_centres(rsrc[1], cache)
        self.fields, self.randoms = {}, {}
        compat = []
        if needs_data:
            for i in (1, 2, 3):
                base = self._load(src[i], f"field vertex {i}", cache)
                key = (id(base), repr(wsrc[i]))
                if key not in built:
                    nb = self._normalized(base)
                    built[key] = self._windowed(nb, wsrc[i])
                self.fields[i] = built[key]
                compat.append(base)
                setattr(self.corr3pcf_data, f"sfc_info{i}", self.fields[i].sfc_info)
        if needs_random:
            for i in (1, 2, 3):
                r = rsrc[i]
                if isinstance(r, str) and r.strip().lower() == "uniform":
                    ref = self.fields.get(i)
                    if ref is None:
                        ref = next(iter(self.fields.values()), None)
                    if ref is None:
                        self.logger.error(f"Cannot resolve the geometry for uniform random vertex {i}. Please provide "
                                          "at least one SFCField input field or an explicit random field.")
                        func_util.safe_exit(1)
                    self.randoms[i] = float(ref.field_mean_density(value_unit="grid"))
                elif r in (None, ""):
                    self.logger.error(f"Missing input for random vertex {i}.")
                    func_util.safe_exit(1)
                else:
                    base = self._load(r, f"random vertex {i}", cache)
                    key = (id(base), repr(wsrc[i]))
                    if key not in built:
                        built[key] = self._windowed(self._normalized(base), wsrc[i])
                    self.randoms[i] = built[key]
                    compat.append(base)
        if compat:
            func_util.validate_sfc_compatibility(compat, SFCField._REQUIRED_ARGV, logger=self.logger,
                                                 label="Corr_3PCF input fields")
        self.geometry = next((f for f in list(self.fields.values()) + list(self.randoms.values())
                              if isinstance(f, SFCField)), None)
        if self.geometry is None:
            self.logger.error("At least one SFCField input is required to define geometry for Corr_3PCF.")
            func_util.safe_exit(1)
        self.rho = float(self.geometry.field_mean_density(value_unit="grid"))
        self._fields_prepared = True

    def _catalogue_random_centres(self, random1, cache):
        """Random vertex-1 centres for r_delta_dd when random_pos1 is not given (corr3pcf.py:825-833, used at
        :1384-1386): the particles of random1's own catalogue with their projection weights in the task's
        normalisation.  A uniform random1 needs none (r_delta_dd = rrr xi23).  A random1 field that cannot
        give its particles ends the task like the reference (logged error, safe_exit)."""
        if random1 in (None, "") or _is_density(random1):
            return None
        if isinstance(random1, str) and random1.strip().lower() == "uniform":
            return None
        base = self._load(random1, "random vertex 1", cache)
        try:
            pdata = self._normalized(base).get_particle_data()
            return self._centre_arrays(pdata["pos"], pdata["projection_weight"], "random1 particle positions",
                                       "random1 particle weight")
        except Exception:
            self.logger.error("For center='particle', random1 could not provide usable particle coordinates. "
                              "Please provide random_pos1 and random_weight1 explicitly.")
            func_util.safe_exit(1)

    # ------------------------------------------------------------ estimator
    def _delta(self, field, random):
        return field - (random if _is_density(random) else random)

    def _grid(self, f):
        return f.device_epsilon(torch.float64)

    def _triplet(self, g1, g2, g3):
        """Per-mu box-random estimate of <g1(c) g2(c + d12) g3(c + d13)> (grid units)."""
        sf = float(self.geometry.scale_factor)
        R1, R2 = self.r12 * sf, self.r13 * sf
        n1 = kernels.interpolate(g1, self._centres, self._phi)  # vertex 1 sits on the centres
        out = np.empty(self.mu_arr.size)
        for it, mu in enumerate(self.mu_arr):
            abc = rotation_draws(self.base_seed + 1, self.n_rot, it)
            out[it] = kernels.mc_triplet_box(R1, R2, float(mu), self._centres, n1, abc, g2, g3, self._phi)
        return out

    def _pair_stats(self, fa, fb, ra, rb, radii, cache):
        L, box = int(self.geometry.L), float(self.geometry.box_size)
        rr = shell_pair_products(ra, rb, radii, L, box, cache)
        da = self._delta_cached(fa, ra, cache)
        db = da if (fb is fa and rb == ra) else self._delta_cached(fb, rb, cache)
        dd = shell_pair_products(da, db, radii, L, box, cache)
        return {"rr": rr, "delta_dd": dd, "xi": dd / rr}

    def _delta_cached(self, f, r, cache):
        key = ("delta", id(f), r if _is_density(r) else id(r))
        if key not in cache:
            cache[key] = self._delta(f, r)
        return cache[key]

    def run(self, save_result=True, overwrite=False):
        t0 = time.perf_counter()
        if not self._fields_prepared:
            self.prepare_input_fields()
        exp = self._expanded_products()
        geo = self.geometry
        L = int(geo.L)
        dev = kernels._dev()
        particle = self.center == "particle"
        if not particle:
            centres = np.random.default_rng(seed=self.base_seed + CENTRE_SEED_OFFSET).uniform(
                0.0, float(L), size=(self.n_box_centers, 3))
            self._centres = kernels.as_device(centres, torch.float64)
        self._phi = kernels.as_device(np.asarray(geo.phi_array, dtype=np.float64), torch.float64)
        # particle centres with a uniform random vertex 1 and no random positions: r_delta_dd = rrr xi23
        # (corr3pcf.py:1236-1244, :1607-1610), which needs both
        r_shortcut = (particle and "r_delta_dd" in exp and self._random_centres is None
                      and _is_density(self.randoms.get(1)))
        want_rrr = "rrr" in exp or r_shortcut
        want_xi23 = "xi23" in exp or r_shortcut
        defer_rrr = want_rrr and want_xi23 and _is_density(self.randoms.get(1))
        results = {}
        cache = {}
        if particle:
            results = self._particle_products(exp, r_shortcut, want_rrr and not defer_rrr, cache)
        else:
            if "ddd" in exp:
                g = [self._grid(self.fields[i]) for i in (1, 2, 3)]
                results["ddd"] = self._triplet(*g)
            if "delta_ddd" in exp:
                d = [self._delta_cached(self.fields[i], self.randoms[i], cache) for i in (1, 2, 3)]
                results["delta_ddd"] = self._triplet(*(self._grid(x) for x in d))
            if "rrr" in exp and not defer_rrr:
                results["rrr"] = self._rrr(cache)
        scale = getattr(self, "_density_scale", 1.0) ** 3
        if not np.isclose(scale, 1.0):
            results = {k: v * scale for k, v in results.items()}
        pair = {}
        r23 = third_side_from_mu(self.r12, self.r13, self.mu_arr)
        if "xi12" in exp:
            pair["xi12"] = {k: v[0] for k, v in self._pair_stats(self.fields[1], self.fields[2], self.randoms[1],
                                                                 self.randoms[2], [self.r12], cache).items()}
        if "xi13" in exp:
            pair["xi13"] = {k: v[0] for k, v in self._pair_stats(self.fields[1], self.fields[3], self.randoms[1],
                                                                 self.randoms[3], [self.r13], cache).items()}
        rr23 = None
        if want_xi23 or want_rrr:
            if want_xi23 or defer_rrr:
                pair["xi23"] = self._pair_stats(self.fields[2], self.fields[3], self.randoms[2], self.randoms[3],
                                                r23, cache)
                rr23 = pair["xi23"]["rr"]
        data = self.corr3pcf_data
        data.mu, data.theta, data.r23 = self.mu_arr, self.theta_arr, r23
        data.r12, data.r13 = self.r12, self.r13
        data.ddd = results.get("ddd")
        data.rrr = results.get("rrr")
        data.delta_ddd = results.get("delta_ddd")
        data.d_delta_dd = results.get("d_delta_dd")
        data.r_delta_dd = results.get("r_delta_dd")
        data.xi12 = pair.get("xi12", {}).get("xi")
        data.xi13 = pair.get("xi13", {}).get("xi")
        xi23 = pair.get("xi23", {}).get("xi")
        data.xi23 = xi23 if "xi23" in exp else None
        rrr = data.rrr
        if want_rrr and rrr is None:
            rrr = float(self.randoms[1]) * rr23 * scale
        data.rrr = rrr if "rrr" in exp else None
        if r_shortcut:
            data.r_delta_dd = rrr * xi23
        if particle and "delta_ddd" in exp:
            data.delta_ddd = data.d_delta_dd - data.r_delta_dd
        if "zeta" in exp:
            data.zeta = data.delta_ddd / rrr
        if "zeta_H" in exp:
            x12, x13, x23 = data.xi12, data.xi13, data.xi23
            data.zeta_H = x12 * x13 + x12 * x23 + x13 * x23
        if "Q" in exp:
            data.Q = data.zeta / data.zeta_H
        data.corr3pcf_info = {"center": self.center, "r12": self.r12, "r13": self.r13, "n_rot": self.n_rot,
                              "n_box_centers": None if particle else self.n_box_centers, "base_seed": self.base_seed,
                              "products": list(self.products), "angle_param": self.angle_param}
        if save_result and self.fout_path:
            data.save_corr3pcf(self.fout_path, overwrite=overwrite)
        self.logger.info(f"The time for task: {time.perf_counter() - t0:.4f} sec")
        return data

    def _particle_estimate(self, centres, g2, g3, rho1):
        """corr3pcf_kernels.py:66-97 per mu: rho1 sum_rot sum_i w_i g2 g3 / (n_rot sum w) (grid units)."""
        pos, w = centres
        sf = float(self.geometry.scale_factor)
        R1, R2 = self.r12 * sf, self.r13 * sf
        pos_d = kernels.as_device(pos * sf, torch.float64)
        w_d = kernels.as_device(w, torch.float64)
        n, wsum = pos.shape[0], float(np.sum(w))
        out = np.empty(self.mu_arr.size)
        for it, mu in enumerate(self.mu_arr):
            abc = rotation_draws(self.base_seed + 1, self.n_rot, it)
            mean = kernels.mc_triplet_box(R1, R2, float(mu), pos_d, w_d, abc, g2, g3, self._phi)
            out[it] = rho1 * mean * n / wsum
        return out

    def _particle_products(self, exp, r_shortcut, want_rrr, cache):
        """corr3pcf.py:1135-1180 with the densities of :1430-1450 (V = L^3 in grid units)."""
        V = float(self.geometry.L) ** 3
        data_c = self._data_centres
        rand_c = self._random_centres if self._random_centres is not None else data_c
        rho1_data = None if data_c is None else float(np.sum(data_c[1])) / V
        rho1_rand = float(np.sum(rand_c[1])) / V if rand_c is not None else None
        out = {}
        if "ddd" in exp:
            out["ddd"] = self._particle_estimate(data_c, self._grid(self.fields[2]), self._grid(self.fields[3]),
                                                 rho1_data)
        if "d_delta_dd" in exp or ("r_delta_dd" in exp and not r_shortcut):
            d2 = self._grid(self._delta_cached(self.fields[2], self.randoms[2], cache))
            d3 = self._grid(self._delta_cached(self.fields[3], self.randoms[3], cache))
            if "d_delta_dd" in exp:
                out["d_delta_dd"] = self._particle_estimate(data_c, d2, d3, rho1_data)
            if "r_delta_dd" in exp and not r_shortcut:
                out["r_delta_dd"] = self._particle_estimate(rand_c, d2, d3, rho1_rand)
        if want_rrr:
            out["rrr"] = self._rrr(cache, rho1=rho1_rand, centres=rand_c)
        return out

    def _rrr(self, cache, rho1=None, centres=None):
        """corr3pcf.py:736-772 shortcuts; three explicit random fields run the triplet kernel (box-random
        centres, or the particle centres with their weights and rho1)."""
        r = [self.randoms[i] for i in (1, 2, 3)]
        dens = [_is_density(x) for x in r]
        rho = [float(x) if _is_density(x) else float(x.field_mean_density(value_unit="grid")) for x in r]
        if rho1 is not None:
            rho[0] = float(rho1)
        L, box = int(self.geometry.L), float(self.geometry.box_size)
        n = self.mu_arr.size
        if sum(dens) >= 2:
            return np.full(n, rho[0] * rho[1] * rho[2])
        if dens[0]:
            r23 = third_side_from_mu(self.r12, self.r13, self.mu_arr)
            return rho[0] * shell_pair_products(r[1], r[2], r23, L, box, cache)
        if dens[1]:
            return np.full(n, rho[1] * shell_pair_products(r[0], r[2], [self.r13], L, box, cache)[0])
        if dens[2]:
            return np.full(n, rho[2] * shell_pair_products(r[0], r[1], [self.r12], L, box, cache)[0])
        if centres is not None:
            return self._particle_estimate(centres, self._grid(r[1]), self._grid(r[2]), rho[0])
        return self._triplet(*(self._grid(x) for x in r))
