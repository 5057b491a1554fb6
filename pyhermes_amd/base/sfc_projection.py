"""SFCProjection: catalogue -> SFCField on the HIP device (drop-in for pyhermes.base.SFCProjection).

prepare_input_fields follows sfc_projection.py:221-300 (FP64 sums and projection
weights on the host, same validation and errors); run() replaces the Numba scatter
(sfc_projection.py:317-324 -> wavelet_grid.py:17-70) with hermes_project.  One
process per GPU: the reference's MPI x-slab decomposition is not used.
"""

import copy
import os
import time

import numpy as np
import torch

from .. import kernels
from ..io.sfc_field import SFCField, normalize_weight_normalization
from ..pipeline import TaskBase
from ..utils import func_util
from ..utils.wavelet_grid import sample_scaling_function


class SFCProjection(TaskBase):
    def __init__(self, param_task=None):
        if param_task is None:
            param_task = {"SFCProjection": {}}
        self.task_name = str(self.__class__.__name__)
        super().__init__(param_task=param_task)
        self.format_params()
        self.sfc_field = None
        self._fields_prepared = False

    def format_params(self):
        tp = self.task_params
        self.fin = copy.deepcopy(tp["fin"])
        self.particle_pos = tp["particle_pos"]
        self.catalog_weight = tp["catalog_weight"]
        self.field_value = tp["field_value"]
        self.weight_normalization = normalize_weight_normalization(tp.get("weight_normalization", "catalog"))
        self.box_size = tp["box_size"]
        self.J = tp["J"]
        self.L = 1 << self.J
        self.wavelet_mode = tp["wavelet_mode"]
        self.wavelet_level = tp["wavelet_level"]
        self.phi_resolution = int(tp["phi_resolution"])
        self.save_particle_data = bool(tp["save_particle_data"])
        self.particle_data_path = tp["particle_data_path"]
        self.threads = int(tp["threads"])
        self.fout_path = tp["fout_path"]
        self.precision = str(tp.get("precision", "fp64")).lower()

    def _load_particle_input(self):
        if self.particle_pos is None:
            path = self.fin.get("path", "")
            if not path:
                self.logger.error("No particle_pos array and no input file path given.")
                func_util.safe_exit(1)
            raise NotImplementedError("pyhermes_amd reads in-memory catalogues; file readers are outside its scope.")
        p_pos = self.particle_pos
        if isinstance(p_pos, torch.Tensor):
            p_pos = p_pos.detach().cpu().numpy()
        if not (isinstance(p_pos, np.ndarray) and p_pos.ndim == 2 and p_pos.shape[1] == 3):
            self.logger.error(
                "Wrong input of particle data! 'particle_pos' must be a 2D array of shape (N, 3), "
                f"but got type={type(p_pos)} with shape={getattr(p_pos, 'shape', None)}.")
            func_util.safe_exit(1)
        self.particle_count = p_pos.shape[0]
        n = self.particle_count
        cw = np.ones(n, dtype=np.float32) if self.catalog_weight is None else self.catalog_weight
        self.fin["catalog_weight_key"] = None if self.catalog_weight is None else "custom"
        self.fin["field_value_key"] = None if self.field_value is None else "custom"
        # an absent field value is the unit value: returned as None, and the sums and weights that would
        # multiply by 1.0 (exact in FP64) use the catalogue weights alone
        factors = {"field_value": None}
        inputs = [("catalog_weight", cw)]
        if self.field_value is not None:
            inputs.append(("field_value", self.field_value))
        for name, value in inputs:
            if np.isscalar(value):
                value = np.full(n, value, dtype=np.float64)
            if not isinstance(value, np.ndarray) or value.ndim != 1 or value.shape[0] != n:
                self.logger.error(f"Wrong input of {name}! Expected a numpy array with shape (N,), but got "
                                  f"type={type(value)} shape={getattr(value, 'shape', None)} while N={n}.")
                func_util.safe_exit(1)
            if not np.all(np.isfinite(value)):
                self.logger.error(f"Wrong input of {name}! Values must all be finite.")
                func_util.safe_exit(1)
            factors[name] = value.astype(np.float64, copy=False)
        return p_pos, factors["catalog_weight"], factors["field_value"]

    def prepare_input_fields(self, particle_pos=None, catalog_weight=None, field_value=None, fin=None):
        if fin is not None:
            self.fin.update(fin)
        if particle_pos is not None:
            self.particle_pos = particle_pos
        if catalog_weight is not None:
            self.catalog_weight = catalog_weight
        if field_value is not None:
            self.field_value = field_value
        self.weight_normalization = normalize_weight_normalization(self.weight_normalization)
        self.L = 1 << self.J
        self.sync_runtime_options(context="SFCProjection runtime configuration")
        self.sfc_field = SFCField(threads=self.threads)
        self.phi_array = sample_scaling_function(self.wavelet_mode, self.wavelet_level)
        self.phi_support = self.phi_array.shape[0] // self.phi_resolution
        self.scale_factor = self.L / self.box_size
        p_pos, cw, fv = self._load_particle_input()
        unit_value = fv is None
        if unit_value:  # read-only view of ones: the attribute keeps its shape without an O(N) array
            fv = np.broadcast_to(np.float32(1.0), (cw.shape[0],))
        self.particle_pos, self.catalog_weight, self.field_value = p_pos, cw, fv
        self.catalog_weight_sum = float(np.sum(cw, dtype=np.float64))
        self.catalog_weight_sq_sum = float(np.sum(cw ** 2, dtype=np.float64))
        if not np.isfinite(self.catalog_weight_sum) or np.isclose(self.catalog_weight_sum, 0.0):
            self.logger.error(f"Catalogue weights must have a finite, non-zero sum, got {self.catalog_weight_sum!r}.")
            func_util.safe_exit(1)
        self.raw_field_weighted_sum = (self.catalog_weight_sum if unit_value
                                       else float(np.sum(cw * fv, dtype=np.float64)))
        if not np.isfinite(self.raw_field_weighted_sum):
            self.logger.error("The product catalog_weight * field_value produced a non-finite summed field value.")
            func_util.safe_exit(1)
        if self.weight_normalization == "raw":
            den = 1.0
        elif self.weight_normalization == "catalog":
            den = self.catalog_weight_sum
        else:
            den = self.raw_field_weighted_sum
            if np.isclose(den, 0.0):
                self.logger.error("weight_normalization='field' requires a finite, non-zero raw_field_weighted_sum.")
                func_util.safe_exit(1)
        self.projection_weight = cw / den if unit_value else cw * fv / den
        self.field_integral = self.raw_field_weighted_sum / den
        if self.save_particle_data:
            path = self.particle_data_path or (os.path.splitext(self.fout_path)[0] + "_particles.npz"
                                               if self.fout_path else "")
            if not path:
                self.logger.error("save_particle_data=True requires either 'particle_data_path' or 'fout_path'.")
                func_util.safe_exit(1)
            os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
            np.savez(path, pos=np.ascontiguousarray(p_pos, dtype=np.float32),
                     catalog_weight=np.ascontiguousarray(cw, dtype=np.float64),
                     field_value=np.ascontiguousarray(fv, dtype=np.float64))
            self.particle_data_path = path
        self._fields_prepared = True

    def run(self, save_result=True, overwrite=False):
        t0 = time.perf_counter()
        if not self._fields_prepared:
            self.prepare_input_fields()
        want64 = self.precision != "fp32"
        try:
            out = kernels.project(self.particle_pos, self.projection_weight, self.J, self.box_size, self.phi_array,
                                  want32=not want64, want64=want64, phi_res=self.phi_resolution)
        except RuntimeError as exc:
            self.logger.error(f"Error in process {self.rank}: {exc}")
            func_util.safe_exit(1)
        info = {
            "fin": copy.deepcopy(self.fin), "particle_count": self.particle_count,
            "catalog_weight_sum": self.catalog_weight_sum, "catalog_weight_sq_sum": self.catalog_weight_sq_sum,
            "raw_field_weighted_sum": self.raw_field_weighted_sum, "field_integral": self.field_integral,
            "box_size": self.box_size, "J": self.J, "L": self.L, "V": self.L ** 3, "scale_factor": self.scale_factor,
            "field_kind": "catalog_field", "weight_normalization": self.weight_normalization,
            "particle_data_retrievable": True, "wavelet_mode": self.wavelet_mode,
            "wavelet_level": self.wavelet_level, "phi_resolution": self.phi_resolution,
            "particle_data_path": self.particle_data_path if self.save_particle_data else "",
            "particle_data_format": "npz" if self.save_particle_data else "",
            "phi_support": self.phi_support, "phi_array": self.phi_array,
        }
        self.sfc_field.sfc_info = info
        self.sfc_field.format_sfc_params()
        self.sfc_field._set_device(out["eps64"] if want64 else out["eps32"])
        if save_result and self.fout_path:
            self.sfc_field.save_sfc_field(self.fout_path, overwrite=overwrite)
        self.logger.info(f"The time for task: {time.perf_counter() - t0:.4f} sec")
        return self.sfc_field
