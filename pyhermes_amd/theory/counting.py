"""Counting: the field density at uniform random points of the box (drop-in for pyhermes.theory.Counting).

counting.py:17-258: the task field (catalogue fields in the task's weight normalization, optionally
convolved with a window) is sampled at ``random_count`` points drawn by
numpy default_rng(seed).uniform(0, box_size, (random_count, 3)) (sampling.py:6-10, the single-process
rank-0 seed of counting.py:215), in physical units (field_density_at_pos, wavelet_grid.py:143-186,
through the HIP interpolation kernel); the result is CountingData.nx.
"""

import time

import numpy as np

from ..io import CountingData, SFCField, WindowFunc, normalize_task_weight_normalization
from ..pipeline import TaskBase
from ..utils import func_util


class Counting(TaskBase):
    def __init__(self, param_task=None):
        if param_task is None:
            param_task = {"Counting": {}}
        self.task_name = str(self.__class__.__name__)
        super().__init__(param_task=param_task)
        self.format_params()
        self._fields_prepared = False

    def format_params(self):
        tp = self.task_params
        self.sfc_field = tp.get("sfc_field", "")
        self.random_count = int(tp["random_count"])
        if self.random_count < 0:
            raise ValueError("Counting random_count must be non-negative.")
        self.seed = int(tp["seed"])
        self.weight_normalization = normalize_task_weight_normalization(tp.get("weight_normalization", "catalog"))
        window = tp.get("window", None)
        self.window = window if (window and window.get("type")) else None
        self.threads = int(tp["threads"])
        self.fout_path = tp["fout_path"]

    def _field_in_task_normalization(self, field):
        """counting.py:52-74: 'unit' for every field, the task normalization for catalogue fields."""
        if self.weight_normalization == "unit":
            return field.with_normalization("unit")
        if getattr(field, "field_kind", None) != "catalog_field":
            return field.copy()
        return field.switch_weight_normalization(self.weight_normalization)

    def prepare_input_fields(self, sfc_field=None, window=None):
        self.counting_data = CountingData()
        src = self.sfc_field if sfc_field is None else sfc_field
        win = self.window if window is None else window
        if isinstance(src, str) and src:
            src = SFCField(data_path=src, threads=self.threads)
        if not isinstance(src, SFCField):
            self.logger.error("Counting needs an SFCField (or a path to one) as 'sfc_field'.")
            func_util.safe_exit(1)
        base = self._field_in_task_normalization(src)
        if win is not None:
            w = win if isinstance(win, WindowFunc) else WindowFunc(win, base.sfc_info, threads=self.threads)
            base = base @ w
        base.format_sfc_params()
        self.field = base
        self.counting_data.sfc_info = base.sfc_info
        self._fields_prepared = True

    def run(self, save_result=True, overwrite=False):
        t0 = time.perf_counter()
        if not self._fields_prepared:
            self.prepare_input_fields()
        box = float(self.field.box_size)
        pos = np.random.default_rng(self.seed).uniform(0.0, box, size=(self.random_count, 3))
        nx = (self.field.field_density_at_pos(pos, value_unit="physical").astype(np.float64, copy=False)
              if self.random_count else np.zeros(0))
        data = self.counting_data
        data.nx = nx
        data.counting_info = {"random_count": self.random_count, "seed": self.seed,
                              "weight_normalization": self.weight_normalization,
                              "window": None if self.window is None else
                              {k: v for k, v in dict(self.window).items() if k != "func"},
                              "fout_path": self.fout_path}
        if save_result and self.fout_path:
            data.save_counting(self.fout_path, overwrite=overwrite)
        self.logger.info(f"The time for task: {time.perf_counter() - t0:.4f} sec")
        return data
