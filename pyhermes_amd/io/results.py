"""Result containers of the correlation tasks (drop-in for pyhermes.io.corr2pcf / corr3pcf /
corr3pcf_multipole), saved in the same NumPy-framed pickle layout."""

from ..utils import func_util
from .base import HermesData


class _Result(HermesData):
    _FIELDS = ()
    _REQUIRED = ()
    _INFO = "info"

    def __init__(self, *args, threads=None, **kwargs):
        data_path = kwargs.pop("data_path", None)
        for f in self._FIELDS:
            setattr(self, f, None)
        setattr(self, self._INFO, {})
        super().__init__(*args, threads=threads, **kwargs)
        if data_path:
            self.load(data_path)

    def _payload(self):
        if not getattr(self, self._INFO):
            self.logger.error(f'The dictionary "{self._INFO}" is empty.')
            func_util.safe_exit(1)
        d = {f: getattr(self, f) for f in self._FIELDS}
        d[self._INFO] = getattr(self, self._INFO)
        return d

    def _restore(self, dataset):
        for key in self._REQUIRED:
            if key not in dataset:
                self.logger.error(f"Failed to load the dataset. The file is missing the '{key}' key.")
                func_util.safe_exit(1)
        for f in self._FIELDS:
            setattr(self, f, dataset.get(f))
        info = dataset.get(self._INFO) or {}
        getattr(self, self._INFO).update(info)


class Corr2PCFData(_Result):
    _FIELDS = ("sfc_info1", "sfc_info2", "sampling_names", "sampling", "dd", "dr", "rd", "delta_dd", "rr", "xi",
               "multipoles", "xi_l")
    _REQUIRED = ("sampling",)
    _INFO = "corr2pcf_info"

    def set_sampling(self, names, arrays):
        self.sampling_names = tuple(names)
        self.sampling = dict(arrays)
        for k, v in arrays.items():
            setattr(self, k, v)

    def save_corr2pcf(self, f_out, single=True, overwrite=False):
        return self.save(f_out, overwrite=overwrite)

    def load_corr2pcf(self, f_in, single=True, trusted=False):
        return self.load(f_in, trusted=trusted)


class Corr3PCFData(_Result):
    _FIELDS = ("sfc_info1", "sfc_info2", "sfc_info3", "theta", "mu", "r23", "ddd", "rrr", "d_delta_dd", "r_delta_dd",
               "delta_ddd", "xi12", "xi13", "xi23", "zeta", "zeta_H", "Q", "r12", "r13")
    _REQUIRED = ("theta", "mu")
    _INFO = "corr3pcf_info"

    def save_corr3pcf(self, f_out, single=True, overwrite=False):
        return self.save(f_out, overwrite=overwrite)

    def load_corr3pcf(self, f_in, single=True, trusted=False):
        return self.load(f_in, trusted=trusted)


class Corr3PCFMultipoleData(_Result):
    """io/corr3pcf_multipole.py:9-93 fields."""
    _FIELDS = ("sfc_info1", "sfc_info2", "sfc_info3", "sample_params", "binning_window12", "binning_window13", "r12",
               "r13", "l", "ddd_l", "rrr_l", "delta_ddd_l", "zeta_l", "zeta_condition")
    _REQUIRED = ("l", "sample_params", "binning_window12", "binning_window13")
    _INFO = "corr3pcf_multipole_info"

    def save_corr3pcf_multipole(self, f_out, single=True, overwrite=False):
        return self.save(f_out, overwrite=overwrite)

    def load_corr3pcf_multipole(self, f_in, single=True, trusted=False):
        return self.load(f_in, trusted=trusted)


class CountingData(_Result):
    """io/counting.py:9-60 fields: the field density at random positions (nx) with the field's sfc_info."""
    _FIELDS = ("sfc_info", "nx")
    _REQUIRED = ("nx",)
    _INFO = "counting_info"

    def save_counting(self, f_out, single=True, overwrite=False):
        return self.save(f_out, overwrite=overwrite)

    def load_counting(self, f_in, single=True, trusted=False):
        return self.load(f_in, trusted=trusted)
