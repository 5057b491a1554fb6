"""Radial profiles of the SH-3PCF edge windows on the device.

An edge window (a binning window of Corr_3PCF_Multipole, or a WindowFunc of type radial_multipole) has the
(l, m) kernel U_l(|k|) Y_lm(k-hat).  EdgeProfile turns one such window -- its radial type, physical lengths
and other_args -- into the hermes_radial_profile the kernels take, per l, dispatched as
radial_multipole_windows.py:267-316 does: thin shells (and Gaussian shells without smoothing) in closed
form at every l, the built-in profiles' monopoles in closed form, everything else from a Hankel table U_l
built on the host (utils/radial_profiles.py) and interpolated on the device.
"""

import numpy as np
import torch

from . import kernels
from .utils.radial_profiles import (BUILTIN_RADIAL_PROFILE_TYPES, _key_of, build_radial_multipole_table,
                                    validate_radial_profile_request)

_CLOSED_FORM_ARGS = {"sphere": ("R", None), "gaussian": ("R", None), "thick_shell": ("R", "delta_R"),
                     "gaussian_shell": ("R_shell", "R_smooth")}


class EdgeProfile:
    def __init__(self, radial_type, len_args, other_args, L, box):
        self.radial_type = str(radial_type).strip().lower()
        self.len_args = {k: float(v) for k, v in dict(len_args or {}).items() if v is not None}
        self.other_args = other_args if other_args else {}
        validate_radial_profile_request(self.radial_type, self.len_args, self.other_args)
        self.L, self.box = int(L), float(box)
        self.g2p = self.L / self.box
        self.k_max = float(np.sqrt(3.0) * self.g2p)
        self.key = (self.radial_type, tuple(sorted(self.len_args.items())), _key_of(self.other_args), self.L,
                    self.box)
        self._profiles = {}

    def __eq__(self, other):
        return isinstance(other, EdgeProfile) and self.key == other.key

    def __hash__(self):
        return hash(self.key)

    @property
    def thin_radius(self):
        """Radius in grid units when the transfer is j_l(2 pi k R) at every l, else None."""
        if self.radial_type == "shell":
            return self.len_args["R"] * self.g2p
        if self.radial_type == "gaussian_shell" and self.len_args["R_smooth"] == 0.0:
            return self.len_args["R_shell"] * self.g2p
        return None

    def profile(self, l, device=None):
        """hermes_radial_profile of degree l (the device table, when there is one, stays referenced here)."""
        l = int(l)
        if l in self._profiles:
            return self._profiles[l][0]
        t = self.radial_type
        table = None
        if t == "shell":
            prof = kernels.radial_profile("shell", self.len_args["R"] * self.g2p)
        elif l == 0 and t in BUILTIN_RADIAL_PROFILE_TYPES:
            a, b = _CLOSED_FORM_ARGS[t]
            prof = kernels.radial_profile(t, self.len_args[a] * self.g2p, self.len_args[b] * self.g2p if b else 0.0)
        elif self.thin_radius is not None:
            prof = kernels.radial_profile("shell", self.thin_radius)
        else:
            k_max, host = build_radial_multipole_table(t, self.len_args, l, self.k_max, profile_config=self.other_args)
            table = kernels.as_device(np.ascontiguousarray(host, dtype=np.float64), torch.float64)
            # table index per cycle per grid cell: k_phys = k_grid * L / box, index = k_phys (n - 1) / k_max
            prof = kernels.radial_profile("table", table=table, k_scale=self.g2p * (host.size - 1) / k_max)
        self._profiles[l] = (prof, table)
        return prof

    def window_full(self, geo_J, P, l, m, dbl=True):
        """(L, L, L) complex kernel of (l, m) for this edge (FP64 by default)."""
        return kernels.window_build_full_profile(geo_J, P, self.profile(l), l, m, dbl=dbl)


def as_edges(values, L, box):
    """A list of EdgeProfile from binning-window parameter dicts (type, len_args, other_args)."""
    return [EdgeProfile(w["type"], w.get("len_args", {}), w.get("other_args", {}), L, box) for w in values]
