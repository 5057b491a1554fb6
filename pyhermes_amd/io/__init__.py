from .results import Corr2PCFData, Corr3PCFData, Corr3PCFMultipoleData, CountingData  # noqa: F401
from .sfc_field import (SFCField, normalize_field_value_normalization,  # noqa: F401
                        normalize_task_weight_normalization, normalize_weight_normalization)
from .window import WindowFunc  # noqa: F401
