from .corr2pcf import Corr_2PCF  # noqa: F401
from .corr3pcf import Corr_3PCF  # noqa: F401
from .corr3pcf_multipole import Corr_3PCF_Multipole  # noqa: F401
from .counting import Counting  # noqa: F401
