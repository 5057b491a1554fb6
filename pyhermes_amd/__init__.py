"""pyhermes_amd -- MI355X-native PyHermes field pipeline.

Drop-in for PyHermes' field pipeline: ``SFCProjection(params).run()``,
``field @ WindowFunc`` and the ``Corr_2PCF`` / ``Corr_3PCF`` /
``Corr_3PCF_Multipole`` tasks, computed by hand-written HIP kernels for gfx950
(``pyhermes_amd/csrc``) through the C ABI in ``include/hermes_amd.h``.
"""

__version__ = "0.1.0"
