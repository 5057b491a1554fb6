from .sfc_projection import SFCProjection  # noqa: F401
