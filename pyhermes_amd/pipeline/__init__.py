from .pipeline import TaskBase  # noqa: F401
