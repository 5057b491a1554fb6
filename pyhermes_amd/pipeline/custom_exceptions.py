"""Exception types raised by the task layer."""


class HermesConfigError(ValueError):
    """Invalid task configuration."""


class HermesDeviceError(RuntimeError):
    """The HIP library or device is unavailable or reported an error."""
