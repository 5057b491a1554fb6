"""Error convention and small helpers (mirrors pyhermes.utils.func_util)."""

import inspect
import sys


def get_fname_info():
    frame = inspect.currentframe().f_back
    return frame.f_globals["__name__"], frame.f_code.co_name


def in_jupyter_notebook():
    try:
        from IPython import get_ipython
    except ImportError:
        return False
    ip = get_ipython()
    return ip is not None and "IPKernelApp" in getattr(ip, "config", {})


def safe_exit(exit_code=1):
    """Abort the job: SystemExit(code) (func_util.py:26-49 semantics under the single-process communicator)."""
    if in_jupyter_notebook():
        sys.exit(exit_code)
    from .mpi_util import MPI
    MPI.COMM_WORLD.Abort(exit_code)


def validate_sfc_compatibility(sfc_list, required_keys, logger=None, label="SFCField inputs"):
    """All fields must share the required sfc_info values; returns the shared values."""
    fields = [c for c in sfc_list if c is not None]
    if not fields:
        return {}
    ref = fields[0]
    bad = []
    for idx, cur in enumerate(fields[1:], start=2):
        for key in required_keys:
            if ref.sfc_info.get(key) != cur.sfc_info.get(key):
                bad.append(f"vertex{idx}.{key}={cur.sfc_info.get(key)} (reference={ref.sfc_info.get(key)})")
    if bad:
        if logger is not None:
            logger.error(f"{label} require matching required parameters. Found mismatches: {', '.join(bad)}")
        safe_exit(1)
    return {key: ref.sfc_info.get(key) for key in required_keys}


def describe_window_action(window):
    if not isinstance(window, dict):
        return str(type(window).__name__)
    return f"type={window.get('type')} len_args={window.get('len_args', {})}"
