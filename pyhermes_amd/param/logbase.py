"""Rank-aware logging (mirrors pyhermes.param.logbase: root-only unless PYHERMES_MPI_LOG_MODE=all)."""

import logging
import os
import sys

_state = {"rank": 0, "size": 1, "mode": os.environ.get("PYHERMES_MPI_LOG_MODE", "root").strip().lower()}


def configure_mpi_logging(rank=None, size=None, mode=None):
    if rank is not None:
        _state["rank"] = int(rank)
    if size is not None:
        _state["size"] = max(1, int(size))
    mode = os.environ.get("PYHERMES_MPI_LOG_MODE", _state["mode"]) if mode is None else mode
    mode = str(mode).strip().lower()
    if mode not in ("root", "all"):
        raise ValueError(f"PYHERMES_MPI_LOG_MODE must be either 'root' or 'all', got {mode!r}.")
    _state["mode"] = mode


class MPIRankFilter(logging.Filter):
    """Routine records from rank 0 only; errors from every rank."""

    def filter(self, record):
        multi = _state["size"] > 1
        record.mpi_rank_prefix = f"[rank {_state['rank']}/{_state['size']}] " if multi and (
            _state["mode"] == "all" or _state["rank"] != 0) else ""
        if not multi or _state["mode"] == "all" or _state["rank"] == 0:
            return True
        return record.levelno >= logging.ERROR


def setup_logger(module_name, class_name=None, level=logging.INFO, stream_handler=True):
    name = module_name + (f":{class_name}" if class_name else "")
    logger = logging.getLogger(name)
    if logger.level != level:  # setLevel clears every logger's cache, so only call it on a change
        logger.setLevel(level)
    if stream_handler and not logger.handlers:
        h = logging.StreamHandler(stream=sys.stdout)
        h.addFilter(MPIRankFilter())
        h.setFormatter(logging.Formatter("%(asctime)s - %(levelname)s - %(name)s - %(mpi_rank_prefix)s%(message)s",
                                         datefmt="%H:%M:%S"))
        logger.addHandler(h)
    return logger
