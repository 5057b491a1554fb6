"""TaskBase: parameter merge and runtime options (mirrors pyhermes.pipeline.pipeline:10-68)."""

from ..param.logbase import configure_mpi_logging, setup_logger
from ..param.parambase import ParamBase
from ..utils import func_util
from ..utils.mpi_util import MPI


def _copy_params(value):
    """Copy nested dicts / lists; arrays, fields and callables are shared, not duplicated."""
    if isinstance(value, dict):
        return {k: _copy_params(v) for k, v in value.items()}
    if isinstance(value, list):
        return [_copy_params(v) for v in value]
    return value


class TaskBase(object):
    def __init__(self, param_task):
        self.comm = MPI.COMM_WORLD
        self.rank = self.comm.Get_rank()
        self.size = self.comm.Get_size()
        self.comm_local = self.comm.Split_type(MPI.COMM_TYPE_SHARED)
        self.size_local = self.comm_local.Get_size()
        configure_mpi_logging(rank=self.rank, size=self.size)
        self.logger = setup_logger(__name__, self.__class__.__name__)
        pb = ParamBase()
        pb.read_default(self.__class__)
        defaults = pb.default_params  # read_default hands this ParamBase its own copy
        user = param_task
        if self.task_name not in list(user.keys()):
            self.logger.error(f"Mismatch of task name, expected <{self.task_name}>, but it was not found in the "
                              f"provided task(s) <{list(user.keys())}>")
            func_util.safe_exit(1)
        pb.recursive_update(default_dict=defaults, new_dict=user, section=self.task_name)
        user[self.task_name] = defaults[self.task_name]
        self.task_params = _copy_params(defaults[self.task_name])
        self.threads = self.task_params.get("threads", None)

    def sync_runtime_options(self, context=None, blank_line=False):
        threads = self.task_params.get("threads", self.threads)
        if threads is not None:
            self.threads = max(1, int(threads))
            self.task_params["threads"] = self.threads
        prefix = f"{context}: " if context else ""
        self.logger.info(f"{prefix}running on {self.size} process(es), one HIP device")

    def run(self):
        pass
