"""Common base of the data containers (mirrors pyhermes.io.base.HermesData)."""

from ..param.logbase import setup_logger
from ..utils import func_util
from ..utils.mpi_util import MPI
from .formats import check_fout, read_numpy_pickle, write_numpy_pickle


class HermesData(object):
    def __init__(self, *args, threads=None, **kwargs):
        self.comm = MPI.COMM_WORLD
        self.rank = self.comm.Get_rank()
        self.logger = setup_logger(__name__, self.__class__.__name__)
        self.saveflag = False
        self.task_params = None
        self.threads = max(1, int(threads)) if threads else None

    # subclasses implement _payload() -> dict and _restore(dict)
    def save(self, f_out, overwrite=False, **_ignored):
        path = check_fout(self, f_out, overwrite, self.logger)
        if not path:
            self.logger.info("No <fout_path> specified, skipping disk output.")
            return None
        with open(path, "wb") as f:
            write_numpy_pickle(f, self._payload())
        self.logger.info(f"Writing data to ---> {path} <---")
        return path

    def load(self, f_in, trusted=False, **_ignored):
        """trusted=False resolves only the data types PyHermes files hold (arrays, scalars, containers,
        paths); True loads any pickle, for files whose origin you trust."""
        try:
            with open(f_in, "rb") as f:
                dataset = read_numpy_pickle(f, trusted=trusted)
        except Exception as exc:
            self.logger.error(f"An error occurred while loading the file '{f_in}': {exc}")
            func_util.safe_exit(1)
        self._restore(dataset)
        return self

    def _payload(self):
        raise NotImplementedError

    def _restore(self, dataset):
        raise NotImplementedError
