"""Single-process communicator with the mpi4py surface the tasks use.

Mirrors the reference's FakeMPI (mpi_util.py:58-169): this build runs one
process per GPU and one GPU, so every collective is the identity and Abort
raises SystemExit(code).
"""

import numpy as np

from ..param.logbase import setup_logger


class _Sentinel:
    def __init__(self, name):
        self.name = name

    def __repr__(self):
        return f"MPI.{self.name}"


class _NullComm:
    def Free(self):
        pass


class _Request:
    def __init__(self, comm, key):
        self._comm, self._key, self.completed = comm, key, False

    def test(self):
        data = self._comm._buffers.get(self._key)
        if data is not None and not self.completed:
            self.completed = True
            return True, data
        return False, None

    def wait(self):
        self.completed = True
        return self._comm._buffers.get(self._key)

    Test, Wait = test, wait


def _first(buf):
    return buf[0] if isinstance(buf, (list, tuple)) else buf


class SingleProcessComm:
    def __init__(self):
        self._buffers = {}
        self._logger = None

    def Get_size(self):
        return 1

    def Get_rank(self):
        return 0

    # object (pickle-based) interface
    def send(self, data, dest, tag=0):
        self._buffers[(dest, tag)] = data

    def isend(self, data, dest, tag=0):
        self.send(data, dest, tag)
        return _Request(self, (dest, tag))

    def recv(self, buf=None, source=0, tag=0, status=None):
        return self._buffers.pop((source, tag), buf)

    def irecv(self, buf=None, source=0, tag=0):
        return _Request(self, (source, tag))

    def Iprobe(self, source=0, tag=0, status=None):
        return (source, tag) in self._buffers

    def bcast(self, data, root=0):
        return data

    def allgather(self, data):
        return [data]

    def reduce(self, sendbuf, recvbuf=None, op=None, root=0):
        return sendbuf

    def allreduce(self, sendbuf, op=None):
        return sendbuf

    def gather(self, sendbuf, recvbuf=None, root=0):
        if recvbuf is not None:
            target = _first(recvbuf)
            if isinstance(target, np.ndarray):
                target[...] = sendbuf
            else:
                target[:] = [sendbuf]
        return [sendbuf]

    def scatter(self, data, root=0):
        return data[0]

    # buffer interface
    def Allreduce(self, sendbuf, recvbuf, op=None):
        np.copyto(np.asarray(_first(recvbuf)), np.asarray(_first(sendbuf)))

    def Scatterv(self, sendbuf, recvbuf, root=0):
        src = np.asarray(_first(sendbuf)).reshape(-1)
        dst = np.asarray(_first(recvbuf))
        dst[...] = src[:dst.size].reshape(dst.shape)

    def Bcast(self, buf, root=0):
        return buf

    def Barrier(self):
        pass

    def Split_type(self, split_type, key=0, info=None):
        return self

    def Split(self, color, key=0):
        return MPI.COMM_NULL if color is MPI.UNDEFINED else self

    def Abort(self, errorcode=1):
        raise SystemExit(errorcode)

    def Free(self):
        pass

    Send, Recv, Gather, Scatter = send, recv, gather, scatter


class MPI:
    COMM_NULL = _NullComm()
    COMM_TYPE_SHARED = _Sentinel("COMM_TYPE_SHARED")
    UNDEFINED = _Sentinel("UNDEFINED")
    SUM = _Sentinel("SUM")
    MAX = _Sentinel("MAX")
    DOUBLE = np.float64
    COMPLEX = np.complex64
    COMPLEX16 = np.complex128
    COMM_WORLD = SingleProcessComm()


def _announce():
    setup_logger(__name__, "MPI").debug("single-process communicator (one process per GPU)")
