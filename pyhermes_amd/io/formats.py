"""On-disk format shared with PyHermes data products: a pickled dict (protocol 4)
stored as the uint8 payload of a .npy array (the layout of pickle_compat.py:81-90),
with path-like metadata written as strings.

Reading (pickle_compat.py:29-58 behaviour): path objects in a payload come back as
strings, including those pickled by Python 3.13, whose ``PosixPath`` and friends
live in the private module ``pathlib._local`` that older Pythons lack.  Unlike the
reference reader, the default reader resolves only the globals such data products
hold -- NumPy arrays, dtypes and scalars, a few builtin containers and path
classes -- and refuses any other callable before it could run; ``trusted=True``
restores the unrestricted pickle semantics for files whose origin you trust.
"""

import _compat_pickle
import io
import os
import pickle
from pathlib import PurePath, PurePosixPath, PureWindowsPath

import numpy as np

# path classes by name, as pickled from any pathlib module (public or Python 3.13's private ones)
_PATH_CLASSES = {
    "Path": PurePath, "PurePath": PurePath,
    "PosixPath": PurePosixPath, "PurePosixPath": PurePosixPath,
    "WindowsPath": PureWindowsPath, "PureWindowsPath": PureWindowsPath,
}
_PATH_MODULES = {"pathlib", "pathlib._local", "pathlib._abc"}

# (module, name) of the non-path globals a data product may reference
_DATA_GLOBALS = {
    ("builtins", "complex"), ("builtins", "set"), ("builtins", "frozenset"), ("builtins", "bytearray"),
    ("builtins", "slice"), ("builtins", "range"), ("collections", "OrderedDict"),
    ("_codecs", "encode"),  # bytes in protocol 2 pickles
    ("numpy", "ndarray"), ("numpy", "dtype"),
}
for _mod in ("numpy.core.multiarray", "numpy._core.multiarray"):
    _DATA_GLOBALS |= {(_mod, "_reconstruct"), (_mod, "scalar")}
for _mod in ("numpy.core.numeric", "numpy._core.numeric"):
    _DATA_GLOBALS.add((_mod, "_frombuffer"))
_DATA_GLOBALS = frozenset(_DATA_GLOBALS)


def _path_string(cls):
    def build(*parts):
        return os.fsdecode(os.fspath(cls(*parts)))
    return build


class _DataUnpickler(pickle.Unpickler):
    def __init__(self, file_obj, trusted):
        super().__init__(file_obj)
        self._trusted = trusted

    def find_class(self, module, name):
        if module in _PATH_MODULES and name in _PATH_CLASSES:
            return _path_string(_PATH_CLASSES[name])
        # protocol 0-2 pickles may carry Python 2 names (__builtin__.complex): check the Python 3 ones,
        # which the base class resolves them to
        key = _compat_pickle.NAME_MAPPING.get((module, name),
                                              (_compat_pickle.IMPORT_MAPPING.get(module, module), name))
        if not self._trusted and key not in _DATA_GLOBALS:
            raise pickle.UnpicklingError(
                f"refusing to load global '{module}.{name}': not a data type of PyHermes files "
                f"(pass trusted=True to load a file whose origin you trust)")
        return super().find_class(module, name)


def _portable(value):
    """Path-like metadata as strings, recursively through dicts, lists, tuples and sets."""
    if isinstance(value, os.PathLike):
        return os.fsdecode(os.fspath(value))
    if isinstance(value, dict):
        return {_portable(k): _portable(v) for k, v in value.items()}
    if isinstance(value, (list, tuple, set, frozenset)):
        return type(value)(_portable(v) for v in value)
    return value


def pickle_loads_compatible(payload, trusted=False):
    """One pickle payload, with path objects (any Python's pathlib) returned as strings."""
    return _portable(_DataUnpickler(io.BytesIO(payload), trusted).load())


def write_numpy_pickle(file_obj, value, protocol=4):
    payload = pickle.dumps(_portable(value), protocol=protocol)
    np.lib.format.write_array(file_obj, np.frombuffer(payload, dtype=np.uint8))


def read_numpy_pickle(file_obj, trusted=False):
    raw = np.lib.format.read_array(file_obj, allow_pickle=False)
    return pickle_loads_compatible(raw.tobytes(), trusted=trusted)


def check_fout(instance, fout_path, overwrite=False, logger=None):
    """Resolve an output path (funcs.py:13-57 semantics: folders get a timestamped name,
    existing files get a _N suffix unless overwrite)."""
    import datetime
    if not fout_path:
        return False
    is_folder = fout_path.endswith(("/", "\\")) or os.path.isdir(fout_path)
    if os.path.isfile(fout_path) and not overwrite:
        base, ext = os.path.splitext(fout_path)
        n = 1
        while os.path.exists(f"{base}_{n}{ext}"):
            n += 1
        fout_path = f"{base}_{n}{ext}"
        if logger:
            logger.warning(f"Output file exists; writing '{fout_path}' instead.")
    if is_folder:
        os.makedirs(fout_path, exist_ok=True)
        stamp = datetime.datetime.now().strftime("%Y%m%d%H%M")
        return os.path.join(fout_path, f"Output_{instance.__class__.__name__}_{stamp}.pkl")
    parent = os.path.dirname(fout_path)
    if parent:
        os.makedirs(parent, exist_ok=True)
    return fout_path
