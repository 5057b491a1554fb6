"""On-disk format shared with PyHermes data products: a pickled dict (protocol 4)
stored as the uint8 payload of a .npy array (the layout of pickle_compat.py:81-90),
with path-like metadata written as strings.  Only load files you trust: the
payload is a pickle."""

import os
import pickle

import numpy as np


def _portable(value):
    if isinstance(value, os.PathLike):
        return os.fsdecode(os.fspath(value))
    if isinstance(value, dict):
        return {_portable(k): _portable(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return type(value)(_portable(v) for v in value)
    return value


def write_numpy_pickle(file_obj, value, protocol=4):
    payload = pickle.dumps(_portable(value), protocol=protocol)
    np.lib.format.write_array(file_obj, np.frombuffer(payload, dtype=np.uint8))


def read_numpy_pickle(file_obj):
    raw = np.lib.format.read_array(file_obj, allow_pickle=False)
    return pickle.loads(raw.tobytes())


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
