"""The NumPy-framed pickle format of SFCField and result files (pickle_compat.py:81-90), its path handling
(pickle_compat.py:29-58, the cases of the reference's tests/test_pickle_compat.py) and the data-only reader."""

import io
import os
import pickle
import sys
import types
from pathlib import Path, PurePosixPath, PureWindowsPath

import numpy as np
import pytest

from pyhermes_amd.io.formats import pickle_loads_compatible, read_numpy_pickle, write_numpy_pickle


def _python313_path_payload(name, parts):
    """Pickle of a Python 3.13 path object: its class lives in the private module pathlib._local,
    which this Python does not have (a stand-in module supplies it only while pickling)."""
    module_name = "pathlib._local"
    module = types.ModuleType(module_name)
    cls = type(name, (object,), {"__module__": module_name, "__reduce__": lambda self: (cls, parts)})
    setattr(module, name, cls)
    saved = sys.modules.get(module_name)
    sys.modules[module_name] = module
    try:
        return pickle.dumps(cls(), protocol=4)
    finally:
        if saved is None:
            sys.modules.pop(module_name, None)
        else:
            sys.modules[module_name] = saved


def test_python313_private_pathlib_pickle_loads_as_string():
    assert "pathlib._local" not in sys.modules
    assert pickle_loads_compatible(_python313_path_payload("PosixPath", ("/", "tmp", "field.pkl"))) == \
        "/tmp/field.pkl"
    assert pickle_loads_compatible(_python313_path_payload("WindowsPath", ("C:\\", "data", "f.pkl"))) == \
        "C:\\data\\f.pkl"


def test_paths_are_strings_on_write_and_on_read():
    dataset = {"sfc_info": {"fin": {"path": Path("/tmp/catalog.npz"), "aux": [PurePosixPath("/a/b")],
                                    "pair": (Path("x"), 1)}},
               "epsilon": np.arange(4, dtype=np.float64)}
    buf = io.BytesIO()
    write_numpy_pickle(buf, dataset)
    buf.seek(0)
    loaded = read_numpy_pickle(buf)
    assert loaded["sfc_info"]["fin"]["path"] == "/tmp/catalog.npz"
    assert loaded["sfc_info"]["fin"]["aux"] == ["/a/b"]
    assert loaded["sfc_info"]["fin"]["pair"] == ("x", 1)
    np.testing.assert_array_equal(loaded["epsilon"], dataset["epsilon"])
    # a payload pickled with live path objects (another writer) also reads back as strings
    raw = pickle.dumps({"p": Path("/tmp/a.npz"), "w": PureWindowsPath("C:/x/y")}, protocol=4)
    assert pickle_loads_compatible(raw) == {"p": "/tmp/a.npz", "w": "C:\\x\\y"}


def test_numpy_framing_layout():
    """The file is a .npy array of uint8 whose bytes are the protocol-4 pickle of the dict."""
    buf = io.BytesIO()
    write_numpy_pickle(buf, {"a": 1})
    buf.seek(0)
    arr = np.lib.format.read_array(buf, allow_pickle=False)
    assert arr.dtype == np.uint8 and arr.ndim == 1
    assert pickle.loads(arr.tobytes()) == {"a": 1}
    assert arr.tobytes()[:2] == b"\x80\x04"


@pytest.mark.parametrize("protocol", [2, 4, 5])
def test_data_types_round_trip(protocol):
    value = {"f64": np.linspace(0, 1, 7), "c128": np.arange(3) * (1 + 2j), "f32_2d": np.ones((2, 3), np.float32).T,
             "i8": np.arange(5, dtype=np.int8), "bools": np.array([True, False]), "text": np.array(["ab", "c"]),
             "rec": np.zeros(2, dtype=[("a", "f8"), ("b", "i4")]), "scalars": [np.float64(1.5), np.int64(-3),
                                                                            np.complex64(1j), np.bool_(True)],
             "py": {"c": 1 + 2j, "s": {1, 2}, "fs": frozenset({3}), "b": b"\x00\xff", "none": None,
                    "t": (1, "x")}, "dtype": np.dtype("<c16")}
    got = pickle_loads_compatible(pickle.dumps(value, protocol=protocol))
    for k in ("f64", "c128", "f32_2d", "i8", "bools", "text", "rec"):
        assert got[k].dtype == value[k].dtype
        np.testing.assert_array_equal(got[k], value[k])
    assert got["scalars"] == value["scalars"] and [type(v) for v in got["scalars"]] == \
        [type(v) for v in value["scalars"]]
    assert got["py"] == value["py"] and got["dtype"] == value["dtype"]


def test_reader_refuses_code_unless_trusted(tmp_path):
    marker = tmp_path / "ran"

    class Payload:  # a pickle that would call os.mkdir(marker) when loaded
        def __reduce__(self):
            return (os.mkdir, (str(marker),))

    raw = pickle.dumps({"epsilon": Payload()}, protocol=4)
    with pytest.raises(pickle.UnpicklingError, match="posix.mkdir|os.mkdir"):
        pickle_loads_compatible(raw)
    assert not marker.exists()
    pickle_loads_compatible(raw, trusted=True)  # the unrestricted pickle semantics, on request
    assert marker.exists()


def test_data_containers_use_the_format(tmp_path):
    from pyhermes_amd.io import Corr2PCFData
    d = Corr2PCFData()
    d.set_sampling(["s"], {"s": np.array([1.0, 2.0])})
    d.xi = np.array([0.5, 0.25])
    d.corr2pcf_info = {"sampling": {"s": [1.0, 2.0]}, "products": ["xi"], "fin": Path("/tmp/cat.npz")}
    path = d.save_corr2pcf(str(tmp_path / "xi.pkl"))
    with open(path, "rb") as f:
        raw = read_numpy_pickle(f)
    np.testing.assert_array_equal(raw["xi"], [0.5, 0.25])
    assert raw["corr2pcf_info"]["fin"] == "/tmp/cat.npz"
    back = Corr2PCFData().load_corr2pcf(path)
    np.testing.assert_array_equal(back.xi, d.xi)
