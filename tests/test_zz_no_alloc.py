"""The allocation-free pair reduction (hermes_pair_reduce_noalloc) against hermes_pair_reduce, through the C ABI.

Both run the same kernels in the same order, the first with its scratch in the library's fixed device array, so
their results agree bit for bit.  Two calls in a row, the second on a smaller input, check that the array's
leftover contents from the first call do not leak into the second; a request beyond the array's size is refused
before anything is enqueued.
"""

import ctypes

import numpy as np
import pytest

from oracle import scaling

pytestmark = pytest.mark.gpu


def _spectra(J, dbl, seed):
    import torch
    from pyhermes_amd import kernels
    L = 1 << J
    rng = np.random.default_rng(seed)
    f1 = rng.normal(size=(L, L, L))
    f2 = f1 + 0.4 * rng.normal(size=(L, L, L))
    td = torch.float64 if dbl else torch.float32
    return kernels.rfftn(kernels.as_device(f1, td)), kernels.rfftn(kernels.as_device(f2, td))


def _call(name, J, dbl, F1, F2, Pd, descs):
    from pyhermes_amd import _lib
    arr = (_lib.WindowDesc * len(descs))(*descs)
    out = np.full(len(descs), np.nan)
    _lib.call(name, _lib.context(), J, 1 if dbl else 0, F1.data_ptr(), F2.data_ptr(), Pd.data_ptr(), arr,
              len(descs), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def _case(phi, J, dbl, descs, seed):
    import torch
    from pyhermes_amd import kernels
    L = 1 << J
    Pd = kernels.as_device(scaling.fourier_power_spectrum(phi, 0, 1, L, 1024), torch.float64)
    F1, F2 = _spectra(J, dbl, seed)
    want = _call("hermes_pair_reduce", J, dbl, F1, F2, Pd, descs)
    got = _call("hermes_pair_reduce_noalloc", J, dbl, F1, F2, Pd, descs)
    return got, want


@pytest.mark.parametrize("dbl", [True, False])
def test_noalloc_radial_pass_twice_then_smaller(phi, dbl):
    """Thin shells and even-l multipoles (the radius-grouped pass): J=7 with 20 shells x l = 0, 2, 4, then
    J=5 with 3 shells; each equal bit for bit to hermes_pair_reduce."""
    from pyhermes_amd import kernels
    print(f"noalloc radial pass dbl={dbl}: J=7 60 windows, then J=5 3 windows", flush=True)
    big = [kernels.window_desc("shell_multipole", R, l=l, m=0) if l else kernels.window_desc("shell", R)
           for R in np.linspace(1.5, 18.0, 20) for l in (0, 2, 4)]
    got, want = _case(phi, 7, dbl, big, 11)
    assert np.all(np.isfinite(want))
    np.testing.assert_array_equal(got, want)
    small = [kernels.window_desc("shell", R) for R in (1.25, 2.5, 4.0)]
    got, want = _case(phi, 5, dbl, small, 12)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("dbl", [True, False])
def test_noalloc_generic_pass_twice_then_smaller(phi, dbl):
    """Other window types (the generic pass, descriptors copied into the scratch): J=6 with rings, spheres and
    gaussian shells, then J=4 with two windows."""
    from pyhermes_amd import kernels
    print(f"noalloc generic pass dbl={dbl}: J=6 24 windows, then J=4 2 windows", flush=True)
    big = [kernels.window_desc("ring", R, H, axis=2) for R in (2.0, 4.0, 6.0, 8.0) for H in (0.5, 2.0, 4.0)]
    big += [kernels.window_desc("sphere", R) for R in (1.5, 3.0, 5.0, 7.0, 9.0, 11.0)]
    big += [kernels.window_desc("gaussian_shell", R, 1.0) for R in (2.0, 4.0, 6.0, 8.0, 10.0, 12.0)]
    got, want = _case(phi, 6, dbl, big, 21)
    assert np.all(np.isfinite(want))
    np.testing.assert_array_equal(got, want)
    small = [kernels.window_desc("gaussian", 1.5), kernels.window_desc("thick_shell", 3.0, 1.0)]
    got, want = _case(phi, 4, dbl, small, 22)
    np.testing.assert_array_equal(got, want)


def test_noalloc_rejects_requests_beyond_its_scratch(phi):
    """1200 distinct shells at J=7 need about 10 MB of partial sums, beyond the fixed 8 MiB: status 2 with a
    message, checked on the host before any launch; the ordinary entry point still serves the request."""
    import torch
    from pyhermes_amd import _lib, kernels
    print("noalloc size limit: J=7 1200 shells", flush=True)
    J, L = 7, 128
    Pd = kernels.as_device(scaling.fourier_power_spectrum(phi, 0, 1, L, 1024), torch.float64)
    F1, F2 = _spectra(J, True, 31)
    descs = [kernels.window_desc("shell", R) for R in np.linspace(1.0, 40.0, 1200)]
    arr = (_lib.WindowDesc * len(descs))(*descs)
    out = np.zeros(len(descs))
    lib = _lib.load()
    rc = lib.hermes_pair_reduce_noalloc(_lib.context(), J, 1, F1.data_ptr(), F2.data_ptr(), Pd.data_ptr(), arr,
                                        len(descs), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 2
    assert b"scratch" in lib.hermes_last_error()
    ok = _call("hermes_pair_reduce", J, True, F1, F2, Pd, descs[:8])
    again = _call("hermes_pair_reduce_noalloc", J, True, F1, F2, Pd, descs[:8])
    np.testing.assert_array_equal(again, ok)
