"""Host-side API behaviour that needs no GPU: package surface, single-process MPI,
parameter handling, validation errors and result I/O (tests/test_no_mpi.py of the
reference is the model for the first group)."""

import numpy as np
import pytest


def test_package_version_and_imports():
    import pyhermes_amd
    import pyhermes_amd.base  # noqa: F401
    import pyhermes_amd.theory  # noqa: F401
    assert pyhermes_amd.__version__


def test_package_data_present():
    from pathlib import Path

    import pyhermes_amd.base
    import pyhermes_amd.theory
    assert (Path(pyhermes_amd.base.__file__).parent / "default_params.json").is_file()
    assert (Path(pyhermes_amd.theory.__file__).parent / "default_params.json").is_file()


def test_single_process_collectives():
    from pyhermes_amd.utils.mpi_util import MPI
    comm = MPI.COMM_WORLD
    assert comm.Get_rank() == 0 and comm.Get_size() == 1
    assert comm.bcast({"value": 1}, root=0) == {"value": 1}
    assert comm.reduce(3, op=MPI.SUM, root=0) == 3
    assert comm.allreduce(3, op=MPI.MAX) == 3
    assert comm.allgather(3) == [3]
    assert comm.Split_type(MPI.COMM_TYPE_SHARED) is comm
    assert comm.Split(0) is comm
    assert comm.Split(MPI.UNDEFINED) is MPI.COMM_NULL
    send = np.arange(4, dtype=np.float64)
    red, sc = np.empty_like(send), np.empty_like(send)
    comm.Allreduce(send, red, op=MPI.SUM)
    comm.Scatterv([send, [4], [0], MPI.DOUBLE], sc, root=0)
    np.testing.assert_array_equal(red, send)
    np.testing.assert_array_equal(sc, send)


def test_safe_exit_raises_system_exit():
    from pyhermes_amd.utils.func_util import safe_exit
    with pytest.raises(SystemExit) as exc:
        safe_exit(7)
    assert exc.value.code == 7


def test_thick_shell_validation_and_custom_override():
    from pyhermes_amd.io import WindowFunc
    from pyhermes_amd.utils.window_params import normalize_binning_window_template
    sfc = {"J": 3, "box_size": 64.0, "phi_resolution": 1024, "wavelet_mode": "db2", "wavelet_level": 10}

    def custom(ki, kj, kk, R_in, R_out):
        return np.exp(-(ki ** 2 + kj ** 2 + kk ** 2) * R_in * R_out)

    params = {"type": "thick_shell", "func": custom, "len_args": {"R_in": 8.0, "R_out": 12.0}}
    norm = normalize_binning_window_template(params)
    assert norm["len_args"] == {"R_in": 8.0, "R_out": 12.0}
    assert norm["los_args"] == {}
    assert norm["kernel_mode"] == "full_rfft"
    params["kernel_mode"] = "octant"
    w = WindowFunc(params, sfc)
    assert w.type == "thick_shell" and w.has_custom_func and w.func is custom
    assert np.all(np.isfinite(w._host_custom_kernel()))
    with pytest.raises(ValueError, match="requires finite len_args"):
        WindowFunc({"type": "thick_shell", "len_args": {"R_in": 8.0, "R_out": 12.0}}, sfc)
    cyl = WindowFunc({"type": "cylinder", "len_args": {"R": 1.0, "H": 2.0}}, sfc)
    assert cyl.los_args == {"nx": 0.0, "ny": 0.0, "nz": 1.0}  # the default line of sight
    diag = WindowFunc({"type": "ring", "len_args": {"R": 1.0, "H": 2.0}, "los_args": [1.0, 1.0, 0.0]}, sfc)
    assert diag.los_args == {"nx": 1.0, "ny": 1.0, "nz": 0.0}
    from pyhermes_amd.io.window import los_axis, los_desc
    axis, n = los_axis({"nx": 2.0, "ny": -2.0, "nz": 1.0})
    assert axis == 3 and np.allclose(n, [2 / 3, -2 / 3, 1 / 3], rtol=0, atol=1e-15)
    assert los_axis({"nx": 0.0, "ny": -3.0, "nz": 0.0}) == (1, None)
    d = los_desc("disk", {"R": 1.0, "H": 2.0}, {"nx": 1.0, "ny": 1.0, "nz": 0.0})
    assert d.axis == 3 and abs(d.los[0] - 0.5 ** 0.5) < 1e-15 and d.los[2] == 0.0
    with pytest.raises(ValueError, match="non-zero"):
        los_axis({"nx": 0.0, "ny": 0.0, "nz": 0.0})
    with pytest.raises(NotImplementedError, match="non-axis"):
        WindowFunc({"type": "disk", "len_args": {"R": 1.0, "H": 2.0}, "los_args": [0.0, 1.0, 1.0],
                    "kernel_mode": "octant"}, sfc)
    WindowFunc({"type": "disk", "len_args": {"R": 1.0, "H": 2.0}, "los_args": [0.0, 0.0, -1.0],
                "kernel_mode": "octant"}, sfc)  # an axis line of sight folds to the octant exactly
    with pytest.raises(ValueError, match="requires finite len_args"):
        WindowFunc({"type": "disk", "len_args": {"R": 1.0}}, sfc)
    with pytest.raises(NotImplementedError):
        WindowFunc({"type": "cubic", "len_args": {"R": 1.0}}, sfc)


def test_custom_octant_kernel_equals_alias_sum(phi):
    """A custom transfer function evaluated on the host follows the 8-alias octant fold."""
    from oracle import scaling
    from oracle import windows as ow
    from pyhermes_amd.io import WindowFunc
    sfc = {"J": 3, "box_size": 64.0, "phi_resolution": 1024, "wavelet_mode": "db2", "wavelet_level": 10}

    def shell(ki, kj, kk, R):
        q = 2 * np.pi * np.sqrt(ki * ki + kj * kj + kk * kk) * R
        return 1.0 if q == 0 else np.sin(q) / q

    w = WindowFunc({"type": "custom_shell", "func": shell, "len_args": {"R": 16.0}, "kernel_mode": "octant"}, sfc)
    got = w._host_custom_kernel()
    P = scaling.fourier_power_spectrum(phi, 0, 1, 8, 1024)
    ref = ow.octant_rfft_kernel(8, P, "shell", 16.0 * 8 / 64.0)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def _mp_params(**over):
    base = {"sfc_field": "unused", "random": "uniform", "l_min": 0, "l_max": 3,
            "sampling": {"mode": "grid", "r12": {"values": [10.0, 20.0]}, "r13": {"min": 10.0, "max": 30.0, "n": 3}},
            "binning_window12": {"type": "shell", "mapping": {"R": "r12"}},
            "binning_window13": {"type": "shell", "mapping": {"R": "r13"}}}
    base.update(over)
    return {"Corr_3PCF_Multipole": base}


def test_multipole_sampling_grid_and_paired():
    from pyhermes_amd.theory import Corr_3PCF_Multipole
    t = Corr_3PCF_Multipole(_mp_params())
    assert len(t.samples) == 6
    np.testing.assert_array_equal(t.r12, [10, 10, 10, 20, 20, 20])
    np.testing.assert_array_equal(t.r13, [10, 20, 30, 10, 20, 30])
    assert [w["len_args"]["R"] for w in t.binning_window13] == [10, 20, 30, 10, 20, 30]
    assert t.summation_backend == "gpu"
    paired = Corr_3PCF_Multipole(_mp_params(sampling={"mode": "paired", "r12": [5.0, 6.0, 7.0], "r13": 9.0}))
    np.testing.assert_array_equal(paired.r12, [5, 6, 7])
    np.testing.assert_array_equal(paired.r13, [9, 9, 9])
    with pytest.raises(ValueError, match="paired"):
        Corr_3PCF_Multipole(_mp_params(sampling={"mode": "paired", "r12": [5.0, 6.0], "r13": [1.0, 2.0, 3.0]}))


def test_multipole_parameter_errors():
    from pyhermes_amd.theory import Corr_3PCF_Multipole
    with pytest.raises(ValueError, match="top-level radius"):
        Corr_3PCF_Multipole(_mp_params(r12=10.0))
    with pytest.raises(ValueError, match="Unsupported product"):
        Corr_3PCF_Multipole(_mp_params(products=["nope"]))
    with pytest.raises(ValueError, match="l_min"):
        Corr_3PCF_Multipole(_mp_params(l_min=3, l_max=1))
    with pytest.raises(ValueError, match="summation_backend"):
        Corr_3PCF_Multipole(_mp_params(summation_backend="tpu"))
    t = Corr_3PCF_Multipole(_mp_params(summation_backend="cpu"))
    assert t.summation_backend == "gpu_fp64"


def _3pcf_params(**over):
    base = {"sfc_field": "unused", "random": "uniform", "r12": 10.0, "r13": 20.0, "angle_param": "theta",
            "theta": {"theta_min": 0.0, "theta_max": 3.0, "n_theta": 4}, "n_rot": 2, "center": "box_random",
            "n_box_centers": 10, "base_seed": 3}
    base.update(over)
    return {"Corr_3PCF": base}


def test_corr3pcf_products_and_angles():
    from pyhermes_amd.theory import Corr_3PCF
    t = Corr_3PCF(_3pcf_params(products=["q"]))
    assert t.products == ["Q"]
    assert t._expanded_products() == ["Q", "zeta", "zeta_H", "delta_ddd", "rrr", "xi12", "xi13", "xi23"]
    t._resolve_angle_sampling()
    np.testing.assert_allclose(t.mu_arr, np.cos(np.linspace(0, 3.0, 4)))
    tp = Corr_3PCF(_3pcf_params(center="particle", products=["zeta"]))
    assert tp._expanded_products() == ["zeta", "delta_ddd", "rrr", "d_delta_dd", "r_delta_dd"]
    with pytest.raises(ValueError, match="provided together"):
        tp._centre_arrays(np.zeros((3, 3)), None, "particle_pos1", "particle_weight1")
    with pytest.raises(ValueError, match="positive sum"):
        tp._centre_arrays(np.zeros((3, 3)), np.zeros(3), "particle_pos1", "particle_weight1")
    with pytest.raises(ValueError, match="Unsupported product"):
        Corr_3PCF(_3pcf_params(products=["d_delta_dd"]))  # a particle-centre product
    with pytest.raises(ValueError, match="monotonic"):
        t2 = Corr_3PCF(_3pcf_params(angle_param="mu", mu=[0.1, 0.5, 0.2]))
        t2._resolve_angle_sampling()


def test_rotation_draws_follow_seed_rule():
    from oracle import mc as omc
    from pyhermes_amd.theory.corr3pcf import rotation_draws
    for it in (0, 1, 5):
        np.testing.assert_array_equal(rotation_draws(44, 6, it), omc.rotation_uniforms(44, 6, it))


def test_corr2pcf_products_and_sampling():
    from pyhermes_amd.theory import Corr_2PCF
    from pyhermes_amd.theory.corr2pcf import expand_products, normalize_products
    assert expand_products(normalize_products("xi")) == ["xi", "delta_dd", "rr"]
    with pytest.raises(ValueError):
        normalize_products(["xx"])
    t = Corr_2PCF({"Corr_2PCF": {"sfc_field": "unused", "random": "uniform",
                                 "sampling": {"s": {"min": 5.0, "max": 25.0, "n": 5}}}})
    t._resolve_sampling()
    np.testing.assert_allclose(t.sampling_arrays["s"], [5, 10, 15, 20, 25])


def test_cached_defaults_and_phi_are_private_copies():
    """Tasks read default_params.json and sample the scaling function once per process; every task must
    still own its parameters, and every caller its phi samples."""
    from pyhermes_amd.theory import Corr_2PCF
    from pyhermes_amd.utils.wavelet_grid import sample_scaling_function
    spec = {"sampling": {"s": {"min": 5.0, "max": 25.0, "n": 5}}}
    a = Corr_2PCF({"Corr_2PCF": {"sfc_field": "unused", "random": "uniform", **spec}})
    a.task_params["products"] = "mutated"
    a.task_params["sampling"]["s"]["n"] = 99
    b = Corr_2PCF({"Corr_2PCF": {"sfc_field": "unused", "random": "uniform", **spec}})
    assert b.task_params["products"] != "mutated"
    assert b.task_params["sampling"]["s"]["n"] == 5
    p1 = sample_scaling_function()
    ref = p1.copy()
    p1[:] = -1.0
    p2 = sample_scaling_function()
    np.testing.assert_array_equal(p2, ref)
    assert abs(ref.sum() / 1024 - 1.0) < 1e-12  # the samples integrate to one at level 10


@pytest.mark.parametrize("norm", ["raw", "catalog", "field"])
def test_absent_field_value_equals_explicit_ones(norm):
    """No field value means the unit value: the sums and projection weights match explicit ones bit for bit."""
    from pyhermes_amd.base import SFCProjection
    rng = np.random.default_rng(3)
    pos = rng.uniform(0.0, 100.0, (5000, 3))
    w = rng.lognormal(0.0, 0.5, 5000).astype(np.float32)

    def prepared(fv):
        t = SFCProjection({"SFCProjection": {"J": 5, "box_size": 100.0, "particle_pos": pos, "catalog_weight": w,
                                             "field_value": fv, "weight_normalization": norm, "fout_path": ""}})
        t.prepare_input_fields()
        return t

    a, b = prepared(None), prepared(np.ones(5000, dtype=np.float32))
    for key in ("catalog_weight_sum", "catalog_weight_sq_sum", "raw_field_weighted_sum", "field_integral"):
        assert getattr(a, key) == getattr(b, key), key
    assert np.array_equal(a.projection_weight, b.projection_weight)
    assert a.field_value.shape == (5000,) and np.all(a.field_value == 1.0)
    assert a.fin["field_value_key"] is None and b.fin["field_value_key"] == "custom"


def test_result_save_load_roundtrip(tmp_path):
    from pyhermes_amd.io import Corr2PCFData, Corr3PCFData
    d = Corr2PCFData()
    d.set_sampling(["s"], {"s": np.array([1.0, 2.0])})
    d.xi = np.array([0.5, 0.25])
    d.corr2pcf_info = {"products": ["xi"]}
    p = str(tmp_path / "xi.pkl")
    d.save_corr2pcf(p, overwrite=True)
    e = Corr2PCFData()
    e.load_corr2pcf(p)
    np.testing.assert_array_equal(e.xi, d.xi)
    c = Corr3PCFData()
    c.zeta = np.arange(3.0)
    c.mu = np.linspace(-1, 1, 3)
    c.theta = np.arccos(c.mu)
    c.corr3pcf_info = {"center": "box_random"}
    q = str(tmp_path / "z.pkl")
    c.save_corr3pcf(q, overwrite=True)
    f = Corr3PCFData()
    f.load_corr3pcf(q)
    np.testing.assert_array_equal(f.zeta, c.zeta)
    np.testing.assert_array_equal(f.mu, c.mu)


def test_corr2pcf_anisotropic_sampling_and_descriptors():
    """(s, mu) and (rp, pi) mappings (corr2pcf.py:267-322): sampling axes, grid order and the descriptors of
    the fused pass, built on the host."""
    from pyhermes_amd import kernels
    from pyhermes_amd.theory import Corr_2PCF
    t = Corr_2PCF({"Corr_2PCF": {"sfc_field": "unused", "random": "uniform", "binning_window": "ring",
                                 "sampling": {"mu": [0.0, 0.5, 1.0], "s": {"min": 10.0, "max": 20.0, "n": 2}}}})
    t._resolve_sampling()
    assert t.sampling_names == ("s", "mu") and t.sampling_shape == (2, 3)
    descs, ells = t._descs(64, 128.0)  # grid units = 0.5 x box units
    assert ells == [] and len(descs) == 6
    want = [(s * np.sqrt(1 - m * m) / 2, s * m / 2) for s in (10.0, 20.0) for m in (0.0, 0.5, 1.0)]
    for d, (R, H) in zip(descs, want):
        assert d.type == kernels.WIN_CODES["ring"] and d.axis == 2
        assert d.a0 == pytest.approx(R, abs=1e-12) and d.a1 == pytest.approx(H, abs=1e-12)
    t = Corr_2PCF({"Corr_2PCF": {"sfc_field": "unused", "random": "uniform",
                                 "binning_window": {"type": "cylinder", "mapping": "rppi_to_RH",
                                                    "los_args": {"nx": 0.0, "ny": -1.0, "nz": 0.0}},
                                 "sampling": {"rp": [5.0, 10.0], "pi": [0.0, 8.0]}}})
    t._resolve_sampling()
    descs, _ = t._descs(64, 64.0)
    assert [(d.a0, d.a1, d.axis) for d in descs] == [(5.0, 0.0, 1), (5.0, 8.0, 1), (10.0, 0.0, 1), (10.0, 8.0, 1)]
    with pytest.raises(ValueError, match="must match mapping"):
        Corr_2PCF({"Corr_2PCF": {"sfc_field": "unused", "random": "uniform", "binning_window": "ring",
                                 "sampling": {"s": [10.0]}}})._resolve_sampling()
    with pytest.raises(ValueError, match="non-negative"):
        Corr_2PCF({"Corr_2PCF": {"sfc_field": "unused", "random": "uniform", "binning_window": "disk",
                                 "sampling": {"s": [-1.0, 2.0], "mu": [0.5]}}})._resolve_sampling()
    with pytest.raises(ValueError, match="multipoles"):
        Corr_2PCF({"Corr_2PCF": {"sfc_field": "unused", "random": "uniform", "binning_window": "ring",
                                 "multipoles": [0, 2], "sampling": {"s": [10.0], "mu": [0.5]}}})._resolve_sampling()


def test_counting_defaults_and_validation():
    """Counting parameters (theory/default_params.json 'Counting', counting.py:27-37)."""
    from pyhermes_amd.theory import Counting
    t = Counting({"Counting": {"sfc_field": "unused"}})
    assert t.random_count == 1000000 and t.seed == 42 and t.window is None
    assert t.weight_normalization == "catalog"
    with pytest.raises(ValueError, match="non-negative"):
        Counting({"Counting": {"sfc_field": "unused", "random_count": -1}})


def test_window_parameter_schema_keys_defaults_and_errors():
    """The window-parameter schema keeps the reference's keys, defaults, error types and messages
    (window_params.py): built-in binning strings, type clean-up, unset lengths, the z line of sight of the
    cylindrical windows, kernel-mode defaults and validation, len_args / los_args coercion."""
    from pyhermes_amd.utils import window_params as wp
    assert wp.normalize_binning_window_template(None) == {
        "type": "shell", "len_args": {"R": None}, "los_args": {}, "other_args": {}, "mapping": "s_to_R",
        "kernel_mode": "octant"}
    ring = wp.normalize_binning_window_template(" Ring ")
    assert ring == {"type": "ring", "len_args": {"R": None, "H": None}, "los_args": {"nx": 0.0, "ny": 0.0, "nz": 1.0},
                    "other_args": {}, "mapping": "smu_to_RH", "kernel_mode": "auto"}
    thick = wp.normalize_binning_window_template({"type": "THICK_SHELL", "len_args": ["R"]})
    assert thick["len_args"] == {"R": None, "delta_R": None} and thick["mapping"] == "s_to_R"
    # an untyped dict becomes a shell without the shell defaults; a func makes it custom
    bare = wp.normalize_binning_window_template({"len_args": {"R": 5.0}})
    assert bare["type"] == "shell" and "mapping" not in bare and bare["kernel_mode"] == "octant"
    f = wp.normalize_binning_window_template({"func": lambda k, **kw: k})
    assert f["type"] == "custom" and f["kernel_mode"] == "full_rfft" and f["los_args"] == {}
    assert wp.normalize_binning_window_template({"type": "disk", "los_args": [1, 0, 0]})["los_args"] == {
        "nx": 1.0, "ny": 0.0, "nz": 0.0}
    with pytest.raises(ValueError, match="Supported built-in strings are 'shell', 'thick_shell', 'ring', 'disk', "
                                         "'cylinder', and 'cylshell'"):
        wp.normalize_binning_window_template("sphere")
    with pytest.raises(TypeError, match="expected dict, string, or None"):
        wp.normalize_binning_window_template(3)
    with pytest.raises(TypeError, match="'func' must be callable"):
        wp.normalize_binning_window_template({"type": "shell", "func": 1})
    with pytest.raises(ValueError, match=r"Unsupported kernel_mode 'bogus'. Supported values are \['auto', "
                                         r"'complex_full_fft', 'complex_rfft', 'full_rfft', 'octant'\]"):
        wp.normalize_binning_window_template({"type": "shell", "kernel_mode": " BOGUS"})
    assert wp.default_kernel_mode("radial_multipole") == "complex_full_fft"
    assert wp.default_kernel_mode("directional_derivative") == "complex_rfft"
    assert wp.default_kernel_mode("sphere") == "octant" and wp.default_kernel_mode("ring", True) == "full_rfft"
    assert wp.normalize_len_args("R") == {"R": None} and wp.normalize_len_args(None) == {}
    with pytest.raises(TypeError, match="len_args entries must be strings."):
        wp.normalize_len_args(["R", 2])
    with pytest.raises(TypeError, match="len_args must be a dict, string, list, tuple, or None."):
        wp.normalize_len_args(4.0)
    with pytest.raises(ValueError, match=r"exactly three values: \[nx, ny, nz\]"):
        wp.normalize_los_args([0, 1])
    with pytest.raises(ValueError, match="must define nx, ny, and nz together"):
        wp.normalize_los_args({"nx": 1.0})
    with pytest.raises(ValueError, match="vector must be non-zero"):
        wp.normalize_los_args({"nx": 0, "ny": 0.0, "nz": -0.0})
    with pytest.raises(TypeError, match="los_args must be a dict, length-3 array, or None."):
        wp.normalize_los_args("z")
    assert wp.normalize_los_args({}, "sphere") == {}
    ser = wp.serialize_window_params({"func": np.sin, "len_args": {"R": np.float64(2.0)}, "v": np.arange(2)})
    assert ser == {"custom_func": {"kind": "callable", "name": "numpy.sin"}, "len_args": {"R": 2.0}, "v": [0, 1]}
