"""Host-side checks that need no GPU: the C ABI library loads and exports
everything include/ope_hip.h declares, fails cleanly without a device, and
the C++ ope:: API (through the `octproengine` Python module) reproduces the
reference's configuration semantics and its generated curves bit for bit."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_c_abi_exports_every_declared_symbol():
    from octproengine_amd import _capi
    header = (ROOT / "include" / "ope_hip.h").read_text()
    declared = set(re.findall(r"OPE_HIP_API\s+[\w\s\*]+?\b(ope_hip_\w+)\s*\(", header))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(str(_capi.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name
    # the ctypes binding covers exactly the declared API
    assert declared == set(_capi.SIGNATURES)


def test_c_abi_errors_without_device_or_with_bad_arguments():
    from octproengine_amd import _capi
    lib = _capi.lib
    assert lib.ope_hip_version().startswith(b"octproengine_amd")
    n = ctypes.c_int(-1)
    assert lib.ope_hip_device_count(ctypes.byref(n)) == 0 and n.value >= 0
    # invalid parameters are rejected before any device is touched
    h = ctypes.c_void_p()
    bad = _capi.make_params(signal_length=1000)
    rc = lib.ope_hip_create(0, ctypes.byref(bad), 2, ctypes.byref(h))
    assert rc == _capi.EUNSUP and not h.value
    assert b"1000" in lib.ope_hip_last_error(None)
    bad = _capi.make_params(signal_length=2048, input_type=4)  # UINT64
    assert lib.ope_hip_create(0, ctypes.byref(bad), 2, ctypes.byref(h)) == _capi.EUNSUP
    bad = _capi.make_params(signal_length=2048, fpn_bscan_count=0)
    assert lib.ope_hip_create(0, ctypes.byref(bad), 2, ctypes.byref(h)) == _capi.EINVAL
    ok = _capi.make_params(signal_length=2048)
    assert lib.ope_hip_create(0, ctypes.byref(ok), 9, ctypes.byref(h)) == _capi.EINVAL
    assert lib.ope_hip_create(0, ctypes.byref(ok), 2, None) == _capi.EINVAL
    if _capi.device_count() == 0:
        rc = lib.ope_hip_create(0, ctypes.byref(ok), 2, ctypes.byref(h))
        assert rc == _capi.ERUNTIME and not h.value
    for fn in ("ope_hip_destroy", "ope_hip_request_fpn", "ope_hip_synchronize"):
        assert getattr(lib, fn)(None) in (_capi.OK, _capi.EINVAL)


@pytest.fixture(scope="module")
def ope():
    from octproengine_amd import octproengine
    return octproengine


def test_module_surface_matches_reference_names(ope):
    for name in ("Processor", "ProcessorConfiguration", "DataParameters", "ResamplingParameters",
                 "WindowingParameters", "DispersionCompensationParameters", "BackgroundRemovalParameters",
                 "PostProcessingParameters", "Backend", "DataType", "InterpolationMethod", "WindowType",
                 "InitializationError", "ConfigurationError", "BufferError", "ProcessingError", "BackendError"):
        assert hasattr(ope, name), name
    assert ope.__version__ == "0.0.1"
    methods = ("initialize stop load_config save_config set_config set_input_parameters set_backend "
               "get_backend set_callback add_output_callback remove_output_callback clear_output_callbacks "
               "get_callback_count process get_next_available_buffer set_resampling_coefficients "
               "set_custom_resampling_curve use_custom_resampling_curve enable_resampling "
               "set_interpolation_method set_dispersion_coefficients set_custom_dispersion_curve "
               "use_custom_dispersion_curve enable_dispersion_compensation set_window_parameters "
               "set_custom_window_curve use_custom_window_curve enable_windowing set_grayscale_range "
               "set_signal_multiplicator_and_addend enable_log_scaling enable_background_removal "
               "set_background_removal_window_size enable_bscan_flip enable_sinusoidal_scan_correction "
               "enable_fixed_pattern_noise_removal enable_post_process_background_subtraction "
               "request_post_process_background_recording set_post_process_background_weight "
               "set_post_process_background_offset has_post_process_background_profile "
               "get_post_process_background_profile_size get_post_process_background_profile "
               "set_post_process_background_profile save_post_process_background_profile_to_file "
               "load_post_process_background_profile_from_file request_fixed_pattern_noise_determination "
               "enable_continuous_fixed_pattern_noise_determination set_fixed_pattern_noise_bscan_count "
               "has_fixed_pattern_noise_profile get_fixed_pattern_noise_profile_size "
               "get_fixed_pattern_noise_profile set_fixed_pattern_noise_profile "
               "save_fixed_pattern_noise_profile_to_file load_fixed_pattern_noise_profile_from_file")
    for m in methods.split():
        assert hasattr(ope.Processor, m), m


def test_configuration_defaults_match_reference(ope):
    c = ope.ProcessorConfiguration()
    # include/processorconfiguration.h:33-172
    assert (c.data.signal_length, c.data.samples_per_buffer, c.data.ascans_per_bscan,
            c.data.bscans_per_buffer) == (1024, 1024 * 512, 512, 1)
    assert c.data.input_data_type == ope.DataType.UINT16 and c.data.get_bit_depth() == 16
    assert not c.resampling.enabled and list(c.resampling.coefficients) == [0, 1, 0, 0]
    assert c.resampling.interpolation_method == ope.InterpolationMethod.LINEAR
    assert c.windowing.window_type == ope.WindowType.HANN
    assert np.isclose(c.windowing.window_center_position, 0.5) and np.isclose(c.windowing.window_fill_factor, 0.95)
    assert c.background_removal.rolling_average_window_size == 64
    pp = c.post_processing
    assert pp.log_scaling and pp.grayscale_min == 30 and pp.grayscale_max == 100
    assert pp.multiplicator == 1 and pp.addend == 0 and pp.fixed_pattern_noise_bscan_count == 1
    assert c.validate()
    c.data.signal_length = 0
    assert not c.validate()


def test_generated_curves_bit_exact_with_reference(ope):
    F = np.load(ROOT / "tests" / "golden" / "reference_curves.npz")
    c = ope.ProcessorConfiguration()
    for n in F["lengths"]:
        c.data.signal_length = int(n)
        for i, co in enumerate(F["params_resample"]):
            c.resampling.coefficients = [float(v) for v in co]
            assert np.array_equal(c.get_generated_resampling_curve(), F[f"resample_n{n}_{i}"])
        for i, d in enumerate(F["params_dispersion"]):
            c.dispersion.coefficients = [float(v) for v in d]
            assert np.array_equal(c.get_generated_dispersion_curve(), F[f"dispersion_n{n}_{i}"])
        for i, (t, ce, f) in enumerate(F["params_window"]):
            c.windowing.window_type = ope.WindowType(int(t))
            c.windowing.window_center_position = float(ce)
            c.windowing.window_fill_factor = float(f)
            assert np.array_equal(c.get_generated_window_curve(), F[f"window_n{n}_{i}"]), (n, t, ce, f)


def test_custom_curves_pad_and_truncate(ope):
    # processorconfiguration.cpp:268-334
    c = ope.ProcessorConfiguration()
    c.data.signal_length = 8
    c.set_custom_resampling_curve(np.arange(5, dtype=np.float32))
    assert list(c.get_custom_resampling_curve()) == [0, 1, 2, 3, 4, 0, 0, 0]
    c.set_custom_window_curve(np.arange(12, dtype=np.float32))
    assert list(c.get_custom_window_curve()) == list(range(8))
    c.data.signal_length = 10
    c.adjust_all_custom_curves()
    assert list(c.get_custom_window_curve()) == list(range(10))  # re-fitted from the original
    assert c.has_custom_resampling_curve() and not c.has_custom_dispersion_curve()
    with pytest.raises(ValueError):
        c.set_custom_dispersion_curve(np.zeros(0, np.float32))


def test_processor_state_before_initialize(ope):
    p = ope.Processor(ope.Backend.CUDA)
    assert p.get_backend() == ope.Backend.CUDA and "CUDA" in repr(p)
    p.set_input_parameters(2048, 1024, 16, ope.DataType.UINT16)
    d = p.config.data
    assert (d.signal_length, d.ascans_per_bscan, d.bscans_per_buffer, d.samples_per_buffer) == (2048, 1024, 16, 2048 * 1024 * 16)
    p.enable_resampling(True)
    p.set_interpolation_method(ope.InterpolationMethod.CUBIC)
    p.set_resampling_coefficients([0.5, 2048.0, -100.0, 50.0])
    p.set_window_parameters(ope.WindowType.HANN, 0.9, 0.85)
    p.set_dispersion_coefficients([0.0, 0.0, 1.0, -3.0], 2.0)
    p.set_grayscale_range(0.0, 80.0)
    c = p.config
    assert c.resampling.enabled and c.resampling.interpolation_method == ope.InterpolationMethod.CUBIC
    assert list(c.resampling.coefficients) == [0.5, 2048.0, -100.0, 50.0]
    assert np.isclose(c.windowing.window_center_position, 0.9) and c.dispersion.factor == 2.0
    assert c.post_processing.grayscale_max == 80.0
    with pytest.raises(RuntimeError):
        p.use_custom_window_curve(True)  # no custom curve yet
    p.set_custom_window_curve(np.ones(2048, np.float32))
    assert p.config.windowing.use_custom_curve
    with pytest.raises(ValueError):
        p.set_fixed_pattern_noise_bscan_count(0)
    # FPN/background requests before initialize are accepted (deferred)
    p.request_fixed_pattern_noise_determination()
    p.request_post_process_background_recording()
    assert not p.has_fixed_pattern_noise_profile()


def test_cpu_backend_is_not_part_of_this_library(ope):
    with pytest.raises(ope.BackendError):
        ope.Processor(ope.Backend.CPU)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 (or a distributed launch)
    exits with status 2 and says why, before touching a device."""
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    for args, env in ((["--gpus", "2"], {}), ([], {"WORLD_SIZE": "4"})):
        import os
        r = subprocess.run([sys.executable, str(root / "bench.py"), *args], capture_output=True, text=True,
                           env={**os.environ, **env}, timeout=120)
        assert r.returncode == 2
        assert "one GPU" in r.stderr
        assert r.stdout == ""
