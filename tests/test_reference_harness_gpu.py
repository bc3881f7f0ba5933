"""The reference's own harness configurations, run through the drop-in API
(`octproengine.Processor(Backend.CUDA)`, i.e. the HIP backend) and checked
against the FP64 oracle on the curves the Processor generated.

  * the CPU-vs-CUDA comparison of tests/test_backend_output_comparison.cpp:
    18-45: 2048 x 1024 x 1 uint16, cubic resampling with coefficients
    {0.5, 2048, -100, 50}, Hann window with centre 0.9 and fill 0.85 (a
    window clipped at the top end), dispersion {0, 0, 1, -3} factor 1,
    rolling-average background removal (window 64, the default), log map
    0..80, no flip, no post-process background, no FPN;
  * one frame of each size class of tests/test_performance_benchmark.cpp:
    26-56: cubic {0.5, 2048, -100, 0}, Hann 0.5 / 0.95, dispersion
    {0, 0, 1, 2}, background removal, log map 30..100.

The reference harness fills its frames with generateSyntheticAScan
(test_backend_output_comparison.cpp:51-84): three Gaussian peaks on a level
of 2000, two sines with a lateral phase, and rand() % 400 - 200 noise.  It
is restated here with a seeded generator in place of rand(); the seeded
fringe generator of this repository is used beside it.

The grayscale map out = (dB - min) / (max - min) is undone before comparing,
so the bar is the one of every other test: 1e-3 dB within 60 dB of the
A-scan peak, and 1e-4 of the peak on the linear magnitude of every bin.
The reference's own comparison passes at 0.05 of the grayscale range."""

from __future__ import annotations

import threading

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import DB_TOL, LIN_TOL, db_error, db_lin_error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ope():
    from octproengine_amd import octproengine
    return octproengine


def harness_ascans(n: int, ascans: int, bscans: int, seed: int) -> np.ndarray:
    """generateTestData (test_backend_output_comparison.cpp:86-98) with a
    seeded integer noise in place of rand() % 400 - 200."""
    rng = np.random.default_rng(seed)
    i = np.arange(n, dtype=np.float64)
    base = (2000.0 + 8000.0 * np.exp(-(((i - n * 0.2) / 50.0) ** 2))
            + 5000.0 * np.exp(-(((i - n * 0.5) / 30.0) ** 2))
            + 3000.0 * np.exp(-(((i - n * 0.7) / 40.0) ** 2)))
    out = np.empty((bscans * ascans, n), dtype=np.uint16)
    for b in range(bscans):
        for a in range(ascans):
            lat = a * 0.1
            v = base + 1000.0 * np.sin(i * 0.3 + lat) + 500.0 * np.sin(i * 0.15 + lat * 0.5)
            v += rng.integers(0, 400, size=n) - 200
            # static_cast<uint16_t>(clamp(v, 0, 65535)): truncation
            out[b * ascans + a] = np.clip(v, 0.0, 65535.0).astype(np.uint16)
    return out


class Collector:
    def __init__(self):
        self.out = None
        self.done = threading.Event()

    def __call__(self, arr):
        self.out = arr.copy()
        self.done.set()


def run_processor(ope, cfg, raw):
    n, a, b = cfg["n"], cfg["a"], cfg["b"]
    p = ope.Processor(ope.Backend.CUDA)
    p.set_input_parameters(n, a, b, ope.DataType.UINT16)
    p.enable_resampling(True)
    p.set_interpolation_method(ope.InterpolationMethod.CUBIC)
    p.set_resampling_coefficients(cfg["resampling"])
    p.enable_windowing(True)
    p.set_window_parameters(ope.WindowType.HANN, cfg["center"], cfg["fill"])
    p.enable_dispersion_compensation(True)
    p.set_dispersion_coefficients(cfg["dispersion"], 1.0)
    p.enable_background_removal(True)
    p.enable_log_scaling(True)
    p.set_grayscale_range(cfg["gmin"], cfg["gmax"])
    p.enable_bscan_flip(False)
    p.enable_post_process_background_subtraction(False)
    p.enable_fixed_pattern_noise_removal(False)
    p.initialize()
    col = Collector()
    p.set_callback(col)
    buf = p.get_next_available_buffer()
    buf[:] = raw.reshape(-1)
    p.process(buf)
    assert col.done.wait(60), "no output callback"
    c = p.config
    params = O.ChainParams(signal_length=n, ascans_per_bscan=a, bscans_per_buffer=b, resampling=True,
                           interpolation=O.CUBIC, windowing=True, dispersion=True, dc_removal=True,
                           dc_window=c.background_removal.rolling_average_window_size, log_scaling=True,
                           grayscale_min=cfg["gmin"], grayscale_max=cfg["gmax"])
    ref = O.process_frame(raw, params, c.get_generated_resampling_curve(), c.get_generated_window_curve(),
                          O.phase_to_phasor(c.get_generated_dispersion_curve()))
    p.stop()
    return col.out.reshape(a * b, n // 2), ref


def compare(cfg, got, ref, label):
    span = cfg["gmax"] - cfg["gmin"]
    to_db = lambda v: np.asarray(v, np.float64) * span + cfg["gmin"]  # noqa: E731  (coeff 1, addend 0)
    e, el = db_error(to_db(got), to_db(ref)), db_lin_error(to_db(got), to_db(ref))
    gray = float(np.max(np.abs(np.asarray(got, np.float64) - ref)))
    print(f"[{label}] dB err within 60 dB of peak {e:.2e}; peak-normalised linear {el:.2e}; "
          f"max grayscale difference {gray:.2e} (reference tolerance 0.05)")
    assert e <= DB_TOL and el <= LIN_TOL and gray <= 0.05


COMPARISON = dict(n=2048, a=1024, b=1, resampling=[0.5, 2048.0, -100.0, 50.0], center=0.9, fill=0.85,
                  dispersion=[0.0, 0.0, 1.0, -3.0], gmin=0.0, gmax=80.0)


@pytest.mark.parametrize("data", ["harness", "fringes"])
def test_reference_comparison_configuration(ope, data):
    cfg = COMPARISON
    raw = (harness_ascans(cfg["n"], cfg["a"], cfg["b"], seed=51) if data == "harness"
           else S.fringes(cfg["n"], cfg["a"] * cfg["b"], seed=52))
    got, ref = run_processor(ope, cfg, raw)
    compare(cfg, got, ref, f"comparison config, {data}")


BENCH = dict(resampling=[0.5, 2048.0, -100.0, 0.0], center=0.5, fill=0.95, dispersion=[0.0, 0.0, 1.0, 2.0],
             gmin=30.0, gmax=100.0)


@pytest.mark.parametrize("n,a,b", [(512, 256, 8), (1024, 512, 1), (2048, 1024, 8), (4096, 256, 1)])
def test_reference_benchmark_configuration(ope, n, a, b):
    cfg = dict(BENCH, n=n, a=a, b=b)
    raw = harness_ascans(n, a, b, seed=53)
    got, ref = run_processor(ope, cfg, raw)
    compare(cfg, got, ref, f"benchmark config {n}x{a}x{b}")
