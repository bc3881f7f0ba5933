"""End-to-end through the drop-in API: `octproengine.Processor(Backend.CUDA)`
(the HIP backend) driven the way the reference's own tests drive it
(python/tests/test_backend_comparison.py:169-245, test_multi_consumer.py):
configure with setters, initialize, acquire a buffer, fill it, process, wait
for the callback.  Outputs are checked against the FP64 oracle on the same
float32 curves the Processor generated."""

from __future__ import annotations

import threading

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import DB_TOL, db_error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ope():
    from octproengine_amd import octproengine
    return octproengine


class Collector:
    def __init__(self, n_expected=1):
        self.outputs = []
        self.n = n_expected
        self.done = threading.Event()
        self.lock = threading.Lock()

    def __call__(self, arr):
        with self.lock:
            self.outputs.append(arr.copy())
            if len(self.outputs) >= self.n:
                self.done.set()


def configure(ope, p, n, a, b, interp=None, dc=False):
    p.set_input_parameters(n, a, b, ope.DataType.UINT16)
    p.enable_resampling(interp is not None)
    if interp is not None:
        p.set_interpolation_method(interp)
        p.set_resampling_coefficients([0.5, float(n), -100.0, 50.0])
    p.enable_windowing(True)
    p.set_window_parameters(ope.WindowType.HANN, 0.5, 0.95)
    p.enable_dispersion_compensation(True)
    p.set_dispersion_coefficients([0.0, 0.0, 1.0, -3.0], 1.0)
    p.enable_background_removal(dc)
    p.enable_log_scaling(True)
    p.set_grayscale_range(0.0, 1.0)  # output in dB


def oracle_for(ope, p, raw, interp, dc=False):
    c = p.config
    n = c.data.signal_length
    params = O.ChainParams(signal_length=n, ascans_per_bscan=c.data.ascans_per_bscan,
                           bscans_per_buffer=c.data.bscans_per_buffer, resampling=interp is not None,
                           interpolation=int(interp) if interp is not None else 0, windowing=True,
                           dispersion=True, dc_removal=dc, dc_window=c.background_removal.rolling_average_window_size,
                           grayscale_min=0.0, grayscale_max=1.0)
    curve = c.get_generated_resampling_curve()
    window = c.get_generated_window_curve()
    phasor = O.phase_to_phasor(c.get_generated_dispersion_curve())
    return O.process_frame(raw, params, curve, window, phasor)


# the Processor tunes its launch at initialisation: LINEAR + DC takes the
# banded DC path in whichever variant the tuner picks
@pytest.mark.parametrize("interp,dc", [("LINEAR", False), ("LINEAR", True), ("CUBIC", True), ("LANCZOS", False)])
def test_processor_single_frame_matches_oracle(ope, interp, dc):
    n, a, b = 2048, 256, 2
    method = getattr(ope.InterpolationMethod, interp)
    p = ope.Processor(ope.Backend.CUDA)
    configure(ope, p, n, a, b, method, dc=dc)
    p.initialize()
    col = Collector()
    p.set_callback(col)
    raw = S.fringes(n, a * b, seed=7)
    buf = p.get_next_available_buffer()
    assert buf.dtype == np.uint16 and buf.size == n * a * b
    buf[:] = raw.reshape(-1)
    p.process(buf)
    assert col.done.wait(30), "no output callback"
    out = col.outputs[0].reshape(a * b, n // 2)
    ref = oracle_for(ope, p, raw, method, dc=dc)
    e = db_error(out, ref)
    print(f"Processor {interp} dc={dc}: dB error near peak {e:.2e}")
    assert e <= DB_TOL
    p.stop()


def test_processor_stream_of_frames_and_multi_consumer(ope):
    """Frames through the 2-buffer pool and both output slots; three
    consumers each see every frame (one thread per callback, as
    test_multi_consumer.cpp).  An output buffer is valid until its slot is
    reused two frames later, so frames are submitted in pairs and each pair
    is consumed before the next."""
    n, a, b, frames = 1024, 128, 4, 12
    p = ope.Processor(ope.Backend.CUDA)
    configure(ope, p, n, a, b, ope.InterpolationMethod.CUBIC)
    p.initialize()
    raws = [S.fringes(n, a * b, seed=100 + i) for i in range(frames)]
    refs = [oracle_for(ope, p, r, ope.InterpolationMethod.CUBIC) for r in raws]
    cols = [Collector(frames) for _ in range(3)]
    ids = [p.add_output_callback(c) for c in cols]
    assert p.get_callback_count() == 3

    def consumed(k):
        return all(len(c.outputs) >= k for c in cols)

    for i in range(0, frames, 2):
        for r in raws[i:i + 2]:
            buf = p.get_next_available_buffer()
            buf[:] = r.reshape(-1)
            p.process(buf)
        for _ in range(600):
            if consumed(i + 2):
                break
            threading.Event().wait(0.05)
        assert consumed(i + 2), f"frames {i}, {i + 1} not delivered"
    for c in cols:
        assert len(c.outputs) == frames
        errs = [db_error(o.reshape(a * b, n // 2), ref) for o, ref in zip(c.outputs, refs)]
        assert max(errs) <= DB_TOL, errs
    assert p.remove_output_callback(ids[0]) and not p.remove_output_callback(12345)
    p.clear_output_callbacks()
    assert p.get_callback_count() == 0
    p.stop()


def test_processor_fpn_and_background_api(ope, tmp_path):
    n, a, b = 1024, 64, 2
    p = ope.Processor(ope.Backend.CUDA)
    configure(ope, p, n, a, b, ope.InterpolationMethod.CUBIC, dc=True)
    p.enable_fixed_pattern_noise_removal(True)
    p.request_fixed_pattern_noise_determination()  # before initialize, as the reference test does
    p.initialize()
    col = Collector(2)
    p.set_callback(col)
    for i in range(2):
        buf = p.get_next_available_buffer()
        buf[:] = S.fringes(n, a * b, seed=300 + i).reshape(-1)
        p.process(buf)
    assert col.done.wait(30)
    assert p.has_fixed_pattern_noise_profile()
    prof = p.get_fixed_pattern_noise_profile()
    assert prof.size == n and p.get_fixed_pattern_noise_profile_size() == n // 2
    f = tmp_path / "fpn.csv"
    p.save_fixed_pattern_noise_profile_to_file(str(f))
    assert f.read_text().splitlines()[0] == "Sample Number;Real;Imag"
    p.set_fixed_pattern_noise_profile(np.zeros(n, np.float32))
    p.load_fixed_pattern_noise_profile_from_file(str(f))
    back = p.get_fixed_pattern_noise_profile()
    assert np.allclose(back, prof, rtol=1e-5, atol=1e-3)  # CSV keeps 6 significant digits
    # post-process background profile
    bg = np.linspace(0, 1, n // 2).astype(np.float32)
    p.set_post_process_background_profile(bg)
    assert p.has_post_process_background_profile()
    assert np.array_equal(p.get_post_process_background_profile(), bg)
    g = tmp_path / "bg.csv"
    p.save_post_process_background_profile_to_file(str(g))
    assert g.read_text().splitlines()[0] == "Sample Number;Sample Value"
    p.load_post_process_background_profile_from_file(str(g))
    assert np.allclose(p.get_post_process_background_profile(), bg, atol=1e-6)
    with pytest.raises(ValueError):
        p.set_post_process_background_profile(np.zeros(7, np.float32))  # must be N/2
    p.stop()


def test_processor_reinitialize_on_dimension_change(ope):
    p = ope.Processor(ope.Backend.CUDA)
    configure(ope, p, 512, 64, 1, ope.InterpolationMethod.LINEAR)
    p.initialize()
    col = Collector(2)
    p.set_callback(col)
    buf = p.get_next_available_buffer()
    buf[:] = S.fringes(512, 64, seed=1).reshape(-1)
    p.process(buf)
    # new dimensions: the next process() re-initializes (processor.cpp:127-133)
    p.set_input_parameters(1024, 32, 1, ope.DataType.UINT16)
    p.set_resampling_coefficients([0.5, 1024.0, -100.0, 50.0])
    p.initialize()
    raw = S.fringes(1024, 32, seed=2)
    buf = p.get_next_available_buffer()
    assert buf.size == 1024 * 32
    buf[:] = raw.reshape(-1)
    p.process(buf)
    assert col.done.wait(30)
    out = col.outputs[-1]
    assert out.size == 32 * 512
    ref = oracle_for(ope, p, raw, ope.InterpolationMethod.LINEAR)
    assert db_error(out.reshape(32, 512), ref) <= DB_TOL
    p.stop()


def test_processor_min_variance_mean_operation(ope):
    p = ope.Processor(ope.Backend.CUDA)
    configure(ope, p, 256, 16, 1, ope.InterpolationMethod.LINEAR)
    rng = np.random.default_rng(3)
    h, w = 64, 40
    rows = (rng.normal(size=(h, w)) + 1j * rng.normal(size=(h, w))).astype(np.complex64)
    inter = np.stack([rows.real, rows.imag], axis=-1).astype(np.float32).reshape(-1)
    got = p.get_minimum_variance_mean(inter, w, h, 8)
    assert np.array_equal(got, O.min_variance_mean(rows, 8))
    p.stop()
