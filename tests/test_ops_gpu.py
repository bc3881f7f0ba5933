"""Individual operations of the backend interface (ope_hip_op_*) against
the oracle's restatement of the CPU backend's operations
(cpu_backend.cpp:514-733)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from octproengine_amd import _capi
    c = _capi.Context(_capi.make_params(signal_length=1024))
    yield c
    c.close()


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.float32])
@pytest.mark.parametrize("shift", [0, 1])
def test_convert_input(ctx, dtype, shift):
    from octproengine_amd import _capi
    raw16 = S.fringes(256, 4, seed=1).reshape(-1)
    raw = {np.uint8: (raw16 >> 8).astype(np.uint8), np.uint16: raw16, np.uint32: raw16.astype(np.uint32) * 65537,
           np.float32: (raw16 / 65535.0).astype(np.float32)}[dtype]
    out = np.zeros(raw.size, np.float32)
    rc = _capi.lib.ope_hip_op_convert_input(ctx._h, raw.ctypes.data, _capi.DTYPE_CODES[np.dtype(dtype)], raw.size,
                                            shift, out.ctypes.data)
    assert rc == 0
    want = O.convert_input(raw, bitshift=bool(shift)).astype(np.float32)
    assert np.array_equal(out, want)


@pytest.mark.parametrize("method", [O.LINEAR, O.CUBIC, O.LANCZOS])
def test_klinearization(ctx, method):
    from octproengine_amd import _capi
    n = 1024
    x = _f(S.fringes(n, 1, seed=2)[0])
    curve = S.resampling_curve(n)
    out = np.zeros(n - 10, np.float32)
    assert _capi.lib.ope_hip_op_klinearization(ctx._h, x.ctypes.data, curve.ctypes.data, method, n, out.size,
                                               out.ctypes.data) == 0
    fn = {O.LINEAR: O.klin_linear, O.CUBIC: O.klin_cubic, O.LANCZOS: O.klin_lanczos}[method]
    want = fn(x.astype(np.float64)[None, :], curve)[0][: out.size]
    err = float(np.max(np.abs(out - want)) / np.max(np.abs(want)))
    print(f"kLinearization method {method}: max rel err {err:.2e}")
    assert err < 2e-6


def test_windowing_and_dispersion(ctx):
    from octproengine_amd import _capi
    n = 512
    x = _f(np.random.default_rng(3).normal(size=n) * 1000)
    w = _f(O.generate_window(n, O.HANN, 0.5, 0.95))
    out = np.zeros(n, np.float32)
    assert _capi.lib.ope_hip_op_windowing(ctx._h, x.ctypes.data, w.ctypes.data, n, n, out.ctypes.data) == 0
    assert np.array_equal(out, x * w)
    phase = _f(O.phase_to_phasor(S.dispersion_phase(n, 5e-6, -2e-9)))
    assert _capi.lib.ope_hip_op_dispersion(ctx._h, x.ctypes.data, phase.ctypes.data, n, n, out.ctypes.data) == 0
    assert np.array_equal(out, x * phase[0::2])


@pytest.mark.parametrize("n", [256, 2048, 4096])
def test_ifft_magnitude(ctx, n):
    from octproengine_amd import _capi
    lines = 5
    x = _f(np.random.default_rng(n).normal(size=(lines, n)))
    out = np.zeros((lines, n), np.float32)
    assert _capi.lib.ope_hip_op_ifft_magnitude(ctx._h, x.ctypes.data, n, lines, out.ctypes.data) == 0
    want = np.abs(np.fft.ifft(x.astype(np.float64), axis=1) * n)
    err = float(np.max(np.abs(out - want)) / np.max(want))
    print(f"ifft magnitude N={n}: max err / peak {err:.2e}")
    assert err < 1e-5
    assert _capi.lib.ope_hip_op_ifft_magnitude(ctx._h, x.ctypes.data, 1000, 1, out.ctypes.data) == _capi.EUNSUP


@pytest.mark.parametrize("lines", [1, 3])
def test_fpn_removal_quirk(ctx, lines):
    """The reference subtracts the mean line from the first min(total/2,
    line_width) values of the flattened lines (cpu_kernels.tpp:461-479)."""
    from octproengine_amd import _capi
    w = 64
    rng = np.random.default_rng(lines)
    z = (rng.normal(size=(lines * w)) + 1j * rng.normal(size=(lines * w))).astype(np.complex64)
    inter = np.stack([z.real, z.imag], axis=-1).astype(np.float32).reshape(-1)
    mean = (rng.normal(size=w) + 1j * rng.normal(size=w)).astype(np.complex64)
    mean_i = np.stack([mean.real, mean.imag], axis=-1).astype(np.float32).reshape(-1)
    out = np.zeros(lines * w, np.float32)
    assert _capi.lib.ope_hip_op_fpn_removal(ctx._h, inter.ctypes.data, mean_i.ctypes.data, w, lines,
                                            out.ctypes.data) == 0
    use = min(lines * w // 2, w)
    want = z.astype(np.complex128)
    want[:use] -= mean[:use]
    assert np.allclose(out, np.abs(want), rtol=2e-7, atol=1e-6)
